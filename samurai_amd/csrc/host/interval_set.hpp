// Interval-compressed cell sets of one level (host side of the MR mesh).
//
// A LevelSet stores the cells of one level as rows keyed by (z, y) -- y only in
// 2D -- each holding sorted, disjoint, non-adjacent half-open x runs [s, e).
// This is the canonical form of samurai's LevelCellArray
// (level_cell_array.hpp:68-256, add_interval_back merging at :465-522), so
// comparing two sets is comparing arrays.  The set operations implement the
// meaning of samurai's lazy subset expressions (subset/*.hpp): n-ary ops work
// at the finest operand level, .on(L) refines or coarsens, expand is the box
// dilation, translate shifts.
#pragma once

#include <array>
#include <cstddef>
#include <cstdint>
#include <vector>

namespace smr
{
    struct Run
    {
        int32_t s;
        int32_t e;
    };

    struct LevelSet
    {
        int dim   = 2;
        int level = 0;
        std::vector<int32_t> ry;  // row y
        std::vector<int32_t> rz;  // row z (0 in 2D)
        std::vector<int64_t> rp;  // row -> first run, size nrows+1
        std::vector<Run> runs;

        LevelSet() { rp.push_back(0); }
        LevelSet(int d, int l)
            : dim(d)
            , level(l)
        {
            rp.push_back(0);
        }

        size_t nrows() const { return ry.size(); }
        // room for `rows` rows and `n` runs, so that pushing them does not
        // reallocate
        void reserve(size_t rows, size_t n)
        {
            ry.reserve(rows);
            rz.reserve(rows);
            rp.reserve(rows + 1);
            runs.reserve(n);
        }
        bool empty() const { return runs.empty(); }
        int64_t count() const;
        bool operator==(const LevelSet& o) const;
        bool operator!=(const LevelSet& o) const { return !(*this == o); }
        void clear()
        {
            ry.clear();
            rz.clear();
            rp.assign(1, 0);
            runs.clear();
        }

        // appends a run; rows must come in (z, y) order and runs in x order
        // (a run overlapping or touching the row's last run extends it)
        void push(int32_t y, int32_t z, int32_t s, int32_t e)
        {
            if (e <= s)
            {
                return;
            }
            if (!ry.empty() && ry.back() == y && rz.back() == z)
            {
                Run& b = runs.back();
                if (s <= b.e)
                {
                    if (e > b.e)
                    {
                        b.e = e;
                    }
                    return;
                }
                runs.push_back({s, e});
                rp.back() = int64_t(runs.size());
            }
            else
            {
                ry.push_back(y);
                rz.push_back(z);
                runs.push_back({s, e});
                rp.push_back(int64_t(runs.size()));
            }
        }

        // row index of (y, z) or -1 (binary search)
        int64_t find_row(int32_t y, int32_t z) const;
        // run index containing x in row r, or -1
        int64_t find_run(int64_t r, int32_t x) const;
        bool contains(int32_t x, int32_t y, int32_t z = 0) const;

        template <class F>
        void for_each_run(F&& f) const
        {
            for (size_t r = 0; r < ry.size(); ++r)
            {
                for (int64_t k = rp[r]; k < rp[r + 1]; ++k)
                {
                    f(runs[size_t(k)].s, runs[size_t(k)].e, ry[r], rz[r]);
                }
            }
        }
    };

    inline bool row_less(int32_t y0, int32_t z0, int32_t y1, int32_t z1)
    {
        return z0 < z1 || (z0 == z1 && y0 < y1);
    }

    LevelSet set_union(const LevelSet& a, const LevelSet& b);
    LevelSet set_intersection(const LevelSet& a, const LevelSet& b);
    LevelSet set_difference(const LevelSet& a, const LevelSet& b);
    LevelSet on(const LevelSet& a, int level);
    LevelSet translate(const LevelSet& a, const std::array<int, 3>& t);
    LevelSet expand(const LevelSet& a, int w);
    // on(expand(a, w), level) in one pass
    LevelSet expand_on(const LevelSet& a, int w, int level);
    LevelSet box(int dim, int level, const std::array<int, 3>& lo, const std::array<int, 3>& hi);
    // a n box(lo, hi) and a \ box(lo, hi) in one pass over a (z bounds ignored in 2D)
    LevelSet clip_box(const LevelSet& a, const std::array<int, 3>& lo, const std::array<int, 3>& hi);
    LevelSet minus_box(const LevelSet& a, const std::array<int, 3>& lo, const std::array<int, 3>& hi);

    // (y, z, s, e) quadruples in any order -> canonical set
    LevelSet from_runs(int dim, int level, std::vector<std::array<int32_t, 4>>& q);

    inline LevelSet set_union(const LevelSet& a, const LevelSet& b, const LevelSet& c)
    {
        return set_union(set_union(a, b), c);
    }
    inline LevelSet set_intersection(const LevelSet& a, const LevelSet& b, const LevelSet& c)
    {
        return set_intersection(set_intersection(a, b), c);
    }
} // namespace smr
