#include "interval_set.hpp"

#include <algorithm>
#include <stdexcept>

namespace smr
{
    int64_t LevelSet::count() const
    {
        int64_t n = 0;
        for (const auto& r : runs)
        {
            n += r.e - r.s;
        }
        return n;
    }

    bool LevelSet::operator==(const LevelSet& o) const
    {
        if (runs.size() != o.runs.size() || ry.size() != o.ry.size())
        {
            return false;
        }
        if (!runs.empty() && level != o.level)
        {
            return false;
        }
        if (ry != o.ry || rz != o.rz || rp != o.rp)
        {
            return false;
        }
        for (size_t k = 0; k < runs.size(); ++k)
        {
            if (runs[k].s != o.runs[k].s || runs[k].e != o.runs[k].e)
            {
                return false;
            }
        }
        return true;
    }

    int64_t LevelSet::find_row(int32_t y, int32_t z) const
    {
        // lookups come in row order more often than not (a boundary walk, the
        // sources of consecutive recipes): try the calling thread's last row of
        // this set and the one after it first.  A candidate is accepted only if
        // it holds (y, z), the one row that does, so a stale memo never misleads.
        thread_local const LevelSet* memo_set = nullptr;
        thread_local int64_t memo_row         = -1;
        const int64_t nr                      = int64_t(ry.size());
        if (memo_set == this)
        {
            for (int64_t c = memo_row; c <= memo_row + 1; ++c)
            {
                if (c >= 0 && c < nr && ry[size_t(c)] == y && rz[size_t(c)] == z)
                {
                    memo_row = c;
                    return c;
                }
            }
        }
        int64_t lo = 0, hi = nr;
        while (lo < hi)
        {
            int64_t mid = (lo + hi) / 2;
            if (row_less(ry[size_t(mid)], rz[size_t(mid)], y, z))
            {
                lo = mid + 1;
            }
            else
            {
                hi = mid;
            }
        }
        if (lo < nr && ry[size_t(lo)] == y && rz[size_t(lo)] == z)
        {
            memo_set = this;
            memo_row = lo;
            return lo;
        }
        return -1;
    }

    int64_t LevelSet::find_run(int64_t r, int32_t x) const
    {
        int64_t lo = rp[size_t(r)], hi = rp[size_t(r) + 1];
        // last run with s <= x
        while (lo < hi)
        {
            int64_t mid = (lo + hi) / 2;
            if (runs[size_t(mid)].s <= x)
            {
                lo = mid + 1;
            }
            else
            {
                hi = mid;
            }
        }
        int64_t k = lo - 1;
        if (k >= rp[size_t(r)] && x < runs[size_t(k)].e)
        {
            return k;
        }
        return -1;
    }

    bool LevelSet::contains(int32_t x, int32_t y, int32_t z) const
    {
        int64_t r = find_row(y, z);
        return r >= 0 && find_run(r, x) >= 0;
    }

    namespace
    {
        enum class Op
        {
            Union,
            Intersection,
            Difference
        };

        void union_runs(const Run* a, size_t na, const Run* b, size_t nb, int32_t y, int32_t z, LevelSet& out)
        {
            size_t p = 0, q = 0;
            bool have = false;
            Run cur{0, 0};
            while (p < na || q < nb)
            {
                Run r;
                if (q >= nb || (p < na && a[p].s <= b[q].s))
                {
                    r = a[p++];
                }
                else
                {
                    r = b[q++];
                }
                if (!have)
                {
                    cur  = r;
                    have = true;
                }
                else if (r.s <= cur.e)
                {
                    cur.e = std::max(cur.e, r.e);
                }
                else
                {
                    out.push(y, z, cur.s, cur.e);
                    cur = r;
                }
            }
            if (have)
            {
                out.push(y, z, cur.s, cur.e);
            }
        }

        void intersect_runs(const Run* a, size_t na, const Run* b, size_t nb, int32_t y, int32_t z, LevelSet& out)
        {
            size_t p = 0, q = 0;
            while (p < na && q < nb)
            {
                int32_t lo = std::max(a[p].s, b[q].s);
                int32_t hi = std::min(a[p].e, b[q].e);
                if (lo < hi)
                {
                    out.push(y, z, lo, hi);
                }
                if (a[p].e < b[q].e)
                {
                    ++p;
                }
                else
                {
                    ++q;
                }
            }
        }

        void difference_runs(const Run* a, size_t na, const Run* b, size_t nb, int32_t y, int32_t z, LevelSet& out)
        {
            size_t q = 0;
            for (size_t p = 0; p < na; ++p)
            {
                int32_t s = a[p].s;
                const int32_t e = a[p].e;
                while (q < nb && b[q].e <= s)
                {
                    ++q;
                }
                for (size_t k = q; k < nb && b[k].s < e; ++k)
                {
                    if (b[k].s > s)
                    {
                        out.push(y, z, s, b[k].s);
                    }
                    s = std::max(s, b[k].e);
                    if (s >= e)
                    {
                        break;
                    }
                }
                if (s < e)
                {
                    out.push(y, z, s, e);
                }
            }
        }

        void copy_row(const LevelSet& a, size_t r, LevelSet& out)
        {
            for (int64_t k = a.rp[r]; k < a.rp[r + 1]; ++k)
            {
                out.push(a.ry[r], a.rz[r], a.runs[size_t(k)].s, a.runs[size_t(k)].e);
            }
        }

        LevelSet combine(const LevelSet& A, const LevelSet& B, Op op)
        {
            const int L = std::max(A.level, B.level);
            LevelSet ta, tb;
            const LevelSet* a = &A;
            const LevelSet* b = &B;
            if (A.level < L)
            {
                ta = on(A, L);
                a  = &ta;
            }
            if (B.level < L)
            {
                tb = on(B, L);
                b  = &tb;
            }
            LevelSet out(A.dim, L);
            size_t i = 0, j = 0;
            const size_t na = a->nrows(), nb = b->nrows();
            if (op == Op::Intersection)
            {
                out.reserve(std::min(na, nb), a->runs.size() + b->runs.size());
            }
            else
            {
                out.reserve(op == Op::Union ? na + nb : na, a->runs.size() + b->runs.size());
            }
            while (i < na || j < nb)
            {
                int which;
                if (i < na && j < nb)
                {
                    if (row_less(a->ry[i], a->rz[i], b->ry[j], b->rz[j]))
                    {
                        which = 1;
                    }
                    else if (row_less(b->ry[j], b->rz[j], a->ry[i], a->rz[i]))
                    {
                        which = 2;
                    }
                    else
                    {
                        which = 3;
                    }
                }
                else
                {
                    which = i < na ? 1 : 2;
                }
                if (which == 1)
                {
                    if (op != Op::Intersection)
                    {
                        copy_row(*a, i, out);
                    }
                    ++i;
                }
                else if (which == 2)
                {
                    if (op == Op::Union)
                    {
                        copy_row(*b, j, out);
                    }
                    ++j;
                }
                else
                {
                    const Run* ra  = a->runs.data() + a->rp[i];
                    const size_t la = size_t(a->rp[i + 1] - a->rp[i]);
                    const Run* rb  = b->runs.data() + b->rp[j];
                    const size_t lb = size_t(b->rp[j + 1] - b->rp[j]);
                    if (op == Op::Union)
                    {
                        union_runs(ra, la, rb, lb, a->ry[i], a->rz[i], out);
                    }
                    else if (op == Op::Intersection)
                    {
                        intersect_runs(ra, la, rb, lb, a->ry[i], a->rz[i], out);
                    }
                    else
                    {
                        difference_runs(ra, la, rb, lb, a->ry[i], a->rz[i], out);
                    }
                    ++i;
                    ++j;
                }
            }
            return out;
        }

        inline int32_t floor_shift(int32_t v, int k)
        {
            return v >> k; // arithmetic shift: floor division by 2^k
        }
    } // namespace

    LevelSet set_union(const LevelSet& a, const LevelSet& b)
    {
        return combine(a, b, Op::Union);
    }
    LevelSet set_intersection(const LevelSet& a, const LevelSet& b)
    {
        return combine(a, b, Op::Intersection);
    }
    LevelSet set_difference(const LevelSet& a, const LevelSet& b)
    {
        return combine(a, b, Op::Difference);
    }

    LevelSet from_runs(int dim, int level, std::vector<std::array<int32_t, 4>>& q)
    {
        if (q.empty())
        {
            return LevelSet(dim, level);
        }
        // Rows are dense integers: bucket the runs by row (counting sort), then
        // order each row's few runs by start.  Falls back to a comparison sort
        // when the row range is much larger than the number of runs.
        int32_t ymin = q[0][0], ymax = q[0][0], zmin = q[0][1], zmax = q[0][1];
        for (const auto& r : q)
        {
            ymin = std::min(ymin, r[0]);
            ymax = std::max(ymax, r[0]);
            zmin = std::min(zmin, r[1]);
            zmax = std::max(zmax, r[1]);
        }
        const int64_t ny = int64_t(ymax) - ymin + 1, nz = int64_t(zmax) - zmin + 1;
        const int64_t nrows = ny * nz;
        if (nrows <= 4 * int64_t(q.size()) + 4096)
        {
            std::vector<int64_t> start(size_t(nrows) + 1, 0);
            for (const auto& r : q)
            {
                start[size_t((int64_t(r[1]) - zmin) * ny + (r[0] - ymin)) + 1]++;
            }
            for (size_t k = 1; k < start.size(); ++k)
            {
                start[k] += start[k - 1];
            }
            std::vector<std::array<int32_t, 2>> byrow(q.size());
            std::vector<int64_t> fill(start.begin(), start.end() - 1);
            for (const auto& r : q)
            {
                const size_t row = size_t((int64_t(r[1]) - zmin) * ny + (r[0] - ymin));
                byrow[size_t(fill[row]++)] = {r[2], r[3]};
            }
            LevelSet out(dim, level);
            out.ry.reserve(size_t(nrows));
            out.runs.reserve(q.size());
            for (int64_t row = 0; row < nrows; ++row)
            {
                auto b = byrow.begin() + start[size_t(row)], e = byrow.begin() + start[size_t(row) + 1];
                if (b == e)
                {
                    continue;
                }
                if (!std::is_sorted(b, e))
                {
                    std::sort(b, e);
                }
                const int32_t y = int32_t(ymin + row % ny), z = int32_t(zmin + row / ny);
                for (auto it = b; it != e; ++it)
                {
                    out.push(y, z, (*it)[0], (*it)[1]);
                }
            }
            return out;
        }
        std::sort(q.begin(),
                  q.end(),
                  [](const std::array<int32_t, 4>& u, const std::array<int32_t, 4>& v)
                  {
                      if (u[1] != v[1])
                      {
                          return u[1] < v[1];
                      }
                      if (u[0] != v[0])
                      {
                          return u[0] < v[0];
                      }
                      return u[2] < v[2];
                  });
        LevelSet out(dim, level);
        for (const auto& r : q)
        {
            out.push(r[0], r[1], r[2], r[3]);
        }
        return out;
    }

    LevelSet on(const LevelSet& a, int level)
    {
        if (level == a.level)
        {
            return a;
        }
        LevelSet out(a.dim, level);
        if (level > a.level)
        {
            const int k     = level - a.level;
            const int32_t f = int32_t(1) << k;
            const int32_t nz = a.dim == 3 ? f : 1;
            out.reserve(a.nrows() * size_t(f) * size_t(nz), a.runs.size() * size_t(f) * size_t(nz));
            size_t r0        = 0;
            while (r0 < a.nrows())
            {
                size_t r1 = r0;
                while (r1 < a.nrows() && a.rz[r1] == a.rz[r0])
                {
                    ++r1;
                }
                for (int32_t dz = 0; dz < nz; ++dz)
                {
                    int32_t z = a.dim == 3 ? a.rz[r0] * f + dz : 0;
                    for (size_t r = r0; r < r1; ++r)
                    {
                        for (int32_t dy = 0; dy < f; ++dy)
                        {
                            for (int64_t k2 = a.rp[r]; k2 < a.rp[r + 1]; ++k2)
                            {
                                out.push(a.ry[r] * f + dy, z, a.runs[size_t(k2)].s * f, a.runs[size_t(k2)].e * f);
                            }
                        }
                    }
                }
                r0 = r1;
            }
            return out;
        }
        const int k = a.level - level;
        if (a.dim == 2)
        {
            // the fine rows of one coarse row are consecutive: merge their
            // coarsened runs (each row's already sorted by start) row group by
            // row group
            std::vector<std::array<int32_t, 2>> buf;
            size_t r0 = 0;
            while (r0 < a.nrows())
            {
                const int32_t y = floor_shift(a.ry[r0], k);
                size_t r1       = r0 + 1;
                while (r1 < a.nrows() && floor_shift(a.ry[r1], k) == y)
                {
                    ++r1;
                }
                buf.clear();
                for (size_t r = r0; r < r1; ++r)
                {
                    const size_t mid = buf.size();
                    for (int64_t q = a.rp[r]; q < a.rp[r + 1]; ++q)
                    {
                        const Run& run = a.runs[size_t(q)];
                        buf.push_back({floor_shift(run.s, k), floor_shift(run.e - 1, k) + 1});
                    }
                    if (mid > 0)
                    {
                        std::inplace_merge(buf.begin(), buf.begin() + std::ptrdiff_t(mid), buf.end());
                    }
                }
                for (const auto& b : buf)
                {
                    out.push(y, 0, b[0], b[1]);
                }
                r0 = r1;
            }
            return out;
        }
        std::vector<std::array<int32_t, 4>> q;
        q.reserve(a.runs.size());
        a.for_each_run(
            [&](int32_t s, int32_t e, int32_t y, int32_t z)
            {
                q.push_back({floor_shift(y, k), a.dim == 3 ? floor_shift(z, k) : 0, floor_shift(s, k), floor_shift(e - 1, k) + 1});
            });
        return from_runs(a.dim, level, q);
    }

    LevelSet translate(const LevelSet& a, const std::array<int, 3>& t)
    {
        LevelSet out = a;
        for (auto& y : out.ry)
        {
            y += t[1];
        }
        if (a.dim == 3)
        {
            for (auto& z : out.rz)
            {
                z += t[2];
            }
        }
        for (auto& r : out.runs)
        {
            r.s += t[0];
            r.e += t[0];
        }
        return out;
    }

    namespace
    {
        // 2D: on(expand(a, w), a.level - k) for k >= 0 in one sweep over the
        // rows.  Coarse row Y collects the input rows y with
        // Y 2^k - w <= y <= Y 2^k + 2^k - 1 + w (consecutive in a); each maps
        // its runs to [floor((s - w) / 2^k), floor((e - 1 + w) / 2^k) + 1),
        // sorted per row, so a merge of the window's rows gives the row.
        LevelSet expand_on_2d(const LevelSet& a, int w, int k)
        {
            LevelSet out(2, a.level - k);
            const size_t n = a.nrows();
            if (n == 0)
            {
                return out;
            }
            const int64_t f = int64_t(1) << k;
            std::vector<std::array<int32_t, 2>> buf;
            size_t lo = 0, hi = 0;
            int64_t y = floor_shift(int32_t(a.ry[0] - w), k);
            for (;;)
            {
                const int64_t ylo = y * f - w, yhi = y * f + f - 1 + w;
                while (hi < n && a.ry[hi] <= yhi)
                {
                    ++hi;
                }
                while (lo < hi && a.ry[lo] < ylo)
                {
                    ++lo;
                }
                if (lo == hi)
                {
                    if (hi == n)
                    {
                        break;
                    }
                    y = floor_shift(int32_t(a.ry[hi] - w), k);
                    continue;
                }
                buf.clear();
                for (size_t r = lo; r < hi; ++r)
                {
                    const size_t mid = buf.size();
                    for (int64_t q = a.rp[r]; q < a.rp[r + 1]; ++q)
                    {
                        const Run& run = a.runs[size_t(q)];
                        buf.push_back({floor_shift(run.s - w, k), floor_shift(run.e - 1 + w, k) + 1});
                    }
                    if (mid > 0)
                    {
                        std::inplace_merge(buf.begin(), buf.begin() + std::ptrdiff_t(mid), buf.end());
                    }
                }
                for (const auto& b : buf)
                {
                    out.push(int32_t(y), 0, b[0], b[1]);
                }
                ++y;
            }
            return out;
        }
    } // namespace

    LevelSet expand(const LevelSet& a, int w)
    {
        if (w == 0 || a.empty())
        {
            return a;
        }
        if (a.dim == 2)
        {
            return expand_on_2d(a, w, 0);
        }
        std::vector<std::array<int32_t, 4>> q;
        const int wz = a.dim == 3 ? w : 0;
        q.reserve(a.runs.size() * size_t(2 * w + 1) * size_t(2 * wz + 1));
        a.for_each_run(
            [&](int32_t s, int32_t e, int32_t y, int32_t z)
            {
                for (int dz = -wz; dz <= wz; ++dz)
                {
                    for (int dy = -w; dy <= w; ++dy)
                    {
                        q.push_back({y + dy, z + dz, s - w, e + w});
                    }
                }
            });
        return from_runs(a.dim, a.level, q);
    }

    LevelSet expand_on(const LevelSet& a, int w, int level)
    {
        const int k = a.level - level;
        if (k < 0)
        {
            return on(expand(a, w), level);
        }
        if (a.dim == 2)
        {
            return expand_on_2d(a, w, k);
        }
        // the box dilation of a run projects onto a box of coarse cells
        std::vector<std::array<int32_t, 4>> q;
        const int wz = a.dim == 3 ? w : 0;
        q.reserve(a.runs.size() * 3);
        a.for_each_run(
            [&](int32_t s, int32_t e, int32_t y, int32_t z)
            {
                const int32_t x0 = floor_shift(s - w, k), x1 = floor_shift(e - 1 + w, k) + 1;
                const int32_t y0 = floor_shift(y - w, k), y1 = floor_shift(y + w, k);
                const int32_t z0 = a.dim == 3 ? floor_shift(z - wz, k) : 0, z1 = a.dim == 3 ? floor_shift(z + wz, k) : 0;
                for (int32_t cz = z0; cz <= z1; ++cz)
                {
                    for (int32_t cy = y0; cy <= y1; ++cy)
                    {
                        q.push_back({cy, cz, x0, x1});
                    }
                }
            });
        return from_runs(a.dim, level, q);
    }

    namespace
    {
        bool row_in_box(const LevelSet& a, size_t r, const std::array<int, 3>& lo, const std::array<int, 3>& hi)
        {
            return a.ry[r] >= lo[1] && a.ry[r] < hi[1] && (a.dim == 2 || (a.rz[r] >= lo[2] && a.rz[r] < hi[2]));
        }
    } // namespace

    LevelSet clip_box(const LevelSet& a, const std::array<int, 3>& lo, const std::array<int, 3>& hi)
    {
        LevelSet out(a.dim, a.level);
        const int32_t zlo = a.dim == 3 ? int32_t(lo[2]) : 0, zhi = a.dim == 3 ? int32_t(hi[2]) : 1;
        if (lo[0] >= hi[0] || lo[1] >= hi[1] || zlo >= zhi)
        {
            return out;
        }
        // first row at or after (y, z) in (z, y) order
        auto first_row = [&](size_t from, int32_t y, int32_t z)
        {
            size_t l = from, h = a.nrows();
            while (l < h)
            {
                const size_t mid = (l + h) / 2;
                if (row_less(a.ry[mid], a.rz[mid], y, z))
                {
                    l = mid + 1;
                }
                else
                {
                    h = mid;
                }
            }
            return l;
        };
        // the rows of the box: one range of rows per plane, found by bisection
        size_t r = first_row(0, int32_t(lo[1]), zlo);
        while (r < a.nrows() && a.rz[r] < zhi)
        {
            if (a.ry[r] < lo[1])
            {
                r = first_row(r, int32_t(lo[1]), a.rz[r]);
                continue;
            }
            if (a.ry[r] >= hi[1])
            {
                r = first_row(r, int32_t(lo[1]), a.rz[r] + 1);
                continue;
            }
            // the runs of the row that meet [lo[0], hi[0]): the first run
            // ending after lo[0] by bisection, then until a run starts at hi[0]
            int64_t k = a.rp[r], ke = a.rp[r + 1];
            int64_t kh = ke;
            while (k < kh)
            {
                const int64_t mid = (k + kh) / 2;
                if (a.runs[size_t(mid)].e <= lo[0])
                {
                    k = mid + 1;
                }
                else
                {
                    kh = mid;
                }
            }
            for (; k < ke && a.runs[size_t(k)].s < hi[0]; ++k)
            {
                const Run& run = a.runs[size_t(k)];
                out.push(a.ry[r], a.rz[r], std::max(run.s, int32_t(lo[0])), std::min(run.e, int32_t(hi[0])));
            }
            ++r;
        }
        return out;
    }

    LevelSet minus_box(const LevelSet& a, const std::array<int, 3>& lo, const std::array<int, 3>& hi)
    {
        LevelSet out(a.dim, a.level);
        out.reserve(a.nrows(), 2 * a.runs.size());
        for (size_t r = 0; r < a.nrows(); ++r)
        {
            const bool in = row_in_box(a, r, lo, hi);
            for (int64_t k = a.rp[r]; k < a.rp[r + 1]; ++k)
            {
                const Run& run = a.runs[size_t(k)];
                if (!in)
                {
                    out.push(a.ry[r], a.rz[r], run.s, run.e);
                    continue;
                }
                out.push(a.ry[r], a.rz[r], run.s, std::min(run.e, int32_t(lo[0])));
                out.push(a.ry[r], a.rz[r], std::max(run.s, int32_t(hi[0])), run.e);
            }
        }
        return out;
    }

    LevelSet box(int dim, int level, const std::array<int, 3>& lo, const std::array<int, 3>& hi)
    {
        LevelSet out(dim, level);
        const int z0 = dim == 3 ? lo[2] : 0, z1 = dim == 3 ? hi[2] : 1;
        for (int z = z0; z < z1; ++z)
        {
            for (int y = lo[1]; y < hi[1]; ++y)
            {
                out.push(y, z, lo[0], hi[0]);
            }
        }
        return out;
    }
} // namespace smr
