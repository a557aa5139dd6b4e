// Oracle (test infrastructure only): dense-grid restatement of samurai's
// subset algebra.  Only tests/, __graft_entry__.smoke() and bench.py's CPU
// baseline leg may use anything under oracle/; the product never links it.
//
// samurai evaluates set expressions lazily over interval traversers
// (include/samurai/subset/*).  Their meaning is pure set algebra on cells of a
// level:
//   * an n-ary union/intersection/difference works at the max level of its
//     operands, coarser operands being refined (subset/nary_set_operator.hpp:
//     102-117, utils::powMinus2 on the row index and interval << shift);
//   * .on(L) refines (all children) or coarsens (a coarse cell is kept when any
//     fine cell projects onto it; subset/projection.hpp:129-177,
//     traversers/last_dim_projection_traverser.hpp:74-100);
//   * translate(s, d) shifts (subset/translation.hpp);
//   * expand(s, w) is the box dilation by w in every direction (native expand,
//     samurai_config.hpp:55; subset/expansion.hpp:139-193 and
//     traversers/last_dim_expansion_traverser.hpp:44-60).
// Here every set is a dense bitmap over a bounding box at its level, which
// makes each of those rules a few lines and independent of the interval
// representation used by the product.
#pragma once

#include <algorithm>
#include <array>
#include <cstdint>
#include <functional>
#include <stdexcept>
#include <vector>

namespace oracle
{
    struct DSet
    {
        int dim   = 2;
        int level = 0;
        std::array<int, 3> lo{0, 0, 0};
        std::array<int, 3> hi{0, 0, 0};
        std::vector<uint8_t> b;

        DSet() = default;
        DSet(int d, int l)
            : dim(d)
            , level(l)
        {
        }

        bool box_empty() const
        {
            for (int d = 0; d < 3; ++d)
            {
                if (hi[d] <= lo[d])
                {
                    return true;
                }
            }
            return false;
        }
        int64_t nx() const { return hi[0] - lo[0]; }
        int64_t ny() const { return hi[1] - lo[1]; }
        int64_t nz() const { return hi[2] - lo[2]; }
        int64_t volume() const { return box_empty() ? 0 : nx() * ny() * nz(); }

        void alloc(const std::array<int, 3>& l, const std::array<int, 3>& h)
        {
            lo = l;
            hi = h;
            if (dim == 2)
            {
                lo[2] = 0;
                hi[2] = 1;
            }
            b.assign(static_cast<size_t>(volume()), 0);
        }
        int64_t off(int x, int y, int z) const
        {
            return ((int64_t(z) - lo[2]) * ny() + (int64_t(y) - lo[1])) * nx() + (int64_t(x) - lo[0]);
        }
        bool inside(int x, int y, int z) const
        {
            return x >= lo[0] && x < hi[0] && y >= lo[1] && y < hi[1] && z >= lo[2] && z < hi[2];
        }
        bool get(int x, int y, int z = 0) const
        {
            return !box_empty() && inside(x, y, z) && b[size_t(off(x, y, z))];
        }
        void set(int x, int y, int z, bool v = true)
        {
            if (!inside(x, y, z))
            {
                throw std::logic_error("DSet::set outside the bounding box");
            }
            b[size_t(off(x, y, z))] = v ? 1 : 0;
        }
        bool empty() const
        {
            for (auto v : b)
            {
                if (v)
                {
                    return false;
                }
            }
            return true;
        }
        int64_t count() const
        {
            int64_t n = 0;
            for (auto v : b)
            {
                n += v;
            }
            return n;
        }
        // Visits the cells in samurai's traversal order: last dimension outermost,
        // x innermost (subset/apply.hpp:12-46).
        template <class F>
        void for_each(F&& f) const
        {
            if (box_empty())
            {
                return;
            }
            for (int z = lo[2]; z < hi[2]; ++z)
            {
                for (int y = lo[1]; y < hi[1]; ++y)
                {
                    const uint8_t* row = &b[size_t(off(lo[0], y, z))];
                    for (int x = lo[0]; x < hi[0]; ++x)
                    {
                        if (row[x - lo[0]])
                        {
                            f(x, y, z);
                        }
                    }
                }
            }
        }
    };

    inline DSet tighten(const DSet& s)
    {
        DSet out(s.dim, s.level);
        if (s.box_empty())
        {
            return out;
        }
        std::array<int, 3> l{s.hi[0], s.hi[1], s.hi[2]};
        std::array<int, 3> h{s.lo[0], s.lo[1], s.lo[2]};
        bool any = false;
        s.for_each(
            [&](int x, int y, int z)
            {
                any  = true;
                l[0] = std::min(l[0], x);
                l[1] = std::min(l[1], y);
                l[2] = std::min(l[2], z);
                h[0] = std::max(h[0], x + 1);
                h[1] = std::max(h[1], y + 1);
                h[2] = std::max(h[2], z + 1);
            });
        if (!any)
        {
            return out;
        }
        out.alloc(l, h);
        s.for_each(
            [&](int x, int y, int z)
            {
                out.set(x, y, z);
            });
        return out;
    }

    inline DSet box_set(int dim, int level, std::array<int, 3> lo, std::array<int, 3> hi)
    {
        DSet s(dim, level);
        s.alloc(lo, hi);
        std::fill(s.b.begin(), s.b.end(), 1);
        return s;
    }

    inline DSet empty_set(int dim, int level)
    {
        return DSet(dim, level);
    }

    // refine by k levels: every cell becomes its 2^(dim*k) descendants.
    inline DSet refine(const DSet& s, int k)
    {
        DSet out(s.dim, s.level + k);
        if (k == 0)
        {
            return s;
        }
        if (s.box_empty())
        {
            return out;
        }
        std::array<int, 3> l = s.lo, h = s.hi;
        for (int d = 0; d < s.dim; ++d)
        {
            l[d] = s.lo[d] * (1 << k);
            h[d] = s.hi[d] * (1 << k);
        }
        out.alloc(l, h);
        for (int z = out.lo[2]; z < out.hi[2]; ++z)
        {
            int zc = (s.dim == 3) ? (z >> k) : z;
            for (int y = out.lo[1]; y < out.hi[1]; ++y)
            {
                int yc = y >> k;
                for (int x = out.lo[0]; x < out.hi[0]; ++x)
                {
                    if (s.b[size_t(s.off(x >> k, yc, zc))])
                    {
                        out.b[size_t(out.off(x, y, z))] = 1;
                    }
                }
            }
        }
        return out;
    }

    // coarsen by k levels: a coarse cell is in the result if any descendant is.
    inline DSet coarsen(const DSet& s, int k)
    {
        DSet out(s.dim, s.level - k);
        if (k == 0)
        {
            return s;
        }
        if (s.level - k < 0)
        {
            throw std::logic_error("coarsen below level 0");
        }
        if (s.box_empty())
        {
            return out;
        }
        std::array<int, 3> l = s.lo, h = s.hi;
        for (int d = 0; d < s.dim; ++d)
        {
            l[d] = s.lo[d] >> k;
            h[d] = ((s.hi[d] - 1) >> k) + 1;
        }
        out.alloc(l, h);
        s.for_each(
            [&](int x, int y, int z)
            {
                out.b[size_t(out.off(x >> k, y >> k, s.dim == 3 ? (z >> k) : z))] = 1;
            });
        return out;
    }

    // samurai's `.on(level)`.
    inline DSet on(const DSet& s, int level)
    {
        if (level >= s.level)
        {
            return refine(s, level - s.level);
        }
        return coarsen(s, s.level - level);
    }

    enum class Op
    {
        Union,
        Intersection,
        Difference
    };

    inline DSet binop(const DSet& a0, const DSet& b0, Op op)
    {
        const int L = std::max(a0.level, b0.level);
        DSet ra, rb;
        if (a0.level != L)
        {
            ra = refine(a0, L - a0.level);
        }
        if (b0.level != L)
        {
            rb = refine(b0, L - b0.level);
        }
        const DSet& a = a0.level != L ? ra : a0;
        const DSet& b = b0.level != L ? rb : b0;
        DSet out(a.dim, L);
        std::array<int, 3> l{}, h{};
        if (op == Op::Union)
        {
            if (a.box_empty())
            {
                return b;
            }
            if (b.box_empty())
            {
                return a;
            }
            for (int d = 0; d < 3; ++d)
            {
                l[d] = std::min(a.lo[d], b.lo[d]);
                h[d] = std::max(a.hi[d], b.hi[d]);
            }
        }
        else if (op == Op::Intersection)
        {
            if (a.box_empty() || b.box_empty())
            {
                return out;
            }
            for (int d = 0; d < 3; ++d)
            {
                l[d] = std::max(a.lo[d], b.lo[d]);
                h[d] = std::min(a.hi[d], b.hi[d]);
                if (h[d] <= l[d])
                {
                    return out;
                }
            }
        }
        else
        {
            if (a.box_empty())
            {
                return out;
            }
            l = a.lo;
            h = a.hi;
        }
        out.alloc(l, h);
        for (int z = out.lo[2]; z < out.hi[2]; ++z)
        {
            for (int y = out.lo[1]; y < out.hi[1]; ++y)
            {
                for (int x = out.lo[0]; x < out.hi[0]; ++x)
                {
                    bool va = a.get(x, y, z);
                    bool vb = b.get(x, y, z);
                    bool v  = op == Op::Union ? (va || vb) : (op == Op::Intersection ? (va && vb) : (va && !vb));
                    if (v)
                    {
                        out.b[size_t(out.off(x, y, z))] = 1;
                    }
                }
            }
        }
        return out;
    }

    inline DSet union_(const DSet& a, const DSet& b)
    {
        return binop(a, b, Op::Union);
    }
    inline DSet intersection(const DSet& a, const DSet& b)
    {
        return binop(a, b, Op::Intersection);
    }
    inline DSet difference(const DSet& a, const DSet& b)
    {
        return binop(a, b, Op::Difference);
    }

    inline DSet translate(const DSet& s, std::array<int, 3> t)
    {
        DSet out = s;
        if (s.box_empty())
        {
            return out;
        }
        for (int d = 0; d < s.dim; ++d)
        {
            out.lo[d] += t[d];
            out.hi[d] += t[d];
        }
        return out;
    }

    // box dilation by w along each of the dim axes (separable).
    inline DSet expand(const DSet& s, int w)
    {
        if (w == 0 || s.box_empty())
        {
            return s;
        }
        DSet cur = s;
        for (int axis = 0; axis < s.dim; ++axis)
        {
            DSet out(s.dim, s.level);
            std::array<int, 3> l = cur.lo, h = cur.hi;
            l[axis] -= w;
            h[axis] += w;
            out.alloc(l, h);
            cur.for_each(
                [&](int x, int y, int z)
                {
                    for (int dw = -w; dw <= w; ++dw)
                    {
                        int c[3] = {x, y, z};
                        c[axis] += dw;
                        out.b[size_t(out.off(c[0], c[1], c[2]))] = 1;
                    }
                });
            cur = std::move(out);
        }
        return cur;
    }

    inline bool same_cells(const DSet& a, const DSet& b)
    {
        if (a.empty() && b.empty())
        {
            return true;
        }
        if (a.level != b.level)
        {
            return false;
        }
        return difference(a, b).empty() && difference(b, a).empty();
    }
} // namespace oracle
