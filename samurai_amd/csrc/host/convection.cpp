#include "convection.hpp"

#include <algorithm>
#include <array>
#include <cmath>
#include <stdexcept>

namespace smr
{
    namespace
    {
        struct Event
        {
            int64_t out, s0, s1;
            double c;
            int8_t dir = 0;
        };

        int64_t base_of(const MRMeshData& m, int level, int32_t x, int32_t y, int32_t z)
        {
            int64_t b = 0;
            if (!m.run_base(level, x, y, z, b))
            {
                throw std::out_of_range("convection: an interface stencil cell is missing from the reference mesh");
            }
            return b;
        }

        // group by output slot, keeping the order of each slot's terms
        ConvTerms group_events(const std::vector<Event>& ev, int64_t n, bool two_sources, bool with_dir = false)
        {
            std::vector<int64_t> cnt(size_t(n) + 1, 0);
            for (const Event& x : ev)
            {
                if (x.out < 0 || x.out >= n || x.s0 < 0 || x.s0 >= n || x.s1 >= n || (two_sources && x.s1 < 0))
                {
                    throw std::out_of_range("convection: a term is outside the field storage");
                }
                ++cnt[size_t(x.out) + 1];
            }
            ConvTerms t;
            std::vector<int64_t> start(size_t(n) + 1, 0);
            for (int64_t i = 0; i < n; ++i)
            {
                start[size_t(i) + 1] = start[size_t(i)] + cnt[size_t(i) + 1];
                if (cnt[size_t(i) + 1] > 0)
                {
                    t.row.push_back(i);
                }
            }
            const size_t nt = ev.size();
            t.src.assign(2 * nt, -1);
            t.coef.assign(nt, 0.0);
            if (with_dir)
            {
                t.dir.assign(nt, 0);
            }
            std::vector<int64_t> pos(start.begin(), start.end() - 1);
            for (const Event& x : ev)
            {
                const int64_t k          = pos[size_t(x.out)]++;
                t.src[size_t(2 * k)]     = x.s0;
                t.src[size_t(2 * k + 1)] = x.s1;
                t.coef[size_t(k)]        = x.c;
                if (with_dir)
                {
                    t.dir[size_t(k)] = x.dir;
                }
            }
            t.ptr.clear();
            t.ptr.reserve(t.row.size() + 1);
            for (int64_t r : t.row)
            {
                t.ptr.push_back(start[size_t(r)]);
            }
            t.ptr.push_back(int64_t(nt));
            return t;
        }

        // h_factor (flux_based_scheme__lin_hom.hpp:62-67)
        double h_factor(int dim, double h_face, double h_cell)
        {
            return std::pow(h_face, dim - 1) / std::pow(h_cell, dim);
        }
    } // namespace

    ConvTerms build_convection_terms(const MRMeshData& m, const double* v)
    {
        const int dim     = m.cfg.dim;
        const auto& cells = m.ids[CELLS];
        const int cmin = ca_min_level(cells), cmax = ca_max_level(cells);
        const int minl = std::max(m.cfg.min_level, cmin - 1), maxl = std::min(m.cfg.max_level, cmax + 1);
        const int zs   = dim == 3 ? 1 : 0;
        std::vector<Event> ev;
        auto add = [&](int64_t out, int64_t s0, int64_t s1, double c)
        {
            if (c != 0.0)
            {
                ev.push_back({out, s0, s1, c});
            }
        };
        for (int d = 0; d < dim; ++d)
        {
            std::array<int, 3> e{0, 0, 0};
            e[size_t(d)] = 1;
            const std::array<int, 3> me{-e[0], -e[1], -e[2]};
            // stencil {left cell, right cell}: the upwind side carries the velocity (convection_lin.hpp:36-77)
            const double fc[2] = {v[d] >= 0 ? v[d] : 0.0, v[d] >= 0 ? 0.0 : v[d]};

            // interfaces between cells of the same level (interface.hpp:35-108)
            for (int level = minl; level <= maxl; ++level)
            {
                const double h  = 1.0 / double(1 << level);
                const double fa = h_factor(dim, h, h);
                double L[2], R[2];
                for (int c = 0; c < 2; ++c)
                {
                    L[c] = fa * fc[c];
                    R[c] = -L[c];
                }
                const LevelSet I = set_intersection(cells[size_t(level)], translate(cells[size_t(level)], me));
                I.for_each_run(
                    [&](int32_t x0, int32_t x1, int32_t y, int32_t z)
                    {
                        const int64_t bp = base_of(m, level, x0, y, z);
                        const int64_t bq = base_of(m, level, x0 + e[0], y + e[1], z + e[2]) + e[0];
                        for (int c = 0; c < 2; ++c)
                        {
                            const int64_t bs = c == 0 ? bp : bq;
                            for (int32_t x = x0; x < x1; ++x)
                            {
                                add(bp + x, bs + x, -1, L[c]);
                            }
                            for (int32_t x = x0; x < x1; ++x)
                            {
                                add(bq + x, bs + x, -1, R[c]);
                            }
                        }
                    });
            }

            // level jumps level -- level+1 (interface.hpp:110-315)
            for (int level = minl; level < maxl && level < m.cfg.max_level; ++level)
            {
                const int lf     = level + 1;
                const double hl  = 1.0 / double(1 << level);
                const double hf  = 1.0 / double(1 << lf);
                auto coarse_base = [&](int32_t x0, int32_t y, int32_t z) { return base_of(m, level, x0 >> 1, y >> 1, z >> zs); };
                // coarse cell on the left, fine cells on the right
                {
                    const double lfac = h_factor(dim, hf, hl), rfac = h_factor(dim, hf, hf);
                    double L[2], R[2];
                    for (int c = 0; c < 2; ++c)
                    {
                        L[c] = lfac * fc[c];
                        R[c] = -rfac * fc[c];
                    }
                    const LevelSet J = on(set_intersection(cells[size_t(level)], translate(cells[size_t(lf)], me)), lf);
                    J.for_each_run(
                        [&](int32_t x0, int32_t x1, int32_t y, int32_t z)
                        {
                            const int32_t n  = x1 - x0;
                            const int64_t bg = base_of(m, lf, x0, y, z);                                // p (ghost in the coarse cell)
                            const int64_t bf = base_of(m, lf, x0 + e[0], y + e[1], z + e[2]) + e[0];    // p + e (fine cell)
                            const int64_t bc = coarse_base(x0, y, z);
                            for (int c = 0; c < 2; ++c)
                            {
                                const int64_t bs = c == 0 ? bg : bf;
                                if (n == 1)
                                {
                                    add(bc + (x0 >> 1), bs + x0, -1, L[c]);
                                    add(bf + x0, bs + x0, -1, R[c]);
                                }
                                else
                                {
                                    for (int32_t i = 0; i < n / 2; ++i)
                                    {
                                        const int32_t xa = x0 + 2 * i;
                                        add(bc + (xa >> 1), bs + xa, bs + xa + 1, L[c]);
                                    }
                                    for (int32_t x = x0; x < x1; ++x)
                                    {
                                        add(bf + x, bs + x, -1, R[c]);
                                    }
                                }
                            }
                        });
                }
                // fine cells on the left, coarse cell on the right
                {
                    const double lfac = h_factor(dim, hf, hf), rfac = h_factor(dim, hf, hl);
                    double L[2], R[2];
                    for (int c = 0; c < 2; ++c)
                    {
                        L[c] = lfac * fc[c];
                        R[c] = -rfac * fc[c];
                    }
                    const LevelSet J = on(set_intersection(cells[size_t(level)], translate(cells[size_t(lf)], e)), lf);
                    J.for_each_run(
                        [&](int32_t x0, int32_t x1, int32_t y, int32_t z)
                        {
                            const int32_t n  = x1 - x0;
                            const int64_t bg = base_of(m, lf, x0, y, z);                                // p (ghost in the coarse cell)
                            const int64_t bf = base_of(m, lf, x0 - e[0], y - e[1], z - e[2]) - e[0];    // p - e (fine cell)
                            const int64_t bc = coarse_base(x0, y, z);
                            for (int c = 0; c < 2; ++c)
                            {
                                // stencil {p - e, p}
                                const int64_t bs = c == 0 ? bf : bg;
                                if (n == 1)
                                {
                                    add(bf + x0, bs + x0, -1, L[c]);
                                    add(bc + (x0 >> 1), bs + x0, -1, R[c]);
                                }
                                else
                                {
                                    for (int32_t x = x0; x < x1; ++x)
                                    {
                                        add(bf + x, bs + x, -1, L[c]);
                                    }
                                    for (int32_t i = 0; i < n / 2; ++i)
                                    {
                                        const int32_t xa = x0 + 2 * i;
                                        add(bc + (xa >> 1), bs + xa, bs + xa + 1, R[c]);
                                    }
                                }
                            }
                        });
                }
            }

            // boundary interfaces (interface.hpp:440-509, boundary.hpp:6-33)
            for (int level = cmin; level <= cmax; ++level)
            {
                const LevelSet& cl = cells[size_t(level)];
                if (cl.empty())
                {
                    continue;
                }
                const double h  = 1.0 / double(1 << level);
                const double fa = h_factor(dim, h, h);
                double F[2]     = {fc[0] * fa, fc[1] * fa};
                // cells whose neighbour across the face is outside the domain:
                // cl minus the domain box shifted by -e (then by +e)
                const int n       = 1 << level;
                const LevelSet B  = minus_box(cl, me, {n + me[0], n + me[1], n + me[2]});
                B.for_each_run(
                    [&](int32_t x0, int32_t x1, int32_t y, int32_t z)
                    {
                        const int64_t bp = base_of(m, level, x0, y, z);
                        const int64_t bo = base_of(m, level, x0 + e[0], y + e[1], z + e[2]) + e[0];
                        for (int c = 0; c < 2; ++c)
                        {
                            const int64_t bs = c == 0 ? bp : bo;
                            for (int32_t x = x0; x < x1; ++x)
                            {
                                add(bp + x, bs + x, -1, F[c]);
                            }
                        }
                    });
                F[0] *= -1;
                F[1] *= -1;
                const LevelSet Bo = minus_box(cl, e, {n + e[0], n + e[1], n + e[2]});
                Bo.for_each_run(
                    [&](int32_t x0, int32_t x1, int32_t y, int32_t z)
                    {
                        const int64_t bp = base_of(m, level, x0, y, z);
                        const int64_t bo = base_of(m, level, x0 - e[0], y - e[1], z - e[2]) - e[0];
                        for (int c = 0; c < 2; ++c)
                        {
                            // stencil {p - e, p}
                            const int64_t bs = c == 0 ? bo : bp;
                            for (int32_t x = x0; x < x1; ++x)
                            {
                                add(bp + x, bs + x, -1, F[c]);
                            }
                        }
                    });
            }
        }

        return group_events(ev, m.nref, false);
    }

    namespace
    {
        // the interfaces of the flux-based operators one by one, in the order
        // the reference applies them (flux_based_scheme__nonlin.hpp:408-590,
        // flux_based_scheme__lin_het.hpp:75-250): each interface adds a term for
        // its left cell and one for its right cell; a boundary interface one
        // for its inner cell.  Events carry the direction.
        std::vector<Event> interface_events(const MRMeshData& m)
        {
            const int dim     = m.cfg.dim;
            const auto& cells = m.ids[CELLS];
            const int cmin = ca_min_level(cells), cmax = ca_max_level(cells);
            const int minl = std::max(m.cfg.min_level, cmin - 1), maxl = std::min(m.cfg.max_level, cmax + 1);
            const int zs   = dim == 3 ? 1 : 0;
            std::vector<Event> ev;
            // one interface: flux * lf into the left cell, (-flux) * rf = flux * (-rf) into the right one
            int8_t dir = 0;
            auto iface = [&](int64_t left, int64_t right, int64_t sl, int64_t sr, double lf, double rf)
            {
                ev.push_back({left, sl, sr, lf, dir});
                ev.push_back({right, sl, sr, -rf, dir});
            };
            for (int d = 0; d < dim; ++d)
            {
                dir = int8_t(d);
                std::array<int, 3> e{0, 0, 0};
                e[size_t(d)] = 1;
                const std::array<int, 3> me{-e[0], -e[1], -e[2]};
                for (int level = minl; level <= maxl; ++level)
                {
                    const double h  = 1.0 / double(1 << level);
                    const double fa = h_factor(dim, h, h);
                    set_intersection(cells[size_t(level)], translate(cells[size_t(level)], me))
                        .for_each_run(
                            [&](int32_t x0, int32_t x1, int32_t y, int32_t z)
                            {
                                const int64_t bp = base_of(m, level, x0, y, z);
                                const int64_t bq = base_of(m, level, x0 + e[0], y + e[1], z + e[2]) + e[0];
                                for (int32_t x = x0; x < x1; ++x)
                                {
                                    iface(bp + x, bq + x, bp + x, bq + x, fa, fa);
                                }
                            });
                }
                for (int level = minl; level < maxl && level < m.cfg.max_level; ++level)
                {
                    const int lf     = level + 1;
                    const double hl  = 1.0 / double(1 << level);
                    const double hf  = 1.0 / double(1 << lf);
                    auto coarse_base = [&](int32_t x0, int32_t y, int32_t z) { return base_of(m, level, x0 >> 1, y >> 1, z >> zs); };
                    // coarse cell on the left: stencil {p (fine ghost in the coarse cell), p + e}
                    {
                        const double lfac = h_factor(dim, hf, hl), rfac = h_factor(dim, hf, hf);
                        on(set_intersection(cells[size_t(level)], translate(cells[size_t(lf)], me)), lf)
                            .for_each_run(
                                [&](int32_t x0, int32_t x1, int32_t y, int32_t z)
                                {
                                    const int64_t bg = base_of(m, lf, x0, y, z);
                                    const int64_t bf = base_of(m, lf, x0 + e[0], y + e[1], z + e[2]) + e[0];
                                    for (int32_t x = x0; x < x1; ++x)
                                    {
                                        iface(coarse_base(x, y, z) + (x >> 1), bf + x, bg + x, bf + x, lfac, rfac);
                                    }
                                });
                    }
                    // fine cell on the left: stencil {p - e, p (fine ghost in the coarse cell)}
                    {
                        const double lfac = h_factor(dim, hf, hf), rfac = h_factor(dim, hf, hl);
                        on(set_intersection(cells[size_t(level)], translate(cells[size_t(lf)], e)), lf)
                            .for_each_run(
                                [&](int32_t x0, int32_t x1, int32_t y, int32_t z)
                                {
                                    const int64_t bg = base_of(m, lf, x0, y, z);
                                    const int64_t bf = base_of(m, lf, x0 - e[0], y - e[1], z - e[2]) - e[0];
                                    for (int32_t x = x0; x < x1; ++x)
                                    {
                                        iface(bf + x, coarse_base(x, y, z) + (x >> 1), bf + x, bg + x, lfac, rfac);
                                    }
                                });
                    }
                }
                // boundary interfaces: the outer ghost holds the boundary value
                for (int level = cmin; level <= cmax; ++level)
                {
                    const LevelSet& cl = cells[size_t(level)];
                    if (cl.empty())
                    {
                        continue;
                    }
                    const double h  = 1.0 / double(1 << level);
                    const double fa = h_factor(dim, h, h);
                    const int n     = 1 << level;
                    minus_box(cl, me, {n + me[0], n + me[1], n + me[2]})
                        .for_each_run(
                            [&](int32_t x0, int32_t x1, int32_t y, int32_t z)
                            {
                                const int64_t bp = base_of(m, level, x0, y, z);
                                const int64_t bo = base_of(m, level, x0 + e[0], y + e[1], z + e[2]) + e[0];
                                for (int32_t x = x0; x < x1; ++x)
                                {
                                    ev.push_back({bp + x, bp + x, bo + x, fa, dir});
                                }
                            });
                    minus_box(cl, e, {n + e[0], n + e[1], n + e[2]})
                        .for_each_run(
                            [&](int32_t x0, int32_t x1, int32_t y, int32_t z)
                            {
                                const int64_t bp = base_of(m, level, x0, y, z);
                                const int64_t bo = base_of(m, level, x0 - e[0], y - e[1], z - e[2]) - e[0];
                                for (int32_t x = x0; x < x1; ++x)
                                {
                                    ev.push_back({bp + x, bo + x, bp + x, -fa, dir});
                                }
                            });
                }
            }
            return ev;
        }
    } // namespace

    int weno5_dense_level(const MRMeshData& m)
    {
        const int dim     = m.cfg.dim;
        const auto& cells = m.ids[CELLS];
        const int cmin = ca_min_level(cells), cmax = ca_max_level(cells);
        if (cmin != cmax || std::max(m.cfg.radius, m.cfg.pred_s) != 3)
        {
            return -1;
        }
        const int level     = cmin;
        const int64_t n     = int64_t(1) << level;
        const int64_t W     = n + 6;
        const int64_t nrows = dim == 3 ? W * W : W;
        if (m.nref != (dim == 3 ? W * W * W : W * W) || cells[size_t(level)].count() != (dim == 3 ? n * n * n : n * n))
        {
            return -1;
        }
        for (int l = 0; l < MAXL; ++l)
        {
            if (l != level && !m.ids[REFERENCE][size_t(l)].empty())
            {
                return -1;
            }
        }
        const LevelSet& ref = m.ids[REFERENCE][size_t(level)];
        if (int64_t(ref.nrows()) != nrows || int64_t(ref.runs.size()) != nrows)
        {
            return -1;
        }
        for (size_t r = 0; r < ref.nrows(); ++r)
        {
            const int64_t y = int64_t(r) % W - 3, z = dim == 3 ? int64_t(r) / W - 3 : 0;
            if (ref.ry[r] != y || ref.rz[r] != z || ref.runs[r].s != -3 || ref.runs[r].e != n + 3)
            {
                return -1;
            }
        }
        if (m.index(level, -3, -3, dim == 3 ? -3 : 0) != 0)
        {
            return -1;
        }
        return level;
    }

    ConvTerms build_weno5_terms(const MRMeshData& m)
    {
        const int dim     = m.cfg.dim;
        const auto& cells = m.ids[CELLS];
        const int cmin = ca_min_level(cells), cmax = ca_max_level(cells);
        if (cmin != cmax)
        {
            throw std::invalid_argument("WENO5 is built for meshes whose cells are all at one level");
        }
        if (std::max(m.cfg.radius, m.cfg.pred_s) < 3)
        {
            throw std::invalid_argument("WENO5 needs ghost layers of width 3 (a mesh of stencil radius 3, max_stencil_size 6)");
        }
        const int level    = cmin;
        const int n        = 1 << level;
        const double h     = 1.0 / double(n);
        const double fa    = h_factor(dim, h, h);
        const LevelSet& cl = cells[size_t(level)];
        // the interfaces in order: f(out, x, y, z, j0, coef, d) for the term of
        // cell `out` whose stencil starts at (x, y, z) + j0 e_d
        auto walk = [&](auto&& f)
        {
            for (int d = 0; d < dim; ++d)
            {
                std::array<int, 3> e{0, 0, 0};
                e[size_t(d)] = 1;
                const std::array<int, 3> me{-e[0], -e[1], -e[2]};
                set_intersection(cl, translate(cl, me))
                    .for_each_run(
                        [&](int32_t x0, int32_t x1, int32_t y, int32_t z)
                        {
                            const int64_t bp = base_of(m, level, x0, y, z);
                            const int64_t bq = base_of(m, level, x0 + e[0], y + e[1], z + e[2]) + e[0];
                            for (int32_t x = x0; x < x1; ++x)
                            {
                                f(bp + x, x, y, z, -2, fa, d, e);
                                f(bq + x, x, y, z, -2, -fa, d, e);
                            }
                        });
                minus_box(cl, me, {n + me[0], n + me[1], n + me[2]})
                    .for_each_run(
                        [&](int32_t x0, int32_t x1, int32_t y, int32_t z)
                        {
                            const int64_t bp = base_of(m, level, x0, y, z);
                            for (int32_t x = x0; x < x1; ++x)
                            {
                                f(bp + x, x, y, z, -2, fa, d, e);
                            }
                        });
                minus_box(cl, e, {n + e[0], n + e[1], n + e[2]})
                    .for_each_run(
                        [&](int32_t x0, int32_t x1, int32_t y, int32_t z)
                        {
                            const int64_t bp = base_of(m, level, x0, y, z);
                            for (int32_t x = x0; x < x1; ++x)
                            {
                                f(bp + x, x, y, z, -3, -fa, d, e);
                            }
                        });
            }
        };
        // two passes (count, then fill each slot's terms in order): no list of
        // events the size of the output
        const int64_t nref = m.nref;
        std::vector<int64_t> pos(size_t(nref) + 1, 0);
        int64_t nt = 0;
        walk([&](int64_t out, int32_t, int32_t, int32_t, int, double, int, const std::array<int, 3>&) {
            ++pos[size_t(out) + 1];
            ++nt;
        });
        ConvTerms t;
        t.ptr.clear();
        for (int64_t i = 0; i < nref; ++i)
        {
            if (pos[size_t(i) + 1] > 0)
            {
                t.row.push_back(i);
            }
            pos[size_t(i) + 1] += pos[size_t(i)];
        }
        for (int64_t r : t.row)
        {
            t.ptr.push_back(pos[size_t(r)]);
        }
        t.ptr.push_back(nt);
        t.src.resize(size_t(6 * nt));
        t.coef.resize(size_t(nt));
        t.dir.resize(size_t(nt));
        walk([&](int64_t out, int32_t x, int32_t y, int32_t z, int j0, double c, int d, const std::array<int, 3>& e) {
            const int64_t k = pos[size_t(out)]++;
            for (int j = 0; j < 6; ++j)
            {
                const int q            = j0 + j;
                t.src[size_t(6 * k + j)] = m.index(level, x + q * e[0], y + q * e[1], z + q * e[2]);
            }
            t.coef[size_t(k)] = c;
            t.dir[size_t(k)]  = int8_t(d);
        });
        return t;
    }

    ConvTerms build_burgers_terms(const MRMeshData& m)
    {
        return group_events(interface_events(m), m.nref, true);
    }

    ConvTerms build_heterogeneous_terms(const MRMeshData& m)
    {
        return group_events(interface_events(m), m.nref, true, true);
    }
} // namespace smr
