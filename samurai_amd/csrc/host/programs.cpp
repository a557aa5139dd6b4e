#include "programs.hpp"
#include "parallel.hpp"

#include <algorithm>
#include <cstdlib>
#include <numeric>
#include <stdexcept>
#include <unordered_map>

namespace smr
{
    // ------------------------------------------------------------------
    // WorkList
    // ------------------------------------------------------------------
    void WorkList::begin_level(int level)
    {
        if (level_run.empty())
        {
            level_run.assign(MAXL + 1, -1);
        }
        if (level_run[size_t(level)] < 0)
        {
            level_run[size_t(level)] = nruns();
        }
    }

    void WorkList::add(const LevelSet& s)
    {
        begin_level(s.level);
        const size_t r0 = size_t(nruns()), nr = s.runs.size();
        iv.resize(iv.size() + 5 * nr);
        pref.resize(pref.size() + nr);
        int32_t* v = iv.data() + 5 * r0;
        int64_t* p = pref.data() + r0;
        for (size_t r = 0; r < s.nrows(); ++r)
        {
            for (int64_t k = s.rp[r]; k < s.rp[r + 1]; ++k)
            {
                const Run& run = s.runs[size_t(k)];
                v[0] = run.s;
                v[1] = run.e;
                v[2] = s.ry[r];
                v[3] = s.rz[r];
                v[4] = s.level;
                v += 5;
                p[1] = p[0] + (run.e - run.s);
                ++p;
            }
        }
    }

    void WorkList::finish()
    {
        if (level_run.empty())
        {
            level_run.assign(MAXL + 1, -1);
        }
        // levels never started begin where the next started level begins
        level_run[MAXL] = nruns();
        for (int l = MAXL - 1; l >= 0; --l)
        {
            if (level_run[size_t(l)] < 0)
            {
                level_run[size_t(l)] = level_run[size_t(l + 1)];
            }
        }
    }

    int64_t WorkList::runs_of_levels(int l0, int l1, int64_t& r0) const
    {
        l0 = std::max(l0, 0);
        l1 = std::min(l1, MAXL);
        if (l1 <= l0)
        {
            r0 = 0;
            return 0;
        }
        r0 = level_run[size_t(l0)];
        return level_run[size_t(l1)] - r0;
    }

    // ------------------------------------------------------------------
    // Recipes
    // ------------------------------------------------------------------
    void Recipes::push(int64_t d, int32_t o, const int64_t* s, int n)
    {
        dst.push_back(d);
        op.push_back(o);
        for (int k = 0; k < n; ++k)
        {
            src.push_back(s[k]);
        }
        src_ptr.push_back(int64_t(src.size()));
    }

    bool Recipes::stage_without_bc_only(size_t k) const
    {
        for (int64_t q = stage_ptr[k]; q < stage_ptr[k + 1]; ++q)
        {
            if (op[size_t(q)] != RECIPE_COPY_NO_BC)
            {
                return false;
            }
        }
        return stage_ptr[k + 1] > stage_ptr[k];
    }

    void Recipes::stage()
    {
        const size_t n = dst.size();
        std::vector<int> st(n, 0);
        std::unordered_map<int64_t, int> last_write, last_read;
        last_write.reserve(n);
        last_read.reserve(src.size());
        int nstages = 0;
        for (size_t r = 0; r < n; ++r)
        {
            int s = 0;
            for (int64_t k = src_ptr[r]; k < src_ptr[r + 1]; ++k)
            {
                auto it = last_write.find(src[size_t(k)]);
                if (it != last_write.end())
                {
                    s = std::max(s, it->second + 1);
                }
            }
            auto w = last_write.find(dst[r]);
            if (w != last_write.end())
            {
                s = std::max(s, w->second + 1);
            }
            auto rd = last_read.find(dst[r]);
            if (rd != last_read.end())
            {
                s = std::max(s, rd->second + 1);
            }
            st[r]               = s;
            last_write[dst[r]]  = s;
            for (int64_t k = src_ptr[r]; k < src_ptr[r + 1]; ++k)
            {
                int& lr = last_read[src[size_t(k)]];
                lr      = std::max(lr, s);
            }
            nstages = std::max(nstages, s + 1);
        }
        // the copies of B.C.-less fields get stages of their own (stage s ->
        // 2s for the other recipes, 2s + 1 for them, empty stages dropped), so
        // that a field with a B.C. skips whole stages (same order of every
        // dependent pair: a later stage stays later)
        {
            std::vector<int> used(size_t(2 * nstages), 0);
            for (size_t r = 0; r < n; ++r)
            {
                st[r]           = 2 * st[r] + (op[r] == RECIPE_COPY_NO_BC ? 1 : 0);
                used[size_t(st[r])] = 1;
            }
            std::vector<int> renum(used.size(), 0);
            int k = 0;
            for (size_t q = 0; q < used.size(); ++q)
            {
                renum[q] = k;
                k += used[q];
            }
            for (size_t r = 0; r < n; ++r)
            {
                st[r] = renum[size_t(st[r])];
            }
            nstages = k;
        }
        std::vector<size_t> order(n);
        std::iota(order.begin(), order.end(), size_t(0));
        std::stable_sort(order.begin(),
                         order.end(),
                         [&](size_t a, size_t b)
                         {
                             return st[a] < st[b];
                         });
        Recipes out;
        out.stage_ptr.assign(size_t(nstages) + 1, 0);
        for (size_t r : order)
        {
            out.push(dst[r], op[r], src.data() + src_ptr[r], int(src_ptr[r + 1] - src_ptr[r]));
            out.stage_ptr[size_t(st[r]) + 1]++;
        }
        for (size_t k = 1; k < out.stage_ptr.size(); ++k)
        {
            out.stage_ptr[k] += out.stage_ptr[k - 1];
        }
        *this = std::move(out);
    }

    namespace
    {
        void add_row_taps(WorkList& w, int8_t table, int8_t dl, int8_t dx, int lo, int hi, int dim, int8_t wlo, int8_t whi)
        {
            const int zlo = dim == 3 ? lo : 0, zhi = dim == 3 ? hi : 0;
            for (int dz = zlo; dz <= zhi; ++dz)
            {
                for (int dy = lo; dy <= hi; ++dy)
                {
                    w.taps.push_back({table, dl, dx, int8_t(dy), int8_t(dz), wlo, whi});
                }
            }
        }

        // the box of domain(level) shifted by t: [t, 2^level + t) on each axis
        struct Bounds
        {
            std::array<int, 3> lo, hi;
        };
        Bounds domain_bounds(int level, const Dir& t = {0, 0, 0})
        {
            const int n = 1 << level;
            return {{t[0], t[1], t[2]}, {n + t[0], n + t[1], n + t[2]}};
        }

        // the cells of domain(level) whose neighbour in `dir` is outside it:
        // the face layer x = n - 1 for dir x = +1, x = 0 for -1 (each axis)
        Bounds domain_layer(int level, const Dir& dir)
        {
            if (std::abs(dir[0]) + std::abs(dir[1]) + std::abs(dir[2]) != 1)
            {
                throw std::logic_error("domain_layer: a cartesian direction is expected");
            }
            Bounds b = domain_bounds(level);
            for (int i = 0; i < 3; ++i)
            {
                if (dir[size_t(i)] > 0)
                {
                    b.lo[size_t(i)] = b.hi[size_t(i)] - 1;
                }
                else if (dir[size_t(i)] < 0)
                {
                    b.hi[size_t(i)] = b.lo[size_t(i)] + 1;
                }
            }
            return b;
        }

        // clip_box(s, domain_layer(level, dir)): the cells of s on the domain face
        // in `dir`; for an x face one run lookup per row instead of a clip of
        // every row
        LevelSet face_layer(const LevelSet& s, int level, const Dir& dir)
        {
            const Bounds layer = domain_layer(level, dir);
            if (dir[0] == 0)
            {
                return clip_box(s, layer.lo, layer.hi);
            }
            const int32_t x = layer.lo[0];
            LevelSet out(s.dim, s.level);
            for (size_t r = 0; r < s.nrows(); ++r)
            {
                const int32_t y = s.ry[r], z = s.rz[r];
                if (y < layer.lo[1] || y >= layer.hi[1] || (s.dim == 3 && (z < layer.lo[2] || z >= layer.hi[2])))
                {
                    continue;
                }
                if (s.find_run(int64_t(r), x) >= 0)
                {
                    out.push(y, z, x, x + 1);
                }
            }
            return out;
        }

        // a n b for a set a much smaller than b: each row of a is looked up in b
        // and its runs intersected with the runs of that row from the first one
        // ending after them (bisection), so b is never walked as a whole
        LevelSet small_intersection(const LevelSet& a, const LevelSet& b)
        {
            if (a.level != b.level)
            {
                throw std::logic_error("small_intersection: sets of different levels");
            }
            LevelSet out(a.dim, a.level);
            if (a.empty() || b.empty())
            {
                return out;
            }
            out.reserve(a.nrows(), a.runs.size());
            for (size_t r = 0; r < a.nrows(); ++r)
            {
                const int64_t rb = b.find_row(a.ry[r], a.rz[r]);
                if (rb < 0)
                {
                    continue;
                }
                const Run* j        = b.runs.data() + b.rp[size_t(rb)];
                const Run* const je = b.runs.data() + b.rp[size_t(rb) + 1];
                for (int64_t i = a.rp[r]; i < a.rp[r + 1]; ++i)
                {
                    const Run x = a.runs[size_t(i)];
                    j           = std::lower_bound(j, je, x.s, [](const Run& q, int32_t v) { return q.e <= v; });
                    for (const Run* k = j; k < je && k->s < x.e; ++k)
                    {
                        out.push(a.ry[r], a.rz[r], std::max(x.s, k->s), std::min(x.e, k->e));
                    }
                }
            }
            return out;
        }

        // update_outer_ghosts(level) (algorithm/update_outer_ghost.hpp:338-432).
        // Every set below involves only cells next to the boundary: the
        // leaves, union cells and reference sets are clipped to the boundary
        // layer (or to the box of a corner) before the set algebra, which gives
        // the same sets as the whole-level expressions (the leaves and union
        // cells lie in the domain) at the cost of the boundary cells only.
        // apply_extrapolation_bc_cells<2k> (bc/apply_field_bc.hpp:124-150) in
        // direction d (Cartesian or diagonal: the B.C.'s line stencil turned
        // towards d is {-(k-1)..k} * d, stencil.hpp:226-231, 498-561): for each
        // cell b of `at` whose ghost b + k d exists, that ghost from the 2k - 1
        // cells b + j d, j = -(k-1)..k-1.
        void extrapolate_layer(const MRMeshData& m, int level, const LevelSet& at, const Dir& d, int k, Recipes& rec)
        {
            at.for_each_run(
                [&](int32_t x0, int32_t x1, int32_t y, int32_t z)
                {
                    int64_t src[5];
                    for (int32_t x = x0; x < x1; ++x)
                    {
                        if (!m.exists(level, x + k * d[0], y + k * d[1], z + k * d[2]))
                        {
                            continue;
                        }
                        for (int j = -(k - 1); j <= k - 1; ++j)
                        {
                            src[j + k - 1] = m.index(level, x + j * d[0], y + j * d[1], z + j * d[2]);
                        }
                        rec.push(m.index(level, x + k * d[0], y + k * d[1], z + k * d[2]), RECIPE_EXTRAPOLATE, src, 2 * k - 1);
                    }
                });
        }

        // update_outer_corners_by_polynomial_extrapolation beyond layer 1
        // (bc/apply_field_bc.hpp:332-464): layers 2..w of the diagonal through
        // each corner cell by extrapolation along the diagonal, then every
        // off-diagonal ghost of layer k (first non-zero axis of d at offset k,
        // the other ones at offsets 1..w) copies the diagonal ghost of layer k.
        void wide_corner(const MRMeshData& m, int level, const LevelSet& cc, const Dir& d, int w, Recipes& rec)
        {
            for (int k = 2; k <= w; ++k)
            {
                extrapolate_layer(m, level, cc, d, k, rec);
            }
            std::vector<int> nz;
            for (int i = 0; i < m.cfg.dim; ++i)
            {
                if (d[size_t(i)] != 0)
                {
                    nz.push_back(i);
                }
            }
            int ncombos = 1;
            for (size_t p = 1; p < nz.size(); ++p)
            {
                ncombos *= w;
            }
            for (int k = 1; k <= w; ++k)
            {
                for (int combo = 0; combo < ncombos; ++combo)
                {
                    Dir delta{0, 0, 0};
                    bool is_source = true;
                    int t          = combo;
                    for (size_t p = 1; p < nz.size(); ++p)
                    {
                        const int g  = t % w;
                        t           /= w;
                        const int dp = g - (k - 1);
                        delta[size_t(nz[p])] += dp * d[size_t(nz[p])];
                        is_source = is_source && dp == 0;
                    }
                    if (is_source)
                    {
                        continue;
                    }
                    cc.for_each_run(
                        [&](int32_t x0, int32_t x1, int32_t y, int32_t z)
                        {
                            for (int32_t x = x0; x < x1; ++x)
                            {
                                const int sx = x + k * d[0], sy = y + k * d[1], sz = z + k * d[2];
                                int64_t s    = m.index(level, sx, sy, sz);
                                rec.push(m.index(level, sx + delta[0], sy + delta[1], sz + delta[2]), RECIPE_COPY, &s, 1);
                            }
                        });
                }
            }
        }

        Recipes outer_recipes(const MRMeshData& m, int level, const std::vector<Dir>& cart)
        {
            Recipes rec;
            const int dim   = m.cfg.dim;
            const int width = std::max(m.cfg.radius, m.cfg.pred_s); // ghost width (mesh_config.hpp:395)
            const auto& cells = m.ids[CELLS];
            const auto& ref   = m.ids[REFERENCE];
            if (level <= m.cfg.max_level && level >= m.cfg.min_level)
            {
                for (size_t c = 0; c < m.corners.size(); ++c)
                {
                    const Dir& d = m.corner_dirs[c];
                    // corner extrapolation, ghost width 1 (bc/apply_field_bc.hpp:315-466):
                    // the diagonal ghost copies the corner cell
                    LevelSet corner = on(m.corners[c], level);
                    LevelSet cc     = small_intersection(corner, cells[size_t(level)]);
                    cc.for_each_run(
                        [&](int32_t x0, int32_t x1, int32_t y, int32_t z)
                        {
                            for (int32_t x = x0; x < x1; ++x)
                            {
                                int64_t s = m.index(level, x, y, z);
                                rec.push(m.index(level, x + d[0], y + d[1], z + d[2]), RECIPE_COPY, &s, 1);
                            }
                        });
                    if (width > 1)
                    {
                        wide_corner(m, level, cc, d, width, rec);
                    }
                    // project_corner_below (algorithm/update_outer_ghost.hpp:267-336)
                    if (level > 0)
                    {
                        LevelSet fine_outer = small_intersection(translate(corner, d), ref[size_t(level)]);
                        for (int dl = 1; dl <= 2 && !fine_outer.empty(); ++dl)
                        {
                            const int pl   = level - dl;
                            LevelSet ghost = small_intersection(on(fine_outer, pl), ref[size_t(pl)]);
                            const int last = (1 << dl) - 1;
                            ghost.for_each_run(
                                [&](int32_t x0, int32_t x1, int32_t y, int32_t z)
                                {
                                    for (int32_t x = x0; x < x1; ++x)
                                    {
                                        int cx = x * (1 << dl) + (d[0] == -1 ? last : 0);
                                        int cy = y * (1 << dl) + (d[1] == -1 ? last : 0);
                                        int cz = dim == 3 ? z * (1 << dl) + (d[2] == -1 ? last : 0) : 0;
                                        int64_t s;
                                        if (m.run_base(level, cx, cy, cz, s))
                                        {
                                            s += cx;
                                            rec.push(m.index(pl, x, y, z), RECIPE_COPY, &s, 1);
                                        }
                                    }
                                });
                            if (pl == 0)
                            {
                                break;
                            }
                        }
                    }
                }
            }
            // the boundary cells of the level on each face (B.C., B.C. prediction, far layers)
            std::vector<LevelSet> faces(cart.size());
            if (level >= m.cfg.min_level)
            {
                for (size_t di = 0; di < cart.size(); ++di)
                {
                    faces[di] = face_layer(cells[size_t(level)], level, cart[di]);
                }
            }
            for (size_t di = 0; di < cart.size(); ++di)
            {
                const Dir& dir = cart[di];
                if (level < m.cfg.max_level)
                {
                    // project_bc, layer 1 (algorithm/update_outer_ghost.hpp:20-152)
                    const Bounds dom  = domain_bounds(level);
                    LevelSet outside  = minus_box(translate(face_layer(m.ids[UNION_CELLS][size_t(level)], level, dir), dir), dom.lo, dom.hi);
                    LevelSet ghosts   = small_intersection(outside, ref[size_t(level)]);
                    std::vector<int64_t> ch;
                    ghosts.for_each_run(
                        [&](int32_t x0, int32_t x1, int32_t y, int32_t z)
                        {
                            for (int32_t x = x0; x < x1; ++x)
                            {
                                ch.clear();
                                for (int cl = level + 1; cl <= level + 2 && cl < MAXL; ++cl)
                                {
                                    const int k = cl - level, f = 1 << k;
                                    const int nz = dim == 3 ? f : 1;
                                    for (int dz = 0; dz < nz; ++dz)
                                    {
                                        for (int dy = 0; dy < f; ++dy)
                                        {
                                            for (int dx = 0; dx < f; ++dx)
                                            {
                                                int64_t b;
                                                int cx = x * f + dx, cy = y * f + dy, cz = dim == 3 ? z * f + dz : 0;
                                                if (m.run_base(cl, cx, cy, cz, b))
                                                {
                                                    ch.push_back(b + cx);
                                                }
                                            }
                                        }
                                    }
                                    if (!ch.empty())
                                    {
                                        break;
                                    }
                                }
                                rec.push(m.index(level, x, y, z), RECIPE_AVERAGE, ch.data(), int(ch.size()));
                            }
                        });
                }
                if (level >= m.cfg.min_level)
                {
                    // Dirichlet<1> on the boundary cells (bc/apply_field_bc.hpp:53-101)
                    faces[di].for_each_run(
                        [&](int32_t x0, int32_t x1, int32_t y, int32_t z)
                        {
                            for (int32_t x = x0; x < x1; ++x)
                            {
                                int64_t s = m.index(level, x, y, z);
                                rec.push(m.index(level, x + dir[0], y + dir[1], z + dir[2]), RECIPE_DIRICHLET, &s, 1);
                            }
                        });
                }
                if (level < m.cfg.max_level && level >= m.cfg.min_level)
                {
                    // predict_bc, order 0 (algorithm/update_outer_ghost.hpp:154-262)
                    // the leaves outside the domain shifted by -dir are those of
                    // the boundary layer in direction dir
                    const Bounds dom   = domain_bounds(level);
                    LevelSet bc_ghosts = minus_box(translate(faces[di], dir), dom.lo, dom.hi);
                    LevelSet ghosts    = small_intersection(on(bc_ghosts, level + 1), ref[size_t(level + 1)]);
                    ghosts.for_each_run(
                        [&](int32_t x0, int32_t x1, int32_t y, int32_t z)
                        {
                            for (int32_t x = x0; x < x1; ++x)
                            {
                                int64_t s = m.index(level, x >> 1, y >> 1, dim == 3 ? (z >> 1) : 0);
                                rec.push(m.index(level + 1, x, y, z), RECIPE_COPY, &s, 1);
                            }
                        });
                }
            }
            // update_further_ghosts_by_polynomial_extrapolation (bc/apply_field_bc.hpp:
            // 499-563, update_outer_ghost.hpp:414-426): the layers the B.C. leaves,
            // extrapolated from the boundary cells.  Layer 1 is the B.C.'s when the
            // field has one (Dirichlet<1>), else a copy (RECIPE_COPY_NO_BC, which
            // the field's B.C. skips).  The part next to the union cells' ghosts
            // has nothing to do at ghost width 1 or on a mesh of one level.
            if (level >= m.cfg.min_level)
            {
                for (size_t di = 0; di < cart.size(); ++di)
                {
                    const Dir& dir        = cart[di];
                    const LevelSet& bdry  = faces[di];
                    if (bdry.empty())
                    {
                        continue;
                    }
                    bdry.for_each_run(
                        [&](int32_t x0, int32_t x1, int32_t y, int32_t z)
                        {
                            for (int32_t x = x0; x < x1; ++x)
                            {
                                int64_t s = m.index(level, x, y, z);
                                rec.push(m.index(level, x + dir[0], y + dir[1], z + dir[2]), RECIPE_COPY_NO_BC, &s, 1);
                            }
                        });
                    for (int k = 2; k <= width; ++k)
                    {
                        extrapolate_layer(m, level, bdry, dir, k, rec);
                    }
                }
            }
            rec.stage();
            return rec;
        }
    } // namespace

    namespace
    {
        // the per-level set computations of the ghost and adapt programs
        struct ProgramSets
        {
            std::vector<LevelSet> proj, pred, det, crit;
            // update_fields from an older mesh: copy, coarsen, refine per level
            std::vector<LevelSet> cp, co, re;
        };

        // task t of the update_fields sets (algorithm/update_fields.hpp:27-127)
        // from the mesh om to the mesh nm
        void uf_task(const MRMeshData& om, const MRMeshData& nm, int task, ProgramSets& sets)
        {
            const int minl = om.cfg.min_level;
            const int kind = task % 3, level = task / 3;
            if (kind == 0 && level >= minl)
            {
                sets.cp[size_t(level)] = set_intersection(om.ids[REFERENCE][size_t(level)], nm.ids[CELLS][size_t(level)]);
            }
            else if (kind == 1 && level >= minl + 1)
            {
                sets.co[size_t(level)] = set_intersection(on(om.ids[CELLS][size_t(level)], level - 1), nm.ids[CELLS][size_t(level - 1)]);
            }
            else if (kind == 2 && level >= minl + 1)
            {
                sets.re[size_t(level)] = set_intersection(on(nm.ids[CELLS][size_t(level)], level - 1), om.ids[CELLS][size_t(level - 1)]);
            }
        }

        // list `which` (0 copy, 1 coarsen, 2 refine) of update_fields from its sets
        void uf_assemble(const MRMeshData& om, int which, WorkList& w, const ProgramSets& sets)
        {
            const int dim = om.cfg.dim, s = om.cfg.pred_s;
            const int minl = om.cfg.min_level, maxl = om.cfg.max_level;
            w      = WorkList();
            w.taps = list_taps(LIST_UF_COPY + which, dim, s);
            if (which == 0)
            {
                for (int level = minl; level <= maxl; ++level)
                {
                    w.add(sets.cp[size_t(level)]);
                }
            }
            else if (which == 1)
            {
                for (int level = minl + 1; level <= maxl; ++level)
                {
                    w.add(sets.co[size_t(level)]);
                }
            }
            else
            {
                for (int level = minl + 1; level <= maxl; ++level)
                {
                    w.add(sets.re[size_t(level)]);
                }
            }
            w.finish();
        }

        int ghost_tasks(const MRMeshData& m)
        {
            return 3 * (m.cfg.max_level + 1);
        }
        int adapt_tasks(const MRMeshData& m)
        {
            return 2 * (m.cfg.max_level + 1);
        }

        // task t of the ghost program: outer recipes, projection set or
        // prediction set of one level (finest levels first)
        void ghost_task(const MRMeshData& m, const std::vector<Dir>& cart, int task, GhostProgram& p, ProgramSets& sets)
        {
            const int maxl = m.cfg.max_level;
            const int kind = task % 3, level = maxl - task / 3;
            if (kind == 0)
            {
                p.outer[size_t(level)] = outer_recipes(m, level, cart);
            }
            else if (kind == 1 && level >= 1)
            {
                sets.proj[size_t(level)] =
                    set_intersection(on(m.ids[REFERENCE][size_t(level)], level - 1), m.ids[PROJ_CELLS][size_t(level - 1)]);
            }
            else if (kind == 2 && level >= 1)
            {
                const LevelSet& ref  = m.ids[REFERENCE][size_t(level)];
                LevelSet pred_ghosts = set_difference(ref, set_union(m.ids[CELLS][size_t(level)], m.ids[PROJ_CELLS][size_t(level)]));
                const Bounds dom = domain_bounds(level);
                sets.pred[size_t(level)] =
                    set_intersection(clip_box(pred_ghosts, dom.lo, dom.hi), m.ids[REFERENCE][size_t(level - 1)]);
            }
        }

        void adapt_task(const MRMeshData& m, int task, ProgramSets& sets)
        {
            const int minl = m.cfg.min_level, maxl = m.cfg.max_level;
            const int l0    = minl > 0 ? minl - 1 : 0;
            const int level = maxl - task / 2;
            if (task % 2 == 0 && level >= l0 && level < maxl)
            {
                LevelSet fine           = set_union(m.ids[CELLS][size_t(level + 1)], m.ids[CELLS][size_t(level + 2)]);
                sets.det[size_t(level)] = set_intersection(m.ids[REFERENCE][size_t(level)], on(fine, level));
            }
            else if (task % 2 == 1 && level >= 1)
            {
                sets.crit[size_t(level)] = set_intersection(on(m.ids[CELLS][size_t(level)], level - 1), m.ids[REFERENCE][size_t(level - 1)]);
            }
        }

        // the five work lists of the ghost and adapt programs from their
        // per-level sets; the lists are independent and built in parallel
        void assemble(const MRMeshData& m, GhostProgram* gp, AdaptProgram* ap, ProgramSets& sets, const MRMeshData* old = nullptr,
                      WorkList* uf = nullptr)
        {
            const int dim = m.cfg.dim, s = m.cfg.pred_s;
            const int minl = m.cfg.min_level, maxl = m.cfg.max_level;
            const int l0   = minl > 0 ? minl - 1 : 0;
            parallel_for(old && uf ? 8 : 5,
                         [&](int task)
                         {
                             if (task >= 5)
                             {
                                 uf_assemble(*old, task - 5, uf[task - 5], sets);
                             }
                             else if (task == 0 && gp)
                             {
                                 // projection (numeric/projection.hpp:15-65): runs at the coarse level
                                 WorkList& w = gp->projection;
                                 w.taps      = list_taps(LIST_PROJECTION, dim, s);
                                 for (int level = 1; level <= maxl; ++level)
                                 {
                                     w.add(sets.proj[size_t(level)]);
                                 }
                                 w.finish();
                             }
                             else if (task == 1 && gp)
                             {
                                 // prediction <s,false> (numeric/prediction.hpp:259-357)
                                 WorkList& w = gp->prediction;
                                 w.taps      = list_taps(LIST_PREDICTION, dim, s);
                                 for (int level = 1; level <= maxl; ++level)
                                 {
                                     w.add(sets.pred[size_t(level)]);
                                 }
                                 w.finish();
                             }
                             else if (task == 2 && ap)
                             {
                                 // compute_detail (mr/operators.hpp:139-540): fine rows, then coarse stencil rows
                                 WorkList& w = ap->detail;
                                 w.taps      = list_taps(LIST_DETAIL, dim, s);
                                 for (int level = l0; level < maxl; ++level)
                                 {
                                     w.add(sets.det[size_t(level)]);
                                 }
                                 w.finish();
                             }
                             else if ((task == 3 || task == 4) && ap)
                             {
                                 // mr_criteria / maximum (mr/criteria.hpp:13-122, mr/operators.hpp:22-96):
                                 // the criteria sets are those of the fine levels [min, max], the
                                 // maximum sets those of [1, max]; both are (cells[l] n reference[l-1]).on(l-1)
                                 WorkList& w = task == 3 ? ap->criteria : ap->maximum;
                                 w.taps      = list_taps(task == 3 ? LIST_CRITERIA : LIST_MAXIMUM, dim, s);
                                 for (int level = task == 3 ? std::max(1, minl) : 1; level <= maxl; ++level)
                                 {
                                     w.add(sets.crit[size_t(level)]);
                                 }
                                 w.finish();
                             }
                         });
        }

        ProgramSets make_sets(const MRMeshData& m)
        {
            ProgramSets sets;
            const size_t n = size_t(m.cfg.max_level) + 1;
            sets.proj.resize(n);
            sets.pred.resize(n);
            sets.det.resize(n);
            sets.crit.resize(n);
            sets.cp.resize(n);
            sets.co.resize(n);
            sets.re.resize(n);
            return sets;
        }
    } // namespace

    std::vector<Tap> list_taps(int kind, int dim, int s)
    {
        WorkList w;
        switch (kind)
        {
            case LIST_PROJECTION:
                // numeric/projection.hpp:15-65: the cell, then the fine rows
                w.taps.push_back({0, 0, 0, 0, 0, 0, 0});
                add_row_taps(w, 0, 1, 0, 0, 1, dim, 0, 1);
                break;
            case LIST_PREDICTION:
                // prediction <s,false> (numeric/prediction.hpp:259-357): the
                // ghost, then the coarse stencil rows
                w.taps.push_back({0, 0, 0, 0, 0, 0, 0});
                add_row_taps(w, 0, -1, int8_t(-s), -s, s, dim, 0, int8_t(2 * s));
                break;
            case LIST_DETAIL:
                // compute_detail (mr/operators.hpp:139-540): fine rows, then coarse stencil rows
                add_row_taps(w, 0, 1, 0, 0, 1, dim, 0, 1);
                add_row_taps(w, 0, 0, int8_t(-s), -s, s, dim, 0, int8_t(2 * s));
                break;
            case LIST_CRITERIA:
            case LIST_MAXIMUM:
                // mr_criteria / maximum (mr/criteria.hpp:13-122, mr/operators.hpp:22-96)
                w.taps.push_back({0, 0, 0, 0, 0, 0, 0});
                add_row_taps(w, 0, 1, 0, 0, 1, dim, 0, 1);
                break;
            case LIST_UF_COPY:
                // update_fields (algorithm/update_fields.hpp:27-127): new cell, old cell
                w.taps.push_back({0, 0, 0, 0, 0, 0, 0});
                w.taps.push_back({1, 0, 0, 0, 0, 0, 0});
                break;
            case LIST_UF_COARSEN:
                w.taps.push_back({0, 0, 0, 0, 0, 0, 0});
                add_row_taps(w, 1, 1, 0, 0, 1, dim, 0, 1);
                break;
            case LIST_UF_REFINE:
                add_row_taps(w, 0, 1, 0, 0, 1, dim, 0, 1);
                add_row_taps(w, 1, 0, int8_t(-s), -s, s, dim, 0, int8_t(2 * s));
                break;
            default: throw std::logic_error("list_taps: unknown list");
        }
        return std::move(w.taps);
    }

    void build_recipes_and_cells(const MRMeshData& m, std::vector<Recipes>& outer, WorkList& cells)
    {
        const auto cart = cartesian_directions(m.cfg.dim);
        const int nl    = m.cfg.max_level + 1;
        outer.assign(size_t(nl), Recipes());
        parallel_for(nl + 1,
                     [&](int task)
                     {
                         if (task == nl)
                         {
                             cells = build_cells_list(m);
                         }
                         else
                         {
                             // the finest levels first: they hold the most boundary cells
                             const int level   = nl - 1 - task;
                             outer[size_t(level)] = outer_recipes(m, level, cart);
                         }
                     });
    }

    GhostProgram build_ghost_program(const MRMeshData& m)
    {
        GhostProgram p;
        const auto cart = cartesian_directions(m.cfg.dim);
        p.outer.resize(size_t(m.cfg.max_level) + 1);
        ProgramSets sets = make_sets(m);
        parallel_for(ghost_tasks(m), [&](int task) { ghost_task(m, cart, task, p, sets); });
        assemble(m, &p, nullptr, sets);
        return p;
    }

    AdaptProgram build_adapt_program(const MRMeshData& m)
    {
        AdaptProgram p;
        ProgramSets sets = make_sets(m);
        parallel_for(adapt_tasks(m), [&](int task) { adapt_task(m, task, sets); });
        assemble(m, nullptr, &p, sets);
        return p;
    }

    void build_programs(const MRMeshData& m, WorkList& cells, GhostProgram& ghost, AdaptProgram& adapt, const MRMeshData* old,
                        WorkList* uf)
    {
        ghost            = GhostProgram();
        adapt            = AdaptProgram();
        const auto cart  = cartesian_directions(m.cfg.dim);
        ghost.outer.resize(size_t(m.cfg.max_level) + 1);
        ProgramSets sets = make_sets(m);
        const int ng = ghost_tasks(m), na = adapt_tasks(m);
        const int nu = old && uf ? 3 * (m.cfg.max_level + 1) : 0;
        if (nu && (old->cfg.max_level != m.cfg.max_level || old->cfg.min_level != m.cfg.min_level))
        {
            throw std::invalid_argument("update_fields between meshes of different level ranges");
        }
        // one pool of tasks for both programs, the cells list and the
        // update_fields sets from the older mesh
        parallel_for(ng + na + 1 + nu,
                     [&](int task)
                     {
                         if (task < ng)
                         {
                             ghost_task(m, cart, task, ghost, sets);
                         }
                         else if (task < ng + na)
                         {
                             adapt_task(m, task - ng, sets);
                         }
                         else if (task == ng + na)
                         {
                             cells = build_cells_list(m);
                         }
                         else
                         {
                             uf_task(*old, m, task - ng - na - 1, sets);
                         }
                     });
        assemble(m, &ghost, &adapt, sets, nu ? old : nullptr, uf);
    }

    WorkList build_cells_list(const MRMeshData& m)
    {
        WorkList w;
        w.walk = m.cfg.dim == 2;
        w.taps.push_back({0, 0, 0, 0, 0, -1, 1});
        w.taps.push_back({0, 0, 0, -1, 0, 0, 0});
        w.taps.push_back({0, 0, 0, 1, 0, 0, 0});
        if (m.cfg.dim == 3)
        {
            w.taps.push_back({0, 0, 0, 0, -1, 0, 0});
            w.taps.push_back({0, 0, 0, 0, 1, 0, 0});
        }
        for (int level = 0; level < MAXL; ++level)
        {
            if (!m.ids[CELLS][size_t(level)].empty())
            {
                w.add(m.ids[CELLS][size_t(level)]);
            }
        }
        w.finish();
        return w;
    }

    void build_update_fields_lists(const MRMeshData& om, const MRMeshData& nm, WorkList& copy, WorkList& coarsen, WorkList& refine)
    {
        ProgramSets sets = make_sets(om);
        parallel_for(3 * (om.cfg.max_level + 1), [&](int task) { uf_task(om, nm, task, sets); });
        WorkList* out[3] = {&copy, &coarsen, &refine};
        parallel_for(3, [&](int k) { uf_assemble(om, k, *out[k], sets); });
    }
} // namespace smr

namespace smr
{
    std::string validate_list(const WorkList& w, const MRMeshData& t0, const MRMeshData& t1)
    {
        for (int64_t r = 0; r < w.nruns(); ++r)
        {
            const int32_t* iv = w.iv.data() + 5 * r;
            const int32_t x0 = iv[0], x1 = iv[1], y = iv[2], z = iv[3];
            const int level  = iv[4];
            for (const Tap& t : w.taps)
            {
                int tl;
                int32_t tx, last, ty, tz;
                if (t.dl == 0)
                {
                    tl = level, tx = x0 + t.dx, last = x1 - 1 + t.dx, ty = y + t.dy, tz = z + t.dz;
                }
                else if (t.dl > 0)
                {
                    tl = level + 1, tx = 2 * x0 + t.dx, last = 2 * (x1 - 1) + t.dx, ty = 2 * y + t.dy, tz = 2 * z + t.dz;
                }
                else
                {
                    tl = level - 1, tx = (x0 >> 1) + t.dx, last = ((x1 - 1) >> 1) + t.dx, ty = (y >> 1) + t.dy, tz = (z >> 1) + t.dz;
                }
                const MRMeshData& m = t.table == 0 ? t0 : t1;
                bool ok             = tl >= 0 && tl < MAXL;
                if (ok)
                {
                    const LevelSet& ref = m.ids[REFERENCE][size_t(tl)];
                    const int64_t row   = ref.find_row(ty, tz);
                    const int64_t k     = row >= 0 ? ref.find_run(row, tx) : -1;
                    ok = k >= 0 && ref.runs[size_t(k)].s <= tx + t.wlo && last + t.whi < ref.runs[size_t(k)].e;
                }
                if (!ok)
                {
                    return "run (level " + std::to_string(level) + "; x " + std::to_string(x0) + ".." + std::to_string(x1) + ", y " +
                           std::to_string(y) + ", z " + std::to_string(z) + ") has a tap outside the reference";
                }
            }
        }
        return "";
    }

    std::string validate_programs(const MRMeshData& m)
    {
        std::string why = validate_list(build_cells_list(m), m, m);
        if (!why.empty())
        {
            return "cells: " + why;
        }
        GhostProgram g = build_ghost_program(m);
        if (!(why = validate_list(g.projection, m, m)).empty())
        {
            return "projection: " + why;
        }
        if (!(why = validate_list(g.prediction, m, m)).empty())
        {
            return "prediction: " + why;
        }
        AdaptProgram a = build_adapt_program(m);
        if (!(why = validate_list(a.detail, m, m)).empty())
        {
            return "detail: " + why;
        }
        if (!(why = validate_list(a.criteria, m, m)).empty())
        {
            return "criteria: " + why;
        }
        if (!(why = validate_list(a.maximum, m, m)).empty())
        {
            return "maximum: " + why;
        }
        return "";
    }
} // namespace smr
