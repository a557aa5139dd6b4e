#include "mr_mesh.hpp"
#include "parallel.hpp"

#include <map>
#include <mutex>

#include <algorithm>
#include <cstdio>
#include <stdexcept>

namespace smr
{
    CellArray empty_cell_array(int dim)
    {
        CellArray ca;
        ca.reserve(MAXL);
        for (int l = 0; l < MAXL; ++l)
        {
            ca.emplace_back(dim, l);
        }
        return ca;
    }

    int ca_min_level(const CellArray& ca)
    {
        for (int l = 0; l < MAXL; ++l)
        {
            if (!ca[size_t(l)].empty())
            {
                return l;
            }
        }
        return MAXL;
    }

    int ca_max_level(const CellArray& ca)
    {
        for (int l = MAXL - 1; l >= 0; --l)
        {
            if (!ca[size_t(l)].empty())
            {
                return l;
            }
        }
        return 0;
    }

    bool ca_equal(const CellArray& a, const CellArray& b)
    {
        for (int l = 0; l < MAXL; ++l)
        {
            if (a[size_t(l)] != b[size_t(l)])
            {
                return false;
            }
        }
        return true;
    }

    // for_each_cartesian_direction (stencil.hpp:300-316)
    std::vector<Dir> cartesian_directions(int dim)
    {
        std::vector<Dir> out;
        for (int d = 0; d < dim; ++d)
        {
            Dir v{0, 0, 0};
            v[size_t(d)] = 1;
            out.push_back(v);
            v[size_t(d)] = -1;
            out.push_back(v);
        }
        return out;
    }

    // for_each_diagonal_direction (stencil.hpp:337-349): static_nested_loop<dim,-1,2>
    // order (last dimension outermost) keeping directions with >1 non-zero entry
    std::vector<Dir> diagonal_directions(int dim)
    {
        std::vector<Dir> out;
        const int zr = dim == 3 ? 1 : 0;
        for (int z = -zr; z <= zr; ++z)
        {
            for (int y = -1; y <= 1; ++y)
            {
                for (int x = -1; x <= 1; ++x)
                {
                    if ((x != 0) + (y != 0) + (z != 0) > 1)
                    {
                        out.push_back({x, y, z});
                    }
                }
            }
        }
        return out;
    }

    bool MRMeshData::run_base(int level, int x, int y, int z, int64_t& base) const
    {
        if (level < 0 || level >= MAXL)
        {
            return false;
        }
        const LevelSet& r = ids[REFERENCE][size_t(level)];
        int64_t row       = r.find_row(y, z);
        if (row < 0)
        {
            return false;
        }
        int64_t k = r.find_run(row, x);
        if (k < 0)
        {
            return false;
        }
        base = ridx[size_t(level)][size_t(k)];
        return true;
    }

    int64_t MRMeshData::index(int level, int x, int y, int z) const
    {
        int64_t base = 0;
        if (!run_base(level, x, y, z, base))
        {
            char buf[160];
            std::snprintf(buf, sizeof(buf), "cell (level %d; %d, %d, %d) is not in the reference mesh", level, x, y, z);
            throw std::out_of_range(buf);
        }
        return base + x;
    }

    bool MRMeshData::exists(int level, int x, int y, int z) const
    {
        int64_t base = 0;
        return run_base(level, x, y, z, base);
    }

    int64_t MRMeshData::nb_cells(int id) const
    {
        if (id == REFERENCE)
        {
            return nref;
        }
        int64_t n = 0;
        for (int l = 0; l < MAXL; ++l)
        {
            n += ids[id][size_t(l)].count();
        }
        return n;
    }

    namespace
    {
        LevelSet domain_at(int dim, int level)
        {
            const int n = 1 << level;
            return box(dim, level, {0, 0, 0}, {n, n, dim == 3 ? n : 1});
        }

        // Mesh_base::construct_union (mesh.hpp:1231-1262)
        void construct_union(MRMeshData& m)
        {
            CellArray& uni = m.ids[UNION_CELLS];
            uni            = empty_cell_array(m.cfg.dim);
            for (int level = m.cfg.max_level; level >= 1; --level)
            {
                uni[size_t(level - 1)] = on(set_union(m.ids[CELLS][size_t(level)], uni[size_t(level)]), level - 1);
            }
        }

        // MRMesh::update_sub_mesh_impl (mr/mesh.hpp:221-455), non-periodic, with
        // Mesh_base::construct_union (mesh.hpp:1231-1262).  The level sets are
        // built as a task graph: the union chain runs beside the per-level
        // ghost expansions; every level of the reference before the
        // projection ghosts is cells_and_ghosts[t] U below1[t+1] U below2[t+2]
        // (the prediction ghosts one and two levels below the cells); the
        // projection ghosts are added level by level, each level depending on
        // the one below.
        void update_sub_mesh(MRMeshData& m)
        {
            const int r = m.cfg.radius, s = m.cfg.pred_s, dim = m.cfg.dim;
            const bool mr = m.cfg.max_level != m.cfg.min_level;
            const CellArray& cells = m.ids[CELLS];
            CellArray cag = empty_cell_array(dim), below1 = empty_cell_array(dim), below2 = empty_cell_array(dim);
            m.ids[UNION_CELLS] = empty_cell_array(dim);
            // phase 1: the union chain | cells_and_ghosts[l] | below2[l]
            parallel_for(1 + 2 * MAXL,
                         [&](int task)
                         {
                             if (task == 0)
                             {
                                 construct_union(m);
                                 return;
                             }
                             const int level    = (task - 1) / 2;
                             const LevelSet& cl = cells[size_t(level)];
                             if (cl.empty())
                             {
                                 return;
                             }
                             if ((task - 1) % 2 == 0)
                             {
                                 cag[size_t(level)] = expand(cl, r);
                             }
                             else if (mr && level - 1 > 0)
                             {
                                 below2[size_t(level)] = expand(on(cl, level - 2), s);
                             }
                         });
            if (mr)
            {
                // phase 2: below1[l] = expand(cells_and_ghosts[l] on l-1, s) within
                // the box [-s, 2^(l-1) + s)
                parallel_for(MAXL,
                             [&](int level)
                             {
                                 if (level < 1 || cells[size_t(level)].empty())
                                 {
                                     return;
                                 }
                                 const int n           = 1 << (level - 1);
                                 below1[size_t(level)] = clip_box(expand(on(cag[size_t(level)], level - 1), s), {-s, -s, -s},
                                                                  {n + s, n + s, n + s});
                             });
            }
            CellArray ref = empty_cell_array(dim);
            // phase 3: the reference before the projection ghosts, per level
            parallel_for(MAXL,
                         [&](int t)
                         {
                             LevelSet x = cag[size_t(t)];
                             if (mr && t + 1 < MAXL && !below1[size_t(t + 1)].empty())
                             {
                                 x = set_union(x, below1[size_t(t + 1)]);
                             }
                             if (mr && t + 2 < MAXL && !below2[size_t(t + 2)].empty())
                             {
                                 x = set_union(x, below2[size_t(t + 2)]);
                             }
                             ref[size_t(t)] = std::move(x);
                         });
            m.ids[CELLS_AND_GHOSTS] = std::move(cag);
            m.ids[PROJ_CELLS]       = empty_cell_array(dim);
            if (mr)
            {
                // phase 4: projection ghosts, level by level on the growing reference
                for (int level = 0; level < MAXL - 1 && level <= ca_max_level(ref); ++level)
                {
                    if (ref[size_t(level)].empty())
                    {
                        continue;
                    }
                    LevelSet expr = set_intersection(ref[size_t(level)], m.ids[UNION_CELLS][size_t(level)]);
                    ref[size_t(level + 1)]           = set_union(ref[size_t(level + 1)], on(expr, level + 1));
                    m.ids[PROJ_CELLS][size_t(level)] = std::move(expr);
                }
            }
            m.ids[REFERENCE] = std::move(ref);
        }

        // Mesh_base::construct_corners (mesh.hpp:913-997).  The corners depend on
        // the domain at max_level only, so they are computed once per (dim,
        // max_level) and shared by every mesh version.
        void construct_corners(MRMeshData& m)
        {
            static std::mutex mu;
            static std::map<std::pair<int, int>, std::pair<std::vector<LevelSet>, std::vector<Dir>>> cache;
            const std::pair<int, int> key{m.cfg.dim, m.cfg.max_level};
            {
                std::lock_guard<std::mutex> lk(mu);
                auto it = cache.find(key);
                if (it != cache.end())
                {
                    m.corners     = it->second.first;
                    m.corner_dirs = it->second.second;
                    return;
                }
            }
            m.corners.clear();
            m.corner_dirs.clear();
            const LevelSet dom = domain_at(m.cfg.dim, m.cfg.max_level);
            for (const Dir& d : diagonal_directions(m.cfg.dim))
            {
                LevelSet u(m.cfg.dim, m.cfg.max_level);
                for (int i = 0; i < m.cfg.dim; ++i)
                {
                    if (d[size_t(i)] != 0)
                    {
                        Dir t{0, 0, 0};
                        t[size_t(i)] = -d[size_t(i)];
                        u            = set_union(u, translate(dom, t));
                    }
                }
                m.corners.push_back(set_difference(dom, u));
                m.corner_dirs.push_back(d);
            }
            std::lock_guard<std::mutex> lk(mu);
            cache.emplace(key, std::make_pair(m.corners, m.corner_dirs));
        }

        // Mesh_base::renumbering / LevelCellArray::update_index
        // (mesh.hpp:893-911, level_cell_array.hpp:774-784)
        void renumbering(MRMeshData& m)
        {
            m.ridx.assign(MAXL, {});
            m.level_base.assign(MAXL + 1, 0);
            int64_t n = 0;
            for (int l = 0; l < MAXL; ++l)
            {
                m.level_base[size_t(l)] = n;
                const LevelSet& r       = m.ids[REFERENCE][size_t(l)];
                auto& idx               = m.ridx[size_t(l)];
                idx.resize(r.runs.size());
                for (size_t k = 0; k < r.runs.size(); ++k)
                {
                    idx[k] = n - r.runs[k].s;
                    n += r.runs[k].e - r.runs[k].s;
                }
            }
            m.level_base[MAXL] = n;
            m.nref             = n;
        }
    } // namespace

    void complete_mesh(MRMeshData& m)
    {
        construct_corners(m);
        renumbering(m);
    }

    MRMeshData build_mesh(const MeshConfig& cfg, const CellArray& cells)
    {
        return build_mesh(cfg, CellArray(cells));
    }

    MRMeshData build_mesh(const MeshConfig& cfg, CellArray&& cells)
    {
        // radius 2 or 3: ghost layers filled by polynomial extrapolation up to
        // stencil 6 (bc/polynomial_extrapolation.hpp:17), built for meshes of one level
        if (cfg.radius != 1 && !(cfg.radius >= 2 && cfg.radius <= 3 && cfg.min_level == cfg.max_level))
        {
            throw std::invalid_argument("a scheme stencil radius of 1 is supported, and of 2 or 3 on meshes with min_level == max_level");
        }
        if (cfg.pred_s < 0 || cfg.pred_s > 1)
        {
            throw std::invalid_argument("prediction stencil radius must be 0 or 1");
        }
        if (cfg.max_level < cfg.min_level)
        {
            throw std::invalid_argument("Max level must be greater than min level.");
        }
        if (cfg.min_level < 1 || cfg.max_level >= MAXL - 1)
        {
            throw std::invalid_argument("levels must satisfy 1 <= min_level <= max_level <= 19");
        }
        MRMeshData m;
        m.cfg = cfg;
        for (auto& v : m.ids)
        {
            v = empty_cell_array(cfg.dim);
        }
        if (cells.size() != size_t(MAXL))
        {
            throw std::invalid_argument("a cell array holds one level set per level");
        }
        m.ids[CELLS] = std::move(cells);
        update_sub_mesh(m);
        construct_corners(m);
        renumbering(m);
        return m;
    }

    MRMeshData build_box_mesh(const MeshConfig& cfg)
    {
        CellArray cells                = empty_cell_array(cfg.dim);
        cells[size_t(cfg.max_level)]   = domain_at(cfg.dim, cfg.max_level);
        MRMeshData m                   = build_mesh(cfg, cells);
        m.graduated                    = true; // a single level
        return m;
    }

    CellArray update_cell_array_from_tag(const MRMeshData& mesh, const uint8_t* tags)
    {
        const int dim = mesh.cfg.dim;
        // per level: cells added by coarsening (parents) and by refinement
        // (children), cells removed by coarsening and by refinement
        CellArray add_m = empty_cell_array(dim), rm_m = empty_cell_array(dim), add_p = empty_cell_array(dim),
                  rm_p = empty_cell_array(dim);
        const CellArray& cells = mesh.ids[CELLS];
        const int start = ca_min_level(cells), end = ca_max_level(cells) + 1;
        // level l only writes add_m[l-1], rm_m[l], add_p[l+1] and rm_p[l]: the
        // levels are independent.  Cells are visited in (z, y, x) order and
        // their actions gathered in stretches of equal action, so every set
        // except the 3D children comes out in canonical order and is pushed
        // directly; the 3D children (rows 2y, 2y+1 of planes 2z, 2z+1) are
        // collected and ordered by from_runs.
        parallel_for(std::max(end - start, 0), [&](int li)
        {
            const int level    = start + li;
            const LevelSet& cl = cells[size_t(level)];
            const bool can_refine = level < mesh.cfg.max_level, can_coarsen = level > mesh.cfg.min_level;
            LevelSet& rmm = rm_m[size_t(level)];
            LevelSet& rmp = rm_p[size_t(level)];
            LevelSet& amm = add_m[size_t(level > 0 ? level - 1 : 0)];
            LevelSet kids2(dim, level + 1);
            std::vector<std::array<int32_t, 4>> kids3;
            std::vector<std::array<int32_t, 2>> row_kids;
            rmm = LevelSet(dim, level);
            rmp = LevelSet(dim, level);
            if (level > 0)
            {
                amm = LevelSet(dim, level - 1);
            }
            for (size_t r = 0; r < cl.nrows(); ++r)
            {
                const int32_t y = cl.ry[r], z = cl.rz[r];
                const bool yz_even = (y % 2 == 0) && (dim == 2 || z % 2 == 0);
                row_kids.clear();
                // action of a stretch: 1 refine, 2 coarsen, 0 none
                auto flush = [&](int act, int32_t xa, int32_t xb)
                {
                    if (act == 1)
                    {
                        rmp.push(y, z, xa, xb);
                        row_kids.push_back({2 * xa, 2 * xb});
                    }
                    else if (act == 2)
                    {
                        rmm.push(y, z, xa, xb);
                        if (yz_even)
                        {
                            // parents of the even x of [xa, xb)
                            const int32_t e0 = xa + (xa & 1), e1 = (xb - 1) - ((xb - 1) & 1);
                            if (e0 <= e1)
                            {
                                amm.push(y >> 1, dim == 3 ? (z >> 1) : 0, e0 >> 1, (e1 >> 1) + 1);
                            }
                        }
                    }
                };
                for (int64_t k = cl.rp[r]; k < cl.rp[r + 1]; ++k)
                {
                    const Run run      = cl.runs[size_t(k)];
                    const int64_t base = mesh.index(level, run.s, y, z) - run.s;
                    int act = -1;
                    int32_t xa = run.s;
                    for (int32_t x = run.s; x < run.e; ++x)
                    {
                        const uint8_t t = tags[base + x];
                        int a           = 0;
                        if ((t & FLAG_REFINE) && can_refine)
                        {
                            a = 1;
                        }
                        else if ((t & FLAG_COARSEN) && !(t & FLAG_KEEP) && can_coarsen)
                        {
                            a = 2;
                        }
                        if (a != act)
                        {
                            flush(act, xa, x);
                            act = a;
                            xa  = x;
                        }
                    }
                    flush(act, xa, run.e);
                }
                if (!row_kids.empty())
                {
                    if (dim == 2)
                    {
                        for (int dy = 0; dy < 2; ++dy)
                        {
                            for (const auto& kx : row_kids)
                            {
                                kids2.push(2 * y + dy, 0, kx[0], kx[1]);
                            }
                        }
                    }
                    else
                    {
                        for (int dz = 0; dz < 2; ++dz)
                        {
                            for (int dy = 0; dy < 2; ++dy)
                            {
                                for (const auto& kx : row_kids)
                                {
                                    kids3.push_back({2 * y + dy, 2 * z + dz, kx[0], kx[1]});
                                }
                            }
                        }
                    }
                }
            }
            if (level + 1 < MAXL)
            {
                add_p[size_t(level + 1)] = dim == 2 ? std::move(kids2) : from_runs(dim, level + 1, kids3);
            }
        });
        CellArray out = empty_cell_array(dim);
        const int lmin = mesh.cfg.min_level, lmax = mesh.cfg.max_level;
        parallel_for(lmax - lmin + 1,
                     [&](int li)
                     {
                         const int level = lmin + li;
                         const size_t l  = size_t(level);
                         out[l] = set_difference(set_union(cells[l], add_m[l], add_p[l]), set_union(rm_m[l], rm_p[l]));
                     });
        return out;
    }

    int make_graduation(CellArray& ca, int dim, int grad_width)
    {
        const int max_level = ca_max_level(ca);
        const int min_level = ca_min_level(ca);
        int nit             = 0;
        bool changed        = true;
        while (changed)
        {
            ++nit;
            CellArray remove = empty_cell_array(dim);
            // list_interval_to_refine_for_graduation (graduation.hpp:245-330):
            // for every fine level f and every coarse level c <= f - 2, the
            // cells of c under on(expand(cells[f], 2w), c).  Coarsening
            // distributes over unions and composes, so the union over f of
            // those projections at level c is P_c = on(P_{c+1}, c) |
            // on(expand(cells[c + 2], 2w), c): one pass from the finest level
            // gives the same sets as the per-fine-level cascades
            // (the projections of each level and the intersections run in
            // parallel over the levels; only the chain of unions is serial)
            const int fmax = ca_max_level(ca) + 1;
            const int cmin = ca_min_level(ca);
            std::vector<LevelSet> proj(MAXL);
            parallel_for(MAXL,
                         [&](int c)
                         {
                             if (c >= cmin && c + 2 <= fmax && !ca[size_t(c + 2)].empty())
                             {
                                 proj[size_t(c)] = expand_on(ca[size_t(c + 2)], 2 * grad_width, c);
                             }
                         });
            for (int c = fmax - 3; c >= cmin; --c)
            {
                const LevelSet& finer = proj[size_t(c + 1)];
                if (!finer.empty())
                {
                    LevelSet p      = on(finer, c);
                    proj[size_t(c)] = proj[size_t(c)].empty() ? std::move(p) : set_union(proj[size_t(c)], p);
                }
            }
            parallel_for(MAXL,
                         [&](int c)
                         {
                             if (c >= cmin && !proj[size_t(c)].empty())
                             {
                                 remove[size_t(c)] = set_intersection(proj[size_t(c)], ca[size_t(c)]);
                             }
                         });
            bool any_change   = false;
            CellArray add_p   = empty_cell_array(dim);
            for (int level = min_level; level < max_level + 1; ++level)
            {
                if (!remove[size_t(level)].empty())
                {
                    any_change               = true;
                    add_p[size_t(level + 1)] = on(remove[size_t(level)], level + 1);
                }
            }
            if (!any_change)
            {
                changed = false;
                continue;
            }
            CellArray new_ca = empty_cell_array(dim);
            const int lo = std::min(ca_min_level(ca), ca_min_level(add_p));
            const int hi = std::max(ca_max_level(ca), ca_max_level(add_p));
            for (int level = lo; level < hi + 1; ++level)
            {
                new_ca[size_t(level)] = set_difference(set_union(ca[size_t(level)], add_p[size_t(level)]), remove[size_t(level)]);
            }
            changed = !ca_equal(new_ca, ca);
            std::swap(new_ca, ca);
        }
        return nit - 1;
    }
} // namespace smr
