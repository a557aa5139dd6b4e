// Oracle (test infrastructure only; parity checker and CPU baseline).
//
// A CPU restatement of samurai v0.33.0's adaptive-MR time step on the unit box:
// MRMesh construction, update_ghost_mr (with Dirichlet<1> and corner/BC
// projection/prediction), the Harten adaptation (detail, criteria, maximum,
// update_cell_array_from_tag, make_graduation, update_fields) and the upwind
// finite-volume step.  Sets are dense bitmaps (dense_set.hpp); every
// floating-point expression is written in the reference's operation order and
// the file is compiled with -ffp-contract=off.  Each function cites the
// reference file:line it follows (paths relative to the samurai source tree).
//
// Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg may load
// this library.  The product (samurai_amd/) never links or calls it.
//
// Pinning: the 2D path is checked against the reference's advection_2d golden
// fixtures (cells exact, values <= 1e-12).  In 2D and 3D, the operators are
// checked against the reference's own known answers and the mathematics
// (tests/test_oracle_pinning.py, tests/test_oracle_invariants.py,
// tests/test_convection.py): the projection/prediction round trip of
// per-axis quadratics (s = 0 and 1), quadratic averages reproduced by
// update_ghost_mr on adapted meshes, details vanishing on quadratics,
// Dirichlet<1> ghosts on uniform and adapted meshes, corner ghosts as the
// reflection of the field, the vertex-ghost cascade of project_corner_below,
// the textbook upwind update in 2D and 3D, the edge-ghost cascade of
// project_corner_below in 3D, graduation and tiling invariants.  Planted
// mistakes (ORACLE_PLANT below) check that these checks catch a swapped
// parity coefficient, a dropped flux, a wrong ghost source or sign.
#include "dense_set.hpp"

#include <cmath>
#include <cstdio>
#include <cstring>
#include <limits>
#include <memory>
#include <string>

// Mistakes planted on purpose, one per value of ORACLE_PLANT, for the mutation
// test of the pinning checks (tests/test_oracle_mutants.py builds each into a
// scratch library and expects the checks to fail).  0, the default, plants
// nothing.
#ifndef ORACLE_PLANT
#define ORACLE_PLANT 0
#endif

namespace oracle
{
    constexpr int kPlant = ORACLE_PLANT;
    constexpr int MAXL = 21; // CellArray max_size 20 -> levels 0..20 (samurai_config.hpp:50)

    enum MeshId
    {
        CELLS            = 0,
        CELLS_AND_GHOSTS = 1,
        PROJ_CELLS       = 2,
        UNION_CELLS      = 3,
        REFERENCE        = 4
    };

    enum Flag : uint8_t
    {
        KEEP    = 1,
        COARSEN = 2,
        REFINE  = 4
    };

    using Dir = std::array<int, 3>;

    // static_nested_loop<dim,-1,2> order (static_algorithm.hpp:18-37): the last
    // dimension is the outermost loop, x the innermost.
    inline std::vector<Dir> all_directions(int dim)
    {
        std::vector<Dir> out;
        if (dim == 2)
        {
            for (int y = -1; y <= 1; ++y)
            {
                for (int x = -1; x <= 1; ++x)
                {
                    if (x != 0 || y != 0)
                    {
                        out.push_back({x, y, 0});
                    }
                }
            }
        }
        else
        {
            for (int z = -1; z <= 1; ++z)
            {
                for (int y = -1; y <= 1; ++y)
                {
                    for (int x = -1; x <= 1; ++x)
                    {
                        if (x != 0 || y != 0 || z != 0)
                        {
                            out.push_back({x, y, z});
                        }
                    }
                }
            }
        }
        return out;
    }

    inline int n_nonzero(const Dir& d)
    {
        return (d[0] != 0) + (d[1] != 0) + (d[2] != 0);
    }

    // for_each_diagonal_direction (stencil.hpp:337-349)
    inline std::vector<Dir> diagonal_directions(int dim)
    {
        std::vector<Dir> out;
        for (auto& d : all_directions(dim))
        {
            if (n_nonzero(d) > 1)
            {
                out.push_back(d);
            }
        }
        return out;
    }

    // for_each_cartesian_direction (stencil.hpp:300-316): +e0, -e0, +e1, -e1, ...
    inline std::vector<Dir> cartesian_directions(int dim)
    {
        std::vector<Dir> out;
        for (int d = 0; d < dim; ++d)
        {
            Dir v{0, 0, 0};
            v[d] = 1;
            out.push_back(v);
            v[d] = -1;
            out.push_back(v);
        }
        return out;
    }

    inline Dir scale(const Dir& d, int k)
    {
        return {d[0] * k, d[1] * k, d[2] * k};
    }

    struct Mesh
    {
        int dim = 2, min_level = 0, max_level = 0;
        int pred_s = 1, radius = 1, grad_width = 1;
        std::vector<DSet> ids[5];
        std::vector<DSet> domain; // m_domain pyramid, levels 0..max_level
        std::vector<DSet> corners;
        std::vector<Dir> corner_dirs;
        // storage index of the reference cells
        std::vector<std::vector<int64_t>> idx;
        std::vector<int64_t> level_base;
        int64_t nref = 0;

        const DSet& get(int id, int l) const { return ids[id][size_t(l)]; }

        int64_t index(int l, int x, int y, int z = 0) const
        {
            const DSet& r = ids[REFERENCE][size_t(l)];
            if (l < 0 || l >= MAXL || !r.get(x, y, z))
            {
                char buf[160];
                std::snprintf(buf, sizeof(buf), "oracle: cell (%d; %d,%d,%d) not in the reference mesh", l, x, y, z);
                throw std::out_of_range(buf);
            }
            return idx[size_t(l)][size_t(r.off(x, y, z))];
        }
        bool exists(int l, int x, int y, int z = 0) const
        {
            return l >= 0 && l < MAXL && ids[REFERENCE][size_t(l)].get(x, y, z);
        }
        int ca_min_level(int id) const
        {
            for (int l = 0; l < MAXL; ++l)
            {
                if (!ids[id][size_t(l)].empty())
                {
                    return l;
                }
            }
            return MAXL;
        }
        int ca_max_level(int id) const
        {
            for (int l = MAXL - 1; l >= 0; --l)
            {
                if (!ids[id][size_t(l)].empty())
                {
                    return l;
                }
            }
            return 0;
        }
    };

    inline std::vector<DSet> empty_ca(int dim)
    {
        std::vector<DSet> v;
        for (int l = 0; l < MAXL; ++l)
        {
            v.push_back(empty_set(dim, l));
        }
        return v;
    }

    inline DSet domain_at(int dim, int level)
    {
        int n = 1 << level;
        return box_set(dim, level, {0, 0, 0}, {n, n, dim == 3 ? n : 1});
    }

    // Mesh_base::construct_union (mesh.hpp:1231-1262)
    static void construct_union(Mesh& m)
    {
        auto& uni                  = m.ids[UNION_CELLS];
        uni                        = empty_ca(m.dim);
        for (int level = m.max_level; level >= 1; --level)
        {
            uni[size_t(level - 1)] = tighten(on(union_(m.ids[CELLS][size_t(level)], uni[size_t(level)]), level - 1));
        }
    }

    // MRMesh::update_sub_mesh_impl (mr/mesh.hpp:221-455), serial non-periodic.
    static void update_sub_mesh(Mesh& m)
    {
        const int r = m.radius, s = m.pred_s;
        std::vector<DSet> cell_list = empty_ca(m.dim);
        for (int l = 0; l < MAXL; ++l)
        {
            if (!m.ids[CELLS][size_t(l)].empty())
            {
                // cells expanded by the stencil radius in x and in every row offset (:240-253)
                cell_list[size_t(l)] = expand(m.ids[CELLS][size_t(l)], r);
            }
        }
        m.ids[CELLS_AND_GHOSTS] = cell_list;
        m.ids[PROJ_CELLS]       = empty_ca(m.dim);

        if (m.max_level != m.min_level)
        {
            for (int level = 0; level < MAXL; ++level)
            {
                const DSet& cl = m.ids[CELLS][size_t(level)];
                if (cl.empty())
                {
                    continue;
                }
                // prediction ghosts one level below (:331-342, :355-361)
                DSet sub_cg = intersection(expand(on(m.ids[CELLS_AND_GHOSTS][size_t(level)], level - 1), s),
                                           expand(m.domain[size_t(level - 1)], s));
                cell_list[size_t(level - 1)] = tighten(union_(cell_list[size_t(level - 1)], on(sub_cg, level - 1)));
                if (level - 1 > 0)
                {
                    DSet sub_c                   = expand(on(cl, level - 2), s);
                    cell_list[size_t(level - 2)] = tighten(union_(cell_list[size_t(level - 2)], sub_c));
                }
            }
            m.ids[REFERENCE] = cell_list;

            // projection ghosts (:432-452); the level loop re-reads the (growing)
            // reference like for_each_level does (algorithm.hpp:52-61).
            for (int level = 0; level < MAXL - 1; ++level)
            {
                if (level > m.ca_max_level(REFERENCE))
                {
                    break;
                }
                if (m.ids[REFERENCE][size_t(level)].empty())
                {
                    continue;
                }
                DSet expr  = tighten(on(intersection(m.ids[REFERENCE][size_t(level)], m.ids[UNION_CELLS][size_t(level)]), level));
                DSet& lcl  = cell_list[size_t(level + 1)];
                lcl        = tighten(union_(lcl, refine(expr, 1)));
                m.ids[REFERENCE][size_t(level + 1)] = lcl;
                m.ids[PROJ_CELLS][size_t(level)]    = expr;
            }
        }
        m.ids[REFERENCE] = cell_list;
    }

    // Mesh_base::construct_corners (mesh.hpp:913-997)
    static void construct_corners(Mesh& m)
    {
        m.corners.clear();
        m.corner_dirs.clear();
        const DSet& dom = m.domain[size_t(m.max_level)];
        for (auto& d : diagonal_directions(m.dim))
        {
            DSet u = empty_set(m.dim, m.max_level);
            for (int i = 0; i < m.dim; ++i)
            {
                if (d[i] != 0)
                {
                    Dir t{0, 0, 0};
                    t[i] = -d[i];
                    u    = union_(u, translate(dom, t));
                }
            }
            m.corners.push_back(tighten(difference(dom, u)));
            m.corner_dirs.push_back(d);
        }
    }

    // Mesh_base::renumbering + LevelCellArray::update_index (mesh.hpp:893-911,
    // level_cell_array.hpp:774-784): cumulative count over (level, z, y, x).
    static void renumbering(Mesh& m)
    {
        m.idx.assign(MAXL, {});
        m.level_base.assign(MAXL + 1, 0);
        int64_t n = 0;
        for (int l = 0; l < MAXL; ++l)
        {
            m.level_base[size_t(l)] = n;
            const DSet& r           = m.ids[REFERENCE][size_t(l)];
            m.idx[size_t(l)].assign(size_t(r.volume()), -1);
            r.for_each(
                [&](int x, int y, int z)
                {
                    m.idx[size_t(l)][size_t(r.off(x, y, z))] = n++;
                });
        }
        m.level_base[MAXL] = n;
        m.nref             = n;
    }

    // Mesh_base(ca, cfg) / Mesh_base(ca, ref_mesh) -> finalize_mesh (mesh.hpp:386-439)
    static Mesh make_mesh_from_cells(const Mesh& cfg, const std::vector<DSet>& cells)
    {
        Mesh m;
        m.dim        = cfg.dim;
        m.min_level  = cfg.min_level;
        m.max_level  = cfg.max_level;
        m.pred_s     = cfg.pred_s;
        m.radius     = cfg.radius;
        m.grad_width = cfg.grad_width;
        m.domain.clear();
        for (int l = 0; l <= m.max_level; ++l)
        {
            m.domain.push_back(domain_at(m.dim, l));
        }
        for (auto& v : m.ids)
        {
            v = empty_ca(m.dim);
        }
        for (int l = 0; l < MAXL; ++l)
        {
            m.ids[CELLS][size_t(l)] = tighten(cells[size_t(l)]);
        }
        construct_union(m);
        update_sub_mesh(m);
        construct_corners(m);
        renumbering(m);
        return m;
    }

    struct Sim
    {
        Mesh mesh;
        std::vector<double> u;
        std::vector<double> detail;
        std::vector<uint8_t> tag;
        // fields adapted together with u (make_MRAdapt(u, v...),
        // mr/adapt.hpp:138-146): their storage and their details
        std::vector<std::vector<double>> extra, extra_detail;
        bool has_bc     = false;
        // mra_config().relative_detail(true) (mr/config.hpp:36-45)
        bool relative_detail = false;
        // the --refine-boundary option (arguments.hpp)
        bool refine_boundary = false;
        double bc_value = 0.0;
        bool poison     = true; // NaN-fill storage nobody wrote (SAMURAI_CHECK_NAN idea)
        int last_harten_iterations = 0;
        // the tags/details of the last harten call, on the mesh they were computed on
        Mesh tag_mesh;
        std::vector<uint8_t> last_tag;
        std::vector<double> last_detail;
    };

    inline double& U(Sim& s, int l, int x, int y, int z = 0)
    {
        return s.u[size_t(s.mesh.index(l, x, y, z))];
    }

    // ------------------------------------------------------------------
    // ghost update (algorithm/update_ghost_mr.hpp:194-237)
    // ------------------------------------------------------------------

    // Dirichlet<1>: u[ghost] = 2*v - u[cell] (bc/dirichlet.hpp:29) on the cells of
    // the boundary region (bc/apply_field_bc.hpp:53-101, bc/bc.hpp:326-363).
    static void apply_field_bc(Sim& s, int level, const Dir& dir)
    {
        if (!s.has_bc)
        {
            return;
        }
        const Mesh& m    = s.mesh;
        const DSet& dom  = m.domain[size_t(m.max_level)];
        DSet region      = difference(dom, translate(dom, scale(dir, -1)));
        DSet bdry        = on(intersection(m.ids[CELLS][size_t(level)], region), level);
        const double val = s.bc_value;
        bdry.for_each(
            [&](int x, int y, int z)
            {
                double uin                                  = U(s, level, x, y, z);
                U(s, level, x + dir[0], y + dir[1], z + dir[2]) = kPlant == 5 ? 2 * val + uin : 2 * val - uin;
            });
    }

    // ghost width of the mesh: max(stencil radius, prediction radius)
    // (mesh_config.hpp:395)
    inline int ghost_width(const Mesh& m)
    {
        return std::max(m.radius, m.pred_s);
    }

    // PolynomialExtrapolation<stencil 2k> (bc/polynomial_extrapolation.hpp:22-74):
    // the ghost after the n = 2k - 1 known values v[0..n), in the reference's
    // operation order (degree n - 1 through them, evaluated one step further)
    inline double extrapolate(const double* v, int n)
    {
        if (n == 1)
        {
            return v[0];
        }
        if (n == 3)
        {
            return kPlant == 11 ? v[0] - v[2] * 3.0 + v[1] * 3.0 : v[0] - v[1] * 3.0 + v[2] * 3.0;
        }
        if (n == 5)
        {
            return v[0] - v[1] * 5.0 + v[2] * (kPlant == 10 ? 9.0 : 1.0E+1) - v[3] * 1.0E+1 + v[4] * 5.0;
        }
        throw std::invalid_argument("oracle: polynomial extrapolation beyond stencil 6");
    }

    // apply_extrapolation_bc_cells<2k> (bc/apply_field_bc.hpp:124-150) in
    // direction d (Cartesian or diagonal; the line stencil of the B.C. turned
    // towards d, stencil.hpp:226-231, 498-561, is {-(k-1)..k} * d): for each
    // cell b of `at` (whose ghost b + k d exists when k > 1), the ghost
    // b + k d from the cells b + j d, j = -(k-1)..k-1.
    static void extrapolate_layer(Sim& s, int level, const DSet& at, const Dir& d, int k)
    {
        const Mesh& m = s.mesh;
        at.for_each(
            [&](int x, int y, int z)
            {
                if (k > 1 && !m.exists(level, x + k * d[0], y + k * d[1], z + k * d[2]))
                {
                    return;
                }
                double v[5];
                for (int j = -(k - 1); j <= k - 1; ++j)
                {
                    v[j + k - 1] = U(s, level, x + j * d[0], y + j * d[1], z + j * d[2]);
                }
                U(s, level, x + k * d[0], y + k * d[1], z + k * d[2]) = extrapolate(v, 2 * k - 1);
            });
    }

    // update_outer_corners_by_polynomial_extrapolation(level, direction)
    // (bc/apply_field_bc.hpp:315-466; polynomial_extrapolation.hpp:63-72):
    // layer k = 1..w of the diagonal through each corner cell by the
    // extrapolation of stencil 2k along the diagonal (layer 1: a copy of the
    // cell), then each off-diagonal cell of layer k (first non-zero axis at
    // offset k, the others at offsets 1..w) copies the diagonal cell of layer k.
    static void corner_extrapolation(Sim& s, int level, size_t c)
    {
        const Mesh& m = s.mesh;
        const Dir& d  = m.corner_dirs[c];
        const int w   = ghost_width(m);
        DSet corner   = on(m.corners[c], level);
        DSet cells    = on(intersection(m.ids[CELLS][size_t(level)], corner), level);
        cells.for_each(
            [&](int x, int y, int z)
            {
                U(s, level, x + d[0], y + d[1], z + d[2]) = kPlant == 4 ? U(s, level, x + d[0], y, z) : U(s, level, x, y, z);
            });
        for (int k = 2; k <= w; ++k)
        {
            extrapolate_layer(s, level, cells, d, k);
        }
        // step 2 (:368-464): offsets (g_p - (k-1)) along the non-first axes
        std::vector<int> nz;
        for (int i = 0; i < m.dim; ++i)
        {
            if (d[i] != 0)
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
                    delta[nz[p]] += dp * d[nz[p]];
                    if (dp != 0)
                    {
                        is_source = false;
                    }
                }
                if (is_source)
                {
                    continue;
                }
                cells.for_each(
                    [&](int x, int y, int z)
                    {
                        const int kk = kPlant == 13 ? k - 1 : k;
                        const int sx = x + kk * d[0], sy = y + kk * d[1], sz = z + kk * d[2];
                        U(s, level, sx + delta[0], sy + delta[1], sz + delta[2]) = U(s, level, sx, sy, sz);
                    });
            }
        }
    }

    // update_further_ghosts_by_polynomial_extrapolation(level, direction)
    // (bc/apply_field_bc.hpp:499-563), first part: the ghost layers beyond
    // those the B.C. fills (Dirichlet<1>: one; no B.C.: none) by the
    // extrapolation of stencil 2 * layer from the boundary cells.  The second
    // part (layers >= 2 next to the union cells' ghosts) has nothing to do
    // where the ghost width is 1 or the mesh has one level, the only meshes
    // built here.
    static void further_ghosts(Sim& s, int level, const Dir& dir)
    {
        const Mesh& m   = s.mesh;
        const int w     = ghost_width(m);
        const int first = s.has_bc ? 2 : 1;
        if (first > w)
        {
            return;
        }
        const DSet& dom = m.domain[size_t(level)];
        DSet bdry       = difference(m.ids[CELLS][size_t(level)], translate(dom, scale(dir, -1)));
        for (int layer = first; layer <= w; ++layer)
        {
            extrapolate_layer(s, level, bdry, dir, layer);
        }
    }

    // project_corner_below (algorithm/update_outer_ghost.hpp:267-336)
    static void project_corner_below(Sim& s, int level, size_t c)
    {
        const Mesh& m = s.mesh;
        const Dir& d  = m.corner_dirs[c];
        if (level == 0)
        {
            return;
        }
        for (int dl = 1; dl <= 2; ++dl)
        {
            int pl                 = level - dl;
            DSet fine_inner        = on(m.corners[c], level);
            DSet fine_outer        = intersection(translate(fine_inner, d), m.ids[REFERENCE][size_t(level)]);
            DSet proj_ghost        = intersection(on(fine_outer, pl), m.ids[REFERENCE][size_t(pl)]);
            const int last         = (1 << dl) - 1;
            proj_ghost.for_each(
                [&](int x, int y, int z)
                {
                    int cx = (x << dl) + (d[0] == -1 ? last : 0);
                    int cy = (y << dl) + (d[1] == -1 ? last : 0);
                    int cz = m.dim == 3 ? ((z << dl) + (d[2] == -1 ? (kPlant == 6 ? 0 : last) : 0)) : 0;
                    if (m.exists(level, cx, cy, cz))
                    {
                        U(s, pl, x, y, z) = U(s, level, cx, cy, cz);
                    }
                });
            if (pl == 0)
            {
                break;
            }
        }
    }

    // project_bc for layer 1 on a box domain (algorithm/update_outer_ghost.hpp:20-152)
    static void project_bc(Sim& s, int pl, const Dir& dir)
    {
        const Mesh& m  = s.mesh;
        DSet outside   = difference(translate(m.ids[UNION_CELLS][size_t(pl)], dir), m.domain[size_t(pl)]);
        DSet ghosts    = on(intersection(outside, m.ids[REFERENCE][size_t(pl)]), pl);
        ghosts.for_each(
            [&](int x, int y, int z)
            {
                U(s, pl, x, y, z) = 0;
            });
        ghosts.for_each(
            [&](int x, int y, int z)
            {
                double& g      = U(s, pl, x, y, z);
                int n_children = 0;
                for (int cl = pl + 1; cl <= pl + 2; ++cl)
                {
                    if (cl >= MAXL)
                    {
                        break;
                    }
                    DSet one = box_set(m.dim, pl, {x, y, z}, {x + 1, y + 1, z + 1});
                    DSet ch  = intersection(on(one, cl), m.ids[REFERENCE][size_t(cl)]);
                    ch.for_each(
                        [&](int cx, int cy, int cz)
                        {
                            g += U(s, cl, cx, cy, cz);
                            n_children++;
                        });
                    if (n_children > 0)
                    {
                        break;
                    }
                }
                if (n_children > 0)
                {
                    g /= n_children;
                }
            });
    }

    // domain_boundary_outer_layer(mesh, level, direction, 1) (boundary.hpp:79-99)
    static DSet outer_layer(const Mesh& m, int level, const Dir& dir)
    {
        const DSet& dom = m.domain[size_t(level)];
        DSet inner      = difference(m.ids[CELLS][size_t(level)], translate(dom, scale(dir, -1)));
        return difference(translate(inner, dir), dom);
    }

    // predict_bc (algorithm/update_outer_ghost.hpp:154-262): order-0 copy to level+1
    static void predict_bc(Sim& s, int pred_level, const Dir& dir)
    {
        const Mesh& m  = s.mesh;
        DSet bc_ghosts = outer_layer(m, pred_level - 1, dir);
        DSet ghosts    = on(intersection(bc_ghosts, m.ids[REFERENCE][size_t(pred_level)]), pred_level);
        ghosts.for_each(
            [&](int x, int y, int z)
            {
                U(s, pred_level, x, y, z) = U(s, pred_level - 1, x >> 1, y >> 1, m.dim == 3 ? (z >> 1) : 0);
            });
    }

    // update_outer_ghosts(level) (algorithm/update_outer_ghost.hpp:338-432)
    static void update_outer_ghosts(Sim& s, int level)
    {
        const Mesh& m = s.mesh;
        if (level <= m.max_level && level >= m.min_level)
        {
            for (size_t c = 0; c < m.corner_dirs.size(); ++c)
            {
                corner_extrapolation(s, level, c);
                project_corner_below(s, level, c);
            }
        }
        for (auto& dir : cartesian_directions(m.dim))
        {
            if (level < m.max_level)
            {
                project_bc(s, level, dir);
            }
            if (level >= m.min_level)
            {
                apply_field_bc(s, level, dir);
            }
            if (level < m.max_level && level >= m.min_level)
            {
                predict_bc(s, level + 1, dir);
            }
        }
        // the layers the B.C. leaves (update_outer_ghost.hpp:414-426)
        if (level >= m.min_level)
        {
            for (auto& dir : cartesian_directions(m.dim))
            {
                further_ghosts(s, level, dir);
            }
        }
    }

    // projection_op_ (numeric/projection.hpp:15-65)
    static double project_value(const Sim& s, int lf, int x, int y, int z)
    {
        const Mesh& m    = s.mesh;
        const double inv = 1.0 / double(1ULL << m.dim);
        double sum       = 0;
        if (m.dim == 2)
        {
            for (int r = 0; r < 2; ++r)
            {
                sum += s.u[size_t(m.index(lf, 2 * x, 2 * y + r))] + s.u[size_t(m.index(lf, 2 * x + 1, 2 * y + r))];
            }
        }
        else
        {
            for (int rz = 0; rz < 2; ++rz)
            {
                for (int ry = 0; ry < 2; ++ry)
                {
                    sum += s.u[size_t(m.index(lf, 2 * x, 2 * y + ry, 2 * z + rz))]
                         + s.u[size_t(m.index(lf, 2 * x + 1, 2 * y + ry, 2 * z + rz))];
                }
            }
        }
        return sum * inv;
    }

    inline std::array<std::array<double, 3>, 2> interp_pair()
    {
        // interp_coeffs<3>(+1), interp_coeffs<3>(-1) (numeric/prediction.hpp:31-35)
        return {{{1. / 8., 1, -1. / 8.}, {-1. / 8., 1, 1. / 8.}}};
    }

    // prediction_op <s,false> (numeric/prediction.hpp:259-357): value of the fine
    // cell (l, x, y, z) from the coarse stencil at l-1.
    static double predict_fine(const Sim& s, int l, int x, int y, int z)
    {
        const Mesh& m = s.mesh;
        const int xc = x >> 1, yc = y >> 1, zc = m.dim == 3 ? (z >> 1) : 0;
        if (m.pred_s == 0)
        {
            return s.u[size_t(m.index(l - 1, xc, yc, zc))];
        }
        auto pair = interp_pair();
        int px = x & 1, py = y & 1, pz = z & 1;
        double v = 0;
        if (m.dim == 2)
        {
            for (int sy = 0; sy < 3; ++sy)
            {
                for (int sx = 0; sx < 3; ++sx)
                {
                    double f     = s.u[size_t(m.index(l - 1, xc + sx - 1, yc + sy - 1))];
                    double coeff = 1.;
                    coeff *= pair[size_t(kPlant == 8 ? 1 - px : px)][size_t(sx)];
                    coeff *= pair[size_t(py)][size_t(sy)];
                    v += f * coeff;
                }
            }
        }
        else
        {
            for (int sz = 0; sz < 3; ++sz)
            {
                for (int sy = 0; sy < 3; ++sy)
                {
                    for (int sx = 0; sx < 3; ++sx)
                    {
                        double f     = s.u[size_t(m.index(l - 1, xc + sx - 1, yc + sy - 1, zc + sz - 1))];
                        double coeff = 1.;
                        coeff *= pair[size_t(px)][size_t(sx)];
                        coeff *= pair[size_t(py)][size_t(sy)];
                        coeff *= pair[size_t(kPlant == 1 ? 1 - pz : pz)][size_t(sz)];
                        v += f * coeff;
                    }
                }
            }
        }
        return v;
    }

    // update_ghost_mr_aggregated (algorithm/update_ghost_mr.hpp:194-237), serial.
    void update_ghost_mr(Sim& s)
    {
        const Mesh& m = s.mesh;
        for (int level = m.max_level; level >= 0; --level)
        {
            update_outer_ghosts(s, level);
            if (level > 0)
            {
                DSet set = on(intersection(m.ids[REFERENCE][size_t(level)], m.ids[PROJ_CELLS][size_t(level - 1)]), level - 1);
                set.for_each(
                    [&](int x, int y, int z)
                    {
                        double v          = project_value(s, level, x, y, z);
                        U(s, level - 1, x, y, z) = v;
                    });
            }
        }
        for (int level = 1; level <= m.max_level; ++level)
        {
            DSet pred_ghosts = difference(m.ids[REFERENCE][size_t(level)],
                                          union_(m.ids[CELLS][size_t(level)], m.ids[PROJ_CELLS][size_t(level)]));
            DSet expr = on(intersection(intersection(pred_ghosts, m.domain[size_t(level)]), m.ids[REFERENCE][size_t(level - 1)]),
                           level);
            expr.for_each(
                [&](int x, int y, int z)
                {
                    double v             = predict_fine(s, level, x, y, z);
                    U(s, level, x, y, z) = v;
                });
        }
    }

    // ------------------------------------------------------------------
    // adaptation (mr/adapt.hpp:147-389)
    // ------------------------------------------------------------------

    // compute_detail_op (mr/operators.hpp:139-540) for the coarse cell (l; x,y,z)
    static void compute_detail(Sim& s, int l, int x, int y, int z)
    {
        const Mesh& m = s.mesh;
        const int lf  = l + 1;
        const int nch = 1 << m.dim;
        int64_t fidx[8];
        for (int c = 0; c < nch; ++c)
        {
            int cx = 2 * x + (c & 1), cy = 2 * y + ((c >> 1) & 1), cz = m.dim == 3 ? 2 * z + ((c >> 2) & 1) : 0;
            fidx[c] = m.index(lf, cx, cy, cz);
        }
        if (m.pred_s == 0)
        {
            double uc = s.u[size_t(m.index(l, x, y, z))];
            for (int c = 0; c < nch; ++c)
            {
                s.detail[size_t(fidx[c])] = s.u[size_t(fidx[c])] - uc;
            }
            return;
        }
        auto pair   = interp_pair();
        auto& ie    = pair[0];
        auto& io    = pair[1];
        double d[8];
        for (int c = 0; c < nch; ++c)
        {
            d[c] = s.u[size_t(fidx[c])];
        }
        if (m.dim == 2)
        {
            for (int kj = 0; kj < 3; ++kj)
            {
                for (int ki = 0; ki < 3; ++ki)
                {
                    double src = s.u[size_t(m.index(l, x + ki - 1, y + kj - 1))];
                    double c4[4] = {ie[size_t(ki)] * ie[size_t(kj)], io[size_t(ki)] * ie[size_t(kj)], ie[size_t(ki)] * io[size_t(kj)],
                                    io[size_t(ki)] * io[size_t(kj)]};
                    for (int k = 0; k < 4; ++k)
                    {
                        d[k] -= c4[k] * src;
                    }
                }
            }
        }
        else
        {
            for (int kk = 0; kk < 3; ++kk)
            {
                for (int kj = 0; kj < 3; ++kj)
                {
                    for (int ki = 0; ki < 3; ++ki)
                    {
                        double src = s.u[size_t(m.index(l, x + ki - 1, y + kj - 1, z + kk - 1))];
                        const double* cx[2] = {ie.data(), io.data()};
                        double c8[8];
                        for (int c = 0; c < 8; ++c)
                        {
                            c8[c] = cx[c & 1][ki] * cx[(c >> 1) & 1][kj] * cx[kPlant == 2 ? 1 - ((c >> 2) & 1) : (c >> 2) & 1][kk];
                        }
                        for (int c = 0; c < 8; ++c)
                        {
                            d[c] -= c8[c] * src;
                        }
                    }
                }
            }
        }
        for (int c = 0; c < nch; ++c)
        {
            s.detail[size_t(fidx[c])] = d[c];
        }
    }

    // mr_criteria_op (mr/criteria.hpp:13-122) for the coarse cell (level; x,y,z)
    static void criteria(Sim& s, int level, int x, int y, int z, double eps, double regularity)
    {
        const Mesh& m          = s.mesh;
        const int fine_level   = level + 1;
        const int nch          = 1 << m.dim;
        const double fine_eps  = std::pow(2.0, regularity) * eps;
        const double coarse_eps = fine_eps / (1 << m.dim);
        int64_t fidx[8];
        for (int c = 0; c < nch; ++c)
        {
            fidx[c] = m.index(fine_level, 2 * x + (c & 1), 2 * y + ((c >> 1) & 1), m.dim == 3 ? 2 * z + ((c >> 2) & 1) : 0);
        }
        // the details of every adapted field (the n_comp loops of
        // criteria.hpp:47-56, 63-72, 96-105)
        std::vector<const std::vector<double>*> comps{&s.detail};
        for (const auto& d : s.extra_detail)
        {
            comps.push_back(&d);
        }
        if (fine_level > m.min_level)
        {
            bool cond = false;
            for (const auto* d : comps)
            {
                if (std::abs((*d)[size_t(m.index(level, x, y, z))]) > coarse_eps)
                {
                    cond = true;
                    break;
                }
            }
            if (!cond)
            {
                bool coarsen_ok = true;
                for (int c = 0; c < nch; ++c)
                {
                    for (const auto* d : comps)
                    {
                        if (std::abs((*d)[size_t(fidx[c])]) > eps)
                        {
                            coarsen_ok = false;
                        }
                    }
                }
                if (coarsen_ok)
                {
                    for (int c = 0; c < nch; ++c)
                    {
                        s.tag[size_t(fidx[c])] = COARSEN;
                    }
                }
            }
        }
        if (fine_level < m.max_level)
        {
            for (int c = 0; c < nch; ++c)
            {
                bool r = kPlant == 9; // plant 9: refine only when every field asks
                for (const auto* d : comps)
                {
                    const bool above = std::abs((*d)[size_t(fidx[c])]) > fine_eps;
                    r                = kPlant == 9 ? (r && above) : (r || above);
                }
                s.tag[size_t(fidx[c])] |= uint8_t(r) * REFINE;
            }
        }
    }

    // maximum_op (mr/operators.hpp:22-96) for the coarse cell (level; x,y,z)
    static void maximum(Sim& s, int level, int x, int y, int z)
    {
        const Mesh& m = s.mesh;
        const int nch = 1 << m.dim;
        int64_t fidx[8];
        for (int c = 0; c < nch; ++c)
        {
            fidx[c] = m.index(level + 1, 2 * x + (c & 1), 2 * y + ((c >> 1) & 1), m.dim == 3 ? 2 * z + ((c >> 2) & 1) : 0);
        }
        int64_t ci = m.index(level, x, y, z);
        bool keep  = false;
        for (int c = 0; c < nch; ++c)
        {
            keep = keep || (s.tag[size_t(fidx[c])] & KEEP);
        }
        if (keep)
        {
            for (int c = 0; c < nch; ++c)
            {
                s.tag[size_t(fidx[c])] |= KEEP;
            }
            s.tag[size_t(ci)] |= KEEP;
            return;
        }
        bool all_coarsen = true;
        for (int c = 0; c < nch; ++c)
        {
            all_coarsen = all_coarsen && (s.tag[size_t(fidx[c])] & COARSEN);
        }
        if (all_coarsen)
        {
            s.tag[size_t(ci)] |= KEEP;
        }
        else
        {
            for (int c = 0; c < nch; ++c)
            {
                s.tag[size_t(fidx[c])] &= uint8_t(~COARSEN);
            }
        }
    }

    // update_cell_array_from_tag (algorithm/graduation.hpp:743-843)
    static std::vector<DSet> update_cell_array_from_tag(const Sim& s)
    {
        const Mesh& m = s.mesh;
        const int dim = m.dim;
        std::vector<DSet> add_m = empty_ca(dim), rm_m = empty_ca(dim), add_p = empty_ca(dim), rm_p = empty_ca(dim);
        // build each as a dense grid over the domain (+ nothing outside: cells are inside)
        for (int l = 0; l < MAXL; ++l)
        {
            int n = 1 << std::min(l, 30);
            if (l <= m.max_level + 1)
            {
                rm_p[size_t(l)] = box_set(dim, l, {0, 0, 0}, {n, n, dim == 3 ? n : 1});
                std::fill(rm_p[size_t(l)].b.begin(), rm_p[size_t(l)].b.end(), 0);
                rm_m[size_t(l)] = rm_p[size_t(l)];
                add_p[size_t(l)] = rm_p[size_t(l)];
                add_m[size_t(l)] = rm_p[size_t(l)];
            }
        }
        const int start = m.ca_min_level(CELLS), end = m.ca_max_level(CELLS) + 1;
        for (int level = start; level < end; ++level)
        {
            m.ids[CELLS][size_t(level)].for_each(
                [&](int x, int y, int z)
                {
                    uint8_t t       = s.tag[size_t(m.index(level, x, y, z))];
                    bool refine     = t & REFINE;
                    bool coarsen_nk = (t & COARSEN) && !(t & KEEP);
                    if (refine && level < m.max_level)
                    {
                        rm_p[size_t(level)].set(x, y, z);
                        for (int c = 0; c < (1 << dim); ++c)
                        {
                            add_p[size_t(level + 1)].set(2 * x + (c & 1), 2 * y + ((c >> 1) & 1), dim == 3 ? 2 * z + ((c >> 2) & 1) : 0);
                        }
                    }
                    else if (coarsen_nk && level > m.min_level)
                    {
                        bool yz_even = (y % 2 == 0) && (dim == 2 || z % 2 == 0);
                        if (x % 2 == 0 && yz_even)
                        {
                            add_m[size_t(level - 1)].set(x >> 1, y >> 1, dim == 3 ? (z >> 1) : 0);
                        }
                        rm_m[size_t(level)].set(x, y, z);
                    }
                });
        }
        std::vector<DSet> out = empty_ca(dim);
        for (int level = m.min_level; level <= m.max_level; ++level)
        {
            DSet a = union_(union_(m.ids[CELLS][size_t(level)], add_m[size_t(level)]), add_p[size_t(level)]);
            DSet r = union_(rm_m[size_t(level)], rm_p[size_t(level)]);
            out[size_t(level)] = tighten(difference(a, r));
        }
        return out;
    }

    static int ca_min(const std::vector<DSet>& ca)
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
    static int ca_max(const std::vector<DSet>& ca)
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
    static bool ca_equal(const std::vector<DSet>& a, const std::vector<DSet>& b)
    {
        for (int l = 0; l < MAXL; ++l)
        {
            if (!same_cells(a[size_t(l)], b[size_t(l)]))
            {
                return false;
            }
        }
        return true;
    }

    // make_graduation (algorithm/graduation.hpp:573-726) with the graduation
    // listing of list_interval_to_refine_for_graduation (:245-330); the
    // contiguous-boundary refinement (:369-497) is a no-op for stencil radius 1.
    static int make_graduation(std::vector<DSet>& ca, int dim, int grad_width)
    {
        const int max_level = ca_max(ca);
        const int min_level = ca_min(ca);
        int nit             = 0;
        bool changed        = true;
        while (changed)
        {
            ++nit;
            std::vector<DSet> remove = empty_ca(dim);
            {
                const int fmax     = ca_max(ca) + 1;
                const int cmin     = ca_min(ca);
                const int min_fine = (cmin + 2) - 1;
                for (int fine_level = fmax; fine_level > min_fine; --fine_level)
                {
                    if (fine_level >= MAXL || ca[size_t(fine_level)].empty())
                    {
                        continue;
                    }
                    DSet proj = on(expand(ca[size_t(fine_level)], 2 * grad_width), fine_level - 2);
                    for (int coarse_level = fine_level - 2;; --coarse_level)
                    {
                        if (!proj.empty())
                        {
                            remove[size_t(coarse_level)] = union_(remove[size_t(coarse_level)], intersection(proj, ca[size_t(coarse_level)]));
                        }
                        if (coarse_level == cmin || proj.empty())
                        {
                            break;
                        }
                        proj = on(proj, coarse_level - 1);
                    }
                }
            }
            bool any_change = false;
            std::vector<DSet> add_p = empty_ca(dim);
            for (int level = min_level; level < max_level + 1; ++level)
            {
                remove[size_t(level)] = tighten(remove[size_t(level)]);
                if (!remove[size_t(level)].empty())
                {
                    any_change = true;
                    add_p[size_t(level + 1)] = refine(remove[size_t(level)], 1);
                }
            }
            if (!any_change)
            {
                changed = false;
                continue;
            }
            std::vector<DSet> new_ca = empty_ca(dim);
            int lo = std::min(ca_min(ca), ca_min(add_p)), hi = std::max(ca_max(ca), ca_max(add_p));
            for (int level = lo; level < hi + 1; ++level)
            {
                new_ca[size_t(level)] = tighten(difference(union_(ca[size_t(level)], add_p[size_t(level)]), remove[size_t(level)]));
            }
            changed = !ca_equal(new_ca, ca);
            std::swap(new_ca, ca);
        }
        return nit - 1;
    }

    static std::vector<double> new_storage(const Sim& s, int64_t n)
    {
        return std::vector<double>(size_t(n), s.poison ? std::numeric_limits<double>::quiet_NaN() : 0.0);
    }

    // prediction_op <s,true> (numeric/prediction.hpp:149-257): children of the
    // coarse cell (l-1; xc,yc,zc) of the new mesh from the old field.
    static void predict_children(const Mesh& om, const std::vector<double>& src, const Mesh& nm, std::vector<double>& dst,
                                 int lc, int xc, int yc, int zc)
    {
        const int dim = om.dim, nch = 1 << dim;
        double vals[8] = {0, 0, 0, 0, 0, 0, 0, 0};
        if (om.pred_s == 0)
        {
            double v = src[size_t(om.index(lc, xc, yc, zc))];
            for (int c = 0; c < nch; ++c)
            {
                vals[c] = v;
            }
        }
        else
        {
            auto pair = interp_pair();
            const int nzs = dim == 3 ? 3 : 1;
            for (int sz = 0; sz < nzs; ++sz)
            {
                for (int sy = 0; sy < 3; ++sy)
                {
                    for (int sx = 0; sx < 3; ++sx)
                    {
                        double f = src[size_t(om.index(lc, xc + sx - 1, yc + sy - 1, dim == 3 ? zc + sz - 1 : 0))];
                        for (int c = 0; c < nch; ++c)
                        {
                            double coeff = 1.;
                            coeff *= pair[size_t(c & 1)][size_t(sx)];
                            coeff *= pair[size_t((c >> 1) & 1)][size_t(sy)];
                            if (dim == 3)
                            {
                                coeff *= pair[size_t(kPlant == 7 ? 1 - ((c >> 2) & 1) : (c >> 2) & 1)][size_t(sz)];
                            }
                            vals[c] += f * coeff;
                        }
                    }
                }
            }
        }
        for (int c = 0; c < nch; ++c)
        {
            int cx = 2 * xc + (c & 1), cy = 2 * yc + ((c >> 1) & 1), cz = dim == 3 ? 2 * zc + ((c >> 2) & 1) : 0;
            dst[size_t(nm.index(lc + 1, cx, cy, cz))] = vals[c];
        }
    }

    // update_fields (algorithm/update_fields.hpp:27-127) for one scalar field.
    static void update_fields(Sim& s, Mesh& nm)
    {
        const Mesh& om          = s.mesh;
        std::vector<double> dst(size_t(nm.nref), 0.0); // make_new_field fills 0 (:56-66)
        for (int level = om.min_level; level <= om.max_level; ++level)
        {
            DSet set = intersection(om.ids[REFERENCE][size_t(level)], nm.ids[CELLS][size_t(level)]);
            set.for_each(
                [&](int x, int y, int z)
                {
                    dst[size_t(nm.index(level, x, y, z))] = s.u[size_t(om.index(level, x, y, z))];
                });
        }
        for (int level = om.min_level + 1; level <= om.max_level; ++level)
        {
            DSet set_coarsen = on(intersection(om.ids[CELLS][size_t(level)], nm.ids[CELLS][size_t(level - 1)]), level - 1);
            set_coarsen.for_each(
                [&](int x, int y, int z)
                {
                    dst[size_t(nm.index(level - 1, x, y, z))] = project_value(s, level, x, y, z);
                });
            DSet set_refine = on(intersection(nm.ids[CELLS][size_t(level)], om.ids[CELLS][size_t(level - 1)]), level - 1);
            set_refine.for_each(
                [&](int x, int y, int z)
                {
                    predict_children(om, s.u, nm, dst, level - 1, x, y, z);
                });
        }
        s.u = std::move(dst);
    }

    // Adapt::harten (mr/adapt.hpp:276-389); returns true at the fixed point.
    bool harten(Sim& s, int ite, double eps, double reg)
    {
        Mesh& m                 = s.mesh;
        const int min_level = m.min_level, max_level = m.max_level, dim = m.dim;
        s.detail.assign(size_t(m.nref), 0.0);
        s.tag.assign(size_t(m.nref), 0);
        for (int l = 0; l < MAXL; ++l)
        {
            m.ids[CELLS][size_t(l)].for_each(
                [&](int x, int y, int z)
                {
                    s.tag[size_t(m.index(l, x, y, z))] = KEEP;
                });
        }
        // update_ghost_mr(m_fields) and compute_detail(m_detail, m_fields)
        // (mr/adapt.hpp:292, 309-318): each extra field is swapped into u and
        // its detail array into detail while its turn lasts
        s.extra_detail.assign(s.extra.size(), std::vector<double>(size_t(m.nref), 0.0));
        for (size_t k = 0; k <= s.extra.size(); ++k)
        {
            if (k > 0)
            {
                std::swap(s.u, s.extra[k - 1]);
                std::swap(s.detail, s.extra_detail[k - 1]);
            }
            update_ghost_mr(s);
            for (int level = (min_level > 0 ? min_level - 1 : 0); level < max_level - ite; ++level)
            {
                DSet below = on(intersection(m.ids[REFERENCE][size_t(level)],
                                             union_(m.ids[CELLS][size_t(level + 1)], m.ids[CELLS][size_t(std::min(level + 2, MAXL - 1))])),
                                level);
                below.for_each(
                    [&](int x, int y, int z)
                    {
                        compute_detail(s, level, x, y, z);
                    });
            }
            if (k > 0)
            {
                std::swap(s.u, s.extra[k - 1]);
                std::swap(s.detail, s.extra_detail[k - 1]);
            }
        }
        // compute_relative_detail (mr/rel_detail.hpp:74-112, mr/adapt.hpp:319-322):
        // the max of |field| over the leaves, folded from the smallest normal
        // double, 1 where it is below the machine epsilon, and every entry of
        // the field's detail array times its inverse
        if (s.relative_detail)
        {
            for (size_t k = 0; k <= s.extra.size(); ++k)
            {
                const std::vector<double>& f = k == 0 ? s.u : s.extra[k - 1];
                double mx                    = std::numeric_limits<double>::min();
                for (int l = 0; l < MAXL; ++l)
                {
                    m.ids[CELLS][size_t(l)].for_each(
                        [&](int x, int y, int z)
                        {
                            mx = std::max(mx, std::abs(f[size_t(m.index(l, x, y, z))]));
                        });
                }
                if (mx < std::numeric_limits<double>::epsilon())
                {
                    mx = 1.0;
                }
                const double inv = kPlant == 15 ? mx : 1. / mx;
                for (double& v : (k == 0 ? s.detail : s.extra_detail[k - 1]))
                {
                    v *= inv;
                }
            }
        }
        for (int level = min_level; level <= max_level - ite; ++level)
        {
            int exponent             = dim * (max_level - level);
            double eps_l             = eps / (1 << exponent);
            double regularity_to_use = reg + dim;
            DSet sub = on(intersection(m.ids[CELLS][size_t(level)], m.ids[REFERENCE][size_t(level - 1)]), level - 1);
            sub.for_each(
                [&](int x, int y, int z)
                {
                    criteria(s, level - 1, x, y, z, eps_l, regularity_to_use);
                });
        }
        // --refine-boundary: keep_boundary_refined (mr/adapt.hpp:245-272,
        // 341-346): the leaves of max_level within the stencil radius of a face
        // are tagged keep (assigned), direction by direction
        if (s.refine_boundary)
        {
            const int n = 1 << max_level, r = m.radius;
            for (int d = 0; d < dim; ++d)
            {
                for (int sgn : {1, -1})
                {
                    m.ids[CELLS][size_t(max_level)].for_each(
                        [&](int x, int y, int z)
                        {
                            const int c = d == 0 ? x : (d == 1 ? y : z);
                            if (sgn > 0 ? c >= n - r : (kPlant != 16 && c < r))
                            {
                                s.tag[size_t(m.index(max_level, x, y, z))] = KEEP;
                            }
                        });
                }
            }
        }
        for (int level = max_level; level > 0; --level)
        {
            DSet sub = on(intersection(m.ids[CELLS][size_t(level)], m.ids[REFERENCE][size_t(level - 1)]), level - 1);
            sub.for_each(
                [&](int x, int y, int z)
                {
                    maximum(s, level - 1, x, y, z);
                });
        }
        s.tag_mesh    = m;
        s.last_tag    = s.tag;
        s.last_detail = s.detail;

        std::vector<DSet> new_ca = update_cell_array_from_tag(s);
        make_graduation(new_ca, dim, m.grad_width);
        if (ca_equal(m.ids[CELLS], new_ca))
        {
            return true;
        }
        Mesh nm = make_mesh_from_cells(m, new_ca);
        for (size_t k = 0; k <= s.extra.size(); ++k)
        {
            if (k > 0)
            {
                std::swap(s.u, s.extra[k - 1]);
            }
            update_fields(s, nm);
            if (k > 0)
            {
                std::swap(s.u, s.extra[k - 1]);
            }
        }
        s.mesh = std::move(nm);
        return false;
    }

    // Adapt::operator()(mra_config&) (mr/adapt.hpp:147-195)
    int adapt(Sim& s, double eps, double reg)
    {
        const int min_level = s.mesh.min_level, max_level = s.mesh.max_level;
        if (min_level == max_level)
        {
            s.last_harten_iterations = 0;
            return 0;
        }
        int i = 0;
        for (; i < max_level - min_level; ++i)
        {
            if (harten(s, i, eps, reg))
            {
                ++i;
                break;
            }
        }
        s.last_harten_iterations = i;
        return i;
    }

    // flux of upwind_op (stencil_field.hpp:91-97)
    inline double flux(double a, double ul, double ur)
    {
        return .5 * a * (ul + ur) + .5 * std::abs(a) * (ul - ur);
    }

    // unp1 = u - dt * upwind(a, u) over mesh[cells] (stencil_field.hpp:37-53,
    // field/field_base.hpp:230-242), then std::swap(u.array(), unp1.array())
    // (demos/FiniteVolume/advection_2d.cpp:142-146).
    void upwind(Sim& s, const double* a, double dt)
    {
        const Mesh& m = s.mesh;
        std::vector<double> unp1 = new_storage(s, m.nref);
        for (int level = 0; level < MAXL; ++level)
        {
            const double dx = 1.0 / (1 << level); // cell_length (cell.hpp:17-20), unit box
            m.ids[CELLS][size_t(level)].for_each(
                [&](int x, int y, int z)
                {
                    double uc = s.u[size_t(m.index(level, x, y, z))];
                    double S;
                    if (m.dim == 2)
                    {
                        double FL = flux(a[0], s.u[size_t(m.index(level, x - 1, y))], uc);
                        double FR = flux(a[0], uc, s.u[size_t(m.index(level, x + 1, y))]);
                        double FD = flux(a[1], s.u[size_t(m.index(level, x, y - 1))], uc);
                        double FU = flux(a[1], uc, s.u[size_t(m.index(level, x, y + 1))]);
                        S         = -FL + FR + -FD + FU;
                    }
                    else
                    {
                        double FL = flux(a[0], s.u[size_t(m.index(level, x - 1, y, z))], uc);
                        double FR = flux(a[0], uc, s.u[size_t(m.index(level, x + 1, y, z))]);
                        double FD = flux(a[1], s.u[size_t(m.index(level, x, y - 1, z))], uc);
                        double FU = flux(a[1], uc, s.u[size_t(m.index(level, x, y + 1, z))]);
                        double FF = flux(a[2], s.u[size_t(m.index(level, x, y, z - 1))], uc);
                        double FB = flux(a[2], uc, s.u[size_t(m.index(level, x, y, z + 1))]);
                        S         = kPlant == 3 ? -FL + FR + -FD + FU + FB : -FL + FR + -FD + FU + -FF + FB;
                    }
                    unp1[size_t(m.index(level, x, y, z))] = uc - dt * (S / dx);
                });
        }
        s.u.swap(unp1);
    }

    // flux of upwind_scalar_burgers_op (stencil_field.hpp:193-216): 0 by
    // default; 0.5 * min(|ul|, |ur|)^2 where (a ul < a ur) and ul ur > 0;
    // 0.5 * max(|ul|, |ur|)^2 where !(a ul < a ur)
    inline double burgers_upwind_flux(double a, double ul, double ur)
    {
        const bool mask1 = (a * ul) < (a * ur);
        const bool mask2 = (ul * ur) > 0.0;
        const double mn  = std::min(std::abs(ul), std::abs(ur));
        const double mx  = std::max(std::abs(ul), std::abs(ur));
        double out       = 0.0;
        if (mask1 && mask2)
        {
            out = .5 * mn * mn;
        }
        if (!mask1)
        {
            out = .5 * mx * mx;
        }
        return out;
    }

    // unp1 = u - dt * upwind_scalar_burgers(k, u) over mesh[cells], 2D
    // (stencil_field.hpp:181-255 with finite_volume, :37-53; demos/FiniteVolume/
    // scalar_burgers_2d.cpp:155-160), then the swap of the arrays
    void upwind_scalar_burgers(Sim& s, const double* k, double dt)
    {
        const Mesh& m = s.mesh;
        std::vector<double> unp1 = new_storage(s, m.nref);
        for (int level = 0; level < MAXL; ++level)
        {
            const double dx = 1.0 / (1 << level);
            m.ids[CELLS][size_t(level)].for_each(
                [&](int x, int y, int z)
                {
                    const double uc = s.u[size_t(m.index(level, x, y, z))];
                    const double FL = burgers_upwind_flux(k[0], s.u[size_t(m.index(level, x - 1, y))], uc);
                    const double FR = burgers_upwind_flux(k[0], uc, s.u[size_t(m.index(level, x + 1, y))]);
                    const double FD = burgers_upwind_flux(k[1], s.u[size_t(m.index(level, x, y - 1))], uc);
                    const double FU = burgers_upwind_flux(k[1], uc, s.u[size_t(m.index(level, x, y + 1))]);
                    const double S  = -FL + FR + -FD + FU;
                    unp1[size_t(m.index(level, x, y, z))] = uc - dt * (S / dx);
                });
        }
        s.u.swap(unp1);
    }

    // README.md:135-154: for_each_interval(mesh, ...) over the cells,
    //   unp1(level, i, j) = u(level, i, j) - dt / dx * (u(level, i, j) - u(level, i - 1, j)
    //                                                 + u(level, i, j) - u(level, i, j - 1));
    // then std::swap(unp1.array(), u.array())
    void upwind_hand_loop(Sim& s, double dt)
    {
        const Mesh& m = s.mesh;
        std::vector<double> unp1 = new_storage(s, m.nref);
        for (int level = 0; level < MAXL; ++level)
        {
            const double dx = 1.0 / (1 << level);
            m.ids[CELLS][size_t(level)].for_each(
                [&](int x, int y, int z)
                {
                    const double uc = s.u[size_t(m.index(level, x, y, z))];
                    const double ul = s.u[size_t(m.index(level, x - 1, y, z))];
                    const double ud = s.u[size_t(m.index(level, x, y - 1, z))];
                    unp1[size_t(m.index(level, x, y, z))] = uc - dt / dx * (uc - ul + uc - ud);
                });
        }
        s.u.swap(unp1);
    }

    // maximal x-runs of a set in traversal order (z, y ascending, then x)
    template <class F>
    void for_each_run(const DSet& s, F&& f)
    {
        if (s.box_empty())
        {
            return;
        }
        for (int z = s.lo[2]; z < s.hi[2]; ++z)
        {
            for (int y = s.lo[1]; y < s.hi[1]; ++y)
            {
                int x = s.lo[0];
                while (x < s.hi[0])
                {
                    if (!s.get(x, y, z))
                    {
                        ++x;
                        continue;
                    }
                    int e = x;
                    while (e < s.hi[0] && s.get(e, y, z))
                    {
                        ++e;
                    }
                    f(x, e, y, z);
                    x = e;
                }
            }
        }
    }

    // h_factor (flux_based_scheme__lin_hom.hpp:62-67)
    inline double h_factor(int dim, double h_face, double h_cell)
    {
        return std::pow(h_face, dim - 1) / std::pow(h_cell, dim);
    }

    // make_convection_upwind(v)(u) (schemes/fv/operators/convection_lin.hpp:15-89)
    // applied explicitly (FV_scheme.hpp:202-210): the output is zero-filled
    // (explicit_FV_scheme.hpp:51-56); then for each direction d the interior
    // interfaces -- same level for every level, then the level jumps in the
    // direction and in the opposite direction (flux_based_scheme__lin_hom.hpp:74-159,
    // interface.hpp:35-315) -- and the boundary interfaces of every level
    // (flux_based_scheme__lin_hom.hpp:166-207, interface.hpp:440-509),
    // accumulated in that order as explicit_flux_based_scheme__lin_hom.hpp:39-119,
    // 233-319 does (per stencil cell c: the left cells of an interval, then its
    // right cells; at a jump the coarse side adds a pair of fine products).
    std::vector<double> convection_upwind(const Sim& s, const double* v)
    {
        const Mesh& m = s.mesh;
        const int dim = m.dim;
        std::vector<double> out(size_t(m.nref), 0.0);
        auto idx = [&](int l, int x, int y, int z) { return size_t(m.index(l, x, y, z)); };
        const auto& cells = m.ids[CELLS];
        const int cmin = m.ca_min_level(CELLS), cmax = m.ca_max_level(CELLS);
        const int minl = std::max(m.min_level, cmin - 1), maxl = std::min(m.max_level, cmax + 1);
        for (int d = 0; d < dim; ++d)
        {
            std::array<int, 3> e{0, 0, 0};
            e[size_t(d)] = 1;
            const std::array<int, 3> me{-e[0], -e[1], -e[2]};
            // coefficients of the stencil {left cell, right cell} (convection_lin.hpp:36-77)
            const double fc[2] = {v[d] >= 0 ? v[d] : 0.0, v[d] >= 0 ? 0.0 : v[d]};
            // same level
            for (int level = minl; level <= maxl; ++level)
            {
                const double h  = 1.0 / (1 << level);
                const double fa = h_factor(dim, h, h);
                double L[2], R[2];
                for (int c = 0; c < 2; ++c)
                {
                    L[c] = fa * fc[c];
                    R[c] = -L[c];
                }
                const DSet I = intersection(cells[size_t(level)], translate(cells[size_t(level)], me));
                for_each_run(I,
                             [&](int x0, int x1, int y, int z)
                             {
                                 for (int c = 0; c < 2; ++c)
                                 {
                                     for (int x = x0; x < x1; ++x)
                                     {
                                         out[idx(level, x, y, z)] += L[c] * s.u[idx(level, x + c * e[0], y + c * e[1], z + c * e[2])];
                                     }
                                     for (int x = x0; x < x1; ++x)
                                     {
                                         out[idx(level, x + e[0], y + e[1], z + e[2])] +=
                                             R[c] * s.u[idx(level, x + c * e[0], y + c * e[1], z + c * e[2])];
                                     }
                                 }
                             });
            }
            // level jumps l -- l+1
            for (int level = minl; level < maxl; ++level)
            {
                if (level >= m.max_level)
                {
                    break;
                }
                const int lf       = level + 1;
                const double hl    = 1.0 / (1 << level);
                const double hf    = 1.0 / (1 << lf);
                const int zsh      = dim == 3 ? 1 : 0;
                auto coarse        = [&](int x, int y, int z) { return idx(level, x >> 1, y >> 1, z >> zsh); };
                auto fine_u        = [&](int x, int y, int z) { return s.u[idx(lf, x, y, z)]; };
                // coarse on the left, fine on the right (interface.hpp:123-203)
                {
                    const double lfac = h_factor(dim, hf, hl), rfac = h_factor(dim, hf, hf);
                    double L[2], R[2];
                    for (int c = 0; c < 2; ++c)
                    {
                        L[c] = lfac * fc[c];
                        R[c] = -rfac * fc[c];
                    }
                    const DSet J = on(intersection(cells[size_t(level)], translate(cells[size_t(lf)], me)), lf);
                    for_each_run(J,
                                 [&](int x0, int x1, int y, int z)
                                 {
                                     const int n = x1 - x0;
                                     for (int c = 0; c < 2; ++c)
                                     {
                                         const int cx = c * e[0], cy = c * e[1], cz = c * e[2];
                                         if (n == 1)
                                         {
                                             out[coarse(x0, y, z)] += L[c] * fine_u(x0 + cx, y + cy, z + cz);
                                             out[idx(lf, x0 + e[0], y + e[1], z + e[2])] += R[c] * fine_u(x0 + cx, y + cy, z + cz);
                                         }
                                         else
                                         {
                                             for (int i = 0; i < n / 2; ++i)
                                             {
                                                 const int xa = x0 + 2 * i;
                                                 out[coarse(xa, y, z)] += L[c] * fine_u(xa + cx, y + cy, z + cz) + L[c] * fine_u(xa + 1 + cx, y + cy, z + cz);
                                             }
                                             for (int x = x0; x < x1; ++x)
                                             {
                                                 out[idx(lf, x + e[0], y + e[1], z + e[2])] += R[c] * fine_u(x + cx, y + cy, z + cz);
                                             }
                                         }
                                     }
                                 });
                }
                // fine on the left, coarse on the right (interface.hpp:205-315)
                {
                    const double lfac = h_factor(dim, hf, hf), rfac = h_factor(dim, hf, hl);
                    double L[2], R[2];
                    for (int c = 0; c < 2; ++c)
                    {
                        L[c] = lfac * fc[c];
                        R[c] = -rfac * fc[c];
                    }
                    const DSet J = on(intersection(cells[size_t(level)], translate(cells[size_t(lf)], e)), lf);
                    for_each_run(J,
                                 [&](int x0, int x1, int y, int z)
                                 {
                                     const int n = x1 - x0;
                                     for (int c = 0; c < 2; ++c)
                                     {
                                         // stencil {p - e, p}
                                         const int cx = (c - 1) * e[0], cy = (c - 1) * e[1], cz = (c - 1) * e[2];
                                         if (n == 1)
                                         {
                                             out[idx(lf, x0 - e[0], y - e[1], z - e[2])] += L[c] * fine_u(x0 + cx, y + cy, z + cz);
                                             out[coarse(x0, y, z)] += R[c] * fine_u(x0 + cx, y + cy, z + cz);
                                         }
                                         else
                                         {
                                             for (int x = x0; x < x1; ++x)
                                             {
                                                 out[idx(lf, x - e[0], y - e[1], z - e[2])] += L[c] * fine_u(x + cx, y + cy, z + cz);
                                             }
                                             for (int i = 0; i < n / 2; ++i)
                                             {
                                                 const int xa = x0 + 2 * i;
                                                 out[coarse(xa, y, z)] += R[c] * fine_u(xa + cx, y + cy, z + cz) + R[c] * fine_u(xa + 1 + cx, y + cy, z + cz);
                                             }
                                         }
                                     }
                                 });
                }
            }
            // boundary interfaces (domain_boundary, boundary.hpp:6-33)
            for (int level = cmin; level <= cmax; ++level)
            {
                const double h  = 1.0 / (1 << level);
                const double fa = h_factor(dim, h, h);
                double F[2]     = {fc[0] * fa, fc[1] * fa};
                const DSet& dom = m.domain[size_t(level)];
                const DSet B    = difference(cells[size_t(level)], translate(dom, me));
                for_each_run(B,
                             [&](int x0, int x1, int y, int z)
                             {
                                 for (int c = 0; c < 2; ++c)
                                 {
                                     for (int x = x0; x < x1; ++x)
                                     {
                                         out[idx(level, x, y, z)] += F[c] * s.u[idx(level, x + c * e[0], y + c * e[1], z + c * e[2])];
                                     }
                                 }
                             });
                F[0] *= -1;
                F[1] *= -1;
                const DSet Bo = difference(cells[size_t(level)], translate(dom, e));
                for_each_run(Bo,
                             [&](int x0, int x1, int y, int z)
                             {
                                 for (int c = 0; c < 2; ++c)
                                 {
                                     const int cx = (c - 1) * e[0], cy = (c - 1) * e[1], cz = (c - 1) * e[2];
                                     for (int x = x0; x < x1; ++x)
                                     {
                                         out[idx(level, x, y, z)] += F[c] * s.u[idx(level, x + cx, y + cy, z + cz)];
                                     }
                                 }
                             });
            }
        }
        return out;
    }

    // the scalar flux of the non-linear make_convection_upwind<Field>()
    // (schemes/fv/operators/convection_nonlin.hpp:24-79): the upwind side of
    // f(u) = u * u, chosen by the sign of the mean of the two values
    inline double burgers_flux(double ul, double ur)
    {
        const double v = 0.5 * (ul + ur);
        return v >= 0 ? ul * ul : ur * ur;
    }

    // make_convection_upwind<Field>()(u) for a scalar field, applied explicitly
    // (explicit_flux_based_scheme__nonlin.hpp:35-86): for each direction, the
    // same-level interfaces of every level, then the level jumps in the
    // direction and in the opposite direction, with the fine ghost inside the
    // coarse cell as stencil value (flux_based_scheme__nonlin.hpp:408-528,
    // copy_stencil_values, :179-186), then the boundary interfaces
    // (:530-590).  Each interface adds flux * left_factor to its left cell and
    // (-flux) * right_factor to its right cell, in that order
    // (process_interior_interfaces, :334-364; process_boundary_interfaces,
    // :366-398); the output starts at zero.
    std::vector<double> convection_upwind_nonlinear(const Sim& s)
    {
        const Mesh& m = s.mesh;
        const int dim = m.dim;
        std::vector<double> out(size_t(m.nref), 0.0);
        auto idx = [&](int l, int x, int y, int z) { return size_t(m.index(l, x, y, z)); };
        const auto& cells = m.ids[CELLS];
        const int cmin = m.ca_min_level(CELLS), cmax = m.ca_max_level(CELLS);
        const int minl = std::max(m.min_level, cmin - 1), maxl = std::min(m.max_level, cmax + 1);
        auto interface = [&](size_t left, size_t right, double ul, double ur, double lf, double rf)
        {
            const double f = burgers_flux(ul, ur);
            out[left] += f * lf;
            out[right] += (-f) * rf;
        };
        for (int d = 0; d < dim; ++d)
        {
            std::array<int, 3> e{0, 0, 0};
            e[size_t(d)] = 1;
            const std::array<int, 3> me{-e[0], -e[1], -e[2]};
            for (int level = minl; level <= maxl; ++level)
            {
                const double h  = 1.0 / (1 << level);
                const double fa = h_factor(dim, h, h);
                const DSet I    = intersection(cells[size_t(level)], translate(cells[size_t(level)], me));
                for_each_run(I,
                             [&](int x0, int x1, int y, int z)
                             {
                                 for (int x = x0; x < x1; ++x)
                                 {
                                     const size_t p = idx(level, x, y, z), q = idx(level, x + e[0], y + e[1], z + e[2]);
                                     interface(p, q, s.u[p], s.u[q], fa, fa);
                                 }
                             });
            }
            for (int level = minl; level < maxl; ++level)
            {
                if (level >= m.max_level)
                {
                    break;
                }
                const int lf    = level + 1;
                const double hl = 1.0 / (1 << level);
                const double hf = 1.0 / (1 << lf);
                const int zsh   = dim == 3 ? 1 : 0;
                auto coarse     = [&](int x, int y, int z) { return idx(level, x >> 1, y >> 1, z >> zsh); };
                // coarse cell on the left: stencil {p (fine ghost in the coarse cell), p + e}
                {
                    const double lfac = h_factor(dim, hf, hl), rfac = h_factor(dim, hf, hf);
                    const DSet J      = on(intersection(cells[size_t(level)], translate(cells[size_t(lf)], me)), lf);
                    for_each_run(J,
                                 [&](int x0, int x1, int y, int z)
                                 {
                                     for (int x = x0; x < x1; ++x)
                                     {
                                         const size_t g = idx(lf, x, y, z), f = idx(lf, x + e[0], y + e[1], z + e[2]);
                                         interface(coarse(x, y, z), f, s.u[g], s.u[f], lfac, rfac);
                                     }
                                 });
                }
                // fine cell on the left: stencil {p - e, p (fine ghost in the coarse cell)}
                {
                    const double lfac = h_factor(dim, hf, hf), rfac = h_factor(dim, hf, hl);
                    const DSet J      = on(intersection(cells[size_t(level)], translate(cells[size_t(lf)], e)), lf);
                    for_each_run(J,
                                 [&](int x0, int x1, int y, int z)
                                 {
                                     for (int x = x0; x < x1; ++x)
                                     {
                                         const size_t f = idx(lf, x - e[0], y - e[1], z - e[2]), g = idx(lf, x, y, z);
                                         interface(f, coarse(x, y, z), s.u[f], s.u[g], lfac, rfac);
                                     }
                                 });
                }
            }
            // boundary interfaces: the outer ghost holds the boundary value
            for (int level = cmin; level <= cmax; ++level)
            {
                const double h  = 1.0 / (1 << level);
                const double fa = h_factor(dim, h, h);
                const DSet& dom = m.domain[size_t(level)];
                for_each_run(difference(cells[size_t(level)], translate(dom, me)),
                             [&](int x0, int x1, int y, int z)
                             {
                                 for (int x = x0; x < x1; ++x)
                                 {
                                     const size_t p = idx(level, x, y, z);
                                     out[p] += burgers_flux(s.u[p], s.u[idx(level, x + e[0], y + e[1], z + e[2])]) * fa;
                                 }
                             });
                for_each_run(difference(cells[size_t(level)], translate(dom, e)),
                             [&](int x0, int x1, int y, int z)
                             {
                                 for (int x = x0; x < x1; ++x)
                                 {
                                     const size_t p = idx(level, x, y, z);
                                     out[p] += (-burgers_flux(s.u[idx(level, x - e[0], y - e[1], z - e[2])], s.u[p])) * fa;
                                 }
                             });
            }
        }
        return out;
    }

    // The WENO5 flux of Jiang & Shu (1996) on five flux values f[0..5) centred
    // on f[2] (schemes/fv/operators/weno_impl.hpp:26-66): the three
    // third-order candidates (eq. 2.8, Table I), their smoothness indicators
    // (3.2-3.4), the non-linear weights from the linear ones 0.1, 0.6, 0.3 and
    // eps = 1e-6 (2.15-2.16, Table II), and the weighted sum (2.10), each in
    // the reference's operation order.  The reference squares with pow(x, 2),
    // which compilers emit as x * x (the correctly rounded square).
    inline double weno5_sq(double x)
    {
        return x * x;
    }

    inline double weno5_flux(const double* f)
    {
        const double eps = 1e-6;
        const double q0  = 1. / 3 * f[0] - 7. / 6 * f[1] + 11. / 6 * f[2];
        const double q1  = -1. / 6 * f[1] + 5. / 6 * f[2] + 1. / 3 * f[3];
        const double q2  = 1. / 3 * f[2] + 5. / 6 * f[3] - 1. / 6 * f[4];
        const double is0 = 13. / 12 * weno5_sq(f[0] - 2 * f[1] + f[2]) + 1. / 4 * weno5_sq(f[0] - 4 * f[1] + 3 * f[2]);
        const double is1 = 13. / 12 * weno5_sq(f[1] - 2 * f[2] + f[3]) + 1. / 4 * weno5_sq(f[1] - f[3]);
        const double is2 = 13. / 12 * weno5_sq(f[2] - 2 * f[3] + f[4]) + 1. / 4 * weno5_sq(3 * f[2] - 4 * f[3] + f[4]);
        const double a0  = 0.1 / weno5_sq(eps + is0);
        const double a1  = (kPlant == 12 ? 0.3 : 0.6) / weno5_sq(eps + is1);
        const double a2  = (kPlant == 12 ? 0.6 : 0.3) / weno5_sq(eps + is2);
        const double sum = a0 + a1 + a2;
        const double w0 = a0 / sum, w1 = a1 / sum, w2 = a2 / sum;
        return w0 * q0 + w1 * q1 + w2 * q2;
    }

    // the flux of make_convection_weno5(velocity) in direction d
    // (schemes/fv/operators/convection_lin.hpp:95-178) from the six stencil
    // values u[0..6) (cells -2..3 along d from the interface's left cell):
    // v_d * u[0..5) for v_d >= 0, else v_d * (u[5], u[4], u[3], u[2], u[1])
    inline double weno5_interface_flux(double vd, const double* u)
    {
        double f[5];
        for (int i = 0; i < 5; ++i)
        {
            f[i] = (vd >= 0 ? u[i] : u[5 - i]) * vd;
        }
        return weno5_flux(f);
    }

    // make_convection_weno5(velocity)(u) on a mesh of one level, applied
    // explicitly like every non-linear flux-based operator
    // (explicit_flux_based_scheme__nonlin.hpp:35-86): for each direction, the
    // interfaces between neighbouring cells (flux * h-factor into the left
    // cell, (-flux) * h-factor into the right one), then the boundary
    // interfaces, the inner cell taking flux * h-factor in the direction and
    // (-flux) * h-factor in the opposite direction with the stencil moved one
    // cell back (interface.hpp:494-508; flux_based_scheme__nonlin.hpp:366-398,
    // 530-590).  Every stencil cell of a boundary cell's interface reaches up
    // to three ghost layers, which the ghost update fills beforehand.
    std::vector<double> convection_weno5(const Sim& s, const double* v)
    {
        const Mesh& m = s.mesh;
        const int dim = m.dim;
        const int cmin = m.ca_min_level(CELLS), cmax = m.ca_max_level(CELLS);
        if (cmin != cmax)
        {
            throw std::invalid_argument("oracle: WENO5 is built for meshes whose cells are all at one level");
        }
        if (ghost_width(m) < 3)
        {
            throw std::invalid_argument("oracle: WENO5 needs a stencil radius of 3");
        }
        std::vector<double> out(size_t(m.nref), 0.0);
        const int level   = cmin;
        const DSet& cells = m.ids[CELLS][size_t(level)];
        const DSet& dom   = m.domain[size_t(level)];
        const double h    = 1.0 / (1 << level);
        const double fa   = h_factor(dim, h, h);
        for (int d = 0; d < dim; ++d)
        {
            std::array<int, 3> e{0, 0, 0};
            e[size_t(d)] = 1;
            const std::array<int, 3> me{-e[0], -e[1], -e[2]};
            // flux of the interface whose stencil starts at cell (x, y, z) + j0 e
            auto flux = [&](int x, int y, int z, int j0)
            {
                double u[6];
                for (int j = 0; j < 6; ++j)
                {
                    const int k = j0 + j;
                    u[j]        = s.u[size_t(m.index(level, x + k * e[0], y + k * e[1], z + k * e[2]))];
                }
                return weno5_interface_flux(v[d], u);
            };
            for_each_run(intersection(cells, translate(cells, me)),
                         [&](int x0, int x1, int y, int z)
                         {
                             for (int x = x0; x < x1; ++x)
                             {
                                 const size_t p = size_t(m.index(level, x, y, z));
                                 const size_t q = size_t(m.index(level, x + e[0], y + e[1], z + e[2]));
                                 const double F = flux(x, y, z, -2);
                                 out[p] += F * fa;
                                 out[q] += (-F) * fa;
                             }
                         });
            for_each_run(difference(cells, translate(dom, me)),
                         [&](int x0, int x1, int y, int z)
                         {
                             for (int x = x0; x < x1; ++x)
                             {
                                 out[size_t(m.index(level, x, y, z))] += flux(x, y, z, -2) * fa;
                             }
                         });
            for_each_run(difference(cells, translate(dom, e)),
                         [&](int x0, int x1, int y, int z)
                         {
                             for (int x = x0; x < x1; ++x)
                             {
                                 out[size_t(m.index(level, x, y, z))] += (-flux(x, y, z, kPlant == 14 ? -2 : -3)) * fa;
                             }
                         });
        }
        return out;
    }

    // make_convection_upwind<Field>(velocity_field)(u) for a scalar field
    // (schemes/fv/operators/convection_lin.hpp:184-264, the LinearHeterogeneous
    // flux): at each interface the velocity is 0.5 * (v[left] + v[right]) of the
    // stencil cells' velocity values (ghosts included, updated beforehand as
    // FV_scheme.hpp:176-196 does for the parameter field), and the coefficient
    // of the upwind side is factor * v_d.  Applied explicitly
    // (explicit_flux_based_scheme__lin_het.hpp:35-110) cell by cell over the
    // same interfaces as the other operators (flux_based_scheme__lin_het.hpp:
    // 75-250): the left cell gets (lf * v_d) * u, the right cell
    // (-(rf * v_d)) * u of the upwind stencil value; the zero coefficient of the
    // other side adds nothing.  vel[d] is the storage of component d.
    std::vector<double> convection_upwind_heterogeneous(const Sim& s, const double* const* vel)
    {
        const Mesh& m = s.mesh;
        const int dim = m.dim;
        std::vector<double> out(size_t(m.nref), 0.0);
        auto idx = [&](int l, int x, int y, int z) { return size_t(m.index(l, x, y, z)); };
        const auto& cells = m.ids[CELLS];
        const int cmin = m.ca_min_level(CELLS), cmax = m.ca_max_level(CELLS);
        const int minl = std::max(m.min_level, cmin - 1), maxl = std::min(m.max_level, cmax + 1);
        for (int d = 0; d < dim; ++d)
        {
            const double* v = vel[d];
            // the coefficient of the upwind side and that side's value
            auto upwind = [&](size_t sl, size_t sr, double factor, double& value)
            {
                const double vd = 0.5 * (v[sl] + v[sr]);
                value           = vd >= 0 ? s.u[sl] : s.u[sr];
                return factor * vd;
            };
            auto interface = [&](size_t left, size_t right, size_t sl, size_t sr, double lf, double rf)
            {
                double value = 0;
                const double cl = upwind(sl, sr, lf, value);
                const double cr = -upwind(sl, sr, rf, value);
                out[left] += cl * value;
                out[right] += cr * value;
            };
            std::array<int, 3> e{0, 0, 0};
            e[size_t(d)] = 1;
            const std::array<int, 3> me{-e[0], -e[1], -e[2]};
            for (int level = minl; level <= maxl; ++level)
            {
                const double h  = 1.0 / (1 << level);
                const double fa = h_factor(dim, h, h);
                for_each_run(intersection(cells[size_t(level)], translate(cells[size_t(level)], me)),
                             [&](int x0, int x1, int y, int z)
                             {
                                 for (int x = x0; x < x1; ++x)
                                 {
                                     const size_t p = idx(level, x, y, z), q = idx(level, x + e[0], y + e[1], z + e[2]);
                                     interface(p, q, p, q, fa, fa);
                                 }
                             });
            }
            for (int level = minl; level < maxl; ++level)
            {
                if (level >= m.max_level)
                {
                    break;
                }
                const int lf    = level + 1;
                const double hl = 1.0 / (1 << level);
                const double hf = 1.0 / (1 << lf);
                const int zsh   = dim == 3 ? 1 : 0;
                auto coarse     = [&](int x, int y, int z) { return idx(level, x >> 1, y >> 1, z >> zsh); };
                {
                    const double lfac = h_factor(dim, hf, hl), rfac = h_factor(dim, hf, hf);
                    for_each_run(on(intersection(cells[size_t(level)], translate(cells[size_t(lf)], me)), lf),
                                 [&](int x0, int x1, int y, int z)
                                 {
                                     for (int x = x0; x < x1; ++x)
                                     {
                                         const size_t g = idx(lf, x, y, z), f = idx(lf, x + e[0], y + e[1], z + e[2]);
                                         interface(coarse(x, y, z), f, g, f, lfac, rfac);
                                     }
                                 });
                }
                {
                    const double lfac = h_factor(dim, hf, hf), rfac = h_factor(dim, hf, hl);
                    for_each_run(on(intersection(cells[size_t(level)], translate(cells[size_t(lf)], e)), lf),
                                 [&](int x0, int x1, int y, int z)
                                 {
                                     for (int x = x0; x < x1; ++x)
                                     {
                                         const size_t f = idx(lf, x - e[0], y - e[1], z - e[2]), g = idx(lf, x, y, z);
                                         interface(f, coarse(x, y, z), f, g, lfac, rfac);
                                     }
                                 });
                }
            }
            for (int level = cmin; level <= cmax; ++level)
            {
                const double h  = 1.0 / (1 << level);
                const double fa = h_factor(dim, h, h);
                const DSet& dom = m.domain[size_t(level)];
                for_each_run(difference(cells[size_t(level)], translate(dom, me)),
                             [&](int x0, int x1, int y, int z)
                             {
                                 for (int x = x0; x < x1; ++x)
                                 {
                                     const size_t p = idx(level, x, y, z);
                                     double value   = 0;
                                     const double c = upwind(p, idx(level, x + e[0], y + e[1], z + e[2]), fa, value);
                                     out[p] += c * value;
                                 }
                             });
                for_each_run(difference(cells[size_t(level)], translate(dom, e)),
                             [&](int x0, int x1, int y, int z)
                             {
                                 for (int x = x0; x < x1; ++x)
                                 {
                                     const size_t p = idx(level, x, y, z);
                                     double value   = 0;
                                     const double c = upwind(idx(level, x - e[0], y - e[1], z - e[2]), p, -fa, value);
                                     out[p] += c * value;
                                 }
                             });
            }
        }
        return out;
    }

    Sim* make_box_sim(int dim, int min_level, int max_level, int pred_s, int radius, int grad_width)
    {
        auto* s           = new Sim();
        Mesh cfg;
        cfg.dim           = dim;
        cfg.min_level     = min_level;
        cfg.max_level     = max_level;
        cfg.pred_s        = pred_s;
        cfg.radius        = radius;
        cfg.grad_width    = grad_width;
        std::vector<DSet> cells = empty_ca(dim);
        cells[size_t(max_level)] = domain_at(dim, max_level); // start_level = max_level (mr/mesh.hpp:510-518)
        s->mesh = make_mesh_from_cells(cfg, cells);
        s->u    = new_storage(*s, s->mesh.nref);
        return s;
    }
} // namespace oracle

// ----------------------------------------------------------------------
// C API used by the Python tests (ctypes).
// ----------------------------------------------------------------------
using namespace oracle;

static thread_local std::string g_err;

#define ORC_TRY(body)                                                                                                                 \
    try                                                                                                                               \
    {                                                                                                                                 \
        body;                                                                                                                         \
    }                                                                                                                                 \
    catch (const std::exception& e)                                                                                                  \
    {                                                                                                                                 \
        g_err = e.what();                                                                                                             \
        return -1;                                                                                                                    \
    }

extern "C"
{
    const char* orc_last_error() { return g_err.c_str(); }

    // scheme stencil radius 1 on every mesh; 2 or 3 (ghost layers filled by
    // polynomial extrapolation up to stencil 6) on meshes of one level
    static bool radius_supported(int radius, int min_level, int max_level)
    {
        return radius == 1 || (radius >= 2 && radius <= 3 && min_level == max_level);
    }

    void* orc_create_box(int dim, int min_level, int max_level, int pred_s, int radius, int grad_width)
    {
        try
        {
            if (dim != 2 && dim != 3)
            {
                throw std::invalid_argument("dim must be 2 or 3");
            }
            if (!radius_supported(radius, min_level, max_level) || grad_width < 1 || pred_s < 0 || pred_s > 1 || max_level < min_level ||
                min_level < 1)
            {
                throw std::invalid_argument("unsupported mesh configuration");
            }
            return make_box_sim(dim, min_level, max_level, pred_s, radius, grad_width);
        }
        catch (const std::exception& e)
        {
            g_err = e.what();
            return nullptr;
        }
    }

    // MRMesh(ca, cfg) from n leaves given as rows (level, x, y, z)
    void* orc_create_from_cells(int dim, int min_level, int max_level, int pred_s, int radius, int grad_width, const int32_t* cells,
                                int64_t n)
    {
        try
        {
            if ((dim != 2 && dim != 3) || !radius_supported(radius, min_level, max_level) || pred_s < 0 || pred_s > 1 || max_level < min_level ||
                min_level < 1)
            {
                throw std::invalid_argument("unsupported mesh configuration");
            }
            Mesh cfg;
            cfg.dim        = dim;
            cfg.min_level  = min_level;
            cfg.max_level  = max_level;
            cfg.pred_s     = pred_s;
            cfg.radius     = radius;
            cfg.grad_width = grad_width;
            std::vector<DSet> ca = empty_ca(dim);
            for (int l = 0; l < MAXL; ++l)
            {
                int n_l = 1 << std::min(l, 30);
                if (l <= max_level)
                {
                    ca[size_t(l)] = box_set(dim, l, {0, 0, 0}, {n_l, n_l, dim == 3 ? n_l : 1});
                    std::fill(ca[size_t(l)].b.begin(), ca[size_t(l)].b.end(), 0);
                }
            }
            for (int64_t k = 0; k < n; ++k)
            {
                const int32_t* r = cells + 4 * k;
                ca[size_t(r[0])].set(r[1], r[2], dim == 3 ? r[3] : 0);
            }
            auto* s = new Sim();
            s->mesh = make_mesh_from_cells(cfg, ca);
            s->u    = new_storage(*s, s->mesh.nref);
            return s;
        }
        catch (const std::exception& e)
        {
            g_err = e.what();
            return nullptr;
        }
    }

    void orc_free(void* p) { delete static_cast<Sim*>(p); }

    void orc_set_poison(void* p, int on) { static_cast<Sim*>(p)->poison = on != 0; }

    // number of cells of a mesh id; REFERENCE gives the storage size.
    int64_t orc_nb_cells(void* p, int id)
    {
        auto* s = static_cast<Sim*>(p);
        if (id == REFERENCE)
        {
            return s->mesh.nref;
        }
        int64_t n = 0;
        for (int l = 0; l < MAXL; ++l)
        {
            n += s->mesh.ids[id][size_t(l)].count();
        }
        return n;
    }

    // cells of a mesh id in for_each_cell order: rows of (level, x, y, z, index)
    int orc_get_cells(void* p, int id, int64_t* out)
    {
        auto* s = static_cast<Sim*>(p);
        ORC_TRY({
            int64_t k = 0;
            for (int l = 0; l < MAXL; ++l)
            {
                s->mesh.ids[id][size_t(l)].for_each(
                    [&](int x, int y, int z)
                    {
                        out[5 * k + 0] = l;
                        out[5 * k + 1] = x;
                        out[5 * k + 2] = y;
                        out[5 * k + 3] = z;
                        out[5 * k + 4] = s->mesh.exists(l, x, y, z) ? s->mesh.index(l, x, y, z) : -1;
                        ++k;
                    });
            }
        });
        return 0;
    }

    int orc_set_cells_values(void* p, const double* v)
    {
        auto* s = static_cast<Sim*>(p);
        ORC_TRY({
            int64_t k = 0;
            for (int l = 0; l < MAXL; ++l)
            {
                s->mesh.ids[CELLS][size_t(l)].for_each(
                    [&](int x, int y, int z)
                    {
                        s->u[size_t(s->mesh.index(l, x, y, z))] = v[k++];
                    });
            }
        });
        return 0;
    }

    int orc_get_cells_values(void* p, double* v)
    {
        auto* s = static_cast<Sim*>(p);
        ORC_TRY({
            int64_t k = 0;
            for (int l = 0; l < MAXL; ++l)
            {
                s->mesh.ids[CELLS][size_t(l)].for_each(
                    [&](int x, int y, int z)
                    {
                        v[k++] = s->u[size_t(s->mesh.index(l, x, y, z))];
                    });
            }
        });
        return 0;
    }

    // another field adapted with the main one (make_MRAdapt(u, v...)), zero
    // storage; returns its number k >= 0
    int orc_add_field(void* p)
    {
        auto* s = static_cast<Sim*>(p);
        s->extra.emplace_back(s->u.size(), 0.0);
        return int(s->extra.size()) - 1;
    }

    // values of extra field k on the leaves / its whole storage
    int orc_set_extra_values(void* p, int k, const double* v)
    {
        auto* s = static_cast<Sim*>(p);
        std::swap(s->u, s->extra[size_t(k)]);
        const int rc = orc_set_cells_values(p, v);
        std::swap(s->u, s->extra[size_t(k)]);
        return rc;
    }

    int orc_get_extra_values(void* p, int k, double* v)
    {
        auto* s = static_cast<Sim*>(p);
        std::swap(s->u, s->extra[size_t(k)]);
        const int rc = orc_get_cells_values(p, v);
        std::swap(s->u, s->extra[size_t(k)]);
        return rc;
    }

    int orc_get_extra_storage(void* p, int k, double* v)
    {
        auto* s = static_cast<Sim*>(p);
        std::memcpy(v, s->extra[size_t(k)].data(), s->extra[size_t(k)].size() * sizeof(double));
        return 0;
    }

    int orc_get_storage(void* p, double* v)
    {
        auto* s = static_cast<Sim*>(p);
        std::memcpy(v, s->u.data(), s->u.size() * sizeof(double));
        return 0;
    }

    int orc_set_storage(void* p, const double* v)
    {
        auto* s = static_cast<Sim*>(p);
        std::memcpy(s->u.data(), v, s->u.size() * sizeof(double));
        return 0;
    }

    void orc_set_dirichlet(void* p, double value)
    {
        auto* s     = static_cast<Sim*>(p);
        s->has_bc   = true;
        s->bc_value = value;
    }

    int orc_update_ghost_mr(void* p)
    {
        ORC_TRY(update_ghost_mr(*static_cast<Sim*>(p)));
        return 0;
    }

    // one MRadaptation(mra_config) call; returns the number of harten calls
    int orc_adapt(void* p, double eps, double reg)
    {
        int n = 0;
        ORC_TRY(n = adapt(*static_cast<Sim*>(p), eps, reg));
        return n;
    }

    // one harten(ite) call: 1 at the fixed point, 0 if the mesh changed
    int orc_harten(void* p, int ite, double eps, double reg)
    {
        int r = 0;
        ORC_TRY(r = harten(*static_cast<Sim*>(p), ite, eps, reg) ? 1 : 0);
        return r;
    }

    // tags/details of the last harten call, over the storage of the mesh they
    // were computed on (size orc_tag_storage_size)
    int64_t orc_tag_storage_size(void* p) { return static_cast<Sim*>(p)->tag_mesh.nref; }
    int orc_get_last_tags(void* p, uint8_t* t, double* d)
    {
        auto* s = static_cast<Sim*>(p);
        std::memcpy(t, s->last_tag.data(), s->last_tag.size());
        std::memcpy(d, s->last_detail.data(), s->last_detail.size() * sizeof(double));
        return 0;
    }

    // update_ghost_mr(u) then unp1 = u - dt*upwind(a,u), swap
    // update_ghost_mr(u), then the README hand loop (with its swap)
    int orc_upwind_hand_loop(void* p, double dt)
    {
        auto* s = static_cast<Sim*>(p);
        ORC_TRY({
            update_ghost_mr(*s);
            upwind_hand_loop(*s, dt);
        });
        return 0;
    }

    // out (storage size) = make_convection_upwind(v)(u) after update_ghost_mr(u)
    int orc_convection_upwind(void* p, const double* v, double* out)
    {
        auto* s = static_cast<Sim*>(p);
        ORC_TRY({
            update_ghost_mr(*s);
            std::vector<double> r = convection_upwind(*s, v);
            std::copy(r.begin(), r.end(), out);
        });
        return 0;
    }

    // out (storage size) = make_convection_upwind<Field>()(u) (scalar,
    // non-linear) after update_ghost_mr(u)
    int orc_convection_upwind_nonlinear(void* p, double* out)
    {
        auto* s = static_cast<Sim*>(p);
        ORC_TRY({
            update_ghost_mr(*s);
            std::vector<double> r = convection_upwind_nonlinear(*s);
            std::copy(r.begin(), r.end(), out);
        });
        return 0;
    }

    // out (storage size) = make_convection_weno5(v)(u) after update_ghost_mr(u)
    int orc_convection_weno5(void* p, const double* v, double* out)
    {
        auto* s = static_cast<Sim*>(p);
        ORC_TRY({
            update_ghost_mr(*s);
            std::vector<double> r = convection_weno5(*s, v);
            std::copy(r.begin(), r.end(), out);
        });
        return 0;
    }

    // --refine-boundary for the following adaptations
    int orc_set_refine_boundary(void* p, int on)
    {
        static_cast<Sim*>(p)->refine_boundary = on != 0;
        return 0;
    }

    // mra_config().relative_detail(on) for the following adaptations
    int orc_set_relative_detail(void* p, int on)
    {
        static_cast<Sim*>(p)->relative_detail = on != 0;
        return 0;
    }

    // the WENO5 flux of five values (diagnostic)
    double orc_weno5_flux(const double* f)
    {
        return weno5_flux(f);
    }

    // out = make_convection_upwind<Field>(velocity_field)(u) after
    // update_ghost_mr(u); vel: dim storages of the velocity components, ghosts
    // already updated
    int orc_convection_upwind_heterogeneous(void* p, const double* const* vel, double* out)
    {
        auto* s = static_cast<Sim*>(p);
        ORC_TRY({
            update_ghost_mr(*s);
            std::vector<double> r = convection_upwind_heterogeneous(*s, vel);
            std::copy(r.begin(), r.end(), out);
        });
        return 0;
    }

    // update_ghost_mr(u); unp1 = u - dt * upwind_scalar_burgers(k, u); swap (2D)
    int orc_upwind_scalar_burgers_step(void* p, const double* k, double dt)
    {
        auto* s = static_cast<Sim*>(p);
        ORC_TRY({
            if (s->mesh.dim != 2)
            {
                throw std::invalid_argument("upwind_scalar_burgers is two-dimensional");
            }
            update_ghost_mr(*s);
            upwind_scalar_burgers(*s, k, dt);
        });
        return 0;
    }

    int orc_upwind_step(void* p, const double* a, double dt)
    {
        auto* s = static_cast<Sim*>(p);
        ORC_TRY({
            update_ghost_mr(*s);
            upwind(*s, a, dt);
        });
        return 0;
    }

    // unp1 = u - dt*upwind(a,u) alone (no ghost update), swap
    int orc_upwind_only(void* p, const double* a, double dt)
    {
        ORC_TRY(upwind(*static_cast<Sim*>(p), a, dt));
        return 0;
    }

    // One projection or prediction operator on the sim's own mesh, as the
    // reference's round-trip test applies them
    // (tests/test_projection_prediction_roundtrip.cpp:139-197), over
    // set = intersection(cells[level], reference[level-1]):
    //   op 0: set.on(level-1).apply_op(projection(u))      (numeric/projection.hpp:15-65)
    //   op 1: set.on(level).apply_op(prediction<s,false>(u)) (numeric/prediction.hpp:259-357)
    //   op 2: set.on(level-1).apply_op(prediction<s,true>(u)) (numeric/prediction.hpp:107-257)
    int orc_apply_level_op(void* p, int op, int level)
    {
        auto* s = static_cast<Sim*>(p);
        try
        {
            const Mesh& m = s->mesh;
            if (level < 1 || level >= MAXL || op < 0 || op > 2)
            {
                throw std::invalid_argument("bad level or operator");
            }
            DSet set = intersection(m.ids[CELLS][size_t(level)], m.ids[REFERENCE][size_t(level - 1)]);
            if (op == 0)
            {
                on(set, level - 1).for_each([&](int x, int y, int z) { U(*s, level - 1, x, y, z) = project_value(*s, level, x, y, z); });
            }
            else if (op == 1)
            {
                // every value is computed from the coarse level before any fine one is written
                std::vector<std::pair<int64_t, double>> out;
                on(set, level).for_each([&](int x, int y, int z) { out.emplace_back(m.index(level, x, y, z), predict_fine(*s, level, x, y, z)); });
                for (auto& iv : out)
                {
                    s->u[size_t(iv.first)] = iv.second;
                }
            }
            else
            {
                std::vector<double> dst = s->u;
                on(set, level - 1).for_each([&](int x, int y, int z) { predict_children(m, s->u, m, dst, level - 1, x, y, z); });
                s->u = std::move(dst);
            }
        }
        catch (const std::exception& e)
        {
            g_err = e.what();
            return -1;
        }
        return 0;
    }

    // apply_field_bc(u, direction) (bc/apply_field_bc.hpp:256-267): the
    // Dirichlet<1> ghosts of one Cartesian direction at every level from
    // min_level to max_level, with the sim's boundary value
    int orc_apply_field_bc_direction(void* p, int dx, int dy, int dz)
    {
        auto* s = static_cast<Sim*>(p);
        ORC_TRY({
            if (std::abs(dx) + std::abs(dy) + std::abs(dz) != 1)
            {
                throw std::invalid_argument("not a Cartesian direction");
            }
            for (int level = s->mesh.min_level; level <= s->mesh.max_level; ++level)
            {
                apply_field_bc(*s, level, Dir{dx, dy, dz});
            }
        });
        return 0;
    }
}
