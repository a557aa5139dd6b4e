// Single-core CPU baseline of the adaptive step (test infrastructure: timed
// by bench.py's cpu_baseline leg and checked against the oracle by
// tests/test_cpu_step.py; never used by the product).
//
// It runs samurai's algorithm as the reference does on the CPU: interval
// sets, not dense bitmaps.  The mesh, its ghost layers, the storage
// numbering, the work lists, update_cell_array_from_tag and make_graduation
// are the product's host interval algebra (samurai_amd/csrc/host); the
// per-cell operators are the product's operator bodies
// (samurai_amd/csrc/device/cell_ops.hpp, host-callable), looped over host
// copies of the lists with their taps resolved on the host.  One thread
// (SMR_HOST_THREADS=1 when timed), compiled -O3 -ffp-contract=off.
//
// The step follows demos/FiniteVolume/advection_2d.cpp:129-153:
// MRadaptation (mr/adapt.hpp:147-389), update_ghost_mr
// (algorithm/update_ghost_mr.hpp:194-237), unp1 = u - dt * upwind(a, u)
// (stencil_field.hpp:16-179), swap.
#include "../samurai_amd/csrc/device/cell_ops.hpp"
#include "../samurai_amd/csrc/host/programs.hpp"

#include <algorithm>
#include <array>
#include <cmath>
#include <cstring>
#include <memory>
#include <stdexcept>
#include <string>
#include <vector>

namespace cpu
{
    using namespace smr;

    // the reference interval table of a mesh, as the device's dev::Table
    struct HostTable
    {
        std::vector<int64_t> key, row_off, rp, idx;
        std::vector<int32_t> rs, re;
        dev::Table t{};

        explicit HostTable(const MRMeshData& m)
        {
            row_off.assign(MAXL + 1, 0);
            rp.push_back(0);
            for (int l = 0; l < MAXL; ++l)
            {
                row_off[size_t(l)] = int64_t(key.size());
                const LevelSet& r  = m.ids[REFERENCE][size_t(l)];
                for (size_t row = 0; row < r.nrows(); ++row)
                {
                    key.push_back((int64_t(r.rz[row]) << 32) + (int64_t(r.ry[row]) + 2147483648LL));
                    for (int64_t k = r.rp[row]; k < r.rp[row + 1]; ++k)
                    {
                        rs.push_back(r.runs[size_t(k)].s);
                        re.push_back(r.runs[size_t(k)].e);
                        idx.push_back(m.ridx[size_t(l)][size_t(k)]);
                    }
                    rp.push_back(int64_t(rs.size()));
                }
            }
            row_off[MAXL] = int64_t(key.size());
            t.row_key     = key.data();
            t.row_off     = row_off.data();
            t.rp          = rp.data();
            t.run_s       = rs.data();
            t.run_e       = re.data();
            t.run_idx     = idx.data();
        }
    };

    // a work list with its taps resolved against tables t0 (and t1)
    struct HostList
    {
        WorkList w;
        std::vector<int32_t> hint;
        std::vector<int8_t> taps;
        std::vector<int64_t> base;
        dev::List l{};

        HostList(WorkList&& wl, const dev::Table& t0, const dev::Table& t1)
            : w(std::move(wl))
        {
            const int64_t nruns = w.nruns(), ncells = w.ncells();
            const int64_t nch   = (ncells + (1 << dev::HINT_SHIFT) - 1) >> dev::HINT_SHIFT;
            hint.assign(size_t(nch) + 1, 0);
            int64_t r = 0;
            for (int64_t j = 0; j < nch; ++j)
            {
                while (w.pref[size_t(r + 1)] <= (j << dev::HINT_SHIFT))
                {
                    ++r;
                }
                hint[size_t(j)] = int32_t(r);
            }
            hint[size_t(nch)] = int32_t(std::max<int64_t>(nruns - 1, 0));
            for (const Tap& tp : w.taps)
            {
                taps.insert(taps.end(), {tp.table, tp.dl, tp.dx, tp.dy, tp.dz, tp.wlo, tp.whi});
            }
            base.assign(size_t(nruns * int64_t(w.taps.size())), 0);
            l.iv     = w.iv.data();
            l.pref   = w.pref.data();
            l.hint   = hint.data();
            l.base   = base.data();
            l.taps   = taps.data();
            l.nruns  = nruns;
            l.ncells = ncells;
            l.ntaps  = int(w.taps.size());
            int err  = 0;
            for (int64_t g = 0; g < nruns * l.ntaps; ++g)
            {
                dev::op_resolve(l, g, t0, t1, &err);
            }
            if (err)
            {
                throw std::out_of_range("a work-list row is missing from the reference interval table");
            }
        }

        void cells_of_levels(int l0, int l1, int64_t& c0, int64_t& c1) const
        {
            int64_t r0       = 0;
            const int64_t nr = w.runs_of_levels(l0, l1, r0);
            c0               = w.pref[size_t(r0)];
            c1               = w.pref[size_t(r0 + nr)];
        }
    };

    // one mesh version with its programs
    struct Image
    {
        MRMeshData m;
        std::unique_ptr<HostTable> table;
        std::unique_ptr<HostList> cells, proj, pred, detail, criteria, maximum;
        std::vector<Recipes> outer;

        explicit Image(MRMeshData&& md)
            : m(std::move(md))
        {
            table = std::make_unique<HostTable>(m);
            WorkList wc;
            GhostProgram g;
            AdaptProgram a;
            build_programs(m, wc, g, a);
            const dev::Table& t = table->t;
            cells               = std::make_unique<HostList>(std::move(wc), t, t);
            proj                = std::make_unique<HostList>(std::move(g.projection), t, t);
            pred                = std::make_unique<HostList>(std::move(g.prediction), t, t);
            detail              = std::make_unique<HostList>(std::move(a.detail), t, t);
            criteria            = std::make_unique<HostList>(std::move(a.criteria), t, t);
            maximum             = std::make_unique<HostList>(std::move(a.maximum), t, t);
            outer               = std::move(g.outer);
        }
    };

    struct Sim
    {
        std::unique_ptr<Image> img;
        std::vector<double> u, det;
        std::vector<uint8_t> tag;
        bool has_bc = false;
        double bc   = 0;

        int dim() const
        {
            return img->m.cfg.dim;
        }

        template <int DIM>
        void ghost_update()
        {
            Image& im      = *img;
            const int maxl = im.m.cfg.max_level, ps = im.m.cfg.pred_s;
            double* f      = u.data();
            for (int level = maxl; level >= 0; --level)
            {
                const Recipes& r = im.outer[size_t(level)];
                for (size_t k = 0; k + 1 < r.stage_ptr.size(); ++k)
                {
                    for (int64_t q = r.stage_ptr[k]; q < r.stage_ptr[k + 1]; ++q)
                    {
                        dev::op_recipe(r.dst.data(), r.op.data(), r.src_ptr.data(), r.src.data(), q, f, has_bc, bc);
                    }
                }
                if (level > 0)
                {
                    int64_t c0, c1;
                    im.proj->cells_of_levels(level - 1, level, c0, c1);
                    for (int64_t c = c0; c < c1; ++c)
                    {
                        dev::op_project<DIM>(im.proj->l, c, f, f);
                    }
                }
            }
            for (int level = 1; level <= maxl; ++level)
            {
                int64_t c0, c1;
                im.pred->cells_of_levels(level, level + 1, c0, c1);
                for (int64_t c = c0; c < c1; ++c)
                {
                    if (ps == 0)
                    {
                        dev::op_predict_fine<DIM, 0>(im.pred->l, c, f);
                    }
                    else
                    {
                        dev::op_predict_fine<DIM, 1>(im.pred->l, c, f);
                    }
                }
            }
        }

        void update_ghost_mr()
        {
            if (dim() == 2)
            {
                ghost_update<2>();
            }
            else
            {
                ghost_update<3>();
            }
        }

        // mr/adapt.hpp:325-339
        static dev::Thresholds thresholds(const MeshConfig& cfg, int ite, double eps, double regularity)
        {
            dev::Thresholds thr{};
            for (int level = cfg.min_level; level <= cfg.max_level - ite; ++level)
            {
                const int exponent = cfg.dim * (cfg.max_level - level);
                const double eps_l = eps / (1 << exponent);
                const double fine  = std::pow(2.0, regularity + cfg.dim) * eps_l;
                thr.eps[level]     = eps_l;
                thr.fine[level]    = fine;
                thr.coarse[level]  = fine / (1 << cfg.dim);
            }
            return thr;
        }

        // one harten call (mr/adapt.hpp:276-389); true at the fixed point
        template <int DIM>
        bool harten_dim(int ite, double eps, double regularity)
        {
            Image& im            = *img;
            const MeshConfig cfg = im.m.cfg;
            const int ps         = cfg.pred_s;
            const size_t n       = size_t(im.m.nref);
            det.assign(n, 0.0);
            tag.assign(n, 0);
            for (int64_t c = 0; c < im.cells->l.ncells; ++c)
            {
                dev::op_tag_init(im.cells->l, c, tag.data());
            }
            ghost_update<DIM>();
            int64_t c0, c1;
            im.detail->cells_of_levels(cfg.min_level > 0 ? cfg.min_level - 1 : 0, cfg.max_level - ite, c0, c1);
            for (int64_t c = c0; c < c1; ++c)
            {
                if (ps == 0)
                {
                    dev::op_detail<DIM, 0>(im.detail->l, c, u.data(), det.data());
                }
                else
                {
                    dev::op_detail<DIM, 1>(im.detail->l, c, u.data(), det.data());
                }
            }
            const dev::Thresholds thr = thresholds(cfg, ite, eps, regularity);
            im.criteria->cells_of_levels(cfg.min_level - 1, cfg.max_level - ite, c0, c1);
            for (int64_t c = c0; c < c1; ++c)
            {
                dev::op_criteria<DIM>(im.criteria->l, c, det.data(), tag.data(), thr.eps, cfg.min_level, cfg.max_level);
            }
            for (int level = cfg.max_level; level > 0; --level)
            {
                im.maximum->cells_of_levels(level - 1, level, c0, c1);
                for (int64_t c = c0; c < c1; ++c)
                {
                    dev::op_maximum<DIM>(im.maximum->l, c, tag.data());
                }
            }
            // mesh update (algorithm/graduation.hpp:743-843, 573-726)
            CellArray ca = update_cell_array_from_tag(im.m, tag.data());
            make_graduation(ca, cfg.dim, cfg.grad_width);
            if (ca_equal(im.m.ids[CELLS], ca))
            {
                return true;
            }
            MRMeshData nm = build_mesh(cfg, ca);
            nm.graduated  = true;
            // update_fields (algorithm/update_fields.hpp:27-127)
            WorkList wcopy, wcoarsen, wrefine;
            build_update_fields_lists(im.m, nm, wcopy, wcoarsen, wrefine);
            auto nimg             = std::make_unique<Image>(std::move(nm));
            const dev::Table& t0  = nimg->table->t;
            const dev::Table& t1  = im.table->t;
            HostList copy(std::move(wcopy), t0, t1), coarsen(std::move(wcoarsen), t0, t1), refine(std::move(wrefine), t0, t1);
            std::vector<double> nu(size_t(nimg->m.nref), 0.0);
            for (int64_t c = 0; c < copy.l.ncells; ++c)
            {
                dev::op_copy(copy.l, c, nu.data(), u.data());
            }
            for (int64_t c = 0; c < coarsen.l.ncells; ++c)
            {
                dev::op_project<DIM>(coarsen.l, c, nu.data(), u.data());
            }
            for (int64_t c = 0; c < refine.l.ncells; ++c)
            {
                if (ps == 0)
                {
                    dev::op_predict_children<DIM, 0>(refine.l, c, nu.data(), u.data());
                }
                else
                {
                    dev::op_predict_children<DIM, 1>(refine.l, c, nu.data(), u.data());
                }
            }
            u.swap(nu);
            img = std::move(nimg);
            return false;
        }

        // Adapt::operator() (mr/adapt.hpp:147-195); the number of harten calls
        int adapt(double eps, double regularity)
        {
            const int minl = img->m.cfg.min_level, maxl = img->m.cfg.max_level;
            if (minl == maxl)
            {
                return 0;
            }
            for (int i = 0; i < maxl - minl; ++i)
            {
                const bool fixed = dim() == 2 ? harten_dim<2>(i, eps, regularity) : harten_dim<3>(i, eps, regularity);
                if (fixed)
                {
                    return i + 1;
                }
            }
            return maxl - minl;
        }

        // update_ghost_mr, unp1 = u - dt * upwind(a, u), swap
        void upwind_step(const double* a, double dt)
        {
            update_ghost_mr();
            std::vector<double> unp1(u.size(), 0.0);
            const dev::List& l = img->cells->l;
            const double a2    = dim() == 3 ? a[2] : 0.0;
            for (int64_t c = 0; c < l.ncells; ++c)
            {
                const dev::Cell cell = dev::locate(l, c);
                const int64_t* b     = dev::bases(l, cell.run);
                if (dim() == 2)
                {
                    unp1[size_t(b[0] + cell.x)] = dev::upwind_value<2>(b, cell.x, cell.level, u.data(), .5 * a[0], .5 * std::fabs(a[0]),
                                                                       .5 * a[1], .5 * std::fabs(a[1]), .5 * a2, .5 * std::fabs(a2), dt);
                }
                else
                {
                    unp1[size_t(b[0] + cell.x)] = dev::upwind_value<3>(b, cell.x, cell.level, u.data(), .5 * a[0], .5 * std::fabs(a[0]),
                                                                       .5 * a[1], .5 * std::fabs(a[1]), .5 * a2, .5 * std::fabs(a2), dt);
                }
            }
            u.swap(unp1);
        }

        // the leaves in traversal order: (level, x, y, z, storage index)
        std::vector<int64_t> cells() const
        {
            std::vector<int64_t> out;
            const MRMeshData& m = img->m;
            for (int l = 0; l < MAXL; ++l)
            {
                m.ids[CELLS][size_t(l)].for_each_run(
                    [&](int32_t x0, int32_t x1, int32_t y, int32_t z)
                    {
                        const int64_t b = m.index(l, x0, y, z) - x0;
                        for (int32_t x = x0; x < x1; ++x)
                        {
                            out.insert(out.end(), {int64_t(l), int64_t(x), int64_t(y), int64_t(z), b + x});
                        }
                    });
            }
            return out;
        }
    };

    thread_local std::string g_error;
} // namespace cpu

extern "C"
{
    const char* cpu_last_error()
    {
        return cpu::g_error.c_str();
    }

    // MRMesh(ca, cfg) from n leaves (level, x, y, z), values zero
    void* cpu_create_from_cells(int dim, int min_level, int max_level, int pred_s, const int32_t* cells, int64_t n)
    {
        try
        {
            smr::MeshConfig cfg;
            cfg.dim       = dim;
            cfg.min_level = min_level;
            cfg.max_level = max_level;
            cfg.pred_s    = pred_s;
            std::vector<std::vector<std::array<int32_t, 4>>> q(smr::MAXL);
            for (int64_t k = 0; k < n; ++k)
            {
                const int32_t* r = cells + 4 * k;
                q[size_t(r[0])].push_back({r[2], dim == 3 ? r[3] : 0, r[1], r[1] + 1});
            }
            smr::CellArray ca = smr::empty_cell_array(dim);
            for (int l = 0; l < smr::MAXL; ++l)
            {
                ca[size_t(l)] = smr::from_runs(dim, l, q[size_t(l)]);
            }
            smr::CellArray g = ca;
            smr::make_graduation(g, dim, cfg.grad_width);
            smr::MRMeshData m = smr::build_mesh(cfg, ca);
            m.graduated       = smr::ca_equal(g, ca);
            auto* s           = new cpu::Sim;
            s->img            = std::make_unique<cpu::Image>(std::move(m));
            s->u.assign(size_t(s->img->m.nref), 0.0);
            return s;
        }
        catch (const std::exception& e)
        {
            cpu::g_error = e.what();
            return nullptr;
        }
    }

    void cpu_free(void* p)
    {
        delete static_cast<cpu::Sim*>(p);
    }

    int64_t cpu_nb_cells(void* p)
    {
        return static_cast<cpu::Sim*>(p)->img->m.nb_cells(smr::CELLS);
    }

    int cpu_get_cells(void* p, int64_t* out)
    {
        const auto c = static_cast<cpu::Sim*>(p)->cells();
        std::memcpy(out, c.data(), c.size() * 8);
        return 0;
    }

    int cpu_set_values(void* p, const double* v)
    {
        auto* s      = static_cast<cpu::Sim*>(p);
        const auto c = s->cells();
        for (size_t k = 0; k < c.size() / 5; ++k)
        {
            s->u[size_t(c[5 * k + 4])] = v[k];
        }
        return 0;
    }

    int cpu_get_values(void* p, double* v)
    {
        auto* s      = static_cast<cpu::Sim*>(p);
        const auto c = s->cells();
        for (size_t k = 0; k < c.size() / 5; ++k)
        {
            v[k] = s->u[size_t(c[5 * k + 4])];
        }
        return 0;
    }

    void cpu_set_dirichlet(void* p, double value)
    {
        auto* s   = static_cast<cpu::Sim*>(p);
        s->has_bc = true;
        s->bc     = value;
    }

    int cpu_adapt(void* p, double eps, double regularity)
    {
        try
        {
            return static_cast<cpu::Sim*>(p)->adapt(eps, regularity);
        }
        catch (const std::exception& e)
        {
            cpu::g_error = e.what();
            return -1;
        }
    }

    int cpu_upwind_step(void* p, const double* a, double dt)
    {
        try
        {
            static_cast<cpu::Sim*>(p)->upwind_step(a, dt);
            return 0;
        }
        catch (const std::exception& e)
        {
            cpu::g_error = e.what();
            return -1;
        }
    }
}
