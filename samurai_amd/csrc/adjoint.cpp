#include "adjoint.hpp"

#include "host/parallel.hpp"

#include <algorithm>
#include <cstring>
#include <iterator>
#include <stdexcept>

namespace smr
{
    namespace
    {
        struct Triplet
        {
            int64_t dst, src;
            double w;
        };

        LinearStage make_stage(std::vector<int64_t>&& dst, std::vector<Triplet>& t, int64_t nref)
        {
            for (int64_t d : dst)
            {
                if (d < 0 || d >= nref)
                {
                    throw std::out_of_range("adjoint: ghost output outside the field storage");
                }
            }
            for (const Triplet& e : t)
            {
                if (e.src < 0 || e.src >= nref || e.dst < 0 || e.dst >= nref)
                {
                    throw std::out_of_range("adjoint: ghost stencil outside the field storage");
                }
            }
            LinearStage st;
            st.dst = std::move(dst);
            std::sort(t.begin(), t.end(), [](const Triplet& a, const Triplet& b) { return a.src != b.src ? a.src < b.src : a.dst < b.dst; });
            for (size_t i = 0; i < t.size();)
            {
                size_t j = i;
                st.tgt.push_back(t[i].src);
                for (; j < t.size() && t[j].src == t[i].src; ++j)
                {
                    st.src.push_back(t[j].dst);
                    st.w.push_back(t[j].w);
                }
                st.ptr.push_back(int64_t(st.src.size()));
                i = j;
            }
            return st;
        }

        // storage base of the row (level, y, z) at x (value index = base + x)
        int64_t row_base(const MRMeshData& m, int level, int32_t x, int32_t y, int32_t z)
        {
            int64_t b = 0;
            if (!m.run_base(level, x, y, z, b))
            {
                throw std::out_of_range("adjoint: a stencil row is missing from the reference table");
            }
            return b;
        }

        // interp_coeffs<3> (numeric/prediction.hpp:31-35), by parity and stencil position
        double pcoef(int parity, int k)
        {
            return k == 1 ? 1.0 : ((parity == 0) == (k == 0) ? 1. / 8. : -1. / 8.);
        }

        void add_scatter_list(DevWorkList& w, const MRMeshData& m, const WorkList& cells, const std::array<int, 3>& off)
        {
            w.host.taps = cells.taps;
            for (int level = 0; level < MAXL; ++level)
            {
                const LevelSet& c = m.ids[CELLS][size_t(level)];
                if (c.empty())
                {
                    continue;
                }
                const LevelSet ghosts = set_difference(m.ids[REFERENCE][size_t(level)], c);
                const LevelSet hit    = set_intersection(c, translate(ghosts, {-off[0], -off[1], -off[2]}));
                if (!hit.empty())
                {
                    w.host.add(hit);
                }
            }
            w.host.finish();
        }

        // storage slots of the ghost neighbours in direction off of the cells of w
        std::vector<int64_t> scatter_targets(const DevWorkList& w, const MRMeshData& m, const std::array<int, 3>& off)
        {
            std::vector<int64_t> t;
            t.reserve(size_t(w.host.ncells()));
            const auto& iv = w.host.iv;
            for (size_t r = 0; r + 4 < iv.size(); r += 5)
            {
                const int32_t x0 = iv[r], x1 = iv[r + 1], y = iv[r + 2], z = iv[r + 3], level = iv[r + 4];
                const int64_t b  = m.index(level, x0 + off[0], y + off[1], z + off[2]) - x0;
                for (int32_t x = x0; x < x1; ++x)
                {
                    t.push_back(b + x);
                }
            }
            std::sort(t.begin(), t.end());
            return t;
        }
    } // namespace

    std::vector<std::vector<std::pair<int, int>>> group_reverse_stages(const std::vector<LinearStage>& ghost)
    {
        auto sorted_unique = [](std::vector<int64_t> v)
        {
            std::sort(v.begin(), v.end());
            v.erase(std::unique(v.begin(), v.end()), v.end());
            return v;
        };
        auto overlap = [](const std::vector<int64_t>& a, const std::vector<int64_t>& b)
        {
            size_t i = 0, j = 0;
            while (i < a.size() && j < b.size())
            {
                if (a[i] == b[j])
                {
                    return true;
                }
                if (a[i] < b[j])
                {
                    ++i;
                }
                else
                {
                    ++j;
                }
            }
            return false;
        };
        auto merged = [](const std::vector<int64_t>& a, const std::vector<int64_t>& b)
        {
            std::vector<int64_t> out;
            out.reserve(a.size() + b.size());
            std::set_union(a.begin(), a.end(), b.begin(), b.end(), std::back_inserter(out));
            return out;
        };
        std::vector<std::vector<std::pair<int, int>>> groups;
        std::vector<int64_t> rs, ws; // slots read / written by the open group
        // the reverse sequence: gather of stage k then zeroing of its outputs,
        // for k from the last forward stage down; an item joins the open
        // group when it neither reads nor writes a slot the group writes and
        // writes no slot the group reads, else it opens the next group
        for (size_t k = ghost.size(); k-- > 0;)
        {
            const LinearStage& g = ghost[k];
            for (int kind = 0; kind < 2; ++kind)
            {
                std::vector<int64_t> r, w;
                if (kind == 0)
                {
                    if (g.tgt.empty())
                    {
                        continue;
                    }
                    r = sorted_unique(g.src);
                    w = sorted_unique(g.tgt);
                }
                else
                {
                    if (g.dst.empty())
                    {
                        continue;
                    }
                    w = sorted_unique(g.dst);
                }
                if (!groups.empty() && !overlap(ws, r) && !overlap(ws, w) && !overlap(rs, w))
                {
                    groups.back().emplace_back(int(k), kind);
                    rs = merged(rs, r);
                    ws = merged(ws, w);
                }
                else
                {
                    groups.push_back({{int(k), kind}});
                    rs = std::move(r);
                    ws = std::move(w);
                }
            }
        }
        return groups;
    }

    std::vector<LinearStage> ghost_linear_stages(const MRMeshData& m, const DevMesh& dm, bool has_bc)
    {
        const MeshConfig& cfg = m.cfg;
        const int dim = cfg.dim, s = cfg.pred_s;
        std::vector<LinearStage> out;
        for (int level = cfg.max_level; level >= 0; --level)
        {
            const Recipes& r = dm.outer[size_t(level)].host;
            const auto& sp   = r.stage_ptr;
            for (size_t k = 0; k + 1 < sp.size(); ++k)
            {
                std::vector<int64_t> dst;
                std::vector<Triplet> t;
                for (int64_t q = sp[k]; q < sp[k + 1]; ++q)
                {
                    const int op    = r.op[size_t(q)];
                    const int64_t d = r.dst[size_t(q)];
                    const int64_t a = r.src_ptr[size_t(q)], b = r.src_ptr[size_t(q) + 1];
                    if ((op == RECIPE_DIRICHLET && !has_bc) || (op == RECIPE_COPY_NO_BC && has_bc))
                    {
                        continue; // not written (op_recipe)
                    }
                    dst.push_back(d);
                    if (op == RECIPE_COPY || op == RECIPE_COPY_NO_BC)
                    {
                        t.push_back({d, r.src[size_t(a)], 1.0});
                    }
                    else if (op == RECIPE_EXTRAPOLATE)
                    {
                        static const double w3[3] = {1.0, -3.0, 3.0}, w5[5] = {1.0, -5.0, 10.0, -10.0, 5.0};
                        const double* w = b - a == 3 ? w3 : w5;
                        for (int64_t e = a; e < b; ++e)
                        {
                            t.push_back({d, r.src[size_t(e)], w[e - a]});
                        }
                    }
                    else if (op == RECIPE_DIRICHLET)
                    {
                        t.push_back({d, r.src[size_t(a)], -1.0});
                    }
                    else
                    {
                        for (int64_t e = a; e < b; ++e)
                        {
                            t.push_back({d, r.src[size_t(e)], 1.0 / double(b - a)});
                        }
                    }
                }
                if (!dst.empty())
                {
                    out.push_back(make_stage(std::move(dst), t, m.nref));
                }
            }
            if (level > 0)
            {
                // projection: dst(l-1, x) = sum over the fine rows of u(l, 2x) + u(l, 2x+1), times 2^-dim
                const WorkList& w = dm.projection.host;
                int64_t r0        = 0;
                const int64_t nr  = w.runs_of_levels(level - 1, level, r0);
                std::vector<int64_t> dst;
                std::vector<Triplet> t;
                const double inv = 1.0 / double(1 << dim);
                const int nrows  = 1 << (dim - 1);
                for (int64_t k = r0; k < r0 + nr; ++k)
                {
                    const int32_t* iv = &w.iv[size_t(5 * k)];
                    const int32_t x0 = iv[0], x1 = iv[1], y = iv[2], z = iv[3];
                    const int lv     = iv[4];
                    const int64_t bd = row_base(m, lv, x0, y, z);
                    int64_t bf[4];
                    for (int q = 0; q < nrows; ++q)
                    {
                        const int dy = q & 1, dz = q >> 1;
                        bf[q]        = row_base(m, lv + 1, 2 * x0, 2 * y + dy, 2 * z + dz);
                    }
                    for (int32_t x = x0; x < x1; ++x)
                    {
                        const int64_t d = bd + x;
                        dst.push_back(d);
                        for (int q = 0; q < nrows; ++q)
                        {
                            t.push_back({d, bf[q] + 2 * x, inv});
                            t.push_back({d, bf[q] + 2 * x + 1, inv});
                        }
                    }
                }
                if (!dst.empty())
                {
                    out.push_back(make_stage(std::move(dst), t, m.nref));
                }
            }
        }
        for (int level = 1; level <= cfg.max_level; ++level)
        {
            // prediction <s,false> of the fine ghosts of this level from level-1
            const WorkList& w = dm.prediction.host;
            int64_t r0        = 0;
            const int64_t nr  = w.runs_of_levels(level, level + 1, r0);
            std::vector<int64_t> dst;
            std::vector<Triplet> t;
            const int nz = dim == 3 ? 2 * s + 1 : 1;
            for (int64_t k = r0; k < r0 + nr; ++k)
            {
                const int32_t* iv = &w.iv[size_t(5 * k)];
                const int32_t x0 = iv[0], x1 = iv[1], y = iv[2], z = iv[3];
                const int lv     = iv[4];
                const int64_t bd = row_base(m, lv, x0, y, z);
                const int32_t yc = y >> 1, zc = z >> 1, xs = (x0 >> 1) - s;
                std::vector<int64_t> bc(size_t(nz * (2 * s + 1)));
                for (int rz = 0; rz < nz; ++rz)
                {
                    for (int ry = 0; ry < 2 * s + 1; ++ry)
                    {
                        const int32_t zz = dim == 3 ? zc + rz - s : 0;
                        bc[size_t(rz * (2 * s + 1) + ry)] = row_base(m, lv - 1, xs, yc + ry - s, zz);
                    }
                }
                const int py = y & 1, pz = z & 1;
                for (int32_t x = x0; x < x1; ++x)
                {
                    const int64_t d = bd + x;
                    dst.push_back(d);
                    const int32_t xc = x >> 1;
                    if (s == 0)
                    {
                        t.push_back({d, bc[0] + xc, 1.0});
                        continue;
                    }
                    const int px = x & 1;
                    for (int rz = 0; rz < nz; ++rz)
                    {
                        for (int ry = 0; ry < 3; ++ry)
                        {
                            for (int sx = 0; sx < 3; ++sx)
                            {
                                double c = pcoef(px, sx) * pcoef(py, ry);
                                if (dim == 3)
                                {
                                    c *= pcoef(pz, rz);
                                }
                                t.push_back({d, bc[size_t(rz * 3 + ry)] + xc - 1 + sx, c});
                            }
                        }
                    }
                }
            }
            if (!dst.empty())
            {
                out.push_back(make_stage(std::move(dst), t, m.nref));
            }
        }
        return out;
    }

    // ------------------------------------------------------------------
    // Adjoint
    // ------------------------------------------------------------------
    Adjoint::Adjoint(Ctx* ctx, std::vector<Mesh*> meshes, const double* dt, const double* bc, int steps, int every)
        : ctx_(ctx), steps_(steps), every_(every)
    {
        ctx_->require_device();
        if (steps <= 0 || every <= 0 || steps % every != 0)
        {
            throw std::invalid_argument("adjoint: steps must be a positive multiple of the checkpoint interval");
        }
        nseg_ = steps / every;
        inst_.resize(meshes.size());
        int64_t off = 0;
        for (size_t i = 0; i < meshes.size(); ++i)
        {
            Instance& in = inst_[i];
            if (!meshes[i] || meshes[i]->ctx != ctx_)
            {
                throw std::invalid_argument("adjoint: null mesh or mesh of another context");
            }
            in.mesh    = meshes[i];
            in.version = in.mesh->version;
            in.dt      = dt[i];
            in.has_bc  = bc != nullptr;
            in.bc      = bc ? bc[i] : 0.0;
            in.nref    = in.mesh->host.nref;
            in.ncells  = in.mesh->host.nb_cells(CELLS);
            in.off     = off;
            off += in.ncells;
            in.mesh->ensure_host_lists();
        }
        std::vector<const Mesh*> sorted(meshes.begin(), meshes.end());
        std::sort(sorted.begin(), sorted.end());
        if (std::adjacent_find(sorted.begin(), sorted.end()) != sorted.end())
        {
            throw std::invalid_argument("adjoint: two instances share a mesh");
        }
        // transposed ghost stages and the ghost-neighbour scatter lists, built in parallel
        parallel_for(int(inst_.size()),
                     [&](int i)
                     {
                         Instance& in  = inst_[size_t(i)];
                         const auto& m = in.mesh->host;
                         in.ghost      = ghost_linear_stages(m, *in.mesh->dev, in.has_bc);
                         in.adj_groups = group_reverse_stages(in.ghost);
                         const std::array<std::array<int, 3>, 6> offs = {
                             {{-1, 0, 0}, {1, 0, 0}, {0, -1, 0}, {0, 1, 0}, {0, 0, -1}, {0, 0, 1}}};
                         // directions whose ghost targets are disjoint share a launch
                         std::vector<int64_t> open;
                         in.scatter_groups.clear();
                         for (int d = 0; d < 2 * m.cfg.dim; ++d)
                         {
                             add_scatter_list(in.wdir[size_t(d)], m, in.mesh->dev->cells.host, offs[size_t(d)]);
                             if (in.wdir[size_t(d)].host.ncells() == 0)
                             {
                                 continue;
                             }
                             std::vector<int64_t> t = scatter_targets(in.wdir[size_t(d)], m, offs[size_t(d)]);
                             std::vector<int64_t> both;
                             std::set_intersection(open.begin(), open.end(), t.begin(), t.end(), std::back_inserter(both));
                             if (!in.scatter_groups.empty() && both.empty())
                             {
                                 in.scatter_groups.back().push_back(d);
                                 std::vector<int64_t> u;
                                 std::set_union(open.begin(), open.end(), t.begin(), t.end(), std::back_inserter(u));
                                 open.swap(u);
                             }
                             else
                             {
                                 in.scatter_groups.push_back({d});
                                 open.swap(t);
                             }
                         }
                     });
        hipStream_t st = ctx_->stream;
        for (Instance& in : inst_)
        {
            const int dim = in.mesh->host.cfg.dim;
            in.blob       = std::make_unique<Blob>();
            Packer p(*in.blob);
            for (int d = 0; d < 2 * dim; ++d)
            {
                p.list(in.wdir[size_t(d)]);
            }
            struct Off
            {
                size_t dst, tgt, ptr, src, w;
            };
            std::vector<Off> offs;
            for (const LinearStage& g : in.ghost)
            {
                Off o;
                o.dst = in.blob->add(g.dst.data(), g.dst.size() * 8);
                o.tgt = in.blob->add(g.tgt.data(), g.tgt.size() * 8);
                o.ptr = in.blob->add(g.ptr.data(), g.ptr.size() * 8);
                o.src = in.blob->add(g.src.data(), g.src.size() * 8);
                o.w   = in.blob->add(g.w.data(), g.w.size() * 8);
                offs.push_back(o);
            }
            p.finish(st);
            for (size_t k = 0; k < in.ghost.size(); ++k)
            {
                const LinearStage& g = in.ghost[k];
                Instance::DevStage ds;
                ds.dst   = reinterpret_cast<const int64_t*>(in.blob->dptr(offs[k].dst));
                ds.tgt   = reinterpret_cast<const int64_t*>(in.blob->dptr(offs[k].tgt));
                ds.ptr   = reinterpret_cast<const int64_t*>(in.blob->dptr(offs[k].ptr));
                ds.src   = reinterpret_cast<const int64_t*>(in.blob->dptr(offs[k].src));
                ds.w     = reinterpret_cast<const double*>(in.blob->dptr(offs[k].w));
                ds.ndst  = int64_t(g.dst.size());
                ds.nrows = int64_t(g.tgt.size());
                in.dghost.push_back(ds);
            }
            for (int d = 0; d < 2 * dim; ++d)
            {
                dev::resolve(in.wdir[size_t(d)].d, in.mesh->dev->table, in.mesh->dev->table, ctx_->d_err, st);
            }
            const size_t n = size_t(std::max<int64_t>(in.nref, 1));
            in.states      = static_cast<double*>(dev_alloc(n * size_t(every_ + 1) * 8, st));
            in.ckpt        = static_cast<double*>(dev_alloc(n * size_t(nseg_) * 8, st));
            in.lam         = static_cast<double*>(dev_alloc(n * 2 * 8, st));
            // defined contents everywhere (ghost slots of the states are read before
            // the first ghost update writes them only by the copies, never by operators)
            hip_check(hipMemsetAsync(in.states, 0, n * size_t(every_ + 1) * 8, st), "hipMemsetAsync(states)");
        }
        // the per-cell velocity-gradient terms of all instances, one array
        ga_off_.assign(inst_.size() + 1, 0);
        for (size_t i = 0; i < inst_.size(); ++i)
        {
            ga_off_[i + 1] = ga_off_[i] + inst_[i].ncells * inst_[i].mesh->host.cfg.dim;
        }
        ga_arena_ = static_cast<double*>(dev_alloc(size_t(std::max<int64_t>(ga_off_.back(), 1)) * 8, st));
        for (size_t i = 0; i < inst_.size(); ++i)
        {
            inst_[i].ga = ga_arena_ + ga_off_[i];
        }
        ctx_->check_resolve();
    }

    Adjoint::~Adjoint()
    {
        hipStream_t s = ctx_->stream;
        for (Instance& in : inst_)
        {
            for (double* p : {in.states, in.ckpt, in.lam})
            {
                if (p)
                {
                    dev_free(p, s);
                }
            }
        }
        dev_free(ga_arena_, s);
        (void)hipStreamSynchronize(s);
    }

    int64_t Adjoint::nb_cells(int i) const
    {
        if (i < 0)
        {
            int64_t n = 0;
            for (const Instance& in : inst_)
            {
                n += in.ncells;
            }
            return n;
        }
        return inst_.at(size_t(i)).ncells;
    }

    void Adjoint::check_meshes() const
    {
        for (const Instance& in : inst_)
        {
            if (in.mesh->version != in.version)
            {
                throw std::runtime_error("adjoint: a mesh changed after the adjoint was created (meshes must stay frozen)");
            }
        }
    }

    void Adjoint::set_velocity(const double* a)
    {
        for (size_t i = 0; i < inst_.size(); ++i)
        {
            for (int d = 0; d < 3; ++d)
            {
                inst_[i].a[d] = a[3 * i + size_t(d)];
            }
        }
        programs_ready_ = false;
    }

    std::unique_ptr<ProgramBuf> Adjoint::make_program(const std::vector<Chain>& chains)
    {
        auto p = std::make_unique<ProgramBuf>(ctx_);
        p->append(chains);
        p->upload();
        return p;
    }

    void Adjoint::build_programs()
    {
        const size_t n = inst_.size();
        fseg_.clear();
        bseg_.clear();
        // the chains of every piece (one forward step, one adjoint step of each
        // parity, checkpoint copies), then one program per checkpoint segment
        std::vector<std::vector<Chain>> fwd, adj[2], save, restore;
        auto per = [&](auto&& f)
        {
            std::vector<Chain> chains(n);
            for (size_t i = 0; i < n; ++i)
            {
                f(inst_[i], chains[i]);
            }
            return chains;
        };
        auto copy_seg = [](const Instance& in, double* dst, const double* src)
        {
            dev::Segment g = make_seg(dev::SEG_COPY_RAW, in.mesh->host.cfg);
            g.out          = dst;
            g.u            = const_cast<double*>(src);
            g.c1           = in.nref;
            return g;
        };
        for (int j = 0; j < every_; ++j)
        {
            fwd.push_back(per(
                [&](Instance& in, Chain& ch)
                {
                    DevMesh& dm         = *in.mesh->dev;
                    const MeshConfig& c = in.mesh->host.cfg;
                    ghost_chain(ch, dm, c, in.state(j), in.has_bc, in.bc);
                    dev::Segment g = list_seg(dev::SEG_UPWIND, c, dm.cells, 0, dm.cells.d.ncells);
                    g.u            = in.state(j);
                    g.out          = in.state(j + 1);
                    g.a0           = in.a[0];
                    g.a1           = in.a[1];
                    g.a2           = c.dim == 3 ? in.a[2] : 0.0;
                    g.dt           = in.dt;
                    push_stage(ch, g);
                }));
            for (int p = 0; p < 2; ++p)
            {
                adj[p].push_back(per(
                    [&](Instance& in, Chain& ch)
                    {
                        DevMesh& dm         = *in.mesh->dev;
                        const MeshConfig& c = in.mesh->host.cfg;
                        double* lam         = in.lam + int64_t(p) * in.nref;
                        double* mu          = in.lam + int64_t(1 - p) * in.nref;
                        dev::Segment z      = make_seg(dev::SEG_FILL_F64, c);
                        z.out               = mu;
                        z.c1                = in.nref;
                        push_stage(ch, z);
                        dev::Segment u = list_seg(dev::SEG_ADJ_UPWIND, c, dm.cells, 0, dm.cells.d.ncells);
                        u.u            = lam;
                        u.out          = mu;
                        u.in2          = in.state(j);
                        u.out2         = in.ga;
                        u.a0           = in.a[0];
                        u.a1           = in.a[1];
                        u.a2           = c.dim == 3 ? in.a[2] : 0.0;
                        u.dt           = in.dt;
                        // the gather writes cells only and the scatters ghost slots only, so
                        // the first group of scatters shares the gather's stage; the
                        // directions of a group hit disjoint ghosts
                        Stage st{u};
                        for (size_t q = 0; q < in.scatter_groups.size(); ++q)
                        {
                            if (q > 0)
                            {
                                ch.push_back(std::move(st));
                                st.clear();
                            }
                            for (int d : in.scatter_groups[q])
                            {
                                DevWorkList& w = in.wdir[size_t(d)];
                                dev::Segment g = list_seg(dev::SEG_ADJ_SCATTER, c, w, 0, w.d.ncells);
                                g.pad          = d;
                                g.u            = lam;
                                g.out          = mu;
                                g.a0           = u.a0;
                                g.a1           = u.a1;
                                g.a2           = u.a2;
                                g.dt           = in.dt;
                                st.push_back(g);
                            }
                        }
                        ch.push_back(std::move(st));
                        for (const auto& grp : in.adj_groups)
                        {
                            Stage st;
                            for (const auto& [k, kind] : grp)
                            {
                                const Instance::DevStage& ds = in.dghost[size_t(k)];
                                if (kind == 0)
                                {
                                    dev::Segment g = make_seg(dev::SEG_ADJ_GATHER, c);
                                    g.rdst         = ds.tgt;
                                    g.rsp          = ds.ptr;
                                    g.rsrc         = ds.src;
                                    g.thr          = ds.w;
                                    g.out          = mu;
                                    g.c1           = ds.nrows;
                                    st.push_back(g);
                                }
                                else
                                {
                                    dev::Segment zr = make_seg(dev::SEG_ADJ_ZERO, c);
                                    zr.rdst         = ds.dst;
                                    zr.out          = mu;
                                    zr.c1           = ds.ndst;
                                    st.push_back(zr);
                                }
                            }
                            ch.push_back(std::move(st));
                        }
                    }));
            }
        }
        for (int sg = 0; sg < nseg_; ++sg)
        {
            save.push_back(per([&](Instance& in, Chain& ch) { push_stage(ch, copy_seg(in, in.ckpt + int64_t(sg) * in.nref, in.state(0))); }));
            restore.push_back(
                per([&](Instance& in, Chain& ch) { push_stage(ch, copy_seg(in, in.state(0), in.ckpt + int64_t(sg) * in.nref)); }));
        }
        const std::vector<Chain> roll = per([&](Instance& in, Chain& ch) { push_stage(ch, copy_seg(in, in.state(0), in.state(every_))); });
        for (int sg = 0; sg < nseg_; ++sg)
        {
            // forward: checkpoint, `every` steps, the last state becomes the first
            auto f = std::make_unique<ProgramBuf>(ctx_);
            f->set_independent_chains(true); // chain i is instance i throughout
            f->append(save[size_t(sg)]);
            for (int j = 0; j < every_; ++j)
            {
                f->append(fwd[size_t(j)]);
            }
            f->append(roll);
            f->upload();
            fseg_.push_back(std::move(f));
            // backward: restore, recompute the states, then the adjoint steps
            // backwards; lam and mu swap every step, from the parity of the
            // number of adjoint steps of the later segments
            auto b = std::make_unique<ProgramBuf>(ctx_);
            b->set_independent_chains(true);
            b->append(restore[size_t(sg)]);
            for (int j = 0; j < every_; ++j)
            {
                b->append(fwd[size_t(j)]);
            }
            int done = (nseg_ - 1 - sg) * every_;
            for (int j = every_ - 1; j >= 0; --j)
            {
                b->append(adj[done % 2][size_t(j)]);
                ++done;
            }
            b->upload();
            bseg_.push_back(std::move(b));
        }
        programs_ready_ = true;
        bound_u0_       = nullptr;
        bound_uT_       = nullptr;
        bound_gT_       = nullptr;
        bound_g0_       = nullptr;
    }

    void Adjoint::bind_io(const double* u0, double* uT, const double* gT, double* g0)
    {
        const size_t n = inst_.size();
        auto per       = [&](auto&& f)
        {
            std::vector<Chain> chains(n);
            for (size_t i = 0; i < n; ++i)
            {
                f(inst_[i], chains[i]);
            }
            return make_program(chains);
        };
        if (u0 && u0 != bound_u0_)
        {
            in_ = per(
                [&](Instance& in, Chain& ch)
                {
                    DevMesh& dm    = *in.mesh->dev;
                    dev::Segment g = list_seg(dev::SEG_CELLS_FROM_COMPACT, in.mesh->host.cfg, dm.cells, 0, dm.cells.d.ncells);
                    g.out          = in.state(0);
                    g.in2          = u0 + in.off;
                    push_stage(ch, g);
                });
            bound_u0_ = u0;
        }
        if (uT && uT != bound_uT_)
        {
            out_ = per(
                [&](Instance& in, Chain& ch)
                {
                    DevMesh& dm    = *in.mesh->dev;
                    dev::Segment g = list_seg(dev::SEG_CELLS_TO_COMPACT, in.mesh->host.cfg, dm.cells, 0, dm.cells.d.ncells);
                    g.out          = uT + in.off;
                    g.u            = in.state(0);
                    push_stage(ch, g);
                });
            bound_uT_ = uT;
        }
        if (gT && gT != bound_gT_)
        {
            lam_in_ = per(
                [&](Instance& in, Chain& ch)
                {
                    DevMesh& dm         = *in.mesh->dev;
                    const MeshConfig& c = in.mesh->host.cfg;
                    dev::Segment z      = make_seg(dev::SEG_FILL_F64, c);
                    z.out               = in.lam;
                    z.c1                = in.nref;
                    dev::Segment zg     = make_seg(dev::SEG_FILL_F64, c);
                    zg.out              = in.ga;
                    zg.c1               = in.ncells * c.dim;
                    ch.push_back(Stage{z, zg});
                    dev::Segment g = list_seg(dev::SEG_CELLS_FROM_COMPACT, c, dm.cells, 0, dm.cells.d.ncells);
                    g.out          = in.lam;
                    g.in2          = gT + in.off;
                    push_stage(ch, g);
                });
            bound_gT_ = gT;
        }
        if (g0 && g0 != bound_g0_)
        {
            // after steps_ adjoint steps the result is in lam parity steps_ % 2
            grad_out_ = per(
                [&](Instance& in, Chain& ch)
                {
                    DevMesh& dm    = *in.mesh->dev;
                    dev::Segment g = list_seg(dev::SEG_CELLS_TO_COMPACT, in.mesh->host.cfg, dm.cells, 0, dm.cells.d.ncells);
                    g.out          = g0 + in.off;
                    g.u            = in.lam + int64_t(steps_ % 2) * in.nref;
                    push_stage(ch, g);
                });
            bound_g0_ = g0;
        }
    }

    void Adjoint::forward(const double* u0, double* uT)
    {
        check_meshes();
        if (!programs_ready_)
        {
            build_programs();
        }
        bind_io(u0, uT, nullptr, nullptr);
        ScopedTimer t(ctx_, "adjoint forward");
        in_->run(tuning_);
        for (int sg = 0; sg < nseg_; ++sg)
        {
            fseg_[size_t(sg)]->run(tuning_);
        }
        out_->run(tuning_);
        ctx_->sync();
    }

    void Adjoint::backward(const double* gT, double* g0, double* ga)
    {
        check_meshes();
        if (!programs_ready_)
        {
            throw std::logic_error("adjoint: backward before forward");
        }
        bind_io(nullptr, nullptr, gT, g0);
        ScopedTimer t(ctx_, "adjoint backward");
        lam_in_->run(tuning_);
        for (int sg = nseg_ - 1; sg >= 0; --sg)
        {
            bseg_[size_t(sg)]->run(tuning_);
        }
        grad_out_->run(tuning_);
        // velocity gradient: per-cell terms of all instances in one copy, summed
        // on the host per instance in cell order
        const size_t total = size_t(ga_off_.back());
        const double* h    = reinterpret_cast<const double*>(ga_host_.acquire(std::max<size_t>(total, 1) * 8));
        if (total > 0)
        {
            hip_check(hipMemcpyAsync(const_cast<double*>(h), ga_arena_, total * 8, hipMemcpyDeviceToHost, ctx_->stream),
                      "hipMemcpyAsync(ga)");
        }
        ctx_->sync();
        parallel_for(int(inst_.size()),
                     [&](int q)
                     {
                         const size_t i     = size_t(q);
                         const Instance& in = inst_[i];
                         const int dim      = in.mesh->host.cfg.dim;
                         const double* hi   = h + ga_off_[i];
                         double acc[3]      = {0, 0, 0};
                         for (int64_t c = 0; c < in.ncells; ++c)
                         {
                             for (int d = 0; d < dim; ++d)
                             {
                                 acc[d] += hi[c * dim + d];
                             }
                         }
                         for (int d = 0; d < 3; ++d)
                         {
                             ga[3 * i + size_t(d)] = d < dim ? acc[d] : 0.0;
                         }
                     });
    }
} // namespace smr
