// extern "C" boundary (include/samurai_amd.h) over the runtime.
#include "../../include/samurai_amd.h"
#include "adjoint.hpp"
#include "batch.hpp"
#include "runtime.hpp"

#include <algorithm>
#include <cstring>
#include <string>

struct smr_ctx
{
    smr::Ctx c;
};
struct smr_mesh
{
    smr::Mesh m;
};
struct smr_field
{
    smr::Field f;
};
struct smr_adapt
{
    smr::AdaptState a;
};
struct smr_batch
{
    std::unique_ptr<smr::Batch> b;
};
struct smr_adjoint
{
    std::unique_ptr<smr::Adjoint> a;
};

namespace
{
    thread_local std::string g_error;

    template <class F>
    int guarded(F&& f)
    {
        try
        {
            return f();
        }
        catch (const std::invalid_argument& e)
        {
            g_error = e.what();
            return SMR_ERR_ARGUMENT;
        }
        catch (const std::out_of_range& e)
        {
            g_error = e.what();
            return SMR_ERR_RANGE;
        }
        catch (const std::exception& e)
        {
            g_error = e.what();
            std::string w = e.what();
            return w.find("has no GPU") != std::string::npos ? SMR_ERR_NO_DEVICE : SMR_ERR_RUNTIME;
        }
        catch (...)
        {
            g_error = "unknown error";
            return SMR_ERR_RUNTIME;
        }
    }

    void require(bool cond, const char* what)
    {
        if (!cond)
        {
            throw std::invalid_argument(what);
        }
    }

    smr::MeshConfig to_cfg(const smr_mesh_config* c)
    {
        require(c != nullptr, "null mesh config");
        require(c->dim == 2 || c->dim == 3, "dim must be 2 or 3");
        smr::MeshConfig cfg;
        cfg.dim        = c->dim;
        cfg.min_level  = c->min_level;
        cfg.max_level  = c->max_level;
        cfg.pred_s     = c->prediction_stencil_radius;
        cfg.radius     = c->max_stencil_radius;
        cfg.grad_width = c->graduation_width;
        require(cfg.grad_width >= 1, "graduation width must be >= 1");
        return cfg;
    }

    smr_mesh_config from_cfg(const smr::MeshConfig& c)
    {
        smr_mesh_config o;
        o.dim                       = c.dim;
        o.min_level                 = c.min_level;
        o.max_level                 = c.max_level;
        o.prediction_stencil_radius = c.pred_s;
        o.max_stencil_radius        = c.radius;
        o.graduation_width          = c.grad_width;
        return o;
    }

    // MRMesh(ca, cfg) (mr/mesh.hpp:386-396) from leaves given as rows
    // (level, x, y, z); graduated is recorded, not enforced
    smr::MRMeshData mesh_from_cells(const smr::MeshConfig& c, const int32_t* cells, int64_t n)
    {
        require(n >= 0, "negative cell count");
        std::vector<std::vector<std::array<int32_t, 4>>> q(smr::MAXL);
        for (int64_t k = 0; k < n; ++k)
        {
            const int32_t* r = cells + 4 * k;
            require(r[0] >= 0 && r[0] < smr::MAXL, "cell level out of range");
            q[size_t(r[0])].push_back({r[2], c.dim == 3 ? r[3] : 0, r[1], r[1] + 1});
        }
        smr::CellArray ca = smr::empty_cell_array(c.dim);
        for (int l = 0; l < smr::MAXL; ++l)
        {
            ca[size_t(l)] = smr::from_runs(c.dim, l, q[size_t(l)]);
        }
        smr::CellArray g = ca;
        smr::make_graduation(g, c.dim, c.grad_width);
        smr::MRMeshData m = smr::build_mesh(c, ca);
        m.graduated       = smr::ca_equal(g, ca);
        return m;
    }

    std::vector<int64_t> cell_storage_index(const smr::MRMeshData& m)
    {
        std::vector<int64_t> idx;
        idx.reserve(size_t(m.nb_cells(smr::CELLS)));
        for (int l = 0; l < smr::MAXL; ++l)
        {
            const smr::LevelSet& s = m.ids[smr::CELLS][size_t(l)];
            s.for_each_run(
                [&](int32_t x0, int32_t x1, int32_t y, int32_t z)
                {
                    const int64_t b = m.index(l, x0, y, z) - x0;
                    for (int32_t x = x0; x < x1; ++x)
                    {
                        idx.push_back(b + x);
                    }
                });
        }
        return idx;
    }
} // namespace

extern "C"
{
    const char* smr_last_error(void)
    {
        return g_error.c_str();
    }

    int smr_ctx_create(int device, smr_ctx** out)
    {
        return guarded(
            [&]
            {
                require(out != nullptr, "null output pointer");
                auto* c     = new smr_ctx();
                c->c.device = device;
                if (device >= 0)
                {
                    try
                    {
                        smr::hip_check(hipSetDevice(device), "hipSetDevice");
                        smr::hip_check(hipStreamCreateWithFlags(&c->c.stream, hipStreamNonBlocking), "hipStreamCreate");
                        // keep freed stream-ordered memory cached in the default pool
                        hipMemPool_t pool = nullptr;
                        smr::hip_check(hipDeviceGetDefaultMemPool(&pool, device), "hipDeviceGetDefaultMemPool");
                        uint64_t threshold = UINT64_MAX;
                        smr::hip_check(hipMemPoolSetAttribute(pool, hipMemPoolAttrReleaseThreshold, &threshold),
                                       "hipMemPoolSetAttribute");
                        smr::hip_check(hipMalloc(reinterpret_cast<void**>(&c->c.d_err), 2 * sizeof(int)), "hipMalloc(err)");
                        smr::hip_check(hipMemsetAsync(c->c.d_err, 0, 2 * sizeof(int), c->c.stream), "hipMemsetAsync(err)");
                        c->c.sync();
                        smr::noalloc_init(device);
                    }
                    catch (...)
                    {
                        delete c;
                        throw;
                    }
                }
                *out = c;
                return 0;
            });
    }

    int smr_ctx_destroy(smr_ctx* ctx)
    {
        return guarded(
            [&]
            {
                delete ctx;
                return 0;
            });
    }

    int smr_ctx_synchronize(smr_ctx* ctx)
    {
        return guarded(
            [&]
            {
                require(ctx != nullptr, "null context");
                if (ctx->c.has_device())
                {
                    ctx->c.sync();
                }
                return 0;
            });
    }

    int smr_mesh_create_box(smr_ctx* ctx, const smr_mesh_config* cfg, smr_mesh** out)
    {
        return guarded(
            [&]
            {
                require(ctx && out, "null argument");
                auto* m   = new smr_mesh();
                m->m.ctx  = &ctx->c;
                try
                {
                    m->m.host = smr::build_box_mesh(to_cfg(cfg));
                }
                catch (...)
                {
                    delete m;
                    throw;
                }
                *out = m;
                return 0;
            });
    }

    int smr_mesh_create_from_cells(smr_ctx* ctx, const smr_mesh_config* cfg, const int32_t* cells, int64_t n, smr_mesh** out)
    {
        return guarded(
            [&]
            {
                require(ctx && out && (cells || n == 0), "null argument");
                auto m   = std::make_unique<smr_mesh>();
                m->m.ctx = &ctx->c;
                m->m.host = mesh_from_cells(to_cfg(cfg), cells, n);
                *out      = m.release();
                return 0;
            });
    }

    int smr_mesh_assign_cells(smr_mesh* mesh, int min_level, int max_level, const int32_t* cells, int64_t n)
    {
        return guarded(
            [&]
            {
                require(mesh && (cells || n == 0), "null argument");
                smr_mesh_config c = from_cfg(mesh->m.host.cfg);
                c.min_level       = min_level;
                c.max_level       = max_level;
                mesh->m.replace(mesh_from_cells(to_cfg(&c), cells, n), nullptr);
                return 0;
            });
    }

    int smr_mesh_swap(smr_mesh* a, smr_mesh* b)
    {
        return guarded(
            [&]
            {
                require(a && b, "null mesh");
                require(a->m.ctx == b->m.ctx, "swap of meshes of different contexts");
                if (a == b)
                {
                    return 0;
                }
                // device tables are rebuilt from the host data on next use;
                // replace() moves both meshes to a new version, so fields on
                // either must be resized before their next use
                smr::MRMeshData ha = std::move(a->m.host);
                a->m.replace(std::move(b->m.host), nullptr);
                b->m.replace(std::move(ha), nullptr);
                return 0;
            });
    }

    int smr_mesh_destroy(smr_mesh* mesh)
    {
        return guarded(
            [&]
            {
                delete mesh;
                return 0;
            });
    }

    int64_t smr_mesh_nb_cells(const smr_mesh* mesh, int mesh_id)
    {
        if (!mesh || mesh_id < 0 || mesh_id >= smr::N_MESH_IDS)
        {
            g_error = "bad mesh or mesh id";
            return SMR_ERR_ARGUMENT;
        }
        return mesh->m.host.nb_cells(mesh_id);
    }

    int smr_mesh_get_config(const smr_mesh* mesh, smr_mesh_config* out)
    {
        return guarded(
            [&]
            {
                require(mesh != nullptr && out != nullptr, "null mesh or output");
                *out = from_cfg(mesh->m.host.cfg);
                return 0;
            });
    }

    int smr_mesh_get_cells(const smr_mesh* mesh, int mesh_id, int64_t* out)
    {
        return guarded(
            [&]
            {
                require(mesh && out && mesh_id >= 0 && mesh_id < smr::N_MESH_IDS, "bad argument");
                const smr::MRMeshData& m = mesh->m.host;
                int64_t k                = 0;
                for (int l = 0; l < smr::MAXL; ++l)
                {
                    m.ids[mesh_id][size_t(l)].for_each_run(
                        [&](int32_t x0, int32_t x1, int32_t y, int32_t z)
                        {
                            int64_t b     = 0;
                            const bool ok = m.run_base(l, x0, y, z, b);
                            for (int32_t x = x0; x < x1; ++x)
                            {
                                int64_t* r = out + 5 * k++;
                                r[0]       = l;
                                r[1]       = x;
                                r[2]       = y;
                                r[3]       = z;
                                r[4]       = ok ? b + x : -1;
                            }
                        });
                }
                return 0;
            });
    }

    int smr_mesh_rebuild_from_device_tags(smr_mesh* mesh, const uint8_t* tags, int64_t n)
    {
        return guarded(
            [&]
            {
                require(mesh && tags, "null argument");
                smr::Mesh& mm = mesh->m;
                mm.ctx->require_device();
                require(n == mm.host.nref, "tag array size differs from the storage size");
                smr::CellArray ca;
                smr::MRMeshData nm;
                nm.nref = -1; // set when the device builds the whole mesh
                if (!smr::device_mesh_update(*mm.ctx, mm.device(), mm.host.cfg, tags, ca, &mm.host, &nm))
                {
                    return 1;
                }
                if (nm.nref < 0)
                {
                    smr::ScopedTimer t3(nullptr, "mesh construction");
                    nm           = smr::build_mesh(mm.host.cfg, ca);
                    nm.graduated = true;
                }
                mm.replace(std::move(nm), nullptr);
                return 0;
            });
    }

    int smr_mesh_rebuild_from_tags(smr_mesh* mesh, const uint8_t* tags, int64_t n)
    {
        return guarded(
            [&]
            {
                require(mesh && tags, "null argument");
                smr::MRMeshData& m = mesh->m.host;
                require(n == m.nref, "tag array size differs from the storage size");
                smr::CellArray ca;
                {
                    smr::ScopedTimer t1(nullptr, "update_cell_array_from_tag");
                    ca = smr::update_cell_array_from_tag(m, tags);
                }
                {
                    smr::ScopedTimer t2(nullptr, "make_graduation");
                    smr::make_graduation(ca, m.cfg.dim, m.cfg.grad_width);
                }
                if (smr::ca_equal(ca, m.ids[smr::CELLS]))
                {
                    return 1;
                }
                smr::MRMeshData nm;
                {
                    smr::ScopedTimer t3(nullptr, "mesh construction");
                    nm           = smr::build_mesh(m.cfg, ca);
                    nm.graduated = true;
                }
                smr::WorkList copy, coarsen, refine;
                smr::build_update_fields_lists(m, nm, copy, coarsen, refine);
                for (const smr::WorkList* w : {&copy, &coarsen, &refine})
                {
                    std::string why = smr::validate_list(*w, nm, m);
                    if (!why.empty())
                    {
                        throw std::out_of_range("update_fields: " + why);
                    }
                }
                mesh->m.replace(std::move(nm), nullptr);
                return 0;
            });
    }

    int smr_upwind_variant_count(void)
    {
        return smr::dev::upwind_variant_count();
    }

    int smr_set_upwind_variant(int id)
    {
        if (id < -1 || id >= smr::dev::upwind_variant_count())
        {
            g_error = "upwind variant out of range";
            return SMR_ERR_ARGUMENT;
        }
        // -1: the default shape
        smr::dev::set_upwind_variant(id);
        return 0;
    }

    int smr_timers_enable(int on)
    {
        smr::timers().enabled = on != 0;
        return 0;
    }

    int smr_timers_reset(void)
    {
        smr::timers().entries.clear();
        return 0;
    }

    int64_t smr_timers_report(char* buf, int64_t size)
    {
        const std::string r = smr::timers().report();
        if (buf && size > 0)
        {
            const size_t n = std::min<size_t>(r.size(), size_t(size - 1));
            std::memcpy(buf, r.data(), n);
            buf[n] = 0;
        }
        return int64_t(r.size());
    }

    void* smr_ctx_stream(smr_ctx* ctx)
    {
        return ctx ? static_cast<void*>(ctx->c.stream) : nullptr;
    }

    int64_t smr_mesh_ghost_update_bytes(const smr_mesh* mesh)
    {
        if (!mesh)
        {
            g_error = "null mesh";
            return SMR_ERR_ARGUMENT;
        }
        int64_t bytes = 0;
        const int rc  = guarded(
            [&]
            {
                const smr::MRMeshData& m = mesh->m.host;
                const int dim            = m.cfg.dim;
                const smr::GhostProgram g = smr::build_ghost_program(m);
                // projection: 2^dim fine values read, one written per ghost
                bytes += 8 * ((int64_t(1) << dim) + 1) * g.projection.ncells();
                // prediction: one write per ghost, and the distinct coarse
                // values of the (2s+1)^dim stencils
                for (int l = 1; l < smr::MAXL; ++l)
                {
                    int64_t r0       = 0;
                    const int64_t nr = g.prediction.runs_of_levels(l, l + 1, r0);
                    if (nr == 0)
                    {
                        continue;
                    }
                    smr::LevelSet fine(dim, l);
                    for (int64_t r = r0; r < r0 + nr; ++r)
                    {
                        const int32_t* v = &g.prediction.iv[size_t(5 * r)];
                        fine.push(v[2], v[3], v[0], v[1]);
                    }
                    bytes += 8 * fine.count();
                    bytes += 8 * smr::expand(smr::on(fine, l - 1), m.cfg.pred_s).count();
                }
                // outer ghosts (boundary condition, corners, BC projection and
                // prediction): each recipe reads its sources and writes one value
                // (for a field with a B.C.: the copies of B.C.-less fields are skipped)
                for (const smr::Recipes& r : g.outer)
                {
                    for (size_t q = 0; q < r.size(); ++q)
                    {
                        if (r.op[q] != smr::RECIPE_COPY_NO_BC)
                        {
                            bytes += 8 * (1 + r.src_ptr[q + 1] - r.src_ptr[q]);
                        }
                    }
                }
                return 0;
            });
        return rc < 0 ? rc : bytes;
    }

    int smr_mesh_device_view(smr_mesh* mesh, smr_mesh_view* out)
    {
        return guarded(
            [&]
            {
                require(mesh && out, "null argument");
                smr::Mesh& mm = mesh->m;
                mm.ctx->require_device();
                smr::DevMesh& dm     = mm.device();
                const smr::dev::List& c = dm.cells.d;
                smr_mesh_view v{};
                v.dim              = mm.host.cfg.dim;
                v.min_level        = mm.host.cfg.min_level;
                v.max_level        = mm.host.cfg.max_level;
                v.row_key          = dm.table.row_key;
                v.row_off          = dm.table.row_off;
                v.row_run          = dm.table.rp;
                v.run_start        = dm.table.run_s;
                v.run_end          = dm.table.run_e;
                v.run_index        = dm.table.run_idx;
                v.cell_runs        = c.iv;
                v.cell_prefix      = c.pref;
                v.cell_hint        = c.hint;
                v.cell_base        = c.base;
                v.cell_base_stride = c.ntaps;
                v.nb_cell_runs     = c.nruns;
                v.nb_cells         = c.ncells;
                *out               = v;
                return 0;
            });
    }

    int64_t smr_mesh_upwind_bytes(const smr_mesh* mesh)
    {
        if (!mesh)
        {
            g_error = "null mesh";
            return SMR_ERR_ARGUMENT;
        }
        const smr::MRMeshData& m = mesh->m.host;
        int64_t bytes            = 0;
        for (int l = 0; l < smr::MAXL; ++l)
        {
            const smr::LevelSet& c = m.ids[smr::CELLS][size_t(l)];
            if (c.empty())
            {
                continue;
            }
            smr::LevelSet read = c;
            for (const smr::Dir& d : smr::cartesian_directions(m.cfg.dim))
            {
                read = smr::set_union(read, smr::translate(c, d));
            }
            bytes += 8 * read.count() + 8 * c.count();
        }
        return bytes;
    }

    int smr_mesh_validate(const smr_mesh* mesh)
    {
        return guarded(
            [&]
            {
                require(mesh != nullptr, "null mesh");
                std::string why = smr::validate_programs(mesh->m.host);
                if (!why.empty())
                {
                    throw std::out_of_range(why);
                }
                return 0;
            });
    }

    int smr_mesh_check_device_lists(smr_mesh* mesh)
    {
        return guarded(
            [&]
            {
                require(mesh != nullptr, "null mesh");
                return smr::check_device_lists(mesh->m);
            });
    }

    int smr_field_create(smr_mesh* mesh, const char* name, smr_field** out)
    {
        return guarded(
            [&]
            {
                require(mesh && out, "null argument");
                auto* f   = new smr_field();
                f->f.mesh = &mesh->m;
                f->f.name = name ? name : "";
                try
                {
                    f->f.allocate(true);
                }
                catch (...)
                {
                    delete f;
                    throw;
                }
                *out = f;
                return 0;
            });
    }

    int smr_field_destroy(smr_field* field)
    {
        return guarded(
            [&]
            {
                delete field;
                return 0;
            });
    }

    int smr_field_resize(smr_field* field)
    {
        return guarded(
            [&]
            {
                require(field != nullptr, "null field");
                if (field->f.version != field->f.mesh->version || field->f.size != field->f.mesh->host.nref)
                {
                    field->f.allocate(true);
                }
                return 0;
            });
    }

    int smr_field_set_cells(smr_field* field, const double* values)
    {
        return guarded(
            [&]
            {
                require(field && values, "null argument");
                smr::Field& f = field->f;
                f.check_current();
                auto idx = cell_storage_index(f.mesh->host);
                std::vector<double> st(size_t(f.size));
                hipStream_t s = f.mesh->ctx->stream;
                smr::hip_check(hipMemcpyAsync(st.data(), f.d, st.size() * 8, hipMemcpyDeviceToHost, s), "hipMemcpyAsync");
                f.mesh->ctx->sync();
                for (size_t k = 0; k < idx.size(); ++k)
                {
                    st[size_t(idx[k])] = values[k];
                }
                smr::hip_check(hipMemcpyAsync(f.d, st.data(), st.size() * 8, hipMemcpyHostToDevice, s), "hipMemcpyAsync");
                f.mesh->ctx->sync();
                return 0;
            });
    }

    const char* smr_field_name(const smr_field* field)
    {
        if (!field)
        {
            g_error = "null field";
            return nullptr;
        }
        return field->f.name.c_str();
    }

    int smr_field_on_mesh(const smr_field* field, const smr_mesh* mesh)
    {
        if (!field || !mesh)
        {
            g_error = "null field or mesh";
            return SMR_ERR_ARGUMENT;
        }
        return field->f.mesh == &mesh->m ? 1 : 0;
    }

    int smr_field_get_cells(smr_field* field, double* values)
    {
        return guarded(
            [&]
            {
                require(field && values, "null argument");
                smr::Field& f = field->f;
                f.check_current();
                auto idx = cell_storage_index(f.mesh->host);
                std::vector<double> st(size_t(f.size));
                smr::hip_check(hipMemcpyAsync(st.data(), f.d, st.size() * 8, hipMemcpyDeviceToHost, f.mesh->ctx->stream), "hipMemcpyAsync");
                f.mesh->ctx->sync();
                for (size_t k = 0; k < idx.size(); ++k)
                {
                    values[k] = st[size_t(idx[k])];
                }
                return 0;
            });
    }

    int smr_field_get_storage(smr_field* field, double* values)
    {
        return guarded(
            [&]
            {
                require(field && values, "null argument");
                smr::Field& f = field->f;
                f.check_current();
                smr::hip_check(hipMemcpyAsync(values, f.d, size_t(f.size) * 8, hipMemcpyDeviceToHost, f.mesh->ctx->stream), "hipMemcpyAsync");
                f.mesh->ctx->sync();
                return 0;
            });
    }

    int smr_field_set_storage(smr_field* field, const double* values)
    {
        return guarded(
            [&]
            {
                require(field && values, "null argument");
                smr::Field& f = field->f;
                f.check_current();
                smr::hip_check(hipMemcpyAsync(f.d, values, size_t(f.size) * 8, hipMemcpyHostToDevice, f.mesh->ctx->stream), "hipMemcpyAsync");
                f.mesh->ctx->sync();
                return 0;
            });
    }

    int64_t smr_field_size(const smr_field* field)
    {
        return field ? field->f.size : -1;
    }

    double* smr_field_device_data(smr_field* field)
    {
        return field ? field->f.d : nullptr;
    }

    int smr_field_attach_dirichlet(smr_field* field, double value)
    {
        return guarded(
            [&]
            {
                require(field != nullptr, "null field");
                field->f.has_bc   = true;
                field->f.bc_value = value;
                return 0;
            });
    }

    int smr_field_swap(smr_field* a, smr_field* b)
    {
        return guarded(
            [&]
            {
                require(a && b, "null field");
                require(a->f.mesh == b->f.mesh, "swap of fields on different meshes");
                std::swap(a->f.d, b->f.d);
                std::swap(a->f.size, b->f.size);
                std::swap(a->f.version, b->f.version);
                return 0;
            });
    }

    int smr_update_ghost_mr(smr_field* field)
    {
        return guarded(
            [&]
            {
                require(field != nullptr, "null field");
                smr::update_ghost_mr(field->f);
                return 0;
            });
    }

    int smr_upwind(const double* a, double dt, smr_field* u, smr_field* unp1)
    {
        return guarded(
            [&]
            {
                require(a && u && unp1, "null argument");
                smr::upwind(a, dt, u->f, unp1->f);
                return 0;
            });
    }

    int smr_adapt_create(smr_field* field, smr_adapt** out)
    {
        return guarded(
            [&]
            {
                require(field && out, "null argument");
                auto* a    = new smr_adapt();
                a->a.field = &field->f;
                a->a.fields.push_back(&field->f);
                *out       = a;
                return 0;
            });
    }

    int smr_adapt_create_fields(smr_field* const* fields, int n, smr_adapt** out)
    {
        return guarded(
            [&]
            {
                require(fields && out && n >= 1, "null argument or no field");
                for (int k = 0; k < n; ++k)
                {
                    require(fields[k] != nullptr, "null field");
                    require(fields[k]->f.mesh == fields[0]->f.mesh, "make_MRAdapt: the fields must live on one mesh");
                    for (int q = 0; q < k; ++q)
                    {
                        require(fields[q] != fields[k], "make_MRAdapt: a field is given twice");
                    }
                }
                auto* a    = new smr_adapt();
                a->a.field = &fields[0]->f;
                for (int k = 0; k < n; ++k)
                {
                    a->a.fields.push_back(&fields[k]->f);
                }
                *out = a;
                return 0;
            });
    }

    int smr_adapt_nb_fields(const smr_adapt* adapt)
    {
        return adapt ? adapt->a.ncomp() : SMR_ERR_ARGUMENT;
    }

    int smr_adapt_destroy(smr_adapt* adapt)
    {
        return guarded(
            [&]
            {
                delete adapt;
                return 0;
            });
    }

    int smr_adapt_run(smr_adapt* adapt, double eps, double regularity)
    {
        return guarded(
            [&]
            {
                require(adapt != nullptr, "null adapt");
                return smr::adapt(adapt->a, eps, regularity);
            });
    }

    int smr_adapt_set_refine_boundary(smr_adapt* adapt, int on)
    {
        return guarded(
            [&]
            {
                require(adapt != nullptr, "null adapt");
                adapt->a.refine_boundary = on != 0;
                return 0;
            });
    }

    int smr_adapt_set_relative_detail(smr_adapt* adapt, int on)
    {
        return guarded(
            [&]
            {
                require(adapt != nullptr, "null adapt");
                adapt->a.relative_detail = on != 0;
                return 0;
            });
    }

    int smr_adapt_harten(smr_adapt* adapt, int ite, double eps, double regularity)
    {
        return guarded(
            [&]
            {
                require(adapt != nullptr, "null adapt");
                return smr::harten(adapt->a, ite, eps, regularity) ? 1 : 0;
            });
    }

    int smr_adapt_tags(smr_adapt* adapt, int ite, double eps, double regularity, uint8_t* tags)
    {
        return guarded(
            [&]
            {
                require(adapt != nullptr, "null adapt");
                require(tags != nullptr, "null tag array");
                require(ite >= 0, "negative harten iteration");
                adapt->a.field->mesh->ctx->require_device();
                smr::adapt_tags(adapt->a, ite, eps, regularity, tags);
                return 0;
            });
    }

    int smr_adapt_tags_noalloc(smr_adapt* adapt, int ite, double eps, double regularity, uint8_t* tags)
    {
        return guarded(
            [&]
            {
                require(adapt != nullptr, "null adapt");
                require(tags != nullptr, "null tag array");
                require(ite >= 0, "negative harten iteration");
                adapt->a.field->mesh->ctx->require_device();
                smr::adapt_tags_noalloc(adapt->a, ite, eps, regularity, tags);
                return 0;
            });
    }

    int64_t smr_adapt_last_size(const smr_adapt* adapt)
    {
        return adapt ? adapt->a.last_size : -1;
    }

    int smr_adapt_last_tags(const smr_adapt* adapt, uint8_t* tags, double* details)
    {
        return guarded(
            [&]
            {
                require(adapt != nullptr, "null adapt");
                adapt->a.copy_last(tags, details, adapt->a.field->mesh->ctx->stream);
                return 0;
            });
    }

    int smr_batch_create(smr_ctx* ctx, int n, smr_field* const* u, smr_field* const* unp1, const double* a, const double* dt,
                         smr_batch** out)
    {
        return guarded(
            [&]
            {
                require(ctx && out && n >= 0 && (n == 0 || (u && unp1 && a && dt)), "null argument");
                ctx->c.require_device();
                std::vector<smr::BatchInstance> inst{size_t(n)};
                for (int i = 0; i < n; ++i)
                {
                    require(u[i] && unp1[i], "null field");
                    smr::BatchInstance& b = inst[size_t(i)];
                    b.u                   = &u[i]->f;
                    b.unp1                = &unp1[i]->f;
                    b.a[0]                = a[3 * i];
                    b.a[1]                = a[3 * i + 1];
                    b.a[2]                = a[3 * i + 2];
                    b.dt                  = dt[i];
                }
                auto* r = new smr_batch();
                try
                {
                    r->b = std::make_unique<smr::Batch>(&ctx->c, std::move(inst));
                }
                catch (...)
                {
                    delete r;
                    throw;
                }
                *out = r;
                return 0;
            });
    }

    int smr_batch_destroy(smr_batch* batch)
    {
        return guarded(
            [&]
            {
                delete batch;
                return 0;
            });
    }

    int smr_batch_adapt(smr_batch* batch, double eps, double regularity)
    {
        return guarded(
            [&]
            {
                require(batch != nullptr, "null batch");
                return batch->b->adapt(eps, regularity);
            });
    }

    int smr_batch_advance(smr_batch* batch)
    {
        return guarded(
            [&]
            {
                require(batch != nullptr, "null batch");
                batch->b->advance();
                return 0;
            });
    }

    int smr_batch_step(smr_batch* batch, double eps, double regularity)
    {
        return guarded(
            [&]
            {
                require(batch != nullptr, "null batch");
                return batch->b->step(eps, regularity);
            });
    }

    int64_t smr_batch_nb_cells(const smr_batch* batch)
    {
        if (!batch)
        {
            g_error = "null batch";
            return -1;
        }
        return batch->b->nb_cells();
    }

    int smr_batch_set_tuning(smr_batch* batch, int block, int cells_per_thread, int mode)
    {
        return guarded(
            [&]
            {
                require(batch != nullptr, "null batch");
                require(block >= 64 && block <= 1024 && block % 64 == 0, "block size must be a multiple of 64 in [64, 1024]");
                require(cells_per_thread >= 0 && cells_per_thread <= 64, "cells per thread must be in [0, 64] (0: wave items)");
                require(mode >= 0 && mode <= 3, "unknown program mode");
                batch->b->set_tuning(smr::Tuning{block, cells_per_thread, mode});
                return 0;
            });
    }

    int smr_batch_get_tuning(const smr_batch* batch, int* block, int* cells_per_thread, int* mode)
    {
        return guarded(
            [&]
            {
                require(batch && block && cells_per_thread && mode, "null argument");
                const smr::Tuning& t = batch->b->tuning();
                *block               = t.block;
                *cells_per_thread    = t.cpt;
                *mode                = t.mode;
                return 0;
            });
    }

    int smr_batch_autotune(smr_batch* batch, int reps, const char* cache_file)
    {
        return guarded(
            [&]
            {
                require(batch != nullptr && reps > 0, "bad argument");
                batch->b->autotune(reps, cache_file ? std::string(cache_file) : std::string());
                return 0;
            });
    }

    int smr_batch_load_tuning(smr_batch* batch, const char* cache_file)
    {
        return guarded(
            [&]
            {
                require(batch != nullptr, "null batch");
                smr::Tuning t;
                if (!cache_file || !smr::Batch::load_tuning(cache_file, t))
                {
                    return 0;
                }
                batch->b->set_tuning(t);
                return 1;
            });
    }

    int smr_adjoint_create(smr_ctx* ctx, int n, smr_mesh* const* meshes, const double* dt, const double* bc_value, int steps,
                           int checkpoint_every, smr_adjoint** out)
    {
        return guarded(
            [&]
            {
                require(ctx && out && n >= 0 && (n == 0 || (meshes && dt)), "null argument");
                ctx->c.require_device();
                std::vector<smr::Mesh*> ms(static_cast<size_t>(n));
                for (int i = 0; i < n; ++i)
                {
                    require(meshes[i] != nullptr, "null mesh");
                    ms[size_t(i)] = &meshes[i]->m;
                }
                auto* r = new smr_adjoint();
                try
                {
                    r->a = std::make_unique<smr::Adjoint>(&ctx->c, std::move(ms), dt, bc_value, steps, checkpoint_every);
                }
                catch (...)
                {
                    delete r;
                    throw;
                }
                *out = r;
                return 0;
            });
    }

    int smr_adjoint_destroy(smr_adjoint* adj)
    {
        return guarded(
            [&]
            {
                delete adj;
                return 0;
            });
    }

    int64_t smr_adjoint_nb_cells(const smr_adjoint* adj, int instance)
    {
        if (!adj || instance >= adj->a->ninstances())
        {
            g_error = "bad adjoint or instance";
            return -1;
        }
        return adj->a->nb_cells(instance);
    }

    int smr_adjoint_set_velocity(smr_adjoint* adj, const double* a)
    {
        return guarded(
            [&]
            {
                require(adj && (a || adj->a->ninstances() == 0), "null argument");
                adj->a->set_velocity(a);
                return 0;
            });
    }

    int smr_adjoint_forward(smr_adjoint* adj, const double* u0, double* uT)
    {
        return guarded(
            [&]
            {
                require(adj && u0 && uT, "null argument");
                adj->a->forward(u0, uT);
                return 0;
            });
    }

    int smr_adjoint_backward(smr_adjoint* adj, const double* grad_uT, double* grad_u0, double* grad_a)
    {
        return guarded(
            [&]
            {
                require(adj && grad_uT && grad_u0 && grad_a, "null argument");
                adj->a->backward(grad_uT, grad_u0, grad_a);
                return 0;
            });
    }

    int smr_adjoint_set_tuning(smr_adjoint* adj, int block, int cells_per_thread, int mode)
    {
        return guarded(
            [&]
            {
                require(adj != nullptr, "null adjoint");
                require(block >= 64 && block <= 1024 && block % 64 == 0, "block size must be a multiple of 64 in [64, 1024]");
                require(cells_per_thread >= 0 && cells_per_thread <= 64, "cells per thread must be in [0, 64] (0: wave items)");
                require(mode >= 0 && mode <= 3, "unknown program mode");
                adj->a->set_tuning(smr::Tuning{block, cells_per_thread, mode});
                return 0;
            });
    }

    int smr_convection_upwind_nonlinear(smr_field* u, smr_field* out)
    {
        return guarded(
            [&]
            {
                require(u && out, "null argument");
                smr::convection_upwind_nonlinear(u->f, out->f);
                return 0;
            });
    }

    int smr_convection_upwind_field(smr_field* const* velocity, smr_field* u, smr_field* out)
    {
        return guarded(
            [&]
            {
                require(velocity && u && out, "null argument");
                const int dim = u->f.mesh->host.cfg.dim;
                smr::Field* v[3] = {nullptr, nullptr, nullptr};
                for (int d = 0; d < dim; ++d)
                {
                    require(velocity[d] != nullptr, "null velocity component");
                    v[d] = &velocity[d]->f;
                }
                smr::convection_upwind_heterogeneous(v, u->f, out->f);
                return 0;
            });
    }

    int smr_convection_weno5(const double* velocity, smr_field* u, smr_field* out)
    {
        return guarded(
            [&]
            {
                require(velocity && u && out, "null argument");
                smr::convection_weno5(velocity, u->f, out->f);
                return 0;
            });
    }

    int smr_convection_upwind(const double* velocity, smr_field* u, smr_field* out)
    {
        return guarded(
            [&]
            {
                require(velocity && u && out, "null argument");
                smr::convection_upwind(velocity, u->f, out->f);
                return 0;
            });
    }

    int smr_mesh_convection_terms(const smr_mesh* mesh, const double* velocity, int64_t* nrows, int64_t* nterms, int64_t* row, int64_t* ptr,
                                  int64_t* src, double* coef)
    {
        return guarded(
            [&]
            {
                require(mesh && nrows && nterms, "null argument");
                const smr::ConvTerms t =
                    velocity ? smr::build_convection_terms(mesh->m.host, velocity) : smr::build_burgers_terms(mesh->m.host);
                const bool fill        = row && ptr && src && coef;
                if (fill)
                {
                    require(*nrows >= int64_t(t.row.size()) && *nterms >= t.nterms(), "output arrays too small");
                    std::copy(t.row.begin(), t.row.end(), row);
                    std::copy(t.ptr.begin(), t.ptr.end(), ptr);
                    std::copy(t.src.begin(), t.src.end(), src);
                    std::copy(t.coef.begin(), t.coef.end(), coef);
                }
                *nrows  = int64_t(t.row.size());
                *nterms = t.nterms();
                return 0;
            });
    }

    int smr_upwind_scalar_burgers(const double* k, double dt, smr_field* u, smr_field* unp1)
    {
        return guarded(
            [&]
            {
                require(k && u && unp1, "null argument");
                smr::upwind_scalar_burgers(k, dt, u->f, unp1->f);
                return 0;
            });
    }

    int smr_upwind_hand_loop(double dt, smr_field* u, smr_field* unp1)
    {
        return guarded(
            [&]
            {
                require(u && unp1, "null field");
                smr::upwind_hand_loop(dt, u->f, unp1->f);
                return 0;
            });
    }

    int smr_field_sub_scaled(smr_field* u, double a, smr_field* v, smr_field* out)
    {
        return guarded(
            [&]
            {
                require(u && v && out, "null field");
                smr::sub_scaled(u->f, a, v->f, out->f);
                return 0;
            });
    }
}
