/* C ABI of the MI355X implementation of samurai's adaptive-MR time step.
 *
 * samurai (hpc-maths/samurai v0.33.0) is a header-only C++ library with no FFI:
 * each entry point below replaces one C++ call of the reference's public API
 * for this path (cited as file:line relative to the samurai source tree).  The
 * C++ drop-in headers in include/samurai_amd/ wrap these functions under the
 * reference's names; the Python tests bind them with ctypes.
 *
 * Conventions: every function returns 0 (or a documented non-negative value)
 * on success and a negative status on failure, with the message available
 * from smr_last_error() on the calling thread.  No C++ exception crosses the
 * ABI.  All device work of a context is ordered on the context's HIP stream; a
 * context, and every object created from it, is used by one thread at a time.
 * Fields are FP64, stored on the device in the reference's storage order
 * (levels ascending, then rows (z, y), then x; mesh.hpp:893-911).
 */
#ifndef SAMURAI_AMD_H
#define SAMURAI_AMD_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct smr_ctx smr_ctx;
typedef struct smr_mesh smr_mesh;
typedef struct smr_field smr_field;
typedef struct smr_adapt smr_adapt;
typedef struct smr_batch smr_batch;
typedef struct smr_adjoint smr_adjoint;

/* MRMeshId (mr/mesh.hpp:25-34) */
enum smr_mesh_id
{
    SMR_CELLS            = 0,
    SMR_CELLS_AND_GHOSTS = 1,
    SMR_PROJ_CELLS       = 2,
    SMR_UNION_CELLS      = 3,
    SMR_REFERENCE        = 4
};

enum smr_status
{
    SMR_OK            = 0,
    SMR_ERR_ARGUMENT  = -1, /* std::invalid_argument in the reference */
    SMR_ERR_RANGE     = -2, /* std::out_of_range (cell absent from the mesh) */
    SMR_ERR_RUNTIME   = -3, /* HIP or other runtime failure */
    SMR_ERR_NO_DEVICE = -4  /* a device call on a host-only context */
};

/* samurai::mesh_config<dim, prediction_stencil_radius> (mesh_config.hpp:16-432),
 * with max_stencil_size(2).disable_minimal_ghost_width() as in
 * demos/FiniteVolume/advection_2d.cpp:101: max_stencil_radius 1, or 2 or 3
 * (max_stencil_size 4 or 6: ghost layers of that width, the far ones filled
 * by polynomial extrapolation) on meshes with min_level == max_level. */
typedef struct smr_mesh_config
{
    int dim;                       /* 2 or 3 */
    int min_level;
    int max_level;
    int prediction_stencil_radius; /* 0 or 1 */
    int max_stencil_radius;        /* 1; 2 or 3 when min_level == max_level */
    int graduation_width;          /* default 1 */
} smr_mesh_config;

/* samurai::initialize (samurai.hpp:47-113): a context on HIP device `device`,
 * or a host-only context (mesh construction only) when device < 0. */
int smr_ctx_create(int device, smr_ctx** out);
/* samurai::finalize (samurai.hpp:100-113).  Every object made from the
 * context (batches, adjoint problems, adaptations, fields, then meshes) must
 * be destroyed before it: their destructors use its stream.  Call it before
 * the process exits, not from a static destructor or an atexit handler that
 * runs after the HIP runtime has begun to shut down. */
int smr_ctx_destroy(smr_ctx* ctx);
int smr_ctx_synchronize(smr_ctx* ctx);
const char* smr_last_error(void);

/* samurai::mra::make_mesh(Box unit square/cube, cfg) (mr/mesh.hpp:481-529) */
int smr_mesh_create_box(smr_ctx* ctx, const smr_mesh_config* cfg, smr_mesh** out);
/* MRMesh(ca, cfg) (mesh.hpp:386-396): leaves given as n rows (level, x, y, z) */
int smr_mesh_create_from_cells(smr_ctx* ctx, const smr_mesh_config* cfg, const int32_t* cells, int64_t n, smr_mesh** out);
/* mesh = MRMesh(ca, cfg) in place, cfg being the mesh's own with the given
 * min and max levels: what samurai::load does to the mesh it is given
 * (io/restart.hpp:449-455, std::swap(mesh, new_mesh)).  Leaves as rows
 * (level, x, y, z); fields on the mesh must be resized afterwards. */
int smr_mesh_assign_cells(smr_mesh* mesh, int min_level, int max_level, const int32_t* cells, int64_t n);
/* std::swap(mesh_a, mesh_b) of two meshes of one context: the contents are
 * exchanged while fields stay attached to their mesh object, as after the
 * reference's `mesh = mra::make_mesh(...)` or load (io/restart.hpp:454); fields
 * on either mesh must be resized before their next use. */
int smr_mesh_swap(smr_mesh* a, smr_mesh* b);
int smr_mesh_destroy(smr_mesh* mesh);
/* mesh.nb_cells(mesh_id) (mesh.hpp:540-551); SMR_REFERENCE gives the storage size */
int64_t smr_mesh_nb_cells(const smr_mesh* mesh, int mesh_id);
/* mesh.cfg() (mesh.hpp:162-163, 642-652): the configuration the mesh was built with */
int smr_mesh_get_config(const smr_mesh* mesh, smr_mesh_config* out);
/* for_each_cell(mesh[mesh_id]) (algorithm.hpp:319-323): n rows of
 * (level, x, y, z, storage index) in traversal order */
int smr_mesh_get_cells(const smr_mesh* mesh, int mesh_id, int64_t* out);
/* the host half of Adapt::harten's mesh update (mr/adapt.hpp:359-386):
 * update_cell_array_from_tag (algorithm/graduation.hpp:743-843) and
 * make_graduation (:573-726) from tags indexed by the reference storage, then
 * MRMesh(new_ca, mesh) (mesh.hpp:398-408).  Returns 1 at the fixed point (mesh
 * unchanged), 0 if the mesh was rebuilt.  Host only: fields on the mesh must
 * be resized afterwards, their values are not transferred. */
int smr_mesh_rebuild_from_tags(smr_mesh* mesh, const uint8_t* tags, int64_t n);
/* smr_mesh_rebuild_from_tags with update_cell_array_from_tag, the
 * make_graduation fixed point and the new mesh's ghost layers and union
 * (MRMesh::update_sub_mesh_impl, mr/mesh.hpp:221-455) run on the device
 * (per-level bitmaps), from a device array of tags; the host receives the
 * new mesh ids as runs only when the leaves change, and renumbers them.  The
 * path every harten call of smr_adapt_run takes.  Returns 1 at the fixed
 * point, 0 if the mesh was rebuilt. */
int smr_mesh_rebuild_from_device_tags(smr_mesh* mesh, const uint8_t* tags, int64_t n);

/* The device image of a mesh version, for user kernels that walk the leaves
 * and look cells up (for_each_interval(mesh, f), algorithm.hpp:75-132, with a
 * device f; include/samurai_amd/device.hpp).  Pointers are device addresses,
 * valid until the mesh changes (adaptation, assignment, swap, load).
 *  - reference interval table (LevelCellArray of mesh[reference] per level,
 *    level_cell_array.hpp:68-256): rows of level l are
 *    [row_off[l], row_off[l+1]) with keys ((int64)z << 32) + (y + 2^31),
 *    ascending; row r owns the x-runs [row_run[r], row_run[r+1]) with
 *    [run_start, run_end) and storage index run_index + x;
 *  - the leaves (mesh[cells]) as x-runs (x0, x1, y, z, level), 5 int32 each,
 *    in traversal order; cell_prefix[r] is the first cell of run r
 *    (cell_prefix[nb_cell_runs] = nb_cells); cell_hint[j] is the run holding
 *    cell 256 j; cell_base[r * cell_base_stride] + x is the storage index of
 *    cell x of run r. */
typedef struct smr_mesh_view
{
    int dim;
    int min_level;
    int max_level;
    const int64_t* row_key;
    const int64_t* row_off; /* 22 entries */
    const int64_t* row_run;
    const int32_t* run_start;
    const int32_t* run_end;
    const int64_t* run_index;
    const int32_t* cell_runs;
    const int64_t* cell_prefix;
    const int32_t* cell_hint;
    const int64_t* cell_base;
    int64_t cell_base_stride;
    int64_t nb_cell_runs;
    int64_t nb_cells;
} smr_mesh_view;
int smr_mesh_device_view(smr_mesh* mesh, smr_mesh_view* out);

/* make_scalar_field<double>(name, mesh) (field/scalar_field.hpp:171-176) */
int smr_field_create(smr_mesh* mesh, const char* name, smr_field** out);
int smr_field_destroy(smr_field* field);
/* field.resize() after a mesh change (field/access_base.hpp:105-119) */
int smr_field_resize(smr_field* field);
/* field.name() (field/field_base.hpp:74-76), or NULL for a null field */
const char* smr_field_name(const smr_field* field);
/* 1 if field lives on mesh, 0 if not (&field.mesh() == &mesh) */
int smr_field_on_mesh(const smr_field* field, const smr_mesh* mesh);
/* values on mesh[cells] in traversal order (like for_each_cell + u[cell]) */
int smr_field_set_cells(smr_field* field, const double* values);
int smr_field_get_cells(smr_field* field, double* values);
/* the whole storage (u.array()) */
int smr_field_get_storage(smr_field* field, double* values);
int smr_field_set_storage(smr_field* field, const double* values);
int64_t smr_field_size(const smr_field* field);
/* device pointer of the storage (for interop with torch or other HIP code) */
double* smr_field_device_data(smr_field* field);
/* make_bc<Dirichlet<1>>(u, value) (bc/bc.hpp:771-795, bc/dirichlet.hpp:29) */
int smr_field_attach_dirichlet(smr_field* field, double value);
/* samurai::swap(u, v) / std::swap(u.array(), v.array()) (field/swap.hpp:12-19) */
int smr_field_swap(smr_field* a, smr_field* b);

/* update_ghost_mr(u) (algorithm/update_ghost_mr.hpp:260-270) */
int smr_update_ghost_mr(smr_field* field);
/* unp1 = u - dt * upwind(a, u) over mesh[cells] (stencil_field.hpp:175-179,
 * field/field_base.hpp:230-242); a has dim components */
int smr_upwind(const double* a, double dt, smr_field* u, smr_field* unp1);

/* make_MRAdapt(u) (mr/adapt.hpp:391-397) */
int smr_adapt_create(smr_field* field, smr_adapt** out);
/* make_MRAdapt(u, v, ...) (mr/adapt.hpp:138-146, 391-397): n fields on one
 * mesh adapted together -- the details of every field join the criteria
 * (mr/criteria.hpp:47-105: a coarse detail of any field above the coarse
 * threshold blocks the coarsening, which needs every field's children below
 * eps; a child is refined when any field's detail is above the fine
 * threshold) and every field is transferred to each new mesh
 * (update_fields, algorithm/update_fields.hpp:27-127). */
int smr_adapt_create_fields(smr_field* const* fields, int n, smr_adapt** out);
/* the number of fields an adaptation adapts */
int smr_adapt_nb_fields(const smr_adapt* adapt);
int smr_adapt_destroy(smr_adapt* adapt);
/* MRadaptation(mra_config().epsilon(eps).regularity(reg)) (mr/adapt.hpp:147-195):
 * returns the number of harten iterations run (>= 0) */
int smr_adapt_run(smr_adapt* adapt, double eps, double regularity);
/* mra_config().relative_detail(on) (mr/config.hpp:36-45, applied at
 * mr/adapt.hpp:319-322 by compute_relative_detail, mr/rel_detail.hpp:74-112):
 * every detail divided by max |field| over the leaves (1 below the machine
 * epsilon) before the criteria, for the following smr_adapt_run,
 * smr_adapt_harten and smr_adapt_tags calls; off by default.
 * smr_adapt_tags_noalloc rejects it (SMR_ERR_ARGUMENT). */
int smr_adapt_set_relative_detail(smr_adapt* adapt, int on);
/* The --refine-boundary option (arguments.hpp; keep_boundary_refined,
 * mr/adapt.hpp:245-272, 341-346): after the criteria, the leaves of
 * max_level within the stencil radius of a face are tagged keep, so the
 * boundary stays at the finest level; for the following harten calls of this
 * adaptation (every entry point); off by default. */
int smr_adapt_set_refine_boundary(smr_adapt* adapt, int on);
/* one Adapt::harten(ite) call (mr/adapt.hpp:276-389): 1 at the fixed point, 0 if
 * the mesh changed */
int smr_adapt_harten(smr_adapt* adapt, int ite, double eps, double regularity);
/* The device half of one Adapt::harten(ite) call (mr/adapt.hpp:276-358:
 * tag init to keep, update_ghost_mr, compute_detail, mr_criteria_op and the
 * maximum_op tag finalization), without the host mesh update: tags
 * (CellFlag bits) into `tags`, a device array of
 * smr_mesh_nb_cells(mesh, SMR_REFERENCE) bytes indexed by the field's
 * storage, ordered on the context's stream (nothing is copied to the host).
 * The field's ghosts are updated; the details go to the adaptation's own
 * buffer.  The mesh is not changed. */
int smr_adapt_tags(smr_adapt* adapt, int ite, double eps, double regularity, uint8_t* tags);
/* No counterpart in the reference.  smr_adapt_tags for callers that cannot
 * allocate (code under HIP graph capture, tight serving loops): it calls no
 * hipMalloc, hipFree or allocating routine and clears with kernels, not
 * hipMemsetAsync.  The details go to a static device array of 2^21 FP64
 * slots per GPU (16 MiB, held until the process ends), so the mesh's
 * reference storage times the number of adapted fields must not exceed 2^21
 * slots (SMR_ERR_ARGUMENT otherwise), and the mesh version's device programs must already exist:
 * run smr_adapt_tags or smr_adapt_run once on it first (SMR_ERR_ARGUMENT
 * otherwise).  Every call on one GPU shares that array: calls that overlap
 * on one GPU (several contexts or threads) race. */
int smr_adapt_tags_noalloc(smr_adapt* adapt, int ite, double eps, double regularity, uint8_t* tags);
/* tags (CellFlag bits, cell_flag.hpp:11-17) and details of the last harten
 * call, indexed by the storage of the mesh they were computed on; details
 * holds smr_adapt_nb_fields arrays of smr_adapt_last_size values, one per
 * field in the order given */
int64_t smr_adapt_last_size(const smr_adapt* adapt);
int smr_adapt_last_tags(const smr_adapt* adapt, uint8_t* tags, double* details);

/* Diagnostics (no counterpart in the reference): checks on the host that
 * every work-list row the mesh's kernels will touch exists and that each
 * run's access window lies inside one reference interval -- the same check
 * the device performs before any operator kernel runs.  0 if valid,
 * SMR_ERR_RANGE otherwise. */
int smr_mesh_validate(const smr_mesh* mesh);
/* Diagnostics (no counterpart in the reference): after a mesh change made by
 * the device, compares the work lists the device built for the mesh's image
 * (projection, prediction, detail, criteria, maximum: runs, cell prefixes,
 * chunk hints, per-level offsets) with those the host builds from the same
 * mesh.  0 if identical, 1 if the image has no device-built lists,
 * SMR_ERR_RUNTIME if they differ. */
int smr_mesh_check_device_lists(smr_mesh* mesh);
/* The HIP stream (hipStream_t) the context orders its work on, for timing
 * with HIP events; NULL for a host-only context. */
void* smr_ctx_stream(smr_ctx* ctx);
/* Algorithmic HBM bytes of one upwind launch on mesh[cells]: 8 bytes per
 * distinct cell read (the leaves and their face neighbours) plus 8 bytes per
 * leaf written. */
int64_t smr_mesh_upwind_bytes(const smr_mesh* mesh);
/* Algorithmic HBM bytes of one update_ghost_mr on the mesh: per projection
 * ghost 8 * (2^dim + 1) (children read, ghost written), per prediction ghost
 * 8 written plus 8 per distinct coarse stencil value read, per outer-ghost
 * recipe 8 * (1 + its sources). */
int64_t smr_mesh_ghost_update_bytes(const smr_mesh* mesh);
/* Phase timers under the reference's ScopedTimer names (timers.hpp:979-1001;
 * "mesh adaptation", "ghost update", "detail computation", "tag cells",
 * "tag finalization", "mesh update", "fields update", "field expressions"),
 * process-wide.  While enabled, every phase synchronizes its stream at entry
 * and exit, so use it for breakdowns, not for throughput.  The report is a JSON
 * object {name: {"seconds": s, "calls": n}} written into buf (truncated to
 * size); returns its full length. */
/* Tuning knob (no counterpart in the reference): selects the launch shape of
 * the upwind kernel, 0 <= id < smr_upwind_variant_count(), or -1 for the
 * default shape.  Every variant computes identical results. */
int smr_upwind_variant_count(void);
int smr_set_upwind_variant(int id);
int smr_timers_enable(int on);
int smr_timers_reset(void);
int64_t smr_timers_report(char* buf, int64_t size);

/* out = u - a * v on the cells: the field expression unp1 = u - dt * f of
 * the flux-based demos (field_base.hpp:230-242 assignment of an expression;
 * per cell u - (a * v)).  All three fields on one mesh; out may alias u. */
int smr_field_sub_scaled(smr_field* u, double a, smr_field* v, smr_field* out);
/* The README's hand-written upwind loop (README.md:135-154): for every cell,
 *   unp1(level, i, j) = u(level, i, j) - dt / dx * (u(level, i, j) - u(level, i - 1, j)
 *                                                + u(level, i, j) - u(level, i, j - 1))
 * with dx = cell_length(level), as for_each_interval(mesh, ...) runs it (2D,
 * velocity (1, 1)).  u's ghosts must be up to date (update_ghost_mr). */
/* unp1 = u - dt * upwind_scalar_burgers(k, u) (stencil_field.hpp:181-255,
 * demos/FiniteVolume/scalar_burgers_2d.cpp:155-160), two-dimensional: the
 * finite-volume combination of upwind (-F_left + F_right - F_down + F_up) / dx
 * with the flux 0.5 * min(|ul|,|ur|)^2 where (k ul < k ur) and ul ur > 0,
 * 0.5 * max(|ul|,|ur|)^2 where !(k ul < k ur), else 0.  k: two values.  As
 * for smr_upwind, u's ghosts must have been updated (update_ghost_mr). */
int smr_upwind_scalar_burgers(const double* k, double dt, smr_field* u, smr_field* unp1);
int smr_upwind_hand_loop(double dt, smr_field* u, smr_field* unp1);
/* out = make_convection_upwind<Field>(velocity)(u): the flux-based linear
 * upwind convection operator (schemes/fv/operators/convection_lin.hpp:15-89)
 * applied explicitly (FV_scheme.hpp:202-210): out is zero-filled and
 * receives the interface fluxes divided by the cell measure.  u's ghosts are
 * updated first (update_ghost_mr, as update_ghosts_if_needed does);
 * velocity: dim values; u and out: two fields on one mesh. */
int smr_convection_upwind(const double* velocity, smr_field* u, smr_field* out);
/* make_convection_upwind<Field>()(u) for a scalar field
 * (schemes/fv/operators/convection_nonlin.hpp:24-79): the non-linear upwind
 * of f(u) = u * u, the flux at an interface being f of the side picked by the
 * sign of the mean of the two values, applied explicitly
 * (explicit_flux_based_scheme__nonlin.hpp:35-86) over the same interfaces as
 * smr_convection_upwind; u's ghosts are updated first. */
int smr_convection_upwind_nonlinear(smr_field* u, smr_field* out);
/* make_convection_upwind<Field>(velocity_field)(u) for a scalar field
 * (schemes/fv/operators/convection_lin.hpp:184-264): the velocity field is
 * given as dim scalar fields on u's mesh (component d in velocity[d], each
 * with its own boundary condition); at an interface the velocity is the mean
 * of the two stencil cells' values, and the upwind side's value is carried
 * (explicit_flux_based_scheme__lin_het.hpp:35-110).  The ghosts of the
 * velocity components, then of u, are updated first (FV_scheme.hpp:176-196). */
int smr_convection_upwind_field(smr_field* const* velocity, smr_field* u, smr_field* out);
/* out = make_convection_weno5<Field>(velocity)(u)
 * (schemes/fv/operators/convection_lin.hpp:95-178, the WENO5 flux of
 * weno_impl.hpp:26-66): six-cell stencils along each direction, applied
 * explicitly like the other flux-based operators.  The mesh holds its cells
 * at one level with ghost layers of width 3 (smr_mesh_config
 * max_stencil_radius = 3, min_level == max_level: the reference's
 * max_stencil_size(6)); the far ghost layers come from polynomial
 * extrapolation.  u's ghosts are updated first; velocity: dim values. */
int smr_convection_weno5(const double* velocity, smr_field* u, smr_field* out);
/* Diagnostic: the ordered contribution lists behind smr_convection_upwind,
 * computed on the host.  Call with null arrays to get *nrows and *nterms;
 * row: nrows storage slots, ptr: nrows+1, src: 2*nterms (second -1 for a
 * single product), coef: nterms.  With a null velocity, the lists behind
 * smr_convection_upwind_nonlinear: term k = coef[k] * F(u[src[2k]], u[src[2k+1]]). */
int smr_mesh_convection_terms(const smr_mesh* mesh, const double* velocity, int64_t* nrows, int64_t* nterms, int64_t* row, int64_t* ptr,
                              int64_t* src, double* coef);

/* ------------------------------------------------------------------------
 * Batched form: many independent instances of the advection demo loop
 * (demos/FiniteVolume/advection_2d.cpp:129-153) advanced together.  No
 * single reference entry point corresponds: each call does, for every
 * instance, what the reference does per mesh with make_MRAdapt(u)(eps, reg)
 * (mr/adapt.hpp:147-203), unp1.resize(), update_ghost_mr(u)
 * (algorithm/update_ghost_mr.hpp:260-270), unp1 = u - dt * upwind(a, u)
 * (stencil_field.hpp:175-179) and swap(u, unp1) (field/swap.hpp:12-19).
 * The results are bit-identical to those per-instance calls.
 * ---------------------------------------------------------------------- */
/* u[i], unp1[i]: the two fields of instance i (same mesh, one mesh per
 * instance); a: n*3 velocities (third ignored in 2D); dt: n time steps */
int smr_batch_create(smr_ctx* ctx, int n, smr_field* const* u, smr_field* const* unp1, const double* a, const double* dt,
                     smr_batch** out);
int smr_batch_destroy(smr_batch* batch);
/* adaptation of every u; returns the total number of harten calls (>= 0) */
int smr_batch_adapt(smr_batch* batch, double eps, double regularity);
/* ghost update, upwind into unp1 and swap, for every instance */
int smr_batch_advance(smr_batch* batch);
/* adapt + advance; returns the total number of harten calls (>= 0) */
int smr_batch_step(smr_batch* batch, double eps, double regularity);
int64_t smr_batch_nb_cells(const smr_batch* batch);
/* launch shape of the stage programs: block size (multiple of 64, <= 1024),
 * consecutive cells per thread (0: one precomputed 64-index wave item per
 * wavefront, blocks of 256), mode (0 one cooperative launch per program,
 * 1 one launch per stage, 2 stage launches replayed from a hipGraph, 3 for
 * programs of independent instances: every run of small stages one launch
 * with one block per instance, large stages grid-wide, replayed from a
 * hipGraph; a program without independent instances runs mode 3 as 2) */
int smr_batch_set_tuning(smr_batch* batch, int block, int cells_per_thread, int mode);
int smr_batch_get_tuning(const smr_batch* batch, int* block, int* cells_per_thread, int* mode);
/* times every launch shape on the advance program of the current state and
 * keeps the fastest; stores it for this GPU architecture in cache_file when
 * non-null */
int smr_batch_autotune(smr_batch* batch, int reps, const char* cache_file);
/* 1 if cache_file holds a launch shape for this GPU architecture (then applied), else 0 */
int smr_batch_load_tuning(smr_batch* batch, const char* cache_file);

/* ------------------------------------------------------------------------
 * Adjoint form: gradients of a loss of the cell values after `steps` steps
 * of update_ghost_mr(u) (algorithm/update_ghost_mr.hpp:194-237) +
 * unp1 = u - dt * upwind(a, u) (stencil_field.hpp:16-97) + swap, on frozen
 * meshes, with respect to the initial cell values and the velocity.  The
 * reference has no adjoint; these entry points differentiate exactly the
 * operators above.  Many instances (one mesh each) are handled together.
 * Cell vectors are compact: the cells of each instance in
 * for_each_cell(mesh[cells]) order, instances concatenated.
 * ---------------------------------------------------------------------- */
/* meshes stay frozen while the object lives; dt: n steps; bc_value: n
 * Dirichlet values (null: no boundary condition); steps must be a multiple of
 * checkpoint_every (one checkpoint of the state per checkpoint_every steps) */
int smr_adjoint_create(smr_ctx* ctx, int n, smr_mesh* const* meshes, const double* dt, const double* bc_value, int steps,
                       int checkpoint_every, smr_adjoint** out);
int smr_adjoint_destroy(smr_adjoint* adj);
/* compact cell count of one instance, or of all instances when instance < 0 */
int64_t smr_adjoint_nb_cells(const smr_adjoint* adj, int instance);
/* n*3 velocities (third ignored in 2D) */
int smr_adjoint_set_velocity(smr_adjoint* adj, const double* a);
/* u0 -> uT; device pointers (compact cell vectors); returns when uT is written */
int smr_adjoint_forward(smr_adjoint* adj, const double* u0, double* uT);
/* grad_uT (device) -> grad_u0 (device) and grad_a (host, n*3) of the last forward */
int smr_adjoint_backward(smr_adjoint* adj, const double* grad_uT, double* grad_u0, double* grad_a);
/* launch shape of the adjoint's stage programs (as smr_batch_set_tuning) */
int smr_adjoint_set_tuning(smr_adjoint* adj, int block, int cells_per_thread, int mode);

#ifdef __cplusplus
}
#endif

#endif
