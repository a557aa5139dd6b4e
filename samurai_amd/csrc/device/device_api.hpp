// Launchers of the MR-step kernels (kernels.hip).  Plain pointers and sizes;
// every launch is asynchronous on the given stream.
#pragma once

#include <hip/hip_runtime_api.h>

#include <cstdint>
#include <vector>

namespace smr::dev
{
    // Flattened reference interval table of one mesh (all levels).
    struct Table
    {
        const int64_t* row_key; // ((int64)z << 32) + (y + 2^31), sorted per level
        const int64_t* row_off; // MAXL+1: first row of each level
        const int64_t* rp;      // first run of each row (+1 sentinel)
        const int32_t* run_s;
        const int32_t* run_e;
        const int64_t* run_idx; // value index = run_idx + x
    };

    // A work list: x-runs (x0, x1, y, z, level), cell prefix, chunk hints and
    // the resolved storage bases of its taps.
    struct List
    {
        const int32_t* iv;
        const int64_t* pref;
        const int32_t* hint; // run holding cell 256*j, j = 0..ceil(ncells/256)
        int64_t* base;       // nruns * ntaps
        int32_t* cellrun;    // the run of every cell, filled by the resolve (null: cells are searched for)
        const int8_t* taps;  // 7 per tap: table, dl, dx, dy, dz, wlo, whi
        const int32_t* tiles; // (run, x0) of every 64-cell tile of a run, runs in order
        // column walks (2D cells list): (first run, x0, number of runs n, first
        // row slot) per walk; next[r] is the run right above run r with the same
        // x range, or -1; walk_rows holds, per walk, walk_stride slots: the
        // storage bases of the n + 2 rows it touches (the row below, the n rows,
        // the row above) from slot 0, then (x0 << 32 | n) and (x1 << 32 | level)
        // in the last two slots, filled on the device from the resolved bases,
        // so one load of a walk's slots gives a wave everything it needs
        const int32_t* walks;
        const int32_t* next;
        int64_t* walk_rows;
        int32_t walk_stride;
        int64_t nruns;
        int64_t ncells;
        int64_t ntiles;
        int64_t nwalks;
        int ntaps;
    };

    struct Thresholds
    {
        double eps[21];
        double fine[21];
        double coarse[21];
    };

    constexpr int HINT_SHIFT = 8; // 256 cells per hint chunk

    // Per-level dense bitmaps of the unit box at levels [lo, hi] (rebuild.hip):
    // level l stores each row (y, or (z, y)) in wpr[l] = max(1, 2^l / 64)
    // 64-bit words from word off[l] (off[lo] = 0, off[hi + 1] = total), x
    // ascending from bit 0.
    struct Bits
    {
        int lo = 0, hi = -1;
        int64_t off[22] = {};
        int64_t wpr[21] = {};
    };

    // One mesh of a batched device mesh update (rebuild_batch.hip): its bitmap
    // layout, its resolved cells list, its tags, and the first word of its
    // bitmaps in the concatenated arrays of the batch.
    struct MeshItem
    {
        Bits b;
        List cells;
        const uint8_t* tag;
        int64_t w0;
        int dim, min_level, max_level;
    };

    void mesh_batch_zero(uint64_t* p, int64_t n, hipStream_t s);
    // out = (old | add) & ~rem over n words (graduation.hpp:832-840)
    void mesh_batch_combine(const uint64_t* old, const uint64_t* add, const uint64_t* rem, uint64_t* out, int64_t n, hipStream_t s);
    // the leaves of every item's cells list (rpref: prefix of their runs)
    void mesh_batch_leaves(const MeshItem* items, const int64_t* rpref, int64_t nruns, int n, uint64_t* bits, hipStream_t s);
    // update_cell_array_from_tag's marks of every item (cpref: prefix of cells)
    void mesh_batch_tag_marks(const MeshItem* items, const int64_t* cpref, int64_t ncells, int n, uint64_t* add, uint64_t* rem, hipStream_t s);
    // one graduation pass of every item (wpref: prefix of bitmap words);
    // flags[k * fstride + pass] = 1 when item k refines a cell
    void mesh_batch_graduation_pass(const MeshItem* items, const int64_t* wpref, int64_t nwords, int n, const uint64_t* ca, uint64_t* p,
                                    uint64_t* out, int w2, int* flags, int fstride, int pass, hipStream_t s);
    // flags[k * fstride + slot] = 1 when item k's bitmaps a and c differ
    void mesh_batch_differ(const int64_t* wpref, int64_t nwords, int n, const uint64_t* a, const uint64_t* c, int* flags, int fstride, int slot,
                           hipStream_t s);
    size_t mesh_batch_scan_temp_bytes(int64_t nwords);
    // runs of the items whose flag `slot` is set: counts and their exclusive sum,
    // then (x0, x1, y, z, level, item) records, at most cap
    void mesh_batch_run_offsets(const MeshItem* items, const int64_t* wpref, int64_t nwords, int n, const uint64_t* bits, const int* flags,
                                int fstride, int slot, int32_t* counts, int32_t* offsets, void* temp, size_t temp_bytes, hipStream_t s);
    void mesh_batch_run_emit(const MeshItem* items, const int64_t* wpref, int64_t nwords, int n, const uint64_t* bits, const int* flags,
                             int fstride, int slot, const int32_t* offsets, int32_t* runs, int64_t cap, hipStream_t s);

    // Per-level dense bitmaps of levels [0, hi] padded by `pad` cells on every
    // side (submesh.hip): level l has n[l] = 2^l + 2 pad cells per axis, a
    // row (Y, or (Z, Y) in 3D, padded indices) holds wpr[l] 64-bit words, x
    // ascending from bit 0 at unpadded x = -pad; level l starts at word off[l].
    struct PBits
    {
        int dim, hi, pad;
        int64_t n[21];
        int64_t wpr[21];
        int64_t off[22];
    };

    inline PBits make_pbits(int dim, int hi, int pad)
    {
        PBits p{};
        p.dim       = dim;
        p.hi        = hi;
        p.pad       = pad;
        int64_t off = 0;
        for (int l = 0; l <= 20; ++l)
        {
            p.n[l]   = (int64_t(1) << l) + 2 * pad;
            p.wpr[l] = (p.n[l] + 63) / 64;
            p.off[l] = off;
            if (l <= hi)
            {
                off += p.wpr[l] * (dim == 3 ? p.n[l] * p.n[l] : p.n[l]);
            }
        }
        p.off[21] = off;
        return p;
    }

    // int32 per run (x0, x1, y, z) and segment starts per set (levels 0..21)
    // of the device mesh construction's output
    constexpr int SM_RUN = 4;
    constexpr int SM_SEG = 22;

    // the sets of the device mesh construction, in the order of the mesh ids
    enum SubmeshSet : int
    {
        SM_CELLS            = 0,
        SM_CELLS_AND_GHOSTS = 1,
        SM_PROJ_CELLS       = 2,
        SM_UNION_CELLS      = 3,
        SM_REFERENCE        = 4,
        SM_NSETS            = 5
    };

    // the five ids of MRMesh::update_sub_mesh_impl from the leaves of levels
    // [b.lo, b.hi] (unpadded), into sets (SM_NSETS x p.off[p.hi + 1] words),
    // with scratch of 2 x p.off[p.hi + 1] words; r: scheme stencil radius, s:
    // prediction radius, mr: max_level > min_level
    void submesh_build(const Bits& b, const uint64_t* leaves, const PBits& p, int r, int s, bool mr, uint64_t* sets, uint64_t* scratch,
                       hipStream_t st);
    size_t submesh_scan_temp_bytes(int64_t n);
    // per-word run counts of nsets consecutive sets and their exclusive sum
    void submesh_run_offsets(const PBits& p, int nsets, const uint64_t* sets, int32_t* counts, int32_t* offsets, void* temp, size_t temp_bytes,
                             hipStream_t s);
    // the runs (x0, x1, y, z), SM_RUN int32 each, of at most cap runs, in
    // (set, level, z, y, x) order
    void submesh_run_emit(const PBits& p, int nsets, const uint64_t* sets, const int32_t* offsets, int32_t* runs, int64_t cap, hipStream_t s);
    // seg[set * SM_SEG + l] = first run of level l of the set (levels past
    // p.hi: the end of the set), for the nsets sets
    void submesh_segments(const PBits& p, int nsets, const int32_t* offsets, const int32_t* counts, int32_t* seg, hipStream_t s);

    // The work lists of build_programs (host/programs.cpp) built on the device
    // from the padded sets of a mesh (SM_NSETS x p.off[p.hi + 1] words) and,
    // for the update_fields lists, of the mesh it replaces (lists.hip): each
    // list is a per-level bitmap expression of the mesh ids, run-length
    // encoded in (level, z, y, x) order into the arrays the host would build.
    enum DevListKind : int
    {
        DL_PROJECTION = 0,
        DL_PREDICTION = 1,
        DL_DETAIL     = 2,
        DL_CRITERIA   = 3,
        DL_MAXIMUM    = 4,
        DL_UF_COPY    = 5,
        DL_UF_COARSEN = 6,
        DL_UF_REFINE  = 7,
        DL_N          = 8
    };

    // where list k's runs go: iv (5 int32 per run), pref (nruns + 1), hint
    // (one per 2^HINT_SHIFT cells, plus one); run0, cell0: the list's first
    // run and cell in the scanned counts
    struct ListOut
    {
        int32_t* iv;
        int64_t* pref;
        int32_t* hint;
        int64_t run0, cell0, nruns, ncells;
    };
    struct ListOuts
    {
        ListOut l[DL_N];
    };

    size_t lists_scan_temp_bytes(int64_t n);
    // words of the first nlists lists of a mesh: every list over the padded
    // words of its own run levels only, one list after the other
    int64_t lists_words(const PBits& p, int minl, int maxl, int nlists);
    // the bitmaps of the first nlists lists (lists_words words; old_sets may
    // be null when nlists <= DL_UF_COPY), per-word packed counts (runs << 32
    // | cells) and their exclusive sum, and table[k * SM_SEG + l]: the packed
    // scan at the first word of level l of list k (levels below the list's:
    // its start, above: its end)
    void lists_build(const PBits& p, int minl, int maxl, const uint64_t* sets, const uint64_t* old_sets, int nlists, uint64_t* bits,
                     uint64_t* counts, uint64_t* scan, void* temp, size_t temp_bytes, uint64_t* table, hipStream_t s);
    // writes the runs, prefixes, hints and trailers of every list into outs
    void lists_emit(const PBits& p, int minl, int maxl, int nlists, const uint64_t* bits, const uint64_t* scan, const ListOuts& outs,
                    hipStream_t s);

    // One mesh of a batched device mesh construction (submesh_batch.hip): its
    // new leaves (unpadded layout b, from word lw0 of the concatenated
    // leaves), its padded layout p (its SM_NSETS sets from word SM_NSETS * pw0
    // of the concatenated sets, its two scratch maps from word 2 * pw0 of the
    // scratch), its stencil radius r, prediction radius s and mr = max_level >
    // min_level, and its min_level.
    struct SubmeshItem
    {
        Bits b;
        PBits p;
        int64_t lw0, pw0;
        int r, s, mr, minl;
    };

    // all DL_N lists of every item (its min_level minl, its max_level p.hi)
    // from the items' padded ids (sets) and the ids of the meshes they replace
    // (old_sets), both laid out like submesh_batch_build's sets: bitmaps
    // (item k's lists_words words from index cp[k] on, nwords = cp[n]),
    // packed counts and their scan, and table[(DL_N * k + j) * SM_SEG + l]:
    // the packed scan at the first word of level l of list j of item k
    void lists_batch_build(const SubmeshItem* items, const int64_t* cp, int64_t nwords, int n, const uint64_t* sets, const uint64_t* old_sets,
                           uint64_t* bits, uint64_t* counts, uint64_t* scan, void* temp, size_t temp_bytes, uint64_t* table, hipStream_t s);
    // the runs and trailers of every list; outs[DL_N * k + j]: list j of item k (device memory)
    void lists_batch_emit(const SubmeshItem* items, const int64_t* cp, int64_t nwords, int n, const uint64_t* bits, const uint64_t* scan,
                          const ListOut* outs, hipStream_t s);

    // submesh_build of every item (ppref: prefix of the items' padded words,
    // pwords = ppref[n])
    void submesh_batch_build(const SubmeshItem* items, const int64_t* ppref, int64_t pwords, int n, const uint64_t* leaves, uint64_t* sets,
                             uint64_t* scratch, hipStream_t st);
    // per-word run counts of the SM_NSETS * pwords words of the sets and
    // their exclusive sum
    void submesh_batch_run_offsets(const SubmeshItem* items, const int64_t* ppref, int64_t pwords, int n, const uint64_t* sets, int32_t* counts,
                                   int32_t* offsets, void* temp, size_t temp_bytes, hipStream_t st);
    // the runs (x0, x1, y, z), SM_RUN int32 each, of at most cap runs, in
    // (item, set, level, z, y, x) order
    void submesh_batch_run_emit(const SubmeshItem* items, const int64_t* ppref, int64_t pwords, int n, const uint64_t* sets, const int32_t* offsets,
                                int32_t* runs, int64_t cap, hipStream_t st);
    // seg[(SM_NSETS * item + set) * SM_SEG + l] = first run of level l of
    // that set of that item (levels past its p.hi: the end of the set)
    void submesh_batch_segments(const SubmeshItem* items, const int64_t* ppref, int n, const int32_t* offsets, const int32_t* counts, int32_t* seg,
                                hipStream_t st);


    // One operator applied to a cell range of one instance: the unit of the
    // fused stage program that advances many independent meshes together.
    enum SegmentOp : int32_t
    {
        SEG_RECIPE        = 0,
        SEG_PROJECT       = 1,
        SEG_PREDICT_FINE  = 2,
        SEG_DETAIL        = 3,
        SEG_CRITERIA      = 4,
        SEG_MAXIMUM       = 5,
        SEG_TAG_INIT      = 6,
        SEG_UPWIND        = 7,
        // update_fields: `out` is the new storage, `u` the old one
        SEG_COPY             = 8,
        SEG_PROJECT_TO       = 9,
        SEG_PREDICT_CHILDREN = 10,
        // storage bases of a list's (run, tap) pairs against tabs[0], tabs[1];
        // the "cells" of the segment are run * ntaps + tap
        SEG_RESOLVE = 11,
        // zero fills of out[c] / tag[c]
        SEG_FILL_F64 = 12,
        SEG_FILL_U8  = 13,
        // out[c] = u[c]
        SEG_COPY_RAW = 14,
        // cells of a list between storage and a compact vector (list order):
        // to: out[c] = u[storage(c)]; from: out[storage(c)] = in2[c]
        SEG_CELLS_TO_COMPACT   = 15,
        SEG_CELLS_FROM_COMPACT = 16,
        // adjoint of a linear stage, as gathers over its transpose (CSR rows r:
        // target rdst[r], entries rsp[r]..rsp[r+1] of (rsrc, thr)):
        // out[rdst[r]] += sum thr[k] * out[rsrc[k]]
        SEG_ADJ_GATHER = 17,
        // out[rdst[r]] = 0 (the outputs of the stage whose adjoint was gathered)
        SEG_ADJ_ZERO = 18,
        // adjoint of the upwind step on the cells list: out = transposed
        // stencil applied to u (zero outside cells); velocity-gradient terms
        // accumulated into out2[dim * c + d] using the forward state in2
        SEG_ADJ_UPWIND = 19,
        // the part of the transposed stencil that lands on ghost neighbours:
        // cells of the list whose neighbour in direction `pad` is a ghost
        SEG_ADJ_SCATTER = 20,
        // err[0] = 1 when a cell's tag changes the cell array (cells list)
        SEG_TAG_CHANGE = 21,
        // the row bases of column walk c of the list (after its resolve)
        SEG_WALK_ROWS = 22
    };

    // operators whose index c is a cell of the segment's work list (the
    // others index recipes, rows, runs x taps, walks or storage slots)
    __host__ __device__ constexpr bool is_cell_list_op(int op)
    {
        return op == SEG_PROJECT || op == SEG_PREDICT_FINE || op == SEG_DETAIL || op == SEG_CRITERIA || op == SEG_MAXIMUM ||
               op == SEG_TAG_INIT || op == SEG_UPWIND || op == SEG_COPY || op == SEG_PROJECT_TO || op == SEG_PREDICT_CHILDREN ||
               op == SEG_CELLS_TO_COMPACT || op == SEG_CELLS_FROM_COMPACT || op == SEG_ADJ_UPWIND || op == SEG_ADJ_SCATTER ||
               op == SEG_TAG_CHANGE;
    }

    // One stage of a chain of small dependent stages run in order by a single
    // block (k_sequence): a range of outer-ghost recipes, of projection or
    // prediction ghosts, or of tag-finalization cells.  Launching each such
    // stage on its own costs more than its few hundred cells.
    enum SeqOp : int32_t
    {
        SEQ_RECIPES = 0,
        SEQ_PROJECT = 1,
        SEQ_PREDICT = 2,
        SEQ_MAXIMUM = 3
    };
    struct SeqItem
    {
        int32_t op;
        int32_t pad;
        int64_t c0, c1;
        const int64_t* dst; // recipes of the stage (SEQ_RECIPES)
        const int32_t* rop;
        const int64_t* sp;
        const int64_t* src;
    };
    // items [i0, i1) of a device array, in order, by one block of 1024
    // threads with a block barrier between items; l_proj, l_pred and l_max are
    // the lists the project, predict and maximum items index
    void sequence(int dim, int ps, const SeqItem* items, int i0, int i1, const List& l_proj, const List& l_pred, const List& l_max, double* u,
                  uint8_t* tag, bool has_bc, double bc_value, hipStream_t s);

    struct Segment
    {
        int32_t op, dim, pred_s, has_bc;
        int32_t min_level, max_level;
        // host use while a program is assembled: bit 0 (thr) and bit 1 (tabs)
        // mark pointers that still hold offsets into the program's aux bytes
        int32_t aux, pad;
        int64_t c0, c1;  // cell (or recipe) range of the list
        List list;       // work list (list ops)
        const int64_t* rdst;  // recipes
        const int32_t* rop;
        const int64_t* rsp;
        const int64_t* rsrc;
        double* u;       // field storage
        double* out;     // unp1 (upwind) or details
        uint8_t* tag;
        const double* thr; // eps/fine/coarse per fine level (criteria)
        const Table* tabs; // resolve: the two tables of the list's taps
        int* err;          // resolve: set to 1 when a pair fails
        const double* in2; // second input (adjoint: forward state; compact input)
        double* out2;      // second output (adjoint: velocity-gradient terms)
        double a0, a1, a2, dt, bc;
    };

    // Runs stages [0, nstages): stage k is the segments [stage_seg[k],
    // stage_seg[k+1]); seg_pref is the running cell count over all segments
    // (segment i covers [seg_pref[i], seg_pref[i+1]), nsegs + 1 entries).
    // Stages are separated by grid barriers (cooperative launch) or, in
    // stage-wise mode, by kernel boundaries.
    // The work of one wavefront in a stage: n chunks of 64 consecutive
    // indices of the segments [seg, end) (consecutive in the program) from
    // index `off` of segment seg on, lanes stepping into the next segments as
    // they pass their ends.  Segments of a stage are packed densely, and a
    // wave knows its first segment without searching for it; n > 1 only
    // within one segment (large fills and copies).
    struct WaveItem
    {
        int32_t seg, off, end, n;
    };

    // operators cheap enough per index that a wave takes several chunks
    __host__ __device__ constexpr bool is_streaming_op(int op)
    {
        return op == SEG_FILL_F64 || op == SEG_FILL_U8 || op == SEG_COPY_RAW;
    }

    struct Program
    {
        const Segment* segs;
        const int64_t* seg_pref;
        const int32_t* stage_seg;
        int nstages;
        // wave items of stage k: [stage_item[k], stage_item[k + 1])
        const WaveItem* items;
        const int64_t* stage_item;
    };

    // maximum co-resident blocks of the fused kernel for a block size
    int program_max_blocks(int block);
    // one cooperative launch; returns false if the runtime refused it
    bool run_program_fused(const Program& p, int block, int nblocks, int cpt, hipStream_t s);
    // stage k alone, grid-wide, one wave item per wavefront
    void run_program_stage(const Program& p, int k, int64_t nitems, hipStream_t s);
    // stage k alone, grid-wide, cpt consecutive cells per thread in blocks of `block`
    void run_program_stage_cells(const Program& p, int k, int64_t cells, int block, int cpt, hipStream_t s);
    // one block of `threads` per chain, each running its chain's items
    // items[off[b] .. off[b+1]): item = segment index << 1 | (block barrier after it)
    void run_program_blocks(const Program& p, const int32_t* off, const int32_t* items, int nchains, int threads, hipStream_t s);

    // threads of a resolve: the (run, tap) pairs, and the cells for the run map
    inline int64_t resolve_range(const List& l)
    {
        const int64_t pairs = l.nruns * l.ntaps;
        return l.cellrun && l.ncells > pairs ? l.ncells : pairs;
    }
    void resolve(const List& l, const Table& t0, const Table& t1, int* err, hipStream_t s);
    // the resolves of up to kResolveJobs lists in one launch: list k against
    // (t_new, t_new), or (t_new, t_old) when old[k] is set
    constexpr int kResolveJobs = 12;
    struct ResolveBatch
    {
        List lists[kResolveJobs];
        int64_t start[kResolveJobs + 1]; // first thread of each list
        int8_t old[kResolveJobs];
        int n;
        Table t_new, t_old;
    };
    void resolve_batch(const ResolveBatch& b, int* err, hipStream_t s);
    // fills l.walk_rows from the resolved bases (after resolve; no-op without walks)
    void walk_rows(const List& l, hipStream_t s);
    // out[row[r]] = ordered sum of the terms of row r (flux-based convection)
    void conv_gather(const int64_t* row, const int64_t* ptr, const int64_t* src, const double* coef, int64_t nrows, const double* u,
                     double* out, hipStream_t s);
    // the same with term k = coef[k] * F(u[src[2k]], u[src[2k+1]]), F the
    // scalar non-linear upwind flux of f(u) = u * u
    void conv_burgers_gather(const int64_t* row, const int64_t* ptr, const int64_t* src, const double* coef, int64_t nrows, const double* u,
                     double* out, hipStream_t s);
    // make_convection_upwind(velocity_field): term k of direction dir[k] is
    // (coef[k] * w) * u[upwind side], w the mean of velocity[dir[k]] over the
    // two stencil slots (velocity[d] for d >= dim may be null)
    void conv_het_gather(const int64_t* row, const int64_t* ptr, const int64_t* src, const double* coef, const int8_t* dir, int64_t nrows,
                         const double* u, const double* const* velocity, double* out, hipStream_t s);
    // make_convection_weno5(velocity): term k of direction dir[k] is
    // coef[k] * F, F the WENO5 flux of velocity v[dir[k]] over the six stencil
    // slots src[6k..6k+6) (v: three host values)
    void conv_weno5_gather(const int64_t* row, const int64_t* ptr, const int64_t* src, const double* coef, const int8_t* dir, int64_t nrows,
                           const double* v, const double* u, double* out, hipStream_t s);
    // --refine-boundary (mr/adapt.hpp:245-272): tag keep on cells [c0, c1) of a
    // list within r of a face of the domain [0, n)^dim
    void keep_boundary(const List& cells, int64_t c0, int64_t c1, uint8_t* tag, int dim, int n, int r, hipStream_t s);
    // mra_config().relative_detail(true) (mr/rel_detail.hpp:74-112): det[0, n)
    // times 1 / max |u| over cells [c0, c1) of the list (mx: one device word)
    void relative_detail(const List& cells, int64_t c0, int64_t c1, const double* u, double* det, int64_t n, unsigned long long* mx,
                         hipStream_t s);
    // the same on a one-level mesh stored as the dense box of side n + 6 (ghost
    // width 3), without lists; fa: the h-factor of the level
    void weno5_dense(int dim, int n, const double* v, double fa, const double* u, double* out, hipStream_t s);
    // launch shape of the upwind kernel (cells per thread, XCD remap, stores)
    int upwind_variant_count();
    void set_upwind_variant(int id);
    void upwind(int dim, const List& l, int64_t c0, int64_t c1, const double* u, double* unp1, double a0, double a1, double a2,
                double dt, hipStream_t s);
    // out = u - a * v on the cells of a list
    void sub_scaled(const List& l, int64_t c0, int64_t c1, const double* u, double a, const double* v, double* out, hipStream_t s);
    // the README's for_each_interval upwind loop (2D, a = (1, 1))
    void hand_loop(const List& l, int64_t c0, int64_t c1, const double* u, double* unp1, double dt, hipStream_t s);
    // unp1 = u - dt * upwind_scalar_burgers(k, u) on cells [c0, c1) of a 2D list
    void upwind_scalar_burgers(const List& l, int64_t c0, int64_t c1, const double* u, double* unp1, double k0, double k1, double dt,
                               hipStream_t s);
    void project(int dim, const List& l, int64_t c0, int64_t c1, double* dst, const double* src, hipStream_t s);
    void predict_fine(int dim, int pred_s, const List& l, int64_t c0, int64_t c1, double* u, hipStream_t s);
    void detail(int dim, int pred_s, const List& l, int64_t c0, int64_t c1, const double* u, double* det, hipStream_t s);
    // ncomp detail arrays (one per adapted field), cstride apart
    void criteria(int dim, const List& l, int64_t c0, int64_t c1, const double* det, uint8_t* tag, const Thresholds& thr, int min_level,
                  int max_level, hipStream_t s, int ncomp = 1, int64_t cstride = 0);
    void maximum(int dim, const List& l, int64_t c0, int64_t c1, uint8_t* tag, hipStream_t s);
    void tag_init(const List& l, int64_t c0, int64_t c1, uint8_t* tag, hipStream_t s);
    // p[0, n) = 0 by a kernel (graph-capturable, unlike hipMemsetAsync)
    void zero_f64(double* p, int64_t n, hipStream_t s);
    void zero_u8(uint8_t* p, int64_t n, hipStream_t s);
    // flag[0] = 1 if a cell's tag changes the cell array (cells list)
    void tag_change(const List& l, int64_t c0, int64_t c1, const uint8_t* tag, int min_level, int max_level, int* flag, hipStream_t s);
    void copy_cells(const List& l, int64_t c0, int64_t c1, double* dst, const double* src, hipStream_t s);
    void predict_children(int dim, int pred_s, const List& l, int64_t c0, int64_t c1, double* dst, const double* src, hipStream_t s);
    // one dependency-free stage of outer-ghost recipes
    // (Dirichlet recipes are skipped when the field has no boundary condition)
    // mesh update of a harten call on bitmaps (rebuild.hip)
    void bits_zero(const Bits& b, uint64_t* bits, hipStream_t s);
    // bits = the leaves of a cells work list
    void bits_from_cells(const List& l, const Bits& b, uint64_t* bits, hipStream_t s);
    // update_cell_array_from_tag's additions and removals of the leaves of l
    void bits_tag_marks(int dim, const List& l, const Bits& b, const uint8_t* tag, int min_level, int max_level, uint64_t* add, uint64_t* rem,
                        hipStream_t s);
    // out = (old | add) & ~rem
    void bits_combine(const Bits& b, const uint64_t* old, const uint64_t* add, const uint64_t* rem, uint64_t* out, hipStream_t s);
    // one pass of make_graduation's fixed point from ca into out (p: scratch
    // bitmap); *flag = 1 if a cell was refined (flag not cleared)
    void bits_graduation_pass(int dim, const Bits& b, const uint64_t* ca, uint64_t* p, uint64_t* out, int w2, int* flag, hipStream_t s);
    // *flag = 1 if a and c differ (flag not cleared)
    void bits_differ(const Bits& b, const uint64_t* a, const uint64_t* c, int* flag, hipStream_t s);
    // run-length encoding: counts and exclusive offsets (one per word of
    // levels [lo, hi]) of the runs of bits, then the runs as (x0, x1, y, z,
    // level) in (level, z, y, x) order
    size_t bits_scan_temp_bytes(const Bits& b);
    void bits_run_offsets(int dim, const Bits& b, const uint64_t* bits, int32_t* counts, int32_t* offsets, void* temp, size_t temp_bytes,
                          hipStream_t s);
    // (runs of rank >= cap are not written)
    void bits_run_emit(int dim, const Bits& b, const uint64_t* bits, const int32_t* offsets, int32_t* runs, int64_t cap, hipStream_t s);

    void recipes(const int64_t* dst, const int32_t* op, const int64_t* src_ptr, const int64_t* src, int64_t r0, int64_t r1, double* u,
                 bool has_bc, double bc_value, hipStream_t s);
} // namespace smr::dev
