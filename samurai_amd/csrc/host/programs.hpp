// Host-built work descriptions of the MR step for one mesh version.
//
// Every bulk operator of the reference is `set.apply_op(op)`: a set
// expression evaluated to x-runs, then a per-run loop that reads and writes
// field storage through `memory_offset(mesh, {level, x, row})` plus an x
// offset.  Here the host evaluates the set expression (interval algebra) into
// a WorkList of runs, and lists the rows the operator touches as Taps; the
// device resolves each (run, tap) to a storage base by binary search in the
// reference interval table, then the operator kernels stream over the cells.
//
// The outer-ghost updates (Dirichlet, corner extrapolation, BC projection and
// prediction) touch O(boundary) cells with data-dependent sources; they are
// resolved on the host into Recipes (dst index, op, source indices) kept in
// the reference's execution order and grouped into dependency-free stages.
#pragma once

#include "mr_mesh.hpp"

#include <cstdint>
#include <string>
#include <vector>

namespace smr
{
    // A row touched by an operator, relative to a work run at level l, row (y, z),
    // starting at x0:
    //   dl =  0: level l,   row (y+dy, z+dz),           first x = x0 + dx
    //   dl = +1: level l+1, row (2y+dy, 2z+dz),         first x = 2*x0 + dx
    //   dl = -1: level l-1, row ((y>>1)+dy, (z>>1)+dz), first x = (x0>>1) + dx
    // table 0 is the mesh the list belongs to, table 1 the previous mesh
    // (update_fields reads the old field through it).  [wlo, whi] is the window
    // the kernel reads or writes around each cell's target x: the device checks
    // that the window of the whole run lies in the one reference run it
    // resolved, so every kernel access stays inside the field storage.
    struct Tap
    {
        int8_t table, dl, dx, dy, dz, wlo, whi;
    };

    struct WorkList
    {
        std::vector<int32_t> iv;        // 5 per run: x0, x1, y, z, level
        std::vector<int64_t> pref{0};   // cell prefix over runs
        std::vector<Tap> taps;
        std::vector<int64_t> level_run; // first run of each level (MAXL+1)
        // column walks for the upwind list: runs stacked at y, y+1, ... with one x range
        bool walk = false;

        int64_t nruns() const { return int64_t(iv.size() / 5); }
        int64_t ncells() const { return pref.back(); }
        void add(const LevelSet& s);
        // marks that the runs of `level` start here (levels must be added in increasing order)
        void begin_level(int level);
        void finish();
        int64_t runs_of_levels(int l0, int l1, int64_t& r0) const; // runs of levels [l0, l1)
    };

    enum RecipeOp : int32_t
    {
        RECIPE_COPY      = 0, // dst = src0
        RECIPE_DIRICHLET = 1, // dst = 2*v - src0
        RECIPE_AVERAGE     = 2, // dst = 0; dst += src_k; dst /= n (if n > 0)
        RECIPE_EXTRAPOLATE = 3, // polynomial extrapolation from n = 3 or 5 sources (stencil 4 or 6)
        RECIPE_COPY_NO_BC  = 4  // dst = src0 for a field without a B.C. (skipped when it has one)
    };

    struct Recipes
    {
        std::vector<int64_t> dst;
        std::vector<int32_t> op;
        std::vector<int64_t> src_ptr{0};
        std::vector<int64_t> src;
        std::vector<int64_t> stage_ptr{0}; // recipes of stage k: [stage_ptr[k], stage_ptr[k+1])

        size_t size() const { return dst.size(); }
        void push(int64_t d, int32_t o, const int64_t* s, int n);
        // reorders the recipes into dependency-free stages, keeping the order of
        // every pair that shares a storage location
        void stage();
        // stage k holds only RECIPE_COPY_NO_BC recipes (nothing for a field with a B.C.)
        bool stage_without_bc_only(size_t k) const;
    };

    // update_ghost_mr (algorithm/update_ghost_mr.hpp:194-237) for one mesh
    struct GhostProgram
    {
        std::vector<Recipes> outer;      // per level: update_outer_ghosts(level)
        WorkList projection;             // per level l: (reference[l] n proj_cells[l-1]).on(l-1)
        WorkList prediction;             // per level l: prediction ghosts
    };

    // harten's per-mesh sets (mr/adapt.hpp:276-389)
    struct AdaptProgram
    {
        WorkList detail;   // runs at level l (coarse): ghosts below cells
        WorkList criteria; // runs at level l-1 (coarse) for fine level l
        WorkList maximum;  // same sets as criteria
    };

    // the work lists of the programs, in the order of the device-built lists
    enum ListKind : int
    {
        LIST_PROJECTION = 0,
        LIST_PREDICTION = 1,
        LIST_DETAIL     = 2,
        LIST_CRITERIA   = 3,
        LIST_MAXIMUM    = 4,
        LIST_UF_COPY    = 5,
        LIST_UF_COARSEN = 6,
        LIST_UF_REFINE  = 7
    };
    // the taps of a list of that kind (dim, prediction radius s)
    std::vector<Tap> list_taps(int kind, int dim, int s);
    // update_outer_ghosts' recipes of every level and the cells list (the
    // parts of the programs the device does not build)
    void build_recipes_and_cells(const MRMeshData& m, std::vector<Recipes>& outer, WorkList& cells);

    GhostProgram build_ghost_program(const MRMeshData& m);
    AdaptProgram build_adapt_program(const MRMeshData& m);
    WorkList build_cells_list(const MRMeshData& m);             // mesh[cells], taps for upwind
    // the three above together, their per-level work in one parallel loop
    // with old and uf: also the update_fields lists (copy, coarsen, refine)
    // from the mesh old to m into uf[0..2], in the same pool of tasks
    void build_programs(const MRMeshData& m, WorkList& cells, GhostProgram& ghost, AdaptProgram& adapt, const MRMeshData* old = nullptr,
                        WorkList* uf = nullptr);
    // update_fields (algorithm/update_fields.hpp:27-127): table 0 = new mesh, 1 = old
    void build_update_fields_lists(const MRMeshData& old_m, const MRMeshData& new_m, WorkList& copy, WorkList& coarsen,
                                   WorkList& refine);

    // Host restatement of the device resolve check: every (run, tap) row exists
    // and its access window lies in one reference run.  Returns "" or a message.
    std::string validate_list(const WorkList& w, const MRMeshData& t0, const MRMeshData& t1);
    // validate_list over every list of the mesh's programs
    std::string validate_programs(const MRMeshData& m);
} // namespace smr
