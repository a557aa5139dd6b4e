// Adjoint of the upwind time step on frozen adapted meshes.
//
// On a frozen mesh one step is linear in the cell values:
//   u_{n+1} = U(a) G u_n,
// with G = update_ghost_mr (outer-ghost recipes, projection, prediction: each
// stage writes ghosts as fixed linear combinations of earlier values) and U(a)
// the upwind update of the cells.  The adjoint step is
//   lam_n = G^T U(a)^T lam_{n+1},
// G^T being G's stages in reverse order, each as a deterministic gather over its
// transposed (target <- sources) lists followed by zeroing its outputs.  The
// velocity gradient accumulates lam_{n+1} . dU/da (u_n) per cell.
//
// Many instances are advanced together with the batched stage programs; the
// forward pass keeps a checkpoint every `every` steps and the backward pass
// recomputes each segment from its checkpoint before running it backwards.
#pragma once

#include "batch.hpp"

#include <array>
#include <map>
#include <memory>
#include <vector>

namespace smr
{
    // u[dst] = sum w u[src] for the (dst, src, w) triplets of one forward stage,
    // stored transposed: row r targets tgt[r] (a forward source) and gathers
    // (src[k], w[k]) for k in [ptr[r], ptr[r+1]) (forward outputs)
    struct LinearStage
    {
        std::vector<int64_t> dst;
        std::vector<int64_t> tgt;
        std::vector<int64_t> ptr{0};
        std::vector<int64_t> src;
        std::vector<double> w;
    };

    // G of one mesh as linear stages, in the forward order of ghost_chain
    std::vector<LinearStage> ghost_linear_stages(const MRMeshData& m, const DevMesh& dm, bool has_bc);

    // G^T as launches: the reverse stages (gather, then zeroing, of each
    // forward stage from the last) in order, consecutive items that touch
    // disjoint slots sharing a launch; item (k, 0) gathers stage k, (k, 1)
    // zeroes its outputs
    std::vector<std::vector<std::pair<int, int>>> group_reverse_stages(const std::vector<LinearStage>& ghost);

    class Adjoint
    {
      public:
        // meshes: frozen adapted meshes (each with a complete device image);
        // dt: n time steps; bc: n Dirichlet values or null (no boundary condition)
        Adjoint(Ctx* ctx, std::vector<Mesh*> meshes, const double* dt, const double* bc, int steps, int every);
        ~Adjoint();

        int ninstances() const { return int(inst_.size()); }
        // compact cell count of instance i, or of all instances (i < 0)
        int64_t nb_cells(int i) const;
        // n*3 velocities (the third is ignored in 2D)
        void set_velocity(const double* a);
        // u0 -> uT after `steps` steps; device pointers to the compact cell
        // vectors of all instances (instance order, for_each_cell order)
        void forward(const double* u0, double* uT);
        // gT = dL/duT -> g0 = dL/du0 (device), ga = dL/da (host, n*3)
        void backward(const double* gT, double* g0, double* ga);
        void set_tuning(const Tuning& t) { tuning_ = t; }
        const Tuning& tuning() const { return tuning_; }
        int steps() const { return steps_; }
        int every() const { return every_; }

      private:
        struct Instance
        {
            Mesh* mesh = nullptr;
            uint64_t version = 0;
            double a[3] = {0, 0, 0};
            double dt = 0;
            bool has_bc = false;
            double bc = 0;
            int64_t nref = 0, ncells = 0, off = 0; // off: first compact cell
            std::vector<LinearStage> ghost;
            std::array<DevWorkList, 6> wdir;
            std::unique_ptr<Blob> blob;
            struct DevStage
            {
                const int64_t* dst;
                const int64_t* tgt;
                const int64_t* ptr;
                const int64_t* src;
                const double* w;
                int64_t ndst, nrows;
            };
            std::vector<DevStage> dghost;
            // the reverse ghost stages grouped into launches: item (k, 0) is the
            // gather of forward stage k, (k, 1) its zeroing; items of a group
            // touch disjoint slots (no read-write or write-write overlap)
            std::vector<std::vector<std::pair<int, int>>> adj_groups;
            // scatter directions grouped into launches (disjoint ghost targets)
            std::vector<std::vector<int>> scatter_groups;
            double* states = nullptr; // (every + 1) * nref
            double* ckpt   = nullptr; // nseg * nref
            double* lam    = nullptr; // 2 * nref
            double* ga     = nullptr; // dim * ncells
            double* state(int j) const { return states + int64_t(j) * nref; }
        };

        Ctx* ctx_;
        std::vector<Instance> inst_;
        int steps_, every_, nseg_;
        // the programs are built once per velocity and replayed: graph replays of
        // the stage launches by default
        Tuning tuning_{256, 0, MODE_BLOCKS};
        bool programs_ready_ = false;
        const double* bound_u0_ = nullptr;
        double* bound_uT_       = nullptr;
        const double* bound_gT_ = nullptr;
        double* bound_g0_       = nullptr;
        // per checkpoint segment: the forward program (checkpoint, steps, roll)
        // and the backward one (restore, recomputed steps, adjoint steps)
        std::vector<std::unique_ptr<ProgramBuf>> fseg_, bseg_;
        std::unique_ptr<ProgramBuf> in_, out_, lam_in_, grad_out_;
        // per-cell velocity-gradient terms of all instances (instance i at ga_off_[i])
        double* ga_arena_ = nullptr;
        std::vector<int64_t> ga_off_;
        PinnedStaging ga_host_;

        void check_meshes() const;
        void build_programs();
        void bind_io(const double* u0, double* uT, const double* gT, double* g0);
        std::unique_ptr<ProgramBuf> make_program(const std::vector<Chain>& chains);
    };
} // namespace smr
