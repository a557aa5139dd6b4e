// Batched MR time steps: many independent adaptive instances (each a mesh
// with its fields u and unp1) advanced together.  Every device phase of a step
// (the Harten adaptation passes, update_fields, update_ghost_mr + upwind) is
// one stage program over all instances: a single persistent cooperative launch,
// or one launch per stage replayed from a hipGraph.  The host mesh rebuilds of
// the instances run in parallel.
#pragma once

#include "runtime.hpp"

#include <cstdint>
#include <memory>
#include <string>
#include <vector>

namespace smr
{
    enum ProgramMode : int
    {
        MODE_FUSED  = 0, // one cooperative launch, grid barriers between stages
        MODE_STAGES = 1, // one launch per stage
        MODE_GRAPH  = 2, // MODE_STAGES captured in a hipGraph, replayed while the program is unchanged
        // chains of independent instances: each run of consecutive small stages
        // is one launch with one block per chain (block barriers between the
        // stages), each large stage a grid-wide launch; captured in a hipGraph.
        // Programs whose chains are not declared independent run as MODE_GRAPH.
        MODE_BLOCKS = 3
    };

    struct Tuning
    {
        int block = 256;
        // grid-wide stage launches: consecutive cells per thread (interval chunk
        // length), each wave searching for its segment; 0: one precomputed wave
        // item (64 indices, its first segment known) per wavefront
        int cpt  = 1;
        int mode = MODE_FUSED;
    };

    using Stage = std::vector<dev::Segment>;
    using Chain = std::vector<Stage>;

    // segment of operator `op` with the mesh configuration filled in
    dev::Segment make_seg(int op, const MeshConfig& cfg);
    // segment over cells [c0, c1) of a work list
    dev::Segment list_seg(int op, const MeshConfig& cfg, const DevWorkList& l, int64_t c0, int64_t c1);
    // appends g as a stage of its own when it is not empty
    void push_stage(Chain& ch, const dev::Segment& g);
    // update_ghost_mr of storage u as a chain of stages
    void ghost_chain(Chain& ch, DevMesh& dm, const MeshConfig& cfg, double* u, bool has_bc, double bc);
    // stage k of the result holds stage k of a and of b
    Chain zip(Chain a, const Chain& b);

    // A stage program being assembled on the host and its device copy.
    class ProgramBuf
    {
      public:
        explicit ProgramBuf(Ctx* ctx) : ctx_(ctx) {}
        ~ProgramBuf();
        ProgramBuf(const ProgramBuf&)            = delete;
        ProgramBuf& operator=(const ProgramBuf&) = delete;

        void clear();
        // appends the chains side by side: new stage k holds stage k of every chain
        void append(const std::vector<Chain>& chains);
        // declares that chain i of every append is the same instance and that
        // no chain reads or writes another chain's data (what MODE_BLOCKS needs)
        void set_independent_chains(bool on) { independent_ = on; }
        bool blocks_capable() const { return independent_ && chains_consistent_ && nchains_ > 0; }
        // launches of one run in MODE_BLOCKS (block launches + grid stages)
        int block_plan_launches() const { return int(plan_.size()); }
        // bytes copied with the program; returns their offset, which a segment
        // stores in its thr / tabs pointer with the matching `aux` bit set
        size_t add_aux(const void* src, size_t n);
        size_t nsegments() const { return segs_.size(); }
        int nstages() const { return int(stage_cells_.size()); }
        int64_t max_stage_cells() const;
        int64_t total_cells() const;
        void upload();
        // upload with only the launch data the shape t uses (no wave items
        // unless t.cpt == 0, no block plan unless t.mode == MODE_BLOCKS); the
        // program must then run with t's launch data
        void upload_for(const Tuning& t);
        void run(const Tuning& t);

      private:
        Ctx* ctx_;
        std::vector<dev::Segment> segs_;
        std::vector<int64_t> seg_pref_{0}; // cumulative cells over all segments (+ total)
        std::vector<int32_t> stage_seg_{0};
        std::vector<int64_t> stage_cells_;
        std::vector<int64_t> stage_max_chain_; // per stage: the most cells one chain has in it
        bool items_built_ = false, plan_built_ = false;
        void upload_with(bool items, bool plan);
        std::vector<dev::WaveItem> witems_;     // wave items of all stages, stage by stage
        std::vector<int64_t> stage_item_{0};    // first wave item of each stage (+ total)
        std::vector<int64_t> stage_nitems_;     // wave items per stage
        // per chain: (segment, stage) in program order
        std::vector<std::vector<std::pair<int32_t, int32_t>>> chain_items_;
        int nchains_            = -1;
        bool chains_consistent_ = true;
        bool independent_       = false;
        struct PlanStep
        {
            bool blocks;   // one block per chain over stages [k0, k1), else stage k0 grid-wide
            int k0, k1;
            size_t off;    // blocks: index of the chain offsets in blk_
        };
        std::vector<PlanStep> plan_;
        std::vector<int32_t> blk_; // chain offsets (nchains + 1 per block step), then the items
        size_t blk_items_ = 0;     // index of the first item in blk_
        const int32_t* dblk_ = nullptr;
        std::vector<char> aux_;
        char* dbuf_ = nullptr;
        size_t dcap_ = 0;
        char* hbuf_ = nullptr; // pinned staging
        size_t hcap_ = 0;
        hipEvent_t copied_ = nullptr;
        dev::Program prog_{};
        // graph of the stage-wise launches of the uploaded program
        hipGraphExec_t gexec_ = nullptr;
        int g_block_ = 0, g_cpt_ = 0, g_mode_ = -1;
        void build_block_plan();
        void build_wave_items();
        void launch_block_plan(const Tuning& t, hipStream_t s) const;
        void launch_stage(const Tuning& t, int k, hipStream_t s) const;
        void launch_stagewise(const Tuning& t, hipStream_t s) const;
        bool uploaded_ = false;
        void drop_graph();
    };

    struct BatchInstance
    {
        Field* u    = nullptr;
        Field* unp1 = nullptr;
        double a[3] = {0, 0, 0};
        double dt   = 0;
    };

    class Batch
    {
      public:
        Batch(Ctx* ctx, std::vector<BatchInstance> inst);
        ~Batch();

        // Adapt::operator()(eps, regularity) on every instance (mr/adapt.hpp:147-195),
        // then unp1 is resized on every changed mesh; returns the total number of harten calls
        int adapt(double eps, double regularity);
        // update_ghost_mr(u); unp1 = u - dt * upwind(a, u); swap(u, unp1) on every instance
        void advance();
        // one time step of the advection demo loop (advection_2d.cpp:129-153)
        int step(double eps, double regularity);

        int64_t nb_cells() const;
        const Tuning& tuning() const { return tuning_; }
        void set_tuning(const Tuning& t);
        // times the advance program for each candidate shape and keeps the fastest;
        // persisted per GPU architecture in `cache_file` when not empty
        Tuning autotune(int reps, const std::string& cache_file);
        static bool load_tuning(const std::string& cache_file, Tuning& t);

      private:
        Ctx* ctx_;
        std::vector<BatchInstance> inst_;
        Tuning tuning_;
        ProgramBuf adapt_prog_, resolve_prog_, uf_prog_;
        // advance programs for the two u/unp1 parities, reused while the instances' storages are unchanged
        std::unique_ptr<ProgramBuf> fv_prog_[2];
        std::vector<const void*> fv_key_[2];
        int fv_next_ = 0;
        // per-instance detail (f64) and tag (u8) storage, concatenated
        double* det_arena_ = nullptr;
        uint8_t* tag_arena_ = nullptr;
        uint8_t* host_tags_ = nullptr;
        int* flags_         = nullptr; // per active instance: tags change the cell array
        int* host_flags_    = nullptr;
        int64_t arena_cap_ = 0, host_cap_ = 0;
        // pinned staging of the new images' host parts
        PinnedStaging staging_;
        // the versions the last adaptation replaced (host meshes, images,
        // update_fields lists), freed while the next launches run
        std::vector<std::shared_ptr<void>> replaced_;

        void release_replaced();
        // the launch shape of the programs built for one adaptation pass
        Tuning pass_tuning() const;
        void check_instances();
        void ensure_images(const std::vector<int>& ids);
        void ensure_arenas(int64_t n);
        ProgramBuf& fv_program();
    };
} // namespace smr
