// Fused stage program: many independent MR instances advanced together.
//
// A stage is a set of segments (operator, instance, cell range) with no
// dependencies between them; consecutive stages are ordered.  The fused
// kernel is one cooperative launch for the whole program, with a grid barrier
// between stages; the stage-wise form launches one kernel per stage (for hipGraph
// capture).  Each thread walks the flattened cells of the stage and runs the
// same per-cell operator bodies as the standalone kernels (cell_ops.hpp).
#include "cell_ops.hpp"

#include <hip/hip_cooperative_groups.h>

namespace smr::dev
{
    namespace
    {
        // a cell-list operator on cell c of the list, located at `cell`
        template <int DIM>
        __device__ __forceinline__ void run_located_dim(const Segment& g, int64_t c, const Cell& cell)
        {
            switch (g.op)
            {
                case SEG_PROJECT: op_project_at<DIM>(g.list, c, cell, g.u, g.u); break;
                case SEG_PREDICT_FINE:
                    if (g.pred_s == 0)
                    {
                        op_predict_fine_at<DIM, 0>(g.list, c, cell, g.u);
                    }
                    else
                    {
                        op_predict_fine_at<DIM, 1>(g.list, c, cell, g.u);
                    }
                    break;
                case SEG_DETAIL:
                    if (g.pred_s == 0)
                    {
                        op_detail_at<DIM, 0>(g.list, c, cell, g.u, g.out);
                    }
                    else
                    {
                        op_detail_at<DIM, 1>(g.list, c, cell, g.u, g.out);
                    }
                    break;
                case SEG_CRITERIA: op_criteria_at<DIM>(g.list, c, cell, g.out, g.tag, g.thr, g.min_level, g.max_level); break;
                case SEG_MAXIMUM: op_maximum_at<DIM>(g.list, c, cell, g.tag); break;
                case SEG_TAG_INIT: op_tag_init_at(g.list, c, cell, g.tag); break;
                case SEG_TAG_CHANGE: op_tag_change_at(g.list, c, cell, g.tag, g.min_level, g.max_level, g.err); break;
                case SEG_COPY: op_copy_at(g.list, c, cell, g.out, g.u); break;
                case SEG_PROJECT_TO: op_project_at<DIM>(g.list, c, cell, g.out, g.u); break;
                case SEG_PREDICT_CHILDREN:
                    if (g.pred_s == 0)
                    {
                        op_predict_children_at<DIM, 0>(g.list, c, cell, g.out, g.u);
                    }
                    else
                    {
                        op_predict_children_at<DIM, 1>(g.list, c, cell, g.out, g.u);
                    }
                    break;
                case SEG_CELLS_TO_COMPACT: op_cells_to_compact_at(g.list, c, cell, g.out, g.u); break;
                case SEG_CELLS_FROM_COMPACT: op_cells_from_compact_at(g.list, c, cell, g.out, g.in2); break;
                case SEG_ADJ_UPWIND: op_adj_upwind_at<DIM>(g.list, c, cell, g.u, g.out, g.in2, g.out2, g.a0, g.a1, g.a2, g.dt); break;
                case SEG_ADJ_SCATTER: op_adj_scatter_at<DIM>(g.list, c, cell, g.pad, g.u, g.out, g.a0, g.a1, g.a2, g.dt); break;
                case SEG_UPWIND:
                {
                    const int64_t* b = bases(g.list, cell.run);
                    g.out[b[0] + cell.x] = upwind_value<DIM>(b, cell.x, cell.level, g.u, .5 * g.a0, .5 * fabs(g.a0), .5 * g.a1,
                                                             .5 * fabs(g.a1), .5 * g.a2, .5 * fabs(g.a2), g.dt);
                    break;
                }
                default: break;
            }
        }

        __device__ __forceinline__ void run_located(const Segment& g, int64_t c, const Cell& cell)
        {
            if (g.dim == 2)
            {
                run_located_dim<2>(g, c, cell);
            }
            else
            {
                run_located_dim<3>(g, c, cell);
            }
        }

        __device__ __forceinline__ void run_cell(const Segment& g, int64_t c)
        {
            if (g.op == SEG_RECIPE)
            {
                op_recipe(g.rdst, g.rop, g.rsp, g.rsrc, c, g.u, g.has_bc, g.bc);
            }
            else if (g.op == SEG_RESOLVE)
            {
                if (c < g.list.nruns * g.list.ntaps)
                {
                    op_resolve(g.list, c, g.tabs[0], g.tabs[1], g.err);
                }
                if (g.list.cellrun && c < g.list.ncells)
                {
                    op_cellrun(g.list, c);
                }
            }
            else if (g.op == SEG_FILL_F64)
            {
                g.out[c] = 0.0;
            }
            else if (g.op == SEG_FILL_U8)
            {
                g.tag[c] = 0;
            }
            else if (g.op == SEG_COPY_RAW)
            {
                g.out[c] = g.u[c];
            }
            else if (g.op == SEG_ADJ_GATHER)
            {
                op_adj_gather(g.rdst, g.rsp, g.rsrc, g.thr, c, g.out);
            }
            else if (g.op == SEG_WALK_ROWS)
            {
                op_walk_rows(g.list, c);
            }
            else if (g.op == SEG_ADJ_ZERO)
            {
                g.out[g.rdst[c]] = 0.0;
            }
            else if (is_cell_list_op(g.op))
            {
                run_located(g, c, locate(g.list, c));
            }
        }

        // wave item `it` of a program: 64 consecutive indices of the segments
        // [it.seg, it.end) from index it.off of segment it.seg on
        __device__ __forceinline__ void run_item(const Program& p, WaveItem it, int lane)
        {
            if (it.n > 1)
            {
                // chunks of one segment of a streaming operator
                const Segment& g = p.segs[it.seg];
                const int64_t c  = g.c0 + it.off + lane;
                for (int q = 0; q < it.n; ++q)
                {
                    if (c + 64 * q < g.c1)
                    {
                        run_cell(g, c + 64 * q);
                    }
                }
                return;
            }
            const int64_t pos = p.seg_pref[it.seg] + it.off + lane;
            if (pos >= p.seg_pref[it.end])
            {
                return;
            }
            int32_t sg = it.seg;
            while (p.seg_pref[sg + 1] <= pos)
            {
                ++sg;
            }
            const Segment& g = p.segs[sg];
            run_cell(g, g.c0 + (pos - p.seg_pref[sg]));
        }

        // the last segment s of [sb, se) with pref[s] - base <= w (w uniform
        // over the wave, all 64 lanes active): 64 prefixes per round trip,
        // narrowed by a ballot, instead of one load per bisection step
        __device__ __forceinline__ int32_t find_segment_wave(const int64_t* pref, int32_t sb, int32_t se, int64_t base, int64_t w)
        {
            const int lane = int(threadIdx.x & 63);
            int32_t lo     = sb;
            int32_t n      = se - sb; // pref[lo] - base <= w throughout
            while (n > 64)
            {
                const int32_t step = (n + 63) >> 6;
                const int32_t idx  = lo + lane * step;
                const bool ok      = idx < lo + n && pref[idx] - base <= w;
                const uint64_t m   = __ballot(ok);
                const int32_t nlo  = lo + (63 - __clzll(m)) * step;
                n                  = min(step, lo + n - nlo);
                lo                 = nlo;
            }
            const bool ok    = lane < n && pref[lo + lane] - base <= w;
            const uint64_t m = __ballot(ok);
            return __builtin_amdgcn_readfirstlane(lo + (63 - __clzll(m)));
        }

        // the cells of one stage, grid-strided; CPT consecutive cells per
        // thread.  The waves step through the stage together (the loop bound
        // is the wave's first cell), so the segment search can use the whole
        // wave; each lane then walks forward from the wave's first segment.
        __device__ __forceinline__ void run_stage(const Program& p, int k, int64_t first_thread, int64_t nthreads, int cpt)
        {
            const int32_t sb = p.stage_seg[k], se = p.stage_seg[k + 1];
            if (se <= sb)
            {
                return;
            }
            const int64_t base  = p.seg_pref[sb];
            const int64_t total = p.seg_pref[se] - base;
            const int lane      = int(threadIdx.x & 63);
            for (int64_t w0 = (first_thread - lane) * cpt; w0 < total; w0 += nthreads * cpt)
            {
                int32_t seg      = find_segment_wave(p.seg_pref, sb, se, base, w0);
                const int64_t i0 = w0 + int64_t(lane) * cpt;
                for (int j = 0; j < cpt; ++j)
                {
                    const int64_t i = i0 + j;
                    if (i >= total)
                    {
                        break;
                    }
                    while (seg + 1 < se && p.seg_pref[seg + 1] - base <= i)
                    {
                        ++seg;
                    }
                    const Segment& g = p.segs[seg];
                    run_cell(g, g.c0 + (i - (p.seg_pref[seg] - base)));
                }
            }
        }

        __global__ void k_program_fused(Program p, int cpt)
        {
            cooperative_groups::grid_group grid = cooperative_groups::this_grid();
            const int64_t tid                  = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
            const int64_t nthreads             = int64_t(gridDim.x) * blockDim.x;
            for (int k = 0; k < p.nstages; ++k)
            {
                run_stage(p, k, tid, nthreads, cpt);
                grid.sync();
            }
        }

        // stage k, cpt consecutive cells per thread, each wave searching for
        // its first segment
        __global__ void k_program_stage_cells(Program p, int k, int cpt)
        {
            const int64_t tid      = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
            const int64_t nthreads = int64_t(gridDim.x) * blockDim.x;
            run_stage(p, k, tid, nthreads, cpt);
        }

        // stage k, one wave item per wavefront (blocks of 256 threads)
        __global__ __launch_bounds__(256) void k_program_stage(Program p, int k)
        {
            const int64_t w = int64_t(blockIdx.x) * 4 + __builtin_amdgcn_readfirstlane(int(threadIdx.x >> 6));
            const int64_t i = p.stage_item[k] + w;
            if (i >= p.stage_item[k + 1])
            {
                return;
            }
            run_item(p, p.items[i], int(threadIdx.x & 63));
        }

        // One block per chain: block b runs the segments items[off[b] ..
        // off[b+1]) in order, all its threads on one segment at a time, with
        // a block barrier after the items whose low bit is set (the last
        // segment of the chain in a stage).  The item list is uniform over
        // the block, so the segment descriptors are scalar loads and every
        // barrier is reached by all threads.  Chains must not touch each
        // other's data: blocks are not ordered among themselves.
        __global__ __launch_bounds__(1024) void k_program_blocks(Program p, const int32_t* __restrict__ off,
                                                                 const int32_t* __restrict__ items)
        {
            const int32_t t0   = off[blockIdx.x], t1 = off[blockIdx.x + 1];
            const int wave     = __builtin_amdgcn_readfirstlane(int(threadIdx.x >> 6));
            const int nwaves   = int(blockDim.x >> 6);
            const int lane     = int(threadIdx.x & 63);
            for (int32_t t = t0; t < t1; ++t)
            {
                const int32_t v  = items[t];
                const int32_t sg = v >> 1;
                const Segment& g = p.segs[sg];
                const int32_t n  = int32_t(g.c1 - g.c0);
                for (int32_t q = wave * 64; q < n; q += nwaves * 64)
                {
                    run_item(p, WaveItem{sg, q, sg + 1, 1}, lane);
                }
                if (v & 1)
                {
                    __syncthreads();
                }
            }
        }
    } // namespace

    void run_program_blocks(const Program& p, const int32_t* off, const int32_t* items, int nchains, int threads, hipStream_t s)
    {
        if (nchains > 0)
        {
            hipLaunchKernelGGL(k_program_blocks, dim3(unsigned(nchains)), dim3(unsigned(threads)), 0, s, p, off, items);
        }
    }

    void run_program_stage(const Program& p, int k, int64_t nitems, hipStream_t s)
    {
        if (nitems <= 0)
        {
            return;
        }
        hipLaunchKernelGGL(k_program_stage, dim3(unsigned((nitems + 3) / 4)), dim3(256), 0, s, p, k);
    }

    int program_max_blocks(int block)
    {
        int per_cu = 0;
        if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, k_program_fused, block, 0) != hipSuccess || per_cu <= 0)
        {
            return 0;
        }
        int dev = 0, ncu = 0;
        (void)hipGetDevice(&dev);
        if (hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess)
        {
            return 0;
        }
        return per_cu * ncu;
    }

    bool run_program_fused(const Program& p, int block, int nblocks, int cpt, hipStream_t s)
    {
        if (p.nstages <= 0)
        {
            return true;
        }
        Program prog = p;
        void* args[] = {&prog, &cpt};
        return hipLaunchCooperativeKernel(reinterpret_cast<const void*>(k_program_fused), dim3(nblocks), dim3(block), args, 0, s) ==
               hipSuccess;
    }

    void run_program_stage_cells(const Program& p, int k, int64_t cells, int block, int cpt, hipStream_t s)
    {
        if (cells <= 0)
        {
            return;
        }
        const int64_t threads = (cells + cpt - 1) / cpt;
        const unsigned grid   = unsigned((threads + block - 1) / block);
        hipLaunchKernelGGL(k_program_stage_cells, dim3(grid), dim3(block), 0, s, p, k, cpt);
    }
} // namespace smr::dev
