// The device mesh construction of submesh.hip (update_sub_mesh_impl,
// mr/mesh.hpp:221-455) for many independent meshes in each launch: the
// padded sets of all meshes are concatenated (mesh k's sets from word
// SM_NSETS * pw0 on), and the single-mesh bodies of submesh_ops.hpp run on
// each mesh's slice.  A block whose words all belong to one mesh (nearly
// every block: a mesh holds hundreds of words or more) stages that mesh's
// layout in LDS; a block that straddles two meshes finds each thread's mesh
// by bisection over the prefix of padded words.  The two level chains (union
// from the finest level down, projection from level 0 up) run as one block
// per mesh, with block barriers between the levels.
#include "submesh_ops.hpp"

#include <hip/hip_runtime.h>
#include <hipcub/hipcub.hpp>

namespace smr::dev
{
    using namespace smops;

    namespace
    {
        constexpr int CB = 1024; // threads of a chain block

        __device__ __forceinline__ uint64_t* item_set(const SubmeshItem& it, uint64_t* sets, int id)
        {
            return sets + SM_NSETS * it.pw0 + int64_t(id) * item_words(it);
        }

        __global__ __launch_bounds__(SB) void k_bs_cells(const SubmeshItem* items, const int64_t* ppref, int n, const uint64_t* leaves,
                                                         uint64_t* sets)
        {
            for_word(items, ppref, n, 1,
                     [&](const SubmeshItem& it, int, int64_t gw) { item_set(it, sets, SM_CELLS)[gw] = cells_word(it.b, leaves + it.lw0, it.p, gw); });
        }

        __global__ __launch_bounds__(SB) void k_bs_dilate(const SubmeshItem* items, const int64_t* ppref, int n, uint64_t* sets)
        {
            for_word(items, ppref, n, 1,
                     [&](const SubmeshItem& it, int, int64_t gw)
                     { item_set(it, sets, SM_CELLS_AND_GHOSTS)[gw] = dilate_word(it.p, item_set(it, sets, SM_CELLS), it.r, gw); });
        }

        // one block per item: clear union[hi], then union[l - 1] for l = hi .. 1
        __global__ __launch_bounds__(CB) void k_bs_union(const SubmeshItem* items, uint64_t* sets)
        {
            __shared__ uint64_t s_raw[kItemWords];
            const SubmeshItem& it = stage_item(s_raw, items[blockIdx.x]);
            __syncthreads();
            const PBits& p = it.p;
            uint64_t* U    = item_set(it, sets, SM_UNION_CELLS);
            for (int64_t i = p.off[p.hi] + threadIdx.x; i < p.off[p.hi + 1]; i += blockDim.x)
            {
                U[i] = 0;
            }
            __syncthreads();
            union_chain(p, item_set(it, sets, SM_CELLS), U, p.hi);
        }

        __global__ __launch_bounds__(SB) void k_bs_coarse(const SubmeshItem* items, const int64_t* ppref, int n, uint64_t* sets, uint64_t* scratch)
        {
            for_word(items, ppref, n, 1,
                     [&](const SubmeshItem& it, int, int64_t gw)
                     {
                         if (it.mr)
                         {
                             uint64_t* K1 = scratch + 2 * it.pw0;
                             coarse_words(it.p, item_set(it, sets, SM_CELLS), item_set(it, sets, SM_CELLS_AND_GHOSTS), K1, K1 + item_words(it), gw);
                         }
                     });
        }

        __global__ __launch_bounds__(SB) void k_bs_reference(const SubmeshItem* items, const int64_t* ppref, int n, uint64_t* sets,
                                                             const uint64_t* scratch)
        {
            for_word(items, ppref, n, 1,
                     [&](const SubmeshItem& it, int, int64_t gw)
                     {
                         const uint64_t* K1 = scratch + 2 * it.pw0;
                         item_set(it, sets, SM_REFERENCE)[gw] =
                             reference_word(it.p, item_set(it, sets, SM_CELLS_AND_GHOSTS), K1, K1 + item_words(it), it.s, it.mr, gw);
                     });
        }

        // one block per item: clear proj, then the projection steps t = 0 .. hi - 1
        __global__ __launch_bounds__(CB) void k_bs_project(const SubmeshItem* items, uint64_t* sets)
        {
            __shared__ uint64_t s_raw[kItemWords];
            const SubmeshItem& it = stage_item(s_raw, items[blockIdx.x]);
            __syncthreads();
            uint64_t* J = item_set(it, sets, SM_PROJ_CELLS);
            for (int64_t i = threadIdx.x; i < item_words(it); i += blockDim.x)
            {
                J[i] = 0;
            }
            __syncthreads();
            if (it.mr)
            {
                project_chain(it.p, item_set(it, sets, SM_REFERENCE), item_set(it, sets, SM_UNION_CELLS), J, it.p.hi);
            }
        }

        // run counts of word gw of the item's SM_NSETS sets, at the word's
        // index in every set (counts are laid out like the sets)
        __global__ __launch_bounds__(SB) void k_bs_run_counts(const SubmeshItem* items, const int64_t* ppref, int n, const uint64_t* sets,
                                                              int32_t* counts)
        {
            for_word(items, ppref, n, 1,
                     [&](const SubmeshItem& it, int, int64_t gw)
                     {
                         const int64_t total = item_words(it), w0 = SM_NSETS * it.pw0;
                         for (int id = 0; id < SM_NSETS; ++id)
                         {
                             counts[w0 + id * total + gw] = run_count_word(it.p, sets + w0 + id * total, gw);
                         }
                     });
        }

        __global__ __launch_bounds__(SB) void k_bs_run_emit(const SubmeshItem* items, const int64_t* ppref, int n, const uint64_t* sets,
                                                            const int32_t* offsets, int32_t* runs, int64_t cap)
        {
            for_word(items, ppref, n, 1,
                     [&](const SubmeshItem& it, int, int64_t gw)
                     {
                         const int64_t total = item_words(it), w0 = SM_NSETS * it.pw0;
                         for (int id = 0; id < SM_NSETS; ++id)
                         {
                             const int64_t w = w0 + id * total + gw;
                             run_emit_word(it.p, sets + w0 + id * total, gw, offsets[w], runs, cap);
                         }
                     });
        }

        // one thread per (item, set, level)
        __global__ __launch_bounds__(SB) void k_bs_segments(const SubmeshItem* items, const int64_t* ppref, int n, const int32_t* offsets,
                                                            const int32_t* counts, int32_t* seg)
        {
            const int64_t i = int64_t(blockIdx.x) * SB + threadIdx.x;
            if (i >= int64_t(n) * SM_NSETS * SM_SEG)
            {
                return;
            }
            const int k           = int(i / (SM_NSETS * SM_SEG));
            const int id          = int(i / SM_SEG % SM_NSETS);
            const SubmeshItem& it = items[k];
            seg[i] = segment_start(it.p, SM_NSETS * it.pw0 + id * item_words(it), int(i % SM_SEG), offsets, counts, SM_NSETS * ppref[n]);
        }
    } // namespace

    void submesh_batch_build(const SubmeshItem* items, const int64_t* ppref, int64_t pwords, int n, const uint64_t* leaves, uint64_t* sets,
                             uint64_t* scratch, hipStream_t st)
    {
        if (pwords <= 0 || n <= 0)
        {
            return;
        }
        const dim3 g(sm_blocks(pwords));
        hipLaunchKernelGGL(k_bs_cells, g, dim3(SB), 0, st, items, ppref, n, leaves, sets);
        hipLaunchKernelGGL(k_bs_dilate, g, dim3(SB), 0, st, items, ppref, n, sets);
        hipLaunchKernelGGL(k_bs_union, dim3(unsigned(n)), dim3(CB), 0, st, items, sets);
        hipLaunchKernelGGL(k_bs_coarse, g, dim3(SB), 0, st, items, ppref, n, sets, scratch);
        hipLaunchKernelGGL(k_bs_reference, g, dim3(SB), 0, st, items, ppref, n, sets, scratch);
        hipLaunchKernelGGL(k_bs_project, dim3(unsigned(n)), dim3(CB), 0, st, items, sets);
    }

    void submesh_batch_run_offsets(const SubmeshItem* items, const int64_t* ppref, int64_t pwords, int n, const uint64_t* sets, int32_t* counts,
                                   int32_t* offsets, void* temp, size_t temp_bytes, hipStream_t st)
    {
        if (pwords <= 0 || n <= 0)
        {
            return;
        }
        hipLaunchKernelGGL(k_bs_run_counts, dim3(sm_blocks(pwords)), dim3(SB), 0, st, items, ppref, n, sets, counts);
        (void)hipcub::DeviceScan::ExclusiveSum(temp, temp_bytes, counts, offsets, int(SM_NSETS * pwords), st);
    }

    void submesh_batch_run_emit(const SubmeshItem* items, const int64_t* ppref, int64_t pwords, int n, const uint64_t* sets, const int32_t* offsets,
                                int32_t* runs, int64_t cap, hipStream_t st)
    {
        if (pwords > 0 && n > 0)
        {
            hipLaunchKernelGGL(k_bs_run_emit, dim3(sm_blocks(pwords)), dim3(SB), 0, st, items, ppref, n, sets, offsets, runs, cap);
        }
    }

    void submesh_batch_segments(const SubmeshItem* items, const int64_t* ppref, int n, const int32_t* offsets, const int32_t* counts, int32_t* seg,
                                hipStream_t st)
    {
        if (n > 0)
        {
            hipLaunchKernelGGL(k_bs_segments, dim3(sm_blocks(int64_t(n) * SM_NSETS * SM_SEG)), dim3(SB), 0, st, items, ppref, n, offsets, counts,
                               seg);
        }
    }
} // namespace smr::dev
