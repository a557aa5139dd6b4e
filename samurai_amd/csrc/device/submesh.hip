// Device half of MRMesh::update_sub_mesh_impl (mr/mesh.hpp:221-455) with
// Mesh_base::construct_union (mesh.hpp:1231-1262), non-periodic, on per-level
// dense bitmaps padded by `pad` cells on every side (the ghosts of a level
// reach at most max(stencil radius, prediction radius) cells outside the
// box).  From the leaves of every level it builds the five mesh ids
//
//   cells_and_ghosts[l] = nestedExpand(cells[l], r)
//   union[l - 1]        = (cells[l] u union[l]).on(l - 1)
//   reference[t]        = cells_and_ghosts[t]
//                       u clip(nestedExpand(cells_and_ghosts[t + 1].on(t), s), [-s, 2^t + s))
//                       u nestedExpand(cells[t + 2].on(t), s)
//   then for t = 0, 1, ...:  proj[t] = reference[t] n union[t],
//                            reference[t + 1] u= proj[t].on(t + 1)
//
// and encodes them as runs (x0, x1, y, z, level, id) for the host, which
// renumbers the reference.  Each set is a word-wise operation: dilation by
// shifted loads and row ORs, .on(l - k) by OR-compressing bit pairs over the
// child rows, .on(l + 1) by spreading bits.
#include "submesh_ops.hpp"

#include <hip/hip_runtime.h>
#include <hipcub/hipcub.hpp>

#include <algorithm>

namespace smr::dev
{
    using namespace smops;

    namespace
    {
        __global__ __launch_bounds__(SB) void k_sm_zero(uint64_t* p, int64_t n)
        {
            const int64_t i = int64_t(blockIdx.x) * SB + threadIdx.x;
            if (i < n)
            {
                p[i] = 0;
            }
        }

        __global__ __launch_bounds__(SB) void k_sm_cells(Bits b, const uint64_t* leaves, PBits p, uint64_t* C)
        {
            const int64_t gw = int64_t(blockIdx.x) * SB + threadIdx.x;
            if (gw < p.off[p.hi + 1])
            {
                C[gw] = cells_word(b, leaves, p, gw);
            }
        }

        __global__ __launch_bounds__(SB) void k_sm_dilate(PBits p, const uint64_t* C, uint64_t* G, int r)
        {
            const int64_t gw = int64_t(blockIdx.x) * SB + threadIdx.x;
            if (gw < p.off[p.hi + 1])
            {
                G[gw] = dilate_word(p, C, r, gw);
            }
        }

        // union[l - 1] for one level l
        __global__ __launch_bounds__(SB) void k_sm_union_level(PBits p, const uint64_t* C, uint64_t* U, int l)
        {
            const int64_t i = int64_t(blockIdx.x) * SB + threadIdx.x;
            if (i < p.off[l] - p.off[l - 1])
            {
                union_word(p, C, U, l - 1, i);
            }
        }

        // union[l - 1] for l = l0 down to 1, by one block (small levels)
        __global__ __launch_bounds__(1024) void k_sm_union_cascade(PBits p, const uint64_t* C, uint64_t* U, int l0)
        {
            union_chain(p, C, U, l0);
        }

        __global__ __launch_bounds__(SB) void k_sm_coarse(PBits p, const uint64_t* C, const uint64_t* G, uint64_t* K1, uint64_t* K2)
        {
            const int64_t gw = int64_t(blockIdx.x) * SB + threadIdx.x;
            if (gw < p.off[p.hi + 1])
            {
                coarse_words(p, C, G, K1, K2, gw);
            }
        }

        __global__ __launch_bounds__(SB) void k_sm_reference(PBits p, const uint64_t* G, const uint64_t* K1, const uint64_t* K2, uint64_t* R, int s,
                                                             int mr)
        {
            const int64_t gw = int64_t(blockIdx.x) * SB + threadIdx.x;
            if (gw < p.off[p.hi + 1])
            {
                R[gw] = reference_word(p, G, K1, K2, s, mr, gw);
            }
        }

        __global__ __launch_bounds__(SB) void k_sm_project_step(PBits p, uint64_t* R, const uint64_t* U, uint64_t* J, int t)
        {
            project_word(p, R, U, J, t, int64_t(blockIdx.x) * SB + threadIdx.x);
        }

        // projection steps t = 0 .. t1 - 1 by one block (small levels)
        __global__ __launch_bounds__(1024) void k_sm_project_cascade(PBits p, uint64_t* R, const uint64_t* U, uint64_t* J, int t1)
        {
            project_chain(p, R, U, J, t1);
        }

        __global__ __launch_bounds__(SB) void k_sm_run_counts(PBits p, int nsets, const uint64_t* sets, int32_t* counts)
        {
            const int64_t total = p.off[p.hi + 1];
            const int64_t i     = int64_t(blockIdx.x) * SB + threadIdx.x;
            if (i < total * nsets)
            {
                counts[i] = run_count_word(p, sets + (i / total) * total, i % total);
            }
        }

        // runs (x0, x1, y, z), unpadded, in (id, level, z, y, x) order
        __global__ __launch_bounds__(SB) void k_sm_run_emit(PBits p, int nsets, const uint64_t* sets, const int32_t* offsets, int32_t* runs,
                                                            int64_t cap)
        {
            const int64_t total = p.off[p.hi + 1];
            const int64_t i     = int64_t(blockIdx.x) * SB + threadIdx.x;
            if (i < total * nsets)
            {
                run_emit_word(p, sets + (i / total) * total, i % total, offsets[i], runs, cap);
            }
        }

        __global__ __launch_bounds__(SB) void k_sm_segments(PBits p, int nsets, const int32_t* offsets, const int32_t* counts, int32_t* seg)
        {
            const int i = int(blockIdx.x) * SB + threadIdx.x;
            if (i < nsets * SM_SEG)
            {
                const int64_t total = p.off[p.hi + 1];
                seg[i]              = segment_start(p, (i / SM_SEG) * total, i % SM_SEG, offsets, counts, total * nsets);
            }
        }
    } // namespace

    void submesh_build(const Bits& b, const uint64_t* leaves, const PBits& p, int r, int s, bool mr, uint64_t* sets, uint64_t* scratch,
                       hipStream_t st)
    {
        const int64_t total = p.off[p.hi + 1];
        if (total <= 0)
        {
            return;
        }
        uint64_t* C = sets + SM_CELLS * total;
        uint64_t* G = sets + SM_CELLS_AND_GHOSTS * total;
        uint64_t* J = sets + SM_PROJ_CELLS * total;
        uint64_t* U = sets + SM_UNION_CELLS * total;
        uint64_t* R = sets + SM_REFERENCE * total;
        hipLaunchKernelGGL(k_sm_cells, dim3(sm_blocks(total)), dim3(SB), 0, st, b, leaves, p, C);
        hipLaunchKernelGGL(k_sm_dilate, dim3(sm_blocks(total)), dim3(SB), 0, st, p, C, G, r);
        // the union chain from the finest level down (union[hi] is empty): a
        // launch per large level, then one block for the small ones
        constexpr int64_t kSmall = 1024; // words of a level a single block takes
        hipLaunchKernelGGL(k_sm_zero, dim3(sm_blocks(total)), dim3(SB), 0, st, U, total);
        int l = p.hi;
        for (; l >= 1 && p.off[l] - p.off[l - 1] > kSmall; --l)
        {
            const int64_t n = p.off[l] - p.off[l - 1];
            hipLaunchKernelGGL(k_sm_union_level, dim3(sm_blocks(n)), dim3(SB), 0, st, p, C, U, l);
        }
        if (l >= 1)
        {
            hipLaunchKernelGGL(k_sm_union_cascade, dim3(1), dim3(1024), 0, st, p, C, U, l);
        }
        uint64_t* K1 = scratch;
        uint64_t* K2 = scratch + total;
        if (mr)
        {
            hipLaunchKernelGGL(k_sm_coarse, dim3(sm_blocks(total)), dim3(SB), 0, st, p, C, G, K1, K2);
        }
        hipLaunchKernelGGL(k_sm_reference, dim3(sm_blocks(total)), dim3(SB), 0, st, p, G, K1, K2, R, s, mr ? 1 : 0);
        hipLaunchKernelGGL(k_sm_zero, dim3(sm_blocks(total)), dim3(SB), 0, st, J, total);
        if (mr)
        {
            // the projection chain from level 0 up: one block for the steps
            // between small levels, then a launch per step
            int t1 = 0;
            while (t1 < p.hi && std::max(p.off[t1 + 1] - p.off[t1], p.off[t1 + 2] - p.off[t1 + 1]) <= kSmall)
            {
                ++t1;
            }
            if (t1 > 0)
            {
                hipLaunchKernelGGL(k_sm_project_cascade, dim3(1), dim3(1024), 0, st, p, R, U, J, t1);
            }
            for (int t = t1; t < p.hi; ++t)
            {
                const int64_t n = std::max(p.off[t + 1] - p.off[t], p.off[t + 2] - p.off[t + 1]);
                hipLaunchKernelGGL(k_sm_project_step, dim3(sm_blocks(n)), dim3(SB), 0, st, p, R, U, J, t);
            }
        }
    }

    size_t submesh_scan_temp_bytes(int64_t n)
    {
        size_t bytes = 0;
        (void)hipcub::DeviceScan::ExclusiveSum(nullptr, bytes, static_cast<const int32_t*>(nullptr), static_cast<int32_t*>(nullptr),
                                               int(n > 0 ? n : 1));
        return bytes;
    }

    void submesh_run_offsets(const PBits& p, int nsets, const uint64_t* sets, int32_t* counts, int32_t* offsets, void* temp, size_t temp_bytes,
                             hipStream_t s)
    {
        const int64_t n = p.off[p.hi + 1] * nsets;
        if (n <= 0)
        {
            return;
        }
        hipLaunchKernelGGL(k_sm_run_counts, dim3(sm_blocks(n)), dim3(SB), 0, s, p, nsets, sets, counts);
        (void)hipcub::DeviceScan::ExclusiveSum(temp, temp_bytes, counts, offsets, int(n), s);
    }

    void submesh_run_emit(const PBits& p, int nsets, const uint64_t* sets, const int32_t* offsets, int32_t* runs, int64_t cap, hipStream_t s)
    {
        const int64_t n = p.off[p.hi + 1] * nsets;
        if (n > 0)
        {
            hipLaunchKernelGGL(k_sm_run_emit, dim3(sm_blocks(n)), dim3(SB), 0, s, p, nsets, sets, offsets, runs, cap);
        }
    }

    void submesh_segments(const PBits& p, int nsets, const int32_t* offsets, const int32_t* counts, int32_t* seg, hipStream_t s)
    {
        if (p.off[p.hi + 1] * nsets > 0)
        {
            hipLaunchKernelGGL(k_sm_segments, dim3(sm_blocks(nsets * SM_SEG)), dim3(SB), 0, s, p, nsets, offsets, counts, seg);
        }
    }
} // namespace smr::dev
