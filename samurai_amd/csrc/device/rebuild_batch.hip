// The device mesh update of rebuild.hip (update_cell_array_from_tag and the
// make_graduation fixed point, algorithm/graduation.hpp:573-843) for many
// independent meshes in each launch: the bitmaps of all instances are
// concatenated (instance k from word wpref[k]), and a thread finds its
// instance by bisection over the prefix of words, cells or runs it works on,
// then runs the single-mesh body of bits_ops.hpp on that instance's slice.
// One block per instance runs the single-block cascade.  Flags are per
// instance: kPasses pass flags, then the "leaves changed" flag.
#include "bits_ops.hpp"

#include <hip/hip_runtime.h>
#include <hipcub/hipcub.hpp>

namespace smr::dev
{
    using namespace bitops;

    namespace
    {
        __global__ __launch_bounds__(RB) void k_zero_words_b(uint64_t* p, int64_t n)
        {
            const int64_t i = int64_t(blockIdx.x) * RB + threadIdx.x;
            if (i < n)
            {
                p[i] = 0;
            }
        }

        __global__ __launch_bounds__(RB) void k_combine_b(const uint64_t* old, const uint64_t* add, const uint64_t* rem, uint64_t* out, int64_t n)
        {
            const int64_t i = int64_t(blockIdx.x) * RB + threadIdx.x;
            if (i < n)
            {
                out[i] = (old[i] | add[i]) & ~rem[i];
            }
        }

        // the instance holding global index i of a prefix over n instances
        __device__ __forceinline__ int find_instance(const int64_t* pref, int n, int64_t i)
        {
            int lo = 0, hi = n - 1;
            while (lo < hi)
            {
                const int mid = (lo + hi + 1) >> 1;
                if (pref[mid] <= i)
                {
                    lo = mid;
                }
                else
                {
                    hi = mid - 1;
                }
            }
            return lo;
        }

        __global__ __launch_bounds__(RB) void k_b_bits_from_runs(const MeshItem* items, const int64_t* rpref, int n, uint64_t* bits)
        {
            const int64_t i = int64_t(blockIdx.x) * RB + threadIdx.x;
            if (i >= rpref[n])
            {
                return;
            }
            const int k         = find_instance(rpref, n, i);
            const MeshItem& it  = items[k];
            bits_from_run(it.cells, it.b, bits + it.w0, i - rpref[k]);
        }

        __global__ __launch_bounds__(RB) void k_b_tag_marks(const MeshItem* items, const int64_t* cpref, int n, uint64_t* add, uint64_t* rem)
        {
            const int64_t i = int64_t(blockIdx.x) * RB + threadIdx.x;
            if (i >= cpref[n])
            {
                return;
            }
            const int k        = find_instance(cpref, n, i);
            const MeshItem& it = items[k];
            if (it.dim == 2)
            {
                tag_marks_cell<2>(it.cells, it.b, it.tag, it.min_level, it.max_level, add + it.w0, rem + it.w0, i - cpref[k]);
            }
            else
            {
                tag_marks_cell<3>(it.cells, it.b, it.tag, it.min_level, it.max_level, add + it.w0, rem + it.w0, i - cpref[k]);
            }
        }

        __global__ __launch_bounds__(RB) void k_b_dilate_project(const MeshItem* items, const int64_t* wpref, int n, const uint64_t* ca,
                                                                 uint64_t* p, int w2)
        {
            const int64_t i = int64_t(blockIdx.x) * RB + threadIdx.x;
            if (i >= wpref[n])
            {
                return;
            }
            const int k        = find_instance(wpref, n, i);
            const MeshItem& it = items[k];
            if (it.dim == 2)
            {
                dilate_project_word<2>(it.b, ca + it.w0, p + it.w0, w2, i - wpref[k]);
            }
            else
            {
                dilate_project_word<3>(it.b, ca + it.w0, p + it.w0, w2, i - wpref[k]);
            }
        }

        // one block per instance: the coarse levels of the cascade
        __global__ __launch_bounds__(1024) void k_b_cascade(const MeshItem* items, uint64_t* p)
        {
            const MeshItem& it = items[blockIdx.x];
            if (it.b.hi - 3 < it.b.lo)
            {
                return;
            }
            if (it.dim == 2)
            {
                cascade_levels<2>(it.b, p + it.w0, it.b.hi - 3);
            }
            else
            {
                cascade_levels<3>(it.b, p + it.w0, it.b.hi - 3);
            }
        }

        __global__ __launch_bounds__(RB) void k_b_grad_apply(const MeshItem* items, const int64_t* wpref, int n, const uint64_t* ca,
                                                             const uint64_t* p, uint64_t* out, int* flags, int fstride, int pass)
        {
            const int64_t i = int64_t(blockIdx.x) * RB + threadIdx.x;
            if (i >= wpref[n])
            {
                return;
            }
            const int k        = find_instance(wpref, n, i);
            const MeshItem& it = items[k];
            int* flag          = flags + int64_t(k) * fstride + pass;
            if (it.dim == 2)
            {
                grad_apply_word<2>(it.b, ca + it.w0, p + it.w0, out + it.w0, flag, i - wpref[k]);
            }
            else
            {
                grad_apply_word<3>(it.b, ca + it.w0, p + it.w0, out + it.w0, flag, i - wpref[k]);
            }
        }

        // flags[k * fstride + slot] = 1 where instance k's a and c differ
        __global__ __launch_bounds__(RB) void k_b_differ(const int64_t* wpref, int n, const uint64_t* a, const uint64_t* c, int* flags,
                                                         int fstride, int slot)
        {
            const int64_t i = int64_t(blockIdx.x) * RB + threadIdx.x;
            if (i < wpref[n] && a[i] != c[i])
            {
                flags[int64_t(find_instance(wpref, n, i)) * fstride + slot] = 1;
            }
        }

        // run starts per word, zero for the instances whose flag `slot` is clear
        __global__ __launch_bounds__(RB) void k_b_run_counts(const MeshItem* items, const int64_t* wpref, int n, const uint64_t* bits,
                                                             const int* flags, int fstride, int slot, int32_t* counts)
        {
            const int64_t i = int64_t(blockIdx.x) * RB + threadIdx.x;
            if (i >= wpref[n])
            {
                return;
            }
            const int k        = find_instance(wpref, n, i);
            const MeshItem& it = items[k];
            int32_t c          = 0;
            if (flags[int64_t(k) * fstride + slot])
            {
                c = it.dim == 2 ? run_count_word<2>(it.b, bits + it.w0, i - wpref[k]) : run_count_word<3>(it.b, bits + it.w0, i - wpref[k]);
            }
            counts[i] = c;
        }

        // runs (x0, x1, y, z, level, instance) in (instance, level, z, y, x) order
        __global__ __launch_bounds__(RB) void k_b_run_emit(const MeshItem* items, const int64_t* wpref, int n, const uint64_t* bits,
                                                           const int* flags, int fstride, int slot, const int32_t* offsets, int32_t* runs,
                                                           int64_t cap)
        {
            const int64_t i = int64_t(blockIdx.x) * RB + threadIdx.x;
            if (i >= wpref[n])
            {
                return;
            }
            const int k        = find_instance(wpref, n, i);
            const MeshItem& it = items[k];
            if (!flags[int64_t(k) * fstride + slot])
            {
                return;
            }
            if (it.dim == 2)
            {
                run_emit_word<2>(it.b, bits + it.w0, i - wpref[k], offsets[i], runs, cap, 6, k);
            }
            else
            {
                run_emit_word<3>(it.b, bits + it.w0, i - wpref[k], offsets[i], runs, cap, 6, k);
            }
        }
    } // namespace

    void mesh_batch_zero(uint64_t* p, int64_t n, hipStream_t s)
    {
        if (n > 0)
        {
            hipLaunchKernelGGL(k_zero_words_b, dim3(blocks_for(n)), dim3(RB), 0, s, p, n);
        }
    }

    void mesh_batch_combine(const uint64_t* old, const uint64_t* add, const uint64_t* rem, uint64_t* out, int64_t n, hipStream_t s)
    {
        if (n > 0)
        {
            hipLaunchKernelGGL(k_combine_b, dim3(blocks_for(n)), dim3(RB), 0, s, old, add, rem, out, n);
        }
    }

    void mesh_batch_leaves(const MeshItem* items, const int64_t* rpref, int64_t nruns, int n, uint64_t* bits, hipStream_t s)
    {
        if (nruns > 0)
        {
            hipLaunchKernelGGL(k_b_bits_from_runs, dim3(blocks_for(nruns)), dim3(RB), 0, s, items, rpref, n, bits);
        }
    }

    void mesh_batch_tag_marks(const MeshItem* items, const int64_t* cpref, int64_t ncells, int n, uint64_t* add, uint64_t* rem, hipStream_t s)
    {
        if (ncells > 0)
        {
            hipLaunchKernelGGL(k_b_tag_marks, dim3(blocks_for(ncells)), dim3(RB), 0, s, items, cpref, n, add, rem);
        }
    }

    void mesh_batch_graduation_pass(const MeshItem* items, const int64_t* wpref, int64_t nwords, int n, const uint64_t* ca, uint64_t* p,
                                    uint64_t* out, int w2, int* flags, int fstride, int pass, hipStream_t s)
    {
        if (nwords <= 0)
        {
            return;
        }
        hipLaunchKernelGGL(k_zero_words_b, dim3(blocks_for(nwords)), dim3(RB), 0, s, p, nwords);
        hipLaunchKernelGGL(k_b_dilate_project, dim3(blocks_for(nwords)), dim3(RB), 0, s, items, wpref, n, ca, p, w2);
        hipLaunchKernelGGL(k_b_cascade, dim3(unsigned(n)), dim3(1024), 0, s, items, p);
        hipLaunchKernelGGL(k_b_grad_apply, dim3(blocks_for(nwords)), dim3(RB), 0, s, items, wpref, n, ca, p, out, flags, fstride, pass);
    }

    void mesh_batch_differ(const int64_t* wpref, int64_t nwords, int n, const uint64_t* a, const uint64_t* c, int* flags, int fstride, int slot,
                           hipStream_t s)
    {
        if (nwords > 0)
        {
            hipLaunchKernelGGL(k_b_differ, dim3(blocks_for(nwords)), dim3(RB), 0, s, wpref, n, a, c, flags, fstride, slot);
        }
    }

    void mesh_batch_run_offsets(const MeshItem* items, const int64_t* wpref, int64_t nwords, int n, const uint64_t* bits, const int* flags,
                                int fstride, int slot, int32_t* counts, int32_t* offsets, void* temp, size_t temp_bytes, hipStream_t s)
    {
        if (nwords <= 0)
        {
            return;
        }
        hipLaunchKernelGGL(k_b_run_counts, dim3(blocks_for(nwords)), dim3(RB), 0, s, items, wpref, n, bits, flags, fstride, slot, counts);
        (void)hipcub::DeviceScan::ExclusiveSum(temp, temp_bytes, counts, offsets, int(nwords), s);
    }

    void mesh_batch_run_emit(const MeshItem* items, const int64_t* wpref, int64_t nwords, int n, const uint64_t* bits, const int* flags,
                             int fstride, int slot, const int32_t* offsets, int32_t* runs, int64_t cap, hipStream_t s)
    {
        if (nwords > 0)
        {
            hipLaunchKernelGGL(k_b_run_emit, dim3(blocks_for(nwords)), dim3(RB), 0, s, items, wpref, n, bits, flags, fstride, slot, offsets, runs,
                               cap);
        }
    }

    size_t mesh_batch_scan_temp_bytes(int64_t nwords)
    {
        size_t bytes = 0;
        (void)hipcub::DeviceScan::ExclusiveSum(nullptr, bytes, static_cast<const int32_t*>(nullptr), static_cast<int32_t*>(nullptr),
                                               int(nwords > 0 ? nwords : 1));
        return bytes;
    }
} // namespace smr::dev
