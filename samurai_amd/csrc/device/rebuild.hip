// Device half of the mesh update of a harten call: update_cell_array_from_tag
// (algorithm/graduation.hpp:743-843) and the make_graduation fixed point
// (:573-726, list_interval_to_refine_for_graduation :245-330) on per-level
// dense bitmaps, then the run-length encoding of the new leaves.
//
// A level l of the unit box has 2^l cells per axis; its bitmap stores each
// row (y, or (z, y) in 3D) in max(1, 2^l / 64) 64-bit words, x ascending from
// bit 0 (a row shorter than 64 cells uses the low bits of one word).  The
// leaf sets of all levels are small next to the HBM (2 MB for a 4096^2
// level), so every set operation of the reference becomes a word-wise
// operation: union |, intersection &, difference & ~, the box dilation of
// nestedExpand a few shifts and row ORs, and .on(l - 1) a pairwise OR with
// bit compaction.
#include "bits_ops.hpp"

#include <hip/hip_runtime.h>
#include <hipcub/hipcub.hpp>

namespace smr::dev
{
    using namespace bitops;

    // ------------------------------------------------------------------
    // kernels (the bodies are in bits_ops.hpp)
    // ------------------------------------------------------------------
    namespace
    {
        __global__ __launch_bounds__(RB) void k_zero_words(uint64_t* p, int64_t n)
        {
            const int64_t i = int64_t(blockIdx.x) * RB + threadIdx.x;
            if (i < n)
            {
                p[i] = 0;
            }
        }

        // the leaves of a cells work list as bits: one thread per run
        __global__ __launch_bounds__(RB) void k_bits_from_runs(List l, Bits b, uint64_t* bits)
        {
            const int64_t r = int64_t(blockIdx.x) * RB + threadIdx.x;
            if (r < l.nruns)
            {
                bits_from_run(l, b, bits, r);
            }
        }

        template <int DIM>
        __global__ __launch_bounds__(RB) void k_tag_marks(List l, Bits b, const uint8_t* tag, int min_level, int max_level, uint64_t* add,
                                                         uint64_t* rem)
        {
            const int64_t c = int64_t(blockIdx.x) * RB + threadIdx.x;
            if (c < l.ncells)
            {
                tag_marks_cell<DIM>(l, b, tag, min_level, max_level, add, rem, c);
            }
        }

        // new = (old | add) & ~rem (graduation.hpp:832-840)
        __global__ __launch_bounds__(RB) void k_combine(const uint64_t* old, const uint64_t* add, const uint64_t* rem, uint64_t* out,
                                                       int64_t n)
        {
            const int64_t i = int64_t(blockIdx.x) * RB + threadIdx.x;
            if (i < n)
            {
                out[i] = (old[i] | add[i]) & ~rem[i];
            }
        }

        template <int DIM>
        __global__ __launch_bounds__(RB) void k_dilate_project(Bits b, const uint64_t* ca, uint64_t* p, int w2)
        {
            const int64_t w = int64_t(blockIdx.x) * RB + threadIdx.x;
            if (w < b.off[b.hi + 1] - b.off[b.lo])
            {
                dilate_project_word<DIM>(b, ca, p, w2, b.off[b.lo] + w);
            }
        }

        // one level c of the cascade, grid-wide (the large levels)
        template <int DIM>
        __global__ __launch_bounds__(RB) void k_project_cascade_level(Bits b, uint64_t* p, int c)
        {
            const int64_t gw = b.off[c] + int64_t(blockIdx.x) * RB + threadIdx.x;
            if (gw < b.off[c + 1])
            {
                cascade_word<DIM>(b, p, c, gw);
            }
        }

        // levels c0 down to lo of the cascade by one block
        template <int DIM>
        __global__ __launch_bounds__(1024) void k_project_cascade(Bits b, uint64_t* p, int c0)
        {
            cascade_levels<DIM>(b, p, c0);
        }

        template <int DIM>
        __global__ __launch_bounds__(RB) void k_grad_apply(Bits b, const uint64_t* ca, const uint64_t* p, uint64_t* out, int* flag)
        {
            const int64_t w = int64_t(blockIdx.x) * RB + threadIdx.x;
            if (w < b.off[b.hi + 1] - b.off[b.lo])
            {
                grad_apply_word<DIM>(b, ca, p, out, flag, b.off[b.lo] + w);
            }
        }

        // *flag = 1 where a and b differ
        __global__ __launch_bounds__(RB) void k_differ(const uint64_t* a, const uint64_t* c, int64_t n, int* flag)
        {
            const int64_t i = int64_t(blockIdx.x) * RB + threadIdx.x;
            if (i < n && a[i] != c[i])
            {
                flag[0] = 1;
            }
        }

        template <int DIM>
        __global__ __launch_bounds__(RB) void k_run_counts(Bits b, const uint64_t* bits, int32_t* counts)
        {
            const int64_t w = int64_t(blockIdx.x) * RB + threadIdx.x;
            if (w < b.off[b.hi + 1] - b.off[b.lo])
            {
                counts[w] = run_count_word<DIM>(b, bits, b.off[b.lo] + w);
            }
        }

        // runs (x0, x1, y, z, level) in (level, z, y, x) order: each run
        // start found in word w is written at offsets[w] + its rank in w
        template <int DIM>
        __global__ __launch_bounds__(RB) void k_run_emit(Bits b, const uint64_t* bits, const int32_t* offsets, int32_t* runs, int64_t cap)
        {
            const int64_t w = int64_t(blockIdx.x) * RB + threadIdx.x;
            if (w < b.off[b.hi + 1] - b.off[b.lo])
            {
                run_emit_word<DIM>(b, bits, b.off[b.lo] + w, offsets[w], runs, cap, 5, 0);
            }
        }
    } // namespace

    // ------------------------------------------------------------------
    // launchers
    // ------------------------------------------------------------------
    void bits_zero(const Bits& b, uint64_t* bits, hipStream_t s)
    {
        const int64_t n = b.off[b.hi + 1];
        if (n > 0)
        {
            hipLaunchKernelGGL(k_zero_words, dim3(blocks_for(n)), dim3(RB), 0, s, bits, n);
        }
    }

    void bits_from_cells(const List& l, const Bits& b, uint64_t* bits, hipStream_t s)
    {
        bits_zero(b, bits, s);
        if (l.nruns > 0)
        {
            hipLaunchKernelGGL(k_bits_from_runs, dim3(blocks_for(l.nruns)), dim3(RB), 0, s, l, b, bits);
        }
    }

    void bits_tag_marks(int dim, const List& l, const Bits& b, const uint8_t* tag, int min_level, int max_level, uint64_t* add, uint64_t* rem,
                        hipStream_t s)
    {
        bits_zero(b, add, s);
        bits_zero(b, rem, s);
        if (l.ncells > 0)
        {
            if (dim == 2)
            {
                hipLaunchKernelGGL(k_tag_marks<2>, dim3(blocks_for(l.ncells)), dim3(RB), 0, s, l, b, tag, min_level, max_level, add, rem);
            }
            else
            {
                hipLaunchKernelGGL(k_tag_marks<3>, dim3(blocks_for(l.ncells)), dim3(RB), 0, s, l, b, tag, min_level, max_level, add, rem);
            }
        }
    }

    void bits_combine(const Bits& b, const uint64_t* old, const uint64_t* add, const uint64_t* rem, uint64_t* out, hipStream_t s)
    {
        const int64_t n = b.off[b.hi + 1];
        if (n > 0)
        {
            hipLaunchKernelGGL(k_combine, dim3(blocks_for(n)), dim3(RB), 0, s, old, add, rem, out, n);
        }
    }

    void bits_graduation_pass(int dim, const Bits& b, const uint64_t* ca, uint64_t* p, uint64_t* out, int w2, int* flag, hipStream_t s)
    {
        const int64_t n = b.off[b.hi + 1] - b.off[b.lo];
        if (n <= 0)
        {
            return;
        }
        bits_zero(b, p, s);
        // the cascade: a grid-wide launch per level of more than kSmall words,
        // then one block for the rest
        constexpr int64_t kSmall = 1024;
        int c = b.hi - 3;
        if (dim == 2)
        {
            hipLaunchKernelGGL(k_dilate_project<2>, dim3(blocks_for(n)), dim3(RB), 0, s, b, ca, p, w2);
            for (; c >= b.lo && b.off[c + 1] - b.off[c] > kSmall; --c)
            {
                hipLaunchKernelGGL(k_project_cascade_level<2>, dim3(blocks_for(b.off[c + 1] - b.off[c])), dim3(RB), 0, s, b, p, c);
            }
            if (c >= b.lo)
            {
                hipLaunchKernelGGL(k_project_cascade<2>, dim3(1), dim3(1024), 0, s, b, p, c);
            }
            hipLaunchKernelGGL(k_grad_apply<2>, dim3(blocks_for(n)), dim3(RB), 0, s, b, ca, p, out, flag);
        }
        else
        {
            hipLaunchKernelGGL(k_dilate_project<3>, dim3(blocks_for(n)), dim3(RB), 0, s, b, ca, p, w2);
            for (; c >= b.lo && b.off[c + 1] - b.off[c] > kSmall; --c)
            {
                hipLaunchKernelGGL(k_project_cascade_level<3>, dim3(blocks_for(b.off[c + 1] - b.off[c])), dim3(RB), 0, s, b, p, c);
            }
            if (c >= b.lo)
            {
                hipLaunchKernelGGL(k_project_cascade<3>, dim3(1), dim3(1024), 0, s, b, p, c);
            }
            hipLaunchKernelGGL(k_grad_apply<3>, dim3(blocks_for(n)), dim3(RB), 0, s, b, ca, p, out, flag);
        }
    }

    void bits_differ(const Bits& b, const uint64_t* a, const uint64_t* c, int* flag, hipStream_t s)
    {
        const int64_t n = b.off[b.hi + 1];
        if (n > 0)
        {
            hipLaunchKernelGGL(k_differ, dim3(blocks_for(n)), dim3(RB), 0, s, a, c, n, flag);
        }
    }

    size_t bits_scan_temp_bytes(const Bits& b)
    {
        const int64_t n = b.off[b.hi + 1] - b.off[b.lo];
        size_t bytes    = 0;
        (void)hipcub::DeviceScan::ExclusiveSum(nullptr, bytes, static_cast<const int32_t*>(nullptr), static_cast<int32_t*>(nullptr),
                                               int(n > 0 ? n : 1));
        return bytes;
    }

    void bits_run_offsets(int dim, const Bits& b, const uint64_t* bits, int32_t* counts, int32_t* offsets, void* temp, size_t temp_bytes,
                          hipStream_t s)
    {
        const int64_t n = b.off[b.hi + 1] - b.off[b.lo];
        if (n <= 0)
        {
            return;
        }
        if (dim == 2)
        {
            hipLaunchKernelGGL(k_run_counts<2>, dim3(blocks_for(n)), dim3(RB), 0, s, b, bits, counts);
        }
        else
        {
            hipLaunchKernelGGL(k_run_counts<3>, dim3(blocks_for(n)), dim3(RB), 0, s, b, bits, counts);
        }
        (void)hipcub::DeviceScan::ExclusiveSum(temp, temp_bytes, counts, offsets, int(n), s);
    }

    void bits_run_emit(int dim, const Bits& b, const uint64_t* bits, const int32_t* offsets, int32_t* runs, int64_t cap, hipStream_t s)
    {
        const int64_t n = b.off[b.hi + 1] - b.off[b.lo];
        if (n <= 0)
        {
            return;
        }
        if (dim == 2)
        {
            hipLaunchKernelGGL(k_run_emit<2>, dim3(blocks_for(n)), dim3(RB), 0, s, b, bits, offsets, runs, cap);
        }
        else
        {
            hipLaunchKernelGGL(k_run_emit<3>, dim3(blocks_for(n)), dim3(RB), 0, s, b, bits, offsets, runs, cap);
        }
    }
} // namespace smr::dev
