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
#include "device_api.hpp"

#include <hip/hip_runtime.h>
#include <hipcub/hipcub.hpp>

namespace smr::dev
{
    namespace
    {
        constexpr int RB = 256;

        inline unsigned blocks_for(int64_t n)
        {
            return unsigned((n + RB - 1) / RB);
        }

        __device__ __forceinline__ int64_t row_words(const Bits& b, int l)
        {
            return b.wpr[l];
        }

        // word index of cell x of row (y, z) at level l
        __device__ __forceinline__ int64_t word_of(const Bits& b, int l, int32_t x, int32_t y, int32_t z)
        {
            const int64_t n = int64_t(1) << l;
            return b.off[l] + ((int64_t(z) * n + y) * b.wpr[l]) + (x >> 6);
        }

        __device__ __forceinline__ void set_bit(uint64_t* bits, const Bits& b, int l, int32_t x, int32_t y, int32_t z)
        {
            atomicOr(reinterpret_cast<unsigned long long*>(bits + word_of(b, l, x, y, z)), 1ull << (x & 63));
        }

        // level of global word index w (levels [b.lo, b.hi])
        __device__ __forceinline__ int level_of_word(const Bits& b, int64_t w)
        {
            int l = b.lo;
            while (l < b.hi && w >= b.off[l + 1])
            {
                ++l;
            }
            return l;
        }

        // valid bits of a row word at level l (rows of fewer than 64 cells)
        __device__ __forceinline__ uint64_t row_mask(int l)
        {
            return l >= 6 ? ~0ull : ((1ull << (1 << l)) - 1);
        }

        // bit i of the result = bit 2i | bit 2i+1 of w (32 bits)
        __device__ __forceinline__ uint64_t pair_or_compress(uint64_t w)
        {
            uint64_t x = (w | (w >> 1)) & 0x5555555555555555ull;
            x          = (x | (x >> 1)) & 0x3333333333333333ull;
            x          = (x | (x >> 2)) & 0x0F0F0F0F0F0F0F0Full;
            x          = (x | (x >> 4)) & 0x00FF00FF00FF00FFull;
            x          = (x | (x >> 8)) & 0x0000FFFF0000FFFFull;
            x          = (x | (x >> 16)) & 0x00000000FFFFFFFFull;
            return x;
        }

        // bits 2i and 2i+1 of the result = bit i of v (v: 32 bits)
        __device__ __forceinline__ uint64_t spread(uint64_t v)
        {
            uint64_t x = v & 0xFFFFFFFFull;
            x          = (x | (x << 16)) & 0x0000FFFF0000FFFFull;
            x          = (x | (x << 8)) & 0x00FF00FF00FF00FFull;
            x          = (x | (x << 4)) & 0x0F0F0F0F0F0F0F0Full;
            x          = (x | (x << 2)) & 0x3333333333333333ull;
            x          = (x | (x << 1)) & 0x5555555555555555ull;
            return x | (x << 1);
        }

        // word wx of row (y, z) at level l, decoded from a global word index
        struct WordPos
        {
            int l;
            int32_t y, z;
            int64_t wx;
        };

        __device__ __forceinline__ WordPos word_pos(const Bits& b, int dim, int64_t w)
        {
            WordPos p;
            p.l             = level_of_word(b, w);
            const int64_t r = w - b.off[p.l];
            const int64_t n = int64_t(1) << p.l;
            p.wx            = r % b.wpr[p.l];
            const int64_t row = r / b.wpr[p.l];
            p.y             = int32_t(dim == 3 ? row % n : row);
            p.z             = int32_t(dim == 3 ? row / n : 0);
            return p;
        }

        // .on(l) of level l + 1 for word wx of row (y, z) at level l: OR over
        // the child rows and the child pairs of each bit
        template <int DIM>
        __device__ __forceinline__ uint64_t on_parent_word(const Bits& b, const uint64_t* src, int l, int32_t y, int32_t z, int64_t wx)
        {
            const int lc     = l + 1;
            const int64_t cw = b.wpr[lc];
            uint64_t acc     = 0;
            const int nz     = DIM == 3 ? 2 : 1;
            for (int dz = 0; dz < nz; ++dz)
            {
                for (int dy = 0; dy < 2; ++dy)
                {
                    const int64_t base = word_of(b, lc, 0, 2 * y + dy, DIM == 3 ? 2 * z + dz : 0);
                    const uint64_t c0  = src[base + 2 * wx];
                    const uint64_t c1  = (2 * wx + 1 < cw) ? src[base + 2 * wx + 1] : 0ull;
                    acc |= pair_or_compress(c0) | (pair_or_compress(c1) << 32);
                }
            }
            return acc;
        }

        // one bit per run: does word w of row (y, z) at level l start a run
        // (a set bit whose left neighbour in the row is clear)
        __device__ __forceinline__ uint64_t run_starts(const uint64_t* bits, int64_t w, int64_t wx)
        {
            const uint64_t cur   = bits[w];
            const uint64_t carry = wx > 0 ? (bits[w - 1] >> 63) : 0ull;
            return cur & ~((cur << 1) | carry);
        }
    } // namespace

    // ------------------------------------------------------------------
    // kernels
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
            if (r >= l.nruns)
            {
                return;
            }
            const int32_t x0 = l.iv[5 * r + 0], x1 = l.iv[5 * r + 1], y = l.iv[5 * r + 2], z = l.iv[5 * r + 3];
            const int lev    = l.iv[5 * r + 4];
            for (int32_t x = x0; x < x1;)
            {
                const int32_t wend = (x | 63) + 1;
                const int32_t e    = x1 < wend ? x1 : wend;
                const int nb       = e - x;
                const uint64_t m   = (nb == 64 ? ~0ull : ((1ull << nb) - 1)) << (x & 63);
                atomicOr(reinterpret_cast<unsigned long long*>(bits + word_of(b, lev, x, y, z)), m);
                x = e;
            }
        }

        // update_cell_array_from_tag's per-cell rules (graduation.hpp:783-820):
        // refine below max_level -> remove the cell, add its 2^dim children;
        // coarsen and not keep above min_level -> remove the cell, and the
        // child of even coordinates adds the parent
        template <int DIM>
        __global__ __launch_bounds__(RB) void k_tag_marks(List l, Bits b, const uint8_t* tag, int min_level, int max_level, uint64_t* add,
                                                         uint64_t* rem)
        {
            const int64_t c = int64_t(blockIdx.x) * RB + threadIdx.x;
            if (c >= l.ncells)
            {
                return;
            }
            // the run of cell c (binary search between the chunk hints)
            const int64_t j = c >> HINT_SHIFT;
            int64_t lo = l.hint[j], hi = l.hint[j + 1];
            while (lo < hi)
            {
                const int64_t mid = (lo + hi + 1) >> 1;
                if (l.pref[mid] <= c)
                {
                    lo = mid;
                }
                else
                {
                    hi = mid - 1;
                }
            }
            const int32_t x  = l.iv[5 * lo + 0] + int32_t(c - l.pref[lo]);
            const int32_t y  = l.iv[5 * lo + 2];
            const int32_t z  = l.iv[5 * lo + 3];
            const int lev    = l.iv[5 * lo + 4];
            const uint8_t t  = tag[l.base[lo * l.ntaps] + x];
            const bool refine            = (t & 4) != 0;
            const bool coarsen_not_keep  = (t & 2) && !(t & 1);
            if (refine && lev < max_level)
            {
                set_bit(rem, b, lev, x, y, z);
                const int nz = DIM == 3 ? 2 : 1;
                for (int dz = 0; dz < nz; ++dz)
                {
                    for (int dy = 0; dy < 2; ++dy)
                    {
                        const int32_t cx = 2 * x;
                        atomicOr(reinterpret_cast<unsigned long long*>(add + word_of(b, lev + 1, cx, 2 * y + dy, DIM == 3 ? 2 * z + dz : 0)),
                                 3ull << (cx & 63));
                    }
                }
            }
            else if (coarsen_not_keep && lev > min_level)
            {
                const bool even = (x % 2 == 0) && (y % 2 == 0) && (DIM < 3 || z % 2 == 0);
                if (even)
                {
                    set_bit(add, b, lev - 1, x >> 1, y >> 1, z >> 1);
                }
                set_bit(rem, b, lev, x, y, z);
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

        // list_interval_to_refine_for_graduation (graduation.hpp:245-330),
        // first half: every word of nestedExpand(ca[f], w2) (box dilation
        // clipped to the domain) at a level f in [lo + 2, hi] is projected two
        // levels down (.on(f - 2): OR over 4x4 blocks) and ORed into p[f - 2]
        // (p zeroed beforehand; at most 16 fine words share a target word)
        template <int DIM>
        __global__ __launch_bounds__(RB) void k_dilate_project(Bits b, const uint64_t* ca, uint64_t* p, int w2)
        {
            const int64_t w = int64_t(blockIdx.x) * RB + threadIdx.x;
            if (w >= b.off[b.hi + 1] - b.off[b.lo])
            {
                return;
            }
            const int64_t gw = b.off[b.lo] + w;
            const WordPos q  = word_pos(b, DIM, gw);
            if (q.l < b.lo + 2)
            {
                return;
            }
            const int32_t n = int32_t(1) << q.l;
            const int64_t W = b.wpr[q.l];
            uint64_t acc    = 0;
            const int zlo = DIM == 3 ? max(0, q.z - w2) : 0, zhi = DIM == 3 ? min(n - 1, q.z + w2) : 0;
            const int ylo = max(0, q.y - w2), yhi = min(n - 1, q.y + w2);
            for (int zz = zlo; zz <= zhi; ++zz)
            {
                for (int yy = ylo; yy <= yhi; ++yy)
                {
                    const int64_t base = word_of(b, q.l, 0, yy, zz);
                    const uint64_t cur = ca[base + q.wx];
                    const uint64_t prv = q.wx > 0 ? ca[base + q.wx - 1] : 0ull;
                    const uint64_t nxt = q.wx + 1 < W ? ca[base + q.wx + 1] : 0ull;
                    uint64_t r         = cur;
                    for (int s = 1; s <= w2; ++s)
                    {
                        r |= (cur << s) | (prv >> (64 - s));
                        r |= (cur >> s) | (nxt << (64 - s));
                    }
                    acc |= r;
                }
            }
            acc &= row_mask(q.l);
            if (acc == 0)
            {
                return;
            }
            const uint64_t m     = pair_or_compress(pair_or_compress(acc));
            const int32_t cx     = int32_t(q.wx * 64) >> 2;
            uint64_t* target     = p + word_of(b, q.l - 2, cx, q.y >> 2, q.z >> 2);
            const uint64_t bits  = m << (cx & 63);
            if ((__atomic_load_n(target, __ATOMIC_RELAXED) & bits) != bits)
            {
                atomicOr(reinterpret_cast<unsigned long long*>(target), bits);
            }
        }

        // second half, by one block: p[c] |= p[c + 1].on(c) from c = hi - 3
        // down to lo, a level at a time (these levels hold 1/64 of the words
        // of the finest one and less), so that p[c] = union over f >= c + 2 of
        // the dilated fine cells of level f projected on c
        template <int DIM>
        __device__ __forceinline__ void cascade_word(const Bits& b, uint64_t* p, int c, int64_t gw)
        {
            const WordPos q  = word_pos(b, DIM, gw);
            const uint64_t v = on_parent_word<DIM>(b, p, c, q.y, q.z, q.wx);
            if (v)
            {
                p[gw] |= v;
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
            for (int c = c0; c >= b.lo; --c)
            {
                const int64_t n0 = b.off[c], n1 = b.off[c + 1];
                for (int64_t gw = n0 + threadIdx.x; gw < n1; gw += 1024)
                {
                    cascade_word<DIM>(b, p, c, gw);
                }
                __syncthreads();
            }
        }

        // one graduation pass (graduation.hpp:629-718): the cells of level c
        // under p[c] are replaced by their children:
        // out[l] = (ca[l] | children(rem[l-1])) & ~rem[l], rem[c] = p[c] & ca[c];
        // *flag = 1 if any cell is refined
        template <int DIM>
        __global__ __launch_bounds__(RB) void k_grad_apply(Bits b, const uint64_t* ca, const uint64_t* p, uint64_t* out, int* flag)
        {
            const int64_t w = int64_t(blockIdx.x) * RB + threadIdx.x;
            if (w >= b.off[b.hi + 1] - b.off[b.lo])
            {
                return;
            }
            const int64_t gw   = b.off[b.lo] + w;
            const WordPos q    = word_pos(b, DIM, gw);
            const uint64_t c   = ca[gw];
            const uint64_t rem = p[gw] & c;
            uint64_t add       = 0;
            if (q.l - 1 >= b.lo)
            {
                // the refined parents of this word at l - 1
                const int pl      = q.l - 1;
                const int64_t pw  = word_of(b, pl, 0, q.y >> 1, q.z >> 1) + (q.wx >> 1);
                const uint64_t pr = p[pw] & ca[pw];
                add               = spread(pr >> (32 * (q.wx & 1)));
            }
            out[gw] = (c | add) & ~rem;
            if (rem != 0)
            {
                flag[0] = 1;
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
            if (w >= b.off[b.hi + 1] - b.off[b.lo])
            {
                return;
            }
            const int64_t gw = b.off[b.lo] + w;
            const WordPos p  = word_pos(b, DIM, gw);
            counts[w]        = __popcll(run_starts(bits, gw, p.wx));
        }

        // runs (x0, x1, y, z, level) in (level, z, y, x) order: each run
        // start found in word w is written at offsets[w] + its rank in w; its
        // end is the first clear bit after it in the row
        template <int DIM>
        __global__ __launch_bounds__(RB) void k_run_emit(Bits b, const uint64_t* bits, const int32_t* offsets, int32_t* runs, int64_t cap)
        {
            const int64_t w = int64_t(blockIdx.x) * RB + threadIdx.x;
            if (w >= b.off[b.hi + 1] - b.off[b.lo])
            {
                return;
            }
            const int64_t gw = b.off[b.lo] + w;
            const WordPos p  = word_pos(b, DIM, gw);
            uint64_t st      = run_starts(bits, gw, p.wx);
            int32_t k        = offsets[w];
            const int64_t W  = b.wpr[p.l];
            const int32_t n  = int32_t(1) << p.l;
            while (st)
            {
                const int i    = __ffsll(static_cast<unsigned long long>(st)) - 1;
                st &= st - 1;
                const int32_t x0 = int32_t(p.wx * 64 + i);
                // first clear bit at or after x0 + 1 in the row
                int64_t wx      = p.wx;
                uint64_t clear  = ~bits[gw] & (i == 63 ? 0ull : (~0ull << (i + 1)));
                int32_t x1      = -1;
                while (true)
                {
                    if (clear)
                    {
                        x1 = int32_t(wx * 64 + (__ffsll(static_cast<unsigned long long>(clear)) - 1));
                        break;
                    }
                    ++wx;
                    if (wx >= W)
                    {
                        break;
                    }
                    clear = ~bits[gw - p.wx + wx];
                }
                if (x1 < 0 || x1 > n)
                {
                    x1 = n;
                }
                if (k >= cap)
                {
                    break;
                }
                int32_t* r = runs + 5 * int64_t(k);
                r[0]       = x0;
                r[1]       = x1;
                r[2]       = p.y;
                r[3]       = p.z;
                r[4]       = p.l;
                ++k;
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
