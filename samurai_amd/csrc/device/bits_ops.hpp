// Bit-level helpers and per-word (per-cell, per-run) bodies of the device
// mesh update on per-level dense bitmaps (rebuild.hip for one mesh,
// rebuild_batch.hip for many): each body works on one mesh given its layout
// and its arrays, so a batched kernel finds its instance and calls the same
// body on that instance's slice.
#pragma once

#include "device_api.hpp"

#include <hip/hip_runtime.h>

namespace smr::dev::bitops
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

        // the bits of run r of a cells work list
        __device__ __forceinline__ void bits_from_run(const List& l, const Bits& b, uint64_t* bits, int64_t r)
        {
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
        __device__ __forceinline__ void tag_marks_cell(const List& l, const Bits& b, const uint8_t* tag, int min_level, int max_level, uint64_t* add,
                                                       uint64_t* rem, int64_t c)
        {
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

        // list_interval_to_refine_for_graduation (graduation.hpp:245-330),
        // first half: every word of nestedExpand(ca[f], w2) (box dilation
        // clipped to the domain) at a level f in [lo + 2, hi] is projected two
        // levels down (.on(f - 2): OR over 4x4 blocks) and ORed into p[f - 2]
        // (p zeroed beforehand; at most 16 fine words share a target word)
        template <int DIM>
        __device__ __forceinline__ void dilate_project_word(const Bits& b, const uint64_t* ca, uint64_t* p, int w2, int64_t gw)
        {
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

        // levels c0 down to lo of the cascade, by the calling block
        template <int DIM>
        __device__ __forceinline__ void cascade_levels(const Bits& b, uint64_t* p, int c0)
        {
            for (int c = c0; c >= b.lo; --c)
            {
                const int64_t n0 = b.off[c], n1 = b.off[c + 1];
                for (int64_t gw = n0 + threadIdx.x; gw < n1; gw += blockDim.x)
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
        __device__ __forceinline__ void grad_apply_word(const Bits& b, const uint64_t* ca, const uint64_t* p, uint64_t* out, int* flag, int64_t gw)
        {
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

        // runs (x0, x1, y, z, level) in (level, z, y, x) order: each run
        // start found in word w is written at offsets[w] + its rank in w; its
        // end is the first clear bit after it in the row
        // (with stride 6, the id `extra` in the sixth field of each run)
        template <int DIM>
        __device__ __forceinline__ void run_emit_word(const Bits& b, const uint64_t* bits, int64_t gw, int64_t k, int32_t* runs, int64_t cap,
                                                      int stride, int32_t extra)
        {
            const WordPos p  = word_pos(b, DIM, gw);
            uint64_t st      = run_starts(bits, gw, p.wx);
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
                int32_t* r = runs + stride * int64_t(k);
                r[0]       = x0;
                r[1]       = x1;
                r[2]       = p.y;
                r[3]       = p.z;
                r[4]       = p.l;
                if (stride == 6)
                {
                    r[5] = extra;
                }
                ++k;
            }
        }
        // run starts in word gw
        template <int DIM>
        __device__ __forceinline__ int run_count_word(const Bits& b, const uint64_t* bits, int64_t gw)
        {
            const WordPos p = word_pos(b, DIM, gw);
            return __popcll(run_starts(bits, gw, p.wx));
        }
} // namespace smr::dev::bitops
