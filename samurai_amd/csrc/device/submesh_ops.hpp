// Bit-level helpers and per-word bodies of the device mesh construction
// (update_sub_mesh_impl on padded per-level bitmaps): submesh.hip runs them
// on one mesh, submesh_batch.hip on many meshes per launch (a thread finds
// its mesh, then calls the same body on that mesh's slice).
#pragma once

#include "device_api.hpp"

#include <hip/hip_runtime.h>

namespace smr::dev::smops
{
        constexpr int SB = 256;

        inline unsigned sm_blocks(int64_t n)
        {
            return unsigned((n + SB - 1) / SB);
        }

        // 64 bits of a row of nw words from bit `start` on (zero outside the row)
        __device__ __forceinline__ uint64_t load64(const uint64_t* row, int64_t nw, int64_t start)
        {
            const int64_t w  = start >> 6;
            const int sh     = int(start & 63);
            const uint64_t a = (w >= 0 && w < nw) ? row[w] : 0ull;
            if (sh == 0)
            {
                return a;
            }
            const uint64_t b = (w + 1 >= 0 && w + 1 < nw) ? row[w + 1] : 0ull;
            return (a >> sh) | (b << (64 - sh));
        }

        // bit i of the result = bit 2i | bit 2i+1 of w (32 bits)
        __device__ __forceinline__ uint64_t pair_or(uint64_t w)
        {
            uint64_t x = (w | (w >> 1)) & 0x5555555555555555ull;
            x          = (x | (x >> 1)) & 0x3333333333333333ull;
            x          = (x | (x >> 2)) & 0x0F0F0F0F0F0F0F0Full;
            x          = (x | (x >> 4)) & 0x00FF00FF00FF00FFull;
            x          = (x | (x >> 8)) & 0x0000FFFF0000FFFFull;
            x          = (x | (x >> 16)) & 0x00000000FFFFFFFFull;
            return x;
        }

        // bits 2i and 2i+1 of the result = bit i of v (the low 32 bits of v)
        __device__ __forceinline__ uint64_t spread32(uint64_t v)
        {
            uint64_t x = v & 0xFFFFFFFFull;
            x          = (x | (x << 16)) & 0x0000FFFF0000FFFFull;
            x          = (x | (x << 8)) & 0x00FF00FF00FF00FFull;
            x          = (x | (x << 4)) & 0x0F0F0F0F0F0F0F0Full;
            x          = (x | (x << 2)) & 0x3333333333333333ull;
            x          = (x | (x << 1)) & 0x5555555555555555ull;
            return x | (x << 1);
        }

        // the 64 fine bits whose parents are the bits of v from bit 0 on, the
        // first fine bit being the second child of its parent when o = 1
        __device__ __forceinline__ uint64_t spread_parents(uint64_t v, int o)
        {
            uint64_t c = spread32(v) >> o;
            if (o)
            {
                c |= ((v >> 32) & 1ull) << 63;
            }
            return c;
        }

        struct PWord
        {
            int l;
            int64_t Y, Z, wx; // padded row, plane and word of the row
        };

        __device__ __forceinline__ PWord pword(const PBits& p, int64_t gw)
        {
            PWord q;
            int l = 0;
            while (l < p.hi && gw >= p.off[l + 1])
            {
                ++l;
            }
            q.l               = l;
            const int64_t r   = gw - p.off[l];
            q.wx              = r % p.wpr[l];
            const int64_t row = r / p.wpr[l];
            q.Y               = p.dim == 3 ? row % p.n[l] : row;
            q.Z               = p.dim == 3 ? row / p.n[l] : 0;
            return q;
        }

        // row (Y, Z) of level l of a set, or null outside the padded level
        __device__ __forceinline__ const uint64_t* prow(const PBits& p, const uint64_t* set, int l, int64_t Y, int64_t Z)
        {
            if (l < 0 || l > p.hi || Y < 0 || Y >= p.n[l] || Z < 0 || (p.dim == 3 ? Z >= p.n[l] : Z != 0))
            {
                return nullptr;
            }
            return set + p.off[l] + (Z * p.n[l] + Y) * p.wpr[l];
        }

        // the bits of word wx of a padded row of level l that lie in the
        // padded row (the last word of a row may hang over its end)
        __device__ __forceinline__ uint64_t tail_mask(const PBits& p, int l, int64_t wx)
        {
            const int64_t rem = p.n[l] - 64 * wx;
            return rem >= 64 ? ~0ull : ((1ull << rem) - 1);
        }

        // bits with unpadded coordinate in [lo, hi) of word wx (padded bit X = x + pad)
        __device__ __forceinline__ uint64_t range_mask(int pad, int64_t wx, int64_t lo, int64_t hi)
        {
            const int64_t a = lo + pad - 64 * wx, b = hi + pad - 64 * wx; // bit range [a, b) of the word
            const int64_t a0 = a < 0 ? 0 : (a > 64 ? 64 : a), b0 = b < 0 ? 0 : (b > 64 ? 64 : b);
            if (b0 <= a0)
            {
                return 0ull;
            }
            const uint64_t hi_m = b0 >= 64 ? ~0ull : ((1ull << b0) - 1);
            const uint64_t lo_m = a0 >= 64 ? 0ull : ~((1ull << a0) - 1);
            return hi_m & lo_m;
        }

        // .on(t) of level t + k for the 64 coarse bits from padded bit X0 of
        // coarse row (Y, Z): OR over the 2^k child rows (planes) and the
        // 2^k-blocks of children of each coarse bit
        __device__ __forceinline__ uint64_t coarse_word(const PBits& p, const uint64_t* set, int t, int k, int64_t Y, int64_t Z, int64_t X0)
        {
            const int lf      = t + k;
            const int64_t f   = int64_t(1) << k;
            const int pad     = p.pad;
            const int64_t yc  = (Y - pad) * f + pad, zc = p.dim == 3 ? (Z - pad) * f + pad : 0;
            const int64_t xc  = (X0 - pad) * f + pad;
            const int nz      = p.dim == 3 ? int(f) : 1;
            uint64_t acc      = 0;
            for (int dz = 0; dz < nz; ++dz)
            {
                for (int dy = 0; dy < int(f); ++dy)
                {
                    const uint64_t* row = prow(p, set, lf, yc + dy, zc + dz);
                    if (!row)
                    {
                        continue;
                    }
                    const int64_t nw = p.wpr[lf];
                    if (k == 1)
                    {
                        acc |= pair_or(load64(row, nw, xc)) | (pair_or(load64(row, nw, xc + 64)) << 32);
                    }
                    else
                    {
                        for (int q = 0; q < 4; ++q)
                        {
                            acc |= pair_or(pair_or(load64(row, nw, xc + 64 * q))) << (16 * q);
                        }
                    }
                }
            }
            return acc;
        }

        // word gw of cells: the leaves of levels [b.lo, b.hi] (unpadded layout) shifted by pad
        __device__ __forceinline__ uint64_t cells_word(const Bits& b, const uint64_t* leaves, const PBits& p, int64_t gw)
        {
            const PWord q   = pword(p, gw);
            const int64_t y = q.Y - p.pad, z = q.Z - p.pad;
            const int64_t n = int64_t(1) << q.l;
            if (q.l >= b.lo && q.l <= b.hi && y >= 0 && y < n && (p.dim == 2 || (z >= 0 && z < n)))
            {
                const uint64_t* row = leaves + b.off[q.l] + ((p.dim == 3 ? z : 0) * n + y) * b.wpr[q.l];
                return load64(row, b.wpr[q.l], 64 * q.wx - p.pad) & range_mask(p.pad, q.wx, 0, n);
            }
            return 0ull;
        }

        // word gw of cells_and_ghosts = nestedExpand(cells, r): OR of the
        // cells within r cells along every axis
        __device__ __forceinline__ uint64_t dilate_word(const PBits& p, const uint64_t* C, int r, int64_t gw)
        {
            const PWord q  = pword(p, gw);
            const int rz   = p.dim == 3 ? r : 0;
            uint64_t acc   = 0;
            for (int dz = -rz; dz <= rz; ++dz)
            {
                for (int dy = -r; dy <= r; ++dy)
                {
                    const uint64_t* row = prow(p, C, q.l, q.Y + dy, q.Z + dz);
                    if (!row)
                    {
                        continue;
                    }
                    for (int dx = -r; dx <= r; ++dx)
                    {
                        acc |= load64(row, p.wpr[q.l], 64 * q.wx + dx);
                    }
                }
            }
            return acc & tail_mask(p, q.l, q.wx);
        }

        // word i of union[t] = (cells[t + 1] | union[t + 1]).on(t)
        __device__ __forceinline__ void union_word(const PBits& p, const uint64_t* C, uint64_t* U, int t, int64_t i)
        {
            const PWord q    = pword(p, p.off[t] + i);
            const uint64_t v = coarse_word(p, C, t, 1, q.Y, q.Z, 64 * q.wx) | coarse_word(p, U, t, 1, q.Y, q.Z, 64 * q.wx);
            U[p.off[t] + i]  = v & tail_mask(p, t, q.wx);
        }

        // union[l - 1] for l = l0 down to 1 by the threads of one block
        // (union[l0] must be clear)
        __device__ __forceinline__ void union_chain(const PBits& p, const uint64_t* C, uint64_t* U, int l0)
        {
            for (int l = l0; l >= 1; --l)
            {
                for (int64_t i = threadIdx.x; i < p.off[l] - p.off[l - 1]; i += blockDim.x)
                {
                    union_word(p, C, U, l - 1, i);
                }
                __syncthreads();
            }
        }

        // word gw of K1[t] = cells_and_ghosts[t + 1].on(t) and K2[t] =
        // cells[t + 2].on(t): the coarse maps the prediction ghosts are dilated from
        __device__ __forceinline__ void coarse_words(const PBits& p, const uint64_t* C, const uint64_t* G, uint64_t* K1, uint64_t* K2, int64_t gw)
        {
            const PWord q    = pword(p, gw);
            const int t      = q.l;
            const uint64_t m = tail_mask(p, t, q.wx);
            K1[gw]           = t + 1 <= p.hi ? coarse_word(p, G, t, 1, q.Y, q.Z, 64 * q.wx) & m : 0ull;
            K2[gw]           = t + 2 <= p.hi ? coarse_word(p, C, t, 2, q.Y, q.Z, 64 * q.wx) & m : 0ull;
        }

        // word gw of the reference before the projection ghosts:
        // cells_and_ghosts[t] | nestedExpand(K1[t], s) clipped to [-s, 2^t + s)
        // | nestedExpand(K2[t], s)
        __device__ __forceinline__ uint64_t reference_word(const PBits& p, const uint64_t* G, const uint64_t* K1, const uint64_t* K2, int s, int mr,
                                                           int64_t gw)
        {
            const PWord q = pword(p, gw);
            const int t   = q.l;
            uint64_t acc  = G[gw];
            if (mr)
            {
                const int sz    = p.dim == 3 ? s : 0;
                const int64_t n = int64_t(1) << t;
                uint64_t b1 = 0, b2 = 0;
                for (int dz = -sz; dz <= sz; ++dz)
                {
                    for (int dy = -s; dy <= s; ++dy)
                    {
                        const uint64_t* r1 = prow(p, K1, t, q.Y + dy, q.Z + dz);
                        const uint64_t* r2 = prow(p, K2, t, q.Y + dy, q.Z + dz);
                        for (int dx = -s; dx <= s; ++dx)
                        {
                            if (r1)
                            {
                                b1 |= load64(r1, p.wpr[t], 64 * q.wx + dx);
                            }
                            if (r2)
                            {
                                b2 |= load64(r2, p.wpr[t], 64 * q.wx + dx);
                            }
                        }
                    }
                }
                // clip_box(..., [-s, 2^t + s)) along every axis
                const int64_t y = q.Y - p.pad, z = q.Z - p.pad;
                const bool in_row = y >= -s && y < n + s && (p.dim == 2 || (z >= -s && z < n + s));
                b1 = in_row ? (b1 & range_mask(p.pad, q.wx, -s, n + s)) : 0ull;
                acc |= b1 | b2;
            }
            return acc & tail_mask(p, t, q.wx);
        }

        // projection step t, word i: proj[t] = reference[t] & union[t] (words
        // of level t), reference[t + 1] |= proj[t].on(t + 1) (words of level t + 1)
        __device__ __forceinline__ void project_word(const PBits& p, uint64_t* R, const uint64_t* U, uint64_t* J, int t, int64_t i)
        {
            const int64_t nwt = p.off[t + 1] - p.off[t];
            const int64_t nwf = p.off[t + 2] - p.off[t + 1];
            if (i < nwt)
            {
                J[p.off[t] + i] = R[p.off[t] + i] & U[p.off[t] + i];
            }
            if (i < nwf)
            {
                const PWord q     = pword(p, p.off[t + 1] + i);
                const int pad     = p.pad;
                // parent row (plane) of the padded row (plane) Y (Z) at level t + 1
                const int64_t Yp  = ((q.Y - pad) >> 1) + pad;
                const int64_t Zp  = p.dim == 3 ? ((q.Z - pad) >> 1) + pad : 0;
                const uint64_t* r = prow(p, R, t, Yp, Zp);
                const uint64_t* u = prow(p, U, t, Yp, Zp);
                if (r && u)
                {
                    // target bit j (unpadded x = 64 wx - pad + j) has the parent (x >> 1)
                    const int64_t x0 = 64 * q.wx - pad;
                    const int64_t p0 = x0 >> 1;
                    const uint64_t v = load64(r, p.wpr[t], p0 + pad) & load64(u, p.wpr[t], p0 + pad);
                    R[p.off[t + 1] + i] |= spread_parents(v, int(x0 & 1)) & tail_mask(p, t + 1, q.wx);
                }
            }
        }

        // set[t - 1].on(t) for the 64 bits of word wx of padded row (Y, Z) at level t
        __device__ __forceinline__ uint64_t refine_word(const PBits& p, const uint64_t* set, int t, int64_t Y, int64_t Z, int64_t wx)
        {
            const int pad     = p.pad;
            const uint64_t* r = prow(p, set, t - 1, ((Y - pad) >> 1) + pad, p.dim == 3 ? ((Z - pad) >> 1) + pad : 0);
            if (!r)
            {
                return 0ull;
            }
            const int64_t x0 = 64 * wx - pad;
            return spread_parents(load64(r, p.wpr[t - 1], (x0 >> 1) + pad), int(x0 & 1));
        }

        // projection steps t = 0 .. t1 - 1 by the threads of one block
        __device__ __forceinline__ void project_chain(const PBits& p, uint64_t* R, const uint64_t* U, uint64_t* J, int t1)
        {
            for (int t = 0; t < t1; ++t)
            {
                const int64_t n = p.off[t + 1] - p.off[t] > p.off[t + 2] - p.off[t + 1] ? p.off[t + 1] - p.off[t] : p.off[t + 2] - p.off[t + 1];
                for (int64_t i = threadIdx.x; i < n; i += blockDim.x)
                {
                    project_word(p, R, U, J, t, i);
                }
                __syncthreads();
            }
        }

        // run starts of word gw of a set (bit set, left neighbour clear)
        __device__ __forceinline__ uint64_t sm_run_starts(const uint64_t* set, int64_t gw, int64_t wx)
        {
            const uint64_t cur   = set[gw];
            const uint64_t carry = wx > 0 ? (set[gw - 1] >> 63) : 0ull;
            return cur & ~((cur << 1) | carry);
        }

        // runs starting in word gw of a set
        __device__ __forceinline__ int run_count_word(const PBits& p, const uint64_t* set, int64_t gw)
        {
            return __popcll(sm_run_starts(set, gw, pword(p, gw).wx));
        }

        // the runs (x0, x1, y, z), unpadded, that start in word gw of a set,
        // from run index k on
        __device__ __forceinline__ void run_emit_word(const PBits& p, const uint64_t* set, int64_t gw, int64_t k, int32_t* runs, int64_t cap)
        {
            const PWord q       = pword(p, gw);
            uint64_t st         = sm_run_starts(set, gw, q.wx);
            const int64_t W     = p.wpr[q.l];
            const int pad       = p.pad;
            while (st)
            {
                const int b   = __ffsll(static_cast<unsigned long long>(st)) - 1;
                st &= st - 1;
                const int64_t X0 = q.wx * 64 + b;
                // first clear bit after X0 in the row (tails past the row are clear)
                int64_t wx     = q.wx;
                uint64_t clear = ~set[gw] & (b == 63 ? 0ull : (~0ull << (b + 1)));
                int64_t X1     = p.n[q.l];
                while (true)
                {
                    if (clear)
                    {
                        X1 = wx * 64 + (__ffsll(static_cast<unsigned long long>(clear)) - 1);
                        break;
                    }
                    ++wx;
                    if (wx >= W)
                    {
                        break;
                    }
                    clear = ~set[gw - q.wx + wx];
                }
                if (X1 > p.n[q.l])
                {
                    X1 = p.n[q.l];
                }
                if (k >= cap)
                {
                    break;
                }
                int32_t* r = runs + SM_RUN * k;
                r[0]       = int32_t(X0 - pad);
                r[1]       = int32_t(X1 - pad);
                r[2]       = int32_t(q.Y - pad);
                r[3]       = int32_t(p.dim == 3 ? q.Z - pad : 0);
                ++k;
            }
        }

        // the first run of level l (levels past hi: the end of the set) of the
        // set whose words start at w0 of the scanned run counts
        __device__ __forceinline__ int32_t segment_start(const PBits& p, int64_t w0, int l, const int32_t* offsets, const int32_t* counts,
                                                         int64_t nw)
        {
            const int64_t w = w0 + p.off[l <= p.hi + 1 ? l : p.hi + 1];
            return w < nw ? offsets[w] : offsets[nw - 1] + counts[nw - 1];
        }
        // ---- many meshes per launch: a prefix over the meshes' padded words
        // (pref[k] = first word of mesh k; times `scale` when every mesh holds
        // `scale` arrays of its words) locates a thread's mesh ----

        static_assert(sizeof(SubmeshItem) % 8 == 0, "items are staged in LDS as 64-bit words");

        // the item holding global index i of the prefix pref (scaled) over n items
        __device__ __forceinline__ int find_item(const int64_t* pref, int n, int64_t i, int64_t scale)
        {
            int lo = 0, hi = n - 1;
            while (lo < hi)
            {
                const int mid = (lo + hi + 1) >> 1;
                if (pref[mid] * scale <= i)
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

        // the item of every index of this block, or -1 when they span two
        __device__ __forceinline__ int block_item(const int64_t* pref, int n, int64_t scale)
        {
            const int64_t end   = pref[n] * scale;
            const int64_t first = int64_t(blockIdx.x) * SB;
            const int64_t last  = (first + SB < end ? first + SB : end) - 1;
            const int a         = find_item(pref, n, first, scale);
            return a == find_item(pref, n, last, scale) ? a : -1;
        }

        constexpr int kItemWords = int(sizeof(SubmeshItem) / 8);

        // copies src into the LDS words dst by the block's threads and returns
        // the copy (read after a barrier)
        __device__ __forceinline__ const SubmeshItem& stage_item(uint64_t* dst, const SubmeshItem& src)
        {
            const uint64_t* s = reinterpret_cast<const uint64_t*>(&src);
            for (int i = threadIdx.x; i < kItemWords; i += blockDim.x)
            {
                dst[i] = s[i];
            }
            return *reinterpret_cast<const SubmeshItem*>(dst);
        }

        // f(item, k, j) for global index blockIdx.x * SB + threadIdx.x < pref[n]
        // * scale: its item k and its index j within that item's scale x
        // words.  A block whose indices all belong to one item (nearly every
        // block) stages that item in LDS; the others find each thread's item
        // by bisection.
        template <class F>
        __device__ __forceinline__ void for_word(const SubmeshItem* items, const int64_t* pref, int n, int64_t scale, F&& f)
        {
            __shared__ uint64_t s_raw[kItemWords];
            const SubmeshItem& s_item = *reinterpret_cast<const SubmeshItem*>(s_raw);
            const int64_t i           = int64_t(blockIdx.x) * SB + threadIdx.x;
            const int kb              = block_item(pref, n, scale);
            if (kb >= 0)
            {
                stage_item(s_raw, items[kb]);
            }
            __syncthreads();
            if (i >= pref[n] * scale)
            {
                return;
            }
            if (kb >= 0)
            {
                f(s_item, kb, i - pref[kb] * scale);
            }
            else
            {
                const int k = find_item(pref, n, i, scale);
                f(items[k], k, i - pref[k] * scale);
            }
        }

        __device__ __forceinline__ int64_t item_words(const SubmeshItem& it)
        {
            return it.p.off[it.p.hi + 1];
        }
} // namespace smr::dev::smops
