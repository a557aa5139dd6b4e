// The work lists of build_programs (the program sets of the ghost update, the
// adaptation and update_fields) on the device, from the padded per-level
// bitmaps of the mesh ids that submesh.hip builds.  Each list is, level by
// level, a word-wise expression of the ids (.on(l - 1) by OR-compressing bit
// pairs of the child rows, .on(l + 1) by spreading the parent bits):
//
//   projection  t in [0, max)        reference[t + 1].on(t) n proj_cells[t]
//   prediction  l in [1, max]        (reference[l] \ (cells[l] u proj_cells[l])) n domain n reference[l - 1].on(l)
//   detail      l in [l0, max)       reference[l] n (cells[l + 1] u cells[l + 2]).on(l)
//   criteria    t in [max(1, min) - 1, max)   cells[t + 1].on(t) n reference[t]
//   maximum     t in [0, max)        the same
//   copy        l in [min, max]      old reference[l] n cells[l]
//   coarsen     t in [min, max)      old cells[t + 1].on(t) n cells[t]
//   refine      t in [min, max)      cells[t + 1].on(t) n old cells[t]
//
// (mr/mesh.hpp, numeric/projection.hpp, numeric/prediction.hpp,
// mr/operators.hpp, algorithm/update_fields.hpp: the same sets as the host's
// ghost_task, adapt_task and uf_task), with l0 = max(min - 1, 0).  The runs
// of every list come out in (level, z, y, x) order with their cell prefix
// and the run holding every 2^HINT_SHIFT-th cell.
#include "submesh_ops.hpp"

#include <hip/hip_runtime.h>
#include <hipcub/hipcub.hpp>

namespace smr::dev
{
    using namespace smops;

    namespace
    {
        // the run levels of list k
        __device__ __forceinline__ bool list_has_level(int k, int t, int minl, int maxl)
        {
            switch (k)
            {
                case DL_PROJECTION: return t < maxl;
                case DL_PREDICTION: return t >= 1 && t <= maxl;
                case DL_DETAIL: return t >= (minl > 0 ? minl - 1 : 0) && t < maxl;
                case DL_CRITERIA: return t >= (minl > 1 ? minl : 1) - 1 && t < maxl;
                case DL_MAXIMUM: return t < maxl;
                case DL_UF_COPY: return t >= minl && t <= maxl;
                default: return t >= minl && t < maxl; // coarsen, refine
            }
        }

        // word gw of list k
        __device__ __forceinline__ uint64_t list_word(const PBits& p, int minl, int maxl, const uint64_t* sets, const uint64_t* old_sets, int k,
                                                      int64_t gw)
        {
            const PWord q = pword(p, gw);
            const int t   = q.l;
            if (!list_has_level(k, t, minl, maxl))
            {
                return 0ull;
            }
            const int64_t total = p.off[p.hi + 1];
            const uint64_t* C   = sets + SM_CELLS * total;
            const uint64_t* J   = sets + SM_PROJ_CELLS * total;
            const uint64_t* R   = sets + SM_REFERENCE * total;
            const int64_t X0    = 64 * q.wx;
            uint64_t v          = 0;
            switch (k)
            {
                case DL_PROJECTION: v = coarse_word(p, R, t, 1, q.Y, q.Z, X0) & J[gw]; break;
                case DL_PREDICTION:
                {
                    // clip to the domain [0, 2^t) on every axis
                    const int64_t n = int64_t(1) << t;
                    const int64_t y = q.Y - p.pad, z = q.Z - p.pad;
                    if (y < 0 || y >= n || (p.dim == 3 && (z < 0 || z >= n)))
                    {
                        return 0ull;
                    }
                    v = R[gw] & ~(C[gw] | J[gw]) & range_mask(p.pad, q.wx, 0, n) & refine_word(p, R, t, q.Y, q.Z, q.wx);
                    break;
                }
                case DL_DETAIL: v = R[gw] & (coarse_word(p, C, t, 1, q.Y, q.Z, X0) | coarse_word(p, C, t, 2, q.Y, q.Z, X0)); break;
                case DL_CRITERIA:
                case DL_MAXIMUM: v = coarse_word(p, C, t, 1, q.Y, q.Z, X0) & R[gw]; break;
                case DL_UF_COPY: v = old_sets[SM_REFERENCE * total + gw] & C[gw]; break;
                case DL_UF_COARSEN: v = coarse_word(p, old_sets + SM_CELLS * total, t, 1, q.Y, q.Z, X0) & C[gw]; break;
                default: v = coarse_word(p, C, t, 1, q.Y, q.Z, X0) & old_sets[SM_CELLS * total + gw]; break;
            }
            return v & tail_mask(p, t, q.wx);
        }

        __global__ __launch_bounds__(SB) void k_list_words(PBits p, int minl, int maxl, const uint64_t* sets, const uint64_t* old_sets, int nlists,
                                                           uint64_t* bits)
        {
            const int64_t total = p.off[p.hi + 1];
            const int64_t i     = int64_t(blockIdx.x) * SB + threadIdx.x;
            if (i < total * nlists)
            {
                bits[i] = list_word(p, minl, maxl, sets, old_sets, int(i / total), i % total);
            }
        }

        // runs starting in the word << 32 | cells of the word
        __global__ __launch_bounds__(SB) void k_list_counts(PBits p, int nlists, const uint64_t* bits, uint64_t* counts)
        {
            const int64_t total = p.off[p.hi + 1];
            const int64_t i     = int64_t(blockIdx.x) * SB + threadIdx.x;
            if (i < total * nlists)
            {
                const int64_t gw = i % total;
                const uint64_t* set = bits + (i / total) * total;
                counts[i] = (uint64_t(run_count_word(p, set, gw)) << 32) | uint64_t(__popcll(set[gw]));
            }
        }

        __global__ __launch_bounds__(SB) void k_list_table(PBits p, int nlists, const uint64_t* counts, const uint64_t* scan, uint64_t* table)
        {
            const int i = int(blockIdx.x) * SB + threadIdx.x;
            if (i >= nlists * SM_SEG)
            {
                return;
            }
            const int64_t total = p.off[p.hi + 1], n = total * nlists;
            const int l         = i % SM_SEG;
            const int64_t w     = (i / SM_SEG) * total + p.off[l <= p.hi + 1 ? l : p.hi + 1];
            table[i]            = w < n ? scan[w] : scan[n - 1] + counts[n - 1];
        }

        // the runs of list o that start in word gw of its bitmap set (scan_i:
        // the packed scan at that word); the thread of word 0 writes the list's
        // trailer (the cell count and the last hint)
        __device__ __forceinline__ void emit_list_word(const PBits& p, const uint64_t* set, int64_t gw, uint64_t scan_i, const ListOut& o)
        {
            const int64_t nchunks = (o.ncells + (int64_t(1) << HINT_SHIFT) - 1) >> HINT_SHIFT;
            if (gw == 0)
            {
                o.pref[o.nruns] = o.ncells;
                o.hint[nchunks] = int32_t(o.nruns > 0 ? o.nruns - 1 : 0);
            }
            const PWord q = pword(p, gw);
            uint64_t st   = sm_run_starts(set, gw, q.wx);
            if (!st)
            {
                return;
            }
            int64_t r        = int64_t(scan_i >> 32) - o.run0;
            const int64_t c0 = int64_t(scan_i & 0xFFFFFFFFull) - o.cell0;
            const int64_t W  = p.wpr[q.l];
            const int pad    = p.pad;
            while (st)
            {
                const int b = __ffsll(static_cast<unsigned long long>(st)) - 1;
                st &= st - 1;
                const int64_t X0 = q.wx * 64 + b;
                // first clear bit after X0 in the row (the bits past the row are clear)
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
                int32_t* v = o.iv + 5 * r;
                v[0]       = int32_t(X0 - pad);
                v[1]       = int32_t(X1 - pad);
                v[2]       = int32_t(q.Y - pad);
                v[3]       = int32_t(p.dim == 3 ? q.Z - pad : 0);
                v[4]       = q.l;
                // cells before this run: those of the words before, then the
                // bits of this word below the run's first bit
                const int64_t a = c0 + __popcll(set[gw] & ((1ull << b) - 1));
                o.pref[r]       = a;
                const int64_t e = a + (X1 - X0);
                for (int64_t j = (a + (int64_t(1) << HINT_SHIFT) - 1) >> HINT_SHIFT; (j << HINT_SHIFT) < e; ++j)
                {
                    o.hint[j] = int32_t(r);
                }
                ++r;
            }
        }

        __global__ __launch_bounds__(SB) void k_list_emit(PBits p, int nlists, const uint64_t* bits, const uint64_t* scan, ListOuts outs)
        {
            const int64_t total = p.off[p.hi + 1];
            const int64_t i     = int64_t(blockIdx.x) * SB + threadIdx.x;
            if (i < total * nlists)
            {
                const int k = int(i / total);
                emit_list_word(p, bits + int64_t(k) * total, i % total, scan[i], outs.l[k]);
            }
        }

        // ---- the lists of many meshes per launch: mesh k's DL_N bitmaps from
        // word DL_N * pw0 of the concatenated bitmaps (list j from DL_N * pw0 +
        // j * words), its ids and old ids from word SM_NSETS * pw0 ----

        __global__ __launch_bounds__(SB) void k_blist_words(const SubmeshItem* items, const int64_t* pp, int n, const uint64_t* sets,
                                                            const uint64_t* old_sets, uint64_t* bits)
        {
            for_word(items, pp, n, DL_N,
                     [&](const SubmeshItem& it, int, int64_t j)
                     {
                         const int64_t total = item_words(it);
                         bits[DL_N * it.pw0 + j] = list_word(it.p, it.minl, it.p.hi, sets + SM_NSETS * it.pw0, old_sets + SM_NSETS * it.pw0,
                                                             int(j / total), j % total);
                     });
        }

        __global__ __launch_bounds__(SB) void k_blist_counts(const SubmeshItem* items, const int64_t* pp, int n, const uint64_t* bits,
                                                             uint64_t* counts)
        {
            for_word(items, pp, n, DL_N,
                     [&](const SubmeshItem& it, int, int64_t j)
                     {
                         const int64_t total = item_words(it), gw = j % total;
                         const uint64_t* set = bits + DL_N * it.pw0 + (j / total) * total;
                         counts[DL_N * it.pw0 + j] = (uint64_t(run_count_word(it.p, set, gw)) << 32) | uint64_t(__popcll(set[gw]));
                     });
        }

        // one thread per (mesh, list, level)
        __global__ __launch_bounds__(SB) void k_blist_table(const SubmeshItem* items, const int64_t* pp, int n, const uint64_t* counts,
                                                            const uint64_t* scan, uint64_t* table)
        {
            const int64_t i = int64_t(blockIdx.x) * SB + threadIdx.x;
            if (i >= int64_t(n) * DL_N * SM_SEG)
            {
                return;
            }
            const int k           = int(i / (DL_N * SM_SEG));
            const int j           = int(i / SM_SEG % DL_N);
            const int l           = int(i % SM_SEG);
            const SubmeshItem& it = items[k];
            const PBits& p        = it.p;
            const int64_t N       = DL_N * pp[n];
            const int64_t w       = DL_N * it.pw0 + j * item_words(it) + p.off[l <= p.hi + 1 ? l : p.hi + 1];
            table[i]              = w < N ? scan[w] : scan[N - 1] + counts[N - 1];
        }

        __global__ __launch_bounds__(SB) void k_blist_emit(const SubmeshItem* items, const int64_t* pp, int n, const uint64_t* bits,
                                                           const uint64_t* scan, const ListOut* outs)
        {
            for_word(items, pp, n, DL_N,
                     [&](const SubmeshItem& it, int k, int64_t j)
                     {
                         const int64_t total = item_words(it);
                         const int list      = int(j / total);
                         emit_list_word(it.p, bits + DL_N * it.pw0 + list * total, j % total, scan[DL_N * it.pw0 + j], outs[k * DL_N + list]);
                     });
        }
    } // namespace

    size_t lists_scan_temp_bytes(int64_t n)
    {
        size_t bytes = 0;
        (void)hipcub::DeviceScan::ExclusiveSum(nullptr, bytes, static_cast<const uint64_t*>(nullptr), static_cast<uint64_t*>(nullptr),
                                               int(n > 0 ? n : 1));
        return bytes;
    }

    void lists_build(const PBits& p, int minl, int maxl, const uint64_t* sets, const uint64_t* old_sets, int nlists, uint64_t* bits,
                     uint64_t* counts, uint64_t* scan, void* temp, size_t temp_bytes, uint64_t* table, hipStream_t s)
    {
        const int64_t n = p.off[p.hi + 1] * nlists;
        if (n <= 0)
        {
            return;
        }
        hipLaunchKernelGGL(k_list_words, dim3(sm_blocks(n)), dim3(SB), 0, s, p, minl, maxl, sets, old_sets, nlists, bits);
        hipLaunchKernelGGL(k_list_counts, dim3(sm_blocks(n)), dim3(SB), 0, s, p, nlists, bits, counts);
        (void)hipcub::DeviceScan::ExclusiveSum(temp, temp_bytes, counts, scan, int(n), s);
        hipLaunchKernelGGL(k_list_table, dim3(sm_blocks(nlists * SM_SEG)), dim3(SB), 0, s, p, nlists, counts, scan, table);
    }

    void lists_emit(const PBits& p, int nlists, const uint64_t* bits, const uint64_t* scan, const ListOuts& outs, hipStream_t s)
    {
        const int64_t n = p.off[p.hi + 1] * nlists;
        if (n > 0)
        {
            hipLaunchKernelGGL(k_list_emit, dim3(sm_blocks(n)), dim3(SB), 0, s, p, nlists, bits, scan, outs);
        }
    }

    void lists_batch_build(const SubmeshItem* items, const int64_t* pp, int64_t pwords, int n, const uint64_t* sets, const uint64_t* old_sets,
                           uint64_t* bits, uint64_t* counts, uint64_t* scan, void* temp, size_t temp_bytes, uint64_t* table, hipStream_t s)
    {
        const int64_t nw = DL_N * pwords;
        if (nw <= 0 || n <= 0)
        {
            return;
        }
        hipLaunchKernelGGL(k_blist_words, dim3(sm_blocks(nw)), dim3(SB), 0, s, items, pp, n, sets, old_sets, bits);
        hipLaunchKernelGGL(k_blist_counts, dim3(sm_blocks(nw)), dim3(SB), 0, s, items, pp, n, bits, counts);
        (void)hipcub::DeviceScan::ExclusiveSum(temp, temp_bytes, counts, scan, int(nw), s);
        hipLaunchKernelGGL(k_blist_table, dim3(sm_blocks(int64_t(n) * DL_N * SM_SEG)), dim3(SB), 0, s, items, pp, n, counts, scan, table);
    }

    void lists_batch_emit(const SubmeshItem* items, const int64_t* pp, int64_t pwords, int n, const uint64_t* bits, const uint64_t* scan,
                          const ListOut* outs, hipStream_t s)
    {
        const int64_t nw = DL_N * pwords;
        if (nw > 0 && n > 0)
        {
            hipLaunchKernelGGL(k_blist_emit, dim3(sm_blocks(nw)), dim3(SB), 0, s, items, pp, n, bits, scan, outs);
        }
    }
} // namespace smr::dev
