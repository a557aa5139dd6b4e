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
        // the run levels [lo, hi) of list k
        __host__ __device__ __forceinline__ void list_levels(int k, int minl, int maxl, int& lo, int& hi)
        {
            switch (k)
            {
                case DL_PROJECTION: lo = 0, hi = maxl; break;
                case DL_PREDICTION: lo = 1, hi = maxl + 1; break;
                case DL_DETAIL: lo = minl > 0 ? minl - 1 : 0, hi = maxl; break;
                case DL_CRITERIA: lo = (minl > 1 ? minl : 1) - 1, hi = maxl; break;
                case DL_MAXIMUM: lo = 0, hi = maxl; break;
                case DL_UF_COPY: lo = minl, hi = maxl + 1; break;
                default: lo = minl, hi = maxl; break; // coarsen, refine
            }
            if (hi < lo)
            {
                hi = lo;
            }
        }

        // the lists of one mesh laid out one after the other, each over the
        // words of its own levels only: list k holds the words [off[lo[k]],
        // off[hi[k]]) of the padded layout, from index base[k] on
        struct ListLayout
        {
            int64_t base[DL_N + 1];
            int lo[DL_N], hi[DL_N];
        };

        __device__ __forceinline__ ListLayout list_layout(const PBits& p, int minl, int maxl, int nlists)
        {
            ListLayout L;
            L.base[0] = 0;
            for (int k = 0; k < DL_N; ++k)
            {
                int lo = 0, hi = 0;
                if (k < nlists)
                {
                    list_levels(k, minl, maxl, lo, hi);
                }
                L.lo[k]       = lo;
                L.hi[k]       = hi;
                L.base[k + 1] = L.base[k] + (p.off[hi] - p.off[lo]);
            }
            return L;
        }

        // the list holding layout index j, and its word of the padded layout
        __device__ __forceinline__ int list_of(const ListLayout& L, const PBits& p, int64_t j, int64_t& gw)
        {
            int k = 0;
            while (k + 1 < DL_N && L.base[k + 1] <= j)
            {
                ++k;
            }
            gw = p.off[L.lo[k]] + (j - L.base[k]);
            return k;
        }

        // the bitmap of list k addressed by padded word (only its levels' words exist)
        __device__ __forceinline__ const uint64_t* list_set(const ListLayout& L, const PBits& p, const uint64_t* bits, int k)
        {
            return bits + L.base[k] - p.off[L.lo[k]];
        }

        // the scan at the first word of level l of list k (levels outside the
        // list's range: its start or its end); n: the scanned words
        __device__ __forceinline__ uint64_t list_level_start(const ListLayout& L, const PBits& p, int k, int l, const uint64_t* counts,
                                                             const uint64_t* scan, int64_t n)
        {
            const int c     = l < L.lo[k] ? L.lo[k] : (l > L.hi[k] ? L.hi[k] : l);
            const int64_t w = L.base[k] + (p.off[c] - p.off[L.lo[k]]);
            return w < n ? scan[w] : (n > 0 ? scan[n - 1] + counts[n - 1] : 0ull);
        }

        // word gw (of a level of the list) of list k
        __device__ __forceinline__ uint64_t list_word(const PBits& p, const uint64_t* sets, const uint64_t* old_sets, int k, int64_t gw)
        {
            const PWord q = pword(p, gw);
            const int t   = q.l;
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
            const ListLayout L = list_layout(p, minl, maxl, nlists);
            const int64_t i    = int64_t(blockIdx.x) * SB + threadIdx.x;
            if (i < L.base[DL_N])
            {
                int64_t gw;
                const int k = list_of(L, p, i, gw);
                bits[i]     = list_word(p, sets, old_sets, k, gw);
            }
        }

        // runs starting in the word << 32 | cells of the word
        __global__ __launch_bounds__(SB) void k_list_counts(PBits p, int minl, int maxl, int nlists, const uint64_t* bits, uint64_t* counts)
        {
            const ListLayout L = list_layout(p, minl, maxl, nlists);
            const int64_t i    = int64_t(blockIdx.x) * SB + threadIdx.x;
            if (i < L.base[DL_N])
            {
                int64_t gw;
                const int k         = list_of(L, p, i, gw);
                const uint64_t* set = list_set(L, p, bits, k);
                counts[i]           = (uint64_t(run_count_word(p, set, gw)) << 32) | uint64_t(__popcll(set[gw]));
            }
        }

        __global__ __launch_bounds__(SB) void k_list_table(PBits p, int minl, int maxl, int nlists, const uint64_t* counts, const uint64_t* scan,
                                                           uint64_t* table)
        {
            const int i = int(blockIdx.x) * SB + threadIdx.x;
            if (i < nlists * SM_SEG)
            {
                const ListLayout L = list_layout(p, minl, maxl, nlists);
                table[i]           = list_level_start(L, p, i / SM_SEG, i % SM_SEG, counts, scan, L.base[DL_N]);
            }
        }

        // the list's trailer: the cell count after the last run, and the last hint
        __device__ __forceinline__ void list_trailer(const ListOut& o)
        {
            const int64_t nchunks = (o.ncells + (int64_t(1) << HINT_SHIFT) - 1) >> HINT_SHIFT;
            o.pref[o.nruns]       = o.ncells;
            o.hint[nchunks]       = int32_t(o.nruns > 0 ? o.nruns - 1 : 0);
        }

        // the runs of list o that start in word gw of its bitmap set (scan_i:
        // the packed scan at that word)
        __device__ __forceinline__ void emit_list_word(const PBits& p, const uint64_t* set, int64_t gw, uint64_t scan_i, const ListOut& o)
        {
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

        // the runs of every list; the threads past the last word write the
        // lists' trailers
        __global__ __launch_bounds__(SB) void k_list_emit(PBits p, int minl, int maxl, int nlists, const uint64_t* bits, const uint64_t* scan,
                                                          ListOuts outs)
        {
            const ListLayout L = list_layout(p, minl, maxl, nlists);
            const int64_t i    = int64_t(blockIdx.x) * SB + threadIdx.x;
            const int64_t n    = L.base[DL_N];
            if (i < n)
            {
                int64_t gw;
                const int k = list_of(L, p, i, gw);
                emit_list_word(p, list_set(L, p, bits, k), gw, scan[i], outs.l[k]);
            }
            else if (i < n + nlists)
            {
                list_trailer(outs.l[i - n]);
            }
        }

        // ---- the lists of many meshes per launch: mesh k's lists (its compact
        // layout) from index cp[k] of the concatenated bitmaps, its ids and old
        // ids from word SM_NSETS * pw0 ----

        __device__ __forceinline__ ListLayout item_layout(const SubmeshItem& it)
        {
            return list_layout(it.p, it.minl, it.p.hi, DL_N);
        }

        __global__ __launch_bounds__(SB) void k_blist_words(const SubmeshItem* items, const int64_t* cp, int n, const uint64_t* sets,
                                                            const uint64_t* old_sets, uint64_t* bits)
        {
            for_word(items, cp, n, 1,
                     [&](const SubmeshItem& it, int k, int64_t j)
                     {
                         const ListLayout L = item_layout(it);
                         int64_t gw;
                         const int list  = list_of(L, it.p, j, gw);
                         bits[cp[k] + j] = list_word(it.p, sets + SM_NSETS * it.pw0, old_sets + SM_NSETS * it.pw0, list, gw);
                     });
        }

        __global__ __launch_bounds__(SB) void k_blist_counts(const SubmeshItem* items, const int64_t* cp, int n, const uint64_t* bits,
                                                             uint64_t* counts)
        {
            for_word(items, cp, n, 1,
                     [&](const SubmeshItem& it, int k, int64_t j)
                     {
                         const ListLayout L = item_layout(it);
                         int64_t gw;
                         const int list      = list_of(L, it.p, j, gw);
                         const uint64_t* set = list_set(L, it.p, bits + cp[k], list);
                         counts[cp[k] + j]   = (uint64_t(run_count_word(it.p, set, gw)) << 32) | uint64_t(__popcll(set[gw]));
                     });
        }

        // one thread per (mesh, list, level)
        __global__ __launch_bounds__(SB) void k_blist_table(const SubmeshItem* items, const int64_t* cp, int n, const uint64_t* counts,
                                                            const uint64_t* scan, uint64_t* table)
        {
            const int64_t i = int64_t(blockIdx.x) * SB + threadIdx.x;
            if (i >= int64_t(n) * DL_N * SM_SEG)
            {
                return;
            }
            const int k           = int(i / (DL_N * SM_SEG));
            const SubmeshItem& it = items[k];
            const ListLayout L    = item_layout(it);
            // the item's counts and scan, read through its first index (the
            // scan keeps counting over the items before it)
            const int64_t N = cp[n];
            const int list  = int(i / SM_SEG % DL_N);
            const int l     = int(i % SM_SEG);
            const int c     = l < L.lo[list] ? L.lo[list] : (l > L.hi[list] ? L.hi[list] : l);
            const int64_t w = cp[k] + L.base[list] + (it.p.off[c] - it.p.off[L.lo[list]]);
            table[i]        = w < N ? scan[w] : (N > 0 ? scan[N - 1] + counts[N - 1] : 0ull);
        }

        __global__ __launch_bounds__(SB) void k_blist_emit(const SubmeshItem* items, const int64_t* cp, int n, const uint64_t* bits,
                                                           const uint64_t* scan, const ListOut* outs)
        {
            for_word(items, cp, n, 1,
                     [&](const SubmeshItem& it, int k, int64_t j)
                     {
                         const ListLayout L = item_layout(it);
                         int64_t gw;
                         const int list = list_of(L, it.p, j, gw);
                         emit_list_word(it.p, list_set(L, it.p, bits + cp[k], list), gw, scan[cp[k] + j], outs[k * DL_N + list]);
                     });
        }

        __global__ __launch_bounds__(SB) void k_blist_trailers(const ListOut* outs, int n)
        {
            const int i = int(blockIdx.x) * SB + threadIdx.x;
            if (i < n)
            {
                list_trailer(outs[i]);
            }
        }
    } // namespace

    size_t lists_scan_temp_bytes(int64_t n)
    {
        size_t bytes = 0;
        (void)hipcub::DeviceScan::ExclusiveSum(nullptr, bytes, static_cast<const uint64_t*>(nullptr), static_cast<uint64_t*>(nullptr),
                                               int(n > 0 ? n : 1));
        return bytes;
    }

    int64_t lists_words(const PBits& p, int minl, int maxl, int nlists)
    {
        int64_t n = 0;
        for (int k = 0; k < nlists; ++k)
        {
            int lo = 0, hi = 0;
            list_levels(k, minl, maxl, lo, hi);
            n += p.off[hi] - p.off[lo];
        }
        return n;
    }

    void lists_build(const PBits& p, int minl, int maxl, const uint64_t* sets, const uint64_t* old_sets, int nlists, uint64_t* bits,
                     uint64_t* counts, uint64_t* scan, void* temp, size_t temp_bytes, uint64_t* table, hipStream_t s)
    {
        const int64_t n = lists_words(p, minl, maxl, nlists);
        if (n > 0)
        {
            hipLaunchKernelGGL(k_list_words, dim3(sm_blocks(n)), dim3(SB), 0, s, p, minl, maxl, sets, old_sets, nlists, bits);
            hipLaunchKernelGGL(k_list_counts, dim3(sm_blocks(n)), dim3(SB), 0, s, p, minl, maxl, nlists, bits, counts);
            (void)hipcub::DeviceScan::ExclusiveSum(temp, temp_bytes, counts, scan, int(n), s);
        }
        hipLaunchKernelGGL(k_list_table, dim3(sm_blocks(nlists * SM_SEG)), dim3(SB), 0, s, p, minl, maxl, nlists, counts, scan, table);
    }

    void lists_emit(const PBits& p, int minl, int maxl, int nlists, const uint64_t* bits, const uint64_t* scan, const ListOuts& outs, hipStream_t s)
    {
        const int64_t n = lists_words(p, minl, maxl, nlists) + nlists;
        hipLaunchKernelGGL(k_list_emit, dim3(sm_blocks(n)), dim3(SB), 0, s, p, minl, maxl, nlists, bits, scan, outs);
    }

    void lists_batch_build(const SubmeshItem* items, const int64_t* cp, int64_t nwords, int n, const uint64_t* sets, const uint64_t* old_sets,
                           uint64_t* bits, uint64_t* counts, uint64_t* scan, void* temp, size_t temp_bytes, uint64_t* table, hipStream_t s)
    {
        if (n <= 0)
        {
            return;
        }
        if (nwords > 0)
        {
            hipLaunchKernelGGL(k_blist_words, dim3(sm_blocks(nwords)), dim3(SB), 0, s, items, cp, n, sets, old_sets, bits);
            hipLaunchKernelGGL(k_blist_counts, dim3(sm_blocks(nwords)), dim3(SB), 0, s, items, cp, n, bits, counts);
            (void)hipcub::DeviceScan::ExclusiveSum(temp, temp_bytes, counts, scan, int(nwords), s);
        }
        hipLaunchKernelGGL(k_blist_table, dim3(sm_blocks(int64_t(n) * DL_N * SM_SEG)), dim3(SB), 0, s, items, cp, n, counts, scan, table);
    }

    void lists_batch_emit(const SubmeshItem* items, const int64_t* cp, int64_t nwords, int n, const uint64_t* bits, const uint64_t* scan,
                          const ListOut* outs, hipStream_t s)
    {
        if (nwords > 0 && n > 0)
        {
            hipLaunchKernelGGL(k_blist_emit, dim3(sm_blocks(nwords)), dim3(SB), 0, s, items, cp, n, bits, scan, outs);
        }
        if (n > 0)
        {
            hipLaunchKernelGGL(k_blist_trailers, dim3(sm_blocks(int64_t(n) * DL_N)), dim3(SB), 0, s, outs, n * DL_N);
        }
    }
} // namespace smr::dev
