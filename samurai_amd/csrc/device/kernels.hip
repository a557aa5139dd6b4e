// MR-step kernels for gfx950 (CDNA4), FP64 on the vector ALU, wave64.
//
// Every kernel streams over the cells of a work list: one thread per cell,
// the run of the cell found by a short binary search between two chunk hints,
// and the storage position of every touched cell obtained as
// (resolved run base of its row) + x, which is exactly the reference's
// memory_offset(...) + i arithmetic.  All floating-point expressions keep the
// reference's operation order; the file is compiled with -ffp-contract=off.
#include "device_api.hpp"

#include <hip/hip_runtime.h>

#include <cstdlib>

namespace smr::dev
{
    namespace
    {
        constexpr int BLOCK = 256;

        __device__ __forceinline__ int64_t row_key(int32_t y, int32_t z)
        {
            return (int64_t(z) << 32) + (int64_t(y) + 2147483648LL);
        }

        // LevelCellArray::get_interval/find (level_cell_array.hpp:669-743,
        // algorithm.hpp:391-467) as two binary searches: the row among the
        // level's rows, then the run holding x among the row's runs.
        __device__ bool find_base(const Table& t, int level, int32_t x, int32_t y, int32_t z, int64_t& base, int32_t& rs, int32_t& re)
        {
            if (level < 0 || level > 20)
            {
                return false;
            }
            const int64_t key = row_key(y, z);
            int64_t lo = t.row_off[level], hi = t.row_off[level + 1];
            const int64_t end = hi;
            while (lo < hi)
            {
                int64_t mid = (lo + hi) >> 1;
                if (t.row_key[mid] < key)
                {
                    lo = mid + 1;
                }
                else
                {
                    hi = mid;
                }
            }
            if (lo >= end || t.row_key[lo] != key)
            {
                return false;
            }
            const int64_t first = t.rp[lo];
            int64_t a = first, b = t.rp[lo + 1];
            while (a < b)
            {
                int64_t mid = (a + b) >> 1;
                if (t.run_s[mid] <= x)
                {
                    a = mid + 1;
                }
                else
                {
                    b = mid;
                }
            }
            const int64_t k = a - 1;
            if (k < first || x >= t.run_e[k])
            {
                return false;
            }
            base = t.run_idx[k];
            rs   = t.run_s[k];
            re   = t.run_e[k];
            return true;
        }

        // Resolves the storage base of every (run, tap) pair.  A pair fails when
        // the tap's row is absent or when the x window the kernels will touch
        // for the whole run is not inside the single reference run found: the
        // host then refuses to launch the operator kernels, so their accesses
        // are always inside the field storage.
        __global__ void k_resolve(List l, Table t0, Table t1, int* err)
        {
            const int64_t gid = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
            if (gid >= l.nruns * l.ntaps)
            {
                return;
            }
            const int64_t r   = gid / l.ntaps;
            const int tp      = int(gid - r * l.ntaps);
            const int32_t x0  = l.iv[5 * r + 0], x1 = l.iv[5 * r + 1], y = l.iv[5 * r + 2], z = l.iv[5 * r + 3];
            const int level   = l.iv[5 * r + 4];
            const int8_t* tap = l.taps + 7 * tp;
            const int dl = tap[1], dx = tap[2], dy = tap[3], dz = tap[4], wlo = tap[5], whi = tap[6];
            int tl;
            int32_t tx, ty, tz, last;
            if (dl == 0)
            {
                tl   = level;
                tx   = x0 + dx;
                last = (x1 - 1) + dx;
                ty   = y + dy;
                tz   = z + dz;
            }
            else if (dl > 0)
            {
                tl   = level + 1;
                tx   = 2 * x0 + dx;
                last = 2 * (x1 - 1) + dx;
                ty   = 2 * y + dy;
                tz   = 2 * z + dz;
            }
            else
            {
                tl   = level - 1;
                tx   = (x0 >> 1) + dx;
                last = ((x1 - 1) >> 1) + dx;
                ty   = (y >> 1) + dy;
                tz   = (z >> 1) + dz;
            }
            int64_t base = 0;
            int32_t rs = 0, re = 0;
            bool ok = find_base(tap[0] == 0 ? t0 : t1, tl, tx, ty, tz, base, rs, re);
            ok      = ok && rs <= tx + wlo && last + whi < re;
            if (!ok)
            {
                base   = INT64_MIN;
                err[0] = 1;
            }
            l.base[gid] = base;
        }

        struct Cell
        {
            int64_t run;
            int32_t x, y, z;
            int level;
        };

        __device__ __forceinline__ Cell locate(const List& l, int64_t c)
        {
            const int64_t j = c >> HINT_SHIFT;
            int64_t lo = l.hint[j], hi = l.hint[j + 1];
            while (lo < hi)
            {
                int64_t mid = (lo + hi + 1) >> 1;
                if (l.pref[mid] <= c)
                {
                    lo = mid;
                }
                else
                {
                    hi = mid - 1;
                }
            }
            Cell cell;
            cell.run   = lo;
            cell.x     = l.iv[5 * lo + 0] + int32_t(c - l.pref[lo]);
            cell.y     = l.iv[5 * lo + 2];
            cell.z     = l.iv[5 * lo + 3];
            cell.level = l.iv[5 * lo + 4];
            return cell;
        }

        __device__ __forceinline__ const int64_t* bases(const List& l, int64_t run)
        {
            return l.base + run * l.ntaps;
        }

        // The hardware deals consecutive workgroups to the 8 XCDs round-robin.
        // Re-map them so that each XCD owns a contiguous range of tiles: rows
        // that are y-neighbours are then read through the same L2.  The grid is
        // launched as a multiple of 8 blocks; `nlogical` tiles exist.
        __device__ __forceinline__ int64_t xcd_tile(unsigned nlogical)
        {
            const unsigned per = gridDim.x >> 3;
            const unsigned t   = (blockIdx.x & 7u) * per + (blockIdx.x >> 3);
            return t < nlogical ? int64_t(t) : -1;
        }

        // flux of upwind_op (stencil_field.hpp:91-97),
        //   .5 * a * (ul + ur) + .5 * |a| * (ul - ur),
        // with the scalar factors (0.5*a, 0.5*|a|) formed once per launch
        __device__ __forceinline__ double flux_h(double ha, double habs, double ul, double ur)
        {
            return ha * (ul + ur) + habs * (ul - ur);
        }

        // unp1 = u - dt * upwind(a, u) on mesh[cells] (stencil_field.hpp:37-53).
        // Each block processes CPT*BLOCK consecutive cells, BLOCK at a time.
        // S / dx with dx = 2^-level is computed as S * 2^level: both are the
        // exact value S*2^level, so the result is bit-identical.
        template <int DIM, int CPT, bool REMAP, bool NT>
        __global__ __launch_bounds__(BLOCK) void k_upwind(List l, int64_t c0, int64_t c1, const double* __restrict__ u,
                                                          double* __restrict__ unp1, double a0, double a1, double a2, double dt,
                                                          unsigned nlogical)
        {
            const int64_t tile = REMAP ? xcd_tile(nlogical) : int64_t(blockIdx.x);
            if (tile < 0 || tile >= int64_t(nlogical))
            {
                return;
            }
            const double ha0 = .5 * a0, hb0 = .5 * fabs(a0);
            const double ha1 = .5 * a1, hb1 = .5 * fabs(a1);
            const double ha2 = .5 * a2, hb2 = .5 * fabs(a2);
            const int64_t first = c0 + tile * (BLOCK * CPT) + threadIdx.x;
#pragma unroll
            for (int k = 0; k < CPT; ++k)
            {
                const int64_t c = first + k * BLOCK;
                if (c >= c1)
                {
                    return;
                }
                const Cell cell    = locate(l, c);
                const int64_t* b   = bases(l, cell.run);
                const double* uc_p = u + b[0] + cell.x;
                const double uc    = uc_p[0];
                const double FL    = flux_h(ha0, hb0, uc_p[-1], uc);
                const double FR    = flux_h(ha0, hb0, uc, uc_p[1]);
                const double FD    = flux_h(ha1, hb1, u[b[1] + cell.x], uc);
                const double FU    = flux_h(ha1, hb1, uc, u[b[2] + cell.x]);
                double S;
                if constexpr (DIM == 2)
                {
                    S = -FL + FR + -FD + FU;
                }
                else
                {
                    const double FF = flux_h(ha2, hb2, u[b[3] + cell.x], uc);
                    const double FB = flux_h(ha2, hb2, uc, u[b[4] + cell.x]);
                    S               = -FL + FR + -FD + FU + -FF + FB;
                }
                const double inv_dx = double(1 << cell.level);
                const double v      = uc - dt * (S * inv_dx);
                if constexpr (NT)
                {
                    __builtin_nontemporal_store(v, unp1 + b[0] + cell.x);
                }
                else
                {
                    unp1[b[0] + cell.x] = v;
                }
            }
        }

        // Upwind on wave tiles: every wavefront owns one tile of <= 64 consecutive
        // cells of a single run, so the run's level, end and resolved row bases
        // are wave-uniform and live in scalar registers; each lane then does five
        // coalesced loads and one store.  Same arithmetic as k_upwind.
        template <int DIM, bool NT>
        __global__ __launch_bounds__(BLOCK) void k_upwind_tiles(List l, int64_t t0, int64_t t1, const double* __restrict__ u,
                                                                double* __restrict__ unp1, double a0, double a1, double a2, double dt,
                                                                unsigned nlogical)
        {
            const int64_t blk = xcd_tile(nlogical);
            if (blk < 0)
            {
                return;
            }
            const int64_t t = t0 + blk * (BLOCK / 64) + (threadIdx.x >> 6);
            if (t >= t1)
            {
                return;
            }
            const int32_t run = __builtin_amdgcn_readfirstlane(l.tiles[2 * t]);
            const int32_t xs  = __builtin_amdgcn_readfirstlane(l.tiles[2 * t + 1]);
            const int32_t x   = xs + int32_t(threadIdx.x & 63);
            const int32_t* iv = l.iv + 5 * int64_t(run);
            if (x >= iv[1])
            {
                return;
            }
            const int64_t* b   = l.base + int64_t(run) * l.ntaps;
            const double ha0 = .5 * a0, hb0 = .5 * fabs(a0);
            const double ha1 = .5 * a1, hb1 = .5 * fabs(a1);
            const double* uc_p = u + b[0] + x;
            const double uc    = uc_p[0];
            const double FL    = flux_h(ha0, hb0, uc_p[-1], uc);
            const double FR    = flux_h(ha0, hb0, uc, uc_p[1]);
            const double FD    = flux_h(ha1, hb1, u[b[1] + x], uc);
            const double FU    = flux_h(ha1, hb1, uc, u[b[2] + x]);
            double S;
            if constexpr (DIM == 2)
            {
                S = -FL + FR + -FD + FU;
            }
            else
            {
                const double ha2 = .5 * a2, hb2 = .5 * fabs(a2);
                const double FF  = flux_h(ha2, hb2, u[b[3] + x], uc);
                const double FB  = flux_h(ha2, hb2, uc, u[b[4] + x]);
                S                = -FL + FR + -FD + FU + -FF + FB;
            }
            const double v = uc - dt * (S * double(1 << iv[4]));
            if constexpr (NT)
            {
                __builtin_nontemporal_store(v, unp1 + b[0] + x);
            }
            else
            {
                unp1[b[0] + x] = v;
            }
        }

        // projection_op_ (numeric/projection.hpp:15-65): taps dst, then the 2^(dim-1) fine rows
        template <int DIM>
        __global__ __launch_bounds__(BLOCK) void k_project(List l, int64_t c0, int64_t c1, double* dst, const double* src)
        {
            const int64_t c = c0 + int64_t(blockIdx.x) * BLOCK + threadIdx.x;
            if (c >= c1)
            {
                return;
            }
            const Cell cell  = locate(l, c);
            const int64_t* b = bases(l, cell.run);
            constexpr int NR = 1 << (DIM - 1);
            constexpr double inv = 1.0 / double(1 << DIM);
            const int32_t xf = 2 * cell.x;
            double sum       = 0;
#pragma unroll
            for (int r = 0; r < NR; ++r)
            {
                sum += src[b[1 + r] + xf] + src[b[1 + r] + xf + 1];
            }
            dst[b[0] + cell.x] = sum * inv;
        }

        __device__ __forceinline__ double pcoef(int parity, int k)
        {
            // interp_coeffs<3>(+1) = {1/8, 1, -1/8}, interp_coeffs<3>(-1) = {-1/8, 1, 1/8}
            // (numeric/prediction.hpp:31-35)
            return k == 1 ? 1.0 : ((parity == 0) == (k == 0) ? 1. / 8. : -1. / 8.);
        }

        // prediction_op <s,false> (numeric/prediction.hpp:259-357): taps dst, then
        // the coarse rows (z outer, y inner) starting at x = (x0>>1) - s
        template <int DIM, int S>
        __global__ __launch_bounds__(BLOCK) void k_predict_fine(List l, int64_t c0, int64_t c1, double* u)
        {
            const int64_t c = c0 + int64_t(blockIdx.x) * BLOCK + threadIdx.x;
            if (c >= c1)
            {
                return;
            }
            const Cell cell  = locate(l, c);
            const int64_t* b = bases(l, cell.run);
            const int32_t xc = cell.x >> 1;
            double v;
            if constexpr (S == 0)
            {
                v = u[b[1] + xc];
            }
            else
            {
                const int px = cell.x & 1, py = cell.y & 1, pz = cell.z & 1;
                v            = 0;
                constexpr int NZ = DIM == 3 ? 3 : 1;
#pragma unroll
                for (int rz = 0; rz < NZ; ++rz)
                {
#pragma unroll
                    for (int ry = 0; ry < 3; ++ry)
                    {
                        const int64_t rb = b[1 + rz * 3 + ry] + xc - 1;
#pragma unroll
                        for (int sx = 0; sx < 3; ++sx)
                        {
                            const double f = u[rb + sx];
                            double coeff   = 1.;
                            coeff *= pcoef(px, sx);
                            coeff *= pcoef(py, ry);
                            if constexpr (DIM == 3)
                            {
                                coeff *= pcoef(pz, rz);
                            }
                            v += f * coeff;
                        }
                    }
                }
            }
            u[b[0] + cell.x] = v;
        }

        // compute_detail_op (mr/operators.hpp:139-540): taps = the 2^(dim-1) fine
        // rows, then the coarse stencil rows (z outer, y inner)
        template <int DIM, int S>
        __global__ __launch_bounds__(BLOCK) void k_detail(List l, int64_t c0, int64_t c1, const double* u, double* det)
        {
            const int64_t c = c0 + int64_t(blockIdx.x) * BLOCK + threadIdx.x;
            if (c >= c1)
            {
                return;
            }
            const Cell cell  = locate(l, c);
            const int64_t* b = bases(l, cell.run);
            constexpr int NR = 1 << (DIM - 1);
            constexpr int NC = 1 << DIM;
            const int32_t x  = cell.x;
            int64_t fi[NC];
            double d[NC];
#pragma unroll
            for (int k = 0; k < NC; ++k)
            {
                fi[k] = b[k >> 1] + 2 * x + (k & 1);
                d[k]  = u[fi[k]];
            }
            if constexpr (S == 0)
            {
                const double uc = u[b[NR] + x];
#pragma unroll
                for (int k = 0; k < NC; ++k)
                {
                    d[k] = d[k] - uc;
                }
            }
            else
            {
                constexpr int NZ = DIM == 3 ? 3 : 1;
#pragma unroll
                for (int kk = 0; kk < NZ; ++kk)
                {
#pragma unroll
                    for (int kj = 0; kj < 3; ++kj)
                    {
                        const int64_t rb = b[NR + kk * 3 + kj] + x - 1;
#pragma unroll
                        for (int ki = 0; ki < 3; ++ki)
                        {
                            const double src = u[rb + ki];
#pragma unroll
                            for (int k = 0; k < NC; ++k)
                            {
                                double coef = pcoef(k & 1, ki) * pcoef((k >> 1) & 1, kj);
                                if constexpr (DIM == 3)
                                {
                                    coef = coef * pcoef((k >> 2) & 1, kk);
                                }
                                d[k] -= coef * src;
                            }
                        }
                    }
                }
            }
#pragma unroll
            for (int k = 0; k < NC; ++k)
            {
                det[fi[k]] = d[k];
            }
        }

        // mr_criteria_op (mr/criteria.hpp:13-122): taps coarse, then fine rows
        template <int DIM>
        __global__ __launch_bounds__(BLOCK) void k_criteria(List l, int64_t c0, int64_t c1, const double* det, uint8_t* tag, Thresholds thr,
                                                            int min_level, int max_level)
        {
            const int64_t c = c0 + int64_t(blockIdx.x) * BLOCK + threadIdx.x;
            if (c >= c1)
            {
                return;
            }
            const Cell cell  = locate(l, c);
            const int64_t* b = bases(l, cell.run);
            constexpr int NC = 1 << DIM;
            const int fl     = cell.level + 1;
            const int32_t x  = cell.x;
            int64_t fi[NC];
#pragma unroll
            for (int k = 0; k < NC; ++k)
            {
                fi[k] = b[1 + (k >> 1)] + 2 * x + (k & 1);
            }
            const double eps = thr.eps[fl], fine_eps = thr.fine[fl], coarse_eps = thr.coarse[fl];
            if (fl > min_level)
            {
                if (!(fabs(det[b[0] + x]) > coarse_eps))
                {
                    bool ok = true;
#pragma unroll
                    for (int k = 0; k < NC; ++k)
                    {
                        ok = ok && !(fabs(det[fi[k]]) > eps);
                    }
                    if (ok)
                    {
#pragma unroll
                        for (int k = 0; k < NC; ++k)
                        {
                            tag[fi[k]] = 2;
                        }
                    }
                }
            }
            if (fl < max_level)
            {
#pragma unroll
                for (int k = 0; k < NC; ++k)
                {
                    tag[fi[k]] |= uint8_t(fabs(det[fi[k]]) > fine_eps) * uint8_t(4);
                }
            }
        }

        // maximum_op (mr/operators.hpp:22-96): taps coarse, then fine rows
        template <int DIM>
        __global__ __launch_bounds__(BLOCK) void k_maximum(List l, int64_t c0, int64_t c1, uint8_t* tag)
        {
            const int64_t c = c0 + int64_t(blockIdx.x) * BLOCK + threadIdx.x;
            if (c >= c1)
            {
                return;
            }
            const Cell cell  = locate(l, c);
            const int64_t* b = bases(l, cell.run);
            constexpr int NC = 1 << DIM;
            const int32_t x  = cell.x;
            int64_t fi[NC];
            uint8_t t[NC];
            bool keep = false;
#pragma unroll
            for (int k = 0; k < NC; ++k)
            {
                fi[k] = b[1 + (k >> 1)] + 2 * x + (k & 1);
                t[k]  = tag[fi[k]];
                keep  = keep || (t[k] & 1);
            }
            const int64_t ci = b[0] + x;
            if (keep)
            {
#pragma unroll
                for (int k = 0; k < NC; ++k)
                {
                    tag[fi[k]] = t[k] | 1;
                }
                tag[ci] |= 1;
                return;
            }
            bool all_coarsen = true;
#pragma unroll
            for (int k = 0; k < NC; ++k)
            {
                all_coarsen = all_coarsen && (t[k] & 2);
            }
            if (all_coarsen)
            {
                tag[ci] |= 1;
            }
            else
            {
#pragma unroll
                for (int k = 0; k < NC; ++k)
                {
                    tag[fi[k]] = t[k] & uint8_t(~2);
                }
            }
        }

        __global__ __launch_bounds__(BLOCK) void k_tag_init(List l, int64_t c0, int64_t c1, uint8_t* tag)
        {
            const int64_t c = c0 + int64_t(blockIdx.x) * BLOCK + threadIdx.x;
            if (c >= c1)
            {
                return;
            }
            const Cell cell = locate(l, c);
            tag[bases(l, cell.run)[0] + cell.x] = 1;
        }

        __global__ __launch_bounds__(BLOCK) void k_copy(List l, int64_t c0, int64_t c1, double* dst, const double* src)
        {
            const int64_t c = c0 + int64_t(blockIdx.x) * BLOCK + threadIdx.x;
            if (c >= c1)
            {
                return;
            }
            const Cell cell  = locate(l, c);
            const int64_t* b = bases(l, cell.run);
            dst[b[0] + cell.x] = src[b[1] + cell.x];
        }

        // prediction_op <s,true> (numeric/prediction.hpp:107-257): taps = the
        // 2^(dim-1) destination rows (children), then the source stencil rows
        template <int DIM, int S>
        __global__ __launch_bounds__(BLOCK) void k_predict_children(List l, int64_t c0, int64_t c1, double* dst, const double* src)
        {
            const int64_t c = c0 + int64_t(blockIdx.x) * BLOCK + threadIdx.x;
            if (c >= c1)
            {
                return;
            }
            const Cell cell  = locate(l, c);
            const int64_t* b = bases(l, cell.run);
            constexpr int NR = 1 << (DIM - 1);
            constexpr int NC = 1 << DIM;
            const int32_t x  = cell.x;
            double v[NC];
            if constexpr (S == 0)
            {
                const double s0 = src[b[NR] + x];
#pragma unroll
                for (int k = 0; k < NC; ++k)
                {
                    v[k] = s0;
                }
            }
            else
            {
#pragma unroll
                for (int k = 0; k < NC; ++k)
                {
                    v[k] = 0;
                }
                constexpr int NZ = DIM == 3 ? 3 : 1;
#pragma unroll
                for (int rz = 0; rz < NZ; ++rz)
                {
#pragma unroll
                    for (int ry = 0; ry < 3; ++ry)
                    {
                        const int64_t rb = b[NR + rz * 3 + ry] + x - 1;
#pragma unroll
                        for (int sx = 0; sx < 3; ++sx)
                        {
                            const double f = src[rb + sx];
#pragma unroll
                            for (int k = 0; k < NC; ++k)
                            {
                                double coeff = 1.;
                                coeff *= pcoef(k & 1, sx);
                                coeff *= pcoef((k >> 1) & 1, ry);
                                if constexpr (DIM == 3)
                                {
                                    coeff *= pcoef((k >> 2) & 1, rz);
                                }
                                v[k] += f * coeff;
                            }
                        }
                    }
                }
            }
#pragma unroll
            for (int k = 0; k < NC; ++k)
            {
                dst[b[k >> 1] + 2 * x + (k & 1)] = v[k];
            }
        }

        __global__ __launch_bounds__(BLOCK) void k_recipes(const int64_t* __restrict__ dst, const int32_t* __restrict__ op,
                                                           const int64_t* __restrict__ sp, const int64_t* __restrict__ src, int64_t r0,
                                                           int64_t r1, double* u, int has_bc, double bc_value)
        {
            const int64_t r = r0 + int64_t(blockIdx.x) * BLOCK + threadIdx.x;
            if (r >= r1)
            {
                return;
            }
            const int o = op[r];
            if (o == 1 && !has_bc)
            {
                return;
            }
            double v;
            if (o == 0)
            {
                v = u[src[sp[r]]];
            }
            else if (o == 1)
            {
                // Dirichlet<1> (bc/dirichlet.hpp:29)
                v = 2 * bc_value - u[src[sp[r]]];
            }
            else
            {
                // BC projection: sum of the children, then divided by their count
                // (algorithm/update_outer_ghost.hpp:78-131)
                v            = 0;
                const int64_t n = sp[r + 1] - sp[r];
                for (int64_t k = sp[r]; k < sp[r + 1]; ++k)
                {
                    v += u[src[k]];
                }
                if (n > 0)
                {
                    v /= int(n);
                }
            }
            u[dst[r]] = v;
        }

        inline unsigned nblocks(int64_t n)
        {
            return unsigned((n + BLOCK - 1) / BLOCK);
        }

        // launch shapes of the upwind kernel: cells per thread, XCD remap,
        // non-temporal stores
        struct UpwindVariant
        {
            int id, cpt;
            bool remap, nt;
        };
        // cpt 0: wave tiles (k_upwind_tiles)
        constexpr UpwindVariant kUpwindVariants[] = {{0, 1, false, false}, {1, 1, true, false}, {2, 2, false, false},
                                                     {3, 2, true, false},  {4, 4, false, false}, {5, 4, true, true},
                                                     {6, 1, false, true},  {7, 8, false, false}, {8, 0, true, true},
                                                     {9, 0, true, false}};
        constexpr int kNumUpwindVariants = 10;
        constexpr int kDefaultUpwindVariant = 5;
        int g_upwind_variant = -1;

        UpwindVariant upwind_variant()
        {
            if (g_upwind_variant < 0)
            {
                const char* e    = std::getenv("SMR_UPWIND_VARIANT");
                int id           = e ? std::atoi(e) : kDefaultUpwindVariant;
                g_upwind_variant = (id >= 0 && id < kNumUpwindVariants) ? id : kDefaultUpwindVariant;
            }
            return kUpwindVariants[g_upwind_variant];
        }
    } // namespace

    int upwind_variant_count()
    {
        return kNumUpwindVariants;
    }

    void set_upwind_variant(int id)
    {
        g_upwind_variant = (id >= 0 && id < kNumUpwindVariants) ? id : kDefaultUpwindVariant;
    }

    void resolve(const List& l, const Table& t0, const Table& t1, int* err, hipStream_t s)
    {
        const int64_t n = l.nruns * l.ntaps;
        if (n > 0)
        {
            hipLaunchKernelGGL(k_resolve, dim3(nblocks(n)), dim3(BLOCK), 0, s, l, t0, t1, err);
        }
    }

    void upwind(int dim, const List& l, int64_t c0, int64_t c1, const double* u, double* unp1, double a0, double a1, double a2, double dt,
                hipStream_t s)
    {
        if (c1 <= c0)
        {
            return;
        }
        const UpwindVariant v = upwind_variant();
        if (v.cpt == 0 && c0 == 0 && c1 == l.ncells)
        {
            const int64_t nblk    = (l.ntiles + BLOCK / 64 - 1) / (BLOCK / 64);
            const unsigned ngrid  = unsigned((nblk + 7) / 8 * 8);
            const unsigned nlog   = unsigned(nblk);
            if (dim == 2)
            {
                if (v.nt)
                {
                    hipLaunchKernelGGL((k_upwind_tiles<2, true>), dim3(ngrid), dim3(BLOCK), 0, s, l, 0, l.ntiles, u, unp1, a0, a1, a2, dt, nlog);
                }
                else
                {
                    hipLaunchKernelGGL((k_upwind_tiles<2, false>), dim3(ngrid), dim3(BLOCK), 0, s, l, 0, l.ntiles, u, unp1, a0, a1, a2, dt, nlog);
                }
            }
            else
            {
                if (v.nt)
                {
                    hipLaunchKernelGGL((k_upwind_tiles<3, true>), dim3(ngrid), dim3(BLOCK), 0, s, l, 0, l.ntiles, u, unp1, a0, a1, a2, dt, nlog);
                }
                else
                {
                    hipLaunchKernelGGL((k_upwind_tiles<3, false>), dim3(ngrid), dim3(BLOCK), 0, s, l, 0, l.ntiles, u, unp1, a0, a1, a2, dt, nlog);
                }
            }
            return;
        }
        const int cpt = v.cpt == 0 ? 4 : v.cpt;
        const int64_t ntiles  = (c1 - c0 + BLOCK * cpt - 1) / (BLOCK * cpt);
        const unsigned ngrid  = v.remap ? unsigned((ntiles + 7) / 8 * 8) : unsigned(ntiles);
        const unsigned nlog   = unsigned(ntiles);
        const dim3 g(ngrid), bl(BLOCK);
#define SMR_UPWIND_LAUNCH(D, C, R, N)                                                                                                  \
    hipLaunchKernelGGL((k_upwind<D, C, R, N>), g, bl, 0, s, l, c0, c1, u, unp1, a0, a1, a2, dt, nlog)
#define SMR_UPWIND_DISPATCH(D)                                                                                                         \
    switch (v.id)                                                                                                                      \
    {                                                                                                                                  \
        case 0: SMR_UPWIND_LAUNCH(D, 1, false, false); break;                                                                         \
        case 1: SMR_UPWIND_LAUNCH(D, 1, true, false); break;                                                                          \
        case 2: SMR_UPWIND_LAUNCH(D, 2, false, false); break;                                                                         \
        case 3: SMR_UPWIND_LAUNCH(D, 2, true, false); break;                                                                          \
        case 4: SMR_UPWIND_LAUNCH(D, 4, false, false); break;                                                                         \
        case 5: SMR_UPWIND_LAUNCH(D, 4, true, true); break;                                                                           \
        case 6: SMR_UPWIND_LAUNCH(D, 1, false, true); break;                                                                          \
        default: SMR_UPWIND_LAUNCH(D, 8, false, false); break;                                                                        \
    }
        if (dim == 2)
        {
            SMR_UPWIND_DISPATCH(2)
        }
        else
        {
            SMR_UPWIND_DISPATCH(3)
        }
#undef SMR_UPWIND_DISPATCH
#undef SMR_UPWIND_LAUNCH
    }

    void project(int dim, const List& l, int64_t c0, int64_t c1, double* dst, const double* src, hipStream_t s)
    {
        if (c1 <= c0)
        {
            return;
        }
        if (dim == 2)
        {
            hipLaunchKernelGGL(k_project<2>, dim3(nblocks(c1 - c0)), dim3(BLOCK), 0, s, l, c0, c1, dst, src);
        }
        else
        {
            hipLaunchKernelGGL(k_project<3>, dim3(nblocks(c1 - c0)), dim3(BLOCK), 0, s, l, c0, c1, dst, src);
        }
    }

    void predict_fine(int dim, int ps, const List& l, int64_t c0, int64_t c1, double* u, hipStream_t s)
    {
        if (c1 <= c0)
        {
            return;
        }
        const dim3 g(nblocks(c1 - c0)), bl(BLOCK);
        if (dim == 2)
        {
            if (ps == 0)
            {
                hipLaunchKernelGGL((k_predict_fine<2, 0>), g, bl, 0, s, l, c0, c1, u);
            }
            else
            {
                hipLaunchKernelGGL((k_predict_fine<2, 1>), g, bl, 0, s, l, c0, c1, u);
            }
        }
        else
        {
            if (ps == 0)
            {
                hipLaunchKernelGGL((k_predict_fine<3, 0>), g, bl, 0, s, l, c0, c1, u);
            }
            else
            {
                hipLaunchKernelGGL((k_predict_fine<3, 1>), g, bl, 0, s, l, c0, c1, u);
            }
        }
    }

    void detail(int dim, int ps, const List& l, int64_t c0, int64_t c1, const double* u, double* det, hipStream_t s)
    {
        if (c1 <= c0)
        {
            return;
        }
        const dim3 g(nblocks(c1 - c0)), bl(BLOCK);
        if (dim == 2)
        {
            if (ps == 0)
            {
                hipLaunchKernelGGL((k_detail<2, 0>), g, bl, 0, s, l, c0, c1, u, det);
            }
            else
            {
                hipLaunchKernelGGL((k_detail<2, 1>), g, bl, 0, s, l, c0, c1, u, det);
            }
        }
        else
        {
            if (ps == 0)
            {
                hipLaunchKernelGGL((k_detail<3, 0>), g, bl, 0, s, l, c0, c1, u, det);
            }
            else
            {
                hipLaunchKernelGGL((k_detail<3, 1>), g, bl, 0, s, l, c0, c1, u, det);
            }
        }
    }

    void criteria(int dim, const List& l, int64_t c0, int64_t c1, const double* det, uint8_t* tag, const Thresholds& thr, int min_level,
                  int max_level, hipStream_t s)
    {
        if (c1 <= c0)
        {
            return;
        }
        const dim3 g(nblocks(c1 - c0)), bl(BLOCK);
        if (dim == 2)
        {
            hipLaunchKernelGGL(k_criteria<2>, g, bl, 0, s, l, c0, c1, det, tag, thr, min_level, max_level);
        }
        else
        {
            hipLaunchKernelGGL(k_criteria<3>, g, bl, 0, s, l, c0, c1, det, tag, thr, min_level, max_level);
        }
    }

    void maximum(int dim, const List& l, int64_t c0, int64_t c1, uint8_t* tag, hipStream_t s)
    {
        if (c1 <= c0)
        {
            return;
        }
        const dim3 g(nblocks(c1 - c0)), bl(BLOCK);
        if (dim == 2)
        {
            hipLaunchKernelGGL(k_maximum<2>, g, bl, 0, s, l, c0, c1, tag);
        }
        else
        {
            hipLaunchKernelGGL(k_maximum<3>, g, bl, 0, s, l, c0, c1, tag);
        }
    }

    void tag_init(const List& l, int64_t c0, int64_t c1, uint8_t* tag, hipStream_t s)
    {
        if (c1 > c0)
        {
            hipLaunchKernelGGL(k_tag_init, dim3(nblocks(c1 - c0)), dim3(BLOCK), 0, s, l, c0, c1, tag);
        }
    }

    void copy_cells(const List& l, int64_t c0, int64_t c1, double* dst, const double* src, hipStream_t s)
    {
        if (c1 > c0)
        {
            hipLaunchKernelGGL(k_copy, dim3(nblocks(c1 - c0)), dim3(BLOCK), 0, s, l, c0, c1, dst, src);
        }
    }

    void predict_children(int dim, int ps, const List& l, int64_t c0, int64_t c1, double* dst, const double* src, hipStream_t s)
    {
        if (c1 <= c0)
        {
            return;
        }
        const dim3 g(nblocks(c1 - c0)), bl(BLOCK);
        if (dim == 2)
        {
            if (ps == 0)
            {
                hipLaunchKernelGGL((k_predict_children<2, 0>), g, bl, 0, s, l, c0, c1, dst, src);
            }
            else
            {
                hipLaunchKernelGGL((k_predict_children<2, 1>), g, bl, 0, s, l, c0, c1, dst, src);
            }
        }
        else
        {
            if (ps == 0)
            {
                hipLaunchKernelGGL((k_predict_children<3, 0>), g, bl, 0, s, l, c0, c1, dst, src);
            }
            else
            {
                hipLaunchKernelGGL((k_predict_children<3, 1>), g, bl, 0, s, l, c0, c1, dst, src);
            }
        }
    }

    void recipes(const int64_t* dst, const int32_t* op, const int64_t* src_ptr, const int64_t* src, int64_t r0, int64_t r1, double* u,
                 bool has_bc, double bc_value, hipStream_t s)
    {
        if (r1 > r0)
        {
            hipLaunchKernelGGL(k_recipes, dim3(nblocks(r1 - r0)), dim3(BLOCK), 0, s, dst, op, src_ptr, src, r0, r1, u, int(has_bc),
                               bc_value);
        }
    }
} // namespace smr::dev
