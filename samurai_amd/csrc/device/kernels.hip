// MR-step kernels for gfx950 (CDNA4), FP64 on the vector ALU, wave64.
//
// The per-cell bodies live in cell_ops.hpp (shared with the fused stage
// program); the kernels here launch one thread per cell of a work list range.
// The file is compiled with -ffp-contract=off.
#include "cell_ops.hpp"

#include <algorithm>
#include <cfloat>
#include <cstdlib>

namespace smr::dev
{
    namespace
    {
        constexpr int BLOCK = 256;

        __global__ void k_resolve_batch(ResolveBatch b, int* err)
        {
            const int64_t gid = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
            if (gid >= b.start[b.n])
            {
                return;
            }
            int k = 0;
            while (k + 1 < b.n && gid >= b.start[k + 1])
            {
                ++k;
            }
            const List& l   = b.lists[k];
            const int64_t i = gid - b.start[k];
            if (i < l.nruns * l.ntaps)
            {
                op_resolve(l, i, b.t_new, b.old[k] ? b.t_old : b.t_new, err);
            }
            if (l.cellrun && i < l.ncells)
            {
                op_cellrun(l, i);
            }
        }

        // the (run, tap) bases and, in the same launch, the run map
        __global__ void k_resolve(List l, Table t0, Table t1, int* err)
        {
            const int64_t gid = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
            if (gid < l.nruns * l.ntaps)
            {
                op_resolve(l, gid, t0, t1, err);
            }
            if (l.cellrun && gid < l.ncells)
            {
                op_cellrun(l, gid);
            }
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

        // unp1 = u - dt * upwind(a, u) on mesh[cells]; each block processes
        // CPT*BLOCK consecutive cells, BLOCK at a time
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
                const Cell cell  = locate(l, c);
                const int64_t* b = bases(l, cell.run);
                const double v   = upwind_value<DIM>(b, cell.x, cell.level, u, ha0, hb0, ha1, hb1, ha2, hb2, dt);
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
            const int64_t* b = l.base + int64_t(run) * l.ntaps;
            const double v   = upwind_value<DIM>(b, x, iv[4], u, .5 * a0, .5 * fabs(a0), .5 * a1, .5 * fabs(a1), .5 * a2, .5 * fabs(a2), dt);
            if constexpr (NT)
            {
                __builtin_nontemporal_store(v, unp1 + b[0] + x);
            }
            else
            {
                unp1[b[0] + x] = v;
            }
        }

        // make_convection_upwind(v)(u): one thread per output slot, adding its
        // ordered contributions (a pair term is (c*u0 + c*u1), added at once)
        __global__ void k_conv_gather(const int64_t* __restrict__ row, const int64_t* __restrict__ ptr, const int64_t* __restrict__ src,
                                      const double* __restrict__ coef, int64_t nrows, const double* __restrict__ u, double* __restrict__ out)
        {
            const int64_t r = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
            if (r >= nrows)
            {
                return;
            }
            double acc = 0;
            for (int64_t k = ptr[r]; k < ptr[r + 1]; ++k)
            {
                const int64_t s1 = src[2 * k + 1];
                if (s1 < 0)
                {
                    acc += coef[k] * u[src[2 * k]];
                }
                else
                {
                    acc += coef[k] * u[src[2 * k]] + coef[k] * u[s1];
                }
            }
            out[row[r]] = acc;
        }

        // scalar non-linear upwind flux of f(u) = u * u
        // (schemes/fv/operators/convection_nonlin.hpp:54-67)
        __device__ __forceinline__ double burgers_flux(double ul, double ur)
        {
            const double v = 0.5 * (ul + ur);
            return v >= 0 ? ul * ul : ur * ur;
        }

        __global__ void k_conv_burgers_gather(const int64_t* __restrict__ row, const int64_t* __restrict__ ptr,
                                              const int64_t* __restrict__ src, const double* __restrict__ coef, int64_t nrows,
                                              const double* __restrict__ u, double* __restrict__ out)
        {
            const int64_t r = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
            if (r >= nrows)
            {
                return;
            }
            double acc = 0;
            for (int64_t k = ptr[r]; k < ptr[r + 1]; ++k)
            {
                acc += burgers_flux(u[src[2 * k]], u[src[2 * k + 1]]) * coef[k];
            }
            out[row[r]] = acc;
        }

        // make_convection_upwind(velocity_field): term k of direction dir[k]
        // is (coef[k] * w) * u[upwind side], w = 0.5 * (v_d[left] + v_d[right])
        // (convection_lin.hpp:210-226)
        __global__ void k_conv_het_gather(const int64_t* __restrict__ row, const int64_t* __restrict__ ptr, const int64_t* __restrict__ src,
                                          const double* __restrict__ coef, const int8_t* __restrict__ dir, int64_t nrows,
                                          const double* __restrict__ u, const double* __restrict__ v0, const double* __restrict__ v1,
                                          const double* __restrict__ v2, double* __restrict__ out)
        {
            const int64_t r = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
            if (r >= nrows)
            {
                return;
            }
            double acc = 0;
            for (int64_t k = ptr[r]; k < ptr[r + 1]; ++k)
            {
                const int d        = dir[k];
                const double* v    = d == 0 ? v0 : (d == 1 ? v1 : v2);
                const int64_t sl   = src[2 * k], sr = src[2 * k + 1];
                const double w     = 0.5 * (v[sl] + v[sr]);
                const double value = w >= 0 ? u[sl] : u[sr];
                acc += (coef[k] * w) * value;
            }
            out[row[r]] = acc;
        }

        // The WENO5 flux of Jiang & Shu (1996) on five flux values centred on
        // f[2] (schemes/fv/operators/weno_impl.hpp:26-66): third-order
        // candidates (eq. 2.8), smoothness indicators (3.2-3.4), non-linear
        // weights from 0.1, 0.6, 0.3 with eps = 1e-6 (2.15-2.16), weighted sum
        // (2.10), in the reference's operation order; its pow(x, 2) is the
        // correctly rounded square x * x.
        __device__ __forceinline__ double weno5_flux(double f0, double f1, double f2, double f3, double f4)
        {
            const double eps = 1e-6;
            const double q0  = 1. / 3 * f0 - 7. / 6 * f1 + 11. / 6 * f2;
            const double q1  = -1. / 6 * f1 + 5. / 6 * f2 + 1. / 3 * f3;
            const double q2  = 1. / 3 * f2 + 5. / 6 * f3 - 1. / 6 * f4;
            const double b0 = f0 - 2 * f1 + f2, c0 = f0 - 4 * f1 + 3 * f2;
            const double b1 = f1 - 2 * f2 + f3, c1 = f1 - f3;
            const double b2 = f2 - 2 * f3 + f4, c2 = 3 * f2 - 4 * f3 + f4;
            const double is0 = 13. / 12 * (b0 * b0) + 1. / 4 * (c0 * c0);
            const double is1 = 13. / 12 * (b1 * b1) + 1. / 4 * (c1 * c1);
            const double is2 = 13. / 12 * (b2 * b2) + 1. / 4 * (c2 * c2);
            const double e0 = eps + is0, e1 = eps + is1, e2 = eps + is2;
            const double a0  = 0.1 / (e0 * e0);
            const double a1  = 0.6 / (e1 * e1);
            const double a2  = 0.3 / (e2 * e2);
            const double sum = a0 + a1 + a2;
            return a0 / sum * q0 + a1 / sum * q1 + a2 / sum * q2;
        }

        // make_convection_weno5(velocity) (convection_lin.hpp:95-178): term k of
        // direction d = dir[k] is coef[k] * F with F the WENO5 flux of
        // v_d * (u[s0..s4]) for v_d >= 0, else of v_d * (u[s5], u[s4], ..., u[s1])
        __global__ __launch_bounds__(BLOCK) void k_conv_weno5_gather(const int64_t* __restrict__ row, const int64_t* __restrict__ ptr,
                                                                     const int64_t* __restrict__ src, const double* __restrict__ coef,
                                                                     const int8_t* __restrict__ dir, int64_t nrows, double v0, double v1,
                                                                     double v2, const double* __restrict__ u, double* __restrict__ out)
        {
            const int64_t r = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
            if (r >= nrows)
            {
                return;
            }
            double acc = 0;
            for (int64_t k = ptr[r]; k < ptr[r + 1]; ++k)
            {
                const int d          = dir[k];
                const double vd      = d == 0 ? v0 : (d == 1 ? v1 : v2);
                const int64_t* s     = src + 6 * k;
                double w[6];
                for (int j = 0; j < 6; ++j)
                {
                    w[j] = u[s[j]];
                }
                double F;
                if (vd >= 0)
                {
                    F = weno5_flux(w[0] * vd, w[1] * vd, w[2] * vd, w[3] * vd, w[4] * vd);
                }
                else
                {
                    F = weno5_flux(w[5] * vd, w[4] * vd, w[3] * vd, w[2] * vd, w[1] * vd);
                }
                acc += F * coef[k];
            }
            out[row[r]] = acc;
        }

        // the WENO5 flux of direction d from six values at base + j * stride
        __device__ __forceinline__ double weno5_at(const double* __restrict__ base, int64_t stride, double vd)
        {
            double w[6];
            for (int j = 0; j < 6; ++j)
            {
                w[j] = base[j * stride];
            }
            return vd >= 0 ? weno5_flux(w[0] * vd, w[1] * vd, w[2] * vd, w[3] * vd, w[4] * vd)
                           : weno5_flux(w[5] * vd, w[4] * vd, w[3] * vd, w[2] * vd, w[1] * vd);
        }

        // make_convection_weno5 on a mesh of one level whose storage is the dense
        // box of side W = n + 6 (ghost width 3, rows (z, y) in order, x fastest):
        // one thread per cell, no lists.  Per direction the cell takes the
        // flux of its left interface times -fa and of its right one times fa,
        // in the order the interface walk gives them: left then right, except
        // at the low boundary, whose boundary interface comes after the
        // interior ones (the same sums as k_conv_weno5_gather, bit for bit).
        template <int DIM>
        __global__ __launch_bounds__(BLOCK) void k_weno5_dense(const double* __restrict__ u, double* __restrict__ out, int n, int64_t W,
                                                               double v0, double v1, double v2, double fa)
        {
            const int64_t t     = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
            const int64_t total = DIM == 3 ? int64_t(n) * n * n : int64_t(n) * n;
            if (t >= total)
            {
                return;
            }
            const int x = int(t % n), y = int((t / n) % n), z = DIM == 3 ? int(t / (int64_t(n) * n)) : 0;
            const int g = 3;
            const int64_t c = DIM == 3 ? ((int64_t(z + g) * W + (y + g)) * W + (x + g)) : (int64_t(y + g) * W + (x + g));
            double acc = 0;
            for (int d = 0; d < DIM; ++d)
            {
                const int64_t stride = d == 0 ? 1 : (d == 1 ? W : W * W);
                const int coord      = d == 0 ? x : (d == 1 ? y : z);
                const double vd      = d == 0 ? v0 : (d == 1 ? v1 : v2);
                const double fl      = weno5_at(u + c - 3 * stride, stride, vd);
                const double fr      = weno5_at(u + c - 2 * stride, stride, vd);
                if (coord == 0)
                {
                    acc += fr * fa;
                    acc += fl * -fa;
                }
                else
                {
                    acc += fl * -fa;
                    acc += fr * fa;
                }
            }
            out[c] = acc;
        }

        // the 2D dense kernel on tiles of 64 x 4 cells: each interface flux is
        // evaluated once (every cell its right and upper interfaces, the tile's
        // first column and row also the ones below and to the left) and shared
        // through LDS; the fluxes and the order of the sums are those of
        // k_weno5_dense<2>, so the results are the same bits.  n >= 64.
        __global__ __launch_bounds__(256) void k_weno5_dense2d_tiled(const double* __restrict__ u, double* __restrict__ out, int n,
                                                                     int64_t W, double v0, double v1, double fa)
        {
            __shared__ double fx[4][65];
            __shared__ double fy[5][64];
            const int tx = int(threadIdx.x) & 63, ty = int(threadIdx.x) >> 6;
            const int x = int(blockIdx.x) * 64 + tx, y = int(blockIdx.y) * 4 + ty;
            const int g     = 3;
            const int64_t c = int64_t(y + g) * W + (x + g);
            fx[ty][tx + 1]  = weno5_at(u + c - 2, 1, v0);
            fy[ty + 1][tx]  = weno5_at(u + c - 2 * W, W, v1);
            if (tx == 0)
            {
                fx[ty][0] = weno5_at(u + c - 3, 1, v0);
            }
            if (ty == 0)
            {
                fy[0][tx] = weno5_at(u + c - 3 * W, W, v1);
            }
            __syncthreads();
            double acc = 0;
            const double flx = fx[ty][tx], frx = fx[ty][tx + 1];
            if (x == 0)
            {
                acc += frx * fa;
                acc += flx * -fa;
            }
            else
            {
                acc += flx * -fa;
                acc += frx * fa;
            }
            const double fly = fy[ty][tx], fry = fy[ty + 1][tx];
            if (y == 0)
            {
                acc += fry * fa;
                acc += fly * -fa;
            }
            else
            {
                acc += fly * -fa;
                acc += fry * fa;
            }
            out[c] = acc;
        }

        __device__ __forceinline__ int64_t readlane64(int64_t v, int lane)
        {
            const uint64_t w  = uint64_t(v);
            const uint32_t lo = uint32_t(__builtin_amdgcn_readlane(int32_t(uint32_t(w)), lane));
            const uint32_t hi = uint32_t(__builtin_amdgcn_readlane(int32_t(uint32_t(w >> 32)), lane));
            return int64_t((uint64_t(hi) << 32) | lo);
        }

        // lane i <- lane i - 1 (CTRL 0x138, wave_shr:1) or lane i + 1 (0x130,
        // wave_shl:1) of a double; lanes shifted in from outside the wave get 0
        template <int CTRL>
        __device__ __forceinline__ double wave_shift(double v)
        {
            const uint64_t w = uint64_t(__double_as_longlong(v));
            const int lo     = __builtin_amdgcn_update_dpp(0, int(uint32_t(w)), CTRL, 0xF, 0xF, false);
            const int hi     = __builtin_amdgcn_update_dpp(0, int(uint32_t(w >> 32)), CTRL, 0xF, 0xF, false);
            return __longlong_as_double(int64_t((uint64_t(uint32_t(hi)) << 32) | uint32_t(lo)));
        }

        // Upwind (2D) by column walks: a wavefront owns 64 columns of a stack of
        // up to 60 runs with one x range at rows y, y+1, ...  Lane k holds the
        // storage base of row k of the walk (the row below, the runs' rows, the
        // row above), read by every lane with readlane, so the walk carries no
        // dependent loads; walking up, the row above becomes the centre row and
        // the centre row the row below, so each row is fetched once per walk.
        // Rows are fetched G at a time ahead of their arithmetic; PIPE: the next
        // G rows are fetched before the arithmetic of the current G.  SHIFT: the
        // x - 1 / x + 1 neighbours come from the centre row's registers of the
        // neighbouring lanes; the two values beyond the strip are loaded by two
        // lanes (`ee`): lane 0 the left one, lane min(last + 1, 63) the right one
        // (last: the last active lane).  Else each lane loads both neighbours.
        // NTL: the rows above are read with non-temporal loads.  Same arithmetic
        // as k_upwind.
        template <int G, bool SHIFT, bool NTL = false>
        struct WalkRows
        {
            double uu[G], el[G], er[SHIFT ? 1 : G];

            // centre rows k + 1 .. k + G of the walk, rows above k + 2 .. k + G + 1
            __device__ __forceinline__ void load(const double* __restrict__ u, int64_t rb, int32_t k, int32_t n, int32_t x, int lane,
                                                 int last, bool active)
            {
                const int edge = min(last + 1, 63);
#pragma unroll
                for (int g = 0; g < G; ++g)
                {
                    uu[g] = 0;
                    el[g] = 0;
                    if constexpr (!SHIFT)
                    {
                        er[g] = 0;
                    }
                    if (G == 1 || k + g < n)
                    {
                        const int64_t bu = readlane64(rb, k + g + 2);
                        const int64_t bc = readlane64(rb, k + g + 1);
                        if (active)
                        {
                            if constexpr (NTL)
                            {
                                uu[g] = __builtin_nontemporal_load(u + bu + x);
                            }
                            else
                            {
                                uu[g] = u[bu + x];
                            }
                        }
                        if constexpr (SHIFT)
                        {
                            // x - 1 of lane 0, or x + 1 of the last active lane
                            if (lane == 0 || lane == edge)
                            {
                                el[g] = u[bc + (lane == 0 ? x - 1 : x - lane + last + 1)];
                            }
                        }
                        else if (active)
                        {
                            el[g] = u[bc + x - 1];
                            er[g] = u[bc + x + 1];
                        }
                    }
                }
            }
        };

        // one column walk t (wave-uniform)
        template <bool NT, int G, bool SHIFT, bool PIPE, bool NTL>
        __device__ __forceinline__ void upwind_walk(const List& l, const double* __restrict__ u, double* __restrict__ unp1, double a0,
                                                    double a1, double dt, int64_t t)
        {
            // one load: the walk's row bases and, in its last two slots, its data
            const int st      = l.walk_stride;
            const int lane    = int(threadIdx.x & 63);
            const int64_t rb  = lane < st ? l.walk_rows[t * st + lane] : 0;
            const int64_t m0  = readlane64(rb, st - 2), m1 = readlane64(rb, st - 1);
            const int32_t xs  = int32_t(m0 >> 32), n = int32_t(m0);
            const int32_t xe  = int32_t(m1 >> 32), level = int32_t(m1);
            const int32_t x   = xs + lane;
            const bool active = x < xe; // every run of the walk has the same x range
            const int last    = min(63, xe - 1 - xs); // last active lane
            const double ha0 = .5 * a0, hb0 = .5 * fabs(a0), ha1 = .5 * a1, hb1 = .5 * fabs(a1);
            const int64_t bd0 = readlane64(rb, 0), bc0 = readlane64(rb, 1);
            double ud = 0, uc = 0;
            if (active)
            {
                ud = u[bd0 + x];
                uc = u[bc0 + x];
            }

            WalkRows<G, SHIFT, NTL> cur;
            if constexpr (PIPE)
            {
                cur.load(u, rb, 0, n, x, lane, last, active);
            }
            for (int32_t k = 0; k < n; k += G)
            {
                WalkRows<G, SHIFT, NTL> nxt;
                if constexpr (PIPE)
                {
                    if (k + G < n)
                    {
                        nxt.load(u, rb, k + G, n, x, lane, last, active);
                    }
                }
                else
                {
                    cur.load(u, rb, k, n, x, lane, last, active);
                }
#pragma unroll
                for (int g = 0; g < G; ++g)
                {
                    if (G == 1 || k + g < n)
                    {
                        const int64_t bc = readlane64(rb, k + g + 1);
                        double ul, ur;
                        if constexpr (SHIFT)
                        {
                            // lane last + 1 (< 64) holds x + 1 of the last active lane in el
                            const double ee = cur.el[g];
                            const double sl = wave_shift<0x138>(uc);
                            const double sr = wave_shift<0x130>(lane == last + 1 ? ee : uc);
                            ul              = lane == 0 ? ee : sl;
                            ur              = lane == 63 ? ee : sr;
                        }
                        else
                        {
                            ul = cur.el[g];
                            ur = cur.er[g];
                        }
                        if (active)
                        {
                            const double v = upwind_vals_2d(uc, ul, ur, ud, cur.uu[g], level, ha0, hb0, ha1, hb1, dt);
                            if constexpr (NT)
                            {
                                __builtin_nontemporal_store(v, unp1 + bc + x);
                            }
                            else
                            {
                                unp1[bc + x] = v;
                            }
                        }
                        ud = uc;
                        uc = cur.uu[g];
                    }
                }
                if constexpr (PIPE)
                {
                    cur = nxt;
                }
            }
        }

        // one wave per walk
        template <bool NT, int G, bool SHIFT, bool PIPE, bool NTL = false>
        __global__ __launch_bounds__(BLOCK) void k_upwind_walk(List l, const double* __restrict__ u, double* __restrict__ unp1, double a0,
                                                               double a1, double dt, unsigned nlogical)
        {
            const int64_t blk = xcd_tile(nlogical);
            if (blk < 0)
            {
                return;
            }
            const int64_t t = blk * (BLOCK / 64) + (threadIdx.x >> 6);
            if (t >= l.nwalks)
            {
                return;
            }
            upwind_walk<NT, G, SHIFT, PIPE, NTL>(l, u, unp1, a0, a1, dt, t);
        }

        __device__ __forceinline__ double readlane_f64(double v, int lane)
        {
            return __longlong_as_double(readlane64(__double_as_longlong(v), lane));
        }

        // Two adjacent strips of one band (walks 2p and 2p + 1: same runs, x
        // ranges [xs, xs + 64) and [xs + 64, ...)) as one 128-column walk: lane
        // k holds columns xs + k (half A) and xs + 64 + k (half B).  The
        // neighbours across the halves' seam come from a readlane, so only the
        // two ends of the 128 columns are loaded apart (one load, two lanes)
        // and the strip edge between the halves costs nothing.  A walk with no
        // adjacent partner runs as a single strip.  G rows ahead, non-temporal
        // loads of the rows above and non-temporal stores, as k_upwind_walk.
        template <int G, int GF = 16>
        __global__ __launch_bounds__(BLOCK) void k_upwind_walk_pair(List l, const double* __restrict__ u, double* __restrict__ unp1,
                                                                    double a0, double a1, double dt, unsigned nlogical)
        {
            const int64_t blk = xcd_tile(nlogical);
            if (blk < 0)
            {
                return;
            }
            const int64_t t = 2 * (blk * (BLOCK / 64) + (threadIdx.x >> 6));
            if (t >= l.nwalks)
            {
                return;
            }
            // the slots of both walks in one load each (walk t + 1 only if it exists)
            const int st      = l.walk_stride;
            const int lane    = int(threadIdx.x & 63);
            const bool has2   = t + 1 < l.nwalks;
            const int64_t rb  = lane < st ? l.walk_rows[t * st + lane] : 0;
            const int64_t rb2 = has2 && lane < st ? l.walk_rows[(t + 1) * st + lane] : 0;
            const int64_t m0  = readlane64(rb, st - 2), m1 = readlane64(rb, st - 1);
            const int32_t xs  = int32_t(m0 >> 32), n = int32_t(m0);
            const int32_t xe  = int32_t(m1 >> 32), level = int32_t(m1);
            // adjacent strips of one stack: walk t + 1 starts 64 columns on, has
            // the same rows (same first row base, same x range end, same count)
            bool pair = false;
            if (has2)
            {
                const int64_t p0 = readlane64(rb2, st - 2), p1 = readlane64(rb2, st - 1);
                pair = int32_t(p0 >> 32) == xs + 64 && int32_t(p0) == n && int32_t(p1 >> 32) == xe &&
                       readlane64(rb2, 1) == readlane64(rb, 1);
            }
            if (!pair)
            {
                upwind_walk<true, GF, true, false, true>(l, u, unp1, a0, a1, dt, t);
                if (t + 1 < l.nwalks)
                {
                    upwind_walk<true, GF, true, false, true>(l, u, unp1, a0, a1, dt, t + 1);
                }
                return;
            }
            const int32_t xa = xs + lane, xb = xs + 64 + lane; // half A is whole: walk 2p + 1 exists
            const bool act_b  = xb < xe;
            const int last_b  = min(63, xe - 1 - xs - 64); // last active lane of half B
            const int edge    = min(last_b + 1, 63);       // lane loading x + 1 of half B's last lane
            const double ha0 = .5 * a0, hb0 = .5 * fabs(a0), ha1 = .5 * a1, hb1 = .5 * fabs(a1);
            const int64_t bd0 = readlane64(rb, 0), bc0 = readlane64(rb, 1);
            double uda = u[bd0 + xa], uca = u[bc0 + xa];
            double udb = 0, ucb = 0;
            if (act_b)
            {
                udb = u[bd0 + xb];
                ucb = u[bc0 + xb];
            }
            for (int32_t k = 0; k < n; k += G)
            {
                double ua[G], ub[G], ee[G];
#pragma unroll
                for (int g = 0; g < G; ++g)
                {
                    ua[g] = 0;
                    ub[g] = 0;
                    ee[g] = 0;
                    if (k + g < n)
                    {
                        const int64_t bu = readlane64(rb, k + g + 2);
                        const int64_t bc = readlane64(rb, k + g + 1);
                        ua[g]            = __builtin_nontemporal_load(u + bu + xa);
                        if (act_b)
                        {
                            ub[g] = __builtin_nontemporal_load(u + bu + xb);
                        }
                        // x - 1 of half A's lane 0, or x + 1 of half B's last lane
                        if (lane == 0 || lane == edge)
                        {
                            ee[g] = u[bc + (lane == 0 ? xs - 1 : xs + 64 + last_b + 1)];
                        }
                    }
                }
#pragma unroll
                for (int g = 0; g < G; ++g)
                {
                    if (k + g < n)
                    {
                        const int64_t bc  = readlane64(rb, k + g + 1);
                        const double a63  = readlane_f64(uca, 63);
                        const double b0   = readlane_f64(ucb, 0);
                        const double sla  = wave_shift<0x138>(uca);
                        const double sra  = wave_shift<0x130>(uca);
                        const double slb  = wave_shift<0x138>(ucb);
                        const double srb  = wave_shift<0x130>(lane == last_b + 1 ? ee[g] : ucb);
                        const double ula  = lane == 0 ? ee[g] : sla;
                        const double ura  = lane == 63 ? b0 : sra;
                        const double ulb  = lane == 0 ? a63 : slb;
                        const double urb  = lane == 63 ? ee[g] : srb;
                        const double va = upwind_vals_2d(uca, ula, ura, uda, ua[g], level, ha0, hb0, ha1, hb1, dt);
                        const double vb = act_b ? upwind_vals_2d(ucb, ulb, urb, udb, ub[g], level, ha0, hb0, ha1, hb1, dt) : 0.0;
                        __builtin_nontemporal_store(va, unp1 + bc + xa);
                        if (act_b)
                        {
                            __builtin_nontemporal_store(vb, unp1 + bc + xb);
                        }
                        uda = uca;
                        uca = ua[g];
                        udb = ucb;
                        ucb = ub[g];
                    }
                }
            }
        }

        __global__ void k_walk_rows(List l)
        {
            const int64_t w = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
            if (w < l.nwalks)
            {
                op_walk_rows(l, w);
            }
        }

#define SMR_CELL_KERNEL(NAME, BODY, ...)                                                                                             \
    __global__ __launch_bounds__(BLOCK) void NAME(List l, int64_t c0, int64_t c1, __VA_ARGS__)                                        \
    {                                                                                                                                  \
        const int64_t c = c0 + int64_t(blockIdx.x) * BLOCK + threadIdx.x;                                                             \
        if (c >= c1)                                                                                                                   \
        {                                                                                                                              \
            return;                                                                                                                    \
        }                                                                                                                              \
        BODY;                                                                                                                          \
    }

        template <int DIM>
        SMR_CELL_KERNEL(k_project, (op_project<DIM>(l, c, dst, src)), double* dst, const double* src)
        template <int DIM, int S>
        SMR_CELL_KERNEL(k_predict_fine, (op_predict_fine<DIM, S>(l, c, u)), double* u)
        template <int DIM, int S>
        SMR_CELL_KERNEL(k_detail, (op_detail<DIM, S>(l, c, u, det)), const double* u, double* det)
        template <int DIM>
        SMR_CELL_KERNEL(k_criteria, (op_criteria<DIM>(l, c, det, tag, thr.eps, min_level, max_level, ncomp, cstride)), const double* det,
                        uint8_t* tag, Thresholds thr, int min_level, int max_level, int ncomp, int64_t cstride)
        template <int DIM>
        SMR_CELL_KERNEL(k_maximum, (op_maximum<DIM>(l, c, tag)), uint8_t* tag)
        SMR_CELL_KERNEL(k_tag_init, (op_tag_init(l, c, tag)), uint8_t* tag)
        SMR_CELL_KERNEL(k_sub_scaled, ({
                            const Cell cell  = locate(l, c);
                            const int64_t i  = bases(l, cell.run)[0] + cell.x;
                            out[i]           = u[i] - a * v[i];
                        }),
                        const double* u, double a, const double* v, double* out)
        SMR_CELL_KERNEL(k_hand_loop, ({
                            const Cell cell         = locate(l, c);
                            const int64_t* b        = bases(l, cell.run);
                            unp1[b[0] + cell.x] = hand_loop_value(b, cell.x, cell.level, u, dt);
                        }),
                        const double* u, double* unp1, double dt)
        SMR_CELL_KERNEL(k_upwind_burgers, ({
                            const Cell cell         = locate(l, c);
                            const int64_t* b        = bases(l, cell.run);
                            unp1[b[0] + cell.x] = burgers_value(b, cell.x, cell.level, u, k0, k1, dt);
                        }),
                        const double* u, double* unp1, double k0, double k1, double dt)
        // one flag for the whole list: one store per wave, from the lowest
        // active lane that saw a change
        SMR_CELL_KERNEL(k_tag_change, ({
                            const uint64_t m = __ballot(tag_changes_mesh(l, c, tag, min_level, max_level));
                            if (m != 0 && int(threadIdx.x & 63) == __ffsll(static_cast<unsigned long long>(m)) - 1)
                            {
                                flag[0] = 1;
                            }
                        }),
                        const uint8_t* tag, int min_level, int max_level, int* flag)
        SMR_CELL_KERNEL(k_copy, (op_copy(l, c, dst, src)), double* dst, const double* src)
        // keep_boundary_refined (mr/adapt.hpp:245-272): keep (assigned) on the
        // cells within r of a face of the level's domain [0, n)^dim
        SMR_CELL_KERNEL(k_keep_boundary, ({
                            const Cell cell = locate(l, c);
                            bool near       = cell.x < r || cell.x >= n - r || cell.y < r || cell.y >= n - r;
                            if (dim == 3)
                            {
                                near = near || cell.z < r || cell.z >= n - r;
                            }
                            if (near)
                            {
                                tag[bases(l, cell.run)[0] + cell.x] = 1;
                            }
                        }),
                        uint8_t* tag, int dim, int n, int r)
        template <int DIM, int S>
        SMR_CELL_KERNEL(k_predict_children, (op_predict_children<DIM, S>(l, c, dst, src)), double* dst, const double* src)
#undef SMR_CELL_KERNEL

        // max |u| over cells [c0, c1) of a list into *mx (atomic max of the bit
        // patterns, which order as the non-negative doubles do; a NaN value is
        // passed over like std::max(m, NaN) does)
        __global__ __launch_bounds__(BLOCK) void k_abs_max(const List l, int64_t c0, int64_t c1, const double* __restrict__ u,
                                                           unsigned long long* mx)
        {
            __shared__ double part[BLOCK / 64];
            double m = 0;
            for (int64_t c = c0 + int64_t(blockIdx.x) * blockDim.x + threadIdx.x; c < c1; c += int64_t(gridDim.x) * blockDim.x)
            {
                const Cell cell = locate(l, c);
                const double a  = fabs(u[bases(l, cell.run)[0] + cell.x]);
                m               = a > m ? a : m;
            }
            for (int off = 32; off > 0; off >>= 1)
            {
                const double o = __shfl_xor(m, off, 64);
                m              = o > m ? o : m;
            }
            if ((threadIdx.x & 63) == 0)
            {
                part[threadIdx.x >> 6] = m;
            }
            __syncthreads();
            if (threadIdx.x == 0)
            {
                for (int w = 1; w < BLOCK / 64; ++w)
                {
                    m = part[w] > m ? part[w] : m;
                }
                atomicMax(mx, static_cast<unsigned long long>(__double_as_longlong(m)));
            }
        }

        // compute_relative_detail's scaling (mr/rel_detail.hpp:89-110): the max
        // folded from the smallest normal double, 1 below the machine epsilon,
        // every detail times its inverse
        __global__ __launch_bounds__(BLOCK) void k_scale_relative(double* __restrict__ d, int64_t n, const unsigned long long* mx)
        {
            const int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
            if (i >= n)
            {
                return;
            }
            double m = __longlong_as_double(static_cast<long long>(*mx));
            m        = m > DBL_MIN ? m : DBL_MIN;
            if (m < DBL_EPSILON)
            {
                m = 1.0;
            }
            const double inv = 1. / m;
            d[i] *= inv;
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
            op_recipe(dst, op, sp, src, r, u, has_bc, bc_value);
        }

        inline unsigned nblocks(int64_t n)
        {
            return unsigned((n + BLOCK - 1) / BLOCK);
        }

        constexpr int SEQ_BLOCK = 1024;

        // the items of a chain of small stages in order; the block barrier
        // orders every write of an item before the reads of the next (global
        // memory written by the block is visible to the block after it)
        template <int DIM, int S>
        __global__ __launch_bounds__(SEQ_BLOCK) void k_sequence(const SeqItem* items, int i0, int i1, List l_proj, List l_pred, List l_max,
                                                                double* u, uint8_t* tag, int has_bc, double bc_value)
        {
            for (int k = i0; k < i1; ++k)
            {
                const SeqItem it = items[k];
                for (int64_t c = it.c0 + threadIdx.x; c < it.c1; c += SEQ_BLOCK)
                {
                    switch (it.op)
                    {
                        case SEQ_RECIPES: op_recipe(it.dst, it.rop, it.sp, it.src, c, u, has_bc, bc_value); break;
                        case SEQ_PROJECT: op_project<DIM>(l_proj, c, u, u); break;
                        case SEQ_PREDICT: op_predict_fine<DIM, S>(l_pred, c, u); break;
                        case SEQ_MAXIMUM: op_maximum<DIM>(l_max, c, tag); break;
                        default: break;
                    }
                }
                __syncthreads();
            }
        }

        // p[0, n) = 0 (a kernel, not hipMemsetAsync, so that a caller may
        // capture the launch in a HIP graph)
        template <class T>
        __global__ __launch_bounds__(BLOCK) void k_zero(T* p, int64_t n)
        {
            const int64_t i = int64_t(blockIdx.x) * BLOCK + threadIdx.x;
            if (i < n)
            {
                p[i] = T(0);
            }
        }

        // launch shapes of the upwind kernel: cells per thread (0: wave tiles,
        // -2: column walks (2D), -1: column walks (2D) or wave tiles (3D) when at
        // least 3/4 of the lanes hold a cell, else 4 cells per thread), XCD
        // remap, non-temporal stores
        // (column walks: rows fetched ahead per wave, x +- 1 by lane shifts)
        struct UpwindVariant
        {
            int id, cpt;
            bool remap, nt;
            int group  = 1;
            bool shift = false;
            bool pipe  = false;
        };
        constexpr UpwindVariant kUpwindVariants[] = {
            {0, 1, false, false},       {1, 1, true, false},        {2, 2, false, false},       {3, 2, true, false},
            {4, 4, false, false},       {5, 4, true, true},         {6, 1, false, true},        {7, 8, false, false},
            {8, 0, true, true},         {9, 0, true, false},        {10, -1, true, true},       {11, -2, true, true},
            {12, -2, true, false},      {13, -2, true, true, 2},    {14, -2, true, true, 4},    {15, -2, true, true, 8},
            {16, -2, true, true, 1, true}, {17, -2, true, true, 4, true}, {18, -2, true, true, 8, true},
            {19, -2, true, true, 16, true}, {20, -2, true, false, 8, true}, {21, -2, true, true, 4, true, true},
            {22, -2, true, true, 8, true, true}, {23, -2, true, true, 2, true, true}, {24, -2, true, true, 16, true},
            {25, -2, true, true, 16, true},   {26, -2, true, true, 16, true},   {27, -2, true, true, 16, true},
            {28, -2, true, true, 16, true}};
        constexpr int kNumUpwindVariants    = 29;
        constexpr int kDefaultUpwindVariant = 10;
        int g_upwind_variant                = -1;

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

    void conv_gather(const int64_t* row, const int64_t* ptr, const int64_t* src, const double* coef, int64_t nrows, const double* u,
                     double* out, hipStream_t s)
    {
        if (nrows > 0)
        {
            hipLaunchKernelGGL(k_conv_gather, dim3(nblocks(nrows)), dim3(BLOCK), 0, s, row, ptr, src, coef, nrows, u, out);
        }
    }

    void conv_burgers_gather(const int64_t* row, const int64_t* ptr, const int64_t* src, const double* coef, int64_t nrows, const double* u,
                     double* out, hipStream_t s)
    {
        if (nrows > 0)
        {
            hipLaunchKernelGGL(k_conv_burgers_gather, dim3(nblocks(nrows)), dim3(BLOCK), 0, s, row, ptr, src, coef, nrows, u, out);
        }
    }

    void conv_het_gather(const int64_t* row, const int64_t* ptr, const int64_t* src, const double* coef, const int8_t* dir, int64_t nrows,
                         const double* u, const double* const* velocity, double* out, hipStream_t s)
    {
        if (nrows > 0)
        {
            hipLaunchKernelGGL(k_conv_het_gather, dim3(nblocks(nrows)), dim3(BLOCK), 0, s, row, ptr, src, coef, dir, nrows, u, velocity[0],
                               velocity[1], velocity[2], out);
        }
    }

    void conv_weno5_gather(const int64_t* row, const int64_t* ptr, const int64_t* src, const double* coef, const int8_t* dir, int64_t nrows,
                           const double* v, const double* u, double* out, hipStream_t s)
    {
        if (nrows > 0)
        {
            hipLaunchKernelGGL(k_conv_weno5_gather, dim3(nblocks(nrows)), dim3(BLOCK), 0, s, row, ptr, src, coef, dir, nrows, v[0], v[1], v[2], u,
                               out);
        }
    }

    void keep_boundary(const List& cells, int64_t c0, int64_t c1, uint8_t* tag, int dim, int n, int r, hipStream_t s)
    {
        if (c1 > c0)
        {
            hipLaunchKernelGGL(k_keep_boundary, dim3(nblocks(c1 - c0)), dim3(BLOCK), 0, s, cells, c0, c1, tag, dim, n, r);
        }
    }

    void relative_detail(const List& cells, int64_t c0, int64_t c1, const double* u, double* det, int64_t n, unsigned long long* mx,
                         hipStream_t s)
    {
        zero_f64(reinterpret_cast<double*>(mx), 1, s);
        if (c1 > c0)
        {
            const int64_t nb = std::min<int64_t>(nblocks(c1 - c0), 1024);
            hipLaunchKernelGGL(k_abs_max, dim3(nb), dim3(BLOCK), 0, s, cells, c0, c1, u, mx);
        }
        if (n > 0)
        {
            hipLaunchKernelGGL(k_scale_relative, dim3(nblocks(n)), dim3(BLOCK), 0, s, det, n, mx);
        }
    }

    void weno5_dense(int dim, int n, const double* v, double fa, const double* u, double* out, hipStream_t s)
    {
        const int64_t W     = int64_t(n) + 6;
        const int64_t total = dim == 3 ? int64_t(n) * n * n : int64_t(n) * n;
        if (dim == 3)
        {
            hipLaunchKernelGGL(k_weno5_dense<3>, dim3(nblocks(total)), dim3(BLOCK), 0, s, u, out, n, W, v[0], v[1], v[2], fa);
        }
        else if (n >= 64 && !std::getenv("SMR_WENO5_UNTILED"))
        {
            hipLaunchKernelGGL(k_weno5_dense2d_tiled, dim3(unsigned(n / 64), unsigned(n / 4)), dim3(256), 0, s, u, out, n, W, v[0], v[1], fa);
        }
        else
        {
            hipLaunchKernelGGL(k_weno5_dense<2>, dim3(nblocks(total)), dim3(BLOCK), 0, s, u, out, n, W, v[0], v[1], v[2], fa);
        }
    }

    void walk_rows(const List& l, hipStream_t s)
    {
        if (l.nwalks > 0)
        {
            hipLaunchKernelGGL(k_walk_rows, dim3(nblocks(l.nwalks)), dim3(BLOCK), 0, s, l);
        }
    }

    void resolve(const List& l, const Table& t0, const Table& t1, int* err, hipStream_t s)
    {
        const int64_t n = resolve_range(l);
        if (n > 0)
        {
            hipLaunchKernelGGL(k_resolve, dim3(nblocks(n)), dim3(BLOCK), 0, s, l, t0, t1, err);
        }
    }

    void resolve_batch(const ResolveBatch& b, int* err, hipStream_t s)
    {
        const int64_t n = b.n > 0 ? b.start[b.n] : 0;
        if (n > 0)
        {
            hipLaunchKernelGGL(k_resolve_batch, dim3(nblocks(n)), dim3(BLOCK), 0, s, b, err);
        }
    }

    void upwind(int dim, const List& l, int64_t c0, int64_t c1, const double* u, double* unp1, double a0, double a1, double a2, double dt,
                hipStream_t s)
    {
        if (c1 <= c0)
        {
            return;
        }
        UpwindVariant v = upwind_variant();
        const bool whole = c0 == 0 && c1 == l.ncells;
        if (v.cpt == -1)
        {
            const bool dense = 4 * l.ncells >= 3 * 64 * l.ntiles;
            v                = !dense ? kUpwindVariants[5] : (dim == 2 && l.nwalks > 0 ? kUpwindVariants[28] : kUpwindVariants[8]);
        }
        if (v.cpt == -2)
        {
            if (dim == 2 && l.nwalks > 0 && whole)
            {
                const int64_t nblk   = (l.nwalks + BLOCK / 64 - 1) / (BLOCK / 64);
                const unsigned ngrid = unsigned((nblk + 7) / 8 * 8);
#define SMR_WALK_LAUNCH(N, G, S, P) \
    hipLaunchKernelGGL((k_upwind_walk<N, G, S, P>), dim3(ngrid), dim3(BLOCK), 0, s, l, u, unp1, a0, a1, dt, unsigned(nblk))
                switch (v.id)
                {
                    case 12: SMR_WALK_LAUNCH(false, 1, false, false); break;
                    case 13: SMR_WALK_LAUNCH(true, 2, false, false); break;
                    case 14: SMR_WALK_LAUNCH(true, 4, false, false); break;
                    case 15: SMR_WALK_LAUNCH(true, 8, false, false); break;
                    case 16: SMR_WALK_LAUNCH(true, 1, true, false); break;
                    case 17: SMR_WALK_LAUNCH(true, 4, true, false); break;
                    case 18: SMR_WALK_LAUNCH(true, 8, true, false); break;
                    case 19: SMR_WALK_LAUNCH(true, 16, true, false); break;
                    case 20: SMR_WALK_LAUNCH(false, 8, true, false); break;
                    case 21: SMR_WALK_LAUNCH(true, 4, true, true); break;
                    case 22: SMR_WALK_LAUNCH(true, 8, true, true); break;
                    case 23: SMR_WALK_LAUNCH(true, 2, true, true); break;
                    case 24:
                        hipLaunchKernelGGL((k_upwind_walk<true, 16, true, false, true>), dim3(ngrid), dim3(BLOCK), 0, s, l, u, unp1, a0, a1,
                                           dt, unsigned(nblk));
                        break;
                    case 25: // adjacent strips in pairs, 8, 4, 16, 16 (8 for single strips) rows ahead
                    case 26:
                    case 27:
                    case 28:
                    {
                        const int64_t np  = (l.nwalks + 1) / 2;
                        const int64_t nb  = (np + BLOCK / 64 - 1) / (BLOCK / 64);
                        const unsigned ng = unsigned((nb + 7) / 8 * 8);
                        if (v.id == 25)
                        {
                            hipLaunchKernelGGL((k_upwind_walk_pair<8>), dim3(ng), dim3(BLOCK), 0, s, l, u, unp1, a0, a1, dt, unsigned(nb));
                        }
                        else if (v.id == 26)
                        {
                            hipLaunchKernelGGL((k_upwind_walk_pair<4>), dim3(ng), dim3(BLOCK), 0, s, l, u, unp1, a0, a1, dt, unsigned(nb));
                        }
                        else if (v.id == 27)
                        {
                            hipLaunchKernelGGL((k_upwind_walk_pair<16>), dim3(ng), dim3(BLOCK), 0, s, l, u, unp1, a0, a1, dt, unsigned(nb));
                        }
                        else
                        {
                            hipLaunchKernelGGL((k_upwind_walk_pair<16, 8>), dim3(ng), dim3(BLOCK), 0, s, l, u, unp1, a0, a1, dt, unsigned(nb));
                        }
                        break;
                    }
                    default: SMR_WALK_LAUNCH(true, 1, false, false); break;
                }
#undef SMR_WALK_LAUNCH
                return;
            }
            v = kUpwindVariants[5];
        }
        if (v.cpt == 0 && c0 == 0 && c1 == l.ncells)
        {
            const int64_t nblk   = (l.ntiles + BLOCK / 64 - 1) / (BLOCK / 64);
            const unsigned ngrid = unsigned((nblk + 7) / 8 * 8);
            const unsigned nlog  = unsigned(nblk);
            const dim3 g(ngrid), bl(BLOCK);
            if (dim == 2)
            {
                if (v.nt)
                {
                    hipLaunchKernelGGL((k_upwind_tiles<2, true>), g, bl, 0, s, l, 0, l.ntiles, u, unp1, a0, a1, a2, dt, nlog);
                }
                else
                {
                    hipLaunchKernelGGL((k_upwind_tiles<2, false>), g, bl, 0, s, l, 0, l.ntiles, u, unp1, a0, a1, a2, dt, nlog);
                }
            }
            else
            {
                if (v.nt)
                {
                    hipLaunchKernelGGL((k_upwind_tiles<3, true>), g, bl, 0, s, l, 0, l.ntiles, u, unp1, a0, a1, a2, dt, nlog);
                }
                else
                {
                    hipLaunchKernelGGL((k_upwind_tiles<3, false>), g, bl, 0, s, l, 0, l.ntiles, u, unp1, a0, a1, a2, dt, nlog);
                }
            }
            return;
        }
        const int cpt        = v.cpt == 0 ? 4 : v.cpt;
        const int64_t ntiles = (c1 - c0 + BLOCK * cpt - 1) / (BLOCK * cpt);
        const unsigned ngrid = v.remap ? unsigned((ntiles + 7) / 8 * 8) : unsigned(ntiles);
        const unsigned nlog  = unsigned(ntiles);
        const dim3 g(ngrid), bl(BLOCK);
#define SMR_UPWIND_LAUNCH(D, C, R, N) hipLaunchKernelGGL((k_upwind<D, C, R, N>), g, bl, 0, s, l, c0, c1, u, unp1, a0, a1, a2, dt, nlog)
#define SMR_UPWIND_DISPATCH(D)                                                                                                         \
    switch (v.id)                                                                                                                      \
    {                                                                                                                                  \
        case 0: SMR_UPWIND_LAUNCH(D, 1, false, false); break;                                                                         \
        case 1: SMR_UPWIND_LAUNCH(D, 1, true, false); break;                                                                          \
        case 2: SMR_UPWIND_LAUNCH(D, 2, false, false); break;                                                                         \
        case 3: SMR_UPWIND_LAUNCH(D, 2, true, false); break;                                                                          \
        case 4: SMR_UPWIND_LAUNCH(D, 4, false, false); break;                                                                         \
        case 6: SMR_UPWIND_LAUNCH(D, 1, false, true); break;                                                                          \
        case 7: SMR_UPWIND_LAUNCH(D, 8, false, false); break;                                                                         \
        default: SMR_UPWIND_LAUNCH(D, 4, true, true); break;                                                                          \
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

#define SMR_DISPATCH_DIM(KERNEL, ...)                                                                                                  \
    if (dim == 2)                                                                                                                      \
    {                                                                                                                                  \
        hipLaunchKernelGGL((KERNEL<2>), dim3(nblocks(c1 - c0)), dim3(BLOCK), 0, s, l, c0, c1, __VA_ARGS__);                          \
    }                                                                                                                                  \
    else                                                                                                                               \
    {                                                                                                                                  \
        hipLaunchKernelGGL((KERNEL<3>), dim3(nblocks(c1 - c0)), dim3(BLOCK), 0, s, l, c0, c1, __VA_ARGS__);                          \
    }
#define SMR_DISPATCH_DIM_S(KERNEL, ...)                                                                                                \
    {                                                                                                                                  \
        const dim3 g(nblocks(c1 - c0)), bl(BLOCK);                                                                                     \
        if (dim == 2 && ps == 0)                                                                                                       \
        {                                                                                                                              \
            hipLaunchKernelGGL((KERNEL<2, 0>), g, bl, 0, s, l, c0, c1, __VA_ARGS__);                                                  \
        }                                                                                                                              \
        else if (dim == 2)                                                                                                             \
        {                                                                                                                              \
            hipLaunchKernelGGL((KERNEL<2, 1>), g, bl, 0, s, l, c0, c1, __VA_ARGS__);                                                  \
        }                                                                                                                              \
        else if (ps == 0)                                                                                                              \
        {                                                                                                                              \
            hipLaunchKernelGGL((KERNEL<3, 0>), g, bl, 0, s, l, c0, c1, __VA_ARGS__);                                                  \
        }                                                                                                                              \
        else                                                                                                                           \
        {                                                                                                                              \
            hipLaunchKernelGGL((KERNEL<3, 1>), g, bl, 0, s, l, c0, c1, __VA_ARGS__);                                                  \
        }                                                                                                                              \
    }

    void project(int dim, const List& l, int64_t c0, int64_t c1, double* dst, const double* src, hipStream_t s)
    {
        if (c1 > c0)
        {
            SMR_DISPATCH_DIM(k_project, dst, src)
        }
    }

    void predict_fine(int dim, int ps, const List& l, int64_t c0, int64_t c1, double* u, hipStream_t s)
    {
        if (c1 > c0)
        {
            SMR_DISPATCH_DIM_S(k_predict_fine, u)
        }
    }

    void detail(int dim, int ps, const List& l, int64_t c0, int64_t c1, const double* u, double* det, hipStream_t s)
    {
        if (c1 > c0)
        {
            SMR_DISPATCH_DIM_S(k_detail, u, det)
        }
    }

    void criteria(int dim, const List& l, int64_t c0, int64_t c1, const double* det, uint8_t* tag, const Thresholds& thr, int min_level,
                  int max_level, hipStream_t s, int ncomp, int64_t cstride)
    {
        if (c1 > c0)
        {
            SMR_DISPATCH_DIM(k_criteria, det, tag, thr, min_level, max_level, ncomp, cstride)
        }
    }

    void maximum(int dim, const List& l, int64_t c0, int64_t c1, uint8_t* tag, hipStream_t s)
    {
        if (c1 > c0)
        {
            SMR_DISPATCH_DIM(k_maximum, tag)
        }
    }

    void sub_scaled(const List& l, int64_t c0, int64_t c1, const double* u, double a, const double* v, double* out, hipStream_t s)
    {
        if (c1 > c0)
        {
            hipLaunchKernelGGL(k_sub_scaled, dim3(nblocks(c1 - c0)), dim3(BLOCK), 0, s, l, c0, c1, u, a, v, out);
        }
    }

    void hand_loop(const List& l, int64_t c0, int64_t c1, const double* u, double* unp1, double dt, hipStream_t s)
    {
        if (c1 > c0)
        {
            hipLaunchKernelGGL(k_hand_loop, dim3(nblocks(c1 - c0)), dim3(BLOCK), 0, s, l, c0, c1, u, unp1, dt);
        }
    }

    void upwind_scalar_burgers(const List& l, int64_t c0, int64_t c1, const double* u, double* unp1, double k0, double k1, double dt,
                               hipStream_t s)
    {
        if (c1 > c0)
        {
            hipLaunchKernelGGL(k_upwind_burgers, dim3(nblocks(c1 - c0)), dim3(BLOCK), 0, s, l, c0, c1, u, unp1, k0, k1, dt);
        }
    }

    void tag_change(const List& l, int64_t c0, int64_t c1, const uint8_t* tag, int min_level, int max_level, int* flag, hipStream_t s)
    {
        if (c1 > c0)
        {
            hipLaunchKernelGGL(k_tag_change, dim3(nblocks(c1 - c0)), dim3(BLOCK), 0, s, l, c0, c1, tag, min_level, max_level, flag);
        }
    }

    void zero_f64(double* p, int64_t n, hipStream_t s)
    {
        if (n > 0)
        {
            hipLaunchKernelGGL(k_zero<double>, dim3(nblocks(n)), dim3(BLOCK), 0, s, p, n);
        }
    }

    void zero_u8(uint8_t* p, int64_t n, hipStream_t s)
    {
        if (n > 0)
        {
            hipLaunchKernelGGL(k_zero<uint8_t>, dim3(nblocks(n)), dim3(BLOCK), 0, s, p, n);
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
        if (c1 > c0)
        {
            SMR_DISPATCH_DIM_S(k_predict_children, dst, src)
        }
    }
#undef SMR_DISPATCH_DIM
#undef SMR_DISPATCH_DIM_S

    void sequence(int dim, int ps, const SeqItem* items, int i0, int i1, const List& l_proj, const List& l_pred, const List& l_max, double* u,
                  uint8_t* tag, bool has_bc, double bc_value, hipStream_t s)
    {
        if (i1 <= i0)
        {
            return;
        }
        const dim3 g(1), b(SEQ_BLOCK);
        if (dim == 2 && ps == 0)
        {
            hipLaunchKernelGGL((k_sequence<2, 0>), g, b, 0, s, items, i0, i1, l_proj, l_pred, l_max, u, tag, int(has_bc), bc_value);
        }
        else if (dim == 2)
        {
            hipLaunchKernelGGL((k_sequence<2, 1>), g, b, 0, s, items, i0, i1, l_proj, l_pred, l_max, u, tag, int(has_bc), bc_value);
        }
        else if (ps == 0)
        {
            hipLaunchKernelGGL((k_sequence<3, 0>), g, b, 0, s, items, i0, i1, l_proj, l_pred, l_max, u, tag, int(has_bc), bc_value);
        }
        else
        {
            hipLaunchKernelGGL((k_sequence<3, 1>), g, b, 0, s, items, i0, i1, l_proj, l_pred, l_max, u, tag, int(has_bc), bc_value);
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
