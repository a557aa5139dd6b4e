// Per-cell bodies of the MR-step operators, shared by the standalone kernels
// (kernels.hip) and the fused stage-program kernel (program.hip), so both
// paths execute the same arithmetic in the same order.  They are host
// callable too: the single-core CPU baseline (oracle/cpu_step.cpp) runs the
// same bodies over host copies of the lists.
//
// A cell is addressed by its index c in a work list (List): its run is found
// between two chunk hints, and every storage position it touches is
// (resolved row base of a tap) + x, exactly the reference's
// memory_offset(...) + i arithmetic.  All floating-point expressions keep the
// reference's operation order (the code is compiled with -ffp-contract=off).
#pragma once

#include "device_api.hpp"

#include <hip/hip_runtime.h>

namespace smr::dev
{
    struct Cell
    {
        int64_t run;
        int32_t x, y, z;
        int level;
    };

    // the run holding cell c: a bisection between the runs of two hints
    __host__ __device__ __forceinline__ int64_t find_run(const List& l, int64_t c)
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
        return lo;
    }

    // the run map of the list (one entry per cell), written by the resolve
    __host__ __device__ __forceinline__ void op_cellrun(const List& l, int64_t c)
    {
        l.cellrun[c] = int32_t(find_run(l, c));
    }

    // cell c of the list: its run from the run map (one load) once the list
    // is resolved, else by bisection
    __host__ __device__ __forceinline__ Cell locate(const List& l, int64_t c)
    {
        const int64_t lo = l.cellrun ? int64_t(l.cellrun[c]) : find_run(l, c);
        Cell cell;
        cell.run   = lo;
        cell.x     = l.iv[5 * lo + 0] + int32_t(c - l.pref[lo]);
        cell.y     = l.iv[5 * lo + 2];
        cell.z     = l.iv[5 * lo + 3];
        cell.level = l.iv[5 * lo + 4];
        return cell;
    }

    __host__ __device__ __forceinline__ const int64_t* bases(const List& l, int64_t run)
    {
        return l.base + run * l.ntaps;
    }

    __host__ __device__ __forceinline__ int64_t row_key(int32_t y, int32_t z)
    {
        return (int64_t(z) << 32) + (int64_t(y) + 2147483648LL);
    }

    // LevelCellArray::get_interval/find (level_cell_array.hpp:669-743,
    // algorithm.hpp:391-467) as two binary searches: the row among the
    // level's rows, then the run holding x among the row's runs.
    __host__ __device__ __forceinline__ bool find_base(const Table& t, int level, int32_t x, int32_t y, int32_t z, int64_t& base, int32_t& rs, int32_t& re)
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
    __host__ __device__ __forceinline__ void op_resolve(const List& l, int64_t gid, const Table& t0, const Table& t1, int* err)
    {
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

    // flux of upwind_op (stencil_field.hpp:91-97),
    //   .5 * a * (ul + ur) + .5 * |a| * (ul - ur),
    // with the scalar factors (0.5*a, 0.5*|a|) formed once per launch
    __host__ __device__ __forceinline__ double flux_h(double ha, double habs, double ul, double ur)
    {
        return ha * (ul + ur) + habs * (ul - ur);
    }

    // flux of upwind_scalar_burgers_op (stencil_field.hpp:193-216)
    __host__ __device__ __forceinline__ double burgers_upwind_flux(double a, double ul, double ur)
    {
        const bool mask1 = (a * ul) < (a * ur);
        const bool mask2 = (ul * ur) > 0.0;
        const double mn  = fmin(fabs(ul), fabs(ur));
        const double mx  = fmax(fabs(ul), fabs(ur));
        double out       = 0.0;
        if (mask1 && mask2)
        {
            out = .5 * mn * mn;
        }
        if (!mask1)
        {
            out = .5 * mx * mx;
        }
        return out;
    }

    // unp1 = u - dt * upwind_scalar_burgers(k, u) (2D; same taps and order as upwind_value<2>)
    __host__ __device__ __forceinline__ double burgers_value(const int64_t* b, int32_t x, int level, const double* __restrict__ u, double k0,
                                                    double k1, double dt)
    {
        const double* uc_p = u + b[0] + x;
        const double uc    = uc_p[0];
        const double FL    = burgers_upwind_flux(k0, uc_p[-1], uc);
        const double FR    = burgers_upwind_flux(k0, uc, uc_p[1]);
        const double FD    = burgers_upwind_flux(k1, u[b[1] + x], uc);
        const double FU    = burgers_upwind_flux(k1, uc, u[b[2] + x]);
        const double S     = -FL + FR + -FD + FU;
        return uc - dt * (S * double(1 << level));
    }

    // interp_coeffs<3>(+1) = {1/8, 1, -1/8}, interp_coeffs<3>(-1) = {-1/8, 1, 1/8}
    // (numeric/prediction.hpp:31-35), indexed by parity (0 even, 1 odd)
    __host__ __device__ __forceinline__ double pcoef(int parity, int k)
    {
        return k == 1 ? 1.0 : ((parity == 0) == (k == 0) ? 1. / 8. : -1. / 8.);
    }

    // unp1 = u - dt * upwind(a, u) (stencil_field.hpp:37-53).  Taps: the cell's
    // row (covering x-1..x+1), then the rows y-1, y+1 (and z-1, z+1).  S / dx with
    // dx = 2^-level is computed as S * 2^level: both are the exact value
    // S*2^level, so the result is bit-identical.
    template <int DIM>
    __host__ __device__ __forceinline__ double upwind_value(const int64_t* b, int32_t x, int level, const double* __restrict__ u, double ha0,
                                                   double hb0, double ha1, double hb1, double ha2, double hb2, double dt)
    {
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
            const double FF = flux_h(ha2, hb2, u[b[3] + x], uc);
            const double FB = flux_h(ha2, hb2, uc, u[b[4] + x]);
            S               = -FL + FR + -FD + FU + -FF + FB;
        }
        return uc - dt * (S * double(1 << level));
    }

    // storage bases of the rows of column walk w: the row below its first run,
    // the rows of its n runs, the row above its last run
    __host__ __device__ __forceinline__ void op_walk_rows(const List& l, int64_t w)
    {
        const int32_t* wk = l.walks + 4 * w;
        int32_t r         = wk[0];
        const int32_t n   = wk[2];
        int64_t* rows     = l.walk_rows + wk[3];
        const int32_t* iv = l.iv + 5 * int64_t(r);
        rows[l.walk_stride - 2] = int64_t((uint64_t(uint32_t(wk[1])) << 32) | uint32_t(n));
        rows[l.walk_stride - 1] = int64_t((uint64_t(uint32_t(iv[1])) << 32) | uint32_t(iv[4]));
        rows[0]           = l.base[int64_t(r) * l.ntaps + 1];
        for (int32_t k = 0; k < n; ++k)
        {
            rows[k + 1] = l.base[int64_t(r) * l.ntaps];
            if (k + 1 == n)
            {
                rows[n + 1] = l.base[int64_t(r) * l.ntaps + 2];
            }
            else
            {
                r = l.next[r];
            }
        }
    }

    // upwind_value<2> on values already loaded: left, right, down, up neighbours
    // of uc (same arithmetic, same order)
    __host__ __device__ __forceinline__ double upwind_vals_2d(double uc, double ul, double ur, double ud, double uu, int level, double ha0,
                                                     double hb0, double ha1, double hb1, double dt)
    {
        const double FL = flux_h(ha0, hb0, ul, uc);
        const double FR = flux_h(ha0, hb0, uc, ur);
        const double FD = flux_h(ha1, hb1, ud, uc);
        const double FU = flux_h(ha1, hb1, uc, uu);
        const double S  = -FL + FR + -FD + FU;
        return uc - dt * (S * double(1 << level));
    }

    // projection_op_ (numeric/projection.hpp:15-65): taps dst, then the 2^(dim-1) fine rows
    template <int DIM>
    __host__ __device__ __forceinline__ void op_project_at(const List& l, int64_t c, const Cell& cell, double* dst, const double* src)
    {
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

    template <int DIM>
    __host__ __device__ __forceinline__ void op_project(const List& l, int64_t c, double* dst, const double* src)
    {
        op_project_at<DIM>(l, c, locate(l, c), dst, src);
    }

    // prediction_op <s,false> (numeric/prediction.hpp:259-357): taps dst, then
    // the coarse rows (z outer, y inner) starting at x = (x0>>1) - s
    template <int DIM, int S>
    __host__ __device__ __forceinline__ void op_predict_fine_at(const List& l, int64_t c, const Cell& cell, double* u)
    {
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

    template <int DIM, int S>
    __host__ __device__ __forceinline__ void op_predict_fine(const List& l, int64_t c, double* u)
    {
        op_predict_fine_at<DIM, S>(l, c, locate(l, c), u);
    }

    // compute_detail_op (mr/operators.hpp:139-540): taps = the 2^(dim-1) fine
    // rows, then the coarse stencil rows (z outer, y inner)
    template <int DIM, int S>
    __host__ __device__ __forceinline__ void op_detail_at(const List& l, int64_t c, const Cell& cell, const double* u, double* det)
    {
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

    template <int DIM, int S>
    __host__ __device__ __forceinline__ void op_detail(const List& l, int64_t c, const double* u, double* det)
    {
        op_detail_at<DIM, S>(l, c, locate(l, c), u, det);
    }

    // mr_criteria_op (mr/criteria.hpp:13-122): taps coarse, then fine rows;
    // thr holds eps, fine and coarse thresholds per fine level (3 x 21).  The
    // details of ncomp components (one per adapted field) are stored one array
    // after the other, cstride apart: a coarse detail of any component above
    // coarse_eps blocks the coarsening, which needs every component of every
    // child at most eps; a child is refined when any of its components is
    // above fine_eps (criteria.hpp:47-56, 63-72, 96-105)
    template <int DIM>
    __host__ __device__ __forceinline__ void op_criteria_at(const List& l, int64_t c, const Cell& cell, const double* det, uint8_t* tag, const double* thr, int min_level,
                                                int max_level, int ncomp = 1, int64_t cstride = 0)
    {
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
        const double eps = thr[fl], fine_eps = thr[21 + fl], coarse_eps = thr[42 + fl];
        if (fl > min_level)
        {
            bool cond = false;
            for (int n = 0; n < ncomp && !cond; ++n)
            {
                cond = fabs(det[n * cstride + b[0] + x]) > coarse_eps;
            }
            if (!cond)
            {
                bool ok = true;
#pragma unroll
                for (int k = 0; k < NC; ++k)
                {
                    for (int n = 0; n < ncomp; ++n)
                    {
                        ok = ok && !(fabs(det[n * cstride + fi[k]]) > eps);
                    }
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
                bool refine = false;
                for (int n = 0; n < ncomp; ++n)
                {
                    refine = refine || fabs(det[n * cstride + fi[k]]) > fine_eps;
                }
                tag[fi[k]] |= uint8_t(refine) * uint8_t(4);
            }
        }
    }

    template <int DIM>
    __host__ __device__ __forceinline__ void op_criteria(const List& l, int64_t c, const double* det, uint8_t* tag, const double* thr, int min_level,
                                                int max_level, int ncomp = 1, int64_t cstride = 0)
    {
        op_criteria_at<DIM>(l, c, locate(l, c), det, tag, thr, min_level, max_level, ncomp, cstride);
    }

    // maximum_op (mr/operators.hpp:22-96): taps coarse, then fine rows
    template <int DIM>
    __host__ __device__ __forceinline__ void op_maximum_at(const List& l, int64_t c, const Cell& cell, uint8_t* tag)
    {
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

    template <int DIM>
    __host__ __device__ __forceinline__ void op_maximum(const List& l, int64_t c, uint8_t* tag)
    {
        op_maximum_at<DIM>(l, c, locate(l, c), tag);
    }

    __host__ __device__ __forceinline__ void op_tag_init_at(const List& l, int64_t c, const Cell& cell, uint8_t* tag)
    {
        tag[bases(l, cell.run)[0] + cell.x] = 1;
    }

    __host__ __device__ __forceinline__ void op_tag_init(const List& l, int64_t c, uint8_t* tag)
    {
        op_tag_init_at(l, c, locate(l, c), tag);
    }

    // README.md:135-154 hand loop over for_each_interval (2D, a = (1, 1)):
    //   unp1 = u - dt / dx * (u - u(i-1) + u - u(j-1)),
    // with dt / dx = dt * 2^level exactly (dx = 2^-level)
    __host__ __device__ __forceinline__ double hand_loop_value(const int64_t* b, int32_t x, int level, const double* __restrict__ u, double dt)
    {
        const double uc = u[b[0] + x];
        const double k  = dt / ldexp(1.0, -level);
        return uc - k * (uc - u[b[0] + x - 1] + uc - u[b[1] + x]);
    }

    // true when the tag of a cell changes the cell array: a refine below
    // max_level or a coarsen without keep above min_level
    // (update_cell_array_from_tag, graduation.hpp:743-843)
    __host__ __device__ __forceinline__ bool tag_changes_mesh_at(const List& l, int64_t c, const Cell& cell, const uint8_t* tag, int min_level, int max_level)
    {
        const uint8_t t = tag[bases(l, cell.run)[0] + cell.x];
        const bool ref  = (t & 4) && cell.level < max_level;
        const bool crs  = (t & 2) && !(t & 1) && cell.level > min_level;
        return ref || crs;
    }

    __host__ __device__ __forceinline__ bool tag_changes_mesh(const List& l, int64_t c, const uint8_t* tag, int min_level, int max_level)
    {
        return tag_changes_mesh_at(l, c, locate(l, c), tag, min_level, max_level);
    }

    // flag[0] = 1 when the cell's tag changes the cell array.  Every lane that
    // sees a change stores: in a stage program one wave can hold cells of
    // several segments (instances), each with its own flag.
    __host__ __device__ __forceinline__ void op_tag_change_at(const List& l, int64_t c, const Cell& cell, const uint8_t* tag, int min_level,
                                                     int max_level, int* flag)
    {
        if (tag_changes_mesh_at(l, c, cell, tag, min_level, max_level))
        {
            flag[0] = 1;
        }
    }

    __host__ __device__ __forceinline__ void op_tag_change(const List& l, int64_t c, const uint8_t* tag, int min_level, int max_level, int* flag)
    {
        if (tag_changes_mesh(l, c, tag, min_level, max_level))
        {
            flag[0] = 1;
        }
    }

    __host__ __device__ __forceinline__ void op_copy_at(const List& l, int64_t c, const Cell& cell, double* dst, const double* src)
    {
        const int64_t* b = bases(l, cell.run);
        dst[b[0] + cell.x] = src[b[1] + cell.x];
    }

    __host__ __device__ __forceinline__ void op_copy(const List& l, int64_t c, double* dst, const double* src)
    {
        op_copy_at(l, c, locate(l, c), dst, src);
    }

    // prediction_op <s,true> (numeric/prediction.hpp:107-257): taps = the
    // 2^(dim-1) destination rows (children), then the source stencil rows
    template <int DIM, int S>
    __host__ __device__ __forceinline__ void op_predict_children_at(const List& l, int64_t c, const Cell& cell, double* dst, const double* src)
    {
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

    template <int DIM, int S>
    __host__ __device__ __forceinline__ void op_predict_children(const List& l, int64_t c, double* dst, const double* src)
    {
        op_predict_children_at<DIM, S>(l, c, locate(l, c), dst, src);
    }

    // one outer-ghost recipe: copy, Dirichlet<1> (bc/dirichlet.hpp:29), the BC
    // projection average (algorithm/update_outer_ghost.hpp:78-131), the
    // polynomial extrapolation of stencil 4 or 6 from 3 or 5 sources
    // (bc/polynomial_extrapolation.hpp:63-72, same operation order), or the
    // layer-1 copy of a field without a B.C.
    __host__ __device__ __forceinline__ void op_recipe(const int64_t* __restrict__ dst, const int32_t* __restrict__ op, const int64_t* __restrict__ sp,
                                              const int64_t* __restrict__ src, int64_t r, double* u, int has_bc, double bc_value)
    {
        const int o = op[r];
        if ((o == 1 && !has_bc) || (o == 4 && has_bc))
        {
            return;
        }
        double v;
        if (o == 0 || o == 4)
        {
            v = u[src[sp[r]]];
        }
        else if (o == 1)
        {
            v = 2 * bc_value - u[src[sp[r]]];
        }
        else if (o == 3)
        {
            const int64_t* s = src + sp[r];
            if (sp[r + 1] - sp[r] == 3)
            {
                v = u[s[0]] - u[s[1]] * 3.0 + u[s[2]] * 3.0;
            }
            else
            {
                v = u[s[0]] - u[s[1]] * 5.0 + u[s[2]] * 10.0 - u[s[3]] * 10.0 + u[s[4]] * 5.0;
            }
        }
        else
        {
            v               = 0;
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
    // ------------------------------------------------------------------
    // Adjoint of the upwind step on a frozen mesh.  With
    //   F(a, ul, ur) = a+ ul + a- ur   (a+ = max(a, 0), a- = min(a, 0)),
    // unp1[c] = u[c] - k S, k = dt 2^level, is linear in u:
    //   d unp1[c] / d u[c]      = 1 - k sum_d |a_d|
    //   d unp1[c] / d u[c - e_d] =  k a_d+,   d unp1[c] / d u[c + e_d] = -k a_d-
    // so the transpose at a cell s gathers lam from its face neighbours
    // (lam is zero outside the cells).  The velocity derivative uses
    //   dF/da = .5 (ul + ur) + .5 sign(a) (ul - ur).
    // ------------------------------------------------------------------
    __host__ __device__ __forceinline__ double dflux_da(double a, double ul, double ur)
    {
        const double sg = a > 0 ? 1.0 : (a < 0 ? -1.0 : 0.0);
        return .5 * (ul + ur) + .5 * sg * (ul - ur);
    }

    template <int DIM>
    __host__ __device__ __forceinline__ void op_adj_upwind_at(const List& l, int64_t c, const Cell& cell, const double* lam, double* mu, const double* un, double* ga,
                                                  double a0, double a1, double a2, double dt)
    {
        const int64_t* b = bases(l, cell.run);
        const int32_t x  = cell.x;
        const double k   = dt * ldexp(1.0, cell.level);
        const int64_t i  = b[0] + x;
        double self      = fabs(a0) + fabs(a1);
        if constexpr (DIM == 3)
        {
            self += fabs(a2);
        }
        double m = (1.0 - k * self) * lam[i];
        m += k * fmax(a0, 0.0) * lam[i + 1] - k * fmin(a0, 0.0) * lam[i - 1];
        m += k * fmax(a1, 0.0) * lam[b[2] + x] - k * fmin(a1, 0.0) * lam[b[1] + x];
        if constexpr (DIM == 3)
        {
            m += k * fmax(a2, 0.0) * lam[b[4] + x] - k * fmin(a2, 0.0) * lam[b[3] + x];
        }
        mu[i] = m;
        const double li = lam[i];
        const double uc = un[i];
        ga[DIM * c + 0] += -k * li * (dflux_da(a0, uc, un[i + 1]) - dflux_da(a0, un[i - 1], uc));
        ga[DIM * c + 1] += -k * li * (dflux_da(a1, uc, un[b[2] + x]) - dflux_da(a1, un[b[1] + x], uc));
        if constexpr (DIM == 3)
        {
            ga[DIM * c + 2] += -k * li * (dflux_da(a2, uc, un[b[4] + x]) - dflux_da(a2, un[b[3] + x], uc));
        }
    }

    template <int DIM>
    __host__ __device__ __forceinline__ void op_adj_upwind(const List& l, int64_t c, const double* lam, double* mu, const double* un, double* ga,
                                                  double a0, double a1, double a2, double dt)
    {
        op_adj_upwind_at<DIM>(l, c, locate(l, c), lam, mu, un, ga, a0, a1, a2, dt);
    }

    // dir: 0 x-1, 1 x+1, 2 y-1, 3 y+1, 4 z-1, 5 z+1 (the neighbour is a ghost)
    template <int DIM>
    __host__ __device__ __forceinline__ void op_adj_scatter_at(const List& l, int64_t c, const Cell& cell, int dir, const double* lam, double* mu, double a0, double a1,
                                                   double a2, double dt)
    {
        const int64_t* b = bases(l, cell.run);
        const int32_t x  = cell.x;
        const double k   = dt * ldexp(1.0, cell.level);
        const double li  = lam[b[0] + x];
        switch (dir)
        {
            case 0: mu[b[0] + x - 1] += k * fmax(a0, 0.0) * li; break;
            case 1: mu[b[0] + x + 1] += -k * fmin(a0, 0.0) * li; break;
            case 2: mu[b[1] + x] += k * fmax(a1, 0.0) * li; break;
            case 3: mu[b[2] + x] += -k * fmin(a1, 0.0) * li; break;
            case 4: mu[b[3] + x] += k * fmax(a2, 0.0) * li; break;
            case 5: mu[b[4] + x] += -k * fmin(a2, 0.0) * li; break;
            default: break;
        }
    }

    template <int DIM>
    __host__ __device__ __forceinline__ void op_adj_scatter(const List& l, int64_t c, int dir, const double* lam, double* mu, double a0, double a1,
                                                   double a2, double dt)
    {
        op_adj_scatter_at<DIM>(l, c, locate(l, c), dir, lam, mu, a0, a1, a2, dt);
    }

    __host__ __device__ __forceinline__ void op_adj_gather(const int64_t* __restrict__ tgt, const int64_t* __restrict__ ptr,
                                                  const int64_t* __restrict__ src, const double* __restrict__ w, int64_t r, double* mu)
    {
        double acc = 0;
        for (int64_t k = ptr[r]; k < ptr[r + 1]; ++k)
        {
            acc += w[k] * mu[src[k]];
        }
        mu[tgt[r]] += acc;
    }

    __host__ __device__ __forceinline__ void op_cells_to_compact_at(const List& l, int64_t c, const Cell& cell, double* out, const double* u)
    {
        out[c]          = u[bases(l, cell.run)[0] + cell.x];
    }

    __host__ __device__ __forceinline__ void op_cells_to_compact(const List& l, int64_t c, double* out, const double* u)
    {
        op_cells_to_compact_at(l, c, locate(l, c), out, u);
    }

    __host__ __device__ __forceinline__ void op_cells_from_compact_at(const List& l, int64_t c, const Cell& cell, double* u, const double* in)
    {
        u[bases(l, cell.run)[0] + cell.x] = in[c];
    }

    __host__ __device__ __forceinline__ void op_cells_from_compact(const List& l, int64_t c, double* u, const double* in)
    {
        op_cells_from_compact_at(l, c, locate(l, c), u, in);
    }
} // namespace smr::dev
