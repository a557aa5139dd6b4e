// Device-side operators over the leaves of a mesh, with samurai's call shape:
//
//   samurai::for_each_interval(mesh, f)            algorithm.hpp:75-132
//
// where f is a __device__ callable (a HIP device lambda or functor) invoked
// once per leaf as f(level, i, index): i is the one-cell x interval
// [x, x + 1) of the leaf (its `index` member is the storage offset of its
// row, value = data[i.index + x], as for samurai's intervals) and index holds
// (y) in 2D, (y, z) in 3D.  Fields are read and written inside f through
// views: u(level, i, j) in 2D, u(level, i, j, k) in 3D, looked up in the
// mesh's reference interval table (two binary searches, as
// LevelCellArray::find, level_cell_array.hpp:669-743), which yields a
// double&.  A lookup of a cell that is not in the reference mesh raises
// std::out_of_range from for_each_interval after the launch, as the
// reference throws from find (level_cell_array.hpp:674-680).
//
// The README's hand-written upwind loop (README.md:135-154) becomes
//
//   auto uv = samurai::device::view(u), vv = samurai::device::view(unp1);
//   auto mv = samurai::device::view(mesh);
//   samurai::for_each_interval(mesh, [=] __device__ (int level, const auto& i, const auto& index) {
//       auto j    = index[0];
//       double dx = mv.cell_length(level);
//       vv(level, i, j) = uv(level, i, j) - dt / dx * (uv(level, i, j) - uv(level, i - 1, j)
//                                                       + uv(level, i, j) - uv(level, i, j - 1));
//   });
//
// Compile the calling code with hipcc (gfx950).  Views are taken after the
// field's host-side writes and before the loop; they stay valid while the
// mesh does not change.  Calls are ordered on the library's stream and the
// loop returns when the kernel has run.
#pragma once

#include "samurai.hpp"

#include <hip/hip_runtime.h>

#include <cstdint>
#include <stdexcept>

namespace samurai_amd
{
    namespace device
    {
        inline void hip_ok(hipError_t e, const char* what)
        {
            if (e != hipSuccess)
            {
                throw std::runtime_error(std::string(what) + ": " + hipGetErrorString(e));
            }
        }

        // one-cell interval [start, end) of a leaf, shifted by i +- n
        struct interval_t
        {
            int start, end, step;
            int64_t index;
            __host__ __device__ interval_t operator+(int n) const
            {
                return {start + n, end + n, step, index};
            }
            __host__ __device__ interval_t operator-(int n) const
            {
                return {start - n, end - n, step, index};
            }
            __host__ __device__ int size() const
            {
                return end - start;
            }
        };

        template <std::size_t dim>
        struct index_t
        {
            int v[dim - 1];
            __host__ __device__ int operator[](std::size_t k) const
            {
                return v[k];
            }
        };

        // where a failed lookup leaves its mark (and a slot a failed lookup
        // can return a reference to)
        struct ErrorSlot
        {
            int missing;
            double sink;
        };

        inline ErrorSlot* error_slot()
        {
            // one per process, released with the process
            static ErrorSlot* slot = []
            {
                ErrorSlot* p = nullptr;
                hip_ok(hipMalloc(reinterpret_cast<void**>(&p), sizeof(ErrorSlot)), "hipMalloc(error slot)");
                hip_ok(hipMemset(p, 0, sizeof(ErrorSlot)), "hipMemset(error slot)");
                return p;
            }();
            return slot;
        }

        struct mesh_view
        {
            smr_mesh_view v;
            ErrorSlot* err;

            __device__ double cell_length(int level) const
            {
                return ldexp(1.0, -level); // cell.hpp:17-20 on the unit box
            }

            // storage index of cell (level; x, y, z), or -1
            __device__ int64_t find(int level, int x, int y, int z) const
            {
                if (level < 0 || level > 20)
                {
                    return -1;
                }
                const int64_t key = (int64_t(z) << 32) + (int64_t(y) + 2147483648LL);
                int64_t lo = v.row_off[level], hi = v.row_off[level + 1];
                const int64_t end = hi;
                while (lo < hi)
                {
                    const int64_t mid = (lo + hi) >> 1;
                    if (v.row_key[mid] < key)
                    {
                        lo = mid + 1;
                    }
                    else
                    {
                        hi = mid;
                    }
                }
                if (lo >= end || v.row_key[lo] != key)
                {
                    return -1;
                }
                const int64_t first = v.row_run[lo];
                int64_t a = first, b = v.row_run[lo + 1];
                while (a < b)
                {
                    const int64_t mid = (a + b) >> 1;
                    if (v.run_start[mid] <= x)
                    {
                        a = mid + 1;
                    }
                    else
                    {
                        b = mid;
                    }
                }
                const int64_t k = a - 1;
                if (k < first || x >= v.run_end[k])
                {
                    return -1;
                }
                return v.run_index[k] + x;
            }
        };

        template <std::size_t dim>
        struct field_view
        {
            double* data;
            mesh_view m;

            __device__ double& at(int level, int x, int y, int z) const
            {
                const int64_t k = m.find(level, x, y, z);
                if (k < 0)
                {
                    m.err->missing = 1;
                    return m.err->sink;
                }
                return data[k];
            }
            __device__ double& operator()(int level, const interval_t& i, int j) const
            {
                static_assert(dim == 2, "u(level, i, j) is the 2D access");
                return at(level, i.start, j, 0);
            }
            __device__ double& operator()(int level, const interval_t& i, int j, int k) const
            {
                static_assert(dim == 3, "u(level, i, j, k) is the 3D access");
                return at(level, i.start, j, k);
            }
        };

        template <std::size_t dim>
        inline mesh_view view(MRMesh<dim>& mesh)
        {
            mesh_view mv;
            check(smr_mesh_device_view(mesh.handle(), &mv.v));
            mv.err = error_slot();
            return mv;
        }

        // the field's device storage; its host mirror is dropped, so host
        // reads after a loop that writes the field see the device values
        template <std::size_t dim>
        inline field_view<dim> view(ScalarField<dim>& u)
        {
            field_view<dim> fv;
            fv.data = smr_field_device_data(u.handle());
            if (!fv.data)
            {
                throw std::runtime_error("field has no device storage");
            }
            fv.m = view(u.mesh());
            u.touched();
            return fv;
        }

        template <std::size_t dim, class F>
        __global__ __launch_bounds__(256) void k_for_each_leaf(smr_mesh_view v, F f)
        {
            const int64_t c = int64_t(blockIdx.x) * 256 + threadIdx.x;
            if (c >= v.nb_cells)
            {
                return;
            }
            // the run of leaf c: binary search between two chunk hints
            const int64_t j = c >> 8;
            int64_t lo = v.cell_hint[j], hi = v.cell_hint[j + 1];
            while (lo < hi)
            {
                const int64_t mid = (lo + hi + 1) >> 1;
                if (v.cell_prefix[mid] <= c)
                {
                    lo = mid;
                }
                else
                {
                    hi = mid - 1;
                }
            }
            const int32_t* r = v.cell_runs + 5 * lo;
            const int x      = r[0] + int(c - v.cell_prefix[lo]);
            interval_t i{x, x + 1, 1, v.cell_base[lo * v.cell_base_stride]};
            index_t<dim> idx;
            idx.v[0] = r[2];
            if constexpr (dim == 3)
            {
                idx.v[1] = r[3];
            }
            f(r[4], i, idx);
        }
    } // namespace device

    // for_each_interval(mesh[cells], f) (algorithm.hpp:75-132) on the device
    template <std::size_t dim, class F>
    void for_each_interval(MRMesh<dim>& mesh, F f)
    {
        smr_mesh_view v;
        check(smr_mesh_device_view(mesh.handle(), &v));
        hipStream_t s           = static_cast<hipStream_t>(smr_ctx_stream(context()));
        device::ErrorSlot* err  = device::error_slot();
        device::hip_ok(hipMemsetAsync(err, 0, sizeof(int), s), "hipMemsetAsync(error slot)");
        if (v.nb_cells > 0)
        {
            const unsigned blocks = unsigned((v.nb_cells + 255) / 256);
            auto kernel           = device::k_for_each_leaf<dim, F>;
            hipLaunchKernelGGL(kernel, dim3(blocks), dim3(256), 0, s, v, f);
            device::hip_ok(hipGetLastError(), "for_each_interval launch");
        }
        int missing = 0;
        device::hip_ok(hipMemcpyAsync(&missing, err, sizeof(int), hipMemcpyDeviceToHost, s), "hipMemcpyAsync(error slot)");
        device::hip_ok(hipStreamSynchronize(s), "hipStreamSynchronize");
        if (missing)
        {
            throw std::out_of_range("for_each_interval: a cell read or written by the operator is not in the mesh");
        }
    }
} // namespace samurai_amd
