// The harten tags of one iteration without any allocation at call time: the
// details (the only intermediate array of the device half of harten) live in
// a device array of the code object, `scratch`, instead of memory the library
// allocates.  For callers that cannot allocate (HIP graph capture, tight
// serving loops).  Every call on a GPU shares that GPU's array: two calls
// that overlap on one GPU race.
#include "../runtime.hpp"

#include <hip/hip_runtime.h>

#include <mutex>
#include <stdexcept>
#include <string>

// details of one harten iteration, one slot per reference storage slot
static __device__ __align__(256) double scratch[smr::kNoAllocSlots];

namespace smr
{
    namespace
    {
        constexpr int kMaxDevices = 64;
        std::mutex g_mu;
        // device address of `scratch` on each GPU (the array's name is a
        // host-side stand-in, not a device address)
        double* g_scratch[kMaxDevices] = {};
    } // namespace

    void noalloc_init(int device)
    {
        if (device < 0 || device >= kMaxDevices)
        {
            throw std::invalid_argument("noalloc_init: device ordinal out of range");
        }
        std::lock_guard<std::mutex> lk(g_mu);
        if (g_scratch[device])
        {
            return;
        }
        int current = -1;
        hip_check(hipGetDevice(&current), "hipGetDevice");
        if (current != device)
        {
            throw std::logic_error("noalloc_init: the device must be current");
        }
        void* p = nullptr;
        hip_check(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)), "hipGetSymbolAddress(scratch)");
        g_scratch[device] = static_cast<double*>(p);
    }

    void adapt_tags_noalloc(AdaptState& st, int ite, double eps, double regularity, uint8_t* tag)
    {
        Field& f = *st.field;
        f.check_current();
        Mesh& m         = *f.mesh;
        const int dev   = m.ctx->device;
        double* det     = nullptr;
        {
            std::lock_guard<std::mutex> lk(g_mu);
            det = (dev >= 0 && dev < kMaxDevices) ? g_scratch[dev] : nullptr;
        }
        if (!det)
        {
            throw std::runtime_error("adapt_tags_noalloc: no scratch address for this device (context not initialised)");
        }
        if (m.host.nref * st.ncomp() > kNoAllocSlots)
        {
            throw std::invalid_argument("adapt_tags_noalloc: " + std::to_string(m.host.nref) + " storage slots x " + std::to_string(st.ncomp())
                                        + " fields exceed the scratch of " + std::to_string(kNoAllocSlots));
        }
        if (st.relative_detail)
        {
            throw std::invalid_argument("adapt_tags_noalloc: the relative detail needs a device word per field; use smr_adapt_tags");
        }
        if (!m.dev || !m.dev->has_ghost || !m.dev->has_adapt)
        {
            throw std::invalid_argument("adapt_tags_noalloc: the mesh's device programs are not built; run the allocating "
                                        "entry point (smr_adapt_tags or smr_adapt_run) once on this mesh version first");
        }
        if (st.fields.empty())
        {
            Field* fs[1] = {&f};
            harten_tags(fs, 1, ite, eps, regularity, tag, det, nullptr, st.refine_boundary);
        }
        else
        {
            harten_tags(st.fields.data(), int(st.fields.size()), ite, eps, regularity, tag, det, nullptr, st.refine_boundary);
        }
    }
} // namespace smr
