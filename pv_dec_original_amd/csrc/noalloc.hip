// noalloc.hip — the device scratch array of the entry points that allocate nothing.
//
// pvlbm_sample_kinds_noalloc stages its sample indices and its output in `scratch`, a file-scope
// device array of PVLBM_NOALLOC_SCRATCH_BYTES that every GPU the library runs on holds until the
// process ends.  Host code never passes the array's name to a kernel or a copy (on the host that
// name is only a stand-in for the symbol): each device's address of it is looked up once, with
// that device current, by hipGetSymbolAddress, and only that address is used.

#include <hip/hip_runtime.h>

#include <mutex>

#include "../../include/pvlbm.h"
#include "kernels.hpp"

namespace pvlbm {

namespace {

static __device__ __align__(256) unsigned char scratch[PVLBM_NOALLOC_SCRATCH_BYTES];

constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
void* g_scratch_ptr[kMaxDevices] = {};

}  // namespace

hipError_t noalloc_scratch(int device, void** ptr) {
    if (!ptr || device < 0 || device >= kMaxDevices) return hipErrorInvalidValue;
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (!g_scratch_ptr[device]) {
        int prev = 0;
        hipError_t e = hipGetDevice(&prev);
        if (e != hipSuccess) return e;
        if (prev != device && (e = hipSetDevice(device)) != hipSuccess) return e;
        void* p = nullptr;
        e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (prev != device) {
            const hipError_t e2 = hipSetDevice(prev);
            if (e == hipSuccess) e = e2;
        }
        if (e != hipSuccess) return e;
        if (!p) return hipErrorInvalidSymbol;
        g_scratch_ptr[device] = p;
    }
    *ptr = g_scratch_ptr[device];
    return hipSuccess;
}

}  // namespace pvlbm
