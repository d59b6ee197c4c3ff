// Fixed device scratch of the entry points that allocate nothing
// (cf_advance_to_noalloc): one file-scope device array of SCRATCH_DOUBLES
// doubles.  Every GPU the library runs on holds its own copy, at its own
// address, from module load until the process ends.  Host code never passes
// the array's name to a kernel or a copy (on the host it is only a stand-in
// for the symbol): it uses the device address that scratch_address looks up
// once per device, with that device current.
#include <hip/hip_runtime.h>

#include <mutex>

#include "cf_kernels.h"

namespace cf {

static __device__ __align__(256) double scratch[SCRATCH_DOUBLES];

namespace {
constexpr int MAX_DEVICES = 64;
std::mutex g_scratch_mu;
double* g_scratch_addr[MAX_DEVICES] = {};
}  // namespace

hipError_t scratch_address(int device, double** out) {
    if (device < 0 || device >= MAX_DEVICES) return hipErrorInvalidDevice;
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (!g_scratch_addr[device]) {
        int prev = -1;
        hipError_t e = hipGetDevice(&prev);
        if (e != hipSuccess) return e;
        if (prev != device && (e = hipSetDevice(device)) != hipSuccess) return e;
        void* p = nullptr;
        e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (prev != device) (void)hipSetDevice(prev);
        if (e != hipSuccess) return e;
        g_scratch_addr[device] = static_cast<double*>(p);
    }
    *out = g_scratch_addr[device];
    return hipSuccess;
}

}  // namespace cf
