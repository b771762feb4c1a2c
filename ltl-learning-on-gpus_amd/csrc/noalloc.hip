// noalloc.hip -- the static device scratch region of ltl_learn_noalloc.
//
// Callers that cannot allocate device memory (code under HIP graph capture,
// tight serving loops, harnesses that forbid hipMalloc) learn within this
// region: one array of LTL_NOALLOC_SCRATCH_BYTES per GPU, held from the
// library's load until the process ends.  Host code never passes the array's
// name to a kernel or a copy (in host code it is a stand-in, not a device
// address): noalloc_scratch() looks up its device address once per device
// with that device current, and the library carves the call's buffers from
// that address.
#include <hip/hip_runtime.h>

#include <mutex>

#include "../../include/ltl_amd.h"
#include "kernels.hpp"

static __device__ __align__(256) unsigned char scratch[LTL_NOALLOC_SCRATCH_BYTES];

namespace ltl {

namespace {
constexpr int kMaxDevices = 64;
std::mutex g_mu;
char *g_ptr[kMaxDevices] = {};
}  // namespace

hipError_t noalloc_scratch(int device, char **ptr) {
    *ptr = nullptr;
    if (device < 0 || device >= kMaxDevices) return hipErrorInvalidDevice;
    std::lock_guard<std::mutex> lock(g_mu);
    if (!g_ptr[device]) {
        void *p = nullptr;
        const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess) return e;
        g_ptr[device] = static_cast<char *>(p);
    }
    *ptr = g_ptr[device];
    return hipSuccess;
}

}  // namespace ltl
