// mgbf_apply_noalloc: mgbf_apply whose workspaces (the multigrid U and T,
// the analysis-grid intermediate of the transfers) live in
// a static device array instead of memory the context allocates, so that a
// call allocates nothing. No counterpart in the reference.
#include <hip/hip_runtime.h>

#include <mutex>
#include <string>

#include "mgbf.h"
#include "mgbf_launch.hpp"

// One copy per device that loads this code object, held until the process
// ends. Host code never takes this name as an address (on the host it is a
// stand-in): the device address comes from hipGetSymbolAddress, per device.
static __device__ __align__(256) float scratch[mgbf::kNoallocScratchFloats];

namespace {

constexpr int kMaxDevices = 64;
std::mutex g_mu;
float* g_scratch[kMaxDevices] = {};

}  // namespace

namespace mgbf {

int noalloc_setup(int device, std::string* err) {
    if (device < 0 || device >= kMaxDevices) {
        *err = "device index beyond the static scratch table";
        return MGBF_ERR_ARG;
    }
    std::lock_guard<std::mutex> lock(g_mu);
    if (g_scratch[device]) return MGBF_OK;
    int cur = -1;
    if (hipGetDevice(&cur) != hipSuccess || cur != device) {
        *err = "internal: the scratch is set up with its device current";
        return MGBF_ERR_HIP;
    }
    void* p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || p == nullptr) {
        *err = std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(e);
        return MGBF_ERR_HIP;
    }
    g_scratch[device] = static_cast<float*>(p);
    return MGBF_OK;
}

float* noalloc_scratch(int device) {
    if (device < 0 || device >= kMaxDevices) return nullptr;
    std::lock_guard<std::mutex> lock(g_mu);
    return g_scratch[device];
}

}  // namespace mgbf

extern "C" int mgbf_apply_noalloc(mgbf_ctx* ctx, int op, const float* in, float* out, void* stream) {
    return mgbf::apply_in_scratch(ctx, op, in, out, stream, &mgbf::noalloc_scratch,
                                  mgbf::kNoallocScratchFloats);
}
