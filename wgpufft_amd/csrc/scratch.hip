// The library's static device scratch: the workspace of
// mifft_execute_static_scratch, for callers that cannot allocate device
// memory.  Each GPU that loads the library holds one copy of the array until
// the process ends; host code never uses the array's name as an address (in
// host code it is only a stand-in for the symbol): the device address comes
// from hipGetSymbolAddress, once per device, with that device current.
#include <mutex>

#include "plan.hpp"

namespace mifft {

static __device__ __align__(256) unsigned char scratch[kStaticScratchBytes];

namespace {
constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
void* g_scratch_ptr[kMaxDevices] = {};
}  // namespace

void* static_scratch(int device) {
    if (device < 0 || device >= kMaxDevices) fail(MIFFT_ERR_INVALID_ENUM, "device index out of range");
    std::lock_guard<std::mutex> g(g_scratch_mu);
    if (!g_scratch_ptr[device]) {
        int cur = -1;
        hip_check(hipGetDevice(&cur), "hipGetDevice");
        if (cur != device) fail(MIFFT_ERR_INVALID_ENUM, "the static scratch is resolved on the current device only");
        void* ptr = nullptr;
        hip_check(hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch)), "hipGetSymbolAddress(scratch)");
        if (!ptr) fail(MIFFT_ERR_HIP, "hipGetSymbolAddress(scratch) returned no address");
        g_scratch_ptr[device] = ptr;
    }
    return g_scratch_ptr[device];
}

}  // namespace mifft
