// noalloc.hip -- the device scratch of tsu_align_batched_noalloc.
//
// One file-scope array per GPU, held for the life of the process: the work
// counters and the queue-order array of the persistent launches, which
// tsu_align_batched allocates per call.  Host code never passes the array's
// name to a kernel or a copy (on the host it is a stand-in, not a device
// address): tsu_create resolves the device address once per device, with
// that device current, and every use goes through that address.
#include <hip/hip_runtime.h>

#include <mutex>

#include "tsu_internal.h"

static __device__ __align__(256) uint8_t scratch[tsu::kNoallocScratchBytes];

namespace {

constexpr int kMaxDevices = 64;
std::mutex g_mu;
void* g_addr[kMaxDevices] = {};

}  // namespace

// Device address of `scratch` on `device`, which must be the current device;
// resolved on the first call for that device and kept.  Returns a
// hipError_t value (0 on success).
int tsu::scratch_address(int device, void** out) {
    if (device < 0 || device >= kMaxDevices || !out) return (int)hipErrorInvalidValue;
    std::lock_guard<std::mutex> lock(g_mu);
    if (!g_addr[device]) {
        void* p = nullptr;
        const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess) return (int)e;
        if (!p) return (int)hipErrorInvalidSymbol;
        g_addr[device] = p;
    }
    *out = g_addr[device];
    return 0;
}
