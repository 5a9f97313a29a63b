// wfa_launch.hip -- host-side launch plumbing of the WFA kernels and the
// streaming copy kernel used to calibrate HBM byte counters.
#include <hip/hip_runtime.h>
#include <stdint.h>

#include <map>
#include <mutex>

#include "wfa_kernel.h"

namespace tsu {

KernelFn pick_nw1(const Shape& sh);
KernelFn pick_nw2(const Shape& sh);
KernelFn pick_nw4(const Shape& sh);
KernelFn pick_nw8(const Shape& sh);
KernelFn pick_nw16(const Shape& sh);
KernelFn pick_reg(const Shape& sh);
KernelFn pick_regmw(const Shape& sh);
int regmw_exchange_words(int nw);
KernelFn pick_regml(const Shape& sh);
int regml_exchange_words(int nw);
KernelFn pick_2p(const Shape& sh);
int64_t lds_bytes_2p(const Shape& sh, const KArgs& a);

// Plain streaming copy (HBM byte-counter calibration).
__global__ __launch_bounds__(256) void copy_kernel(const uint4* __restrict__ src, uint4* __restrict__ dst, int64_t n) {
    for (int64_t i = blockIdx.x * 256ll + threadIdx.x; i < n; i += (int64_t)gridDim.x * 256) dst[i] = src[i];
}

// ---------------------------------------------------------------------------
// host-side launch plumbing
// ---------------------------------------------------------------------------
int32_t ring_stride_entries(const Shape& sh, int32_t W) {
    return sh.off16 ? ring_stride<int16_t>(W) : ring_stride<int32_t>(W);
}

int64_t ring_bytes(const Shape& sh, int32_t W, int32_t rm, int32_t ri) {
    // rm M slots, ri I slots, ri D slots and one permanent NULL slot
    return (int64_t)(rm + 2 * ri + 1) * ring_stride_entries(sh, W) * (sh.off16 ? 2 : 4);
}

int64_t lds_bytes(const Shape& sh, const KArgs& a) {
    if (sh.two_piece) return lds_bytes_2p(sh, a);
    if (sh.reg_c > 0) {
        // sequences + misc (one wave), + the edge exchange area (several waves
        // of 8 chunks), + the exchange area and the int16 M ring (16 chunks)
        // (one wave with the int16 ring: 4 misc words + 3 slots of 64*reg_c + 2 entries)
        if (sh.nw == 1 && sh.lring) return 4ll * (2 * a.seq_words + 4) + ((3 * (64 * sh.reg_c + 2) * 2 + 3) & ~3);
        const int ex = sh.nw == 1 ? 16 : sh.reg_c == 16 ? regml_exchange_words(sh.nw) : regmw_exchange_words(sh.nw);
        return 4ll * (2 * a.seq_words + ex);
    }
    int64_t b = lds_head_bytes(a.seq_words, a.rm + 2 * a.ri, sh.nw);
    if (sh.lds_ring) b += ring_bytes(sh, a.W, a.rm, a.ri);
    return b;
}

static KernelFn pick(const Shape& sh) {
    if (sh.two_piece) return pick_2p(sh);
    if (sh.reg_c > 0) return sh.nw == 1 ? pick_reg(sh) : sh.reg_c == 16 ? pick_regml(sh) : pick_regmw(sh);
    switch (sh.nw) {
    case 1: return pick_nw1(sh);
    case 2: return pick_nw2(sh);
    case 4: return pick_nw4(sh);
    case 8: return pick_nw8(sh);
    case 16: return pick_nw16(sh);
    default: return nullptr;
    }
}

int64_t max_lds_for(const Shape& sh) {
    KernelFn f = pick(sh);
    if (!f) return 0;
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess) return 0;
    int v = 0;
    if (hipDeviceGetAttribute(&v, hipDeviceAttributeMaxSharedMemoryPerBlock, dev) != hipSuccess) return 0;
    return v;
}

// Occupancy answers are cached per (device, kernel, LDS bytes): the server
// plans a launch for every flight, and the runtime query is not free.
namespace {
struct OccKey {
    int dev;
    KernelFn f;
    int64_t lds;
    bool operator<(const OccKey& o) const {
        if (dev != o.dev) return dev < o.dev;
        if (f != o.f) return reinterpret_cast<uintptr_t>(f) < reinterpret_cast<uintptr_t>(o.f);
        return lds < o.lds;
    }
};
std::mutex g_occ_mu;
std::map<OccKey, int> g_occ;
}  // namespace

int occupancy(const Shape& sh, int64_t lds, int* blocks_per_cu) {
    KernelFn f = pick(sh);
    if (!f) return (int)hipErrorInvalidValue;
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess) return (int)hipErrorInvalidDevice;
    const OccKey key{dev, f, lds};
    {
        std::lock_guard<std::mutex> g(g_occ_mu);
        auto it = g_occ.find(key);
        if (it != g_occ.end()) { *blocks_per_cu = it->second; return 0; }
    }
    if (lds > 65536) {
        hipError_t e = hipFuncSetAttribute(reinterpret_cast<const void*>(f), hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
        if (e != hipSuccess) return (int)e;
    }
    int n = 0;
    hipError_t e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&n, reinterpret_cast<const void*>(f), sh.nw * 64, (size_t)lds);
    if (e != hipSuccess) return (int)e;
    *blocks_per_cu = n;
    std::lock_guard<std::mutex> g(g_occ_mu);
    g_occ[key] = n;
    return 0;
}

int launch_wfa(const Shape& sh, const KArgs& a, int64_t grid, void* stream) {
    KernelFn f = pick(sh);
    if (!f) return (int)hipErrorInvalidValue;
    if (grid <= 0) return 0;
    if (grid > 0x7fffffff) return (int)hipErrorInvalidValue;
    const int64_t lds = lds_bytes(sh, a);
    if (lds > 65536) {
        hipError_t e = hipFuncSetAttribute(reinterpret_cast<const void*>(f), hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
        if (e != hipSuccess) return (int)e;
    }
    (void)hipGetLastError();  // every earlier call was checked where it was made
    hipLaunchKernelGGL(f, dim3((unsigned)grid), dim3(sh.nw * 64), (size_t)lds, (hipStream_t)stream, a);
    return (int)hipGetLastError();
}

int launch_copy(const void* src, void* dst, int64_t bytes, void* stream) {
    const int64_t n = bytes / 16;
    if (n <= 0) return 0;
    int64_t grid = (n + 255) / 256;
    if (grid > 65536) grid = 65536;
    (void)hipGetLastError();
    hipLaunchKernelGGL(copy_kernel, dim3((unsigned)grid), dim3(256), 0, (hipStream_t)stream,
                       reinterpret_cast<const uint4*>(src), reinterpret_cast<uint4*>(dst), n);
    return (int)hipGetLastError();
}

}  // namespace tsu
