// tsu_host.h -- host-only internals shared by api.cpp and serve.cpp.
#pragma once

#include <hip/hip_runtime_api.h>

#include <algorithm>
#include <climits>
#include <cstdlib>
#include <memory>
#include <string>
#include <vector>

#include "tsu_internal.h"

using tsu::KArgs;
using tsu::KPair;
using tsu::Shape;

struct tsu_ctx {
    int device = 0;
    int32_t x = 0, o = 0, e = 0;        // penalties as given
    int32_t xs = 0, os = 0, es = 0, g = 1;  // divided by gcd
    int32_t rm = 0, ri = 0;
    bool two_piece = false;             // gap-affine 2-piece penalties
    int32_t o2 = 0, e2 = 0;             // second gap piece as given
    int32_t o2s = 0, e2s = 0, r2 = 0;   // divided by gcd; I2/D2 ring depth
    hipStream_t stream = nullptr;
    hipStream_t stream2 = nullptr;      // batched form: launches beside those on `stream`
    hipStream_t stream3 = nullptr;      // batched form: a third stream of launches
    hipStream_t stream_io = nullptr;    // batched form: score copies while launches run
    hipEvent_t ev0 = nullptr, ev1 = nullptr;
    int cu_count = 0;
    int64_t max_lds = 0;
    char name[256] = {0};
    std::string err;
    std::vector<tsu_tune_entry> tune;   // per-bucket shapes (autotuner)
    // backtrace memory of tsu_align_cigar (tsu_set_cigar_memory)
    int32_t cigar_mode = TSU_CIGAR_AUTO;
    int32_t cigar_ck_every = 0;         // 0 = chosen per launch
    int64_t cigar_budget = 0;           // 0 = half the free device memory
    int32_t poll_limit = 0;             // neighbour-sync polls before a pair fails (0 = default)
    // scratch reused across calls
    int32_t* d_queue = nullptr;
    void* scratch = nullptr;            // this device's no-allocation scratch (tsu::scratch_address)
    void* h_stage = nullptr;            // pinned staging for large device-to-host copies
    size_t h_stage_bytes = 0;
    // backtrace state of tsu_align_cigar (records, step bounds, step offsets,
    // checkpoints; then records, step bounds and step offsets of the launch
    // on the second stream), kept between calls and grown when a launch
    // needs more: freeing tens of GB of written device memory can take a
    // second.  Released by tsu_set_cigar_memory and tsu_destroy.
    void* bt_buf[7] = {};
    size_t bt_bytes[7] = {};
};

struct tsu_batch {
    tsu_ctx* ctx = nullptr;        // compared, never dereferenced after creation
    int device = 0;                // freeing does not need the context
    uint32_t* d_words = nullptr;
    KPair* d_pairs = nullptr;
    int32_t* d_scores = nullptr;
    int32_t n = 0;
    int64_t n_words = 0;
    std::vector<tsu_pair> pairs;   // host copy of the descriptors
    int32_t max_len = 0;
    double cells = 0.0;
};

// Host buffers written once right after allocation (the CIGAR run pool):
// from 8 MiB on, 2 MiB aligned and advised as transparent huge pages, so the
// first writes take a fault per 2 MiB instead of per 4 KiB page; smaller
// ones plain malloc.  Either is released with free.
struct HugeHostFree {
    void operator()(void* p) const { std::free(p); }
};
void* alloc_host_huge(size_t bytes);

struct tsu_cigars {
    std::unique_ptr<uint32_t, HugeHostFree> runs;  // run pool as copied from the device (not zero-filled)
    std::vector<int64_t> off;
    std::vector<int32_t> n;
    int64_t state_bytes = 0;       // largest backtrace state of one launch (device)
    int32_t record_launches = 0;   // launches with records for every score
    int32_t checkpoint_launches = 0;
    uint32_t kernels = 0;          // TSU_KERNEL_* bits of the launches that recorded
};

namespace tsu_host {
// TSU_KERNEL_* bit of a backtrace launch of this shape.
uint32_t kernel_bit(const Shape& sh);
}  // namespace tsu_host

namespace tsu_host {

// Kernel arguments with the context's scaled penalties and ring depths.
KArgs base_args(const tsu_ctx* ctx);
// LDS words per staged sequence for sequences up to max_len bases.
int32_t seq_words_for(const tsu_ctx* ctx, int32_t max_len);
// Marks an LDS-kernel shape whose sequences are longer than LDS holds: that
// kernel then reads them from HBM (max_seq_len; seq_words_for is then 0).
void place_sequences(const tsu_ctx* ctx, Shape& sh, int32_t max_len);
// Length bucket bounds on max(p, t) and the bucket of a length.
int bucket_count();
int32_t bucket_bound(int i);
int bucket_of(int32_t len);
// First diagonal capacity tried for a bucket (widened on overflow).
int32_t initial_capacity(const tsu_ctx* ctx, int32_t bound);
// Capacity no pair of lengths (p, t) can exceed.
int32_t capacity_full(int32_t p, int32_t t);
// Highest capacity the batched and served forms widen to for pairs of at
// most max_len bases: capacity_full, or with 2-piece penalties (whose kernel
// keeps its five rings in LDS) the widest window whose rings fit the LDS.
int32_t capacity_max(const tsu_ctx* ctx, int32_t max_len);
// Persistent-launch shape for capacity W (ring in HBM when LDS is too small).
Shape batched_shape(const tsu_ctx* ctx, int32_t W, int32_t max_len, int nw_hint);
// Tuned waves per pair for a bucket (0 when the autotuner has not run).
int tuned_waves(const tsu_ctx* ctx, int32_t bound);
// Tuned initial capacity for a bucket (0 when none was tuned).
int32_t tuned_capacity(const tsu_ctx* ctx, int32_t bound);
// Resident workgroups of a shape on the whole device.
int64_t resident_grid(const tsu_ctx* ctx, const Shape& sh, const KArgs& a, int* status);
int fail(tsu_ctx* ctx, int status, const std::string& msg);
// The context's penalties divide to (2, 3, 1), the case with a compile-time ring.
bool fixed_242(const tsu_ctx* ctx);
// Chunks of 64 diagonals of the register-resident kernel for capacity W (0: LDS kernel).
int reg_chunks(const tsu_ctx* ctx, int32_t W);
// Longest sequence staged in LDS (both sequences of a pair are, next to the
// kernel's own state); longer ones are read from HBM by the LDS kernel.
int32_t max_seq_len(const tsu_ctx* ctx);
// Longest sequence a pair may hold: kMaxSeqLen (int32 offsets and capacities
// of up to 2^29 diagonals), or max_seq_len with 2-piece penalties, whose
// kernel stages both sequences in LDS.
constexpr int32_t kMaxSeqLen = 1 << 28;
int32_t seq_len_limit(const tsu_ctx* ctx);
// Waves of the multi-wave register kernel for capacity W (0: not eligible).
int reg_waves(const tsu_ctx* ctx, int32_t W);
// Chunks (1, 2, 4, 8, 12 or 16) of the 2-piece single-wave register kernel
// for capacity W: extensions (e1, e2) of (1, 1), (2, 1) or (3, 1) after the
// gcd, sequences up to 16000 bases, its M ring in LDS (0: the LDS-ring
// 2-piece kernel).
int reg2p_chunks(const tsu_ctx* ctx, int32_t W, int32_t max_len);
// Waves of the 16-chunk multi-wave kernel (older M scores in an int16 LDS
// ring) for capacity W and sequences up to max_len (0: not eligible).
int ml_waves(const tsu_ctx* ctx, int32_t W, int32_t max_len);

// Probe of the batched and served forms: the single-wave register kernel
// runs a pair up to kProbeCap scaled scores (its 512-diagonal window holds
// the range that far), and the furthest h + v it reached gives the capacity
// the pair is placed at (probe_capacity).
constexpr int32_t kProbeCap = 240;
// margin: the factor of spare range (kProbeMargin for throughput; the
// server passes kProbeMarginServed, since a re-run adds to its requests'
// latency).
constexpr double kProbeMargin = 1.1;
constexpr double kProbeMarginServed = 1.2;
int32_t probe_capacity(const tsu_ctx* ctx, const tsu_pair& pr, int32_t progress, double margin = kProbeMargin);
// Pairs of a length bucket are probed when the register kernels take them
// (their sequences, up to max_len bases, fit in LDS) and the bucket's first
// capacity (W) needs several waves.
inline bool probe_bucket(const tsu_ctx* ctx, int32_t W, int32_t tuned_cap, int tuned_waves, int32_t max_len) {
    return fixed_242(ctx) && tuned_cap <= 0 && tuned_waves <= 1 && W > 1024 && max_len <= max_seq_len(ctx);
}

// Reorders idx by key(i) descending, keeping the order of equal keys (a
// counting sort over the key range, O(n + range), or a comparison sort when
// the range is much wider than n).  Keys must be non-negative.
template <typename KeyFn>
void order_desc(std::vector<int32_t>& idx, KeyFn key) {
    if (idx.size() < 2) return;
    int64_t lo = INT64_MAX, hi = INT64_MIN;
    for (int32_t i : idx) { const int64_t k = key(i); lo = k < lo ? k : lo; hi = k > hi ? k : hi; }
    const int64_t range = hi - lo + 1;
    if (range > 4 * (int64_t)idx.size() + 65536) {
        std::stable_sort(idx.begin(), idx.end(), [&](int32_t a, int32_t b) { return key(a) > key(b); });
        return;
    }
    std::vector<int64_t> start((size_t)range + 1, 0);
    for (int32_t i : idx) ++start[(size_t)(hi - key(i)) + 1];
    for (int64_t r = 1; r <= range; ++r) start[(size_t)r] += start[(size_t)r - 1];
    std::vector<int32_t> out(idx.size());
    for (int32_t i : idx) out[(size_t)start[(size_t)(hi - key(i))]++] = i;
    idx.swap(out);
}

// Longest first (p + t descending): the order persistent launches pop pairs in.
void order_longest_first(std::vector<int32_t>& idx, const tsu_pair* pairs);

}  // namespace tsu_host
