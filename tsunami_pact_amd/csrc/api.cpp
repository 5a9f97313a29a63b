// api.cpp -- host side of libtsunami_hip.so: contexts, device batches, the
// plain / batched / CIGAR alignment entry points and the autotuner.
//
// Replaces the inline host code of the reference's main() (wfa_gpu.cu:43-148):
// argument handling, cudaMalloc/cudaMemcpy, the <<<N,32>>> launch, kernel-only
// timing and the score copy.  Unlike the reference, device memory is sized for
// the workgroups that are resident at once (the rings live in LDS), not for
// all N pairs (wfa_gpu.cu:100-108 allocate num_wf*wf_length*N shorts x3).
#include <hip/hip_runtime_api.h>
#include <sys/mman.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <map>
#include <memory>
#include <mutex>
#include <new>
#include <numeric>
#include <string>
#include <thread>
#include <vector>

#include "tsu_host.h"

using tsu_host::fail;

namespace tsu_host {

int fail(tsu_ctx* ctx, int status, const std::string& msg) {
    if (ctx) ctx->err = msg;
    return status;
}

}  // namespace tsu_host

namespace {

int64_t gcd64(int64_t a, int64_t b) {
    while (b) { int64_t t = a % b; a = b; b = t; }
    return a < 0 ? -a : a;
}

double now_ms() {
    return std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now().time_since_epoch()).count();
}

int hip_fail(tsu_ctx* ctx, hipError_t e, const char* what) {
    return fail(ctx, e == hipErrorOutOfMemory ? TSU_ENOMEM : TSU_EHIP, std::string(what) + ": " + hipGetErrorString(e));
}

#define HIPCHK(call, what)                                   \
    do {                                                     \
        hipError_t _e = (call);                              \
        if (_e != hipSuccess) return hip_fail(ctx, _e, what); \
    } while (0)

// Device buffer that frees itself.
template <typename T>
struct DevBuf {
    T* p = nullptr;
    size_t n = 0;
    ~DevBuf() { if (p) (void)hipFree(p); }
    hipError_t alloc(size_t count) {
        if (p) { (void)hipFree(p); p = nullptr; }
        n = count;
        if (count == 0) return hipSuccess;
        return hipMalloc(reinterpret_cast<void**>(&p), count * sizeof(T));
    }
    // at least count elements, keeping the buffer when it is large enough
    hipError_t reserve(size_t count) { return p && n >= count ? hipSuccess : alloc(count); }
};

// Diagonal capacity that holds every k-range up to score bound s (scaled):
// |k| <= (s - o) / e + 1 on each side (+ the union's +-1).
int32_t capacity_for_score(const tsu_ctx* ctx, int64_t s_scaled) {
    int64_t half = (s_scaled - ctx->os) / ctx->es + 2;
    // with two gap pieces the range grows with the faster-growing piece
    if (ctx->two_piece) half = std::max<int64_t>(half, (s_scaled - ctx->o2s) / ctx->e2s + 2);
    if (half < 4) half = 4;
    int64_t w = 2 * half + 1;
    return (int32_t)std::min<int64_t>(w, 1 << 26);
}

}  // namespace

// Diagonal capacity no pair of lengths (p, t) can exceed: the range lies in
// [-p-1, t+1].
int32_t tsu_host::capacity_full(int32_t p, int32_t t) { return p + t + 3; }
using tsu_host::capacity_full;

void* alloc_host_huge(size_t bytes) {
    constexpr size_t kHuge = 2u << 20;
    if (bytes < 4 * kHuge) return std::malloc(std::max<size_t>(bytes, 1));  // small: no 2 MiB rounding
    const size_t n = (bytes + kHuge - 1) / kHuge * kHuge;
    void* p = std::aligned_alloc(kHuge, n);
    if (p) (void)madvise(p, n, MADV_HUGEPAGE);  // advisory: plain pages where not granted
    return p;
}

// A pair overflowed the widest window its form may use: the caller's error.
static int overflow_at(tsu_ctx* ctx, int32_t W, int32_t max_len) {
    if (W < capacity_full(max_len, max_len))  // 2-piece: the LDS holds no wider rings
        return fail(ctx, TSU_EKOVERFLOW, "a pair's 2-piece wavefront range does not fit the CU's LDS (" +
                                             std::to_string(W) + " diagonals)");
    return fail(ctx, TSU_EINTERNAL, "overflow at full capacity");
}

// ---------------------------------------------------------------------------
// context
// ---------------------------------------------------------------------------
extern "C" int tsu_create(tsu_ctx** out, int device, int32_t mismatch, int32_t gap_opening, int32_t gap_extension) {
    if (!out) return TSU_EINVAL;
    *out = nullptr;
    if (mismatch <= 0 || gap_opening < 0 || gap_extension <= 0) return TSU_EINVAL;
    if (mismatch > (1 << 20) || gap_opening > (1 << 20) || gap_extension > (1 << 20)) return TSU_EINVAL;
    tsu_ctx* ctx = new (std::nothrow) tsu_ctx();
    if (!ctx) return TSU_ENOMEM;
    ctx->device = device;
    ctx->x = mismatch; ctx->o = gap_opening; ctx->e = gap_extension;
    ctx->g = (int32_t)gcd64(gcd64(mismatch, gap_opening), gap_extension);
    ctx->xs = mismatch / ctx->g; ctx->os = gap_opening / ctx->g; ctx->es = gap_extension / ctx->g;
    ctx->rm = std::max(ctx->xs, ctx->os + ctx->es) + 1;
    ctx->ri = ctx->es + 1;
    if (ctx->rm > 4096) { delete ctx; return TSU_EINVAL; }
    int ndev = 0;
    hipError_t e = hipGetDeviceCount(&ndev);
    if (e != hipSuccess || ndev <= 0) { delete ctx; return TSU_EHIP; }
    if (device < 0 || device >= ndev) { delete ctx; return TSU_EINVAL; }
    if ((e = hipSetDevice(device)) != hipSuccess) { delete ctx; return TSU_EHIP; }
    hipDeviceProp_t prop;
    if ((e = hipGetDeviceProperties(&prop, device)) != hipSuccess) { delete ctx; return TSU_EHIP; }
    // the no-allocation entry point's scratch: its address on this device,
    // resolved once per device with the device current
    if (tsu::scratch_address(device, &ctx->scratch) != 0) { delete ctx; return TSU_EHIP; }
    ctx->cu_count = prop.multiProcessorCount;
    ctx->max_lds = (int64_t)prop.sharedMemPerBlock;
    std::snprintf(ctx->name, sizeof(ctx->name), "%.180s (%.60s)", prop.name, prop.gcnArchName);
    if (hipStreamCreateWithFlags(&ctx->stream, hipStreamNonBlocking) != hipSuccess ||
        hipStreamCreateWithFlags(&ctx->stream2, hipStreamNonBlocking) != hipSuccess ||
        hipStreamCreateWithFlags(&ctx->stream3, hipStreamNonBlocking) != hipSuccess ||
        hipStreamCreateWithFlags(&ctx->stream_io, hipStreamNonBlocking) != hipSuccess ||
        hipEventCreate(&ctx->ev0) != hipSuccess || hipEventCreate(&ctx->ev1) != hipSuccess ||
        hipMalloc(reinterpret_cast<void**>(&ctx->d_queue), 64 * sizeof(int32_t)) != hipSuccess) {
        tsu_destroy(ctx);
        return TSU_EHIP;
    }
    *out = ctx;
    return TSU_OK;
}

extern "C" int tsu_create_2p(tsu_ctx** out, int device, int32_t mismatch, int32_t gap_opening1, int32_t gap_extension1,
                             int32_t gap_opening2, int32_t gap_extension2) {
    if (!out) return TSU_EINVAL;
    *out = nullptr;
    if (gap_opening2 < 0 || gap_extension2 <= 0 || gap_opening2 > (1 << 20) || gap_extension2 > (1 << 20)) return TSU_EINVAL;
    int r = tsu_create(out, device, mismatch, gap_opening1, gap_extension1);
    if (r != TSU_OK) return r;
    tsu_ctx* ctx = *out;
    ctx->two_piece = true;
    ctx->o2 = gap_opening2; ctx->e2 = gap_extension2;
    ctx->g = (int32_t)gcd64(gcd64(gcd64(gcd64(mismatch, gap_opening1), gap_extension1), gap_opening2), gap_extension2);
    ctx->xs = mismatch / ctx->g; ctx->os = gap_opening1 / ctx->g; ctx->es = gap_extension1 / ctx->g;
    ctx->o2s = gap_opening2 / ctx->g; ctx->e2s = gap_extension2 / ctx->g;
    ctx->rm = std::max(std::max(ctx->xs, ctx->os + ctx->es), ctx->o2s + ctx->e2s) + 1;
    ctx->ri = ctx->es + 1;
    ctx->r2 = ctx->e2s + 1;
    if (ctx->rm > 4096) { tsu_destroy(ctx); *out = nullptr; return TSU_EINVAL; }
    return TSU_OK;
}

// Frees the backtrace state the context keeps between tsu_align_cigar calls.
static void bt_release(tsu_ctx* ctx) {
    for (int i = 0; i < 7; ++i) {
        if (ctx->bt_buf[i]) (void)hipFree(ctx->bt_buf[i]);
        ctx->bt_buf[i] = nullptr;
        ctx->bt_bytes[i] = 0;
    }
}

// Backtrace state buffer i with at least `bytes`, grown (never shrunk) in
// the context.  The launches that used the old buffer have completed.
static hipError_t bt_reserve(tsu_ctx* ctx, int i, size_t bytes, void** p) {
    if (ctx->bt_bytes[i] < bytes) {
        if (ctx->bt_buf[i]) (void)hipFree(ctx->bt_buf[i]);
        ctx->bt_buf[i] = nullptr;
        ctx->bt_bytes[i] = 0;
        hipError_t e = hipMalloc(&ctx->bt_buf[i], bytes);
        if (e != hipSuccess) { ctx->bt_buf[i] = nullptr; return e; }
        ctx->bt_bytes[i] = bytes;
    }
    *p = ctx->bt_buf[i];
    return hipSuccess;
}

extern "C" void tsu_destroy(tsu_ctx* ctx) {
    if (!ctx) return;
    (void)hipSetDevice(ctx->device);
    if (ctx->stream) (void)hipStreamSynchronize(ctx->stream);
    if (ctx->stream2) (void)hipStreamSynchronize(ctx->stream2);
    if (ctx->stream3) (void)hipStreamSynchronize(ctx->stream3);
    if (ctx->stream_io) (void)hipStreamSynchronize(ctx->stream_io);
    bt_release(ctx);
    if (ctx->d_queue) (void)hipFree(ctx->d_queue);
    if (ctx->h_stage) (void)hipHostFree(ctx->h_stage);
    if (ctx->ev0) (void)hipEventDestroy(ctx->ev0);
    if (ctx->ev1) (void)hipEventDestroy(ctx->ev1);
    if (ctx->stream) (void)hipStreamDestroy(ctx->stream);
    if (ctx->stream2) (void)hipStreamDestroy(ctx->stream2);
    if (ctx->stream3) (void)hipStreamDestroy(ctx->stream3);
    if (ctx->stream_io) (void)hipStreamDestroy(ctx->stream_io);
    delete ctx;
}

extern "C" const char* tsu_last_error(const tsu_ctx* ctx) { return ctx ? ctx->err.c_str() : "no context"; }

extern "C" int tsu_device_info(tsu_ctx* ctx, int32_t* cu_count, int32_t* lds_bytes, char* name, int32_t name_cap) {
    if (!ctx) return TSU_EINVAL;
    if (cu_count) *cu_count = ctx->cu_count;
    if (lds_bytes) *lds_bytes = (int32_t)ctx->max_lds;
    if (name && name_cap > 0) std::snprintf(name, (size_t)name_cap, "%s", ctx->name);
    return TSU_OK;
}

extern "C" int tsu_sync(tsu_ctx* ctx) {
    if (!ctx) return TSU_EINVAL;
    HIPCHK(hipSetDevice(ctx->device), "hipSetDevice");
    HIPCHK(hipStreamSynchronize(ctx->stream), "hipStreamSynchronize");
    return TSU_OK;
}

// ---------------------------------------------------------------------------
// batches
// ---------------------------------------------------------------------------
static int batch_upload(tsu_ctx* ctx, const uint32_t* words, int64_t n_words, const tsu_pair* pairs, int32_t n,
                        tsu_batch** out) {
    *out = nullptr;
    for (int32_t i = 0; i < n; ++i) {
        const tsu_pair& p = pairs[i];
        if (p.p_len < 0 || p.t_len < 0 || p.p_off < 0 || p.t_off < 0 ||
            p.p_off + tsu_packed_words(p.p_len) > n_words || p.t_off + tsu_packed_words(p.t_len) > n_words)
            return fail(ctx, TSU_EINVAL, "pair " + std::to_string(i) + " lies outside the packed arena");
        if (p.p_len > tsu_host::seq_len_limit(ctx) || p.t_len > tsu_host::seq_len_limit(ctx))
            return fail(ctx, TSU_EINVAL, "pair " + std::to_string(i) + ": a sequence is longer than " +
                                             std::to_string(tsu_host::seq_len_limit(ctx)) + " bases" +
                                             (ctx->two_piece ? " (the 2-piece kernel stages both sequences in LDS)" : ""));
    }
    tsu_batch* b = new (std::nothrow) tsu_batch();
    if (!b) return fail(ctx, TSU_ENOMEM, "host allocation");
    b->ctx = ctx;
    b->device = ctx->device;
    b->n = n;
    b->n_words = n_words;
    try { b->pairs.assign(pairs, pairs + n); } catch (...) { delete b; return fail(ctx, TSU_ENOMEM, "host allocation"); }
    for (int32_t i = 0; i < n; ++i) {
        b->max_len = std::max(b->max_len, std::max(pairs[i].p_len, pairs[i].t_len));
        b->cells += (double)pairs[i].p_len * (double)pairs[i].t_len;
    }
    hipError_t e;
    if ((e = hipSetDevice(ctx->device)) != hipSuccess ||
        (e = hipMalloc(reinterpret_cast<void**>(&b->d_words), (size_t)std::max<int64_t>(n_words, 1) * 4)) != hipSuccess ||
        (e = hipMalloc(reinterpret_cast<void**>(&b->d_pairs), (size_t)std::max(n, 1) * sizeof(KPair))) != hipSuccess ||
        (e = hipMalloc(reinterpret_cast<void**>(&b->d_scores), (size_t)std::max(n, 1) * sizeof(int32_t))) != hipSuccess) {
        tsu_batch_free(b);
        return hip_fail(ctx, e, "hipMalloc (batch)");
    }
    if ((n_words > 0 && (e = hipMemcpy(b->d_words, words, (size_t)n_words * 4, hipMemcpyHostToDevice)) != hipSuccess) ||
        (n > 0 && (e = hipMemcpy(b->d_pairs, pairs, (size_t)n * sizeof(KPair), hipMemcpyHostToDevice)) != hipSuccess)) {
        tsu_batch_free(b);
        return hip_fail(ctx, e, "hipMemcpy H2D (batch)");
    }
    *out = b;
    return TSU_OK;
}

extern "C" int tsu_batch_create(tsu_ctx* ctx, const uint32_t* words, int64_t n_words, const tsu_pair* pairs,
                                int32_t n_pairs, tsu_batch** out) {
    if (!ctx || !out || n_pairs < 0 || n_words < 0 || (n_pairs > 0 && !pairs) || (n_words > 0 && !words))
        return fail(ctx, TSU_EINVAL, "tsu_batch_create: bad arguments");
    return batch_upload(ctx, words, n_words, pairs, n_pairs, out);
}

extern "C" int tsu_batch_from_dataset(tsu_ctx* ctx, const tsu_dataset* ds, int64_t first, int64_t count,
                                      tsu_batch** out) {
    if (!ctx || !ds || !out || first < 0 || count < 0 || first + count > (int64_t)ds->p_len.size() || count > INT32_MAX)
        return fail(ctx, TSU_EINVAL, "tsu_batch_from_dataset: bad range");
    std::vector<uint32_t> words;
    std::vector<tsu_pair> pairs;
    int r = tsu_pack_dataset(ds, first, count, words, pairs);
    if (r != TSU_OK) return fail(ctx, r, "packing failed (non-ACGT base?)");
    return batch_upload(ctx, words.data(), (int64_t)words.size(), pairs.data(), (int32_t)count, out);
}

extern "C" int32_t tsu_batch_size(const tsu_batch* b) { return b ? b->n : 0; }

extern "C" int tsu_batch_device(const tsu_batch* b, const uint32_t** d_words, int64_t* n_words,
                                const tsu_pair** d_pairs, int32_t** d_scores) {
    if (!b) return TSU_EINVAL;
    if (d_words) *d_words = b->d_words;
    if (n_words) *n_words = b->n_words;
    if (d_pairs) *d_pairs = reinterpret_cast<const tsu_pair*>(b->d_pairs);
    if (d_scores) *d_scores = b->d_scores;
    return TSU_OK;
}

extern "C" int64_t tsu_batch_bytes(const tsu_batch* b) {
    if (!b) return 0;
    return b->n_words * 4 + (int64_t)b->n * (int64_t)(sizeof(KPair) + sizeof(int32_t));
}

extern "C" void tsu_batch_free(tsu_batch* b) {
    if (!b) return;
    (void)hipSetDevice(b->device);
    if (b->d_words) (void)hipFree(b->d_words);
    if (b->d_pairs) (void)hipFree(b->d_pairs);
    if (b->d_scores) (void)hipFree(b->d_scores);
    delete b;
}

// ---------------------------------------------------------------------------
// shapes and launches
// ---------------------------------------------------------------------------
KArgs tsu_host::base_args(const tsu_ctx* ctx) {
    KArgs a;
    std::memset(&a, 0, sizeof(a));
    a.x = ctx->xs; a.o = ctx->os; a.e = ctx->es; a.g = ctx->g;
    a.rm = ctx->rm; a.ri = ctx->ri;
    a.o2 = ctx->o2s; a.e2 = ctx->e2s; a.r2 = ctx->r2;
    a.poll_limit = ctx->poll_limit;
    return a;
}

int32_t tsu_host::seq_words_for(const tsu_ctx* ctx, int32_t max_len) {
    return max_len > max_seq_len(ctx) ? 0 : (int32_t)((max_len + 15) / 16 + 2);
}

void tsu_host::place_sequences(const tsu_ctx* ctx, Shape& sh, int32_t max_len) {
    sh.gseq = sh.reg_c == 0 && !sh.two_piece && max_len > max_seq_len(ctx);
}

// Longest sequence staged in LDS: both sequences of a pair, next to the LDS
// kernel's slot limits and misc words.
int32_t tsu_host::max_seq_len(const tsu_ctx* ctx) {
    const int64_t words = ctx->max_lds / 4 - 2 * (ctx->rm + 2 * ctx->ri + 1) - 16 - 4;
    return (int32_t)std::max<int64_t>(0, (words / 2 - 2) * 16);
}

int32_t tsu_host::seq_len_limit(const tsu_ctx* ctx) { return ctx->two_piece ? max_seq_len(ctx) : kMaxSeqLen; }

using tsu_host::base_args;
using tsu_host::seq_words_for;

namespace {

// Waves per pair for a diagonal capacity: enough chunks of 64 diagonals per
// wave to amortise the per-step bookkeeping, at most 16 waves.
int default_waves(int32_t W) {
    const int chunks = (W + 63) / 64;
    int nw = 1;
    while (nw < 16 && chunks >= nw * 2 * 4) nw *= 2;
    return nw;
}

}  // namespace

int32_t tsu_host::capacity_max(const tsu_ctx* ctx, int32_t max_len) {
    const int32_t full = capacity_full(max_len, max_len);
    if (!ctx->two_piece) return full;
    Shape s2{1, max_len <= 32000, true, false, true};
    s2.two_piece = true;
    KArgs a = base_args(ctx);
    a.seq_words = seq_words_for(ctx, max_len);
    int32_t lo = 0, hi = full;  // lo fits (or is 0), hi the highest candidate
    while (lo < hi) {
        const int32_t mid = lo + (hi - lo + 1) / 2;
        a.W = mid;
        if (tsu::lds_bytes(s2, a) <= ctx->max_lds) lo = mid; else hi = mid - 1;
    }
    return std::max<int32_t>(lo, 1);
}

int64_t tsu_host::resident_grid(const tsu_ctx* ctx, const Shape& sh, const KArgs& a, int* status) {
    int per_cu = 0;
    const int r = tsu::occupancy(sh, tsu::lds_bytes(sh, a), &per_cu);
    *status = r;
    if (r != 0 || per_cu <= 0) return 0;
    return (int64_t)per_cu * ctx->cu_count;
}

// Device arrays a launch reads and writes: packed words, descriptors and
// per-pair scores (a batch's own, or a caller's).
struct DevIO {
    const uint32_t* words;
    const KPair* pairs;
    int32_t* scores;
};

static DevIO io_of(const tsu_batch* b) { return DevIO{b->d_words, b->d_pairs, b->d_scores}; }

// Launches one shape on `stream` (queue counter `queue`, events e0/e1 around
// the kernel) without waiting; an HBM ring, when the shape needs one, is left
// in `ring` and must outlive the kernel.  *launched is false when there was
// nothing to launch.
static int launch_shape(tsu_ctx* ctx, const DevIO& io, Shape sh, int32_t W, const int32_t* d_order, int32_t n_entries,
                        int32_t max_len, int64_t grid_cap, KArgs* extra, hipStream_t stream, int32_t* queue,
                        hipEvent_t e0, hipEvent_t e1, DevBuf<uint8_t>& ring, bool* launched) {
    *launched = false;
    if (max_len > tsu_host::max_seq_len(ctx)) {
        // only the LDS kernel reads sequences in place
        if (sh.reg_c > 0 || sh.two_piece)
            return fail(ctx, TSU_EINTERNAL, "a register or 2-piece launch was given sequences longer than LDS holds");
        sh.gseq = true;
        sh.off16 = false;
    }
    KArgs a = extra ? *extra : base_args(ctx);
    a.x = ctx->xs; a.o = ctx->os; a.e = ctx->es; a.g = ctx->g; a.rm = ctx->rm; a.ri = ctx->ri;
    a.words = io.words;
    a.pairs = io.pairs;
    a.order = d_order;
    a.n = n_entries;
    a.scores = io.scores;
    a.W = W;
    a.seq_words = seq_words_for(ctx, max_len);
    a.queue = queue;
    const int64_t lds = tsu::lds_bytes(sh, a);
    if (lds > ctx->max_lds) return fail(ctx, TSU_EINVAL, "shape does not fit in LDS");
    int64_t grid = n_entries;
    if (sh.queue) {
        int per_cu = 0;
        int r = tsu::occupancy(sh, lds, &per_cu);
        if (r != 0) return hip_fail(ctx, (hipError_t)r, "occupancy query");
        if (per_cu <= 0) return fail(ctx, TSU_EINVAL, "shape has zero occupancy");
        grid = std::min<int64_t>((int64_t)per_cu * ctx->cu_count, n_entries);
        if (grid_cap > 0) grid = std::min(grid, grid_cap);
        HIPCHK(hipMemsetAsync(queue, 0, sizeof(int32_t), stream), "hipMemsetAsync (queue)");
        if (!sh.lds_ring) {
            a.ring_stride = tsu::ring_bytes(sh, W, ctx->rm, ctx->ri);
            HIPCHK(ring.alloc((size_t)(a.ring_stride * std::max<int64_t>(grid, 1))), "hipMalloc (ring)");
            a.ring_g = ring.p;
        }
    } else if (!sh.lds_ring || sh.bt) {
        return fail(ctx, TSU_EINVAL, "plain launches keep the ring in LDS");
    }
    if (grid <= 0) return TSU_OK;
    HIPCHK(hipEventRecord(e0, stream), "hipEventRecord");
    int r = tsu::launch_wfa(sh, a, grid, stream);
    if (r != 0) return hip_fail(ctx, (hipError_t)r, "kernel launch");
    HIPCHK(hipEventRecord(e1, stream), "hipEventRecord");
    *launched = true;
    return TSU_OK;
}

// Runs one launch of `shape` over the queue entries `order` (or pairs [0,n)
// when order is null and the launch is plain) on the context's stream and
// waits for it.  Scores land in b->d_scores.
static int run_shape(tsu_ctx* ctx, const DevIO& io, Shape sh, int32_t W, const int32_t* d_order, int32_t n_entries,
                     int32_t max_len, int64_t grid_cap, KArgs* extra, double* kernel_ms) {
    DevBuf<uint8_t> ring;
    bool launched = false;
    if (kernel_ms) *kernel_ms = 0;
    int r = launch_shape(ctx, io, sh, W, d_order, n_entries, max_len, grid_cap, extra, ctx->stream, ctx->d_queue,
                         ctx->ev0, ctx->ev1, ring, &launched);
    if (r != TSU_OK || !launched) return r;
    HIPCHK(hipEventSynchronize(ctx->ev1), "hipEventSynchronize");
    float ms = 0.f;
    HIPCHK(hipEventElapsedTime(&ms, ctx->ev0, ctx->ev1), "hipEventElapsedTime");
    if (kernel_ms) *kernel_ms = ms;
    return TSU_OK;
}

bool tsu_host::fixed_242(const tsu_ctx* ctx) { return !ctx->two_piece && ctx->xs == 2 && ctx->os == 3 && ctx->es == 1; }

// Register-resident kernel window (chunks of 64 diagonals) for capacity W,
// or 0 when the LDS kernel is used (other penalties, or W above 512).
int tsu_host::reg_chunks(const tsu_ctx* ctx, int32_t W) {
    if (!tsu_host::fixed_242(ctx) || W > 512) return 0;
    int c = 1;
    while (64 * c < W) c *= 2;
    return c;
}

int tsu_host::reg2p_chunks(const tsu_ctx* ctx, int32_t W, int32_t max_len) {
    const int e = 10 * ctx->es + ctx->e2s;
    if (!ctx->two_piece || (e != 11 && e != 21 && e != 31) || W > 1024 || max_len > 16000) return 0;
    int c = 1;
    while (64 * c < W) c = c < 8 ? 2 * c : c + 4;  // 1, 2, 4, 8, 12, 16
    Shape sh{1, true, true, false, true};
    sh.two_piece = true;
    sh.reg_c = c;
    sh.reg_e = e;
    KArgs a = base_args(ctx);
    a.W = 64 * c;
    a.seq_words = seq_words_for(ctx, max_len);
    return tsu::lds_bytes(sh, a) <= ctx->max_lds ? c : 0;
}

int tsu_host::reg_waves(const tsu_ctx* ctx, int32_t W) {
    if (!tsu_host::fixed_242(ctx) || W <= 512 || W > 16 * 512) return 0;
    int nw = 2;
    while (512 * nw < W) nw *= 2;
    return nw;
}

// Capacities above this use the 16-chunk waves with the int16 LDS M ring
// rather than 8-chunk waves with the whole ring in registers: measured
// faster from 1,025 diagonals on (the 4-wave, neighbour-synchronised
// instance fits 5 pairs per CU), slower at 850.
constexpr int32_t kMlMinCapacity = 1024;

int tsu_host::ml_waves(const tsu_ctx* ctx, int32_t W, int32_t max_len) {
    if (!tsu_host::fixed_242(ctx) || W <= kMlMinCapacity || W > 16 * 1024 || max_len > 32000) return 0;
    int nw = 4;
    while (1024 * nw < W) nw *= 2;
    return nw;
}

// Register kernels when the penalties and the capacity allow: one wave for
// up to 512 diagonals, 2-16 waves of 8 chunks for up to 8192, 4-16 waves of
// 16 chunks (older M scores in LDS) for up to 16384 (and the two sequences
// fit in LDS).
static void use_register_kernel(const tsu_ctx* ctx, int32_t W, int32_t max_len, Shape& sh) {
    if (max_len > tsu_host::max_seq_len(ctx)) return;  // the register kernels stage both sequences in LDS
    Shape r = sh;
    if (int c = tsu_host::reg_chunks(ctx, W)) {
        r.reg_c = c;
        r.nw = 1;
        // the int16 LDS ring variant (doubled offsets fit int16 up to 16000
        // bases) measured faster only with 8 chunks on short sequences; with
        // narrower windows or long, extend-heavy pairs its ring traffic costs
        // more than the DPP moves it saves
        r.lring = c == 8 && max_len <= 4096;
    } else if (int nwl = tsu_host::ml_waves(ctx, W, max_len)) {
        r.reg_c = 16;
        r.nw = nwl;
        r.dbl = max_len <= tsu::kMlDblMaxLen;
    } else if (int nw = tsu_host::reg_waves(ctx, W)) {
        r.reg_c = 8;
        r.nw = nw;
    } else {
        return;
    }
    KArgs a = base_args(ctx);
    a.W = W;
    a.seq_words = seq_words_for(ctx, max_len);
    if (tsu::lds_bytes(r, a) <= ctx->max_lds) sh = r;
}

static Shape plain_shape(const tsu_ctx* ctx, int32_t W, int32_t max_len, bool* fits) {
    if (ctx->two_piece) {
        Shape s2{1, max_len <= 32000, true, false, false};
        s2.two_piece = true;
        s2.reg_c = tsu_host::reg2p_chunks(ctx, W, max_len);
        s2.reg_e = 10 * ctx->es + ctx->e2s;
        KArgs a = base_args(ctx);
        a.W = W;
        a.seq_words = seq_words_for(ctx, max_len);
        *fits = tsu::lds_bytes(s2, a) <= ctx->max_lds;
        return s2;
    }
    Shape sh{default_waves(W), max_len <= 32000, true, false, false};
    sh.fixed_242 = tsu_host::fixed_242(ctx);
    use_register_kernel(ctx, W, max_len, sh);
    tsu_host::place_sequences(ctx, sh, max_len);
    KArgs a = base_args(ctx);
    a.W = W;
    a.seq_words = seq_words_for(ctx, max_len);
    *fits = tsu::lds_bytes(sh, a) <= ctx->max_lds;
    return sh;
}

// Lowest of n scores: only a negative one (a status) calls for the per-pair
// checks after a call.
static int32_t min_score(const int32_t* s, int32_t n) {
    int32_t m = INT32_MAX;
#pragma omp simd reduction(min : m)
    for (int32_t i = 0; i < n; ++i) m = s[i] < m ? s[i] : m;
    return m;
}

// (below) with min32 non-null the copied bytes are int32 scores and *min32
// receives their lowest; *synced_ms (if given) the host clock once the first
// slice (all of it up to 32 MiB) is in pinned memory
static int copy_to_host(tsu_ctx* ctx, void* dst, const void* src, size_t bytes, int32_t* min32 = nullptr,
                        double* synced_ms = nullptr);

extern "C" int tsu_align_plain(tsu_ctx* ctx, tsu_batch* b, int32_t wf_length, int32_t* scores, tsu_timing* t) {
    if (!ctx || !b || !scores || b->ctx != ctx) return fail(ctx, TSU_EINVAL, "tsu_align_plain: bad arguments");
    if (wf_length < 8) wf_length = 8;
    // a capacity beyond p+t+3 can never be used
    int32_t W = std::min<int64_t>(wf_length, 2ll * b->max_len + 3);
    W = std::max(W, 8);
    HIPCHK(hipSetDevice(ctx->device), "hipSetDevice");
    const double w0 = now_ms();
    bool fits = false;
    Shape sh = plain_shape(ctx, W, b->max_len, &fits);
    if (sh.two_piece && !fits)
        return fail(ctx, TSU_EINVAL, "tsu_align_plain: 2-piece rings for wf_length " + std::to_string(W) + " do not fit in LDS");
    if (!fits) {
        // ring too wide for LDS: same one-workgroup-per-pair work, ring in HBM,
        // resident workgroups fed from the queue
        sh.queue = true;
        sh.lds_ring = false;
        sh.nw = 16;
    }
    // the launch, and the scores' copy queued behind it on the same stream:
    // one wait for both (the HBM ring, if any, outlives the wait)
    DevBuf<uint8_t> ring;
    bool launched = false;
    int r = launch_shape(ctx, io_of(b), sh, W, nullptr, b->n, b->max_len, 0, nullptr, ctx->stream, ctx->d_queue,
                         ctx->ev0, ctx->ev1, ring, &launched);
    if (r != TSU_OK) return r;
    int32_t lowest = 0;
    double synced = w0;
    r = copy_to_host(ctx, scores, b->d_scores, (size_t)b->n * sizeof(int32_t), &lowest, &synced);
    if (r != TSU_OK) return r;
    // wall: launch until the scores are in pinned host memory; d2h: from
    // there into the caller's array
    double wall = synced - w0;
    double d2h = now_ms() - synced;
    double kms = 0;
    if (launched) {
        float ms = 0.f;
        HIPCHK(hipEventElapsedTime(&ms, ctx->ev0, ctx->ev1), "hipEventElapsedTime");
        kms = ms;
    }
    int launches = 1;
    const bool clean = lowest >= 0;
    if (!clean && (sh.reg_c == 16 || sh.lring || (sh.two_piece && sh.reg_c))) {
        // the kernels with an int16 LDS ring stop past their score limit with
        // an overflow status: such pairs get the LDS kernel at the same
        // capacity, so whether a pair fits depends on the capacity and not on
        // that limit (the re-run's launch counts in the wall time)
        const double r0 = now_ms();
        std::vector<int32_t> again;
        for (int32_t i = 0; i < b->n; ++i)
            if (scores[i] == TSU_EKOVERFLOW) again.push_back(i);
        if (!again.empty()) {
            Shape s2{default_waves(W), b->max_len <= 32000, true, false, true};
            s2.fixed_242 = sh.fixed_242;
            if (sh.two_piece) {  // the 2-piece LDS-ring kernel
                s2.nw = 1;
                s2.two_piece = true;
            }
            tsu_host::place_sequences(ctx, s2, b->max_len);
            KArgs a2 = base_args(ctx);
            a2.W = W;
            a2.seq_words = seq_words_for(ctx, b->max_len);
            if (tsu::lds_bytes(s2, a2) > ctx->max_lds) { s2.lds_ring = false; s2.nw = 16; }
            DevBuf<int32_t> d_order;
            HIPCHK(d_order.alloc(again.size()), "hipMalloc (order)");
            HIPCHK(hipMemcpy(d_order.p, again.data(), again.size() * sizeof(int32_t), hipMemcpyHostToDevice), "hipMemcpy (order)");
            double kms2 = 0;
            r = run_shape(ctx, io_of(b), s2, W, d_order.p, (int32_t)again.size(), b->max_len, 0, nullptr, &kms2);
            if (r != TSU_OK) return r;
            kms += kms2;
            ++launches;
            wall += now_ms() - r0;
            const double d0 = now_ms();
            r = copy_to_host(ctx, scores, b->d_scores, (size_t)b->n * sizeof(int32_t));
            if (r != TSU_OK) return r;
            d2h += now_ms() - d0;
        }
    }
    if (t) {
        std::memset(t, 0, sizeof(*t));
        t->kernel_ms = kms; t->wall_ms = wall; t->d2h_ms = d2h;
        t->pairs = b->n; t->cells = b->cells; t->launches = launches;
    }
    if (clean) return TSU_OK;
    for (int32_t i = 0; i < b->n; ++i) {
        if (scores[i] == TSU_EKOVERFLOW) return fail(ctx, TSU_EKOVERFLOW, "pair " + std::to_string(i) + ": diagonal range exceeded wf_length " + std::to_string(wf_length));
    }
    for (int32_t i = 0; i < b->n; ++i) {
        if (scores[i] < 0) return fail(ctx, scores[i], "pair " + std::to_string(i) + ": " + tsu_strerror(scores[i]));
    }
    return TSU_OK;
}

extern "C" int tsu_align_fixed(tsu_ctx* ctx, const char* patterns, const char* texts, int32_t num_couples,
                               int32_t seq_len, int32_t wf_length, int32_t* scores, tsu_timing* t) {
    if (!ctx || num_couples < 0 || seq_len < 0 || (num_couples > 0 && (!patterns || !texts || !scores)))
        return fail(ctx, TSU_EINVAL, "tsu_align_fixed: bad arguments");
    const int64_t wps = tsu_packed_words(seq_len);
    std::vector<uint32_t> words;
    std::vector<tsu_pair> pairs((size_t)num_couples);
    try { words.assign((size_t)(2 * wps * (int64_t)num_couples), 0u); } catch (...) { return fail(ctx, TSU_ENOMEM, "host allocation"); }
    int bad = 0;
#pragma omp parallel for schedule(static, 1024) reduction(+:bad)
    for (int32_t i = 0; i < num_couples; ++i) {
        tsu_pair& pr = pairs[(size_t)i];
        pr.p_off = 2 * wps * (int64_t)i; pr.t_off = pr.p_off + wps;
        pr.p_len = seq_len; pr.t_len = seq_len;
        if (tsu_pack_seq(patterns + (int64_t)i * seq_len, seq_len, words.data() + pr.p_off) != TSU_OK) bad++;
        if (tsu_pack_seq(texts + (int64_t)i * seq_len, seq_len, words.data() + pr.t_off) != TSU_OK) bad++;
    }
    if (bad) return fail(ctx, TSU_EBADINPUT, "non-ACGT base in input");
    const double h0 = now_ms();
    tsu_batch* b = nullptr;
    int r = batch_upload(ctx, words.data(), (int64_t)words.size(), pairs.data(), num_couples, &b);
    if (r != TSU_OK) return r;
    const double h1 = now_ms();
    r = tsu_align_plain(ctx, b, wf_length, scores, t);
    if (t) t->h2d_ms = h1 - h0;
    tsu_batch_free(b);
    return r;
}

// ---------------------------------------------------------------------------
// batched form: buckets, persistent queue launches, overflow widening
// ---------------------------------------------------------------------------
namespace {

struct Bucket {
    int32_t bound;                // max(p, t) of the bucket
    std::vector<int32_t> idx;     // pair indices, longest first
};

// Bucket bounds on max(p, t).
// (16383: the longest sequence the doubled-offset 16-chunk kernel takes)
const int32_t kBucketBounds[] = {128, 256, 512, 1024, 2048, 4096, 8192, 16383, 32000, 1 << 26};

}  // namespace

void tsu_host::order_longest_first(std::vector<int32_t>& idx, const tsu_pair* pairs) {
    order_desc(idx, [&](int32_t i) -> int64_t { return (int64_t)pairs[(size_t)i].p_len + pairs[(size_t)i].t_len; });
}

int tsu_host::bucket_count() { return (int)(sizeof(kBucketBounds) / sizeof(kBucketBounds[0])); }
int32_t tsu_host::bucket_bound(int i) { return kBucketBounds[i]; }
int tsu_host::bucket_of(int32_t len) {
    int bi = 0;
    while (bi < bucket_count() - 1 && len > kBucketBounds[bi]) ++bi;
    return bi;
}

// Initial capacity for a bucket: enough for about 15% divergence at the
// bucket's length (score ~ 0.15 * L * max(x, o+e)), clamped to what the
// lengths allow and to the bucket's length itself (pairs up to 15%
// divergence need at most ~0.85 L diagonals; a window no wider than the
// length keeps more pairs resident per CU).  With the register kernels
// available (4/6/2) it is also capped at their widest window (16384
// diagonals, 8192 for sequences over 32000 bases): most pairs fit, and the
// few that overflow are re-run wider, which costs less than running every
// pair at the widest capacity.
int32_t tsu_host::initial_capacity(const tsu_ctx* ctx, int32_t bound) {
    const double per_base = 0.15 * (double)std::max(ctx->xs, ctx->os + ctx->es);
    const int64_t s_est = (int64_t)std::ceil(per_base * bound) + ctx->os + ctx->es;
    int32_t W = capacity_for_score(ctx, s_est);
    W = std::min<int32_t>(W, capacity_full(bound, bound));
    W = std::min<int32_t>(W, std::max<int32_t>(bound, 64));
    if (fixed_242(ctx)) W = std::min<int32_t>(W, bound <= 32000 ? 16 * 1024 : 16 * 512);
    return std::max<int32_t>(W, 16);
}

Shape tsu_host::batched_shape(const tsu_ctx* ctx, int32_t W, int32_t max_len, int nw_hint) {
    if (ctx->two_piece) {  // the 2-piece kernels, fed from the queue
        Shape s2{1, max_len <= 32000, true, false, true};
        s2.two_piece = true;
        s2.reg_c = tsu_host::reg2p_chunks(ctx, W, max_len);
        s2.reg_e = 10 * ctx->es + ctx->e2s;
        return s2;
    }
    Shape sh{nw_hint > 0 ? nw_hint : default_waves(W), max_len <= 32000, true, false, true};
    sh.fixed_242 = tsu_host::fixed_242(ctx);
    // a tuned hint above 1 asks for the LDS kernel with that many waves
    if (nw_hint <= 1) {
        use_register_kernel(ctx, W, max_len, sh);
        if (sh.reg_c) return sh;
    }
    tsu_host::place_sequences(ctx, sh, max_len);
    KArgs a = base_args(ctx);
    a.W = W;
    a.seq_words = seq_words_for(ctx, max_len);
    if (tsu::lds_bytes(sh, a) > ctx->max_lds) {
        sh.lds_ring = false;
        if (nw_hint <= 0) sh.nw = 16;
    }
    return sh;
}

int32_t tsu_host::tuned_capacity(const tsu_ctx* ctx, int32_t bound) {
    for (const tsu_tune_entry& te : ctx->tune)
        if (te.bucket_bound == bound) return te.capacity;
    return 0;
}

int tsu_host::tuned_waves(const tsu_ctx* ctx, int32_t bound) {
    for (const tsu_tune_entry& te : ctx->tune)
        if (te.bucket_bound == bound) return te.waves_per_pair;
    return 0;
}

using tsu_host::batched_shape;
using tsu_host::initial_capacity;

static const tsu_tune_entry* tune_lookup_in(const std::vector<tsu_tune_entry>& table, int32_t bound) {
    for (const tsu_tune_entry& te : table)
        if (te.bucket_bound == bound) return &te;
    return nullptr;
}

static const tsu_tune_entry* tune_lookup(const tsu_ctx* ctx, int32_t bound) { return tune_lookup_in(ctx->tune, bound); }

// Probe of the batched form: pairs whose bucket would run on a multi-wave
// window are first aligned by the single-wave register kernel (512
// diagonals) up to kProbeCap scaled scores.  Those that finish are done; for
// the others, how far that score took them along both sequences (h + v)
// predicts the final score, s_end ~ kProbeCap (p + t) / (h + v), and so the
// diagonal range they need, about 2 s_end.  Each then runs in the narrowest
// window predicted to hold it with kProbeMargin to spare, instead of its
// bucket's widest: a low-error pair of 10 kbp runs in a window of 1,024
// diagonals (several pairs per CU) rather than alone on a CU in 16,384.  The
// prediction only places a pair; a pair that overflows is re-run wider, so
// scores do not depend on it.  (On 400 pairs of 4-20 kbp at 1-15%, the ratio
// of the true to the predicted score had median 1.01, 95th percentile 1.16,
// maximum 1.29.)  A margin of 1.05-1.1 measured fastest on 500,000 pairs of
// 100-20,000 bp at 1-15%: about 1.5% below 1.2 (more pairs in the narrower
// windows outweigh the extra re-runs), 1.0 and 1.3 slower.  The served form
// keeps 1.2: there a re-run waits for its flight to end and adds to its
// request's latency (the two margins measured alike there, within noise).

// Capacity a probed pair gets: the predicted range with the margin, rounded
// to the window of the kernel that will take it (64..512 one wave, 1024 two
// waves, 4096 / 8192 / 16384 the 16-chunk kernel), so pairs fall into a few
// launch groups.  Beyond 16,384 diagonals only the LDS kernel with its ring
// in device memory remains, several times slower per diagonal: a pair goes
// there directly only when even its unmargined prediction is well beyond
// 16,384, and otherwise tries the 16-chunk kernel's widest window first.
int32_t tsu_host::probe_capacity(const tsu_ctx* ctx, const tsu_pair& pr, int32_t progress, double margin) {
    const double s_end = (double)kProbeCap * (double)(pr.p_len + pr.t_len) / (double)std::max(progress, 1);
    const double need = std::min(2.0 * (s_end - ctx->os) / ctx->es + 3.0, 1e9);
    const int64_t w = (int64_t)std::ceil(margin * need);
    int64_t c;
    if (w <= 512) {
        for (c = 64; c < w; c *= 2) {}
    } else if (w <= 1024) {
        c = 1024;
    } else if (w <= 16384 || need <= 1.25 * 16384) {
        for (c = 4096; c < std::min<int64_t>(w, 16384); c *= 2) {}
    } else {
        c = (w + 16383) / 16384 * 16384;
    }
    // (a launch clamps it to what its longest pair can use)
    return (int32_t)std::min<int64_t>(c, 1 << 26);
}

// The probe launch over `ids` (reordered longest first; d_order holds at
// least ids.size() entries).  On return sc[i] holds the score of each pair
// it finished, kProbeStopped for a pair stopped at kProbeCap (prog[i] then
// holds how far along both sequences it got) or TSU_EKOVERFLOW where the
// 512-diagonal window overflowed; sc and prog hold b->n entries.
static int run_probe(tsu_ctx* ctx, tsu_batch* b, std::vector<int32_t>& ids, int32_t* d_order, std::vector<int32_t>& sc,
                     std::vector<int32_t>& prog, double* kms) {
    tsu_host::order_longest_first(ids, b->pairs.data());
    int32_t max_len = 0;
    for (int32_t i : ids) max_len = std::max(max_len, std::max(b->pairs[(size_t)i].p_len, b->pairs[(size_t)i].t_len));
    DevBuf<int32_t> d_prog;
    HIPCHK(d_prog.alloc((size_t)std::max(b->n, 1)), "hipMalloc (probe)");
    KArgs pa = base_args(ctx);
    pa.probe_cap = tsu_host::kProbeCap;
    pa.probe_out = d_prog.p;
    Shape ps{1, max_len <= 32000, true, false, true};
    ps.fixed_242 = true;
    ps.reg_c = 8;
    ps.probe = true;
    HIPCHK(hipMemcpyAsync(d_order, ids.data(), ids.size() * sizeof(int32_t), hipMemcpyHostToDevice, ctx->stream),
           "hipMemcpyAsync (order)");
    int r = run_shape(ctx, io_of(b), ps, 512, d_order, (int32_t)ids.size(), max_len, 0, &pa, kms);
    if (r != TSU_OK) return r;
    sc.assign((size_t)b->n, 0);
    prog.assign((size_t)b->n, 0);
    HIPCHK(hipMemcpy(sc.data(), b->d_scores, (size_t)b->n * sizeof(int32_t), hipMemcpyDeviceToHost), "hipMemcpy D2H (scores)");
    HIPCHK(hipMemcpy(prog.data(), d_prog.p, (size_t)b->n * sizeof(int32_t), hipMemcpyDeviceToHost), "hipMemcpy D2H (probe)");
    return TSU_OK;
}

// Aligns the pairs listed in `sel` (pair indices) with persistent launches
// per group of pairs (a length bucket, or a length bucket and predicted
// window after the probe), widening the capacity of pairs that overflow.
// Scores are left in b->d_scores and copied to `scores` (full batch order)
// for the selected pairs.
static int batched_core(tsu_ctx* ctx, tsu_batch* b, const std::vector<int32_t>& sel, int32_t* scores,
                        tsu_timing* t, bool use_tune) {
    const int nb = (int)(sizeof(kBucketBounds) / sizeof(kBucketBounds[0]));
    std::vector<Bucket> buckets((size_t)nb);
    for (int i = 0; i < nb; ++i) buckets[(size_t)i].bound = kBucketBounds[i];
    for (int32_t i : sel) {
        const tsu_pair& pr = b->pairs[(size_t)i];
        const int32_t L = std::max(pr.p_len, pr.t_len);
        int bi = 0;
        while (bi < nb - 1 && L > kBucketBounds[bi]) ++bi;
        buckets[(size_t)bi].idx.push_back(i);
    }
    DevBuf<int32_t> d_order;
    HIPCHK(d_order.alloc(std::max<size_t>(sel.size(), 1)), "hipMalloc (order)");
    double kms_total = 0;
    int launches = 0, retried = 0;
    std::vector<int32_t> tmp((size_t)b->n);

    // One launch group: pairs of one length bucket at one first capacity.
    struct Group {
        int32_t W;        // first capacity
        int32_t max_len;
        int nw_hint;      // tuned waves per pair (0: default shape)
        std::vector<int32_t> idx;
        double cost = 0;  // estimated wave-steps: sum over pairs of score x waves
        int stream = 0;
    };
    // waves a pair takes at capacity W (windows of 1024 diagonals per wave
    // above one wave's 512), as a weight of its estimated score
    auto waves_of = [](int32_t W) { return std::max(0.25, (double)W / 1024.0); };
    std::vector<Group> groups;
    auto max_len_of = [&](const std::vector<int32_t>& ids) {
        int32_t m = 0;
        for (int32_t i : ids) m = std::max(m, std::max(b->pairs[(size_t)i].p_len, b->pairs[(size_t)i].t_len));
        return m;
    };
    // buckets whose pairs are probed: the register kernels take them (2/3/1
    // penalties), their first window would hold more than 1,024 diagonals and
    // no capacity was tuned for them (a tuned capacity is used as given)
    std::vector<int32_t> probe_ids;
    std::vector<int> probed_bucket((size_t)nb, 0);
    for (int bi = nb - 1; bi >= 0; --bi) {
        Bucket& bk = buckets[(size_t)bi];
        if (bk.idx.empty()) continue;
        const int32_t max_len = max_len_of(bk.idx);
        const tsu_tune_entry* te = use_tune ? tune_lookup(ctx, bk.bound) : nullptr;
        int32_t W = (te && te->capacity > 0) ? te->capacity : initial_capacity(ctx, std::min(bk.bound, max_len));
        W = std::min<int32_t>(W, tsu_host::capacity_max(ctx, max_len));
        const int nw_hint = te ? te->waves_per_pair : 0;
        if (tsu_host::probe_bucket(ctx, W, te ? te->capacity : 0, nw_hint, max_len)) {
            probed_bucket[(size_t)bi] = 1;
            probe_ids.insert(probe_ids.end(), bk.idx.begin(), bk.idx.end());
            continue;
        }
        Group g{W, max_len, nw_hint, bk.idx};
        // scores unknown: about a quarter of the length (a few % divergence)
        for (int32_t i : g.idx) g.cost += 0.25 * (b->pairs[(size_t)i].p_len + b->pairs[(size_t)i].t_len) / 2 * waves_of(W);
        groups.push_back(std::move(g));
    }
    if (!probe_ids.empty()) {
        double kms = 0;
        std::vector<int32_t> prog;
        int r = run_probe(ctx, b, probe_ids, d_order.p, tmp, prog, &kms);
        if (r != TSU_OK) return r;
        kms_total += kms;
        ++launches;
        // the unfinished pairs, grouped by bucket and predicted capacity
        std::map<std::pair<int, int32_t>, std::vector<int32_t>> by;
        std::map<std::pair<int, int32_t>, double> by_cost;
        for (int bi = 0; bi < nb; ++bi) {
            if (!probed_bucket[(size_t)bi]) continue;
            for (int32_t i : buckets[(size_t)bi].idx) {
                const int32_t v = tmp[(size_t)i];
                if (v >= 0) { scores[i] = v; continue; }
                if (v != tsu::kProbeStopped) {
                    // the probe window overflowed: the bucket's own capacity
                    const int32_t ml = std::max(b->pairs[(size_t)i].p_len, b->pairs[(size_t)i].t_len);
                    const std::pair<int, int32_t> key{bi, std::min<int32_t>(
                        initial_capacity(ctx, std::min(buckets[(size_t)bi].bound, ml)), tsu_host::capacity_max(ctx, ml))};
                    by[key].push_back(i);
                    by_cost[key] += 0.25 * ml * waves_of(key.second);
                    continue;
                }
                const tsu_pair& pr = b->pairs[(size_t)i];
                const std::pair<int, int32_t> key{bi, tsu_host::probe_capacity(ctx, pr, prog[(size_t)i])};
                by[key].push_back(i);
                by_cost[key] += (double)tsu_host::kProbeCap * (pr.p_len + pr.t_len) / std::max(prog[(size_t)i], 1) *
                                waves_of(key.second);
            }
        }
        for (auto& kv : by) {
            Group g{kv.first.second, max_len_of(kv.second), 0, std::move(kv.second)};
            g.cost = by_cost[kv.first];
            groups.push_back(std::move(g));
        }
    }
    // widest capacity first (longest pairs first among equals): the widest
    // launches hold the longest single pairs, whose tails overlap nothing,
    // so they start first; each group goes to the stream with the least
    // estimated work so far (launches of one stream run one after another)
    std::stable_sort(groups.begin(), groups.end(), [](const Group& x, const Group& y) {
        return x.W != y.W ? x.W > y.W : x.max_len > y.max_len;
    });
    {
        double load[3] = {0, 0, 0};
        for (Group& g : groups) {
            int best = 0;
            for (int k = 1; k < 3; ++k)
                if (load[k] < load[best]) best = k;
            g.stream = best;
            load[best] += g.cost;
        }
    }
    // every group ordered longest first, all ordered up front (one thread per
    // group), so no sort sits between two launches
#pragma omp parallel for schedule(dynamic, 1)
    for (int gi = 0; gi < (int)groups.size(); ++gi) tsu_host::order_longest_first(groups[(size_t)gi].idx, b->pairs.data());
    // Every group's first launch is issued at once, over three streams: a
    // persistent launch holds the CUs until its queue drains, and
    // the next launches' workgroups take the CUs it frees (or the room its
    // workgroups leave on a CU), so the long launches' tails overlap the
    // other groups' work instead of idling the chip between launches.  When
    // a launch ends, its pairs that overflowed are re-launched at twice the
    // capacity on the same stream while the others still run.
    hipStream_t streams[3] = {ctx->stream, ctx->stream2, ctx->stream3};
    struct Run {
        std::vector<int32_t> ids;
        int32_t W, max_len;
        int nw_hint;
        int32_t* order;      // this group's region of d_order
        int32_t* counter;    // this group's work counter
        hipStream_t stream;
        hipEvent_t e0 = nullptr, e1 = nullptr;
        DevBuf<uint8_t> ring;
        bool done = false;
        ~Run() { if (e0) (void)hipEventDestroy(e0); if (e1) (void)hipEventDestroy(e1); }
    };
    std::vector<std::unique_ptr<Run>> runs;
    DevBuf<int32_t> d_counters;
    HIPCHK(d_counters.alloc(std::max<size_t>(groups.size(), 1)), "hipMalloc (counters)");
    // on every return the streams are drained before the launches' buffers are freed
    struct Drain {
        hipStream_t* s;
        ~Drain() { for (int i = 0; i < 3; ++i) (void)hipStreamSynchronize(s[i]); }
    } drain{streams};
    std::vector<int32_t> all_ids;
    all_ids.reserve(sel.size());
    for (const Group& g : groups) all_ids.insert(all_ids.end(), g.idx.begin(), g.idx.end());
    if (!all_ids.empty())
        HIPCHK(hipMemcpy(d_order.p, all_ids.data(), all_ids.size() * sizeof(int32_t), hipMemcpyHostToDevice), "hipMemcpy (order)");
    auto start_run = [&](std::unique_ptr<Run> rn) -> int {
        HIPCHK(hipEventCreate(&rn->e0), "hipEventCreate");
        HIPCHK(hipEventCreate(&rn->e1), "hipEventCreate");
        bool launched = false;
        const Shape sh = batched_shape(ctx, rn->W, rn->max_len, rn->nw_hint);
        int r = launch_shape(ctx, io_of(b), sh, rn->W, rn->order, (int32_t)rn->ids.size(), rn->max_len, 0, nullptr,
                             rn->stream, rn->counter, rn->e0, rn->e1, rn->ring, &launched);
        if (r != TSU_OK) return r;
        if (!launched) return fail(ctx, TSU_EINTERNAL, "empty launch group");
        ++launches;
        runs.push_back(std::move(rn));
        return TSU_OK;
    };
    size_t off = 0;
    for (size_t gi = 0; gi < groups.size(); ++gi) {
        Group& g = groups[gi];
        auto rn = std::make_unique<Run>();
        rn->max_len = g.max_len;
        rn->W = std::min<int32_t>(g.W, tsu_host::capacity_max(ctx, g.max_len));
        rn->nw_hint = g.nw_hint;
        rn->order = d_order.p + off;
        rn->counter = d_counters.p + gi;
        rn->stream = streams[g.stream];
        off += g.idx.size();
        rn->ids = std::move(g.idx);
        int r = start_run(std::move(rn));
        if (r != TSU_OK) return r;
    }
    for (size_t open = runs.size(); open > 0;) {
        bool progressed = false;
        for (size_t ri = 0; ri < runs.size(); ++ri) {
            Run* rn = runs[ri].get();
            if (rn->done) continue;
            const hipError_t q = hipEventQuery(rn->e1);
            if (q == hipErrorNotReady) continue;
            if (q != hipSuccess) return hip_fail(ctx, q, "hipEventQuery");
            rn->done = true;
            --open;
            progressed = true;
            float ms = 0.f;
            HIPCHK(hipEventElapsedTime(&ms, rn->e0, rn->e1), "hipEventElapsedTime");
            kms_total += ms;
            // scores live at their pair index
            HIPCHK(hipMemcpyAsync(tmp.data(), b->d_scores, (size_t)b->n * sizeof(int32_t), hipMemcpyDeviceToHost,
                                  ctx->stream_io), "hipMemcpyAsync D2H (scores)");
            HIPCHK(hipStreamSynchronize(ctx->stream_io), "hipStreamSynchronize");
            std::vector<int32_t> again;
            for (int32_t i : rn->ids) {
                scores[i] = tmp[(size_t)i];
                if (tmp[(size_t)i] == TSU_EKOVERFLOW) again.push_back(i);
            }
            if (again.empty()) continue;
            if (rn->W >= tsu_host::capacity_max(ctx, rn->max_len)) return overflow_at(ctx, rn->W, rn->max_len);
            retried += (int)again.size();
            auto nx = std::make_unique<Run>();
            nx->W = (int32_t)std::min<int64_t>(2ll * rn->W, tsu_host::capacity_max(ctx, rn->max_len));
            nx->max_len = rn->max_len;
            nx->nw_hint = 0;  // wider re-runs take the default shape of their capacity
            nx->order = rn->order;      // the finished launch's region and counter
            nx->counter = rn->counter;
            nx->stream = rn->stream;
            nx->ids = std::move(again);
            HIPCHK(hipMemcpyAsync(nx->order, nx->ids.data(), nx->ids.size() * sizeof(int32_t), hipMemcpyHostToDevice,
                                  nx->stream), "hipMemcpyAsync (order)");
            int r = start_run(std::move(nx));
            if (r != TSU_OK) return r;
            ++open;
        }
        if (!progressed) std::this_thread::sleep_for(std::chrono::microseconds(50));
    }
    if (t) { t->kernel_ms += kms_total; t->launches += launches; t->retried += retried; }
    return TSU_OK;
}

extern "C" int tsu_align_batched(tsu_ctx* ctx, tsu_batch* b, int32_t* scores, tsu_timing* t) {
    if (!ctx || !b || !scores || b->ctx != ctx) return fail(ctx, TSU_EINVAL, "tsu_align_batched: bad arguments");
    HIPCHK(hipSetDevice(ctx->device), "hipSetDevice");
    if (t) std::memset(t, 0, sizeof(*t));
    std::vector<int32_t> sel((size_t)b->n);
    std::iota(sel.begin(), sel.end(), 0);
    const double w0 = now_ms();
    int r = batched_core(ctx, b, sel, scores, t, true);
    const double w1 = now_ms();
    if (t) { t->wall_ms = w1 - w0; t->pairs = b->n; t->cells = b->cells; }
    if (r != TSU_OK) return r;
    if (min_score(scores, b->n) >= 0) return TSU_OK;
    for (int32_t i = 0; i < b->n; ++i)
        if (scores[i] < 0) return fail(ctx, scores[i], "pair " + std::to_string(i) + ": " + tsu_strerror(scores[i]));
    return TSU_OK;
}

// Batched form without device allocation: the caller's device arrays in and
// out, the work counters and queue order in the device's scratch array
// (tsu::scratch_address), every launch on the context's stream.  Buckets and
// capacities as tsu_align_batched; overflowing pairs are re-run wider, one
// launch after another on the same stream; a pair whose capacity would need
// a ring in device memory is reported as overflowing.
extern "C" int tsu_align_batched_noalloc(tsu_ctx* ctx, const uint32_t* d_words, int64_t n_words,
                                         const tsu_pair* d_pairs, const tsu_pair* pairs, int32_t n_pairs,
                                         int32_t* d_scores, int32_t* scores, tsu_timing* t) {
    if (!ctx) return TSU_EINVAL;
    if (n_pairs < 0 || n_words < 0 ||
        (n_pairs > 0 && (!d_words || !d_pairs || !pairs || !d_scores || !scores || n_words == 0)))
        return fail(ctx, TSU_EINVAL, "tsu_align_batched_noalloc: bad arguments");
    if (n_pairs > tsu::kNoallocMaxPairs)
        return fail(ctx, TSU_EINVAL, "tsu_align_batched_noalloc: " + std::to_string(n_pairs) + " pairs; at most " +
                                         std::to_string(tsu::kNoallocMaxPairs) +
                                         " fit its device scratch (one queue entry per pair)");
    if (!ctx->scratch) return fail(ctx, TSU_EINTERNAL, "tsu_align_batched_noalloc: no scratch address");
    int32_t max_all = 0;
    double cells = 0;
    for (int32_t i = 0; i < n_pairs; ++i) {
        const tsu_pair& p = pairs[i];
        if (p.p_len < 0 || p.t_len < 0 || p.p_off < 0 || p.t_off < 0 ||
            p.p_off + tsu_packed_words(p.p_len) > n_words || p.t_off + tsu_packed_words(p.t_len) > n_words)
            return fail(ctx, TSU_EINVAL, "pair " + std::to_string(i) + " lies outside the packed arena");
        max_all = std::max(max_all, std::max(p.p_len, p.t_len));
        cells += (double)p.p_len * (double)p.t_len;
    }
    if (max_all > tsu_host::seq_len_limit(ctx))
        return fail(ctx, TSU_EINVAL, "a sequence is longer than " + std::to_string(tsu_host::seq_len_limit(ctx)) + " bases");
    HIPCHK(hipSetDevice(ctx->device), "hipSetDevice");
    if (t) std::memset(t, 0, sizeof(*t));
    const double w0 = now_ms();
    const DevIO io{d_words, reinterpret_cast<const KPair*>(d_pairs), d_scores};
    int32_t* d_queue = static_cast<int32_t*>(ctx->scratch);
    int32_t* d_order = reinterpret_cast<int32_t*>(static_cast<uint8_t*>(ctx->scratch) + tsu::kNoallocCounterBytes);
    const int nb = tsu_host::bucket_count();
    std::vector<std::vector<int32_t>> bk((size_t)nb);
    for (int32_t i = 0; i < n_pairs; ++i)
        bk[(size_t)tsu_host::bucket_of(std::max(pairs[i].p_len, pairs[i].t_len))].push_back(i);
    // one region of the order array per bucket, longest pairs first
    std::vector<int32_t> order;
    order.reserve((size_t)n_pairs);
    std::vector<size_t> base((size_t)nb, 0);
    for (int bi = 0; bi < nb; ++bi) {
        tsu_host::order_longest_first(bk[(size_t)bi], pairs);
        base[(size_t)bi] = order.size();
        order.insert(order.end(), bk[(size_t)bi].begin(), bk[(size_t)bi].end());
    }
    if (n_pairs > 0)
        HIPCHK(hipMemcpyAsync(d_order, order.data(), order.size() * sizeof(int32_t), hipMemcpyHostToDevice, ctx->stream),
               "hipMemcpyAsync (order)");
    double kms_total = 0;
    int launches = 0, retried = 0;
    int32_t no_room = -1;  // a pair whose capacity needs a ring in device memory
    DevBuf<uint8_t> unused_ring;  // stays empty: every launch here keeps its ring in LDS
    for (int bi = nb - 1; bi >= 0; --bi) {
        std::vector<int32_t> todo = bk[(size_t)bi];
        if (todo.empty()) continue;
        int32_t max_len = 0;
        for (int32_t i : todo) max_len = std::max(max_len, std::max(pairs[i].p_len, pairs[i].t_len));
        const tsu_tune_entry* te = tune_lookup(ctx, tsu_host::bucket_bound(bi));
        int32_t W = (te && te->capacity > 0) ? te->capacity
                                             : initial_capacity(ctx, std::min(tsu_host::bucket_bound(bi), max_len));
        const int nw_hint = te ? te->waves_per_pair : 0;
        int32_t* region = d_order + base[(size_t)bi];
        for (int attempt = 0; !todo.empty(); ++attempt) {
            W = std::min<int32_t>(W, tsu_host::capacity_max(ctx, max_len));
            const Shape sh = batched_shape(ctx, W, max_len, attempt == 0 ? nw_hint : 0);
            if (!sh.lds_ring) {
                for (int32_t i : todo) scores[i] = TSU_EKOVERFLOW;
                no_room = todo[0];
                break;
            }
            if (attempt > 0)
                HIPCHK(hipMemcpyAsync(region, todo.data(), todo.size() * sizeof(int32_t), hipMemcpyHostToDevice,
                                      ctx->stream), "hipMemcpyAsync (order)");
            bool launched = false;
            int r = launch_shape(ctx, io, sh, W, region, (int32_t)todo.size(), max_len, 0, nullptr, ctx->stream, d_queue,
                                 ctx->ev0, ctx->ev1, unused_ring, &launched);
            if (r != TSU_OK) return r;
            if (launched) {
                HIPCHK(hipEventSynchronize(ctx->ev1), "hipEventSynchronize");
                float ms = 0.f;
                HIPCHK(hipEventElapsedTime(&ms, ctx->ev0, ctx->ev1), "hipEventElapsedTime");
                kms_total += ms;
                ++launches;
            }
            HIPCHK(hipMemcpy(scores, d_scores, (size_t)n_pairs * sizeof(int32_t), hipMemcpyDeviceToHost),
                   "hipMemcpy D2H (scores)");
            std::vector<int32_t> again;
            for (int32_t i : todo)
                if (scores[i] == TSU_EKOVERFLOW) again.push_back(i);
            if (!again.empty()) {
                if (W >= tsu_host::capacity_max(ctx, max_len)) return overflow_at(ctx, W, max_len);
                retried += (int)again.size();
                W = (int32_t)std::min<int64_t>(2ll * W, tsu_host::capacity_max(ctx, max_len));
            }
            todo.swap(again);
        }
    }
    const double w1 = now_ms();
    if (t) {
        t->kernel_ms = kms_total; t->wall_ms = w1 - w0; t->launches = launches; t->retried = retried;
        t->pairs = n_pairs; t->cells = cells;
    }
    if (no_room >= 0)
        return fail(ctx, TSU_EKOVERFLOW, "pair " + std::to_string(no_room) +
                                             ": its capacity needs a ring in device memory, which this entry point "
                                             "does not allocate (tsu_align_batched aligns it)");
    for (int32_t i = 0; i < n_pairs; ++i)
        if (scores[i] < 0) return fail(ctx, scores[i], "pair " + std::to_string(i) + ": " + tsu_strerror(scores[i]));
    return TSU_OK;
}

// Autotuner: for each populated bucket, time the candidate shapes (initial
// capacity, or waves per pair of the LDS kernel) on a sample of the bucket
// and keep the fastest.
extern "C" int tsu_autotune(tsu_ctx* ctx, tsu_batch* b, int32_t sample_pairs) {
    if (!ctx || !b || b->ctx != ctx) return fail(ctx, TSU_EINVAL, "tsu_autotune: bad arguments");
    HIPCHK(hipSetDevice(ctx->device), "hipSetDevice");
    if (sample_pairs <= 0) sample_pairs = 4096;
    if (ctx->two_piece) { ctx->tune.clear(); return TSU_OK; }  // one shape: nothing to tune
    const int nb = (int)(sizeof(kBucketBounds) / sizeof(kBucketBounds[0]));
    std::vector<std::vector<int32_t>> members((size_t)nb);
    for (int32_t i = 0; i < b->n; ++i) {
        const int32_t L = std::max(b->pairs[(size_t)i].p_len, b->pairs[(size_t)i].t_len);
        int bi = 0;
        while (bi < nb - 1 && L > kBucketBounds[bi]) ++bi;
        members[(size_t)bi].push_back(i);
    }
    std::vector<tsu_tune_entry> table;
    std::vector<int32_t> scratch((size_t)b->n);
    for (int bi = 0; bi < nb; ++bi) {
        std::vector<int32_t>& m = members[(size_t)bi];
        if (m.empty()) continue;
        // deterministic stride sample
        std::vector<int32_t> sample;
        const size_t step = std::max<size_t>(1, m.size() / (size_t)sample_pairs);
        for (size_t j = 0; j < m.size() && sample.size() < (size_t)sample_pairs; j += step) sample.push_back(m[j]);
        // Candidates: where the register kernels take the bucket's initial
        // capacity (4/6/2 penalties), that capacity and half of it (more
        // pairs per CU; pairs that overflow are re-run at the full one, and
        // that time is counted), and, where that capacity needs several
        // waves, capacity 0: each pair placed by the probe; otherwise the
        // LDS kernel with 1-16 waves.
        int32_t max_len = 0;
        for (int32_t i : sample) max_len = std::max(max_len, std::max(b->pairs[(size_t)i].p_len, b->pairs[(size_t)i].t_len));
        const int32_t W0 = initial_capacity(ctx, std::min(kBucketBounds[bi], max_len));
        std::vector<tsu_tune_entry> cands;
        if (batched_shape(ctx, W0, max_len, 1).reg_c > 0) {
            if (W0 > 1024) cands.push_back(tsu_tune_entry{1, 0, kBucketBounds[bi], 0});
            cands.push_back(tsu_tune_entry{1, W0, kBucketBounds[bi], 0});
            if (W0 / 2 >= 256) cands.push_back(tsu_tune_entry{1, W0 / 2, kBucketBounds[bi], 0});
        } else {
            for (int nw : {1, 2, 4, 8, 16}) cands.push_back(tsu_tune_entry{nw, 0, kBucketBounds[bi], 0});
        }
        double best = 1e300;
        const tsu_tune_entry* pick = nullptr;
        for (const tsu_tune_entry& probe : cands) {
            ctx->tune.clear();
            ctx->tune.push_back(probe);
            tsu_timing tt;
            std::memset(&tt, 0, sizeof(tt));
            // wall time: the launches of one call overlap, so the sum of
            // their durations overstates a shape that splits its pairs
            const double p0 = now_ms();
            int r = batched_core(ctx, b, sample, scratch.data(), &tt, true);
            tt.wall_ms = now_ms() - p0;
            // a shape that cannot run here (no room in LDS, no occupancy, no
            // device memory for its ring) is skipped; any other failure is a
            // device or internal error and ends the call with no table
            if (r == TSU_EINVAL || r == TSU_ENOMEM) { (void)hipGetLastError(); continue; }
            if (r != TSU_OK) { ctx->tune.clear(); return r; }
            if (tt.wall_ms < best) { best = tt.wall_ms; pick = &probe; }
        }
        if (pick) table.push_back(*pick);
    }
    ctx->tune = table;
    return TSU_OK;
}

extern "C" int tsu_tune_table(const tsu_ctx* ctx, tsu_tune_entry* entries, int32_t cap, int32_t* n) {
    if (!ctx || !n) return TSU_EINVAL;
    *n = (int32_t)ctx->tune.size();
    for (int32_t i = 0; i < *n && i < cap; ++i) entries[i] = ctx->tune[(size_t)i];
    return TSU_OK;
}

// Replaces the context's table (one saved from tsu_tune_table, or a window
// pinned per bucket).  Each entry names one bucket by its bound; waves_per_pair
// is 0 or 1 (the automatic shape) or 2-16 (the LDS kernel with that many
// waves per pair); capacity 0 lets the bucket be placed automatically.
extern "C" int tsu_set_tune_table(tsu_ctx* ctx, const tsu_tune_entry* entries, int32_t n) {
    if (!ctx || n < 0 || n > 64 || (n > 0 && !entries)) return fail(ctx, TSU_EINVAL, "tsu_set_tune_table: bad arguments");
    const int nb = (int)(sizeof(kBucketBounds) / sizeof(kBucketBounds[0]));
    std::vector<tsu_tune_entry> table;
    for (int32_t k = 0; k < n; ++k) {
        const tsu_tune_entry& e = entries[k];
        const bool bucket = std::find(kBucketBounds, kBucketBounds + nb, e.bucket_bound) != kBucketBounds + nb;
        const int w = e.waves_per_pair;
        if (!bucket || (w != 0 && w != 1 && w != 2 && w != 4 && w != 8 && w != 16) || e.capacity < 0 ||
            e.capacity > (1 << 26) || (w > 1 && e.capacity != 0) || tune_lookup_in(table, e.bucket_bound))
            return fail(ctx, TSU_EINVAL, "tsu_set_tune_table: entry " + std::to_string(k) +
                                             " (a bucket bound of the batched form, listed once; waves 0, 1, 2, 4, 8 "
                                             "or 16; capacity 0..2^26, 0 with waves above 1)");
        table.push_back(tsu_tune_entry{w, e.capacity, e.bucket_bound, 0});
    }
    ctx->tune = table;
    return TSU_OK;
}

// ---------------------------------------------------------------------------
// CIGAR form
// ---------------------------------------------------------------------------
// Device-to-host copy into pageable memory, queued behind the work on the
// context's stream: slices go through a pinned staging buffer of two halves
// kept by the context; while host threads move one slice on (touching the
// destination's fresh pages in parallel), the next slice's DMA fills the
// other half.  (A pageable hipMemcpy goes through the runtime's own staging
// in small pieces: the scores of a million pairs took about 1 ms that way.)
static int copy_to_host(tsu_ctx* ctx, void* dst, const void* src, size_t bytes, int32_t* min32, double* synced_ms) {
    constexpr size_t kSlice = 32u << 20;
    if (min32) *min32 = INT32_MAX;
    if (bytes == 0) return TSU_OK;
    const size_t want = 2 * kSlice;
    if (ctx->h_stage_bytes < want) {
        if (ctx->h_stage) (void)hipHostFree(ctx->h_stage);
        ctx->h_stage = nullptr;
        ctx->h_stage_bytes = 0;
        HIPCHK(hipHostMalloc(&ctx->h_stage, want, hipHostMallocDefault), "hipHostMalloc (staging)");
        ctx->h_stage_bytes = want;
    }
    uint8_t* half[2] = {static_cast<uint8_t*>(ctx->h_stage), static_cast<uint8_t*>(ctx->h_stage) + kSlice};
    const uint8_t* from = static_cast<const uint8_t*>(src);
    const size_t slices = (bytes + kSlice - 1) / kSlice;
    HIPCHK(hipMemcpyAsync(half[0], from, std::min(kSlice, bytes), hipMemcpyDeviceToHost, ctx->stream), "hipMemcpyAsync D2H");
    for (size_t k = 0; k < slices; ++k) {
        HIPCHK(hipStreamSynchronize(ctx->stream), "hipStreamSynchronize");  // slice k is in half[k & 1]
        if (k == 0 && synced_ms) *synced_ms = now_ms();
        if (k + 1 < slices) {
            const size_t off1 = (k + 1) * kSlice;
            HIPCHK(hipMemcpyAsync(half[(k + 1) & 1], from + off1, std::min(kSlice, bytes - off1), hipMemcpyDeviceToHost,
                                  ctx->stream), "hipMemcpyAsync D2H");
        }
        const size_t off = k * kSlice;
        const size_t n = std::min(kSlice, bytes - off);
        const uint8_t* st = half[k & 1];
        // each thread also takes the lowest int32 of its part while that is
        // in its own cache (slices and parts are multiples of 64 bytes)
        auto part_min = [&](size_t b0, size_t b1) -> int32_t {
            return min_score(reinterpret_cast<const int32_t*>(static_cast<uint8_t*>(dst) + off + b0),
                             (int32_t)((b1 - b0) / 4));
        };
        if (n <= (1u << 20)) {  // small: one thread
            std::memcpy(static_cast<uint8_t*>(dst) + off, st, n);
            if (min32) *min32 = std::min(*min32, part_min(0, n));
            continue;
        }
        const size_t part = (n + 16 * 64 - 1) / (16 * 64) * 64;
        int32_t m = INT32_MAX;
#pragma omp parallel for schedule(static) reduction(min : m)
        for (int64_t i = 0; i < 16; ++i) {
            const size_t b0 = std::min(n, (size_t)i * part), b1 = std::min(n, b0 + part);
            if (b1 > b0) {
                std::memcpy(static_cast<uint8_t*>(dst) + off + b0, st + b0, b1 - b0);
                if (min32) m = std::min(m, part_min(b0, b1));
            }
        }
        if (min32) *min32 = std::min(*min32, m);
    }
    return TSU_OK;
}

namespace {

// Chunks of the single-wave recording kernel for capacity W (0: none): the
// register windows of up to 512 diagonals, and 1024 (16 chunks), which only
// the recording instance takes.
int bt_reg_chunks(const tsu_ctx* ctx, int32_t W) {
    if (int c = tsu_host::reg_chunks(ctx, W)) return c;
    return tsu_host::fixed_242(ctx) && W <= 1024 ? 16 : 0;
}

// Kernel family of a CIGAR launch at capacity W for a pair of length len
// and scaled score s: 1 the single-wave register kernel, 2 + waves the
// 16-chunk kernel, 0 the LDS kernel.
int bt_family(const tsu_ctx* ctx, int32_t W, int32_t len, int64_t s) {
    if (bt_reg_chunks(ctx, W) > 0 && len <= 16000 && s <= 16000) return 1;
    if (int nw = tsu_host::ml_waves(ctx, W, len)) return 2 + nw;
    return 0;
}

// Highest scaled score a pair of lengths (p, t) can have: min(p, t)
// mismatches and one gap of |p - t|.
int64_t score_bound(const tsu_ctx* ctx, int32_t p, int32_t t) {
    const int64_t d = std::abs((int64_t)p - t);
    return (int64_t)ctx->xs * std::min(p, t) + (d ? ctx->os + ctx->es * d : 0);
}

// Highest scaled score whose diagonal range fits a window of W diagonals
// (-1 if none does).
int64_t score_for_capacity(const tsu_ctx* ctx, int32_t W) {
    int64_t lo = -1, hi = 1ll << 30;
    while (hi - lo > 1) {
        const int64_t mid = lo + (hi - lo) / 2;
        if (capacity_for_score(ctx, mid) <= W) lo = mid; else hi = mid;
    }
    return lo;
}

// One length bucket of the record pass, at its capacity.
struct RecordPlan {
    std::vector<int32_t> ids;
    Shape sh;
    KArgs a;
    int32_t W, max_len, step_cap, win;
    int64_t recs, per_wg;
    bool at_window;  // the step bound is the window's: a pair past it needs a wider one
};

// The LDS-kernel launch for pairs wider than every register window.
struct WidePlan {
    Shape sh;
    KArgs a;
    int32_t W, max_len, step_cap;
    int64_t recs, per_wg, runs;
};

// That launch on the second stream, beside the record launches.
struct WideAsync {
    DevBuf<int32_t> d_ids;
    DevBuf<uint8_t> ring;
    hipEvent_t e0 = nullptr, e1 = nullptr;
    bool launched = false;
    hipStream_t s = nullptr;
    ~WideAsync() {
        if (s) (void)hipStreamSynchronize(s);
        if (e0) (void)hipEventDestroy(e0);
        if (e1) (void)hipEventDestroy(e1);
    }
};

// One tsu_align_cigar call.  Three paths, each over the pairs the ones
// before it left: record launches without a score pass (register kernels,
// per length bucket at the batched form's capacity), one LDS-kernel launch
// over the whole diagonal range for pairs wider than every register window,
// and scores first (batched form) for the rest, which size the backtrace
// state exactly.  The walks append their runs to one pool on the device.
class CigarCall {
public:
    CigarCall(tsu_ctx* c, tsu_batch* bt, int32_t* sc, tsu_timing* tm) : ctx(c), b(bt), scores(sc), t(tm) {}

    int run(tsu_cigars** out) {
        int r = setup();
        if (ctx->two_piece) {
            if (r == TSU_OK) r = two_piece_pass();
        } else {
            if (r == TSU_OK && ctx->cigar_mode != TSU_CIGAR_CHECKPOINTS && tsu_host::fixed_242(ctx)) r = record_pass();
            if (r == TSU_OK) r = wide_pass();
            if (r == TSU_OK) r = score_first_pass();
        }
        if (r == TSU_OK) r = collect(out);
        return r;
    }

private:
    int setup();
    int64_t bt_grid(const Shape& sh, const KArgs& a, int64_t n, int64_t per_wg) const;
    int launch_bt(const Shape& sh, KArgs& a, int32_t W, const std::vector<int32_t>& ids, int32_t max_len, int64_t recs,
                  int32_t step_cap, int64_t per_wg);
    int grow_runs(int64_t more);
    WidePlan plan_wide(const std::vector<int32_t>& ids) const;
    int settle_wide(const std::vector<int32_t>& ids);
    int launch_wide_async(WidePlan& wp, const std::vector<int32_t>& ids);
    int plan_records(std::vector<RecordPlan>& plans, int64_t& runs_need);
    void plan_group(int32_t W, bool placed, std::vector<int32_t>& ids, std::vector<RecordPlan>& plans, int64_t& runs_need);
    int settle_plan(const RecordPlan& pl);
    int record_pass();
    int wide_pass();
    int score_first_pass();
    int two_piece_pass();
    int collect(tsu_cigars** out);

    tsu_ctx* ctx;
    tsu_batch* b;
    int32_t* scores;
    tsu_timing* t;
    double w0 = 0;
    std::unique_ptr<tsu_cigars> cg;  // handed to the caller at the end
    int32_t unit = 1;                // a walk reserves 2 * (s / unit) + 2 run words for its scaled score s
    DevBuf<uint32_t> d_runs;
    DevBuf<unsigned long long> d_used;
    DevBuf<int64_t> d_run_off;
    DevBuf<int32_t> d_run_n;
    DevBuf<int32_t> d_order;
    int64_t runs_cap = 0;  // words in d_runs
    double kms_total = 0;
    int launches = 0;
    int64_t budget = 0;
    std::vector<int32_t> sc2;      // scores the backtrace launches returned
    std::vector<uint8_t> done;
    std::vector<uint8_t> scored;   // scored by the probe: records from the score-first path
    std::vector<int32_t> wide;     // pairs that overflowed the widest register window
    std::vector<int32_t> order;    // pairs of the score-first path
    WideAsync wide_async;          // declared after the buffers its launch uses: synchronised first
};

int CigarCall::setup() {
    HIPCHK(hipSetDevice(ctx->device), "hipSetDevice");
    if (t) std::memset(t, 0, sizeof(*t));
    w0 = now_ms();
    cg.reset(new (std::nothrow) tsu_cigars());
    if (!cg) return fail(ctx, TSU_ENOMEM, "host allocation");
    cg->off.assign((size_t)b->n, 0);
    cg->n.assign((size_t)b->n, 0);
    unit = std::min(ctx->xs, ctx->os + ctx->es);
    HIPCHK(d_used.alloc(1), "hipMalloc");
    HIPCHK(d_run_off.alloc((size_t)std::max(b->n, 1)), "hipMalloc");
    HIPCHK(d_run_n.alloc((size_t)std::max(b->n, 1)), "hipMalloc");
    HIPCHK(d_order.alloc((size_t)std::max(b->n, 1)), "hipMalloc");
    HIPCHK(hipMemsetAsync(d_used.p, 0, sizeof(unsigned long long), ctx->stream), "hipMemsetAsync");
    size_t free_b = 0, total_b = 0;
    HIPCHK(hipMemGetInfo(&free_b, &total_b), "hipMemGetInfo");
    // the state this context already holds counts as free
    for (int i = 0; i < 7; ++i) free_b += ctx->bt_bytes[i];
    budget = ctx->cigar_budget > 0 ? ctx->cigar_budget : (int64_t)(0.5 * (double)free_b);
    sc2.assign((size_t)b->n, 0);
    done.assign((size_t)b->n, 0);
    scored.assign((size_t)b->n, 0);
    return TSU_OK;
}

// One backtrace launch of `ids` with `recs` record words and `step_cap`
// scores of state per workgroup (per_wg bytes in all).
// Workgroups of a backtrace launch of n pairs: every pair that fits on the
// CUs at once, within the budget (0: the shape has no occupancy).
int64_t CigarCall::bt_grid(const Shape& sh, const KArgs& a, int64_t n, int64_t per_wg) const {
    int per_cu = 0;
    if (tsu::occupancy(sh, tsu::lds_bytes(sh, a), &per_cu) != 0 || per_cu <= 0) return 0;
    const int64_t grid = std::min<int64_t>((int64_t)per_cu * ctx->cu_count, n);
    return std::max<int64_t>(1, std::min<int64_t>(grid, budget / std::max<int64_t>(per_wg, 1)));
}

int CigarCall::launch_bt(const Shape& sh, KArgs& a, int32_t W, const std::vector<int32_t>& ids, int32_t max_len,
                         int64_t recs, int32_t step_cap, int64_t per_wg) {
    const int64_t grid = bt_grid(sh, a, (int64_t)ids.size(), per_wg);
    if (grid <= 0) return fail(ctx, TSU_EINTERNAL, "backtrace shape has no occupancy");
    cg->state_bytes = std::max(cg->state_bytes, per_wg * grid);
    if (sh.ck) ++cg->checkpoint_launches; else ++cg->record_launches;
    cg->kernels |= tsu_host::kernel_bit(sh);
    void *rec = nullptr, *slo = nullptr, *soff = nullptr, *ckb = nullptr;
    HIPCHK(bt_reserve(ctx, 0, (size_t)(recs * grid) * 4, &rec), "hipMalloc (records)");
    HIPCHK(bt_reserve(ctx, 1, (size_t)((int64_t)step_cap * grid) * 4, &slo), "hipMalloc (steps)");
    HIPCHK(bt_reserve(ctx, 2, (size_t)((int64_t)step_cap * grid) * 8, &soff), "hipMalloc (steps)");
    if (sh.ck) {
        a.ck_stride = (int64_t)a.ck_cap * a.ck_bytes;
        HIPCHK(bt_reserve(ctx, 3, (size_t)(a.ck_stride * grid), &ckb), "hipMalloc (checkpoints)");
        a.ck_g = static_cast<uint8_t*>(ckb);
    }
    HIPCHK(hipMemcpyAsync(d_order.p, ids.data(), ids.size() * sizeof(int32_t), hipMemcpyHostToDevice, ctx->stream), "hipMemcpyAsync");
    a.rec_g = static_cast<uint32_t*>(rec); a.rec_cap = recs;
    a.step_lo_g = static_cast<int32_t*>(slo); a.step_off_g = static_cast<int64_t*>(soff); a.step_cap = step_cap;
    a.runs = d_runs.p; a.runs_cap = std::max<int64_t>(runs_cap, 1); a.runs_used = d_used.p;
    a.run_off = d_run_off.p; a.run_n = d_run_n.p;
    double kms = 0;
    int r = run_shape(ctx, io_of(b), sh, W, d_order.p, (int32_t)ids.size(), max_len, grid, &a, &kms);
    if (r != TSU_OK) return r;
    kms_total += kms;
    ++launches;
    return TSU_OK;
}

// Makes room in the run pool for `more` words after what was reserved so far.
int CigarCall::grow_runs(int64_t more) {
    unsigned long long used0 = 0;
    HIPCHK(hipMemcpy(&used0, d_used.p, sizeof(used0), hipMemcpyDeviceToHost), "hipMemcpy D2H");
    const int64_t need = (int64_t)used0 + more;
    if (need > runs_cap) {
        DevBuf<uint32_t> grown;
        HIPCHK(grown.alloc((size_t)need), "hipMalloc (runs)");
        if (used0) HIPCHK(hipMemcpy(grown.p, d_runs.p, (size_t)used0 * 4, hipMemcpyDeviceToDevice), "hipMemcpy D2D");
        std::swap(grown.p, d_runs.p);
        std::swap(grown.n, d_runs.n);
        runs_cap = need;
    }
    return TSU_OK;
}

// Pairs wider than every register window (they overflowed the 16,384
// diagonals of the 16-chunk kernel): one LDS-kernel launch over the whole
// diagonal range of their lengths, records sized for the highest score the
// lengths allow, when that fits the budget (a few long divergent pairs, each
// alone on a CU: a score pass first would double their time).
WidePlan CigarCall::plan_wide(const std::vector<int32_t>& ids) const {
    WidePlan wp{};
    int64_t s_len = 0;
    wp.max_len = 0;
    wp.runs = 0;
    for (int32_t i : ids) {
        const tsu_pair& pr = b->pairs[(size_t)i];
        wp.max_len = std::max(wp.max_len, std::max(pr.p_len, pr.t_len));
        const int64_t sb = score_bound(ctx, pr.p_len, pr.t_len);
        s_len = std::max(s_len, sb);
        wp.runs += 2 * (sb / unit) + 2;
    }
    wp.W = capacity_full(wp.max_len, wp.max_len);
    wp.step_cap = (int32_t)std::min<int64_t>(s_len + 2, 1 << 30);
    wp.recs = 64;
    for (int64_t s = 0; s <= s_len; ++s)
        wp.recs += std::min<int64_t>(wp.W, 2 * (std::max<int64_t>(0, s - ctx->os) / ctx->es + 2) + 1);
    wp.a = base_args(ctx);
    wp.a.W = wp.W;
    wp.a.seq_words = seq_words_for(ctx, wp.max_len);
    wp.sh = Shape{default_waves(wp.W), wp.max_len <= 32000, true, true, true};
    wp.sh.fixed_242 = tsu_host::fixed_242(ctx);
    tsu_host::place_sequences(ctx, wp.sh, wp.max_len);
    if (tsu::lds_bytes(wp.sh, wp.a) > ctx->max_lds) { wp.sh.lds_ring = false; wp.sh.nw = 16; }
    wp.per_wg = wp.recs * 4 + (int64_t)wp.step_cap * 12 +
                (wp.sh.lds_ring ? 0 : tsu::ring_bytes(wp.sh, wp.W, ctx->rm, ctx->ri));
    return wp;
}

int CigarCall::settle_wide(const std::vector<int32_t>& ids) {
    for (int32_t i : ids) {
        const int32_t v = sc2[(size_t)i];
        if (v >= 0) { scores[i] = v; done[(size_t)i] = 1; }
        else if (v != TSU_EKOVERFLOW && v != TSU_ECIGAR)
            return fail(ctx, v, "pair " + std::to_string(i) + ": " + tsu_strerror(v));
    }
    return TSU_OK;
}

// The wide launch on the second stream, beside the main stream's launches,
// with state of its own (context slots 4-6; the run pool was reserved for
// these pairs up front).  Leaves wide_async.launched false when that state
// cannot be allocated now: the launch then runs afterwards.
int CigarCall::launch_wide_async(WidePlan& wp, const std::vector<int32_t>& ids) {
    int per_cu = 0;
    if (tsu::occupancy(wp.sh, tsu::lds_bytes(wp.sh, wp.a), &per_cu) != 0 || per_cu <= 0) return TSU_OK;
    int64_t grid = std::min<int64_t>((int64_t)per_cu * ctx->cu_count, (int64_t)ids.size());
    grid = std::max<int64_t>(1, std::min<int64_t>(grid, budget / std::max<int64_t>(wp.per_wg, 1)));
    void *rec = nullptr, *slo = nullptr, *soff = nullptr;
    if (bt_reserve(ctx, 4, (size_t)(wp.recs * grid) * 4, &rec) != hipSuccess ||
        bt_reserve(ctx, 5, (size_t)((int64_t)wp.step_cap * grid) * 4, &slo) != hipSuccess ||
        bt_reserve(ctx, 6, (size_t)((int64_t)wp.step_cap * grid) * 8, &soff) != hipSuccess ||
        wide_async.d_ids.alloc(ids.size()) != hipSuccess) {
        (void)hipGetLastError();  // out of memory now: run it afterwards instead
        return TSU_OK;
    }
    HIPCHK(hipEventCreate(&wide_async.e0), "hipEventCreate");
    HIPCHK(hipEventCreate(&wide_async.e1), "hipEventCreate");
    HIPCHK(hipMemcpyAsync(wide_async.d_ids.p, ids.data(), ids.size() * sizeof(int32_t), hipMemcpyHostToDevice, ctx->stream2),
           "hipMemcpyAsync");
    KArgs a = wp.a;
    a.rec_g = static_cast<uint32_t*>(rec); a.rec_cap = wp.recs;
    a.step_lo_g = static_cast<int32_t*>(slo); a.step_off_g = static_cast<int64_t*>(soff); a.step_cap = wp.step_cap;
    a.runs = d_runs.p; a.runs_cap = std::max<int64_t>(runs_cap, 1); a.runs_used = d_used.p;
    a.run_off = d_run_off.p; a.run_n = d_run_n.p;
    wide_async.s = ctx->stream2;
    int r = launch_shape(ctx, io_of(b), wp.sh, wp.W, wide_async.d_ids.p, (int32_t)ids.size(), wp.max_len, grid, &a, ctx->stream2,
                         ctx->d_queue + 16, wide_async.e0, wide_async.e1, wide_async.ring, &wide_async.launched);
    if (r != TSU_OK) return r;
    cg->state_bytes = std::max(cg->state_bytes, wp.per_wg * grid);
    ++cg->record_launches;
    cg->kernels |= tsu_host::kernel_bit(wp.sh);
    ++launches;
    return TSU_OK;
}

// Record launches without a score pass, per length bucket of the batched
// form at its capacity, where a register kernel records: the state is sized
// for the highest score the window (or the lengths) allow.  A pair whose
// window overflows, or whose score passes that step bound, is left to the
// paths after.
//
// Buckets whose first window needs several waves are probed first, as in the
// batched form: a pair the probe finished has its score, and its records come
// from the score-first path's launches at its exact capacity; the others are
// grouped by bucket and predicted window, so a low-error pair of 10 kbp
// records in 4,096 diagonals (four waves, a quarter of the record state)
// rather than its bucket's 16,384.
int CigarCall::plan_records(std::vector<RecordPlan>& plans, int64_t& runs_need) {
    const int nb = tsu_host::bucket_count();
    // one pass for every pair's bucket and each bucket's count and longest
    // sequence, then the buckets' lists in pair order
    std::vector<uint8_t> bof((size_t)b->n);
    std::vector<int64_t> cnt((size_t)nb, 0);
    std::vector<int32_t> bmax((size_t)nb, 0);
    for (int32_t i = 0; i < b->n; ++i) {
        const int32_t L = std::max(b->pairs[(size_t)i].p_len, b->pairs[(size_t)i].t_len);
        const int bi = tsu_host::bucket_of(L);
        bof[(size_t)i] = (uint8_t)bi;
        ++cnt[(size_t)bi];
        bmax[(size_t)bi] = std::max(bmax[(size_t)bi], L);
    }
    std::vector<std::vector<int32_t>> bk((size_t)nb);
    for (int bi = 0; bi < nb; ++bi) bk[(size_t)bi].reserve((size_t)cnt[(size_t)bi]);
    for (int32_t i = 0; i < b->n; ++i) bk[bof[(size_t)i]].push_back(i);
    runs_need = 0;
    struct Group {
        int32_t W;    // first capacity
        bool placed;  // by the probe
        std::vector<int32_t> ids;
    };
    std::vector<Group> groups;
    std::vector<int32_t> probe_ids;
    std::vector<int> probed((size_t)nb, 0);
    for (int bi = nb - 1; bi >= 0; --bi) {
        std::vector<int32_t>& ids = bk[(size_t)bi];
        if (ids.empty()) continue;
        const int32_t max_len = bmax[(size_t)bi];
        const int32_t bound = tsu_host::bucket_bound(bi);
        const tsu_tune_entry* te = tune_lookup(ctx, bound);
        if (te && te->waves_per_pair > 1) continue;  // tuned for the LDS kernel
        int32_t W = (te && te->capacity > 0) ? te->capacity : initial_capacity(ctx, std::min(bound, max_len));
        W = std::min<int32_t>(W, capacity_full(max_len, max_len));
        if (tsu_host::probe_bucket(ctx, W, te ? te->capacity : 0, te ? te->waves_per_pair : 0, max_len)) {
            probed[(size_t)bi] = 1;
            probe_ids.insert(probe_ids.end(), ids.begin(), ids.end());
            continue;
        }
        groups.push_back(Group{W, false, std::move(ids)});
    }
    if (!probe_ids.empty()) {
        std::vector<int32_t> sc, prog;
        double kms = 0;
        int r = run_probe(ctx, b, probe_ids, d_order.p, sc, prog, &kms);
        if (r != TSU_OK) return r;
        kms_total += kms;
        ++launches;
        std::map<std::pair<int, int32_t>, std::vector<int32_t>> by;
        for (int bi = 0; bi < nb; ++bi) {
            if (!probed[(size_t)bi]) continue;
            for (int32_t i : bk[(size_t)bi]) {
                const tsu_pair& pr = b->pairs[(size_t)i];
                const int32_t v = sc[(size_t)i];
                if (v >= 0) { scores[i] = v; scored[(size_t)i] = 1; continue; }
                const int32_t ml = std::max(pr.p_len, pr.t_len);
                // a pair that overflowed the probe's window starts at its bucket's capacity
                const int32_t W = v == tsu::kProbeStopped
                                      ? tsu_host::probe_capacity(ctx, pr, prog[(size_t)i])
                                      : initial_capacity(ctx, std::min(tsu_host::bucket_bound(bi), ml));
                by[{bi, W}].push_back(i);
            }
        }
        for (auto& kv : by) groups.push_back(Group{kv.first.second, true, std::move(kv.second)});
    }
    // the groups of the 16-wave window share one launch: a pair fills a CU
    // in either, and one launch leaves one tail instead of one per bucket
    {
        Group merged{0, false, {}};
        std::vector<Group> rest;
        for (Group& g : groups) {
            int32_t ml = 0;
            for (int32_t i : g.ids) ml = std::max(ml, std::max(b->pairs[(size_t)i].p_len, b->pairs[(size_t)i].t_len));
            if (tsu_host::ml_waves(ctx, std::min<int32_t>(g.W, capacity_full(ml, ml)), ml) == 16) {
                merged.W = std::max(merged.W, g.W);
                merged.ids.insert(merged.ids.end(), g.ids.begin(), g.ids.end());
            } else {
                rest.push_back(std::move(g));
            }
        }
        if (!merged.ids.empty()) rest.push_back(std::move(merged));
        groups.swap(rest);
    }
    for (Group& g : groups) plan_group(g.W, g.placed, g.ids, plans, runs_need);
    // widest windows first, longest pairs first among equals
    std::stable_sort(plans.begin(), plans.end(), [](const RecordPlan& x, const RecordPlan& y) {
        return x.win != y.win ? x.win > y.win : x.max_len > y.max_len;
    });
    return TSU_OK;
}

// The record launch of one group at first capacity W, when a register kernel
// records it (placed: W is the probe's prediction).
void CigarCall::plan_group(int32_t W, bool placed, std::vector<int32_t>& ids, std::vector<RecordPlan>& plans,
                           int64_t& runs_need) {
    int32_t max_len = 0;
    int64_t s_len = 0;
    for (int32_t i : ids) {
        const tsu_pair& pr = b->pairs[(size_t)i];
        max_len = std::max(max_len, std::max(pr.p_len, pr.t_len));
        s_len = std::max(s_len, score_bound(ctx, pr.p_len, pr.t_len));
    }
    W = std::min<int32_t>(W, capacity_full(max_len, max_len));
    // the single-wave recorder holds sequences of up to 16,000 bases:
    // longer pairs of a narrow window take the 16-chunk kernel's narrowest
    if (max_len > 16000 && W <= kMlMinCapacity) W = std::min<int32_t>(4 * 1024, capacity_full(max_len, max_len));
    Shape sh{default_waves(W), max_len <= 32000, true, true, true};
    sh.fixed_242 = true;
    // the kernel's own window: 64 diagonals per chunk
    int32_t win = 0;
    const int c1 = bt_reg_chunks(ctx, W);
    if (c1 > 0 && max_len <= 16000) {
        sh.nw = 1;
        sh.reg_c = c1;
        sh.lring = true;
        win = 64 * c1;
    } else if (int nwl = tsu_host::ml_waves(ctx, W, max_len)) {
        sh.nw = nwl;
        sh.reg_c = 16;
        win = 1024 * nwl;
    } else if (placed && W > 16 * 1024) {
        // predicted beyond every register window: straight to the launch
        // over the whole diagonal range (room in the run pool for their bound)
        for (int32_t i : ids) {
            const tsu_pair& pr = b->pairs[(size_t)i];
            runs_need += 2 * (score_bound(ctx, pr.p_len, pr.t_len) / unit) + 2;
            wide.push_back(i);
        }
        return;
    } else {
        return;  // the LDS kernel records per score's width: sized from scores
    }
    // a window holding every diagonal of the longest pair never overflows
    int64_t s_cap = win >= capacity_full(max_len, max_len) ? s_len : std::min(s_len, score_for_capacity(ctx, win));
    if (s_cap < 0 || (sh.nw == 1 && s_cap > 16000)) return;
    KArgs a = base_args(ctx);
    a.W = W;
    a.seq_words = seq_words_for(ctx, max_len);
    int per_cu = 0;
    if (tsu::occupancy(sh, tsu::lds_bytes(sh, a), &per_cu) != 0 || per_cu <= 0) return;
    // the launch's state (records of every resident pair) stays within
    // the budget: fewer steps if needed, unless that leaves less than a
    // third of the window's scores
    const int64_t resident = std::min<int64_t>((int64_t)per_cu * ctx->cu_count, (int64_t)ids.size());
    const int64_t step_bytes = (int64_t)win * 4 + 12;
    const int64_t fit = budget / (resident * step_bytes) - 2;
    const int64_t s_win = s_cap;
    if (fit < s_cap) {
        if (fit * 3 < s_cap) return;
        s_cap = fit;
    }
    const int32_t step_cap = (int32_t)(s_cap + 2);
    const int64_t recs = (int64_t)step_cap * win;
    if (sh.nw > 1) sh.dbl = max_len <= tsu::kMlDblMaxLen && s_cap <= tsu::kMlDblMaxScore;
    const int64_t per_wg = recs * 4 + (int64_t)step_cap * 12;
    // pairs of the widest window may also be walked by the launch for
    // wider pairs, beside later launches: room for their whole bound
    for (int32_t i : ids) {
        const tsu_pair& pr = b->pairs[(size_t)i];
        const int64_t sb = score_bound(ctx, pr.p_len, pr.t_len);
        runs_need += 2 * ((win >= 16 * 1024 ? sb : std::min(s_cap + 1, sb)) / unit) + 2;
    }
    tsu_host::order_longest_first(ids, b->pairs.data());
    plans.push_back(RecordPlan{std::move(ids), sh, a, W, max_len, step_cap, win, recs, per_wg,
                               s_cap == s_win && s_win < s_len});
}

int CigarCall::settle_plan(const RecordPlan& pl) {
    for (int32_t i : pl.ids) {
        const int32_t v = sc2[(size_t)i];
        if (v >= 0) { scores[i] = v; done[(size_t)i] = 1; }
        else if (pl.win >= 16 * 1024 && (v == TSU_EKOVERFLOW || (v == TSU_ECIGAR && pl.at_window)))
            wide.push_back(i);
        else if (v != TSU_EKOVERFLOW && v != TSU_ECIGAR)
            return fail(ctx, v, "pair " + std::to_string(i) + ": " + tsu_strerror(v));
    }
    return TSU_OK;
}

// 1) The record pass.  The widest windows go first: their pairs that need
// more are then known, and run on the second stream beside the other
// launches.
int CigarCall::record_pass() {
    std::vector<RecordPlan> plans;
    int64_t runs_need = 0;
    int r0 = plan_records(plans, runs_need);
    if (r0 != TSU_OK) return r0;
    if (plans.empty() && wide.empty()) return TSU_OK;
    runs_cap = std::max<int64_t>(runs_need, 1);
    HIPCHK(d_runs.alloc((size_t)runs_cap), "hipMalloc (runs)");
    size_t k = 0;
    for (; k < plans.size() && plans[k].win >= 16 * 1024; ++k) {
        int r = launch_bt(plans[k].sh, plans[k].a, plans[k].W, plans[k].ids, plans[k].max_len, plans[k].recs,
                          plans[k].step_cap, plans[k].per_wg);
        if (r != TSU_OK) return r;
    }
    if (k > 0) {
        HIPCHK(hipMemcpy(sc2.data(), b->d_scores, (size_t)b->n * sizeof(int32_t), hipMemcpyDeviceToHost), "hipMemcpy D2H");
        for (size_t j = 0; j < k; ++j) {
            int r = settle_plan(plans[j]);
            if (r != TSU_OK) return r;
        }
    }
    if (!wide.empty() && k < plans.size()) {
        // the state of the launches after it is reserved before the wide
        // launch starts: freeing a buffer to grow it waits for the whole
        // device, so growing one later would stall them until it ends
        size_t need[3] = {0, 0, 0};
        for (size_t j = k; j < plans.size(); ++j) {
            const int64_t grid = bt_grid(plans[j].sh, plans[j].a, (int64_t)plans[j].ids.size(), plans[j].per_wg);
            need[0] = std::max(need[0], (size_t)(plans[j].recs * grid) * 4);
            need[1] = std::max(need[1], (size_t)((int64_t)plans[j].step_cap * grid) * 4);
            need[2] = std::max(need[2], (size_t)((int64_t)plans[j].step_cap * grid) * 8);
        }
        for (int i = 0; i < 3; ++i) {
            void* p = nullptr;
            HIPCHK(bt_reserve(ctx, i, need[i], &p), "hipMalloc (records)");
        }
        WidePlan wp = plan_wide(wide);
        if (wp.per_wg * (int64_t)wide.size() <= budget) {
            int r = launch_wide_async(wp, wide);
            if (r != TSU_OK) return r;
        }
    }
    for (size_t j = k; j < plans.size(); ++j) {
        int r = launch_bt(plans[j].sh, plans[j].a, plans[j].W, plans[j].ids, plans[j].max_len, plans[j].recs,
                          plans[j].step_cap, plans[j].per_wg);
        if (r != TSU_OK) return r;
    }
    if (wide_async.launched) {
        HIPCHK(hipStreamSynchronize(ctx->stream2), "hipStreamSynchronize");
        float ms = 0.f;
        HIPCHK(hipEventElapsedTime(&ms, wide_async.e0, wide_async.e1), "hipEventElapsedTime");
        kms_total += ms;
    }
    HIPCHK(hipMemcpy(sc2.data(), b->d_scores, (size_t)b->n * sizeof(int32_t), hipMemcpyDeviceToHost), "hipMemcpy D2H");
    for (size_t j = k; j < plans.size(); ++j) {
        int r = settle_plan(plans[j]);
        if (r != TSU_OK) return r;
    }
    if (wide_async.launched) {
        int r = settle_wide(wide);
        if (r != TSU_OK) return r;
        wide.clear();
    }
    return TSU_OK;
}

// 2) Pairs wider than every register window, when they did not run beside
// the record launches.
int CigarCall::wide_pass() {
    if (wide.empty()) return TSU_OK;
    WidePlan wp = plan_wide(wide);
    if (wp.per_wg * (int64_t)wide.size() > budget) return TSU_OK;  // scores first instead
    int r = grow_runs(wp.runs);
    if (r != TSU_OK) return r;
    r = launch_bt(wp.sh, wp.a, wp.W, wide, wp.max_len, wp.recs, wp.step_cap, wp.per_wg);
    if (r != TSU_OK) return r;
    HIPCHK(hipMemcpy(sc2.data(), b->d_scores, (size_t)b->n * sizeof(int32_t), hipMemcpyDeviceToHost), "hipMemcpy D2H");
    return settle_wide(wide);
}

// 3) The other pairs: scores first (batched form), which size the backtrace
// state exactly, then launches grouped by score.
int CigarCall::score_first_pass() {
    std::vector<int32_t> need;  // pairs without a score yet
    for (int32_t i = 0; i < b->n; ++i)
        if (!done[(size_t)i]) {
            order.push_back(i);
            if (!scored[(size_t)i]) need.push_back(i);
        }
    if (order.empty()) return TSU_OK;
    int r = need.empty() ? TSU_OK : batched_core(ctx, b, need, scores, t, true);
    if (r != TSU_OK) return r;
    for (int32_t i : order)
        if (scores[i] < 0) return fail(ctx, scores[i], "pair " + std::to_string(i) + ": " + tsu_strerror(scores[i]));
    tsu_host::order_desc(order, [&](int32_t i) -> int64_t { return scores[i]; });
    // their runs go after what the record launches used
    int64_t more = 0;
    for (int32_t i : order) more += 2ll * (scores[i] / ctx->g / unit) + 2;
    r = grow_runs(more);
    if (r != TSU_OK) return r;
    // group pairs into launches of similar score so buffers fit the widest
    size_t pos = 0;
    while (pos < order.size()) {
        const int32_t s0 = scores[order[pos]] / ctx->g;  // widest in this group (sorted desc)
        int32_t max_len = 0;
        size_t end = pos;
        // all pairs with score within this group share W and buffers
        const int32_t W = std::max<int32_t>(capacity_for_score(ctx, s0), 16);
        const int32_t steps = s0 + 2;
        KArgs a = base_args(ctx);
        // the group ends where the kernel family of a pair's own capacity
        // changes: a few pairs that need the LDS kernel do not take the
        // register kernels' pairs with them
        int family = -1;
        for (end = pos; end < order.size(); ++end) {
            const tsu_pair& pr = b->pairs[(size_t)order[end]];
            const int32_t len = std::max(pr.p_len, pr.t_len);
            const int32_t si = scores[order[end]] / ctx->g;
            const int f = bt_family(ctx, std::max<int32_t>(capacity_for_score(ctx, si), 16), len, si);
            if (family < 0) family = f;
            else if (f != family) break;
            max_len = std::max(max_len, len);
            if (end - pos >= 65536) break;
        }
        a.W = W;
        a.seq_words = seq_words_for(ctx, max_len);
        const int64_t group_n = (int64_t)(end - pos);
        Shape sh{default_waves(W), max_len <= 32000, true, true, true};
        sh.fixed_242 = tsu_host::fixed_242(ctx);
        int64_t recs = 0;
        int32_t step_cap = steps;
        const int c1 = bt_reg_chunks(ctx, W);
        if (c1 > 0 && max_len <= 16000 && s0 <= 16000) {
            // single-wave kernel with the int16 LDS ring: records at a fixed
            // place per step and chunk, (s * chunks + c) * 64 + lane
            sh.nw = 1;
            sh.reg_c = c1;
            sh.lring = true;
            recs = (int64_t)steps * c1 * 64;
        } else if (int nwl = tsu_host::ml_waves(ctx, W, max_len)) {
            // 16-chunk multi-wave kernel: records at a fixed place per step and
            // chunk, (s * chunks + j) * 64 + lane
            sh.nw = nwl;
            sh.reg_c = 16;
            sh.dbl = max_len <= tsu::kMlDblMaxLen && s0 <= tsu::kMlDblMaxScore;
            recs = (int64_t)steps * nwl * 16 * 64;
        } else {
            for (int64_t s = 0; s <= s0; ++s)
                recs += std::min<int64_t>(W, 2 * (std::max<int64_t>(0, s - ctx->os) / ctx->es + 2) + 1);
            recs += 64;
            tsu_host::place_sequences(ctx, sh, max_len);
            if (tsu::lds_bytes(sh, a) > ctx->max_lds) { sh.lds_ring = false; sh.nw = 16; }
        }
        int64_t per_wg = recs * 4 + (int64_t)steps * 12 + (sh.lds_ring || sh.reg_c ? 0 : tsu::ring_bytes(sh, W, ctx->rm, ctx->ri));
        // checkpoints instead of records when records do not fit one pair per CU
        const bool ck = ctx->cigar_mode == TSU_CIGAR_CHECKPOINTS ||
                        (ctx->cigar_mode == TSU_CIGAR_AUTO && s0 > 0 &&
                         per_wg * std::min<int64_t>(group_n, ctx->cu_count) > budget);
        if (ck) {
            sh = Shape{std::max(4, default_waves(W)), max_len <= 32000, true, true, true};
            sh.fixed_242 = tsu_host::fixed_242(ctx);
            sh.ck = true;
            tsu_host::place_sequences(ctx, sh, max_len);
            if (tsu::lds_bytes(sh, a) > ctx->max_lds) { sh.lds_ring = false; sh.nw = 16; }
            const int32_t nslot = ctx->rm + 2 * ctx->ri;
            const int32_t osz = sh.off16 ? 2 : 4;
            // window K balancing snapshots (s0 / K of nslot * W cells) against
            // records (K * W words)
            int32_t K = ctx->cigar_ck_every;
            if (K <= 0) K = (int32_t)std::max(8.0, std::round(std::sqrt((double)s0 * nslot * osz / 4.0)));
            K = std::max<int32_t>(1, std::min<int32_t>(K, std::max<int32_t>(s0, 1)));
            a.ck_every = K;
            a.ck_cap = (std::max<int32_t>(s0, 1) - 1) / K + 1;
            a.ck_head = (nslot * 8 + 15) / 16 * 16;
            a.ck_bytes = a.ck_head + ((int64_t)nslot * W * osz + 15) / 16 * 16;
            recs = (int64_t)K * W + 64;
            step_cap = K + 1;
            per_wg = recs * 4 + (int64_t)step_cap * 12 + (int64_t)a.ck_cap * a.ck_bytes +
                     (sh.lds_ring ? 0 : tsu::ring_bytes(sh, W, ctx->rm, ctx->ri));
        }
        std::vector<int32_t> ids(order.begin() + (ptrdiff_t)pos, order.begin() + (ptrdiff_t)end);
        r = launch_bt(sh, a, W, ids, max_len, recs, step_cap, per_wg);
        if (r != TSU_OK) return r;
        pos = end;
    }
    return TSU_OK;
}

// 2-piece penalties: scores first (batched form), then record launches on
// the 2-piece kernel in groups of falling score, each at the capacity of its
// highest score, records for the whole range of every score up to it.  (No
// checkpointed variant: the records are what CIGARs of 2-piece pairs use.)
int CigarCall::two_piece_pass() {
    for (int32_t i = 0; i < b->n; ++i) order.push_back(i);
    if (order.empty()) return TSU_OK;
    int r = batched_core(ctx, b, order, scores, t, true);
    if (r != TSU_OK) return r;
    for (int32_t i : order)
        if (scores[i] < 0) return fail(ctx, scores[i], "pair " + std::to_string(i) + ": " + tsu_strerror(scores[i]));
    tsu_host::order_desc(order, [&](int32_t i) -> int64_t { return scores[i]; });
    // a walk reserves 2 * (s / unit) + 2 runs, unit the cheapest operation
    const int32_t unit2 = std::min(ctx->xs, std::min(ctx->os + ctx->es, ctx->o2s + ctx->e2s));
    int64_t more = 0;
    for (int32_t i : order) more += 2ll * (scores[i] / ctx->g / unit2) + 2;
    r = grow_runs(more);
    if (r != TSU_OK) return r;
    // the kernel (and register window) a pair's own capacity takes: groups
    // end where it changes, so the widest pairs do not take the others onto
    // the LDS-ring kernel
    auto family = [&](int32_t i) {
        const tsu_pair& pr = b->pairs[(size_t)i];
        const int32_t Wi = std::max<int32_t>(capacity_for_score(ctx, scores[i] / ctx->g), 16);
        return tsu_host::reg2p_chunks(ctx, Wi, std::max(pr.p_len, pr.t_len));
    };
    size_t pos = 0;
    while (pos < order.size()) {
        const int32_t s0 = scores[order[pos]] / ctx->g;  // highest score of this group (sorted desc)
        const int32_t W = std::max<int32_t>(capacity_for_score(ctx, s0), 16);
        const int f0 = family(order[pos]);
        size_t end = pos + 1;
        while (end < order.size() && end - pos < 65536 && family(order[end]) == f0) ++end;
        int32_t max_len = 0;
        for (size_t j = pos; j < end; ++j) {
            const tsu_pair& pr = b->pairs[(size_t)order[j]];
            max_len = std::max(max_len, std::max(pr.p_len, pr.t_len));
        }
        KArgs a = base_args(ctx);
        a.W = W;
        a.seq_words = seq_words_for(ctx, max_len);
        Shape sh{1, max_len <= 32000, true, true, true};  // records, queue-fed
        sh.two_piece = true;
        const int32_t steps = s0 + 2;
        int64_t recs = 64;
        if (int c = tsu_host::reg2p_chunks(ctx, W, max_len)) {
            // the register kernel: records at a fixed place per score and chunk
            sh.reg_c = c;
            sh.reg_e = 10 * ctx->es + ctx->e2s;
            recs = (int64_t)steps * c * 64;
        } else {
            if (tsu::lds_bytes(sh, a) > ctx->max_lds)
                return fail(ctx, TSU_EKOVERFLOW, "pair " + std::to_string(order[pos]) +
                                                     ": its 2-piece wavefront range does not fit the CU's LDS");
            for (int64_t sc = 0; sc <= s0; ++sc) recs += std::min<int64_t>(W, capacity_for_score(ctx, sc));
        }
        const int64_t per_wg = recs * 4 + (int64_t)steps * 12;
        std::vector<int32_t> ids(order.begin() + (ptrdiff_t)pos, order.begin() + (ptrdiff_t)end);
        r = launch_bt(sh, a, W, ids, max_len, recs, steps, per_wg);
        if (r != TSU_OK) return r;
        pos = end;
    }
    return TSU_OK;
}

// The runs to the host; the score-first path's pairs must score the same in
// their backtrace launch as in the score pass.
int CigarCall::collect(tsu_cigars** out) {
    HIPCHK(hipMemcpy(sc2.data(), b->d_scores, (size_t)b->n * sizeof(int32_t), hipMemcpyDeviceToHost), "hipMemcpy D2H");
    unsigned long long used = 0;
    HIPCHK(hipMemcpy(&used, d_used.p, sizeof(used), hipMemcpyDeviceToHost), "hipMemcpy D2H");
    cg->runs.reset(static_cast<uint32_t*>(alloc_host_huge((size_t)std::max<unsigned long long>(used, 1) * 4)));
    if (!cg->runs) return fail(ctx, TSU_ENOMEM, "host allocation");
    if (used) {
        int r2 = copy_to_host(ctx, cg->runs.get(), d_runs.p, (size_t)used * 4);
        if (r2 != TSU_OK) return r2;
    }
    if (b->n) {
        HIPCHK(hipMemcpy(cg->off.data(), d_run_off.p, (size_t)b->n * 8, hipMemcpyDeviceToHost), "hipMemcpy D2H");
        HIPCHK(hipMemcpy(cg->n.data(), d_run_n.p, (size_t)b->n * 4, hipMemcpyDeviceToHost), "hipMemcpy D2H");
    }
    const double w1 = now_ms();
    if (t) { t->kernel_ms += kms_total; t->launches += launches; t->wall_ms = w1 - w0; t->pairs = b->n; t->cells = b->cells; }
    for (int32_t i : order) {
        if (sc2[(size_t)i] != scores[i]) {
            const int st = sc2[(size_t)i] < 0 ? sc2[(size_t)i] : TSU_EINTERNAL;
            return fail(ctx, st, "pair " + std::to_string(i) + ": backtrace pass returned " + std::to_string(sc2[(size_t)i]) +
                                     " (score pass " + std::to_string(scores[i]) + ")");
        }
    }
    *out = cg.release();
    return TSU_OK;
}

}  // namespace

extern "C" int tsu_align_cigar(tsu_ctx* ctx, tsu_batch* b, int32_t* scores, tsu_cigars** out, tsu_timing* t) {
    if (!ctx || !b || !scores || !out || b->ctx != ctx) return fail(ctx, TSU_EINVAL, "tsu_align_cigar: bad arguments");
    *out = nullptr;
    CigarCall call(ctx, b, scores, t);
    return call.run(out);
}

extern "C" int tsu_set_cigar_memory(tsu_ctx* ctx, int32_t mode, int32_t checkpoint_every, int64_t budget_bytes) {
    if (!ctx) return TSU_EINVAL;
    if (mode < TSU_CIGAR_AUTO || mode > TSU_CIGAR_CHECKPOINTS || checkpoint_every < 0 || budget_bytes < 0)
        return fail(ctx, TSU_EINVAL, "tsu_set_cigar_memory: bad arguments");
    ctx->cigar_mode = mode;
    ctx->cigar_ck_every = checkpoint_every;
    ctx->cigar_budget = budget_bytes;
    HIPCHK(hipSetDevice(ctx->device), "hipSetDevice");
    bt_release(ctx);  // the next call sizes its state for the new setting
    return TSU_OK;
}

extern "C" int tsu_set_poll_limit(tsu_ctx* ctx, int32_t polls) {
    if (!ctx) return TSU_EINVAL;
    if (polls < 0) return fail(ctx, TSU_EINVAL, "tsu_set_poll_limit: negative limit");
    ctx->poll_limit = polls;
    return TSU_OK;
}

extern "C" int tsu_cigars_stats(const tsu_cigars* c, int64_t* state_bytes, int32_t* record_launches,
                                int32_t* checkpoint_launches) {
    if (!c) return TSU_EINVAL;
    if (state_bytes) *state_bytes = c->state_bytes;
    if (record_launches) *record_launches = c->record_launches;
    if (checkpoint_launches) *checkpoint_launches = c->checkpoint_launches;
    return TSU_OK;
}

uint32_t tsu_host::kernel_bit(const Shape& sh) {
    if (sh.ck) return TSU_KERNEL_LDS_CK;
    if (sh.two_piece) return sh.reg_c ? TSU_KERNEL_2P_REG : TSU_KERNEL_2P;
    if (sh.reg_c > 0 && sh.nw == 1) return TSU_KERNEL_REG1;
    if (sh.reg_c == 16) return sh.nw <= 4 ? TSU_KERNEL_ML4 : sh.nw == 8 ? TSU_KERNEL_ML8 : TSU_KERNEL_ML16;
    if (sh.reg_c > 0) return TSU_KERNEL_REG1;
    return TSU_KERNEL_LDS;
}

extern "C" int tsu_cigars_kernels(const tsu_cigars* c, uint32_t* kernels) {
    if (!c || !kernels) return TSU_EINVAL;
    *kernels = c->kernels;
    return TSU_OK;
}

extern "C" int tsu_cigars_get(const tsu_cigars* c, int64_t i, const uint32_t** runs, int32_t* n_runs) {
    if (!c || i < 0 || i >= (int64_t)c->n.size()) return TSU_EINVAL;
    if (runs) *runs = c->runs.get() + c->off[(size_t)i];
    if (n_runs) *n_runs = c->n[(size_t)i];
    return TSU_OK;
}

extern "C" int tsu_cigars_string(const tsu_cigars* c, int64_t i, char* buf, int32_t cap) {
    if (!c || i < 0 || i >= (int64_t)c->n.size()) return TSU_EINVAL;
    static const char ops[4] = {'M', 'X', 'I', 'D'};
    std::string s;
    const uint32_t* r = c->runs.get() + c->off[(size_t)i];
    for (int32_t j = 0; j < c->n[(size_t)i]; ++j) {
        s += std::to_string(r[j] >> 2);
        s += ops[r[j] & 3u];
    }
    if (buf && cap > 0) std::snprintf(buf, (size_t)cap, "%s", s.c_str());
    return (int)s.size();
}

extern "C" void tsu_cigars_free(tsu_cigars* c) { delete c; }

// ---------------------------------------------------------------------------
// calibration
// ---------------------------------------------------------------------------
extern "C" int tsu_calibrate_copy(tsu_ctx* ctx, int64_t bytes, double* kernel_ms) {
    if (!ctx || bytes <= 0) return TSU_EINVAL;
    HIPCHK(hipSetDevice(ctx->device), "hipSetDevice");
    bytes = bytes / 16 * 16;
    DevBuf<uint8_t> src, dst;
    HIPCHK(src.alloc((size_t)bytes), "hipMalloc");
    HIPCHK(dst.alloc((size_t)bytes), "hipMalloc");
    HIPCHK(hipMemsetAsync(src.p, 1, (size_t)bytes, ctx->stream), "hipMemsetAsync");
    HIPCHK(hipEventRecord(ctx->ev0, ctx->stream), "hipEventRecord");
    int r = tsu::launch_copy(src.p, dst.p, bytes, ctx->stream);
    if (r != 0) return hip_fail(ctx, (hipError_t)r, "copy launch");
    HIPCHK(hipEventRecord(ctx->ev1, ctx->stream), "hipEventRecord");
    HIPCHK(hipEventSynchronize(ctx->ev1), "hipEventSynchronize");
    float ms = 0;
    HIPCHK(hipEventElapsedTime(&ms, ctx->ev0, ctx->ev1), "hipEventElapsedTime");
    if (kernel_ms) *kernel_ms = ms;
    return TSU_OK;
}
