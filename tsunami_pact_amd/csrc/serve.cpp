// serve.cpp -- continuous-batching alignment server on one GPU.
//
// Requests (1..N pairs each) are packed on the submitting thread; their pairs
// join FIFO queues keyed by length bucket and capacity.  One worker thread
// owns the device side:
//   - pairs of long-pair buckets (whose first capacity needs several waves
//     per pair) first go through the probe stage: the waiting ones are
//     aligned together by the single-wave kernel up to a score cap on a
//     stream of their own (tsu_host::probe_capacity); those that finish are
//     done, the others join the queue of their bucket at the capacity the
//     probe predicts, so a low-error long pair runs in a narrow window
//     (several pairs per CU) rather than alone on a CU in the widest one;
//   - a launch slot that frees up takes the queue whose oldest waiting pair
//     is the oldest, and merges up to max_batch_pairs of its pairs from any
//     number of requests into one flight: their sequences are copied into
//     one arena, one persistent queue-fed launch aligns them, the scores come
//     back async;
//   - device memory is one pre-allocated region of hbm_budget_bytes carved
//     by a first-fit allocator; a flight (or a probe batch) takes fewer pairs
//     (or waits) when its bytes (sequences, descriptors, scores, order, ring
//     spill) do not fit what is free;
//   - up to max_inflight flights run at once on separate streams, beside the
//     probe stage, so pairs that arrive while others are on the GPU are
//     launched as soon as a slot and memory free up, and launches of
//     different queues overlap;
//   - a finished flight scatters its scores to the requests (a request is
//     done when its last pair is) and releases its memory; pairs whose
//     diagonal range overflowed are re-launched at twice their capacity on
//     the same stream first.
// Each pair's result depends only on that pair, so a request's scores are the
// same alone and in any mix.
#include <algorithm>
#include <chrono>
#include <condition_variable>
#include <cstring>
#include <deque>
#include <map>
#include <mutex>
#include <new>
#include <thread>
#include <unordered_map>
#include <vector>

#include "tsu_host.h"

namespace {

int64_t mono_ns() {
    return std::chrono::duration_cast<std::chrono::nanoseconds>(std::chrono::steady_clock::now().time_since_epoch()).count();
}

int64_t align256(int64_t b) { return (b + 255) / 256 * 256; }

// First-fit allocator over one device region.
struct Pool {
    uint8_t* base = nullptr;
    int64_t size = 0;
    int64_t used = 0;
    std::map<int64_t, int64_t> free_;  // offset -> bytes
    void init(uint8_t* b, int64_t s) { base = b; size = s; free_.clear(); free_[0] = s; used = 0; }
    int64_t alloc(int64_t bytes) {
        bytes = align256(std::max<int64_t>(bytes, 256));
        for (auto it = free_.begin(); it != free_.end(); ++it) {
            if (it->second >= bytes) {
                const int64_t off = it->first;
                const int64_t rest = it->second - bytes;
                free_.erase(it);
                if (rest > 0) free_[off + bytes] = rest;
                used += bytes;
                return off;
            }
        }
        return -1;
    }
    void release(int64_t off, int64_t bytes) {
        bytes = align256(std::max<int64_t>(bytes, 256));
        used -= bytes;
        auto it = free_.emplace(off, bytes).first;
        auto nx = std::next(it);
        if (nx != free_.end() && it->first + it->second == nx->first) { it->second += nx->second; free_.erase(nx); }
        if (it != free_.begin()) {
            auto pv = std::prev(it);
            if (pv->first + pv->second == it->first) { pv->second += it->second; free_.erase(it); }
        }
    }
};

struct Req {
    uint64_t ticket = 0;
    std::vector<uint32_t> words;
    std::vector<tsu_pair> pairs;
    std::vector<int32_t> scores;
    int32_t max_len = 0;
    int32_t remaining = 0;  // pairs without a score yet
    int64_t submit_ns = 0, admit_ns = 0, done_ns = 0;
    int status = TSU_OK;
    bool done = false;
};

// One waiting (or in-flight) pair: pair i of request r, to be aligned at
// capacity W (0: its length bucket's first capacity).
struct Entry {
    Req* r;
    int32_t i;
    int32_t W;
};

// Launches one flight holds at once (length buckets, or probe windows).
constexpr int kMaxLaunches = 16;

struct LaunchPlan {
    int32_t first = 0, count = 0;  // range of the flight's order array
    int32_t W = 0;
    int32_t max_len = 0;
    Shape sh{};
    int64_t grid = 0;
    bool probe = false;
};

struct Slot {
    hipStream_t stream = nullptr;
    hipEvent_t ev = nullptr;
    bool active = false;
    std::vector<Entry> ents;         // flight pair k is ents[k]
    int32_t n = 0;
    int64_t off = -1, bytes = 0;
    uint32_t* d_words = nullptr;
    KPair* d_pairs = nullptr;
    int32_t* d_scores = nullptr;
    int32_t* d_order = nullptr;
    int32_t* d_queue = nullptr;  // one counter per launch
    int32_t* d_probe = nullptr;  // probe launch: furthest h + v per pair
    bool probing = false;        // this slot runs the probe stage
    uint8_t* d_ring = nullptr;
    int64_t ring_bytes = 0;
    std::vector<LaunchPlan> plan;
    std::vector<int32_t> order;       // host copy of the launch order (pair indices)
    std::vector<tsu_pair> fpairs;     // flight descriptors (host copy)
    // pinned staging
    uint32_t* h_words = nullptr; int64_t h_words_cap = 0;
    tsu_pair* h_pairs = nullptr; int64_t h_pairs_cap = 0;
    int32_t* h_order = nullptr; int64_t h_order_cap = 0;
    int32_t* h_scores = nullptr; int64_t h_scores_cap = 0;
    int32_t* h_probe = nullptr; int64_t h_probe_cap = 0;
};

template <typename T>
bool grow_pinned(T*& p, int64_t& cap, int64_t need) {
    if (need <= cap) return true;
    if (p) (void)hipHostFree(p);
    p = nullptr;
    cap = 0;
    const int64_t n = std::max<int64_t>(need, 1024) * 3 / 2;
    if (hipHostMalloc(reinterpret_cast<void**>(&p), (size_t)n * sizeof(T), hipHostMallocDefault) != hipSuccess) return false;
    cap = n;
    return true;
}

}  // namespace

struct tsu_server {
    tsu_ctx* ctx = nullptr;
    tsu_serve_cfg cfg{};
    std::mutex mu;
    std::condition_variable cv;
    std::map<int64_t, std::deque<Entry>> pending;  // waiting pairs per (bucket, capacity), FIFO
    std::deque<Entry> probe_wait;                  // long pairs waiting for the probe stage
    int64_t n_pending = 0;                         // pairs in pending and probe_wait
    Slot probe_slot;
    std::unordered_map<uint64_t, Req*> reqs;
    std::vector<Slot> slots;
    Pool pool;
    uint8_t* d_base = nullptr;
    uint64_t next_ticket = 1;
    bool stop = false;
    int fatal = TSU_OK;
    std::thread worker;
};

namespace {

// Queue key of a pair: its length bucket and capacity (0: the bucket's).
int64_t queue_key(int bucket, int32_t W) { return ((int64_t)bucket << 32) | (uint32_t)W; }

// First capacity and tuned waves of a length bucket for pairs up to max_len.
int32_t bucket_capacity(const tsu_ctx* ctx, int bucket, int32_t max_len, int* waves) {
    const int32_t bound = tsu_host::bucket_bound(bucket);
    const int32_t tcap = tsu_host::tuned_capacity(ctx, bound);
    if (waves) *waves = tsu_host::tuned_waves(ctx, bound);
    return std::min(tcap > 0 ? tcap : tsu_host::initial_capacity(ctx, std::min(bound, max_len)),
                    tsu_host::capacity_max(ctx, max_len));
}

// Pairs of this bucket go through the probe stage first.
bool probed_bucket(const tsu_ctx* ctx, int bucket, int32_t max_len) {
    const int32_t bound = tsu_host::bucket_bound(bucket);
    int waves = 0;
    const int32_t W = bucket_capacity(ctx, bucket, max_len, &waves);
    return tsu_host::probe_bucket(ctx, W, tsu_host::tuned_capacity(ctx, bound), waves, max_len);
}

// Plans the flight's one launch (its pairs share one queue: one length
// bucket, one capacity; or the probe) and returns its device bytes.  The
// ring region covers the launch and every re-launch of overflowing pairs up
// to the full capacity of the longest pair, so widening never fails.
int64_t plan_flight(tsu_server* s, Slot& sl, int32_t W_entry, int* status) {
    tsu_ctx* ctx = s->ctx;
    sl.order.resize((size_t)sl.n);
    for (int32_t i = 0; i < sl.n; ++i) sl.order[(size_t)i] = i;
    tsu_host::order_longest_first(sl.order, sl.fpairs.data());
    sl.plan.clear();
    LaunchPlan lp;
    lp.first = 0;
    lp.count = sl.n;
    for (const tsu_pair& pr : sl.fpairs) lp.max_len = std::max(lp.max_len, std::max(pr.p_len, pr.t_len));
    const int bucket = tsu_host::bucket_of(lp.max_len);
    int waves = 0;
    const int32_t W0 = bucket_capacity(ctx, bucket, lp.max_len, &waves);
    const int32_t full = tsu_host::capacity_max(ctx, lp.max_len);
    KArgs a = tsu_host::base_args(ctx);
    a.seq_words = tsu_host::seq_words_for(ctx, lp.max_len);
    if (sl.probing) {
        lp.probe = true;
        lp.W = 512;
        lp.sh = Shape{1, lp.max_len <= 32000, true, false, true};
        lp.sh.fixed_242 = true;
        lp.sh.reg_c = 8;
        lp.sh.probe = true;
    } else {
        lp.W = W_entry > 0 ? std::min(W_entry, full) : W0;
        lp.sh = tsu_host::batched_shape(ctx, lp.W, lp.max_len, W_entry > 0 ? 0 : waves);
    }
    a.W = lp.W;
    lp.grid = std::min<int64_t>(tsu_host::resident_grid(ctx, lp.sh, a, status), lp.count);
    if (*status != 0 || lp.grid <= 0) { *status = TSU_EINTERNAL; return -1; }
    int64_t ring = 0;
    if (!sl.probing) {
        if (!lp.sh.lds_ring) ring = tsu::ring_bytes(lp.sh, lp.W, ctx->rm, ctx->ri) * lp.grid;
        for (int32_t W2 = lp.W; W2 < full;) {
            W2 = (int32_t)std::min<int64_t>(2ll * W2, full);
            const Shape sh2 = tsu_host::batched_shape(ctx, W2, lp.max_len, 0);
            if (sh2.lds_ring) continue;
            KArgs a2 = a;
            a2.W = W2;
            const int64_t g2 = std::min<int64_t>(tsu_host::resident_grid(ctx, sh2, a2, status), lp.count);
            if (*status != 0 || g2 <= 0) { *status = TSU_EINTERNAL; return -1; }
            ring = std::max(ring, tsu::ring_bytes(sh2, W2, ctx->rm, ctx->ri) * g2);
        }
    }
    sl.plan.push_back(lp);
    sl.ring_bytes = ring;
    int64_t words = 0;
    for (const tsu_pair& pr : sl.fpairs) words = std::max<int64_t>(words, std::max(pr.p_off + tsu_packed_words(pr.p_len), pr.t_off + tsu_packed_words(pr.t_len)));
    *status = 0;
    return align256(words * 4) + align256((int64_t)sl.n * (int64_t)sizeof(KPair)) + 3 * align256((int64_t)sl.n * 4) +
           align256((int64_t)kMaxLaunches * 4 + 4) + align256(ring);
}

int launch_plan(tsu_server* s, Slot& sl) {
    tsu_ctx* ctx = s->ctx;
    for (size_t li = 0; li < sl.plan.size(); ++li) {
        const LaunchPlan& lp = sl.plan[li];
        KArgs a = tsu_host::base_args(ctx);
        a.words = sl.d_words;
        a.pairs = sl.d_pairs;
        a.order = sl.d_order + lp.first;
        a.n = lp.count;
        a.scores = sl.d_scores;
        a.W = lp.W;
        a.seq_words = tsu_host::seq_words_for(ctx, lp.max_len);
        a.queue = sl.d_queue + li;
        if (lp.probe) {
            a.probe_cap = tsu_host::kProbeCap;
            a.probe_out = sl.d_probe;
        }
        if (!lp.sh.lds_ring) {
            a.ring_g = sl.d_ring;
            a.ring_stride = tsu::ring_bytes(lp.sh, lp.W, ctx->rm, ctx->ri);
        }
        if (hipMemsetAsync(sl.d_queue + li, 0, sizeof(int32_t), sl.stream) != hipSuccess) return TSU_EHIP;
        if (tsu::launch_wfa(lp.sh, a, lp.grid, sl.stream) != 0) return TSU_EHIP;
    }
    return TSU_OK;
}

// Records the result (score or negative status) of pair i of request r; the
// request is done when its last pair has one.
void finish_entry(Req* r, int32_t i, int32_t v, int64_t now) {
    r->scores[(size_t)i] = v;
    if (v < 0 && r->status == TSU_OK) r->status = v;
    if (--r->remaining == 0) {
        r->done = true;
        r->done_ns = now;
    }
}

// A flight that could not be launched or did not complete: its pairs get
// `status` and its memory is released.
void fail_flight(tsu_server* s, Slot& sl, int status) {
    const int64_t now = mono_ns();
    for (const Entry& en : sl.ents) finish_entry(en.r, en.i, status, now);
    sl.ents.clear();
    if (sl.off >= 0) s->pool.release(sl.off, sl.bytes);
    sl.off = -1;
    sl.active = false;
    s->cv.notify_all();
}

// Builds and launches a flight in slot sl: from the queue whose oldest
// waiting pair is the oldest (or, for the probe slot, from the pairs waiting
// for the probe), up to max_batch_pairs of its pairs, fewer when the flight
// would not fit the free memory.  Returns true when something was launched.
bool admit(tsu_server* s, Slot& sl) {
    sl.ents.clear();
    sl.fpairs.clear();
    std::deque<Entry>* qp = nullptr;
    auto best = s->pending.end();
    if (sl.probing) {
        if (!s->probe_wait.empty()) qp = &s->probe_wait;
    } else {
        for (auto it = s->pending.begin(); it != s->pending.end(); ++it) {
            if (it->second.empty()) continue;
            if (!qp || it->second.front().r->submit_ns < qp->front().r->submit_ns) { qp = &it->second; best = it; }
        }
    }
    if (!qp) return false;
    std::deque<Entry>& q = *qp;
    const int32_t W_entry = q.front().W;
    int32_t take = (int32_t)std::min<int64_t>((int64_t)q.size(), sl.probing ? 4 * s->cfg.max_batch_pairs : s->cfg.max_batch_pairs);
    int64_t words = 0, bytes = 0;
    for (;;) {
        // flight arena: each pair's pattern and text words, in order
        sl.fpairs.resize((size_t)take);
        words = 0;
        for (int32_t k = 0; k < take; ++k) {
            const Entry& en = q[(size_t)k];
            const tsu_pair& pr = en.r->pairs[(size_t)en.i];
            tsu_pair f = pr;
            f.p_off = words;
            words += tsu_packed_words(pr.p_len);
            f.t_off = words;
            words += tsu_packed_words(pr.t_len);
            sl.fpairs[(size_t)k] = f;
        }
        sl.n = take;
        int st = 0;
        bytes = plan_flight(s, sl, W_entry, &st);
        if (st != 0 || bytes < 0) {
            s->fatal = TSU_EINTERNAL;
            return false;
        }
        const int64_t off = s->pool.alloc(bytes);
        if (off >= 0) {
            sl.off = off;
            break;
        }
        if (take == 1) return false;  // waits for memory
        take /= 2;
    }
    sl.ents.assign(q.begin(), q.begin() + take);
    q.erase(q.begin(), q.begin() + take);
    s->n_pending -= take;
    if (q.empty() && best != s->pending.end()) s->pending.erase(best);
    const int32_t n = take;
    sl.bytes = bytes;
    sl.active = false;
    uint8_t* p = s->pool.base + sl.off;
    auto carve = [&](int64_t b) { uint8_t* q = p; p += align256(b); return q; };
    sl.d_words = reinterpret_cast<uint32_t*>(carve(words * 4));
    sl.d_pairs = reinterpret_cast<KPair*>(carve((int64_t)n * (int64_t)sizeof(KPair)));
    sl.d_scores = reinterpret_cast<int32_t*>(carve((int64_t)n * 4));
    sl.d_order = reinterpret_cast<int32_t*>(carve((int64_t)n * 4));
    sl.d_probe = reinterpret_cast<int32_t*>(carve((int64_t)n * 4));
    sl.d_queue = reinterpret_cast<int32_t*>(carve((int64_t)kMaxLaunches * 4 + 4));
    sl.d_ring = sl.ring_bytes ? carve(sl.ring_bytes) : nullptr;
    // stage into pinned memory, copy, launch, copy back
    if (!grow_pinned(sl.h_words, sl.h_words_cap, std::max<int64_t>(words, 1)) ||
        !grow_pinned(sl.h_pairs, sl.h_pairs_cap, n) || !grow_pinned(sl.h_order, sl.h_order_cap, n) ||
        !grow_pinned(sl.h_scores, sl.h_scores_cap, n) || !grow_pinned(sl.h_probe, sl.h_probe_cap, n)) {
        s->fatal = TSU_ENOMEM;
        fail_flight(s, sl, TSU_ENOMEM);
        return false;
    }
    for (int32_t k = 0; k < n; ++k) {
        const Entry& en = sl.ents[(size_t)k];
        const tsu_pair& pr = en.r->pairs[(size_t)en.i];
        const tsu_pair& f = sl.fpairs[(size_t)k];
        std::memcpy(sl.h_words + f.p_off, en.r->words.data() + pr.p_off, (size_t)tsu_packed_words(pr.p_len) * 4);
        std::memcpy(sl.h_words + f.t_off, en.r->words.data() + pr.t_off, (size_t)tsu_packed_words(pr.t_len) * 4);
    }
    std::memcpy(sl.h_pairs, sl.fpairs.data(), (size_t)n * sizeof(tsu_pair));
    std::memcpy(sl.h_order, sl.order.data(), (size_t)n * 4);
    const int64_t now = mono_ns();
    for (const Entry& en : sl.ents)
        if (en.r->admit_ns == 0) en.r->admit_ns = now;
    bool ok = hipMemcpyAsync(sl.d_words, sl.h_words, (size_t)words * 4, hipMemcpyHostToDevice, sl.stream) == hipSuccess &&
              hipMemcpyAsync(sl.d_pairs, sl.h_pairs, (size_t)n * sizeof(KPair), hipMemcpyHostToDevice, sl.stream) == hipSuccess &&
              hipMemcpyAsync(sl.d_order, sl.h_order, (size_t)n * 4, hipMemcpyHostToDevice, sl.stream) == hipSuccess;
    ok = ok && launch_plan(s, sl) == TSU_OK;
    ok = ok && hipMemcpyAsync(sl.h_scores, sl.d_scores, (size_t)n * 4, hipMemcpyDeviceToHost, sl.stream) == hipSuccess;
    if (sl.probing)
        ok = ok && hipMemcpyAsync(sl.h_probe, sl.d_probe, (size_t)n * 4, hipMemcpyDeviceToHost, sl.stream) == hipSuccess;
    ok = ok && hipEventRecord(sl.ev, sl.stream) == hipSuccess;
    if (!ok) {
        // nothing of this flight may still run when its memory is reused
        (void)hipStreamSynchronize(sl.stream);
        s->fatal = TSU_EHIP;
        fail_flight(s, sl, TSU_EHIP);
        return false;
    }
    sl.active = true;
    return true;
}

// The probe batch finished: pairs it aligned are done; the others join the
// queue of their bucket at the predicted capacity (at the bucket's first
// capacity when the probe window overflowed).
void retire_probe(tsu_server* s, Slot& sl) {
    tsu_ctx* ctx = s->ctx;
    const int64_t now = mono_ns();
    for (int32_t k = 0; k < sl.n; ++k) {
        const Entry& en = sl.ents[(size_t)k];
        const int32_t v = sl.h_scores[k];
        if (v >= 0) { finish_entry(en.r, en.i, v, now); continue; }
        const tsu_pair& pr = sl.fpairs[(size_t)k];
        const int bucket = tsu_host::bucket_of(std::max(pr.p_len, pr.t_len));
        const int32_t W = v == tsu::kProbeStopped ? tsu_host::probe_capacity(ctx, pr, sl.h_probe[k], tsu_host::kProbeMarginServed) : 0;
        s->pending[queue_key(bucket, W)].push_back(Entry{en.r, en.i, W});
        ++s->n_pending;
    }
    s->pool.release(sl.off, sl.bytes);
    sl.off = -1;
    sl.ents.clear();
    sl.active = false;
    s->cv.notify_all();
}

// A flight's launch finished: pairs that overflowed go again at twice the
// capacity they had, on the same stream; when none is left, the scores are
// handed out and the memory released.
void retire(tsu_server* s, Slot& sl) {
    if (sl.probing) { retire_probe(s, sl); return; }
    tsu_ctx* ctx = s->ctx;
    std::vector<int32_t> again;
    int32_t max_len = 0;
    for (int32_t i = 0; i < sl.n; ++i) {
        if (sl.h_scores[i] != TSU_EKOVERFLOW) continue;
        again.push_back(i);
        max_len = std::max(max_len, std::max(sl.fpairs[(size_t)i].p_len, sl.fpairs[(size_t)i].t_len));
    }
    const int32_t W = sl.plan.empty() ? 0 : sl.plan[0].W;
    if (!again.empty() && W < tsu_host::capacity_max(ctx, max_len)) {
        LaunchPlan lp;
        lp.first = 0;
        lp.count = (int32_t)again.size();
        lp.W = (int32_t)std::min<int64_t>(2ll * W, tsu_host::capacity_max(ctx, max_len));
        lp.max_len = max_len;
        lp.sh = tsu_host::batched_shape(ctx, lp.W, max_len, 0);
        KArgs a = tsu_host::base_args(ctx);
        a.W = lp.W;
        a.seq_words = tsu_host::seq_words_for(ctx, max_len);
        int st = 0;
        lp.grid = std::min<int64_t>(tsu_host::resident_grid(ctx, lp.sh, a, &st), lp.count);
        // the flight reserved the ring of every widening up to full capacity
        if (!lp.sh.lds_ring && st == 0)
            lp.grid = std::min<int64_t>(lp.grid, sl.ring_bytes / std::max<int64_t>(1, tsu::ring_bytes(lp.sh, lp.W, ctx->rm, ctx->ri)));
        if (st == 0 && lp.grid > 0) {
            sl.order = again;
            std::memcpy(sl.h_order, again.data(), again.size() * 4);
            sl.plan.assign(1, lp);
            bool ok = hipMemcpyAsync(sl.d_order, sl.h_order, again.size() * 4, hipMemcpyHostToDevice, sl.stream) == hipSuccess;
            ok = ok && launch_plan(s, sl) == TSU_OK;
            ok = ok && hipMemcpyAsync(sl.h_scores, sl.d_scores, (size_t)sl.n * 4, hipMemcpyDeviceToHost, sl.stream) == hipSuccess &&
                 hipEventRecord(sl.ev, sl.stream) == hipSuccess;
            if (ok) return;  // still in flight
            (void)hipStreamSynchronize(sl.stream);  // before the memory is reused
            s->fatal = TSU_EHIP;
        } else if (st != 0) {
            s->fatal = TSU_EINTERNAL;
        }
    }
    const int64_t now = mono_ns();
    for (int32_t k = 0; k < sl.n; ++k) finish_entry(sl.ents[(size_t)k].r, sl.ents[(size_t)k].i, sl.h_scores[k], now);
    s->pool.release(sl.off, sl.bytes);
    sl.off = -1;
    sl.ents.clear();
    sl.active = false;
    s->cv.notify_all();
}

// The flight slots and the probe slot.
std::vector<Slot*> all_slots(tsu_server* s) {
    std::vector<Slot*> v;
    for (Slot& sl : s->slots) v.push_back(&sl);
    v.push_back(&s->probe_slot);
    return v;
}

void worker_loop(tsu_server* s) {
    (void)hipSetDevice(s->ctx->device);
    std::unique_lock<std::mutex> lk(s->mu);
    for (;;) {
        bool progressed = false;
        bool any_active = false;
        auto poll = [&](Slot& sl) {
            if (!sl.active) return;
            const hipError_t q = hipEventQuery(sl.ev);
            if (q == hipSuccess) { retire(s, sl); progressed = true; }
            else if (q != hipErrorNotReady) {
                (void)hipStreamSynchronize(sl.stream);
                s->fatal = TSU_EHIP;
                fail_flight(s, sl, TSU_EHIP);
                progressed = true;
            }
            if (sl.active) any_active = true;
        };
        poll(s->probe_slot);
        for (Slot& sl : s->slots) poll(sl);
        if (s->fatal == TSU_OK) {
            // the probe stage first: its pairs feed the queues the slots take from
            if (!s->probe_slot.active && !s->probe_wait.empty() && admit(s, s->probe_slot)) {
                progressed = true;
                any_active = true;
            }
            for (Slot& sl : s->slots) {
                if (sl.active || s->pending.empty()) continue;
                if (admit(s, sl)) { progressed = true; any_active = true; }
                else break;
            }
        }
        if (s->fatal != TSU_OK) {
            // fail everything still queued; in-flight work is waited for below
            const int64_t now = mono_ns();
            for (auto& kv : s->pending) {
                for (const Entry& en : kv.second) finish_entry(en.r, en.i, s->fatal, now);
            }
            s->pending.clear();
            for (const Entry& en : s->probe_wait) finish_entry(en.r, en.i, s->fatal, now);
            s->probe_wait.clear();
            s->n_pending = 0;
            s->cv.notify_all();
        }
        if (s->stop && s->n_pending == 0 && !any_active) break;
        if (!progressed) {
            if (any_active) s->cv.wait_for(lk, std::chrono::microseconds(50));
            else s->cv.wait(lk);
        }
    }
}

}  // namespace

extern "C" int tsu_serve_open(tsu_ctx* ctx, const tsu_serve_cfg* cfg, tsu_server** out) {
    if (!ctx || !cfg || !out) return TSU_EINVAL;
    *out = nullptr;
    if (cfg->hbm_budget_bytes < (1 << 20) || cfg->max_inflight < 1 || cfg->max_inflight > 64 || cfg->max_batch_pairs < 1)
        return tsu_host::fail(ctx, TSU_EINVAL, "tsu_serve_open: budget >= 1 MiB, 1 <= max_inflight <= 64, max_batch_pairs >= 1");
    if (hipSetDevice(ctx->device) != hipSuccess) return tsu_host::fail(ctx, TSU_EHIP, "hipSetDevice");
    tsu_server* s = new (std::nothrow) tsu_server();
    if (!s) return TSU_ENOMEM;
    s->ctx = ctx;
    s->cfg = *cfg;
    if (hipMalloc(reinterpret_cast<void**>(&s->d_base), (size_t)cfg->hbm_budget_bytes) != hipSuccess) {
        delete s;
        return tsu_host::fail(ctx, TSU_ENOMEM, "tsu_serve_open: cannot reserve the HBM budget");
    }
    s->pool.init(s->d_base, cfg->hbm_budget_bytes);
    s->slots.resize((size_t)cfg->max_inflight);
    s->probe_slot.probing = true;
    for (Slot* sl : all_slots(s)) {
        if (hipStreamCreateWithFlags(&sl->stream, hipStreamNonBlocking) != hipSuccess ||
            hipEventCreateWithFlags(&sl->ev, hipEventDisableTiming) != hipSuccess) {
            tsu_serve_close(s);
            return tsu_host::fail(ctx, TSU_EHIP, "tsu_serve_open: stream creation");
        }
    }
    s->worker = std::thread(worker_loop, s);
    *out = s;
    return TSU_OK;
}

static int enqueue(tsu_server* s, Req* r, uint64_t* ticket);

extern "C" int tsu_serve_submit(tsu_server* s, const char* const* patterns, const int32_t* p_len,
                                const char* const* texts, const int32_t* t_len, int32_t n, uint64_t* ticket) {
    if (!s || !ticket || n < 0 || (n > 0 && (!patterns || !p_len || !texts || !t_len))) return TSU_EINVAL;
    Req* r = new (std::nothrow) Req();
    if (!r) return TSU_ENOMEM;
    r->submit_ns = mono_ns();
    try {
        r->pairs.resize((size_t)n);
        int64_t w = 0;
        for (int32_t i = 0; i < n; ++i) {
            const int32_t lmax = tsu_host::seq_len_limit(s->ctx);
            if (p_len[i] < 0 || t_len[i] < 0 || p_len[i] > lmax || t_len[i] > lmax) { delete r; return TSU_EINVAL; }
            r->pairs[(size_t)i] = tsu_pair{w, w + tsu_packed_words(p_len[i]), p_len[i], t_len[i]};
            w += tsu_packed_words(p_len[i]) + tsu_packed_words(t_len[i]);
            r->max_len = std::max(r->max_len, std::max(p_len[i], t_len[i]));
        }
        r->words.assign((size_t)w, 0u);
        for (int32_t i = 0; i < n; ++i) {
            if (tsu_pack_seq(patterns[i], p_len[i], r->words.data() + r->pairs[(size_t)i].p_off) != TSU_OK ||
                tsu_pack_seq(texts[i], t_len[i], r->words.data() + r->pairs[(size_t)i].t_off) != TSU_OK) {
                delete r;
                return TSU_EBADINPUT;
            }
        }
    } catch (...) {
        delete r;
        return TSU_ENOMEM;
    }
    return enqueue(s, r, ticket);
}

extern "C" int tsu_serve_submit_packed(tsu_server* s, const uint32_t* words, int64_t n_words, const tsu_pair* pairs,
                                       int32_t n, uint64_t* ticket) {
    if (!s || !ticket || n < 0 || n_words < 0 || (n > 0 && (!words || !pairs))) return TSU_EINVAL;
    const int32_t lmax = tsu_host::seq_len_limit(s->ctx);
    for (int32_t i = 0; i < n; ++i) {
        const tsu_pair& pr = pairs[i];
        if (pr.p_len < 0 || pr.t_len < 0 || pr.p_len > lmax || pr.t_len > lmax || pr.p_off < 0 || pr.t_off < 0 ||
            pr.p_off + tsu_packed_words(pr.p_len) > n_words || pr.t_off + tsu_packed_words(pr.t_len) > n_words)
            return TSU_EINVAL;
    }
    Req* r = new (std::nothrow) Req();
    if (!r) return TSU_ENOMEM;
    r->submit_ns = mono_ns();
    try {
        r->words.assign(words, words + n_words);
        r->pairs.assign(pairs, pairs + n);
    } catch (...) {
        delete r;
        return TSU_ENOMEM;
    }
    for (int32_t i = 0; i < n; ++i) r->max_len = std::max(r->max_len, std::max(pairs[i].p_len, pairs[i].t_len));
    return enqueue(s, r, ticket);
}

// Queues a packed request (or completes an empty one) under the server lock.
static int enqueue(tsu_server* s, Req* r, uint64_t* ticket) {
    const int32_t n = (int32_t)r->pairs.size();
    (void)hipSetDevice(s->ctx->device);
    std::lock_guard<std::mutex> lk(s->mu);
    if (s->stop) { delete r; return TSU_EINVAL; }
    if (s->fatal != TSU_OK) { delete r; return s->fatal; }
    r->ticket = s->next_ticket++;
    *ticket = r->ticket;
    s->reqs[r->ticket] = r;
    if (n == 0) {
        r->done = true;
        r->admit_ns = r->done_ns = mono_ns();
    } else {
        // a request that can never fit the budget is refused up front: a
        // flight of all its pairs at once must fit
        Slot trial;
        trial.fpairs = r->pairs;
        trial.n = n;
        int st = 0;
        const int64_t need = plan_flight(s, trial, 0, &st);
        if (st != 0 || need > s->pool.size) {
            s->reqs.erase(r->ticket);
            delete r;
            return st != 0 ? TSU_EINTERNAL : TSU_EBUDGET;
        }
        r->scores.assign((size_t)n, 0);
        r->remaining = n;
        for (int32_t i = 0; i < n; ++i) {
            const tsu_pair& pr = r->pairs[(size_t)i];
            const int32_t ml = std::max(pr.p_len, pr.t_len);
            const int bucket = tsu_host::bucket_of(ml);
            if (probed_bucket(s->ctx, bucket, ml)) s->probe_wait.push_back(Entry{r, i, 0});
            else s->pending[queue_key(bucket, 0)].push_back(Entry{r, i, 0});
        }
        s->n_pending += n;
    }
    s->cv.notify_all();
    return TSU_OK;
}

static int take_result(tsu_server* s, Req* r, int32_t* scores, tsu_req_stats* st) {
    if (scores && !r->scores.empty()) std::memcpy(scores, r->scores.data(), r->scores.size() * 4);
    if (st) {
        st->submit_ns = r->submit_ns;
        st->admit_ns = r->admit_ns;
        st->done_ns = r->done_ns;
        st->status = r->status;
        st->n_pairs = (int32_t)r->pairs.size();
    }
    const int status = r->status;
    s->reqs.erase(r->ticket);
    delete r;
    return status;
}

extern "C" int tsu_serve_poll(tsu_server* s, uint64_t ticket, int32_t* scores, tsu_req_stats* st) {
    if (!s) return TSU_EINVAL;
    std::lock_guard<std::mutex> lk(s->mu);
    auto it = s->reqs.find(ticket);
    if (it == s->reqs.end()) return TSU_EINVAL;
    if (!it->second->done) return TSU_EAGAIN;
    return take_result(s, it->second, scores, st);
}

extern "C" int tsu_serve_wait(tsu_server* s, uint64_t ticket, int32_t* scores, tsu_req_stats* st) {
    if (!s) return TSU_EINVAL;
    std::unique_lock<std::mutex> lk(s->mu);
    auto it = s->reqs.find(ticket);
    if (it == s->reqs.end()) return TSU_EINVAL;
    Req* r = it->second;
    s->cv.wait(lk, [&] { return r->done; });
    return take_result(s, r, scores, st);
}

extern "C" int tsu_serve_close(tsu_server* s) {
    if (!s) return TSU_EINVAL;
    {
        std::lock_guard<std::mutex> lk(s->mu);
        s->stop = true;
    }
    s->cv.notify_all();
    if (s->worker.joinable()) s->worker.join();
    (void)hipSetDevice(s->ctx->device);
    for (Slot* sl : all_slots(s)) {
        if (sl->stream) (void)hipStreamSynchronize(sl->stream);
        if (sl->ev) (void)hipEventDestroy(sl->ev);
        if (sl->stream) (void)hipStreamDestroy(sl->stream);
        if (sl->h_words) (void)hipHostFree(sl->h_words);
        if (sl->h_pairs) (void)hipHostFree(sl->h_pairs);
        if (sl->h_order) (void)hipHostFree(sl->h_order);
        if (sl->h_scores) (void)hipHostFree(sl->h_scores);
        if (sl->h_probe) (void)hipHostFree(sl->h_probe);
    }
    if (s->d_base) (void)hipFree(s->d_base);
    const int status = s->fatal;
    for (auto& kv : s->reqs) delete kv.second;
    delete s;
    return status;
}
