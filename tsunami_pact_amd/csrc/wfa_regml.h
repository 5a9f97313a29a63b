// wfa_regml.h -- multi-wave WFA with the two newest M scores, I and D in
// registers and the two older M scores in LDS.
//
// With the scaled penalties 2/3/1 the score-s step reads M[s-2] (mismatch,
// same diagonal), M[s-4] (gap open, diagonals k-1 and k+1) and I/D[s-1].
// Here a lane keeps, per chunk of 64 diagonals, M of the two newest scores
// (a ring of two VGPRs by the parity of s) and I and D (updated in place):
// four VGPRs per chunk instead of the seven of a full register ring, so a
// wave holds 16 chunks (1024 diagonals) and 16 waves hold 16384.  M[s-2],
// once read for the mismatch, is written to an int16 LDS ring of three
// slots (the scores s-2, s-3 and s-4 of the current step), from which the
// gap-open values M[s-4][k-1] and M[s-4][k+1] are read straight at their
// diagonals, across chunk and wave boundaries alike.  The slot written at
// step s (score s-2) is never the one read (score s-4); the waves of a pair
// synchronise with their neighbours only (below), which orders the writes
// before the reads two steps later.
//
// Global chunk j belongs to wave j mod NW, register slot j / NW, as in the
// multi-wave register kernel (wfa_regmw.h).  I and D cross chunk edges
// through LDS: each computed chunk publishes lane 63 of I and lane 0 of D
// for its parity of s.  The chunks computed at score s are those that meet
// the wavefront limits of s (WFA2-lib's lo / hi, wfa_gpu.cu:426-468), which
// every wave derives from the same scalar recurrence, so the waves exchange
// no activity masks.  The limits only grow, so a chunk, once computed, is
// computed at every later score; the registers, edge words and ring slots of
// a chunk not yet computed hold NULL.
//
// NULL is kRingNull in registers and -32768 in the int16 ring: its low half
// is the ring's NULL, so an M register (a valid offset or exactly NULL) goes
// to the ring by plain truncation.  An I or D value fed only by the ring's
// NULL grows by one per score, so the kernel gives up
// (TSU_EKOVERFLOW, the host re-runs the pair on another kernel) beyond
// kMlMaxScore scaled scores, where such a value would still be negative;
// offsets must fit int16, so sequences are at most 32000 bases.
//
// With DBL the offsets are kept doubled, as in the single-wave kernels: a
// doubled offset is the bit position of its base in the packed words, which
// saves the shifts around every extend.  Doubled offsets up to 2t fit the
// int16 ring for sequences of at most 16383 bases, and a doubled value fed
// only by the ring's NULL grows by two per score, so that variant gives up
// past kMlMaxScoreDbl = 16383 scaled scores (still negative: -32768 + 2s).
// A pair whose sequences hold at most 8190 bases cannot score more.
//
// Recurrences, bounds test, extend and end test are those of the other
// kernels (wfa_gpu.cu:253-331,610-631); the trim (wfa_gpu.cu:633-719) is
// replaced by NULLing out-of-bounds cells where they are computed (below).
//
// With BT every computed chunk also writes its 64 cell records (wfa_bt.h)
// to global memory at ((s * NC + j) * 64 + lane) of the workgroup's region
// (j the global chunk index), one coalesced 256-byte store per chunk and
// step; after the end cell is reached, wave 0 walks them into CIGAR runs.
#pragma once

#include "wfa_bt.h"
#include "wfa_regmw.h"

namespace tsu {

constexpr int kMlMaxScore = 30000;
constexpr int kMlMaxScoreDbl = kMlDblMaxScore;
// Register NULL of the kernels with an int16 M ring: large and negative, low
// 16 bits 0x8000 (the int16 ring's NULL).
constexpr int32_t kRingNull = (int32_t)0xC0008000u;

// LDS layout (int32 words) after the interleaved sequence words.
template <int NW, int C>
struct MlLayout {
    static constexpr int NC = NW * C;                       // chunks in the window
    static constexpr int W = 64 * NC;                       // diagonals in the window
    static constexpr int EI = 0;                            // [2][NC+2] I lane 63 per chunk, by parity of s
    static constexpr int ED = EI + 2 * (NC + 2);            // [2][NC+2] D lane 0 per chunk
    static constexpr int MISC = ED + 2 * (NC + 2);          // [4] queue index
    static constexpr int PG = MISC + 4;                     // [NW] neighbour sync: last score each wave finished
    static constexpr int EV = PG + NW;                      // [4] neighbour sync: first event, score * 4 + kind
    static constexpr int SK = EV + 4;                       // [64+NC] write sink of the lanes that publish nothing
    static constexpr int MR = (SK + 64 + NC + 3) & ~3;      // int16 [3][W+2] M ring, NULL guards at both ends
    static constexpr int MR_STRIDE = W + 2;                 // int16 entries per slot
    static constexpr int WORDS = MR + (3 * MR_STRIDE + 1) / 2;
};

// Neighbour synchronisation.  Wave w's step s reads, besides its own
// registers, only what the waves owning its neighbour chunks (w-1 and w+1,
// cyclically) wrote at steps s-1 and s-3: their edge lanes and their ring
// entries.  So instead of a workgroup barrier per score, each wave publishes
// the last score it finished and starts score s+1 once both neighbours have
// finished s.  Neighbours are then never more than one score apart, which is
// what the double-buffered edges and the 3-slot M ring need: a wave at s+1
// overwrites only data of scores its neighbours have finished reading.  What needs the whole window is turned
// into per-wave events kept as a minimum (score * 4 + kind): the end cell
// (kind 0), a non-NULL cell at the window edge (1), the score limit (2) and
// an internal error (3).  Every wave goes on until it has finished the
// lowest recorded score, so an event at a lower score in a slower wave is
// never missed, and the pair's result is that minimum.  The exact trim is
// replaced by NULLing every out-of-bounds I or D cell where it is computed.
// Cells can differ from the trimmed computation, but neither the score nor
// the CIGAR walked over the records does: the oracle's orc_wfa_nulled
// restates this variant and its tests pin both to the trimmed WFA on every
// pair of up to 3 bases and on ragged random pairs.  With BT, the records of
// a score past the end cell (a wave may start one more score before it sees
// the event) are written but never walked.  Spins are bounded (kMlSpinCap sleeps, or KArgs::poll_limit);
// past it the pair fails with TSU_EINTERNAL instead of hanging.
constexpr int kEvEnd = 0, kEvEdge = 1, kEvMax = 2, kEvInternal = 3;
constexpr int kMlSpinCap = 1 << 20;

typedef __attribute__((address_space(3))) const int32_t lds_ci32;

template <int NW, int C, bool QUEUE, bool BT = false, bool DBL = false>
__global__ __launch_bounds__(NW * 64) void wfa_regml_kernel(KArgs a) {
    constexpr int U = DBL ? 2 : 1;  // offset unit: 2 when offsets are kept doubled
    static_assert(NW >= 2 && NW <= 16, "2..16 waves");
    static_assert(C >= 1 && C <= 16, "16 chunk bits per wave");
    constexpr int32_t NUL = kRingNull;
    constexpr int NC = NW * C;
    constexpr int W = 64 * NC;
    constexpr int NT = NW * 64;
    using L = MlLayout<NW, C>;
    extern __shared__ __align__(16) uint32_t smem[];
    const int tid = threadIdx.x;
    const int lane = tid & 63;
    const int w = __builtin_amdgcn_readfirstlane(tid >> 6);
    uint32_t* sS = smem;  // interleaved pattern / text words
    int32_t* X = reinterpret_cast<int32_t*>(smem + 2 * a.seq_words);
    int16_t* MR = reinterpret_cast<int16_t*>(X + L::MR);
    const bool lds_at_zero = (uint32_t)(uintptr_t)sS == 0u;
    const int wl = w == 0 ? NW - 1 : w - 1;  // owner of chunk j-1 (slot r-1 when w == 0)
    const int wr = w == NW - 1 ? 0 : w + 1;  // owner of chunk j+1 (slot r+1 when w == NW-1)
    // per computed chunk every lane stores once: lane 63 its I edge, lane 0
    // its D edge, the others into a sink (cheaper than masking the store)
    const int pub_off = lane == 0 ? L::ED - L::EI : 0;
    const int poll_limit = a.poll_limit > 0 ? a.poll_limit : kMlSpinCap;
    uint32_t* rec = nullptr;
    if constexpr (BT) rec = a.rec_g + (int64_t)blockIdx.x * a.rec_cap;

    for (int iter = 0;; ++iter) {
        int idx;
        if constexpr (QUEUE) {
            if (tid == 0) X[L::MISC] = atomicAdd(a.queue, 1);
            __syncthreads();
            idx = __builtin_amdgcn_readfirstlane(X[L::MISC]);
            __syncthreads();
        } else {
            if (iter > 0) break;
            idx = blockIdx.x;
        }
        if (idx >= a.n) break;
        const int pi = a.order ? __builtin_amdgcn_readfirstlane(a.order[idx]) : idx;
        const KPair pr = a.pairs[pi];
        const int p = __builtin_amdgcn_readfirstlane(pr.p_len);
        const int t = __builtin_amdgcn_readfirstlane(pr.t_len);
        const int pw = (p + 15) >> 4, tw = (t + 15) >> 4;
        for (int i = tid; i < a.seq_words; i += NT) {
            sS[2 * i] = i < pw ? a.words[pr.p_off + i] : 0u;
            sS[2 * i + 1] = i < tw ? a.words[pr.t_off + i] : 0u;
        }
        for (int i = tid; i < L::WORDS; i += NT) {
            int32_t v = NUL;                                          // I / D edges
            if (i >= L::MISC) v = 0;
            if (i >= L::MR) v = (int32_t)0x80008000u;                 // two int16 NULLs
            if (i >= L::PG && i < L::EV) v = 0;                       // progress: score 0 done
            if (i == L::EV) v = 0x7fffffff;                           // no event yet
            if (i < L::MISC || i >= L::PG) X[i] = v;
        }
        __syncthreads();

        int kmin = max(-(W / 2), -p - 1);
        int kmax = kmin + W - 1;
        if (kmax > t + 1) { kmax = t + 1; kmin = max(kmax - W + 1, -p - 1); }
        const int tp = t - p;
        const int jtp = tp - kmin >= 0 && tp - kmin < W ? (tp - kmin) >> 6 : -1;
        // chunk bit of the end diagonal in its owner wave, and its lane
        const uint32_t endm = jtp >= 0 && jtp % NW == w ? 1u << (jtp / NW) : 0u;
        const int ltp = (tp - kmin) & 63;
        const int64_t bound_u = 2ll * min(p, t) + (p != t ? 3ll + (int64_t)abs(p - t) : 0);
        const int s_bound = (int)min<int64_t>(bound_u, 0x3fffffff);

        int32_t Mr[2][C], Ir[C], Dr[C];
#pragma unroll
        for (int c = 0; c < C; ++c) {
            Mr[0][c] = NUL;
            Mr[1][c] = NUL;
            Ir[c] = NUL;
            Dr[c] = NUL;
        }
        // Window diagonal limits (j = k - kmin) of the cells that may be
        // non-NULL, for the scaled penalties 2/3/1: I/D at s within
        // [min(lo M[s-4], lo I/D[s-1]) - 1, max(hi M[s-4], hi I/D[s-1]) + 1],
        // M at s the union of that and M at s-2; mt/mb hold the M limits of
        // the last two scores of each parity ([par][0] newest), it/ib those
        // of I/D at s-1.
        constexpr int RNONE = 1 << 20;  // limits of an empty range: top -RNONE, bottom +RNONE
        int mt[2][2], mb[2][2];
        mt[0][0] = -kmin; mb[0][0] = -kmin;  // score 0: M[0][0] only
        mt[0][1] = -RNONE; mb[0][1] = RNONE;
        mt[1][0] = -RNONE; mb[1][0] = RNONE;
        mt[1][1] = -RNONE; mb[1][1] = RNONE;
        int it = -RNONE, ib = RNONE;
        const uint32_t w0 = w == 0 ? 1u : 0u, wlast = w == NW - 1 ? 1u : 0u;

        // score 0: M[0][0] = 0, extended (every lane computes it)
        const int ext0 = lcp_interleaved(sS, 0, 0, min(p, t));
        {
            const int j0 = (0 - kmin) >> 6, l0 = (0 - kmin) & 63;
            if (j0 % NW == w) {
                const int r0 = j0 / NW;
#pragma unroll
                for (int r = 0; r < C; ++r)
                    if (r == r0) Mr[0][r] = lane == l0 ? U * ext0 : NUL;
            }
        }
        __syncthreads();
        int s = 0;
        int l3 = 0;  // s mod 3

        // One score step; PAR = s & 1 selects the register slot holding
        // M[s-2], which receives M[s].  Returns 0 to go on, 1 once an event
        // is recorded at a score this wave has finished, or a negative
        // status (the score limits, reached by every wave at the same score,
        // or a neighbour wait past its bound).
        auto step = [&](auto PC) __attribute__((always_inline)) -> int {
            constexpr int PAR = decltype(PC)::value;
            {
                // both neighbours have finished score s; stop once an event
                // is recorded at a score this wave has finished
                for (int spins = 0;; ++spins) {
                    const int ev = __builtin_amdgcn_readfirstlane(
                        __hip_atomic_load(X + L::EV, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP));
                    if ((ev >> 2) <= s) return 1;
                    const int g0 = __builtin_amdgcn_readfirstlane(
                        __hip_atomic_load(X + L::PG + wl, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP));
                    const int g1 = __builtin_amdgcn_readfirstlane(
                        __hip_atomic_load(X + L::PG + wr, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP));
                    if (g0 >= s && g1 >= s) break;
                    if (spins >= poll_limit) {
                        atomicMin(X + L::EV, ((s + 1) << 2) | kEvInternal);  // uniform: one lane's atomic
                        return TSU_EINTERNAL;
                    }
                    __builtin_amdgcn_s_sleep(2);  // about 128 clocks: 0.5% faster than 64 on a mixed batch; 256 within noise of it, 512 slower
                }
                __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup");
            }
            ++s;
            // every wave reaches the score limits at the same score, so with
            // neighbour synchronisation too they need no event
            if (s > s_bound) return TSU_EINTERNAL;
            if (s > (DBL ? kMlMaxScoreDbl : kMlMaxScore)) return TSU_EKOVERFLOW;
            if constexpr (BT) {
                if (s >= a.step_cap) return TSU_ECIGAR;
            }
            l3 = l3 == 2 ? 0 : l3 + 1;
            const int rd = l3 == 0 ? 2 : l3 - 1;   // ring slot of s-4
            const int wrs = l3 == 2 ? 0 : l3 + 1;  // ring slot of s-2 (formerly s-5)
            // opaque per-step copies of the window base and the lane: the
            // per-chunk diagonals are recomputed each step rather than held
            // in VGPRs across the score loop
            int pb = 64 * w;
            int ln = lane;
            asm volatile("" : "+s"(pb));
            asm volatile("" : "+v"(ln));
            const int kl = U * (kmin + pb + ln);  // diagonal of this lane in chunk 0 (in offset units)
            // limits at s; this wave computes its chunks r (global chunk
            // r NW + w) that meet the M range
            it = max(mt[PAR][1], it) + 1;
            ib = min(mb[PAR][1], ib) - 1;
            mt[PAR][1] = mt[PAR][0];
            mb[PAR][1] = mb[PAR][0];
            mt[PAR][0] = max(mt[PAR][0], it);
            mb[PAR][0] = min(mb[PAR][0], ib);
            uint32_t act = 0;
            {
                const int jlo = max(mb[PAR][0], 0) >> 6, jhi = min(mt[PAR][0], W - 1) >> 6;
                const int rlo = (jlo + NW - 1 - w) / NW;  // >= 0: jlo >= 0, w < NW
                const int rhi = jhi >= w ? (jhi - w) / NW : -1;
                if (mb[PAR][0] <= mt[PAR][0] && rlo <= rhi) act = (2u << rhi) - (1u << rlo);
            }
            // edges of score s-1, read through one VGPR address per step
            int ea = 4 * (2 * a.seq_words + L::EI + (PAR ^ 1) * (NC + 2) + w);
            asm volatile("" : "+v"(ea));
            lds_ci32* ei1 = reinterpret_cast<lds_ci32*>((uintptr_t)(uint32_t)ea);
            lds_ci32* ed1 = ei1 + (L::ED - L::EI);
            int32_t* epub = (lane == 0 || lane == 63) ? X + L::EI + PAR * (NC + 2) + w + 1 + pub_off : X + L::SK + lane;
            uint32_t em = endm;  // opaque per step: no per-chunk predicates held across the loop
            asm volatile("" : "+s"(em));
            // records of this step: chunk j at (s * NC + j) * 64 (uniform base,
            // 32-bit lane offset pb + ln)
            uint32_t* recs = nullptr;
            const uint32_t rl = (uint32_t)(pb + ln);
            if constexpr (BT) recs = rec + (int64_t)s * NC * 64;
            const int16_t* mrd = MR + rd * L::MR_STRIDE + pb + ln;        // entry of k-1
            int16_t* mwr = MR + wrs * L::MR_STRIDE + pb + ln + 1;         // entry of k
            uint32_t edge = 0;
            int32_t mend = NUL;
#pragma unroll
            for (int c = 0; c < C; ++c) {
                if (act & (1u << c)) {
                    // one past the last offset in bounds: min(k + p + 1, t + 1)
                    // (>= 0 on every diagonal of the window, kmin >= -p-1)
                    const int mt1 = min(kl + U * (p + 1 + 64 * NW * c), U * (t + 1));
                    const int32_t mx = Mr[PAR][c];
                    mwr[64 * NW * c] = (int16_t)mx;  // exactly NULL or an offset below 32768
                    const int32_t mol = mrd[64 * NW * c];
                    const int32_t mor = mrd[64 * NW * c + 2];
                    const int32_t fil = ei1[c * NW];
                    const int32_t fdr = ed1[c * NW + 2];
                    const int32_t iel = dpp_shr1(fil, Ir[c]);
                    const int32_t der = dpp_shl1(fdr, Dr[c]);
                    uint32_t rb = 0;  // backtrace record bits of the gap sources
                    if constexpr (BT) rb = (iel >= mol ? 4u : 0u) | (der >= mor ? 8u : 0u);
                    int32_t I = max(mol, iel) + U;
                    int32_t D = max(mor, der);
                    // out-of-bounds I / D cells are NULL where computed
                    I = (uint32_t)I < (uint32_t)mt1 ? I : NUL;
                    D = (uint32_t)D < (uint32_t)mt1 ? D : NUL;
                    const int32_t X = mx + U;
                    // M before the bounds test: non-NULL where M, I or D is,
                    // or a mismatch from a non-NULL M[s-2]
                    const int32_t Mp = max(max(D, X), I);
                    const bool valid = (uint32_t)Mp < (uint32_t)mt1;
                    if constexpr (BT) {
                        const uint32_t r = rb | (valid ? (((uint32_t)Mp >> (U - 1)) << 4) | ((X >= D && X >= I) ? 1u : (D >= I ? 2u : 3u)) : 0u);
                        recs[(uint32_t)(64 * NW * c) + rl] = r;
                    }
                    int32_t M = valid ? Mp : NUL;
                    if constexpr (DBL) {
                        const int lim2 = mt1 - 2 - M;  // 2 x bases left
                        if (valid && lim2 > 0) M += extend2((uint32_t)(M - kl - 128 * NW * c), (uint32_t)M, lim2);
                    } else {
                        const int lim1 = mt1 - M;  // bases left + 1
                        if (valid && lim1 > 1) M += extend_len((uint32_t)(M - kl - 64 * NW * c), (uint32_t)M, lim1 - 1);
                    }
                    Mr[PAR][c] = M;
                    Ir[c] = I;
                    Dr[c] = D;
                    epub[c * NW] = lane == 0 ? D : I;
                    if (c == 0) edge |= (uint32_t)(vote(Mp >= 0) & 1ull) & w0;
                    if (c == C - 1) edge |= (uint32_t)(vote(Mp >= 0) >> 63) & wlast;
                    if (em & (1u << c)) mend = __builtin_amdgcn_readlane(M, ltp);
                }
            }
            if (mend >= U * t) atomicMin(X + L::EV, (s << 2) | kEvEnd);  // uniform conditions and values
            if (edge) atomicMin(X + L::EV, (s << 2) | kEvEdge);
            // this step's ring entries and edges before the progress word
            __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup");
            if (lane == 0) __hip_atomic_store(X + L::PG + w, s, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
            return 0;
        };

        int r = (p == t && ext0 == t) ? 1 : 0;  // 1: score 0 (identical sequences)
        while (r == 0) {
            r = step(std::integral_constant<int, 1>{});
            if (r == 0) r = step(std::integral_constant<int, 0>{});
        }
        int score = r;
        {
            __syncthreads();  // every wave has stopped: the first event is final
            const int ev = __builtin_amdgcn_readfirstlane(X[L::EV]);
            const int kind = ev & 3;
            if (p == t && ext0 == t) score = 0;
            else if ((ev >> 2) <= s) score = kind == kEvEnd ? (ev >> 2) : kind == kEvInternal ? TSU_EINTERNAL : TSU_EKOVERFLOW;
            // else: no event at or below the score this wave reached, so it
            // stopped on a score limit (r holds that status)
        }
        if constexpr (BT) {
            __syncthreads();  // every wave's records written
            if (w == 0) {     // wave 0 walks (bt_walk_wave: a block of records per round trip)
                int st = !lds_at_zero ? TSU_EINTERNAL : score < 0 ? score : 0;
                int64_t out_off = 0;
                int32_t out_n = 0;
                if (st == 0) {
                    auto record = [&](int ss, int kk, uint32_t& rr) __attribute__((always_inline)) -> bool {
                        if (ss < 0 || ss > s) return false;
                        const int j = kk - kmin;
                        if (j < 0 || j >= W) return false;
                        rr = rec[((int64_t)ss * NC + (j >> 6)) * 64 + (j & 63)];
                        return true;
                    };
                    st = bt_walk_wave(record, score, p, t, 2, 3, 1, a, &out_off, &out_n);
                }
                if (lane == 0) {
                    a.run_off[pi] = out_off;
                    a.run_n[pi] = out_n;
                    a.scores[pi] = st ? st : score * a.g;
                }
            }
        } else {
            if (tid == 0) a.scores[pi] = !lds_at_zero ? TSU_EINTERNAL : score < 0 ? score : score * a.g;
        }
        __syncthreads();  // LDS reused by the next pair of this workgroup
    }
}

}  // namespace tsu
