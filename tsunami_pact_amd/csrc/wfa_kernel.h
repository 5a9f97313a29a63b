#pragma once
// wfa_kernels.hip -- gap-affine WaveFront Alignment kernels for MI355X (gfx950).
//
// One workgroup of NW wave64s owns one pair at a time.  Per score step s it
// computes the I, D and M wavefronts over the diagonal range [lo, hi], the
// 64 lanes of each wave owning 64 consecutive diagonals of a chunk and the
// waves striding over chunks.  The extend of each new M cell is fused into
// the same pass: the lane that produced M[s][k] extends it straight away by
// comparing 32 bases per iteration of the 2-bit packed sequences, which sit
// in LDS for the whole pair.
//
// Semantics (reference wfa_gpu.cu, which follows WFA2-lib):
//   sources M[s-x], M[s-o-e], I[s-e], D[s-e] and null steps   wfa_gpu.cu:347-422
//   range union lo/hi                                          wfa_gpu.cu:429-439
//   I/D exist iff one of their sources does                    wfa_gpu.cu:455-468
//   I = max(M_o[k-1], I_e[k-1]) + 1; D = max(M_o[k+1], D_e[k+1]);
//   M = max(D, M_x[k] + 1, I), NULL when h > t or v > p         wfa_gpu.cu:610-631
//   trim_ends of M, I, D                                       wfa_gpu.cu:633-719
//   extend + end test M[s][t-p] >= t                          wfa_gpu.cu:253-331
// Reads outside a wavefront's range return NULL (the reference's init_ends,
// wfa_gpu.cu:470-606): every ring slot holds NULL outside its current range,
// which is maintained by clearing stale and trimmed cells after each step.
//
// Design points that differ from the reference on purpose:
//   - penalties are divided by g = gcd(x, o, e), so no odd null steps run;
//     wavefront s' here is WFA2-lib's wavefront g*s';
//   - the rings (M: max(x,o+e)+1 slots, I/D: e+1 slots) live in LDS (or in a
//     per-workgroup HBM region when W is too wide), never per-pair in HBM;
//   - int16 or int32 offsets, chosen per launch from the sequence lengths;
//     out-of-text I values are clamped to t+1 and negative values are stored
//     as NULL, neither of which changes any in-bounds cell;
//   - a k-range wider than the capacity W is reported (TSU_EKOVERFLOW), never
//     written out of bounds.
//
// Backtrace (BT): each computed cell writes one 32-bit record (format and
// walk in wfa_bt.h), stored step after step with the step's lowest diagonal
// and first record kept per step; thread 0 walks them back from the end cell
// into run-length CIGARs.
//
// Checkpointed backtrace (CK, with BT): records for every score take
// O(score * W) words per pair.  Instead, the forward pass records nothing
// and saves the ring (every slot's live range and limits) after every
// ck_every-th score; the backward pass then restores the snapshot below the
// walk's current score, recomputes that window of ck_every scores with
// records, and lets the walk continue through it (checkpointing as in
// Grice, Hughey & Speck, "Reduced space sequence alignment", CABIOS 1997).
// The recomputed cells are the forward pass's cells, so the CIGAR is the one
// full records give, for O((score / ck_every + ck_every) * W) words and
// about twice the score pass's work.
#include <hip/hip_runtime.h>
#include <stdint.h>

#include <utility>

#include "tsu_internal.h"
#include "wfa_bt.h"

namespace tsu {

template <typename OffT> struct Null;
template <> struct Null<int16_t> { static constexpr int32_t v = -16384; };
template <> struct Null<int32_t> { static constexpr int32_t v = -1073741824; };

constexpr int kSmemInts = 16;  // misc uniform values in LDS

// Word reader of one 2-bit packed sequence (16 bases per word, LSB first).
// G false: the copy staged in LDS, with two zero words of padding after the
// last base, so index j+2 is always in bounds.  G true: the packed words in
// HBM (sequences too long for LDS), read through the caches; words past the
// sequence's own read as 0, so no read leaves it.  Bases past the end never
// decide a result: every comparison is cut at the shorter remaining length.
template <bool G>
struct SeqWords {
    const uint32_t* w;
    int n;  // words of the sequence (G only)
    __device__ __forceinline__ uint32_t operator[](int j) const {
        if constexpr (G) return j < n ? w[j] : 0u;
        else return w[j];
    }
};

// Longest common prefix of pattern[v..] and text[h..], at most lim bases.
template <bool G>
__device__ __forceinline__ int lcp_packed(SeqWords<G> P, SeqWords<G> T, int v, int h, int lim) {
    int n = 0;
    while (n < lim) {
        const int a = v + n, b = h + n;
        const int ja = a >> 4, sa = (a & 15) << 1;
        const int jb = b >> 4, sb = (b & 15) << 1;
        const uint32_t p0 = P[ja], p1 = P[ja + 1], p2 = P[ja + 2];
        const uint32_t t0 = T[jb], t1 = T[jb + 1], t2 = T[jb + 2];
        const uint32_t xl = __builtin_amdgcn_alignbit(p1, p0, sa) ^ __builtin_amdgcn_alignbit(t1, t0, sb);
        if (xl) { n += __builtin_ctz(xl) >> 1; break; }
        const uint32_t xh = __builtin_amdgcn_alignbit(p2, p1, sa) ^ __builtin_amdgcn_alignbit(t2, t1, sb);
        if (xh) { n += 16 + (__builtin_ctz(xh) >> 1); break; }
        n += 32;
    }
    return n < lim ? n : lim;
}

__device__ __forceinline__ int uni(int v) { return __builtin_amdgcn_readfirstlane(v); }

// Entries per ring slot: W diagonals, 2 NULL guards and one write-only cell
// (index W+2) that idle lanes store to, padded to 16 bytes.
template <typename OffT>
__host__ __device__ __forceinline__ int ring_stride(int W) {
    constexpr int per16 = 16 / (int)sizeof(OffT);
    return (W + 3 + per16 - 1) / per16 * per16;
}

// LDS bytes before the ring: two sequences, slot limits, reduction area and
// misc values, rounded up to 16 bytes.
__host__ __device__ __forceinline__ int64_t lds_head_bytes(int seq_words, int nslot, int nw) {
    (void)nw;
    const int64_t b = 4ll * (2 * seq_words + 2 * (nslot + 1) + kSmemInts);
    return (b + 15) / 16 * 16;
}

constexpr int cmax(int a, int b) { return a > b ? a : b; }

// Penalty specialisation: XS > 0 fixes the scaled penalties (x, o, e) =
// (XS, OS, ES) at compile time; XS == 0 reads them from the arguments.
template <int XS, int OS, int ES>
struct Pen {
    static constexpr bool fixed = XS > 0;
    static constexpr int rm = fixed ? cmax(XS, OS + ES) + 1 : 0;  // M ring depth
    static constexpr int ri = fixed ? ES + 1 : 0;                  // I/D ring depth
};

template <int NW, typename OffT, bool LDS_RING, bool BT, bool QUEUE, int XS, int OS, int ES, bool CK = false,
          bool GSEQ = false>
__global__ __launch_bounds__(NW * 64) void wfa_kernel(KArgs a) {
    static_assert(!CK || (BT && QUEUE), "checkpointed backtrace runs in persistent backtrace launches");
    static_assert(!GSEQ || sizeof(OffT) == 4, "sequences read from HBM are longer than int16 offsets hold");
    using PP = Pen<XS, OS, ES>;
    extern __shared__ __align__(16) uint32_t smem[];
    constexpr int NT = NW * 64;
    constexpr int32_t NUL = Null<OffT>::v;
    const int tid = threadIdx.x;
    const int lane = tid & 63;
    const int wave = uni(tid >> 6);

    const int rm = PP::fixed ? PP::rm : a.rm;
    const int ri = PP::fixed ? PP::ri : a.ri;
    const int nslot = rm + 2 * ri;      // M slots, I slots, D slots
    const int null_slot = nslot;        // one more slot that always holds NULL
    uint32_t* sP = smem;
    uint32_t* sT = smem + a.seq_words;
    int2* sLim = reinterpret_cast<int2*>(smem + 2 * a.seq_words);  // (lo, hi) per slot (runtime penalties)
    int32_t* sMisc = reinterpret_cast<int32_t*>(sLim + nslot + 1);  // kSmemInts
    const int stride = ring_stride<OffT>(a.W);
    OffT* ring;
    if constexpr (LDS_RING) {
        ring = reinterpret_cast<OffT*>(reinterpret_cast<uint8_t*>(smem) + lds_head_bytes(a.seq_words, nslot, NW));
    } else {
        ring = reinterpret_cast<OffT*>(a.ring_g + (int64_t)blockIdx.x * a.ring_stride);
    }
    uint32_t* rec = nullptr;
    int32_t* stepLo = nullptr;
    int64_t* stepOff = nullptr;
    if constexpr (BT) {
        rec = a.rec_g + (int64_t)blockIdx.x * a.rec_cap;
        stepLo = a.step_lo_g + (int64_t)blockIdx.x * a.step_cap;
        stepOff = a.step_off_g + (int64_t)blockIdx.x * a.step_cap;
    }
    uint8_t* ckb = nullptr;
    if constexpr (CK) ckb = a.ck_g + (int64_t)blockIdx.x * a.ck_stride;

    const int x = PP::fixed ? XS : a.x;
    const int o = PP::fixed ? OS : a.o;
    const int e = PP::fixed ? ES : a.e;
    const int oe = o + e;
    for (int iter = 0;; ++iter) {
        int idx;
        if constexpr (QUEUE) {
            if (tid == 0) sMisc[0] = atomicAdd(a.queue, 1);
            __syncthreads();
            idx = uni(sMisc[0]);
            __syncthreads();
        } else {
            if (iter > 0) break;
            idx = blockIdx.x;
        }
        if (idx >= a.n) break;
        const int pi = a.order ? uni(a.order[idx]) : idx;
        const KPair pr = a.pairs[pi];
        const int p = uni(pr.p_len), t = uni(pr.t_len);
        const int pw = (p + 15) >> 4, tw = (t + 15) >> 4;

        // ---- stage sequences (or read them in place), clear the ring --------
        SeqWords<GSEQ> P{sP, pw}, T{sT, tw};
        if constexpr (GSEQ) {
            P.w = a.words + pr.p_off;
            T.w = a.words + pr.t_off;
        } else {
            for (int i = tid; i < a.seq_words; i += NT) {
                sP[i] = i < pw ? a.words[pr.p_off + i] : 0u;
                sT[i] = i < tw ? a.words[pr.t_off + i] : 0u;
            }
        }
        auto fill_ring = [&]() __attribute__((always_inline)) {
            // slots are padded to a multiple of 16 bytes: fill with uint4
            const int64_t nvec = (int64_t)(nslot + 1) * stride * (int64_t)sizeof(OffT) / 16;
            const uint32_t nw32 = sizeof(OffT) == 2 ? ((uint32_t)(uint16_t)NUL | ((uint32_t)(uint16_t)NUL << 16))
                                                    : (uint32_t)NUL;
            const uint4 fill = make_uint4(nw32, nw32, nw32, nw32);
            uint4* r4 = reinterpret_cast<uint4*>(ring);
            for (int64_t i = tid; i < nvec; i += NT) r4[i] = fill;
        };
        fill_ring();
        for (int i = tid; i <= nslot; i += NT) sLim[i] = make_int2(1, 0);
        __syncthreads();

        // slot limits (lo, hi) in LDS, lo > hi = null; the slot of score s is
        // s mod depth (compile-time depths when the penalties are fixed)
        auto lim_get = [&](int slot) __attribute__((always_inline)) -> int2 {
            const int2 v = sLim[slot];
            return make_int2(uni(v.x), uni(v.y));
        };
        auto lim_set = [&](int slot, int2 v) __attribute__((always_inline)) {
            if (tid == 0) sLim[slot] = v;
        };

        // diagonal window [kmin, kmax] of the capacity; slot index = k - kmin + 1
        int kmin = max(-(a.W / 2), -p - 1);
        int kmax = kmin + a.W - 1;
        if (kmax > t + 1) { kmax = t + 1; kmin = max(kmax - a.W + 1, -p - 1); }
        const int tp = t - p;
        // scaled upper bound of the optimum (all mismatches plus one gap)
        const int64_t bound_u = (int64_t)x * min(p, t) + (p != t ? (int64_t)o + (int64_t)e * abs(p - t) : 0);
        const int s_bound = (int)min<int64_t>(bound_u, 0x3fffffff);

        int status = 0;
        int score = -1;
        int64_t rec_next = 0;
        // CK: scores (sb, s] of the current window have records; the forward
        // pass records nothing
        int sb = 0;
        bool rec_on = !CK;

        // ---- score 0: M[0][0] = 0, extended --------------------------------
        if (tid == 0) {
            const int ext = lcp_packed(P, T, 0, 0, min(p, t));
            ring[0 * stride + (0 - kmin + 1)] = (OffT)ext;
            sMisc[1] = (p == t && ext == t) ? 1 : 0;
            if constexpr (BT && !CK) { rec[0] = 0u; stepLo[0] = 0; stepOff[0] = 0; }
        }
        lim_set(0, make_int2(0, 0));
        if constexpr (BT && !CK) rec_next = 1;
        __syncthreads();
        if (uni(sMisc[1])) score = 0;

        auto clear_range = [&](OffT* sl, int from, int to) __attribute__((always_inline)) {
            for (int k = from + tid; k <= to; k += NT) sl[k - kmin + 1] = (OffT)NUL;
        };
        // NULL every cell of [alo, ahi] outside [klo, khi] (klo > khi: all)
        auto clear_except = [&](OffT* sl, int alo, int ahi, int klo, int khi) __attribute__((always_inline)) {
            if (alo > ahi) return;
            if (klo > khi) { clear_range(sl, alo, ahi); return; }
            if (alo < klo) clear_range(sl, alo, min(ahi, klo - 1));
            if (ahi > khi) clear_range(sl, max(alo, khi + 1), ahi);
        };
        auto in_bounds = [&](int32_t v, int k) __attribute__((always_inline)) -> bool {
            return (uint32_t)v <= (uint32_t)t && (uint32_t)(v - k) <= (uint32_t)p;
        };

        int s = 0;
        int rcM = 0, rcI = 0;  // s mod rm, s mod ri
        // One score step.  Returns false when the pair is finished.
        auto step = [&](int) __attribute__((always_inline)) -> bool {
            ++s;
            if (s > s_bound) { status = TSU_EINTERNAL; return false; }
            rcM = (rcM + 1 == rm) ? 0 : rcM + 1;
            rcI = (rcI + 1 == ri) ? 0 : rcI + 1;
            // slots of scores s-x, s-o-e, s-e; a negative score maps to a slot
            // of a future score, which still holds the null initial range
            int jx = rcM - x; if (jx < 0) jx += rm;
            int jo = rcM - oe; if (jo < 0) jo += rm;
            int je = rcI - e; if (je < 0) je += ri;
            const int je_i = rm + je, je_d = rm + ri + je;
            const int slM = rcM, slI = rm + rcI, slD = rm + ri + rcI;
            const int2 Lx = lim_get(jx), Lo = lim_get(jo), Li = lim_get(je_i), Ld = lim_get(je_d);
            const int2 Om = lim_get(slM), Oi = lim_get(slI), Od = lim_get(slD);
            auto push = [&](int2 nm, int2 ni, int2 nd) __attribute__((always_inline)) {
                lim_set(slM, nm); lim_set(slI, ni); lim_set(slD, nd);
            };
            const bool nx = Lx.x > Lx.y, no = Lo.x > Lo.y, ni = Li.x > Li.y, nd = Ld.x > Ld.y;
            OffT* Mt = ring + slM * stride;
            OffT* It = ring + slI * stride;
            OffT* Dt = ring + slD * stride;
            if (nx && no && ni && nd) {
                // null step (wfa_gpu.cu:417-422): all three outputs are null
                clear_except(Mt, Om.x, Om.y, 1, 0);
                clear_except(It, Oi.x, Oi.y, 1, 0);
                clear_except(Dt, Od.x, Od.y, 1, 0);
                if constexpr (BT) {
                    if (rec_on) {
                        if (s - sb >= a.step_cap) { status = TSU_ECIGAR; return false; }
                        if (tid == 0) { stepLo[s - sb] = 0; stepOff[s - sb] = rec_next; }
                    }
                }
                __syncthreads();
                push(make_int2(1, 0), make_int2(1, 0), make_int2(1, 0));
                __syncthreads();
                return true;
            }
            int lo = 0x3fffffff, hi = -0x3fffffff;
            if (!nx) { lo = Lx.x; hi = Lx.y; }
            if (!no) { lo = min(lo, Lo.x - 1); hi = max(hi, Lo.y + 1); }
            if (!ni) { lo = min(lo, Li.x + 1); hi = max(hi, Li.y + 1); }
            if (!nd) { lo = min(lo, Ld.x - 1); hi = max(hi, Ld.y - 1); }
            if (lo < kmin || hi > kmax) { status = TSU_EKOVERFLOW; return false; }
            const int width = hi - lo + 1;
            if constexpr (BT) {
                if (rec_on && (s - sb >= a.step_cap || rec_next + width > a.rec_cap)) { status = TSU_ECIGAR; return false; }
            }
            const bool hasI = !no || !ni;
            const bool hasD = !no || !nd;
            const OffT* Mx = ring + (nx ? null_slot : jx) * stride;
            const OffT* Mo = ring + (no ? null_slot : jo) * stride;
            const OffT* Ie = ring + (ni ? null_slot : je_i) * stride;
            const OffT* De = ring + (nd ? null_slot : je_d) * stride;
            uint32_t* recs = nullptr;
            if constexpr (BT) recs = rec + rec_next;

            // ---- compute + extend, 64 diagonals per wave per chunk ----------
            // The pass spans [clo, chi]: [lo, hi] plus whatever the target
            // slots held before (usually inside [lo, hi]); cells outside
            // [lo, hi] are written NULL, so no separate clear is needed.
            int clo = lo, chi = hi;
            if (Om.x <= Om.y) { clo = min(clo, Om.x); chi = max(chi, Om.y); }
            if (Oi.x <= Oi.y) { clo = min(clo, Oi.x); chi = max(chi, Oi.y); }
            if (Od.x <= Od.y) { clo = min(clo, Od.x); chi = max(chi, Od.y); }
            for (int kb = clo + wave * 64; kb <= chi; kb += NT) {
                const int k = kb + lane;
                const bool act = k <= chi;
                const bool inr = k >= lo && k <= hi;
                const int j = min(k, chi) - kmin + 1;  // idle lanes read a valid cell
                const int32_t mol = Mo[j - 1];
                const int32_t mor = Mo[j + 1];
                const int32_t iel = Ie[j - 1];
                const int32_t der = De[j + 1];
                const int32_t mx = Mx[j];
                const int32_t ipre = max(mol, iel);
                const int32_t I = ipre >= 0 ? min(ipre + 1, t + 1) : NUL;
                const int32_t D = max(mor, der);
                const int32_t X = mx >= 0 ? mx + 1 : NUL;
                int32_t M = max(D, max(X, I));
                if ((uint32_t)M > (uint32_t)t || (uint32_t)(M - k) > (uint32_t)p || !inr) M = NUL;
                if constexpr (BT) {
                    if (rec_on && inr) {
                        uint32_t r = 0;
                        if (M >= 0) {
                            const uint32_t mw = (X >= D && X >= I) ? 1u : (D >= I ? 2u : 3u);
                            r = ((uint32_t)M << 4) | mw;
                        }
                        if (iel >= mol) r |= 4u;
                        if (der >= mor) r |= 8u;
                        recs[k - lo] = r;
                    }
                }
                // extend in uniform rounds of 32 bases; lanes that are done
                // keep their count and read a harmless word
                {
                    const int v = M - k;
                    const int lim = (act && M >= 0) ? min(p - v, t - M) : 0;
                    bool go = lim > 0;
                    int n = 0;
                    while (__ballot(go)) {
                        const int pa = go ? v + n : 0;
                        const int pb = go ? M + n : 0;
                        const int ja = pa >> 4, sa = (pa & 15) << 1;
                        const int jb = pb >> 4, sb = (pb & 15) << 1;
                        const uint32_t p0 = P[ja], p1 = P[ja + 1], p2 = P[ja + 2];
                        const uint32_t t0 = T[jb], t1 = T[jb + 1], t2 = T[jb + 2];
                        const uint32_t xl = __builtin_amdgcn_alignbit(p1, p0, sa) ^ __builtin_amdgcn_alignbit(t1, t0, sb);
                        const uint32_t xh = __builtin_amdgcn_alignbit(p2, p1, sa) ^ __builtin_amdgcn_alignbit(t2, t1, sb);
                        // first differing bit of the 64-bit window (64 when equal)
                        const uint64_t xx = ((uint64_t)xh << 32) | xl;
                        const int m = (xx ? (int)__builtin_ctzll(xx) : 64) >> 1;
                        n = go ? n + m : n;
                        go = go && m == 32 && n < lim;
                    }
                    M += min(n, lim);
                }
                const int js = act ? j : a.W + 2;  // idle lanes: write-only cell
                Mt[js] = (OffT)M;
                It[js] = (OffT)(inr ? I : NUL);
                Dt[js] = (OffT)(inr ? D : NUL);
                // flags read after the barrier: in-bounds state of the six
                // boundary cells (lo, hi of M, I, D) and the end test at t-p
                if (k == lo || k == hi) {
                    const int side = k == hi ? 1 : 0;
                    if (k == lo) sMisc[2] = M >= 0, sMisc[4] = in_bounds(I, k), sMisc[6] = in_bounds(D, k);
                    if (side) sMisc[3] = M >= 0, sMisc[5] = in_bounds(I, k), sMisc[7] = in_bounds(D, k);
                }
                if (k == tp && inr) sMisc[8] = M >= t;
            }
            __syncthreads();  // every cell of step s is written

            // ---- trim ends (wfa_gpu.cu:633-719) ----------------------------
            // common case: the six boundary cells are in bounds -> no trim
            int tloM = lo, thiM = hi, tloI = hasI ? lo : 1, thiI = hasI ? hi : 0, tloD = hasD ? lo : 1, thiD = hasD ? hi : 0;
            {
                // bits 0-1: M at lo/hi, 2-3: I, 4-5: D out of bounds
                uint32_t bb = 0;
                if (!uni(sMisc[2])) bb |= 1u;
                if (!uni(sMisc[3])) bb |= 2u;
                if (hasI && !uni(sMisc[4])) bb |= 4u;
                if (hasI && !uni(sMisc[5])) bb |= 8u;
                if (hasD && !uni(sMisc[6])) bb |= 16u;
                if (hasD && !uni(sMisc[7])) bb |= 32u;
                if (bb) {
                    // first in-bounds cell from lo upwards, then last from hi
                    // downwards; trimmed ends are short, so each walk usually
                    // takes one 64-cell round (every wave computes the same)
                    auto scan = [&](const OffT* sl, int& f, int& l) __attribute__((always_inline)) {
                        f = 1; l = 0;
                        int kb = lo;
                        for (; kb <= hi; kb += 64) {
                            const int k = kb + lane;
                            const bool in = k <= hi && in_bounds((int32_t)sl[min(k, hi) - kmin + 1], k);
                            const uint64_t b = __ballot(in);
                            if (b) { f = kb + (int)__builtin_ctzll(b); break; }
                        }
                        if (kb > hi) return;  // no cell in bounds: empty range
                        // cell f is in bounds, so this walk stops at f at the latest
                        for (int kt = hi;; kt -= 64) {
                            const int k = kt - lane;
                            const bool in = k >= f && in_bounds((int32_t)sl[max(k, f) - kmin + 1], k);
                            const uint64_t b = __ballot(in);
                            if (b) { l = kt - (int)__builtin_ctzll(b); break; }
                        }
                    };
                    if (bb & 3u) scan(Mt, tloM, thiM);
                    if (bb & 12u) scan(It, tloI, thiI);
                    if (bb & 48u) scan(Dt, tloD, thiD);
                }
            }
            // end test: M[s][t-p] >= t (wfa_gpu.cu:288-331)
            bool fin = false;
            if (tp >= tloM && tp <= thiM) fin = uni(sMisc[8]) != 0;
            // cells outside the kept ranges return to NULL (M cells trimmed
            // away are NULL already)
            // (the pass left NULL outside [lo, hi]; I/D cells trimmed away
            // inside it still hold out-of-bounds values)
            if (tloI != lo || thiI != hi) clear_except(It, lo, hi, tloI, thiI);
            if (tloD != lo || thiD != hi) clear_except(Dt, lo, hi, tloD, thiD);
            push(make_int2(tloM, thiM), make_int2(tloI, thiI), make_int2(tloD, thiD));
            if constexpr (BT) {
                if (rec_on) {
                    if (tid == 0) { stepLo[s - sb] = lo; stepOff[s - sb] = rec_next; }
                    rec_next += width;
                }
            }
            __syncthreads();
            if (fin) { score = s; return false; }
            return true;
        };

        // CK snapshot c: limits of every slot, then W cells per slot (the
        // live range of each; the ring is NULL elsewhere)
        auto ck_save = [&](int c) __attribute__((always_inline)) {
            uint8_t* base = ckb + (int64_t)c * a.ck_bytes;
            int2* cl = reinterpret_cast<int2*>(base);
            OffT* cc = reinterpret_cast<OffT*>(base + a.ck_head);
            for (int i = 0; i < nslot; ++i) {
                const int2 L = lim_get(i);
                if (tid == 0) cl[i] = L;
                for (int k = L.x + tid; k <= L.y; k += NT) cc[(int64_t)i * a.W + (k - kmin)] = ring[i * stride + (k - kmin + 1)];
            }
            __syncthreads();  // the next step overwrites the oldest slots
        };
        auto ck_load = [&](int c) __attribute__((always_inline)) {
            const uint8_t* base = ckb + (int64_t)c * a.ck_bytes;
            const int2* cl = reinterpret_cast<const int2*>(base);
            const OffT* cc = reinterpret_cast<const OffT*>(base + a.ck_head);
            __syncthreads();  // the previous window's readers are done
            fill_ring();
            __syncthreads();
            for (int i = 0; i < nslot; ++i) {
                const int2 L = cl[i];
                const int lo = uni(L.x), hi = uni(L.y);
                if (tid == 0) sLim[i] = make_int2(lo, hi);
                for (int k = lo + tid; k <= hi; k += NT) ring[i * stride + (k - kmin + 1)] = cc[(int64_t)i * a.W + (k - kmin)];
            }
            __syncthreads();
        };

        if (score < 0) {
            if constexpr (CK) {
                ck_save(0);
                while (step(0)) {
                    if (s % a.ck_every == 0) {
                        if (s / a.ck_every >= a.ck_cap) { status = TSU_ECIGAR; break; }
                        ck_save(s / a.ck_every);
                    }
                }
            } else {
                while (step(0)) {}
            }
        }

        // ---- outputs --------------------------------------------------------
        if constexpr (CK) {
            // backward pass: window w covers scores (w*K, min((w+1)*K, score)]
            int32_t st = status;
            int64_t out_off = 0;
            int32_t out_n = 0;
            BtWalk wk;
            if (st == 0 && tid == 0) st = bt_begin(wk, score, p, t, x, o, e, a);
            if (tid == 0) sMisc[9] = st;
            __syncthreads();
            st = uni(sMisc[9]);
            const int K = a.ck_every;
            int w = score > 0 ? (score - 1) / K : 0;
            while (st == 0 && score > 0) {
                ck_load(w);
                s = w * K; sb = s;
                rcM = s % rm; rcI = s % ri;
                rec_next = 0;
                rec_on = true;
                const int s_end = min((w + 1) * K, score);
                while (s < s_end && step(0)) {}
                if (status) { st = status; break; }
                if (tid == 0) {
                    auto record = [&](int ss, int kk, uint32_t& r) __attribute__((always_inline)) -> bool {
                        if (ss <= sb || ss > s) return false;
                        const int l0 = stepLo[ss - sb];
                        const int64_t o0 = stepOff[ss - sb];
                        const int64_t o1 = (ss < s) ? stepOff[ss - sb + 1] : rec_next;
                        const int64_t at = o0 + (kk - l0);
                        if (kk < l0 || at >= o1) return false;
                        r = rec[at];
                        return true;
                    };
                    int rr = bt_run(wk, record, sb);
                    // 1: continue in the window below the walk's score
                    sMisc[9] = rr < 0 ? rr : 0;
                    sMisc[10] = rr == 1 ? (wk.ss - 1) / K : -1;
                }
                __syncthreads();
                st = uni(sMisc[9]);
                const int nw_ = uni(sMisc[10]);
                __syncthreads();
                if (st || nw_ < 0) break;
                w = nw_;
            }
            if (tid == 0) {
                if (st == 0) st = bt_end(wk, &out_off, &out_n);
                a.run_off[pi] = out_off;
                a.run_n[pi] = out_n;
                a.scores[pi] = st ? st : score * a.g;
            }
        } else if constexpr (BT) {
            if (tid == 0) {
                int32_t st = status;
                int64_t out_off = 0;
                int32_t out_n = 0;
                if (st == 0) {
                    auto record = [&](int ss, int kk, uint32_t& r) __attribute__((always_inline)) -> bool {
                        if (ss < 0 || ss > s) return false;
                        const int l0 = stepLo[ss];
                        const int64_t o0 = stepOff[ss];
                        const int64_t o1 = (ss < s) ? stepOff[ss + 1] : rec_next;
                        const int64_t at = o0 + (kk - l0);
                        if (kk < l0 || at >= o1) return false;
                        r = rec[at];
                        return true;
                    };
                    st = bt_walk(record, score, p, t, x, o, e, a, &out_off, &out_n);
                }
                a.run_off[pi] = out_off;
                a.run_n[pi] = out_n;
                a.scores[pi] = st ? st : score * a.g;
            }
        } else {
            if (tid == 0) a.scores[pi] = status ? status : score * a.g;
        }
        __syncthreads();  // LDS reused by the next pair of this workgroup
    }
}

typedef void (*KernelFn)(KArgs);

template <int NW, typename OffT, bool LDS, bool BT, bool Q, bool G = false>
inline KernelFn fn(bool fixed_242) {
    // scaled penalties (2, 3, 1) are the reference's 4/6/2 divided by 2
    return fixed_242 ? wfa_kernel<NW, OffT, LDS, BT, Q, 2, 3, 1, false, G> : wfa_kernel<NW, OffT, LDS, BT, Q, 0, 0, 0, false, G>;
}

template <int NW, typename OffT, bool LDS, bool G = false>
inline KernelFn fn_ck(bool fixed_242) {
    return fixed_242 ? wfa_kernel<NW, OffT, LDS, true, true, 2, 3, 1, true, G>
                     : wfa_kernel<NW, OffT, LDS, true, true, 0, 0, 0, true, G>;
}

// Instances reading the sequences from HBM (int32 offsets only).
template <int NW>
KernelFn pick_gseq(const Shape& sh) {
    const bool f = sh.fixed_242;
    if (sh.off16) return nullptr;
    if (sh.ck) {
        if (!sh.queue || !sh.bt || NW < 4) return nullptr;
        return sh.lds_ring ? fn_ck<NW, int32_t, true, true>(f) : fn_ck<NW, int32_t, false, true>(f);
    }
    if (!sh.queue) {
        if (!sh.lds_ring || sh.bt) return nullptr;
        return fn<NW, int32_t, true, false, false, true>(f);
    }
    if (sh.lds_ring) return sh.bt ? fn<NW, int32_t, true, true, true, true>(f) : fn<NW, int32_t, true, false, true, true>(f);
    return sh.bt ? fn<NW, int32_t, false, true, true, true>(f) : fn<NW, int32_t, false, false, true, true>(f);
}

template <int NW>
KernelFn pick_nw_impl(const Shape& sh) {
    const bool f = sh.fixed_242;
    if (sh.gseq) return pick_gseq<NW>(sh);
    if (sh.ck) {
        // checkpointed backtrace: persistent launches of 4 or more waves
        if (!sh.queue || !sh.bt || NW < 4) return nullptr;
        if (sh.off16) return sh.lds_ring ? fn_ck<NW, int16_t, true>(f) : fn_ck<NW, int16_t, false>(f);
        return sh.lds_ring ? fn_ck<NW, int32_t, true>(f) : fn_ck<NW, int32_t, false>(f);
    }
    if (!sh.queue) {
        // plain launches: ring in LDS, no backtrace
        if (!sh.lds_ring || sh.bt) return nullptr;
        return sh.off16 ? fn<NW, int16_t, true, false, false>(f) : fn<NW, int32_t, true, false, false>(f);
    }
    if (sh.off16) {
        if (sh.lds_ring) return sh.bt ? fn<NW, int16_t, true, true, true>(f) : fn<NW, int16_t, true, false, true>(f);
        return sh.bt ? fn<NW, int16_t, false, true, true>(f) : fn<NW, int16_t, false, false, true>(f);
    }
    if (sh.lds_ring) return sh.bt ? fn<NW, int32_t, true, true, true>(f) : fn<NW, int32_t, true, false, true>(f);
    return sh.bt ? fn<NW, int32_t, false, true, true>(f) : fn<NW, int32_t, false, false, true>(f);
}


}  // namespace tsu
