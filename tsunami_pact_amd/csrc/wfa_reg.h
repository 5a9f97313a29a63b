// wfa_reg.h -- register-resident WFA kernel for short and medium pairs.
//
// One wave64 per pair.  Diagonal k lives in lane (k - kmin) % 64 of chunk
// (k - kmin) / 64 for the whole alignment: the M ring (max(x, o+e)+1 scores)
// and the I / D wavefronts (updated in place, which needs e = 1 after gcd
// scaling) are arrays of VGPRs indexed by compile-time chunk and slot.  The
// k-1 / k+1 neighbours come from DPP wave shifts (wave_shr:1 / wave_shl:1)
// whose boundary lane is taken from the adjacent chunk, rotated into place
// by a DPP wave rotate (or by v_readlane for the in-place I).  The score loop
// is unrolled over the M ring's depth so every slot index is a constant.
//
// All diagonals of the C*64-wide window are computed implicitly: a chunk is
// computed when it meets the wavefront limits WFA2-lib's range [lo, hi]
// would have (a scalar recurrence per score), and a diagonal outside that
// range has only NULL sources, so it comes out NULL exactly as the
// range-limited computation leaves it.  Trim-ends (wfa_gpu.cu:633-719)
// reduce to: NULL every I/D cell below the first and above the last
// in-bounds cell of that wavefront; M cells outside the bounds are NULL
// already.  The common case (no out-of-bounds I/D value at all) is detected
// from one ballot per chunk of the maximum before the bounds test.
//
// Sequences sit in LDS as interleaved 2-bit packed words (pattern word i at
// 2i, text word i at 2i+1), so both are addressed from LDS offset 0.
//
// Recurrences, bounds test, end test and extend are those of the LDS kernel
// (wfa_gpu.cu:253-331,610-631).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "wfa_kernel.h"

namespace tsu {

constexpr int32_t kRegNull = -1073741824;

// Step masks: for the M, I and D wavefronts of one score together, bit c of
//   bits 0-7    chunk c holds a non-NULL value
//   bits 8-15   lane 0 of chunk c holds one (it feeds D of chunk c-1)
//   bits 16-23  lane 63 of chunk c holds one (it feeds I of chunk c+1)

__device__ __forceinline__ uint64_t vote(bool b) { return __builtin_amdgcn_ballot_w64(b); }

__device__ __forceinline__ int32_t dpp_shr1(int32_t fill_lane0, int32_t v) {
    // lane l receives lane l-1; lane 0 keeps fill_lane0
    return __builtin_amdgcn_update_dpp(fill_lane0, v, 0x138, 0xF, 0xF, false);
}
__device__ __forceinline__ int32_t dpp_shl1(int32_t fill_lane63, int32_t v) {
    // lane l receives lane l+1; lane 63 keeps fill_lane63
    return __builtin_amdgcn_update_dpp(fill_lane63, v, 0x130, 0xF, 0xF, false);
}
__device__ __forceinline__ int32_t dpp_ror1(int32_t v) {
    // lane l receives lane l-1 mod 64 (lane 0 gets lane 63); every lane has
    // a source, so no old value is kept
    return __builtin_amdgcn_mov_dpp(v, 0x13C, 0xF, 0xF, false);
}
__device__ __forceinline__ int32_t dpp_rol1(int32_t v) {
    // lane l receives lane l+1 mod 64 (lane 63 gets lane 0)
    return __builtin_amdgcn_mov_dpp(v, 0x134, 0xF, 0xF, false);
}

// Compare 32 bases of pattern and text from the interleaved words at byte
// offsets bp / bt (word j at 8j, text words 4 bytes further), shifted by sa /
// sb bits; returns the number of equal bits (64 when all 32 bases match).
typedef __attribute__((address_space(3))) const uint32_t lds_cu32;
__device__ __forceinline__ int cmp32(uint32_t bp, uint32_t bt, uint32_t sa, uint32_t sb) {
    // The interleaved words start the kernel's dynamic LDS, which is LDS
    // address 0 (the kernel has no static LDS; it checks this and reports
    // TSU_EINTERNAL otherwise), so no base is added to the byte offsets.
    lds_cu32* pp = reinterpret_cast<lds_cu32*>((uintptr_t)bp);
    lds_cu32* tt = reinterpret_cast<lds_cu32*>((uintptr_t)bt);
    const uint32_t p0 = pp[0], p1 = pp[2], p2 = pp[4];
    const uint32_t t0 = tt[1], t1 = tt[3], t2 = tt[5];
    const uint32_t xl = __builtin_amdgcn_alignbit(p1, p0, sa) ^ __builtin_amdgcn_alignbit(t1, t0, sb);
    const uint32_t xh = __builtin_amdgcn_alignbit(p2, p1, sa) ^ __builtin_amdgcn_alignbit(t2, t1, sb);
    const uint64_t xx = ((uint64_t)xh << 32) | xl;
    return xx ? (int)__builtin_ctzll(xx) : 64;
}

// Extension on doubled offsets (2 x bases, the unit the register kernels
// keep offsets in): v2 = 2v, h2 = 2h, lim2 = 2 x the most bases allowed
// (>= 2).  Returns 2 x the bases pattern[v..] and text[h..] agree on, at
// most lim2.  16 bases first (most diagonals stop there), then rounds of
// 16 over the next words; a doubled offset is also the bit
// position of its base in the packed words, so it is the funnel shift
// (alignbit takes it modulo 32).  Reads stay below word (v+lim-1)/16 + 2 of
// each sequence, inside the two padding words.
__device__ __forceinline__ lds_cu32* pp2(uint32_t b) { return reinterpret_cast<lds_cu32*>((uintptr_t)b); }
__device__ __forceinline__ int extend2(uint32_t v2, uint32_t h2, int lim2) {
    const uint32_t bp = (v2 >> 5) << 3, bt = (h2 >> 5) << 3;
    lds_cu32* pp = reinterpret_cast<lds_cu32*>((uintptr_t)bp);
    lds_cu32* tt = reinterpret_cast<lds_cu32*>((uintptr_t)bt);
    const uint32_t p0 = pp[0], p1 = pp[2], t0 = tt[1], t1 = tt[3];
    const uint32_t ap = __builtin_amdgcn_alignbit(p1, p0, v2), at = __builtin_amdgcn_alignbit(t1, t0, h2);
    const uint32_t x = ap ^ at;
    // matched bits (2 per base), clamped to lim2 at once: 32 then means 16
    // bases matched and lim2 >= 32 (at lim2 == 32 the next round reads one
    // word more than needed, still below the bound above, and is clamped)
    int n2 = (int)min(x ? (uint32_t)__builtin_ctz(x) : 32u, (uint32_t)lim2);
    if (n2 == 32) {
        const uint32_t xh = __builtin_amdgcn_alignbit(pp[4], p1, v2) ^ __builtin_amdgcn_alignbit(tt[5], t1, h2);
        int m = xh ? (int)__builtin_ctz(xh) : 32;
        n2 += m;
        if (m == 32 && n2 < lim2) {
            // 16 bases per round from here (long runs are rare; fewer live
            // registers than a 32-base round)
            uint32_t bp2 = bp + 16, bt2 = bt + 16;  // words w+2 (w: the word of base v / h)
            do {
                const uint32_t y = __builtin_amdgcn_alignbit(pp2(bp2 + 8)[0], pp2(bp2)[0], v2) ^
                                   __builtin_amdgcn_alignbit(pp2(bt2 + 8)[1], pp2(bt2)[1], h2);
                m = y ? (int)__builtin_ctz(y) : 32;
                n2 += m;
                bp2 += 8;
                bt2 += 8;
            } while (m == 32 && n2 < lim2);
        }
        n2 = min(n2, lim2);
    }
    return n2 & ~1;  // lim2 is even
}

// The same on plain offsets: bases pattern[v..] and text[h..] agree on, at
// most lim (>= 1).
__device__ __forceinline__ int extend_len(uint32_t v, uint32_t h, int lim) {
    return extend2(v << 1, h << 1, lim << 1) >> 1;
}

// Longest common prefix of pattern[v..] and text[h..] (at most lim bases)
// over the interleaved layout; the same 32-bases-per-round compare as the
// extend in the kernel.
__device__ __forceinline__ int lcp_interleaved(const uint32_t* S, int v, int h, int lim) {
    const uint32_t* pp = S + 2 * (v >> 4);
    const uint32_t* tt = S + 2 * (h >> 4) + 1;
    const uint32_t sa = (uint32_t)v << 1, sb = (uint32_t)h << 1;
    const int lim2 = 2 * lim;
    int n2 = 0;
    for (;;) {
        const uint32_t xl = __builtin_amdgcn_alignbit(pp[2], pp[0], sa) ^ __builtin_amdgcn_alignbit(tt[2], tt[0], sb);
        const uint32_t xh = __builtin_amdgcn_alignbit(pp[4], pp[2], sa) ^ __builtin_amdgcn_alignbit(tt[4], tt[2], sb);
        const uint64_t xx = ((uint64_t)xh << 32) | xl;
        const int m = xx ? (int)__builtin_ctzll(xx) : 64;
        n2 += m;
        if (m != 64 || n2 >= lim2) break;
        pp += 4;
        tt += 4;
    }
    return min(n2 >> 1, lim);
}

// With PROBE the kernel stops after a.probe_cap scores: a pair that has not
// reached its end cell by then gets kProbeStopped, and a.probe_out[pair]
// receives the largest h + v over the valid M cells of its last RM scores
// (how far along both sequences that score took it).
template <int C, int XS, int OS, int ES, bool QUEUE, bool PROBE = false>
__global__ __launch_bounds__(64) void wfa_reg_kernel(KArgs a) {
    static_assert(ES == 1, "in-place I/D needs e == 1 after gcd scaling");
    static_assert(C >= 1 && C <= 8, "window of 64..512 diagonals");
    constexpr int RM = (XS > OS + ES ? XS : OS + ES) + 1;
    constexpr int32_t NUL = kRegNull;
    constexpr int W = 64 * C;
    extern __shared__ __align__(16) uint32_t smem[];
    const int lane = threadIdx.x;
    const int lane2 = 2 * lane;
    uint32_t* sS = smem;  // interleaved pattern / text words
    int32_t* sMisc = reinterpret_cast<int32_t*>(sS + 2 * a.seq_words);
    // the extend addresses the words from LDS address 0 (see cmp32)
    const bool lds_at_zero = (uint32_t)(uintptr_t)sS == 0u;

    for (int iter = 0;; ++iter) {
        int idx;
        if constexpr (QUEUE) {
            if (lane == 0) sMisc[0] = atomicAdd(a.queue, 1);
            __syncthreads();
            idx = __builtin_amdgcn_readfirstlane(sMisc[0]);
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
        for (int i = lane; i < a.seq_words; i += 64) {
            sS[2 * i] = i < pw ? a.words[pr.p_off + i] : 0u;
            sS[2 * i + 1] = i < tw ? a.words[pr.t_off + i] : 0u;
        }
        __syncthreads();

        int kmin = max(-(W / 2), -p - 1);
        int kmax = kmin + W - 1;
        if (kmax > t + 1) { kmax = t + 1; kmin = max(kmax - W + 1, -p - 1); }
        const int tp = t - p;
        // chunk and lane of the end diagonal t - p (a chunk index >= C or < 0
        // never matches: the pair then ends by overflow at the window edge)
        const int ctp = (tp - kmin) >= 0 ? (tp - kmin) >> 6 : -1;
        const int ltp = (tp - kmin) & 63;
        const int64_t bound_u = (int64_t)XS * min(p, t) + (p != t ? (int64_t)OS + (int64_t)ES * abs(p - t) : 0);
        const int s_bound = (int)min<int64_t>(bound_u, 0x3fffffff);

        int32_t Mr[RM][C], Ir[C], Dr[C];
#pragma unroll
        for (int c = 0; c < C; ++c) {
#pragma unroll
            for (int r = 0; r < RM; ++r) Mr[r][c] = NUL;
            Ir[c] = NUL;
            Dr[c] = NUL;
        }
        // Window edges that are not bounds of the alignment: a non-NULL cell
        // in the window's first (last) lane then means the range may leave
        // the window.  Diagonals -p-1 and t+1 hold no cell in bounds.
        const uint32_t edges = __builtin_amdgcn_readfirstlane((kmin > -p - 1 ? 1u : 0u) | (kmax < t + 1 ? 2u : 0u));
        // Ranges of window diagonals (j = k - kmin) that may hold a non-NULL
        // cell, as WFA2-lib's wavefront limits (wfa_gpu.cu:426-468) with
        // e = 1: I/D at s within [min(lo M[s-o-e], lo I/D[s-1]) - 1,
        // max(hi M[s-o-e], hi I/D[s-1]) + 1], M at s the union of that and
        // M[s-x].  The ranges only grow (trimmed cells are NULLed but the
        // limits stay), so the range at s also covers every chunk whose ring
        // slot still holds a non-NULL M of score s-RM.  mtop/mbot hold the M
        // limits per ring slot, it/ib those of I/D at s-1.
        constexpr int RNONE = 1 << 20;  // limits of an empty range: top -RNONE, bottom +RNONE
        int mtop[RM], mbot[RM];
#pragma unroll
        for (int r = 0; r < RM; ++r) { mtop[r] = -RNONE; mbot[r] = RNONE; }
        mtop[0] = -kmin; mbot[0] = -kmin;  // score 0: M[0][0] only
        int it = -RNONE, ib = RNONE;

        // score 0: M[0][0] = 0, extended (every lane computes it)
        const int ext0 = lcp_interleaved(sS, 0, 0, min(p, t));
        {
            const int c0 = (0 - kmin) >> 6, l0 = (0 - kmin) & 63;
#pragma unroll
            for (int c = 0; c < C; ++c)
                if (c == c0) Mr[0][c] = lane == l0 ? 2 * ext0 : NUL;
        }
        int s = 0;

        // One score step whose M slot is PH (= s mod RM).  Returns 0 to go
        // on, the score when the end cell is reached, or a negative status.
        auto step = [&](auto PHC) __attribute__((always_inline)) -> int {
            constexpr int PH = decltype(PHC)::value;
            constexpr int SX = (PH - XS + 4 * RM) % RM;
            constexpr int SO = (PH - OS - ES + 4 * RM) % RM;
            ++s;
            if (s > s_bound) return TSU_EINTERNAL;
            if constexpr (PROBE) {
                if (s > a.probe_cap) return kProbeStopped;
            }
            // an opaque zero per step keeps the per-chunk bounds below from
            // being hoisted out of the score loop (16 VGPRs at C = 8): they
            // are recomputed, which costs less than the occupancy
            int zero_s;
            asm volatile("s_mov_b32 %0, 0" : "=s"(zero_s));
            // limits at s; chunk c computes when it meets the M range
            it = max(mtop[SO], it) + 1;
            ib = min(mbot[SO], ib) - 1;
            mtop[PH] = max(mtop[SX], it);
            mbot[PH] = min(mbot[SX], ib);
            const int jlo = max(mbot[PH], 0), jhi = min(mtop[PH], W - 1);
            const uint32_t act = jlo <= jhi ? (2u << (jhi >> 6)) - (1u << (jlo >> 6)) : 0u;
            uint32_t eg = edges;  // opaque per step: scalar tests below
            asm volatile("" : "+s"(eg));
            uint32_t edge = 0;     // a non-NULL cell in an edge lane of the window
            uint64_t oob = 0;      // some I/D cell out of bounds
            int32_t mend = NUL;    // M[s][t-p]
            int32_t carryI = NUL;  // old I of the previous chunk's lane 63
#pragma unroll
            for (int c = 0; c < C; ++c) {
                if (act & (1u << c)) {
                    // offsets are kept doubled: y2 = 2y.  A value y >= 0 on
                    // diagonal k is in bounds iff y <= t and y - k <= p, i.e.
                    // y2 < mt2 = 2 min(k+p+1, t+1) (>= 0 since k >= -p-1)
                    const int k2 = 2 * (kmin + zero_s + 64 * c) + lane2;
                    const int mt2 = min(k2 + 2 * (p + 1), 2 * (t + 1));
                    const int32_t mo = Mr[SO][c];
                    const int32_t mol = dpp_shr1(c > 0 ? dpp_ror1(Mr[SO][c > 0 ? c - 1 : 0]) : NUL, mo);
                    const int32_t mor = dpp_shl1(c < C - 1 ? dpp_rol1(Mr[SO][c < C - 1 ? c + 1 : c]) : NUL, mo);
                    const int32_t iel = dpp_shr1(carryI, Ir[c]);
                    const int32_t der = dpp_shl1(c < C - 1 ? dpp_rol1(Dr[c < C - 1 ? c + 1 : c]) : NUL, Dr[c]);
                    carryI = __builtin_amdgcn_readlane(Ir[c], 63);
                    // NULL sources stay negative through the +1s below (at
                    // most s_bound of them), so I, D and X need no NULL test;
                    // M is put back to the canonical NULL by the bounds test.
                    // I is not clamped: an I above t is out of bounds and
                    // only produces out-of-bounds values, as in WFA2-lib.
                    const int32_t I = max(mol, iel) + 2;
                    const int32_t D = max(mor, der);
                    const int32_t Mp = max(max(D, Mr[SX][c] + 2), I);  // M before the bounds test
                    const bool valid = (uint32_t)Mp < (uint32_t)mt2;
                    int32_t M = valid ? Mp : NUL;
                    // Mp is out of bounds where I or D is, or the mismatch
                    // (then trimming finds nothing and changes nothing)
                    oob |= vote(Mp >= mt2);
                    // Mp is non-NULL where M, I or D is, or a mismatch from a
                    // non-NULL M[s-x] (an edge hit x scores earlier)
                    if (c == 0) edge |= (uint32_t)(vote(Mp >= 0) & 1ull) & eg;
                    if (c == C - 1) edge |= (uint32_t)(vote(Mp >= 0) >> 63) & (eg >> 1);
                    const int lim2 = mt2 - 2 - M;  // 2 min(p - v, t - M) for a valid M
                    // extend from the packed words (v >= 0 for every valid M)
                    if (valid && lim2 > 0) M += extend2((uint32_t)(M - k2), (uint32_t)M, lim2);
                    Mr[PH][c] = M;
                    Ir[c] = I;
                    Dr[c] = D;
                    if (c == ctp) mend = __builtin_amdgcn_readlane(M, ltp);
                } else {
                    carryI = NUL;  // an idle chunk holds only NULL
                }
            }
            if (oob) {
                // exact trim: NULL I/D cells outside [first, last] in-bounds
                // cell, then test the edge lanes again
                auto trim = [&](int32_t (&R)[C]) __attribute__((always_inline)) {
                    int f = 0x3fffffff, l = -1;
#pragma unroll
                    for (int c = 0; c < C; ++c) {
                        const int k2 = 2 * (kmin + 64 * c) + lane2;
                        const int mt2 = min(k2 + 2 * (p + 1), 2 * (t + 1));
                        const uint64_t inb = vote((uint32_t)R[c] < (uint32_t)mt2);
                        if (inb) {
                            f = min(f, 64 * c + (int)__builtin_ctzll(inb));
                            l = max(l, 64 * c + 63 - (int)__builtin_clzll(inb));
                        }
                    }
#pragma unroll
                    for (int c = 0; c < C; ++c) {
                        const int w = 64 * c + lane;
                        if (w < f || w > l) R[c] = NUL;
                    }
                };
                trim(Ir);
                trim(Dr);
                edge = (uint32_t)(vote(max(max(Mr[PH][0], Ir[0]), Dr[0]) >= 0) & 1ull) & eg;
                edge |= (uint32_t)(vote(max(max(Mr[PH][C - 1], Ir[C - 1]), Dr[C - 1]) >= 0) >> 63) & (eg >> 1);
            }
            if (mend >= 2 * t) return s;
            // A non-NULL cell at a window edge could extend the range beyond
            // it (lane 0 of chunk 0, lane 63 of chunk C-1)
            if (edge) return TSU_EKOVERFLOW;
            return 0;
        };

        int r = (p == t && ext0 == t) ? 1 : 0;  // 1: score 0 (identical sequences)
        while (r == 0) {
            // phases 1, 2, ..., RM-1, 0 (s = 1, 2, ... mod RM)
            [&]<int... I>(std::integer_sequence<int, I...>) __attribute__((always_inline)) {
                ((r = (r == 0 ? step(std::integral_constant<int, (I + 1) % RM>{}) : r)), ...);
            }(std::make_integer_sequence<int, RM>{});
        }
        const int score = (r > 0 && s == 0) ? 0 : r;  // r > 0 is the score s of the last step
        if constexpr (PROBE) {
            if (r == kProbeStopped) {
                // largest 2h - k = h + v of a valid M cell in the ring (the
                // offsets are doubled: 2 M - k2 = 2 (h + v))
                int best = -1;
#pragma unroll
                for (int c = 0; c < C; ++c) {
                    const int k2 = 2 * (kmin + 64 * c) + lane2;
#pragma unroll
                    for (int q = 0; q < RM; ++q)
                        if (Mr[q][c] >= 0) best = max(best, 2 * Mr[q][c] - k2);
                }
#pragma unroll
                for (int d = 32; d >= 1; d >>= 1) best = max(best, __shfl_xor(best, d));
                if (lane == 0) a.probe_out[pi] = best >> 1;
            }
        }
        if (lane == 0) a.scores[pi] = !lds_at_zero ? TSU_EINTERNAL : score < 0 ? score : score * a.g;
        __syncthreads();
    }
}

}  // namespace tsu
