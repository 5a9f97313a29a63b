// wfa2p_reg.h -- gap-affine 2-piece WFA, single-wave register kernel.
//
// The 2-piece step of wfa2p.hip (WFA2-lib's affine-2p one: five wavefronts
// M, I1, D1, I2, D2) for small gap extensions after the gcd, E1 = e1 and
// E2 = e2 fixed at compile time ((2, 1) holds WFA2-lib's default 2-piece
// penalties 4/6/2 + 24/1), laid out as the 1-piece single-wave kernel
// (wfa_reg1l.h): one wave64 per pair, a window of 64 C diagonals, chunk c of
// lane l holding diagonal kmin + 64 c + l, offsets kept doubled (the bit
// position of a base in the packed words), the extend fused into the step.
//
//   I1, D1 at s read I1, D1 at s-E1: E1 register slots per chunk, slot
//   s mod E1 holding s-E1 until it receives s; I2, D2 likewise with E2
//   slots (one slot is read before it is written); the score loop is
//   unrolled over lcm(E1, E2) scores so every slot index is a constant;
//   the k-1 / k+1 neighbours come by DPP lane shifts (a chunk's lane 0 takes
//   the previous chunk's old lane 63, lane 63 the next chunk's lane 0);
//   M is read at s-x, s-o1-e1 and s-o2-e2 (up to 25 scores back for
//   4/6/2 + 24/1): every M goes to an int16 LDS ring of RM = max(x, o1+e1,
//   o2+e2) + 1 slots of W + 2 entries (NULL guards at both ends) once it is
//   extended, and is read from there at its own diagonal and at k-1 / k+1.
//
// Chunks are chosen by a range that covers WFA2-lib's wavefront limits and
// only grows: R(s) = R(s-1), joined with R(s-E1) widened by one diagonal
// each side once a first-piece gap can open (s >= o1+e1) and with R(s-E2)
// widened by one once a second-piece gap can (s >= o2+e2).  Every source of
// a score-s cell lies in R(s-1), R(s-E1) or R(s-E2) widened so (each range
// holds the ranges of the scores before it), so R(s) holds the new limits;
// computing cells outside the limits gives NULL there (their sources are
// NULL), and as R never shrinks, a ring slot's entries from RM scores
// earlier are overwritten wherever they were set.  Out-of-bounds M are NULLed by the
// bounds test; when a cell is out of bounds the four I/D wavefronts are
// trimmed at their ends as WFA2-lib's trim does (interior cells keep their
// values, as there).  A non-NULL cell in an edge lane of a window whose
// edge is not a sequence bound reports TSU_EKOVERFLOW (the range could
// leave the window); a doubled value fed only by NULLs grows by at most two
// per score, so the kernel gives up past kP2RegMaxScore with the same status
// (the host then runs the pair on wfa2p.hip).
//
// With BT every computed chunk writes its 64 records (wfa_bt.h, 2-piece
// format: origin of M among X, D2, D1, I2, I1 with ties in that order, four
// extension bits, M before extension) at ((s C + c) 64 + lane) of the
// workgroup's region, and lane 0 walks them with the nine backtrace types.
#pragma once

#include <utility>

#include "wfa_bt.h"
#include "wfa_reg.h"

namespace tsu {

// doubled NULL-fed values (-32768 + 2 per score) stay negative this long
constexpr int kP2RegMaxScore = 16000;

constexpr int gcd_c(int a, int b) { return b ? gcd_c(b, a % b) : a; }

template <int C, int E1, int E2, bool QUEUE, bool BT = false>
__global__ __launch_bounds__(64) void wfa2p_reg_kernel(KArgs a) {
    static_assert(C >= 1 && C <= 16, "window of 64..1024 diagonals");
    static_assert(E1 >= 1 && E1 <= 3 && E2 >= 1 && E2 <= 3, "gap extensions of 1..3 after the gcd");
    constexpr int EM = E1 > E2 ? E1 : E2;              // ranges kept: R(s-1) .. R(s-EM)
    constexpr int UNROLL = E1 * E2 / gcd_c(E1, E2);  // scores per loop iteration
    constexpr int32_t NUL = kRegNull + 0x8000;  // low 16 bits: -32768
    constexpr int W = 64 * C;
    constexpr int MRS = W + 2;
    extern __shared__ __align__(16) uint32_t smem[];
    const int lane = threadIdx.x;
    const int lane2 = 2 * lane;
    uint32_t* sS = smem;  // interleaved pattern / text words
    int32_t* sMisc = reinterpret_cast<int32_t*>(sS + 2 * a.seq_words);
    int16_t* MR = reinterpret_cast<int16_t*>(sMisc + 4);
    const bool lds_at_zero = (uint32_t)(uintptr_t)sS == 0u;
    const int RM = a.rm;
    const int x = a.x, oe1 = a.o + E1, oe2 = a.o2 + E2;
    uint32_t* rec = nullptr;
    if constexpr (BT) rec = a.rec_g + (int64_t)blockIdx.x * a.rec_cap;

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
        uint32_t* mr32 = reinterpret_cast<uint32_t*>(MR);
        for (int i = lane; i < (RM * MRS + 1) / 2; i += 64) mr32[i] = 0x80008000u;  // int16 NULLs
        __syncthreads();

        int kmin = max(-(W / 2), -p - 1);
        int kmax = kmin + W - 1;
        if (kmax > t + 1) { kmax = t + 1; kmin = max(kmax - W + 1, -p - 1); }
        const int tp = t - p;
        const int ctp = (tp - kmin) >= 0 ? (tp - kmin) >> 6 : -1;
        const int ltp = (tp - kmin) & 63;
        const uint32_t endm = ctp >= 0 && ctp < C ? 1u << ctp : 0u;  // chunk of the end cell
        const int d = abs(p - t);
        const int64_t tail = p != t ? min((int64_t)a.o + (int64_t)E1 * d, (int64_t)a.o2 + (int64_t)E2 * d) : 0;
        const int s_bound = (int)min<int64_t>((int64_t)x * min(p, t) + tail, 0x3fffffff);

        int32_t I1r[E1][C], D1r[E1][C], I2r[E2][C], D2r[E2][C];
#pragma unroll
        for (int c = 0; c < C; ++c) {
#pragma unroll
            for (int e = 0; e < E1; ++e) { I1r[e][c] = NUL; D1r[e][c] = NUL; }
#pragma unroll
            for (int e = 0; e < E2; ++e) { I2r[e][c] = NUL; D2r[e][c] = NUL; }
        }
        // per chunk: one past the last doubled offset in bounds,
        // min(2(k + p + 1), 2(t + 1)), and two less (floored at 0)
        int32_t MT2[C], MT2m2[C];
#pragma unroll
        for (int c = 0; c < C; ++c) {
            MT2[c] = min(2 * (kmin + 64 * c) + lane2 + 2 * (p + 1), 2 * (t + 1));
            MT2m2[c] = max(MT2[c] - 2, 0);
        }
        const uint32_t edges = __builtin_amdgcn_readfirstlane((kmin > -p - 1 ? 1u : 0u) | (kmax < t + 1 ? 2u : 0u));

        // R(s-1-i) in window diagonals j = k - kmin, i < EM
        constexpr int RNONE = 1 << 20;
        int rlo[EM], rhi[EM];
        rlo[0] = -kmin; rhi[0] = -kmin;  // score 0: M[0][0] only
#pragma unroll
        for (int i = 1; i < EM; ++i) { rlo[i] = RNONE; rhi[i] = -RNONE; }

        // score 0: M[0][0] = 0, extended, into ring slot 0
        const int ext0 = lcp_interleaved(sS, 0, 0, min(p, t));
        if (lane == 0) MR[(0 - kmin) + 1] = (int16_t)(2 * ext0);
        int s = 0;
        int gs = 0;  // ring slot of score s

        // One score step; P1 = s mod E1 and P2 = s mod E2 select the register
        // slots holding I1/D1 at s-E1 and I2/D2 at s-E2, which receive s.
        auto step = [&](auto P1C, auto P2C) __attribute__((always_inline)) -> int {
            constexpr int P1 = decltype(P1C)::value;
            constexpr int P2 = decltype(P2C)::value;
            ++s;
            if (s > s_bound) return TSU_EINTERNAL;
            if (s > kP2RegMaxScore) return TSU_EKOVERFLOW;
            if constexpr (BT) {
                if (s >= a.step_cap) return TSU_ECIGAR;
            }
            gs = gs + 1 == RM ? 0 : gs + 1;
            int gx = gs - x; gx += gx < 0 ? RM : 0;    // s - x
            int g1 = gs - oe1; g1 += g1 < 0 ? RM : 0;  // s - o1 - e1
            int g2 = gs - oe2; g2 += g2 < 0 ? RM : 0;  // s - o2 - e2
            int zero_s;
            asm volatile("s_mov_b32 %0, 0" : "=s"(zero_s));
            const int nl2 = zero_s - lane2;
            const int nkb = zero_s - 2 * kmin;
            uint32_t em = __builtin_amdgcn_readfirstlane(endm);
            asm volatile("" : "+s"(em));
            uint32_t eg = edges;
            asm volatile("" : "+s"(eg));
            // R(s)
            int nlo = rlo[0], nhi = rhi[0];
            if (s >= oe1) { nlo = min(nlo, rlo[E1 - 1] - 1); nhi = max(nhi, rhi[E1 - 1] + 1); }
            if (s >= oe2) { nlo = min(nlo, rlo[E2 - 1] - 1); nhi = max(nhi, rhi[E2 - 1] + 1); }
#pragma unroll
            for (int i = EM - 1; i > 0; --i) { rlo[i] = rlo[i - 1]; rhi[i] = rhi[i - 1]; }
            rlo[0] = nlo; rhi[0] = nhi;
            const int jlo = max(nlo, 0), jhi = min(nhi, W - 1);
            const uint32_t act = jlo <= jhi ? (2u << (jhi >> 6)) - (1u << (jlo >> 6)) : 0u;
            const int16_t* rx = MR + gx * MRS + lane + 1;  // entry of diagonal j
            const int16_t* r1 = MR + g1 * MRS + lane;      // entry of j-1 (j+1 two further)
            const int16_t* r2 = MR + g2 * MRS + lane;
            int16_t* rw = MR + gs * MRS + lane + 1;
            uint32_t* recs = nullptr;
            if constexpr (BT) recs = rec + (int64_t)s * C * 64;
            uint64_t oob = 0;
            uint32_t edge = 0;
            int32_t mend = NUL;
            int32_t m0 = NUL, mlast = NUL;  // M of the edge chunks (for the edge test after a trim)
            // the previous chunk's old I1 / I2 rotated by one lane (its lane
            // 63 in lane 0); the computed chunks are contiguous, so they are
            // NULL before the first of them and written by each
            int32_t rI1 = NUL, rI2 = NUL;
#pragma unroll
            for (int c = 0; c < C; ++c) {
                if (act & (1u << c)) {
                    const int mt2 = MT2[c];
                    const int32_t mx = rx[64 * c];
                    const int32_t m1l = r1[64 * c], m1r = r1[64 * c + 2];
                    const int32_t m2l = r2[64 * c], m2r = r2[64 * c + 2];
                    const int32_t i1n = dpp_shr1(rI1, I1r[P1][c]);
                    const int32_t i2n = dpp_shr1(rI2, I2r[P2][c]);
                    if (c < C - 1) {
                        rI1 = dpp_ror1(I1r[P1][c]);
                        rI2 = dpp_ror1(I2r[P2][c]);
                    }
                    const int32_t d1n =
                        dpp_shl1(c < C - 1 ? dpp_rol1(D1r[P1][c < C - 1 ? c + 1 : c]) : NUL, D1r[P1][c]);
                    const int32_t d2n =
                        dpp_shl1(c < C - 1 ? dpp_rol1(D2r[P2][c < C - 1 ? c + 1 : c]) : NUL, D2r[P2][c]);
                    const int32_t I1 = max(m1l, i1n) + 2;
                    const int32_t D1 = max(m1r, d1n);
                    const int32_t I2 = max(m2l, i2n) + 2;
                    const int32_t D2 = max(m2r, d2n);
                    const int32_t X = mx + 2;
                    const int32_t Mp = max(max(X, max(D1, D2)), max(I1, I2));  // M before the bounds test
                    const bool valid = (uint32_t)Mp < (uint32_t)mt2;
                    int32_t M = valid ? Mp : NUL;
                    oob |= vote(Mp >= mt2);
                    if (c == 0) edge |= (uint32_t)(vote(Mp >= 0) & 1ull) & eg;
                    if (c == C - 1) edge |= (uint32_t)(vote(Mp >= 0) >> 63) & (eg >> 1);
                    if constexpr (BT) {
                        uint32_t r = (i1n >= m1l ? 8u : 0u) | (d1n >= m1r ? 16u : 0u) | (i2n >= m2l ? 32u : 0u) |
                                     (d2n >= m2r ? 64u : 0u);
                        if (valid) {
                            const uint32_t org = (X >= D2 && X >= D1 && X >= I2 && X >= I1) ? 1u
                                                 : (D2 >= D1 && D2 >= I2 && D2 >= I1) ? 2u
                                                 : (D1 >= I2 && D1 >= I1) ? 3u
                                                 : (I2 >= I1) ? 4u : 5u;
                            r |= (((uint32_t)Mp >> 1) << 7) | org;
                        }
                        recs[64 * c + lane] = r;
                    }
                    const int mt2m2 = MT2m2[c];
                    int sc = nkb - 128 * c;
                    asm volatile("" : "+s"(sc));
                    if ((uint32_t)M < (uint32_t)mt2m2)  // valid, with bases left
                        M += extend2((uint32_t)M + (uint32_t)nl2 + (uint32_t)sc, (uint32_t)M, mt2m2 - M);
                    rw[64 * c] = (int16_t)M;  // exactly NULL or an offset below 32768
                    I1r[P1][c] = I1;
                    D1r[P1][c] = D1;
                    I2r[P2][c] = I2;
                    D2r[P2][c] = D2;
                    if (c == 0) m0 = M;
                    if (c == C - 1) mlast = M;
                    if (em & (1u << c)) mend = __builtin_amdgcn_readlane(M, ltp);
                }
            }
            if (oob) {
                // trim the I/D ends (wfa_gpu.cu:633-719): NULL cells outside
                // [first, last] in-bounds cell of each, then test the edge
                // lanes again
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
                trim(I1r[P1]);
                trim(D1r[P1]);
                trim(I2r[P2]);
                trim(D2r[P2]);
                edge = 0;
                edge |= (uint32_t)(vote(max(max(m0, max(I1r[P1][0], D1r[P1][0])), max(I2r[P2][0], D2r[P2][0])) >= 0) &
                                   1ull) &
                        eg;
                edge |= (uint32_t)(vote(max(max(mlast, max(I1r[P1][C - 1], D1r[P1][C - 1])),
                                            max(I2r[P2][C - 1], D2r[P2][C - 1])) >= 0) >> 63) &
                        (eg >> 1);
            }
            if (mend >= 2 * t) return s;
            if (edge) return TSU_EKOVERFLOW;
            return 0;
        };

        // UNROLL scores per iteration: score s = UNROLL n + i + 1 takes the
        // slots (i + 1) mod E1 and (i + 1) mod E2
        auto steps = [&]<int... I>(std::integer_sequence<int, I...>) __attribute__((always_inline)) -> int {
            int rr = 0;
            ((rr == 0 ? (rr = step(std::integral_constant<int, (I + 1) % E1>{}, std::integral_constant<int, (I + 1) % E2>{}))
                      : 0),
             ...);
            return rr;
        };
        int r = (p == t && ext0 == t) ? 1 : 0;  // 1: score 0 (identical sequences)
        while (r == 0) r = steps(std::make_integer_sequence<int, UNROLL>{});
        const int score = (r > 0 && s == 0) ? 0 : r;
        if constexpr (BT) {
            __syncthreads();  // every lane's records written
            if (lane == 0) {
                int st = !lds_at_zero ? TSU_EINTERNAL : score < 0 ? score : 0;
                int64_t out_off = 0;
                int32_t out_n = 0;
                if (st == 0) {
                    auto record = [&](int ss, int kk, uint32_t& rr) __attribute__((always_inline)) -> bool {
                        if (ss <= 0 || ss > s) return false;
                        const int j = kk - kmin;
                        if (j < 0 || j >= W) return false;
                        rr = rec[((int64_t)ss * C + (j >> 6)) * 64 + (j & 63)];
                        return true;
                    };
                    st = bt2p_walk(record, score, p, t, x, a.o, E1, a.o2, E2, a, &out_off, &out_n);
                }
                a.run_off[pi] = out_off;
                a.run_n[pi] = out_n;
                a.scores[pi] = st ? st : score * a.g;
            }
        } else {
            if (lane == 0) a.scores[pi] = !lds_at_zero ? TSU_EINTERNAL : score < 0 ? score : score * a.g;
        }
        __syncthreads();  // LDS reused by the next pair of this workgroup
    }
}

template <int C, int E1, int E2>
KernelFn pick_2p_reg_c(const Shape& sh) {
    if (sh.bt) return sh.queue && !sh.ck ? wfa2p_reg_kernel<C, E1, E2, true, true> : nullptr;
    return sh.queue ? wfa2p_reg_kernel<C, E1, E2, true> : wfa2p_reg_kernel<C, E1, E2, false>;
}

// Instances for chunks 1, 2, 4, 8, 12 and 16 (one translation unit per
// (E1, E2), so they compile in parallel).
template <int E1, int E2>
KernelFn pick_2p_reg_impl(const Shape& sh) {
    switch (sh.reg_c) {
    case 1: return pick_2p_reg_c<1, E1, E2>(sh);
    case 2: return pick_2p_reg_c<2, E1, E2>(sh);
    case 4: return pick_2p_reg_c<4, E1, E2>(sh);
    case 8: return pick_2p_reg_c<8, E1, E2>(sh);
    case 12: return pick_2p_reg_c<12, E1, E2>(sh);
    case 16: return pick_2p_reg_c<16, E1, E2>(sh);
    default: return nullptr;
    }
}

}  // namespace tsu
