// wfa_reg.h -- register-resident WFA kernel for short and medium pairs.
//
// One wave64 per pair.  Diagonal k lives in lane (k - kmin) % 64 of chunk
// (k - kmin) / 64 for the whole alignment: the M ring (max(x, o+e)+1 scores)
// and the I / D wavefronts (updated in place, which needs e = 1 after gcd
// scaling) are arrays of VGPRs indexed by compile-time chunk and slot.  The
// k-1 / k+1 neighbours come from DPP wave shifts (wave_shr:1 / wave_shl:1)
// with the boundary lane taken from the adjacent chunk by v_readlane.  The
// score loop is unrolled over the M ring's depth so every slot index is a
// constant.
//
// All diagonals of the C*64-wide window are computed implicitly: a chunk is
// computed only when one of its sources can be non-NULL (a scalar bit mask
// per slot, dilated by one chunk for the k+-1 neighbours), and a diagonal
// outside WFA2-lib's range [lo, hi] has only NULL sources, so it comes out
// NULL exactly as the range-limited computation leaves it.  Trim-ends
// (wfa_gpu.cu:633-719) reduce to: NULL every I/D cell below the first and
// above the last in-bounds cell of that wavefront; M cells outside the
// bounds are NULL already.  The common case (no out-of-bounds I/D value at
// all) is detected from one ballot per chunk.
//
// Recurrences, bounds test, end test and extend are those of the LDS kernel
// (wfa_gpu.cu:253-331,610-631).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "wfa_kernel.h"

namespace tsu {

constexpr int32_t kRegNull = -1073741824;
// per-step chunk flag bits (chunk c adds c to the M, I and D positions)
constexpr uint64_t kBitM = 1ull << 1, kBitI = 1ull << 17, kBitD = 1ull << 33;

__device__ __forceinline__ uint64_t vote(bool b) { return __builtin_amdgcn_ballot_w64(b); }

__device__ __forceinline__ int32_t dpp_shr1(int32_t fill_lane0, int32_t v) {
    // lane l receives lane l-1; lane 0 keeps fill_lane0
    return __builtin_amdgcn_update_dpp(fill_lane0, v, 0x138, 0xF, 0xF, false);
}
__device__ __forceinline__ int32_t dpp_shl1(int32_t fill_lane63, int32_t v) {
    // lane l receives lane l+1; lane 63 keeps fill_lane63
    return __builtin_amdgcn_update_dpp(fill_lane63, v, 0x130, 0xF, 0xF, false);
}

template <int C, int XS, int OS, int ES, bool QUEUE>
__global__ __launch_bounds__(64) void wfa_reg_kernel(KArgs a) {
    static_assert(ES == 1, "in-place I/D needs e == 1 after gcd scaling");
    static_assert(C >= 1 && C <= 16, "window of 64..1024 diagonals (16 flag bits per kind)");
    constexpr int RM = (XS > OS + ES ? XS : OS + ES) + 1;
    constexpr int32_t NUL = kRegNull;
    constexpr int W = 64 * C;
    constexpr uint32_t ALL = C == 32 ? 0xffffffffu : ((1u << C) - 1u);
    extern __shared__ __align__(16) uint32_t smem[];
    const int lane = threadIdx.x;
    uint32_t* sP = smem;
    uint32_t* sT = smem + a.seq_words;
    int32_t* sMisc = reinterpret_cast<int32_t*>(sT + a.seq_words);

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
            sP[i] = i < pw ? a.words[pr.p_off + i] : 0u;
            sT[i] = i < tw ? a.words[pr.t_off + i] : 0u;
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
        uint32_t amM[RM];   // chunks holding a non-NULL M value, per slot
#pragma unroll
        for (int r = 0; r < RM; ++r) amM[r] = 0;
        uint32_t amI = 0, amD = 0;

        // score 0: M[0][0] = 0, extended (every lane computes it)
        const int ext0 = lcp_packed(sP, sT, 0, 0, min(p, t));
        {
            const int c0 = (0 - kmin) >> 6, l0 = (0 - kmin) & 63;
#pragma unroll
            for (int c = 0; c < C; ++c)
                if (c == c0) Mr[0][c] = lane == l0 ? ext0 : NUL;
            amM[0] = 1u << c0;
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
            const uint32_t srcI = amM[SO] | amI;
            const uint32_t srcD = amM[SO] | amD;
            const uint32_t act = (amM[SX] | srcI | (srcI << 1) | srcD | (srcD >> 1) | amM[PH]) & ALL;
            uint64_t nn = 0;  // chunk flags: M at bit 1 + c, I at 17 + c, D at 33 + c
            uint64_t oob = 0;
            int32_t mend = NUL;  // M[s][t-p]
            int32_t carryI = NUL;  // old I of the previous chunk's lane 63
#pragma unroll
            for (int c = 0; c < C; ++c) {
                if (act & (1u << c)) {
                    const int k = kmin + 64 * c + lane;
                    // a value y >= 0 on diagonal k is in bounds iff y <= t and
                    // y - k <= p, i.e. y < mt1; mt1 >= 0 since k >= -p-1
                    const int mt1 = min(k + p + 1, t + 1);
                    const int32_t mo = Mr[SO][c];
                    const int32_t mol = dpp_shr1(c > 0 ? __builtin_amdgcn_readlane(Mr[SO][c > 0 ? c - 1 : 0], 63) : NUL, mo);
                    const int32_t mor = dpp_shl1(c < C - 1 ? __builtin_amdgcn_readlane(Mr[SO][c < C - 1 ? c + 1 : c], 0) : NUL, mo);
                    const int32_t iel = dpp_shr1(carryI, Ir[c]);
                    const int32_t der = dpp_shl1(c < C - 1 ? __builtin_amdgcn_readlane(Dr[c < C - 1 ? c + 1 : c], 0) : NUL, Dr[c]);
                    carryI = __builtin_amdgcn_readlane(Ir[c], 63);
                    // NULL sources stay negative through the +1s below (at
                    // most s_bound of them), so I, D and X need no NULL test;
                    // M is put back to the canonical NULL by the bounds test
                    const int32_t I = min(max(mol, iel) + 1, t + 1);
                    const int32_t D = max(mor, der);
                    int32_t M = max(max(D, Mr[SX][c] + 1), I);
                    const bool valid = (uint32_t)M < (uint32_t)mt1;
                    M = valid ? M : NUL;
                    const int lim = mt1 - 1 - M;  // min(p - v, t - M) for a valid M
                    if (valid && lim > 0) {
                        // extend: 32 bases per round from the 2-bit packed
                        // sequences; alignbit takes the shift modulo 32
                        const int v = M - k;
                        const uint32_t* pp = sP + (v >> 4);
                        const uint32_t* tt = sT + (M >> 4);
                        const uint32_t sa = (uint32_t)v << 1, sb = (uint32_t)M << 1;
                        const int lim2 = 2 * lim;
                        int n2 = 0;  // matched bits, 2 per base
                        for (;;) {
                            const uint32_t p0 = pp[0], p1 = pp[1], p2 = pp[2];
                            const uint32_t t0 = tt[0], t1 = tt[1], t2 = tt[2];
                            const uint32_t xl = __builtin_amdgcn_alignbit(p1, p0, sa) ^ __builtin_amdgcn_alignbit(t1, t0, sb);
                            const uint32_t xh = __builtin_amdgcn_alignbit(p2, p1, sa) ^ __builtin_amdgcn_alignbit(t2, t1, sb);
                            const uint64_t xx = ((uint64_t)xh << 32) | xl;
                            const int m = xx ? (int)__builtin_ctzll(xx) : 64;
                            n2 += m;
                            if (m != 64 || n2 >= lim2) break;
                            pp += 2;
                            tt += 2;
                        }
                        M += min(n2 >> 1, lim);
                    }
                    Mr[PH][c] = M;
                    Ir[c] = I;
                    Dr[c] = D;
                    // chunk flags as one scalar integer pinned to an SGPR, so
                    // the compiler carries no lane booleans across the chunk
                    // branches (two vector ops per flag); no flag uses bit 0,
                    // whose zero-extension is also lowered through a VGPR
                    uint64_t fl = (vote(M >= 0) ? kBitM : 0ull) | (vote(I >= 0) ? kBitI : 0ull) |
                                  (vote(D >= 0) ? kBitD : 0ull);
                    asm volatile("" : "+s"(fl));
                    nn |= fl << c;
                    if (c == ctp) mend = __builtin_amdgcn_readlane(M, ltp);
                    // I <= t + 1 and D <= t by construction: out of bounds iff >= mt1
                    oob |= vote(max(I, D) >= mt1);
                } else {
                    carryI = NUL;  // an idle chunk holds only NULL
                }
            }
            uint32_t nnM = (uint32_t)(nn >> 1) & ALL;
            uint32_t nnI = (uint32_t)(nn >> 17) & ALL;
            uint32_t nnD = (uint32_t)(nn >> 33) & ALL;
            if (oob) {
                // exact trim: NULL I/D cells outside [first, last] in-bounds cell
                auto trim = [&](int32_t (&R)[C], uint32_t& nn) __attribute__((always_inline)) {
                    int f = 0x3fffffff, l = -1;
#pragma unroll
                    for (int c = 0; c < C; ++c) {
                        const int k = kmin + 64 * c + lane;
                        const int mt1 = min(k + p + 1, t + 1);
                        const uint64_t inb = vote((uint32_t)R[c] < (uint32_t)mt1);
                        if (inb) {
                            f = min(f, 64 * c + (int)__builtin_ctzll(inb));
                            l = max(l, 64 * c + 63 - (int)__builtin_clzll(inb));
                        }
                    }
                    nn = 0;
#pragma unroll
                    for (int c = 0; c < C; ++c) {
                        const int w = 64 * c + lane;
                        if (w < f || w > l) R[c] = NUL;
                        if (vote(R[c] >= 0)) nn |= 1u << c;
                    }
                };
                trim(Ir, nnI);
                trim(Dr, nnD);
            }
            amM[PH] = nnM;
            amI = nnI;
            amD = nnD;
            if (mend >= t) return s;
            // a non-NULL cell at a window edge could extend the range beyond it
            uint32_t edge = 0;
            {
                const int32_t lo_any = max(Mr[PH][0], max(Ir[0], Dr[0]));
                const int32_t hi_any = max(Mr[PH][C - 1], max(Ir[C - 1], Dr[C - 1]));
                edge = (__builtin_amdgcn_readlane(lo_any, 0) >= 0 ? 1u : 0u) |
                       (__builtin_amdgcn_readlane(hi_any, 63) >= 0 ? 2u : 0u);
            }
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
        if (lane == 0) a.scores[pi] = score < 0 ? score : score * a.g;
        __syncthreads();
    }
}

}  // namespace tsu
