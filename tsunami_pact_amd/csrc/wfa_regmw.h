// wfa_regmw.h -- register-resident WFA with several waves per pair.
//
// The single-wave register kernel (wfa_reg.h) holds a window of at most
// 512 diagonals.  Here a workgroup of NW wave64s holds NW*C*64 diagonals
// (C chunks of 64 per wave): global chunk j belongs to wave j mod NW, in its
// register slot j / NW.  Chunks are dealt round-robin so that every wave has
// a near-equal share of any contiguous diagonal range, and the waves of a
// pair stay equally busy as its wavefront widens.
//
// A chunk's k-1 / k+1 neighbours across its edges live in other waves, so
// after each score step every wave publishes, for all its chunks, lane 0 and
// lane 63 of M, lane 63 of I and lane 0 of D to LDS, together with its step
// mask (the same any / lane-0 / lane-63 chunk flags as the single-wave
// kernel) and the step's flags (out-of-bounds I/D seen, end reached,
// window edge reached).  One workgroup barrier per score step separates the
// writes of step s from the reads of step s+1; the edge slots of M rotate
// with the M ring and those of I and D with the parity of s, so a fast wave
// never overwrites a value a slow wave has still to read.
//
// Recurrences, bounds test, extend, trim and end test are those of the
// single-wave kernel (wfa_gpu.cu:253-331,610-631,633-719); the trim, when
// some I/D cell is out of bounds, reduces the first/last in-bounds diagonal
// over the waves with LDS min/max and two more barriers.
#pragma once

#include "wfa_reg.h"

namespace tsu {

// LDS layout (int32 words) after the interleaved sequence words.
template <int NW, int C, int RM>
struct MwLayout {
    static constexpr int NC = NW * C;                  // chunks in the window
    static constexpr int EM = 0;                       // [RM][NC+2][2] M lane 0 / lane 63 per chunk, one guard chunk each end
    static constexpr int EI = EM + RM * (NC + 2) * 2;  // [2][NC+2] I lane 63 per chunk, by parity of the score
    static constexpr int ED = EI + 2 * (NC + 2);       // [2][NC+2] D lane 0 per chunk
    static constexpr int FW = ED + 2 * (NC + 2);       // [RM][NW] step mask of each wave
    static constexpr int FL = FW + RM * NW;            // [3] step flags, rotating
    static constexpr int TR = FL + 3;                  // [4] trim bounds: I first, I last, D first, D last
    static constexpr int MISC = TR + 4;                // [4] queue index
    static constexpr int WORDS = MISC + 4;
};

constexpr uint32_t kFlOob = 1, kFlEnd = 2, kFlEdge = 4;
constexpr int32_t kBig = 0x3fffffff;

template <int NW, int C, int XS, int OS, int ES, bool QUEUE>
__global__ __launch_bounds__(NW * 64) void wfa_regmw_kernel(KArgs a) {
    static_assert(ES == 1, "in-place I/D needs e == 1 after gcd scaling");
    static_assert(NW >= 2 && NW <= 16, "2..16 waves");
    static_assert(C >= 1 && C <= 8, "8-bit chunk fields per wave");
    constexpr int RM = (XS > OS + ES ? XS : OS + ES) + 1;
    constexpr int32_t NUL = kRegNull;
    constexpr int NC = NW * C;
    constexpr int W = 64 * NC;
    constexpr int NT = NW * 64;
    constexpr uint32_t ALL = (1u << C) - 1u;
    using L = MwLayout<NW, C, RM>;
    extern __shared__ __align__(16) uint32_t smem[];
    const int tid = threadIdx.x;
    const int lane = tid & 63;
    const int w = __builtin_amdgcn_readfirstlane(tid >> 6);
    uint32_t* sS = smem;  // interleaved pattern / text words
    int32_t* X = reinterpret_cast<int32_t*>(smem + 2 * a.seq_words);
    const bool lds_at_zero = (uint32_t)(uintptr_t)sS == 0u;
    // per-wave bases: entry (slot, chunk j = r*NW + w) of EM sits at
    // EMw + (slot*(NC+2) + r*NW + 1)*2 + side; its left neighbour's hi at
    // EMw + (slot*(NC+2) + r*NW)*2 + 1, its right neighbour's lo at
    // EMw + (slot*(NC+2) + r*NW + 2)*2 (guards make every offset >= 0)
    int32_t* EMw = X + L::EM + 2 * w;
    int32_t* EIw = X + L::EI + w;
    int32_t* EDw = X + L::ED + w;
    const int wl = w == 0 ? NW - 1 : w - 1;  // owner of chunk j-1 (slot r-1 when w == 0)
    const int wr = w == NW - 1 ? 0 : w + 1;  // owner of chunk j+1 (slot r+1 when w == NW-1)

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
        for (int i = tid; i < L::MISC; i += NT) {
            int32_t v = NUL;                                     // edge values
            if (i >= L::FW) v = 0;                               // masks, flags
            if (i >= L::TR) v = ((i - L::TR) & 1) ? -kBig : kBig;  // trim bounds (min, max, min, max)
            X[i] = v;
        }
        __syncthreads();

        int kmin = max(-(W / 2), -p - 1);
        int kmax = kmin + W - 1;
        if (kmax > t + 1) { kmax = t + 1; kmin = max(kmax - W + 1, -p - 1); }
        const int tp = t - p;
        // owner wave, slot and lane of the end diagonal t - p (none when it
        // lies outside the window: the pair then ends by overflow)
        const int jtp = tp - kmin >= 0 && tp - kmin < W ? (tp - kmin) >> 6 : -1;
        const int wtp = jtp >= 0 ? jtp % NW : -1;
        const int rtp = jtp >= 0 ? jtp / NW : -1;
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
        uint32_t F[RM];  // this wave's step masks of the last RM scores
#pragma unroll
        for (int r = 0; r < RM; ++r) F[r] = 0;

        // Publishes the edge lanes of every chunk of this wave for the score
        // in M slot PH (I/D parity par) and the wave's step mask.
        auto publish = [&](auto PHC, int par, uint32_t Fs) __attribute__((always_inline)) {
            constexpr int PH = decltype(PHC)::value;
            int32_t* ei = EIw + par * (NC + 2);
            int32_t* ed = EDw + par * (NC + 2);
            if (lane == 0) {
#pragma unroll
                for (int r = 0; r < C; ++r) {
                    EMw[(PH * (NC + 2) + r * NW + 1) * 2] = Mr[PH][r];
                    ed[r * NW + 1] = Dr[r];
                }
                X[L::FW + PH * NW + w] = (int32_t)Fs;
            }
            if (lane == 63) {
#pragma unroll
                for (int r = 0; r < C; ++r) {
                    EMw[(PH * (NC + 2) + r * NW + 1) * 2 + 1] = Mr[PH][r];
                    ei[r * NW + 1] = Ir[r];
                }
            }
        };

        // score 0: M[0][0] = 0, extended (every lane computes it)
        const int ext0 = lcp_interleaved(sS, 0, 0, min(p, t));
        {
            const int j0 = (0 - kmin) >> 6, l0 = (0 - kmin) & 63;
            if (j0 % NW == w) {
                const int r0 = j0 / NW;
#pragma unroll
                for (int r = 0; r < C; ++r)
                    if (r == r0) Mr[0][r] = lane == l0 ? ext0 : NUL;
                F[0] = (1u << r0) | (l0 == 0 ? 1u << (8 + r0) : 0u) | (l0 == 63 ? 1u << (16 + r0) : 0u);
            }
            publish(std::integral_constant<int, 0>{}, 0, F[0]);
        }
        __syncthreads();
        int s = 0;
        int fidx = 0;  // flag word of score s+1 (s+1 mod 3)

        // One score step whose M slot is PH (= s mod RM).  Returns 0 to go
        // on, the score when the end cell is reached, or a negative status.
        // Every wave takes the same branches: they depend only on values
        // read from LDS after a barrier.
        auto step = [&](auto PHC) __attribute__((always_inline)) -> int {
            constexpr int PH = decltype(PHC)::value;
            constexpr int SX = (PH - XS + 4 * RM) % RM;
            constexpr int SO = (PH - OS - ES + 4 * RM) % RM;
            constexpr int P1 = (PH - 1 + RM) % RM;
            ++s;
            if (s > s_bound) return TSU_EINTERNAL;
            fidx = fidx + 1 == 3 ? 0 : fidx + 1;
            // an opaque zero per step: the per-chunk bounds are recomputed
            // rather than held in VGPRs across the score loop
            int zero_s;
            asm volatile("s_mov_b32 %0, 0" : "=s"(zero_s));
            const int par = s & 1, par1 = par ^ 1;
            // neighbour waves' sources (scores s-1 and s-o-e) for the spills
            const uint32_t srcL = (uint32_t)__builtin_amdgcn_readfirstlane(X[L::FW + P1 * NW + wl]) |
                                  (uint32_t)__builtin_amdgcn_readfirstlane(X[L::FW + SO * NW + wl]);
            const uint32_t srcR = (uint32_t)__builtin_amdgcn_readfirstlane(X[L::FW + P1 * NW + wr]) |
                                  (uint32_t)__builtin_amdgcn_readfirstlane(X[L::FW + SO * NW + wr]);
            uint32_t hiL = (srcL >> 16) & 0xffu;
            uint32_t loR = (srcR >> 8) & 0xffu;
            if (w == 0) hiL <<= 1;
            if (w == NW - 1) loR >>= 1;
            const uint32_t src = F[P1] | F[SO];
            const uint32_t act = ((src | F[SX] | F[PH]) | hiL | loR) & ALL;
            const int32_t* ei1 = EIw + par1 * (NC + 2);
            const int32_t* ed1 = EDw + par1 * (NC + 2);
            uint32_t acc = 0;   // step mask, fields one bit up
            uint64_t oob = 0;
            int32_t mend = NUL;
#pragma unroll
            for (int c = 0; c < C; ++c) {
                if (act & (1u << c)) {
                    const int k = kmin + zero_s + 64 * (c * NW + w) + lane;
                    const int mt1 = min(k + p + 1, t + 1);
                    const int32_t mo = Mr[SO][c];
                    // neighbour chunks' edge lanes from LDS (broadcast reads)
                    const int32_t fml = EMw[(SO * (NC + 2) + c * NW) * 2 + 1];
                    const int32_t fmr = EMw[(SO * (NC + 2) + c * NW + 2) * 2];
                    const int32_t fil = ei1[c * NW];
                    const int32_t fdr = ed1[c * NW + 2];
                    const int32_t mol = dpp_shr1(fml, mo);
                    const int32_t mor = dpp_shl1(fmr, mo);
                    const int32_t iel = dpp_shr1(fil, Ir[c]);
                    const int32_t der = dpp_shl1(fdr, Dr[c]);
                    const int32_t I = max(mol, iel) + 1;
                    const int32_t D = max(mor, der);
                    int32_t M = max(max(D, Mr[SX][c] + 1), I);
                    const bool valid = (uint32_t)M < (uint32_t)mt1;
                    // lanes that may hold a non-NULL M, I or D (M before the
                    // bounds test is their maximum)
                    const uint64_t b = vote(M >= 0);
                    M = valid ? M : NUL;
                    const int lim = mt1 - 1 - M;
                    if (valid && lim > 0) M += extend_len((uint32_t)(M - k), (uint32_t)M, lim);
                    Mr[PH][c] = M;
                    Ir[c] = I;
                    Dr[c] = D;
                    uint32_t fl = (b ? 2u : 0u) | ((uint32_t)(b & 1ull) << 9) | ((uint32_t)(b >> 63) << 17);
                    asm volatile("" : "+s"(fl));
                    acc |= fl << c;
                    if (c == rtp && w == wtp) mend = __builtin_amdgcn_readlane(M, ltp);
                    oob |= vote(max(I, D) >= mt1);
                }
            }
            uint32_t Fs = acc >> 1;
            uint32_t flags = (oob ? kFlOob : 0u) | (mend >= t ? kFlEnd : 0u);
            if (w == 0 && ((Fs >> 8) & 1u)) flags |= kFlEdge;                // lane 0 of chunk 0
            if (w == NW - 1 && ((Fs >> (16 + C - 1)) & 1u)) flags |= kFlEdge;  // lane 63 of the last chunk
            if (lane == 0) {
                if (flags) atomicOr(reinterpret_cast<uint32_t*>(X + L::FL + fidx), flags);
                if (w == 0) X[L::FL + (fidx + 1 == 3 ? 0 : fidx + 1)] = 0;  // the next score's word
            }
            F[PH] = Fs;
            publish(PHC, par, Fs);
            __syncthreads();
            const uint32_t all = (uint32_t)__builtin_amdgcn_readfirstlane(X[L::FL + fidx]);
            if (all & kFlEnd) return s;
            if (all & kFlOob) {
                // exact trim (wfa_gpu.cu:633-719): first / last in-bounds I and
                // D over the whole window, cells outside them NULL
                int b4[4] = {kBig, -kBig, kBig, -kBig};
#pragma unroll
                for (int c = 0; c < C; ++c) {
                    const int k = kmin + 64 * (c * NW + w) + lane;
                    const int mt1 = min(k + p + 1, t + 1);
                    const int base = 64 * (c * NW + w);
                    const uint64_t bi = vote((uint32_t)Ir[c] < (uint32_t)mt1);
                    const uint64_t bd = vote((uint32_t)Dr[c] < (uint32_t)mt1);
                    if (bi) {
                        b4[0] = min(b4[0], base + (int)__builtin_ctzll(bi));
                        b4[1] = max(b4[1], base + 63 - (int)__builtin_clzll(bi));
                    }
                    if (bd) {
                        b4[2] = min(b4[2], base + (int)__builtin_ctzll(bd));
                        b4[3] = max(b4[3], base + 63 - (int)__builtin_clzll(bd));
                    }
                }
                if (lane == 0) {
                    atomicMin(X + L::TR + 0, b4[0]);
                    atomicMax(X + L::TR + 1, b4[1]);
                    atomicMin(X + L::TR + 2, b4[2]);
                    atomicMax(X + L::TR + 3, b4[3]);
                }
                __syncthreads();
                const int fI = __builtin_amdgcn_readfirstlane(X[L::TR + 0]);
                const int lI = __builtin_amdgcn_readfirstlane(X[L::TR + 1]);
                const int fD = __builtin_amdgcn_readfirstlane(X[L::TR + 2]);
                const int lD = __builtin_amdgcn_readfirstlane(X[L::TR + 3]);
                uint32_t m2 = 0;
#pragma unroll
                for (int c = 0; c < C; ++c) {
                    const int pos = 64 * (c * NW + w) + lane;
                    if (pos < fI || pos > lI) Ir[c] = NUL;
                    if (pos < fD || pos > lD) Dr[c] = NUL;
                    const uint64_t b = vote(max(max(Mr[PH][c], Ir[c]), Dr[c]) >= 0);
                    m2 |= ((b ? 2u : 0u) | ((uint32_t)(b & 1ull) << 9) | ((uint32_t)(b >> 63) << 17)) << c;
                }
                Fs = m2 >> 1;
                F[PH] = Fs;
                publish(PHC, par, Fs);
                __syncthreads();
                if (tid == 0) {
                    X[L::TR + 0] = kBig;
                    X[L::TR + 1] = -kBig;
                    X[L::TR + 2] = kBig;
                    X[L::TR + 3] = -kBig;
                }
                // the window edges after the trim
                const uint32_t f0 = (uint32_t)__builtin_amdgcn_readfirstlane(X[L::FW + PH * NW + 0]);
                const uint32_t f1 = (uint32_t)__builtin_amdgcn_readfirstlane(X[L::FW + PH * NW + NW - 1]);
                if (((f0 >> 8) & 1u) || ((f1 >> (16 + C - 1)) & 1u)) return TSU_EKOVERFLOW;
                return 0;
            }
            if (all & kFlEdge) return TSU_EKOVERFLOW;
            return 0;
        };

        int r = (p == t && ext0 == t) ? 1 : 0;  // 1: score 0 (identical sequences)
        while (r == 0) {
            [&]<int... I>(std::integer_sequence<int, I...>) __attribute__((always_inline)) {
                ((r = (r == 0 ? step(std::integral_constant<int, (I + 1) % RM>{}) : r)), ...);
            }(std::make_integer_sequence<int, RM>{});
        }
        const int score = (r > 0 && s == 0) ? 0 : r;
        if (tid == 0) a.scores[pi] = !lds_at_zero ? TSU_EINTERNAL : score < 0 ? score : score * a.g;
        __syncthreads();  // LDS reused by the next pair of this workgroup
    }
}

// LDS words of the exchange area (host side sizing).
template <int NW, int C, int RM>
constexpr int regmw_exchange_words() { return MwLayout<NW, C, RM>::WORDS; }

}  // namespace tsu
