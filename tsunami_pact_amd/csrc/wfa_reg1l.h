// wfa_reg1l.h -- single-wave register kernel with the older M scores in an
// int16 LDS ring.
//
// The same one-wave-per-pair kernel as wfa_reg.h (windows of 64..512
// diagonals, offsets kept doubled, fused extend), for the scaled penalties
// 2/3/1, with chunks chosen by the wavefront limits instead of per-chunk
// activity masks (scalar arithmetic per score instead of per chunk; the
// kernel issues as many scalar as vector instructions), and the M ring
// split as in the 16-chunk
// multi-wave kernel (wfa_regml.h): a lane keeps M of the two newest scores
// (two VGPRs per chunk, by the parity of s) and I and D; M[s-2], once read
// for the mismatch, goes to an int16 LDS ring of three slots, from which the
// gap-open values M[s-4][k-1] and M[s-4][k+1] are read at their diagonals.
// That replaces the four DPP moves per chunk and step that fetch the k-1 /
// k+1 neighbours and their chunk-edge fills from registers, and halves the
// state (4 instead of 7 VGPRs per chunk).  One wave owns the ring, and LDS
// operations of a wave complete in order, so no barrier is needed.
//
// Doubled offsets up to 2t must fit int16 (sequences of at most 16000
// bases; the host uses wfa_reg.h beyond).  NULL is -32768 in the ring and
// -2^30 + 2^15 in registers, whose low half is the ring's NULL, so an M
// (exactly NULL or an offset below 32768) goes to the ring as it is; a
// doubled value fed only by the ring's NULL grows by two per score, so the
// kernel gives up past kMl1MaxScore scaled scores with TSU_EKOVERFLOW.
//
// With BT every computed chunk also writes its 64 cell records (wfa_bt.h)
// at ((s * C + c) * 64 + lane) of the workgroup's region, one coalesced
// 256-byte store per chunk and step, and lane 0 walks them into CIGAR runs
// once the end cell is reached (the same records and walk as the 16-chunk
// kernel's, so the CIGARs are the same).
#pragma once

#include "wfa_regml.h"

namespace tsu {

// -32768 + 2 * kMl1MaxScore stays negative
constexpr int kMl1MaxScore = 16000;

// The plain 512-diagonal instance is held to 7 waves per SIMD (72 VGPRs, no
// spills): left alone the compiler takes 73, which the 8-register
// allocation granule turns into 80 and 6 waves.  The queue-fed instance
// spills under the same bound and keeps 6.
template <int C, bool QUEUE, bool BT = false>
__global__ __launch_bounds__(64, (C == 8 && !QUEUE && !BT) ? 7 : 1) void wfa_reg1l_kernel(KArgs a) {
    static_assert(C >= 1 && C <= 16, "window of 64..1024 diagonals");
    constexpr int32_t NUL = kRegNull + 0x8000;  // low 16 bits: -32768
    constexpr int W = 64 * C;
    constexpr int MRS = W + 2;  // int16 entries per ring slot, NULL guards at both ends
    extern __shared__ __align__(16) uint32_t smem[];
    const int lane = threadIdx.x;
    const int lane2 = 2 * lane;
    uint32_t* sS = smem;  // interleaved pattern / text words
    int32_t* sMisc = reinterpret_cast<int32_t*>(sS + 2 * a.seq_words);
    int16_t* MR = reinterpret_cast<int16_t*>(sMisc + 4);
    const bool lds_at_zero = (uint32_t)(uintptr_t)sS == 0u;
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
        for (int i = lane; i < (3 * MRS + 1) / 2; i += 64) mr32[i] = 0x80008000u;  // int16 NULLs
        __syncthreads();

        int kmin = max(-(W / 2), -p - 1);
        int kmax = kmin + W - 1;
        if (kmax > t + 1) { kmax = t + 1; kmin = max(kmax - W + 1, -p - 1); }
        const int tp = t - p;
        const int ctp = (tp - kmin) >= 0 ? (tp - kmin) >> 6 : -1;
        const int ltp = (tp - kmin) & 63;
        const uint32_t endm = ctp >= 0 && ctp < C ? 1u << ctp : 0u;  // chunk of the end cell
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
        // per chunk: one past the last doubled offset in bounds,
        // min(2(k + p + 1), 2(t + 1)), and two less (floored at 0: the only
        // diagonal with no cell in bounds, k = -p-1, holds 0 in both), kept
        // across the score loop.  A valid M below MT2m2 has bases left to
        // extend over, MT2m2 - M doubled.
        int32_t MT2[C], MT2m2[C];
#pragma unroll
        for (int c = 0; c < C; ++c) {
            MT2[c] = min(2 * (kmin + 64 * c) + lane2 + 2 * (p + 1), 2 * (t + 1));
            MT2m2[c] = max(MT2[c] - 2, 0);
        }
        // Window edges that are not bounds of the alignment: a non-NULL cell
        // in the window's first (last) lane then means the range may leave
        // the window.  Diagonals -p-1 and t+1 hold no cell in bounds.
        const uint32_t edges = __builtin_amdgcn_readfirstlane((kmin > -p - 1 ? 1u : 0u) | (kmax < t + 1 ? 2u : 0u));

        // Ranges of window diagonals (j = k - kmin) that may hold a non-NULL
        // cell, as WFA2-lib's wavefront limits (wfa_gpu.cu:426-468) for the
        // scaled penalties 2/3/1: I/D at s within [min(lo M[s-4], lo I/D[s-1])
        // - 1, max(hi M[s-4], hi I/D[s-1]) + 1], M at s the union of that
        // and M at s-2.  The ranges only grow (trimmed cells are NULLed but
        // the limits stay), so the range at s covers every chunk that held a
        // non-NULL cell at s-5 .. s-1 too, and its ring entries get
        // overwritten.  mt/mb hold the M limits of the last two scores of
        // each parity ([par][0] newest), it/ib those of I/D at s-1.
        constexpr int RNONE = 1 << 20;  // limits of an empty range: top -RNONE, bottom +RNONE
        int mt[2][2], mb[2][2];
        mt[0][0] = -kmin; mb[0][0] = -kmin;  // score 0: M[0][0] only
        mt[0][1] = -RNONE; mb[0][1] = RNONE;
        mt[1][0] = -RNONE; mb[1][0] = RNONE;
        mt[1][1] = -RNONE; mb[1][1] = RNONE;
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
        int l3 = 0;  // s mod 3

        // One score step; PAR = s & 1 selects the register slot holding
        // M[s-2], which receives M[s].  Returns 0 to go on, the score when
        // the end cell is reached, or a negative status.
        auto step = [&](auto PC) __attribute__((always_inline)) -> int {
            constexpr int PAR = decltype(PC)::value;
            ++s;
            if (s > s_bound) return TSU_EINTERNAL;
            if (s > kMl1MaxScore) return TSU_EKOVERFLOW;
            if constexpr (BT) {
                if (s >= a.step_cap) return TSU_ECIGAR;
            }
            l3 = l3 == 2 ? 0 : l3 + 1;
            const int rd = l3 == 0 ? 2 : l3 - 1;   // ring slot of s-4
            const int wrs = l3 == 2 ? 0 : l3 + 1;  // ring slot of s-2 (formerly s-5)
            // an opaque zero per step keeps the per-chunk bounds from being
            // hoisted out of the score loop into VGPRs
            int zero_s;
            asm volatile("s_mov_b32 %0, 0" : "=s"(zero_s));
            // the doubled pattern offset of a lane in chunk c is
            // 2v = M - 2k = M + nl2 + (nkb - 128 c); nl2 is recomputed per
            // step, so no per-chunk sum of it is held in a VGPR
            const int nl2 = zero_s - lane2;
            const int nkb = zero_s - 2 * kmin;
            // opaque per step: scalar bit tests per chunk
            uint32_t em = __builtin_amdgcn_readfirstlane(endm);
            asm volatile("" : "+s"(em));
            uint32_t eg = edges;
            asm volatile("" : "+s"(eg));
            // limits at s; chunk c computes when it meets the M range
            it = max(mt[PAR][1], it) + 1;
            ib = min(mb[PAR][1], ib) - 1;
            mt[PAR][1] = mt[PAR][0];
            mb[PAR][1] = mb[PAR][0];
            mt[PAR][0] = max(mt[PAR][0], it);
            mb[PAR][0] = min(mb[PAR][0], ib);
            const int jlo = max(mb[PAR][0], 0), jhi = min(mt[PAR][0], W - 1);
            const uint32_t act = jlo <= jhi ? (2u << (jhi >> 6)) - (1u << (jlo >> 6)) : 0u;
            const int16_t* mrd = MR + rd * MRS + lane;       // entry of k-1
            int16_t* mwr = MR + wrs * MRS + lane + 1;        // entry of k
            uint32_t* recs = nullptr;                        // records of this step
            if constexpr (BT) recs = rec + (int64_t)s * C * 64;
            uint64_t oob = 0;
            uint32_t edge = 0;  // a non-NULL cell in an edge lane of the window
            int32_t mend = NUL;
            // the previous chunk's old I rotated by one lane (its lane 63 in
            // lane 0); the computed chunks are contiguous, so it is NULL
            // before the first of them and written by each
            int32_t rotI = NUL;
#pragma unroll
            for (int c = 0; c < C; ++c) {
                if (act & (1u << c)) {
                    const int mt2 = MT2[c];
                    const int32_t mx = Mr[PAR][c];
                    mwr[64 * c] = (int16_t)mx;  // exactly NULL or an offset below 32768
                    const int32_t mol = mrd[64 * c];
                    const int32_t mor = mrd[64 * c + 2];
                    const int32_t iel = dpp_shr1(rotI, Ir[c]);
                    const int32_t der = dpp_shl1(c < C - 1 ? dpp_rol1(Dr[c < C - 1 ? c + 1 : c]) : NUL, Dr[c]);
                    if (c < C - 1) rotI = dpp_ror1(Ir[c]);
                    const int32_t I = max(mol, iel) + 2;
                    const int32_t D = max(mor, der);
                    const int32_t X = mx + 2;
                    const int32_t Mp = max(max(D, X), I);  // M before the bounds test
                    const bool valid = (uint32_t)Mp < (uint32_t)mt2;
                    int32_t M = valid ? Mp : NUL;
                    // Mp is out of bounds where I or D is, or the mismatch
                    // (then trimming finds nothing and changes nothing)
                    oob |= vote(Mp >= mt2);
                    // Mp is non-NULL where M, I or D is, or a mismatch from a
                    // non-NULL M[s-2] (an edge hit two scores earlier)
                    if (c == 0) edge |= (uint32_t)(vote(Mp >= 0) & 1ull) & eg;
                    if (c == C - 1) edge |= (uint32_t)(vote(Mp >= 0) >> 63) & (eg >> 1);
                    if constexpr (BT) {
                        // origin of M (ties M > D > I), gap extension bits, M before extension
                        const uint32_t rb = (iel >= mol ? 4u : 0u) | (der >= mor ? 8u : 0u);
                        recs[64 * c + lane] = rb | (valid ? (((uint32_t)Mp >> 1) << 4) |
                                                                ((X >= D && X >= I) ? 1u : (D >= I ? 2u : 3u))
                                                          : 0u);
                    }
                    const int mt2m2 = MT2m2[c];
                    int sc = nkb - 128 * c;  // one scalar operand: a three-way add
                    asm volatile("" : "+s"(sc));
                    if ((uint32_t)M < (uint32_t)mt2m2)  // valid, with bases left
                        M += extend2((uint32_t)M + (uint32_t)nl2 + (uint32_t)sc, (uint32_t)M, mt2m2 - M);
                    Mr[PAR][c] = M;
                    Ir[c] = I;
                    Dr[c] = D;
                    if (em & (1u << c)) mend = __builtin_amdgcn_readlane(M, ltp);
                }
            }
            if (oob) {
                // exact trim (wfa_gpu.cu:633-719): NULL I/D cells outside
                // [first, last] in-bounds cell, then test the edge lanes again
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
                edge = 0;
                edge |= (uint32_t)(vote(max(max(Mr[PAR][0], Ir[0]), Dr[0]) >= 0) & 1ull) & eg;
                edge |= (uint32_t)(vote(max(max(Mr[PAR][C - 1], Ir[C - 1]), Dr[C - 1]) >= 0) >> 63) & (eg >> 1);
            }
            if (mend >= 2 * t) return s;
            // a non-NULL cell at a window edge could extend the range beyond it
            if (edge) return TSU_EKOVERFLOW;
            return 0;
        };

        int r = (p == t && ext0 == t) ? 1 : 0;  // 1: score 0 (identical sequences)
        while (r == 0) {
            r = step(std::integral_constant<int, 1>{});
            if (r == 0) r = step(std::integral_constant<int, 0>{});
        }
        const int score = (r > 0 && s == 0) ? 0 : r;
        if constexpr (BT) {
            __syncthreads();  // every lane's records written
            if (lane == 0) {
                int st = !lds_at_zero ? TSU_EINTERNAL : score < 0 ? score : 0;
                int64_t out_off = 0;
                int32_t out_n = 0;
                if (st == 0) {
                    auto record = [&](int ss, int kk, uint32_t& rr) __attribute__((always_inline)) -> bool {
                        if (ss < 0 || ss > s) return false;
                        const int j = kk - kmin;
                        if (j < 0 || j >= W) return false;
                        rr = rec[((int64_t)ss * C + (j >> 6)) * 64 + (j & 63)];
                        return true;
                    };
                    st = bt_walk(record, score, p, t, 2, 3, 1, a, &out_off, &out_n);
                }
                a.run_off[pi] = out_off;
                a.run_n[pi] = out_n;
                a.scores[pi] = st ? st : score * a.g;
            }
        } else {
            if (lane == 0) a.scores[pi] = !lds_at_zero ? TSU_EINTERNAL : score < 0 ? score : score * a.g;
        }
        __syncthreads();
    }
}

}  // namespace tsu
