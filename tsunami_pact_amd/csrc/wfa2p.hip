// wfa2p.hip -- gap-affine 2-piece WFA kernel (one wave64 per pair).
//
// Gap-affine 2-piece scoring (a gap of length L costs min(o1 + L e1,
// o2 + L e2)): five wavefronts M, I1, D1, I2, D2.  The reference declares
// them (common.h:52-58,80-87) but computes only one gap piece; the step
// follows WFA2-lib's affine-2p one, restated in oracle/wfa2p_oracle.c:
//
//   I1 = max(M[s-o1-e1][k-1], I1[s-e1][k-1]) + 1   I2 likewise with o2, e2
//   D1 = max(M[s-o1-e1][k+1], D1[s-e1][k+1])       D2 likewise
//   M  = max(M[s-x][k] + 1, I1, I2, D1, D2), NULL out of bounds
//
// with the range union, null steps, existence of I/D from their sources and
// trim-ends of the gap-affine kernels (wfa_gpu.cu:347-468,610-719).
//
// Layout (LDS, one pair per workgroup): the interleaved packed sequences at
// address 0 (as every kernel here, so the extend addresses them directly),
// 8 misc words, one (lo, hi) pair per ring slot, then the rings: M keeps
// max(x, o1+e1, o2+e2)+1 scores, I1/D1 e1+1, I2/D2 e2+1, plus one slot that
// stays NULL (the source of scores below 0); each slot holds W+2 entries,
// diagonal k at k - kmin + 1 with NULL guards at both ends.  Entries are
// int16 for sequences up to 32000 bases, int32 beyond.  Stored values are
// canonical: a negative value is stored as NULL and an I above t as t+1,
// which keeps every in-bounds cell and the out-of-bounds classification of
// every other cell (so int16 NULLs cannot drift towards valid offsets).  Every
// slot holds NULL outside its current range: a step writes its outputs over
// the union of the new range and the old contents' ranges, and a trim writes
// NULL over the cells it cuts.  One wave owns everything, and LDS operations
// of a wave complete in order, so no barrier is needed within a pair.
// Penalties are runtime values (divided by their gcd on the host).
//
// With BT every cell of the range [lo, hi] of a non-null score also writes
// its record (wfa_bt.h, 2-piece format) to global memory, the scores' cells
// packed one after another (score s from step_off[s], first diagonal
// step_lo[s]), and lane 0 walks them into CIGAR runs once the end cell is
// reached.
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "wfa_bt.h"
#include "wfa_reg.h"

namespace tsu {

template <typename OffT> struct P2Null;
template <> struct P2Null<int16_t> { static constexpr int32_t v = -16384; };
template <> struct P2Null<int32_t> { static constexpr int32_t v = -1073741824; };

struct P2Src {
    int lo, hi;  // lo > hi: null
};

template <typename OffT, bool QUEUE, bool BT = false>
__global__ __launch_bounds__(64) void wfa2p_kernel(KArgs a) {
    constexpr int32_t NUL = P2Null<OffT>::v;
    extern __shared__ __align__(16) uint32_t smem[];
    const int lane = threadIdx.x;
    uint32_t* sS = smem;
    int32_t* misc = reinterpret_cast<int32_t*>(smem + 2 * a.seq_words);
    const int RM = a.rm, R1 = a.ri, R2 = a.r2;
    const int NS = RM + 2 * R1 + 2 * R2;  // slot NS is the permanent NULL slot
    int2* lim = reinterpret_cast<int2*>(misc + 8);
    const int W = a.W;          // diagonal capacity
    const int STR = W + 2;
    OffT* ring = reinterpret_cast<OffT*>(lim + NS + 1);
    const bool lds_at_zero = (uint32_t)(uintptr_t)sS == 0u;
    const int x = a.x, oe1 = a.o + a.e, oe2 = a.o2 + a.e2, e1 = a.e, e2 = a.e2;
    uint32_t* rec = nullptr;
    int32_t* stepLo = nullptr;
    int64_t* stepOff = nullptr;
    if constexpr (BT) {
        rec = a.rec_g + (int64_t)blockIdx.x * a.rec_cap;
        stepLo = a.step_lo_g + (int64_t)blockIdx.x * a.step_cap;
        stepOff = a.step_off_g + (int64_t)blockIdx.x * a.step_cap;
    }

    for (int iter = 0;; ++iter) {
        int idx;
        if constexpr (QUEUE) {
            if (lane == 0) misc[0] = atomicAdd(a.queue, 1);
            __syncthreads();
            idx = __builtin_amdgcn_readfirstlane(misc[0]);
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
        for (int i = lane; i < (NS + 1) * STR; i += 64) ring[i] = (OffT)NUL;
        for (int i = lane; i <= NS; i += 64) lim[i] = make_int2(1, 0);
        __syncthreads();

        int kmin = max(-(W / 2), -p - 1);
        int kmax = kmin + W - 1;
        if (kmax > t + 1) { kmax = t + 1; kmin = max(kmax - W + 1, -p - 1); }
        const int tp = t - p;
        const int d = abs(p - t);
        const int64_t tail = p != t ? min((int64_t)a.o + (int64_t)e1 * d, (int64_t)a.o2 + (int64_t)e2 * d) : 0;
        const int s_bound = (int)min<int64_t>((int64_t)x * min(p, t) + tail, 0x3fffffff);
        auto R = [&](int g) __attribute__((always_inline)) { return ring + g * STR + (1 - kmin); };  // indexed by k
        auto getlim = [&](int g) __attribute__((always_inline)) {
            const int2 v = lim[g];
            return P2Src{__builtin_amdgcn_readfirstlane(v.x), __builtin_amdgcn_readfirstlane(v.y)};
        };

        // score 0: M[0][0], extended
        const int ext0 = lcp_interleaved(sS, 0, 0, min(p, t));
        if (lane == 0) { R(0)[0] = (OffT)ext0; lim[0] = make_int2(0, 0); }
        __syncthreads();
        int status = (p == t && ext0 == t) ? 1 : 0;  // 1: score 0
        int s = 0;
        int64_t rec_next = 1;  // record 0: score 0 (never read)
        if constexpr (BT) {
            if (lane == 0) { rec[0] = 0u; stepLo[0] = 0; stepOff[0] = 0; }
        }
        int jm = 0, j1 = 0, j2 = 0;  // slots of score s in the M, I1/D1 and I2/D2 rings
        while (status == 0) {
            ++s;
            if (s > s_bound) { status = TSU_EINTERNAL; break; }
            jm = jm + 1 == RM ? 0 : jm + 1;
            j1 = j1 + 1 == R1 ? 0 : j1 + 1;
            j2 = j2 + 1 == R2 ? 0 : j2 + 1;
            // source slots (the NULL slot below score 0)
            const int gx = s >= x ? (jm >= x ? jm - x : jm - x + RM) : NS;
            const int go1 = s >= oe1 ? (jm >= oe1 ? jm - oe1 : jm - oe1 + RM) : NS;
            const int go2 = s >= oe2 ? (jm >= oe2 ? jm - oe2 : jm - oe2 + RM) : NS;
            const int s1 = j1 + 1 == R1 ? 0 : j1 + 1;  // score s-e1
            const int s2 = j2 + 1 == R2 ? 0 : j2 + 1;  // score s-e2
            const int gi1 = s >= e1 ? RM + s1 : NS, gd1 = s >= e1 ? RM + R1 + s1 : NS;
            const int gi2 = s >= e2 ? RM + 2 * R1 + s2 : NS, gd2 = s >= e2 ? RM + 2 * R1 + R2 + s2 : NS;
            const P2Src Lx = getlim(gx), Lo1 = getlim(go1), Lo2 = getlim(go2);
            const P2Src Li1 = getlim(gi1), Ld1 = getlim(gd1), Li2 = getlim(gi2), Ld2 = getlim(gd2);
            // output slots and what they still hold (scores s-RM, s-R1, s-R2)
            const int gout[5] = {jm, RM + j1, RM + R1 + j1, RM + 2 * R1 + j2, RM + 2 * R1 + R2 + j2};
            int clo = 0x3fffffff, chi = -0x3fffffff;
#pragma unroll
            for (int c = 0; c < 5; ++c) {
                const P2Src o = getlim(gout[c]);
                if (o.lo <= o.hi) { clo = min(clo, o.lo); chi = max(chi, o.hi); }
            }
            const bool nx = Lx.lo > Lx.hi, no1 = Lo1.lo > Lo1.hi, no2 = Lo2.lo > Lo2.hi;
            const bool ni1 = Li1.lo > Li1.hi, nd1 = Ld1.lo > Ld1.hi, ni2 = Li2.lo > Li2.hi, nd2 = Ld2.lo > Ld2.hi;
            int lo = 0x3fffffff, hi = -0x3fffffff;
            if (!nx) { lo = min(lo, Lx.lo); hi = max(hi, Lx.hi); }
            if (!no1) { lo = min(lo, Lo1.lo - 1); hi = max(hi, Lo1.hi + 1); }
            if (!no2) { lo = min(lo, Lo2.lo - 1); hi = max(hi, Lo2.hi + 1); }
            if (!ni1) { lo = min(lo, Li1.lo + 1); hi = max(hi, Li1.hi + 1); }
            if (!ni2) { lo = min(lo, Li2.lo + 1); hi = max(hi, Li2.hi + 1); }
            if (!nd1) { lo = min(lo, Ld1.lo - 1); hi = max(hi, Ld1.hi - 1); }
            if (!nd2) { lo = min(lo, Ld2.lo - 1); hi = max(hi, Ld2.hi - 1); }
            const bool null_step = lo > hi;  // every source null (or an empty union)
            if (!null_step && (lo < kmin || hi > kmax)) { status = TSU_EKOVERFLOW; break; }
            uint32_t* recs = nullptr;
            if constexpr (BT) {
                const int64_t width = null_step ? 0 : hi - lo + 1;
                if (s >= a.step_cap || rec_next + width > a.rec_cap) { status = TSU_ECIGAR; break; }
                if (lane == 0) { stepLo[s] = null_step ? 0 : lo; stepOff[s] = rec_next; }
                recs = rec + rec_next - lo;  // indexed by k
                rec_next += width;
            }
            const bool h_i1 = !null_step && (!no1 || !ni1), h_d1 = !null_step && (!no1 || !nd1);
            const bool h_i2 = !null_step && (!no2 || !ni2), h_d2 = !null_step && (!no2 || !nd2);
            if (!null_step) { clo = min(clo, lo); chi = max(chi, hi); }
            const OffT* Mx = R(gx);
            const OffT* Mo1 = R(go1);
            const OffT* Mo2 = R(go2);
            const OffT* I1e = R(gi1);
            const OffT* D1e = R(gd1);
            const OffT* I2e = R(gi2);
            const OffT* D2e = R(gd2);
            OffT* oM = R(gout[0]);
            OffT* oI1 = R(gout[1]);
            OffT* oD1 = R(gout[2]);
            OffT* oI2 = R(gout[3]);
            OffT* oD2 = R(gout[4]);
            int32_t mend = NUL;
            for (int cb = clo; cb <= chi; cb += 64) {
                const int k = cb + lane;
                const bool act = k <= chi;
                const int kk = act ? k : chi;  // idle lanes read a cell of the region
                const bool in = k >= lo && k <= hi;
                const int32_t mo1l = Mo1[kk - 1], i1el = I1e[kk - 1], mo2l = Mo2[kk - 1], i2el = I2e[kk - 1];
                const int32_t mo1r = Mo1[kk + 1], d1er = D1e[kk + 1], mo2r = Mo2[kk + 1], d2er = D2e[kk + 1];
                const int32_t i1p = max(mo1l, i1el);
                const int32_t i2p = max(mo2l, i2el);
                int32_t i1 = i1p >= 0 ? min(i1p + 1, t + 1) : NUL;
                int32_t i2 = i2p >= 0 ? min(i2p + 1, t + 1) : NUL;
                int32_t d1 = max(mo1r, d1er);
                int32_t d2 = max(mo2r, d2er);
                const int32_t mxv = Mx[kk];
                const int32_t xv = mxv >= 0 ? mxv + 1 : NUL;
                int32_t m = max(max(xv, max(i1, i2)), max(d1, d2));
                const bool valid = in && (uint32_t)m <= (uint32_t)t && (uint32_t)(m - k) <= (uint32_t)p;
                if constexpr (BT) {
                    if (in) {
                        // origin of M (ties X > D2 > D1 > I2 > I1), extension bits, M before extension
                        uint32_t r = (i1el >= mo1l ? 8u : 0u) | (d1er >= mo1r ? 16u : 0u) |
                                     (i2el >= mo2l ? 32u : 0u) | (d2er >= mo2r ? 64u : 0u);
                        if (valid) {
                            const uint32_t org = (xv >= d2 && xv >= d1 && xv >= i2 && xv >= i1) ? 1u
                                                 : (d2 >= d1 && d2 >= i2 && d2 >= i1) ? 2u
                                                 : (d1 >= i2 && d1 >= i1) ? 3u
                                                 : (i2 >= i1) ? 4u : 5u;
                            r |= ((uint32_t)m << 7) | org;
                        }
                        recs[k] = r;
                    }
                }
                if (!valid) m = NUL;
                if (!in || !h_i1) i1 = NUL;
                if (!in || !h_d1) d1 = NUL;
                if (!in || !h_i2) i2 = NUL;
                if (!in || !h_d2) d2 = NUL;
                if (valid) {
                    const int lm = min(p - (m - k), t - m);
                    if (lm > 0) m += extend_len((uint32_t)(m - k), (uint32_t)m, lm);
                }
                if (act) {
                    oM[k] = (OffT)m;
                    oI1[k] = (OffT)i1;
                    oD1[k] = (OffT)d1;
                    oI2[k] = (OffT)i2;
                    oD2[k] = (OffT)d2;
                }
                if (tp >= cb && tp < cb + 64) mend = __builtin_amdgcn_readlane(m, tp - cb);
            }
            if (null_step) {
#pragma unroll
                for (int c = 0; c < 5; ++c)
                    if (lane == 0) lim[gout[c]] = make_int2(1, 0);
                __syncthreads();
                continue;
            }
            if (tp >= lo && tp <= hi && mend >= t) { status = 2; break; }
            // trim-ends on each wavefront that exists: first / last in-bounds
            // cell, NULL over the cells cut
            const bool has[5] = {true, h_i1, h_d1, h_i2, h_d2};
#pragma unroll
            for (int c = 0; c < 5; ++c) {
                OffT* Rc = R(gout[c]);
                int f = hi + 1, l = lo - 1;
                if (has[c]) {
                    for (int kb = lo; kb <= hi; kb += 64) {
                        const int k = kb + lane;
                        const int32_t v = Rc[min(k, hi)];
                        const bool ok = k <= hi && (uint32_t)v <= (uint32_t)t && (uint32_t)(v - k) <= (uint32_t)p;
                        const uint64_t b = __ballot(ok);
                        if (b) { f = kb + (int)__builtin_ctzll(b); break; }
                    }
                    if (f <= hi) {
                        for (int kt = hi;; kt -= 64) {
                            const int k = kt - lane;
                            const int32_t v = Rc[max(k, f)];
                            const bool ok = k >= f && (uint32_t)v <= (uint32_t)t && (uint32_t)(v - k) <= (uint32_t)p;
                            const uint64_t b = __ballot(ok);
                            if (b) { l = kt - (int)__builtin_ctzll(b); break; }
                        }
                    }
                    for (int k = lo + lane; k < f; k += 64) Rc[k] = (OffT)NUL;
                    for (int k = max(l + 1, lo) + lane; k <= hi; k += 64) Rc[k] = (OffT)NUL;
                }
                if (lane == 0) lim[gout[c]] = has[c] && f <= l ? make_int2(f, l) : make_int2(1, 0);
            }
            __syncthreads();
        }
        const int score = status == 1 ? 0 : status == 2 ? s : status;
        if constexpr (BT) {
            __syncthreads();  // every lane's records written
            if (lane == 0) {
                int st = !lds_at_zero ? TSU_EINTERNAL : score < 0 ? score : 0;
                int64_t out_off = 0;
                int32_t out_n = 0;
                if (st == 0) {
                    auto record = [&](int ss, int kk, uint32_t& rr) __attribute__((always_inline)) -> bool {
                        if (ss <= 0 || ss > s) return false;
                        const int l0 = stepLo[ss];
                        const int64_t o0 = stepOff[ss];
                        const int64_t o1 = ss < s ? stepOff[ss + 1] : rec_next;
                        const int64_t at = o0 + (kk - l0);
                        if (kk < l0 || at >= o1) return false;
                        rr = rec[at];
                        return true;
                    };
                    st = bt2p_walk(record, score, p, t, x, a.o, e1, a.o2, e2, a, &out_off, &out_n);
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

// the register kernel's instances, per (e1, e2) after the gcd (wfa2p_reg_eXY.hip)
KernelFn pick_2p_reg_11(const Shape& sh);
KernelFn pick_2p_reg_21(const Shape& sh);
KernelFn pick_2p_reg_31(const Shape& sh);

KernelFn pick_2p(const Shape& sh) {
    if (sh.reg_c > 0) {  // the single-wave register kernel
        switch (sh.reg_e) {
        case 11: return pick_2p_reg_11(sh);
        case 21: return pick_2p_reg_21(sh);
        case 31: return pick_2p_reg_31(sh);
        default: return nullptr;
        }
    }
    if (sh.bt) {  // CIGAR launches: records, queue-fed
        if (sh.ck || !sh.queue) return nullptr;
        return sh.off16 ? wfa2p_kernel<int16_t, true, true> : wfa2p_kernel<int32_t, true, true>;
    }
    if (sh.off16) return sh.queue ? wfa2p_kernel<int16_t, true> : wfa2p_kernel<int16_t, false>;
    return sh.queue ? wfa2p_kernel<int32_t, true> : wfa2p_kernel<int32_t, false>;
}

// LDS bytes of the 2-piece kernels: sequences, misc, slot limits, rings
// (the register kernel: sequences, misc and the int16 M ring).
int64_t lds_bytes_2p(const Shape& sh, const KArgs& a) {
    if (sh.reg_c > 0) {
        const int64_t ring = (int64_t)a.rm * (64ll * sh.reg_c + 2) * 2;
        return 4 * (2 * (int64_t)a.seq_words + 4) + ((ring + 3) & ~3ll);
    }
    const int64_t ns = (int64_t)a.rm + 2 * a.ri + 2 * a.r2 + 1;
    const int64_t ring = ns * ((int64_t)a.W + 2) * (sh.off16 ? 2 : 4);
    return 4 * (2 * (int64_t)a.seq_words + 8) + 8 * ns + ((ring + 3) & ~3ll);
}

}  // namespace tsu
