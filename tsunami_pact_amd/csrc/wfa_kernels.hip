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
// Backtrace (BT): each computed cell writes one 32-bit record
//   bits 0-1  origin of M[s][k] before extension: 1 X, 2 D, 3 I (0 = NULL)
//             (ties resolved M > D > I, with D = max(D_open, D_ext) and
//              I = max(I_open, I_ext): the piggy-backed order of common.h:67-77)
//   bit  2    I[s][k] came from I[s-e][k-1] (extension) rather than M_o
//   bit  3    D[s][k] came from D[s-e][k+1] (extension) rather than M_o
//   bits 4-31 M[s][k] before extension
// and thread 0 walks them back from the end cell into run-length CIGARs.
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "tsu_internal.h"

namespace tsu {

template <typename OffT> struct Null;
template <> struct Null<int16_t> { static constexpr int32_t v = -16384; };
template <> struct Null<int32_t> { static constexpr int32_t v = -1073741824; };

constexpr int kSmemInts = 16;  // misc uniform values in LDS

// Longest common prefix of pattern[v..] and text[h..], at most lim bases.
// P and T are 2-bit packed (16 bases per word, LSB first) with two zero
// words of padding after the last base, so index j+2 is always in bounds.
__device__ __forceinline__ int lcp_packed(const uint32_t* __restrict__ P, const uint32_t* __restrict__ T,
                                          int v, int h, int lim) {
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

// Entries per ring slot: W diagonals + 2 NULL guards, padded to 16 bytes.
template <typename OffT>
__host__ __device__ __forceinline__ int ring_stride(int W) {
    constexpr int per16 = 16 / (int)sizeof(OffT);
    return (W + 2 + per16 - 1) / per16 * per16;
}

// LDS bytes before the ring: two sequences, slot limits, reduction area and
// misc values, rounded up to 16 bytes.
__host__ __device__ __forceinline__ int64_t lds_head_bytes(int seq_words, int nslot, int nw) {
    const int64_t b = 4ll * (2 * seq_words + 2 * nslot + nw * 8 + kSmemInts);
    return (b + 15) / 16 * 16;
}

template <int NW, typename OffT, bool LDS_RING, bool BT, bool QUEUE>
__global__ __launch_bounds__(NW * 64) void wfa_kernel(KArgs a) {
    extern __shared__ __align__(16) uint32_t smem[];
    constexpr int NT = NW * 64;
    constexpr int32_t NUL = Null<OffT>::v;
    const int tid = threadIdx.x;
    const int lane = tid & 63;
    const int wave = tid >> 6;

    const int rm = a.rm, ri = a.ri;
    const int nslot = rm + 2 * ri;
    uint32_t* sP = smem;
    uint32_t* sT = smem + a.seq_words;
    int32_t* sLo = reinterpret_cast<int32_t*>(smem + 2 * a.seq_words);
    int32_t* sHi = sLo + nslot;
    int32_t* sRed = sHi + nslot;           // NW * 8
    int32_t* sMisc = sRed + NW * 8;        // kSmemInts
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

    const int x = a.x, o = a.o, e = a.e;
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

        // ---- stage sequences, clear the ring --------------------------------
        for (int i = tid; i < a.seq_words; i += NT) {
            sP[i] = i < pw ? a.words[pr.p_off + i] : 0u;
            sT[i] = i < tw ? a.words[pr.t_off + i] : 0u;
        }
        {
            // ring slots are padded to a multiple of 16 bytes: fill with uint4
            const int64_t nvec = (int64_t)nslot * stride * (int64_t)sizeof(OffT) / 16;
            uint32_t nw32 = sizeof(OffT) == 2 ? ((uint32_t)(uint16_t)NUL | ((uint32_t)(uint16_t)NUL << 16))
                                              : (uint32_t)NUL;
            uint4 fill = make_uint4(nw32, nw32, nw32, nw32);
            uint4* r4 = reinterpret_cast<uint4*>(ring);
            for (int64_t i = tid; i < nvec; i += NT) r4[i] = fill;
        }
        for (int i = tid; i < nslot; i += NT) { sLo[i] = 1; sHi[i] = 0; }
        __syncthreads();

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

        // ---- score 0: M[0][0] = 0, extended --------------------------------
        if (tid == 0) {
            const int ext = lcp_packed(sP, sT, 0, 0, min(p, t));
            ring[0 * stride + (0 - kmin + 1)] = (OffT)ext;
            sLo[0] = 0; sHi[0] = 0;
            sMisc[1] = (p == t && ext == t) ? 1 : 0;
            if constexpr (BT) { rec[0] = 0u; stepLo[0] = 0; stepOff[0] = 0; }
        }
        if constexpr (BT) rec_next = 1;
        __syncthreads();
        if (uni(sMisc[1])) score = 0;

        int s = 0;
        while (score < 0) {
            ++s;
            if (s > s_bound) { status = TSU_EINTERNAL; break; }
            const int sm = s - x, so = s - o - e, se = s - e;
            const int slM = s % rm, slI = rm + s % ri, slD = rm + ri + s % ri;
            int lo_x = 1, hi_x = 0, lo_o = 1, hi_o = 0, lo_i = 1, hi_i = 0, lo_d = 1, hi_d = 0;
            if (sm >= 0) { lo_x = uni(sLo[sm % rm]); hi_x = uni(sHi[sm % rm]); }
            if (so >= 0) { lo_o = uni(sLo[so % rm]); hi_o = uni(sHi[so % rm]); }
            if (se >= 0) {
                lo_i = uni(sLo[rm + se % ri]); hi_i = uni(sHi[rm + se % ri]);
                lo_d = uni(sLo[rm + ri + se % ri]); hi_d = uni(sHi[rm + ri + se % ri]);
            }
            const bool nx = lo_x > hi_x, no = lo_o > hi_o, ni = lo_i > hi_i, nd = lo_d > hi_d;
            const int olo_M = uni(sLo[slM]), ohi_M = uni(sHi[slM]);
            const int olo_I = uni(sLo[slI]), ohi_I = uni(sHi[slI]);
            const int olo_D = uni(sLo[slD]), ohi_D = uni(sHi[slD]);
            OffT* Mt = ring + slM * stride;
            OffT* It = ring + slI * stride;
            OffT* Dt = ring + slD * stride;
            auto clear_range = [&](OffT* sl, int from, int to) {
                for (int k = from + tid; k <= to; k += NT) sl[k - kmin + 1] = (OffT)NUL;
            };
            auto clear_except = [&](OffT* sl, int alo, int ahi, int klo, int khi) {
                if (alo > ahi) return;
                if (klo > khi) { clear_range(sl, alo, ahi); return; }
                clear_range(sl, alo, min(ahi, klo - 1));
                clear_range(sl, max(alo, khi + 1), ahi);
            };
            if (nx && no && ni && nd) {
                // null step (wfa_gpu.cu:417-422): all three outputs are null
                clear_except(Mt, olo_M, ohi_M, 1, 0);
                clear_except(It, olo_I, ohi_I, 1, 0);
                clear_except(Dt, olo_D, ohi_D, 1, 0);
                if (tid == 0) {
                    sLo[slM] = 1; sHi[slM] = 0; sLo[slI] = 1; sHi[slI] = 0; sLo[slD] = 1; sHi[slD] = 0;
                }
                if constexpr (BT) {
                    if (s >= a.step_cap) { status = TSU_ECIGAR; break; }
                    if (tid == 0) { stepLo[s] = 0; stepOff[s] = rec_next; }
                }
                __syncthreads();
                continue;
            }
            int lo = 0, hi = 0;
            {
                bool have = false;
                if (!nx) { lo = lo_x; hi = hi_x; have = true; }
                if (!no) { lo = have ? min(lo, lo_o - 1) : lo_o - 1; hi = have ? max(hi, hi_o + 1) : hi_o + 1; have = true; }
                if (!ni) { lo = have ? min(lo, lo_i + 1) : lo_i + 1; hi = have ? max(hi, hi_i + 1) : hi_i + 1; have = true; }
                if (!nd) { lo = have ? min(lo, lo_d - 1) : lo_d - 1; hi = have ? max(hi, hi_d - 1) : hi_d - 1; }
            }
            if (lo < kmin || hi > kmax) { status = TSU_EKOVERFLOW; break; }
            const int width = hi - lo + 1;
            if constexpr (BT) {
                if (s >= a.step_cap || rec_next + width > a.rec_cap) { status = TSU_ECIGAR; break; }
            }
            const bool hasI = !no || !ni;
            const bool hasD = !no || !nd;
            const OffT* Mx = ring + (nx ? 0 : (sm % rm)) * stride;
            const OffT* Mo = ring + (no ? 0 : (so % rm)) * stride;
            const OffT* Ie = ring + (ni ? rm : (rm + se % ri)) * stride;
            const OffT* De = ring + (nd ? rm + ri : (rm + ri + se % ri)) * stride;
            uint32_t* recs = nullptr;
            if constexpr (BT) recs = rec + rec_next;

            int fM = 0x7fffffff, lM = -0x7fffffff, fI = 0x7fffffff, lI = -0x7fffffff, fD = 0x7fffffff, lD = -0x7fffffff;
            int fin = 0;
            for (int kb = lo + wave * 64; kb <= hi; kb += NW * 64) {
                const int k = kb + lane;
                const bool act = k <= hi;
                int32_t M = NUL, I = NUL, D = NUL;
                bool inM = false, inI = false, inD = false, end = false;
                if (act) {
                    const int j = k - kmin + 1;
                    const int32_t mol = no ? NUL : (int32_t)Mo[j - 1];
                    const int32_t mor = no ? NUL : (int32_t)Mo[j + 1];
                    const int32_t iel = ni ? NUL : (int32_t)Ie[j - 1];
                    const int32_t der = nd ? NUL : (int32_t)De[j + 1];
                    const int32_t mx = nx ? NUL : (int32_t)Mx[j];
                    const int32_t ipre = max(mol, iel);
                    I = ipre >= 0 ? min(ipre + 1, t + 1) : NUL;
                    D = max(mor, der);
                    const int32_t X = mx >= 0 ? mx + 1 : NUL;
                    M = max(D, max(X, I));
                    if ((uint32_t)M > (uint32_t)t || (uint32_t)(M - k) > (uint32_t)p) M = NUL;
                    if constexpr (BT) {
                        uint32_t r = 0;
                        if (M >= 0) {
                            const uint32_t mw = (X >= D && X >= I) ? 1u : (D >= I ? 2u : 3u);
                            r = ((uint32_t)M << 4) | mw;
                        }
                        if (iel >= mol) r |= 4u;
                        if (der >= mor) r |= 8u;
                        recs[k - lo] = r;
                    }
                    if (M >= 0) {
                        const int v = M - k;
                        M += lcp_packed(sP, sT, v, M, min(p - v, t - M));
                        inM = true;
                        end = (k == tp) && (M >= t);
                    }
                    inI = (uint32_t)I <= (uint32_t)t && (uint32_t)(I - k) <= (uint32_t)p;
                    inD = (uint32_t)D <= (uint32_t)t && (uint32_t)(D - k) <= (uint32_t)p;
                    Mt[j] = (OffT)M;
                    if (hasI) It[j] = (OffT)I;
                    if (hasD) Dt[j] = (OffT)D;
                }
                const uint64_t bM = __ballot(inM), bI = __ballot(inI), bD = __ballot(inD);
                if (bM) { if (fM == 0x7fffffff) fM = kb + __builtin_ctzll(bM); lM = kb + 63 - __builtin_clzll(bM); }
                if (bI) { if (fI == 0x7fffffff) fI = kb + __builtin_ctzll(bI); lI = kb + 63 - __builtin_clzll(bI); }
                if (bD) { if (fD == 0x7fffffff) fD = kb + __builtin_ctzll(bD); lD = kb + 63 - __builtin_clzll(bD); }
                if (__ballot(end)) fin = 1;
            }
            if constexpr (NW > 1) {
                if (lane == 0) {
                    int32_t* r = sRed + wave * 8;
                    r[0] = fM; r[1] = lM; r[2] = fI; r[3] = lI; r[4] = fD; r[5] = lD; r[6] = fin;
                }
                __syncthreads();
                fM = 0x7fffffff; lM = -0x7fffffff; fI = 0x7fffffff; lI = -0x7fffffff; fD = 0x7fffffff; lD = -0x7fffffff; fin = 0;
                for (int w = 0; w < NW; ++w) {
                    const int32_t* r = sRed + w * 8;
                    fM = min(fM, uni(r[0])); lM = max(lM, uni(r[1]));
                    fI = min(fI, uni(r[2])); lI = max(lI, uni(r[3]));
                    fD = min(fD, uni(r[4])); lD = max(lD, uni(r[5]));
                    fin |= uni(r[6]);
                }
            }
            // trimmed ranges (empty -> lo=1, hi=0)
            int tloM = 1, thiM = 0, tloI = 1, thiI = 0, tloD = 1, thiD = 0;
            if (fM <= lM) { tloM = fM; thiM = lM; }
            if (hasI && fI <= lI) { tloI = fI; thiI = lI; }
            if (hasD && fD <= lD) { tloD = fD; thiD = lD; }
            __syncthreads();  // every cell of step s is written before clearing
            clear_except(Mt, olo_M, ohi_M, tloM, thiM);
            if (hasI) clear_except(It, lo, hi, tloI, thiI);
            clear_except(It, olo_I, ohi_I, tloI, thiI);
            if (hasD) clear_except(Dt, lo, hi, tloD, thiD);
            clear_except(Dt, olo_D, ohi_D, tloD, thiD);
            if (tid == 0) {
                sLo[slM] = tloM; sHi[slM] = thiM;
                sLo[slI] = tloI; sHi[slI] = thiI;
                sLo[slD] = tloD; sHi[slD] = thiD;
                if constexpr (BT) { stepLo[s] = lo; stepOff[s] = rec_next; }
            }
            if constexpr (BT) rec_next += width;
            __syncthreads();
            if (fin) score = s;
        }

        // ---- outputs --------------------------------------------------------
        if constexpr (BT) {
            if (tid == 0) {
                int32_t st = status;
                int64_t out_off = 0;
                int32_t out_n = 0;
                if (st == 0) {
                    const int unit = min(x, o + e);
                    const int64_t need = 2 * (int64_t)(score / unit) + 2;
                    const unsigned long long base = atomicAdd(a.runs_used, (unsigned long long)need);
                    if ((int64_t)base + need > a.runs_cap) {
                        st = TSU_ECIGAR;
                    } else {
                        uint32_t* out = a.runs + base;
                        int64_t nr = 0;
                        uint32_t cur_op = 4, cur_len = 0;
                        bool full = false;
                        auto put = [&](uint32_t run) {
                            if (nr >= need) { full = true; return; }
                            out[need - 1 - nr] = run;
                            ++nr;
                        };
                        auto emit = [&](uint32_t op, int n) {
                            if (n <= 0) return;
                            if (op == cur_op) { cur_len += (uint32_t)n; return; }
                            if (cur_len) put((cur_len << 2) | cur_op);
                            cur_op = op; cur_len = (uint32_t)n;
                        };
                        auto record = [&](int ss, int kk, uint32_t& r) -> bool {
                            if (ss < 0 || ss > s) return false;
                            const int l0 = stepLo[ss];
                            const int64_t o0 = stepOff[ss];
                            const int64_t o1 = (ss < s) ? stepOff[ss + 1] : rec_next;
                            const int64_t at = o0 + (kk - l0);
                            if (kk < l0 || at >= o1) return false;
                            r = rec[at];
                            return true;
                        };
                        int k = tp, off = t, ss = score, type = 0;  // 0 M, 1 I, 2 D
                        int v = p, h = t;
                        while (v > 0 && h > 0 && ss > 0) {
                            uint32_t r;
                            if (!record(ss, k, r)) { st = TSU_EINTERNAL; break; }
                            if (type == 0) {
                                const int pre = (int)(r >> 4);
                                const uint32_t mw = r & 3u;
                                if (mw == 0 || pre > off) { st = TSU_EINTERNAL; break; }
                                emit(0, off - pre);
                                off = pre; v = off - k; h = off;
                                if (v <= 0 || h <= 0) break;
                                if (mw == 1) { emit(1, 1); ss -= x; --off; v = off - k; h = off; continue; }
                                type = (mw == 2) ? 2 : 1;
                            }
                            if (type == 1) {
                                const bool ext = (r & 4u) != 0;
                                emit(2, 1); --k; --off;
                                ss -= ext ? e : (o + e);
                                type = ext ? 1 : 0;
                            } else {
                                const bool ext = (r & 8u) != 0;
                                emit(3, 1); ++k;
                                ss -= ext ? e : (o + e);
                                type = ext ? 2 : 0;
                            }
                            v = off - k; h = off;
                        }
                        if (st == 0) {
                            if (type == 0) {
                                if (v > 0 && h > 0) { const int nm = min(v, h); emit(0, nm); v -= nm; h -= nm; }
                                emit(3, v);
                                emit(2, h);
                            } else if (v != 0 || h != 0) {
                                st = TSU_EINTERNAL;
                            }
                        }
                        if (st == 0) {
                            if (cur_len) put((cur_len << 2) | cur_op);
                            if (full) {
                                st = TSU_ECIGAR;
                            } else {
                                out_off = (int64_t)base + need - nr;
                                out_n = (int32_t)nr;
                            }
                        }
                    }
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

// Plain streaming copy (HBM byte-counter calibration).
__global__ __launch_bounds__(256) void copy_kernel(const uint4* __restrict__ src, uint4* __restrict__ dst, int64_t n) {
    for (int64_t i = blockIdx.x * 256ll + threadIdx.x; i < n; i += (int64_t)gridDim.x * 256) dst[i] = src[i];
}

// ---------------------------------------------------------------------------
// host-side launch plumbing
// ---------------------------------------------------------------------------
int32_t ring_stride_entries(const Shape& sh, int32_t W) {
    return sh.off16 ? ring_stride<int16_t>(W) : ring_stride<int32_t>(W);
}

int64_t ring_bytes(const Shape& sh, int32_t W, int32_t rm, int32_t ri) {
    return (int64_t)(rm + 2 * ri) * ring_stride_entries(sh, W) * (sh.off16 ? 2 : 4);
}

int64_t lds_bytes(const Shape& sh, const KArgs& a) {
    int64_t b = lds_head_bytes(a.seq_words, a.rm + 2 * a.ri, sh.nw);
    if (sh.lds_ring) b += ring_bytes(sh, a.W, a.rm, a.ri);
    return b;
}

typedef void (*KernelFn)(KArgs);

template <int NW, typename OffT, bool LDS, bool BT, bool Q>
static KernelFn fn() { return wfa_kernel<NW, OffT, LDS, BT, Q>; }

template <int NW>
static KernelFn pick_nw(const Shape& sh) {
    if (!sh.queue) {
        // plain launches: ring in LDS, no backtrace
        if (!sh.lds_ring || sh.bt) return nullptr;
        return sh.off16 ? fn<NW, int16_t, true, false, false>() : fn<NW, int32_t, true, false, false>();
    }
    if (sh.off16) {
        if (sh.lds_ring) return sh.bt ? fn<NW, int16_t, true, true, true>() : fn<NW, int16_t, true, false, true>();
        return sh.bt ? fn<NW, int16_t, false, true, true>() : fn<NW, int16_t, false, false, true>();
    }
    if (sh.lds_ring) return sh.bt ? fn<NW, int32_t, true, true, true>() : fn<NW, int32_t, true, false, true>();
    return sh.bt ? fn<NW, int32_t, false, true, true>() : fn<NW, int32_t, false, false, true>();
}

static KernelFn pick(const Shape& sh) {
    switch (sh.nw) {
    case 1: return pick_nw<1>(sh);
    case 2: return pick_nw<2>(sh);
    case 4: return pick_nw<4>(sh);
    case 8: return pick_nw<8>(sh);
    case 16: return pick_nw<16>(sh);
    default: return nullptr;
    }
}

int64_t max_lds_for(const Shape& sh) {
    KernelFn f = pick(sh);
    if (!f) return 0;
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess) return 0;
    int v = 0;
    if (hipDeviceGetAttribute(&v, hipDeviceAttributeMaxSharedMemoryPerBlock, dev) != hipSuccess) return 0;
    return v;
}

int occupancy(const Shape& sh, int64_t lds, int* blocks_per_cu) {
    KernelFn f = pick(sh);
    if (!f) return (int)hipErrorInvalidValue;
    if (lds > 65536) {
        hipError_t e = hipFuncSetAttribute(reinterpret_cast<const void*>(f), hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
        if (e != hipSuccess) return (int)e;
    }
    int n = 0;
    hipError_t e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&n, reinterpret_cast<const void*>(f), sh.nw * 64, (size_t)lds);
    if (e != hipSuccess) return (int)e;
    *blocks_per_cu = n;
    return 0;
}

int launch_wfa(const Shape& sh, const KArgs& a, int64_t grid, void* stream) {
    KernelFn f = pick(sh);
    if (!f) return (int)hipErrorInvalidValue;
    if (grid <= 0) return 0;
    if (grid > 0x7fffffff) return (int)hipErrorInvalidValue;
    const int64_t lds = lds_bytes(sh, a);
    if (lds > 65536) {
        hipError_t e = hipFuncSetAttribute(reinterpret_cast<const void*>(f), hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
        if (e != hipSuccess) return (int)e;
    }
    hipLaunchKernelGGL(f, dim3((unsigned)grid), dim3(sh.nw * 64), (size_t)lds, (hipStream_t)stream, a);
    return (int)hipGetLastError();
}

int launch_copy(const void* src, void* dst, int64_t bytes, void* stream) {
    const int64_t n = bytes / 16;
    if (n <= 0) return 0;
    int64_t grid = (n + 255) / 256;
    if (grid > 65536) grid = 65536;
    hipLaunchKernelGGL(copy_kernel, dim3((unsigned)grid), dim3(256), 0, (hipStream_t)stream,
                       reinterpret_cast<const uint4*>(src), reinterpret_cast<uint4*>(dst), n);
    return (int)hipGetLastError();
}

}  // namespace tsu
