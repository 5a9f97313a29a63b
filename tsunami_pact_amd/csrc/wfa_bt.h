// wfa_bt.h -- backtrace walk shared by the kernels that record cells.
//
// Every computed cell has one 32-bit record:
//   bits 0-1  origin of M[s][k] before extension: 1 X, 2 D, 3 I (0 = NULL)
//             (ties resolved M > D > I, with D = max(D_open, D_ext) and
//              I = max(I_open, I_ext): the piggy-backed order of common.h:67-77)
//   bit  2    I[s][k] came from I[s-e][k-1] (extension) rather than M_o
//   bit  3    D[s][k] came from D[s-e][k+1] (extension) rather than M_o
//   bits 4-31 M[s][k] before extension
// One thread walks them back from the end cell (score, k = t - p, offset t)
// and writes run-length CIGAR runs (run = length << 2 | op, op 0 M, 1 X,
// 2 I, 3 D) into a pool, reserving 2 * score / min(x, o+e) + 2 runs with
// one atomic add (a bound on the runs of any alignment of that score).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "tsu_internal.h"

namespace tsu {

// Walk state.  A walk may run in pieces: bt_run stops before it needs a
// record at a score <= floor, so a kernel that keeps only the records of a
// window of scores (recomputed from a checkpoint) refills the window below
// and resumes.  bt_walk is the one-piece form.
struct BtWalk {
    uint32_t* out;
    int64_t base, need, nr;
    uint32_t cur_op, cur_len;
    bool full;
    bool writer = true;  // this lane stores the runs (one lane of a wave-wide walk)
    int k, off, ss, type, v, h;  // type 0 M, 1 I, 2 D
    int x, o, e;

    __device__ __forceinline__ void put(uint32_t run) {
        if (nr >= need) { full = true; return; }
        if (writer) out[need - 1 - nr] = run;
        ++nr;
    }
    __device__ __forceinline__ void emit(uint32_t op, int n) {
        if (n <= 0) return;
        if (op == cur_op) { cur_len += (uint32_t)n; return; }
        if (cur_len) put((cur_len << 2) | cur_op);
        cur_op = op; cur_len = (uint32_t)n;
    }
};

// Reserves the walk's runs (one atomic add) and sets it at the end cell.
// Penalties are the scaled ones of the score loop.  Returns 0 or TSU_ECIGAR.
__device__ __forceinline__ int bt_begin(BtWalk& w, int score, int p, int t, int x, int o, int e, const KArgs& a) {
    const int unit = min(x, o + e);
    w.need = 2 * (int64_t)(score / unit) + 2;
    w.base = (int64_t)atomicAdd(a.runs_used, (unsigned long long)w.need);
    if (w.base + w.need > a.runs_cap) return TSU_ECIGAR;
    w.out = a.runs + w.base;
    w.nr = 0;
    w.cur_op = 4; w.cur_len = 0;
    w.full = false;
    w.k = t - p; w.off = t; w.ss = score; w.type = 0;
    w.v = p; w.h = t;
    w.x = x; w.o = o; w.e = e;
    return 0;
}

// record(ss, kk, r) returns false when (ss, kk) has no record.  Returns 0
// when the walk has reached its start (row or column 0, or score 0), 1 when
// it stopped before a record at a score <= floor, or a negative status.
template <typename RecFn>
__device__ __forceinline__ int bt_run(BtWalk& w, RecFn record, int floor) {
    while (w.v > 0 && w.h > 0 && w.ss > 0) {
        if (w.ss <= floor) return 1;
        uint32_t r;
        if (!record(w.ss, w.k, r)) return TSU_EINTERNAL;
        if (w.type == 0) {
            const int pre = (int)(r >> 4);
            const uint32_t mw = r & 3u;
            if (mw == 0 || pre > w.off) return TSU_EINTERNAL;
            w.emit(0, w.off - pre);
            w.off = pre; w.v = w.off - w.k; w.h = w.off;
            if (w.v <= 0 || w.h <= 0) break;
            if (mw == 1) { w.emit(1, 1); w.ss -= w.x; --w.off; w.v = w.off - w.k; w.h = w.off; continue; }
            w.type = (mw == 2) ? 2 : 1;
        }
        if (w.type == 1) {
            const bool ext = (r & 4u) != 0;
            w.emit(2, 1); --w.k; --w.off;
            w.ss -= ext ? w.e : (w.o + w.e);
            w.type = ext ? 1 : 0;
        } else {
            const bool ext = (r & 8u) != 0;
            w.emit(3, 1); ++w.k;
            w.ss -= ext ? w.e : (w.o + w.e);
            w.type = ext ? 2 : 0;
        }
        w.v = w.off - w.k; w.h = w.off;
    }
    return 0;
}

// Emits the leading gap / match runs of a finished walk and locates its runs
// in a.runs (*out_off, *out_n).  Returns 0 or a negative status.
__device__ __forceinline__ int bt_end(BtWalk& w, int64_t* out_off, int32_t* out_n) {
    if (w.type == 0) {
        if (w.v > 0 && w.h > 0) { const int nm = min(w.v, w.h); w.emit(0, nm); w.v -= nm; w.h -= nm; }
        w.emit(3, w.v);
        w.emit(2, w.h);
    } else if (w.v != 0 || w.h != 0) {
        return TSU_EINTERNAL;
    }
    if (w.cur_len) w.put((w.cur_len << 2) | w.cur_op);
    if (w.full) return TSU_ECIGAR;
    *out_off = w.base + w.need - w.nr;
    *out_n = (int32_t)w.nr;
    return 0;
}

// One-piece walk over records that are all present.  Returns 0 or a
// negative status; on success *out_off / *out_n locate the runs in a.runs.
template <typename RecFn>
__device__ __forceinline__ int bt_walk(RecFn record, int score, int p, int t, int x, int o, int e, const KArgs& a,
                                       int64_t* out_off, int32_t* out_n) {
    BtWalk w;
    int st = bt_begin(w, score, p, t, x, o, e, a);
    if (st) return st;
    st = bt_run(w, record, 0);
    if (st) return st;
    return bt_end(w, out_off, out_n);
}

// Wave-wide form of bt_walk, for records that sit at a fixed place per score
// and diagonal.  The walk is a chain of dependent reads, one per step, each
// a round trip to device memory; here the 64 lanes of one wave load the
// 8-score x 8-diagonal block of records around the walk's current cell in
// one round trip (scores ss-7..ss, diagonals k-3..k+4), and the walk reads
// them from there until a step leaves the block: a mismatch moves 2 scores
// down the same diagonal, a gap 1 to 4 scores and one diagonal, so a block
// serves several steps.  The walk's state is the same on every lane (each
// value read is broadcast from its lane); lane 0 reserves and writes the
// runs.  load(ss, kk, r) is record's per-lane form: it reads nothing and
// returns false where (ss, kk) has no record.  Call it from all 64 lanes of
// the wave.
template <typename LoadFn>
__device__ __forceinline__ int bt_walk_wave(LoadFn load, int score, int p, int t, int x, int o, int e, const KArgs& a,
                                            int64_t* out_off, int32_t* out_n) {
    const int lane = (int)(threadIdx.x & 63);
    BtWalk w;
    const int unit = min(x, o + e);
    w.need = 2 * (int64_t)(score / unit) + 2;
    unsigned long long b0 = 0;
    if (lane == 0) b0 = atomicAdd(a.runs_used, (unsigned long long)w.need);
    const uint32_t blo = __builtin_amdgcn_readfirstlane((uint32_t)b0);
    const uint32_t bhi = __builtin_amdgcn_readfirstlane((uint32_t)(b0 >> 32));
    w.base = (int64_t)(((uint64_t)bhi << 32) | blo);
    if (w.base + w.need > a.runs_cap) return TSU_ECIGAR;
    w.out = a.runs + w.base;
    w.nr = 0;
    w.cur_op = 4; w.cur_len = 0;
    w.full = false;
    w.writer = lane == 0;
    w.k = t - p; w.off = t; w.ss = score; w.type = 0;
    w.v = p; w.h = t;
    w.x = x; w.o = o; w.e = e;
    bool have = false;    // a block is loaded
    int bs = 0, bk = 0;   // its top score and its diagonal k (column 3)
    uint32_t blk = 0;     // this lane's record: score bs - lane/8, diagonal bk - 3 + lane%8
    uint64_t bvalid = 0;  // lanes whose record exists
    auto record = [&](int ss, int kk, uint32_t& r) __attribute__((always_inline)) -> bool {
        int ds = bs - ss, dk = kk - bk + 3;
        if (!have || ds < 0 || ds > 7 || dk < 0 || dk > 7) {
            bs = ss; bk = kk; have = true;
            uint32_t v = 0;
            const bool ok = load(ss - (lane >> 3), kk - 3 + (lane & 7), v);
            blk = ok ? v : 0u;
            bvalid = __ballot(ok);
            ds = 0; dk = 3;
        }
        const int li = __builtin_amdgcn_readfirstlane(8 * ds + dk);
        if (!((bvalid >> li) & 1ull)) return false;
        r = (uint32_t)__builtin_amdgcn_readlane((int)blk, li);
        return true;
    };
    int st = bt_run(w, record, 0);
    if (st) return st;
    return bt_end(w, out_off, out_n);
}

// Gap-affine 2-piece records (wfa2p.hip), one 32-bit word per computed cell:
//   bits 0-2  origin of M[s][k] before extension: 1 X, 2 D2, 3 D1, 4 I2,
//             5 I1 (0 = NULL); ties resolved in that order, each gap piece
//             being max(open, ext): the piggy-backed order of the nine types
//             of common.h:67-77 (M 9, D2 8/7, D1 6/5, I2 4/3, I1 2/1)
//   bit  3    I1[s][k] came from I1[s-e1][k-1] (extension) rather than M_o1
//   bit  4    D1[s][k] came from D1[s-e1][k+1]
//   bit  5    I2[s][k] came from I2[s-e2][k-1]
//   bit  6    D2[s][k] came from D2[s-e2][k+1]
//   bits 7-31 M[s][k] before extension
// The walk is bt_walk's with five components; runs are reserved for
// 2 * score / min(x, o1+e1, o2+e2) + 2.
template <typename RecFn>
__device__ __forceinline__ int bt2p_walk(RecFn record, int score, int p, int t, int x, int o1, int e1, int o2, int e2,
                                         const KArgs& a, int64_t* out_off, int32_t* out_n) {
    BtWalk w;
    int st = bt_begin(w, score, p, t, x, min(o1 + e1, o2 + e2), 0, a);
    if (st) return st;
    int comp = 0;  // 0 M, 1 I1, 2 D1, 3 I2, 4 D2
    while (w.v > 0 && w.h > 0 && w.ss > 0) {
        uint32_t r;
        if (!record(w.ss, w.k, r)) return TSU_EINTERNAL;
        if (comp == 0) {
            const int pre = (int)(r >> 7);
            const uint32_t org = r & 7u;
            if (org == 0 || org > 5 || pre > w.off) return TSU_EINTERNAL;
            w.emit(0, w.off - pre);
            w.off = pre; w.v = w.off - w.k; w.h = w.off;
            if (w.v <= 0 || w.h <= 0) break;
            if (org == 1) { w.emit(1, 1); w.ss -= x; --w.off; w.v = w.off - w.k; w.h = w.off; continue; }
            comp = org == 2 ? 4 : org == 3 ? 2 : org == 4 ? 3 : 1;
        }
        const bool piece2 = comp >= 3;
        const int oo = piece2 ? o2 : o1, ee = piece2 ? e2 : e1;
        if (comp == 1 || comp == 3) {
            const bool ext = (r >> (piece2 ? 5 : 3)) & 1u;
            w.emit(2, 1); --w.k; --w.off;
            w.ss -= ext ? ee : oo + ee;
            if (!ext) comp = 0;
        } else {
            const bool ext = (r >> (piece2 ? 6 : 4)) & 1u;
            w.emit(3, 1); ++w.k;
            w.ss -= ext ? ee : oo + ee;
            if (!ext) comp = 0;
        }
        w.v = w.off - w.k; w.h = w.off;
    }
    w.type = comp == 0 ? 0 : 1;
    return bt_end(w, out_off, out_n);
}

}  // namespace tsu
