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

// record(ss, kk, r) returns false when (ss, kk) has no record.  Penalties are
// the scaled ones of the score loop.  Returns 0 or a negative status; on
// success *out_off / *out_n locate the runs in a.runs.
template <typename RecFn>
__device__ __forceinline__ int bt_walk(RecFn record, int score, int p, int t, int x, int o, int e, const KArgs& a,
                                       int64_t* out_off, int32_t* out_n) {
    int st = 0;
    const int unit = min(x, o + e);
    const int64_t need = 2 * (int64_t)(score / unit) + 2;
    const unsigned long long base = atomicAdd(a.runs_used, (unsigned long long)need);
    if ((int64_t)base + need > a.runs_cap) return TSU_ECIGAR;
    uint32_t* out = a.runs + base;
    int64_t nr = 0;
    uint32_t cur_op = 4, cur_len = 0;
    bool full = false;
    auto put = [&](uint32_t run) __attribute__((always_inline)) {
        if (nr >= need) { full = true; return; }
        out[need - 1 - nr] = run;
        ++nr;
    };
    auto emit = [&](uint32_t op, int n) __attribute__((always_inline)) {
        if (n <= 0) return;
        if (op == cur_op) { cur_len += (uint32_t)n; return; }
        if (cur_len) put((cur_len << 2) | cur_op);
        cur_op = op; cur_len = (uint32_t)n;
    };
    int k = t - p, off = t, ss = score, type = 0;  // 0 M, 1 I, 2 D
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
        if (full) return TSU_ECIGAR;
        *out_off = (int64_t)base + need - nr;
        *out_n = (int32_t)nr;
    }
    return st;
}

}  // namespace tsu
