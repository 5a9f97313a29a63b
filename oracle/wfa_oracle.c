/*
 * wfa_oracle.c -- CPU oracle for the gap-affine WaveFront Alignment hot path.
 *
 * TEST INFRASTRUCTURE ONLY.  Nothing in the product (the HIP library, the
 * wfa_gpu / generate_dataset CLIs, the Python package) links, loads or calls
 * this file.  It is used by tests/ (to check results), by __graft_entry__.smoke()
 * and by bench.py's cpu_baseline leg (timed as a "port" baseline).
 *
 * Contents
 *   orc_gotoh()      textbook Gotoh (1982) affine-gap global DP, O(p*t) time,
 *                    O(t) memory.  Independent of any WFA code; pins scores.
 *   orc_wfa()        exact gap-affine WFA with WFA2-lib semantics, unscaled
 *                    penalties (odd-score null steps included, exactly as the
 *                    reference processes them), optional CIGAR backtrace.
 *   orc_wfa_nulled() the same WFA with every out-of-bounds I, D and M cell
 *                    NULL where it is computed and no trim: the semantics of
 *                    the neighbour-synchronised GPU kernel, which has no
 *                    whole-window view per score.  Tests pin its score and
 *                    CIGAR to orc_wfa's (and the score to Gotoh's).
 *   orc_wfa_prio()   orc_wfa with caller-given backtrace priorities: lets the
 *                    tests show that the priority order is pinned.
 *   orc_wfa_batch()  OpenMP loop over many pairs (CPU baseline; mirrors the
 *                    reference's check loop, wfa_gpu.cu:150-166, but reuses one
 *                    scratch state per thread).
 *
 * Semantics followed (reference file:line):
 *   offsets are text positions h; diagonal k = h - v; the end cell is
 *     k = tlen - plen, offset = tlen                      wfa_gpu.cu:214-215
 *   OFFSET_NULL for cells outside a wavefront's [lo,hi]   common.h:6,
 *                                                         wfa_gpu.cu:470-606
 *   score loop: extend M[s], test termination, s++, compute
 *                                                         wfa_gpu.cu:244-345
 *   sources M[s-x], M[s-o-e], I[s-e], D[s-e]; null-step when all are null
 *                                                         wfa_gpu.cu:347-422
 *   range union  lo = min(lo_x, lo_o-1, lo_i+1, lo_d-1),
 *                hi = max(hi_x, hi_o+1, hi_i+1, hi_d-1)   wfa_gpu.cu:429-439
 *   I/D wavefront exists iff one of its two sources does wfa_gpu.cu:455-468
 *   recurrences  I = max(M_o[k-1], I_e[k-1]) + 1
 *                D = max(M_o[k+1], D_e[k+1])
 *                M = max(D, M_x[k] + 1, I), NULL if h > t or v > p
 *                (unsigned test, as WFA2-lib; the reference's v-test at
 *                 wfa_gpu.cu:629 is off by 2*lo and is not followed)
 *                                                         wfa_gpu.cu:610-631
 *   trim_ends on M, I and D                               wfa_gpu.cu:633-719
 *   score = s at termination (WFA2-lib reports -s)        wfa_gpu.cu:165,335
 *   backtrace tie-break priority (piggy-backed max)        common.h:67-77
 *     M(9) > D_ext(6) > D_open(5) > I_ext(2) > I_open(1)
 *   CIGAR ops 'M' match, 'X' mismatch, 'I' consumes text, 'D' consumes pattern
 *                                                         wfa_gpu.cu:616-623
 *
 * The algorithm of the backtrace is the published one of smarco/WFA2-lib
 * (wavefront_backtrace_affine; the submodule at .gitmodules:1-3 is empty in
 * the reference, version unpinned): from the end cell, take the piggy-backed
 * max over the candidate sources of the current cell; in M emit the matches
 * down to that max, then emit the operation of the winning source and move
 * to it; finish with leading matches / deletions / insertions.
 *
 * Parity pins (tests/test_oracle.py, tests/test_oracle_ties.py): Gotoh ==
 * brute-force enumeration on tiny inputs; closed forms; WFA score == Gotoh
 * on every fixture; CIGARs re-score to the score and consume both
 * sequences; the WFA2-lib README's published example (score 24,
 * "3M1X4M1D7M1I9M1X6M"); the choice among co-optimal CIGARs by twenty pairs
 * whose CIGARs were derived by hand from the published backtrace, each
 * decided by one comparison of the priority order, so that every other
 * order of the five source types (orc_wfa_prio) changes at least one of
 * them, and by gap placement in homopolymer runs.  What stays unpinned:
 * agreement with a WFA2-lib build on details its published algorithm does
 * not fix (WFA2-lib is not available offline).
 */
#include <stdint.h>
#include <stdlib.h>
#include <string.h>
#include <limits.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#define ORC_NULL ((int32_t)(INT32_MIN / 2))
#define ORC_MAX(a, b) ((a) >= (b) ? (a) : (b))
#define ORC_MIN(a, b) ((a) <= (b) ? (a) : (b))

enum { ORC_OK = 0, ORC_EINVAL = -1, ORC_ENOMEM = -4, ORC_EINTERNAL = -9, ORC_ECIGAR = -10 };

/* Backtrace priorities of the five source kinds, in the order
 * { M (mismatch), D_ext, D_open, I_ext, I_open }: the values of WFA2-lib's
 * backtrace_type that the reference restates in common.h:67-77
 * (bt_M 9, bt_D1_ext 6, bt_D1_open 5, bt_I1_ext 2, bt_I1_open 1). */
static const int ORC_PRIO[5] = { 9, 6, 5, 2, 1 };

typedef struct {
    int64_t cells;        /* sum over computed (non-null) scores of (hi-lo+1) */
    int64_t extend_words; /* 32-base window compares in the extend */
    int64_t steps;        /* score steps executed (incl. null steps) */
    int64_t max_width;    /* widest computed range (hi-lo+1) */
    int64_t min_lo;       /* lowest diagonal of any computed range (before trims) */
    int64_t max_hi;       /* highest diagonal of any computed range (before trims) */
} orc_counters;

/* ------------------------------------------------------------------------ */
/* Gotoh affine-gap global alignment (match 0, mismatch x, gap o + L*e).     */
/* ------------------------------------------------------------------------ */
int64_t orc_gotoh(const char* p, int plen, const char* t, int tlen, int x, int o, int e)
{
    if (plen < 0 || tlen < 0) return ORC_EINVAL;
    const int64_t INF = (int64_t)1 << 60;
    /* H[i][j]: best cost aligning p[0,i) with t[0,j).
     * Ins[i][j]: best cost whose last op is an insertion (consumes t[j-1]).
     * Del[i][j]: best cost whose last op is a deletion (consumes p[i-1]).
     * Rows are rolled: Hp/Dp = row i-1, Hc/Dc = row i; Ins runs along the row. */
    int64_t *Hp = (int64_t*)malloc(sizeof(int64_t) * (size_t)(tlen + 1));
    int64_t *Hc = (int64_t*)malloc(sizeof(int64_t) * (size_t)(tlen + 1));
    int64_t *Dp = (int64_t*)malloc(sizeof(int64_t) * (size_t)(tlen + 1));
    int64_t *Dc = (int64_t*)malloc(sizeof(int64_t) * (size_t)(tlen + 1));
    if (!Hp || !Hc || !Dp || !Dc) { free(Hp); free(Hc); free(Dp); free(Dc); return ORC_ENOMEM; }
    Hp[0] = 0; Dp[0] = INF;
    for (int j = 1; j <= tlen; ++j) { Hp[j] = o + (int64_t)e * j; Dp[j] = INF; }
    for (int i = 1; i <= plen; ++i) {
        Dc[0] = o + (int64_t)e * i;
        Hc[0] = Dc[0];
        int64_t ins = INF;
        for (int j = 1; j <= tlen; ++j) {
            int64_t del = ORC_MIN(Hp[j] + o + e, Dp[j] + e);
            ins = ORC_MIN(Hc[j - 1] + o + e, ins + e);
            int64_t h = Hp[j - 1] + (p[i - 1] == t[j - 1] ? 0 : x);
            if (del < h) h = del;
            if (ins < h) h = ins;
            Dc[j] = del;
            Hc[j] = h;
        }
        int64_t* tmp;
        tmp = Hp; Hp = Hc; Hc = tmp;
        tmp = Dp; Dp = Dc; Dc = tmp;
    }
    int64_t r = Hp[tlen];
    free(Hp); free(Hc); free(Dp); free(Dc);
    return r;
}

/* ------------------------------------------------------------------------ */
/* WFA with WFA2-lib semantics.                                              */
/* ------------------------------------------------------------------------ */
typedef struct {
    int lo, hi;       /* lo > hi means null */
    int32_t* off;     /* off[k - lo], k in [lo, hi] (pre-trim allocation) */
    int alo;          /* allocated lo (off is indexed from alo) */
    int ahi;
} orc_wf;

typedef struct {
    orc_wf* M; orc_wf* I; orc_wf* D;
    int cap;          /* number of score slots allocated */
} orc_state;

static inline int wf_null(const orc_wf* w) { return w->lo > w->hi; }

static inline int32_t wf_get(const orc_wf* w, int k)
{
    if (k < w->lo || k > w->hi) return ORC_NULL;
    return w->off[k - w->alo];
}

static void wf_free(orc_wf* w) { free(w->off); w->off = NULL; w->lo = 1; w->hi = 0; w->alo = 1; w->ahi = 0; }

static int state_reserve(orc_state* st, int s)
{
    if (s < st->cap) return 0;
    int ncap = st->cap ? st->cap : 64;
    while (ncap <= s) ncap *= 2;
    orc_wf* nM = (orc_wf*)realloc(st->M, sizeof(orc_wf) * (size_t)ncap);
    if (!nM) return -1;
    st->M = nM;
    orc_wf* nI = (orc_wf*)realloc(st->I, sizeof(orc_wf) * (size_t)ncap);
    if (!nI) return -1;
    st->I = nI;
    orc_wf* nD = (orc_wf*)realloc(st->D, sizeof(orc_wf) * (size_t)ncap);
    if (!nD) return -1;
    st->D = nD;
    for (int i = st->cap; i < ncap; ++i) {
        st->M[i].off = st->I[i].off = st->D[i].off = NULL;
        st->M[i].lo = st->I[i].lo = st->D[i].lo = 1;
        st->M[i].hi = st->I[i].hi = st->D[i].hi = 0;
        st->M[i].alo = st->I[i].alo = st->D[i].alo = 1;
        st->M[i].ahi = st->I[i].ahi = st->D[i].ahi = 0;
    }
    st->cap = ncap;
    return 0;
}

static void state_release(orc_state* st, int upto)
{
    for (int i = 0; i < upto && i < st->cap; ++i) { wf_free(&st->M[i]); wf_free(&st->I[i]); wf_free(&st->D[i]); }
}

static void state_destroy(orc_state* st)
{
    state_release(st, st->cap);
    free(st->M); free(st->I); free(st->D);
    st->M = st->I = st->D = NULL; st->cap = 0;
}

static int wf_alloc(orc_wf* w, int lo, int hi)
{
    free(w->off);
    w->off = (int32_t*)malloc(sizeof(int32_t) * (size_t)(hi - lo + 1));
    if (!w->off) return -1;
    w->lo = lo; w->hi = hi; w->alo = lo; w->ahi = hi;
    return 0;
}

/* trim_ends: wfa_gpu.cu:636-719 (WFA2-lib wavefront_compute_trim_ends). */
static void wf_trim(orc_wf* w, int plen, int tlen)
{
    if (wf_null(w)) return;
    int k;
    const int lo = w->lo;
    for (k = w->hi; k >= lo; --k) {
        int32_t off = w->off[k - w->alo];
        uint32_t h = (uint32_t)off;
        uint32_t v = (uint32_t)(off - k);
        if (h <= (uint32_t)tlen && v <= (uint32_t)plen) break;
    }
    w->hi = k;
    const int hi = w->hi;
    for (k = w->lo; k <= hi; ++k) {
        int32_t off = w->off[k - w->alo];
        uint32_t h = (uint32_t)off;
        uint32_t v = (uint32_t)(off - k);
        if (h <= (uint32_t)tlen && v <= (uint32_t)plen) break;
    }
    w->lo = k;
}

/* Extend one diagonal: number of matching characters from (v, h). */
static inline int orc_lcp(const char* p, int plen, const char* t, int tlen, int v, int h, int64_t* words)
{
    int lim = ORC_MIN(plen - v, tlen - h);
    int n = 0;
    while (n < lim && p[v + n] == t[h + n]) ++n;
    if (words && lim > 0) *words += (n < lim) ? (n / 32 + 1) : ((lim + 31) / 32);
    return n;
}

static void orc_reverse(char* a, int n)
{
    for (int i = 0, j = n - 1; i < j; ++i, --j) { char c = a[i]; a[i] = a[j]; a[j] = c; }
}

/*
 * Backtrace (WFA2-lib wavefront_backtrace_affine, end-to-end, M end component).
 * Returns number of ops written to ops (forward order) or ORC_ECIGAR.
 */
static int orc_backtrace(const orc_state* st, const char* p, int plen, const char* t, int tlen,
                         int x, int o, int e, int score_end, char* ops, int ops_cap, const int* prio)
{
    enum { T_M = 0, T_I = 1, T_D = 2 };
    /* priority of each source kind in the piggy-backed max (common.h:67-77) */
    const int BT_M = prio[0], BT_D_EXT = prio[1], BT_D_OPEN = prio[2], BT_I_EXT = prio[3], BT_I_OPEN = prio[4];
    const int64_t PNULL = ((int64_t)ORC_NULL) * 16;
    int n = 0;
    int k = tlen - plen;
    int32_t offset = tlen;
    int v = plen, h = tlen;
    int score = score_end;
    int type = T_M;
#define PIGGY(off, ty) ((((int64_t)(off)) * 16) | (ty))
#define EMIT(c) do { if (n >= ops_cap) return ORC_ECIGAR; ops[n++] = (c); } while (0)
    while (v > 0 && h > 0 && score > 0) {
        const int s_mis = score - x;
        const int s_open = score - o - e;
        const int s_ext = score - e;
        int64_t best = PNULL;
        if (type == T_M) {
            if (s_mis >= 0) { int64_t c = PIGGY(wf_get(&st->M[s_mis], k) + 1, BT_M); if (c > best) best = c; }
            if (s_open >= 0) {
                int64_t c1 = PIGGY(wf_get(&st->M[s_open], k - 1) + 1, BT_I_OPEN); if (c1 > best) best = c1;
                int64_t c2 = PIGGY(wf_get(&st->M[s_open], k + 1), BT_D_OPEN); if (c2 > best) best = c2;
            }
            if (s_ext >= 0) {
                int64_t c3 = PIGGY(wf_get(&st->I[s_ext], k - 1) + 1, BT_I_EXT); if (c3 > best) best = c3;
                int64_t c4 = PIGGY(wf_get(&st->D[s_ext], k + 1), BT_D_EXT); if (c4 > best) best = c4;
            }
        } else if (type == T_I) {
            if (s_open >= 0) { int64_t c = PIGGY(wf_get(&st->M[s_open], k - 1) + 1, BT_I_OPEN); if (c > best) best = c; }
            if (s_ext >= 0) { int64_t c = PIGGY(wf_get(&st->I[s_ext], k - 1) + 1, BT_I_EXT); if (c > best) best = c; }
        } else {
            if (s_open >= 0) { int64_t c = PIGGY(wf_get(&st->M[s_open], k + 1), BT_D_OPEN); if (c > best) best = c; }
            if (s_ext >= 0) { int64_t c = PIGGY(wf_get(&st->D[s_ext], k + 1), BT_D_EXT); if (c > best) best = c; }
        }
        if (best < 0) return ORC_EINTERNAL;
        const int32_t max_off = (int32_t)(best >> 4);
        const int bt = (int)(best & 15);
        if (type == T_M) {
            int nm = offset - max_off;
            if (nm < 0) return ORC_EINTERNAL;
            for (int i = 0; i < nm; ++i) EMIT('M');
            offset = max_off;
            v = offset - k; h = offset;
            if (v <= 0 || h <= 0) break;
        }
        if (bt == BT_M) { score = s_mis; type = T_M; EMIT('X'); --offset; }
        else if (bt == BT_I_OPEN) { score = s_open; type = T_M; EMIT('I'); --k; --offset; }
        else if (bt == BT_I_EXT) { score = s_ext; type = T_I; EMIT('I'); --k; --offset; }
        else if (bt == BT_D_OPEN) { score = s_open; type = T_M; EMIT('D'); ++k; }
        else if (bt == BT_D_EXT) { score = s_ext; type = T_D; EMIT('D'); ++k; }
        else return ORC_EINTERNAL;
        v = offset - k; h = offset;
    }
    if (type == T_M) {
        if (v > 0 && h > 0) {
            int nm = ORC_MIN(v, h);
            for (int i = 0; i < nm; ++i) EMIT('M');
            v -= nm; h -= nm;
        }
        while (v > 0) { EMIT('D'); --v; }
        while (h > 0) { EMIT('I'); --h; }
    } else if (v != 0 || h != 0) {
        return ORC_EINTERNAL;
    }
#undef PIGGY
#undef EMIT
    orc_reverse(ops, n);
    (void)p; (void)t;
    return n;
}

static int orc_wfa_impl(orc_state* st, const char* p, int plen, const char* t, int tlen,
                        int x, int o, int e, int want_cigar, char* ops, int ops_cap, int* ops_len,
                        orc_counters* cnt, int null_oob, const int* prio)
{
    if (plen < 0 || tlen < 0 || x <= 0 || o < 0 || e <= 0) return ORC_EINVAL;
    const int ring = ORC_MAX(x, o + e) + 1;          /* wfa_gpu.cu:83-84 */
    const int align_k = tlen - plen;                  /* wfa_gpu.cu:214 */
    const int align_off = tlen;                       /* wfa_gpu.cu:215 */
    /* Upper bound of the optimum: all-mismatch diagonal plus one gap. */
    const int64_t s_bound = (int64_t)x * ORC_MIN(plen, tlen) + (plen != tlen ? (int64_t)o + (int64_t)e * abs(plen - tlen) : 0);
    int64_t* ew = cnt ? &cnt->extend_words : NULL;

    if (state_reserve(st, 0)) return ORC_ENOMEM;
    if (wf_alloc(&st->M[0], 0, 0)) return ORC_ENOMEM;
    st->M[0].off[0] = 0;
    wf_free(&st->I[0]); wf_free(&st->D[0]);
    int s = 0;
    int score = -1;
    for (;;) {
        if (cnt) cnt->steps++;
        orc_wf* M = &st->M[s];
        if (!wf_null(M)) {
            /* EXTEND (wfa_gpu.cu:253-287) */
            for (int k = M->lo; k <= M->hi; ++k) {
                int32_t off = M->off[k - M->alo];
                if (off < 0) continue;
                int v = off - k;
                M->off[k - M->alo] = off + orc_lcp(p, plen, t, tlen, v, off, ew);
            }
            /* TERMINATE (wfa_gpu.cu:288-331) */
            if (M->lo <= align_k && align_k <= M->hi && M->off[align_k - M->alo] >= align_off) {
                score = s;
                break;
            }
        }
        ++s;
        if (s > s_bound) return ORC_EINTERNAL;
        if (state_reserve(st, s)) return ORC_ENOMEM;
        if (!want_cigar && s - ring - 1 >= 0) {
            int old = s - ring - 1;
            wf_free(&st->M[old]); wf_free(&st->I[old]); wf_free(&st->D[old]);
        }
        const int s_mis = s - x, s_open = s - o - e, s_ext = s - e;
        static const orc_wf NULLWF = { 1, 0, NULL, 1, 0 };
        const orc_wf* Mx = s_mis >= 0 ? &st->M[s_mis] : &NULLWF;
        const orc_wf* Mo = s_open >= 0 ? &st->M[s_open] : &NULLWF;
        const orc_wf* Ie = s_ext >= 0 ? &st->I[s_ext] : &NULLWF;
        const orc_wf* De = s_ext >= 0 ? &st->D[s_ext] : &NULLWF;
        const int nx = wf_null(Mx), no = wf_null(Mo), ni = wf_null(Ie), nd = wf_null(De);
        orc_wf* oM = &st->M[s];
        orc_wf* oI = &st->I[s];
        orc_wf* oD = &st->D[s];
        wf_free(oM); wf_free(oI); wf_free(oD);
        if (nx && no && ni && nd) continue;            /* null step: wfa_gpu.cu:417-422 */
        /* range union: wfa_gpu.cu:429-439 */
        int lo = 1, hi = -1, have = 0;
        if (!nx) { lo = Mx->lo; hi = Mx->hi; have = 1; }
        if (!no) { if (!have || lo > Mo->lo - 1) lo = Mo->lo - 1; if (!have || hi < Mo->hi + 1) hi = Mo->hi + 1; have = 1; }
        if (!ni) { if (!have || lo > Ie->lo + 1) lo = Ie->lo + 1; if (!have || hi < Ie->hi + 1) hi = Ie->hi + 1; have = 1; }
        if (!nd) { if (!have || lo > De->lo - 1) lo = De->lo - 1; if (!have || hi < De->hi - 1) hi = De->hi - 1; have = 1; }
        if (lo > hi) continue;
        const int hasI = !no || !ni;                   /* wfa_gpu.cu:455-461 */
        const int hasD = !no || !nd;                   /* wfa_gpu.cu:462-468 */
        if (wf_alloc(oM, lo, hi)) return ORC_ENOMEM;
        if (hasI && wf_alloc(oI, lo, hi)) return ORC_ENOMEM;
        if (hasD && wf_alloc(oD, lo, hi)) return ORC_ENOMEM;
        if (cnt) {
            cnt->cells += hi - lo + 1;
            if (hi - lo + 1 > cnt->max_width) cnt->max_width = hi - lo + 1;
            if (lo < cnt->min_lo) cnt->min_lo = lo;
            if (hi > cnt->max_hi) cnt->max_hi = hi;
        }
        /* COMPUTE-NEXT (wfa_gpu.cu:610-631) */
        for (int k = lo; k <= hi; ++k) {
            int32_t ins = ORC_MAX(wf_get(Mo, k - 1), wf_get(Ie, k - 1)) + 1;
            int32_t del = ORC_MAX(wf_get(Mo, k + 1), wf_get(De, k + 1));
            int32_t mis = wf_get(Mx, k) + 1;
            if (null_oob) {
                if ((uint32_t)ins > (uint32_t)tlen || (uint32_t)(ins - k) > (uint32_t)plen) ins = ORC_NULL;
                if ((uint32_t)del > (uint32_t)tlen || (uint32_t)(del - k) > (uint32_t)plen) del = ORC_NULL;
            }
            int32_t m = ORC_MAX(del, ORC_MAX(mis, ins));
            if ((uint32_t)m > (uint32_t)tlen || (uint32_t)(m - k) > (uint32_t)plen) m = ORC_NULL;
            if (hasI) oI->off[k - lo] = ins;
            if (hasD) oD->off[k - lo] = del;
            oM->off[k - lo] = m;
        }
        /* TRIM ENDS (wfa_gpu.cu:633-719) */
        if (!null_oob) {
            wf_trim(oM, plen, tlen);
            if (hasI) wf_trim(oI, plen, tlen);
            if (hasD) wf_trim(oD, plen, tlen);
        }
    }
    if (want_cigar) {
        int n = orc_backtrace(st, p, plen, t, tlen, x, o, e, score, ops, ops_cap, prio ? prio : ORC_PRIO);
        if (n < 0) return n;
        if (ops_len) *ops_len = n;
    }
    return score;
}

/*
 * Align one pair.  Returns the optimal gap-affine score (>= 0) or a negative
 * error.  With want_cigar, writes the operation string (no terminator) into
 * ops[0, *ops_len); ops_cap must be >= plen + tlen.
 */
int orc_wfa(const char* p, int plen, const char* t, int tlen, int x, int o, int e,
            int want_cigar, char* ops, int ops_cap, int* ops_len, orc_counters* cnt)
{
    orc_state st = { NULL, NULL, NULL, 0 };
    int r = orc_wfa_impl(&st, p, plen, t, tlen, x, o, e, want_cigar, ops, ops_cap, ops_len, cnt, 0, NULL);
    state_destroy(&st);
    return r;
}

/*
 * One pair with out-of-bounds cells NULLed where computed instead of
 * trimmed (the neighbour-synchronised GPU kernel's semantics); the CIGAR, if
 * asked for, is walked over those cells.
 */
int orc_wfa_nulled(const char* p, int plen, const char* t, int tlen, int x, int o, int e,
                   int want_cigar, char* ops, int ops_cap, int* ops_len)
{
    orc_state st = { NULL, NULL, NULL, 0 };
    int r = orc_wfa_impl(&st, p, plen, t, tlen, x, o, e, want_cigar, ops, ops_cap, ops_len, NULL, 1, NULL);
    state_destroy(&st);
    return r;
}

/*
 * One pair with the backtrace priorities given by the caller (prio[5] in the
 * order M, D_ext, D_open, I_ext, I_open; five distinct positive values below
 * 16).  Used only by the tests that check that every other priority order
 * changes some hand-derived CIGAR.
 */
int orc_wfa_prio(const char* p, int plen, const char* t, int tlen, int x, int o, int e,
                 const int* prio, char* ops, int ops_cap, int* ops_len)
{
    if (!prio) return ORC_EINVAL;
    for (int i = 0; i < 5; ++i) {
        if (prio[i] <= 0 || prio[i] >= 16) return ORC_EINVAL;
        for (int j = 0; j < i; ++j) if (prio[i] == prio[j]) return ORC_EINVAL;
    }
    orc_state st = { NULL, NULL, NULL, 0 };
    int r = orc_wfa_impl(&st, p, plen, t, tlen, x, o, e, 1, ops, ops_cap, ops_len, NULL, 0, prio);
    state_destroy(&st);
    return r;
}

/*
 * Many pairs under OpenMP (CPU baseline).  seqs holds all characters; pair i
 * is pattern seqs[p_off[i], +p_len[i]) and text seqs[t_off[i], +t_len[i]).
 * threads <= 0 uses the OpenMP default.  Returns the number of pairs whose
 * alignment failed (their score slot receives the negative error code).
 * When want_cigar, ops for pair i go to ops[ops_off[i], +p_len[i]+t_len[i])
 * and ops_len[i] receives the count.
 */
int orc_wfa_batch(const char* seqs, const int64_t* p_off, const int32_t* p_len,
                  const int64_t* t_off, const int32_t* t_len, int n,
                  int x, int o, int e, int threads, int32_t* scores,
                  int want_cigar, char* ops, const int64_t* ops_off, int32_t* ops_len,
                  orc_counters* total)
{
    int failed = 0;
    int64_t cells = 0, words = 0, steps = 0, maxw = 0, minlo = 0, maxhi = 0;
#ifdef _OPENMP
    if (threads > 0) omp_set_num_threads(threads);
#else
    (void)threads;
#endif
#pragma omp parallel reduction(+:failed, cells, words, steps) reduction(max:maxw, maxhi) reduction(min:minlo)
    {
        orc_state st = { NULL, NULL, NULL, 0 };
        orc_counters c = { 0, 0, 0, 0, 0, 0 };
#pragma omp for schedule(dynamic, 16)
        for (int i = 0; i < n; ++i) {
            int olen = 0;
            int r = orc_wfa_impl(&st, seqs + p_off[i], p_len[i], seqs + t_off[i], t_len[i], x, o, e,
                                 want_cigar, want_cigar ? ops + ops_off[i] : NULL,
                                 p_len[i] + t_len[i], &olen, total ? &c : NULL, 0, NULL);
            scores[i] = r;
            if (want_cigar && ops_len) ops_len[i] = olen;
            if (r < 0) failed++;
            /* release wavefront memory between pairs but keep the slot table */
            state_release(&st, st.cap);
        }
        state_destroy(&st);
        cells += c.cells; words += c.extend_words; steps += c.steps;
        if (c.max_width > maxw) maxw = c.max_width;
        if (c.min_lo < minlo) minlo = c.min_lo;
        if (c.max_hi > maxhi) maxhi = c.max_hi;
    }
    if (total) {
        total->cells = cells; total->extend_words = words; total->steps = steps; total->max_width = maxw;
        total->min_lo = minlo; total->max_hi = maxhi;
    }
    return failed;
}

/* Number of threads OpenMP would use. */
int orc_max_threads(void)
{
#ifdef _OPENMP
    return omp_get_max_threads();
#else
    return 1;
#endif
}
