/*
 * wfa2p_oracle.c -- CPU oracle for gap-affine 2-piece WaveFront Alignment.
 *
 * TEST INFRASTRUCTURE ONLY.  Nothing in the product links, loads or calls
 * this file; it is used by tests/ (to check results) and by bench.py's
 * cpu_baseline leg.
 *
 * Gap-affine 2-piece scoring: match 0, mismatch x, a gap of length L costs
 * min(o1 + L*e1, o2 + L*e2).  The reference declares the matrices and
 * penalties of this model but never uses them (common.h:52-58,80-87: I2/D2
 * wavefronts, gap_open2/gap_ext2; the backtrace types I2/D2 open/extend in
 * common.h:67-77).  The algorithm followed is smarco/WFA2-lib's affine-2p
 * wavefront step (wavefront_compute_affine2p; the submodule at .gitmodules:1-3
 * is empty in the reference, version unpinned), i.e. the gap-affine step of
 * wfa_oracle.c with two gap pieces:
 *
 *   sources   M[s-x], M[s-o1-e1], M[s-o2-e2], I1/D1[s-e1], I2/D2[s-e2];
 *             null step when all are null
 *   range     lo = min(lo_x, lo_o1-1, lo_o2-1, lo_i1+1, lo_i2+1, lo_d1-1, lo_d2-1)
 *             hi = max(hi_x, hi_o1+1, hi_o2+1, hi_i1+1, hi_i2+1, hi_d1-1, hi_d2-1)
 *   I1 = max(M_o1[k-1], I1_e1[k-1]) + 1     I2 = max(M_o2[k-1], I2_e2[k-1]) + 1
 *   D1 = max(M_o1[k+1], D1_e1[k+1])         D2 = max(M_o2[k+1], D2_e2[k+1])
 *   M  = max(M_x[k] + 1, I1, I2, D1, D2), NULL if h > t or v > p
 *   I1 exists iff M_o1 or I1_e1 does (likewise I2, D1, D2); trim_ends on all
 *   five; end when M[s][t-p] >= t after the extend.
 *
 * orc_gotoh2p() is the textbook 2-piece Gotoh DP (five matrices, O(p*t)
 * time, O(t) memory), independent of the WFA code; it pins the scores.
 *
 * orc_wfa2p_cigar() keeps every wavefront and walks back from the end cell
 * as WFA2-lib's affine-2p backtrace does (wavefront_backtrace_affine): at an
 * M cell the candidates are the mismatch M[s-x][k]+1, the opens
 * M[s-o-e][k-1]+1 / M[s-o-e][k+1] and the extensions I[s-e][k-1]+1 /
 * D[s-e][k+1] of both pieces, compared as offset * 16 + type (largest offset
 * first, ties to the larger type) with the types of common.h:67-77: M 9,
 * D2 ext 8, D2 open 7, D1 ext 6, D1 open 5, I2 ext 4, I2 open 3, I1 ext 2,
 * I1 open 1; in an I or D chain only that component's open and extend.
 * orc_wfa2p_prio() takes the nine type values from the caller (mutation
 * tests).  Pinned by: the CIGAR's 2-piece cost equals orc_gotoh2p's score;
 * with the second piece never cheaper, and with two equal pieces, it equals
 * the pinned 1-piece CIGAR of orc_wfa; hand-derived tie cases in
 * tests/tie_cases.py, each decided by one comparison of the order above.
 */
#include <stdint.h>
#include <stdlib.h>
#include <string.h>
#include <limits.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#define P2_NULL ((int32_t)(INT32_MIN / 2))
#define P2_MAX(a, b) ((a) >= (b) ? (a) : (b))
#define P2_MIN(a, b) ((a) <= (b) ? (a) : (b))

enum { P2_OK = 0, P2_EINVAL = -1, P2_ENOMEM = -4, P2_EINTERNAL = -9 };

/* Gotoh 2-piece global DP. */
int64_t orc_gotoh2p(const char* p, int plen, const char* t, int tlen, int x, int o1, int e1, int o2, int e2)
{
    if (plen < 0 || tlen < 0) return P2_EINVAL;
    const int64_t INF = (int64_t)1 << 60;
    const size_t n = (size_t)tlen + 1;
    int64_t* Hp = (int64_t*)malloc(sizeof(int64_t) * n);
    int64_t* Hc = (int64_t*)malloc(sizeof(int64_t) * n);
    int64_t* D1p = (int64_t*)malloc(sizeof(int64_t) * n);
    int64_t* D1c = (int64_t*)malloc(sizeof(int64_t) * n);
    int64_t* D2p = (int64_t*)malloc(sizeof(int64_t) * n);
    int64_t* D2c = (int64_t*)malloc(sizeof(int64_t) * n);
    if (!Hp || !Hc || !D1p || !D1c || !D2p || !D2c) {
        free(Hp); free(Hc); free(D1p); free(D1c); free(D2p); free(D2c);
        return P2_ENOMEM;
    }
    /* row 0: only insertions (gaps in the pattern of length j) */
    Hp[0] = 0; D1p[0] = INF; D2p[0] = INF;
    for (int j = 1; j <= tlen; ++j) {
        Hp[j] = P2_MIN(o1 + (int64_t)e1 * j, o2 + (int64_t)e2 * j);
        D1p[j] = INF; D2p[j] = INF;
    }
    for (int i = 1; i <= plen; ++i) {
        D1c[0] = o1 + (int64_t)e1 * i;
        D2c[0] = o2 + (int64_t)e2 * i;
        Hc[0] = P2_MIN(D1c[0], D2c[0]);
        int64_t ins1 = INF, ins2 = INF;
        for (int j = 1; j <= tlen; ++j) {
            const int64_t d1 = P2_MIN(Hp[j] + o1 + e1, D1p[j] + e1);
            const int64_t d2 = P2_MIN(Hp[j] + o2 + e2, D2p[j] + e2);
            ins1 = P2_MIN(Hc[j - 1] + o1 + e1, ins1 + e1);
            ins2 = P2_MIN(Hc[j - 1] + o2 + e2, ins2 + e2);
            int64_t h = Hp[j - 1] + (p[i - 1] == t[j - 1] ? 0 : x);
            h = P2_MIN(h, P2_MIN(P2_MIN(d1, d2), P2_MIN(ins1, ins2)));
            D1c[j] = d1; D2c[j] = d2; Hc[j] = h;
        }
        int64_t* tmp;
        tmp = Hp; Hp = Hc; Hc = tmp;
        tmp = D1p; D1p = D1c; D1c = tmp;
        tmp = D2p; D2p = D2c; D2c = tmp;
    }
    const int64_t r = Hp[tlen];
    free(Hp); free(Hc); free(D1p); free(D1c); free(D2p); free(D2c);
    return r;
}

/* A wavefront: offsets on [lo, hi] (lo > hi: null). */
typedef struct { int lo, hi, alo; int32_t* off; } p2_wf;

static inline int p2_null(const p2_wf* w) { return w->lo > w->hi; }
static inline int32_t p2_get(const p2_wf* w, int k)
{
    return (k < w->lo || k > w->hi) ? P2_NULL : w->off[k - w->alo];
}
static void p2_free(p2_wf* w) { free(w->off); w->off = NULL; w->lo = 1; w->hi = 0; w->alo = 1; }
static int p2_alloc(p2_wf* w, int lo, int hi)
{
    free(w->off);
    w->off = (int32_t*)malloc(sizeof(int32_t) * (size_t)(hi - lo + 1));
    if (!w->off) return -1;
    w->lo = lo; w->hi = hi; w->alo = lo;
    return 0;
}
/* trim_ends (wfa_gpu.cu:633-719 for one component) */
static void p2_trim(p2_wf* w, int plen, int tlen)
{
    if (p2_null(w)) return;
    int k;
    for (k = w->hi; k >= w->lo; --k) {
        const int32_t off = w->off[k - w->alo];
        if ((uint32_t)off <= (uint32_t)tlen && (uint32_t)(off - k) <= (uint32_t)plen) break;
    }
    w->hi = k;
    for (k = w->lo; k <= w->hi; ++k) {
        const int32_t off = w->off[k - w->alo];
        if ((uint32_t)off <= (uint32_t)tlen && (uint32_t)(off - k) <= (uint32_t)plen) break;
    }
    w->lo = k;
}

enum { C_M = 0, C_I1, C_D1, C_I2, C_D2, C_N };

/*
 * Score of one pair under gap-affine 2-piece penalties (>= 0), or a negative
 * error.  cells (optional) receives the sum of computed range widths.  With
 * keep, every score has its own slot and *kept receives the wavefronts
 * (ring = s_bound + 1 slots of C_N components, slot s), for the caller to
 * free with p2_release.
 */
static int p2_core(const char* p, int plen, const char* t, int tlen, int x, int o1, int e1, int o2, int e2,
                   int64_t* cells, int keep, p2_wf** kept, int* kept_ring)
{
    if (plen < 0 || tlen < 0 || x <= 0 || o1 < 0 || e1 <= 0 || o2 < 0 || e2 <= 0) return P2_EINVAL;
    const int align_k = tlen - plen, align_off = tlen;
    const int64_t gap_tail = plen != tlen ? P2_MIN((int64_t)o1 + (int64_t)e1 * abs(plen - tlen),
                                                   (int64_t)o2 + (int64_t)e2 * abs(plen - tlen)) : 0;
    const int64_t s_bound = (int64_t)x * P2_MIN(plen, tlen) + gap_tail;
    if (keep && s_bound > (1 << 24)) return P2_EINVAL;
    const int ring = keep ? (int)s_bound + 1 : P2_MAX(x, P2_MAX(o1 + e1, o2 + e2)) + 1;
    /* wavefronts of the last `ring` scores, slot s % ring */
    p2_wf* W = (p2_wf*)calloc((size_t)ring * C_N, sizeof(p2_wf));
    if (!W) return P2_ENOMEM;
    for (int i = 0; i < ring * C_N; ++i) { W[i].lo = 1; W[i].hi = 0; W[i].alo = 1; }
#define WF(sc, c) (&W[((sc) % ring) * C_N + (c)])
    static const p2_wf NULLWF = { 1, 0, 1, NULL };
    int r = P2_EINTERNAL;
    if (p2_alloc(WF(0, C_M), 0, 0)) { r = P2_ENOMEM; goto out; }
    WF(0, C_M)->off[0] = 0;
    for (int s = 0;; ++s) {
        p2_wf* M = WF(s, C_M);
        if (!p2_null(M)) {
            for (int k = M->lo; k <= M->hi; ++k) {
                int32_t off = M->off[k - M->alo];
                if (off < 0) continue;
                int v = off - k, h = off;
                while (v < plen && h < tlen && p[v] == t[h]) { ++v; ++h; }
                M->off[k - M->alo] = h;
            }
            if (M->lo <= align_k && align_k <= M->hi && M->off[align_k - M->alo] >= align_off) { r = s; break; }
        }
        const int sn = s + 1;
        if (sn > s_bound) { r = P2_EINTERNAL; break; }
        const p2_wf* Mx = sn - x >= 0 ? WF(sn - x, C_M) : &NULLWF;
        const p2_wf* Mo1 = sn - o1 - e1 >= 0 ? WF(sn - o1 - e1, C_M) : &NULLWF;
        const p2_wf* Mo2 = sn - o2 - e2 >= 0 ? WF(sn - o2 - e2, C_M) : &NULLWF;
        const p2_wf* I1e = sn - e1 >= 0 ? WF(sn - e1, C_I1) : &NULLWF;
        const p2_wf* D1e = sn - e1 >= 0 ? WF(sn - e1, C_D1) : &NULLWF;
        const p2_wf* I2e = sn - e2 >= 0 ? WF(sn - e2, C_I2) : &NULLWF;
        const p2_wf* D2e = sn - e2 >= 0 ? WF(sn - e2, C_D2) : &NULLWF;
        /* the output slots held score sn - ring, no longer a source */
        p2_wf* out[C_N];
        for (int c = 0; c < C_N; ++c) { out[c] = WF(sn, c); p2_free(out[c]); }
        const int nx = p2_null(Mx), no1 = p2_null(Mo1), no2 = p2_null(Mo2);
        const int ni1 = p2_null(I1e), nd1 = p2_null(D1e), ni2 = p2_null(I2e), nd2 = p2_null(D2e);
        if (nx && no1 && no2 && ni1 && nd1 && ni2 && nd2) continue;  /* null step */
        int lo = INT_MAX, hi = INT_MIN;
#define UNION(cond, l, h) do { if (!(cond)) { if ((l) < lo) lo = (l); if ((h) > hi) hi = (h); } } while (0)
        UNION(nx, Mx->lo, Mx->hi);
        UNION(no1, Mo1->lo - 1, Mo1->hi + 1);
        UNION(no2, Mo2->lo - 1, Mo2->hi + 1);
        UNION(ni1, I1e->lo + 1, I1e->hi + 1);
        UNION(ni2, I2e->lo + 1, I2e->hi + 1);
        UNION(nd1, D1e->lo - 1, D1e->hi - 1);
        UNION(nd2, D2e->lo - 1, D2e->hi - 1);
#undef UNION
        if (lo > hi) continue;
        const int has[C_N] = { 1, !no1 || !ni1, !no1 || !nd1, !no2 || !ni2, !no2 || !nd2 };
        for (int c = 0; c < C_N; ++c)
            if (has[c] && p2_alloc(out[c], lo, hi)) { r = P2_ENOMEM; goto out; }
        if (cells) *cells += hi - lo + 1;
        for (int k = lo; k <= hi; ++k) {
            const int32_t i1 = P2_MAX(p2_get(Mo1, k - 1), p2_get(I1e, k - 1)) + 1;
            const int32_t i2 = P2_MAX(p2_get(Mo2, k - 1), p2_get(I2e, k - 1)) + 1;
            const int32_t d1 = P2_MAX(p2_get(Mo1, k + 1), p2_get(D1e, k + 1));
            const int32_t d2 = P2_MAX(p2_get(Mo2, k + 1), p2_get(D2e, k + 1));
            int32_t m = P2_MAX(P2_MAX(p2_get(Mx, k) + 1, P2_MAX(i1, i2)), P2_MAX(d1, d2));
            if ((uint32_t)m > (uint32_t)tlen || (uint32_t)(m - k) > (uint32_t)plen) m = P2_NULL;
            out[C_M]->off[k - lo] = m;
            if (has[C_I1]) out[C_I1]->off[k - lo] = i1;
            if (has[C_D1]) out[C_D1]->off[k - lo] = d1;
            if (has[C_I2]) out[C_I2]->off[k - lo] = i2;
            if (has[C_D2]) out[C_D2]->off[k - lo] = d2;
        }
        for (int c = 0; c < C_N; ++c) if (has[c]) p2_trim(out[c], plen, tlen);
    }
out:
    if (keep && r >= 0) {
        *kept = W;
        *kept_ring = ring;
    } else {
        for (int i = 0; i < ring * C_N; ++i) free(W[i].off);
        free(W);
    }
#undef WF
    return r;
}

static void p2_release(p2_wf* W, int ring)
{
    for (int i = 0; i < ring * C_N; ++i) free(W[i].off);
    free(W);
}

int orc_wfa2p(const char* p, int plen, const char* t, int tlen, int x, int o1, int e1, int o2, int e2,
              int64_t* cells)
{
    return p2_core(p, plen, t, tlen, x, o1, e1, o2, e2, cells, 0, NULL, NULL);
}

/* backtrace type values of common.h:67-77, in the order of prio[] below */
static const int P2_PRIO[9] = { 9, 8, 7, 6, 5, 4, 3, 2, 1 };
enum { B_M = 0, B_D2E, B_D2O, B_D1E, B_D1O, B_I2E, B_I2O, B_I1E, B_I1O };

/*
 * Walk back from the end cell (WFA2-lib wavefront_backtrace_affine with two
 * gap pieces, end-to-end).  prio[] gives the type value of M, D2 ext, D2
 * open, D1 ext, D1 open, I2 ext, I2 open, I1 ext, I1 open.  Returns the
 * number of ops written (forward order) or a negative error.
 */
static int p2_backtrace(const p2_wf* W, int ring, int plen, int tlen, int x, int o1, int e1, int o2, int e2,
                        int score_end, const int* prio, char* ops, int ops_cap)
{
#define WF(sc, c) (&W[((sc) % ring) * C_N + (c)])
#define PIGGY(off, ty) ((((int64_t)(off)) * 16) | (int64_t)(ty))
#define EMIT(ch) do { if (n >= ops_cap) return P2_EINVAL; ops[n++] = (ch); } while (0)
    const int64_t PNULL = ((int64_t)P2_NULL) * 16;
    int n = 0;
    int k = tlen - plen;
    int32_t offset = tlen;
    int v = plen, h = tlen;
    int score = score_end;
    int comp = C_M;
    while (v > 0 && h > 0 && score > 0) {
        const int sx = score - x, so1 = score - o1 - e1, se1 = score - e1, so2 = score - o2 - e2, se2 = score - e2;
        int64_t best = PNULL;
        int bk = -1;
#define CAND(cond, sc, c, kk, add, b)                                                  \
        do {                                                                           \
            if (cond) {                                                                \
                const int64_t cv = PIGGY(p2_get(WF(sc, c), kk) + (add), prio[b]);      \
                if (cv > best) { best = cv; bk = (b); }                                \
            }                                                                          \
        } while (0)
        if (comp == C_M) CAND(sx >= 0, sx, C_M, k, 1, B_M);
        if (comp == C_M || comp == C_I1) {
            CAND(so1 >= 0, so1, C_M, k - 1, 1, B_I1O);
            CAND(se1 >= 0, se1, C_I1, k - 1, 1, B_I1E);
        }
        if (comp == C_M || comp == C_D1) {
            CAND(so1 >= 0, so1, C_M, k + 1, 0, B_D1O);
            CAND(se1 >= 0, se1, C_D1, k + 1, 0, B_D1E);
        }
        if (comp == C_M || comp == C_I2) {
            CAND(so2 >= 0, so2, C_M, k - 1, 1, B_I2O);
            CAND(se2 >= 0, se2, C_I2, k - 1, 1, B_I2E);
        }
        if (comp == C_M || comp == C_D2) {
            CAND(so2 >= 0, so2, C_M, k + 1, 0, B_D2O);
            CAND(se2 >= 0, se2, C_D2, k + 1, 0, B_D2E);
        }
#undef CAND
        if (best < 0 || bk < 0) return P2_EINTERNAL;
        const int32_t max_off = (int32_t)(best >> 4);
        if (comp == C_M) {
            const int nm = offset - max_off;
            if (nm < 0) return P2_EINTERNAL;
            for (int i = 0; i < nm; ++i) EMIT('M');
            offset = max_off;
            v = offset - k; h = offset;
            if (v <= 0 || h <= 0) break;
        }
        switch (bk) {
        case B_M:   score = sx;  comp = C_M;  EMIT('X'); --offset; break;
        case B_I1O: score = so1; comp = C_M;  EMIT('I'); --k; --offset; break;
        case B_I1E: score = se1; comp = C_I1; EMIT('I'); --k; --offset; break;
        case B_I2O: score = so2; comp = C_M;  EMIT('I'); --k; --offset; break;
        case B_I2E: score = se2; comp = C_I2; EMIT('I'); --k; --offset; break;
        case B_D1O: score = so1; comp = C_M;  EMIT('D'); ++k; break;
        case B_D1E: score = se1; comp = C_D1; EMIT('D'); ++k; break;
        case B_D2O: score = so2; comp = C_M;  EMIT('D'); ++k; break;
        case B_D2E: score = se2; comp = C_D2; EMIT('D'); ++k; break;
        default: return P2_EINTERNAL;
        }
        v = offset - k; h = offset;
    }
    if (comp == C_M) {
        if (v > 0 && h > 0) {
            const int nm = P2_MIN(v, h);
            for (int i = 0; i < nm; ++i) EMIT('M');
            v -= nm; h -= nm;
        }
        while (v > 0) { EMIT('D'); --v; }
        while (h > 0) { EMIT('I'); --h; }
    } else if (v != 0 || h != 0) {
        return P2_EINTERNAL;
    }
    for (int i = 0, j = n - 1; i < j; ++i, --j) { const char c = ops[i]; ops[i] = ops[j]; ops[j] = c; }
    return n;
#undef WF
#undef PIGGY
#undef EMIT
}

/*
 * Score and CIGAR ops ('M', 'X', 'I', 'D', forward order) of one pair under
 * 2-piece penalties; prio NULL for common.h's backtrace types, else nine
 * distinct values in 1..15 (see p2_backtrace).  Returns the score or a
 * negative error; *ops_len receives the number of ops.
 */
int orc_wfa2p_prio(const char* p, int plen, const char* t, int tlen, int x, int o1, int e1, int o2, int e2,
                   const int* prio, char* ops, int ops_cap, int* ops_len)
{
    if (!ops || !ops_len) return P2_EINVAL;
    if (prio) {
        for (int i = 0; i < 9; ++i) {
            if (prio[i] < 1 || prio[i] > 15) return P2_EINVAL;
            for (int j = 0; j < i; ++j)
                if (prio[i] == prio[j]) return P2_EINVAL;
        }
    }
    p2_wf* W = NULL;
    int ring = 0;
    const int r = p2_core(p, plen, t, tlen, x, o1, e1, o2, e2, NULL, 1, &W, &ring);
    if (r < 0) return r;
    const int n = p2_backtrace(W, ring, plen, tlen, x, o1, e1, o2, e2, r, prio ? prio : P2_PRIO, ops, ops_cap);
    p2_release(W, ring);
    if (n < 0) return n;
    *ops_len = n;
    return r;
}

/* Many pairs under OpenMP; returns the number of failed pairs. */
int orc_wfa2p_batch(const char* seqs, const int64_t* p_off, const int32_t* p_len,
                    const int64_t* t_off, const int32_t* t_len, int n,
                    int x, int o1, int e1, int o2, int e2, int threads, int32_t* scores, int64_t* cells)
{
    int failed = 0;
    int64_t cl = 0;
#ifdef _OPENMP
    if (threads > 0) omp_set_num_threads(threads);
#else
    (void)threads;
#endif
#pragma omp parallel for schedule(dynamic, 16) reduction(+:failed, cl)
    for (int i = 0; i < n; ++i) {
        int64_t c = 0;
        const int r = orc_wfa2p(seqs + p_off[i], p_len[i], seqs + t_off[i], t_len[i], x, o1, e1, o2, e2, &c);
        scores[i] = r;
        cl += c;
        if (r < 0) failed++;
    }
    if (cells) *cells = cl;
    return failed;
}
