// cpu_check.cpp -- the CPU side of `wfa_gpu ... check=1`.
//
// The reference's check mode (wfa_gpu.cu:150-185) aligns every pair again on
// the CPU with WFA2-lib under OpenMP, prints "CPU Time" / "Estimated GCUPS SW"
// and compares the scores.  WFA2-lib is not part of this build, so the check
// uses this score-only gap-affine WFA: penalties divided by their gcd, one
// ring of full-width diagonal arrays per thread (diagonals -p-1 .. t+1, NULL
// outside the live range), extend by 8-byte compares.  It is a verification
// mode of the CLI only; the library never computes alignments on the CPU.
#include "cpu_check.h"

#include <algorithm>
#include <cstdint>
#include <cstring>
#include <vector>
#ifdef _OPENMP
#include <omp.h>
#endif

namespace {

constexpr int32_t kNull = INT32_MIN / 2;

struct Ring {
    int32_t rm = 0, ri = 0, width = 0, base = 0;   // base: index of diagonal 0
    std::vector<int32_t> M, I, D;                   // rm / ri / ri slots of width
    std::vector<int32_t> lo, hi;                    // per slot: M [0,rm), I [rm, rm+ri), D [rm+ri, rm+2ri)
};

int64_t gcd64(int64_t a, int64_t b) { while (b) { int64_t t = a % b; a = b; b = t; } return a; }

inline int lcp_bytes(const char* p, int plen, const char* t, int tlen, int v, int h) {
    const int lim = std::min(plen - v, tlen - h);
    int n = 0;
    while (n + 8 <= lim) {
        uint64_t a, b;
        std::memcpy(&a, p + v + n, 8);
        std::memcpy(&b, t + h + n, 8);
        const uint64_t x = a ^ b;
        if (x) return n + (__builtin_ctzll(x) >> 3);
        n += 8;
    }
    while (n < lim && p[v + n] == t[h + n]) ++n;
    return n;
}

// Score of one pair; -1 on an internal error.
int64_t wfa_score(Ring& r, const char* p, int plen, const char* t, int tlen, int x, int o, int e) {
    const int width = plen + tlen + 5;
    if (r.width < width) {
        r.width = width;
        r.M.assign((size_t)r.rm * width, kNull);
        r.I.assign((size_t)r.ri * width, kNull);
        r.D.assign((size_t)r.ri * width, kNull);
    } else {
        std::fill(r.M.begin(), r.M.end(), kNull);
        std::fill(r.I.begin(), r.I.end(), kNull);
        std::fill(r.D.begin(), r.D.end(), kNull);
    }
    const int W = r.width;
    r.base = plen + 2;
    r.lo.assign((size_t)(r.rm + 2 * r.ri), 1);
    r.hi.assign((size_t)(r.rm + 2 * r.ri), 0);
    auto slotM = [&](int s) { return r.M.data() + (size_t)(s % r.rm) * W + r.base; };
    auto slotI = [&](int s) { return r.I.data() + (size_t)(s % r.ri) * W + r.base; };
    auto slotD = [&](int s) { return r.D.data() + (size_t)(s % r.ri) * W + r.base; };
    const int tp = tlen - plen;
    int32_t* m0 = slotM(0);
    m0[0] = lcp_bytes(p, plen, t, tlen, 0, 0);
    r.lo[0] = 0; r.hi[0] = 0;
    if (tp == 0 && m0[0] >= tlen) return 0;
    const int64_t bound = (int64_t)x * std::min(plen, tlen) + (plen != tlen ? o + (int64_t)e * std::abs(plen - tlen) : 0);
    for (int s = 1; s <= bound; ++s) {
        const int sm = s - x, so = s - o - e, se = s - e;
        auto rng = [&](int slot, int sc, int& lo, int& hi) {
            if (sc < 0) { lo = 1; hi = 0; return; }
            lo = r.lo[(size_t)slot]; hi = r.hi[(size_t)slot];
        };
        int lx, hx, lo_o, ho, li, hi_i, ld, hd;
        rng(sm >= 0 ? sm % r.rm : 0, sm, lx, hx);
        rng(so >= 0 ? so % r.rm : 0, so, lo_o, ho);
        rng(se >= 0 ? r.rm + se % r.ri : 0, se, li, hi_i);
        rng(se >= 0 ? r.rm + r.ri + se % r.ri : 0, se, ld, hd);
        const bool nx = lx > hx, no = lo_o > ho, ni = li > hi_i, nd = ld > hd;
        int32_t* Mt = slotM(s);
        int32_t* It = slotI(s);
        int32_t* Dt = slotD(s);
        const size_t sM = (size_t)(s % r.rm), sI = (size_t)(r.rm + s % r.ri), sD = (size_t)(r.rm + r.ri + s % r.ri);
        // clear what the reused slots held
        for (int k = r.lo[sM]; k <= r.hi[sM]; ++k) Mt[k] = kNull;
        for (int k = r.lo[sI]; k <= r.hi[sI]; ++k) It[k] = kNull;
        for (int k = r.lo[sD]; k <= r.hi[sD]; ++k) Dt[k] = kNull;
        r.lo[sM] = r.lo[sI] = r.lo[sD] = 1;
        r.hi[sM] = r.hi[sI] = r.hi[sD] = 0;
        if (nx && no && ni && nd) continue;
        int lo = INT32_MAX, hi = INT32_MIN;
        if (!nx) { lo = std::min(lo, lx); hi = std::max(hi, hx); }
        if (!no) { lo = std::min(lo, lo_o - 1); hi = std::max(hi, ho + 1); }
        if (!ni) { lo = std::min(lo, li + 1); hi = std::max(hi, hi_i + 1); }
        if (!nd) { lo = std::min(lo, ld - 1); hi = std::max(hi, hd - 1); }
        const int32_t* Mx = nx ? nullptr : slotM(sm);
        const int32_t* Mo = no ? nullptr : slotM(so);
        const int32_t* Ie = ni ? nullptr : slotI(se);
        const int32_t* De = nd ? nullptr : slotD(se);
        const bool hasI = !no || !ni, hasD = !no || !nd;
        bool done = false;
        for (int k = lo; k <= hi; ++k) {
            const int32_t mol = Mo ? Mo[k - 1] : kNull, mor = Mo ? Mo[k + 1] : kNull;
            const int32_t iel = Ie ? Ie[k - 1] : kNull, der = De ? De[k + 1] : kNull;
            const int32_t ins = std::max(mol, iel) + 1;
            const int32_t del = std::max(mor, der);
            const int32_t mis = (Mx ? Mx[k] : kNull) + 1;
            int32_t m = std::max(del, std::max(mis, ins));
            if ((uint32_t)m > (uint32_t)tlen || (uint32_t)(m - k) > (uint32_t)plen) m = kNull;
            if (m >= 0) {
                m += lcp_bytes(p, plen, t, tlen, m - k, m);
                if (k == tp && m >= tlen) done = true;
            }
            Mt[k] = m;
            if (hasI) It[k] = ins < 0 ? kNull : ins;
            if (hasD) Dt[k] = del;
        }
        // trim ends (wfa_gpu.cu:633-719); values outside the kept range -> NULL
        auto trim = [&](int32_t* w, size_t slot, bool has) {
            if (!has) return;
            int a = lo, b = hi;
            auto in = [&](int k) { return (uint32_t)w[k] <= (uint32_t)tlen && (uint32_t)(w[k] - k) <= (uint32_t)plen; };
            while (b >= a && !in(b)) { w[b] = kNull; --b; }
            while (a <= b && !in(a)) { w[a] = kNull; ++a; }
            r.lo[slot] = a; r.hi[slot] = b;
            if (a > b) { r.lo[slot] = 1; r.hi[slot] = 0; }
        };
        trim(Mt, sM, true);
        trim(It, sI, hasI);
        trim(Dt, sD, hasD);
        if (done) return s;
    }
    return -1;
}

}  // namespace

int cpu_check_scores(const tsu_dataset* ds, int32_t x, int32_t o, int32_t e, std::vector<int32_t>& out, int* threads_used) {
    const int64_t n = tsu_dataset_size(ds);
    out.assign((size_t)n, 0);
    const int64_t g = gcd64(gcd64(x, o), e);
    const int xs = (int)(x / g), os = (int)(o / g), es = (int)(e / g);
    int failed = 0;
    int nt = 1;
#pragma omp parallel reduction(+:failed)
    {
#ifdef _OPENMP
#pragma omp single
        nt = omp_get_num_threads();
#endif
        Ring r;
        r.rm = std::max(xs, os + es) + 1;
        r.ri = es + 1;
#pragma omp for schedule(dynamic, 64)
        for (int64_t i = 0; i < n; ++i) {
            const char *p, *t;
            int32_t pl, tl;
            tsu_dataset_pair(ds, i, &p, &pl, &t, &tl);
            const int64_t s = wfa_score(r, p, pl, t, tl, xs, os, es);
            if (s < 0) { failed++; out[(size_t)i] = -1; }
            else out[(size_t)i] = (int32_t)(s * g);
        }
    }
    if (threads_used) *threads_used = nt;
    return failed;
}
