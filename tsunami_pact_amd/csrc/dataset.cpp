// dataset.cpp -- host-side inputs of the aligner: the sequences.txt format,
// the seeded pair generator and the 2-bit packer.
//
// Format (reference: wfa_gpu.cu:71-81 reads it, generate_dataset.cpp:39-41,68-69
// writes it): three integers "num len err", then for each pair a pattern
// token and a text token separated by white space.  The reference reads each
// token into a fixed slot of len characters; a token longer than len (its own
// generator prints buffers that are not NUL-terminated, generate_dataset.cpp:
// 26-27,68-69, so heap bytes can follow the sequence) therefore contributes
// its first len characters.  That rule is kept.  Shorter tokens are kept whole, which makes
// the same format carry pairs of unequal length (indels).
#include "tsu_internal.h"

#include <algorithm>
#include <cmath>
#include <cstddef>
#include <cstdio>
#include <cstring>
#include <new>
#include <string>
#include <vector>
#ifdef _OPENMP
#include <omp.h>
#endif

namespace {

// Base codes: A=0, C=1, G=2, T=3; 0xFF for anything else.
struct BaseTable {
    uint8_t code[256];
    BaseTable() {
        std::memset(code, 0xFF, sizeof(code));
        code[(unsigned char)'A'] = 0;
        code[(unsigned char)'C'] = 1;
        code[(unsigned char)'G'] = 2;
        code[(unsigned char)'T'] = 3;
    }
};
const BaseTable kBases;
const char kAlphabet[4] = {'A', 'C', 'G', 'T'};

// splitmix64 seeds a xoshiro256** stream per pair.
inline uint64_t splitmix64(uint64_t& x) {
    uint64_t z = (x += 0x9E3779B97F4A7C15ull);
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}

struct Xoshiro {
    uint64_t s[4];
    Xoshiro(uint64_t seed, uint64_t stream) {
        uint64_t x = seed ^ (0xD1B54A32D192ED03ull * (stream + 1));
        for (int i = 0; i < 4; ++i) s[i] = splitmix64(x);
    }
    static inline uint64_t rotl(uint64_t v, int k) { return (v << k) | (v >> (64 - k)); }
    uint64_t next() {
        const uint64_t r = rotl(s[1] * 5, 7) * 9;
        const uint64_t t = s[1] << 17;
        s[2] ^= s[0]; s[3] ^= s[1]; s[1] ^= s[2]; s[0] ^= s[3];
        s[2] ^= t; s[3] = rotl(s[3], 45);
        return r;
    }
    // uniform in [0, n)
    uint64_t below(uint64_t n) { return (uint64_t)(((unsigned __int128)next() * n) >> 64); }
    double unit() { return (double)(next() >> 11) * (1.0 / 9007199254740992.0); }
};

void generate_pair(const tsu_gen_cfg& cfg, uint64_t index, std::string& pat, std::string& txt,
                   std::vector<uint8_t>& mark) {
    Xoshiro rng(cfg.seed, index);
    int32_t L = cfg.len_min;
    if (cfg.len_max > cfg.len_min) {
        const double a = std::log((double)cfg.len_min), b = std::log((double)cfg.len_max);
        double v = std::floor(std::exp(a + rng.unit() * (b - a)));
        L = (int32_t)std::min<double>(std::max<double>(v, cfg.len_min), cfg.len_max);
    }
    double rate = cfg.err_min;
    if (cfg.err_max > cfg.err_min) rate = cfg.err_min + rng.unit() * (cfg.err_max - cfg.err_min);
    // edit_mode 2 counts like the reference's integer arithmetic,
    // len * percent / 100 rounded down (generate_dataset.cpp:22-23)
    int64_t edits = cfg.edit_mode == 2 ? (int64_t)L * (int64_t)std::llround(rate * 100.0) / 100
                                       : (int64_t)std::llround(rate * (double)L);
    if (edits < 0) edits = 0;
    if (edits > L) edits = L;
    pat.resize((size_t)L);
    for (int32_t j = 0; j < L; ++j) pat[(size_t)j] = kAlphabet[rng.below(4)];
    // distinct positions: Floyd's sampling over a bitmap
    mark.assign((size_t)L, 0);
    for (int64_t j = L - edits; j < L; ++j) {
        uint64_t r = rng.below((uint64_t)j + 1);
        if (mark[(size_t)r]) mark[(size_t)j] = 1; else mark[(size_t)r] = 1;
    }
    txt.clear();
    txt.reserve((size_t)L + (size_t)edits + 1);
    for (int32_t j = 0; j < L; ++j) {
        const char c = pat[(size_t)j];
        if (!mark[(size_t)j]) { txt.push_back(c); continue; }
        const int type = cfg.edit_mode != 0 ? 0 : (int)rng.below(3);
        if (type == 0) {                 // substitution to a different base
            const int cur = kBases.code[(unsigned char)c];
            txt.push_back(kAlphabet[(cur + 1 + (int)rng.below(3)) & 3]);
        } else if (type == 1) {          // insertion of a uniform base before c
            txt.push_back(kAlphabet[rng.below(4)]);
            txt.push_back(c);
        }                                 // type 2: deletion of c
    }
}

}  // namespace

int tsu_pack_seq(const char* seq, int32_t len, uint32_t* out) {
    const int64_t nw = ((int64_t)len + 15) / 16;
    for (int64_t w = 0; w < nw; ++w) {
        uint32_t word = 0;
        const int64_t base = w * 16;
        const int m = (int)std::min<int64_t>(16, len - base);
        for (int i = 0; i < m; ++i) {
            const uint8_t c = kBases.code[(unsigned char)seq[base + i]];
            if (c > 3) return TSU_EBADINPUT;
            word |= (uint32_t)c << (2 * i);
        }
        out[w] = word;
    }
    return TSU_OK;
}

bool tsu_is_acgt(const char* seq, int64_t len, int64_t* bad_pos) {
    for (int64_t i = 0; i < len; ++i)
        if (kBases.code[(unsigned char)seq[i]] > 3) { if (bad_pos) *bad_pos = i; return false; }
    return true;
}

extern "C" const char* tsu_strerror(int status) {
    switch (status) {
    case TSU_OK: return "ok";
    case TSU_EINVAL: return "invalid argument";
    case TSU_EBADINPUT: return "bad input (malformed file or non-ACGT base)";
    case TSU_EKOVERFLOW: return "diagonal range exceeded wf_length";
    case TSU_ENOMEM: return "out of memory";
    case TSU_EHIP: return "HIP runtime error";
    case TSU_EBUDGET: return "request exceeds the memory budget";
    case TSU_EAGAIN: return "not ready";
    case TSU_EIO: return "I/O error";
    case TSU_EINTERNAL: return "internal error";
    case TSU_ECIGAR: return "backtrace buffers exhausted";
    default: return "unknown status";
    }
}

extern "C" int64_t tsu_packed_words(int32_t len) { return len <= 0 ? 0 : ((int64_t)len + 15) / 16; }

extern "C" int tsu_pack(const char* seq, int32_t len, uint32_t* out) {
    if (len < 0 || (len > 0 && (!seq || !out))) return TSU_EINVAL;
    return tsu_pack_seq(seq, len, out);
}

// ---------------------------------------------------------------------------
// datasets
// ---------------------------------------------------------------------------
static void set_err(char* err, int32_t cap, const std::string& m) {
    if (err && cap > 0) { std::snprintf(err, (size_t)cap, "%s", m.c_str()); }
}

extern "C" int tsu_dataset_read(const char* path, tsu_dataset** out, char* err, int32_t err_cap) {
    if (!path || !out) return TSU_EINVAL;
    *out = nullptr;
    FILE* fp = std::fopen(path, "rb");
    if (!fp) { set_err(err, err_cap, std::string("cannot open ") + path); return TSU_EIO; }
    std::vector<char> buf;
    {
        char tmp[1 << 16];
        size_t n;
        while ((n = std::fread(tmp, 1, sizeof(tmp), fp)) > 0) buf.insert(buf.end(), tmp, tmp + n);
    }
    std::fclose(fp);
    size_t pos = 0;
    const size_t end = buf.size();
    auto skip_ws = [&]() { while (pos < end && (buf[pos] == ' ' || buf[pos] == '\n' || buf[pos] == '\r' || buf[pos] == '\t' || buf[pos] == '\v' || buf[pos] == '\f')) ++pos; };
    auto next_token = [&](size_t& b, size_t& e) -> bool {
        skip_ws();
        if (pos >= end) return false;
        b = pos;
        while (pos < end && !(buf[pos] == ' ' || buf[pos] == '\n' || buf[pos] == '\r' || buf[pos] == '\t' || buf[pos] == '\v' || buf[pos] == '\f')) ++pos;
        e = pos;
        return true;
    };
    long long hdr[3];
    for (int i = 0; i < 3; ++i) {
        size_t b, e;
        if (!next_token(b, e)) { set_err(err, err_cap, "truncated header (expected: num len err)"); return TSU_EBADINPUT; }
        std::string tok(&buf[b], e - b);
        char* endp = nullptr;
        hdr[i] = std::strtoll(tok.c_str(), &endp, 10);
        if (!endp || *endp != '\0') { set_err(err, err_cap, "header field is not an integer: " + tok); return TSU_EBADINPUT; }
    }
    if (hdr[0] < 0 || hdr[0] > (1ll << 31) - 1) { set_err(err, err_cap, "bad pair count in header"); return TSU_EBADINPUT; }
    const int64_t num = hdr[0];
    const int64_t L = hdr[1];
    tsu_dataset* ds = new (std::nothrow) tsu_dataset();
    if (!ds) return TSU_ENOMEM;
    ds->hdr_num = (int32_t)hdr[0];
    ds->hdr_len = (int32_t)hdr[1];
    ds->hdr_err = (int32_t)hdr[2];
    try {
        ds->p_off.resize((size_t)num); ds->t_off.resize((size_t)num);
        ds->p_len.resize((size_t)num); ds->t_len.resize((size_t)num);
        ds->chars.reserve(buf.size());
    } catch (...) { delete ds; return TSU_ENOMEM; }
    for (int64_t i = 0; i < num; ++i) {
        for (int which = 0; which < 2; ++which) {
            size_t b, e;
            if (!next_token(b, e)) {
                set_err(err, err_cap, "file ends at pair " + std::to_string(i) + " of " + std::to_string(num));
                delete ds; return TSU_EBADINPUT;
            }
            int64_t n = (int64_t)(e - b);
            if (L > 0 && n > L) n = L;          // fixed-slot rule (wfa_gpu.cu:78-81)
            if (n > INT32_MAX) { set_err(err, err_cap, "token too long"); delete ds; return TSU_EBADINPUT; }
            int64_t bad = 0;
            if (!tsu_is_acgt(&buf[b], n, &bad)) {
                set_err(err, err_cap, "pair " + std::to_string(i) + (which ? " text" : " pattern") +
                                          ": non-ACGT character at position " + std::to_string(bad));
                delete ds; return TSU_EBADINPUT;
            }
            const int64_t off = (int64_t)ds->chars.size();
            ds->chars.insert(ds->chars.end(), buf.begin() + (std::ptrdiff_t)b, buf.begin() + (std::ptrdiff_t)(b + (size_t)n));
            if (which == 0) { ds->p_off[(size_t)i] = off; ds->p_len[(size_t)i] = (int32_t)n; }
            else { ds->t_off[(size_t)i] = off; ds->t_len[(size_t)i] = (int32_t)n; }
        }
    }
    *out = ds;
    return TSU_OK;
}

extern "C" int tsu_dataset_write(const tsu_dataset* ds, const char* path, int32_t error_rate_field) {
    if (!ds || !path) return TSU_EINVAL;
    FILE* fp = std::fopen(path, "wb");
    if (!fp) return TSU_EIO;
    int32_t maxlen = 0;
    for (size_t i = 0; i < ds->p_len.size(); ++i) maxlen = std::max(maxlen, std::max(ds->p_len[i], ds->t_len[i]));
    std::fprintf(fp, "%lld\n%d\n%d\n", (long long)ds->p_len.size(), maxlen, error_rate_field);
    for (size_t i = 0; i < ds->p_len.size(); ++i) {
        std::fwrite(ds->chars.data() + ds->p_off[i], 1, (size_t)ds->p_len[i], fp);
        std::fputc('\n', fp);
        std::fwrite(ds->chars.data() + ds->t_off[i], 1, (size_t)ds->t_len[i], fp);
        std::fputc('\n', fp);
    }
    const bool ok = std::fflush(fp) == 0;
    std::fclose(fp);
    return ok ? TSU_OK : TSU_EIO;
}

extern "C" int tsu_dataset_generate(const tsu_gen_cfg* cfg, int64_t n_pairs, tsu_dataset** out) {
    if (!cfg || !out || n_pairs < 0 || n_pairs > INT32_MAX) return TSU_EINVAL;
    if (cfg->len_min < 1 || cfg->len_max < cfg->len_min) return TSU_EINVAL;
    if (!(cfg->err_min >= 0.0) || !(cfg->err_max >= cfg->err_min) || cfg->err_max > 1.0) return TSU_EINVAL;
    if (cfg->edit_mode < 0 || cfg->edit_mode > 2) return TSU_EINVAL;
    *out = nullptr;
    tsu_dataset* ds = new (std::nothrow) tsu_dataset();
    if (!ds) return TSU_ENOMEM;
    try {
        ds->p_off.resize((size_t)n_pairs); ds->t_off.resize((size_t)n_pairs);
        ds->p_len.resize((size_t)n_pairs); ds->t_len.resize((size_t)n_pairs);
    } catch (...) { delete ds; return TSU_ENOMEM; }
    // Chunks are generated into local buffers, then concatenated in order.
    const int64_t chunk = 4096;
    const int64_t nchunks = (n_pairs + chunk - 1) / chunk;
    std::vector<std::vector<char>> local((size_t)nchunks);
    bool oom = false;
#pragma omp parallel
    {
        std::string pat, txt;
        std::vector<uint8_t> mark;
#pragma omp for schedule(dynamic, 1)
        for (int64_t c = 0; c < nchunks; ++c) {
            std::vector<char>& lb = local[(size_t)c];
            const int64_t lo = c * chunk, hi = std::min(n_pairs, lo + chunk);
            try {
                for (int64_t i = lo; i < hi; ++i) {
                    generate_pair(*cfg, (uint64_t)i, pat, txt, mark);
                    ds->p_off[(size_t)i] = (int64_t)lb.size();
                    ds->p_len[(size_t)i] = (int32_t)pat.size();
                    lb.insert(lb.end(), pat.begin(), pat.end());
                    ds->t_off[(size_t)i] = (int64_t)lb.size();
                    ds->t_len[(size_t)i] = (int32_t)txt.size();
                    lb.insert(lb.end(), txt.begin(), txt.end());
                }
            } catch (...) {
#pragma omp critical
                oom = true;
            }
        }
    }
    if (oom) { delete ds; return TSU_ENOMEM; }
    std::vector<int64_t> base((size_t)nchunks + 1, 0);
    for (int64_t c = 0; c < nchunks; ++c) base[(size_t)c + 1] = base[(size_t)c] + (int64_t)local[(size_t)c].size();
    try { ds->chars.resize((size_t)base[(size_t)nchunks]); } catch (...) { delete ds; return TSU_ENOMEM; }
#pragma omp parallel for schedule(dynamic, 1)
    for (int64_t c = 0; c < nchunks; ++c) {
        std::memcpy(ds->chars.data() + base[(size_t)c], local[(size_t)c].data(), local[(size_t)c].size());
        std::vector<char>().swap(local[(size_t)c]);
        const int64_t lo = c * chunk, hi = std::min(n_pairs, lo + chunk);
        for (int64_t i = lo; i < hi; ++i) {
            ds->p_off[(size_t)i] += base[(size_t)c];
            ds->t_off[(size_t)i] += base[(size_t)c];
        }
    }
    *out = ds;
    return TSU_OK;
}

extern "C" int tsu_dataset_from_strings(const char* const* patterns, const int32_t* p_len,
                                        const char* const* texts, const int32_t* t_len, int64_t n,
                                        tsu_dataset** out) {
    if (!out || n < 0 || n > INT32_MAX || (n > 0 && (!patterns || !p_len || !texts || !t_len))) return TSU_EINVAL;
    *out = nullptr;
    tsu_dataset* ds = new (std::nothrow) tsu_dataset();
    if (!ds) return TSU_ENOMEM;
    try {
        int64_t total = 0;
        for (int64_t i = 0; i < n; ++i) {
            if (p_len[i] < 0 || t_len[i] < 0) { delete ds; return TSU_EINVAL; }
            total += p_len[i] + t_len[i];
        }
        ds->chars.reserve((size_t)total);
        ds->p_off.resize((size_t)n); ds->t_off.resize((size_t)n);
        ds->p_len.resize((size_t)n); ds->t_len.resize((size_t)n);
        for (int64_t i = 0; i < n; ++i) {
            if (!tsu_is_acgt(patterns[i], p_len[i], nullptr) || !tsu_is_acgt(texts[i], t_len[i], nullptr)) {
                delete ds; return TSU_EBADINPUT;
            }
            ds->p_off[(size_t)i] = (int64_t)ds->chars.size();
            ds->p_len[(size_t)i] = p_len[i];
            ds->chars.insert(ds->chars.end(), patterns[i], patterns[i] + p_len[i]);
            ds->t_off[(size_t)i] = (int64_t)ds->chars.size();
            ds->t_len[(size_t)i] = t_len[i];
            ds->chars.insert(ds->chars.end(), texts[i], texts[i] + t_len[i]);
        }
    } catch (...) { delete ds; return TSU_ENOMEM; }
    *out = ds;
    return TSU_OK;
}

extern "C" int64_t tsu_dataset_size(const tsu_dataset* ds) { return ds ? (int64_t)ds->p_len.size() : 0; }

extern "C" int tsu_dataset_pair(const tsu_dataset* ds, int64_t i, const char** pattern, int32_t* p_len,
                                const char** text, int32_t* t_len) {
    if (!ds || i < 0 || i >= (int64_t)ds->p_len.size()) return TSU_EINVAL;
    if (pattern) *pattern = ds->chars.data() + ds->p_off[(size_t)i];
    if (p_len) *p_len = ds->p_len[(size_t)i];
    if (text) *text = ds->chars.data() + ds->t_off[(size_t)i];
    if (t_len) *t_len = ds->t_len[(size_t)i];
    return TSU_OK;
}

extern "C" int tsu_dataset_lengths(const tsu_dataset* ds, int32_t* p_len, int32_t* t_len) {
    if (!ds || !p_len || !t_len) return TSU_EINVAL;
    std::memcpy(p_len, ds->p_len.data(), ds->p_len.size() * sizeof(int32_t));
    std::memcpy(t_len, ds->t_len.data(), ds->t_len.size() * sizeof(int32_t));
    return TSU_OK;
}

extern "C" int tsu_dataset_header(const tsu_dataset* ds, int32_t* num, int32_t* len, int32_t* err) {
    if (!ds) return TSU_EINVAL;
    if (num) *num = ds->hdr_num;
    if (len) *len = ds->hdr_len;
    if (err) *err = ds->hdr_err;
    return TSU_OK;
}

extern "C" void tsu_dataset_free(tsu_dataset* ds) { delete ds; }

// Packs pairs [first, first+count) into one arena: pattern words then text
// words of each pair, in order.
int tsu_pack_dataset(const tsu_dataset* ds, int64_t first, int64_t count, std::vector<uint32_t>& words,
                     std::vector<tsu_pair>& pairs) {
    pairs.resize((size_t)count);
    int64_t nw = 0;
    for (int64_t j = 0; j < count; ++j) {
        const size_t i = (size_t)(first + j);
        pairs[(size_t)j].p_len = ds->p_len[i];
        pairs[(size_t)j].t_len = ds->t_len[i];
        pairs[(size_t)j].p_off = nw;
        nw += tsu_packed_words(ds->p_len[i]);
        pairs[(size_t)j].t_off = nw;
        nw += tsu_packed_words(ds->t_len[i]);
    }
    try { words.assign((size_t)nw, 0u); } catch (...) { return TSU_ENOMEM; }
    int bad = 0;
#pragma omp parallel for schedule(static, 1024) reduction(+:bad)
    for (int64_t j = 0; j < count; ++j) {
        const size_t i = (size_t)(first + j);
        const tsu_pair& pr = pairs[(size_t)j];
        if (tsu_pack_seq(ds->chars.data() + ds->p_off[i], pr.p_len, words.data() + pr.p_off) != TSU_OK) bad++;
        if (tsu_pack_seq(ds->chars.data() + ds->t_off[i], pr.t_len, words.data() + pr.t_off) != TSU_OK) bad++;
    }
    return bad ? TSU_EBADINPUT : TSU_OK;
}
