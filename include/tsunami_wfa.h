/*
 * tsunami_wfa.h -- C-ABI of the MI355X (gfx950) gap-affine WaveFront
 * Alignment library (libtsunami_hip.so).
 *
 * The reference (TSUNAMI, wfa_gpu.cu) has no library boundary: its host
 * main() reads sequences.txt, allocates device buffers, launches one CUDA
 * block of 32 threads per pair and copies back int16 scores.  Each entry
 * point below replaces one piece of that inline host code; the replaced
 * reference lines are cited next to each declaration.
 *
 * Conventions
 *   - every function returns an int status: TSU_OK (0) or a negative TSU_E*;
 *     tsu_last_error() gives a message for the last failing call on a context;
 *   - the caller owns every host array it passes; the context owns device
 *     memory; no exception crosses this boundary;
 *   - one context per host thread (a server object is internally locked);
 *   - scores are the gap-affine penalty of the optimal end-to-end alignment
 *     (match 0, mismatch x, gap of length L costs o + L*e), i.e. the value the
 *     reference prints as "-(cigar->score)" from WFA2-lib (wfa_gpu.cu:165).
 *     A per-pair score slot holds a negative TSU_E* when that pair failed.
 *   - sequences are DNA over {A,C,G,T} (upper case), 2-bit packed 16 bases
 *     per 32-bit word, base i of a sequence at bits [2*(i%16), 2*(i%16)+2) of
 *     word i/16 (A=0, C=1, G=2, T=3).
 */
#ifndef TSUNAMI_WFA_H
#define TSUNAMI_WFA_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- status codes (wfa_gpu.cu:13-31 printed and exited instead) -------- */
#define TSU_OK            0
#define TSU_EINVAL       -1   /* bad argument */
#define TSU_EBADINPUT    -2   /* malformed file or non-ACGT base */
#define TSU_EKOVERFLOW   -3   /* a pair's diagonal range exceeded wf_length
                                 (the reference silently overwrites memory,
                                 wfa_gpu.cu:217,618-626) */
#define TSU_ENOMEM       -4   /* host or device allocation failed */
#define TSU_EHIP         -5   /* HIP runtime error */
#define TSU_EBUDGET      -6   /* request larger than the server's HBM budget */
#define TSU_EAGAIN       -7   /* result not ready yet (non-blocking poll) */
#define TSU_EIO          -8   /* file could not be opened / written */
#define TSU_EINTERNAL    -9   /* internal invariant violated (a bug) */
#define TSU_ECIGAR      -10   /* CIGAR / backtrace buffers exhausted */

typedef struct tsu_ctx tsu_ctx;
typedef struct tsu_batch tsu_batch;
typedef struct tsu_dataset tsu_dataset;
typedef struct tsu_cigars tsu_cigars;
typedef struct tsu_server tsu_server;

/* One pair inside a packed arena: word offsets (32-bit words) and lengths in
 * bases.  Replaces the fixed slots pattern+i*L / text+i*L (wfa_gpu.cu:76-81). */
typedef struct {
    int64_t p_off;
    int64_t t_off;
    int32_t p_len;
    int32_t t_len;
} tsu_pair;

/* Timing of one alignment call. */
typedef struct {
    double kernel_ms;  /* HIP-event time of the alignment kernels, recorded on
                          the stream they run on */
    double wall_ms;    /* host clock around launch + synchronise: the window
                          the reference times (wfa_gpu.cu:116-125); the plain
                          form waits once for the kernel and the scores' copy
                          to pinned host memory queued behind it, so its
                          window also holds that copy (4 MB per million
                          pairs) */
    double h2d_ms;     /* host->device copies made by this call (0 when the
                          batch is already resident) */
    double d2h_ms;     /* device->host copy of scores (wfa_gpu.cu:128) */
    int64_t pairs;
    double cells;      /* sum of p_len*t_len: "equivalent cells" of the GCUPS
                          figure (wfa_gpu.cu:130-133) */
    int32_t launches;  /* kernel launches issued */
    int32_t retried;   /* pairs re-run with a wider diagonal capacity */
} tsu_timing;

/* ---- context ------------------------------------------------------------ */

/* Replaces the penalty parsing and cudaDeviceReset of wfa_gpu.cu:59-61,96.
 * Penalties must satisfy mismatch > 0, gap_opening >= 0, gap_extension > 0. */
int tsu_create(tsu_ctx** out, int device, int32_t mismatch, int32_t gap_opening,
               int32_t gap_extension);
/* Gap-affine 2-piece penalties: a gap of length L costs
 * min(gap_opening1 + L*gap_extension1, gap_opening2 + L*gap_extension2).
 * The reference declares this model (common.h:52-58,80-87: gap_open2,
 * gap_ext2 and the I2/D2 wavefronts) but never computes it, so this entry
 * point has no counterpart there.  Such a context runs every form (one wave
 * per pair, all five wavefronts in LDS; CIGARs follow the nine backtrace
 * types of common.h:67-77); a pair whose five rings at its capacity exceed
 * the CU's LDS fails with TSU_EKOVERFLOW. */
int tsu_create_2p(tsu_ctx** out, int device, int32_t mismatch, int32_t gap_opening1,
                  int32_t gap_extension1, int32_t gap_opening2, int32_t gap_extension2);
void tsu_destroy(tsu_ctx* ctx);
const char* tsu_last_error(const tsu_ctx* ctx);
const char* tsu_strerror(int status);
/* Device facts the library sizes launches with (CUs, LDS per workgroup). */
int tsu_device_info(tsu_ctx* ctx, int32_t* cu_count, int32_t* lds_bytes, char* name, int32_t name_cap);

/* ---- packing (host) ------------------------------------------------------ */

/* Words needed to hold len bases: (len + 15) / 16. */
int64_t tsu_packed_words(int32_t len);
/* Packs len bases of seq into out[0, tsu_packed_words(len)).  Returns
 * TSU_EBADINPUT on a character outside {A,C,G,T}.  The reference compares
 * raw bytes (wfa_gpu.cu:735,276); its generator emits only ACGT
 * (generate_dataset.cpp:25,50,63-65). */
int tsu_pack(const char* seq, int32_t len, uint32_t* out);

/* ---- host datasets: sequences.txt and the seeded generator --------------- */

/* Reads a sequences.txt file: header "num len err" then per pair a pattern
 * token and a text token (wfa_gpu.cu:71-81, generate_dataset.cpp:39-41,68-69).
 * A token longer than the header length keeps its first len characters, as
 * the reference's fixed slots do; shorter tokens are kept whole (variable-
 * length files written by tsu_dataset_write).  err (may be NULL) receives a
 * message on failure. */
int tsu_dataset_read(const char* path, tsu_dataset** out, char* err, int32_t err_cap);
/* Writes the same format; the header length is the longest token. */
int tsu_dataset_write(const tsu_dataset* ds, const char* path, int32_t error_rate_field);

/* Seeded generator (replaces generate_dataset.cpp:11-75, which is
 * time-seeded and substitution-only).  Pattern length is len_min when
 * len_min == len_max, else floor(exp(U(ln len_min, ln len_max))); the edit
 * rate is err_min, or U(err_min, err_max); edits = round(rate * len) at
 * distinct pattern positions; each edit is a substitution (to a different
 * base), an insertion (uniform base) or a deletion with probability 1/3
 * (edit_mode 0), or a substitution only (edit_mode 1).  edit_mode 2 is the
 * reference generator's model: substitutions only, len * percent / 100
 * rounded down of them, with percent = rate * 100 rounded to a whole number
 * (generate_dataset.cpp:20-23,56-66).  Pair i depends only on (seed, i), so
 * the result is independent of the thread count. */
typedef struct {
    uint64_t seed;
    int32_t len_min;
    int32_t len_max;
    double err_min;
    double err_max;
    int32_t edit_mode;
    int32_t reserved;
} tsu_gen_cfg;
int tsu_dataset_generate(const tsu_gen_cfg* cfg, int64_t n_pairs, tsu_dataset** out);
/* Builds a dataset from caller strings (copied). */
int tsu_dataset_from_strings(const char* const* patterns, const int32_t* p_len,
                             const char* const* texts, const int32_t* t_len, int64_t n,
                             tsu_dataset** out);
int64_t tsu_dataset_size(const tsu_dataset* ds);
int tsu_dataset_pair(const tsu_dataset* ds, int64_t i, const char** pattern, int32_t* p_len,
                     const char** text, int32_t* t_len);
/* Lengths of all pairs: p_len[i], t_len[i] for i < tsu_dataset_size(ds). */
int tsu_dataset_lengths(const tsu_dataset* ds, int32_t* p_len, int32_t* t_len);
/* Header fields as read (or 0 for generated data). */
int tsu_dataset_header(const tsu_dataset* ds, int32_t* num, int32_t* len, int32_t* err);
void tsu_dataset_free(tsu_dataset* ds);

/* ---- device-resident batches (wfa_gpu.cu:96-111: cudaMalloc + H2D) ------- */

/* Sequences of up to about 327,000 bases (the exact bound follows from the
 * device's LDS size) are staged in LDS while their pair is aligned; a pair
 * with a longer one runs on the LDS kernel reading its packed words in place
 * from HBM, as the reference reads its sequences from global memory
 * (wfa_gpu.cu:196-197,266-276), in every form.  A sequence may hold at most
 * 2^28 bases; 2-piece contexts stage both sequences in LDS and refuse
 * longer ones.  Batches (and server requests) beyond these limits are refused
 * with TSU_EINVAL. */

/* Packs pairs [first, first+count) of ds and copies them to HBM. */
int tsu_batch_from_dataset(tsu_ctx* ctx, const tsu_dataset* ds, int64_t first, int64_t count,
                           tsu_batch** out);
/* Uploads an already packed arena. */
int tsu_batch_create(tsu_ctx* ctx, const uint32_t* words, int64_t n_words, const tsu_pair* pairs,
                     int32_t n_pairs, tsu_batch** out);
int32_t tsu_batch_size(const tsu_batch* b);
/* The batch's device arrays: packed words (n_words 32-bit words), pair
 * descriptors and per-pair score slots (any pointer may be NULL).  They stay
 * valid until tsu_batch_free.  The reference keeps its device buffers inside
 * main() (wfa_gpu.cu:96-111); this hands a resident batch to callers' own
 * device code or to tsu_align_batched_noalloc. */
int tsu_batch_device(const tsu_batch* b, const uint32_t** d_words, int64_t* n_words,
                     const tsu_pair** d_pairs, int32_t** d_scores);
/* Bytes of HBM the batch holds (packed words + descriptors + score slots). */
int64_t tsu_batch_bytes(const tsu_batch* b);
/* May be called before or after tsu_destroy of the batch's context. */
void tsu_batch_free(tsu_batch* b);

/* ---- alignment ----------------------------------------------------------- */

/* Plain form: one workgroup per pair, diagonal capacity wf_length as the
 * reference's argv[5] (wfa_gpu.cu:63,217).  Replaces the launch of
 * wfa_dispatcher<<<N,32>>> and the score copy (wfa_gpu.cu:113-128,191-194).
 * scores[i] receives pair i's score; a pair whose diagonal range exceeds the
 * capacity gets TSU_EKOVERFLOW and the call returns TSU_EKOVERFLOW.
 * wf_length is the least capacity used: with penalties that divide to
 * 2/3/1 (4/6/2 and its multiples) the register kernels round it up to their
 * window, 64 * 2^j diagonals (wf_length 400 runs in 512), so a pair whose
 * range is wider than wf_length but fits that window succeeds; with other
 * penalties the capacity is wf_length itself.  The reference writes past its
 * ring instead when a range exceeds wf_length (wfa_gpu.cu:217,618-626). */
int tsu_align_plain(tsu_ctx* ctx, tsu_batch* b, int32_t wf_length, int32_t* scores, tsu_timing* t);

/* Reference-shaped convenience: num_couples pairs of seq_len bases in the
 * reference's fixed-slot layout (patterns + i*seq_len, texts + i*seq_len),
 * exactly what wfa_gpu.cu:76-128 hands to the GPU.  Packs, uploads, aligns
 * (plain form), copies scores back. */
int tsu_align_fixed(tsu_ctx* ctx, const char* patterns, const char* texts, int32_t num_couples,
                    int32_t seq_len, int32_t wf_length, int32_t* scores, tsu_timing* t);

/* Batched form: persistent kernels fed from a device work queue, pairs
 * ordered by length into buckets, per-bucket shape (waves per pair, ring
 * placement) from the autotuner; diagonal capacity chosen per bucket and
 * widened automatically for pairs that overflow.  No reference counterpart:
 * the reference has only the plain launch (wfa_gpu.cu:113-120). */
typedef struct {
    int32_t waves_per_pair;   /* 0 = autotuned */
    int32_t capacity;         /* diagonals, 0 = automatic */
    int32_t bucket_bound;     /* max pattern length of a bucket, 0 = all */
    int32_t reserved;
} tsu_tune_entry;
int tsu_align_batched(tsu_ctx* ctx, tsu_batch* b, int32_t* scores, tsu_timing* t);
/* The batched form on the caller's device arrays with no device allocation.
 * No counterpart in the reference: it is for callers that cannot allocate
 * device memory during a call (serving loops with a fixed footprint,
 * harnesses that forbid hipMalloc).  Inputs: d_words (n_words packed words)
 * and d_pairs (n_pairs descriptors) on the context's device, and the same
 * descriptors on the host (pairs), which bucket the work.  Outputs: d_scores
 * (device, n_pairs slots) and scores (host, n_pairs slots, the same values).
 * Scores equal tsu_align_batched's.  The work counters and the queue order
 * live in one device array per GPU, allocated with the library's code and
 * held until the process ends (TSU_NOALLOC_MAX_PAIRS * 4 + 256 bytes):
 * more than TSU_NOALLOC_MAX_PAIRS pairs are refused with TSU_EINVAL.  Every
 * call on a GPU shares that array, so calls that overlap on one GPU (from
 * several threads or contexts) race and must be serialised by the caller.
 * No hipMalloc or hipFree is made; every launch keeps its ring in LDS, so a
 * pair whose capacity would need a ring in device memory (wider than
 * 16,384 diagonals with 4/6/2, or than LDS holds with other penalties) gets
 * TSU_EKOVERFLOW, as does the call (tsu_align_batched aligns such pairs).
 * The call is synchronous. */
#define TSU_NOALLOC_MAX_PAIRS 1048576
int tsu_align_batched_noalloc(tsu_ctx* ctx, const uint32_t* d_words, int64_t n_words, const tsu_pair* d_pairs,
                              const tsu_pair* pairs, int32_t n_pairs, int32_t* d_scores, int32_t* scores,
                              tsu_timing* t);
/* Runs the autotuner on a sample of b and stores the per-bucket table in the
 * context (also used by later tsu_align_batched calls).  Optional. */
int tsu_autotune(tsu_ctx* ctx, tsu_batch* b, int32_t sample_pairs);
int tsu_tune_table(const tsu_ctx* ctx, tsu_tune_entry* entries, int32_t cap, int32_t* n);
/* Replaces the context's per-bucket table: a table saved from
 * tsu_tune_table, or a first window pinned for a bucket (capacity > 0 runs
 * the bucket's pairs in that window first, also in the CIGAR form, instead of
 * placing each pair by the probe).  Entries name buckets by bucket_bound
 * (one of 128, 256, 512, 1024, 2048, 4096, 8192, 16383, 32000, 67108864),
 * each at most once; waves_per_pair 0, 1, 2, 4, 8 or 16 (above 1: the LDS
 * kernel with that many waves per pair, capacity 0).  n = 0 clears the
 * table.  TSU_EINVAL on any other entry, leaving the table as it was.  No
 * reference counterpart (the reference has no tuning). */
int tsu_set_tune_table(tsu_ctx* ctx, const tsu_tune_entry* entries, int32_t n);

/* Scores and CIGARs.  Backtrace follows WFA2-lib's piggy-backed priority
 * (common.h:67-77).  Pairs whose length bucket runs on a register kernel are
 * recorded in one pass, with state sized for the highest score the kernel's
 * window holds; pairs wider than the widest register window are recorded in
 * one launch over their whole diagonal range when that fits the budget; the
 * others get their scores first (the batched form), which size the state
 * exactly.  With 2-piece penalties every pair gets its score first, then
 * records on the 2-piece kernel.  Each CIGAR is a run list, run = (length << 2) | op with
 * op 0 'M' match, 1 'X' mismatch, 2 'I' insertion (consumes text), 3 'D'
 * deletion (consumes pattern). */
int tsu_align_cigar(tsu_ctx* ctx, tsu_batch* b, int32_t* scores, tsu_cigars** out, tsu_timing* t);
int tsu_cigars_get(const tsu_cigars* c, int64_t i, const uint32_t** runs, int32_t* n_runs);
/* Writes "3M1X4M..." (NUL-terminated); returns chars needed (excluding NUL)
 * or a negative status. */
int tsu_cigars_string(const tsu_cigars* c, int64_t i, char* buf, int32_t cap);
void tsu_cigars_free(tsu_cigars* c);

/* Backtrace memory of tsu_align_cigar: low-memory backtrace for long reads
 * (the reference reserves the backtrace types and packed-CIGAR fields,
 * common.h:10,17-23,48-49, but computes no CIGAR).  Per-cell records for
 * every score take O(score * W) words per pair; ring checkpoints every K
 * scores take O((score / K + K) * W) and recompute each window of K scores
 * once, giving the same CIGARs.
 *   mode             TSU_CIGAR_AUTO: records when a launch can hold them for
 *                    at least one pair per CU within the budget, else
 *                    checkpoints; TSU_CIGAR_RECORDS / TSU_CIGAR_CHECKPOINTS
 *                    force one
 *   checkpoint_every K, 0 = about sqrt(score * ring slots) per launch
 *   budget_bytes     device memory one launch may hold for backtrace state,
 *                    0 = half the free device memory
 * The backtrace state stays allocated in the context between calls (freeing
 * tens of GB of device memory can take a second) and grows when a launch
 * needs more; this call and tsu_destroy release it. */
#define TSU_CIGAR_AUTO         0
#define TSU_CIGAR_RECORDS      1
#define TSU_CIGAR_CHECKPOINTS  2
int tsu_set_cigar_memory(tsu_ctx* ctx, int32_t mode, int32_t checkpoint_every, int64_t budget_bytes);
/* Waves of one pair that synchronise with their neighbours (score-only and
 * CIGAR launches of 1,025-16,384 diagonals) poll a neighbour's progress word at
 * most this many times per score before the pair fails with TSU_EINTERNAL
 * instead of waiting for ever; 0 restores the default (2^20 polls, far above
 * any legitimate wait).  No reference counterpart (the reference's kernel
 * has one warp per pair); for callers that want a tighter bound and for
 * tests of that failure path. */
int tsu_set_poll_limit(tsu_ctx* ctx, int32_t polls);

/* What the call that made c used: the largest backtrace state of one launch
 * (device bytes) and the launches of each kind.  Any pointer may be NULL. */
int tsu_cigars_stats(const tsu_cigars* c, int64_t* state_bytes, int32_t* record_launches,
                     int32_t* checkpoint_launches);
/* Which kernels recorded the backtraces of the call that made c, as a mask
 * of TSU_KERNEL_* bits.  No reference counterpart (the reference computes no
 * CIGAR on the GPU); for callers and tests that check which path ran. */
#define TSU_KERNEL_REG1    0x01u  /* one wave per pair, wavefronts in registers */
#define TSU_KERNEL_ML4     0x02u  /* 16 chunks per wave, 4 waves per pair */
#define TSU_KERNEL_ML8     0x04u  /* 16 chunks per wave, 8 waves per pair */
#define TSU_KERNEL_ML16    0x08u  /* 16 chunks per wave, 16 waves per pair */
#define TSU_KERNEL_LDS     0x10u  /* rings in LDS (or HBM), records for every score */
#define TSU_KERNEL_LDS_CK  0x20u  /* rings in LDS (or HBM), ring checkpoints */
#define TSU_KERNEL_2P      0x40u  /* gap-affine 2-piece kernel, records for every score */
#define TSU_KERNEL_2P_REG  0x80u  /* gap-affine 2-piece single-wave register kernel ((e1, e2) small) */
int tsu_cigars_kernels(const tsu_cigars* c, uint32_t* kernels);

/* ---- serving (no reference counterpart) ---------------------------------- */

typedef struct {
    int64_t hbm_budget_bytes;  /* device memory the server may hold for
                                  resident requests (sequences + state) */
    int32_t max_inflight;      /* launches in flight at once (>= 1) */
    int32_t max_batch_pairs;   /* pairs merged into one launch (>= 1) */
} tsu_serve_cfg;

typedef struct {
    int64_t submit_ns;  /* CLOCK_MONOTONIC at tsu_serve_submit */
    int64_t admit_ns;   /* when its pairs were launched */
    int64_t done_ns;    /* when its scores were on the host */
    int32_t status;
    int32_t n_pairs;
} tsu_req_stats;

int tsu_serve_open(tsu_ctx* ctx, const tsu_serve_cfg* cfg, tsu_server** out);
/* Copies the request (n pairs given as char strings) and queues it.  Fails
 * with TSU_EBUDGET if the request alone exceeds the budget. */
int tsu_serve_submit(tsu_server* s, const char* const* patterns, const int32_t* p_len,
                     const char* const* texts, const int32_t* t_len, int32_t n, uint64_t* ticket);
/* The same for pairs already packed (tsu_pack layout): pairs[i] index
 * words[0, n_words), which are copied.  No reference counterpart (the
 * reference has no server); for clients that keep sequences packed. */
int tsu_serve_submit_packed(tsu_server* s, const uint32_t* words, int64_t n_words, const tsu_pair* pairs,
                            int32_t n, uint64_t* ticket);
/* Non-blocking: TSU_EAGAIN until done; then copies n scores and forgets the
 * ticket. */
int tsu_serve_poll(tsu_server* s, uint64_t ticket, int32_t* scores, tsu_req_stats* st);
int tsu_serve_wait(tsu_server* s, uint64_t ticket, int32_t* scores, tsu_req_stats* st);
int tsu_serve_close(tsu_server* s);

/* ---- measurement helpers ------------------------------------------------- */

/* Copies bytes of device memory with a plain streaming kernel and returns its
 * HIP-event time: the calibration kernel for HBM byte counters. */
int tsu_calibrate_copy(tsu_ctx* ctx, int64_t bytes, double* kernel_ms);
/* Synchronises the context's stream. */
int tsu_sync(tsu_ctx* ctx);

#ifdef __cplusplus
}
#endif
#endif /* TSUNAMI_WFA_H */
