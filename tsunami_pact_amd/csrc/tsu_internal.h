// tsu_internal.h -- structures shared by the host side of libtsunami_hip.so
// and the launchers of the HIP kernels.
#pragma once

#include <stdint.h>
#include <string>
#include <vector>

#include "../../include/tsunami_wfa.h"

struct tsu_dataset {
    std::vector<char> chars;
    std::vector<int64_t> p_off, t_off;
    std::vector<int32_t> p_len, t_len;
    int32_t hdr_num = 0, hdr_len = 0, hdr_err = 0;
};

int tsu_pack_seq(const char* seq, int32_t len, uint32_t* out);
bool tsu_is_acgt(const char* seq, int64_t len, int64_t* bad_pos);
int tsu_pack_dataset(const tsu_dataset* ds, int64_t first, int64_t count, std::vector<uint32_t>& words,
                     std::vector<tsu_pair>& pairs);

namespace tsu {

// Kernel-side view of a pair (same layout as tsu_pair).
struct KPair {
    int64_t p_off;
    int64_t t_off;
    int32_t p_len;
    int32_t t_len;
};
static_assert(sizeof(KPair) == sizeof(tsu_pair), "KPair must mirror tsu_pair");

// Arguments of every alignment kernel.
struct KArgs {
    const uint32_t* words;     // packed arena (HBM)
    const KPair* pairs;        // descriptors (HBM)
    const int32_t* order;      // queue entry -> pair index (nullable: identity)
    int32_t n;                 // queue entries / pairs
    int32_t* scores;           // per pair: score (unscaled) or negative status
    int32_t x, o, e, g;        // penalties divided by g = gcd(x, o, e)
    int32_t W;                 // diagonal capacity of a ring slot
    int32_t seq_words;         // LDS words reserved per sequence
    int32_t rm, ri;            // ring depths: M keeps max(x, o+e)+1 scores, I/D keep e+1
    int32_t o2, e2, r2;        // gap-affine 2-piece: second gap piece (divided by g), I2/D2 ring depth e2+1
    int32_t* queue;            // work counter (persistent launches)
    // ring in global memory (when it does not fit in LDS): one region per workgroup
    uint8_t* ring_g;
    int64_t ring_stride;       // bytes per workgroup
    // backtrace (CIGAR) state: one region per workgroup
    uint32_t* rec_g;           // per-cell records
    int64_t rec_cap;           // records per workgroup
    int32_t* step_lo_g;        // per score step: lowest diagonal recorded
    int64_t* step_off_g;       // per score step: offset of its first record
    int32_t step_cap;          // score steps per workgroup
    uint32_t* runs;            // CIGAR run pool
    int64_t runs_cap;
    unsigned long long* runs_used;
    int64_t* run_off;          // per pair
    int32_t* run_n;            // per pair
    // checkpointed backtrace: ring snapshots every ck_every score steps, one
    // region of ck_cap snapshots (ck_bytes each) per workgroup; the records
    // above (rec_cap, step_cap) then hold one window of ck_every steps
    uint8_t* ck_g;
    int64_t ck_stride;         // bytes per workgroup
    int64_t ck_bytes;          // bytes per snapshot: slot limits (ck_head bytes), then W cells per slot
    int32_t ck_head;
    int32_t ck_every;
    int32_t ck_cap;
    // neighbour-synchronised kernels: polls before a pair fails (0 = default)
    int32_t poll_limit;
    // probe launches (single-wave kernel): stop after probe_cap scaled
    // scores; a pair not finished by then gets kProbeStopped and, in
    // probe_out, the largest h + v of its last scores
    int32_t probe_cap;
    int32_t* probe_out;
};

// Score slot of a pair a probe launch stopped before its end (internal: the
// host reads it back and aligns the pair again; never returned to callers).
constexpr int32_t kProbeStopped = -1000;

struct Shape {
    int nw;          // waves per pair (1, 2, 4, 8, 16)
    bool off16;      // int16 offsets in the ring
    bool lds_ring;   // ring in LDS (else global memory)
    bool bt;         // record backtrace + emit CIGAR
    bool queue;      // persistent launch fed by a work counter
    bool fixed_242 = false;  // scaled penalties are (2, 3, 1): compile-time ring
    int reg_c = 0;           // > 0: register-resident kernel, reg_c chunks of 64 diagonals per wave (nw waves)
    bool lring = false;      // one wave: the older M scores in an int16 LDS ring (sequences up to 16000 bases)
    bool two_piece = false;  // gap-affine 2-piece kernel (one wave per pair, rings in LDS)
    bool ck = false;         // backtrace from ring checkpoints (bt and queue set, LDS kernel)
    bool dbl = false;        // 16-chunk kernel with doubled offsets (sequences up to kMlDblMaxLen bases)
    bool probe = false;      // single-wave register kernel stopping at KArgs::probe_cap
    bool gseq = false;       // LDS kernel reading the sequences from HBM (too long for LDS; int32 offsets)
    int reg_e = 0;           // 2-piece register kernel: 10 e1 + e2 after the gcd (11, 21 or 31)
};

// Doubled-offset 16-chunk kernel: longest sequence (2t fits int16) and the
// highest scaled score it runs to (-32768 + 2s, the ring's NULL after s
// scores, stays negative).
constexpr int32_t kMlDblMaxLen = 16383;
constexpr int32_t kMlDblMaxScore = 16383;

// Device scratch of tsu_align_batched_noalloc (a file-scope array per GPU):
// work counters, then one queue-order entry per pair.
constexpr int32_t kNoallocMaxPairs = TSU_NOALLOC_MAX_PAIRS;
constexpr size_t kNoallocCounterBytes = 256;
constexpr size_t kNoallocScratchBytes = kNoallocCounterBytes + 4 * (size_t)kNoallocMaxPairs;
// Device address of that array on `device` (current); hipError_t value.
int scratch_address(int device, void** out);

// LDS bytes a launch of this shape needs.
int64_t lds_bytes(const Shape& sh, const KArgs& a);
// Bytes of one ring (all slots) for W diagonals.
int64_t ring_bytes(const Shape& sh, int32_t W, int32_t rm, int32_t ri);
int32_t ring_stride_entries(const Shape& sh, int32_t W);
// Launches the kernel of the given shape on stream with grid blocks.
// Returns a hipError_t value (0 on success).
int launch_wfa(const Shape& sh, const KArgs& a, int64_t grid, void* stream);
// Streaming copy used to calibrate HBM byte counters.
int launch_copy(const void* src, void* dst, int64_t bytes, void* stream);
// Max dynamic LDS bytes the kernel of this shape may use (after raising the
// attribute); 0 on error.
int64_t max_lds_for(const Shape& sh);
// Resident workgroups of this shape per CU for a given LDS size.
int occupancy(const Shape& sh, int64_t lds, int* blocks_per_cu);

}  // namespace tsu
