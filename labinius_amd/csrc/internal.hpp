// Internal declarations shared by the host translation units.
#pragma once

#include <hip/hip_runtime.h>

#include <array>
#include <atomic>
#include <cmath>
#include <cstdint>
#include <string>
#include <thread>
#include <vector>

#include "../../include/labinius.h"
#include "../../include/labinius_debug.h"
#include "blake3.hpp"
#include "device.hpp"
#include "ring.hpp"
#include "threads.hpp"

namespace lbn {

void set_error(const std::string &msg);
int fail(int code, const std::string &msg);

#define HIP_TRY(x)                                                                        \
    do {                                                                                  \
        hipError_t e_ = (x);                                                              \
        if (e_ != hipSuccess)                                                             \
            return fail(e_ == hipErrorOutOfMemory ? LBN_ERR_OOM : LBN_ERR_HIP,            \
                        std::string(#x) + ": " + hipGetErrorString(e_));                  \
    } while (0)

inline size_t round_up(size_t x, size_t m) { return (x + m - 1) / m * m; }



template <typename T>
int dalloc(T **p, size_t count) {
    HIP_TRY(hipMalloc((void **)p, (count > 0 ? count : 1) * sizeof(T)));
    return LBN_OK;
}

// Waits for a stream by polling it: a round's calls wait for their own short work, and a polling
// wait resumes the caller as soon as the stream drains, where a blocking one may first yield the
// thread.
inline hipError_t stream_wait(hipStream_t s) {
    hipError_t e;
    while ((e = hipStreamQuery(s)) == hipErrorNotReady) {
    }
    return e;
}

inline float elapsed(hipEvent_t a, hipEvent_t b) {
    float ms = 0;
    if (hipEventElapsedTime(&ms, a, b) != hipSuccess) return 0;
    return ms;
}

// Params as the host uses them (scheme/params.rs:62-68, 254-287).
struct Shape {
    uint32_t wl = 0, cl = 0, rows = 0;
    size_t n = 0, r = 0, column_len = 0, len_ring = 0;
    std::vector<uint32_t> primes;  // base first
    uint32_t dropped_bits = 0;
    uint32_t opening = 0;
};
int make_shape(const lbn_params *p, Shape *s);

// xorshift64* (rng.rs:1-20)
struct Rng {
    uint64_t s;
    explicit Rng(uint64_t seed) : s(seed * 0x9E3779B97F4A7C15ull | 1) {}
    uint64_t next() {
        uint64_t x = s;
        x ^= x >> 12;
        x ^= x << 25;
        x ^= x >> 27;
        s = x;
        return x * 0x2545F4914F6CDD1Dull;
    }
    uint32_t below(uint32_t n) { return (uint32_t)(((next() >> 32) * (uint64_t)n) >> 32); }
};

// CommitmentKey::random (key.rs:48-88): a[k][i][j], i = 32 b + p drawn b, j, p.
std::vector<int16_t> expand_key(const std::vector<uint32_t> &primes, size_t len_ring,
                                const uint8_t seed[32]);

// The transcript (challenge.rs:66-160).
struct Transcript {
    Blake3 state = Blake3::hasher();
    uint64_t counter = 0;
    void absorb_bytes(const void *b, size_t n) {
        state.update_u64(n);
        state.update(b, n);
    }
    void absorb_u64(uint64_t x) { state.update_u64(x); }
    Blake3::Output reader(const char *label, size_t label_len);
    void fill(const char *label, size_t label_len, uint8_t *out, size_t n) {
        reader(label, label_len).read(0, out, n);
    }
};

uint64_t sample_short_challenge(Transcript &t, uint32_t weight, double bound, lbn_challenge *out);
// n challenges in a row (FoldingChallenges::derive), the same as n calls of the above
void sample_short_challenges(Transcript &t, size_t n, uint32_t weight, double bound, lbn_challenge *out);

}  // namespace lbn

struct lbn_pp;
struct lbn_pool;
namespace lbn {
// the device pool (pool.cpp): a buffer of at least `bytes` (same_tag: it last held a layout with
// this tag), and its return
int pool_alloc(lbn_pool *pool, uint64_t bytes, uint64_t tag, void **out, bool *same_tag);
void pool_free(lbn_pool *pool, void *p);
uint64_t pool_rounded(uint64_t bytes);
void pool_set_cap(lbn_pool *pool, uint64_t cap);
// verifier.rs:41-57
void absorb_parameters(Transcript &t, const lbn_pp *pp);
// verifier.rs:85-99
void derive_point(Transcript &t, const Shape &s, lbn_f162 *p0, lbn_f162 *p1);
}  // namespace lbn

struct lbn_transcript {
    lbn::Transcript t;
};

// PublicParameters plus everything the kernels read, resident on one device.
struct lbn_pp {
    std::atomic<int> refs{1};
    lbn_params params{};
    lbn::Shape shape;
    uint8_t seed[32];
    std::vector<int16_t> key;  // [limb][len_ring][648], centered
    int device = 0;
    bool host_only = false;  // the key and shape only (a verifier on a host without a GPU)
    // K index lengths: ring elements and columns padded to a step (the A and challenge fragments)
    // and to a power of two (the GEMMs' K order, device.hpp); the GEMMs' steps gemm_kbs(kpad)
    uint32_t lpad = 0, rpad = 0, kpad_c = 0, kpad_f = 0, kb_c = 0, kb_f = 0;
    // commit GEMM: A' resident, ntiles x kb_c steps
    uint32_t ntiles = 0;
    int8_t *d_afrag = nullptr;
    uint16_t *d_lut = nullptr, *d_ee = nullptr;
    int16_t *d_out = nullptr;
    lbn::TileDesc *d_tiles = nullptr;
    lbn::FinishLimb *d_finish = nullptr;
    uint16_t *d_slot648 = nullptr, *d_qcs = nullptr;
    // fold GEMM (base limb)
    uint32_t ntiles_f = 0;
    uint16_t *d_lut_f = nullptr, *d_ee_f = nullptr, *d_expo_f = nullptr, *d_pw_f = nullptr;
    int16_t *d_out_f = nullptr;
    lbn::TileDesc *d_tiles_f = nullptr;
    uint32_t cond_f = 0, mult_f = 0;
    lbn::InttConsts intt{};
    uint32_t *d_twi = nullptr;
    // every limb's inverse transform (the bit-dropped digits)
    std::vector<lbn::InttConsts> intt_limb;
    std::vector<uint32_t *> d_twi_limb;
    std::vector<void *> allocations;
};

struct lbn_dwitness {
    lbn_prover *owner = nullptr;
    lbn_pool *pool = nullptr;  // d_w came from this pool (else from hipMalloc)
    uint64_t *d_w = nullptr;
    size_t n = 0;
    size_t cols = 0;  // its own Params::columns(); the key (row length, moduli) is the prover's
};

// The batched commit's workspace (grow-only), its captured graph and the autotuned schedules.
struct lbn_batch_ws {
    size_t cap_cols = 0;
    uint8_t *d_bytes = nullptr;
    uint32_t *d_yacc = nullptr, *d_raw = nullptr;
    int16_t *d_comp = nullptr;
    const uint64_t **d_cols = nullptr;
    const uint64_t **h_cols = nullptr;  // pinned
    hipGraphExec_t graph = nullptr;
    hipGraph_t graph_def = nullptr;
    size_t graph_cols = 0, last_cols = 0;
    uint32_t graph_t = 0, graph_min_splits = 0, graph_splits = 1;  // the graph's schedule: tiles a round, K ranges asked for and run
    std::vector<std::array<uint32_t, 3>> tuned;  // (width class, sched_t, K ranges asked for)
    hipEvent_t ev[2] = {};
    float ms = 0;
    uint32_t sched_t = 8, splits = 1;  // the last batch's schedule and K ranges
};

// The cross-field switch's workspace (grow-only): eq(r_hi), the batched eq image a, the folded
// trace, and the small tables and outputs.
struct lbn_switch_ws {
    size_t cap_n = 0;
    uint64_t *d_eq = nullptr, *d_a = nullptr, *d_p = nullptr, *d_small = nullptr;
    uint64_t *h_small = nullptr;  // pinned staging of the host's side of d_small (the round trips)
};

// The wire encoder's device scratch (grow-only, wire.cpp).
struct lbn_wire_ws {
    int16_t *d_in = nullptr;  // a host message uploaded for encoding
    size_t cap_in = 0;        // coefficients
    uint32_t *d_staged = nullptr;  // the encoder's words, a u32 per coefficient ([group][lane])
    uint16_t *d_escw = nullptr;    // escapes' raw-bits words, [group][lane]
    uint32_t *d_cnt = nullptr;     // words pushed inside each group's block of 64, then block totals
    uint64_t *d_offs = nullptr;    // words pushed before each block
    size_t cap_groups = 0;
    uint16_t *d_stream = nullptr;
    size_t cap_words = 0;
    uint32_t *d_pack = nullptr;  // a bit-packed object
    size_t cap_pack = 0;         // u32 words
    uint32_t *d_stat = nullptr;  // RANS_STAT words
    unsigned long long *d_hist = nullptr;  // 65537 counts
    lbn::RansSym *d_sym = nullptr;         // 65537 symbols
    hipEvent_t ev[5] = {};
    float ms[3] = {};  // the last encode: total, statistics, the rANS kernel
};

// The device verifier's scratch (verify.cpp), allocated at its first use.
struct lbn_verify_ws {
    bool ready = false;
    int16_t *d_v = nullptr;        // the folded witness
    uint64_t *d_w = nullptr;       // its bit planes as F162 columns
    uint8_t *d_bytes = nullptr;    // their commit layout
    uint32_t *d_yacc = nullptr, *d_raw = nullptr;
    int16_t *d_comp = nullptr;     // A times each plane, decomposed
    uint64_t *d_u = nullptr;       // the planes' row values
    uint64_t *d_stat = nullptr;    // [0] squared norm, [1] largest |v|
    int16_t *d_fcomp = nullptr;    // a commitment to fold
    int16_t *d_fout = nullptr;     // the folded commitment
    int64_t *d_facc = nullptr;
    uint32_t *d_pw = nullptr;      // every limb's root powers
    uint16_t *d_p3 = nullptr;      // POW3_SLOT_EXP
    lbn_challenge *d_ch = nullptr;
    size_t cap_v = 0, cap_fcomp = 0, cap_ch = 0;
};

struct lbn_opening {
    lbn_prover *owner = nullptr;
    uint8_t *d_bytes_f = nullptr;  // the witness in the fold's byte layout
};

namespace lbn {
void batch_free(lbn_batch_ws *b);
// bd/digits.rs:121-182 on the device into pr->d_bd_out (dropped.cpp)
int bd_digits_device(lbn_prover *pr, const Shape &s, const uint32_t *d_raw);
void wire_free(lbn_wire_ws &w);
// The commit GEMM's arguments over ncol columns of a layout into y ([ncol][limb][plane][648]);
// sched_t 8 and no minimum of K ranges (prover.cpp).
GemmArgs commit_gemm_args(const lbn_pp *pp, const uint8_t *bytes, uint32_t *y, size_t ncol);
// Prepass, GEMM and finish of ncol columns (column_len F162 each, at d_w) under the prover's key,
// into d_comp ([c][4][limb][162]) and d_raw ([limb][c][648]) on the prover's stream (prover.cpp).
// d_bytes holds bytes_size(ncol, kpad_c) bytes whose padding reads as zero; d_yacc ncol x
// YACC_PLANES x limbs x 648 words.
int commit_columns_device(lbn_prover *pr, const uint64_t *d_w, size_t ncol, uint8_t *d_bytes, uint32_t *d_yacc,
                          uint32_t *d_raw, int16_t *d_comp);
// the batched commitment's workspace sized for cols columns in advance (batch.cpp)
int commit_batch_reserve(lbn_prover *pr, size_t cols);
// a resident witness of n F162 for the prover (from its pool when it has one), uploaded from w
int witness_alloc_upload(lbn_prover *pr, const lbn_f162 *w, size_t n, size_t cols, lbn_dwitness **out);
void verify_free(lbn_verify_ws &w);
}

struct lbn_prover {
    lbn_pp *pp = nullptr;
    hipStream_t stream = nullptr;
    lbn_pool *pool = nullptr;  // where uploads and fold layouts come from (NULL: hipMalloc / hipFree)
    // workspace
    uint8_t *d_bytes_c = nullptr;
    std::vector<uint8_t *> free_bytes_f;  // recycled fold layouts
    uint32_t *d_yacc = nullptr, *d_raw = nullptr;
    int16_t *d_comp = nullptr;
    lbn_challenge *d_ch = nullptr;
    int8_t *d_chat = nullptr;     // the challenges' transform, ntiles_f x rpad / 64 steps
    int8_t *d_afrag_f = nullptr;  // the fold's A', ntiles_f x kb_f steps
    uint32_t *d_vhat = nullptr;
    int16_t *d_v = nullptr;
    uint64_t *d_p0 = nullptr, *d_eq = nullptr, *d_eqpart = nullptr, *d_u = nullptr;
    uint32_t *d_eqf = nullptr;  // the row evaluation's eq operands (row_eval_scratch_u32)
    uint64_t *d_p1 = nullptr, *d_eqpart1 = nullptr, *d_t = nullptr;  // the MLE claim (d_p1 inside d_p0's block)
    uint64_t *h_pts = nullptr;       // pinned staging of p0 and p1 (one upload)
    lbn_challenge *h_ch = nullptr;   // pinned staging of the folding challenges
    uint64_t *d_wscratch = nullptr;  // host-witness uploads
    // uploads from pageable host memory, one share a thread: its stream, two pinned chunk buffers
    // and the events of their copies (made at the first large upload)
    struct Upload {
        hipStream_t st = nullptr;
        uint8_t *buf[2] = {};
        hipEvent_t ev[2] = {};
    };
    std::vector<Upload> up;
    uint8_t *h_out = nullptr;  // pinned staging of the outputs copied back to host memory (grown on demand)
    size_t h_out_cap = 0;
    // the bit-dropped digits (grow-only): raw rows from the host, centered coefficients, top + digits
    uint32_t *d_bd_raw = nullptr;
    int16_t *d_bd_cen = nullptr;
    uint16_t *d_bd_out = nullptr;
    size_t bd_cap = 0;       // coefficients x limbs
    bool committed = false;  // d_raw holds the raw rows of a commitment
    bool bd_ready = false;   // d_bd_out holds the top and digits of that commitment
    bool folded = false;     // d_v holds the folded witness of a fold
    hipEvent_t ev[12] = {};
    float timings[11] = {};
    lbn_batch_ws *batch = nullptr;
    lbn_switch_ws sw;
    lbn_wire_ws wire;
    lbn_verify_ws vfy;
};
