/*
 * labinius on MI355X: the C ABI of the commit / evaluate / fold prover.
 *
 * Each entry point replaces one entry point of the reference's Rust API (paths relative to the
 * reference checkout, crates/pcs/src/...). Signatures use plain pointers and sizes only, so that
 * a Rust `extern "C"` block (or ctypes, or cgo) binds them unchanged.
 *
 * Conventions
 *   - Every function returns an int status (LBN_OK = 0) unless noted; no exception or panic
 *     crosses the ABI. Where the reference panics (wrong witness length, a point of the wrong
 *     shape, a mismatched opening), these return LBN_ERR_SHAPE / LBN_ERR_WITNESS_LENGTH.
 *     lbn_last_error() gives a thread-local message for the last failure.
 *   - An lbn_pp is immutable after creation and may be shared by provers; an lbn_prover is
 *     single-threaded (it owns one HIP stream), as the Rust Prover takes &mut self.
 *   - An lbn_opening is move-only: lbn_fold and lbn_recycle consume it.
 *   - Layouts
 *       F162            3 little-endian u64, bits 0..161, the top 30 bits zero (fields/scalar.rs:10)
 *       components      [column c][row t][limb k][slot s] int16, centered: entry (t, c) of the
 *                       4 x r VerticallyAlignedMatrix, one R_162 element per limb in primes()
 *                       order (key.rs:226-233, ring/element.rs:74-117)
 *       raw             [limb k][column c][648] u32 in [0, q_k) (key.rs:279-280)
 *       key             [limb k][ring element i][slot j] int16, centered (key.rs:135-141)
 *       folded witness  [ring element i][coefficient 648] int16, centered mod the base prime
 *                       (commitment.rs:253-279)
 *       folded comm.    [row t][limb k][slot s] int16, centered (commitment.rs:283-286)
 */
#ifndef LABINIUS_H
#define LABINIUS_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

enum {
    LBN_OK = 0,
    /* ParamError, scheme/params.rs:22-33 */
    LBN_ERR_COLUMNS_EXCEED_WITNESS = 1,
    LBN_ERR_TOO_FEW_COLUMNS = 2,
    LBN_ERR_COLUMN_TOO_SHORT = 3,
    LBN_ERR_DUPLICATE_MODULUS = 4,
    LBN_ERR_BASE_IS_ALSO_EXTRA = 5,
    LBN_ERR_UNKNOWN_MODULUS = 6,     /* not one of ring/mod.rs:87-95 */
    /* WitnessError::WrongLength, scheme/witness.rs:12-15 */
    LBN_ERR_WITNESS_LENGTH = 7,
    /* a reference assert!: mismatched shapes, points, openings */
    LBN_ERR_SHAPE = 8,
    LBN_ERR_HIP = 9,
    LBN_ERR_OOM = 10,
    LBN_ERR_UNSUPPORTED = 11,        /* e.g. Opening::Recursive (LaBRADOR stays out of scope) */
    LBN_ERR_ARG = 12,
    /* VerificationError::Rejected, scheme/commitment.rs:449-451 */
    LBN_REJECTED = 13,
    LBN_ERR_NO_DEVICE = 14,
    /* WireError, wire/mod.rs:67-74 */
    LBN_WIRE_TRUNCATED = 15,         /* the byte string ran out before the object did */
    LBN_WIRE_MALFORMED = 16,         /* not a coding of any object of the expected shape */
    /* the serving loop: a request that alone needs more device memory than the server's budget */
    LBN_ERR_ADMISSION = 17
};

enum { LBN_OPENING_CLEAR = 0, LBN_OPENING_BIT_DROPPED = 1, LBN_OPENING_RECURSIVE = 2 };

typedef struct { uint64_t w[3]; } lbn_f162;

/* ShortChallenge, challenge.rs:231-235 */
typedef struct {
    uint8_t positions[32];
    uint32_t signs;
    uint32_t weight;
} lbn_challenge;

/* Params, scheme/params.rs:62-68 (moduli by prime: 2917 3889 4861 9721 12637 17497 19441) */
typedef struct {
    uint32_t witness_log_len;
    uint32_t column_log_len;
    uint32_t base;
    uint32_t n_extra;
    uint32_t extra[6];
    uint32_t opening;
    uint32_t dropped_bits;
} lbn_params;

typedef struct lbn_pp lbn_pp;
typedef struct lbn_prover lbn_prover;
typedef struct lbn_opening lbn_opening;
typedef struct lbn_dwitness lbn_dwitness;
typedef struct lbn_transcript lbn_transcript;

const char *lbn_last_error(void);

/* Params::with_base's checks, scheme/params.rs:172-203; returns a ParamError code or LBN_OK */
int lbn_params_check(const lbn_params *params);

/* PublicParameters::from_seed, scheme/params.rs:320-339: expands A (key.rs:48-88) and uploads it
 * to the current HIP device in the layout the commit kernel streams; refcounted. */
int lbn_pp_create(const lbn_params *params, const uint8_t matrix_seed[32], lbn_pp **out);
void lbn_pp_release(lbn_pp *pp);
/* The same public parameters without any device state: the key and the shape, for a verifier on a
 * host without a GPU (Verifier::new, scheme/verifier.rs:28-37, needs only these). A prover cannot
 * be created on it (LBN_ERR_NO_DEVICE). */
int lbn_pp_create_host(const lbn_params *params, const uint8_t matrix_seed[32], lbn_pp **out);
/* CommitmentKey::row for every limb, key.rs:139-141 ([limb][i][648] int16, centered) */
int lbn_pp_key(const lbn_pp *pp, int16_t *a, size_t len);

/* Witness::random, scheme/witness.rs:66-87 (host only) */
int lbn_witness_random(const lbn_params *params, const uint8_t seed[32], lbn_f162 *out, size_t n);

/* Prover::new, scheme/prover.rs:31-53: the device workspace, one stream, and a warm-up round over
 * a zero witness that runs every kernel once. */
int lbn_prover_create(lbn_pp *pp, lbn_prover **out);
void lbn_prover_release(lbn_prover *prover);

/* A witness resident in HBM (Witness::from_elements, scheme/witness.rs:54-62, uploaded once).
 * lbn_dwitness_device_ptr exposes the device copy (n x 24 bytes) for callers that fill it on the
 * device themselves. */
int lbn_witness_upload(lbn_prover *prover, const lbn_f162 *w, size_t n, lbn_dwitness **out);
void *lbn_dwitness_device_ptr(lbn_dwitness *w);
void lbn_dwitness_release(lbn_dwitness *w);

/* Prover::commit, scheme/prover.rs:55-91 (Clear mode; the commitment matrix and the raw rows).
 * The opening keeps the witness's bit layout the fold reads on the device. Either output pointer
 * may be NULL. lbn_commit takes a host witness (uploaded on every call); lbn_commit_resident a
 * resident one. */
int lbn_commit(lbn_prover *prover, const lbn_f162 *w, size_t n, int16_t *components,
               uint32_t *raw, lbn_opening **opening);
int lbn_commit_resident(lbn_prover *prover, const lbn_dwitness *w, int16_t *components,
                        uint32_t *raw, lbn_opening **opening);

/* Prover::fold, scheme/prover.rs:120-133: consumes the opening; v is len_ring x 648 int16. */
int lbn_fold(lbn_prover *prover, lbn_opening *opening, const lbn_challenge *challenges, size_t r,
             int16_t *v);
/* Prover::recycle, scheme/prover.rs:116-118 */
int lbn_recycle(lbn_prover *prover, lbn_opening *opening);
/* Prover::fold (scheme/prover.rs:120-131) of a witness that stayed resident in HBM since its
 * commitment, without an opening: the fold's layout is rebuilt from the witness (what a batched
 * commitment, which keeps none, needs). Same result as lbn_fold. */
int lbn_fold_resident(lbn_prover *prover, const lbn_dwitness *w, const lbn_challenge *challenges, size_t r,
                      int16_t *v);

/* Witness::row_evaluate / mle_evaluate, scheme/witness.rs:94-105: u has 2^column_log_len F162;
 * p0 has row_log_len coordinates, p1 column_log_len. */
int lbn_row_evaluate(lbn_prover *prover, const lbn_dwitness *w, const lbn_f162 *p0, size_t p0_len,
                     lbn_f162 *u);
/* eval.rs:185-193 (claim): t = sum_j eq(p1)_j u_j from the row values, on the host */
int lbn_claim(const lbn_f162 *u, size_t r, const lbn_f162 *p1, size_t p1_len, lbn_f162 *t);
int lbn_mle_evaluate(lbn_prover *prover, const lbn_dwitness *w, const lbn_f162 *p0, size_t p0_len,
                     const lbn_f162 *p1, size_t p1_len, lbn_f162 *t);

/* No counterpart in the reference: lbn_row_evaluate for callers that cannot allocate device memory
 * (code under HIP graph capture, tight serving loops, harnesses that forbid hipMalloc). It needs
 * no prover and allocates nothing: d_w (n F162), d_p0 (p0_len F162) and d_u (n / 2^p0_len F162,
 * overwritten) are device pointers, the work is queued on `stream` (a hipStream_t, NULL for the
 * default stream) and the call returns without waiting. The eq tensor and its operands live in a
 * static device array of 3.7 MB per device, so p0_len is at most 15 (LBN_ERR_SHAPE above); calls that overlap
 * on one device share that array and race. u equals lbn_row_evaluate's. */
int lbn_row_evaluate_noalloc(const lbn_f162 *d_w, size_t n, const lbn_f162 *d_p0, size_t p0_len, lbn_f162 *d_u,
                             void *stream);

/* A device memory pool (no counterpart in the reference; its Prover recycles one workspace,
 * key.rs:185-191, scheme/prover.rs:116-131). A prover given a pool takes its resident witnesses
 * (lbn_witness_upload, lbn_witness_upload_params) and its fold layouts from it and gives them back
 * on release, never calling hipFree while the pool lives; cap_bytes bounds what the pool obtains
 * from the device (LBN_ERR_OOM past it). Thread-safe; provers on several threads may share it. */
typedef struct lbn_pool lbn_pool;
int lbn_pool_create(uint64_t cap_bytes, lbn_pool **out);
void lbn_pool_release(lbn_pool *pool);  /* after every prover and witness that uses it */
int lbn_pool_stats(lbn_pool *pool, uint64_t *obtained, uint64_t *in_use, uint64_t *peak_in_use, uint64_t *mallocs,
                   uint64_t *reuses);
int lbn_prover_set_pool(lbn_prover *prover, lbn_pool *pool);

/* Device time of each stage of the prover's last commit / fold / row evaluation, in ms, measured
 * with HIP events on the prover's stream:
 *   [0] commit total  [1] commit layout pass  [2] commit GEMM  [3] commit finish
 *   [4] fold total    [5] challenge transform [6] fold GEMM    [7] inverse NTT
 *   [8] eval total    [9] both points' eq sub-tables
 *   [10] eq(p0) with its operands, the row values and the claim */
int lbn_prover_timings(const lbn_prover *prover, float *ms, size_t n);

/* Batched commitment (no counterpart in the reference, which commits one witness per call; each
 * result equals Prover::commit, scheme/prover.rs:55, of that witness alone). A witness of other
 * Params that share this prover's key (the same row_log_len and moduli, so the same A, key.rs:48)
 * is uploaded with lbn_witness_upload_params; lbn_commit_batch commits `count` of them in one
 * launch sequence (layout, GEMM, finish; captured once as a HIP graph and replayed when the same
 * total width comes twice in a row): components[i] ([cols_i][4][limb][162]) and raw[i]
 * ([limb][cols_i][648]) may be NULL. The GEMM's schedule is autotuned per power-of-two class of
 * the total column count. */
int lbn_witness_upload_params(lbn_prover *prover, const lbn_params *params, const lbn_f162 *w, size_t n,
                              lbn_dwitness **out);
int lbn_commit_batch(lbn_prover *prover, const lbn_dwitness *const *ws, size_t count, int16_t *const *components,
                     uint32_t *const *raw);
/* The last batch's schedule (tiles per XCD round; the GEMM's K ranges, whose results are added)
 * and its device ms. */
int lbn_batch_info(const lbn_prover *prover, uint32_t *sched_t, uint32_t *k_ranges, float *ms);

/* The serving loop (no counterpart in the reference, whose callers run one round at a time): a
 * server commits, evaluates and folds requests of many sizes that share one key (the same
 * row_log_len and moduli as `key`, so the same A), with continuous batching. A scheduler thread
 * admits queued requests in arrival order while their device memory fits the budget (the
 * witness and the fold's layout, lbn_server_request_bytes, from the server's device pool), uploads
 * them, and commits every request it admitted in one batched commitment; requests that arrive
 * while others are on the GPU join the next batch. `workers` threads, each with its own provers
 * and HIP streams, then run each request's round: the bit-dropped digits when the mode drops bits
 * (scheme/prover.rs:69-72), the evaluation point (verifier.rs:61-100), the row values and the
 * claim (witness.rs:94-105), the folding challenges (commitment.rs:217-228) and the fold of the
 * still-resident witness (prover.rs:120-131), overlapping on the device with each other and with
 * the next batch. Each result is the request's own round, the same as alone (its transcript
 * starts from its own domain). serial = 1 runs the same stream one request at a time, unbatched,
 * with nothing in flight beside it: the comparison. */
typedef struct lbn_server lbn_server;
typedef struct {
    uint64_t hbm_budget;        /* device bytes the server may hold: its own workspaces plus requests */
    uint32_t workers;           /* round threads (0: 4) */
    uint32_t serial;            /* 1: one request at a time, no batching, no overlap */
    uint32_t max_batch_columns; /* columns of one batched commitment (0: 16384) */
    uint32_t n_sizes;           /* witness_log_len of the sizes to set up in advance (others on first use) */
    uint32_t sizes[8];
} lbn_server_config;
/* One request's outputs, in the caller's buffers, sized by its Params (include/labinius.h layouts):
 * components (Clear) and raw always, top and digits when it drops bits; the point (p0: row_log_len,
 * p1: column_log_len F162), the row values u (columns F162), the claim (1 F162), the challenges
 * (columns) and the folded witness v (len_ring x 648). */
typedef struct {
    int16_t *components;
    uint32_t *raw;
    uint16_t *top, *digits;
    lbn_f162 *p0, *p1, *u, *claim;
    lbn_challenge *challenges;
    int16_t *v;
} lbn_server_outputs;
typedef struct {
    int status;                /* LBN_OK or the error the request met (it alone fails) */
    double arrival_s, admitted_s, committed_s, done_s;  /* seconds since the server started */
    uint64_t bytes;            /* the device memory it held */
    uint32_t batch;            /* which batched commitment took it */
} lbn_server_info;
typedef struct {
    uint64_t requests, failed, batches, batched_requests, batched_columns;
    uint64_t reserved_peak, fixed_bytes, pool_obtained, pool_mallocs, pool_reuses;
} lbn_server_stats;
enum { LBN_SERVER_FAIL_FOLD = 1 };  /* test hook: the request's fold is given an invalid challenge */
int lbn_server_create(const lbn_params *key, const uint8_t matrix_seed[32], const lbn_server_config *config,
                      lbn_server **out);
/* Queues a request; w (n F162) and the output buffers stay the caller's and must stay valid until
 * lbn_server_wait returns for the ticket. LBN_ERR_ADMISSION when the request alone exceeds the
 * budget, LBN_ERR_SHAPE when its Params do not share the server's key. */
int lbn_server_submit(lbn_server *server, const lbn_params *params, const lbn_f162 *w, size_t n,
                      const uint8_t *domain, size_t domain_len, const lbn_server_outputs *outputs, uint32_t flags,
                      uint64_t *ticket);
/* Blocks until the request is done; returns LBN_OK (the request's own status is in info). */
int lbn_server_wait(lbn_server *server, uint64_t ticket, lbn_server_info *info);
int lbn_server_stats_get(lbn_server *server, lbn_server_stats *stats);
/* A new budget for admissions from now on (requests already admitted keep their memory). */
int lbn_server_set_budget(lbn_server *server, uint64_t hbm_budget);
/* Finishes every queued request, then stops the threads and frees the device memory. */
void lbn_server_destroy(lbn_server *server);
/* Device bytes a request of these Params holds from admission to the end of its fold. */
uint64_t lbn_server_request_bytes(const lbn_params *params);

/* Transcript, challenge.rs:66-160 */
lbn_transcript *lbn_transcript_new(const uint8_t *domain, size_t len);
lbn_transcript *lbn_transcript_clone(const lbn_transcript *t);
void lbn_transcript_free(lbn_transcript *t);
void lbn_transcript_absorb_bytes(lbn_transcript *t, const uint8_t *b, size_t len);
void lbn_transcript_absorb_u64(lbn_transcript *t, uint64_t x);
void lbn_transcript_fill(lbn_transcript *t, const uint8_t *label, size_t label_len, uint8_t *out,
                         size_t len);
/* sample_short_challenge, challenge.rs:517-533; returns the number of attempts */
uint64_t lbn_sample_short_challenge(lbn_transcript *t, uint32_t weight, double bound,
                                    lbn_challenge *out);

/* The verifier (scheme/verifier.rs), on the host. */
/* verifier.rs:61-100 */
int lbn_derive_evaluation_point(const lbn_pp *pp, lbn_transcript *t, const int16_t *components,
                                lbn_f162 *p0, lbn_f162 *p1);
/* FoldingChallenges::derive, commitment.rs:217-228, over a RowEvaluation (commitment.rs:320-331) */
int lbn_derive_folding_challenges(const lbn_pp *pp, lbn_transcript *t, const lbn_f162 *u, size_t r,
                                  lbn_challenge *out);
/* verifier.rs:114-146 */
int lbn_fold_commitment(const lbn_pp *pp, const int16_t *components, const lbn_challenge *c,
                        size_t r, int16_t *out);
/* verifier.rs:149-155 */
int lbn_fold_row_evaluation(const lbn_f162 *u, const lbn_challenge *c, size_t r, lbn_f162 *out);
/* verifier.rs:158-174: LBN_OK or LBN_REJECTED */
int lbn_verify_evaluation(const lbn_pp *pp, const lbn_f162 *p1, size_t p1_len,
                          const lbn_f162 *claimed, const lbn_f162 *u, size_t r);
/* verifier.rs:178-231 (Clear mode): LBN_OK or LBN_REJECTED */
int lbn_verify_opening_clear(const lbn_pp *pp, const int16_t *folded_commitment, const int16_t *v,
                             size_t len_ring, const lbn_f162 *p0, size_t p0_len,
                             const lbn_f162 *folded_row_value);

/* Bit-dropped commitments (Opening::BitDropped, scheme/params.rs:73,231-237), on the host.
 * top is [c * 648 + i] u16, digits is [limb - 1][c * 648 + i] u16 (bd/mod.rs:7-14). */
/* bd/digits.rs:121-182 (drop_bits, called by Prover::commit, scheme/prover.rs:69-72): from the raw
 * rows ([limb][c][648] u32 as lbn_commit returns them) */
int lbn_drop_bits(const lbn_params *params, const uint32_t *raw, uint16_t *top, uint16_t *digits);
/* The same drop_bits (bd/digits.rs:121-182) on the device: every limb's inverse NTT, then the
 * Garner digits, on the prover's stream. raw is host memory as for lbn_drop_bits; params must
 * have the prover's moduli. The outputs equal lbn_drop_bits's. */
int lbn_drop_bits_device(lbn_prover *pr, const lbn_params *params, const uint32_t *raw, uint16_t *top,
                         uint16_t *digits);
/* Prover::commit in bit-dropped mode (scheme/prover.rs:69-72): drop_bits on the device over the raw
 * rows of the prover's last lbn_commit / lbn_commit_resident, which are still in HBM (the prover's
 * parameters drop bits; LBN_ERR_ARG before any commitment). */
int lbn_commit_drop_bits(lbn_prover *pr, uint16_t *top, uint16_t *digits);
/* wire/mod.rs:386-397 (pack_dropped); with out == NULL only *len is set */
int lbn_pack_dropped(const lbn_params *params, const uint16_t *top, const uint16_t *digits, uint8_t *out,
                     size_t cap, size_t *len);
/* verifier.rs:61-100, the CommitmentValue::Dropped arm */
int lbn_derive_evaluation_point_dropped(const lbn_pp *pp, lbn_transcript *t, const uint16_t *top,
                                        const uint16_t *digits, lbn_f162 *p0, lbn_f162 *p1);
/* bd/residual.rs:77-130: the squared CRT norm of A v - sum_j c_j col_j (as two u64 halves) */
int lbn_bd_residual(const lbn_pp *pp, const uint16_t *top, const uint16_t *digits, const lbn_challenge *c,
                    size_t r, const int16_t *v, size_t len_ring, uint64_t *normsq_lo, uint64_t *normsq_hi);
/* The cross-field switch (switch.rs). B128 values are two u64, the low word first.
 * switch::prove / prove_bits (switch.rs:63-100, 137-164) on the device: w is the lifted trace
 * (Witness::lifted, elements [lo, hi, 0]; the third word is read as zero) resident on pr, r_hi has
 * witness_log_len coordinates, r_lo none (a claim on the words) or 7 (on their bits). Out: the 128
 * partial evaluations, the witness_log_len sumcheck rounds [e0, einf], the opened value, and the
 * point (EvaluationPoint::msb_first) at which the commitment to w must open to it. t advances as
 * the reference's transcript does. */
int lbn_switch_prove(lbn_prover *pr, const lbn_dwitness *w, const uint64_t *r_lo, size_t r_lo_len,
                     const uint64_t *r_hi, size_t r_hi_len, const uint64_t claim[2], lbn_transcript *t,
                     uint64_t *partials, lbn_f162 *rounds, lbn_f162 *opened, lbn_f162 *p0, lbn_f162 *p1);
/* switch::verify / verify_bits (switch.rs:79-86, 110-121, 166-209) on the host: LBN_OK with the
 * point, or LBN_REJECTED. */
int lbn_switch_verify(const lbn_params *params, const uint64_t *r_lo, size_t r_lo_len, const uint64_t *r_hi,
                      size_t r_hi_len, const uint64_t claim[2], const uint64_t *partials, const lbn_f162 *rounds,
                      const lbn_f162 *opened, lbn_transcript *t, lbn_f162 *p0, lbn_f162 *p1);
/* Params::bd_cap, scheme/params.rs:250-252 and bd/mod.rs:71-74 (0 if the parameters drop no bits) */
uint64_t lbn_bd_cap(const lbn_params *params);
/* verifier.rs:232-275 (verify_folded_opening_bd): LBN_OK or LBN_REJECTED */
int lbn_verify_opening_bit_dropped(const lbn_pp *pp, const uint16_t *top, const uint16_t *digits,
                                   const lbn_challenge *c, size_t r, const int16_t *v, size_t len_ring,
                                   const lbn_f162 *p0, size_t p0_len, const lbn_f162 *folded_row_value);

/* The verifier's heavy steps on a prover's device (the prover is only the device context: its key,
 * stream and scratch; nothing of its own commitment or fold is read or changed). Same results as
 * the host entry points above:
 * verifier.rs:186-224 and eval.rs:202-249 (as lbn_verify_opening_clear): the folded witness's norm,
 * A v decomposed against the folded commitment, and its row check; LBN_OK or LBN_REJECTED */
int lbn_verify_opening_clear_device(lbn_prover *prover, const int16_t *folded_commitment, const int16_t *v,
                                    size_t len_ring, const lbn_f162 *p0, size_t p0_len,
                                    const lbn_f162 *folded_row_value);
/* Verifier::fold_commitment, fold.rs:799-873 (as lbn_fold_commitment) */
int lbn_fold_commitment_device(lbn_prover *prover, const int16_t *components, const lbn_challenge *c, size_t r,
                               int16_t *out);

/* Wire forms (wire/mod.rs, wire/rans.rs). Packers and the encoder write `out` (capacity `cap`)
 * and set *len; with out == NULL only *len is set. Unpackers and the decoder return LBN_OK,
 * LBN_WIRE_TRUNCATED or LBN_WIRE_MALFORMED exactly where the reference returns Ok,
 * WireError::Truncated or WireError::Malformed. */
/* Commitment::wire_bytes: commitment_bytes (wire/mod.rs:473-476) or bd::bytes (bd/mod.rs:70-72) */
int lbn_commitment_wire_bytes(const lbn_params *params, size_t *len);
/* pack_commitment (wire/mod.rs:226-243) of the prover's last commitment, packed on the device from
 * the matrix (or, with bit-dropped parameters, from the digits, wire/mod.rs:386-397, computed
 * there if lbn_commit_drop_bits has not) that the commit left in HBM */
int lbn_commit_pack(lbn_prover *prover, uint8_t *out, size_t cap, size_t *len);
/* pack_commitment (wire/mod.rs:226-243) of host components, Clear parameters */
int lbn_pack_commitment(const lbn_params *params, const int16_t *components, uint8_t *out, size_t cap,
                        size_t *len);
/* unpack_commitment (wire/mod.rs:246-301), the matrix arm; components as lbn_commit returns them */
int lbn_unpack_commitment(const lbn_params *params, const uint8_t *bytes, size_t len, int16_t *components);
/* unpack_dropped (wire/mod.rs:399-431) */
int lbn_unpack_dropped(const lbn_params *params, const uint8_t *bytes, size_t len, uint16_t *top,
                       uint16_t *digits);
/* pack_f162 / unpack_f162 (wire/mod.rs:445-470): 162 bits an element (the row evaluation,
 * pack_row_evaluation / unpack_row_evaluation, wire/mod.rs:434-442) */
int lbn_pack_f162(const lbn_f162 *x, size_t n, uint8_t *out, size_t cap, size_t *len);
int lbn_unpack_f162(const uint8_t *bytes, size_t len, size_t n, lbn_f162 *x);
/* An upper bound on the encoded size of `count` ring elements. */
size_t lbn_wire_encode_bound(size_t count);
/* encode (rans.rs:18-99) of the prover's last folded witness (Prover::fold's output, coefficient
 * representation), on the device over the copy the fold left in HBM */
int lbn_fold_encode(lbn_prover *prover, uint16_t base_q, uint8_t *out, size_t cap, size_t *len);
/* encode (rans.rs:18-99) of a host message of `count` ring elements ([count][648] int16), on the
 * device of the prover; ntt is the header's representation flag */
int lbn_wire_encode_device(lbn_prover *prover, const int16_t *v, size_t count, int ntt, uint16_t base_q,
                           uint8_t *out, size_t cap, size_t *len);
/* decode (rans.rs:101-196), on the host: v holds cap_count ring elements; *count and *ntt are set
 * from the header (LBN_ERR_SHAPE if count > cap_count) */
int lbn_wire_decode(const uint8_t *bytes, size_t len, int16_t *v, size_t cap_count, uint32_t *count, int *ntt);
/* Device time of the last encode, in ms: [0] total [1] min / max / counts [2] the rANS kernel */
int lbn_wire_timings(const lbn_prover *prover, float *ms, size_t n);

/* BLAKE3 as the transcript and the seed expansions use it: mode 0 hash, 1 keyed (key = 32 bytes),
 * 2 derive_key (key = NUL-terminated context). */
int lbn_blake3(int mode, const uint8_t *key, const uint8_t *in, size_t len, uint8_t *out,
               size_t out_len);
/* F162 product, fields/scalar.rs:92-116 */
void lbn_f162_mul(const lbn_f162 *a, const lbn_f162 *b, lbn_f162 *out);

#ifdef __cplusplus
}
#endif

#endif
