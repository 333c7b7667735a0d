// Many witnesses of uneven sizes committed in one launch.
//
// Every witness whose Params share the prover's key (the row length 2^row_log_len and the moduli;
// CommitmentKey::random, key.rs:48, depends on nothing else) is a list of independent columns, so
// a batch is one commitment over the union of their columns: the layout pass reads column c
// through a table of column addresses, and the GEMM, the finish and the outputs run over all
// columns at once. The sequence (layout, GEMM, finish) is captured once per batch width in
// a HIP graph and replayed; the GEMM's schedule (tiles per XCD round, K ranges) is chosen per
// width class by timing the candidates on the device the first time that class is seen.
#include <algorithm>
#include <cstring>

#include "internal.hpp"
#include "launch.hpp"

namespace lbn {

void batch_free(lbn_batch_ws *b) {
    if (!b) return;
    if (b->graph) (void)hipGraphExecDestroy(b->graph);
    if (b->graph_def) (void)hipGraphDestroy(b->graph_def);
    void *bufs[] = {b->d_bytes, b->d_yacc, b->d_raw, b->d_comp, (void *)b->d_cols};
    for (void *p : bufs)
        if (p) (void)hipFree(p);
    if (b->h_cols) (void)hipHostFree((void *)b->h_cols);
    for (hipEvent_t &e : b->ev)
        if (e) (void)hipEventDestroy(e);
    delete b;
}

static int batch_reserve(lbn_prover *pr, size_t cols) {
    if (!pr->batch) {
        pr->batch = new lbn_batch_ws();
        for (hipEvent_t &e : pr->batch->ev) HIP_TRY(hipEventCreate(&e));
    }
    lbn_batch_ws *b = pr->batch;
    if (cols <= b->cap_cols) return LBN_OK;
    const lbn_pp *pp = pr->pp;
    const size_t limbs = pp->shape.primes.size();
    const size_t cap = std::max(cols, b->cap_cols * 2);
    void *old[] = {b->d_bytes, b->d_yacc, b->d_raw, b->d_comp, (void *)b->d_cols};
    for (void *p : old)
        if (p) (void)hipFree(p);
    if (b->h_cols) (void)hipHostFree((void *)b->h_cols);
    b->d_bytes = nullptr;
    b->d_yacc = b->d_raw = nullptr;
    b->d_comp = nullptr;
    b->d_cols = nullptr;
    b->h_cols = nullptr;
    b->cap_cols = 0;
    if (b->graph) (void)hipGraphExecDestroy(b->graph);
    if (b->graph_def) (void)hipGraphDestroy(b->graph_def);
    b->graph = nullptr;
    b->graph_def = nullptr;
    b->graph_cols = 0;
    const size_t nb = bytes_size((uint32_t)cap, pp->kpad_c);
    int rc;
    if ((rc = dalloc(&b->d_bytes, nb))) return rc;
    HIP_TRY(hipMemset(b->d_bytes, 0, nb));  // padding inputs are never written and must read as zero
    if ((rc = dalloc(&b->d_yacc, cap * YACC_PLANES * limbs * N))) return rc;
    if ((rc = dalloc(&b->d_raw, cap * limbs * N))) return rc;
    if ((rc = dalloc(&b->d_comp, cap * 4 * limbs * N162))) return rc;
    if ((rc = dalloc(&b->d_cols, cap))) return rc;
    HIP_TRY(hipHostMalloc((void **)&b->h_cols, cap * sizeof(uint64_t *), hipHostMallocDefault));
    b->cap_cols = cap;
    return LBN_OK;
}

// The kernel sequence of one batched commitment over R columns: the GEMM with sched_t tiles per
// XCD round and at least min_splits K ranges (its count of them to *splits).
static int batch_launch(lbn_prover *pr, size_t R, uint32_t sched_t, uint32_t min_splits, bool with_prepass,
                        uint32_t *splits) {
    const lbn_pp *pp = pr->pp;
    const Shape &s = pp->shape;
    const size_t limbs = s.primes.size();
    lbn_batch_ws *b = pr->batch;
    hipStream_t st = pr->stream;
    if (with_prepass)
        HIP_TRY(launch_prepass(st, nullptr, b->d_cols, (uint32_t)s.column_len, (uint32_t)s.len_ring, (uint32_t)R,
                               b->d_bytes, pp->kpad_c, nullptr, 0));
    GemmArgs g = commit_gemm_args(pp, b->d_bytes, b->d_yacc, R);
    g.sched_t = sched_t;
    g.splits = min_splits;
    HIP_TRY(launch_tile_gemm(st, g, splits));
    HIP_TRY(launch_commit_finish(st, b->d_yacc, (uint32_t)limbs, pp->d_finish, pp->d_slot648, pp->d_qcs, b->d_raw,
                                 b->d_comp, (uint32_t)R));
    return LBN_OK;
}

// The autotuner's key: the width's power-of-two class (a serving loop's batches rarely repeat a
// width exactly; schedules change slowly with it).
static uint32_t width_class(size_t R) {
    uint32_t c = 0;
    while (((size_t)1 << c) < R) c++;
    return c;
}

// Candidate schedules, timed on this device for this width class: tiles per XCD round (the rest of
// an XCD's 32 blocks are column blocks) and the launcher's K ranges or twice as many (more, shorter
// items for a narrow batch). The result is the tiles per round and the ranges asked for.
static int autotune(lbn_prover *pr, size_t R, uint32_t *best_t, uint32_t *best_s) {
    lbn_batch_ws *b = pr->batch;
    for (const auto &e : b->tuned)
        if (e[0] == width_class(R)) {
            *best_t = e[1];
            *best_s = e[2];
            return LBN_OK;
        }
    static const uint32_t ts[] = {4, 8, 16};
    float best = 1e30f;
    *best_t = 8;
    *best_s = 0;
    int rc;
    uint32_t auto_splits = 1;
    if ((rc = batch_launch(pr, R, 8, 0, true, &auto_splits))) return rc;  // the layout, and a warm-up
    const uint32_t ps[] = {0, 2 * auto_splits};
    for (uint32_t t : ts)
        for (uint32_t p : ps) {
            HIP_TRY(hipEventRecord(b->ev[0], pr->stream));
            if ((rc = batch_launch(pr, R, t, p, false, nullptr))) return rc;
            HIP_TRY(hipEventRecord(b->ev[1], pr->stream));
            HIP_TRY(hipEventSynchronize(b->ev[1]));
            const float ms = elapsed(b->ev[0], b->ev[1]);
            if (ms < best) {
                best = ms;
                *best_t = t;
                *best_s = p;
            }
        }
    b->tuned.push_back({width_class(R), *best_t, *best_s});
    return LBN_OK;
}

int commit_batch_reserve(lbn_prover *pr, size_t cols) { return batch_reserve(pr, cols); }

}  // namespace lbn

using namespace lbn;

extern "C" {

int lbn_witness_upload_params(lbn_prover *pr, const lbn_params *params, const lbn_f162 *w, size_t n,
                              lbn_dwitness **out) {
    if (!pr || !params || !w || !out) return fail(LBN_ERR_ARG, "null argument");
    Shape s;
    int rc = make_shape(params, &s);
    if (rc) return rc;
    const Shape &k = pr->pp->shape;
    if (s.rows != k.rows || s.primes != k.primes)
        return fail(LBN_ERR_SHAPE, "the witness's parameters do not share this prover's key");
    if (n != s.n) return fail(LBN_ERR_WITNESS_LENGTH, "the witness is not witness_len() elements");
    return witness_alloc_upload(pr, w, n, s.r, out);
}

int lbn_commit_batch(lbn_prover *pr, const lbn_dwitness *const *ws, size_t count, int16_t *const *components,
                     uint32_t *const *raw) {
    if (!pr || !ws || count == 0) return fail(LBN_ERR_ARG, "null argument");
    const lbn_pp *pp = pr->pp;
    const Shape &s = pp->shape;
    const size_t limbs = s.primes.size(), clen = s.column_len;
    size_t R = 0;
    for (size_t i = 0; i < count; i++) {
        if (!ws[i] || !ws[i]->d_w) return fail(LBN_ERR_ARG, "null witness");
        if (ws[i]->owner->pp->device != pp->device || ws[i]->n != ws[i]->cols * clen ||
            ws[i]->owner->pp->shape.rows != s.rows || ws[i]->owner->pp->shape.primes != s.primes)
            return fail(LBN_ERR_SHAPE, "a witness does not share this prover's key");
        R += ws[i]->cols;
    }
    int rc = batch_reserve(pr, R);
    if (rc) return rc;
    lbn_batch_ws *b = pr->batch;
    hipStream_t st = pr->stream;
    size_t c = 0;
    for (size_t i = 0; i < count; i++)
        for (size_t j = 0; j < ws[i]->cols; j++) b->h_cols[c++] = ws[i]->d_w + 3 * j * clen;
    HIP_TRY(hipMemcpyAsync((void *)b->d_cols, (const void *)b->h_cols, R * sizeof(uint64_t *),
                           hipMemcpyHostToDevice, st));
    uint32_t t = 8, p = 0;
    if ((rc = autotune(pr, R, &t, &p))) return rc;
    // a width seen twice in a row is captured once as a graph and replayed; any other runs its
    // three launches directly
    const bool replay = b->last_cols == R;
    b->last_cols = R;
    if (!replay) {
        HIP_TRY(hipEventRecord(b->ev[0], st));
        if ((rc = batch_launch(pr, R, t, p, true, &b->splits))) return rc;
        HIP_TRY(hipEventRecord(b->ev[1], st));
    } else if (!b->graph || b->graph_cols != R || b->graph_t != t || b->graph_min_splits != p) {
        if (b->graph) (void)hipGraphExecDestroy(b->graph);
        if (b->graph_def) (void)hipGraphDestroy(b->graph_def);
        b->graph = nullptr;
        b->graph_def = nullptr;
        HIP_TRY(hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal));
        rc = batch_launch(pr, R, t, p, true, &b->graph_splits);
        hipGraph_t gdef = nullptr;
        const hipError_t ec = hipStreamEndCapture(st, &gdef);
        if (rc) {
            if (gdef) (void)hipGraphDestroy(gdef);
            return rc;
        }
        HIP_TRY(ec);
        b->graph_def = gdef;
        HIP_TRY(hipGraphInstantiate(&b->graph, gdef, nullptr, nullptr, 0));
        b->graph_cols = R;
        b->graph_t = t;
        b->graph_min_splits = p;
    }
    if (replay) {
        HIP_TRY(hipEventRecord(b->ev[0], st));
        HIP_TRY(hipGraphLaunch(b->graph, st));
        HIP_TRY(hipEventRecord(b->ev[1], st));
        b->splits = b->graph_splits;
    }
    // outputs: components [R][4][limb][162] are contiguous per witness; raw [limb][R][648] per limb
    c = 0;
    for (size_t i = 0; i < count; i++) {
        const size_t ci = ws[i]->cols;
        if (components && components[i])
            HIP_TRY(hipMemcpyAsync(components[i], b->d_comp + c * 4 * limbs * N162, ci * 4 * limbs * N162 * 2,
                                   hipMemcpyDeviceToHost, st));
        if (raw && raw[i])
            for (size_t k = 0; k < limbs; k++)
                HIP_TRY(hipMemcpyAsync(raw[i] + k * ci * N, b->d_raw + (k * R + c) * N, ci * N * 4,
                                       hipMemcpyDeviceToHost, st));
        c += ci;
    }
    HIP_TRY(hipStreamSynchronize(st));
    b->ms = elapsed(b->ev[0], b->ev[1]);
    b->sched_t = t;
    return LBN_OK;
}

int lbn_batch_info(const lbn_prover *pr, uint32_t *sched_t, uint32_t *splits, float *ms) {
    if (!pr || !pr->batch) return fail(LBN_ERR_ARG, "no batch has run on this prover");
    if (sched_t) *sched_t = pr->batch->sched_t;
    if (splits) *splits = pr->batch->splits;
    if (ms) *ms = pr->batch->ms;
    return LBN_OK;
}

}  // extern "C"
