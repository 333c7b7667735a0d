// The serving loop (lbn_server): continuous batching of commit -> evaluate -> fold requests of
// many sizes under one key.
//
//   submit   -> pending (arrival order)
//   scheduler: admits from the front while the request's device bytes fit the budget, uploads the
//              admitted witnesses into the device pool, commits them in one batched commitment
//              (lbn_commit_batch on the committer, the key's own prover) -> ready
//   workers  : each takes a committed request and runs the rest of its round on its own prover
//              for that size (its own HIP stream): bit-dropped digits, point, row values, claim,
//              challenges, the fold of the resident witness; then gives the witness back to the
//              pool and the bytes back to the budget
//
// The public matrix A stays in HBM as the committer's and every prover's GEMM tables; a request's
// witness stays resident from admission to the end of its fold. A request that fails (a library
// error on its own data, or the test hook) reports its status; the others go on.
#include <chrono>
#include <condition_variable>
#include <cstring>
#include <deque>
#include <map>
#include <mutex>
#include <thread>
#include <unordered_map>

#include "internal.hpp"

namespace {

struct Request {
    uint64_t ticket = 0;
    lbn_params params{};
    lbn::Shape s;
    const lbn_f162 *w = nullptr;
    std::vector<uint8_t> domain;
    lbn_server_outputs out{};
    uint32_t flags = 0;
    uint64_t bytes = 0;
    lbn_dwitness *dw = nullptr;
    int status = LBN_OK;
    double t_arrival = 0, t_admitted = 0, t_committed = 0, t_done = 0;
    uint32_t batch = 0;
    bool done = false;
};

double seconds_since(std::chrono::steady_clock::time_point t0) {
    return std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
}

}  // namespace

struct lbn_server {
    lbn_server_config cfg{};
    uint8_t seed[32];
    lbn_params key{};
    lbn::Shape ks;
    lbn_pool *pool = nullptr;
    lbn_pp *key_pp = nullptr;
    lbn_prover *committer = nullptr;
    std::map<uint32_t, lbn_pp *> dev_pp;                     // per witness_log_len (the fold's tables)
    std::map<std::pair<uint32_t, uint32_t>, lbn_pp *> host_pp;  // per (witness_log_len, dropped bits)
    struct Worker {
        std::thread th;
        std::map<uint32_t, lbn_prover *> provers;
    };
    std::vector<Worker> workers;
    std::mutex mu;  // the queues, the budget, the tables above
    std::condition_variable cv_sched, cv_work, cv_done;
    std::deque<Request *> pending, ready;
    std::unordered_map<uint64_t, Request *> all;
    uint64_t next_ticket = 1;
    uint64_t reserved = 0, fixed = 0;
    uint32_t inflight = 0;
    bool stop_sched = false, stop_work = false;
    std::thread sched;
    lbn_server_stats st{};
    std::chrono::steady_clock::time_point t0 = std::chrono::steady_clock::now();
};

namespace lbn {

static lbn_params size_params(const lbn_params &key, uint32_t wl, uint32_t dropped) {
    lbn_params p = key;
    p.witness_log_len = wl;
    p.column_log_len = wl - (key.witness_log_len - key.column_log_len);
    p.opening = dropped ? LBN_OPENING_BIT_DROPPED : LBN_OPENING_CLEAR;
    p.dropped_bits = dropped;
    return p;
}

static uint64_t request_bytes(const Shape &s) {
    return pool_rounded(s.n * sizeof(lbn_f162)) + pool_rounded(bytes_size((uint32_t)s.len_ring, gemm_kpad(s.r)));
}

// The device pp of a size (created on first use; immutable and shared by the workers' provers).
static int size_pp(lbn_server *sv, uint32_t wl, lbn_pp **out) {
    std::lock_guard<std::mutex> lk(sv->mu);
    auto it = sv->dev_pp.find(wl);
    if (it != sv->dev_pp.end()) {
        *out = it->second;
        return LBN_OK;
    }
    const lbn_params p = size_params(sv->key, wl, 0);
    lbn_pp *pp = nullptr;
    const int rc = lbn_pp_create(&p, sv->seed, &pp);
    if (rc) return rc;
    sv->dev_pp[wl] = pp;
    *out = pp;
    return LBN_OK;
}

static int size_host_pp(lbn_server *sv, const lbn_params &p, lbn_pp **out) {
    std::lock_guard<std::mutex> lk(sv->mu);
    const auto k = std::make_pair(p.witness_log_len, p.opening == LBN_OPENING_BIT_DROPPED ? p.dropped_bits : 0u);
    auto it = sv->host_pp.find(k);
    if (it != sv->host_pp.end()) {
        *out = it->second;
        return LBN_OK;
    }
    lbn_pp *pp = nullptr;
    const int rc = lbn_pp_create_host(&p, sv->seed, &pp);
    if (rc) return rc;
    sv->host_pp[k] = pp;
    *out = pp;
    return LBN_OK;
}

static int worker_prover(lbn_server *sv, lbn_server::Worker &wk, uint32_t wl, lbn_prover **out) {
    auto it = wk.provers.find(wl);
    if (it != wk.provers.end()) {
        *out = it->second;
        return LBN_OK;
    }
    lbn_pp *pp = nullptr;
    int rc = size_pp(sv, wl, &pp);
    if (rc) return rc;
    lbn_prover *pr = nullptr;
    if ((rc = lbn_prover_create(pp, &pr))) return rc;
    if ((rc = lbn_prover_set_pool(pr, sv->pool))) {
        lbn_prover_release(pr);
        return rc;
    }
    wk.provers[wl] = pr;
    *out = pr;
    return LBN_OK;
}

// Everything after the commitment, as one request's round.
static int run_round(lbn_server *sv, lbn_server::Worker &wk, Request *r) {
    const Shape &s = r->s;
    lbn_prover *pr = nullptr;
    int rc = worker_prover(sv, wk, s.wl, &pr);
    if (rc) return rc;
    lbn_pp *hp = nullptr;
    if ((rc = size_host_pp(sv, r->params, &hp))) return rc;
    const lbn_server_outputs &o = r->out;
    if (s.dropped_bits && (rc = lbn_drop_bits_device(pr, &r->params, o.raw, o.top, o.digits))) return rc;
    lbn_transcript *t = lbn_transcript_new(r->domain.data(), r->domain.size());
    if (!t) return fail(LBN_ERR_OOM, "transcript");
    rc = s.dropped_bits ? lbn_derive_evaluation_point_dropped(hp, t, o.top, o.digits, o.p0, o.p1)
                        : lbn_derive_evaluation_point(hp, t, o.components, o.p0, o.p1);
    if (!rc) rc = lbn_row_evaluate(pr, r->dw, o.p0, s.rows, o.u);
    if (!rc) rc = lbn_claim(o.u, s.r, o.p1, s.cl, o.claim);
    if (!rc) rc = lbn_derive_folding_challenges(hp, t, o.u, s.r, o.challenges);
    lbn_transcript_free(t);
    if (rc) return rc;
    if (r->flags & LBN_SERVER_FAIL_FOLD) o.challenges[0].weight = 33;  // the library refuses it
    return lbn_fold_resident(pr, r->dw, o.challenges, s.r, o.v);
}

static void finish(lbn_server *sv, Request *r, int rc) {
    if (r->dw) {
        lbn_dwitness_release(r->dw);
        r->dw = nullptr;
    }
    {
        std::lock_guard<std::mutex> lk(sv->mu);
        r->status = rc;
        r->t_done = seconds_since(sv->t0);
        r->done = true;
        sv->reserved -= r->bytes;
        sv->inflight--;
        if (rc) sv->st.failed++;
    }
    sv->cv_done.notify_all();
    sv->cv_sched.notify_all();
}

static void worker_loop(lbn_server *sv, size_t idx) {
    lbn_server::Worker &wk = sv->workers[idx];
    for (;;) {
        Request *r = nullptr;
        {
            std::unique_lock<std::mutex> lk(sv->mu);
            sv->cv_work.wait(lk, [&] { return sv->stop_work || !sv->ready.empty(); });
            if (sv->ready.empty()) return;
            r = sv->ready.front();
            sv->ready.pop_front();
        }
        int rc;
        try {
            rc = run_round(sv, wk, r);
        } catch (const std::exception &) {
            rc = LBN_ERR_OOM;
        }
        finish(sv, r, rc);
    }
}

static bool admissible(lbn_server *sv) {
    if (sv->pending.empty()) return false;
    if (sv->cfg.serial && sv->inflight > 0) return false;
    return sv->reserved + sv->pending.front()->bytes + sv->fixed <= sv->cfg.hbm_budget;
}

static void sched_loop(lbn_server *sv) {
    std::vector<Request *> batch;
    std::vector<const lbn_dwitness *> dws;
    std::vector<int16_t *> comps;
    std::vector<uint32_t *> raws;
    for (;;) {
        batch.clear();
        {
            std::unique_lock<std::mutex> lk(sv->mu);
            sv->cv_sched.wait(lk, [&] { return (sv->stop_sched && sv->pending.empty()) || admissible(sv); });
            if (sv->pending.empty()) return;  // stopping, nothing left
            size_t cols = 0;
            while (admissible(sv)) {
                Request *r = sv->pending.front();
                if (!batch.empty() && (sv->cfg.serial || cols + r->s.r > sv->cfg.max_batch_columns)) break;
                sv->pending.pop_front();
                sv->reserved += r->bytes;
                sv->inflight++;
                cols += r->s.r;
                batch.push_back(r);
            }
            sv->st.reserved_peak = std::max(sv->st.reserved_peak, sv->reserved + sv->fixed);
        }
        dws.clear();
        comps.clear();
        raws.clear();
        std::vector<Request *> live;
        for (Request *r : batch) {
            const int rc = lbn_witness_upload_params(sv->committer, &r->params, r->w, r->s.n, &r->dw);
            r->t_admitted = seconds_since(sv->t0);
            if (rc) {
                finish(sv, r, rc);
                continue;
            }
            live.push_back(r);
            dws.push_back(r->dw);
            comps.push_back(r->out.components);
            raws.push_back(r->out.raw);
        }
        if (live.empty()) continue;
        const int rc = lbn_commit_batch(sv->committer, dws.data(), dws.size(), comps.data(), raws.data());
        const double t = seconds_since(sv->t0);
        uint32_t bno;
        {
            std::lock_guard<std::mutex> lk(sv->mu);
            bno = (uint32_t)sv->st.batches++;
            sv->st.batched_requests += live.size();
            for (Request *r : live) sv->st.batched_columns += r->s.r;
        }
        for (Request *r : live) {
            r->t_committed = t;
            r->batch = bno;
            if (rc) finish(sv, r, rc);
        }
        if (rc) continue;
        {
            std::lock_guard<std::mutex> lk(sv->mu);
            for (Request *r : live) sv->ready.push_back(r);
        }
        sv->cv_work.notify_all();
    }
}

static void free_server(lbn_server *sv) {
    for (auto &w : sv->workers)
        for (auto &kv : w.provers) lbn_prover_release(kv.second);
    if (sv->committer) lbn_prover_release(sv->committer);
    for (auto &kv : sv->dev_pp) lbn_pp_release(kv.second);
    for (auto &kv : sv->host_pp) lbn_pp_release(kv.second);
    if (sv->key_pp) lbn_pp_release(sv->key_pp);
    for (auto &kv : sv->all) delete kv.second;
    if (sv->pool) lbn_pool_release(sv->pool);
    delete sv;
}

// Runs every kernel and autotunes every batch width class the server will meet: one batched
// commitment of zero witnesses per power of two from the smallest size's columns up to
// max_batch_columns, and a round of each size on every worker's prover.
static int warm_up(lbn_server *sv) {
    if (sv->cfg.n_sizes == 0) return LBN_OK;
    uint32_t smallest = sv->cfg.sizes[0];
    for (uint32_t i = 0; i < sv->cfg.n_sizes; i++) smallest = std::min(smallest, sv->cfg.sizes[i]);
    const lbn_params p = size_params(sv->key, smallest, 0);
    Shape s;
    int rc = make_shape(&p, &s);
    if (rc) return rc;
    std::vector<lbn_f162> zero(s.n);
    std::memset(zero.data(), 0, zero.size() * sizeof(lbn_f162));
    lbn_dwitness *dw = nullptr;
    if ((rc = lbn_witness_upload_params(sv->committer, &p, zero.data(), zero.size(), &dw))) return rc;
    for (size_t width = s.r; width <= sv->cfg.max_batch_columns && !rc; width *= 2) {
        std::vector<const lbn_dwitness *> ws(width / s.r, dw);
        rc = lbn_commit_batch(sv->committer, ws.data(), ws.size(), nullptr, nullptr);
    }
    lbn_dwitness_release(dw);
    return rc;
}

}  // namespace lbn

using namespace lbn;

extern "C" {

uint64_t lbn_server_request_bytes(const lbn_params *params) {
    Shape s;
    if (make_shape(params, &s)) return 0;
    return request_bytes(s);
}

int lbn_server_create(const lbn_params *key, const uint8_t matrix_seed[32], const lbn_server_config *config,
                      lbn_server **out) {
    if (!key || !matrix_seed || !config || !out) return fail(LBN_ERR_ARG, "null argument");
    *out = nullptr;
    if (config->n_sizes > 8) return fail(LBN_ERR_ARG, "at most 8 sizes");
    lbn_server *sv = new lbn_server();
    sv->cfg = *config;
    if (!sv->cfg.workers) sv->cfg.workers = 4;
    if (!sv->cfg.max_batch_columns) sv->cfg.max_batch_columns = 16384;
    std::memcpy(sv->seed, matrix_seed, 32);
    sv->key = *key;
    sv->key.opening = LBN_OPENING_CLEAR;
    sv->key.dropped_bits = 0;
    int rc = make_shape(&sv->key, &sv->ks);
    size_t free0 = 0, total = 0, free1 = 0;
    if (!rc && hipMemGetInfo(&free0, &total) != hipSuccess) rc = fail(LBN_ERR_NO_DEVICE, "hipMemGetInfo");
    if (!rc) rc = lbn_pp_create(&sv->key, matrix_seed, &sv->key_pp);
    if (!rc) rc = lbn_prover_create(sv->key_pp, &sv->committer);
    if (!rc) rc = commit_batch_reserve(sv->committer, sv->cfg.max_batch_columns);
    sv->workers.resize(sv->cfg.workers);
    for (uint32_t i = 0; i < sv->cfg.n_sizes && !rc; i++) {
        const lbn_params p = size_params(sv->key, sv->cfg.sizes[i], 0);
        Shape s;
        if ((rc = make_shape(&p, &s))) break;
        lbn_pp *pp = nullptr;
        if ((rc = size_pp(sv, sv->cfg.sizes[i], &pp))) break;
        for (auto &wk : sv->workers) {
            lbn_prover *pr = nullptr;
            if ((rc = worker_prover(sv, wk, sv->cfg.sizes[i], &pr))) break;
        }
    }
    // the provers draw from the pool from here on (their warm-up layouts go back to the device)
    if (!rc) rc = lbn_pool_create(sv->cfg.hbm_budget, &sv->pool);
    if (!rc) rc = lbn_prover_set_pool(sv->committer, sv->pool);
    for (auto &wk : sv->workers)
        for (auto &kv : wk.provers)
            if (!rc) rc = lbn_prover_set_pool(kv.second, sv->pool);
    // what the server holds of its own: the key's and every prover's tables and workspaces
    if (!rc && hipMemGetInfo(&free1, &total) != hipSuccess) rc = fail(LBN_ERR_NO_DEVICE, "hipMemGetInfo");
    if (!rc) {
        sv->fixed = free0 > free1 ? free0 - free1 : 0;
        if (sv->fixed >= sv->cfg.hbm_budget) rc = fail(LBN_ERR_ADMISSION, "the server's own workspaces exceed the budget");
    }
    if (!rc) pool_set_cap(sv->pool, sv->cfg.hbm_budget - sv->fixed);
    if (!rc) rc = warm_up(sv);
    if (rc) {
        free_server(sv);
        return rc;
    }
    sv->st.fixed_bytes = sv->fixed;
    sv->t0 = std::chrono::steady_clock::now();
    try {
        for (size_t i = 0; i < sv->workers.size(); i++) sv->workers[i].th = std::thread(worker_loop, sv, i);
        sv->sched = std::thread(sched_loop, sv);
    } catch (const std::exception &) {
        {
            std::lock_guard<std::mutex> lk(sv->mu);
            sv->stop_sched = sv->stop_work = true;
        }
        sv->cv_work.notify_all();
        sv->cv_sched.notify_all();
        for (auto &w : sv->workers)
            if (w.th.joinable()) w.th.join();
        if (sv->sched.joinable()) sv->sched.join();
        free_server(sv);
        return fail(LBN_ERR_OOM, "server threads");
    }
    *out = sv;
    return LBN_OK;
}

int lbn_server_submit(lbn_server *sv, const lbn_params *params, const lbn_f162 *w, size_t n, const uint8_t *domain,
                      size_t domain_len, const lbn_server_outputs *o, uint32_t flags, uint64_t *ticket) {
    if (!sv || !params || !w || !o || !ticket || (!domain && domain_len)) return fail(LBN_ERR_ARG, "null argument");
    Shape s;
    int rc = make_shape(params, &s);
    if (rc) return rc;
    if (s.opening == LBN_OPENING_RECURSIVE) return fail(LBN_ERR_UNSUPPORTED, "Opening::Recursive");
    if (s.rows != sv->ks.rows || s.primes != sv->ks.primes)
        return fail(LBN_ERR_SHAPE, "the request's parameters do not share the server's key");
    if (n != s.n) return fail(LBN_ERR_WITNESS_LENGTH, "the witness is not witness_len() elements");
    if (!o->components || !o->raw || !o->p0 || !o->p1 || !o->u || !o->claim || !o->challenges || !o->v ||
        (s.dropped_bits && (!o->top || !o->digits)))
        return fail(LBN_ERR_ARG, "a missing output buffer");
    const uint64_t bytes = request_bytes(s);
    {
        std::lock_guard<std::mutex> lk(sv->mu);
        if (bytes + sv->fixed > sv->cfg.hbm_budget)
            return fail(LBN_ERR_ADMISSION, "the request alone needs more device memory than the budget");
    }
    if (s.r > sv->cfg.max_batch_columns)
        return fail(LBN_ERR_ADMISSION, "the request has more columns than one batched commitment takes");
    Request *r;
    try {
        r = new Request();
        r->domain.assign(domain, domain + domain_len);
    } catch (const std::exception &) {
        return fail(LBN_ERR_OOM, "request");
    }
    r->params = *params;
    r->s = s;
    r->w = w;
    r->out = *o;
    r->flags = flags;
    r->bytes = bytes;
    {
        std::lock_guard<std::mutex> lk(sv->mu);
        r->ticket = sv->next_ticket++;
        r->t_arrival = seconds_since(sv->t0);
        sv->all[r->ticket] = r;
        sv->pending.push_back(r);
        sv->st.requests++;
    }
    sv->cv_sched.notify_all();
    *ticket = r->ticket;
    return LBN_OK;
}

int lbn_server_wait(lbn_server *sv, uint64_t ticket, lbn_server_info *info) {
    if (!sv) return fail(LBN_ERR_ARG, "null argument");
    std::unique_lock<std::mutex> lk(sv->mu);
    auto it = sv->all.find(ticket);
    if (it == sv->all.end()) return fail(LBN_ERR_ARG, "unknown ticket");
    Request *r = it->second;
    sv->cv_done.wait(lk, [&] { return r->done; });
    if (info) {
        info->status = r->status;
        info->arrival_s = r->t_arrival;
        info->admitted_s = r->t_admitted;
        info->committed_s = r->t_committed;
        info->done_s = r->t_done;
        info->bytes = r->bytes;
        info->batch = r->batch;
    }
    sv->all.erase(it);
    delete r;
    return LBN_OK;
}

int lbn_server_set_budget(lbn_server *sv, uint64_t hbm_budget) {
    if (!sv) return fail(LBN_ERR_ARG, "null argument");
    {
        std::lock_guard<std::mutex> lk(sv->mu);
        if (hbm_budget <= sv->fixed) return fail(LBN_ERR_ADMISSION, "the server's own workspaces exceed the budget");
        sv->cfg.hbm_budget = hbm_budget;
        pool_set_cap(sv->pool, hbm_budget - sv->fixed);
    }
    sv->cv_sched.notify_all();
    return LBN_OK;
}

int lbn_server_stats_get(lbn_server *sv, lbn_server_stats *out) {
    if (!sv || !out) return fail(LBN_ERR_ARG, "null argument");
    {
        std::lock_guard<std::mutex> lk(sv->mu);
        *out = sv->st;
    }
    return lbn_pool_stats(sv->pool, &out->pool_obtained, nullptr, nullptr, &out->pool_mallocs, &out->pool_reuses);
}

void lbn_server_destroy(lbn_server *sv) {
    if (!sv) return;
    {
        std::lock_guard<std::mutex> lk(sv->mu);
        sv->stop_sched = true;
    }
    sv->cv_sched.notify_all();
    if (sv->sched.joinable()) sv->sched.join();  // every queued request is committed
    {
        std::lock_guard<std::mutex> lk(sv->mu);
        sv->stop_work = true;
    }
    sv->cv_work.notify_all();
    for (auto &w : sv->workers)
        if (w.th.joinable()) w.th.join();  // and its round run
    free_server(sv);
}

}  // extern "C"
