// The cross-field switch (switch.rs): the prover's passes on the device over a lifted B128 trace
// that is already resident, the transcript on the host, and the verifier on the host.
#include <algorithm>
#include <cstring>
#include <vector>

#include "internal.hpp"
#include "launch.hpp"

namespace lbn {
namespace {

struct B128 {
    uint64_t w[2];
};

B128 bmul(const B128 &a, const B128 &b) {
    B128 o;
    b128_mul(a.w[0], a.w[1], b.w[0], b.w[1], o.w[0], o.w[1]);
    return o;
}

// eq(r)[j] with coordinate i taking bit i of j (the table the device multiplies its parts of)
std::vector<B128> eq_b128_lsb(const B128 *r, int n) {
    std::vector<B128> t((size_t)1 << n, B128{{0, 0}});
    t[0] = B128{{1, 0}};
    for (int k = 0; k < n; k++) {
        const size_t len = (size_t)1 << k;
        for (size_t b = 0; b < len; b++) {
            const B128 hi = bmul(t[b], r[k]);
            t[b + len] = hi;
            t[b] = B128{{t[b].w[0] ^ hi.w[0], t[b].w[1] ^ hi.w[1]}};
        }
    }
    return t;
}

// crossfield.rs:85-99 / 66-78: coordinate 0 the most significant bit of the index
std::vector<B128> eq_b128_msb(const B128 *r, int n) {
    std::vector<B128> rev(r, r + n);
    std::reverse(rev.begin(), rev.end());
    return eq_b128_lsb(rev.data(), n);
}

std::vector<F162> eq_f162_msb(const F162 *r, int n) {
    std::vector<F162> t((size_t)1 << n, F162{{0, 0, 0}});
    t[0] = F162{{1, 0, 0}};
    for (int k = 0; k < n; k++) {
        const size_t len = (size_t)1 << k;
        const F162 &rk = r[n - 1 - k];  // the last coordinate takes bit 0
        for (size_t b = 0; b < len; b++) {
            const F162 hi = f162_mul(t[b], rk);
            t[b + len] = hi;
            t[b] = f162_add(t[b], hi);
        }
    }
    return t;
}

// challenge.rs:113-149
void absorb_b128(Transcript &t, const B128 *x, size_t n) {
    t.absorb_u64(n);
    for (size_t i = 0; i < n; i++) {
        t.absorb_u64(x[i].w[0]);
        t.absorb_u64(x[i].w[1]);
    }
}
void absorb_f162(Transcript &t, const F162 *x, size_t n) {
    t.absorb_u64(n);
    for (size_t i = 0; i < n; i++)
        for (int k = 0; k < 3; k++) t.absorb_u64(x[i].w[k]);
}
uint64_t le64(const uint8_t *b) {
    uint64_t x = 0;
    for (int k = 0; k < 8; k++) x |= (uint64_t)b[k] << (8 * k);
    return x;
}
std::vector<F162> sample_f162(Transcript &t, const char *label, size_t n) {
    std::vector<uint8_t> b(24 * n);
    t.fill(label, strlen(label), b.data(), b.size());
    std::vector<F162> out(n);
    for (size_t i = 0; i < n; i++)
        out[i] = F162{{le64(&b[24 * i]), le64(&b[24 * i + 8]), le64(&b[24 * i + 16]) & ((1ull << 34) - 1)}};
    return out;
}

// switch.rs:129-134
void statement(Transcript &t, const B128 *r_lo, size_t r_lo_len, const B128 *r_hi, size_t l, const B128 &claim) {
    t.absorb_bytes("labinius/switch", 15);
    absorb_b128(t, r_lo, r_lo_len);
    absorb_b128(t, r_hi, l);
    absorb_b128(t, &claim, 1);
}

// EvaluationPoint::msb_first (scheme/commitment.rs:163-174)
void msb_first(const std::vector<F162> &r_pp, size_t cl, lbn_f162 *p0, lbn_f162 *p1) {
    const size_t l = r_pp.size();
    for (size_t i = 0; i < cl; i++) std::memcpy(p1[i].w, r_pp[cl - 1 - i].w, 24);
    for (size_t i = 0; i < l - cl; i++) std::memcpy(p0[i].w, r_pp[l - 1 - i].w, 24);
}

// crossfield.rs:9-62,106-115: a tensor of 128 F162 (index i = the B128 basis element x^i)
F162 transparent_coeff(const B128 *r_hi, const std::vector<F162> &r_pp, const std::vector<F162> &batch) {
    std::vector<F162> t(128, F162{{0, 0, 0}}), hs(128);
    t[0] = F162{{1, 0, 0}};
    for (size_t round = 0; round < r_pp.size(); round++) {
        const B128 &h = r_hi[round];
        // scale_horizontal: Horner over h's bits from the top, x acting as the shift with the
        // x^128 = x^7 + x^2 + x + 1 wrap
        std::fill(hs.begin(), hs.end(), F162{{0, 0, 0}});
        for (int k = 127; k >= 0; k--) {
            const F162 top = hs[127];
            for (int i = 127; i > 0; i--) hs[i] = hs[i - 1];
            hs[0] = F162{{0, 0, 0}};
            for (int tap : {0, 1, 2, 7}) hs[tap] = f162_add(hs[tap], top);
            if ((h.w[k / 64] >> (k % 64)) & 1)
                for (int i = 0; i < 128; i++) hs[i] = f162_add(hs[i], t[i]);
        }
        for (int i = 0; i < 128; i++) t[i] = f162_add(f162_add(t[i], f162_mul(t[i], r_pp[round])), hs[i]);
    }
    F162 acc{{0, 0, 0}};
    for (int i = 0; i < 128; i++) acc = f162_add(acc, f162_mul(t[i], batch[i]));
    return acc;
}

int check_claim_shape(size_t r_lo_len) {
    if (r_lo_len != 0 && r_lo_len != 7) return fail(LBN_ERR_SHAPE, "r_lo has 0 coordinates (words) or 7 (bits)");
    return LBN_OK;
}

int ilog2(size_t n) {
    int l = 0;
    while (((size_t)1 << l) < n) l++;
    return l;
}

// The prover's switch workspace (grow-only).
// words of d_small (and of its pinned host side): parts, psi, the partials, the round messages
constexpr size_t SW_SMALL = (size_t)4 * 256 * 2 + 16 * 256 * 3 + 256 + 6 * 32;

int sw_reserve(lbn_prover *pr, size_t n, size_t np) {
    lbn_switch_ws &w = pr->sw;
    int rc;
    if (n > w.cap_n) {
        void *old[] = {w.d_eq, w.d_a, w.d_p};
        for (void *p : old)
            if (p) (void)hipFree(p);
        w.d_eq = w.d_a = w.d_p = nullptr;
        w.cap_n = 0;
        if ((rc = dalloc(&w.d_eq, 2 * n))) return rc;
        if ((rc = dalloc(&w.d_a, 3 * n))) return rc;
        if ((rc = dalloc(&w.d_p, 3 * (n / 2 + 1)))) return rc;
        w.cap_n = n;
    }
    if (!w.d_small) {
        // parts (up to 4 x 256 B128), psi (16 x 256 F162), v (128 B128), one message per round
        if ((rc = dalloc(&w.d_small, SW_SMALL))) return rc;
    }
    if (!w.h_small && hipHostMalloc((void **)&w.h_small, SW_SMALL * sizeof(uint64_t), hipHostMallocDefault) != hipSuccess) {
        w.h_small = nullptr;
        return fail(LBN_ERR_OOM, "the switch's pinned staging");
    }
    (void)np;
    return LBN_OK;
}

}  // namespace
}  // namespace lbn

using namespace lbn;

extern "C" {

int lbn_switch_prove(lbn_prover *pr, const lbn_dwitness *wt, const uint64_t *r_lo, size_t r_lo_len,
                     const uint64_t *r_hi, size_t r_hi_len, const uint64_t claim[2], lbn_transcript *tr,
                     uint64_t *partials, lbn_f162 *rounds, lbn_f162 *opened, lbn_f162 *p0, lbn_f162 *p1) {
    if (!pr || !wt || !r_hi || !claim || !tr || !partials || !rounds || !opened || !p0 || !p1 || (r_lo_len && !r_lo))
        return fail(LBN_ERR_ARG, "null argument");
    if (wt->owner != pr) return fail(LBN_ERR_ARG, "the witness belongs to another prover");
    int rc = check_claim_shape(r_lo_len);
    if (rc) return rc;
    const size_t n = wt->n, l = (size_t)ilog2(n), cl = (size_t)ilog2(wt->cols);
    if (((size_t)1 << l) != n || n < 256 || n > ((size_t)1 << 30))
        return fail(LBN_ERR_SHAPE, "the trace is 2^l words, 8 <= l <= 30");
    if (r_hi_len != l) return fail(LBN_ERR_SHAPE, "r_hi has witness_log_len coordinates");
    const size_t np = (l + 7) / 8;
    if ((rc = sw_reserve(pr, n, np))) return rc;
    lbn_switch_ws &ws = pr->sw;
    uint64_t *d_parts = ws.d_small, *d_psi = d_parts + 4 * 256 * 2, *d_v = d_psi + 16 * 256 * 3, *d_msg = d_v + 256;
    // the same layout in pinned host memory: every copy is asynchronous and every wait a poll of the
    // stream (a copy to or from pageable memory goes through the runtime's bounce buffer, and the
    // sumcheck makes one round trip a round)
    uint64_t *h_parts = ws.h_small, *h_psi = h_parts + 4 * 256 * 2, *h_v = h_psi + 16 * 256 * 3, *h_msg = h_v + 256;
    hipStream_t st = pr->stream;
    Transcript &t = tr->t;
    const B128 *rl = (const B128 *)r_lo, *rh = (const B128 *)r_hi;
    const B128 cb{{claim[0], claim[1]}};
    statement(t, rl, r_lo_len, rh, l, cb);

    // eq(r_hi): bit b of j takes coordinate l - 1 - b; part c = bits 8c .. 8c + 7
    uint64_t *parts = h_parts;
    std::memset(parts, 0, np * 256 * 2 * sizeof(uint64_t));
    for (size_t c = 0; c < np; c++) {
        const int nb = (int)std::min<size_t>(8, l - 8 * c);
        std::vector<B128> rc8(nb);
        for (int b = 0; b < nb; b++) rc8[b] = rh[l - 1 - (8 * c + b)];
        const std::vector<B128> e = eq_b128_lsb(rc8.data(), nb);
        for (size_t m = 0; m < e.size(); m++) {
            parts[2 * (c * 256 + m)] = e[m].w[0];
            parts[2 * (c * 256 + m) + 1] = e[m].w[1];
        }
    }
    HIP_TRY(hipMemcpyAsync(d_parts, parts, np * 256 * 2 * 8, hipMemcpyHostToDevice, st));
    HIP_TRY(launch_switch_eq(st, d_parts, (uint32_t)np, (uint32_t)n, ws.d_eq));
    // the partial evaluations
    HIP_TRY(launch_fill_u32(st, (uint32_t *)d_v, 2 * 256, 0));
    HIP_TRY(launch_switch_partials(st, wt->d_w, ws.d_eq, (uint32_t)n, d_v));
    HIP_TRY(hipMemcpyAsync(h_v, d_v, 128 * 16, hipMemcpyDeviceToHost, st));
    HIP_TRY(stream_wait(st));
    std::memcpy(partials, h_v, 128 * 16);
    absorb_b128(t, (const B128 *)partials, 128);
    const std::vector<F162> r7 = sample_f162(t, "labinius/switch-batch", 7);
    const std::vector<F162> batch = eq_f162_msb(r7.data(), 7);
    // psi[c][m]: sum over the bits b of m of batch[8c + b] (crossfield.rs:117-131)
    uint64_t *psi = h_psi;
    std::memset(psi, 0, 16 * 256 * 3 * sizeof(uint64_t));
    for (int c = 0; c < 16; c++)
        for (int m = 1; m < 256; m++) {
            const int b = __builtin_ctz(m);
            for (int k = 0; k < 3; k++)
                psi[3 * (c * 256 + m) + k] = psi[3 * (c * 256 + (m ^ (1 << b))) + k] ^ batch[8 * c + b].w[k];
        }
    HIP_TRY(hipMemcpyAsync(d_psi, psi, 16 * 256 * 3 * 8, hipMemcpyHostToDevice, st));
    HIP_TRY(launch_switch_psi(st, ws.d_eq, (uint32_t)n, d_psi, ws.d_a));
    // the sumcheck: a against the lifted trace (its third word read as zero), the folded trace in
    // d_p. Round i's fold runs fused with round i + 1's message; the last fold (two elements) is
    // the host's.
    std::vector<F162> r_pp(l);
    HIP_TRY(launch_fill_u32(st, (uint32_t *)d_msg, 12 * l, 0));
    HIP_TRY(launch_switch_msg(st, ws.d_a, wt->d_w, (uint32_t)(n / 2), 0, d_msg));
    size_t half = n / 2;
    for (size_t round = 0; round < l; round++) {
        const uint64_t *msg = h_msg;
        HIP_TRY(hipMemcpyAsync(h_msg, d_msg + 6 * round, 48, hipMemcpyDeviceToHost, st));
        HIP_TRY(stream_wait(st));
        std::memcpy(rounds[2 * round].w, msg, 24);
        std::memcpy(rounds[2 * round + 1].w, msg + 3, 24);
        absorb_f162(t, (const F162 *)msg, 2);
        r_pp[round] = sample_f162(t, "labinius/switch-round", 1)[0];
        if (round + 1 < l) {
            HIP_TRY(launch_switch_fold_msg(st, ws.d_a, round == 0 ? wt->d_w : ws.d_p, ws.d_p, (uint32_t)half,
                                           r_pp[round].w, round == 0 ? 0 : 1, d_msg + 6 * (round + 1)));
            half /= 2;
        }
    }
    // the last fold: p[0] + r (p[0] + p[1])
    F162 pl[2];
    HIP_TRY(hipMemcpyAsync(h_msg, ws.d_p, 48, hipMemcpyDeviceToHost, st));
    HIP_TRY(stream_wait(st));
    std::memcpy(pl, h_msg, 48);
    const F162 op = f162_add(pl[0], f162_mul(r_pp[l - 1], f162_add(pl[0], pl[1])));
    std::memcpy(opened->w, op.w, 24);
    msb_first(r_pp, cl, p0, p1);
    return LBN_OK;
}

int lbn_switch_verify(const lbn_params *params, const uint64_t *r_lo, size_t r_lo_len, const uint64_t *r_hi,
                      size_t r_hi_len, const uint64_t claim[2], const uint64_t *partials, const lbn_f162 *rounds,
                      const lbn_f162 *opened, lbn_transcript *tr, lbn_f162 *p0, lbn_f162 *p1) {
    if (!params || !r_hi || !claim || !partials || !rounds || !opened || !tr || !p0 || !p1 || (r_lo_len && !r_lo))
        return fail(LBN_ERR_ARG, "null argument");
    Shape s;
    int rc = make_shape(params, &s);
    if (rc) return rc;
    if ((rc = check_claim_shape(r_lo_len))) return rc;
    const size_t l = s.wl;
    if (r_hi_len != l) return fail(LBN_ERR_SHAPE, "r_hi has witness_log_len coordinates");
    Transcript &t = tr->t;
    const B128 *rl = (const B128 *)r_lo, *rh = (const B128 *)r_hi, *v = (const B128 *)partials;
    const B128 cb{{claim[0], claim[1]}};
    statement(t, rl, r_lo_len, rh, l, cb);
    // the claim from the partial evaluations: weights x^k (words) or eq(r_lo, k) (bits)
    std::vector<B128> wts(128, B128{{0, 0}});
    if (r_lo_len == 0) {
        for (int k = 0; k < 128; k++) wts[k].w[k / 64] = 1ull << (k % 64);
    } else {
        wts = eq_b128_msb(rl, 7);
    }
    B128 acc{{0, 0}};
    for (int k = 0; k < 128; k++) {
        const B128 pr = bmul(v[k], wts[k]);
        acc.w[0] ^= pr.w[0];
        acc.w[1] ^= pr.w[1];
    }
    if (acc.w[0] != cb.w[0] || acc.w[1] != cb.w[1])
        return fail(LBN_REJECTED, "the partial evaluations do not sum to the claim");
    absorb_b128(t, v, 128);
    const std::vector<F162> r7 = sample_f162(t, "labinius/switch-batch", 7);
    const std::vector<F162> batch = eq_f162_msb(r7.data(), 7);
    // slice_sum (crossfield.rs:478-490)
    F162 s_cur{{0, 0, 0}};
    for (int k = 0; k < 128; k++) {
        F162 u{{0, 0, 0}};
        for (int i = 0; i < 128; i++)
            if ((v[i].w[k / 64] >> (k % 64)) & 1) u.w[i / 64] |= 1ull << (i % 64);
        s_cur = f162_add(s_cur, f162_mul(u, batch[k]));
    }
    std::vector<F162> r_pp(l);
    for (size_t round = 0; round < l; round++) {
        const F162 e0{{rounds[2 * round].w[0], rounds[2 * round].w[1], rounds[2 * round].w[2]}};
        const F162 ei{{rounds[2 * round + 1].w[0], rounds[2 * round + 1].w[1], rounds[2 * round + 1].w[2]}};
        const F162 msg[2] = {e0, ei};
        absorb_f162(t, msg, 2);
        const F162 r = sample_f162(t, "labinius/switch-round", 1)[0];
        r_pp[round] = r;
        // s = e0 + r (e0 + e1 + einf) + r^2 einf with e1 = s + e0 (crossfield.rs:520-525)
        const F162 e1 = f162_add(s_cur, e0);
        s_cur = f162_add(f162_add(e0, f162_mul(r, f162_add(f162_add(e0, e1), ei))), f162_mul(f162_mul(r, r), ei));
    }
    const F162 d = transparent_coeff(rh, r_pp, batch);
    if (!(d.w[0] | d.w[1] | d.w[2])) return fail(LBN_REJECTED, "the switch left the opened value unbound");
    const F162 op{{opened->w[0], opened->w[1], opened->w[2]}};
    const F162 dv = f162_mul(d, op);
    if (dv.w[0] != s_cur.w[0] || dv.w[1] != s_cur.w[1] || dv.w[2] != s_cur.w[2])
        return fail(LBN_REJECTED, "sumcheck final check failed");
    msb_first(r_pp, s.cl, p0, p1);
    return LBN_OK;
}

}  // extern "C"
