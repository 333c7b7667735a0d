// The verifier's heavy steps on the device: A v of the folded witness checked against the folded
// commitment, with its norm and its row check (verifier.rs:186-224, eval.rs:202-249), and the fold
// of a commitment (fold.rs:799-873). A v runs through the commit path itself: v + half splits into
// binary planes u_b, each a column the commit GEMM takes as it takes a witness column, and
// A v = sum_b 2^b A u_b - half A J (J the all-ones ring element) holds after the decomposition too,
// which is linear. The results equal lbn_verify_opening_clear's and lbn_fold_commitment's.
#include <cmath>
#include <cstring>
#include <vector>

#include "internal.hpp"
#include "launch.hpp"

namespace lbn {

void verify_free(lbn_verify_ws &w) {
    void *bufs[] = {w.d_v,    w.d_w,    w.d_bytes, w.d_yacc, w.d_raw, w.d_comp, w.d_u,
                    w.d_stat, w.d_fcomp, w.d_fout, w.d_facc, w.d_pw,  w.d_p3,   w.d_ch};
    for (void *b : bufs)
        if (b) (void)hipFree(b);
    w = lbn_verify_ws{};
}

namespace {

constexpr size_t MAX_PLANES = 15;               // q < 2^15
constexpr size_t MAX_COLS = MAX_PLANES + 2;     // the planes, J, v mod 2

uint32_t plane_bits(uint32_t q) { return 32 - (uint32_t)__builtin_clz(q - 1); }  // u = v + half < q

FoldCommConsts fold_consts(const Shape &s, std::vector<uint32_t> *pw) {
    FoldCommConsts fc{};
    fc.limbs = (uint32_t)s.primes.size();
    uint32_t off = 0;
    for (size_t k = 0; k < s.primes.size(); k++) {
        const Prime &P = prime(s.primes[k]);
        fc.q[k] = P.q;
        fc.cond[k] = P.quad ? CONDUCTOR_QUAD : CONDUCTOR;
        fc.mult[k] = P.quad ? 2 : 4;
        fc.pw_off[k] = off;
        off += fc.cond[k];
        if (pw) pw->insert(pw->end(), P.pw.begin(), P.pw.begin() + fc.cond[k]);
    }
    return fc;
}

int verify_alloc(lbn_prover *pr) {
    lbn_verify_ws &w = pr->vfy;
    const lbn_pp *pp = pr->pp;
    const Shape &s = pp->shape;
    const size_t limbs = s.primes.size();
    if (limbs > (size_t)GARNER_MAX) return fail(LBN_ERR_SHAPE, "too many moduli");
    int rc;
    const size_t nbytes = bytes_size((uint32_t)MAX_COLS, pp->kpad_c);
    if ((rc = dalloc(&w.d_v, s.len_ring * N))) return rc;
    if ((rc = dalloc(&w.d_w, MAX_COLS * s.column_len * 3))) return rc;
    if ((rc = dalloc(&w.d_bytes, nbytes))) return rc;
    HIP_TRY(hipMemset(w.d_bytes, 0, nbytes));  // the layout's padding must read as zero
    if ((rc = dalloc(&w.d_yacc, MAX_COLS * YACC_PLANES * limbs * N))) return rc;
    if ((rc = dalloc(&w.d_raw, limbs * MAX_COLS * N))) return rc;
    if ((rc = dalloc(&w.d_comp, MAX_COLS * 4 * limbs * N162))) return rc;
    if ((rc = dalloc(&w.d_u, MAX_COLS * 3))) return rc;
    if ((rc = dalloc(&w.d_stat, 2))) return rc;
    if ((rc = dalloc(&w.d_fout, 4 * limbs * N162))) return rc;
    if ((rc = dalloc(&w.d_facc, 4 * limbs * N162))) return rc;
    std::vector<uint32_t> pw;
    (void)fold_consts(s, &pw);
    if ((rc = dalloc(&w.d_pw, pw.size()))) return rc;
    if ((rc = dalloc(&w.d_p3, N162))) return rc;
    HIP_TRY(hipMemcpy(w.d_pw, pw.data(), pw.size() * sizeof(uint32_t), hipMemcpyHostToDevice));
    HIP_TRY(hipMemcpy(w.d_p3, tables().pow3_slot_exp, N162 * sizeof(uint16_t), hipMemcpyHostToDevice));
    w.ready = true;
    return LBN_OK;
}

// The verifier's scratch, allocated at its first use (all of it, or none).
int verify_ready(lbn_prover *pr) {
    if (pr->vfy.ready) return LBN_OK;
    const int rc = verify_alloc(pr);
    if (rc) verify_free(pr->vfy);
    return rc;
}

}  // namespace
}  // namespace lbn

using namespace lbn;

extern "C" {

int lbn_verify_opening_clear_device(lbn_prover *pr, const int16_t *fc, const int16_t *v, size_t len_ring,
                                    const lbn_f162 *p0, size_t p0_len, const lbn_f162 *folded_row) {
    if (!pr || !fc || !v || !p0 || !folded_row) return fail(LBN_ERR_ARG, "null argument");
    const Shape &s = pr->pp->shape;
    if (len_ring != s.len_ring || p0_len != s.rows) return LBN_REJECTED;
    int rc = verify_ready(pr);
    if (rc) return rc;
    lbn_verify_ws &w = pr->vfy;
    hipStream_t st = pr->stream;
    const size_t limbs = s.primes.size();
    const uint32_t q0 = s.primes[0], half = (q0 - 1) / 2, nb = plane_bits(q0);
    const size_t ncol = nb + 2;
    HIP_TRY(hipMemcpyAsync(w.d_v, v, len_ring * N * sizeof(int16_t), hipMemcpyHostToDevice, st));
    HIP_TRY(hipMemcpyAsync(pr->d_p0, p0, s.rows * sizeof(lbn_f162), hipMemcpyHostToDevice, st));
    HIP_TRY(launch_fill_u32(st, (uint32_t *)w.d_stat, 4, 0));
    HIP_TRY(launch_verify_planes(st, w.d_v, (uint32_t)len_ring, (int)half, nb, w.d_w,
                                 (unsigned long long *)w.d_stat));
    if ((rc = commit_columns_device(pr, w.d_w, ncol, w.d_bytes, w.d_yacc, w.d_raw, w.d_comp))) return rc;
    HIP_TRY(launch_eq(st, pr->d_p0, s.rows, pr->d_eqpart, pr->d_eq));
    HIP_TRY(launch_fill_u32(st, (uint32_t *)w.d_u, ncol * 6, 0));
    HIP_TRY(launch_row_eval_f4(st, w.d_w, pr->d_eq, pr->d_eqf, (uint32_t)s.column_len, (uint32_t)ncol, w.d_u));
    std::vector<int16_t> comp(ncol * 4 * limbs * N162);
    uint64_t stat[2], u[3 * MAX_COLS];
    HIP_TRY(hipMemcpyAsync(comp.data(), w.d_comp, comp.size() * sizeof(int16_t), hipMemcpyDeviceToHost, st));
    HIP_TRY(hipMemcpyAsync(stat, w.d_stat, sizeof(stat), hipMemcpyDeviceToHost, st));
    HIP_TRY(hipMemcpyAsync(u, w.d_u, ncol * 3 * sizeof(uint64_t), hipMemcpyDeviceToHost, st));
    HIP_TRY(hipStreamSynchronize(st));
    // verifier.rs:186-203: centered modulo the base, and the squared norm under the fold cap
    const uint64_t cap = (uint64_t)std::ceil(108.0 * (double)(s.n / 4 * N));  // scheme/params.rs:263-265
    if (stat[1] > half || stat[0] > cap) return LBN_REJECTED;
    // verifier.rs:205-224: A v = sum_b 2^b A u_b - half A J, decomposed, equals the folded commitment
    for (size_t t = 0; t < 4; t++)
        for (size_t k = 0; k < limbs; k++) {
            const int64_t q = s.primes[k];
            for (int sl = 0; sl < N162; sl++) {
                auto at = [&](size_t c) { return (int64_t)comp[((c * 4 + t) * limbs + k) * N162 + sl]; };
                int64_t a = -(int64_t)half * at(nb);
                for (uint32_t b = 0; b < nb; b++) a += at(b) * ((int64_t)1 << b);
                a %= q;
                if (a < 0) a += q;
                if (a > (q - 1) / 2) a -= q;
                if (a != fc[(t * limbs + k) * N162 + sl]) return LBN_REJECTED;
            }
        }
    // eval.rs:202-249: eq(p0) . (v mod 2), the row value of the last column, equals u'
    const uint64_t *par = u + 3 * (ncol - 1);
    return std::memcmp(par, folded_row->w, 24) == 0 ? LBN_OK : LBN_REJECTED;
}

int lbn_fold_commitment_device(lbn_prover *pr, const int16_t *comp, const lbn_challenge *ch, size_t r, int16_t *out) {
    if (!pr || !comp || !ch || !out) return fail(LBN_ERR_ARG, "null argument");
    const Shape &s = pr->pp->shape;
    if (r != s.r) return fail(LBN_ERR_SHAPE, "one challenge per column");
    int rc = verify_ready(pr);
    if (rc) return rc;
    lbn_verify_ws &w = pr->vfy;
    hipStream_t st = pr->stream;
    const size_t limbs = s.primes.size(), n = r * 4 * limbs * N162;
    if (n > w.cap_fcomp) {
        if (w.d_fcomp) (void)hipFree(w.d_fcomp);
        w.d_fcomp = nullptr;
        w.cap_fcomp = 0;
        if ((rc = dalloc(&w.d_fcomp, n))) return rc;
        w.cap_fcomp = n;
    }
    if (r > w.cap_ch) {
        if (w.d_ch) (void)hipFree(w.d_ch);
        w.d_ch = nullptr;
        w.cap_ch = 0;
        if ((rc = dalloc(&w.d_ch, r))) return rc;
        w.cap_ch = r;
    }
    HIP_TRY(hipMemcpyAsync(w.d_fcomp, comp, n * sizeof(int16_t), hipMemcpyHostToDevice, st));
    HIP_TRY(hipMemcpyAsync(w.d_ch, ch, r * sizeof(lbn_challenge), hipMemcpyHostToDevice, st));
    HIP_TRY(launch_fold_comm(st, (const lbn_challenge_dev *)w.d_ch, (uint32_t)r, w.d_fcomp, fold_consts(s, nullptr),
                             w.d_pw, w.d_p3, (long long *)w.d_facc, w.d_fout));
    HIP_TRY(hipMemcpyAsync(out, w.d_fout, 4 * limbs * N162 * sizeof(int16_t), hipMemcpyDeviceToHost, st));
    HIP_TRY(hipStreamSynchronize(st));
    return LBN_OK;
}

}  // extern "C"
