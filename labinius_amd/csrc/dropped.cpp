// Bit-dropped commitments on the host: the prover's (top, digits) from the raw rows, their wire
// packing (what the transcript absorbs), and the verifier's residual check.
#include <algorithm>
#include <cmath>
#include <cstring>
#include <thread>
#include <vector>

#include "internal.hpp"
#include "launch.hpp"

namespace lbn {
namespace {

// wire/mod.rs:77-79
uint32_t residue_bits(uint32_t q) { return 32 - (uint32_t)__builtin_clz(q - 1); }

// bd/mod.rs:58-64
uint32_t top_bits(uint32_t q0, uint32_t d) {
    const uint32_t bound = ((q0 - 1) + (1u << (d - 1))) >> d;
    return 32 - (uint32_t)__builtin_clz(bound);
}

// bd/mod.rs:71-74 (f64, the same evaluation order)
uint64_t bd_cap(size_t cols, uint32_t d) {
    const double spread = (double)((1ull << (2 * d)) - 1) / 12.0;
    return (uint64_t)std::ceil(4.0 * (double)(N * cols * 28) * spread);
}

// Mixed-radix (Garner) digits over the primes: x = m_0 + q_0 (m_1 + q_1 (m_2 + ...)).
struct Garner {
    std::vector<uint64_t> q;
    std::vector<std::vector<uint64_t>> inv;  // inv[k][m] = q_m^-1 mod q_k, m < k
    explicit Garner(const std::vector<uint32_t> &primes) {
        for (uint32_t p : primes) q.push_back(p);
        inv.resize(q.size());
        for (size_t k = 0; k < q.size(); k++)
            for (size_t m = 0; m < k; m++) inv[k].push_back(inv_mod(q[m] % q[k], q[k]));
    }
    // residues r_k in [0, q_k) -> digits m_k
    void digits(const uint64_t *r, uint64_t *m) const {
        for (size_t k = 0; k < q.size(); k++) {
            uint64_t x = r[k];
            for (size_t j = 0; j < k; j++) x = (x + q[k] - m[j] % q[k]) % q[k] * inv[k][j] % q[k];
            m[k] = x;
        }
    }
};

int check_dropped(const Shape &s) {
    if (s.dropped_bits == 0) return fail(LBN_ERR_SHAPE, "the parameters do not drop bits");
    if (s.primes.size() < 2) return fail(LBN_ERR_SHAPE, "a bit-dropped commitment needs two or more moduli");
    if (s.dropped_bits >= 16) return fail(LBN_ERR_SHAPE, "dropped bits out of range");
    return LBN_OK;
}

// The packed (top, digits) of wire/mod.rs:386-397: per coefficient the top, then each digit,
// least significant bit first, the last byte zero-padded.
std::vector<uint8_t> pack(const Shape &s, const uint16_t *top, const uint16_t *digits) {
    const size_t count = s.r * N, limbs = s.primes.size();
    const uint32_t tb = top_bits(s.primes[0], s.dropped_bits);
    std::vector<uint32_t> bits(limbs);
    bits[0] = tb;
    size_t per = tb;
    for (size_t k = 1; k < limbs; k++) per += bits[k] = residue_bits(s.primes[k]);
    std::vector<uint8_t> out;
    out.reserve((count * per + 7) / 8);
    uint64_t acc = 0;
    uint32_t n = 0;
    for (size_t i = 0; i < count; i++)
        for (size_t k = 0; k < limbs; k++) {
            const uint64_t v = k == 0 ? top[i] : digits[(k - 1) * count + i];
            acc |= (v & ((1ull << bits[k]) - 1)) << n;
            n += bits[k];
            while (n >= 8) {
                out.push_back((uint8_t)acc);
                acc >>= 8;
                n -= 8;
            }
        }
    if (n > 0) out.push_back((uint8_t)acc);
    return out;
}

// The prover's bit-dropped scratch for n = limbs x coefficients (grow-only).
int bd_reserve(lbn_prover *pr, size_t n) {
    if (n <= pr->bd_cap) return LBN_OK;
    void *old[] = {pr->d_bd_raw, pr->d_bd_cen, pr->d_bd_out};
    for (void *p : old)
        if (p) (void)hipFree(p);
    pr->d_bd_raw = nullptr;
    pr->d_bd_cen = nullptr;
    pr->d_bd_out = nullptr;
    pr->bd_cap = 0;
    int rc;
    if ((rc = dalloc(&pr->d_bd_raw, n))) return rc;
    if ((rc = dalloc(&pr->d_bd_cen, n))) return rc;
    if ((rc = dalloc(&pr->d_bd_out, n))) return rc;
    pr->bd_cap = n;
    return LBN_OK;
}

}  // namespace

// bd/digits.rs:121-182 on the device: the raw rows d_raw ([limb][c][648], on the device) of a
// commitment of shape s -> every limb's inverse transform -> Garner -> top and digits in
// pr->d_bd_out ([c * 648 + i] top, then [limb - 1][c * 648 + i] digits), on the prover's stream.
int bd_digits_device(lbn_prover *pr, const Shape &s, const uint32_t *d_raw) {
    const lbn_pp *pp = pr->pp;
    const size_t limbs = s.primes.size(), count = s.r * N;
    if (limbs > (size_t)GARNER_MAX) return fail(LBN_ERR_SHAPE, "too many moduli for the device digits");
    if (pp->intt_limb.size() != limbs) return fail(LBN_ERR_SHAPE, "the moduli differ from the prover's");
    int rc = bd_reserve(pr, limbs * count);
    if (rc) return rc;
    hipStream_t st = pr->stream;
    for (size_t k = 0; k < limbs; k++)
        HIP_TRY(launch_intt(st, d_raw + k * count, (uint32_t)s.r, pp->intt_limb[k], pp->d_twi_limb[k],
                            pr->d_bd_cen + k * count));
    GarnerConsts gc{};
    gc.limbs = (uint32_t)limbs;
    gc.d = s.dropped_bits;
    gc.round = 1u << (s.dropped_bits - 1);
    for (size_t k = 0; k < limbs; k++) {
        gc.q[k] = s.primes[k];
        for (size_t j = 0; j < k; j++) gc.inv[k][j] = (uint32_t)inv_mod(s.primes[j] % s.primes[k], s.primes[k]);
    }
    HIP_TRY(launch_garner(st, pr->d_bd_cen, gc, (uint32_t)count, pr->d_bd_out, pr->d_bd_out + count));
    return LBN_OK;
}

namespace {

// bd_digits_device, then top and digits to the host
int drop_bits_device(lbn_prover *pr, const Shape &s, const uint32_t *d_raw, uint16_t *top, uint16_t *digits) {
    int rc = bd_digits_device(pr, s, d_raw);
    if (rc) return rc;
    const size_t limbs = s.primes.size(), count = s.r * N;
    hipStream_t st = pr->stream;
    HIP_TRY(hipMemcpyAsync(top, pr->d_bd_out, count * sizeof(uint16_t), hipMemcpyDeviceToHost, st));
    HIP_TRY(hipMemcpyAsync(digits, pr->d_bd_out + count, (limbs - 1) * count * sizeof(uint16_t),
                           hipMemcpyDeviceToHost, st));
    HIP_TRY(hipStreamSynchronize(st));
    return LBN_OK;
}

}  // namespace
}  // namespace lbn

using namespace lbn;

extern "C" {

static int drop_bits_host(const lbn_params *params, const uint32_t *raw, uint16_t *top, uint16_t *digits);

int lbn_drop_bits(const lbn_params *params, const uint32_t *raw, uint16_t *top, uint16_t *digits) {
    try {
        return drop_bits_host(params, raw, top, digits);
    } catch (const std::exception &) {
        return fail(LBN_ERR_OOM, "the digits' working memory");
    }
}

static int drop_bits_host(const lbn_params *params, const uint32_t *raw, uint16_t *top, uint16_t *digits) {
    if (!params || !raw || !top || !digits) return fail(LBN_ERR_ARG, "null argument");
    Shape s;
    int rc = make_shape(params, &s);
    if (rc) return rc;
    if ((rc = check_dropped(s))) return rc;
    const size_t limbs = s.primes.size(), count = s.r * N;
    const Garner g(s.primes);
    const uint32_t d = s.dropped_bits, round = 1u << (d - 1);
    std::vector<const Prime *> P(limbs);
    for (size_t k = 0; k < limbs; k++) P[k] = &prime(s.primes[k]);
    // columns are independent: a share of them per thread
    auto work = [&](size_t c0, size_t c1) {
        std::vector<uint64_t> v(N), r(limbs), m(limbs);
        std::vector<uint32_t> res(limbs * N);
        for (size_t c = c0; c < c1; c++) {
            // bd/digits.rs:121-132: the column's coefficients per limb, any representative mod q
            for (size_t k = 0; k < limbs; k++) {
                const uint32_t *y = raw + (k * s.r + c) * N;
                for (int u = 0; u < N; u++) v[u] = y[u];
                intt(*P[k], v.data());
                for (int u = 0; u < N; u++) res[k * N + u] = (uint32_t)(v[u] % P[k]->q);
            }
            // bd/digits.rs:134-182
            for (int u = 0; u < N; u++) {
                for (size_t k = 0; k < limbs; k++) r[k] = res[k * N + u];
                g.digits(r.data(), m.data());
                const size_t i = c * N + u;
                top[i] = (uint16_t)((m[0] + round) >> d);
                for (size_t k = 1; k < limbs; k++) digits[(k - 1) * count + i] = (uint16_t)m[k];
            }
        }
    };
    const size_t nt = std::min<size_t>(std::max<size_t>(1, s.r / 64), std::max(1u, std::thread::hardware_concurrency()));
    if (!run_ranges(s.r, nt, work)) return fail(LBN_ERR_OOM, "the digits' working memory");
    return LBN_OK;
}

int lbn_drop_bits_device(lbn_prover *pr, const lbn_params *params, const uint32_t *raw, uint16_t *top,
                         uint16_t *digits) {
    if (!pr || !params || !raw || !top || !digits) return fail(LBN_ERR_ARG, "null argument");
    Shape s;
    int rc = make_shape(params, &s);
    if (rc) return rc;
    if ((rc = check_dropped(s))) return rc;
    if (s.primes != pr->pp->shape.primes) return fail(LBN_ERR_SHAPE, "the moduli differ from the prover's");
    const size_t limbs = s.primes.size(), count = s.r * N;
    if (limbs > (size_t)GARNER_MAX) return fail(LBN_ERR_SHAPE, "too many moduli for the device digits");
    if ((rc = bd_reserve(pr, limbs * count))) return rc;
    HIP_TRY(hipMemcpyAsync(pr->d_bd_raw, raw, limbs * count * sizeof(uint32_t), hipMemcpyHostToDevice, pr->stream));
    pr->bd_ready = false;  // d_bd_out is about to hold the digits of these rows
    return drop_bits_device(pr, s, pr->d_bd_raw, top, digits);
}

int lbn_commit_drop_bits(lbn_prover *pr, uint16_t *top, uint16_t *digits) {
    if (!pr || !top || !digits) return fail(LBN_ERR_ARG, "null argument");
    const Shape &s = pr->pp->shape;
    int rc = check_dropped(s);
    if (rc) return rc;
    if (!pr->committed) return fail(LBN_ERR_ARG, "the prover has made no commitment");
    if ((rc = drop_bits_device(pr, s, pr->d_raw, top, digits))) return rc;
    pr->bd_ready = true;
    return LBN_OK;
}

int lbn_pack_dropped(const lbn_params *params, const uint16_t *top, const uint16_t *digits, uint8_t *out,
                     size_t cap, size_t *len) {
    if (!params || !top || !digits) return fail(LBN_ERR_ARG, "null argument");
    Shape s;
    int rc = make_shape(params, &s);
    if (rc) return rc;
    if ((rc = check_dropped(s))) return rc;
    const std::vector<uint8_t> b = pack(s, top, digits);
    if (len) *len = b.size();
    if (!out) return LBN_OK;
    if (cap < b.size()) return fail(LBN_ERR_SHAPE, "output buffer too small");
    std::memcpy(out, b.data(), b.size());
    return LBN_OK;
}

int lbn_derive_evaluation_point_dropped(const lbn_pp *pp, lbn_transcript *tr, const uint16_t *top,
                                        const uint16_t *digits, lbn_f162 *p0, lbn_f162 *p1) {
    if (!pp || !tr || !top || !digits || !p0 || !p1) return fail(LBN_ERR_ARG, "null argument");
    const Shape &s = pp->shape;
    int rc = check_dropped(s);
    if (rc) return rc;
    Transcript &t = tr->t;
    absorb_parameters(t, pp);
    t.absorb_bytes("labinius/commitment", 19);
    t.absorb_u64(s.r);
    t.absorb_bytes("labinius/dropped-commitment", 27);
    const std::vector<uint8_t> b = pack(s, top, digits);
    t.absorb_bytes(b.data(), b.size());
    derive_point(t, s, p0, p1);
    return LBN_OK;
}

int lbn_bd_residual(const lbn_pp *pp, const uint16_t *top, const uint16_t *digits, const lbn_challenge *ch,
                    size_t r, const int16_t *v, size_t len_ring, uint64_t *normsq_lo, uint64_t *normsq_hi) {
    if (!pp || !top || !digits || !ch || !v || !normsq_lo || !normsq_hi) return fail(LBN_ERR_ARG, "null argument");
    const Shape &s = pp->shape;
    int rc = check_dropped(s);
    if (rc) return rc;
    if (r != s.r || len_ring != s.len_ring) return fail(LBN_ERR_SHAPE, "one challenge per column, len_ring elements");
    const size_t limbs = s.primes.size(), count = s.r * N;
    std::vector<std::vector<uint32_t>> cen(limbs, std::vector<uint32_t>(N));
    std::vector<uint64_t> fc(N), y(N), x(N), cv(N), pr(N), av(N);
    for (size_t k = 0; k < limbs; k++) {
        const Prime &P = prime(s.primes[k]);
        const uint64_t q = P.q;
        // the column rebuilt from (top, digits) mod q (simd/bd.rs:5-37): (top << d) + q_0 (m_1 + q_1 (...))
        std::fill(fc.begin(), fc.end(), 0);
        for (size_t j = 0; j < s.r; j++) {
            for (int u = 0; u < N; u++) {
                const size_t i = j * N + u;
                uint64_t inner = digits[(limbs - 2) * count + i];
                for (int m = (int)limbs - 3; m >= 0; m--) inner = (digits[(size_t)m * count + i] + inner * s.primes[m + 1]) % q;
                x[u] = (((uint64_t)top[i] << s.dropped_bits) + inner * s.primes[0]) % q;
            }
            ntt(P, x.data());
            // c_j(-X^4) (fold.rs:287-300) and its transform
            std::fill(cv.begin(), cv.end(), 0);
            for (uint32_t e = 0; e < ch[j].weight; e++) {
                const uint32_t pos = ch[j].positions[e];
                const int sign = (((ch[j].signs >> e) & 1) ? -1 : 1) * ((pos & 1) ? -1 : 1);
                cv[4 * pos] = sign > 0 ? 1 : q - 1;
            }
            ntt(P, cv.data());
            if (P.quad) {
                mul_quad_slots(P, cv.data(), x.data(), pr.data());
                for (int u = 0; u < N; u++) fc[u] = P.md(fc[u] + pr[u]);
            } else {
                for (int u = 0; u < N; u++) fc[u] = P.md(fc[u] + cv[u] * x[u]);
            }
        }
        // A_k v
        std::fill(y.begin(), y.end(), 0);
        for (size_t i = 0; i < len_ring; i++) {
            for (int u = 0; u < N; u++) x[u] = reduce_signed(v[i * N + u], q);
            ntt(P, x.data());
            const int16_t *ak = pp->key.data() + (k * s.len_ring + i) * N;
            for (int u = 0; u < N; u++) av[u] = reduce_signed(ak[u], q);
            if (P.quad) {
                mul_quad_slots(P, av.data(), x.data(), pr.data());
                for (int u = 0; u < N; u++) y[u] = P.md(y[u] + pr[u]);
            } else {
                for (int u = 0; u < N; u++) y[u] = P.md(y[u] + av[u] * x[u]);
            }
        }
        for (int u = 0; u < N; u++) x[u] = (y[u] + q - fc[u]) % q;
        intt(P, x.data());
        for (int u = 0; u < N; u++) cen[k][u] = (uint32_t)(x[u] % q);
    }
    // Garner::normsq (bd/digits.rs): the CRT value centered mod prod q, squared, summed
    const Garner g(s.primes);
    unsigned __int128 modulus = 1, normsq = 0;
    for (uint32_t q : s.primes) modulus *= q;
    std::vector<uint64_t> rr(limbs), m(limbs);
    for (int i = 0; i < N; i++) {
        for (size_t k = 0; k < limbs; k++) rr[k] = cen[k][i];
        g.digits(rr.data(), m.data());
        unsigned __int128 value = 0;
        for (int k = (int)limbs - 1; k >= 0; k--) value = value * s.primes[k] + m[k];
        const unsigned __int128 mag = value > modulus / 2 ? modulus - value : value;
        normsq += mag * mag;
    }
    *normsq_lo = (uint64_t)normsq;
    *normsq_hi = (uint64_t)(normsq >> 64);
    return LBN_OK;
}

int lbn_verify_opening_bit_dropped(const lbn_pp *pp, const uint16_t *top, const uint16_t *digits,
                                   const lbn_challenge *ch, size_t r, const int16_t *v, size_t len_ring,
                                   const lbn_f162 *p0, size_t p0_len, const lbn_f162 *folded_row) {
    if (!pp || !top || !digits || !ch || !v || !p0 || !folded_row) return fail(LBN_ERR_ARG, "null argument");
    const Shape &s = pp->shape;
    int rc = check_dropped(s);
    if (rc) return rc;
    if (len_ring != s.len_ring || r != s.r || p0_len != s.rows) return LBN_REJECTED;
    // verifier.rs:252-265: the norm and range of v as in the Clear mode
    const int32_t half = (int32_t)((s.primes[0] - 1) / 2);
    uint64_t normsq = 0;
    int32_t worst = 0;
    for (size_t i = 0; i < len_ring * N; i++) {
        const int32_t x = v[i], ax = x < 0 ? -x : x;
        worst = ax > worst ? ax : worst;
        normsq += (uint64_t)((int64_t)x * x);
    }
    const uint64_t cap = (uint64_t)std::ceil(108.0 * (double)(s.n / 4 * N));
    if (worst > half || normsq > cap) return LBN_REJECTED;
    // verifier.rs:267-272
    uint64_t lo = 0, hi = 0;
    if ((rc = lbn_bd_residual(pp, top, digits, ch, r, v, len_ring, &lo, &hi))) return rc;
    if (hi != 0 || lo > bd_cap(s.r, s.dropped_bits)) return LBN_REJECTED;
    // verifier.rs:274, eval.rs:202-249
    std::vector<F162> ps(p0_len);
    for (size_t k = 0; k < p0_len; k++) std::memcpy(ps[k].w, p0[k].w, 24);
    const std::vector<F162> eq = eq_table(ps.data(), (int)p0_len);
    F162 acc{{0, 0, 0}};
    for (size_t m = 0; m < len_ring; m++)
        for (int k = 0; k < 4; k++) {
            F162 e{{0, 0, 0}};
            for (int b = 0; b < 162; b++)
                if (v[m * N + 4 * b + k] & 1) e.w[b >> 6] |= 1ull << (b & 63);
            acc = f162_add(acc, f162_mul(eq[4 * m + k], e));
        }
    F162 want;
    std::memcpy(want.w, folded_row->w, 24);
    return f162_eq(acc, want) ? LBN_OK : LBN_REJECTED;
}

uint64_t lbn_bd_cap(const lbn_params *params) {
    Shape s;
    if (!params || make_shape(params, &s) || s.dropped_bits == 0) return 0;
    return bd_cap(s.r, s.dropped_bits);
}

}  // extern "C"
