// The host side of the protocol: parameter checks, the seed expansions, the Fiat-Shamir
// transcript and its short-challenge sampler, and the verifier's checks. All of it stays on the
// CPU, as in the reference; only commit / fold / row evaluation run on the GPU.
#include <cstring>

#include <thread>

#include "internal.hpp"

namespace lbn {

static thread_local std::string g_last_error;

void set_error(const std::string &msg) { g_last_error = msg; }

int fail(int code, const std::string &msg) {
    g_last_error = msg;
    return code;
}

// Params::with_base's checks (scheme/params.rs:172-203) plus the modulus set (ring/mod.rs:87-95).
int make_shape(const lbn_params *p, Shape *s) {
    if (!p) return fail(LBN_ERR_ARG, "null params");
    if (p->column_log_len > p->witness_log_len)
        return fail(LBN_ERR_COLUMNS_EXCEED_WITNESS, "column_log_len exceeds witness_log_len");
    if (p->column_log_len < 1) return fail(LBN_ERR_TOO_FEW_COLUMNS, "column_log_len must be at least 1");
    if (p->witness_log_len - p->column_log_len < 7)
        return fail(LBN_ERR_COLUMN_TOO_SHORT, "a column must hold at least 2^7 F162 elements");
    if (!is_modulus(p->base)) return fail(LBN_ERR_UNKNOWN_MODULUS, "unknown base modulus");
    if (p->n_extra > 6) return fail(LBN_ERR_ARG, "too many extra moduli");
    for (uint32_t i = 0; i < p->n_extra; i++) {
        if (!is_modulus(p->extra[i])) return fail(LBN_ERR_UNKNOWN_MODULUS, "unknown extra modulus");
        if (p->extra[i] == p->base)
            return fail(LBN_ERR_BASE_IS_ALSO_EXTRA, "the base modulus is listed again as an extra one");
        for (uint32_t k = 0; k < i; k++)
            if (p->extra[k] == p->extra[i])
                return fail(LBN_ERR_DUPLICATE_MODULUS, "a modulus is listed twice");
    }
    if (p->witness_log_len > 30) return fail(LBN_ERR_ARG, "witness_log_len above 30");
    s->wl = p->witness_log_len;
    s->cl = p->column_log_len;
    s->rows = s->wl - s->cl;
    s->n = size_t(1) << s->wl;
    s->r = size_t(1) << s->cl;
    s->column_len = size_t(1) << s->rows;
    s->len_ring = s->column_len / 4;
    s->primes.assign(1, p->base);
    for (uint32_t i = 0; i < p->n_extra; i++) s->primes.push_back(p->extra[i]);
    s->opening = p->opening;
    s->dropped_bits = p->opening == LBN_OPENING_BIT_DROPPED ? p->dropped_bits : 0;
    // bd/digits.rs:141-182 keeps digits and the rounded top in u16 and needs a second modulus for the
    // digits: every entry point rejects other bit-dropped shapes (a reference assert!)
    if (p->opening == LBN_OPENING_BIT_DROPPED && (s->dropped_bits < 1 || s->dropped_bits > 15))
        return fail(LBN_ERR_SHAPE, "dropped bits out of range (1 to 15)");
    if (p->opening == LBN_OPENING_BIT_DROPPED && s->primes.size() < 2)
        return fail(LBN_ERR_SHAPE, "a bit-dropped commitment needs two or more moduli");
    return LBN_OK;
}

std::vector<int16_t> expand_key(const std::vector<uint32_t> &primes, size_t len_ring,
                                const uint8_t seed[32]) {
    // scheme/params.rs:321-322: the u64 seed is the first 8 bytes of BLAKE3(matrix_seed)
    uint8_t d[32];
    Blake3 h = Blake3::hasher();
    h.update(seed, 32);
    h.finalize().read(0, d, 32);
    uint64_t s64;
    std::memcpy(&s64, d, 8);
    std::vector<int16_t> a(primes.size() * len_ring * N);
    for (size_t k = 0; k < primes.size(); k++) {
        const uint32_t q = primes[k];
        const int16_t half = (int16_t)((q - 1) / 2);
        Rng rng(s64 ^ (0x9E3779B9ull * (uint64_t)(k + 1)));
        int16_t *ak = a.data() + k * len_ring * N;
        for (size_t b = 0; b < len_ring / 32; b++)
            for (int j = 0; j < N; j++)
                for (int p = 0; p < 32; p++) ak[(32 * b + p) * N + j] = (int16_t)((int16_t)rng.below(q) - half);
    }
    return a;
}

Blake3::Output Transcript::reader(const char *label, size_t label_len) {
    Blake3 c = state;
    c.update_u64(counter);
    c.update_u64(label_len);
    c.update(label, label_len);
    counter++;
    return c.finalize();
}

// ---------------------------------------------------------------------------------------------
// the short-challenge sampler (challenge.rs:163-221, 259-269, 332-419, 478-533)
// ---------------------------------------------------------------------------------------------

namespace {

struct ByteStream {
    Xof x;
    uint8_t buf[256];
    int pos = 256;
    explicit ByteStream(const Blake3::Output &o) : x(o) {}
    uint8_t byte() {
        if (pos == 256) {
            x.fill(buf, 256);
            pos = 0;
        }
        return buf[pos++];
    }
    uint16_t u16() {
        if (pos + 2 <= 256) {
            uint16_t v = (uint16_t)(buf[pos] | buf[pos + 1] << 8);
            pos += 2;
            return v;
        }
        uint8_t lo = byte(), hi = byte();
        return (uint16_t)(lo | hi << 8);
    }
    // uniform below n by rejection on u16 (challenge.rs:198-221)
    uint16_t below(uint32_t n) {
        const uint32_t limit = (65536u / n) * n;
        const uint64_t magic = (1ull << 32) / n + 1;
        for (;;) {
            uint32_t r = u16();
            if (r < limit) return (uint16_t)(r - (uint32_t)(((uint64_t)r * magic) >> 32) * n);
        }
    }
};

struct Phases {
    double re[N162][81], im[N162][81];
    Phases() {
        int units[N162], n = 0;
        for (int x = 1; x < 243; x++)
            if (x % 3) units[n++] = x;
        for (int p = 0; p < N162; p++)
            for (int k = 0; k < 81; k++) {
                const int m = (p * units[k]) % 243;
                const double angle = 2.0 * 3.141592653589793 * (double)m / 243.0;
                re[p][k] = std::cos(angle);
                im[p][k] = std::sin(angle);
            }
    }
};

const Phases &phases() {
    static const Phases ph;
    return ph;
}

// every one of the first 81 roots has |c(zeta^u)|^2 <= bound_sq; terms added in position order
bool within(const lbn_challenge &c, double bound_sq) {
    const Phases &ph = phases();
    for (int k = 0; k < 81; k++) {
        double re = 0.0, im = 0.0;
        for (uint32_t i = 0; i < c.weight; i++) {
            const int p = c.positions[i];
            const bool neg = (c.signs >> i) & 1;
            re = re + (neg ? -ph.re[p][k] : ph.re[p][k]);
            im = im + (neg ? -ph.im[p][k] : ph.im[p][k]);
        }
        const double r2 = re * re;
        const double i2 = im * im;
        if (r2 + i2 > bound_sq) return false;
    }
    return true;
}

void sign(lbn_challenge &c) {
    static const Blake3 key = Blake3::derive_key("labinius 2026 challenge signs v1");
    Blake3 h = key;
    h.update_u64(c.weight);
    h.update(c.positions, c.weight);
    uint8_t d[4];
    h.finalize().read(0, d, 4);
    const uint32_t bits = (uint32_t)d[0] | (uint32_t)d[1] << 8 | (uint32_t)d[2] << 16 | (uint32_t)d[3] << 24;
    c.signs = bits & (uint32_t)((1ull << c.weight) - 1);
}

void attempt(ByteStream &x, uint32_t weight, lbn_challenge &out) {
    uint8_t perm[N162];
    for (int i = 0; i < N162; i++) perm[i] = (uint8_t)i;
    std::memset(&out, 0, sizeof out);
    out.weight = weight;
    for (uint32_t i = 0; i < weight; i++) {
        const uint32_t j = i + x.below(N162 - i);
        std::swap(perm[i], perm[j]);
    }
    bool seen[N162] = {};
    for (uint32_t i = 0; i < weight; i++) seen[perm[i]] = true;
    int n = 0;
    for (int p = 0; p < N162; p++)
        if (seen[p]) out.positions[n++] = (uint8_t)p;
}

}  // namespace

namespace {

uint64_t sample_from(Blake3::Output reader, uint32_t weight, double bound, lbn_challenge *out) {
    const double bound_sq = bound * bound + 1e-12;
    ByteStream x(reader);
    for (uint64_t attempts = 1;; attempts++) {
        attempt(x, weight, *out);
        sign(*out);
        if (within(*out, bound_sq)) return attempts;
    }
}

}  // namespace

uint64_t sample_short_challenge(Transcript &t, uint32_t weight, double bound, lbn_challenge *out) {
    return sample_from(t.reader("short-challenge", 15), weight, bound, out);
}

// n challenges in a row: challenge j reads its own stream (the transcript's state, its counter
// value, the label), so the streams are taken in order and the rejection sampling runs on threads.
void sample_short_challenges(Transcript &t, size_t n, uint32_t weight, double bound, lbn_challenge *out) {
    std::vector<Blake3::Output> readers;
    readers.reserve(n);
    for (size_t j = 0; j < n; j++) readers.push_back(t.reader("short-challenge", 15));
    const size_t nt = std::min<size_t>(std::max<size_t>(1, n / 16), std::max(1u, std::thread::hardware_concurrency()));
    auto work = [&](size_t j0, size_t j1) {
        for (size_t j = j0; j < j1; j++) sample_from(readers[j], weight, bound, &out[j]);
    };
    (void)run_ranges(n, nt, work);  // sampling allocates nothing: no share can throw
}

// ---------------------------------------------------------------------------------------------
// the verifier's pieces
// ---------------------------------------------------------------------------------------------

// verifier.rs:85-99: one F162 per variable from one read, limb 2 masked to 34 bits
void derive_point(Transcript &t, const Shape &s, lbn_f162 *p0, lbn_f162 *p1) {
    const size_t nv = s.wl;
    std::vector<uint8_t> bytes(24 * nv);
    t.fill("labinius/evaluation-point", 25, bytes.data(), bytes.size());
    for (size_t e = 0; e < nv; e++) {
        lbn_f162 x;
        std::memcpy(x.w, bytes.data() + 24 * e, 24);
        x.w[2] &= (1ull << 34) - 1;
        if (e < s.rows) p0[e] = x;
        else p1[e - s.rows] = x;
    }
}

void absorb_parameters(Transcript &t, const lbn_pp *pp) {
    const Shape &s = pp->shape;
    t.absorb_bytes("labinius/parameters", 19);
    t.absorb_u64(s.wl);
    t.absorb_u64(s.cl);
    for (uint32_t q : s.primes) t.absorb_u64(q);
    t.absorb_u64(s.opening == LBN_OPENING_RECURSIVE ? 1 : 0);
    if (s.dropped_bits > 0) {
        t.absorb_bytes("labinius/dropped-bits", 21);
        t.absorb_u64(s.dropped_bits);
    }
    t.absorb_bytes(pp->seed, 32);
}

static F162 to_f162(const lbn_challenge &c) {
    F162 x{{0, 0, 0}};
    for (uint32_t i = 0; i < c.weight; i++) x.w[c.positions[i] >> 6] |= 1ull << (c.positions[i] & 63);
    return x;
}

}  // namespace lbn

using namespace lbn;

extern "C" {

const char *lbn_last_error(void) { return lbn::g_last_error.c_str(); }

int lbn_params_check(const lbn_params *params) {
    Shape s;
    return make_shape(params, &s);
}

int lbn_witness_random(const lbn_params *params, const uint8_t seed[32], lbn_f162 *out, size_t n) {
    Shape s;
    int rc = make_shape(params, &s);
    if (rc) return rc;
    if (n != s.n) return fail(LBN_ERR_WITNESS_LENGTH, "the witness is not witness_len() elements");
    // scheme/witness.rs:66-87: a keyed BLAKE3 XOF of "labinius/witness", 24 bytes per element
    Blake3 h = Blake3::keyed(seed);
    h.update("labinius/witness", 16);
    Blake3::Output o = h.finalize();
    std::vector<uint8_t> buf(24 * 4096);
    for (size_t i0 = 0; i0 < n; i0 += 4096) {
        const size_t cnt = std::min<size_t>(4096, n - i0);
        o.read(24 * i0, buf.data(), 24 * cnt);
        for (size_t i = 0; i < cnt; i++) {
            std::memcpy(out[i0 + i].w, buf.data() + 24 * i, 24);
            out[i0 + i].w[2] &= (1ull << 34) - 1;
        }
    }
    return LBN_OK;
}

lbn_transcript *lbn_transcript_new(const uint8_t *domain, size_t len) {
    lbn_transcript *t = new lbn_transcript();
    t->t.state.update_u64(len);
    t->t.state.update(domain, len);
    return t;
}

lbn_transcript *lbn_transcript_clone(const lbn_transcript *t) { return new lbn_transcript(*t); }
void lbn_transcript_free(lbn_transcript *t) { delete t; }
void lbn_transcript_absorb_bytes(lbn_transcript *t, const uint8_t *b, size_t len) { t->t.absorb_bytes(b, len); }
void lbn_transcript_absorb_u64(lbn_transcript *t, uint64_t x) { t->t.absorb_u64(x); }
void lbn_transcript_fill(lbn_transcript *t, const uint8_t *label, size_t label_len, uint8_t *out, size_t len) {
    t->t.fill((const char *)label, label_len, out, len);
}

uint64_t lbn_sample_short_challenge(lbn_transcript *t, uint32_t weight, double bound, lbn_challenge *out) {
    if (weight > 32 || weight > N162) return 0;
    return sample_short_challenge(t->t, weight, bound, out);
}

int lbn_derive_evaluation_point(const lbn_pp *pp, lbn_transcript *tr, const int16_t *comp, lbn_f162 *p0,
                                lbn_f162 *p1) {
    if (!pp || !tr || !comp) return fail(LBN_ERR_ARG, "null argument");
    const Shape &s = pp->shape;
    if (s.dropped_bits > 0)
        return fail(LBN_ERR_SHAPE, "a bit-dropped key commits to (top, digits): lbn_derive_evaluation_point_dropped");
    Transcript &t = tr->t;
    absorb_parameters(t, pp);
    t.absorb_bytes("labinius/commitment", 19);
    t.absorb_u64(s.r);
    // absorb_elements per column: the count 4, then the column's 4 entries with no per-element
    // prefix, as one contiguous stream, so that the hasher takes whole subtrees on threads
    const size_t limbs = s.primes.size(), per = 2 * N162 * limbs, col = 8 + 4 * per;
    std::vector<uint8_t> buf;
    try {
        buf.resize(s.r * col);
    } catch (const std::exception &) {
        return fail(LBN_ERR_OOM, "the commitment's transcript bytes");
    }
    for (size_t c = 0; c < s.r; c++) {
        uint8_t *d = buf.data() + c * col;
        const uint64_t four = 4;
        std::memcpy(d, &four, 8);  // little-endian u64
        std::memcpy(d + 8, comp + c * 4 * limbs * N162, 4 * per);  // little-endian i16, rows 0..3
    }
    t.state.update(buf.data(), buf.size());
    derive_point(t, s, p0, p1);
    return LBN_OK;
}

int lbn_derive_folding_challenges(const lbn_pp *pp, lbn_transcript *tr, const lbn_f162 *u, size_t r,
                                  lbn_challenge *out) {
    if (!pp || !tr || !u || !out) return fail(LBN_ERR_ARG, "null argument");
    if (r != pp->shape.r) return fail(LBN_ERR_SHAPE, "one row value per column");
    Transcript &t = tr->t;
    t.absorb_bytes("labinius/row-evaluation", 23);
    t.absorb_bytes(u, 24 * r);
    try {
        sample_short_challenges(t, r, 28, 12.0, out);
    } catch (const std::exception &) {  // the readers' list; the transcript has absorbed u
        return fail(LBN_ERR_OOM, "the challenges' streams");
    }
    return LBN_OK;
}

int lbn_fold_commitment(const lbn_pp *pp, const int16_t *comp, const lbn_challenge *ch, size_t r, int16_t *out) {
    if (!pp || !comp || !ch || !out) return fail(LBN_ERR_ARG, "null argument");
    const Shape &s = pp->shape;
    if (r != s.r) return fail(LBN_ERR_SHAPE, "one challenge per column");
    const Tables &tb = tables();
    const size_t limbs = s.primes.size();
    for (size_t k = 0; k < limbs; k++) {
        const Prime &P = prime(s.primes[k]);
        const uint64_t q = P.q;
        const uint32_t cond = P.quad ? CONDUCTOR_QUAD : CONDUCTOR, mult = P.quad ? 2 : 4;
        std::vector<int64_t> acc(4 * N162, 0);
        for (size_t j = 0; j < r; j++) {
            for (int sl = 0; sl < N162; sl++) {
                // c_j(-theta^{v_s}) (fold.rs:799-873 through the unit-challenge slot table)
                const uint64_t v = tb.pow3_slot_exp[sl];
                int64_t cv = 0;
                for (uint32_t i = 0; i < ch[j].weight; i++) {
                    const uint32_t p = ch[j].positions[i];
                    const int co = (1 - 2 * (int)((ch[j].signs >> i) & 1)) * ((p & 1) ? -1 : 1);
                    cv += co * (int64_t)P.pw[(mult * v * p) % cond];
                }
                const int64_t cs = center(reduce_signed(cv, q), q);
                for (int row = 0; row < 4; row++) {
                    int64_t &a = acc[row * N162 + sl];
                    a = (a + cs * comp[((j * 4 + row) * limbs + k) * N162 + sl]) % (int64_t)q;
                }
            }
        }
        for (int row = 0; row < 4; row++)
            for (int sl = 0; sl < N162; sl++)
                out[(row * limbs + k) * N162 + sl] = center(reduce_signed(acc[row * N162 + sl], q), q);
    }
    return LBN_OK;
}

int lbn_fold_row_evaluation(const lbn_f162 *u, const lbn_challenge *ch, size_t r, lbn_f162 *out) {
    if (!u || !ch || !out) return fail(LBN_ERR_ARG, "null argument");
    F162 acc{{0, 0, 0}};
    for (size_t j = 0; j < r; j++) {
        F162 uj;
        std::memcpy(uj.w, u[j].w, 24);
        acc = f162_add(acc, f162_mul(to_f162(ch[j]), uj));
    }
    std::memcpy(out->w, acc.w, 24);
    return LBN_OK;
}

static F162 claim(const lbn_f162 *u, const lbn_f162 *p1, size_t cl) {
    std::vector<F162> ps(cl);
    for (size_t k = 0; k < cl; k++) std::memcpy(ps[k].w, p1[k].w, 24);
    std::vector<F162> eq = eq_table(ps.data(), (int)cl);
    F162 acc{{0, 0, 0}};
    for (size_t j = 0; j < eq.size(); j++) {
        F162 uj;
        std::memcpy(uj.w, u[j].w, 24);
        acc = f162_add(acc, f162_mul(eq[j], uj));
    }
    return acc;
}

int lbn_verify_evaluation(const lbn_pp *pp, const lbn_f162 *p1, size_t p1_len, const lbn_f162 *claimed,
                          const lbn_f162 *u, size_t r) {
    if (!pp || !p1 || !claimed || !u) return fail(LBN_ERR_ARG, "null argument");
    if (r != pp->shape.r || p1_len != pp->shape.cl) return LBN_REJECTED;
    F162 t = claim(u, p1, p1_len);
    F162 c;
    std::memcpy(c.w, claimed->w, 24);
    return f162_eq(t, c) ? LBN_OK : LBN_REJECTED;
}

int lbn_verify_opening_clear(const lbn_pp *pp, const int16_t *fc, const int16_t *v, size_t len_ring,
                             const lbn_f162 *p0, size_t p0_len, const lbn_f162 *folded_row) {
    if (!pp || !fc || !v || !p0 || !folded_row) return fail(LBN_ERR_ARG, "null argument");
    const Shape &s = pp->shape;
    if (len_ring != s.len_ring || p0_len != s.rows) return LBN_REJECTED;
    // verifier.rs:186-203: centered modulo the base, and the squared norm under the fold cap
    const int32_t half = (int32_t)((s.primes[0] - 1) / 2);
    uint64_t normsq = 0;
    int32_t worst = 0;
    for (size_t i = 0; i < len_ring * N; i++) {
        const int32_t x = v[i], ax = x < 0 ? -x : x;
        worst = ax > worst ? ax : worst;
        normsq += (uint64_t)((int64_t)x * x);
    }
    const uint64_t cap = (uint64_t)std::ceil(108.0 * (double)(s.n / 4 * N));  // scheme/params.rs:263-265
    if (worst > half || normsq > cap) return LBN_REJECTED;
    // verifier.rs:205-224: A v per limb, decomposed, equals the folded commitment
    const size_t limbs = s.primes.size();
    std::vector<uint64_t> x(N), av(N), pr(N), y(N);
    std::vector<uint32_t> y32(N);
    std::vector<int16_t> d(4 * N162);
    for (size_t k = 0; k < limbs; k++) {
        const Prime &P = prime(s.primes[k]);
        const uint64_t q = P.q;
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
        for (int u = 0; u < N; u++) y32[u] = (uint32_t)y[u];
        decompose(P, y32.data(), d.data());
        for (int row = 0; row < 4; row++)
            for (int sl = 0; sl < N162; sl++)
                if (d[row * N162 + sl] != fc[(row * limbs + k) * N162 + sl]) return LBN_REJECTED;
    }
    // eval.rs:202-249: eq(p0) . (v mod 2) == u', element 4m + k = component k of v_m
    std::vector<F162> ps(p0_len);
    for (size_t k = 0; k < p0_len; k++) std::memcpy(ps[k].w, p0[k].w, 24);
    std::vector<F162> eq = eq_table(ps.data(), (int)p0_len);
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

int lbn_blake3(int mode, const uint8_t *key, const uint8_t *in, size_t len, uint8_t *out, size_t out_len) {
    Blake3 h = mode == 0 ? Blake3::hasher()
               : mode == 1 ? Blake3::keyed(key)
                           : Blake3::derive_key(std::string((const char *)key));
    h.update(in, len);
    h.finalize().read(0, out, out_len);
    return LBN_OK;
}

void lbn_f162_mul(const lbn_f162 *a, const lbn_f162 *b, lbn_f162 *out) {
    F162 x, y;
    std::memcpy(x.w, a->w, 24);
    std::memcpy(y.w, b->w, 24);
    F162 z = f162_mul(x, y);
    std::memcpy(out->w, z.w, 24);
}

}  // extern "C"
