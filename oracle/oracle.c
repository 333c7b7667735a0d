/*
 * The oracle: a plain-C, single-threaded CPU restatement of labinius's commit / evaluate / fold
 * path and of the host-side protocol pieces the path's parity tests need (BLAKE3, the transcript,
 * the short-challenge sampler, the verifier's checks).
 *
 * TEST INFRASTRUCTURE ONLY. Nothing in the product links, loads or calls this file; only the
 * tests, `__graft_entry__.smoke()` and the `cpu_baseline` leg of `bench.py` do, and only as the
 * checker of (or the CPU time beside) the HIP path.
 *
 * Every function cites the reference file:line it restates (paths relative to the reference
 * checkout). Arithmetic is exact and fully reduced (u32 in [0, q)), following crates/pcs/src/
 * scalar.rs: there is no lazy reduction anywhere, so any representative the HIP kernels keep
 * internally has to come out identical here after the final reduction.
 *
 * Pinning (what checks each part; the tests under tests/ carry them out):
 *   - trees, psi, SLOT_EXP, POW3 tables: the definitions and the reference's compile-time asserts
 *     (units mod 1944 each once, psi of order exactly 1944, omega^2+omega+1 = 0, ...);
 *   - ntt/ntt_quad: Horner evaluation (eval_at / eval_quad_at) and the ring homomorphism against
 *     the schoolbook mul_mod_phi; intt/intt_quad: round trips;
 *   - decomposition: the defining identity y(psi^u) = sum_k psi^{uk} y_k(theta^u) brute-forced;
 *   - commit: the i64 sum of scalar NTTs (tests/commit.rs:86-101) and, independently, the
 *     coefficient-domain ring product sum_i a_i w_i;
 *   - fold: the coefficient-domain sum sum_j c_j(-X^4) w_j via mul_mod_phi;
 *   - F162 mul: the bit-by-bit schoolbook of tests/fields.rs:33-48 is this implementation;
 *     x has order 243, distributivity;
 *   - eq/row/mle: eq_naive / row_naive / mle_naive of tests/eval.rs:15-46;
 *   - BLAKE3: published test vectors of the spec (hash of "", "abc", keyed and derive-key of the
 *     empty input) plus agreement with the product's independently structured implementation at
 *     every input length around the chunk and tree boundaries;
 *   - A expansion: the state transition is checked to be linear with order 2^64 - 1 over GF(2)
 *     (a full-period xorshift, as the published generator), the draws uniform over the centred
 *     range; which full-period triple, the seed mixing and the draw order stay unpinned;
 *   - challenge sampler: the statistics the reference's tests/challenge.rs asserts (uniform
 *     positions, balanced signs, separation by absorbed data, acceptance at bound 15), plus
 *     shape, bound and determinism; its exact byte stream stays unpinned;
 *   - Witness::random and the transcript framing: parity unpinned (no reference execution is
 *     possible in this container); restated byte by byte;
 *   - wire formats: the reference's own wire tests (tests/wire.rs) restated case for case (exact
 *     sizes, the one-word-per-lane stream of 16 + 5 + 4 * 64 + 2 bytes, round trips on honest and
 *     adversarial messages, which error each broken stream gets), known answers worked from the
 *     format's description (bit order, a hand-computed largest-remainder quantisation), and a
 *     decoder written from the rANS definition in the tests (slot search, no tables). The rANS
 *     step here uses the exact quotient x / f, the device the reference's reciprocal form.
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#define N 648
#define N162 162
#define CONDUCTOR 1944
#define CONDUCTOR_QUAD 972

/* ============================================================================================
 * constants: crates/pcs/src/params.rs
 * ========================================================================================== */

/* params.rs:35-46 */
static uint64_t pow_mod(uint64_t b, uint64_t e, uint64_t q) {
    uint64_t r = 1;
    b %= q;
    while (e > 0) {
        if (e & 1) r = r * b % q;
        b = b * b % q;
        e >>= 1;
    }
    return r;
}

static uint64_t inv_mod(uint64_t a, uint64_t q) { return pow_mod(a, q - 2, q); }

/* params.rs:53-61 */
static uint32_t find_psi(uint32_t q) {
    for (uint64_t x = 2;; x++)
        if (pow_mod(x, 1944, q) == 1 && pow_mod(x, 972, q) != 1 && pow_mod(x, 648, q) != 1)
            return (uint32_t)x;
}

/* params.rs:291-299 */
static uint32_t find_psi972(uint32_t q) {
    for (uint64_t x = 2;; x++)
        if (pow_mod(x, 972, q) == 1 && pow_mod(x, 486, q) != 1 && pow_mod(x, 324, q) != 1)
            return (uint32_t)x;
}

static const int RADIX[7] = {2, 2, 2, 3, 3, 3, 3};
static const int SUBRINGS[8] = {1, 2, 4, 8, 24, 72, 216, 648};
static const int DEGREE[8] = {648, 324, 162, 81, 27, 9, 3, 1};
static const int RADIX_Q[6] = {2, 2, 3, 3, 3, 3};
static const int SUBRINGS_Q[7] = {1, 2, 4, 12, 36, 108, 324};
static const int DEGREE_Q[7] = {648, 324, 162, 54, 18, 6, 2};

/* params.rs:64-70 */
static uint32_t subring_exp(int level, int k) {
    if (level == 1) return k == 0 ? 324 : 1620;
    uint32_t p = (uint32_t)RADIX[level - 1];
    return (subring_exp(level - 1, k / (int)p) + CONDUCTOR * ((uint32_t)k % p)) / p;
}
/* params.rs:74-76 */
static uint32_t twiddle_exp(int level, int k) { return subring_exp(level, k) / (uint32_t)RADIX[level]; }

/* params.rs:305-311 */
static uint32_t subring_exp_quad(int level, int k) {
    if (level == 1) return k == 0 ? 162 : 810;
    uint32_t p = (uint32_t)RADIX_Q[level - 1];
    return (subring_exp_quad(level - 1, k / (int)p) + CONDUCTOR_QUAD * ((uint32_t)k % p)) / p;
}
/* params.rs:316-318 */
static uint32_t twiddle_exp_quad(int level, int k) {
    return subring_exp_quad(level, k) / (uint32_t)RADIX_Q[level];
}

/* params.rs:266-268 */
int orc_is_quad(uint32_t q) { return (q - 1) % CONDUCTOR_QUAD == 0 && (q - 1) % CONDUCTOR != 0; }

static uint16_t SLOT_EXP[N];          /* params.rs:78-89 */
static uint16_t QUAD_SLOT_EXP[324];   /* params.rs:320-332 */
static uint16_t POW3_SLOT_EXP[N162];  /* ring/mod.rs:149-176 */
static uint16_t SLOT_648[4][N162];    /* ring/mod.rs:178-179 */
static uint16_t QUAD_CLASS_SLOT[2][N162]; /* params.rs:357-401 */
static int tables_ready = 0;

static void init_tables(void) {
    if (tables_ready) return;
    for (int j = 0; j < N; j++) SLOT_EXP[j] = (uint16_t)subring_exp(7, j);
    for (int j = 0; j < 324; j++) QUAD_SLOT_EXP[j] = (uint16_t)subring_exp_quad(6, j);
    int pos[486];
    for (int v = 0; v < 486; v++) pos[v] = -1;
    int n = 0;
    for (int j = 0; j < N; j++) {
        int u = SLOT_EXP[j], v = u % 486, t = (u - v) / 486;
        if (pos[v] < 0) {
            pos[v] = n;
            POW3_SLOT_EXP[n] = (uint16_t)v;
            n++;
        }
        SLOT_648[t][pos[v]] = (uint16_t)j;
    }
    for (int s = 0; s < N162; s++) {
        int v = POW3_SLOT_EXP[s];
        for (int j = 0; j < 324; j++) {
            if (QUAD_SLOT_EXP[j] == v) QUAD_CLASS_SLOT[0][s] = (uint16_t)j;
            else if (QUAD_SLOT_EXP[j] == v + 486) QUAD_CLASS_SLOT[1][s] = (uint16_t)j;
        }
    }
    tables_ready = 1;
}

void orc_tables(uint16_t *slot_exp, uint16_t *quad_slot_exp, uint16_t *pow3, uint16_t *slot648,
                uint16_t *quad_class_slot) {
    init_tables();
    memcpy(slot_exp, SLOT_EXP, sizeof SLOT_EXP);
    memcpy(quad_slot_exp, QUAD_SLOT_EXP, sizeof QUAD_SLOT_EXP);
    memcpy(pow3, POW3_SLOT_EXP, sizeof POW3_SLOT_EXP);
    memcpy(slot648, SLOT_648, sizeof SLOT_648);
    memcpy(quad_class_slot, QUAD_CLASS_SLOT, sizeof QUAD_CLASS_SLOT);
}

uint32_t orc_psi(uint32_t q) { return orc_is_quad(q) ? find_psi972(q) : find_psi(q); }

/* ============================================================================================
 * per-prime tables: every twiddle of both trees, precomputed once per prime
 * ========================================================================================== */

typedef struct {
    uint32_t q;
    uint64_t bm;           /* floor(2^40 / q): x mod q for x < 2^40 with one multiply */
    int quad;
    uint64_t psi, w, w2, z6, inv2, inv3, det;
    uint32_t tw[7][648];   /* tw[level][k] = psi^twiddle_exp(level, k) */
    uint32_t twi[7][648];  /* inverses */
    uint32_t leafc[324];   /* quadratic tree: LEAF_C[j] = psi'^QUAD_SLOT_EXP[j], params.rs:444-452 */
} prime_tab;

#define MAX_PRIMES 8
static prime_tab PT[MAX_PRIMES];
static int n_pt = 0;

static const prime_tab *prime(uint32_t q) {
    for (int i = 0; i < n_pt; i++)
        if (PT[i].q == q) return &PT[i];
    prime_tab *t = &PT[n_pt++];
    memset(t, 0, sizeof *t);
    t->q = q;
    t->bm = (1ull << 40) / q;
    t->quad = orc_is_quad(q);
    if (t->quad) {
        t->psi = find_psi972(q);
        t->w = pow_mod(t->psi, 324, q);     /* params.rs:414 */
        t->z6 = pow_mod(t->psi, 162, q);    /* params.rs:416 */
        for (int l = 1; l <= 5; l++)
            for (int k = 0; k < SUBRINGS_Q[l]; k++) {
                t->tw[l][k] = (uint32_t)pow_mod(t->psi, twiddle_exp_quad(l, k), q);
                t->twi[l][k] = (uint32_t)inv_mod(t->tw[l][k], q);
            }
    } else {
        t->psi = find_psi(q);
        t->w = pow_mod(t->psi, 648, q);     /* params.rs:122 */
        t->z6 = pow_mod(t->psi, 324, q);    /* params.rs:124 */
        for (int l = 1; l <= 6; l++)
            for (int k = 0; k < SUBRINGS[l]; k++) {
                t->tw[l][k] = (uint32_t)pow_mod(t->psi, twiddle_exp(l, k), q);
                t->twi[l][k] = (uint32_t)inv_mod(t->tw[l][k], q);
            }
    }
    if (t->quad)
        for (int j = 0; j < 324; j++) t->leafc[j] = (uint32_t)pow_mod(t->psi, QUAD_SLOT_EXP[j], q);
    t->w2 = t->w * t->w % q;
    t->inv2 = inv_mod(2, q);
    t->inv3 = inv_mod(3, q);
    t->det = inv_mod((2 * t->z6 + q - 1) % q, q);
    return t;
}

/* x mod q for x < 2^40 (Barrett with a 40-bit magic; the quotient estimate is low by at most 1) */
static inline uint64_t md(const prime_tab *t, uint64_t x) {
    uint64_t r = x - (uint64_t)(((unsigned __int128)x * t->bm) >> 40) * t->q;
    return r >= t->q ? r - t->q : r;
}

/* ============================================================================================
 * the exact scalar transforms: crates/pcs/src/scalar.rs, recursion/limbs.rs:76-127
 * ========================================================================================== */

/* scalar.rs:9-32: a * b in Z_q[X]/(X^648 - X^324 + 1) */
void orc_mul_mod_phi(const uint32_t *a, const uint32_t *b, uint32_t q, uint32_t *out) {
    uint64_t c[2 * N];
    memset(c, 0, sizeof c);
    uint64_t qq = q;
    for (int i = 0; i < N; i++) {
        if (a[i] == 0) continue;
        for (int j = 0; j < N; j++) c[i + j] += (uint64_t)a[i] * b[j];
        /* keep the running sums bounded: fold every row */
        if ((i & 63) == 63)
            for (int j = 0; j < 2 * N; j++) c[j] %= qq;
    }
    for (int k = 2 * N - 1; k >= N; k--) {
        uint64_t v = c[k] % qq;
        c[k] = 0;
        c[k - 324] += v;
        c[k - 648] += qq * qq - v;
    }
    for (int i = 0; i < N; i++) out[i] = (uint32_t)(c[i] % qq);
}

/* scalar.rs:35-84: forward NTT on the splitting tree, out[j] = a(psi^SLOT_EXP[j]) */
static void ntt_split(const prime_tab *t, uint64_t *v) {
    uint64_t q = t->q, w = t->w, w2 = t->w2, z6 = t->z6;
    for (int i = 0; i < 324; i++) {
        uint64_t a0 = v[i], a1 = v[i + 324], tt = md(t, a1 * z6);
        v[i] = md(t, a0 + tt);
        v[i + 324] = md(t, a0 + a1 + q - tt);
    }
    for (int level = 1; level <= 6; level++) {
        int n = DEGREE[level], p = RADIX[level], m = n / p;
        for (int k = 0; k < SUBRINGS[level]; k++) {
            int base = k * n;
            uint64_t zeta = t->tw[level][k];
            if (p == 2) {
                for (int i = 0; i < m; i++) {
                    uint64_t a0 = v[base + i], tt = md(t, v[base + m + i] * zeta);
                    v[base + i] = md(t, a0 + tt);
                    v[base + m + i] = md(t, a0 + q - tt);
                }
            } else {
                uint64_t zeta2 = md(t, zeta * zeta);
                for (int i = 0; i < m; i++) {
                    uint64_t a0 = v[base + i];
                    uint64_t t1 = md(t, v[base + m + i] * zeta);
                    uint64_t t2 = md(t, v[base + 2 * m + i] * zeta2);
                    v[base + i] = md(t, a0 + t1 + t2);
                    v[base + m + i] = md(t, a0 + w * t1 + w2 * t2);
                    v[base + 2 * m + i] = md(t, a0 + w2 * t1 + w * t2);
                }
            }
        }
    }
}

/* scalar.rs:185-234: forward transform on the quadratic-slot tree */
static void ntt_quad(const prime_tab *t, uint64_t *v) {
    uint64_t q = t->q, w = t->w, w2 = t->w2, z6 = t->z6;
    for (int i = 0; i < 324; i++) {
        uint64_t a0 = v[i], a1 = v[i + 324], tt = md(t, a1 * z6);
        v[i] = md(t, a0 + tt);
        v[i + 324] = md(t, a0 + a1 + q - tt);
    }
    for (int level = 1; level <= 5; level++) {
        int n = DEGREE_Q[level], p = RADIX_Q[level], m = n / p;
        for (int k = 0; k < SUBRINGS_Q[level]; k++) {
            int base = k * n;
            uint64_t zeta = t->tw[level][k];
            if (p == 2) {
                for (int i = 0; i < m; i++) {
                    uint64_t a0 = v[base + i], tt = md(t, v[base + m + i] * zeta);
                    v[base + i] = md(t, a0 + tt);
                    v[base + m + i] = md(t, a0 + q - tt);
                }
            } else {
                uint64_t zeta2 = md(t, zeta * zeta);
                for (int i = 0; i < m; i++) {
                    uint64_t a0 = v[base + i];
                    uint64_t t1 = md(t, v[base + m + i] * zeta);
                    uint64_t t2 = md(t, v[base + 2 * m + i] * zeta2);
                    v[base + i] = md(t, a0 + t1 + t2);
                    v[base + m + i] = md(t, a0 + w * t1 + w2 * t2);
                    v[base + 2 * m + i] = md(t, a0 + w2 * t1 + w * t2);
                }
            }
        }
    }
}

/* scalar.rs:123-177 (split) and recursion/limbs.rs:76-127 (quadratic): the exact inverses */
static void intt_any(const prime_tab *t, uint64_t *u) {
    uint64_t q = t->q, w = t->w, w2 = t->w2;
    int top = t->quad ? 5 : 6;
    for (int level = top; level >= 1; level--) {
        int n = t->quad ? DEGREE_Q[level] : DEGREE[level];
        int p = t->quad ? RADIX_Q[level] : RADIX[level];
        int nk = t->quad ? SUBRINGS_Q[level] : SUBRINGS[level];
        int m = n / p;
        for (int k = 0; k < nk; k++) {
            int base = k * n;
            uint64_t zi = t->twi[level][k];
            if (p == 2) {
                for (int i = 0; i < m; i++) {
                    uint64_t y0 = u[base + i], y1 = u[base + m + i];
                    u[base + i] = (y0 + y1) % q * t->inv2 % q;
                    u[base + m + i] = (y0 + q - y1) % q * t->inv2 % q * zi % q;
                }
            } else {
                uint64_t zi2 = zi * zi % q;
                for (int i = 0; i < m; i++) {
                    uint64_t y0 = u[base + i], y1 = u[base + m + i], y2 = u[base + 2 * m + i];
                    uint64_t a0 = (y0 + y1 + y2) % q * t->inv3 % q;
                    uint64_t t1 = (y0 + w2 * y1 + w * y2) % q * t->inv3 % q;
                    uint64_t t2 = (y0 + w * y1 + w2 * y2) % q * t->inv3 % q;
                    u[base + i] = a0;
                    u[base + m + i] = t1 * zi % q;
                    u[base + 2 * m + i] = t2 * zi2 % q;
                }
            }
        }
    }
    for (int i = 0; i < 324; i++) {
        uint64_t y0 = u[i], y1 = u[i + 324];
        uint64_t a1 = (y0 + q - y1) % q * t->det % q;
        uint64_t a0 = (y0 + q - t->z6 * a1 % q) % q;
        u[i] = a0;
        u[i + 324] = a1;
    }
}

/* The transform of limb q (split or quadratic tree), fully reduced. */
void orc_ntt(uint32_t q, const uint32_t *a, uint32_t *out) {
    init_tables();
    const prime_tab *t = prime(q);
    uint64_t v[N];
    for (int i = 0; i < N; i++) v[i] = a[i] % q;
    if (t->quad) ntt_quad(t, v);
    else ntt_split(t, v);
    for (int i = 0; i < N; i++) out[i] = (uint32_t)v[i];
}

void orc_intt(uint32_t q, const uint32_t *a, uint32_t *out) {
    init_tables();
    const prime_tab *t = prime(q);
    uint64_t v[N];
    for (int i = 0; i < N; i++) v[i] = a[i] % q;
    intt_any(t, v);
    for (int i = 0; i < N; i++) out[i] = (uint32_t)v[i];
}

/* scalar.rs:87-95: a(psi^u) by Horner */
uint32_t orc_eval_at(uint32_t q, const uint32_t *a, uint32_t u) {
    uint64_t x = pow_mod(find_psi(q), u, q), acc = 0;
    for (int i = N - 1; i >= 0; i--) acc = (acc * x + a[i]) % q;
    return (uint32_t)acc;
}

/* scalar.rs:237-248: a mod (X^2 - psi'^u) by Horner */
void orc_eval_quad_at(uint32_t q, const uint32_t *a, uint32_t u, uint32_t *r01) {
    uint64_t c = pow_mod(find_psi972(q), u, q), r0 = 0, r1 = 0;
    for (int i = N - 1; i >= 0; i--) {
        uint64_t n0 = (r1 * c + a[i]) % q, n1 = r0;
        r0 = n0;
        r1 = n1;
    }
    r01[0] = (uint32_t)r0;
    r01[1] = (uint32_t)r1;
}

/* scalar.rs:254-265: leaf-wise product of two quadratic-slot transforms */
static void mul_quad_slots(const prime_tab *t, const uint64_t *a, const uint64_t *b, uint64_t *c) {
    for (int j = 0; j < 324; j++) {
        uint64_t cj = t->leafc[j];
        uint64_t a0 = a[2 * j], a1 = a[2 * j + 1], b0 = b[2 * j], b1 = b[2 * j + 1];
        c[2 * j] = md(t, a0 * b0 + cj * md(t, a1 * b1));
        c[2 * j + 1] = md(t, a0 * b1 + a1 * b0);
    }
}

/* ============================================================================================
 * the 648 -> 4 x 162 decomposition: ring/mod.rs:22-55, 217-312 (split), 363-417 (quadratic)
 * ========================================================================================== */

static int16_t center(uint64_t x, uint64_t q) {
    x %= q;
    return x > (q - 1) / 2 ? (int16_t)((int64_t)x - (int64_t)q) : (int16_t)x;
}

/* out[k][s], centered, for a fully reduced y */
void orc_decompose(uint32_t q, const uint32_t *y, int16_t *out) {
    init_tables();
    const prime_tab *t = prime(q);
    uint64_t qq = q;
    if (!t->quad) {
        /* Y_k(v) = 4^-1 psi^{-vk} sum_t i^{-tk} E_t, i = psi^486 */
        uint64_t inv4 = inv_mod(4, qq), ii = pow_mod(t->psi, 486, qq);
        uint64_t ipow[4] = {1, ii, ii * ii % qq, ii * ii % qq * ii % qq};
        for (int s = 0; s < N162; s++) {
            uint64_t v = POW3_SLOT_EXP[s];
            for (int k = 0; k < 4; k++) {
                uint64_t acc = 0;
                for (int tt = 0; tt < 4; tt++) {
                    uint64_t e = y[SLOT_648[tt][s]] % qq;
                    /* i^{-tk} = i^{(4 - tk mod 4) mod 4} */
                    acc += e * ipow[(4 - (tt * k) % 4) % 4] % qq;
                }
                uint64_t tw = inv4 * pow_mod(t->psi, (CONDUCTOR - (v * k) % CONDUCTOR) % CONDUCTOR, qq) % qq;
                out[k * N162 + s] = center(acc % qq * tw, qq);
            }
        }
    } else {
        /* scalar.rs:287-309 */
        uint64_t inv2 = t->inv2;
        for (int s = 0; s < N162; s++) {
            uint64_t v = POW3_SLOT_EXP[s];
            int jp = QUAD_CLASS_SLOT[0][s], jm = QUAD_CLASS_SLOT[1][s];
            uint64_t ipv = pow_mod(t->psi, (CONDUCTOR_QUAD - v) % CONDUCTOR_QUAD, qq);
            for (int k = 0; k < 2; k++) {
                uint64_t ep = y[2 * jp + k] % qq, em = y[2 * jm + k] % qq;
                out[k * N162 + s] = center((ep + em) % qq * inv2, qq);
                out[(k + 2) * N162 + s] = center((ep + qq - em) % qq * inv2 % qq * ipv, qq);
            }
        }
    }
}

/* ============================================================================================
 * BLAKE3 (the public specification: compression, chunk chaining, the binary tree and the XOF),
 * written here as an incremental hasher with a chaining-value stack.
 * ========================================================================================== */

static const uint32_t B3_IV[8] = {0x6A09E667, 0xBB67AE85, 0x3C6EF372, 0xA54FF53A,
                                  0x510E527F, 0x9B05688C, 0x1F83D9AB, 0x5BE0CD19};
static const int B3_PERM[16] = {2, 6, 3, 10, 7, 0, 4, 13, 1, 11, 12, 5, 9, 14, 15, 8};
enum { CHUNK_START = 1, CHUNK_END = 2, PARENT = 4, ROOT = 8, KEYED_HASH = 16,
       DERIVE_KEY_CONTEXT = 32, DERIVE_KEY_MATERIAL = 64 };

static uint32_t rotr(uint32_t x, int n) { return (x >> n) | (x << (32 - n)); }

static void b3_g(uint32_t *s, int a, int b, int c, int d, uint32_t mx, uint32_t my) {
    s[a] = s[a] + s[b] + mx;
    s[d] = rotr(s[d] ^ s[a], 16);
    s[c] = s[c] + s[d];
    s[b] = rotr(s[b] ^ s[c], 12);
    s[a] = s[a] + s[b] + my;
    s[d] = rotr(s[d] ^ s[a], 8);
    s[c] = s[c] + s[d];
    s[b] = rotr(s[b] ^ s[c], 7);
}

static void b3_compress(const uint32_t cv[8], const uint32_t block[16], uint64_t counter,
                        uint32_t block_len, uint32_t flags, uint32_t out[16]) {
    uint32_t s[16], m[16], t[16];
    for (int i = 0; i < 8; i++) s[i] = cv[i];
    for (int i = 0; i < 4; i++) s[8 + i] = B3_IV[i];
    s[12] = (uint32_t)counter;
    s[13] = (uint32_t)(counter >> 32);
    s[14] = block_len;
    s[15] = flags;
    memcpy(m, block, sizeof m);
    for (int r = 0; r < 7; r++) {
        b3_g(s, 0, 4, 8, 12, m[0], m[1]);
        b3_g(s, 1, 5, 9, 13, m[2], m[3]);
        b3_g(s, 2, 6, 10, 14, m[4], m[5]);
        b3_g(s, 3, 7, 11, 15, m[6], m[7]);
        b3_g(s, 0, 5, 10, 15, m[8], m[9]);
        b3_g(s, 1, 6, 11, 12, m[10], m[11]);
        b3_g(s, 2, 7, 8, 13, m[12], m[13]);
        b3_g(s, 3, 4, 9, 14, m[14], m[15]);
        if (r < 6) {
            for (int i = 0; i < 16; i++) t[i] = m[B3_PERM[i]];
            memcpy(m, t, sizeof m);
        }
    }
    for (int i = 0; i < 8; i++) {
        out[i] = s[i] ^ s[i + 8];
        out[i + 8] = s[i + 8] ^ cv[i];
    }
}

static void words_le(const uint8_t *b, uint32_t *w, int n) {
    for (int i = 0; i < n; i++)
        w[i] = (uint32_t)b[4 * i] | (uint32_t)b[4 * i + 1] << 8 | (uint32_t)b[4 * i + 2] << 16 |
               (uint32_t)b[4 * i + 3] << 24;
}

typedef struct {
    uint32_t key[8];
    uint32_t flags;
    uint32_t cv[8];
    uint64_t chunk_counter;
    uint8_t block[64];
    uint32_t block_len;
    uint32_t blocks_done;
    uint32_t stack[64][8];
    int stack_len;
} b3;

/* the output of a node not yet finalised: enough to squeeze root bytes or take a chaining value */
typedef struct {
    uint32_t cv[8];
    uint32_t block[16];
    uint64_t counter;
    uint32_t block_len;
    uint32_t flags;
} b3_out;

static void b3_init(b3 *h, const uint32_t key[8], uint32_t flags) {
    memset(h, 0, sizeof *h);
    memcpy(h->key, key, 32);
    memcpy(h->cv, key, 32);
    h->flags = flags;
}

static uint32_t chunk_start_flag(const b3 *h) { return h->blocks_done == 0 ? CHUNK_START : 0; }

static void chunk_output(const b3 *h, b3_out *o) {
    memcpy(o->cv, h->cv, 32);
    uint8_t blk[64];
    memset(blk, 0, 64);
    memcpy(blk, h->block, h->block_len);
    words_le(blk, o->block, 16);
    o->counter = h->chunk_counter;
    o->block_len = h->block_len;
    o->flags = h->flags | chunk_start_flag(h) | CHUNK_END;
}

static void parent_output(const uint32_t *l, const uint32_t *r, const uint32_t *key,
                          uint32_t flags, b3_out *o) {
    memcpy(o->cv, key, 32);
    memcpy(o->block, l, 32);
    memcpy(o->block + 8, r, 32);
    o->counter = 0;
    o->block_len = 64;
    o->flags = PARENT | flags;
}

static void out_cv(const b3_out *o, uint32_t cv[8]) {
    uint32_t w[16];
    b3_compress(o->cv, o->block, o->counter, o->block_len, o->flags, w);
    memcpy(cv, w, 32);
}

static void b3_update(b3 *h, const uint8_t *in, size_t len) {
    while (len > 0) {
        if (h->block_len == 64) {
            uint32_t w[16], o[16];
            words_le(h->block, w, 16);
            if (h->blocks_done == 15) {
                /* the chunk is full: finish it only once more input is known to follow */
            } else {
                b3_compress(h->cv, w, h->chunk_counter, 64, h->flags | chunk_start_flag(h), o);
                memcpy(h->cv, o, 32);
                h->blocks_done++;
                h->block_len = 0;
                continue;
            }
            /* 16th block of a chunk with more input following: close the chunk, push its CV */
            b3_out co;
            chunk_output(h, &co);
            uint32_t cv[8];
            out_cv(&co, cv);
            uint64_t total = h->chunk_counter + 1;
            while ((total & 1) == 0) {
                b3_out po;
                h->stack_len--;
                parent_output(h->stack[h->stack_len], cv, h->key, h->flags, &po);
                out_cv(&po, cv);
                total >>= 1;
            }
            memcpy(h->stack[h->stack_len++], cv, 32);
            memcpy(h->cv, h->key, 32);
            h->chunk_counter++;
            h->block_len = 0;
            h->blocks_done = 0;
        }
        size_t take = 64 - h->block_len;
        if (take > len) take = len;
        memcpy(h->block + h->block_len, in, take);
        h->block_len += (uint32_t)take;
        in += take;
        len -= take;
    }
}

static void b3_final_out(const b3 *h, b3_out *o) {
    chunk_output(h, o);
    for (int i = h->stack_len - 1; i >= 0; i--) {
        uint32_t cv[8];
        out_cv(o, cv);
        parent_output(h->stack[i], cv, h->key, h->flags, o);
    }
}

/* root output bytes [pos, pos + len) */
static void b3_squeeze(const b3_out *o, uint64_t pos, uint8_t *dst, size_t len) {
    while (len > 0) {
        uint64_t blk = pos / 64;
        uint32_t w[16];
        b3_compress(o->cv, o->block, blk, o->block_len, o->flags | ROOT, w);
        uint8_t bytes[64];
        for (int i = 0; i < 16; i++) {
            bytes[4 * i] = (uint8_t)w[i];
            bytes[4 * i + 1] = (uint8_t)(w[i] >> 8);
            bytes[4 * i + 2] = (uint8_t)(w[i] >> 16);
            bytes[4 * i + 3] = (uint8_t)(w[i] >> 24);
        }
        size_t off = (size_t)(pos % 64), take = 64 - off;
        if (take > len) take = len;
        memcpy(dst, bytes + off, take);
        dst += take;
        pos += take;
        len -= take;
    }
}

static void b3_new(b3 *h) { b3_init(h, B3_IV, 0); }
static void b3_new_keyed(b3 *h, const uint8_t key[32]) {
    uint32_t k[8];
    words_le(key, k, 8);
    b3_init(h, k, KEYED_HASH);
}
static void b3_new_derive(b3 *h, const char *context) {
    b3 c;
    b3_init(&c, B3_IV, DERIVE_KEY_CONTEXT);
    b3_update(&c, (const uint8_t *)context, strlen(context));
    b3_out o;
    b3_final_out(&c, &o);
    uint8_t ck[32];
    b3_squeeze(&o, 0, ck, 32);
    uint32_t k[8];
    words_le(ck, k, 8);
    b3_init(h, k, DERIVE_KEY_MATERIAL);
}

/* mode 0 hash, 1 keyed (key = 32 bytes), 2 derive_key (key = NUL-terminated context) */
void orc_blake3(int mode, const uint8_t *key, const uint8_t *in, size_t len, uint8_t *out,
                size_t out_len) {
    b3 h;
    if (mode == 0) b3_new(&h);
    else if (mode == 1) b3_new_keyed(&h, key);
    else b3_new_derive(&h, (const char *)key);
    b3_update(&h, in, len);
    b3_out o;
    b3_final_out(&h, &o);
    b3_squeeze(&o, 0, out, out_len);
}

/* ============================================================================================
 * randomness and inputs: rng.rs:1-20, key.rs:48-88, scheme/params.rs:320-324,
 * scheme/witness.rs:64-87
 * ========================================================================================== */

typedef struct { uint64_t s; } rng_t;
static void rng_new(rng_t *r, uint64_t seed) { r->s = seed * 0x9E3779B97F4A7C15ull | 1; }
/* xorshift64*'s state transition (rng.rs:10-14), exported so that the tests can check it is the
 * full-period linear map of the published generator */
uint64_t orc_xorshift_step(uint64_t x) {
    x ^= x >> 12;
    x ^= x << 25;
    x ^= x >> 27;
    return x;
}
static uint64_t rng_next(rng_t *r) {
    r->s = orc_xorshift_step(r->s);
    return r->s * 0x2545F4914F6CDD1Dull;
}
static uint32_t rng_below(rng_t *r, uint32_t n) {
    return (uint32_t)(((rng_next(r) >> 32) * (uint64_t)n) >> 32);
}

/* scheme/params.rs:321-322 */
uint64_t orc_matrix_seed(const uint8_t seed[32]) {
    uint8_t d[32];
    orc_blake3(0, 0, seed, 32, d, 32);
    uint64_t s = 0;
    for (int i = 7; i >= 0; i--) s = s << 8 | d[i];
    return s;
}

/* key.rs:63-81: a[k][i][j] = A_k[i][j], i = 32 b + p, drawn b-major, then j, then p */
void orc_key_expand(const uint32_t *primes, int limbs, int len_ring, uint64_t seed, int16_t *a) {
    for (int k = 0; k < limbs; k++) {
        uint32_t q = primes[k];
        int16_t half = (int16_t)((q - 1) / 2);
        rng_t r;
        rng_new(&r, seed ^ (0x9E3779B9ull * (uint64_t)(k + 1)));
        int16_t *ak = a + (size_t)k * len_ring * N;
        for (int b = 0; b < len_ring / 32; b++)
            for (int j = 0; j < N; j++)
                for (int p = 0; p < 32; p++)
                    ak[(size_t)(32 * b + p) * N + j] = (int16_t)((int16_t)rng_below(&r, q) - half);
    }
}

/* scheme/witness.rs:66-87: n elements as 3 u64 each */
void orc_witness_random(const uint8_t seed[32], size_t n, uint64_t *w) {
    b3 h;
    b3_new_keyed(&h, seed);
    b3_update(&h, (const uint8_t *)"labinius/witness", 16);
    b3_out o;
    b3_final_out(&h, &o);
    uint8_t buf[24 * 256];
    for (size_t i0 = 0; i0 < n; i0 += 256) {
        size_t cnt = n - i0 < 256 ? n - i0 : 256;
        b3_squeeze(&o, 24 * i0, buf, 24 * cnt);
        for (size_t i = 0; i < cnt; i++) {
            for (int k = 0; k < 3; k++) {
                uint64_t x = 0;
                for (int b = 7; b >= 0; b--) x = x << 8 | buf[24 * i + 8 * k + b];
                w[3 * (i0 + i) + k] = x;
            }
            w[3 * (i0 + i) + 2] &= (1ull << 34) - 1;
        }
    }
}

/* f162.rs:32-41: coefficient 4m + k of ring element r is bit m of element 4r + k */
void orc_lift(const uint64_t *elems, uint32_t *c) {
    for (int m = 0; m < 162; m++)
        for (int k = 0; k < 4; k++) c[4 * m + k] = (uint32_t)(elems[3 * k + m / 64] >> (m % 64)) & 1;
}

/* ============================================================================================
 * the commitment: tests/commit.rs:86-101 (split), scalar.rs:254-265 (quadratic), key.rs:192-234
 * ========================================================================================== */

/* raw[k][c][j] in [0, q_k); comp[c][t][k][s] centered (the column-major 4 x r matrix, each entry
 * holding one R_162 element per limb) */
void orc_commit(const uint32_t *primes, int limbs, int len_ring, int cols, const int16_t *a,
                const uint64_t *w, uint32_t *raw, int16_t *comp) {
    init_tables();
    for (int k = 0; k < limbs; k++) {
        const prime_tab *t = prime(primes[k]);
        uint64_t q = t->q;
        const int16_t *ak = a + (size_t)k * len_ring * N;
        for (int c = 0; c < cols; c++) {
            uint64_t y[N];
            memset(y, 0, sizeof y);
            for (int i = 0; i < len_ring; i++) {
                uint32_t lift[N];
                orc_lift(w + 3 * ((size_t)c * len_ring * 4 + 4 * (size_t)i), lift);
                uint64_t v[N];
                for (int j = 0; j < N; j++) v[j] = lift[j];
                uint64_t av[N];
                for (int j = 0; j < N; j++)
                    av[j] = (uint64_t)(((int64_t)ak[(size_t)i * N + j] % (int64_t)q + (int64_t)q) % (int64_t)q);
                if (t->quad) {
                    ntt_quad(t, v);
                    uint64_t pr[N];
                    mul_quad_slots(t, av, v, pr);
                    for (int j = 0; j < N; j++) y[j] = md(t, y[j] + pr[j]);
                } else {
                    ntt_split(t, v);
                    for (int j = 0; j < N; j++) y[j] = md(t, y[j] + v[j] * av[j]);
                }
            }
            uint32_t *rk = raw + ((size_t)k * cols + c) * N;
            for (int j = 0; j < N; j++) rk[j] = (uint32_t)y[j];
            int16_t d[4 * N162];
            orc_decompose(primes[k], rk, d);
            for (int tt = 0; tt < 4; tt++)
                for (int s = 0; s < N162; s++)
                    comp[(((size_t)c * 4 + tt) * limbs + k) * N162 + s] = d[tt * N162 + s];
        }
    }
}

/* ============================================================================================
 * the fold: fold.rs:281-393 (embed, challenge transforms), 605-677 (fold_witness)
 * ========================================================================================== */

typedef struct {
    uint8_t positions[32];
    uint32_t signs;
    uint32_t weight;
} orc_challenge;

/* fold.rs:287-300: c(-X^4), coefficient of X^{4m} = (-1)^m (1 - 2 sign) */
void orc_embed(const orc_challenge *c, int32_t *e) {
    memset(e, 0, N * sizeof(int32_t));
    for (uint32_t i = 0; i < c->weight; i++) {
        int m = c->positions[i];
        int co = 1 - 2 * (int)((c->signs >> i) & 1);
        e[4 * m] = (m % 2 == 0) ? co : -co;
    }
}

/* v (len_ring x 648 centered coefficients mod the base prime) = sum_j c_j(-X^4) W_j, computed as
 * the reference does: transform the embedded challenges and the kept witness modulo the base,
 * multiply slot-wise, sum over the columns, invert. */
void orc_fold(uint32_t q, int len_ring, int cols, const uint64_t *w, const orc_challenge *ch,
              int16_t *v) {
    init_tables();
    const prime_tab *t = prime(q);
    uint64_t qq = q;
    uint64_t *chat = (uint64_t *)malloc(sizeof(uint64_t) * N * cols);
    for (int j = 0; j < cols; j++) {
        int32_t e[N];
        orc_embed(&ch[j], e);
        uint64_t *cj = chat + (size_t)j * N;
        for (int i = 0; i < N; i++) cj[i] = (uint64_t)((e[i] % (int64_t)qq + (int64_t)qq) % (int64_t)qq);
        if (t->quad) {
            ntt_quad(t, cj);
            /* fold.rs:346-352: the leaf image is a scalar; write it into both rows */
            for (int l = 0; l < 324; l++) cj[2 * l + 1] = cj[2 * l];
        } else {
            ntt_split(t, cj);
        }
    }
    for (int i = 0; i < len_ring; i++) {
        uint64_t acc[N];
        memset(acc, 0, sizeof acc);
        for (int j = 0; j < cols; j++) {
            uint32_t lift[N];
            orc_lift(w + 3 * ((size_t)j * len_ring * 4 + 4 * (size_t)i), lift);
            uint64_t x[N];
            for (int u = 0; u < N; u++) x[u] = lift[u];
            if (t->quad) ntt_quad(t, x);
            else ntt_split(t, x);
            const uint64_t *cj = chat + (size_t)j * N;
            for (int u = 0; u < N; u++) acc[u] = md(t, acc[u] + cj[u] * x[u]);
        }
        intt_any(t, acc);
        for (int u = 0; u < N; u++) v[(size_t)i * N + u] = center(acc[u], qq);
    }
    free(chat);
}

/* orc_fold restricted to the output ring elements `rows` (each output element depends only on
 * the ring elements with the same index in every column): the CPU baseline times a sample. */
void orc_fold_rows(uint32_t q, int len_ring, int cols, const uint64_t *w, const orc_challenge *ch,
                   const int32_t *rows, int n_rows, int16_t *v) {
    init_tables();
    const prime_tab *t = prime(q);
    uint64_t qq = q;
    uint64_t *chat = (uint64_t *)malloc(sizeof(uint64_t) * N * cols);
    for (int j = 0; j < cols; j++) {
        int32_t e[N];
        orc_embed(&ch[j], e);
        uint64_t *cj = chat + (size_t)j * N;
        for (int i = 0; i < N; i++) cj[i] = (uint64_t)((e[i] % (int64_t)qq + (int64_t)qq) % (int64_t)qq);
        if (t->quad) {
            ntt_quad(t, cj);
            for (int l = 0; l < 324; l++) cj[2 * l + 1] = cj[2 * l];
        } else {
            ntt_split(t, cj);
        }
    }
    for (int r = 0; r < n_rows; r++) {
        int i = rows[r];
        uint64_t acc[N];
        memset(acc, 0, sizeof acc);
        for (int j = 0; j < cols; j++) {
            uint32_t lift[N];
            orc_lift(w + 3 * ((size_t)j * len_ring * 4 + 4 * (size_t)i), lift);
            uint64_t x[N];
            for (int u = 0; u < N; u++) x[u] = lift[u];
            if (t->quad) ntt_quad(t, x);
            else ntt_split(t, x);
            const uint64_t *cj = chat + (size_t)j * N;
            for (int u = 0; u < N; u++) acc[u] = md(t, acc[u] + cj[u] * x[u]);
        }
        intt_any(t, acc);
        for (int u = 0; u < N; u++) v[(size_t)r * N + u] = center(acc[u], qq);
    }
    free(chat);
}

/* The same v straight from the definition: the ring products c_j(-X^4) w_{j,i} in
 * Z[X]/(X^648 - X^324 + 1) by sparse convolution, summed over Z, then centered mod q. Used to
 * check sampled output elements at sizes where orc_fold is too slow. */
void orc_fold_direct(uint32_t q, int len_ring, int cols, const uint64_t *w,
                     const orc_challenge *ch, const int32_t *rows, int n_rows, int16_t *v) {
    for (int r = 0; r < n_rows; r++) {
        int i = rows[r];
        int64_t acc[2 * N];
        memset(acc, 0, sizeof acc);
        for (int j = 0; j < cols; j++) {
            uint32_t lift[N];
            orc_lift(w + 3 * ((size_t)j * len_ring * 4 + 4 * (size_t)i), lift);
            int32_t e[N];
            orc_embed(&ch[j], e);
            for (int a = 0; a < N; a += 4) {
                if (!e[a]) continue;
                for (int b = 0; b < N; b++)
                    if (lift[b]) acc[a + b] += e[a];
            }
        }
        for (int k = 2 * N - 1; k >= N; k--) {
            acc[k - 324] += acc[k];
            acc[k - 648] -= acc[k];
            acc[k] = 0;
        }
        for (int u = 0; u < N; u++) {
            int64_t x = acc[u] % (int64_t)q;
            if (x < 0) x += q;
            v[(size_t)r * N + u] = center((uint64_t)x, q);
        }
    }
}

/* ============================================================================================
 * F162 = GF(2)[x]/(x^162 + x^81 + 1): fields/scalar.rs:92-116, the schoolbook of
 * tests/fields.rs:33-48
 * ========================================================================================== */

static int bit(const uint64_t *v, int i) { return (int)((v[i / 64] >> (i % 64)) & 1); }
static void flip(uint64_t *v, int i) { v[i / 64] ^= 1ull << (i % 64); }

void orc_f162_mul(const uint64_t *a, const uint64_t *b, uint64_t *out) {
    uint64_t p[6] = {0};
    for (int i = 0; i < 162; i++) {
        if (!bit(a, i)) continue;
        /* p ^= b << i */
        int ws = i / 64, bs = i % 64;
        for (int k = 0; k < 3; k++) {
            p[k + ws] ^= b[k] << bs;
            if (bs) p[k + ws + 1] ^= b[k] >> (64 - bs);
        }
    }
    for (int j = 323; j >= 162; j--)
        if (bit(p, j)) {
            flip(p, j);
            flip(p, j - 162);
            flip(p, j - 81);
        }
    out[0] = p[0];
    out[1] = p[1];
    out[2] = p[2];
}

/* eval.rs:51-63 */
void orc_eq_table(const uint64_t *ps, int n, uint64_t *t) {
    t[0] = 1;
    t[1] = 0;
    t[2] = 0;
    size_t len = 1;
    for (int k = 0; k < n; k++) {
        const uint64_t *p = ps + 3 * k;
        uint64_t p1[3] = {p[0] ^ 1, p[1], p[2]};
        for (size_t b = 0; b < len; b++) {
            uint64_t x[3] = {t[3 * b], t[3 * b + 1], t[3 * b + 2]};
            orc_f162_mul(x, p, t + 3 * (b + len));
            orc_f162_mul(x, p1, t + 3 * b);
        }
        len *= 2;
    }
}

/* eval.rs:171-182: u_j = sum_i eq(p0, i) W[i + wdim j] */
void orc_row_evaluate(const uint64_t *w, size_t n, const uint64_t *p0, int rows, uint64_t *u) {
    size_t wdim = (size_t)1 << rows;
    uint64_t *eq = (uint64_t *)malloc(sizeof(uint64_t) * 3 * wdim);
    orc_eq_table(p0, rows, eq);
    for (size_t j = 0; j < n / wdim; j++) {
        uint64_t acc[3] = {0, 0, 0};
        for (size_t i = 0; i < wdim; i++) {
            uint64_t pr[3];
            orc_f162_mul(eq + 3 * i, w + 3 * (i + wdim * j), pr);
            acc[0] ^= pr[0];
            acc[1] ^= pr[1];
            acc[2] ^= pr[2];
        }
        memcpy(u + 3 * j, acc, 24);
    }
    free(eq);
}

/* eval.rs:185-193: u . eq(p1) */
void orc_claim(const uint64_t *u, const uint64_t *p1, int cols_log, uint64_t *t) {
    size_t r = (size_t)1 << cols_log;
    uint64_t *eq = (uint64_t *)malloc(sizeof(uint64_t) * 3 * r);
    orc_eq_table(p1, cols_log, eq);
    uint64_t acc[3] = {0, 0, 0};
    for (size_t j = 0; j < r; j++) {
        uint64_t pr[3];
        orc_f162_mul(eq + 3 * j, u + 3 * j, pr);
        acc[0] ^= pr[0];
        acc[1] ^= pr[1];
        acc[2] ^= pr[2];
    }
    memcpy(t, acc, 24);
    free(eq);
}

/* ============================================================================================
 * the transcript and the short challenges: challenge.rs:66-221, 231-311, 332-533
 * ========================================================================================== */

typedef struct {
    b3 state;
    uint64_t counter;
} orc_transcript;

static void t_update(orc_transcript *t, const void *p, size_t n) { b3_update(&t->state, (const uint8_t *)p, n); }
static void t_u64(orc_transcript *t, uint64_t x) {
    uint8_t b[8];
    for (int i = 0; i < 8; i++) b[i] = (uint8_t)(x >> (8 * i));
    t_update(t, b, 8);
}

size_t orc_transcript_size(void) { return sizeof(orc_transcript); }

/* challenge.rs:73-78 */
void orc_transcript_new(orc_transcript *t, const uint8_t *domain, size_t len) {
    b3_new(&t->state);
    t->counter = 0;
    t_u64(t, len);
    t_update(t, domain, len);
}
/* challenge.rs:81-84 */
void orc_absorb_bytes(orc_transcript *t, const uint8_t *b, size_t len) {
    t_u64(t, len);
    t_update(t, b, len);
}
/* challenge.rs:87-89 */
void orc_absorb_u64(orc_transcript *t, uint64_t x) { t_u64(t, x); }

/* challenge.rs:153-160: the XOF a derivation reads */
static void t_reader(orc_transcript *t, const char *label, b3_out *o) {
    orc_transcript c = *t;
    t_u64(&c, t->counter);
    size_t n = strlen(label);
    t_u64(&c, n);
    t_update(&c, label, n);
    t->counter++;
    b3_final_out(&c.state, o);
}

/* challenge.rs:109-111 */
void orc_fill(orc_transcript *t, const char *label, uint8_t *out, size_t len) {
    b3_out o;
    t_reader(t, label, &o);
    b3_squeeze(&o, 0, out, len);
}

typedef struct {
    b3_out o;
    uint64_t pos;
    uint8_t buf[256];
    int bpos;
} xof_t;

static uint8_t xof_byte(xof_t *x) {
    if (x->bpos == 256) {
        b3_squeeze(&x->o, x->pos, x->buf, 256);
        x->pos += 256;
        x->bpos = 0;
    }
    return x->buf[x->bpos++];
}
/* challenge.rs:189-196 */
static uint16_t xof_u16(xof_t *x) {
    if (x->bpos + 2 <= 256) {
        uint16_t v = (uint16_t)(x->buf[x->bpos] | x->buf[x->bpos + 1] << 8);
        x->bpos += 2;
        return v;
    }
    uint8_t lo = xof_byte(x), hi = xof_byte(x);
    return (uint16_t)(lo | hi << 8);
}
/* challenge.rs:198-221 */
static uint16_t xof_below(xof_t *x, uint32_t n) {
    uint32_t limit = (65536u / n) * n;
    uint64_t magic = (1ull << 32) / n + 1;
    for (;;) {
        uint32_t r = xof_u16(x);
        if (r < limit) {
            uint32_t qt = (uint32_t)(((uint64_t)r * magic) >> 32);
            return (uint16_t)(r - qt * n);
        }
    }
}

/* challenge.rs:259-269 */
static void challenge_signed(orc_challenge *c) {
    b3 h;
    b3_new_derive(&h, "labinius 2026 challenge signs v1");
    uint8_t wb[8];
    for (int i = 0; i < 8; i++) wb[i] = (uint8_t)((uint64_t)c->weight >> (8 * i));
    b3_update(&h, wb, 8);
    b3_update(&h, c->positions, c->weight);
    b3_out o;
    b3_final_out(&h, &o);
    uint8_t d[4];
    b3_squeeze(&o, 0, d, 4);
    uint32_t bits = (uint32_t)d[0] | (uint32_t)d[1] << 8 | (uint32_t)d[2] << 16 | (uint32_t)d[3] << 24;
    uint32_t mask = (uint32_t)((1ull << c->weight) - 1);
    c->signs = bits & mask;
}

/* challenge.rs:332-360: phase table over the first 81 units */
static double PH_RE[N162][81], PH_IM[N162][81];
static int ph_ready = 0;
static void init_phase(void) {
    if (ph_ready) return;
    int units[N162], n = 0;
    for (int x = 1; x < 243; x++)
        if (x % 3 != 0) units[n++] = x;
    for (int p = 0; p < N162; p++)
        for (int k = 0; k < 81; k++) {
            int m = (p * units[k]) % 243;
            double angle = 2.0 * 3.141592653589793 * (double)m / 243.0;
            PH_RE[p][k] = cos(angle);
            PH_IM[p][k] = sin(angle);
        }
    ph_ready = 1;
}

/* challenge.rs:362-419: every root's |c(zeta^u)|^2 <= bound_sq, accumulated term by term */
static int within(const orc_challenge *c, double bound_sq) {
    init_phase();
    for (int k = 0; k < 81; k++) {
        double re = 0.0, im = 0.0;
        for (uint32_t i = 0; i < c->weight; i++) {
            int p = c->positions[i];
            if ((c->signs >> i) & 1) {
                re = re + (-PH_RE[p][k]);
                im = im + (-PH_IM[p][k]);
            } else {
                re = re + PH_RE[p][k];
                im = im + PH_IM[p][k];
            }
        }
        volatile double a = re * re, b = im * im;
        double sq = a + b;
        if (sq > bound_sq) return 0;
    }
    return 1;
}

double orc_canonical_norm_sq(const orc_challenge *c) {
    init_phase();
    double best = 0.0;
    for (int k = 0; k < 81; k++) {
        double re = 0.0, im = 0.0;
        for (uint32_t i = 0; i < c->weight; i++) {
            int p = c->positions[i];
            double s = ((c->signs >> i) & 1) ? -1.0 : 1.0;
            re += s * PH_RE[p][k];
            im += s * PH_IM[p][k];
        }
        double sq = re * re + im * im;
        if (sq > best) best = sq;
    }
    return best;
}

/* challenge.rs:478-506 */
static void attempt(xof_t *x, uint32_t weight, orc_challenge *out) {
    uint8_t perm[N162];
    for (int i = 0; i < N162; i++) perm[i] = (uint8_t)i;
    memset(out, 0, sizeof *out);
    out->weight = weight;
    for (uint32_t i = 0; i < weight; i++) {
        uint32_t j = i + xof_below(x, (uint32_t)(N162 - i));
        uint8_t tmp = perm[i];
        perm[i] = perm[j];
        perm[j] = tmp;
    }
    uint64_t seen[3] = {0, 0, 0};
    for (uint32_t i = 0; i < weight; i++) seen[perm[i] >> 6] |= 1ull << (perm[i] & 63);
    int n = 0;
    for (int wd = 0; wd < 3; wd++)
        for (int b = 0; b < 64; b++)
            if ((seen[wd] >> b) & 1) out->positions[n++] = (uint8_t)(64 * wd + b);
}

/* challenge.rs:517-533; returns the number of attempts */
uint64_t orc_sample_short_challenge(orc_transcript *t, uint32_t weight, double bound,
                                    orc_challenge *out) {
    double bound_sq = bound * bound + 1e-12;
    xof_t x;
    t_reader(t, "short-challenge", &x.o);
    x.pos = 0;
    x.bpos = 256;
    uint64_t attempts = 0;
    for (;;) {
        attempts++;
        attempt(&x, weight, out);
        challenge_signed(out);
        if (within(out, bound_sq)) return attempts;
    }
}

/* ============================================================================================
 * the protocol glue: scheme/verifier.rs:41-100, scheme/commitment.rs:217-228, 320-331
 * ========================================================================================== */

typedef struct {
    uint32_t witness_log_len, column_log_len;
    uint32_t n_primes;
    uint32_t primes[8];
    uint32_t dropped_bits;
    uint32_t recursion;
    uint8_t matrix_seed[32];
} orc_params;

/* verifier.rs:41-57 */
static void absorb_parameters(orc_transcript *t, const orc_params *p) {
    orc_absorb_bytes(t, (const uint8_t *)"labinius/parameters", 19);
    orc_absorb_u64(t, p->witness_log_len);
    orc_absorb_u64(t, p->column_log_len);
    for (uint32_t i = 0; i < p->n_primes; i++) orc_absorb_u64(t, p->primes[i]);
    orc_absorb_u64(t, p->recursion ? 1 : 0);
    if (p->dropped_bits > 0) {
        orc_absorb_bytes(t, (const uint8_t *)"labinius/dropped-bits", 21);
        orc_absorb_u64(t, p->dropped_bits);
    }
    orc_absorb_bytes(t, p->matrix_seed, 32);
}

/* verifier.rs:61-100 for a Clear commitment (comp laid out as in orc_commit) */
void orc_derive_evaluation_point(orc_transcript *t, const orc_params *p, const int16_t *comp,
                                 uint64_t *point) {
    absorb_parameters(t, p);
    orc_absorb_bytes(t, (const uint8_t *)"labinius/commitment", 19);
    size_t cols = (size_t)1 << p->column_log_len;
    orc_absorb_u64(t, cols);
    int limbs = (int)p->n_primes;
    size_t per = 2 * N162 * (size_t)limbs;
    uint8_t *buf = (uint8_t *)malloc(per);
    for (size_t c = 0; c < cols; c++) {
        t_u64(t, 4);
        for (int row = 0; row < 4; row++) {
            for (int k = 0; k < limbs; k++)
                for (int s = 0; s < N162; s++) {
                    int16_t x = comp[((c * 4 + row) * limbs + k) * N162 + s];
                    buf[2 * (k * N162 + s)] = (uint8_t)x;
                    buf[2 * (k * N162 + s) + 1] = (uint8_t)((uint16_t)x >> 8);
                }
            t_update(t, buf, per);
        }
    }
    free(buf);
    int nv = (int)(p->witness_log_len);
    uint8_t *bytes = (uint8_t *)malloc(24 * (size_t)nv);
    orc_fill(t, "labinius/evaluation-point", bytes, 24 * (size_t)nv);
    for (int e = 0; e < nv; e++) {
        for (int k = 0; k < 3; k++) {
            uint64_t x = 0;
            for (int b = 7; b >= 0; b--) x = x << 8 | bytes[24 * e + 8 * k + b];
            point[3 * e + k] = x;
        }
        point[3 * e + 2] &= (1ull << 34) - 1;
    }
    free(bytes);
}

/* commitment.rs:217-228, 320-331 */
void orc_derive_folding_challenges(orc_transcript *t, const uint64_t *u, size_t cols,
                                   orc_challenge *out) {
    orc_absorb_bytes(t, (const uint8_t *)"labinius/row-evaluation", 23);
    uint8_t *bytes = (uint8_t *)malloc(24 * cols);
    for (size_t j = 0; j < cols; j++)
        for (int k = 0; k < 3; k++)
            for (int b = 0; b < 8; b++) bytes[24 * j + 8 * k + b] = (uint8_t)(u[3 * j + k] >> (8 * b));
    orc_absorb_bytes(t, bytes, 24 * cols);
    free(bytes);
    for (size_t j = 0; j < cols; j++) orc_sample_short_challenge(t, 28, 12.0, &out[j]);
}

/* ============================================================================================
 * the verifier's checks: verifier.rs:114-231, fold.rs:799-873, eval.rs:196-249
 * ========================================================================================== */

/* The challenge's value at R_162 slot s: slot SLOT_648[0][s] of its transform (split), or the
 * scalar of leaf QUAD_CLASS_SLOT[0][s] (quadratic); fold.rs:799-830. */
static void challenge_slots(const prime_tab *t, const orc_challenge *c, uint64_t *out) {
    int32_t e[N];
    orc_embed(c, e);
    uint64_t v[N];
    for (int i = 0; i < N; i++) v[i] = (uint64_t)((e[i] % (int64_t)t->q + (int64_t)t->q) % (int64_t)t->q);
    if (t->quad) {
        ntt_quad(t, v);
        for (int s = 0; s < N162; s++) out[s] = v[2 * QUAD_CLASS_SLOT[0][s]];
    } else {
        ntt_split(t, v);
        for (int s = 0; s < N162; s++) out[s] = v[SLOT_648[0][s]];
    }
}

/* verifier.rs:114-146: out[t][k][s] = sum_j c_j . C_j[t][k][s] mod q_k, centered */
void orc_fold_commitment(const uint32_t *primes, int limbs, int cols, const int16_t *comp,
                         const orc_challenge *ch, int16_t *out) {
    init_tables();
    for (int k = 0; k < limbs; k++) {
        const prime_tab *t = prime(primes[k]);
        int64_t q = t->q;
        int64_t acc[4][N162];
        memset(acc, 0, sizeof acc);
        for (int j = 0; j < cols; j++) {
            uint64_t cs[N162];
            challenge_slots(t, &ch[j], cs);
            for (int row = 0; row < 4; row++)
                for (int s = 0; s < N162; s++) {
                    int64_t x = comp[(((size_t)j * 4 + row) * limbs + k) * N162 + s];
                    acc[row][s] = (acc[row][s] + (int64_t)cs[s] * x) % q;
                }
        }
        for (int row = 0; row < 4; row++)
            for (int s = 0; s < N162; s++) {
                int64_t x = acc[row][s] % q;
                if (x < 0) x += q;
                out[(row * limbs + k) * N162 + s] = center((uint64_t)x, (uint64_t)q);
            }
    }
}

/* eval.rs:196-200 */
void orc_fold_binary(const uint64_t *u, const orc_challenge *ch, size_t cols, uint64_t *out) {
    uint64_t acc[3] = {0, 0, 0};
    for (size_t j = 0; j < cols; j++) {
        uint64_t c[3] = {0, 0, 0};
        for (uint32_t i = 0; i < ch[j].weight; i++) c[ch[j].positions[i] >> 6] |= 1ull << (ch[j].positions[i] & 63);
        uint64_t pr[3];
        orc_f162_mul(c, u + 3 * j, pr);
        acc[0] ^= pr[0];
        acc[1] ^= pr[1];
        acc[2] ^= pr[2];
    }
    memcpy(out, acc, 24);
}

/* verifier.rs:205-224: A v per limb (v of len_ring ring elements, any int16), decomposed:
 * out[row][limb][slot], centered */
void orc_a_times_v(const uint32_t *primes, int limbs, int len_ring, const int16_t *a, const int16_t *v,
                   int16_t *out) {
    init_tables();
    for (int k = 0; k < limbs; k++) {
        const prime_tab *t = prime(primes[k]);
        uint64_t q = t->q;
        uint64_t y[N];
        memset(y, 0, sizeof y);
        for (int i = 0; i < len_ring; i++) {
            uint64_t x[N], av[N];
            for (int u = 0; u < N; u++) x[u] = (uint64_t)((v[(size_t)i * N + u] % (int64_t)q + (int64_t)q) % (int64_t)q);
            for (int u = 0; u < N; u++)
                av[u] = (uint64_t)((a[((size_t)k * len_ring + i) * N + u] % (int64_t)q + (int64_t)q) % (int64_t)q);
            if (t->quad) {
                ntt_quad(t, x);
                uint64_t pr[N];
                mul_quad_slots(t, av, x, pr);
                for (int u = 0; u < N; u++) y[u] = (y[u] + pr[u]) % q;
            } else {
                ntt_split(t, x);
                for (int u = 0; u < N; u++) y[u] = (y[u] + x[u] * av[u]) % q;
            }
        }
        uint32_t y32[N];
        for (int u = 0; u < N; u++) y32[u] = (uint32_t)y[u];
        int16_t d[4 * N162];
        orc_decompose(primes[k], y32, d);
        for (int row = 0; row < 4; row++)
            for (int s = 0; s < N162; s++) out[(row * limbs + k) * N162 + s] = d[row * N162 + s];
    }
}

/* verifier.rs:178-231 for the Clear mode: returns 1 on acceptance. fold_cap per
 * scheme/params.rs:263-265. */
int orc_verify_folded_opening(const orc_params *p, const int16_t *a, const int16_t *folded_comm,
                              const int16_t *v, const uint64_t *p0, const uint64_t *u_folded) {
    init_tables();
    int limbs = (int)p->n_primes;
    int rows = (int)(p->witness_log_len - p->column_log_len);
    int len_ring = (1 << rows) / 4;
    int32_t half = (int32_t)((p->primes[0] - 1) / 2);
    uint64_t normsq = 0;
    int32_t worst = 0;
    for (size_t i = 0; i < (size_t)len_ring * N; i++) {
        int32_t x = v[i];
        int32_t ax = x < 0 ? -x : x;
        if (ax > worst) worst = ax;
        normsq += (uint64_t)((int64_t)x * x);
    }
    double capd = ceil(108.0 * (double)(((size_t)1 << p->witness_log_len) / 4 * N));
    if (worst > half || normsq > (uint64_t)capd) return 0;
    int16_t *d = (int16_t *)malloc(sizeof(int16_t) * 4 * N162 * (size_t)limbs);
    orc_a_times_v(p->primes, limbs, len_ring, a, v, d);
    int same = memcmp(d, folded_comm, sizeof(int16_t) * 4 * N162 * (size_t)limbs) == 0;
    free(d);
    if (!same) return 0;
    /* eval.rs:202-249: eq(p0) . (v mod 2) == u' with element 4m + k = component k of v_m */
    size_t wdim = (size_t)1 << rows;
    uint64_t *eq = (uint64_t *)malloc(sizeof(uint64_t) * 3 * wdim);
    orc_eq_table(p0, rows, eq);
    uint64_t acc[3] = {0, 0, 0};
    for (int m = 0; m < len_ring; m++)
        for (int k = 0; k < 4; k++) {
            uint64_t e[3] = {0, 0, 0};
            for (int b = 0; b < 162; b++)
                if (v[(size_t)m * N + 4 * b + k] & 1) e[b / 64] |= 1ull << (b % 64);
            uint64_t pr[3];
            orc_f162_mul(eq + 3 * (4 * (size_t)m + k), e, pr);
            acc[0] ^= pr[0];
            acc[1] ^= pr[1];
            acc[2] ^= pr[2];
        }
    free(eq);
    return acc[0] == u_folded[0] && acc[1] == u_folded[1] && acc[2] == u_folded[2];
}

/* ============================================================================================
 * bit-dropped commitments: bd/mod.rs:58-77, bd/digits.rs:121-182, simd/bd.rs:5-37,
 * wire/mod.rs:77-79,386-397, bd/residual.rs:77-130, bd/digits.rs (Garner::normsq),
 * verifier.rs:61-100 (Dropped arm), verifier.rs:232-275
 * ========================================================================================== */

/* wire/mod.rs:77-79 */
uint32_t orc_residue_bits(uint32_t q) { return 32 - (uint32_t)__builtin_clz(q - 1); }

/* bd/mod.rs:58-64 */
uint32_t orc_top_bits(uint32_t q0, uint32_t d) {
    uint32_t bound = ((q0 - 1) + (1u << (d - 1))) >> d;
    return 32 - (uint32_t)__builtin_clz(bound);
}

/* bd/mod.rs:71-74, evaluated in f64 in the same order */
uint64_t orc_bd_cap(uint32_t cols, uint32_t d) {
    double spread = (double)((1ull << (2 * d)) - 1) / 12.0;
    return (uint64_t)ceil(4.0 * (double)((uint64_t)N * cols * 28) * spread);
}

/* bd/digits.rs:121-182: per limb the column's coefficients (the inverse transform of its raw row,
 * any representative mod q), then Garner's mixed radix over the primes: digits[k-1][i] = x_k, and
 * the base residue rounded to its top bits. raw is [k][c][648]; top and each digit plane are
 * [c * 648 + i]. */
void orc_drop_bits(const uint32_t *primes, int limbs, int cols, const uint32_t *raw, uint32_t d,
                   uint16_t *top, uint16_t *digits) {
    init_tables();
    size_t count = (size_t)cols * N;
    uint32_t *res = (uint32_t *)malloc(sizeof(uint32_t) * count * limbs);
    for (int k = 0; k < limbs; k++)
        for (int c = 0; c < cols; c++)
            orc_intt(primes[k], raw + ((size_t)k * cols + c) * N, res + (size_t)k * count + (size_t)c * N);
    uint32_t round = 1u << (d - 1);
    for (size_t i = 0; i < count; i++) {
        uint64_t mixed[8];
        uint32_t base = res[i] % primes[0];
        top[i] = (uint16_t)((base + round) >> d);
        mixed[0] = base;
        for (int k = 1; k < limbs; k++) {
            uint64_t q = primes[k];
            uint64_t x = res[(size_t)k * count + i] % q;
            for (int m = 0; m < k; m++) x = (x + q - mixed[m] % q) % q * inv_mod(primes[m] % q, q) % q;
            mixed[k] = x;
            digits[(size_t)(k - 1) * count + i] = (uint16_t)x;
        }
    }
    free(res);
}

/* wire/mod.rs:386-397: per coefficient the top (top_bits) then each digit (residue_bits of its
 * prime), least significant bit first, the last byte zero-padded. Returns the byte count. */
size_t orc_pack_dropped(const uint32_t *primes, int limbs, int cols, uint32_t d, const uint16_t *top,
                        const uint16_t *digits, uint8_t *out) {
    size_t count = (size_t)cols * N, len = 0;
    uint64_t acc = 0;
    uint32_t n = 0, tb = orc_top_bits(primes[0], d);
    for (size_t i = 0; i < count; i++) {
        for (int k = 0; k < limbs; k++) {
            uint32_t bits = k == 0 ? tb : orc_residue_bits(primes[k]);
            uint64_t value = k == 0 ? top[i] : digits[(size_t)(k - 1) * count + i];
            acc |= (value & ((1ull << bits) - 1)) << n;
            n += bits;
            while (n >= 8) {
                out[len++] = (uint8_t)acc;
                acc >>= 8;
                n -= 8;
            }
        }
    }
    if (n > 0) out[len++] = (uint8_t)acc;
    return len;
}

/* simd/bd.rs:5-37: (top << d) + q0 (x_1 + q1 (x_2 + ...)) mod q, centered */
static int32_t bd_combine(const uint32_t *primes, int limbs, uint32_t d, const uint16_t *top,
                          const uint16_t *digits, size_t count, size_t i, uint32_t q) {
    uint64_t inner = digits[(size_t)(limbs - 2) * count + i];
    for (int m = limbs - 3; m >= 0; m--) inner = (digits[(size_t)m * count + i] + inner * primes[m + 1]) % q;
    uint64_t r = (((uint64_t)top[i] << d) + inner * primes[0]) % q;
    return r > (q - 1) / 2 ? (int32_t)r - (int32_t)q : (int32_t)r;
}

/* slot-wise product of two transforms mod q: a product per slot (split) or per leaf (quadratic) */
static void mul_slots(const prime_tab *t, const uint64_t *a, const uint64_t *b, uint64_t *c) {
    if (t->quad)
        mul_quad_slots(t, a, b, c);
    else
        for (int u = 0; u < N; u++) c[u] = a[u] * b[u] % t->q;
}

static void forward_mod(const prime_tab *t, const int64_t *x, uint64_t *v) {
    for (int u = 0; u < N; u++) v[u] = (uint64_t)((x[u] % (int64_t)t->q + (int64_t)t->q) % (int64_t)t->q);
    if (t->quad)
        ntt_quad(t, v);
    else
        ntt_split(t, v);
}

/* bd/residual.rs:77-130 with Garner::normsq (bd/digits.rs): per limb k,
 * z_k = A_k v - sum_j c_j(-X^4) col_j mod q_k in coefficients, col_j rebuilt from (top, digits);
 * then sum over the 648 coefficients of the square of the CRT value centered mod prod q.
 * Returns the norm as two u64 halves. */
void orc_bd_residual(const uint32_t *primes, int limbs, int len_ring, int cols, const int16_t *a,
                     uint32_t d, const uint16_t *top, const uint16_t *digits, const orc_challenge *ch,
                     const int16_t *v, uint64_t *normsq_lo, uint64_t *normsq_hi) {
    init_tables();
    size_t count = (size_t)cols * N;
    uint32_t (*cen)[N] = (uint32_t(*)[N])malloc(sizeof(uint32_t) * N * limbs);
    for (int k = 0; k < limbs; k++) {
        const prime_tab *t = prime(primes[k]);
        uint64_t q = t->q, fc[N], y[N], pr[N], x[N], cv[N];
        int64_t tmp[N];
        memset(fc, 0, sizeof fc);
        memset(y, 0, sizeof y);
        for (int j = 0; j < cols; j++) {
            int32_t e[N];
            orc_embed(&ch[j], e);
            for (int u = 0; u < N; u++) tmp[u] = e[u];
            forward_mod(t, tmp, cv);
            for (int u = 0; u < N; u++)
                tmp[u] = bd_combine(primes, limbs, d, top, digits, count, (size_t)j * N + u, (uint32_t)q);
            forward_mod(t, tmp, x);
            mul_slots(t, cv, x, pr);
            for (int u = 0; u < N; u++) fc[u] = (fc[u] + pr[u]) % q;
        }
        for (int i = 0; i < len_ring; i++) {
            uint64_t av[N];
            for (int u = 0; u < N; u++) {
                tmp[u] = v[(size_t)i * N + u];
                av[u] = (uint64_t)((a[((size_t)k * len_ring + i) * N + u] % (int64_t)q + (int64_t)q) % (int64_t)q);
            }
            forward_mod(t, tmp, x);
            mul_slots(t, av, x, pr);
            for (int u = 0; u < N; u++) y[u] = (y[u] + pr[u]) % q;
        }
        uint32_t z[N];
        for (int u = 0; u < N; u++) z[u] = (uint32_t)((y[u] + q - fc[u]) % q);
        orc_intt(primes[k], z, cen[k]);
    }
    unsigned __int128 modulus = 1, normsq = 0;
    for (int k = 0; k < limbs; k++) modulus *= primes[k];
    for (int i = 0; i < N; i++) {
        uint64_t mixed[8];
        for (int k = 0; k < limbs; k++) {
            uint64_t q = primes[k], x = cen[k][i] % q;
            for (int m = 0; m < k; m++) x = (x + q - mixed[m] % q) % q * inv_mod(primes[m] % q, q) % q;
            mixed[k] = x;
        }
        unsigned __int128 value = 0;
        for (int k = limbs - 1; k >= 0; k--) value = value * primes[k] + mixed[k];
        unsigned __int128 mag = value > modulus / 2 ? modulus - value : value;
        normsq += mag * mag;
    }
    free(cen);
    *normsq_lo = (uint64_t)normsq;
    *normsq_hi = (uint64_t)(normsq >> 64);
}

/* verifier.rs:61-100, the Dropped arm */
void orc_derive_evaluation_point_dropped(orc_transcript *t, const orc_params *p, const uint8_t *packed,
                                         size_t packed_len, uint64_t *point) {
    absorb_parameters(t, p);
    orc_absorb_bytes(t, (const uint8_t *)"labinius/commitment", 19);
    orc_absorb_u64(t, (uint64_t)1 << p->column_log_len);
    orc_absorb_bytes(t, (const uint8_t *)"labinius/dropped-commitment", 27);
    orc_absorb_bytes(t, packed, packed_len);
    int nv = (int)(p->witness_log_len);
    uint8_t *bytes = (uint8_t *)malloc(24 * (size_t)nv);
    orc_fill(t, "labinius/evaluation-point", bytes, 24 * (size_t)nv);
    for (int e = 0; e < nv; e++) {
        for (int k = 0; k < 3; k++) {
            uint64_t x = 0;
            for (int b = 7; b >= 0; b--) x = x << 8 | bytes[24 * e + 8 * k + b];
            point[3 * e + k] = x;
        }
        point[3 * e + 2] &= (1ull << 34) - 1;
    }
    free(bytes);
}

/* verifier.rs:232-275: the norm checks on v, the residual against bd_cap, the binary check.
 * Returns 1 on acceptance. */
int orc_verify_opening_bd(const orc_params *p, const int16_t *a, const uint16_t *top, const uint16_t *digits,
                          const orc_challenge *ch, const int16_t *v, const uint64_t *p0,
                          const uint64_t *u_folded) {
    init_tables();
    int limbs = (int)p->n_primes;
    int rows = (int)(p->witness_log_len - p->column_log_len);
    int len_ring = (1 << rows) / 4;
    int cols = 1 << p->column_log_len;
    int32_t half = (int32_t)((p->primes[0] - 1) / 2);
    uint64_t normsq = 0;
    int32_t worst = 0;
    for (size_t i = 0; i < (size_t)len_ring * N; i++) {
        int32_t x = v[i];
        int32_t ax = x < 0 ? -x : x;
        if (ax > worst) worst = ax;
        normsq += (uint64_t)((int64_t)x * x);
    }
    double capd = ceil(108.0 * (double)(((size_t)1 << p->witness_log_len) / 4 * N));
    if (worst > half || normsq > (uint64_t)capd) return 0;
    uint64_t lo, hi;
    orc_bd_residual(p->primes, limbs, len_ring, cols, a, p->dropped_bits, top, digits, ch, v, &lo, &hi);
    if (hi != 0 || lo > orc_bd_cap((uint32_t)cols, p->dropped_bits)) return 0;
    size_t wdim = (size_t)1 << rows;
    uint64_t *eq = (uint64_t *)malloc(sizeof(uint64_t) * 3 * wdim);
    orc_eq_table(p0, rows, eq);
    uint64_t acc[3] = {0, 0, 0};
    for (int m = 0; m < len_ring; m++)
        for (int k = 0; k < 4; k++) {
            uint64_t e[3] = {0, 0, 0};
            for (int b = 0; b < 162; b++)
                if (v[(size_t)m * N + 4 * b + k] & 1) e[b / 64] |= 1ull << (b % 64);
            uint64_t pr[3];
            orc_f162_mul(eq + 3 * (4 * (size_t)m + k), e, pr);
            acc[0] ^= pr[0];
            acc[1] ^= pr[1];
            acc[2] ^= pr[2];
        }
    free(eq);
    return acc[0] == u_folded[0] && acc[1] == u_folded[1] && acc[2] == u_folded[2];
}

/* ============================================================================================
 * the cross-field switch: switch.rs:40-209, fields/crossfield.rs:9-115 and 327-560,
 * fields/sumcheck.rs:100-145. B128 = GF(2)[x] / (x^128 + x^7 + x^2 + x + 1) (scalar.rs:4, 41-63)
 * as two u64, bit i the coefficient of x^i; F162 as in orc_f162_mul.
 * ========================================================================================== */

/* bit by bit: the 255-bit product, then x^128 = x^7 + x^2 + x + 1 from the top down */
void orc_b128_mul(const uint64_t *a, const uint64_t *b, uint64_t *out) {
    uint64_t p[4] = {0, 0, 0, 0};
    for (int i = 0; i < 128; i++) {
        if (!((a[i / 64] >> (i % 64)) & 1)) continue;
        int ws = i / 64, bs = i % 64;
        for (int k = 0; k < 2; k++) {
            p[k + ws] ^= b[k] << bs;
            if (bs) p[k + ws + 1] ^= b[k] >> (64 - bs);
        }
    }
    for (int j = 254; j >= 128; j--)
        if ((p[j / 64] >> (j % 64)) & 1) {
            const int t[5] = {j, j - 128, j - 127, j - 126, j - 121};
            for (int u = 0; u < 5; u++) p[t[u] / 64] ^= 1ull << (t[u] % 64);
        }
    out[0] = p[0];
    out[1] = p[1];
}

/* crossfield.rs:85-99: eq(r, j) with coordinate 0 the most significant bit of j, doubled in place
 * from the back */
void orc_eq_b128(const uint64_t *r, int n, uint64_t *t) {
    t[0] = 1;
    t[1] = 0;
    for (int round = 0; round < n; round++)
        for (long i = (1L << round) - 1; i >= 0; i--) {
            uint64_t hi[2];
            orc_b128_mul(t + 2 * i, r + 2 * round, hi);
            const uint64_t x0 = t[2 * i], x1 = t[2 * i + 1];
            t[2 * (2 * i + 1)] = hi[0];
            t[2 * (2 * i + 1) + 1] = hi[1];
            t[2 * (2 * i)] = x0 ^ hi[0];
            t[2 * (2 * i) + 1] = x1 ^ hi[1];
        }
}

/* crossfield.rs:66-78: the same order over F162 */
void orc_eq_f162_msb(const uint64_t *r, int n, uint64_t *t) {
    t[0] = 1;
    t[1] = t[2] = 0;
    for (int round = 0; round < n; round++)
        for (long i = (1L << round) - 1; i >= 0; i--) {
            uint64_t hi[3], x[3] = {t[3 * i], t[3 * i + 1], t[3 * i + 2]};
            orc_f162_mul(x, r + 3 * round, hi);
            for (int k = 0; k < 3; k++) {
                t[3 * (2 * i + 1) + k] = hi[k];
                t[3 * (2 * i) + k] = x[k] ^ hi[k];
            }
        }
}

/* switch.rs:40-46: sum_j trace[j] eq(r, j) */
void orc_b128_mle(const uint64_t *trace, int l, const uint64_t *r, uint64_t *out) {
    const size_t n = (size_t)1 << l;
    uint64_t *eq = (uint64_t *)malloc(16 * n);
    orc_eq_b128(r, l, eq);
    uint64_t acc[2] = {0, 0};
    for (size_t j = 0; j < n; j++) {
        uint64_t pr[2];
        orc_b128_mul(trace + 2 * j, eq + 2 * j, pr);
        acc[0] ^= pr[0];
        acc[1] ^= pr[1];
    }
    free(eq);
    out[0] = acc[0];
    out[1] = acc[1];
}

/* switch.rs:49-62: bit k of word j at index k + 128 j */
void orc_bit_mle(const uint64_t *trace, int l, const uint64_t *r_lo, const uint64_t *r_hi, uint64_t *out) {
    const size_t n = (size_t)1 << l;
    uint64_t eq_lo[2 * 128];
    uint64_t *eq_hi = (uint64_t *)malloc(16 * n);
    orc_eq_b128(r_lo, 7, eq_lo);
    orc_eq_b128(r_hi, l, eq_hi);
    uint64_t acc[2] = {0, 0};
    for (size_t j = 0; j < n; j++) {
        uint64_t bits[2] = {0, 0}, pr[2];
        for (int k = 0; k < 128; k++)
            if ((trace[2 * j + k / 64] >> (k % 64)) & 1) {
                bits[0] ^= eq_lo[2 * k];
                bits[1] ^= eq_lo[2 * k + 1];
            }
        orc_b128_mul(bits, eq_hi + 2 * j, pr);
        acc[0] ^= pr[0];
        acc[1] ^= pr[1];
    }
    free(eq_hi);
    out[0] = acc[0];
    out[1] = acc[1];
}

/* crossfield.rs:9-62,106-115: a 128 x F162 tensor (horizontal index = the B128 basis x^i) */
static void tensor_shift_x(uint64_t c[128][3]) {
    uint64_t top[3] = {c[127][0], c[127][1], c[127][2]};
    memmove(c[1], c[0], 127 * 24);
    c[0][0] = c[0][1] = c[0][2] = 0;
    const int taps[4] = {0, 1, 2, 7};
    for (int u = 0; u < 4; u++)
        for (int w = 0; w < 3; w++) c[taps[u]][w] ^= top[w];
}

void orc_transparent_coeff(const uint64_t *r_hi, const uint64_t *r_pp, int l, const uint64_t *batch, uint64_t *out) {
    static uint64_t t[128][3], vs[128][3], hs[128][3];
    memset(t, 0, sizeof t);
    t[0][0] = 1;
    for (int round = 0; round < l; round++) {
        const uint64_t *v = r_pp + 3 * round, *h = r_hi + 2 * round;
        for (int i = 0; i < 128; i++) orc_f162_mul(t[i], v, vs[i]);
        memset(hs, 0, sizeof hs);
        for (int k = 127; k >= 0; k--) {
            tensor_shift_x(hs);
            if ((h[k / 64] >> (k % 64)) & 1)
                for (int i = 0; i < 128; i++)
                    for (int w = 0; w < 3; w++) hs[i][w] ^= t[i][w];
        }
        for (int i = 0; i < 128; i++)
            for (int w = 0; w < 3; w++) t[i][w] ^= vs[i][w] ^ hs[i][w];
    }
    uint64_t acc[3] = {0, 0, 0};
    for (int i = 0; i < 128; i++) {
        uint64_t pr[3];
        orc_f162_mul(t[i], batch + 3 * i, pr);
        for (int w = 0; w < 3; w++) acc[w] ^= pr[w];
    }
    memcpy(out, acc, 24);
}

/* challenge.rs:113-149: length-prefixed absorbs; F162 samples cut to 34 top bits */
void orc_absorb_b128(orc_transcript *t, const uint64_t *xs, size_t n) {
    t_u64(t, (uint64_t)n);
    for (size_t i = 0; i < 2 * n; i++) t_u64(t, xs[i]);
}

void orc_absorb_f162(orc_transcript *t, const uint64_t *xs, size_t n) {
    t_u64(t, (uint64_t)n);
    for (size_t i = 0; i < 3 * n; i++) t_u64(t, xs[i]);
}

void orc_sample_f162(orc_transcript *t, const char *label, size_t n, uint64_t *out) {
    uint8_t *b = (uint8_t *)malloc(24 * n + 1);
    orc_fill(t, label, b, 24 * n);
    for (size_t i = 0; i < 3 * n; i++) {
        uint64_t x = 0;
        for (int k = 0; k < 8; k++) x |= (uint64_t)b[8 * i + k] << (8 * k);
        out[i] = (i % 3 == 2) ? (x & ((1ull << 34) - 1)) : x;
    }
    free(b);
}

void orc_sample_b128(orc_transcript *t, const char *label, size_t n, uint64_t *out) {
    uint8_t *b = (uint8_t *)malloc(16 * n + 1);
    orc_fill(t, label, b, 16 * n);
    for (size_t i = 0; i < 2 * n; i++) {
        uint64_t x = 0;
        for (int k = 0; k < 8; k++) x |= (uint64_t)b[8 * i + k] << (8 * k);
        out[i] = x;
    }
    free(b);
}

/* switch.rs:129-134 */
static void switch_statement(orc_transcript *t, const uint64_t *r_lo, int r_lo_len, const uint64_t *r_hi, int l,
                             const uint64_t *claim) {
    orc_absorb_bytes(t, (const uint8_t *)"labinius/switch", 15);
    orc_absorb_b128(t, r_lo, (size_t)r_lo_len);
    orc_absorb_b128(t, r_hi, (size_t)l);
    orc_absorb_b128(t, claim, 1);
}

/* EvaluationPoint::msb_first (commitment.rs:163-174): the first cl coordinates reversed are p1,
 * the rest reversed p0 */
static void msb_first(const uint64_t *r_pp, int l, int cl, uint64_t *p0, uint64_t *p1) {
    for (int i = 0; i < cl; i++) memcpy(p1 + 3 * i, r_pp + 3 * (cl - 1 - i), 24);
    for (int i = 0; i < l - cl; i++) memcpy(p0 + 3 * i, r_pp + 3 * (l - 1 - i), 24);
}

/* switch.rs:137-164 with crossfield.rs:410-476 and sumcheck.rs:100-145, by plain loops. trace is
 * 2^l B128; r_lo_len is 0 (a claim on the words) or 7 (on their bits, weighed in the verifier);
 * partials 128 B128, rounds l x [e0, einf], opened and the point's p0 (l - cl) and p1 (cl) F162. */
void orc_switch_prove(const uint64_t *trace, int l, int cl, const uint64_t *r_lo, int r_lo_len, const uint64_t *r_hi,
                      const uint64_t *claim, orc_transcript *t, uint64_t *partials, uint64_t *rounds,
                      uint64_t *opened, uint64_t *p0, uint64_t *p1) {
    const size_t n = (size_t)1 << l;
    switch_statement(t, r_lo, r_lo_len, r_hi, l, claim);
    uint64_t *eq = (uint64_t *)malloc(16 * n);
    orc_eq_b128(r_hi, l, eq);
    /* crossfield.rs:154-158: v[k] = sum_j bit_k(trace[j]) eq_hi[j] */
    memset(partials, 0, 16 * 128);
    for (size_t j = 0; j < n; j++)
        for (int k = 0; k < 128; k++)
            if ((trace[2 * j + k / 64] >> (k % 64)) & 1) {
                partials[2 * k] ^= eq[2 * j];
                partials[2 * k + 1] ^= eq[2 * j + 1];
            }
    orc_absorb_b128(t, partials, 128);
    uint64_t r7[3 * 7], batch[3 * 128];
    orc_sample_f162(t, "labinius/switch-batch", 7, r7);
    orc_eq_f162_msb(r7, 7, batch);
    /* a[j] = sum_k bit_k(eq_hi[j]) batch[k] (crossfield.rs:285-325); p[j] = from_b128(trace[j]) */
    uint64_t *a = (uint64_t *)calloc(3 * n, 8), *p = (uint64_t *)calloc(3 * n, 8);
    for (size_t j = 0; j < n; j++) {
        for (int k = 0; k < 128; k++)
            if ((eq[2 * j + k / 64] >> (k % 64)) & 1)
                for (int w = 0; w < 3; w++) a[3 * j + w] ^= batch[3 * k + w];
        p[3 * j] = trace[2 * j];
        p[3 * j + 1] = trace[2 * j + 1];
    }
    uint64_t *r_pp = (uint64_t *)malloc(24 * (size_t)(l > 0 ? l : 1));
    size_t half = n;
    for (int round = 0; round < l; round++) {
        half /= 2;
        uint64_t e0[3] = {0, 0, 0}, einf[3] = {0, 0, 0};
        for (size_t j = 0; j < half; j++) {
            uint64_t pr[3], da[3], dp[3];
            orc_f162_mul(a + 3 * j, p + 3 * j, pr);
            for (int w = 0; w < 3; w++) {
                e0[w] ^= pr[w];
                da[w] = a[3 * j + w] ^ a[3 * (j + half) + w];
                dp[w] = p[3 * j + w] ^ p[3 * (j + half) + w];
            }
            orc_f162_mul(da, dp, pr);
            for (int w = 0; w < 3; w++) einf[w] ^= pr[w];
        }
        uint64_t msg[6] = {e0[0], e0[1], e0[2], einf[0], einf[1], einf[2]};
        memcpy(rounds + 6 * round, msg, 48);
        orc_absorb_f162(t, msg, 2);
        uint64_t *r = r_pp + 3 * round;
        orc_sample_f162(t, "labinius/switch-round", 1, r);
        for (size_t j = 0; j < half; j++) {
            uint64_t da[3], dp[3], ma[3], mp[3];
            for (int w = 0; w < 3; w++) {
                da[w] = a[3 * j + w] ^ a[3 * (j + half) + w];
                dp[w] = p[3 * j + w] ^ p[3 * (j + half) + w];
            }
            orc_f162_mul(r, da, ma);
            orc_f162_mul(r, dp, mp);
            for (int w = 0; w < 3; w++) {
                a[3 * j + w] ^= ma[w];
                p[3 * j + w] ^= mp[w];
            }
        }
    }
    memcpy(opened, p, 24);
    msb_first(r_pp, l, cl, p0, p1);
    free(eq);
    free(a);
    free(p);
    free(r_pp);
}

/* switch.rs:166-209, crossfield.rs:477-541: 1 accepts (and fills the point), 0 rejects */
int orc_switch_verify(int l, int cl, const uint64_t *r_lo, int r_lo_len, const uint64_t *r_hi, const uint64_t *claim,
                      const uint64_t *partials, const uint64_t *rounds, const uint64_t *opened, orc_transcript *t,
                      uint64_t *p0, uint64_t *p1) {
    switch_statement(t, r_lo, r_lo_len, r_hi, l, claim);
    /* the weights: x^k for a claim on words, eq(r_lo, k) for one on bits (switch.rs:114-121) */
    uint64_t wts[2 * 128];
    if (r_lo_len == 0) {
        memset(wts, 0, sizeof wts);
        for (int k = 0; k < 128; k++) wts[2 * k + k / 64] = 1ull << (k % 64);
    } else {
        orc_eq_b128(r_lo, 7, wts);
    }
    uint64_t acc[2] = {0, 0};
    for (int k = 0; k < 128; k++) {
        uint64_t pr[2];
        orc_b128_mul(partials + 2 * k, wts + 2 * k, pr);
        acc[0] ^= pr[0];
        acc[1] ^= pr[1];
    }
    if (acc[0] != claim[0] || acc[1] != claim[1]) return 0;
    orc_absorb_b128(t, partials, 128);
    uint64_t r7[3 * 7], batch[3 * 128];
    orc_sample_f162(t, "labinius/switch-batch", 7, r7);
    orc_eq_f162_msb(r7, 7, batch);
    /* slice_sum (crossfield.rs:478-490): u_k = sum_i bit_k(v_i) x^i, s = sum_k from_b128(u_k) batch[k] */
    uint64_t s[3] = {0, 0, 0};
    for (int k = 0; k < 128; k++) {
        uint64_t u[3] = {0, 0, 0}, pr[3];
        for (int i = 0; i < 128; i++)
            if ((partials[2 * i + k / 64] >> (k % 64)) & 1) u[i / 64] |= 1ull << (i % 64);
        orc_f162_mul(u, batch + 3 * k, pr);
        for (int w = 0; w < 3; w++) s[w] ^= pr[w];
    }
    uint64_t *r_pp = (uint64_t *)malloc(24 * (size_t)(l > 0 ? l : 1));
    for (int round = 0; round < l; round++) {
        const uint64_t *e0 = rounds + 6 * round, *einf = rounds + 6 * round + 3;
        orc_absorb_f162(t, rounds + 6 * round, 2);
        uint64_t *r = r_pp + 3 * round;
        orc_sample_f162(t, "labinius/switch-round", 1, r);
        /* s = e0 + r (e0 + e1 + einf) + r^2 einf with e1 = s + e0 */
        uint64_t sum[3], t1[3], r2[3], t2[3];
        for (int w = 0; w < 3; w++) sum[w] = e0[w] ^ (s[w] ^ e0[w]) ^ einf[w];
        orc_f162_mul(r, sum, t1);
        orc_f162_mul(r, r, r2);
        orc_f162_mul(r2, einf, t2);
        for (int w = 0; w < 3; w++) s[w] = e0[w] ^ t1[w] ^ t2[w];
    }
    uint64_t d[3], dv[3];
    orc_transparent_coeff(r_hi, r_pp, l, batch, d);
    int ok = (d[0] | d[1] | d[2]) != 0;
    orc_f162_mul(d, opened, dv);
    ok = ok && dv[0] == s[0] && dv[1] == s[1] && dv[2] == s[2];
    if (ok) msb_first(r_pp, l, cl, p0, p1);
    free(r_pp);
    return ok;
}

/* ============================================================================================
 * wire formats (wire/mod.rs, wire/rans.rs): the bit-packed commitment and row evaluation, the
 * rANS-coded folded witness. Errors: 0 ok, 1 Truncated, 2 Malformed (wire/mod.rs:67-74).
 * ========================================================================================= */

enum { W_OK = 0, W_TRUNCATED = 1, W_MALFORMED = 2, W_CAPACITY = 3 };
#define W_SCALE 4096u      /* wire/mod.rs:57-58 */
#define W_RANS_LOW 65536u  /* wire/mod.rs:60 */
#define W_HEADER 16
#define W_LANES 64         /* rans.rs:12 */
#define W_MAX_ELEMENTS (1u << 24)

/* The writer of wire/mod.rs:86-138: fields least significant bit first, the last byte padded. */
typedef struct {
    uint8_t *out;
    size_t len;
    uint64_t acc;
    uint32_t n;
} bitw;

static void bw_put(bitw *w, uint64_t value, uint32_t bits) {
    w->acc |= (value & ((1ull << bits) - 1)) << w->n;
    w->n += bits;
    while (w->n >= 8) {
        w->out[w->len++] = (uint8_t)w->acc;
        w->acc >>= 8;
        w->n -= 8;
    }
}

static void bw_put_wide(bitw *w, uint64_t value, uint32_t bits) {
    if (bits <= 32) {
        bw_put(w, value, bits);
    } else {
        bw_put(w, value, 32);
        bw_put(w, value >> 32, bits - 32);
    }
}

/* Elias gamma of v >= 1: k zeros, the leading one, the k bits below it (wire/mod.rs:122-129) */
static void bw_gamma(bitw *w, uint64_t v) {
    uint32_t k = 63 - (uint32_t)__builtin_clzll(v);
    bw_put(w, 0, k);
    bw_put(w, 1, 1);
    bw_put_wide(w, v & ((1ull << k) - 1), k);
}

static size_t bw_finish(bitw *w) {
    if (w->n > 0) w->out[w->len++] = (uint8_t)w->acc;
    return w->len;
}

/* The reader of wire/mod.rs:141-218 as a bit cursor: a read of `bits` fails (Truncated) exactly
 * when fewer remain, which is when the reference's refilled accumulator comes up short; bytes
 * consumed = ceil(bits read / 8) (its pos - n / 8). */
typedef struct {
    const uint8_t *b;
    size_t len;
    uint64_t bit;
} bitr;

static int br_get(bitr *r, uint32_t bits, uint64_t *v) {
    if (r->bit + bits > 8 * (uint64_t)r->len) return W_TRUNCATED;
    uint64_t x = 0;
    for (uint32_t i = 0; i < bits; i++, r->bit++) x |= (uint64_t)((r->b[r->bit >> 3] >> (r->bit & 7)) & 1) << i;
    *v = x;
    return W_OK;
}

static int br_gamma(bitr *r, uint64_t *v) {
    uint32_t k = 0;
    uint64_t b;
    for (;;) {
        int e = br_get(r, 1, &b);
        if (e) return e;
        if (b) break;
        if (++k > 32) return W_MALFORMED;
    }
    uint64_t low;
    int e = br_get(r, k, &low);
    if (e) return e;
    *v = (1ull << k) | low;
    return W_OK;
}

static size_t br_consumed(const bitr *r) { return (size_t)((r->bit + 7) / 8); }

/* wire/mod.rs:220-225 (commitment_bytes) */
size_t orc_commitment_bytes(const uint32_t *primes, int limbs, int cols) {
    uint64_t bits = 0;
    for (int k = 0; k < limbs; k++) bits += orc_residue_bits(primes[k]);
    return (size_t)((4ull * cols * N162 * bits + 7) / 8);
}

/* wire/mod.rs:226-243: comp [c][t][k][s] centered; the entries in VerticallyAlignedMatrix::iter
 * order (column by column, rows 0..3), each limb's 162 slots as residues in [0, q) */
size_t orc_pack_commitment(const uint32_t *primes, int limbs, int cols, const int16_t *comp, uint8_t *out) {
    bitw w = {out, 0, 0, 0};
    for (size_t e = 0; e < 4 * (size_t)cols; e++)
        for (int k = 0; k < limbs; k++) {
            const int32_t q = (int32_t)primes[k];
            const uint32_t bits = orc_residue_bits(primes[k]);
            for (int s = 0; s < N162; s++) {
                int32_t v = comp[(e * limbs + k) * N162 + s] % q;
                if (v < 0) v += q;
                bw_put(&w, (uint64_t)v, bits);
            }
        }
    return bw_finish(&w);
}

/* wire/mod.rs:246-301 (the Clear arm) */
int orc_unpack_commitment(const uint32_t *primes, int limbs, int cols, const uint8_t *bytes, size_t len,
                          int16_t *comp) {
    if (len != orc_commitment_bytes(primes, limbs, cols)) return W_MALFORMED;
    bitr r = {bytes, len, 0};
    for (size_t e = 0; e < 4 * (size_t)cols; e++)
        for (int k = 0; k < limbs; k++) {
            const uint32_t q = primes[k], bits = orc_residue_bits(q), half = (q - 1) / 2;
            for (int s = 0; s < N162; s++) {
                uint64_t v;
                int err = br_get(&r, bits, &v);
                if (err) return err;
                if (v >= q) return W_MALFORMED;
                comp[(e * limbs + k) * N162 + s] = (int16_t)(v > half ? (int32_t)v - (int32_t)q : (int32_t)v);
            }
        }
    return W_OK;
}

/* wire/mod.rs:399-431 */
int orc_unpack_dropped(const uint32_t *primes, int limbs, int cols, uint32_t d, const uint8_t *bytes, size_t len,
                       uint16_t *top, uint16_t *digits) {
    const size_t count = (size_t)cols * N;
    uint64_t per = orc_top_bits(primes[0], d);
    for (int k = 1; k < limbs; k++) per += orc_residue_bits(primes[k]);
    if (d == 0 || len != (count * per + 7) / 8) return W_MALFORMED;
    const uint32_t tb = orc_top_bits(primes[0], d), bound = ((primes[0] - 1) + (1u << (d - 1))) >> d;
    bitr r = {bytes, len, 0};
    for (size_t i = 0; i < count; i++) {
        uint64_t v;
        int err = br_get(&r, tb, &v);
        if (err) return err;
        if (v > bound) return W_MALFORMED;
        top[i] = (uint16_t)v;
        for (int k = 1; k < limbs; k++) {
            if ((err = br_get(&r, orc_residue_bits(primes[k]), &v))) return err;
            if (v >= primes[k]) return W_MALFORMED;
            digits[(size_t)(k - 1) * count + i] = (uint16_t)v;
        }
    }
    return W_OK;
}

/* wire/mod.rs:445-453: 64 + 64 + 34 bits an element */
size_t orc_pack_f162(const uint64_t *x, size_t n, uint8_t *out) {
    bitw w = {out, 0, 0, 0};
    for (size_t i = 0; i < n; i++) {
        bw_put_wide(&w, x[3 * i], 64);
        bw_put_wide(&w, x[3 * i + 1], 64);
        bw_put_wide(&w, x[3 * i + 2], 34);
    }
    return bw_finish(&w);
}

/* wire/mod.rs:456-470 */
int orc_unpack_f162(const uint8_t *bytes, size_t len, size_t n, uint64_t *x) {
    if (len != (n * 162 + 7) / 8) return W_MALFORMED;
    bitr r = {bytes, len, 0};
    for (size_t i = 0; i < n; i++) {
        int err;
        if ((err = br_get(&r, 64, &x[3 * i])) || (err = br_get(&r, 64, &x[3 * i + 1])) ||
            (err = br_get(&r, 34, &x[3 * i + 2])))
            return err;
    }
    return W_OK;
}

/* rans.rs:220-223 */
static uint32_t raw_bits(uint32_t length) { return length <= 1 ? 0 : 32 - (uint32_t)__builtin_clz(length - 1); }

typedef struct {
    uint64_t key;
    size_t i;
} ranked;

/* larger key first, then the smaller index */
static int ranked_cmp(const void *a, const void *b) {
    const ranked *x = (const ranked *)a, *y = (const ranked *)b;
    if (x->key != y->key) return x->key > y->key ? -1 : 1;
    return x->i < y->i ? -1 : (x->i > y->i);
}

/* rans.rs:236-271: counts[0 .. n - 1) the symbols, counts[n - 1] the escape -> freq summing to
 * 4096 (or all zero for an empty message) */
void orc_rans_quantize(const uint64_t *counts, size_t n, uint32_t *freq) {
    const size_t esc = n - 1;
    uint64_t *w = (uint64_t *)malloc(n * sizeof(uint64_t));
    ranked *rk = (ranked *)malloc(n * sizeof(ranked));
    memcpy(w, counts, n * sizeof(uint64_t));
    size_t occ = 0;
    for (size_t i = 0; i < esc; i++)
        if (w[i]) rk[occ++] = (ranked){w[i], i};
    if (occ > W_SCALE - 1) {
        qsort(rk, occ, sizeof(ranked), ranked_cmp);
        for (size_t j = W_SCALE - 1; j < occ; j++) {
            w[esc] += w[rk[j].i];
            w[rk[j].i] = 0;
        }
    }
    uint64_t total = 0;
    occ = 0;
    for (size_t i = 0; i < n; i++) {
        total += w[i];
        freq[i] = 0;
        if (w[i]) occ++;
    }
    if (occ) {
        const uint64_t spare = W_SCALE - occ;
        uint64_t used = 0;
        size_t m = 0;
        for (size_t i = 0; i < n; i++) {
            if (!w[i]) continue;
            const uint64_t share = w[i] * spare / total;
            freq[i] = 1 + (uint32_t)share;
            used += share;
            rk[m++] = (ranked){w[i] * spare - share * total, i};
        }
        qsort(rk, m, sizeof(ranked), ranked_cmp);
        for (uint64_t j = 0; j < spare - used; j++) freq[rk[j].i]++;
    }
    free(rk);
    free(w);
}

/* A bound on orc_rans_encode's output for `count` ring elements. */
size_t orc_rans_bound(size_t count) { return W_HEADER + (65537 * 25 + 7) / 8 + 4 * W_LANES + 4 * count * N + 2; }

/* The rANS step of rans.rs:306-333 written with the exact quotient: x -> (x / f) 2^12 + x mod f +
 * start, after emitting the low word when x >= f 2^20. (The reference's reciprocal and bias
 * compute the same quotient.) */
static void put_symbol(uint32_t *x, uint16_t *words, size_t *top, uint32_t f, uint32_t start) {
    if ((*x >> 20) >= f) {
        words[(*top)++] = (uint16_t)*x;
        *x >>= 16;
    }
    *x = ((*x / f) << 12) + *x % f + start;
}

static void put_bits(uint32_t *x, uint16_t *words, size_t *top, uint32_t value, uint32_t bits) {
    if (!bits) return;
    if ((*x >> (32 - bits)) != 0) {
        words[(*top)++] = (uint16_t)*x;
        *x >>= 16;
    }
    *x = (*x << bits) | value;
}

static void le32(uint8_t *p, uint32_t v) {
    for (int i = 0; i < 4; i++) p[i] = (uint8_t)(v >> (8 * i));
}
static uint32_t rd32(const uint8_t *p) { return p[0] | (uint32_t)p[1] << 8 | (uint32_t)p[2] << 16 | (uint32_t)p[3] << 24; }

/* rans.rs:18-99: v is `count` ring elements of 648 i16, flattened */
size_t orc_rans_encode(const int16_t *v, size_t count, int ntt, uint16_t base_q, uint8_t *out) {
    const size_t coefs = count * N;
    out[0] = 1;
    out[1] = ntt ? 1 : 0;
    out[2] = (uint8_t)base_q;
    out[3] = (uint8_t)(base_q >> 8);
    le32(out + 4, (uint32_t)count);
    if (coefs == 0) {
        le32(out + 8, 0);
        le32(out + 12, 0);
        return W_HEADER;
    }
    int32_t low = 32767, high = -32768;
    for (size_t i = 0; i < coefs; i++) {
        if (v[i] < low) low = v[i];
        if (v[i] > high) high = v[i];
    }
    const uint32_t length = (uint32_t)(high - low + 1);
    le32(out + 8, (uint32_t)low);
    le32(out + 12, length);
    uint64_t *counts = (uint64_t *)calloc(length + 1, sizeof(uint64_t));
    uint32_t *freq = (uint32_t *)malloc((length + 1) * sizeof(uint32_t));
    uint32_t *start = (uint32_t *)malloc((length + 1) * sizeof(uint32_t));
    for (size_t i = 0; i < coefs; i++) counts[v[i] - low]++;
    orc_rans_quantize(counts, length + 1, freq);
    bitw w = {out + W_HEADER, 0, 0, 0};
    for (uint32_t s = 0; s <= length; s++) bw_gamma(&w, (uint64_t)freq[s] + 1);
    size_t len = W_HEADER + bw_finish(&w);
    uint32_t acc = 0;
    uint64_t escaped = 0;
    for (uint32_t s = 0; s <= length; s++) {
        start[s] = acc;
        acc += freq[s];
        if (s < length && !freq[s]) escaped += counts[s];
    }
    const uint32_t raw = raw_bits(length);
    uint32_t x[W_LANES];
    for (int k = 0; k < W_LANES; k++) x[k] = W_RANS_LOW;
    uint16_t *words = (uint16_t *)malloc((coefs + escaped + 1) * sizeof(uint16_t));
    size_t top = 0;
    for (size_t g = (coefs + W_LANES - 1) / W_LANES; g-- > 0;) {
        const size_t lanes = coefs - g * W_LANES < W_LANES ? coefs - g * W_LANES : W_LANES;
        for (size_t k = lanes; k-- > 0;) {
            const uint32_t s = (uint32_t)(v[g * W_LANES + k] - low);
            if (!freq[s]) {
                put_bits(&x[k], words, &top, s, raw);
                put_symbol(&x[k], words, &top, freq[length], start[length]);
            } else {
                put_symbol(&x[k], words, &top, freq[s], start[s]);
            }
        }
    }
    for (int k = 0; k < W_LANES; k++, len += 4) le32(out + len, x[k]);
    for (size_t i = 0; i < top; i++, len += 2) {
        out[len] = (uint8_t)words[top - 1 - i];
        out[len + 1] = (uint8_t)(words[top - 1 - i] >> 8);
    }
    free(words);
    free(start);
    free(freq);
    free(counts);
    return len;
}

/* The element count a coded message announces (0 for a header too short to hold one). */
uint32_t orc_rans_count(const uint8_t *bytes, size_t len) { return len >= W_HEADER ? rd32(bytes + 4) : 0; }

/* rans.rs:101-196 with the scalar Decoder of 336-388; v holds cap ring elements */
int orc_rans_decode(const uint8_t *bytes, size_t len, int16_t *v, size_t cap, uint32_t *count_out, int *ntt) {
    if (len < W_HEADER) return W_TRUNCATED;
    if (bytes[0] != 1 || bytes[1] > 1) return W_MALFORMED;
    const uint32_t count = rd32(bytes + 4), length = rd32(bytes + 12);
    const int32_t low = (int32_t)rd32(bytes + 8);
    *ntt = bytes[1];
    *count_out = count;
    if (count == 0) return len == W_HEADER && length == 0 ? W_OK : W_MALFORMED;
    if (length == 0 || length > (1u << 16) || count > W_MAX_ELEMENTS || low < -32768 ||
        (int64_t)low + length - 1 > 32767)
        return W_MALFORMED;
    if (count > cap) return W_CAPACITY;
    int err = W_OK;
    uint32_t *freq = (uint32_t *)malloc((length + 1) * sizeof(uint32_t));
    uint32_t *table = (uint32_t *)malloc(W_SCALE * 3 * sizeof(uint32_t)); /* f, slot - start, symbol */
    bitr r = {bytes + W_HEADER, len - W_HEADER, 0};
    uint64_t total = 0;
    for (uint32_t s = 0; s <= length && !err; s++) {
        uint64_t g;
        if ((err = br_gamma(&r, &g))) break;
        total += g - 1;
        if (total > W_SCALE) err = W_MALFORMED;
        freq[s] = (uint32_t)(g - 1);
    }
    if (!err && total != W_SCALE) err = W_MALFORMED;
    if (!err) {
        uint32_t acc = 0;
        for (uint32_t s = 0; s <= length; s++) {
            for (uint32_t slot = acc; slot < acc + freq[s]; slot++) {
                table[3 * slot] = freq[s];
                table[3 * slot + 1] = slot - acc;
                table[3 * slot + 2] = s;
            }
            acc += freq[s];
        }
        const size_t used = W_HEADER + br_consumed(&r);
        const uint8_t *rest = bytes + used;
        const size_t rest_len = len - used;
        uint32_t x[W_LANES];
        if (rest_len < 4 * W_LANES) err = W_TRUNCATED;
        for (int k = 0; k < W_LANES && !err; k++) {
            x[k] = rd32(rest + 4 * k);
            if (x[k] < W_RANS_LOW) err = W_MALFORMED;
        }
        const uint8_t *stream = rest + 4 * W_LANES;
        const size_t stream_len = err ? 0 : rest_len - 4 * W_LANES;
        if (!err && stream_len % 2) err = W_MALFORMED;
        const size_t nwords = stream_len / 2, coefs = (size_t)count * N;
        const uint32_t raw = raw_bits(length);
        size_t pos = 0;
        for (size_t g = 0; g * W_LANES < coefs && !err; g++) {
            const size_t lanes = coefs - g * W_LANES < W_LANES ? coefs - g * W_LANES : W_LANES;
            for (size_t k = 0; k < lanes && !err; k++) {
                const uint32_t *e = table + 3 * (x[k] & (W_SCALE - 1));
                uint32_t y = e[0] * (x[k] >> 12) + e[1], s = e[2];
                uint32_t wd = pos < nwords ? (uint32_t)stream[2 * pos] | (uint32_t)stream[2 * pos + 1] << 8 : 0;
                if (y < W_RANS_LOW) {
                    x[k] = (y << 16) | wd;
                    pos++;
                } else {
                    x[k] = y;
                }
                if (s == length) {
                    s = 0;
                    if (raw) {
                        s = x[k] & ((1u << raw) - 1);
                        y = x[k] >> raw;
                        wd = pos < nwords ? (uint32_t)stream[2 * pos] | (uint32_t)stream[2 * pos + 1] << 8 : 0;
                        if (y < W_RANS_LOW) {
                            x[k] = (y << 16) | wd;
                            pos++;
                        } else {
                            x[k] = y;
                        }
                    }
                    if (s >= length) err = W_MALFORMED;
                }
                v[g * W_LANES + k] = (int16_t)(low + (int32_t)s);
            }
            if (!err && pos > nwords) err = W_TRUNCATED;
        }
        if (!err && 2 * pos != stream_len) err = W_MALFORMED;
        for (int k = 0; k < W_LANES && !err; k++)
            if (x[k] != W_RANS_LOW) err = W_MALFORMED;
    }
    free(table);
    free(freq);
    return err;
}

/* rans.rs:500-524 */
double orc_entropy_bytes(const int16_t *v, size_t n) {
    if (!n) return 0.0;
    uint64_t *c = (uint64_t *)calloc(65536, sizeof(uint64_t));
    for (size_t i = 0; i < n; i++) c[(uint16_t)(v[i] + 32768)]++;
    double bits = 0.0;
    for (int i = 0; i < 65536; i++)
        if (c[i]) {
            const double p = (double)c[i] / (double)n;
            bits += -(double)c[i] * log2(p);
        }
    free(c);
    return bits / 8.0;
}
