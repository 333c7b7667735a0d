// Ring constants and exact scalar arithmetic on the host (see ring.hpp).
#include "ring.hpp"

#include <cstring>
#include <map>
#include <memory>
#include <mutex>

namespace lbn {

uint64_t pow_mod(uint64_t b, uint64_t e, uint64_t q) {
    uint64_t r = 1;
    b %= q;
    for (; e; e >>= 1) {
        if (e & 1) r = r * b % q;
        b = b * b % q;
    }
    return r;
}

uint64_t inv_mod(uint64_t a, uint64_t q) { return pow_mod(a, q - 2, q); }

bool is_modulus(uint32_t q) {
    for (uint32_t m : MODULI)
        if (m == q) return true;
    return false;
}

bool is_quad(uint32_t q) { return (q - 1) % CONDUCTOR_QUAD == 0 && (q - 1) % CONDUCTOR != 0; }

// The splitting tree (params.rs:1-16, 28-33): Phi_6 split, two radix-2 levels, four radix-3
// levels. A sub-ring at level l is Z_q[X]/(X^{DEGREE[l]} - psi^e); child s of a radix-p split has
// exponent (e + 1944 s) / p and sits at block offset s * n / p.
static const int RADIX[7] = {2, 2, 2, 3, 3, 3, 3};
static const int SUBRINGS[8] = {1, 2, 4, 8, 24, 72, 216, 648};
static const int DEGREE[8] = {648, 324, 162, 81, 27, 9, 3, 1};
// The quadratic-slot tree (params.rs:230-287): the second radix-2 level removed, conductor 972.
static const int RADIX_Q[6] = {2, 2, 3, 3, 3, 3};
static const int SUBRINGS_Q[7] = {1, 2, 4, 12, 36, 108, 324};
static const int DEGREE_Q[7] = {648, 324, 162, 54, 18, 6, 2};

static uint32_t sub_exp(int level, int k, bool quad) {
    if (level == 1) return quad ? (k == 0 ? 162u : 810u) : (k == 0 ? 324u : 1620u);
    uint32_t p = (uint32_t)(quad ? RADIX_Q[level - 1] : RADIX[level - 1]);
    uint32_t cond = quad ? CONDUCTOR_QUAD : CONDUCTOR;
    return (sub_exp(level - 1, k / (int)p, quad) + cond * ((uint32_t)k % p)) / p;
}

static Tables build_tables() {
    Tables t;
    std::memset(&t, 0, sizeof t);
    for (int j = 0; j < N; j++) t.slot_exp[j] = (uint16_t)sub_exp(7, j, false);
    for (int j = 0; j < 324; j++) t.quad_slot_exp[j] = (uint16_t)sub_exp(6, j, true);
    int pos[486];
    for (int &p : pos) p = -1;
    int n = 0;
    for (int j = 0; j < N; j++) {
        int u = t.slot_exp[j], v = u % 486, tt = (u - v) / 486;
        if (pos[v] < 0) {
            pos[v] = n;
            t.pow3_slot_exp[n++] = (uint16_t)v;
        }
        t.slot_648[tt][pos[v]] = (uint16_t)j;
    }
    for (int s = 0; s < N162; s++) {
        int v = t.pow3_slot_exp[s];
        for (int j = 0; j < 324; j++) {
            if (t.quad_slot_exp[j] == v) t.quad_class_slot[0][s] = (uint16_t)j;
            if (t.quad_slot_exp[j] == v + 486) t.quad_class_slot[1][s] = (uint16_t)j;
        }
    }
    for (int k = 0; k < 8; k++) t.subring_exp3[k] = sub_exp(3, k, false);
    for (int k = 0; k < 4; k++) t.subring_exp_q2[k] = sub_exp(2, k, true);
    return t;
}

const Tables &tables() {
    static const Tables t = build_tables();
    return t;
}

static uint32_t smallest_root(uint32_t q, uint32_t order, uint32_t d1, uint32_t d2) {
    for (uint64_t x = 2;; x++)
        if (pow_mod(x, order, q) == 1 && pow_mod(x, d1, q) != 1 && pow_mod(x, d2, q) != 1)
            return (uint32_t)x;
}

static std::unique_ptr<Prime> build_prime(uint32_t q) {
    auto p = std::make_unique<Prime>();
    const Tables &t = tables();
    p->q = q;
    p->quad = is_quad(q);
    p->bm = (1ull << 40) / q;
    uint32_t cond;
    if (p->quad) {
        cond = CONDUCTOR_QUAD;
        p->root = smallest_root(q, 972, 486, 324);  // params.rs:291-299
        p->w = pow_mod(p->root, 324, q);
        p->z6 = pow_mod(p->root, 162, q);
    } else {
        cond = CONDUCTOR;
        p->root = smallest_root(q, 1944, 972, 648);  // params.rs:53-61
        p->w = pow_mod(p->root, 648, q);
        p->z6 = pow_mod(p->root, 324, q);
    }
    int top = p->quad ? 5 : 6;
    for (int l = 1; l <= top; l++) {
        int nk = p->quad ? SUBRINGS_Q[l] : SUBRINGS[l];
        int rad = p->quad ? RADIX_Q[l] : RADIX[l];
        p->tw[l].resize(nk);
        p->twi[l].resize(nk);
        for (int k = 0; k < nk; k++) {
            uint32_t e = sub_exp(l, k, p->quad) / (uint32_t)rad;
            p->tw[l][k] = (uint32_t)pow_mod(p->root, e, q);
            p->twi[l][k] = (uint32_t)inv_mod(p->tw[l][k], q);
        }
    }
    if (p->quad) {
        p->leafc.resize(324);
        for (int j = 0; j < 324; j++) p->leafc[j] = (uint32_t)pow_mod(p->root, t.quad_slot_exp[j], q);
    }
    p->pw.resize(cond);
    for (uint32_t e = 0; e < cond; e++) p->pw[e] = (uint32_t)pow_mod(p->root, e, q);
    p->w2 = p->w * p->w % q;
    p->inv2 = inv_mod(2, q);
    p->inv3 = inv_mod(3, q);
    p->det = inv_mod((2 * p->z6 + q - 1) % q, q);
    return p;
}

const Prime &prime(uint32_t q) {
    static std::mutex mu;
    static std::map<uint32_t, std::unique_ptr<Prime>> cache;
    std::lock_guard<std::mutex> g(mu);
    auto it = cache.find(q);
    if (it == cache.end()) it = cache.emplace(q, build_prime(q)).first;
    return *it->second;
}

void ntt(const Prime &p, uint64_t *v) {
    const uint64_t q = p.q;
    // level 0: children X^324 - zeta6 and X^324 - zeta6^-1, zeta6^-1 = 1 - zeta6
    for (int i = 0; i < 324; i++) {
        uint64_t a0 = v[i], a1 = v[i + 324], t = p.md(a1 * p.z6);
        v[i] = p.md(a0 + t);
        v[i + 324] = p.md(a0 + a1 + q - t);
    }
    const int top = p.quad ? 5 : 6;
    for (int level = 1; level <= top; level++) {
        const int n = p.quad ? DEGREE_Q[level] : DEGREE[level];
        const int rad = p.quad ? RADIX_Q[level] : RADIX[level];
        const int m = n / rad, nk = p.quad ? SUBRINGS_Q[level] : SUBRINGS[level];
        for (int k = 0; k < nk; k++) {
            uint64_t *b = v + (size_t)k * n;
            const uint64_t z = p.tw[level][k];
            if (rad == 2) {
                for (int i = 0; i < m; i++) {
                    uint64_t t = p.md(b[m + i] * z);
                    uint64_t a0 = b[i];
                    b[i] = p.md(a0 + t);
                    b[m + i] = p.md(a0 + q - t);
                }
            } else {
                const uint64_t z2 = p.md(z * z);
                for (int i = 0; i < m; i++) {
                    uint64_t a0 = b[i], t1 = p.md(b[m + i] * z), t2 = p.md(b[2 * m + i] * z2);
                    b[i] = p.md(a0 + t1 + t2);
                    b[m + i] = p.md(a0 + p.w * t1 + p.w2 * t2);
                    b[2 * m + i] = p.md(a0 + p.w2 * t1 + p.w * t2);
                }
            }
        }
    }
}

void intt(const Prime &p, uint64_t *u) {
    const uint64_t q = p.q;
    const int top = p.quad ? 5 : 6;
    for (int level = top; level >= 1; level--) {
        const int n = p.quad ? DEGREE_Q[level] : DEGREE[level];
        const int rad = p.quad ? RADIX_Q[level] : RADIX[level];
        const int m = n / rad, nk = p.quad ? SUBRINGS_Q[level] : SUBRINGS[level];
        for (int k = 0; k < nk; k++) {
            uint64_t *b = u + (size_t)k * n;
            const uint64_t zi = p.twi[level][k];
            if (rad == 2) {
                for (int i = 0; i < m; i++) {
                    uint64_t y0 = b[i], y1 = b[m + i];
                    b[i] = p.md(p.md(y0 + y1) * p.inv2);
                    b[m + i] = p.md(p.md(p.md(y0 + q - y1) * p.inv2) * zi);
                }
            } else {
                const uint64_t zi2 = p.md(zi * zi);
                for (int i = 0; i < m; i++) {
                    uint64_t y0 = b[i], y1 = b[m + i], y2 = b[2 * m + i];
                    uint64_t a0 = p.md(p.md(y0 + y1 + y2) * p.inv3);
                    uint64_t t1 = p.md(p.md(y0 + p.w2 * y1 + p.w * y2) * p.inv3);
                    uint64_t t2 = p.md(p.md(y0 + p.w * y1 + p.w2 * y2) * p.inv3);
                    b[i] = a0;
                    b[m + i] = p.md(t1 * zi);
                    b[2 * m + i] = p.md(t2 * zi2);
                }
            }
        }
    }
    for (int i = 0; i < 324; i++) {
        uint64_t y0 = u[i], y1 = u[i + 324];
        uint64_t a1 = p.md(p.md(y0 + q - y1) * p.det);
        u[i] = p.md(y0 + q - p.md(p.z6 * a1));
        u[i + 324] = a1;
    }
}

void mul_quad_slots(const Prime &p, const uint64_t *a, const uint64_t *b, uint64_t *c) {
    for (int j = 0; j < 324; j++) {
        uint64_t a0 = a[2 * j], a1 = a[2 * j + 1], b0 = b[2 * j], b1 = b[2 * j + 1];
        c[2 * j] = p.md(a0 * b0 + p.leafc[j] * p.md(a1 * b1));
        c[2 * j + 1] = p.md(a0 * b1 + a1 * b0);
    }
}

void decompose(const Prime &p, const uint32_t *y, int16_t *out) {
    const Tables &t = tables();
    const uint64_t q = p.q;
    if (!p.quad) {
        // Y_k(v) = 4^-1 psi^{-vk} sum_t i^{-tk} E_t with i = psi^486: a radix-4 inverse DFT with
        // the single product i (E_1 - E_3) (ring/mod.rs:274-312).
        const uint64_t inv4 = inv_mod(4, q), ii = p.pw[486];
        for (int s = 0; s < N162; s++) {
            uint64_t e0 = y[t.slot_648[0][s]], e1 = y[t.slot_648[1][s]];
            uint64_t e2 = y[t.slot_648[2][s]], e3 = y[t.slot_648[3][s]];
            uint64_t a = e0 + e2, d0 = e0 + q - e2, c = e1 + e3, d1 = e1 + q - e3;
            uint64_t id = p.md(d1 * ii);
            uint64_t m[4] = {a + c, d0 + q - id, a + 2 * q - c, d0 + id};
            const uint64_t v = t.pow3_slot_exp[s];
            for (int k = 0; k < 4; k++) {
                uint64_t tw = p.md(inv4 * p.pw[(CONDUCTOR - (v * k) % CONDUCTOR) % CONDUCTOR]);
                out[k * N162 + s] = center(p.md(p.md(m[k]) * tw), q);
            }
        }
    } else {
        // the 2-point butterfly over the two leaves of a class (ring/mod.rs:383-417)
        for (int s = 0; s < N162; s++) {
            const uint64_t v = t.pow3_slot_exp[s];
            const int jp = t.quad_class_slot[0][s], jm = t.quad_class_slot[1][s];
            const uint64_t tw = p.md(p.inv2 * p.pw[(CONDUCTOR_QUAD - v) % CONDUCTOR_QUAD]);
            for (int k = 0; k < 2; k++) {
                uint64_t ep = y[2 * jp + k], em = y[2 * jm + k];
                out[k * N162 + s] = center(p.md((ep + em) * p.inv2), q);
                out[(k + 2) * N162 + s] = center(p.md((ep + q - em) * tw), q);
            }
        }
    }
}

// Carry-less 64 x 64 -> 128 with the CPU's carry-less multiply instruction (PCLMULQDQ).
typedef long long clmul_v2 __attribute__((vector_size(16)));
__attribute__((target("pclmul"))) static void clmul64_hw(uint64_t a, uint64_t b, uint64_t &lo, uint64_t &hi) {
    const clmul_v2 x = {(long long)a, 0}, y = {(long long)b, 0};
    const clmul_v2 r = __builtin_ia32_pclmulqdq128(x, y, 0);
    lo = (uint64_t)r[0];
    hi = (uint64_t)r[1];
}

// The same by a 4-bit window, for a CPU without the instruction.
static void clmul64_sw(uint64_t a, uint64_t b, uint64_t &lo, uint64_t &hi) {
    uint64_t tlo[16], thi[16];
    tlo[0] = thi[0] = 0;
    for (int i = 1; i < 16; i++) {
        // tab[i] = i(x) * b(x), up to 67 bits
        uint64_t l = 0, hh = 0;
        for (int k = 0; k < 4; k++)
            if ((i >> k) & 1) {
                l ^= b << k;
                if (k) hh ^= b >> (64 - k);
            }
        tlo[i] = l;
        thi[i] = hh;
    }
    lo = hi = 0;
    for (int s = 60; s >= 0; s -= 4) {
        // (hi:lo) <<= 4, then xor the window's product
        if (s != 60) {
            hi = (hi << 4) | (lo >> 60);
            lo <<= 4;
        }
        unsigned idx = (unsigned)(a >> s) & 15;
        lo ^= tlo[idx];
        hi ^= thi[idx];
    }
}

static const bool has_clmul = __builtin_cpu_supports("pclmul");

static inline void clmul64(uint64_t a, uint64_t b, uint64_t &lo, uint64_t &hi) {
    if (has_clmul)
        clmul64_hw(a, b, lo, hi);
    else
        clmul64_sw(a, b, lo, hi);
}

F162 f162_mul(const F162 &a, const F162 &b) {
    uint64_t p[6] = {0, 0, 0, 0, 0, 0};
    for (int i = 0; i < 3; i++)
        for (int j = 0; j < 3; j++) {
            uint64_t lo, hi;
            clmul64(a.w[i], b.w[j], lo, hi);
            p[i + j] ^= lo;
            p[i + j + 1] ^= hi;
        }
    // p = L + x^162 H: x^162 = x^81 + 1, and x^81 H = (H mod x^81) x^81 + (H >> 81)(x^81 + 1)
    uint64_t H[3];
    H[0] = (p[2] >> 34) | (p[3] << 30);
    H[1] = (p[3] >> 34) | (p[4] << 30);
    H[2] = (p[4] >> 34) | (p[5] << 30);
    // H has at most 161 bits
    uint64_t L[3] = {p[0], p[1], p[2] & ((1ull << 34) - 1)};
    // H_lo = H mod x^81; H2 = H >> 81
    uint64_t Hlo[2] = {H[0], H[1] & ((1ull << 17) - 1)};
    uint64_t H2[2] = {(H[1] >> 17) | (H[2] << 47), H[2] >> 17};
    // r = L + H + (H_lo + H2) x^81 + H2
    uint64_t S[2] = {Hlo[0] ^ H2[0], Hlo[1] ^ H2[1]};  // < 81 bits
    uint64_t r[3];
    r[0] = L[0] ^ H[0] ^ H2[0];
    r[1] = L[1] ^ H[1] ^ H2[1];
    r[2] = L[2] ^ H[2];
    // S << 81: bits 81.. -> word 1 bit 17
    r[1] ^= S[0] << 17;
    r[2] ^= (S[0] >> 47) | (S[1] << 17);
    F162 out = {{r[0], r[1], r[2]}};
    return out;
}

std::vector<F162> eq_table(const F162 *ps, int n) {
    std::vector<F162> t(size_t(1) << n);
    t[0] = F162{{1, 0, 0}};
    size_t len = 1;
    for (int k = 0; k < n; k++) {
        F162 p1 = ps[k];
        p1.w[0] ^= 1;
        for (size_t b = 0; b < len; b++) {
            F162 x = t[b];
            t[b + len] = f162_mul(x, ps[k]);
            t[b] = f162_mul(x, p1);
        }
        len *= 2;
    }
    return t;
}

}  // namespace lbn
