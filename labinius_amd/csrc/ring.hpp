// Ring constants and exact scalar arithmetic on the host: the two NTT trees of R_648 =
// Z_q[X]/(X^648 - X^324 + 1), the 648 -> 4 x 162 decomposition, and F162 = GF(2)[x]/(x^162 +
// x^81 + 1). The HIP kernels never walk these trees; the host uses them to build the tables the
// kernels read and, on the verifier side, to recompute A v.
#pragma once

#include <array>
#include <cstdint>
#include <vector>

namespace lbn {

constexpr int N = 648;
constexpr int N162 = 162;
constexpr uint32_t CONDUCTOR = 1944;
constexpr uint32_t CONDUCTOR_QUAD = 972;

// The seven moduli (ring/mod.rs:87-124).
constexpr uint32_t MODULI[7] = {2917, 3889, 4861, 9721, 12637, 17497, 19441};

uint64_t pow_mod(uint64_t b, uint64_t e, uint64_t q);
uint64_t inv_mod(uint64_t a, uint64_t q);
bool is_modulus(uint32_t q);
// q = 1 mod 972 but not mod 1944: R_648 ends in 324 quadratic leaves (params.rs:266-268)
bool is_quad(uint32_t q);

struct Tables {
    uint16_t slot_exp[N];            // params.rs:78-89
    uint16_t quad_slot_exp[324];     // params.rs:320-332
    uint16_t pow3_slot_exp[N162];    // ring/mod.rs:173-176
    uint16_t slot_648[4][N162];      // ring/mod.rs:178-179
    uint16_t quad_class_slot[2][N162];  // params.rs:394-401
    uint32_t subring_exp3[8];        // exponent of the 8 level-3 subrings of the splitting tree
    uint32_t subring_exp_q2[4];      // exponent of the 4 level-2 subrings of the quadratic tree
};
const Tables &tables();

// Everything a prime needs: the root (psi, a primitive 1944-th root, or psi', a primitive 972-nd
// one for a quadratic prime), the per-level twiddles of its tree and their inverses.
struct Prime {
    uint32_t q;
    bool quad;
    uint64_t root, w, w2, z6, inv2, inv3, det;
    uint64_t bm;  // floor(2^40 / q)
    std::vector<uint32_t> tw[7], twi[7];
    std::vector<uint32_t> leafc;  // quadratic leaves: psi'^QUAD_SLOT_EXP[j]
    std::vector<uint32_t> pw;     // root^e for e < conductor
    uint64_t md(uint64_t x) const {
        uint64_t r = x - (uint64_t)(((unsigned __int128)x * bm) >> 40) * q;
        return r >= q ? r - q : r;
    }
};
const Prime &prime(uint32_t q);

// Forward transform of a fully reduced vector on q's tree (scalar.rs:35-84 / 185-234).
void ntt(const Prime &p, uint64_t *v);
// Exact inverse (scalar.rs:123-177, recursion/limbs.rs:76-127).
void intt(const Prime &p, uint64_t *v);
// Leaf-wise product of two quadratic-slot transforms (scalar.rs:254-265).
void mul_quad_slots(const Prime &p, const uint64_t *a, const uint64_t *b, uint64_t *c);
// The four R_162 components of a fully reduced y, centered: out[k * 162 + s] (ring/mod.rs:274-417).
void decompose(const Prime &p, const uint32_t *y, int16_t *out);

inline int16_t center(uint64_t x, uint64_t q) {
    x %= q;
    return x > (q - 1) / 2 ? (int16_t)((int64_t)x - (int64_t)q) : (int16_t)x;
}
inline uint64_t reduce_signed(int64_t x, uint64_t q) {
    int64_t r = x % (int64_t)q;
    return (uint64_t)(r < 0 ? r + (int64_t)q : r);
}

// F162 = GF(2)[x]/(x^162 + x^81 + 1), three little-endian u64 (fields/scalar.rs:92-116).
struct F162 {
    uint64_t w[3];
};
F162 f162_mul(const F162 &a, const F162 &b);
inline F162 f162_add(const F162 &a, const F162 &b) { return {{a.w[0] ^ b.w[0], a.w[1] ^ b.w[1], a.w[2] ^ b.w[2]}}; }
inline bool f162_eq(const F162 &a, const F162 &b) { return a.w[0] == b.w[0] && a.w[1] == b.w[1] && a.w[2] == b.w[2]; }
// eval.rs:51-63
std::vector<F162> eq_table(const F162 *ps, int n);

}  // namespace lbn
