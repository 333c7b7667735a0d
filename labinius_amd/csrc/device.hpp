// Shapes and descriptors shared by the host orchestration (prover.cpp) and the HIP kernels
// (kernels.hip).
//
// The commitment y = sum_i A_i o NTT(w_i) and the fold's NTT(v) = sum_j c_j o NTT(W_j) are both
// computed without transforming any single ring element. The splitting tree's first three
// radix-2 levels map a binary w to 8 sub-rings of degree 81 whose coefficient s is a function of
// the 8 bits w[81 t + s] (t < 8), so one 256-entry table per sub-ring gives it ("the byte s");
// the quadratic tree's first two levels give 4 sub-rings of degree 162 whose coefficient
// 81 a + s is a function of the even (a = 0) or odd (a = 1) bits of the same byte s. What is left
// is, per 81-slot block, an 81 x 81 linear map E (slot u <- sub-ring coefficient n), so
//
//     y[u] = sum_n E[u][n] sum_i A[i][u] x_i[n] = sum_(n, i) A'[(n, i)][u] x_i[n],
//     A'[(n, i)][u] = E[u][n] A[i][u] mod q,
//
// an integer GEMM with K = (n, i), 81 x the ring elements (commit) or columns (fold), run on int8
// MFMA with both operands split into two int8 digits (UNSIGNED_DIGITS_Q below). A "tile" is one
// such 96 x batch (81 or 82 rows used) GEMM with its own A' and LUT: 8 per splitting limb, 16 per
// quadratic limb (4 sub-rings x 2 halves of the A rows x 2 parities of the inputs). The fold is
// the same GEMM with A = the challenges' transform, 8 tiles on the base.
//
// A tile's A' is shared by every batch element (column for the commit, ring element for the
// fold), so the GEMM's N runs over batch elements: a "group" is 256 of them = 8 fragments of 32.
// The long K of an output is split into ranges of at most KB_EXACT steps whose results are added
// into y (u32 atomics of residues, reduced mod q by y's readers).
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

namespace lbn {

constexpr int TILE = 96;         // padded M of a tile (81 rows used)
constexpr int TILE_USED = 81;
constexpr int KSTEP = 64;        // K per step of the GEMM loop: one v_mfma_i32_16x16x64_i8
// The operands are staged in 32 x 32 fragment order (lane l: row or input l & 31, K 16 (l >> 5)
// of a 32-K half); the GEMM reads them as 16 x 64 pieces.
constexpr int BLK = 32;          // staged fragment block (32 rows or inputs x 32 K)
constexpr int NBLK = TILE / BLK; // 3 blocks per tile side
constexpr int FRAG = 1024;       // bytes of one staged 32 x 32 int8 fragment (64 lanes x 16 B)
constexpr int A_STEP = 2 * 2 * NBLK * FRAG;  // A fragments per K step: [kh][digit][mb]
constexpr int GROUP_COLS = 256;              // batch elements of a group (the GEMM's N per item)
constexpr int GROUP_FRAGS = GROUP_COLS / BLK;
constexpr int TILE_ROWS = 82;               // rows a tile may use (a quadratic commit tile's first half: 82)
constexpr int EP = 88;                      // E rows padded to 176 bytes (zero from n = 81)
constexpr int EE_TILE = TILE_ROWS * EP;      // E of a tile, [m 82][n 88] u16

// The K order: index k = n kpad + i for sub-ring coefficient n < 81 and i < kpad, the ring
// element (commit) or column (fold) padded to a power of two of at least one step, so that a
// step holds one n and runs of steps of one n are a power of two.
inline uint32_t gemm_kpad(uint64_t len) {
    uint32_t p = KSTEP;
    while (p < len) p <<= 1;
    return p;
}
__host__ __device__ inline uint32_t gemm_kbs(uint32_t kpad) { return TILE_USED * kpad / KSTEP; }

// The input bytes in MFMA B-fragment order. Byte n of batch element b at index i lives at
//
//   ((((g * 8 + jg) * kbs + kb) * 2 + kh) * 64 + lane) * 16 + (k & 15)
//
// with k = n kpad + i, g = b / 256, j = b % 256, jg = j / 32, lane = (j & 31) + 32 ((k >> 4) & 1),
// kh = (k >> 5) & 1, kb = k / 64, kbs = gemm_kbs(kpad): one 16-byte load per lane gives a lane
// its B fragment.
__host__ __device__ inline uint64_t byte_offset(uint32_t b, uint32_t n, uint32_t i, uint32_t kpad) {
    const uint32_t g = b / GROUP_COLS, j = b % GROUP_COLS;
    const uint64_t k = (uint64_t)n * kpad + i;
    const uint32_t lane = (j & 31) + 32 * (uint32_t)((k >> 4) & 1);
    return ((((uint64_t)g * GROUP_FRAGS + (j >> 5)) * gemm_kbs(kpad) + (k >> 6)) * 2 + ((k >> 5) & 1)) * 1024 +
           lane * 16 + (k & 15);
}
inline uint64_t bytes_size(uint32_t batches, uint32_t kpad) {
    return (uint64_t)((batches + GROUP_COLS - 1) / GROUP_COLS) * GROUP_FRAGS * gemm_kbs(kpad) * 2 * 1024;
}

// One tile of a GEMM launch.
struct TileDesc {
    uint64_t a_off;     // byte offset of the tile's A' fragments: [kb][kh 2][digit 2][mb 3][lane 64][16]
    uint32_t lut_off;   // element offset (u16) of the tile's two 256-entry LUTs
    uint32_t ee_off;    // element offset (u16) of the tile's E, [m 82][n 88] (A' is built from it)
    uint32_t out_off;   // element offset (i16) of the tile's 96 output slots (-1 = padding row)
    uint32_t q;
    uint32_t par;       // sub-ring coefficient n reads LUT (par + n) & 1
    uint32_t y_off;     // word offset of the tile's limb inside one batch's output rows
};

struct GemmArgs {
    const TileDesc *tiles;
    const int8_t *afrag;      // all tiles' A' fragments
    const uint16_t *lut;      // all tiles' LUTs (low byte x_l, high byte x_h, both int8)
    const int16_t *out;       // all tiles' output slots
    const uint8_t *bytes;     // the input bytes, fragment order (byte_offset)
    uint32_t kbs;             // K steps of the layouts, gemm_kbs(kpad)
    uint32_t run_shift;       // log2(kpad / 64): step kb holds sub-ring coefficient kb >> run_shift
    uint32_t a_kb_stride;     // bytes of A' fragments per K step of a tile (A_STEP)
    uint32_t *y;              // output rows: every (batch element, tile output) has its own word
    uint64_t y_batch;         // words per batch element in y
    uint32_t batches;
    uint32_t groups;          // ceil(batches / 256)
    uint32_t ntiles;
    uint32_t sched_t;         // tiles per XCD round (4 .. 32, a power of two); 32 / sched_t column blocks
    // set by the launcher: K in `splits` ranges of kps steps (the last may be shorter); column
    // blocks (group, range) = groups x splits; the tile sets and column-block sets of the XCD
    // rounds; the work items
    uint32_t splits, kps, cblocks, tsets, csets, items;
};

// Operands mod q are split into two int8 digits, x = x_l + 128 x_h: for q < UNSIGNED_DIGITS_Q
// the residue in [0, q) gives 0 <= x_l, x_h < 128; for larger q (up to 19441) the centred residue
// gives -64 <= x_l < 64 and |x_h| <= 76.
constexpr uint32_t UNSIGNED_DIGITS_Q = 16384;
// The int32 accumulators are exact while K x 32258 < 2^31: the largest class, mid = A_l x_h +
// A_h x_l, adds at most 2 x 127 x 127 = 32258 per K index (non-negative digits; the centred ones
// add at most 2 x 64 x 76). The launcher splits longer K into ranges of at most KB_EXACT steps.
constexpr uint32_t KB_EXACT = 2147483647u / (KSTEP * 32258u);  // 1040 steps, K = 66 560

// The commit GEMM's output is y[c][limb][plane][648]. A splitting limb's 8 tiles write each of
// the 648 slots, into plane 0. A quadratic limb's slot 2j + t is written by the two tiles of
// input parity b = 0 and 1, each into plane b; the finish adds the planes. With one K range a
// word has one writer, which stores its residue; with several, y is cleared first and each range
// adds its residue in [0, q), so a word is below splits x q (far below 2^32) and its reader
// reduces it mod q.
constexpr int YACC_PLANES = 2;

// Per-limb constants of the commit finish (decomposition) kernel.
struct FinishLimb {
    uint32_t q;
    uint32_t quad;
    uint32_t ii;            // psi^486 (splitting limbs)
    uint32_t inv2;          // quadratic limbs
    uint16_t twist[4][162]; // splitting: 4^-1 psi^{-v_s k}; quadratic: [0] = 2^-1 psi'^{-v_s}
};

// Constants of one limb's inverse transform (the fold's last step on the base limb; every limb's
// for the bit-dropped digits).
struct InttConsts {
    uint32_t q;
    uint32_t quad;
    uint32_t w, w2, z6, inv2, inv3, det;
    uint32_t twi_off[7];    // offset of level l's inverse twiddles in the twi array
    uint32_t twi_count;     // the array's length (at most INTT_TWI_MAX)
    float qinv;             // 1 / q, for mulmod_f
};
// The inverse twiddles of one limb: one per sub-block of each level, 327 (splitting) or 487
// (quadratic) in all.
constexpr uint32_t INTT_TWI_MAX = 512;

// Mixed-radix (Garner) digits of the bit-dropped commitment (bd/digits.rs:134-182): x = m_0 +
// q_0 (m_1 + q_1 (m_2 + ...)) from the residues r_k, with inv[k][j] = q_j^-1 mod q_k.
constexpr int GARNER_MAX = 8;
struct GarnerConsts {
    uint32_t limbs;
    uint32_t d, round;  // dropped bits, 2^(d - 1)
    uint32_t q[GARNER_MAX];
    uint32_t inv[GARNER_MAX][GARNER_MAX];
};

// B128 = GF(2)[x] / (x^128 + x^7 + x^2 + x + 1) (fields/scalar.rs:4, 41-63), low word first: the
// 256-bit carry-less product by a comb over a's bits, then the high half H folded as
// H (1 + x + x^2 + x^7) twice (its top 7 bits spill once more).
__host__ __device__ inline void b128_mul(uint64_t a0, uint64_t a1, uint64_t b0, uint64_t b1, uint64_t &o0,
                                         uint64_t &o1) {
    uint64_t p0 = 0, p1 = 0, p2 = 0, p3 = 0;
    for (int j = 0; j < 64; j++) {
        const uint64_t s0 = b0 << j, s1 = (b1 << j) | (j ? b0 >> (64 - j) : 0), s2 = j ? b1 >> (64 - j) : 0;
        const uint64_t m0 = 0ull - ((a0 >> j) & 1), m1 = 0ull - ((a1 >> j) & 1);
        p0 ^= s0 & m0;
        p1 ^= (s1 & m0) ^ (s0 & m1);
        p2 ^= (s2 & m0) ^ (s1 & m1);
        p3 ^= s2 & m1;
    }
    const uint64_t w0 = p2 ^ (p2 << 1) ^ (p2 << 2) ^ (p2 << 7);
    const uint64_t w1 = p3 ^ ((p3 << 1) | (p2 >> 63)) ^ ((p3 << 2) | (p2 >> 62)) ^ ((p3 << 7) | (p2 >> 57));
    const uint64_t th = (p3 >> 63) ^ (p3 >> 62) ^ (p3 >> 57);
    o0 = p0 ^ w0 ^ th ^ (th << 1) ^ (th << 2) ^ (th << 7);
    o1 = p1 ^ w1;
}

// The F162 value passed by value to the sumcheck's fold.
struct F162v {
    uint64_t w[3];
};

// The verifier's fold of a commitment (fold.rs:799-873): per limb, the challenge c_j evaluated at
// slot s as sum_i (-1)^sign_i (-1)^p_i pw[(mult v_s p_i) mod cond] with v_s = POW3_SLOT_EXP[s] and
// pw the limb's root powers (at pw_off in one array).
struct FoldCommConsts {
    uint32_t limbs;
    uint32_t q[GARNER_MAX], cond[GARNER_MAX], mult[GARNER_MAX], pw_off[GARNER_MAX];
};

// A bit-packed wire object (wire/mod.rs:226-243, 386-397): a sequence of records, each the same
// run of segments, each `count` fields of `bits` bits, least significant bit first with nothing
// between them. Field i of segment l of record r is the u16 at src[off + r rec_stride + i stride]
// (a centered int16 taken mod q when q != 0, else the value's low `bits` bits).
constexpr int PACK_SEGS = 8;
struct PackSeg {
    uint32_t count, bits, q, bit0;  // bit0: the segment's first bit inside a record
    uint64_t off, rec_stride;
    uint32_t stride, pad;
};
struct PackDesc {
    uint32_t nseg, rec_bits;
    uint64_t total_bits;
    PackSeg seg[PACK_SEGS];
};

// A symbol of the rANS encoder (the step of rans.rs:306-317). With f its frequency and s its
// start, the state x < 2^32 first sheds its low 16 bits when x > fm = f 2^20 - 1 (never for
// f = 4096), then becomes x + q (4096 - f) + s with q = x / f: floor(x * rcp) in f64, where rcp is
// 1/f rounded up (exact for a power of two), exact for x < 2^32 since x rcp overshoots x / f by
// less than 2^-30 and the product rounds monotonically. cs = (4096 - f) | s << 16; an
// unoccupied symbol (f = 0, folded into the escape) has cs = 4096.
struct alignas(16) RansSym {
    double rcp;
    uint32_t fm, cs;
};
constexpr uint32_t RANS_LANES = 64;       // rans.rs:12
constexpr uint32_t RANS_LOW = 1u << 16;   // wire/mod.rs:60
constexpr uint32_t RANS_LDS_SYMS = 8192;  // symbol tables up to this size are read from LDS
constexpr uint32_t RANS_ESCAPED = 4096;   // cs of an unoccupied symbol
// the encoder's status words: [0] min, [1] max (int), [2..3] words pushed (u64), [4..68) states
constexpr int RANS_STAT = 4 + RANS_LANES;

}  // namespace lbn
