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
//     y[u] = sum_n E[u][n] * Z[u][n],        Z[u][n] = sum_i A[i][u] * x_i[n],
//
// where Z is an integer GEMM (K = ring elements) run on int8 MFMA with both operands split into
// two signed 7-bit digits. A "tile" is one such 96 x 96 (81 x 81 padded) GEMM: 8 per splitting
// limb, 16 per quadratic limb (4 sub-rings x 2 halves of the A rows x 2 parities of the inputs).
// The fold is the same GEMM with K = columns, A = the challenges' transform, 8 tiles on the base.
#pragma once

#include <cstdint>

namespace lbn {

constexpr int TILE = 96;         // padded M and N of a tile (81 rows / inputs used)
constexpr int TILE_USED = 81;
constexpr int KSTEP = 64;        // K per step of the GEMM loop: two v_mfma_i32_32x32x32_i8
constexpr int BLK = 32;          // MFMA block (32 x 32)
constexpr int NBLK = TILE / BLK; // 3 blocks per tile side
constexpr int FRAG = 1024;       // bytes of one 32 x 32 int8 operand fragment (64 lanes x 16 B)
constexpr int A_STEP = 2 * 2 * NBLK * FRAG;  // A fragments per K step: [kh][digit][mb]
constexpr int B_TILE = 2 * 2 * NBLK * FRAG;  // x digits per K step: [kh][digit][nb]
constexpr int EPI_TILE = NBLK * NBLK * 64 * 16 * 3;  // epilogue weights per tile

// One tile of a GEMM launch.
struct TileDesc {
    uint64_t a_off;     // byte offset of the tile's A fragments: [kb][kh 2][digit 2][mb 3][lane 64][16]
    uint32_t lut_off;   // element offset (u16) of the tile's two 256-entry LUTs
    uint32_t epi_off;   // element offset (u16) of the tile's epilogue weights [mb][nb][lane][16][3]
    uint32_t out_off;   // element offset (i16) of the tile's 96 output slots (-1 = padding row)
    uint32_t q;
    uint32_t par;       // input column n reads LUT (par + n) & 1
    uint32_t y_off;     // word offset of the tile's limb inside one batch's output rows
};

struct GemmArgs {
    const TileDesc *tiles;
    const int8_t *afrag;      // all tiles' A fragments
    const uint16_t *lut;      // all tiles' LUTs (low byte x_l, high byte x_h, both int8)
    const uint16_t *epi;      // all tiles' epilogue weights (E, 128 E, 16384 E mod q)
    const int16_t *out;       // all tiles' output slots
    const uint8_t *bytes;     // bytes[batch][n < 81][kpad]
    uint64_t bytes_batch;     // batch stride of bytes
    uint32_t kpad;            // K padded to a multiple of KSTEP (the n stride of bytes)
    uint32_t kb_per_split;    // K steps per split (blockIdx.z)
    uint32_t kb_total;
    uint32_t a_kb_stride;     // bytes of A fragments per K step of a tile (A_STEP)
    uint32_t *y;              // output rows
    uint64_t y_batch;         // words per batch in y
    int accumulate;           // 1: atomicAdd (commit, several contributions), 0: store (fold)
    uint32_t ntiles;          // blockIdx.x = batch * ntiles + tile
};

// Per-limb constants of the commit finish (decomposition) kernel.
struct FinishLimb {
    uint32_t q;
    uint32_t quad;
    uint32_t ii;            // psi^486 (splitting limbs)
    uint32_t inv2;          // quadratic limbs
    uint16_t twist[4][162]; // splitting: 4^-1 psi^{-v_s k}; quadratic: [0] = 2^-1 psi'^{-v_s}
};

// Constants of the base limb's inverse transform (the fold's last step).
struct InttConsts {
    uint32_t q;
    uint32_t quad;
    uint32_t w, w2, z6, inv2, inv3, det;
    uint32_t twi_off[7];    // offset of level l's inverse twiddles in the twi array
};

}  // namespace lbn
