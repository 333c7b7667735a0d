// HIP kernels for gfx950 (MI355X): the commit / fold / row-evaluation path.
//
//   prepass_kernel        witness (F162, 24 B each) -> the two byte layouts the GEMMs stream
//   tile_gemm_kernel      the int8-MFMA tile GEMM behind both the commitment and the fold
//   afold_kernel          a GEMM's A' = E o A (the key's once, the challenges' per fold)
//   commit_finish_kernel  y mod q -> raw rows and the 648 -> 4 x 162 decomposition
//   chat_kernel           the challenges' transform c_j(-X^4) at every slot, as MFMA A fragments
//   intt_kernel           a limb's inverse transform -> centered coefficients (the folded witness)
//   garner_kernel         the bit-dropped commitment's mixed-radix digits and top values
//   eq_kernel             the eq tensor over F162
//   row_eval_f4_kernel    u_j = sum_i eq_i W[i + wdim j] over F162, a GF(2) product on FP4 MFMA
//
// The algebra of the tile GEMM is in device.hpp.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <type_traits>

#include "device.hpp"
#include "launch.hpp"

typedef int v4i __attribute__((ext_vector_type(4)));
typedef int v16i __attribute__((ext_vector_type(16)));
typedef int v8i __attribute__((ext_vector_type(8)));
typedef float v4f __attribute__((ext_vector_type(4)));
typedef unsigned short u16x2 __attribute__((ext_vector_type(2)));
typedef short s16x2 __attribute__((ext_vector_type(2)));

namespace lbn {

// =============================================================================================
// layout pass
// =============================================================================================

// The byte of ring element i of column c at position n < 81: bit t is coefficient 81 t + n, and
// coefficient 4 m + k of a ring element is bit m of its F162 element k (f162.rs:1-16). Written
// twice in fragment order (byte_offset): the commit's (batch c, K = (n, i)) and the fold's
// (batch i, K = (n, c)). A block covers 16 ring elements x 16 columns, one thread per (ring
// element, column), and stages the bytes in LDS, ring elements fastest for the commit's stores
// and then, in the same 22 KB, columns fastest for the fold's, so that each store gathers its 16
// bytes with one aligned 16-byte read; the 16 stores of one n are 256 contiguous bytes.
constexpr int PRE_I = 16, PRE_C = 16, PRE_ROW = PRE_I * PRE_C + 16;

__global__ __launch_bounds__(PRE_I * PRE_C) void prepass_kernel(const uint64_t *__restrict__ w,
                                                                const uint64_t *const *__restrict__ cols,
                                                                uint32_t column_len, uint32_t L, uint32_t r,
                                                                uint8_t *__restrict__ bytes_c, uint32_t kpad_c,
                                                                uint8_t *__restrict__ bytes_f, uint32_t kpad_f) {
    __shared__ __attribute__((aligned(16))) uint8_t stc[TILE_USED][PRE_ROW];  // [n][c * 16 + i], then [n][i * 16 + c]
    const int il = threadIdx.x & (PRE_I - 1), cl = threadIdx.x / PRE_I;
    const uint32_t i0 = blockIdx.y * PRE_I, c0 = blockIdx.x * PRE_C;
    const uint32_t i = i0 + il, c = c0 + cl;
    uint32_t wd[4][6];
    if (i < L && c < r) {
        // column c is at w + c * column_len elements, or wherever the column table says
        const uint64_t *e = (cols ? cols[c] : w + 3 * (size_t)c * column_len) + 12 * (size_t)i;
#pragma unroll
        for (int k = 0; k < 4; k++)
#pragma unroll
            for (int h = 0; h < 3; h++) {
                const uint64_t x = e[3 * k + h];
                wd[k][2 * h] = (uint32_t)x;
                wd[k][2 * h + 1] = (uint32_t)(x >> 32);
            }
    } else {
#pragma unroll
        for (int k = 0; k < 4; k++)
#pragma unroll
            for (int h = 0; h < 6; h++) wd[k][h] = 0;
    }
    // Byte n's bit t is coefficient 81 t + n, bit (81 t + n) >> 2 of element (81 t + n) & 3. For the
    // word g of bytes 4 g .. 4 g + 3, bit 8 i + t is bit g of the run of element (t + i) & 3 that
    // starts at bit 20 t + ((t + i) >> 2): so the 21 words are the transpose of the 32 runs (one
    // funnel shift each), a 32 x 32 bit transpose in five exchange stages.
    uint32_t x[32];
#pragma unroll
    for (int i = 0; i < 4; i++)
#pragma unroll
        for (int t = 0; t < 8; t++) {
            const int k = (t + i) & 3, p0 = 20 * t + ((t + i) >> 2), w = p0 >> 5, o = p0 & 31;
            x[8 * i + t] = o ? __builtin_amdgcn_alignbit(wd[k][w + 1], wd[k][w], o) : wd[k][w];
        }
    // rows k and k + j (bit j of k clear) exchange row k's columns with bit j set for row k + j's
    // columns without it; the half- and byte-wide stages are one byte permute a word
#pragma unroll
    for (int k = 0; k < 16; k++) {
        const uint32_t a = x[k], b = x[k + 16];
        x[k] = __builtin_amdgcn_perm(b, a, 0x05040100u);       // a's low half, b's low half above
        x[k + 16] = __builtin_amdgcn_perm(b, a, 0x07060302u);  // a's high half, b's high half above
    }
#pragma unroll
    for (int k0 = 0; k0 < 32; k0 += 16)
#pragma unroll
        for (int k = k0; k < k0 + 8; k++) {
            const uint32_t a = x[k], b = x[k + 8];
            x[k] = __builtin_amdgcn_perm(b, a, 0x06020400u);      // bytes a0, b0, a2, b2
            x[k + 8] = __builtin_amdgcn_perm(b, a, 0x07030501u);  // bytes a1, b1, a3, b3
        }
    auto stage = [&](auto jc, uint32_t mlo) {
        constexpr int j = decltype(jc)::value;
#pragma unroll
        for (int k0 = 0; k0 < 32; k0 += 2 * j)
#pragma unroll
            for (int k = k0; k < k0 + j; k++) {
                const uint32_t a = x[k], b = x[k + j];
                // a keeps its columns without bit j (mlo) and takes b's mlo columns shifted up;
                // b keeps its columns with bit j and takes a's shifted down
                x[k] = __builtin_amdgcn_bitop3_b32(a, b << j, mlo, 0xE4);
                x[k + j] = __builtin_amdgcn_bitop3_b32(b, a >> j, ~mlo, 0xE4);
            }
    };
    stage(std::integral_constant<int, 4>{}, 0x0F0F0F0Fu);
    stage(std::integral_constant<int, 2>{}, 0x33333333u);
    stage(std::integral_constant<int, 1>{}, 0x55555555u);
#pragma unroll
    for (int n = 0; n < TILE_USED; n++) stc[n][cl * PRE_I + il] = (uint8_t)(x[n >> 2] >> (8 * (n & 3)));
    __syncthreads();
    // commit: for each (c, n) the 16 bytes of ring elements i0 .. i0 + 15 (i0 is a multiple of 16);
    // the 16 columns of one n are consecutive lanes of a fragment, 256 contiguous bytes, so
    // consecutive threads take them
    if (bytes_c)
        for (int u = threadIdx.x; u < PRE_C * TILE_USED; u += PRE_I * PRE_C) {
            const int cc = u % PRE_C, n = u / PRE_C;
            if (c0 + cc >= r) continue;
            *(uint4 *)(bytes_c + byte_offset(c0 + cc, n, i0, kpad_c)) = *(const uint4 *)&stc[n][cc * PRE_I];
        }
    if (!bytes_f) return;
    __syncthreads();
#pragma unroll
    for (int n = 0; n < TILE_USED; n++) stc[n][il * PRE_C + cl] = (uint8_t)(x[n >> 2] >> (8 * (n & 3)));
    __syncthreads();
    // fold: for each (i, n) the 16 bytes of columns c0 .. c0 + 15 (c0 is a multiple of 16), the 16
    // ring elements of one n by consecutive threads
    for (int u = threadIdx.x; u < PRE_I * TILE_USED; u += PRE_I * PRE_C) {
        const int ii = u % PRE_I, n = u / PRE_I;
        if (i0 + ii >= L) continue;
        *(uint4 *)(bytes_f + byte_offset(i0 + ii, n, c0, kpad_f)) = *(const uint4 *)&stc[n][ii * PRE_C];
    }
}

// =============================================================================================
// the tile GEMM
// =============================================================================================

// One tile over one group of 256 batch elements and one range of K steps:
// y[b][out[m]] (+)= sum_k A'[k][m] x_b[k] mod q, with k = (n, i) as in device.hpp.
//
// Eight waves, wave w owning inputs 32 w .. 32 w + 31 as two 16-input blocks. A' (both digits) is
// staged per K step in LDS and shared by the eight waves. Each wave builds its own B fragments
// in registers: two 16-byte loads per lane give the 16 bytes of its two inputs' K range, and the
// tile's LUT (in LDS) turns every byte into its two digits. A product is kept as three int32
// accumulators per element: ll = A_l x_l, mid = A_l x_h + A_h x_l, hh = A_h x_h.
//
// The products are v_mfma_i32_16x16x64_i8. Rows 0..79 are five 16-row blocks with four digit
// products each; rows 80 and 81 (81 = 5 x 16 + 1, and a quadratic commit tile's first half uses
// 82) share a sixth block whose rows are [A80_l, A80_h, A81_l, A81_h], so two products (with x_l
// and with x_h) give all eight of their digit products. A step is 44 MFMAs of 32 K ops a wave,
// where 96 rows of 32 x 32 x 32 blocks took 24 of 64 K: 17% less MFMA work.
//
// Items are ordered so that the 32 blocks an XCD runs at once are t tiles x 32/t column blocks
// (groups of one K range): each tile's A' range then comes from L2 for 32/t blocks and each
// group's bytes for t.
constexpr int GEMM_THREADS = 512;
// A' is staged per K step in an LDS buffer of A_STEP bytes: 768 16-byte words, the first 512 one
// a thread, the other 256 one a thread of waves 0-3. A' streams from HBM, so its copies are issued
// A_PRE steps earlier than the step barriers need them, into A_PRE + 3 buffers.
constexpr int A_LDS = A_STEP;
constexpr int A_PRE = 4, A_BUFS = A_PRE + 3;
static_assert(A_STEP == 768 * 16 && GEMM_THREADS == 512, "two copies a thread of waves 0-3, one of waves 4-7");
// LDS (the kernel's only LDS is the dynamic block, at address 0):
//   [0, 64 K) replicated LUT | plain LUT (1 KB) | the tile's output slots (96 i16) | A_BUFS A' buffers
// The replicated LUT holds entry (byte, parity, copy) at byte * 256 + parity * 128 + copy * 4: at
// address 0 a lookup's address is one byte permute of the input word and the lane's own bits.
constexpr int LUTR_OFF = 0, LUT_OFF = 65536, OUT_OFF = LUT_OFF + 1024, ABUF_OFF = OUT_OFF + 256;
constexpr int GEMM_LDS = ABUF_OFF + A_BUFS * A_LDS;
static_assert(GEMM_LDS <= 160 * 1024, "the GEMM's LDS fits a CU");

// ll + 2^7 mid + 2^14 hh mod q as a centred residue, from the three int32 accumulators of a
// product. The sum is exact in f64 (|.| < 2^46); t = rint(z / q) is within 2^-17 of z / q's
// nearest integer, so the exact remainder z - t q (one fma) lies in [-(q + 1) / 2, (q + 1) / 2].
__device__ __forceinline__ int mod_q_classes(int ll, int mid, int hh, double qd, double qinv) {
    const double z = __builtin_fma(16384.0, (double)hh, __builtin_fma(128.0, (double)mid, (double)ll));
    const double t = __builtin_rint(z * qinv);
    return (int)__builtin_fma(-t, qd, z);
}

// The same for non-negative classes (the digits of q < UNSIGNED_DIGITS_Q), offset by
// M = 1.5 x 2^52: the double whose bits are M's with a u32 x in the low word is M + x, and every
// sum below stays an exact integer in [2^52, 2^53), where the unit is 1. So D = M + z costs one
// or two conversions instead of three, and with k0 = M / q rounded to a double, fma(D, 1/q, -k0)
// is z / q within q ulp(k0) / 2 <= 3/4 of a unit of q: the remainder M + r = fma(-t, q, D) is
// exact with |r| <= (q + 1) / 2, and r is the low word of its bits (no conversion back). Every f64
// instruction and conversion issues at half the rate of a 32-bit add. For K <= KB_U32 steps of
// 64, ll + 2^7 mid <= K x 127^2 x 257 < 2^32 is one u32 sum.
constexpr uint32_t KB_U32 = 4294967295u / (KSTEP * 127u * 127u * 257u);  // 16 steps, K = 1024
constexpr uint64_t M_BITS = 0x4338000000000000ull;  // 1.5 x 2^52
constexpr double M_MAGIC = 6755399441055744.0;
__device__ __forceinline__ double m_plus(uint32_t x) { return __builtin_bit_cast(double, M_BITS | x); }
__device__ __forceinline__ int mod_q_offset(double D, double qd, double qinv, double k0) {
    const double t = __builtin_rint(__builtin_fma(D, qinv, -k0));
    return (int)(uint32_t)__builtin_bit_cast(uint64_t, __builtin_fma(-t, qd, D));
}
__device__ __forceinline__ int mod_q_short(int ll, int mid, int hh, double qd, double qinv, double k0) {
    const double D = m_plus((uint32_t)ll + ((uint32_t)mid << 7));
    return mod_q_offset(__builtin_fma(16384.0, (double)(uint32_t)hh, D), qd, qinv, k0);
}
__device__ __forceinline__ int mod_q_unsigned(int ll, int mid, int hh, double qd, double qinv, double k0) {
    const double D = __builtin_fma(128.0, (double)(uint32_t)mid, m_plus((uint32_t)ll));
    return mod_q_offset(__builtin_fma(16384.0, (double)(uint32_t)hh, D), qd, qinv, k0);
}

// x mod q in [0, q) for |x| < 2^52, the same way
__device__ __forceinline__ uint32_t mod_q_i64(long long x, double qd, double qinv) {
    const double z = (double)x;
    const double t = __builtin_rint(z * qinv);
    const int r = (int)__builtin_fma(-t, qd, z);
    return (uint32_t)(r < 0 ? r + (int)qd : r);
}

__host__ __device__ inline uint32_t gemm_items(uint32_t tsets, uint32_t csets) {
    return 256 * (uint32_t)(((uint64_t)tsets * csets + 7) / 8);
}

// The block order: item -> (tile, column block). The 32 items an XCD runs at once (item & 7 is
// the XCD a block lands on; the grid is a multiple of 8) are one unit of t tiles x 32/t column
// blocks; units go round-robin over the XCDs, tile sets fastest. Column block cb is group
// cb % groups of K range cb / groups, so the column blocks of a unit share their K range (and so
// the tiles' A' ranges) unless the unit straddles two ranges.
__device__ __forceinline__ bool gemm_item_of(const GemmArgs &g, uint32_t item, uint32_t &tile, uint32_t &cb) {
    const uint32_t sh = __builtin_ctz(g.sched_t), xcd = item & 7, k = item >> 3;
    const uint32_t unit = (k >> 5) * 8 + xcd, within = k & 31;
    const uint32_t ts = unit % g.tsets, cs = unit / g.tsets;
    tile = (ts << sh) + (within & (g.sched_t - 1));
    cb = (cs << (5 - sh)) + (within >> sh);
    return tile < g.ntiles && cb < g.cblocks;
}

// A tile descriptor read by vector loads, as scalar (wave-uniform) values (the kernel writes y,
// so the compiler does not read the descriptor with scalar loads).
__device__ __forceinline__ TileDesc uniform_desc(const TileDesc &t) {
    TileDesc u;
    u.a_off = ((uint64_t)__builtin_amdgcn_readfirstlane((uint32_t)(t.a_off >> 32)) << 32) |
              (uint32_t)__builtin_amdgcn_readfirstlane((uint32_t)t.a_off);
    u.lut_off = __builtin_amdgcn_readfirstlane(t.lut_off);
    u.ee_off = __builtin_amdgcn_readfirstlane(t.ee_off);
    u.out_off = __builtin_amdgcn_readfirstlane(t.out_off);
    u.q = __builtin_amdgcn_readfirstlane(t.q);
    u.par = __builtin_amdgcn_readfirstlane(t.par);
    u.y_off = __builtin_amdgcn_readfirstlane(t.y_off);
    return u;
}

// s_waitcnt immediates (this family's encoding: vmcnt in bits 3:0 and 15:14, expcnt 6:4, lgkmcnt
// 11:8): wait until at most n vector-memory operations of this wave are outstanding, nothing else.
// Copies into LDS by buffer loads count there and are not waited for by a barrier; a wave waits
// for its own before the barrier that publishes them.
constexpr int waitcnt_vm(int n) { return (n & 15) | (7 << 4) | (15 << 8) | ((n >> 4) << 14); }

// A workgroup barrier that waits for this wave's LDS accesses and for all but its last N
// vector-memory operations. __syncthreads() waits for every one (its fence covers copies into
// LDS still in flight); the GEMM's loop keeps its A' copies in flight across step barriers. The
// memory clobber keeps the compiler's LDS accesses on their side of the barrier.
template <int N>
__device__ __forceinline__ void barrier_vm() {
    static_assert(N >= 0 && N < 64, "vmcnt holds 0 .. 63");
    asm volatile("s_waitcnt vmcnt(%0) lgkmcnt(0)\n\ts_barrier" ::"n"(N) : "memory");
}

// The fourth word of a raw buffer descriptor (no stride, no format conversion) on this family.
constexpr int BUF_RSRC_WORD3 = 0x00020000;

// A raw buffer descriptor over [p, p + bytes) for a block-uniform p (read from the first lane, so
// the compiler keeps the descriptor in scalar registers).
__device__ __forceinline__ __amdgpu_buffer_rsrc_t uniform_rsrc(const void *p, uint32_t bytes) {
    const uint64_t v = (uint64_t)p;
    const uint64_t u = ((uint64_t)__builtin_amdgcn_readfirstlane((uint32_t)(v >> 32)) << 32) |
                       (uint32_t)__builtin_amdgcn_readfirstlane((uint32_t)v);
    return __builtin_amdgcn_make_buffer_rsrc((void *)u, (short)0, (int)__builtin_amdgcn_readfirstlane(bytes),
                                             BUF_RSRC_WORD3);
}

// The LDS word at byte address a. The GEMM's dynamic LDS block is the kernel's only LDS, so it
// starts at address 0 and the LUT offsets above are LDS addresses.
typedef __attribute__((address_space(3))) const uint32_t lds_u32;
__device__ __forceinline__ lds_u32 *lds_word(uint32_t a) { return (lds_u32 *)(uintptr_t)(a + LUTR_OFF); }

// The work of one item for the waves of one phase (WG 0: waves 0-3, 1: waves 4-7), compiled
// apart so that the compiler counts each phase's own outstanding loads exactly (the two phases
// issue different numbers of A' copies a step). The kernel has set up the LUT and the tile's
// output slots in LDS; the first barrier here publishes them.
template <bool REP, int WG>
__device__ __forceinline__ void gemm_item(const GemmArgs &g, const TileDesc &td, uint32_t grp, uint32_t kb0,
                                          uint32_t kbs, uint8_t *smem) {
    const uint16_t *lut = (const uint16_t *)(smem + LUT_OFF);  // [parity 2][byte 256]
    const int16_t *outs = (const int16_t *)(smem + OUT_OFF);   // [row 96]
    uint8_t *abuf = smem + ABUF_OFF;                            // [A_BUFS][A_LDS]
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    constexpr int wgrp = WG;

    // v_mfma_i32_16x16x64_i8 operands: lane holds row (A) or input (B) c16 = lane & 15 of a 16-block
    // at K group kg = lane >> 4, i.e. K 16 kg .. 16 kg + 15 of the step's 64. A wave owns 32 inputs
    // as two 16-blocks nb, so a lane holds inputs jl and jl + 16. The staged layouts are in
    // 32 x 32 x 32 fragment order (lane l: row or input l & 31, K 32 kh + 16 (l >> 5) + byte), so
    // the 16-byte piece of (row or input x, K group kg) sits at lane (x & 31) + 32 (kg & 1) of
    // half kh = kg >> 1 there.
    const int c16 = lane & 15, kg = lane >> 4;
    const int jl = 32 * wave + c16;  // this lane's input of block nb = 0 within the group
    // A wave whose 32 batch elements all lie past the last (a narrow commitment's group) issues no
    // products and no lookups. It still issues every load and copy and waits for its input bytes
    // where the others use them (an empty asm statement reads them), so that the step barriers'
    // load counts hold for it too.
    const bool active = __builtin_amdgcn_readfirstlane(grp * GROUP_COLS + 32 * wave) < g.batches;
    const int off32 = c16 + 32 * (kg & 1), kh = kg >> 1;
    // A step holds one sub-ring coefficient n, whose bytes read LUT (par + n) & 1: byte 0 of a
    // lookup address is that parity and this lane's LUT copy, byte 1 comes from the input word
    const uint32_t copy_sel = (uint32_t)((lane & 31) << 2);
    auto par_of = [&](uint32_t kb) { return (td.par + ((kb0 + kb) >> g.run_shift)) & 1u; };

    // the (group, wave) byte fragments: [kb][kh 2][lane 64][16]; this lane's two pieces a step. Both
    // streams are raw buffer loads: the region's base in the (uniform) descriptor, the step in the
    // scalar offset, the lane's fixed offset in the vector one, so a step's address is no VALU work.
    const uint32_t wave_u = __builtin_amdgcn_readfirstlane(wave);
    const uint8_t *bbase = g.bytes + ((uint64_t)grp * GROUP_FRAGS + wave_u) * g.kbs * 2 * 64 * 16;
    const __amdgpu_buffer_rsrc_t brs = uniform_rsrc(bbase, g.kbs * 2048u);
    const uint32_t boff = (kh * 64 + off32) * 16;
    const __amdgpu_buffer_rsrc_t ars = uniform_rsrc(g.afrag + td.a_off, g.kbs * g.a_kb_stride);

    v4i raw[2], a0[5], a1[5], ax, bl[2], bh[2];
    auto consume_raw = [&]() { asm volatile("" ::"v"(raw[0]), "v"(raw[1])); };
    auto load_raw = [&](uint32_t kb) {
#pragma unroll
        for (int nb = 0; nb < 2; nb++)
            raw[nb] = __builtin_amdgcn_raw_buffer_load_b128(brs, boff + 256 * nb, (kb0 + kb) * 2048u, 0);
    };
    // the operands of step kb from its input bytes (raw) through the LUT
    auto build = [&](uint32_t kb) {
        const uint32_t par = par_of(kb);
        const uint32_t lsel = copy_sel | (par << 7);
        const uint16_t *L = lut + (par << 8);
        uint32_t v[2][16];
        // all 32 lookups first, so that their latencies overlap
#pragma unroll
        for (int nb = 0; nb < 2; nb++) {
            const uint32_t w4[4] = {(uint32_t)raw[nb][0], (uint32_t)raw[nb][1], (uint32_t)raw[nb][2],
                                    (uint32_t)raw[nb][3]};
#pragma unroll
            for (int d = 0; d < 4; d++)
#pragma unroll
                for (int e = 0; e < 4; e++) {
                    const uint32_t r = w4[d];
                    v[nb][4 * d + e] = REP ? *lds_word(__builtin_amdgcn_perm(r, lsel, 0x0c0c0400u + 0x100u * e))
                                           : L[(r >> (8 * e)) & 255];
                }
        }
#pragma unroll
        for (int nb = 0; nb < 2; nb++) {
            int ol[4], oh[4];
#pragma unroll
            for (int d = 0; d < 4; d++) {
                const uint32_t p01 = __builtin_amdgcn_perm(v[nb][4 * d + 1], v[nb][4 * d], 0x05010400u);
                const uint32_t p23 = __builtin_amdgcn_perm(v[nb][4 * d + 3], v[nb][4 * d + 2], 0x05010400u);
                ol[d] = (int)__builtin_amdgcn_perm(p23, p01, 0x05040100u);
                oh[d] = (int)__builtin_amdgcn_perm(p23, p01, 0x07060302u);
            }
            bl[nb] = (v4i){ol[0], ol[1], ol[2], ol[3]};
            bh[nb] = (v4i){oh[0], oh[1], oh[2], oh[3]};
        }
    };
    // A' goes from global memory straight into an LDS buffer (buffer loads to LDS, no registers): a
    // step's 768 16-byte words, word w by thread w % 512, a wave 64 consecutive ones (the words
    // from 512 on by waves 0-3).
    auto dma_a = [&](uint32_t kb, int buf) {
#pragma unroll
        for (int i = 0; i < 2; i++)
            if (i == 0 || WG == 0)
                __builtin_amdgcn_raw_ptr_buffer_load_lds(
                    ars, (__attribute__((address_space(3))) void *)(abuf + buf * A_LDS + (64 * wave_u + GEMM_THREADS * i) * 16),
                    16, (tid + GEMM_THREADS * i) * 16, (kb0 + kb) * g.a_kb_stride, 0, 0);
    };
    // this lane's A' pieces in a staged step ([kh][digit][mb 3][lane 64][16]): rows 16 mb16 + c16
    // (mb16 < 5, rows 0..79) and, for the sixth block, rows 80 and 81: lane c16 < 4 holds digit
    // c16 & 1 of row 80 + (c16 >> 1); the other lanes read row 82, which every tile's A' keeps zero
    // (a tile uses at most 82 rows)
    const int aoff = (kh * 2 * NBLK * 64 + off32) * 16;
    const int xdig = c16 & 1, xrow = c16 < 4 ? (c16 >> 1) & 1 : 2;
    const int xoff = ((kh * 2 + xdig) * NBLK * 64 + 2 * 64 + 16 + xrow + 32 * (kg & 1)) * 16;
    // ab is the staged step's LDS byte address (block-uniform, the kernel's LDS starting at 0):
    // a lane's two base addresses cost one add each, the pieces' offsets are immediates
    auto read_a0 = [&](uint32_t ab) {
        typedef __attribute__((address_space(3))) const uint8_t lds_u8;
        ab = __builtin_amdgcn_readfirstlane(ab);  // computed by the scalar unit
        lds_u8 *pa = (lds_u8 *)(uintptr_t)(ab + (uint32_t)aoff);
#pragma unroll
        for (int mb = 0; mb < 5; mb++) a0[mb] = *(const __attribute__((address_space(3))) v4i *)(pa + ((mb >> 1) * 64 + 16 * (mb & 1)) * 16);
#pragma unroll
        for (int mb = 0; mb < 5; mb++)
            a1[mb] = *(const __attribute__((address_space(3))) v4i *)(pa + NBLK * FRAG + ((mb >> 1) * 64 + 16 * (mb & 1)) * 16);
        ax = *(const __attribute__((address_space(3))) v4i *)((lds_u8 *)(uintptr_t)(ab + (uint32_t)xoff));
    };

    // accumulators: [mb16 5][nb 2][class 3] 16 x 16 blocks, and the sixth block's [nb][B digit];
    // the first step's products start them from zero (an inline constant, no clearing moves)
    v4i acc[5][2][3], xacc[2][2];

    // Set-up: the bytes of step 0, A'(0 .. A_PRE) and waves 4-7's part of A'(A_PRE + 1), then (after
    // the barrier that publishes them and the LUT) A'(0)'s pieces and step 0's operands.
    load_raw(0);
#pragma unroll
    for (int st = 0; st <= A_PRE + 1; st++)
        if (st <= A_PRE || wgrp) dma_a(min((uint32_t)st, kbs - 1), st);
    __builtin_amdgcn_s_waitcnt(0);  // this thread's copies into LDS are complete
    __syncthreads();                // the LUT, A'(0) and waves 4-7's part of A'(1)
    if (active) {
        read_a0(ABUF_OFF);
        build(0);
    } else {
        consume_raw();
    }

    // The two waves of a SIMD (w and w + 4) run each step's halves, the products of step kb and
    // the lookups of step kb + 1, in opposite phase: waves 0-3 meet the step barrier between the
    // two, waves 4-7 before the products, so one wave's lookups run while the other's products
    // are in the MFMA pipe. Barrier kb publishes A'(kb + 1). A thread stages step
    // kb + 1 + A_PRE (waves 0-3) or kb + 2 + A_PRE (waves 4-7) into the buffer of step kb - 2
    // (waves 0-3) or kb - 1 (waves 4-7), which every wave has read: with three buffers and no
    // A_PRE that is the opposite-phase schedule's only free buffer, and A_PRE more buffers move
    // each copy A_PRE steps earlier. A wave issues the next step's input bytes and then its A'
    // copy just before its products, which read no LDS; the bytes (the older loads) are waited for
    // before the lookups, the copy only before the barrier that publishes it, A_PRE steps later.
    // The next step's A' pieces (both digits, and the sixth block's) are read ahead, at the end of
    // the step before.
    auto step = [&](uint32_t kb, auto first) {
        constexpr bool FIRST = decltype(first)::value;
        const v4i z = {0, 0, 0, 0};
        auto c = [&](const v4i &x) -> v4i { return FIRST ? z : x; };
        // Per step a thread of waves 0-3 issues 2 byte loads and 2 copies, one of waves 4-7 2 and 1:
        // the copy of A'(kb + 1) has 4 A_PRE (waves 0-3, issued in step kb - A_PRE) or 3 A_PRE
        // (waves 4-7, step kb - 1 - A_PRE) loads after it. The counts hold only if every wave
        // issues these loads every step and reads the bytes (an idle wave through consume_raw;
        // with no reader the compiler would move its byte loads into the branch that reads them).
        if (wgrp) barrier_vm<3 * A_PRE>();  // this wave's copy of A'(kb + 1) is in LDS
        load_raw(min(kb + 1, kbs - 1));
        dma_a(min(kb + 1 + wgrp + A_PRE, kbs - 1), (int)((kb + 1 + wgrp + A_PRE) % A_BUFS));
        // the wave issuing products goes first at the SIMD's arbiter, the other wave's lookups
        // fill in around them (the wave's own priority; the stream's stays the default)
        if (active) {
            __builtin_amdgcn_s_setprio(1);
            // The digit classes in turn over all ten (row block, input block) pairs, so that the two
            // products sharing a mid accumulator are 24 MFMAs apart: a dependent MFMA waits for its
            // accumulator in the pipe, and 16 x 16 x 64 products need about 16 in between to run at
            // full rate.
#pragma unroll
            for (int mb = 0; mb < 5; mb++)
#pragma unroll
                for (int nb = 0; nb < 2; nb++)
                    acc[mb][nb][0] = __builtin_amdgcn_mfma_i32_16x16x64_i8(a0[mb], bl[nb], c(acc[mb][nb][0]), 0, 0, 0);
#pragma unroll
            for (int mb = 0; mb < 5; mb++)
#pragma unroll
                for (int nb = 0; nb < 2; nb++)
                    acc[mb][nb][1] = __builtin_amdgcn_mfma_i32_16x16x64_i8(a0[mb], bh[nb], c(acc[mb][nb][1]), 0, 0, 0);
#pragma unroll
            for (int nb = 0; nb < 2; nb++) {
                xacc[nb][0] = __builtin_amdgcn_mfma_i32_16x16x64_i8(ax, bl[nb], c(xacc[nb][0]), 0, 0, 0);
                xacc[nb][1] = __builtin_amdgcn_mfma_i32_16x16x64_i8(ax, bh[nb], c(xacc[nb][1]), 0, 0, 0);
            }
#pragma unroll
            for (int mb = 0; mb < 5; mb++)
#pragma unroll
                for (int nb = 0; nb < 2; nb++)
                    acc[mb][nb][2] = __builtin_amdgcn_mfma_i32_16x16x64_i8(a1[mb], bh[nb], c(acc[mb][nb][2]), 0, 0, 0);
#pragma unroll
            for (int mb = 0; mb < 5; mb++)
#pragma unroll
                for (int nb = 0; nb < 2; nb++)
                    acc[mb][nb][1] = __builtin_amdgcn_mfma_i32_16x16x64_i8(a1[mb], bl[nb], acc[mb][nb][1], 0, 0, 0);
            __builtin_amdgcn_s_setprio(0);
        }
        if (!wgrp) barrier_vm<4 * A_PRE>();  // this wave's copy of A'(kb + 1) is in LDS
        if (active) {
            build(kb + 1);                                    // its bytes were loaded before the products
            read_a0(ABUF_OFF + ((kb + 1) % A_BUFS) * A_LDS);  // A'(kb + 1) is complete: both phases are past barrier kb
        } else {
            consume_raw();
        }
    };
    step(0, std::true_type{});
    __builtin_amdgcn_sched_barrier(0);  // the peeled first step stays apart from the loop
    for (uint32_t kb = 1; kb < kbs; kb++) step(kb, std::false_type{});
    __builtin_amdgcn_s_waitcnt(0);  // the last (clamped) copies into LDS are complete

    // The epilogue: each of the lane's products ll + 2^7 mid + 2^14 hh mod q, as a residue in
    // [0, q), to its output word. The D block of a 16 x 16 product holds, in register r, row
    // 4 kg + r and column c16. In the sixth block, rows 0..3 (lanes of K group 0) are row 80's and
    // row 81's digit products: with B's low digit [l80, h80, l81, h81] x l, with its high digit x h.
    const uint32_t q = td.q;
    const double qd = (double)q, qinv = 1.0 / qd, k0 = M_MAGIC * qinv;
    auto epilogue = [&](auto reduce) {
#pragma unroll
        for (int nb = 0; nb < 2; nb++) {
            const uint32_t batch = grp * GROUP_COLS + jl + 16 * nb;
            if (batch >= g.batches) continue;
            uint32_t *yb = g.y + (size_t)batch * g.y_batch + td.y_off;
            auto put = [&](int s, int r) {
                if (s < 0) return;  // a padding row
                const uint32_t u = (uint32_t)(r < 0 ? r + (int)q : r);
                if (g.splits == 1)  // block-uniform
                    yb[s] = u;
                else
                    atomicAdd(yb + s, u);
            };
#pragma unroll
            for (int mb = 0; mb < 5; mb++)
#pragma unroll
                for (int rr = 0; rr < 4; rr++)
                    put(outs[16 * mb + 4 * kg + rr], reduce(acc[mb][nb][0][rr], acc[mb][nb][1][rr], acc[mb][nb][2][rr]));
            if (kg == 0) {
                const v4i xl = xacc[nb][0], xh = xacc[nb][1];
                put(outs[80], reduce(xl[0], xh[0] + xl[1], xh[1]));
                put(outs[81], reduce(xl[2], xh[2] + xl[3], xh[3]));
            }
        }
    };
    if (q < UNSIGNED_DIGITS_Q && kbs <= KB_U32)  // block-uniform
        epilogue([&](int ll, int mid, int hh) { return mod_q_short(ll, mid, hh, qd, qinv, k0); });
    else if (q < UNSIGNED_DIGITS_Q)
        epilogue([&](int ll, int mid, int hh) { return mod_q_unsigned(ll, mid, hh, qd, qinv, k0); });
    else
        epilogue([&](int ll, int mid, int hh) { return mod_q_classes(ll, mid, hh, qd, qinv); });
}

// REP: the LUT is replicated 32 times (4-byte entries, copy = lane & 31) so that the 32 lanes of a
// half-wave read distinct banks; its set-up (eight 16-byte stores a thread) pays from 8 K steps.
template <bool REP>
__global__ __launch_bounds__(GEMM_THREADS, 1) void tile_gemm_kernel(GemmArgs g) {
    extern __shared__ __attribute__((aligned(16))) uint8_t smem[];
    uint32_t tile, cb;
    if (!gemm_item_of(g, blockIdx.x, tile, cb)) return;  // block-uniform
    const TileDesc td = uniform_desc(g.tiles[tile]);
    const uint32_t grp = cb % g.groups, kb0 = (cb / g.groups) * g.kps;
    const uint32_t kbs = min(g.kps, g.kbs - kb0);  // this item's steps kb0 .. kb0 + kbs - 1
    const int tid = threadIdx.x;
    if (tid < TILE) ((int16_t *)(smem + OUT_OFF))[tid] = g.out[td.out_off + tid];
    // the tile's LUT into LDS: entry (parity tid >> 8, byte tid & 255), replicated or plain
    const uint32_t lv = g.lut[td.lut_off + tid];
    if (REP) {
        // its 32 copies: eight 16-byte stores of one value
        const uint4 v4 = make_uint4(lv, lv, lv, lv);
        uint8_t *row = smem + LUTR_OFF + (tid & 255) * 256 + (tid >> 8) * 128;
#pragma unroll
        for (int c = 0; c < 8; c++) *(uint4 *)(row + 16 * c) = v4;
    } else {
        ((uint16_t *)(smem + LUT_OFF))[tid] = (uint16_t)lv;
    }
    if (__builtin_amdgcn_readfirstlane(tid >> 8))  // waves 4-7
        gemm_item<REP, 1>(g, td, grp, kb0, kbs, smem);
    else
        gemm_item<REP, 0>(g, td, grp, kb0, kbs, smem);
}

// A' fragments (device.hpp): step kb' = n 2^run_shift + kb of a tile holds
// E[m][n] A[64 kb + k][m] mod q as two digits, from the tile's A fragments (src_steps steps of
// the same layout, src_tile bytes a tile; zero from there on) and its E. A thread per
// (step, kh, mb, lane) of a tile: 16 values, both digits.
constexpr int AF_THREADS = 2 * NBLK * 64;  // 384
__device__ __forceinline__ int digits_of(uint32_t r, uint32_t q, int &hi) {
    if (q < UNSIGNED_DIGITS_Q) {
        hi = (int)(r >> 7);
        return (int)(r & 127);
    }
    const int c = r > (q - 1) / 2 ? (int)r - (int)q : (int)r;
    const int lo = ((c + 64) & 127) - 64;
    hi = (c - lo) >> 7;  // exact: c - lo is a multiple of 128
    return lo;
}
__global__ __launch_bounds__(AF_THREADS) void afold_kernel(const int8_t *__restrict__ src, uint64_t src_tile,
                                                           uint32_t src_steps, const TileDesc *__restrict__ tiles,
                                                           const uint16_t *__restrict__ ee, uint32_t run_shift,
                                                           int8_t *__restrict__ dst) {
    const uint32_t kbp = blockIdx.x, t = blockIdx.y, tid = threadIdx.x;
    const uint32_t kh = tid / (NBLK * 64), mb = (tid / 64) % NBLK, lane = tid % 64;
    const uint32_t n = kbp >> run_shift, kb = kbp & ((1u << run_shift) - 1);
    const TileDesc td = tiles[t];
    const uint32_t q = td.q, m = 32 * mb + (lane & 31);
    const uint32_t e = m < TILE_ROWS ? ee[td.ee_off + m * EP + n] : 0;
    const size_t piece = (size_t)((kh * 2) * NBLK + mb) * FRAG + lane * 16;
    uint32_t lo[4] = {0, 0, 0, 0}, hi[4] = {0, 0, 0, 0};
    if (kb < src_steps && e != 0) {
        const int8_t *s = src + t * src_tile + (size_t)kb * A_STEP + piece;
        const uint4 sl = *(const uint4 *)s, sh = *(const uint4 *)(s + NBLK * FRAG);
        const uint32_t wl[4] = {sl.x, sl.y, sl.z, sl.w}, wh[4] = {sh.x, sh.y, sh.z, sh.w};
        const float qf = 1.0f / (float)q;
#pragma unroll
        for (int w = 0; w < 4; w++)
#pragma unroll
            for (int b = 0; b < 4; b++) {
                const int xl = (int8_t)(wl[w] >> (8 * b)), xh = (int8_t)(wh[w] >> (8 * b));
                // |x| < q and e < q (q <= 19441): |x e| < 2^29, and the float quotient is within
                // q 2^-23 of p / q, so the remainder is within one q of [0, q)
                const int p = (xl + 128 * xh) * (int)e;
                int r = p - (int)__builtin_rintf((float)p * qf) * (int)q;
                r += r < 0 ? (int)q : 0;
                int dh;
                const int dl = digits_of((uint32_t)r, q, dh);
                lo[w] |= (uint32_t)(uint8_t)dl << (8 * b);
                hi[w] |= (uint32_t)(uint8_t)dh << (8 * b);
            }
    }
    int8_t *o = dst + td.a_off + (size_t)kbp * A_STEP + piece;
    *(uint4 *)o = make_uint4(lo[0], lo[1], lo[2], lo[3]);
    *(uint4 *)(o + NBLK * FRAG) = make_uint4(hi[0], hi[1], hi[2], hi[3]);
}

// =============================================================================================
// the commitment's finish: y mod q, raw rows, and the decomposition
// =============================================================================================

__device__ __forceinline__ uint32_t mulmod(uint32_t a, uint32_t b, uint32_t q) {
    return (uint32_t)(((uint64_t)a * b) % q);
}

__device__ __forceinline__ int16_t centered(uint32_t x, uint32_t q) {
    return x > (q - 1) / 2 ? (int16_t)((int)x - (int)q) : (int16_t)x;
}

// block (c, limb), thread s < 162 (ring/mod.rs:274-312 split, 383-417 quadratic)
__global__ void commit_finish_kernel(const uint32_t *__restrict__ yacc, uint32_t limbs,
                                     const FinishLimb *__restrict__ fl,
                                     const uint16_t *__restrict__ slot648,
                                     const uint16_t *__restrict__ qcs, uint32_t *__restrict__ raw,
                                     int16_t *__restrict__ comp, uint32_t cols) {
    const uint32_t c = blockIdx.x, k = blockIdx.y, s = threadIdx.x;
    if (s >= 162) return;
    const FinishLimb &f = fl[k];
    const uint32_t q = f.q;
    const uint32_t *y = yacc + ((size_t)c * limbs + k) * (YACC_PLANES * 648);
    uint32_t *rw = raw + ((size_t)k * cols + c) * 648;
    int16_t *out = comp + (size_t)c * 4 * limbs * 162;  // [t][limb][s]
    if (!f.quad) {
        uint32_t e[4];
#pragma unroll
        for (int t = 0; t < 4; t++) {
            const uint32_t j = slot648[t * 162 + s];
            e[t] = y[j] % q;
            rw[j] = e[t];
        }
        const uint32_t a = e[0] + e[2], d0 = e[0] + q - e[2], cc = e[1] + e[3], d1 = e[1] + q - e[3];
        const uint32_t id = mulmod(d1, f.ii, q);
        const uint32_t m[4] = {(a + cc) % q, (d0 + q - id) % q, (a + 2 * q - cc) % q, (d0 + id) % q};
#pragma unroll
        for (int t = 0; t < 4; t++)
            out[((size_t)t * limbs + k) * 162 + s] = centered(mulmod(m[t], f.twist[t][s], q), q);
    } else {
        const uint32_t jp = qcs[s], jm = qcs[162 + s];
#pragma unroll
        for (int kk = 0; kk < 2; kk++) {
            const uint32_t ep = (y[2 * jp + kk] + y[648 + 2 * jp + kk]) % q;
            const uint32_t em = (y[2 * jm + kk] + y[648 + 2 * jm + kk]) % q;
            rw[2 * jp + kk] = ep;
            rw[2 * jm + kk] = em;
            out[((size_t)kk * limbs + k) * 162 + s] = centered(mulmod(ep + em, f.inv2, q), q);
            out[((size_t)(kk + 2) * limbs + k) * 162 + s] = centered(mulmod(ep + q - em, f.twist[0][s], q), q);
        }
    }
}

// =============================================================================================
// the fold's challenge transform
// =============================================================================================

// The challenges' transform in A fragment order (afold_kernel folds E into it for the fold GEMM):
// tile tau, K = column c, M = row m; the value at a row is
// c(-X^4) at the row's root, sum_i (-1)^{p_i} (1 - 2 s_i) root^{mult expo[m] p_i} (fold.rs:287-300
// embeds c as the coefficient (-1)^m c_m of X^{4m}). One thread per (tile, column, row).
// A block per (16 challenges, tile): the challenges and the root powers staged in LDS; thread
// (row m, quarter g4) computes challenges 4 g4 .. 4 g4 + 3 of the 16 for row m. The 16 challenges
// of a K group are 16 consecutive bytes of a row's fragment, so each thread writes its 4 low and 4
// high digits as one word each.
constexpr int CH_GROUP = 16, CH_THREADS = TILE * CH_GROUP / 4;  // 384
constexpr uint32_t CH_PW_MAX = 1944;                           // the largest conductor
__global__ __launch_bounds__(CH_THREADS) void chat_kernel(const lbn_challenge_dev *__restrict__ ch, uint32_t r,
                                                          uint32_t rpad, const uint16_t *__restrict__ expo,
                                                          const uint16_t *__restrict__ pw, uint32_t cond,
                                                          uint32_t mult, uint32_t q, int8_t *__restrict__ afrag) {
    __shared__ uint16_t pws[CH_PW_MAX];
    __shared__ int16_t cps[CH_GROUP][32][2];  // a challenge's (position, signed coefficient) pairs
    const uint32_t c0 = blockIdx.x * CH_GROUP, tau = blockIdx.y, tid = threadIdx.x;
    for (uint32_t i = tid; i < cond; i += CH_THREADS) pws[i] = pw[i];
    for (uint32_t x = tid; x < CH_GROUP * 32; x += CH_THREADS) {
        const uint32_t k = x / 32, i = x % 32, c = c0 + k;
        int16_t pos = 0, co = 0;  // past the weight (and for padding columns): coefficient 0
        if (c < r && i < ch[c].weight) {
            pos = ch[c].positions[i];
            co = (int16_t)((1 - 2 * (int)((ch[c].signs >> i) & 1)) * ((pos & 1) ? -1 : 1));
        }
        cps[k][i][0] = pos;
        cps[k][i][1] = co;
    }
    __syncthreads();
    const uint32_t m = tid % TILE, g4 = tid / TILE;
    const uint32_t ex = expo[tau * TILE + m];
    uint32_t lo = 0, hi = 0;
    if (ex != 0xFFFF) {
        // the root exponent of position p is a p mod cond, a = mult ex mod cond: with a p < cond
        // 2^8 < 2^19 and rc = floor(2^32 / cond) + 1 < 2^22, (a p rc) >> 32 (one 24-bit high
        // product) is floor(a p / cond) or one more, so one correction gives the remainder
        const uint32_t a = (mult * ex) % cond;
        const uint32_t rc = (uint32_t)(0x100000000ull / cond) + 1;
#pragma unroll
        for (int b = 0; b < 4; b++) {
            const int16_t(*cp)[2] = cps[4 * g4 + b];
            int acc = 0;
            // all 32 pairs (coefficient 0 past the weight): independent terms whose reads overlap
#pragma unroll
            for (uint32_t i = 0; i < 32; i++) {
                const uint32_t ap = __umul24(a, (uint32_t)cp[i][0]);
                int e = (int)(ap - __umul24(__umulhi(ap, rc), cond));
                e += e < 0 ? (int)cond : 0;
                acc += cp[i][1] * (int)pws[e];
            }
            int v = acc % (int)q;
            if (v < 0) v += q;
            // the digits of the tile's operands (UNSIGNED_DIGITS_Q): of v in [0, q) below it, of
            // the centred residue above
            int xl, xh;
            if (q < UNSIGNED_DIGITS_Q) {
                xl = v & 127;
                xh = v >> 7;
            } else {
                const int x = v > (int)(q - 1) / 2 ? v - (int)q : v;
                xl = ((x + 64) & 127) - 64;
                xh = (x - xl) / 128;
            }
            lo |= (uint32_t)(uint8_t)(int8_t)xl << (8 * b);
            hi |= (uint32_t)(uint8_t)(int8_t)xh << (8 * b);
        }
    }
    const uint32_t kbs = rpad / KSTEP;
    const uint32_t kb = c0 / KSTEP, kk = c0 % KSTEP, kh = kk / 32, g2 = (kk % 32) / 16;  // byte kk % 16 = 0
    const uint32_t mb = m / 32, lanev = 32 * g2 + (m % 32);
    int8_t *base = afrag + ((size_t)tau * kbs + kb) * (size_t)A_STEP;
    ((uint32_t *)(base + (((kh * 2 + 0) * 3 + mb) * 64 + lanev) * 16))[g4] = lo;
    ((uint32_t *)(base + (((kh * 2 + 1) * 3 + mb) * 64 + lanev) * 16))[g4] = hi;
}

// =============================================================================================
// the inverse transform of the base prime (scalar.rs:123-177, recursion/limbs.rs:76-127)
// =============================================================================================

// a b mod q for a < 4q, b < q, q <= 19441: the product (< 2^31) by a 24-bit multiply, its quotient
// by q from the float reciprocal (relative error below 2^-22, so within one of the true quotient),
// and at most one correction either way.
__device__ __forceinline__ uint32_t mulmod_f(uint32_t a, uint32_t b, uint32_t q, float qinv) {
    const uint32_t p = __umul24(a, b);
    const uint32_t t = (uint32_t)((float)p * qinv);
    int r = (int)(p - __umul24(t, q));
    r += r < 0 ? (int)q : 0;
    r -= r >= (int)q ? (int)q : 0;
    return (uint32_t)r;
}
// x mod q for x < 2q
__device__ __forceinline__ uint32_t sub_q(uint32_t x, uint32_t q) { return x >= q ? x - q : x; }

__global__ __launch_bounds__(256) void intt_kernel(const uint32_t *__restrict__ vhat, InttConsts ic,
                                                   const uint32_t *__restrict__ twi,
                                                   int16_t *__restrict__ v) {
    __shared__ uint32_t u[648];
    __shared__ uint32_t tw[INTT_TWI_MAX];  // every level's twiddles, read once with the input
    const uint32_t i = blockIdx.x;
    const uint32_t q = ic.q;
    const float qf = ic.qinv;
    // the fold GEMM's K ranges each added a residue in [0, q)
    for (int x = threadIdx.x; x < 648; x += 256) u[x] = vhat[(size_t)i * 648 + x] % q;
    for (uint32_t x = threadIdx.x; x < ic.twi_count; x += 256) tw[x] = twi[x];
    __syncthreads();
    const int radix_s[7] = {2, 2, 2, 3, 3, 3, 3};
    const int degree_s[8] = {648, 324, 162, 81, 27, 9, 3, 1};
    const int radix_q[6] = {2, 2, 3, 3, 3, 3};
    const int degree_q[7] = {648, 324, 162, 54, 18, 6, 2};
    const int top = ic.quad ? 5 : 6;
    for (int level = top; level >= 1; level--) {
        const int n = ic.quad ? degree_q[level] : degree_s[level];
        const int rad = ic.quad ? radix_q[level] : radix_s[level];
        const int m = n / rad;
        const int nbf = 648 / rad;
        for (int b = threadIdx.x; b < nbf; b += 256) {
            const int k = b / m, j = b % m, base = k * n;
            const uint32_t zi = tw[ic.twi_off[level] + k];
            if (rad == 2) {
                const uint32_t y0 = u[base + j], y1 = u[base + m + j];
                u[base + j] = mulmod_f(y0 + y1, ic.inv2, q, qf);
                u[base + m + j] = mulmod_f(mulmod_f(y0 + q - y1, ic.inv2, q, qf), zi, q, qf);
            } else {
                const uint32_t y0 = u[base + j], y1 = u[base + m + j], y2 = u[base + 2 * m + j];
                const uint32_t a0 = mulmod_f(y0 + y1 + y2, ic.inv3, q, qf);
                const uint32_t t1 =
                    mulmod_f(y0 + mulmod_f(ic.w2, y1, q, qf) + mulmod_f(ic.w, y2, q, qf), ic.inv3, q, qf);
                const uint32_t t2 =
                    mulmod_f(y0 + mulmod_f(ic.w, y1, q, qf) + mulmod_f(ic.w2, y2, q, qf), ic.inv3, q, qf);
                u[base + j] = a0;
                u[base + m + j] = mulmod_f(t1, zi, q, qf);
                u[base + 2 * m + j] = mulmod_f(t2, mulmod_f(zi, zi, q, qf), q, qf);
            }
        }
        __syncthreads();
    }
    for (int j = threadIdx.x; j < 324; j += 256) {
        const uint32_t y0 = u[j], y1 = u[j + 324];
        const uint32_t a1 = mulmod_f(y0 + q - y1, ic.det, q, qf);
        const uint32_t a0 = sub_q(y0 + q - mulmod_f(ic.z6, a1, q, qf), q);
        v[(size_t)i * 648 + j] = centered(a0, q);
        v[(size_t)i * 648 + j + 324] = centered(a1, q);
    }
}

// =============================================================================================
// F162 on the device: a carry-less 162 x 162 product (the comb method: for each bit position j
// of a word, b << j is formed once and XORed in at every word of a with bit j set), then the
// reduction x^162 = x^81 + 1 (the same two-step fold as the host's f162_mul).
// =============================================================================================

__device__ __forceinline__ uint64_t shfl_xor_u64(uint64_t v, int m) {
    const int lo = __shfl_xor((int)(uint32_t)v, m, 64), hi = __shfl_xor((int)(uint32_t)(v >> 32), m, 64);
    return ((uint64_t)(uint32_t)hi << 32) | (uint32_t)lo;
}

__device__ void f162_mul_dev(const uint64_t *a, const uint64_t *b, uint64_t *out) {
    uint64_t p0 = 0, p1 = 0, p2 = 0, p3 = 0, p4 = 0, p5 = 0;
    const uint64_t a0 = a[0], a1 = a[1], a2 = a[2], b0 = b[0], b1 = b[1], b2 = b[2];
    for (int j = 0; j < 64; j++) {
        const uint64_t s0 = b0 << j;
        const uint64_t s1 = (b1 << j) | (j ? b0 >> (64 - j) : 0);
        const uint64_t s2 = (b2 << j) | (j ? b1 >> (64 - j) : 0);
        const uint64_t s3 = j ? b2 >> (64 - j) : 0;
        const uint64_t m0 = 0ull - ((a0 >> j) & 1), m1 = 0ull - ((a1 >> j) & 1), m2 = 0ull - ((a2 >> j) & 1);
        p0 ^= s0 & m0;
        p1 ^= (s1 & m0) ^ (s0 & m1);
        p2 ^= (s2 & m0) ^ (s1 & m1) ^ (s0 & m2);
        p3 ^= (s3 & m0) ^ (s2 & m1) ^ (s1 & m2);
        p4 ^= (s3 & m1) ^ (s2 & m2);
        p5 ^= s3 & m2;
    }
    const uint64_t H0 = (p2 >> 34) | (p3 << 30), H1 = (p3 >> 34) | (p4 << 30);
    const uint64_t H2w = (p4 >> 34) | (p5 << 30);
    const uint64_t G0 = (H1 >> 17) | (H2w << 47), G1 = H2w >> 17;
    const uint64_t S0 = H0 ^ G0, S1 = (H1 & ((1ull << 17) - 1)) ^ G1;
    out[0] = p0 ^ H0 ^ G0;
    out[1] = p1 ^ H1 ^ G1 ^ (S0 << 17);
    out[2] = (p2 & ((1ull << 34) - 1)) ^ H2w ^ (S0 >> 47) ^ (S1 << 17);
}

// a b over F162 by P consecutive lanes (P a power of two dividing 64): lane sub = lane % P takes the
// bits j of a with j / (64 / P) == sub in every word, the partial products are XOR-summed across
// the P lanes, and every lane gets the reduced product. The comb's dependency chain is P times
// shorter than f162_mul_dev's.
template <int P>
__device__ void f162_mul_lanes(const uint64_t *a, const uint64_t *b, uint64_t *out) {
    constexpr int J = 64 / P;
    const int j0 = (int)(threadIdx.x % P) * J;
    uint64_t p0 = 0, p1 = 0, p2 = 0, p3 = 0, p4 = 0, p5 = 0;
    const uint64_t a0 = a[0], a1 = a[1], a2 = a[2], b0 = b[0], b1 = b[1], b2 = b[2];
    for (int jj = 0; jj < J; jj++) {
        const int j = j0 + jj;
        const uint64_t s0 = b0 << j;
        const uint64_t s1 = (b1 << j) | (j ? b0 >> (64 - j) : 0);
        const uint64_t s2 = (b2 << j) | (j ? b1 >> (64 - j) : 0);
        const uint64_t s3 = j ? b2 >> (64 - j) : 0;
        const uint64_t m0 = 0ull - ((a0 >> j) & 1), m1 = 0ull - ((a1 >> j) & 1), m2 = 0ull - ((a2 >> j) & 1);
        p0 ^= s0 & m0;
        p1 ^= (s1 & m0) ^ (s0 & m1);
        p2 ^= (s2 & m0) ^ (s1 & m1) ^ (s0 & m2);
        p3 ^= (s3 & m0) ^ (s2 & m1) ^ (s1 & m2);
        p4 ^= (s3 & m1) ^ (s2 & m2);
        p5 ^= s3 & m2;
    }
#pragma unroll
    for (int m = 1; m < P; m <<= 1) {
        p0 ^= shfl_xor_u64(p0, m);
        p1 ^= shfl_xor_u64(p1, m);
        p2 ^= shfl_xor_u64(p2, m);
        p3 ^= shfl_xor_u64(p3, m);
        p4 ^= shfl_xor_u64(p4, m);
        p5 ^= shfl_xor_u64(p5, m);
    }
    const uint64_t H0 = (p2 >> 34) | (p3 << 30), H1 = (p3 >> 34) | (p4 << 30);
    const uint64_t H2w = (p4 >> 34) | (p5 << 30);
    const uint64_t G0 = (H1 >> 17) | (H2w << 47), G1 = H2w >> 17;
    const uint64_t S0 = H0 ^ G0, S1 = (H1 & ((1ull << 17) - 1)) ^ G1;
    out[0] = p0 ^ H0 ^ G0;
    out[1] = p1 ^ H1 ^ G1 ^ (S0 << 17);
    out[2] = (p2 & ((1ull << 34) - 1)) ^ H2w ^ (S0 >> 47) ^ (S1 << 17);
}

// eq(p0)[b] = prod_k (bit k of b ? p_k : 1 + p_k) (eval.rs:51-63), as a product of sub-tables over
// groups of 8 coordinates: part[c][x] for the coordinates 8c .. 8c + 7 (x < 256), built as a
// product tree of depth 3 (pairs of coordinates, quads, then the 256 entries), each product on a
// group of lanes: 8 for the 16 pairs and the 32 quads, 4 for the 256 entries (1024 threads).
// Blocks 0 .. nb0 - 1 build point 0's sub-tables (one a block), the others point 1's (p1 may be
// null with nb1 = 0).
__global__ __launch_bounds__(1024) void eq_part_kernel(const uint64_t *__restrict__ pa, uint32_t rows_a,
                                                       uint64_t *__restrict__ part_a, uint32_t nb0,
                                                       const uint64_t *__restrict__ pb, uint32_t rows_b,
                                                       uint64_t *__restrict__ part_b) {
    __shared__ uint64_t pr[4][4][3];   // pair j (coordinates 2j, 2j + 1), 2 bits
    __shared__ uint64_t qd[2][16][3];  // quad j (coordinates 4j .. 4j + 3), 4 bits
    const bool second = blockIdx.x >= nb0;  // block-uniform
    const uint32_t c = second ? blockIdx.x - nb0 : blockIdx.x, x = threadIdx.x;
    const uint64_t *p0 = second ? pb : pa;
    const uint32_t rows = second ? rows_b : rows_a;
    uint64_t *part = second ? part_b : part_a;
    const uint32_t k0 = 8 * c, nk = min(8u, rows - k0);
    // a coordinate past the last one contributes the factor 1 (its bit is then always 0)
    auto factor = [&](uint32_t k, uint32_t bit, uint64_t *f) {
        if (k >= nk) {
            f[0] = bit ? 0 : 1;
            f[1] = f[2] = 0;
            return;
        }
        f[0] = p0[3 * (k0 + k)] ^ (bit ? 0 : 1);
        f[1] = p0[3 * (k0 + k) + 1];
        f[2] = p0[3 * (k0 + k) + 2];
    };
    if (x < 16 * 8) {  // whole waves: 8 lanes a product
        const uint32_t m = x / 8, j = m >> 2, v = m & 3;
        uint64_t f[3], g2[3], r[3];
        factor(2 * j, v & 1, f);
        factor(2 * j + 1, v >> 1, g2);
        f162_mul_lanes<8>(f, g2, r);
        if (x % 8 == 0) pr[j][v][0] = r[0], pr[j][v][1] = r[1], pr[j][v][2] = r[2];
    }
    __syncthreads();
    if (x < 32 * 8) {
        const uint32_t m = x / 8, j = m >> 4, v = m & 15;
        uint64_t r[3];
        f162_mul_lanes<8>(pr[2 * j][v & 3], pr[2 * j + 1][v >> 2], r);
        if (x % 8 == 0) qd[j][v][0] = r[0], qd[j][v][1] = r[1], qd[j][v][2] = r[2];
    }
    __syncthreads();
    const uint32_t e = x / 4;  // entry < 256; its 4 lanes stay together, so each wave is whole
    uint64_t acc[3];
    f162_mul_lanes<4>(qd[0][e & 15], qd[1][e >> 4], acc);
    if (x % 4 != 0 || e >= (1u << nk)) return;
    uint64_t *o = part + 3 * ((size_t)c * 256 + e);
    o[0] = acc[0];
    o[1] = acc[1];
    o[2] = acc[2];
}

// entry b on 4 lanes (EQ_LANES): the product of its sub-table entries
constexpr int EQ_LANES = 4;
__global__ void eq_kernel(const uint64_t *__restrict__ part, uint32_t rows, uint64_t *__restrict__ eq) {
    const uint32_t id = blockIdx.x * blockDim.x + threadIdx.x, b = id / EQ_LANES;
    const bool live = b < (1u << rows);  // the grid is whole waves; lanes past the end compute entry 0
    const uint32_t bb = live ? b : 0;
    const uint64_t *f = part + 3 * (bb & 255);
    uint64_t acc[3] = {f[0], f[1], f[2]};
    for (uint32_t c = 1; 8 * c < rows; c++) {
        const uint64_t *g2 = part + 3 * ((size_t)c * 256 + ((bb >> (8 * c)) & 255));
        uint64_t t[3];
        f162_mul_lanes<EQ_LANES>(acc, g2, t);
        acc[0] = t[0];
        acc[1] = t[1];
        acc[2] = t[2];
    }
    if (!live || id % EQ_LANES) return;
    eq[3 * (size_t)b] = acc[0];
    eq[3 * (size_t)b + 1] = acc[1];
    eq[3 * (size_t)b + 2] = acc[2];
}

// t = sum_j eq1_j u_j over F162 (eval.rs:185-193): a product per thread, XOR-reduced through the
// wave and the block, then into t (cleared first) with atomic XORs.
__global__ __launch_bounds__(256) void claim_kernel(const uint64_t *__restrict__ eq1, const uint64_t *__restrict__ u,
                                                    uint32_t n, unsigned long long *__restrict__ t) {
    uint64_t a0 = 0, a1 = 0, a2 = 0;
    for (uint32_t j = blockIdx.x * 256 + threadIdx.x; j < n; j += gridDim.x * 256) {
        uint64_t pr[3];
        f162_mul_dev(eq1 + 3 * (size_t)j, u + 3 * (size_t)j, pr);
        a0 ^= pr[0];
        a1 ^= pr[1];
        a2 ^= pr[2];
    }
#pragma unroll
    for (int o = 32; o; o >>= 1) {
        a0 ^= shfl_xor_u64(a0, o);
        a1 ^= shfl_xor_u64(a1, o);
        a2 ^= shfl_xor_u64(a2, o);
    }
    __shared__ uint64_t part[4][3];
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    if (lane == 0) {
        part[wave][0] = a0;
        part[wave][1] = a1;
        part[wave][2] = a2;
    }
    __syncthreads();
    if (threadIdx.x < 3) {
        const uint64_t x = part[0][threadIdx.x] ^ part[1][threadIdx.x] ^ part[2][threadIdx.x] ^ part[3][threadIdx.x];
        atomicXor(&t[threadIdx.x], (unsigned long long)x);
    }
}


// ---------------------------------------------------------------------------------------------
// The row evaluation as a GF(2) product on FP4 MFMA.
//
// u_j = sum_i eq_i W_ij over GF(2^162) is, before the reduction by x^162 + x^81 + 1, the
// polynomial P_j = sum_{b,k} x^{b+k} S_j[b][k] with S_j[b][k] = parity of the number of i whose
// W_ij has bit b and eq_i has bit k. That count is an integer matrix product, bits(W_j)^T
// bits(eq) over K = the row's elements, which v_mfma_f32_16x16x128_f8f6f4 computes exactly with
// FP4 (e2m1) operands: a W bit is the nibble 0.5, an eq bit the nibble 2.0, so a product is 1 and
// the f32 sums are exact integers below 2^24. Both sides are padded from 162 to 176 bits (11
// tiles of 16). A chunk is 128 elements (one K step); per chunk the block
//   - transposes the column's 128 x 192-bit chunk into bit planes P[b][q] (bit b of elements
//     32 q .. 32 q + 31), by 16 x 16 bit transposes on 16-lane rows (DPP butterflies);
//   - stages the chunk's eq operands (precomputed once per evaluation by eqfrag_kernel, the same
//     for every column) in LDS;
//   - multiplies: wave w owns the W tiles B = w, w + 4, w + 8 (rows 16 B .. 16 B + 15), masks
//     each lane's plane word into its 32 operand nibbles, and runs them against the 11 eq tiles.
// The f32 counts' parities are then summed along the anti-diagonals b + k (XOR) and reduced.
// ---------------------------------------------------------------------------------------------
constexpr int RF_CHUNK = 128, RF_TILES = 11, RF_BITS = 16 * RF_TILES;  // 176
constexpr int RF_FRAG_U32 = RF_TILES * 64 * 4;                          // u32 of eq operands per chunk

// The W operand of a lane takes its plane word's bits by masks, with no spreading: register r,
// nibble j holds bit 4 j + r of the word (element 4 j + r of the lane's 32) at the nibble's bit r
// for r < 3 (the values 0.5, 1.0, 2.0) and, shifted down, at bit 0 for r = 3 (0.5). The eq
// operand's register r, nibble j then is eq's bit for the same element with the value that makes
// each product 1: 2.0, 1.0, 0.5, 2.0 (0x4, 0x2, 0x1, 0x4) for r = 0 .. 3.

// eq operands: frag[c][N][lane][4 u32]; register r, nibble j (bits 4 j .. 4 j + 3) of a lane is
// that value when bit 16 N + (lane & 15) of eq_{128 c + 32 (lane >> 4) + 4 j + r} is set.
__global__ __launch_bounds__(256) void eqfrag_kernel(const uint64_t *__restrict__ eq, uint32_t nch,
                                                     uint32_t *__restrict__ frag) {
    const uint32_t id = blockIdx.x * 256 + threadIdx.x;
    if (id >= nch * RF_TILES * 64) return;
    const uint32_t lane = id & 63, nt = (id >> 6) % RF_TILES, c = id / (64 * RF_TILES);
    const uint32_t bit = 16 * nt + (lane & 15);
    constexpr uint32_t nibv[4] = {0x4u, 0x2u, 0x1u, 0x4u};
    uint32_t r4[4] = {0, 0, 0, 0};
    if (bit < 162) {
        const uint64_t *e = eq + 3 * ((size_t)RF_CHUNK * c + 32 * (lane >> 4)) + (bit >> 6);
#pragma unroll
        for (int r = 0; r < 4; r++)
#pragma unroll
            for (int j = 0; j < 8; j++)
                r4[r] |= (uint32_t)((e[3 * (4 * j + r)] >> (bit & 63)) & 1) * (nibv[r] << (4 * j));
    }
    uint4 *o = (uint4 *)frag + (size_t)id;
    *o = make_uint4(r4[0], r4[1], r4[2], r4[3]);
}

// eq(p) entry b from its sub-tables: part[k][(b >> 8 k) & 255] multiplied over the tables, L lanes
// a product (whole waves of lanes with live entries).
template <int L>
__device__ __forceinline__ void eq_entry(const uint64_t *__restrict__ part, uint32_t rows, uint32_t b, uint64_t acc[3]) {
    const uint64_t *f = part + 3 * (b & 255);
    acc[0] = f[0];
    acc[1] = f[1];
    acc[2] = f[2];
    for (uint32_t k = 1; 8 * k < rows; k++) {
        const uint64_t *g2 = part + 3 * ((size_t)k * 256 + ((b >> (8 * k)) & 255));
        uint64_t t[3];
        f162_mul_lanes<L>(acc, g2, t);
        acc[0] = t[0];
        acc[1] = t[1];
        acc[2] = t[2];
    }
}

// One block per 128-element chunk of eq(p0): the chunk's entries (EQ_LANES lanes each, into LDS and
// eq), then the chunk's FP4 operands for the row evaluation (eqfrag_kernel's layout); the blocks
// also clear the two output regions the evaluation accumulates into by XOR.
constexpr int EF_THREADS = RF_CHUNK * EQ_LANES;  // 512
__global__ __launch_bounds__(EF_THREADS) void eq_frag_kernel(const uint64_t *__restrict__ part, uint32_t rows,
                                                             uint64_t *__restrict__ eq, uint32_t *__restrict__ frag,
                                                             uint32_t *__restrict__ z0, uint32_t n0,
                                                             uint32_t *__restrict__ z1, uint32_t n1) {
    __shared__ uint64_t es[RF_CHUNK][3];
    const uint32_t c = blockIdx.x, tid = threadIdx.x, e = tid / EQ_LANES, b = RF_CHUNK * c + e;
    uint64_t acc[3];
    eq_entry<EQ_LANES>(part, rows, b, acc);
    if (tid % EQ_LANES == 0) {
        es[e][0] = acc[0];
        es[e][1] = acc[1];
        es[e][2] = acc[2];
        eq[3 * (size_t)b] = acc[0];
        eq[3 * (size_t)b + 1] = acc[1];
        eq[3 * (size_t)b + 2] = acc[2];
    }
    for (uint32_t w = c * EF_THREADS + tid; w < n0; w += gridDim.x * EF_THREADS) z0[w] = 0;
    for (uint32_t w = c * EF_THREADS + tid; w < n1; w += gridDim.x * EF_THREADS) z1[w] = 0;
    __syncthreads();
    constexpr uint32_t nibv[4] = {0x4u, 0x2u, 0x1u, 0x4u};  // as eqfrag_kernel
    for (uint32_t f = tid; f < RF_TILES * 64; f += EF_THREADS) {
        const uint32_t lane = f & 63, nt = f >> 6, bit = 16 * nt + (lane & 15);
        uint32_t r4[4] = {0, 0, 0, 0};
        if (bit < 162) {
            const uint32_t e0 = 32 * (lane >> 4), wd = bit >> 6, sh = bit & 63;
#pragma unroll
            for (int r = 0; r < 4; r++)
#pragma unroll
                for (int j = 0; j < 8; j++)
                    r4[r] |= (uint32_t)((es[e0 + 4 * j + r][wd] >> sh) & 1) * (nibv[r] << (4 * j));
        }
        ((uint4 *)frag)[(size_t)c * RF_TILES * 64 + f] = make_uint4(r4[0], r4[1], r4[2], r4[3]);
    }
}

// The claim with eq(p1) built from its sub-tables in place: t ^= sum_j eq(p1)_j u_j, CL_LANES lanes
// a product (the grid is whole groups; a group past n adds nothing).
constexpr int CL_LANES = 8;
__global__ __launch_bounds__(256) void claim_parts_kernel(const uint64_t *__restrict__ part1, uint32_t rows1,
                                                          const uint64_t *__restrict__ u, uint32_t n,
                                                          unsigned long long *__restrict__ t) {
    uint64_t a0 = 0, a1 = 0, a2 = 0;
    const uint32_t groups = gridDim.x * 256 / CL_LANES;
    for (uint32_t g = (blockIdx.x * 256 + threadIdx.x) / CL_LANES; g < (n + groups - 1) / groups * groups;
         g += groups) {
        const uint32_t j = g < n ? g : 0;  // whole groups of lanes run the products
        uint64_t e[3], pr[3];
        eq_entry<CL_LANES>(part1, rows1, j, e);
        f162_mul_lanes<CL_LANES>(e, u + 3 * (size_t)j, pr);
        if (g < n && threadIdx.x % CL_LANES == 0) {
            a0 ^= pr[0];
            a1 ^= pr[1];
            a2 ^= pr[2];
        }
    }
#pragma unroll
    for (int o = 32; o; o >>= 1) {
        a0 ^= shfl_xor_u64(a0, o);
        a1 ^= shfl_xor_u64(a1, o);
        a2 ^= shfl_xor_u64(a2, o);
    }
    __shared__ uint64_t part[4][3];
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    if (lane == 0) {
        part[wave][0] = a0;
        part[wave][1] = a1;
        part[wave][2] = a2;
    }
    __syncthreads();
    if (threadIdx.x < 3) {
        const uint64_t x = part[0][threadIdx.x] ^ part[1][threadIdx.x] ^ part[2][threadIdx.x] ^ part[3][threadIdx.x];
        atomicXor(&t[threadIdx.x], (unsigned long long)x);
    }
}


// lanes l and l ^ D of each 16-lane row exchange x, by DPP: quad_perm for D = 1, 2; for D = 4 the
// quad mirror (l ^ 3) then the half-row mirror (l ^ 7); row_ror:8 for D = 8
template <int D>
__device__ __forceinline__ uint32_t dpp_xor(uint32_t x) {
    if constexpr (D == 1) return (uint32_t)__builtin_amdgcn_mov_dpp((int)x, 0xB1, 0xF, 0xF, false);
    if constexpr (D == 2) return (uint32_t)__builtin_amdgcn_mov_dpp((int)x, 0x4E, 0xF, 0xF, false);
    if constexpr (D == 4)
        return (uint32_t)__builtin_amdgcn_mov_dpp(__builtin_amdgcn_mov_dpp((int)x, 0x1B, 0xF, 0xF, false), 0x141,
                                                  0xF, 0xF, false);
    return (uint32_t)__builtin_amdgcn_mov_dpp((int)x, 0x128, 0xF, 0xF, false);
}
// one stage of two 16 x 16 bit transposes (the two 16-bit halves) over a 16-lane row: lanes without
// bit D take the partner's positions without it into theirs with it (rotated left by D), lanes
// with bit D the partner's positions with it into theirs without it (rotated right by D); rot and
// keep are the lane's rotation (32 - D or D) and the positions it keeps.
template <int D>
__device__ __forceinline__ uint32_t bt16_stage(uint32_t v, uint32_t rot, uint32_t keep) {
    const uint32_t y = dpp_xor<D>(v);
    // keep ? v : rotated y, bit by bit: one v_bitop3 (truth table 0xE4 over v, y', keep)
    return __builtin_amdgcn_bitop3_b32(v, __builtin_amdgcn_alignbit(y, y, rot), keep, 0xE4);
}
// a lane's (rot, keep) for stage D with positions-without-D mask m
template <int D>
__device__ __forceinline__ void bt16_lane(uint32_t x16, uint32_t m, uint32_t &rot, uint32_t &keep) {
    rot = (x16 & D) ? D : 32 - D;
    keep = (x16 & D) ? ~m : m;
}


// An FP4 operand is the instruction's first four registers (32 nibbles); it reads no more, so the
// other four are left undefined rather than zeroed.
__device__ __forceinline__ v8i fp4_operand(v4i x) { return __builtin_shufflevector(x, x, 0, 1, 2, 3, -1, -1, -1, -1); }

// A block takes one column and a span of chunks, RF_G chunks a phase: the phase's W words and eq
// operands are loaded one phase ahead into registers; each 16-lane row transposes 16 elements x
// one 32-bit word (two 16 x 16 transposes) by DPP butterflies into the bit planes, kept in LDS as
// P[q][b] = bits b of elements 32 q .. 32 q + 31 (two 16-bit halves). In the products, the eq
// operands are read from LDS RF_LOOK tiles ahead of their MFMAs.
constexpr int RF_G = 2, RF_LOOK = 4;
__global__ __launch_bounds__(256, 2) void row_eval_f4_kernel(const uint32_t *__restrict__ w,
                                                             const uint4 *__restrict__ frag, uint32_t wdim,
                                                             uint32_t cols, uint32_t span,
                                                             unsigned long long *__restrict__ u) {
    __shared__ v4i ebuf[RF_G][RF_TILES * 64];  // vector types: struct copies would go through scratch
    __shared__ uint32_t planes[RF_G][4][RF_BITS];
    __shared__ uint64_t red[4][6];
    const uint32_t tid = threadIdx.x, lane = tid & 63, wave = tid >> 6, x16 = tid & 15;
    const uint32_t j = blockIdx.x, nch = wdim / RF_CHUNK;
    const uint32_t ch0 = blockIdx.y * span, ch1 = min(ch0 + span, nch);
    const uint32_t *wc = w + (size_t)j * wdim * 6;  // the column as u32 words, 6 per element
    uint32_t r8, k8, r4, k4, r2, k2, r1, k1;
    bt16_lane<8>(x16, 0x00FF00FFu, r8, k8);
    bt16_lane<4>(x16, 0x0F0F0F0Fu, r4, k4);
    bt16_lane<2>(x16, 0x33333333u, r2, k2);
    bt16_lane<1>(x16, 0x55555555u, r1, k1);
    v4f acc[3][RF_TILES];
#pragma unroll
    for (int a = 0; a < 3; a++)
#pragma unroll
        for (int n = 0; n < RF_TILES; n++) acc[a][n] = (v4f){0, 0, 0, 0};
    // per chunk, this thread's three transposition tasks (g, s) = (t / 6, t % 6), t = 16 p + tid / 16:
    // word s of element 16 g + x16; and three of the chunk's 704 eq operands
    uint32_t word[RF_G][3];
    v4i ef[RF_G][3];
    // A chunk past the span (its MFMAs are skipped) and operand slots past the 704 (never stored)
    // load from clamped addresses inside the column, with no zero fill.
    // The loads are a chunk's uniform base plus this thread's fixed offsets.
    uint32_t woff[3], foff[3];
#pragma unroll
    for (int p = 0; p < 3; p++) {
        const uint32_t t = 16 * p + (tid >> 4), g = t / 6, sw = t % 6;
        woff[p] = (16 * g + x16) * 6 + sw;
        foff[p] = min(tid + 256 * p, (uint32_t)(RF_TILES * 64 - 1));
    }
    auto fetch = [&](uint32_t ph) {
#pragma unroll
        for (int c = 0; c < RF_G; c++) {
            const uint32_t ch = min(ph + c, ch1 - 1);
            const uint32_t *wch = wc + (size_t)ch * RF_CHUNK * 6;
            const v4i *fch = (const v4i *)frag + (size_t)ch * RF_TILES * 64;
#pragma unroll
            for (int p = 0; p < 3; p++) {
                word[c][p] = wch[woff[p]];
                ef[c][p] = fch[foff[p]];
            }
        }
    };
    // the MFMAs of one phase for a wave with NB W tiles
    auto mma = [&](auto nbc, uint32_t ph) {
        constexpr int NB = decltype(nbc)::value;
#pragma unroll
        for (int c = 0; c < RF_G; c++) {
            if (ph + c >= ch1) break;  // block-uniform
            v4i bq[RF_LOOK];
#pragma unroll
            for (int n = 0; n < RF_LOOK; n++) bq[n] = ebuf[c][n * 64 + lane];
            v8i aop[NB];
#pragma unroll
            for (int a = 0; a < NB; a++) {
                const uint32_t pw = planes[c][lane >> 4][16 * (wave + 4 * a) + (lane & 15)];
                const v4i m4 = {(int)(pw & 0x11111111u), (int)(pw & 0x22222222u), (int)(pw & 0x44444444u),
                                (int)((pw >> 3) & 0x11111111u)};
                aop[a] = fp4_operand(m4);
            }
#pragma unroll
            for (int n = 0; n < RF_TILES; n++) {
                const v4i b4 = bq[n % RF_LOOK];
                if (n + RF_LOOK < RF_TILES) bq[n % RF_LOOK] = ebuf[c][(n + RF_LOOK) * 64 + lane];
                const v8i bop = fp4_operand(b4);
#pragma unroll
                for (int a = 0; a < NB; a++)
                    acc[a][n] =
                        __builtin_amdgcn_mfma_scale_f32_16x16x128_f8f6f4(aop[a], bop, acc[a][n], 4, 4, 0, 0, 0, 0);
            }
        }
    };
    if (ch0 < ch1) fetch(ch0);
    for (uint32_t ph = ch0; ph < ch1; ph += RF_G) {
        __syncthreads();  // the previous phase's planes and eq operands are no longer read
#pragma unroll
        for (int c = 0; c < RF_G; c++)
#pragma unroll
            for (int p = 0; p < 3; p++) {
                const uint32_t f = tid + 256 * p;
                if (f < RF_TILES * 64) ebuf[c][f] = ef[c][p];
                const uint32_t t = 16 * p + (tid >> 4), g = t / 6, sw = t % 6;
                uint32_t v = word[c][p];
                v = bt16_stage<8>(v, r8, k8);
                v = bt16_stage<4>(v, r4, k4);
                v = bt16_stage<2>(v, r2, k2);
                v = bt16_stage<1>(v, r1, k1);
                // low half: bit 32 s + x16 of elements 16 g .. 16 g + 15; high half: bit 32 s + 16 + x16
                // (planes of bits 162 .. 175 are whatever the words hold there: their counts are
                // dropped below)
                uint16_t *pl = (uint16_t *)&planes[c][g >> 1][0] + (g & 1);
                const uint32_t b0 = 32 * sw + x16, b1 = b0 + 16;
                pl[2 * b0] = (uint16_t)v;
                if (b1 < RF_BITS) pl[2 * b1] = (uint16_t)(v >> 16);
            }
        __syncthreads();
        fetch(ph + RF_G);  // the next phase's, or (past the span) clamped loads nobody uses
        if (wave == 3)
            mma(std::integral_constant<int, 2>{}, ph);
        else
            mma(std::integral_constant<int, 3>{}, ph);
    }
    // the counts' parities along the anti-diagonals: lane l holds S[16 B + 4 (l >> 4) + r][16 N + (l & 15)]
    const int nb = wave == 3 ? 2 : 3;
    uint64_t pl[6] = {0, 0, 0, 0, 0, 0};
#pragma unroll
    for (int a = 0; a < 3; a++) {
        if (a >= nb) continue;
        const int B = (int)wave + 4 * a;
#pragma unroll
        for (int n = 0; n < RF_TILES; n++) {
            uint64_t nib = 0;
#pragma unroll
            for (int r = 0; r < 4; r++)  // rows of W bits 162 .. 175 (tile 10) are padding
                if (16 * B + 4 * (int)(lane >> 4) + r < 162) nib |= (uint64_t)((uint32_t)acc[a][n][r] & 1u) << r;
            const int pos = 16 * (B + n);  // B is wave-uniform, so the word is too
            pl[pos >> 6] ^= nib << (pos & 63);
        }
    }
    const uint32_t o = 4 * (lane >> 4) + (lane & 15);  // the lane's offset b + k - 16 (B + N) - r
    uint64_t p[6];
#pragma unroll
    for (int q = 5; q >= 0; q--) p[q] = (pl[q] << o) | (o && q ? pl[q - 1] >> (64 - o) : 0);
#pragma unroll
    for (int q = 0; q < 6; q++)
#pragma unroll
        for (int m = 32; m >= 1; m >>= 1) p[q] ^= shfl_xor_u64(p[q], m);
    if (lane == 0)
#pragma unroll
        for (int q = 0; q < 6; q++) red[wave][q] = p[q];
    __syncthreads();
    if (tid == 0 && ch0 < ch1) {
        uint64_t t[6];
#pragma unroll
        for (int q = 0; q < 6; q++) t[q] = red[0][q] ^ red[1][q] ^ red[2][q] ^ red[3][q];
        const uint64_t H0 = (t[2] >> 34) | (t[3] << 30), H1 = (t[3] >> 34) | (t[4] << 30);
        const uint64_t H2w = (t[4] >> 34) | (t[5] << 30);
        const uint64_t G0 = (H1 >> 17) | (H2w << 47), G1 = H2w >> 17;
        const uint64_t S0 = H0 ^ G0, S1 = (H1 & ((1ull << 17) - 1)) ^ G1;
        atomicXor(&u[3 * (size_t)j], (unsigned long long)(t[0] ^ H0 ^ G0));
        atomicXor(&u[3 * (size_t)j + 1], (unsigned long long)(t[1] ^ H1 ^ G1 ^ (S0 << 17)));
        atomicXor(&u[3 * (size_t)j + 2],
                  (unsigned long long)((t[2] & ((1ull << 34) - 1)) ^ H2w ^ (S0 >> 47) ^ (S1 << 17)));
    }
}

// An 8 x 8 bit matrix in a u64 (byte t = row t), transposed.
__device__ __forceinline__ uint64_t transpose8x8(uint64_t x) {
    // byte t, bit p -> byte p, bit t
    uint64_t t = (x ^ (x >> 7)) & 0x00AA00AA00AA00AAull;
    x ^= t ^ (t << 7);
    t = (x ^ (x >> 14)) & 0x0000CCCC0000CCCCull;
    x ^= t ^ (t << 14);
    t = (x ^ (x >> 28)) & 0x00000000F0F0F0F0ull;
    x ^= t ^ (t << 28);
    return x;
}


// =============================================================================================
// the cross-field switch (switch.rs:137-164, fields/crossfield.rs:149-476, fields/sumcheck.rs:100-145)
// =============================================================================================

// eq(r_hi)[j] over B128 as the product of byte-indexed part tables: parts[c][m] covers bits
// 8c .. 8c + 7 of j (the host builds them from the coordinates in the order they take the bits).
__global__ void switch_eq_kernel(const uint64_t *__restrict__ parts, uint32_t np, uint32_t n, uint64_t *__restrict__ eq) {
    const uint32_t j = blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= n) return;
    uint64_t e0 = parts[2 * (j & 255)], e1 = parts[2 * (j & 255) + 1];
    for (uint32_t c = 1; c < np; c++) {
        const uint64_t *f = parts + 2 * ((size_t)c * 256 + ((j >> (8 * c)) & 255));
        b128_mul(e0, e1, f[0], f[1], e0, e1);
    }
    eq[2 * (size_t)j] = e0;
    eq[2 * (size_t)j + 1] = e1;
}

// v[k] = sum_j bit_k(trace[j]) eq[j] (crossfield.rs:149-203), the trace read from a lifted witness
// (F162 [lo, hi, 0]). Per chunk of 64 j: the 256-entry XOR tables of each group of 8 eq values
// (from two 16-entry nibble tables), the 8 x 128 bit matrix of each group's trace words transposed
// into per-bit bytes, then one lookup per (k, group). Thread t owns k = t & 127 for half the groups.
constexpr int SW_CHUNK = 64;
__global__ __launch_bounds__(256) void switch_partials_kernel(const uint64_t *__restrict__ w,
                                                              const uint64_t *__restrict__ eq, uint32_t n,
                                                              uint32_t span, uint64_t *__restrict__ v) {
    __shared__ uint64_t T[8][256][2];
    __shared__ uint64_t NT[8][2][16][2];
    __shared__ uint64_t ws[SW_CHUNK][2];
    __shared__ uint8_t by[8][128];
    __shared__ uint64_t red[128][2];
    const int t = threadIdx.x;
    const uint32_t nch = n / SW_CHUNK, c0 = blockIdx.x * span, c1 = min(c0 + span, nch);
    uint64_t acc0 = 0, acc1 = 0;
    for (uint32_t ch = c0; ch < c1; ch++) {
        const size_t i0 = (size_t)ch * SW_CHUNK;
        __syncthreads();  // the previous chunk's tables and bytes are no longer read
        if (t < 2 * SW_CHUNK) ws[t >> 1][t & 1] = w[3 * (i0 + (t >> 1)) + (t & 1)];
        {
            const int g = t >> 5, h = (t >> 4) & 1, x = t & 15;
            const uint64_t *e = eq + 2 * (i0 + 8 * g + 4 * h);
            uint64_t a0 = 0, a1 = 0;
#pragma unroll
            for (int k = 0; k < 4; k++) {
                const uint64_t msk = 0ull - (uint64_t)((x >> k) & 1);
                a0 ^= e[2 * k] & msk;
                a1 ^= e[2 * k + 1] & msk;
            }
            NT[g][h][x][0] = a0;
            NT[g][h][x][1] = a1;
        }
        __syncthreads();
        for (int e = t; e < 8 * 256; e += 256) {
            const int g = e >> 8, m = e & 255;
            T[g][m][0] = NT[g][0][m & 15][0] ^ NT[g][1][m >> 4][0];
            T[g][m][1] = NT[g][0][m & 15][1] ^ NT[g][1][m >> 4][1];
        }
        if (t < 128) {
            // group g, byte column c: byte c of the group's 8 words, then byte p = bits 8c + p
            const int g = t >> 4, c = t & 15, wd = c >> 3, sh = 8 * (c & 7);
            uint64_t x = 0;
#pragma unroll
            for (int s = 0; s < 8; s++) x |= ((ws[8 * g + s][wd] >> sh) & 255) << (8 * s);
            x = transpose8x8(x);
#pragma unroll
            for (int p = 0; p < 8; p++) by[g][8 * c + p] = (uint8_t)(x >> (8 * p));
        }
        __syncthreads();
        const int k = t & 127, gh = t >> 7;
#pragma unroll
        for (int q = 0; q < 4; q++) {
            const int g = 4 * gh + q;
            const uint64_t *e = T[g][by[g][k]];
            acc0 ^= e[0];
            acc1 ^= e[1];
        }
    }
    __syncthreads();
    if (t >= 128) {
        red[t - 128][0] = acc0;
        red[t - 128][1] = acc1;
    }
    __syncthreads();
    if (t < 128) {
        atomicXor((unsigned long long *)&v[2 * t], (unsigned long long)(acc0 ^ red[t][0]));
        atomicXor((unsigned long long *)&v[2 * t + 1], (unsigned long long)(acc1 ^ red[t][1]));
    }
}

// a[j] = sum_k bit_k(eq[j]) batch[k] (crossfield.rs:285-325): 16 byte tables of 256 F162
// (psi[c][m] = sum over the bits b of m of batch[8c + b]), staged in LDS, one lookup per byte.
__global__ __launch_bounds__(256) void switch_psi_kernel(const uint64_t *__restrict__ eq, uint32_t n,
                                                         const uint64_t *__restrict__ psi, uint64_t *__restrict__ a) {
    extern __shared__ __attribute__((aligned(16))) uint64_t tab[];  // [16][256][3]
    for (int i = threadIdx.x; i < 16 * 256 * 3; i += blockDim.x) tab[i] = psi[i];
    __syncthreads();
    for (uint32_t j = blockIdx.x * blockDim.x + threadIdx.x; j < n; j += gridDim.x * blockDim.x) {
        const uint64_t x[2] = {eq[2 * (size_t)j], eq[2 * (size_t)j + 1]};
        uint64_t s0 = 0, s1 = 0, s2 = 0;
#pragma unroll
        for (int c = 0; c < 16; c++) {
            const uint64_t *e = tab + 3 * (c * 256 + ((x[c >> 3] >> (8 * (c & 7))) & 255));
            s0 ^= e[0];
            s1 ^= e[1];
            s2 ^= e[2];
        }
        a[3 * (size_t)j] = s0;
        a[3 * (size_t)j + 1] = s1;
        a[3 * (size_t)j + 2] = s2;
    }
}

// One sumcheck message (sumcheck.rs:100-123): e0 = sum_{j < half} a[j] p[j] and
// einf = sum (a[j] + a[j + half]) (p[j] + p[j + half]). top = 0 reads p's third word as zero (the
// lifted trace). Per block an XOR reduction, then one atomic XOR per word into out[6].
__device__ __forceinline__ void msg_terms(const uint64_t *x0, const uint64_t *x1, const uint64_t *y0, const uint64_t *y1,
                                          uint64_t *e) {
    const uint64_t dx[3] = {x0[0] ^ x1[0], x0[1] ^ x1[1], x0[2] ^ x1[2]};
    const uint64_t dy[3] = {y0[0] ^ y1[0], y0[1] ^ y1[1], y0[2] ^ y1[2]};
    uint64_t m0[3], m1[3];
    f162_mul_dev(x0, y0, m0);
    f162_mul_dev(dx, dy, m1);
#pragma unroll
    for (int w = 0; w < 3; w++) {
        e[w] ^= m0[w];
        e[3 + w] ^= m1[w];
    }
}

__device__ __forceinline__ void block_xor6(uint64_t *e, uint64_t (*red)[6], uint64_t *out) {
#pragma unroll
    for (int w = 0; w < 6; w++) red[threadIdx.x][w] = e[w];
    __syncthreads();
    for (int s = 128; s > 0; s >>= 1) {
        if ((int)threadIdx.x < s)
#pragma unroll
            for (int w = 0; w < 6; w++) red[threadIdx.x][w] ^= red[threadIdx.x + s][w];
        __syncthreads();
    }
    if (threadIdx.x < 6) atomicXor((unsigned long long *)&out[threadIdx.x], (unsigned long long)red[0][threadIdx.x]);
}

__global__ __launch_bounds__(256) void switch_msg_kernel(const uint64_t *__restrict__ a, const uint64_t *__restrict__ p,
                                                         uint32_t half, int top, uint64_t *__restrict__ out) {
    __shared__ uint64_t red[256][6];
    uint64_t e[6] = {0, 0, 0, 0, 0, 0};
    for (uint32_t j = blockIdx.x * blockDim.x + threadIdx.x; j < half; j += gridDim.x * blockDim.x) {
        const uint64_t *q0 = p + 3 * (size_t)j, *q1 = p + 3 * ((size_t)j + half);
        const uint64_t y0[3] = {q0[0], q0[1], top ? q0[2] : 0}, y1[3] = {q1[0], q1[1], top ? q1[2] : 0};
        msg_terms(a + 3 * (size_t)j, a + 3 * ((size_t)j + half), y0, y1, e);
    }
    block_xor6(e, red, out);
}

// The fold by the round's challenge r (sumcheck.rs:125-145) fused with the next round's message.
// Multiplying by the fixed r is F2-linear: r d = sum_b T_b[byte b of d] over the 21 bytes of d,
// T_b[m] = r (m x^{8b}), built per block in LDS from the basis r x^i (i < 162). With h = half / 2,
// thread j folds the pairs (j, j + half) and (j + h, j + h + half) into j and j + h (in place;
// p is read from pin, the lifted trace in the first round, and written to pout), then adds the
// next message's terms for the pair (j, j + h). The grid is at most one block per CU, striding.
constexpr int RT_BYTES = 21, RT_LDS = RT_BYTES * 256 * 24;
constexpr int RT_TAIL = 256 * 6 * 8;  // after T: the basis (168 x 24 B), then the reduction (256 x 48 B)
static_assert(168 * 24 <= RT_TAIL && RT_LDS + RT_TAIL <= 160 * 1024, "fold tables in LDS");
__device__ __forceinline__ void mul_r(const uint64_t (*T)[3], const uint64_t *d, uint64_t *o) {
    uint64_t s0 = 0, s1 = 0, s2 = 0;
#pragma unroll
    for (int b = 0; b < RT_BYTES; b++) {
        const uint64_t *e = T[b * 256 + ((d[b >> 3] >> (8 * (b & 7))) & 255)];
        s0 ^= e[0];
        s1 ^= e[1];
        s2 ^= e[2];
    }
    o[0] = s0;
    o[1] = s1;
    o[2] = s2;
}

__global__ __launch_bounds__(256) void switch_fold_msg_kernel(uint64_t *__restrict__ a, const uint64_t *pin,
                                                              uint64_t *pout, uint32_t half, F162v r, int top,
                                                              uint64_t *__restrict__ out) {
    extern __shared__ __attribute__((aligned(16))) uint64_t lds[];
    uint64_t(*T)[3] = (uint64_t(*)[3])lds;                    // [21 * 256][3]
    uint64_t(*B)[3] = (uint64_t(*)[3])(lds + RT_BYTES * 256 * 3);  // [168][3], then the reduction
    uint64_t(*red)[6] = (uint64_t(*)[6])(lds + RT_BYTES * 256 * 3);
    const int t = threadIdx.x;
    if (t < 168) {
        // r x^t: r shifted by t, reduced by x^162 = x^81 + 1 (t >= 162: zero, past the field)
        uint64_t p6[6] = {0, 0, 0, 0, 0, 0};
        if (t < 162) {
            const int ws = t >> 6, bs = t & 63;
#pragma unroll
            for (int k = 0; k < 3; k++) {
                p6[k + ws] ^= r.w[k] << bs;
                if (bs) p6[k + ws + 1] ^= r.w[k] >> (64 - bs);
            }
        }
        const uint64_t H0 = (p6[2] >> 34) | (p6[3] << 30), H1 = (p6[3] >> 34) | (p6[4] << 30);
        const uint64_t H2w = (p6[4] >> 34) | (p6[5] << 30);
        const uint64_t G0 = (H1 >> 17) | (H2w << 47), G1 = H2w >> 17;
        const uint64_t S0 = H0 ^ G0, S1 = (H1 & ((1ull << 17) - 1)) ^ G1;
        B[t][0] = p6[0] ^ H0 ^ G0;
        B[t][1] = p6[1] ^ H1 ^ G1 ^ (S0 << 17);
        B[t][2] = (p6[2] & ((1ull << 34) - 1)) ^ H2w ^ (S0 >> 47) ^ (S1 << 17);
    }
    __syncthreads();
    for (int e = t; e < RT_BYTES * 256; e += 256) {
        const int b = e >> 8, m = e & 255;
        uint64_t s0 = 0, s1 = 0, s2 = 0;
#pragma unroll
        for (int k = 0; k < 8; k++) {
            const uint64_t msk = 0ull - (uint64_t)((m >> k) & 1);
            s0 ^= B[8 * b + k][0] & msk;
            s1 ^= B[8 * b + k][1] & msk;
            s2 ^= B[8 * b + k][2] & msk;
        }
        T[e][0] = s0;
        T[e][1] = s1;
        T[e][2] = s2;
    }
    __syncthreads();  // T is built and B read: B's area becomes the reduction buffer
    const uint32_t h = half / 2;
    uint64_t e6[6] = {0, 0, 0, 0, 0, 0};
    for (uint32_t j = blockIdx.x * blockDim.x + t; j < h; j += gridDim.x * blockDim.x) {
        uint64_t fa[2][3], fp[2][3];
#pragma unroll
        for (int u = 0; u < 2; u++) {
            const size_t i0 = (size_t)j + u * h, i1 = i0 + half;
            const uint64_t x0[3] = {a[3 * i0], a[3 * i0 + 1], a[3 * i0 + 2]};
            const uint64_t da[3] = {x0[0] ^ a[3 * i1], x0[1] ^ a[3 * i1 + 1], x0[2] ^ a[3 * i1 + 2]};
            const uint64_t y0[3] = {pin[3 * i0], pin[3 * i0 + 1], top ? pin[3 * i0 + 2] : 0};
            const uint64_t dp[3] = {y0[0] ^ pin[3 * i1], y0[1] ^ pin[3 * i1 + 1], top ? y0[2] ^ pin[3 * i1 + 2] : 0};
            uint64_t ma[3], mp[3];
            mul_r(T, da, ma);
            mul_r(T, dp, mp);
#pragma unroll
            for (int w = 0; w < 3; w++) {
                fa[u][w] = x0[w] ^ ma[w];
                fp[u][w] = y0[w] ^ mp[w];
            }
        }
#pragma unroll
        for (int u = 0; u < 2; u++) {
            const size_t i0 = (size_t)j + u * h;
#pragma unroll
            for (int w = 0; w < 3; w++) {
                a[3 * i0 + w] = fa[u][w];
                pout[3 * i0 + w] = fp[u][w];
            }
        }
        msg_terms(fa[0], fa[1], fp[0], fp[1], e6);
    }
    block_xor6(e6, red, out);
}

// =============================================================================================
// the bit-dropped digits
// =============================================================================================

// bd/digits.rs:134-182, one thread per coefficient i of count: the limbs' centered coefficients
// cen[k][i] -> residues r_k in [0, q_k) -> mixed-radix digits m_k, top = (m_0 + 2^(d-1)) >> d.
__global__ void garner_kernel(const int16_t *__restrict__ cen, GarnerConsts gc, uint32_t count,
                              uint16_t *__restrict__ top, uint16_t *__restrict__ digits) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= count) return;
    uint32_t m[GARNER_MAX];
#pragma unroll
    for (int k = 0; k < GARNER_MAX; k++) {
        if (k >= (int)gc.limbs) break;
        const uint32_t qk = gc.q[k];
        const int c = cen[(size_t)k * count + i];
        uint32_t x = c < 0 ? (uint32_t)(c + (int)qk) : (uint32_t)c;
#pragma unroll
        for (int j = 0; j < k; j++) x = (x + qk - m[j] % qk) % qk * gc.inv[k][j] % qk;
        m[k] = x;
    }
    top[i] = (uint16_t)((m[0] + gc.round) >> gc.d);
#pragma unroll
    for (int k = 1; k < GARNER_MAX; k++)
        if (k < (int)gc.limbs) digits[(size_t)(k - 1) * count + i] = (uint16_t)m[k];
}

// One v_mfma_i32_16x16x64_i8 with the operand and result lane layouts the tile GEMM assumes:
// A[m][k] row-major 16 x 64 and B[k][n] row-major 64 x 16; lane l holds A row l & 15 and B
// column l & 15 at K 16 (l >> 4) .. + 15; D register r is row 4 (l >> 4) + r, column l & 15.
__global__ void mfma_probe_kernel(const int8_t *__restrict__ A, const int8_t *__restrict__ B,
                                  int32_t *__restrict__ D) {
    const int lane = threadIdx.x;
    int8_t ab[16], bb[16];
#pragma unroll
    for (int j = 0; j < 16; j++) {
        ab[j] = A[(lane & 15) * 64 + 16 * (lane >> 4) + j];
        bb[j] = B[(16 * (lane >> 4) + j) * 16 + (lane & 15)];
    }
    v4i a, b;
    __builtin_memcpy(&a, ab, 16);
    __builtin_memcpy(&b, bb, 16);
    v4i acc = {};
    acc = __builtin_amdgcn_mfma_i32_16x16x64_i8(a, b, acc, 0, 0, 0);
#pragma unroll
    for (int r = 0; r < 4; r++) D[(4 * (lane >> 4) + r) * 16 + (lane & 15)] = acc[r];
}

// The int8 MFMA rate: each wave issues iters x 16 independent v_mfma_i32_16x16x64_i8 (fewer chains
// leave the pipe waiting on accumulators). mode 0: constant operands; 1: random bytes per lane,
// both operands switching between two sets every product; 2: operands like the tile GEMM's, low
// digits in [0, 128) and high digits in [0, 64), A switching every product and B every eight.
// Operands that toggle more bits draw more power and run at a lower clock.
__device__ __forceinline__ int rate_hash(uint32_t x) {
    x *= 2654435761u;
    x ^= x >> 15;
    x *= 0x5bd1e995u;
    x ^= x >> 13;
    return (int)x;
}
__global__ __launch_bounds__(256) void mfma_rate_kernel(int mode, uint32_t iters, int32_t *__restrict__ sink) {
    const uint32_t t = blockIdx.x * blockDim.x + threadIdx.x;
    v4i a0 = {0x01010101, 0x01010101, 0x01010101, 0x01010101}, b0 = a0, a1 = a0, b1 = a0;
    if (mode) {
        a0 = (v4i){rate_hash(16 * t), rate_hash(16 * t + 1), rate_hash(16 * t + 2), rate_hash(16 * t + 3)};
        b0 = (v4i){rate_hash(16 * t + 4), rate_hash(16 * t + 5), rate_hash(16 * t + 6), rate_hash(16 * t + 7)};
        a1 = (v4i){rate_hash(16 * t + 8), rate_hash(16 * t + 9), rate_hash(16 * t + 10), rate_hash(16 * t + 11)};
        b1 = (v4i){rate_hash(16 * t + 12), rate_hash(16 * t + 13), rate_hash(16 * t + 14), rate_hash(16 * t + 15)};
    }
    if (mode == 2) {
        a0 &= 0x7F7F7F7F;
        b0 &= 0x7F7F7F7F;
        a1 &= 0x3F3F3F3F;
        b1 &= 0x3F3F3F3F;
    }
    v4i acc[16];
#pragma unroll
    for (int i = 0; i < 16; i++) acc[i] = (v4i){};
    if (mode == 2) {
        for (uint32_t it = 0; it < iters; it++) {
#pragma unroll
            for (int i = 0; i < 16; i++)
                acc[i] = __builtin_amdgcn_mfma_i32_16x16x64_i8((i & 1) ? a1 : a0, (i & 8) ? b1 : b0, acc[i], 0, 0, 0);
        }
    } else {
        for (uint32_t it = 0; it < iters; it++) {
#pragma unroll
            for (int i = 0; i < 16; i++)
                acc[i] = __builtin_amdgcn_mfma_i32_16x16x64_i8((i & 1) ? a1 : a0, (i & 1) ? b1 : b0, acc[i], 0, 0, 0);
        }
    }
    int s = 0;
#pragma unroll
    for (int i = 0; i < 16; i++)
#pragma unroll
        for (int r = 0; r < 4; r++) s ^= acc[i][r];
    sink[t] = s;
}

// One v_mfma_f32_16x16x128_f8f6f4 with both operands in FP4 (e2m1, format code 4): lane l's A and
// B operand registers are the 16 bytes at A + 16 l and B + 16 l (its 32 nibbles, low nibble of
// byte 0 first), and its 4 result registers go to D + 4 l. scaled = 0 passes E8M0 scale 0 to both
// sides (the unscaled encoding), scaled = 1 passes 127 (2^0).
__global__ void mfma_f4_probe_kernel(const uint32_t *__restrict__ A, const uint32_t *__restrict__ B,
                                     float *__restrict__ D, int scaled) {
    const int lane = threadIdx.x;
    v8i a = {}, b = {};
#pragma unroll
    for (int j = 0; j < 4; j++) {
        a[j] = (int)A[4 * lane + j];
        b[j] = (int)B[4 * lane + j];
    }
    v4f acc = {};
    if (scaled)
        acc = __builtin_amdgcn_mfma_scale_f32_16x16x128_f8f6f4(a, b, acc, 4, 4, 0, 127, 0, 127);
    else
        acc = __builtin_amdgcn_mfma_scale_f32_16x16x128_f8f6f4(a, b, acc, 4, 4, 0, 0, 0, 0);
#pragma unroll
    for (int r = 0; r < 4; r++) D[4 * lane + r] = acc[r];
}

// The FP4 MFMA rate: 16 independent v_mfma_f32_16x16x128_f8f6f4 chains a wave on operands of
// nibbles 0 / 0.5 (a 0-1 pattern: constant, or random per lane and alternating between two sets).
__global__ __launch_bounds__(256) void mfma_f4_rate_kernel(int random, uint32_t iters, float *__restrict__ sink) {
    const uint32_t t = blockIdx.x * blockDim.x + threadIdx.x;
    v8i a0 = {0x11111111, 0x11111111, 0x11111111, 0x11111111, 0, 0, 0, 0}, b0 = a0, a1 = a0, b1 = a0;
    if (random) {
#pragma unroll
        for (int j = 0; j < 4; j++) {
            a0[j] = rate_hash(16 * t + j) & 0x11111111;
            b0[j] = rate_hash(16 * t + 4 + j) & 0x11111111;
            a1[j] = rate_hash(16 * t + 8 + j) & 0x11111111;
            b1[j] = rate_hash(16 * t + 12 + j) & 0x11111111;
        }
    }
    v4f acc[16];
#pragma unroll
    for (int i = 0; i < 16; i++) acc[i] = (v4f){};
    for (uint32_t it = 0; it < iters; it++) {
#pragma unroll
        for (int i = 0; i < 16; i++)
            acc[i] = __builtin_amdgcn_mfma_scale_f32_16x16x128_f8f6f4((i & 1) ? a1 : a0, (i & 1) ? b1 : b0, acc[i], 4,
                                                                      4, 0, 0, 0, 0);
    }
    float s = 0;
#pragma unroll
    for (int i = 0; i < 16; i++)
#pragma unroll
        for (int r = 0; r < 4; r++) s += acc[i][r];
    sink[t] = s;
}

__global__ void fill_u32_kernel(uint32_t *p, size_t n, uint32_t v) {
    const size_t id = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (id < n) p[id] = v;
}

// =============================================================================================
// the verifier on the device (verifier.rs:186-224, fold.rs:799-873)
// =============================================================================================

// The folded witness v ([len_ring][648] int16) as F162 columns for the commit path: column b < nb
// holds bit b of u = v + half, column nb all ones (J), column nb + 1 bit 0 of v; so that
// A v = sum_b 2^b A u_b - half A J. Element e = 4 r + kk of a column holds coefficient 4 m + kk of
// ring element r at bit m (f162.rs:32-41). One thread an element; the squared norm and the largest
// |v| (verifier.rs:186-203) are gathered on the way.
__global__ __launch_bounds__(256) void verify_planes_kernel(const int16_t *__restrict__ v, uint32_t len_ring,
                                                            int half, uint32_t nb, uint64_t *__restrict__ w,
                                                            unsigned long long *__restrict__ stat) {
    const uint64_t column_len = 4ull * len_ring, e = (uint64_t)blockIdx.x * 256 + threadIdx.x;
    uint64_t sq = 0;
    int mx = 0;
    if (e < column_len) {
        const int16_t *vr = v + (e >> 2) * 648 + (e & 3);
#pragma unroll
        for (int word = 0; word < 3; word++) {
            uint64_t pl[15] = {}, par = 0;
            const int m1 = word < 2 ? 64 : 34;
#pragma unroll 2
            for (int i = 0; i < m1; i++) {
                const int x = vr[4 * (64 * word + i)];
                sq += (uint64_t)((int64_t)x * x);
                mx = max(mx, x < 0 ? -x : x);
                const uint32_t u = (uint32_t)(x + half);
#pragma unroll
                for (int b = 0; b < 15; b++) pl[b] |= (uint64_t)((u >> b) & 1) << i;
                par |= (uint64_t)(x & 1) << i;
            }
#pragma unroll
            for (int b = 0; b < 15; b++)
                if ((uint32_t)b < nb) w[((uint64_t)b * column_len + e) * 3 + word] = pl[b];
            w[((uint64_t)nb * column_len + e) * 3 + word] = word < 2 ? ~0ull : (1ull << 34) - 1;
            w[((uint64_t)(nb + 1) * column_len + e) * 3 + word] = par;
        }
    }
#pragma unroll
    for (int o = 32; o; o >>= 1) {
        sq += __shfl_xor(sq, o);
        mx = max(mx, __shfl_xor(mx, o));
    }
    if ((threadIdx.x & 63) == 0) {
        atomicAdd(&stat[0], (unsigned long long)sq);
        atomicMax(&stat[1], (unsigned long long)mx);
    }
}

// acc[t][k][s] += sum over this block's columns j of c_j(slot s of limb k) x comp[j][t][k][s]
// (fold.rs:799-873): block (column chunk, limb), thread s.
constexpr int FC_CHUNK = 32;
__global__ __launch_bounds__(192) void fold_comm_kernel(const lbn_challenge_dev *__restrict__ ch, uint32_t r,
                                                        const int16_t *__restrict__ comp, FoldCommConsts fc,
                                                        const uint32_t *__restrict__ pw,
                                                        const uint16_t *__restrict__ p3,
                                                        unsigned long long *__restrict__ acc) {
    const uint32_t s = threadIdx.x, k = blockIdx.y, j0 = blockIdx.x * FC_CHUNK;
    if (s >= 162) return;
    const uint32_t q = fc.q[k], cond = fc.cond[k], mv = fc.mult[k] * p3[s], limbs = fc.limbs;
    const uint32_t *pwk = pw + fc.pw_off[k];
    int64_t a[4] = {0, 0, 0, 0};
    for (uint32_t j = j0; j < min(j0 + FC_CHUNK, r); j++) {
        const lbn_challenge_dev c = ch[j];
        int64_t cv = 0;
        for (uint32_t i = 0; i < c.weight; i++) {
            const uint32_t p = c.positions[i];
            const int64_t t = pwk[(mv * p) % cond];
            cv += (((c.signs >> i) & 1) ^ (p & 1)) ? -t : t;
        }
        int64_t cs = cv % (int64_t)q;  // centered
        if (cs < 0) cs += q;
        if (cs > (int64_t)(q - 1) / 2) cs -= q;
#pragma unroll
        for (int t = 0; t < 4; t++) a[t] += cs * comp[(((uint64_t)j * 4 + t) * limbs + k) * 162 + s];
    }
#pragma unroll
    for (int t = 0; t < 4; t++) atomicAdd(&acc[(t * limbs + k) * 162 + s], (unsigned long long)a[t]);
}

__global__ void fold_comm_finish_kernel(const long long *__restrict__ acc, FoldCommConsts fc,
                                        int16_t *__restrict__ out) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x, n = 4 * fc.limbs * 162;
    if (i >= n) return;
    const uint32_t k = (i / 162) % fc.limbs;
    const int64_t q = fc.q[k];
    int64_t x = acc[i] % q;
    if (x < 0) x += q;
    out[i] = (int16_t)(x > (q - 1) / 2 ? x - q : x);
}

// =============================================================================================
// wire formats (wire/mod.rs, wire/rans.rs)
// =============================================================================================

// Output word w of a bit-packed object: the fields that overlap bits [32 w, 32 w + 32), found from
// the record / segment / field that holds bit 32 w and walked forward. Bits past the object's end
// are zero (the writer's padding of the last byte).
__global__ __launch_bounds__(256) void wire_pack_kernel(PackDesc d, const uint16_t *__restrict__ src,
                                                        uint32_t *__restrict__ out, uint64_t words) {
    const uint64_t w = (uint64_t)blockIdx.x * 256 + threadIdx.x;
    if (w >= words) return;
    const uint64_t p0 = 32 * w, p1 = min(p0 + 32, d.total_bits);
    uint64_t rec = p0 / d.rec_bits;
    const uint32_t off = (uint32_t)(p0 - rec * d.rec_bits);
    uint32_t l = 0;
    while (l + 1 < d.nseg && off >= d.seg[l + 1].bit0) l++;
    uint32_t idx = (off - d.seg[l].bit0) / d.seg[l].bits;
    uint64_t fs = rec * d.rec_bits + d.seg[l].bit0 + (uint64_t)idx * d.seg[l].bits;
    uint64_t acc = 0;
    while (fs < p1) {
        const PackSeg &sg = d.seg[l];
        const uint32_t raw = src[sg.off + rec * sg.rec_stride + (uint64_t)idx * sg.stride];
        uint32_t v;
        if (sg.q) {
            int r = (int)(int16_t)raw % (int)sg.q;  // rem_euclid
            v = (uint32_t)(r < 0 ? r + (int)sg.q : r);
        } else {
            v = raw & ((1u << sg.bits) - 1);
        }
        acc |= fs >= p0 ? (uint64_t)v << (fs - p0) : (uint64_t)v >> (p0 - fs);
        fs += sg.bits;
        if (++idx == sg.count) {
            idx = 0;
            if (++l == d.nseg) {
                l = 0;
                rec++;
            }
        }
    }
    out[w] = (uint32_t)acc;
}

constexpr int WIRE_THREADS = 256;
constexpr uint32_t HIST_LDS = 12288;  // histograms up to this many bins are gathered in LDS

__global__ void wire_init_kernel(int *stat, unsigned long long *hist, uint32_t nhist) {
    const uint32_t id = blockIdx.x * blockDim.x + threadIdx.x;
    for (uint32_t i = id; i < nhist; i += gridDim.x * blockDim.x) hist[i] = 0;
    if (id == 0) {
        stat[0] = 32767;
        stat[1] = -32768;
        stat[2] = 0;
        stat[3] = 0;
    }
}

// min and max of the message (rans.rs:36-43), 8 coefficients a 16-byte load
__global__ __launch_bounds__(WIRE_THREADS) void wire_minmax_kernel(const int16_t *__restrict__ v, uint64_t nvec,
                                                                   int *stat) {
    int lo = 32767, hi = -32768;
    for (uint64_t i = (uint64_t)blockIdx.x * WIRE_THREADS + threadIdx.x; i < nvec;
         i += (uint64_t)gridDim.x * WIRE_THREADS) {
        const uint4 q = ((const uint4 *)v)[i];
        const uint32_t w4[4] = {q.x, q.y, q.z, q.w};
#pragma unroll
        for (int d = 0; d < 4; d++) {
            const int a = (int16_t)(w4[d] & 0xffff), b = (int16_t)(w4[d] >> 16);
            lo = min(lo, min(a, b));
            hi = max(hi, max(a, b));
        }
    }
#pragma unroll
    for (int o = 32; o; o >>= 1) {
        lo = min(lo, __shfl_xor(lo, o));
        hi = max(hi, __shfl_xor(hi, o));
    }
    __shared__ int sl[WIRE_THREADS / 64], sh[WIRE_THREADS / 64];
    if ((threadIdx.x & 63) == 0) {
        sl[threadIdx.x >> 6] = lo;
        sh[threadIdx.x >> 6] = hi;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        for (int i = 1; i < WIRE_THREADS / 64; i++) {
            lo = min(lo, sl[i]);
            hi = max(hi, sh[i]);
        }
        atomicMin(stat, lo);
        atomicMax(stat + 1, hi);
    }
}

// counts[c - low] (rans.rs:48-53): per block in LDS when the occupied range fits, else straight
// into the global counts
__global__ __launch_bounds__(WIRE_THREADS) void wire_hist_kernel(const int16_t *__restrict__ v, uint64_t nvec,
                                                                 const int *stat, unsigned long long *hist) {
    __shared__ uint32_t h[HIST_LDS];
    const int low = stat[0];
    const uint32_t length = (uint32_t)(stat[1] - low + 1);
    const bool lds = length <= HIST_LDS;
    if (lds) {
        for (uint32_t i = threadIdx.x; i < length; i += WIRE_THREADS) h[i] = 0;
        __syncthreads();
    }
    for (uint64_t i = (uint64_t)blockIdx.x * WIRE_THREADS + threadIdx.x; i < nvec;
         i += (uint64_t)gridDim.x * WIRE_THREADS) {
        const uint4 q = ((const uint4 *)v)[i];
        const uint32_t w4[4] = {q.x, q.y, q.z, q.w};
#pragma unroll
        for (int d = 0; d < 8; d++) {
            const uint32_t bin = (uint32_t)((int)(int16_t)(w4[d >> 1] >> (16 * (d & 1))) - low);
            if (lds)
                atomicAdd(&h[bin], 1u);
            else
                atomicAdd(&hist[bin], 1ull);
        }
    }
    if (lds) {
        __syncthreads();
        for (uint32_t i = threadIdx.x; i < length; i += WIRE_THREADS)
            if (h[i]) atomicAdd(&hist[i], (unsigned long long)h[i]);
    }
}

// The interleaved rANS encoder (rans.rs:67-98) as one wavefront: lane k is rANS lane k and keeps
// its state in a register. The groups run backwards, and all that stays on the wave's serial path
// is each lane's state step: what a group emits is staged at a fixed place, for the scan and the
// scatter below to put in stream order. Lane k's staged word of group g carries the low 16 bits
// the lane may shed for its symbol (bit 16: it did); bit 17 flags an escape's raw-bits word, kept
// in escw. Groups go in blocks of 32, each loaded a block ahead; a block with no escaped symbol
// (the honest case) takes the short step.
template <bool B>
struct step_flag {
    static constexpr bool value = B;
};

template <bool LDS>
__global__ __launch_bounds__(64) void rans_encode_kernel(const int16_t *__restrict__ v, uint64_t coefs,
                                                         const RansSym *__restrict__ sym, uint32_t nsym, int low,
                                                         uint32_t raw, uint32_t *__restrict__ staged,
                                                         uint16_t *__restrict__ escw, uint32_t *__restrict__ stat) {
    extern __shared__ __attribute__((aligned(16))) uint8_t smem[];
    const uint32_t k = threadIdx.x;
    const RansSym *T = sym;
    if (LDS) {
        RansSym *ls = (RansSym *)smem;
        for (uint32_t i = k; i < nsym; i += 64) ls[i] = sym[i];
        __syncthreads();
        T = ls;
    }
    const RansSym esc = T[nsym - 1];
    constexpr int U = 32;
    const uint64_t G = (coefs + 63) / 64, B = (G + U - 1) / U;
    int cur[U], nxt[U];  // a register each (int16 pairs packed in one would wait for both loads)
    // block b's coefficients; every block below the last one is whole (the last holds group G - 1)
    auto load = [&](uint64_t b, int *dst, bool last) {
        const int16_t *p = v + b * U * 64 + k;
#pragma unroll
        for (int u = 0; u < U; u++) dst[u] = last ? v[min((b * U + u) * 64 + k, coefs - 1)] : p[u * 64];
    };
    uint32_t x = RANS_LOW;
    // one group's step; ESC: some lane may hold an escaped symbol; PART: some lanes may be past
    // the message's end
    auto step = [&](auto esc_t, auto part_t, uint64_t g, int c, RansSym e) {
        constexpr bool ESC = decltype(esc_t)::value, PART = decltype(part_t)::value;
        const bool act = !PART || g * 64 + k < coefs;
        bool n1 = false;
        if (ESC) {
            if (act && (e.cs & 0xffff) == RANS_ESCAPED) {  // put_bits (rans.rs:320-328)
                n1 = (x >> (32 - raw)) != 0;
                escw[g * 64 + k] = (uint16_t)x;
                if (n1) x >>= 16;
                x = (x << raw) | (uint32_t)(c - low);
                e = esc;
            }
        }
        const bool n2 = act && x > e.fm;  // put_symbol (rans.rs:306-317)
        const uint32_t w = (x & 0xffff) | (n2 ? 1u << 16 : 0u) | (n1 ? 1u << 17 : 0u);
        const uint32_t xr = n2 ? x >> 16 : x;
        const uint32_t q = (uint32_t)((double)xr * e.rcp);
        const uint32_t xn = __umul24(q, e.cs & 0xffff) + xr + (e.cs >> 16);
        x = act ? xn : x;
        if (!PART || g < G) staged[g * 64 + k] = w;
    };
    using yes = step_flag<true>;
    using no = step_flag<false>;
    load(B - 1, nxt, true);
    for (uint64_t b = B; b-- > 0;) {
#pragma unroll
        for (int u = 0; u < U; u++) cur[u] = nxt[u];
        if (b > 0) load(b - 1, nxt, false);
        // branch-free lookups (a lane past the end holds the last coefficient, a valid symbol)
        RansSym sy[U];
        bool escaped = false;
#pragma unroll
        for (int u = 0; u < U; u++) sy[u] = T[(uint32_t)(cur[u] - low)];
#pragma unroll
        for (int u = 0; u < U; u++) escaped |= (sy[u].cs & 0xffff) == RANS_ESCAPED;
        if (b == B - 1) {  // the last groups: lanes past the end, and groups past it
#pragma unroll
            for (int u = U - 1; u >= 0; u--) step(yes{}, yes{}, b * U + u, cur[u], sy[u]);
        } else if (__builtin_amdgcn_ballot_w64(escaped)) {
#pragma unroll
            for (int u = U - 1; u >= 0; u--) step(yes{}, no{}, b * U + u, cur[u], sy[u]);
        } else {
#pragma unroll
            for (int u = U - 1; u >= 0; u--) step(no{}, no{}, b * U + u, cur[u], sy[u]);
        }
    }
    stat[4 + k] = x;
}

// Per block of 64 consecutive groups: the words each group pushed, their sums over the block's
// later groups (pushed first: the groups run backwards) and the block's total.
__global__ __launch_bounds__(256) void rans_count_kernel(const uint32_t *__restrict__ staged, uint64_t G,
                                                         uint32_t *__restrict__ loc, uint32_t *__restrict__ bsum) {
    __shared__ uint32_t c[64];
    const uint32_t wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const uint64_t g0 = (uint64_t)blockIdx.x * 64;
    for (uint32_t i = 0; i < 16; i++) {
        const uint32_t gl = wave * 16 + i;
        uint32_t n = 0;
        if (g0 + gl < G) {
            const uint32_t w = staged[(g0 + gl) * 64 + lane];
            n = __popcll(__builtin_amdgcn_ballot_w64((w >> 16) & 1)) +
                __popcll(__builtin_amdgcn_ballot_w64((w >> 17) & 1));
        }
        if (lane == 0) c[gl] = n;
    }
    __syncthreads();
    if (wave == 0) {
        const uint32_t own = c[lane];
        uint32_t x = own;  // inclusive suffix sum
#pragma unroll
        for (int o = 1; o < 64; o <<= 1) {
            const uint32_t y = __shfl_down(x, o, 64);
            if (lane + o < 64) x += y;
        }
        if (g0 + lane < G) loc[g0 + lane] = x - own;
        if (lane == 0) bsum[blockIdx.x] = x;
    }
}

// boff[b] = the words pushed before block b's groups, i.e. by the blocks after it, and the total
// (stat[2..3]); one block: each thread sums a run of blocks in push order, the block scans the
// sums, and each thread writes its run's offsets.
constexpr int SCAN_THREADS = 1024;
__global__ __launch_bounds__(SCAN_THREADS) void rans_scan_kernel(const uint32_t *__restrict__ bsum, uint64_t nblk,
                                                                 uint64_t *__restrict__ boff,
                                                                 uint32_t *__restrict__ stat) {
    __shared__ uint64_t part[SCAN_THREADS];
    const uint32_t t = threadIdx.x;
    const uint64_t span = (nblk + SCAN_THREADS - 1) / SCAN_THREADS, r0 = min((uint64_t)t * span, nblk),
                   r1 = min(r0 + span, nblk);
    uint64_t sum = 0;
    for (uint64_t r = r0; r < r1; r++) sum += bsum[nblk - 1 - r];  // r counts blocks in push order
    part[t] = sum;
    __syncthreads();
    for (uint32_t d = 1; d < SCAN_THREADS; d <<= 1) {  // inclusive scan (Hillis-Steele)
        const uint64_t add = t >= d ? part[t - d] : 0;
        __syncthreads();
        part[t] += add;
        __syncthreads();
    }
    uint64_t run = part[t] - sum;
    for (uint64_t r = r0; r < r1; r++) {
        const uint64_t b = nblk - 1 - r;
        boff[b] = run;
        run += bsum[b];
    }
    if (t == SCAN_THREADS - 1) {
        stat[2] = (uint32_t)part[t];
        stat[3] = (uint32_t)(part[t] >> 32);
    }
}

// The stream (rans.rs:93-97): the words in push order reversed. A wave takes one group: the words
// of lanes above k were pushed before lane k's, an escape's raw-bits word before its symbol's.
__global__ __launch_bounds__(256) void rans_scatter_kernel(const uint32_t *__restrict__ staged,
                                                           const uint16_t *__restrict__ escw,
                                                           const uint32_t *__restrict__ loc,
                                                           const uint64_t *__restrict__ boff,
                                                           const uint32_t *__restrict__ stat, uint64_t G,
                                                           uint16_t *__restrict__ stream) {
    const uint64_t g = (uint64_t)blockIdx.x * 4 + (threadIdx.x >> 6);
    if (g >= G) return;
    const uint32_t k = threadIdx.x & 63;
    const uint32_t w = staged[g * 64 + k];
    const bool n2 = (w >> 16) & 1, n1 = (w >> 17) & 1;
    const uint64_t m1 = __builtin_amdgcn_ballot_w64(n1), m2 = __builtin_amdgcn_ballot_w64(n2);
    const uint64_t above = k == 63 ? 0ull : ~0ull << (k + 1);
    const uint64_t top = (uint64_t)stat[2] | (uint64_t)stat[3] << 32;
    const uint64_t p = boff[g >> 6] + loc[g] + __popcll(m1 & above) + __popcll(m2 & above);
    if (n1) stream[top - 1 - p] = escw[g * 64 + k];
    if (n2) stream[top - 1 - (p + (n1 ? 1 : 0))] = (uint16_t)w;
}

// =============================================================================================
// launchers
// =============================================================================================

#define LBN_LAUNCH(...)                                                     \
    do {                                                                    \
        hipLaunchKernelGGL(__VA_ARGS__);                                    \
        hipError_t e_ = hipGetLastError();                                  \
        if (e_ != hipSuccess) return e_;                                    \
    } while (0)

hipError_t launch_mfma_rate(hipStream_t s, int mode, uint32_t blocks, uint32_t iters, int32_t *sink) {
    LBN_LAUNCH(mfma_rate_kernel, dim3(blocks), dim3(256), 0, s, mode, iters, sink);
    return hipSuccess;
}

hipError_t launch_mfma_probe(hipStream_t s, const int8_t *A, const int8_t *B, int32_t *D) {
    LBN_LAUNCH(mfma_probe_kernel, dim3(1), dim3(64), 0, s, A, B, D);
    return hipSuccess;
}

hipError_t launch_mfma_f4_probe(hipStream_t s, const uint32_t *A, const uint32_t *B, float *D, int scaled) {
    LBN_LAUNCH(mfma_f4_probe_kernel, dim3(1), dim3(64), 0, s, A, B, D, scaled);
    return hipSuccess;
}

hipError_t launch_mfma_f4_rate(hipStream_t s, int random, uint32_t blocks, uint32_t iters, float *sink) {
    LBN_LAUNCH(mfma_f4_rate_kernel, dim3(blocks), dim3(256), 0, s, random, iters, sink);
    return hipSuccess;
}

hipError_t launch_fill_u32(hipStream_t s, uint32_t *p, size_t n, uint32_t v) {
    if (!n) return hipSuccess;
    LBN_LAUNCH(fill_u32_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, p, n, v);
    return hipSuccess;
}

hipError_t launch_prepass(hipStream_t s, const uint64_t *w, const uint64_t *const *cols, uint32_t column_len,
                          uint32_t L, uint32_t r, uint8_t *bytes_c, uint32_t kpad_c, uint8_t *bytes_f,
                          uint32_t kpad_f) {
    dim3 grid((r + PRE_C - 1) / PRE_C, (L + PRE_I - 1) / PRE_I);
    LBN_LAUNCH(prepass_kernel, grid, dim3(PRE_I * PRE_C), 0, s, w, cols, column_len, L, r, bytes_c, kpad_c, bytes_f,
               kpad_f);
    return hipSuccess;
}

uint32_t gemm_kb_chunk = 0;

// The K ranges of a launch: at least the exact minimum (and the caller's g.splits), then the
// fewest that minimise a full grid's makespan on 256 CUs, ceil(items / 256) rounds of
// ceil(kbs / splits) steps.
static uint32_t gemm_splits(const GemmArgs &g) {
    const uint32_t cap = gemm_kb_chunk ? gemm_kb_chunk : KB_EXACT;
    const uint32_t s_min = std::max((g.kbs + cap - 1) / cap, std::max(g.splits, 1u));
    const uint64_t base = (uint64_t)g.ntiles * g.groups;
    const uint32_t s_max = std::max(s_min, std::min<uint32_t>(g.kbs, s_min + 2 * (uint32_t)((512 + base - 1) / base)));
    uint32_t best = s_min;
    uint64_t best_cost = ~0ull;
    for (uint32_t sp = s_min; sp <= s_max; sp++) {
        const uint64_t cost = ((base * sp + 255) / 256) * ((g.kbs + sp - 1) / sp);
        if (cost < best_cost) {
            best_cost = cost;
            best = sp;
        }
    }
    return best;
}

hipError_t launch_tile_gemm(hipStream_t s, const GemmArgs &g, uint32_t *splits_out) {
    static const hipError_t attr0 = hipFuncSetAttribute((const void *)tile_gemm_kernel<false>,
                                                        hipFuncAttributeMaxDynamicSharedMemorySize, GEMM_LDS);
    static const hipError_t attr1 = hipFuncSetAttribute((const void *)tile_gemm_kernel<true>,
                                                        hipFuncAttributeMaxDynamicSharedMemorySize, GEMM_LDS);
    if (attr0 != hipSuccess) return attr0;
    if (attr1 != hipSuccess) return attr1;
    if (g.groups == 0 || g.ntiles == 0 || g.kbs == 0) return hipSuccess;
    GemmArgs a = g;
    if (a.sched_t < 4 || a.sched_t > 32 || (a.sched_t & (a.sched_t - 1))) a.sched_t = 8;
    const uint32_t sp = gemm_splits(a);
    a.kps = (a.kbs + sp - 1) / sp;
    a.splits = (a.kbs + a.kps - 1) / a.kps;  // no empty range
    a.cblocks = a.groups * a.splits;
    a.tsets = (a.ntiles + a.sched_t - 1) / a.sched_t;
    a.csets = (a.cblocks + 32 / a.sched_t - 1) / (32 / a.sched_t);
    a.items = gemm_items(a.tsets, a.csets);
    if (splits_out) *splits_out = a.splits;
    // several K ranges add their residues into y, which starts from zero
    if (a.splits > 1) {
        const hipError_t e = launch_fill_u32(s, a.y, (size_t)a.batches * a.y_batch, 0);
        if (e != hipSuccess) return e;
    }
    if (a.kps >= 8)
        LBN_LAUNCH(tile_gemm_kernel<true>, dim3(a.items), dim3(GEMM_THREADS), GEMM_LDS, s, a);
    else
        LBN_LAUNCH(tile_gemm_kernel<false>, dim3(a.items), dim3(GEMM_THREADS), GEMM_LDS, s, a);
    return hipSuccess;
}

hipError_t launch_afold(hipStream_t s, const int8_t *src, uint64_t src_tile, uint32_t src_steps,
                        const TileDesc *tiles, uint32_t ntiles, const uint16_t *ee, uint32_t kpad, int8_t *dst) {
    if (!ntiles || kpad < KSTEP || (kpad & (kpad - 1))) return ntiles ? hipErrorInvalidValue : hipSuccess;
    LBN_LAUNCH(afold_kernel, dim3(gemm_kbs(kpad), ntiles), dim3(AF_THREADS), 0, s, src, src_tile, src_steps, tiles, ee,
               (uint32_t)__builtin_ctz(kpad / KSTEP), dst);
    return hipSuccess;
}

hipError_t launch_commit_finish(hipStream_t s, const uint32_t *yacc, uint32_t limbs,
                                const FinishLimb *fl, const uint16_t *slot648, const uint16_t *qcs,
                                uint32_t *raw, int16_t *comp, uint32_t cols) {
    LBN_LAUNCH(commit_finish_kernel, dim3(cols, limbs), dim3(192), 0, s, yacc, limbs, fl, slot648, qcs, raw,
               comp, cols);
    return hipSuccess;
}

hipError_t launch_chat(hipStream_t s, const lbn_challenge_dev *ch, uint32_t r, uint32_t rpad,
                       const uint16_t *expo, const uint16_t *pw, uint32_t cond, uint32_t mult,
                       uint32_t q, int8_t *afrag, uint32_t ntiles) {
    // rpad is a multiple of KSTEP, so of CH_GROUP; cond <= CH_PW_MAX (CONDUCTOR, CONDUCTOR_QUAD)
    if (cond > CH_PW_MAX || rpad % CH_GROUP) return hipErrorInvalidValue;
    LBN_LAUNCH(chat_kernel, dim3(rpad / CH_GROUP, ntiles), dim3(CH_THREADS), 0, s, ch, r, rpad, expo, pw, cond, mult,
               q, afrag);
    return hipSuccess;
}

hipError_t launch_intt(hipStream_t s, const uint32_t *vhat, uint32_t L, const InttConsts &ic,
                       const uint32_t *twi, int16_t *v) {
    LBN_LAUNCH(intt_kernel, dim3(L), dim3(256), 0, s, vhat, ic, twi, v);
    return hipSuccess;
}

hipError_t launch_garner(hipStream_t s, const int16_t *cen, const GarnerConsts &gc, uint32_t count,
                         uint16_t *top, uint16_t *digits) {
    if (count == 0) return hipSuccess;
    LBN_LAUNCH(garner_kernel, dim3((count + 255) / 256), dim3(256), 0, s, cen, gc, count, top, digits);
    return hipSuccess;
}

hipError_t launch_switch_eq(hipStream_t s, const uint64_t *parts, uint32_t np, uint32_t n, uint64_t *eq) {
    LBN_LAUNCH(switch_eq_kernel, dim3((n + 255) / 256), dim3(256), 0, s, parts, np, n, eq);
    return hipSuccess;
}

hipError_t launch_switch_partials(hipStream_t s, const uint64_t *w, const uint64_t *eq, uint32_t n, uint64_t *v) {
    // n is a multiple of 64 (a trace of 2^l words, l >= 8); about 2 blocks per CU
    const uint32_t nch = n / SW_CHUNK, span = std::max(1u, (nch + 511) / 512), blocks = (nch + span - 1) / span;
    LBN_LAUNCH(switch_partials_kernel, dim3(blocks), dim3(256), 0, s, w, eq, n, span, v);
    return hipSuccess;
}

hipError_t launch_switch_psi(hipStream_t s, const uint64_t *eq, uint32_t n, const uint64_t *psi, uint64_t *a) {
    static const hipError_t attr = hipFuncSetAttribute((const void *)switch_psi_kernel,
                                                       hipFuncAttributeMaxDynamicSharedMemorySize, 16 * 256 * 24);
    if (attr != hipSuccess) return attr;
    const uint32_t blocks = std::min(256u, (n + 255) / 256);
    LBN_LAUNCH(switch_psi_kernel, dim3(blocks), dim3(256), 16 * 256 * 24, s, eq, n, psi, a);
    return hipSuccess;
}

hipError_t launch_switch_msg(hipStream_t s, const uint64_t *a, const uint64_t *p, uint32_t half, int top, uint64_t *out) {
    const uint32_t blocks = std::min(1024u, (half + 255) / 256);
    LBN_LAUNCH(switch_msg_kernel, dim3(blocks), dim3(256), 0, s, a, p, half, top, out);
    return hipSuccess;
}

hipError_t launch_switch_fold_msg(hipStream_t s, uint64_t *a, const uint64_t *pin, uint64_t *pout, uint32_t half,
                                  const uint64_t r[3], int top, uint64_t *out) {
    static const hipError_t attr = hipFuncSetAttribute((const void *)switch_fold_msg_kernel,
                                                       hipFuncAttributeMaxDynamicSharedMemorySize, RT_LDS + RT_TAIL);
    if (attr != hipSuccess) return attr;
    F162v rv{{r[0], r[1], r[2]}};
    const uint32_t h = half / 2, blocks = std::max(1u, std::min(256u, (h + 255) / 256));
    LBN_LAUNCH(switch_fold_msg_kernel, dim3(blocks), dim3(256), RT_LDS + RT_TAIL, s, a, pin, pout, half, rv, top, out);
    return hipSuccess;
}

hipError_t launch_eq(hipStream_t s, const uint64_t *p0, uint32_t rows, uint64_t *part, uint64_t *eq) {
    const uint32_t n = 1u << rows;
    LBN_LAUNCH(eq_part_kernel, dim3((rows + 7) / 8), dim3(1024), 0, s, p0, rows, part, (rows + 7) / 8,
               (const uint64_t *)nullptr, 0u, (uint64_t *)nullptr);
    LBN_LAUNCH(eq_kernel, dim3(((uint64_t)n * EQ_LANES + 255) / 256), dim3(256), 0, s, part, rows, eq);
    return hipSuccess;
}

size_t row_eval_scratch_u32(uint32_t wdim) { return (size_t)(wdim / RF_CHUNK) * RF_FRAG_U32; }

hipError_t launch_row_eval_f4(hipStream_t s, const uint64_t *w, const uint64_t *eq, uint32_t *frag, uint32_t wdim,
                              uint32_t cols, uint64_t *u) {
    // wdim = 2^row_log_len >= 128: whole chunks. A block per (column, span of chunks); the chunks
    // are split over the grid's y until there are about 1024 blocks, each span at most 2^17
    // chunks (the f32 counts stay below 2^24).
    const uint32_t nch = wdim / RF_CHUNK;
    LBN_LAUNCH(eqfrag_kernel, dim3((nch * RF_TILES * 64 + 255) / 256), dim3(256), 0, s, eq, nch, frag);
    return launch_row_eval_frag(s, w, frag, wdim, cols, u);
}

hipError_t launch_row_eval_frag(hipStream_t s, const uint64_t *w, const uint32_t *frag, uint32_t wdim, uint32_t cols,
                                uint64_t *u) {
    const uint32_t nch = wdim / RF_CHUNK;
    const uint32_t gy_want = std::max(1u, std::min(nch, (1024 + cols - 1) / cols));
    uint32_t span = std::min((nch + gy_want - 1) / gy_want, 1u << 17);
    span = (span + RF_G - 1) / RF_G * RF_G;  // whole phases
    const uint32_t gy = (nch + span - 1) / span;
    LBN_LAUNCH(row_eval_f4_kernel, dim3(cols, gy), dim3(256), 0, s, (const uint32_t *)w, (const uint4 *)frag, wdim,
               cols, span, (unsigned long long *)u);
    return hipSuccess;
}

hipError_t launch_eq_parts(hipStream_t s, const uint64_t *p0, uint32_t rows0, uint64_t *part0, const uint64_t *p1,
                           uint32_t rows1, uint64_t *part1) {
    const uint32_t nb0 = (rows0 + 7) / 8, nb1 = p1 ? (rows1 + 7) / 8 : 0;
    if (nb0 + nb1 == 0) return hipSuccess;
    LBN_LAUNCH(eq_part_kernel, dim3(nb0 + nb1), dim3(1024), 0, s, p0, rows0, part0, nb0, p1, rows1, part1);
    return hipSuccess;
}

hipError_t launch_eq_frag(hipStream_t s, const uint64_t *part, uint32_t rows, uint64_t *eq, uint32_t *frag,
                          uint32_t *z0, uint32_t n0, uint32_t *z1, uint32_t n1) {
    // 2^rows >= 128: whole chunks, a block each
    LBN_LAUNCH(eq_frag_kernel, dim3((1u << rows) / RF_CHUNK), dim3(EF_THREADS), 0, s, part, rows, eq, frag, z0, n0,
               z1, n1);
    return hipSuccess;
}

hipError_t launch_claim_parts(hipStream_t s, const uint64_t *part1, uint32_t rows1, const uint64_t *u, uint32_t n,
                              uint64_t *t) {
    const uint32_t blocks = std::max(1u, std::min(256u, (n * CL_LANES + 255) / 256));
    LBN_LAUNCH(claim_parts_kernel, dim3(blocks), dim3(256), 0, s, part1, rows1, u, n, (unsigned long long *)t);
    return hipSuccess;
}

hipError_t launch_wire_pack(hipStream_t s, const PackDesc &d, const uint16_t *src, uint32_t *out) {
    const uint64_t words = (d.total_bits + 31) / 32;
    if (words == 0) return hipSuccess;
    LBN_LAUNCH(wire_pack_kernel, dim3((uint32_t)((words + 255) / 256)), dim3(256), 0, s, d, src, out, words);
    return hipSuccess;
}

hipError_t launch_wire_stats(hipStream_t s, const int16_t *v, uint64_t coefs, int *stat, unsigned long long *hist,
                             uint32_t nhist) {
    // coefs is a multiple of 648, so of 8: whole 16-byte vectors
    const uint64_t nvec = coefs / 8;
    LBN_LAUNCH(wire_init_kernel, dim3(64), dim3(256), 0, s, stat, hist, nhist);
    if (nvec == 0) return hipSuccess;
    const uint32_t blocks = (uint32_t)std::min<uint64_t>(512, (nvec + 4 * WIRE_THREADS - 1) / (4 * WIRE_THREADS));
    LBN_LAUNCH(wire_minmax_kernel, dim3(blocks), dim3(WIRE_THREADS), 0, s, v, nvec, stat);
    LBN_LAUNCH(wire_hist_kernel, dim3(blocks), dim3(WIRE_THREADS), 0, s, v, nvec, (const int *)stat, hist);
    return hipSuccess;
}

hipError_t launch_rans_encode(hipStream_t s, const int16_t *v, uint64_t coefs, const RansSym *sym, uint32_t nsym,
                              int low, uint32_t raw, uint32_t *staged, uint16_t *escw, uint32_t *stat) {
    if (nsym <= RANS_LDS_SYMS) {
        static const hipError_t attr = hipFuncSetAttribute((const void *)rans_encode_kernel<true>,
                                                           hipFuncAttributeMaxDynamicSharedMemorySize,
                                                           RANS_LDS_SYMS * sizeof(RansSym));
        if (attr != hipSuccess) return attr;
        LBN_LAUNCH(rans_encode_kernel<true>, dim3(1), dim3(64), nsym * sizeof(RansSym), s, v, coefs, sym, nsym, low,
                   raw, staged, escw, stat);
    } else {
        LBN_LAUNCH(rans_encode_kernel<false>, dim3(1), dim3(64), 0, s, v, coefs, sym, nsym, low, raw, staged, escw,
                   stat);
    }
    return hipSuccess;
}

hipError_t launch_rans_scatter(hipStream_t s, const uint32_t *staged, const uint16_t *escw, uint32_t *cnt,
                               uint64_t *offs, uint32_t *stat, uint64_t groups, uint16_t *stream) {
    if (groups == 0) return hipSuccess;
    // cnt: the per-group sums inside their block, then the block totals; offs: the block offsets
    const uint64_t nblk = (groups + 63) / 64;
    uint32_t *bsum = cnt + groups;
    LBN_LAUNCH(rans_count_kernel, dim3((uint32_t)nblk), dim3(256), 0, s, staged, groups, cnt, bsum);
    LBN_LAUNCH(rans_scan_kernel, dim3(1), dim3(SCAN_THREADS), 0, s, (const uint32_t *)bsum, nblk, offs, stat);
    LBN_LAUNCH(rans_scatter_kernel, dim3((uint32_t)((groups + 3) / 4)), dim3(256), 0, s, staged, escw,
               (const uint32_t *)cnt, (const uint64_t *)offs, (const uint32_t *)stat, groups, stream);
    return hipSuccess;
}

hipError_t launch_verify_planes(hipStream_t s, const int16_t *v, uint32_t len_ring, int half, uint32_t nb,
                                uint64_t *w, unsigned long long *stat) {
    const uint64_t column_len = 4ull * len_ring;
    LBN_LAUNCH(verify_planes_kernel, dim3((uint32_t)((column_len + 255) / 256)), dim3(256), 0, s, v, len_ring, half, nb,
               w, stat);
    return hipSuccess;
}

hipError_t launch_fold_comm(hipStream_t s, const lbn_challenge_dev *ch, uint32_t r, const int16_t *comp,
                            const FoldCommConsts &fc, const uint32_t *pw, const uint16_t *p3, long long *acc,
                            int16_t *out) {
    const uint32_t n = 4 * fc.limbs * 162;
    LBN_LAUNCH(fill_u32_kernel, dim3((2 * n + 255) / 256), dim3(256), 0, s, (uint32_t *)acc, (size_t)(2 * n), 0u);
    if (r) LBN_LAUNCH(fold_comm_kernel, dim3((r + FC_CHUNK - 1) / FC_CHUNK, fc.limbs), dim3(192), 0, s, ch, r, comp, fc,
                      pw, p3, (unsigned long long *)acc);
    LBN_LAUNCH(fold_comm_finish_kernel, dim3((n + 255) / 256), dim3(256), 0, s, (const long long *)acc, fc, out);
    return hipSuccess;
}

hipError_t launch_claim(hipStream_t s, const uint64_t *eq1, const uint64_t *u, uint32_t n, uint64_t *t) {
    const uint32_t blocks = std::max(1u, std::min(64u, (n + 255) / 256));
    LBN_LAUNCH(claim_kernel, dim3(blocks), dim3(256), 0, s, eq1, u, n, (unsigned long long *)t);
    return hipSuccess;
}

}  // namespace lbn
