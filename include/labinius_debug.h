/*
 * Test-only introspection of liblabinius.so. No counterpart in the reference: it exists so that
 * the CPU test suite can replay the GEMM tiles' tables (built on the host) without a device.
 */
#ifndef LABINIUS_DEBUG_H
#define LABINIUS_DEBUG_H

#include "labinius.h"

#ifdef __cplusplus
extern "C" {
#endif

/* which: 0 commit A fragments (int8), 1 commit LUTs (u16), 2 commit E tables [n 81][m 96] (u16),
 * 3 commit output slots (i16), 4 commit tile descriptors, 5 fold LUTs, 6 fold E tables,
 * 7 fold output slots, 8 fold root exponents (u16), 9 fold tile descriptors, 10 finish constants,
 * 11 the base root's powers (u16), 12 inverse-transform constants, 13 inverse twiddles (u32),
 * 14 the expanded key A ([limb][len_ring][648] int16).
 * With dst == NULL only *bytes is set. */
int lbn_debug_tables(const lbn_params *params, const uint8_t seed[32], int which, void *dst, size_t cap,
                     size_t *bytes);

/* One v_mfma_i32_16x16x64_i8 on the device with the operand and result lane layouts the tile GEMM
 * uses: a is 16 x 64 (row-major [m][k]), b is 64 x 16 ([k][n]), d receives the 16 x 16 int32 product. */
int lbn_debug_mfma_probe(const int8_t *a, const int8_t *b, int32_t *d);

/* The tile GEMMs' longest K range in steps of 64 (0: the default, the largest the int32
 * accumulators hold exactly). Tests set a short one to run many K ranges (added into the output)
 * on small shapes. Process-wide; not for concurrent use with running provers. */
int lbn_debug_gemm_chunk(uint32_t kb);

/* Measurement, no counterpart in the reference: the device's int8 MFMA rate in TOP/s
 * (v_mfma_i32_16x16x64_i8, 16 independent chains a wave, all CUs, after a warm-up of about 10 ms),
 * on constant operands (operands = 0), on random bytes switching every product (1), or on
 * operands like the tile GEMM's: random digits in [0, 128) and [0, 64), B switching every eighth
 * product (2). Operands that toggle more bits draw more power and clock lower. */
int lbn_debug_mfma_rate(int operands, double *tops);

/* One v_mfma_f32_16x16x128_f8f6f4 with FP4 (e2m1) operands, given as each lane's operand registers:
 * a and b are 64 lanes x 16 bytes, d receives 64 lanes x 4 floats (the result registers); scaled
 * selects E8M0 scales 127 (2^0) instead of 0. No counterpart in the reference: it pins the lane
 * layouts the row evaluation's FP4 products assume. */
int lbn_debug_mfma_f4_probe(const uint8_t *a, const uint8_t *b, float *d, int scaled);

/* Measurement, no counterpart in the reference: the device's FP4 MFMA rate in TFLOP/s
 * (v_mfma_f32_16x16x128_f8f6f4, 16 independent chains a wave, all CUs, 0 / 0.5 operands). */
int lbn_debug_mfma_f4_rate(int random_operands, double *tflops);

#ifdef __cplusplus
}
#endif

#endif
