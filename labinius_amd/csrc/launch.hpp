// Host-callable launchers of the HIP kernels (kernels.hip). Each returns the launch status.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

#include "device.hpp"

namespace lbn {

// Same layout as lbn_challenge (include/labinius.h).
struct lbn_challenge_dev {
    uint8_t positions[32];
    uint32_t signs;
    uint32_t weight;
};

hipError_t launch_mfma_probe(hipStream_t s, const int8_t *A, const int8_t *B, int32_t *D);
hipError_t launch_mfma_rate(hipStream_t s, int mode, uint32_t blocks, uint32_t iters, int32_t *sink);
hipError_t launch_mfma_f4_probe(hipStream_t s, const uint32_t *A, const uint32_t *B, float *D, int scaled);
hipError_t launch_mfma_f4_rate(hipStream_t s, int random, uint32_t blocks, uint32_t iters, float *sink);
hipError_t launch_fill_u32(hipStream_t s, uint32_t *p, size_t n, uint32_t v);
// w + c * column_len (cols == NULL) or cols[c] is column c; either output may be NULL
// the byte layouts for K index lengths padded to kpad_c (ring elements) and kpad_f (columns)
hipError_t launch_prepass(hipStream_t s, const uint64_t *w, const uint64_t *const *cols, uint32_t column_len,
                          uint32_t L, uint32_t r, uint8_t *bytes_c, uint32_t kpad_c, uint8_t *bytes_f,
                          uint32_t kpad_f);
// One block per work item; K in ranges of at most gemm_kb_chunk steps (0: KB_EXACT), at least
// g.splits of them (0: no minimum), whose count goes to *splits (may be NULL). With several, y
// (batches x y_batch words) is cleared first and the ranges add into it.
hipError_t launch_tile_gemm(hipStream_t s, const GemmArgs &g, uint32_t *splits);
extern uint32_t gemm_kb_chunk;  // tests only (lbn_debug_gemm_chunk)
// The A' fragments of ntiles tiles (device.hpp; tiles[t].a_off into dst, E at tiles[t].ee_off) from
// their A fragments, src_steps steps a tile, src_tile bytes apart; kpad a power of two >= 64.
hipError_t launch_afold(hipStream_t s, const int8_t *src, uint64_t src_tile, uint32_t src_steps,
                        const TileDesc *tiles, uint32_t ntiles, const uint16_t *ee, uint32_t kpad, int8_t *dst);
hipError_t launch_commit_finish(hipStream_t s, const uint32_t *yacc, uint32_t limbs,
                                const FinishLimb *fl, const uint16_t *slot648, const uint16_t *qcs,
                                uint32_t *raw, int16_t *comp, uint32_t cols);
hipError_t launch_chat(hipStream_t s, const lbn_challenge_dev *ch, uint32_t r, uint32_t rpad,
                       const uint16_t *expo, const uint16_t *pw, uint32_t cond, uint32_t mult,
                       uint32_t q, int8_t *afrag, uint32_t ntiles);
hipError_t launch_intt(hipStream_t s, const uint32_t *vhat, uint32_t L, const InttConsts &ic,
                       const uint32_t *twi, int16_t *v);
hipError_t launch_garner(hipStream_t s, const int16_t *cen, const GarnerConsts &gc, uint32_t count,
                         uint16_t *top, uint16_t *digits);
hipError_t launch_eq(hipStream_t s, const uint64_t *p0, uint32_t rows, uint64_t *part, uint64_t *eq);
// the cross-field switch's passes (B128 as 2 u64, F162 as 3 u64)
hipError_t launch_switch_eq(hipStream_t s, const uint64_t *parts, uint32_t np, uint32_t n, uint64_t *eq);
hipError_t launch_switch_partials(hipStream_t s, const uint64_t *w, const uint64_t *eq, uint32_t n, uint64_t *v);
hipError_t launch_switch_psi(hipStream_t s, const uint64_t *eq, uint32_t n, const uint64_t *psi, uint64_t *a);
hipError_t launch_switch_msg(hipStream_t s, const uint64_t *a, const uint64_t *p, uint32_t half, int top, uint64_t *out);
// round i's fold by r (arrays of 2 half -> half) and round i + 1's message (half >= 2)
hipError_t launch_switch_fold_msg(hipStream_t s, uint64_t *a, const uint64_t *pin, uint64_t *pout, uint32_t half,
                                  const uint64_t r[3], int top, uint64_t *out);
// u_j = sum_i eq_i w[i + wdim j] over F162 (u cleared first) by the FP4-MFMA GF(2) product; frag
// holds row_eval_scratch_u32(wdim) u32 of scratch (the eq operands)
hipError_t launch_row_eval_f4(hipStream_t s, const uint64_t *w, const uint64_t *eq, uint32_t *frag, uint32_t wdim,
                              uint32_t cols, uint64_t *u);
size_t row_eval_scratch_u32(uint32_t wdim);
// t (3 u64, cleared first) ^= sum_j eq1_j u_j over F162, j < n
hipError_t launch_claim(hipStream_t s, const uint64_t *eq1, const uint64_t *u, uint32_t n, uint64_t *t);
// The evaluation's fused launches: both points' eq sub-tables in one launch (p1 may be null); eq(p0)
// with the row evaluation's FP4 operands, clearing z0 (n0 words) and z1 (n1 words); the row values
// from those operands; the claim with eq(p1) built from its sub-tables (rows1 <= 24).
hipError_t launch_eq_parts(hipStream_t s, const uint64_t *p0, uint32_t rows0, uint64_t *part0, const uint64_t *p1,
                           uint32_t rows1, uint64_t *part1);
hipError_t launch_eq_frag(hipStream_t s, const uint64_t *part, uint32_t rows, uint64_t *eq, uint32_t *frag,
                          uint32_t *z0, uint32_t n0, uint32_t *z1, uint32_t n1);
hipError_t launch_row_eval_frag(hipStream_t s, const uint64_t *w, const uint32_t *frag, uint32_t wdim, uint32_t cols,
                                uint64_t *u);
hipError_t launch_claim_parts(hipStream_t s, const uint64_t *part1, uint32_t rows1, const uint64_t *u, uint32_t n,
                              uint64_t *t);
// the device verifier: the folded witness's bit planes (stat: squared norm, largest |v|, added
// to); the fold of a commitment (acc: 4 x limbs x 162 scratch)
hipError_t launch_verify_planes(hipStream_t s, const int16_t *v, uint32_t len_ring, int half, uint32_t nb,
                                uint64_t *w, unsigned long long *stat);
hipError_t launch_fold_comm(hipStream_t s, const lbn_challenge_dev *ch, uint32_t r, const int16_t *comp,
                            const FoldCommConsts &fc, const uint32_t *pw, const uint16_t *p3, long long *acc,
                            int16_t *out);
// wire formats: a bit-packed object into ceil(total_bits / 32) words; the encoder's statistics
// (stat: min, max; hist: nhist u64 counts, cleared first); the one-wave rANS encoder, which stages
// each group's words at fixed places; their scatter into the stream
hipError_t launch_wire_pack(hipStream_t s, const PackDesc &d, const uint16_t *src, uint32_t *out);
hipError_t launch_wire_stats(hipStream_t s, const int16_t *v, uint64_t coefs, int *stat, unsigned long long *hist,
                             uint32_t nhist);
hipError_t launch_rans_encode(hipStream_t s, const int16_t *v, uint64_t coefs, const RansSym *sym, uint32_t nsym,
                              int low, uint32_t raw, uint32_t *staged, uint16_t *escw, uint32_t *stat);
// the staged words' counts (cnt: groups + ceil(groups / 64) u32), their scan (offs: ceil(groups / 64)
// u64) and the stream
hipError_t launch_rans_scatter(hipStream_t s, const uint32_t *staged, const uint16_t *escw, uint32_t *cnt,
                               uint64_t *offs, uint32_t *stat, uint64_t groups, uint16_t *stream);

}  // namespace lbn
