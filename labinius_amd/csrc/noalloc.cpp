// The row evaluation without a prover and without allocating (lbn_row_evaluate_noalloc).
//
// The kernels of lbn_row_evaluate need scratch beyond their inputs and output: the eq tensor's
// 8-coordinate sub-tables, the eq tensor itself and its FP4 operands. Here that scratch is one
// file-scope device array, `scratch`, sized for row_log_len <= ROW_NOALLOC_MAX_ROWS; each device
// the library runs on holds its own copy until the process ends. Its address on a device is looked up once, with that
// device current, by hipGetSymbolAddress, and kept per device; host code uses only that address.
// Every call on one device shares the array, so calls that overlap on one device race (the
// caller orders them, for instance on one stream).
#include <mutex>

#include "internal.hpp"
#include "launch.hpp"

namespace lbn {

constexpr uint32_t ROW_NOALLOC_MAX_ROWS = 15;
constexpr size_t ROW_NOALLOC_PART_WORDS = 3 * 256 * ((ROW_NOALLOC_MAX_ROWS + 7) / 8);
constexpr size_t ROW_NOALLOC_EQ_WORDS = 3 * ((size_t)1 << ROW_NOALLOC_MAX_ROWS);
// the FP4 product's eq operands, u32 (row_eval_scratch_u32 of 2^15 elements), as u64 words
constexpr size_t ROW_NOALLOC_FRAG_WORDS = ((size_t)1 << ROW_NOALLOC_MAX_ROWS) / 128 * 11 * 64 * 4 / 2;
constexpr size_t ROW_NOALLOC_WORDS = ROW_NOALLOC_PART_WORDS + ROW_NOALLOC_EQ_WORDS + ROW_NOALLOC_FRAG_WORDS;  // 3.68 MB

static __device__ __align__(256) uint64_t scratch[ROW_NOALLOC_WORDS];

constexpr int MAX_DEVICES = 64;
static std::once_flag scratch_once[MAX_DEVICES];
static uint64_t *scratch_ptr[MAX_DEVICES];
static hipError_t scratch_err[MAX_DEVICES];

// The device address of `scratch` on the current device, looked up on the first call there.
static int scratch_address(uint64_t **out) {
    int dev = -1;
    HIP_TRY(hipGetDevice(&dev));
    if (dev < 0 || dev >= MAX_DEVICES) return fail(LBN_ERR_NO_DEVICE, "device index out of range");
    std::call_once(scratch_once[dev], [dev] {
        void *p = nullptr;
        scratch_err[dev] = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        scratch_ptr[dev] = scratch_err[dev] == hipSuccess ? (uint64_t *)p : nullptr;
    });
    if (scratch_err[dev] != hipSuccess || !scratch_ptr[dev])
        return fail(LBN_ERR_HIP, std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(scratch_err[dev]));
    *out = scratch_ptr[dev];
    return LBN_OK;
}

}  // namespace lbn

using namespace lbn;

extern "C" {

int lbn_row_evaluate_noalloc(const lbn_f162 *d_w, size_t n, const lbn_f162 *d_p0, size_t p0_len, lbn_f162 *d_u,
                             void *stream) {
    if (!d_w || !d_p0 || !d_u) return fail(LBN_ERR_ARG, "null argument");
    if (p0_len < 7) return fail(LBN_ERR_SHAPE, "a row of fewer than 2^7 elements (Params needs wl - cl >= 7)");
    if (p0_len > ROW_NOALLOC_MAX_ROWS)
        return fail(LBN_ERR_SHAPE, "row_log_len above 15, the largest the static scratch holds");
    const size_t wdim = (size_t)1 << p0_len;
    if (n == 0 || n % wdim != 0) return fail(LBN_ERR_SHAPE, "the witness is not a whole number of rows of 2^len(p0)");
    const size_t cols = n / wdim;
    if (cols > 0xFFFFFFFFull) return fail(LBN_ERR_SHAPE, "too many columns");
    uint64_t *base = nullptr;
    int rc = scratch_address(&base);
    if (rc) return rc;
    uint64_t *part = base, *eq = base + ROW_NOALLOC_PART_WORDS;
    uint32_t *frag = (uint32_t *)(eq + ROW_NOALLOC_EQ_WORDS);
    if (row_eval_scratch_u32((uint32_t)wdim) > 2 * ROW_NOALLOC_FRAG_WORDS)
        return fail(LBN_ERR_SHAPE, "the eq operands do not fit the static scratch");
    hipStream_t st = (hipStream_t)stream;
    HIP_TRY(launch_eq(st, (const uint64_t *)d_p0, (uint32_t)p0_len, part, eq));
    HIP_TRY(launch_fill_u32(st, (uint32_t *)d_u, cols * 6, 0));
    HIP_TRY(launch_row_eval_f4(st, (const uint64_t *)d_w, eq, frag, (uint32_t)wdim, (uint32_t)cols, (uint64_t *)d_u));
    return LBN_OK;
}

}  // extern "C"
