// PublicParameters on the device, the Prover's workspace, and the commit / fold / row-evaluation
// orchestration (scheme/params.rs:320-339, scheme/prover.rs, scheme/witness.rs:94-105).
#include <algorithm>
#include <cstring>

#include "internal.hpp"
#include "launch.hpp"
#include "threads.hpp"

namespace lbn {


// ---------------------------------------------------------------------------------------------
// the tables of one GEMM tile (see device.hpp)
// ---------------------------------------------------------------------------------------------

struct HostTile {
    int rows[TILE];     // A row (commit) or -1
    int out[TILE];      // output slot or -1
    uint32_t expo[TILE];  // fold: exponent of the row's root, 0xFFFF for padding
    uint16_t lut[2][256];
    std::vector<uint32_t> E;  // [TILE][TILE] mod q
    uint32_t par = 0;
};

// The two int8 digits (low byte, high byte) of x mod q, x = x_l + 128 x_h (mod q): for q below
// UNSIGNED_DIGITS_Q the residue in [0, q) split into 0 <= x_l < 128 and 0 <= x_h < 128; above it
// the centred residue split into -64 <= x_l < 64 and |x_h| <= 76 (q <= 19441). The GEMM's
// result is the same either way (it is exact and reduced mod q); non-negative digits toggle fewer
// bits in the MFMA's multipliers than sign-extended ones, which lowers its power draw and so
// raises the clock it holds (int8 MFMA rate on such operands measured 5% higher).
static uint16_t digits(int64_t x, uint32_t q) {
    int xl, xh;
    if (q < UNSIGNED_DIGITS_Q) {
        const int r = (int)reduce_signed(x, q);
        xl = r & 127;
        xh = r >> 7;
    } else {
        const int c = center(reduce_signed(x, q), q);
        xl = ((c + 64) & 127) - 64;
        xh = (c - xl) / 128;
    }
    return (uint16_t)((uint8_t)(int8_t)xl | (uint16_t)(uint8_t)(int8_t)xh << 8);
}

// Splitting limb: sub-ring beta < 8 of level 3 is Z_q[X]/(X^81 - psi^e), so its coefficient n is
// sum_t w[81 t + n] psi^{e t} (the LUT on byte n), and slot u of the sub-ring evaluates it at
// psi^{SLOT_EXP[u]}: E[m][n] = psi^{SLOT_EXP[81 beta + m] n} (params.rs:1-16, scalar.rs:35-84).
static HostTile split_tile(const Prime &P, int beta) {
    const Tables &t = tables();
    HostTile h;
    h.E.assign(TILE * TILE, 0);
    const uint32_t e = t.subring_exp3[beta];
    for (int b = 0; b < 256; b++) {
        uint64_t x = 0;
        for (int k = 0; k < 8; k++)
            if ((b >> k) & 1) x += P.pw[(e * k) % CONDUCTOR];
        h.lut[0][b] = h.lut[1][b] = digits((int64_t)(x % P.q), P.q);
    }
    for (int m = 0; m < TILE; m++) {
        const bool ok = m < TILE_USED;
        h.rows[m] = ok ? 81 * beta + m : -1;
        h.out[m] = ok ? 81 * beta + m : -1;
        h.expo[m] = ok ? t.slot_exp[81 * beta + m] : 0xFFFF;
        if (!ok) continue;
        for (int n = 0; n < TILE_USED; n++)
            h.E[m * TILE + n] = P.pw[((uint64_t)t.slot_exp[81 * beta + m] * n) % CONDUCTOR];
    }
    return h;
}

// The LUTs of a quadratic limb's level-2 sub-ring k = Z_q[X]/(X^162 - psi'^e): its coefficient
// 81 a + n is sum_{t<4} w[162 t + 81 a + n] psi'^{e t}, i.e. bits 2 t + a of byte n.
static void quad_luts(const Prime &P, int k, uint16_t lut[2][256]) {
    const uint32_t e = tables().subring_exp_q2[k];
    for (int a = 0; a < 2; a++)
        for (int b = 0; b < 256; b++) {
            uint64_t x = 0;
            for (int tt = 0; tt < 4; tt++)
                if ((b >> (2 * tt + a)) & 1) x += P.pw[(e * tt) % CONDUCTOR_QUAD];
            lut[a][b] = digits((int64_t)(x % P.q), P.q);
        }
}

// Quadratic limb, commit: sub-ring k, half hh of its 81 leaves, input parity b. Leaf j's row
// 2j + b is sum_{s = b mod 2} x[s] c_j^{(s - b)/2} with c_j = psi'^{QUAD_SLOT_EXP[j]}; the
// leaf product (scalar.rs:254-265) sends A row 2j + rbit times that into output 2j + (rbit ^ b),
// with the extra factor c_j when rbit = b = 1.
static HostTile quad_commit_tile(const Prime &P, int k, int hh, int b) {
    const Tables &t = tables();
    HostTile h;
    h.E.assign(TILE * TILE, 0);
    h.par = (uint32_t)b;
    quad_luts(P, k, h.lut);
    const int start = 81 * k + 41 * hh, count = hh ? 40 : 41;
    for (int m = 0; m < TILE; m++) {
        h.rows[m] = h.out[m] = -1;
        h.expo[m] = 0xFFFF;
        if (m >= 2 * count) continue;
        const int rho = 2 * start + m, j = rho / 2, rbit = rho & 1;
        h.rows[m] = rho;
        h.out[m] = 2 * j + (rbit ^ b);
        for (int n = 0; n < TILE_USED; n++) {
            const int s = 81 * ((b + n) & 1) + n;
            const uint64_t ex = (uint64_t)t.quad_slot_exp[j] * ((s - b) / 2 + (rbit & b));
            h.E[m * TILE + n] = P.pw[ex % CONDUCTOR_QUAD];
        }
    }
    return h;
}

// Quadratic base, fold: sub-ring k, parity b; rows are its 81 leaves (the challenge's leaf image
// is a scalar, fold.rs:338-352), output 2j + b.
static HostTile quad_fold_tile(const Prime &P, int k, int b) {
    const Tables &t = tables();
    HostTile h;
    h.E.assign(TILE * TILE, 0);
    h.par = (uint32_t)b;
    quad_luts(P, k, h.lut);
    for (int m = 0; m < TILE; m++) {
        h.rows[m] = h.out[m] = -1;
        h.expo[m] = 0xFFFF;
        if (m >= TILE_USED) continue;
        const int j = 81 * k + m;
        h.out[m] = 2 * j + b;
        h.expo[m] = t.quad_slot_exp[j];
        for (int n = 0; n < TILE_USED; n++) {
            const int s = 81 * ((b + n) & 1) + n;
            h.E[m * TILE + n] = P.pw[((uint64_t)t.quad_slot_exp[j] * ((s - b) / 2)) % CONDUCTOR_QUAD];
        }
    }
    return h;
}

// E of the rows a tile may use, [m < 82][n < 88] (mod q, zero from n = 81): the GEMM's epilogue
// reads a row eight n at a time. Rows from 82 on are padding in every tile (the builders above use
// m < 81, or m < 82 for a quadratic commit tile's first half).
static void pack_ee(const HostTile &h, uint16_t *dst) {
    for (int m = 0; m < TILE_ROWS; m++)
        for (int n = 0; n < TILE_USED; n++) dst[m * EP + n] = (uint16_t)h.E[m * TILE + n];
}

// The A operand of a commit tile in MFMA lane order: [kb][kh][digit][mb][lane][16 bytes], lane
// holding row 32 mb + (lane & 31) and K = 64 kb + 32 kh + 16 (lane >> 5) + byte.
static void pack_key(const HostTile &h, const int16_t *a, size_t len_ring, uint32_t kbs, uint32_t q, int8_t *dst) {
    for (uint32_t kb = 0; kb < kbs; kb++)
        for (int kh = 0; kh < 2; kh++)
            for (int mb = 0; mb < NBLK; mb++)
                for (int lane = 0; lane < 64; lane++) {
                    const int m = 32 * mb + (lane & 31), rho = h.rows[m];
                    for (int byte = 0; byte < 16; byte++) {
                        const size_t i = (size_t)KSTEP * kb + 32 * kh + 16 * (lane >> 5) + byte;
                        int64_t x = 0;
                        if (rho >= 0 && i < len_ring) x = a[i * N + rho];
                        const uint16_t d = digits(x, q);
                        int8_t *o = dst + (size_t)kb * A_STEP + (size_t)((kh * 2) * NBLK + mb) * FRAG + lane * 16 + byte;
                        o[0] = (int8_t)(uint8_t)(d & 255);
                        o[NBLK * FRAG] = (int8_t)(uint8_t)(d >> 8);
                    }
                }
}

template <typename T>
static int upload(lbn_pp *pp, T **dst, const T *src, size_t count) {
    HIP_TRY(hipMalloc((void **)dst, std::max<size_t>(1, count) * sizeof(T)));
    pp->allocations.push_back(*dst);
    if (count) HIP_TRY(hipMemcpy(*dst, src, count * sizeof(T), hipMemcpyHostToDevice));
    return LBN_OK;
}

static void pp_free(lbn_pp *pp) {
    for (void *p : pp->allocations) (void)hipFree(p);
    pp->allocations.clear();
    delete pp;
}

// Everything the GEMMs read, built on the host.
struct HostTables {
    uint32_t lpad = 0, rpad = 0, kpad_c = 0, kpad_f = 0, kb_c = 0, kb_f = 0;
    std::vector<int8_t> afrag;  // the commit tiles' A fragments, lpad / 64 steps a tile
    std::vector<uint16_t> lut, ee;
    std::vector<int16_t> out;
    std::vector<TileDesc> desc;
    std::vector<FinishLimb> finish;
    std::vector<uint16_t> lut_f, ee_f, expo_f, pw_f;
    std::vector<int16_t> out_f;
    std::vector<TileDesc> desc_f;
    uint32_t cond_f = 0, mult_f = 0;
    InttConsts intt{};
    std::vector<uint32_t> twi;
    std::vector<InttConsts> intt_limb;
    std::vector<std::vector<uint32_t>> twi_limb;
};

// One limb's inverse transform constants and its twiddles (7 levels, concatenated).
static void intt_tables(const Prime &P, InttConsts &ic, std::vector<uint32_t> &twi) {
    ic.q = P.q;
    ic.quad = P.quad;
    ic.w = (uint32_t)P.w;
    ic.w2 = (uint32_t)P.w2;
    ic.z6 = (uint32_t)P.z6;
    ic.inv2 = (uint32_t)P.inv2;
    ic.inv3 = (uint32_t)P.inv3;
    ic.det = (uint32_t)P.det;
    ic.qinv = 1.0f / (float)P.q;
    twi.clear();
    for (int l = 0; l < 7; l++) {
        ic.twi_off[l] = (uint32_t)twi.size();
        twi.insert(twi.end(), P.twi[l].begin(), P.twi[l].end());
    }
    ic.twi_count = (uint32_t)twi.size();  // <= INTT_TWI_MAX, checked where the tables are built
}

static void build_host_tables(const Shape &s, const std::vector<int16_t> &key, HostTables &H) {
    const size_t limbs = s.primes.size();
    H.lpad = (uint32_t)round_up(s.len_ring, KSTEP);
    H.rpad = (uint32_t)round_up(s.r, KSTEP);
    H.kpad_c = gemm_kpad(s.len_ring);
    H.kpad_f = gemm_kpad(s.r);
    H.kb_c = gemm_kbs(H.kpad_c);
    H.kb_f = gemm_kbs(H.kpad_f);

    // commit tiles
    std::vector<HostTile> tiles;
    std::vector<uint32_t> tile_limb;
    for (size_t k = 0; k < limbs; k++) {
        const Prime &P = prime(s.primes[k]);
        if (P.quad) {
            for (int sub = 0; sub < 4; sub++)
                for (int hh = 0; hh < 2; hh++)
                    for (int b = 0; b < 2; b++) {
                        tiles.push_back(quad_commit_tile(P, sub, hh, b));
                        tile_limb.push_back((uint32_t)k);
                    }
        } else {
            for (int beta = 0; beta < 8; beta++) {
                tiles.push_back(split_tile(P, beta));
                tile_limb.push_back((uint32_t)k);
            }
        }
    }
    const size_t a_tile = (size_t)(H.lpad / KSTEP) * A_STEP;  // A; A' has kb_c steps a tile
    const size_t ee_tile = (size_t)EE_TILE;
    H.afrag.assign(a_tile * tiles.size(), 0);
    H.lut.assign(512 * tiles.size(), 0);
    H.ee.assign(ee_tile * tiles.size(), 0);
    H.out.assign(TILE * tiles.size(), 0);
    H.desc.resize(tiles.size());
    for (size_t ti = 0; ti < tiles.size(); ti++) {
        const uint32_t k = tile_limb[ti], q = s.primes[k];
        pack_key(tiles[ti], key.data() + (size_t)k * s.len_ring * N, s.len_ring, H.lpad / KSTEP, q,
                 H.afrag.data() + ti * a_tile);
        std::memcpy(H.lut.data() + 512 * ti, tiles[ti].lut, 1024);
        pack_ee(tiles[ti], H.ee.data() + ee_tile * ti);
        for (int m = 0; m < TILE; m++) H.out[TILE * ti + m] = (int16_t)tiles[ti].out[m];
        H.desc[ti] = TileDesc{ti * H.kb_c * A_STEP, (uint32_t)(512 * ti), (uint32_t)(ee_tile * ti), (uint32_t)(TILE * ti),
                              q, tiles[ti].par, (uint32_t)((YACC_PLANES * k + tiles[ti].par) * N)};
    }

    // the commitment's finish
    const Tables &t = tables();
    H.finish.resize(limbs);
    for (size_t k = 0; k < limbs; k++) {
        const Prime &P = prime(s.primes[k]);
        FinishLimb &f = H.finish[k];
        std::memset(&f, 0, sizeof f);
        f.q = P.q;
        f.quad = P.quad;
        f.inv2 = (uint32_t)P.inv2;
        if (!P.quad) {
            f.ii = P.pw[486];
            const uint64_t inv4 = inv_mod(4, P.q);
            for (int kk = 0; kk < 4; kk++)
                for (int sl = 0; sl < N162; sl++) {
                    const uint64_t v = t.pow3_slot_exp[sl];
                    f.twist[kk][sl] = (uint16_t)(inv4 * P.pw[(CONDUCTOR - (v * kk) % CONDUCTOR) % CONDUCTOR] % P.q);
                }
        } else {
            for (int sl = 0; sl < N162; sl++) {
                const uint64_t v = t.pow3_slot_exp[sl];
                f.twist[0][sl] = (uint16_t)(P.inv2 * P.pw[(CONDUCTOR_QUAD - v) % CONDUCTOR_QUAD] % P.q);
            }
        }
    }

    // fold tiles over the base limb
    const Prime &B = prime(s.primes[0]);
    std::vector<HostTile> ft;
    if (B.quad) {
        for (int sub = 0; sub < 4; sub++)
            for (int b = 0; b < 2; b++) ft.push_back(quad_fold_tile(B, sub, b));
    } else {
        for (int beta = 0; beta < 8; beta++) ft.push_back(split_tile(B, beta));
    }
    const size_t af_tile = (size_t)H.kb_f * A_STEP;  // A' (the challenges' transform has rpad / 64 steps)
    H.lut_f.assign(512 * ft.size(), 0);
    H.ee_f.assign(ee_tile * ft.size(), 0);
    H.expo_f.assign(TILE * ft.size(), 0);
    H.out_f.assign(TILE * ft.size(), 0);
    H.desc_f.resize(ft.size());
    for (size_t ti = 0; ti < ft.size(); ti++) {
        std::memcpy(H.lut_f.data() + 512 * ti, ft[ti].lut, 1024);
        pack_ee(ft[ti], H.ee_f.data() + ee_tile * ti);
        for (int m = 0; m < TILE; m++) {
            H.out_f[TILE * ti + m] = (int16_t)ft[ti].out[m];
            H.expo_f[TILE * ti + m] = (uint16_t)ft[ti].expo[m];
        }
        H.desc_f[ti] = TileDesc{ti * af_tile, (uint32_t)(512 * ti), (uint32_t)(ee_tile * ti),
                                (uint32_t)(TILE * ti), B.q, ft[ti].par, 0};
    }
    H.cond_f = B.quad ? CONDUCTOR_QUAD : CONDUCTOR;
    H.mult_f = B.quad ? 2 : 4;
    H.pw_f.assign(B.pw.begin(), B.pw.end());

    // every limb's inverse transform (the base limb's is the fold's)
    H.intt_limb.resize(limbs);
    H.twi_limb.resize(limbs);
    for (size_t k = 0; k < limbs; k++) intt_tables(prime(s.primes[k]), H.intt_limb[k], H.twi_limb[k]);
    H.intt = H.intt_limb[0];
    H.twi = H.twi_limb[0];
}

// The commit tiles' A' (E folded into A, device.hpp), resident: built on the device from the
// A fragments, which are then freed.
static int build_afrag(lbn_pp *pp, const HostTables &H) {
    const size_t steps = H.lpad / KSTEP;
    int8_t *src = nullptr;
    HIP_TRY(hipMalloc((void **)&src, std::max<size_t>(1, H.afrag.size())));
    hipError_t e = hipMemcpy(src, H.afrag.data(), H.afrag.size(), hipMemcpyHostToDevice);
    if (e == hipSuccess) e = hipMalloc((void **)&pp->d_afrag, (size_t)pp->ntiles * pp->kb_c * A_STEP);
    if (e == hipSuccess) {
        pp->allocations.push_back(pp->d_afrag);
        e = launch_afold(nullptr, src, steps * A_STEP, (uint32_t)steps, pp->d_tiles, pp->ntiles, pp->d_ee, pp->kpad_c,
                         pp->d_afrag);
    }
    if (e == hipSuccess) e = hipStreamSynchronize(nullptr);
    (void)hipFree(src);
    HIP_TRY(e);
    return LBN_OK;
}

static int build_device(lbn_pp *pp) {
    HostTables H;
    build_host_tables(pp->shape, pp->key, H);
    for (const InttConsts &ic : H.intt_limb)
        if (ic.twi_count > INTT_TWI_MAX) return fail(LBN_ERR_HIP, "inverse twiddles exceed the kernel's table");
    pp->lpad = H.lpad;
    pp->rpad = H.rpad;
    pp->kpad_c = H.kpad_c;
    pp->kpad_f = H.kpad_f;
    pp->kb_c = H.kb_c;
    pp->kb_f = H.kb_f;
    pp->ntiles = (uint32_t)H.desc.size();
    pp->ntiles_f = (uint32_t)H.desc_f.size();
    pp->cond_f = H.cond_f;
    pp->mult_f = H.mult_f;
    pp->intt = H.intt;
    const Tables &t = tables();
    int rc;
    if ((rc = upload(pp, &pp->d_lut, H.lut.data(), H.lut.size()))) return rc;
    if ((rc = upload(pp, &pp->d_ee, H.ee.data(), H.ee.size()))) return rc;
    if ((rc = upload(pp, &pp->d_out, H.out.data(), H.out.size()))) return rc;
    if ((rc = upload(pp, &pp->d_tiles, H.desc.data(), H.desc.size()))) return rc;
    if ((rc = build_afrag(pp, H))) return rc;
    if ((rc = upload(pp, &pp->d_finish, H.finish.data(), H.finish.size()))) return rc;
    if ((rc = upload(pp, &pp->d_slot648, &t.slot_648[0][0], 4 * N162))) return rc;
    if ((rc = upload(pp, &pp->d_qcs, &t.quad_class_slot[0][0], 2 * N162))) return rc;
    if ((rc = upload(pp, &pp->d_lut_f, H.lut_f.data(), H.lut_f.size()))) return rc;
    if ((rc = upload(pp, &pp->d_ee_f, H.ee_f.data(), H.ee_f.size()))) return rc;
    if ((rc = upload(pp, &pp->d_out_f, H.out_f.data(), H.out_f.size()))) return rc;
    if ((rc = upload(pp, &pp->d_expo_f, H.expo_f.data(), H.expo_f.size()))) return rc;
    if ((rc = upload(pp, &pp->d_tiles_f, H.desc_f.data(), H.desc_f.size()))) return rc;
    if ((rc = upload(pp, &pp->d_pw_f, H.pw_f.data(), H.pw_f.size()))) return rc;
    if ((rc = upload(pp, &pp->d_twi, H.twi.data(), H.twi.size()))) return rc;
    pp->intt_limb = H.intt_limb;
    pp->d_twi_limb.assign(H.twi_limb.size(), nullptr);
    for (size_t k = 0; k < H.twi_limb.size(); k++)
        if ((rc = upload(pp, &pp->d_twi_limb[k], H.twi_limb[k].data(), H.twi_limb[k].size()))) return rc;
    return LBN_OK;
}

// ---------------------------------------------------------------------------------------------
// the prover
// ---------------------------------------------------------------------------------------------


// ---------------------------------------------------------------------------------------------
// uploads from pageable host memory
// ---------------------------------------------------------------------------------------------
// A large upload (a witness at 2^24 is 403 MB) is split over UP_SHARES threads, each copying its
// chunks into two pinned buffers in turn and sending them on its own stream: the copies to
// pinned memory run on several cores and the DMAs overlap them, where one pageable hipMemcpy
// reaches about 30 GB/s. Smaller uploads take the plain copy.
constexpr size_t UP_CHUNK = size_t(8) << 20, UP_MIN = size_t(32) << 20, UP_SHARES = 4;

static void upload_free(std::vector<lbn_prover::Upload> &up) {
    for (auto &u : up) {
        for (int b = 0; b < 2; b++) {
            if (u.ev[b]) (void)hipEventDestroy(u.ev[b]);
            if (u.buf[b]) (void)hipHostFree(u.buf[b]);
        }
        if (u.st) (void)hipStreamDestroy(u.st);
    }
    up.clear();
}

static bool upload_staging(lbn_prover *pr) {
    if (!pr->up.empty()) return true;
    std::vector<lbn_prover::Upload> up(UP_SHARES);
    bool ok = true;
    for (auto &u : up) {
        ok = ok && hipStreamCreate(&u.st) == hipSuccess;
        for (int b = 0; b < 2 && ok; b++)
            ok = hipHostMalloc((void **)&u.buf[b], UP_CHUNK, hipHostMallocDefault) == hipSuccess &&
                 hipEventCreateWithFlags(&u.ev[b], hipEventDisableTiming) == hipSuccess;
    }
    if (!ok) {
        upload_free(up);
        (void)hipGetLastError();  // the failed allocation is not reported by the next call
        return false;
    }
    pr->up = std::move(up);
    return true;
}

// dst (device) <- src (pageable host memory), complete when this returns
static int upload_host(lbn_prover *pr, void *dst, const void *src, size_t bytes) {
    int dev = 0;
    if (bytes < UP_MIN || !upload_staging(pr) || hipGetDevice(&dev) != hipSuccess)
        return hipMemcpy(dst, src, bytes, hipMemcpyHostToDevice) == hipSuccess ? LBN_OK
                                                                              : fail(LBN_ERR_HIP, "witness upload");
    const size_t chunks = (bytes + UP_CHUNK - 1) / UP_CHUNK;
    std::atomic<bool> ok{true};
    const bool ran = run_shares(UP_SHARES, [&](size_t k) {
        lbn_prover::Upload &u = pr->up[k];
        if (hipSetDevice(dev) != hipSuccess) {
            ok = false;
            return;
        }
        const size_t c0 = chunks * k / UP_SHARES, c1 = chunks * (k + 1) / UP_SHARES;
        for (size_t c = c0; c < c1 && ok; c++) {
            const int b = (int)((c - c0) & 1);
            if (c - c0 >= 2 && hipEventSynchronize(u.ev[b]) != hipSuccess) ok = false;  // the buffer's last DMA
            const size_t off = c * UP_CHUNK, n = std::min(UP_CHUNK, bytes - off);
            if (!ok) break;
            std::memcpy(u.buf[b], (const uint8_t *)src + off, n);
            if (hipMemcpyAsync((uint8_t *)dst + off, u.buf[b], n, hipMemcpyHostToDevice, u.st) != hipSuccess ||
                hipEventRecord(u.ev[b], u.st) != hipSuccess)
                ok = false;
        }
        if (hipStreamSynchronize(u.st) != hipSuccess) ok = false;
    });
    return ran && ok ? LBN_OK : fail(LBN_ERR_HIP, "witness upload");
}

// Outputs copied back to the caller's host memory: sent into one pinned buffer of the prover,
// then copied out on several threads once the stream is done. A copy to pageable memory goes
// through the runtime's bounce buffer and touches the destination's pages on one thread
// (2.5 ms for the 17 MB of a round's outputs at 2^24). Under DOWN_MIN bytes, or when no pinned
// memory can be had, the outputs take the plain copy.
struct HostOut {
    void *dst;
    const void *src;
    size_t n;
};
constexpr size_t DOWN_MIN = size_t(1) << 20, DOWN_SLICE = size_t(2) << 20;

static hipError_t download_host(lbn_prover *pr, hipStream_t st, const HostOut *o, size_t count) {
    size_t total = 0;
    for (size_t i = 0; i < count; i++) total += (o[i].n + 255) / 256 * 256;
    bool staged = total >= DOWN_MIN;
    if (staged && pr->h_out_cap < total) {
        if (pr->h_out) (void)hipHostFree(pr->h_out);
        pr->h_out = nullptr;
        pr->h_out_cap = 0;
        if (hipHostMalloc((void **)&pr->h_out, total, hipHostMallocDefault) == hipSuccess)
            pr->h_out_cap = total;
        else {
            pr->h_out = nullptr;
            (void)hipGetLastError();  // the failed allocation is not reported by the next call
            staged = false;
        }
    }
    size_t off = 0;
    for (size_t i = 0; i < count; i++) {
        hipError_t e = hipMemcpyAsync(staged ? pr->h_out + off : o[i].dst, o[i].src, o[i].n, hipMemcpyDeviceToHost, st);
        if (e != hipSuccess) return e;
        off += (o[i].n + 255) / 256 * 256;
    }
    const hipError_t e = stream_wait(st);
    if (e != hipSuccess || !staged) return e;
    off = 0;
    for (size_t i = 0; i < count; i++) {
        const uint8_t *from = pr->h_out + off;
        uint8_t *to = (uint8_t *)o[i].dst;
        const size_t nt = std::min<size_t>(4, (o[i].n + DOWN_SLICE - 1) / DOWN_SLICE);
        (void)run_ranges(o[i].n, nt, [&](size_t lo, size_t hi) { std::memcpy(to + lo, from + lo, hi - lo); });
        off += (o[i].n + 255) / 256 * 256;
    }
    return hipSuccess;
}

static void prover_free(lbn_prover *pr) {
    void *bufs[] = {pr->d_bytes_c, pr->d_yacc, pr->d_raw, pr->d_comp, pr->d_ch, pr->d_chat, pr->d_afrag_f, pr->d_vhat,
                    pr->d_v, pr->d_p0, pr->d_eq, pr->d_eqpart, pr->d_eqf, pr->d_u, pr->d_wscratch,
                    pr->d_eqpart1, pr->d_t,
                    pr->d_bd_raw, pr->d_bd_cen, pr->d_bd_out, pr->sw.d_eq, pr->sw.d_a, pr->sw.d_p,
                    pr->sw.d_small};
    for (void *b : bufs)
        if (b) (void)hipFree(b);
    if (pr->h_pts) (void)hipHostFree(pr->h_pts);
    if (pr->h_ch) (void)hipHostFree(pr->h_ch);
    if (pr->h_out) (void)hipHostFree(pr->h_out);
    if (pr->sw.h_small) (void)hipHostFree(pr->sw.h_small);
    for (uint8_t *b : pr->free_bytes_f) (void)hipFree(b);  // a pooled prover keeps none
    batch_free(pr->batch);
    wire_free(pr->wire);
    verify_free(pr->vfy);
    for (hipEvent_t &e : pr->ev)
        if (e) (void)hipEventDestroy(e);
    upload_free(pr->up);
    if (pr->stream) (void)hipStreamDestroy(pr->stream);
    if (pr->pp && --pr->pp->refs == 0) pp_free(pr->pp);
    delete pr;
}

static size_t bytes_f_size(const lbn_pp *pp) { return bytes_size((uint32_t)pp->shape.len_ring, pp->kpad_f); }

// The fold's layout buffer: recycled by the prover, or from its pool. Padding columns are never
// written by the layout pass and must read as zero: a fresh buffer, or a pooled one that last held
// another layout (another ring element or column count), is cleared.
static int take_bytes_f(lbn_prover *pr, uint8_t **out) {
    const size_t sz = bytes_f_size(pr->pp);
    if (pr->pool) {
        const uint64_t tag = ((uint64_t)pr->pp->shape.r << 32) | pr->pp->shape.len_ring;
        bool same = false;
        void *p = nullptr;
        int rc = pool_alloc(pr->pool, sz, tag, &p, &same);
        if (rc) return rc;
        *out = (uint8_t *)p;
        if (!same) {
            const hipError_t e = hipMemsetAsync(*out, 0, sz, pr->stream);
            if (e != hipSuccess) {
                pool_free(pr->pool, p);
                return fail(LBN_ERR_HIP, std::string("layout clear: ") + hipGetErrorString(e));
            }
        }
        return LBN_OK;
    }
    if (!pr->free_bytes_f.empty()) {
        *out = pr->free_bytes_f.back();
        pr->free_bytes_f.pop_back();
        return LBN_OK;
    }
    HIP_TRY(hipMalloc((void **)out, sz));
    const hipError_t e = hipMemset(*out, 0, sz);
    if (e != hipSuccess) {
        (void)hipFree(*out);
        return fail(LBN_ERR_HIP, std::string("layout clear: ") + hipGetErrorString(e));
    }
    return LBN_OK;
}

// A layout buffer back to where take_bytes_f found it.
static void give_bytes_f(lbn_prover *pr, uint8_t *b) {
    if (!b) return;
    if (pr->pool)
        pool_free(pr->pool, b);
    else
        pr->free_bytes_f.push_back(b);
}

// Returns a taken layout buffer on every early return, unless the opening took it.
struct BytesGuard {
    lbn_prover *pr;
    uint8_t *b;
    ~BytesGuard() { give_bytes_f(pr, b); }
};

// The commit GEMM over ncol columns of a layout (bytes) into y ([ncol][limb][plane][648]).
GemmArgs commit_gemm_args(const lbn_pp *pp, const uint8_t *bytes, uint32_t *y, size_t ncol) {
    GemmArgs g{};
    g.tiles = pp->d_tiles;
    g.afrag = pp->d_afrag;
    g.lut = pp->d_lut;
    g.out = pp->d_out;
    g.bytes = bytes;
    g.kbs = pp->kb_c;
    g.run_shift = (uint32_t)__builtin_ctz(pp->kpad_c / KSTEP);
    g.a_kb_stride = A_STEP;
    g.y = y;
    g.y_batch = YACC_PLANES * pp->shape.primes.size() * N;
    g.batches = (uint32_t)ncol;
    g.groups = (uint32_t)((ncol + GROUP_COLS - 1) / GROUP_COLS);
    g.ntiles = pp->ntiles;
    g.sched_t = 8;
    return g;
}

static int commit_device(lbn_prover *pr, const uint64_t *d_w, int16_t *comp, uint32_t *raw, lbn_opening **opening) {
    lbn_pp *pp = pr->pp;
    const Shape &s = pp->shape;
    const size_t limbs = s.primes.size();
    uint8_t *bytes_f = nullptr;  // the fold's layout, only when an opening is asked for
    if (opening) {
        int rc = take_bytes_f(pr, &bytes_f);
        if (rc) return rc;
    }
    BytesGuard guard{pr, bytes_f};
    hipStream_t st = pr->stream;
    HIP_TRY(hipEventRecord(pr->ev[0], st));
    HIP_TRY(launch_prepass(st, d_w, nullptr, (uint32_t)s.column_len, (uint32_t)s.len_ring, (uint32_t)s.r,
                           pr->d_bytes_c, pp->kpad_c, bytes_f, pp->kpad_f));
    HIP_TRY(hipEventRecord(pr->ev[1], st));
    GemmArgs g = commit_gemm_args(pp, pr->d_bytes_c, pr->d_yacc, s.r);
    HIP_TRY(launch_tile_gemm(st, g, nullptr));
    HIP_TRY(hipEventRecord(pr->ev[2], st));
    HIP_TRY(launch_commit_finish(st, pr->d_yacc, (uint32_t)limbs, pp->d_finish, pp->d_slot648, pp->d_qcs, pr->d_raw,
                                 pr->d_comp, (uint32_t)s.r));
    HIP_TRY(hipEventRecord(pr->ev[3], st));
    HostOut outs[2];
    size_t nout = 0;
    if (comp) outs[nout++] = {comp, pr->d_comp, s.r * 4 * limbs * N162 * sizeof(int16_t)};
    if (raw) outs[nout++] = {raw, pr->d_raw, limbs * s.r * N * sizeof(uint32_t)};
    HIP_TRY(download_host(pr, st, outs, nout));
    pr->committed = true;
    pr->bd_ready = false;
    pr->timings[0] = elapsed(pr->ev[0], pr->ev[3]);
    pr->timings[1] = elapsed(pr->ev[0], pr->ev[1]);
    pr->timings[2] = elapsed(pr->ev[1], pr->ev[2]);
    pr->timings[3] = elapsed(pr->ev[2], pr->ev[3]);
    if (opening) {
        lbn_opening *op = new lbn_opening();
        op->owner = pr;
        op->d_bytes_f = bytes_f;
        *opening = op;
        guard.b = nullptr;  // the opening holds it now
    }
    return LBN_OK;
}

int commit_columns_device(lbn_prover *pr, const uint64_t *d_w, size_t ncol, uint8_t *d_bytes, uint32_t *d_yacc,
                          uint32_t *d_raw, int16_t *d_comp) {
    lbn_pp *pp = pr->pp;
    const Shape &s = pp->shape;
    const size_t limbs = s.primes.size();
    hipStream_t st = pr->stream;
    HIP_TRY(launch_prepass(st, d_w, nullptr, (uint32_t)s.column_len, (uint32_t)s.len_ring, (uint32_t)ncol, d_bytes,
                           pp->kpad_c, nullptr, 0));
    GemmArgs g = commit_gemm_args(pp, d_bytes, d_yacc, ncol);
    HIP_TRY(launch_tile_gemm(st, g, nullptr));
    HIP_TRY(launch_commit_finish(st, d_yacc, (uint32_t)limbs, pp->d_finish, pp->d_slot648, pp->d_qcs, d_raw, d_comp,
                                 (uint32_t)ncol));
    return LBN_OK;
}

static int fold_device(lbn_prover *pr, lbn_opening *op, const lbn_challenge *ch, size_t r, int16_t *v) {
    lbn_pp *pp = pr->pp;
    const Shape &s = pp->shape;
    hipStream_t st = pr->stream;
    std::memcpy(pr->h_ch, ch, r * sizeof(lbn_challenge));  // pinned staging: the copy is asynchronous
    HIP_TRY(hipMemcpyAsync(pr->d_ch, pr->h_ch, r * sizeof(lbn_challenge), hipMemcpyHostToDevice, st));
    HIP_TRY(hipEventRecord(pr->ev[4], st));
    HIP_TRY(launch_chat(st, (const lbn_challenge_dev *)pr->d_ch, (uint32_t)r, pp->rpad, pp->d_expo_f, pp->d_pw_f,
                        pp->cond_f, pp->mult_f, s.primes[0], pr->d_chat, pp->ntiles_f));
    HIP_TRY(launch_afold(st, pr->d_chat, (size_t)(pp->rpad / KSTEP) * A_STEP, pp->rpad / KSTEP, pp->d_tiles_f,
                         pp->ntiles_f, pp->d_ee_f, pp->kpad_f, pr->d_afrag_f));
    HIP_TRY(hipEventRecord(pr->ev[5], st));
    GemmArgs g{};
    g.tiles = pp->d_tiles_f;
    g.afrag = pr->d_afrag_f;
    g.lut = pp->d_lut_f;
    g.out = pp->d_out_f;
    g.bytes = op->d_bytes_f;
    g.kbs = pp->kb_f;
    g.run_shift = (uint32_t)__builtin_ctz(pp->kpad_f / KSTEP);
    g.a_kb_stride = A_STEP;
    g.y = pr->d_vhat;
    g.y_batch = N;
    g.batches = (uint32_t)s.len_ring;
    g.groups = (uint32_t)((s.len_ring + GROUP_COLS - 1) / GROUP_COLS);
    g.ntiles = pp->ntiles_f;
    g.sched_t = 8;
    HIP_TRY(launch_tile_gemm(st, g, nullptr));
    HIP_TRY(hipEventRecord(pr->ev[6], st));
    HIP_TRY(launch_intt(st, pr->d_vhat, (uint32_t)s.len_ring, pp->intt, pp->d_twi, pr->d_v));
    HIP_TRY(hipEventRecord(pr->ev[7], st));
    const HostOut out = {v, pr->d_v, s.len_ring * N * sizeof(int16_t)};
    HIP_TRY(download_host(pr, st, &out, v ? 1 : 0));
    pr->folded = true;
    pr->timings[4] = elapsed(pr->ev[4], pr->ev[7]);
    pr->timings[5] = elapsed(pr->ev[4], pr->ev[5]);
    pr->timings[6] = elapsed(pr->ev[5], pr->ev[6]);
    pr->timings[7] = elapsed(pr->ev[6], pr->ev[7]);
    return LBN_OK;
}

// eq(p0) and the row values u into pr->d_u, on the prover's stream (events 8..10)
// The row values (and, with p1, the claim's inputs): p0 and p1 in one upload from pinned staging,
// both points' eq sub-tables in one launch, then eq(p0) with the FP4 operands (clearing u and t),
// then the row values.
static int row_eval_launch(lbn_prover *pr, const uint64_t *d_w, const lbn_f162 *p0, const lbn_f162 *p1) {
    lbn_pp *pp = pr->pp;
    const Shape &s = pp->shape;
    hipStream_t st = pr->stream;
    std::memcpy(pr->h_pts, p0, s.rows * sizeof(lbn_f162));
    if (p1) std::memcpy(pr->h_pts + 3 * (size_t)s.rows, p1, s.cl * sizeof(lbn_f162));
    HIP_TRY(hipMemcpyAsync(pr->d_p0, pr->h_pts, (s.rows + (p1 ? s.cl : 0)) * sizeof(lbn_f162), hipMemcpyHostToDevice,
                           st));
    HIP_TRY(hipEventRecord(pr->ev[8], st));
    HIP_TRY(launch_eq_parts(st, pr->d_p0, s.rows, pr->d_eqpart, p1 ? pr->d_p1 : nullptr, s.cl, pr->d_eqpart1));
    HIP_TRY(hipEventRecord(pr->ev[9], st));
    HIP_TRY(launch_eq_frag(st, pr->d_eqpart, s.rows, pr->d_eq, pr->d_eqf, (uint32_t *)pr->d_u, (uint32_t)(s.r * 6),
                           (uint32_t *)pr->d_t, p1 ? 6u : 0u));
    HIP_TRY(launch_row_eval_frag(st, d_w, pr->d_eqf, (uint32_t)s.column_len, (uint32_t)s.r, pr->d_u));
    return LBN_OK;
}

static void row_eval_timings(lbn_prover *pr) {
    pr->timings[8] = elapsed(pr->ev[8], pr->ev[10]);
    pr->timings[9] = elapsed(pr->ev[8], pr->ev[9]);
    pr->timings[10] = elapsed(pr->ev[9], pr->ev[10]);
}

static int row_eval_device(lbn_prover *pr, const uint64_t *d_w, const lbn_f162 *p0, lbn_f162 *u) {
    int rc = row_eval_launch(pr, d_w, p0, nullptr);
    if (rc) return rc;
    hipStream_t st = pr->stream;
    HIP_TRY(hipEventRecord(pr->ev[10], st));
    HIP_TRY(hipMemcpyAsync(u, pr->d_u, pr->pp->shape.r * sizeof(lbn_f162), hipMemcpyDeviceToHost, st));
    HIP_TRY(stream_wait(st));
    row_eval_timings(pr);
    return LBN_OK;
}

// The row values, then the claim t = sum_j eq(p1)_j u_j (eval.rs:185-193) on the device: only t
// comes back.
static int mle_eval_device(lbn_prover *pr, const uint64_t *d_w, const lbn_f162 *p0, const lbn_f162 *p1,
                           lbn_f162 *t) {
    const Shape &s = pr->pp->shape;
    int rc = row_eval_launch(pr, d_w, p0, p1);
    if (rc) return rc;
    hipStream_t st = pr->stream;
    HIP_TRY(launch_claim_parts(st, pr->d_eqpart1, s.cl, pr->d_u, (uint32_t)s.r, pr->d_t));
    HIP_TRY(hipEventRecord(pr->ev[10], st));
    HIP_TRY(hipMemcpyAsync(t, pr->d_t, sizeof(lbn_f162), hipMemcpyDeviceToHost, st));
    HIP_TRY(stream_wait(st));
    row_eval_timings(pr);
    return LBN_OK;
}

int witness_alloc_upload(lbn_prover *pr, const lbn_f162 *w, size_t n, size_t cols, lbn_dwitness **out) {
    lbn_dwitness *d = new lbn_dwitness();
    d->owner = pr;
    d->n = n;
    d->cols = cols;
    const size_t bytes = n * sizeof(lbn_f162);
    if (pr->pool) {
        bool same = false;
        void *p = nullptr;
        const int rc = pool_alloc(pr->pool, bytes, 0, &p, &same);
        if (rc) {
            delete d;
            return rc;
        }
        d->pool = pr->pool;
        d->d_w = (uint64_t *)p;
    } else if (hipMalloc((void **)&d->d_w, bytes) != hipSuccess) {
        delete d;
        return fail(LBN_ERR_OOM, "witness allocation");
    }
    const int rc = upload_host(pr, d->d_w, w, bytes);
    if (rc) {
        lbn_dwitness_release(d);
        return rc;
    }
    *out = d;
    return LBN_OK;
}

}  // namespace lbn

using namespace lbn;

extern "C" {

int lbn_pp_create(const lbn_params *params, const uint8_t matrix_seed[32], lbn_pp **out) {
    if (!out || !matrix_seed) return fail(LBN_ERR_ARG, "null argument");
    *out = nullptr;
    Shape s;
    int rc = make_shape(params, &s);
    if (rc) return rc;
    if (s.opening == LBN_OPENING_RECURSIVE)
        return fail(LBN_ERR_UNSUPPORTED, "Opening::Recursive (LaBRADOR) is outside this prover");
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev == 0)
        return fail(LBN_ERR_NO_DEVICE, "no HIP device: the prover runs on the GPU only");
    lbn_pp *pp = new lbn_pp();
    pp->params = *params;
    pp->shape = s;
    std::memcpy(pp->seed, matrix_seed, 32);
    pp->key = expand_key(s.primes, s.len_ring, matrix_seed);
    (void)hipGetDevice(&pp->device);
    rc = build_device(pp);
    if (rc) {
        pp_free(pp);
        return rc;
    }
    *out = pp;
    return LBN_OK;
}

int lbn_pp_create_host(const lbn_params *params, const uint8_t matrix_seed[32], lbn_pp **out) {
    if (!out || !matrix_seed) return fail(LBN_ERR_ARG, "null argument");
    *out = nullptr;
    Shape s;
    int rc = make_shape(params, &s);
    if (rc) return rc;
    if (s.opening == LBN_OPENING_RECURSIVE)
        return fail(LBN_ERR_UNSUPPORTED, "Opening::Recursive (LaBRADOR) is outside this prover");
    lbn_pp *pp = new lbn_pp();
    pp->params = *params;
    pp->shape = s;
    std::memcpy(pp->seed, matrix_seed, 32);
    pp->key = expand_key(s.primes, s.len_ring, matrix_seed);
    pp->device = -1;
    pp->host_only = true;
    *out = pp;
    return LBN_OK;
}

void lbn_pp_release(lbn_pp *pp) {
    if (pp && --pp->refs == 0) pp_free(pp);
}

int lbn_pp_key(const lbn_pp *pp, int16_t *a, size_t len) {
    if (!pp || !a) return fail(LBN_ERR_ARG, "null argument");
    if (len != pp->key.size()) return fail(LBN_ERR_SHAPE, "wrong key buffer length");
    std::memcpy(a, pp->key.data(), len * sizeof(int16_t));
    return LBN_OK;
}

int lbn_prover_create(lbn_pp *pp, lbn_prover **out) {
    if (!pp || !out) return fail(LBN_ERR_ARG, "null argument");
    *out = nullptr;
    if (pp->host_only) return fail(LBN_ERR_NO_DEVICE, "public parameters created without a device");
    (void)hipSetDevice(pp->device);
    lbn_prover *pr = new lbn_prover();
    pr->pp = pp;
    pp->refs++;
    const Shape &s = pp->shape;
    const size_t limbs = s.primes.size();
    int rc = LBN_OK;
#define TRY_RC(x)              \
    do {                       \
        rc = (x);              \
        if (rc) {              \
            prover_free(pr);   \
            return rc;         \
        }                      \
    } while (0)
    TRY_RC(hipStreamCreateWithFlags(&pr->stream, hipStreamNonBlocking) == hipSuccess ? LBN_OK
                                                                                   : fail(LBN_ERR_HIP, "stream"));
    for (hipEvent_t &e : pr->ev)
        TRY_RC(hipEventCreate(&e) == hipSuccess ? LBN_OK : fail(LBN_ERR_HIP, "event"));
    const size_t bc = bytes_size((uint32_t)s.r, pp->kpad_c);
    TRY_RC(dalloc(&pr->d_bytes_c, bc));
    TRY_RC(hipMemset(pr->d_bytes_c, 0, bc) == hipSuccess ? LBN_OK : fail(LBN_ERR_HIP, "memset"));
    TRY_RC(dalloc(&pr->d_yacc, s.r * YACC_PLANES * limbs * N));
    TRY_RC(dalloc(&pr->d_raw, s.r * limbs * N));
    TRY_RC(dalloc(&pr->d_comp, s.r * 4 * limbs * N162));
    TRY_RC(dalloc(&pr->d_ch, s.r));
    TRY_RC(dalloc(&pr->d_chat, (size_t)pp->ntiles_f * (pp->rpad / KSTEP) * A_STEP));
    TRY_RC(dalloc(&pr->d_afrag_f, (size_t)pp->ntiles_f * pp->kb_f * A_STEP));
    TRY_RC(dalloc(&pr->d_vhat, s.len_ring * N));
    TRY_RC(dalloc(&pr->d_v, s.len_ring * N));
    TRY_RC(dalloc(&pr->d_p0, 3 * ((size_t)s.rows + s.cl)));
    pr->d_p1 = pr->d_p0 + 3 * (size_t)s.rows;
    TRY_RC(hipHostMalloc((void **)&pr->h_pts, 3 * ((size_t)s.rows + s.cl) * sizeof(uint64_t), hipHostMallocDefault) ==
                   hipSuccess
               ? LBN_OK
               : fail(LBN_ERR_OOM, "pinned staging"));
    TRY_RC(hipHostMalloc((void **)&pr->h_ch, s.r * sizeof(lbn_challenge), hipHostMallocDefault) == hipSuccess
               ? LBN_OK
               : fail(LBN_ERR_OOM, "pinned staging"));
    TRY_RC(dalloc(&pr->d_eq, 3 * s.column_len));
    TRY_RC(dalloc(&pr->d_eqpart, 3 * 256 * (size_t)((s.rows + 7) / 8)));
    TRY_RC(dalloc(&pr->d_eqf, row_eval_scratch_u32((uint32_t)s.column_len)));
    TRY_RC(dalloc(&pr->d_eqpart1, 3 * 256 * (size_t)((s.cl + 7) / 8)));
    TRY_RC(dalloc(&pr->d_t, 3));
    TRY_RC(dalloc(&pr->d_u, 3 * s.r));

    // warm-up (scheme/prover.rs:39-49): a zero witness through every kernel once
    {
        std::vector<lbn_f162> zero(s.n);
        std::memset(zero.data(), 0, zero.size() * sizeof(lbn_f162));
        lbn_dwitness *dw = nullptr;
        TRY_RC(lbn_witness_upload(pr, zero.data(), zero.size(), &dw));
        lbn_opening *op = nullptr;
        rc = lbn_commit_resident(pr, dw, nullptr, nullptr, &op);
        if (!rc && s.dropped_bits > 0) {
            std::vector<uint16_t> top(s.r * N), digits((s.primes.size() - 1) * s.r * N);
            rc = lbn_commit_drop_bits(pr, top.data(), digits.data());
        }
        pr->committed = false;  // the warm-up's commitment is not the caller's
        std::vector<lbn_f162> p0(s.rows), u(s.r);
        std::memset(p0.data(), 0, p0.size() * sizeof(lbn_f162));
        if (!rc) rc = lbn_row_evaluate(pr, dw, p0.data(), p0.size(), u.data());
        std::vector<lbn_challenge> ch(s.r);
        lbn_transcript *t = lbn_transcript_new((const uint8_t *)"labinius/warm-up", 16);
        for (size_t j = 0; j < s.r; j++) sample_short_challenge(t->t, 28, 12.0, &ch[j]);
        lbn_transcript_free(t);
        if (!rc) rc = lbn_fold(pr, op, ch.data(), ch.size(), nullptr);
        else if (op) lbn_recycle(pr, op);
        if (!rc) {  // the wire encoder's kernels (a fold of zeros)
            size_t len = 0;
            rc = lbn_fold_encode(pr, (uint16_t)s.primes[0], nullptr, 0, &len);
        }
        pr->committed = false;
        pr->bd_ready = false;
        pr->folded = false;
        lbn_dwitness_release(dw);
        if (rc) {
            prover_free(pr);
            return rc;
        }
    }
#undef TRY_RC
    *out = pr;
    return LBN_OK;
}

void lbn_prover_release(lbn_prover *pr) {
    if (pr) prover_free(pr);
}

int lbn_witness_upload(lbn_prover *pr, const lbn_f162 *w, size_t n, lbn_dwitness **out) {
    if (!pr || !w || !out) return fail(LBN_ERR_ARG, "null argument");
    if (n != pr->pp->shape.n) return fail(LBN_ERR_WITNESS_LENGTH, "the witness is not witness_len() elements");
    return witness_alloc_upload(pr, w, n, pr->pp->shape.r, out);
}

void *lbn_dwitness_device_ptr(lbn_dwitness *w) { return w ? (void *)w->d_w : nullptr; }

void lbn_dwitness_release(lbn_dwitness *w) {
    if (!w) return;
    if (w->pool)
        pool_free(w->pool, w->d_w);
    else if (w->d_w)
        (void)hipFree(w->d_w);
    delete w;
}

int lbn_commit_resident(lbn_prover *pr, const lbn_dwitness *w, int16_t *comp, uint32_t *raw, lbn_opening **opening) {
    if (!pr || !w) return fail(LBN_ERR_ARG, "null argument");
    if (w->n != pr->pp->shape.n)
        return fail(LBN_ERR_SHAPE, "the witness was built for other parameters");
    if (opening) *opening = nullptr;
    return commit_device(pr, w->d_w, comp, raw, opening);
}

int lbn_commit(lbn_prover *pr, const lbn_f162 *w, size_t n, int16_t *comp, uint32_t *raw, lbn_opening **opening) {
    if (!pr || !w) return fail(LBN_ERR_ARG, "null argument");
    if (n != pr->pp->shape.n) return fail(LBN_ERR_SHAPE, "the witness was built for other parameters");
    if (!pr->d_wscratch) {
        if (hipMalloc((void **)&pr->d_wscratch, n * sizeof(lbn_f162)) != hipSuccess)
            return fail(LBN_ERR_OOM, "witness scratch");
    }
    HIP_TRY(hipStreamSynchronize(pr->stream));  // nothing of an earlier call still reads the scratch
    const int rc = upload_host(pr, pr->d_wscratch, w, n * sizeof(lbn_f162));
    if (rc) return rc;
    if (opening) *opening = nullptr;
    return commit_device(pr, pr->d_wscratch, comp, raw, opening);
}

int lbn_fold(lbn_prover *pr, lbn_opening *op, const lbn_challenge *ch, size_t r, int16_t *v) {
    if (!pr || !op || !ch) return fail(LBN_ERR_ARG, "null argument");
    if (op->owner != pr) return fail(LBN_ERR_SHAPE, "the opening belongs to another prover");
    if (r != pr->pp->shape.r) return fail(LBN_ERR_SHAPE, "one challenge per chunk");
    for (size_t j = 0; j < r; j++)
        if (ch[j].weight > 32) return fail(LBN_ERR_ARG, "challenge weight above 32");
    int rc = fold_device(pr, op, ch, r, v);
    give_bytes_f(pr, op->d_bytes_f);
    delete op;
    return rc;
}

int lbn_fold_resident(lbn_prover *pr, const lbn_dwitness *w, const lbn_challenge *ch, size_t r, int16_t *v) {
    if (!pr || !w || !ch) return fail(LBN_ERR_ARG, "null argument");
    const Shape &s = pr->pp->shape;
    if (w->n != s.n) return fail(LBN_ERR_SHAPE, "the witness was built for other parameters");
    if (r != s.r) return fail(LBN_ERR_SHAPE, "one challenge per column");
    for (size_t j = 0; j < r; j++)
        if (ch[j].weight > 32) return fail(LBN_ERR_ARG, "challenge weight above 32");
    lbn_opening op;
    op.owner = pr;
    int rc = take_bytes_f(pr, &op.d_bytes_f);
    if (rc) return rc;
    const hipError_t e = launch_prepass(pr->stream, w->d_w, nullptr, (uint32_t)s.column_len, (uint32_t)s.len_ring,
                                        (uint32_t)s.r, nullptr, 0, op.d_bytes_f, pr->pp->kpad_f);
    rc = e == hipSuccess ? fold_device(pr, &op, ch, r, v) : fail(LBN_ERR_HIP, hipGetErrorString(e));
    give_bytes_f(pr, op.d_bytes_f);
    return rc;
}

int lbn_recycle(lbn_prover *pr, lbn_opening *op) {
    if (!pr || !op) return fail(LBN_ERR_ARG, "null argument");
    if (op->owner != pr) return fail(LBN_ERR_SHAPE, "the opening belongs to another prover");
    give_bytes_f(pr, op->d_bytes_f);
    delete op;
    return LBN_OK;
}

int lbn_row_evaluate(lbn_prover *pr, const lbn_dwitness *w, const lbn_f162 *p0, size_t p0_len, lbn_f162 *u) {
    if (!pr || !w || !p0 || !u) return fail(LBN_ERR_ARG, "null argument");
    if (w->n != pr->pp->shape.n || p0_len != pr->pp->shape.rows)
        return fail(LBN_ERR_SHAPE, "the evaluation point does not match the witness");
    return row_eval_device(pr, w->d_w, p0, u);
}

int lbn_claim(const lbn_f162 *u, size_t r, const lbn_f162 *p1, size_t p1_len, lbn_f162 *t) {
    if (!u || !p1 || !t) return fail(LBN_ERR_ARG, "null argument");
    if (r != ((size_t)1 << p1_len)) return fail(LBN_ERR_SHAPE, "2^len(p1) row values");
    std::vector<F162> ps(p1_len);
    for (size_t k = 0; k < p1_len; k++) std::memcpy(ps[k].w, p1[k].w, 24);
    const std::vector<F162> eq = eq_table(ps.data(), (int)p1_len);
    F162 acc{{0, 0, 0}};
    for (size_t j = 0; j < r; j++) {
        F162 x;
        std::memcpy(x.w, u[j].w, 24);
        acc = f162_add(acc, f162_mul(eq[j], x));
    }
    std::memcpy(t->w, acc.w, 24);
    return LBN_OK;
}

int lbn_mle_evaluate(lbn_prover *pr, const lbn_dwitness *w, const lbn_f162 *p0, size_t p0_len, const lbn_f162 *p1,
                     size_t p1_len, lbn_f162 *t) {
    if (!pr || !w || !p0 || !p1 || !t) return fail(LBN_ERR_ARG, "null argument");
    const Shape &s = pr->pp->shape;
    if (w->n != s.n || p0_len != s.rows || p1_len != s.cl)
        return fail(LBN_ERR_SHAPE, "the evaluation point does not match the witness");
    return mle_eval_device(pr, w->d_w, p0, p1, t);
}

// Test-only (include/labinius_debug.h): one MFMA with the tile GEMM's lane layouts.
int lbn_debug_mfma_probe(const int8_t *a, const int8_t *b, int32_t *d) {
    if (!a || !b || !d) return fail(LBN_ERR_ARG, "null argument");
    int8_t *da = nullptr, *db = nullptr;
    int32_t *dd = nullptr;
    int rc = LBN_OK;
    if (hipMalloc((void **)&da, 1024) != hipSuccess || hipMalloc((void **)&db, 1024) != hipSuccess ||
        hipMalloc((void **)&dd, 1024) != hipSuccess) {
        rc = fail(LBN_ERR_OOM, "probe buffers");
    } else if (hipMemcpy(da, a, 1024, hipMemcpyHostToDevice) != hipSuccess ||
               hipMemcpy(db, b, 1024, hipMemcpyHostToDevice) != hipSuccess ||
               launch_mfma_probe(nullptr, da, db, dd) != hipSuccess ||
               hipMemcpy(d, dd, 1024, hipMemcpyDeviceToHost) != hipSuccess) {
        rc = fail(LBN_ERR_HIP, "probe launch");
    }
    if (da) (void)hipFree(da);
    if (db) (void)hipFree(db);
    if (dd) (void)hipFree(dd);
    return rc;
}

// Test-only (include/labinius_debug.h): one FP4 MFMA on raw lane registers.
int lbn_debug_mfma_f4_probe(const uint8_t *a, const uint8_t *b, float *d, int scaled) {
    if (!a || !b || !d) return fail(LBN_ERR_ARG, "null argument");
    uint32_t *da = nullptr, *db = nullptr;
    float *dd = nullptr;
    int rc = LBN_OK;
    if (hipMalloc((void **)&da, 1024) != hipSuccess || hipMalloc((void **)&db, 1024) != hipSuccess ||
        hipMalloc((void **)&dd, 1024) != hipSuccess) {
        rc = fail(LBN_ERR_OOM, "probe buffers");
    } else if (hipMemcpy(da, a, 1024, hipMemcpyHostToDevice) != hipSuccess ||
               hipMemcpy(db, b, 1024, hipMemcpyHostToDevice) != hipSuccess ||
               launch_mfma_f4_probe(nullptr, da, db, dd, scaled) != hipSuccess ||
               hipMemcpy(d, dd, 1024, hipMemcpyDeviceToHost) != hipSuccess) {
        rc = fail(LBN_ERR_HIP, "probe launch");
    }
    if (da) (void)hipFree(da);
    if (db) (void)hipFree(db);
    if (dd) (void)hipFree(dd);
    return rc;
}

// Measurement (include/labinius_debug.h): the FP4 MFMA rate in TFLOP/s.
int lbn_debug_mfma_f4_rate(int random_operands, double *tflops) {
    if (!tflops) return fail(LBN_ERR_ARG, "null argument");
    const uint32_t blocks = 1024, iters = 4000;
    float *sink = nullptr;
    hipEvent_t e0 = nullptr, e1 = nullptr;
    int rc = LBN_OK;
    float ms = 0;
    if (hipMalloc((void **)&sink, (size_t)blocks * 256 * sizeof(float)) != hipSuccess) return fail(LBN_ERR_OOM, "rate sink");
    if (hipEventCreate(&e0) != hipSuccess || hipEventCreate(&e1) != hipSuccess ||
        launch_mfma_f4_rate(nullptr, random_operands, blocks, 16, sink) != hipSuccess ||
        hipEventRecord(e0, nullptr) != hipSuccess ||
        launch_mfma_f4_rate(nullptr, random_operands, blocks, iters, sink) != hipSuccess ||
        hipEventRecord(e1, nullptr) != hipSuccess || hipEventSynchronize(e1) != hipSuccess ||
        hipEventElapsedTime(&ms, e0, e1) != hipSuccess) {
        rc = fail(LBN_ERR_HIP, "rate launch");
    } else {
        *tflops = (double)blocks * 4 * iters * 16 * (16.0 * 16 * 128 * 2) / (ms * 1e-3) / 1e12;
    }
    if (e0) (void)hipEventDestroy(e0);
    if (e1) (void)hipEventDestroy(e1);
    (void)hipFree(sink);
    return rc;
}

// Test-only (include/labinius_debug.h): the GEMMs' longest K range in steps, 0 for the exact-range default.
int lbn_debug_gemm_chunk(uint32_t kb) {
    if (kb > KB_EXACT) return fail(LBN_ERR_ARG, "a K chunk beyond the accumulators' exact range");
    gemm_kb_chunk = kb;
    return LBN_OK;
}

// Measurement (include/labinius_debug.h): the int8 MFMA rate in TOP/s on constant or random operands.
int lbn_debug_mfma_rate(int operands, double *tops) {
    if (!tops) return fail(LBN_ERR_ARG, "null argument");
    if (operands < 0 || operands > 2) return fail(LBN_ERR_ARG, "operands: 0, 1 or 2");
    // 4 waves a block, 4 blocks a CU; a warm-up of about 10 ms first, so that the clock has
    // settled where this load holds it
    const uint32_t blocks = 1024, iters = 8000, warm = 20000;
    int32_t *sink = nullptr;
    hipEvent_t e0 = nullptr, e1 = nullptr;
    int rc = LBN_OK;
    float ms = 0;
    if (hipMalloc((void **)&sink, (size_t)blocks * 256 * sizeof(int32_t)) != hipSuccess) return fail(LBN_ERR_OOM, "rate sink");
    if (hipEventCreate(&e0) != hipSuccess || hipEventCreate(&e1) != hipSuccess ||
        launch_mfma_rate(nullptr, operands, blocks, warm, sink) != hipSuccess ||
        hipEventRecord(e0, nullptr) != hipSuccess ||
        launch_mfma_rate(nullptr, operands, blocks, iters, sink) != hipSuccess ||
        hipEventRecord(e1, nullptr) != hipSuccess || hipEventSynchronize(e1) != hipSuccess ||
        hipEventElapsedTime(&ms, e0, e1) != hipSuccess) {
        rc = fail(LBN_ERR_HIP, "rate launch");
    } else {
        *tops = (double)blocks * 4 * iters * 16 * (16.0 * 16 * 64 * 2) / (ms * 1e-3) / 1e12;
    }
    if (e0) (void)hipEventDestroy(e0);
    if (e1) (void)hipEventDestroy(e1);
    (void)hipFree(sink);
    return rc;
}

// Test-only introspection (include/labinius_debug.h): the host-built GEMM tables, no device.
static int copy_out(void *dst, size_t cap, const void *src, size_t bytes) {
    if (!dst) return LBN_OK;
    if (cap < bytes) return fail(LBN_ERR_SHAPE, "debug buffer too small");
    std::memcpy(dst, src, bytes);
    return LBN_OK;
}

int lbn_debug_tables(const lbn_params *params, const uint8_t seed[32], int which, void *dst, size_t cap,
                     size_t *bytes) {
    Shape s;
    int rc = make_shape(params, &s);
    if (rc) return rc;
    std::vector<int16_t> key = expand_key(s.primes, s.len_ring, seed);
    HostTables H;
    build_host_tables(s, key, H);
    const void *src = nullptr;
    size_t n = 0;
    switch (which) {
        case 0: src = H.afrag.data(); n = H.afrag.size(); break;
        case 1: src = H.lut.data(); n = H.lut.size() * 2; break;
        case 2: src = H.ee.data(); n = H.ee.size() * 2; break;
        case 3: src = H.out.data(); n = H.out.size() * 2; break;
        case 4: src = H.desc.data(); n = H.desc.size() * sizeof(TileDesc); break;
        case 5: src = H.lut_f.data(); n = H.lut_f.size() * 2; break;
        case 6: src = H.ee_f.data(); n = H.ee_f.size() * 2; break;
        case 7: src = H.out_f.data(); n = H.out_f.size() * 2; break;
        case 8: src = H.expo_f.data(); n = H.expo_f.size() * 2; break;
        case 9: src = H.desc_f.data(); n = H.desc_f.size() * sizeof(TileDesc); break;
        case 10: src = H.finish.data(); n = H.finish.size() * sizeof(FinishLimb); break;
        case 11: src = H.pw_f.data(); n = H.pw_f.size() * 2; break;
        case 12: src = &H.intt; n = sizeof(InttConsts); break;
        case 13: src = H.twi.data(); n = H.twi.size() * 4; break;
        case 14: src = key.data(); n = key.size() * 2; break;
        default: return fail(LBN_ERR_ARG, "unknown table");
    }
    if (bytes) *bytes = n;
    return copy_out(dst, cap, src, n);
}

int lbn_prover_timings(const lbn_prover *pr, float *ms, size_t n) {
    if (!pr || !ms) return fail(LBN_ERR_ARG, "null argument");
    for (size_t i = 0; i < n && i < 11; i++) ms[i] = pr->timings[i];
    return LBN_OK;
}

}  // extern "C"
