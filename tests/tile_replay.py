"""A numpy replay of the HIP tile GEMMs from the host-built tables (lbn_debug_tables), reading every
fragment in the order the kernels read it. It checks the tables and layouts on the CPU; the GPU
tests check the kernels themselves. Test infrastructure only."""
import ctypes

import numpy as np

N = 648
TILE, KSTEP, NBLK, FRAG = 96, 64, 3, 1024
A_STEP = 2 * 2 * NBLK * FRAG
EP = 88  # E rows padded to 176 bytes
TILE_ROWS = 82  # rows a tile may use
EE_TILE = TILE_ROWS * EP
YACC_PLANES = 2  # commit output planes: one per input parity of a quadratic limb
UNSIGNED_DIGITS_Q = 16384  # below it operands are split into non-negative digits, above it centred ones
DESC = np.dtype([("a_off", "<u8"), ("lut_off", "<u4"), ("ee_off", "<u4"), ("out_off", "<u4"), ("q", "<u4"),
                 ("par", "<u4"), ("y_off", "<u4")])


def debug(lbn, params, seed, which, dtype):
    L = lbn.lib()
    n = ctypes.c_size_t()
    s = (ctypes.c_uint8 * 32).from_buffer_copy(bytes(seed))
    assert L.lbn_debug_tables(ctypes.byref(params.c()), s, which, None, 0, ctypes.byref(n)) == 0
    buf = np.zeros(n.value, np.uint8)
    assert L.lbn_debug_tables(ctypes.byref(params.c()), s, which, buf.ctypes.data_as(ctypes.c_void_p),
                              ctypes.c_size_t(n.value), None) == 0
    return buf.view(dtype)


def ring_bytes(w, cols, L):
    """b[c][i][n]: bit t = coefficient 81 t + n of ring element i of column c."""
    e = w.reshape(cols, L, 4, 3)
    bits = np.zeros((cols, L, N), np.uint8)
    for m in range(162):
        word, sh = m // 64, m % 64
        for k in range(4):
            bits[:, :, 4 * m + k] = (e[:, :, k, word] >> np.uint64(sh)) & np.uint64(1)
    out = np.zeros((cols, L, 81), np.uint8)
    for t in range(8):
        out |= (bits[:, :, 81 * t:81 * t + 81] << t).astype(np.uint8)
    return out


def a_matrix(afrag, kbs):
    """[kb][kh][d][mb][lane][16] -> A[d][k][m]: lane holds row 32 mb + (lane & 31) and
    K = 64 kb + 32 kh + 16 (lane >> 5) + byte (the A operand of v_mfma_i32_32x32x32_i8)"""
    f = afrag.reshape(kbs, 2, 2, NBLK, 64, 16).astype(np.int64)
    A = np.zeros((2, kbs * KSTEP, TILE), np.int64)
    for lane in range(64):
        row = lane & 31
        for kh in range(2):
            k0 = 32 * kh + 16 * (lane >> 5)
            for mb in range(NBLK):
                for kb in range(kbs):
                    A[:, kb * KSTEP + k0:kb * KSTEP + k0 + 16, 32 * mb + row] = f[kb, kh, :, mb, lane, :]
    return A


def ee_matrix(ee):
    """[m 82][n 88] (zero from n = 81) -> E[m][n] (mod q) for m < 96, the padding rows zero"""
    E = np.zeros((TILE, 81), np.int64)
    t = ee[:TILE_ROWS * EP].reshape(TILE_ROWS, EP).astype(np.int64)
    assert not t[:, 81:].any()
    E[:TILE_ROWS] = t[:, :81]
    return E


def x_digits(bytes_kn, lut, par):
    """bytes [K][81] -> X[d][K][96] through the tile's LUT pair"""
    K = bytes_kn.shape[0]
    X = np.zeros((2, K, TILE), np.int64)
    for n in range(81):
        v = lut[((par + n) & 1) * 256 + bytes_kn[:, n].astype(np.int64)]
        X[0, :, n] = (v & 255).astype(np.uint8).view(np.int8)
        X[1, :, n] = (v >> 8).astype(np.uint8).view(np.int8)
    return X


def tile_rows(A, X, E, q):
    """y[m] = sum_n E[m][n] Z[m][n] mod q with Z = A^T X from the digit products"""
    Z = (A[0].T @ X[0]) + 128 * (A[0].T @ X[1] + A[1].T @ X[0]) + 16384 * (A[1].T @ X[1])
    Z = Z[:, :81] % q
    return (E * Z).sum(axis=1) % q


def replay_commit(lbn, params, seed, w):
    p = params
    cols, L, limbs = p.columns(), p.len_ring(), len(p.primes())
    lpad = -(-L // KSTEP) * KSTEP
    kbs = lpad // KSTEP
    afrag = debug(lbn, p, seed, 0, np.int8)
    lut = debug(lbn, p, seed, 1, np.uint16)
    ee = debug(lbn, p, seed, 2, np.uint16)
    out = debug(lbn, p, seed, 3, np.int16)
    desc = debug(lbn, p, seed, 4, DESC)
    rb = ring_bytes(w, cols, L)
    y = np.zeros((cols, YACC_PLANES * limbs * N), np.int64)
    written = np.zeros(y.shape, bool)
    a_tile = kbs * A_STEP
    for ti, d in enumerate(desc):
        # the host's A fragments, lpad / 64 steps a tile (a_off places the device's A', E folded in)
        A = a_matrix(afrag[ti * a_tile:(ti + 1) * a_tile], kbs)
        E = ee_matrix(ee[d["ee_off"]:d["ee_off"] + EE_TILE])
        lt = lut[d["lut_off"]:d["lut_off"] + 512].astype(np.int64)
        o = out[d["out_off"]:d["out_off"] + TILE]
        for c in range(cols):
            bk = np.zeros((lpad, 81), np.uint8)
            bk[:L] = rb[c]
            X = x_digits(bk, lt, int(d["par"]))
            rows = tile_rows(A, X, E, int(d["q"]))
            for m in range(TILE):
                if o[m] >= 0:
                    # every word has exactly one (tile, row) writer
                    assert not written[c, d["y_off"] + o[m]]
                    written[c, d["y_off"] + o[m]] = True
                    y[c, d["y_off"] + o[m]] = rows[m]
    raw = np.zeros((limbs, cols, N), np.uint32)
    for k, q in enumerate(p.primes()):
        planes = y[:, YACC_PLANES * k * N:YACC_PLANES * (k + 1) * N].reshape(cols, YACC_PLANES, N)
        raw[k] = (planes.sum(axis=1) % q).astype(np.uint32)
    return raw


def replay_fold(lbn, params, seed, w, challenges):
    p = params
    cols, L = p.columns(), p.len_ring()
    q = p.primes()[0]
    rpad = -(-cols // KSTEP) * KSTEP
    kbs = rpad // KSTEP
    lut = debug(lbn, p, seed, 5, np.uint16)
    ee = debug(lbn, p, seed, 6, np.uint16)
    out = debug(lbn, p, seed, 7, np.int16)
    expo = debug(lbn, p, seed, 8, np.uint16)
    desc = debug(lbn, p, seed, 9, DESC)
    pw = debug(lbn, p, seed, 11, np.uint16).astype(np.int64)
    cond = len(pw)
    mult = 2 if cond == 972 else 4
    rb = ring_bytes(w, cols, L)  # [c][i][n]
    vhat = np.zeros((L, N), np.int64)
    for d in desc:
        # the chat kernel: A[c][m] = sum_i coef_i root^{mult expo[m] p_i} mod q, two digits
        A = np.zeros((2, rpad, TILE), np.int64)
        ex = expo[d["out_off"]:d["out_off"] + TILE].astype(np.int64)
        for c, ch in enumerate(challenges):
            for m in range(TILE):
                if ex[m] == 0xFFFF:
                    continue
                acc = 0
                for i in range(ch.weight):
                    pp = ch.positions[i]
                    co = (1 - 2 * ((ch.signs >> i) & 1)) * (-1 if pp & 1 else 1)
                    acc += co * int(pw[(mult * int(ex[m]) * pp) % cond])
                x = acc % q
                if q < UNSIGNED_DIGITS_Q:  # the residue's digits, both in [0, 128)
                    xl = x & 127
                else:  # the centred residue's
                    x = x - q if x > (q - 1) // 2 else x
                    xl = ((x + 64) & 127) - 64
                A[0, c, m], A[1, c, m] = xl, (x - xl) // 128
        E = ee_matrix(ee[d["ee_off"]:d["ee_off"] + EE_TILE])
        lt = lut[d["lut_off"]:d["lut_off"] + 512].astype(np.int64)
        o = out[d["out_off"]:d["out_off"] + TILE]
        for i in range(L):
            bk = np.zeros((rpad, 81), np.uint8)
            bk[:cols] = rb[:, i, :]
            X = x_digits(bk, lt, int(d["par"]))
            rows = tile_rows(A, X, E, q)
            for m in range(TILE):
                if o[m] >= 0:
                    vhat[i, o[m]] = rows[m]
    return vhat
