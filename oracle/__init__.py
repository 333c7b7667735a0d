"""ctypes view of the oracle (oracle/oracle.c): a CPU restatement of labinius's commit / evaluate /
fold path and the host protocol pieces its parity tests need.

Test infrastructure only: the tests, ``__graft_entry__.smoke()`` and the ``cpu_baseline`` leg of
``bench.py`` may use it, and only as the checker (or as the CPU time reported beside the HIP
path). The product never imports it.
"""
import ctypes
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
# ORACLE_LIB names another build of the oracle (a sanitizer build) that tests load instead
LIB_PATH = os.environ.get("ORACLE_LIB") or os.path.join(HERE, "liboracle.so")

N = 648
N162 = 162


def build(force=False):
    """Compile oracle/oracle.c with the committed Makefile (not when ORACLE_LIB names another build)."""
    if os.environ.get("ORACLE_LIB"):
        return
    if force or not os.path.exists(LIB_PATH) or (
        os.path.getmtime(LIB_PATH) < os.path.getmtime(os.path.join(HERE, "oracle.c"))
    ):
        subprocess.run(["make", "-s", "-B", "-C", HERE], check=True)


class Challenge(ctypes.Structure):
    """The ShortChallenge layout (challenge.rs:231-235): sorted positions, sign bits, weight."""

    _fields_ = [("positions", ctypes.c_uint8 * 32), ("signs", ctypes.c_uint32), ("weight", ctypes.c_uint32)]


class OParams(ctypes.Structure):
    _fields_ = [
        ("witness_log_len", ctypes.c_uint32),
        ("column_log_len", ctypes.c_uint32),
        ("n_primes", ctypes.c_uint32),
        ("primes", ctypes.c_uint32 * 8),
        ("dropped_bits", ctypes.c_uint32),
        ("recursion", ctypes.c_uint32),
        ("matrix_seed", ctypes.c_uint8 * 32),
    ]


_lib = None


def lib():
    global _lib
    if _lib is None:
        build()
        _lib = ctypes.CDLL(LIB_PATH)
        _lib.orc_transcript_size.restype = ctypes.c_size_t
        _lib.orc_matrix_seed.restype = ctypes.c_uint64
        _lib.orc_eval_at.restype = ctypes.c_uint32
        _lib.orc_psi.restype = ctypes.c_uint32
        _lib.orc_sample_short_challenge.restype = ctypes.c_uint64
        _lib.orc_canonical_norm_sq.restype = ctypes.c_double
        _lib.orc_verify_folded_opening.restype = ctypes.c_int
        _lib.orc_is_quad.restype = ctypes.c_int
        _lib.orc_residue_bits.restype = ctypes.c_uint32
        _lib.orc_top_bits.restype = ctypes.c_uint32
        _lib.orc_bd_cap.restype = ctypes.c_uint64
        _lib.orc_pack_dropped.restype = ctypes.c_size_t
        _lib.orc_verify_opening_bd.restype = ctypes.c_int
        for name in ("orc_commitment_bytes", "orc_pack_commitment", "orc_pack_f162", "orc_rans_bound",
                     "orc_rans_encode"):
            getattr(_lib, name).restype = ctypes.c_size_t
        _lib.orc_rans_count.restype = ctypes.c_uint32
        _lib.orc_entropy_bytes.restype = ctypes.c_double
    return _lib


def _p(a):
    return a.ctypes.data_as(ctypes.c_void_p)


def is_quad(q):
    return bool(lib().orc_is_quad(ctypes.c_uint32(q)))


def psi(q):
    return lib().orc_psi(ctypes.c_uint32(q))


def tables():
    se = np.zeros(N, np.uint16)
    qse = np.zeros(324, np.uint16)
    p3 = np.zeros(N162, np.uint16)
    s648 = np.zeros((4, N162), np.uint16)
    qcs = np.zeros((2, N162), np.uint16)
    lib().orc_tables(_p(se), _p(qse), _p(p3), _p(s648), _p(qcs))
    return dict(SLOT_EXP=se, QUAD_SLOT_EXP=qse, POW3_SLOT_EXP=p3, SLOT_648=s648, QUAD_CLASS_SLOT=qcs)


def ntt(q, a):
    a = np.ascontiguousarray(a, np.uint32)
    out = np.zeros(N, np.uint32)
    lib().orc_ntt(ctypes.c_uint32(q), _p(a), _p(out))
    return out


def intt(q, a):
    a = np.ascontiguousarray(a, np.uint32)
    out = np.zeros(N, np.uint32)
    lib().orc_intt(ctypes.c_uint32(q), _p(a), _p(out))
    return out


def eval_at(q, a, u):
    a = np.ascontiguousarray(a, np.uint32)
    return lib().orc_eval_at(ctypes.c_uint32(q), _p(a), ctypes.c_uint32(u))


def eval_quad_at(q, a, u):
    a = np.ascontiguousarray(a, np.uint32)
    r = np.zeros(2, np.uint32)
    lib().orc_eval_quad_at(ctypes.c_uint32(q), _p(a), ctypes.c_uint32(u), _p(r))
    return int(r[0]), int(r[1])


def mul_mod_phi(a, b, q):
    a = np.ascontiguousarray(a, np.uint32)
    b = np.ascontiguousarray(b, np.uint32)
    out = np.zeros(N, np.uint32)
    lib().orc_mul_mod_phi(_p(a), _p(b), ctypes.c_uint32(q), _p(out))
    return out


def decompose(q, y):
    y = np.ascontiguousarray(y, np.uint32)
    out = np.zeros((4, N162), np.int16)
    lib().orc_decompose(ctypes.c_uint32(q), _p(y), _p(out))
    return out


def blake3(data, out_len=32, key=None, context=None):
    data = bytes(data)
    buf = (ctypes.c_uint8 * max(1, len(data))).from_buffer_copy(data or b"\0")
    out = (ctypes.c_uint8 * out_len)()
    if context is not None:
        mode, k = 2, ctypes.create_string_buffer(context.encode())
    elif key is not None:
        mode, k = 1, (ctypes.c_uint8 * 32).from_buffer_copy(bytes(key))
    else:
        mode, k = 0, None
    lib().orc_blake3(ctypes.c_int(mode), k, buf, ctypes.c_size_t(len(data)), out, ctypes.c_size_t(out_len))
    return bytes(out)


def matrix_seed(seed):
    return lib().orc_matrix_seed((ctypes.c_uint8 * 32).from_buffer_copy(bytes(seed)))


def xorshift_step(x):
    """rng.rs:10-14: the xorshift64* state transition."""
    lib().orc_xorshift_step.restype = ctypes.c_uint64
    return lib().orc_xorshift_step(ctypes.c_uint64(int(x)))


def key_expand(primes, len_ring, seed64):
    pr = np.array(primes, np.uint32)
    a = np.zeros((len(primes), len_ring, N), np.int16)
    lib().orc_key_expand(_p(pr), ctypes.c_int(len(primes)), ctypes.c_int(len_ring), ctypes.c_uint64(seed64), _p(a))
    return a


def witness_random(seed, n):
    w = np.zeros((n, 3), np.uint64)
    lib().orc_witness_random((ctypes.c_uint8 * 32).from_buffer_copy(bytes(seed)), ctypes.c_size_t(n), _p(w))
    return w


def lift(elems4):
    e = np.ascontiguousarray(elems4, np.uint64)
    out = np.zeros(N, np.uint32)
    lib().orc_lift(_p(e), _p(out))
    return out


def commit(primes, len_ring, cols, a, w):
    pr = np.array(primes, np.uint32)
    a = np.ascontiguousarray(a, np.int16)
    w = np.ascontiguousarray(w, np.uint64)
    raw = np.zeros((len(primes), cols, N), np.uint32)
    comp = np.zeros((cols, 4, len(primes), N162), np.int16)
    lib().orc_commit(_p(pr), ctypes.c_int(len(primes)), ctypes.c_int(len_ring), ctypes.c_int(cols), _p(a), _p(w),
                     _p(raw), _p(comp))
    return raw, comp


def challenges_array(chs):
    """Challenges as the oracle's struct (accepts any object with the same three fields)."""
    arr = (Challenge * len(chs))()
    for i, c in enumerate(chs):
        if isinstance(c, Challenge):
            arr[i] = c
        else:
            ctypes.memmove(ctypes.byref(arr[i]), ctypes.byref(c), ctypes.sizeof(Challenge))
    return arr


def embed(ch):
    e = np.zeros(N, np.int32)
    lib().orc_embed(ctypes.byref(ch), _p(e))
    return e


def fold(q, len_ring, cols, w, chs):
    w = np.ascontiguousarray(w, np.uint64)
    v = np.zeros((len_ring, N), np.int16)
    lib().orc_fold(ctypes.c_uint32(q), ctypes.c_int(len_ring), ctypes.c_int(cols), _p(w), challenges_array(chs), _p(v))
    return v


def fold_rows(q, len_ring, cols, w, chs, rows):
    w = np.ascontiguousarray(w, np.uint64)
    rows = np.ascontiguousarray(rows, np.int32)
    v = np.zeros((len(rows), N), np.int16)
    lib().orc_fold_rows(ctypes.c_uint32(q), ctypes.c_int(len_ring), ctypes.c_int(cols), _p(w), challenges_array(chs),
                        _p(rows), ctypes.c_int(len(rows)), _p(v))
    return v


def fold_direct(q, len_ring, cols, w, chs, rows):
    w = np.ascontiguousarray(w, np.uint64)
    rows = np.ascontiguousarray(rows, np.int32)
    v = np.zeros((len(rows), N), np.int16)
    lib().orc_fold_direct(ctypes.c_uint32(q), ctypes.c_int(len_ring), ctypes.c_int(cols), _p(w),
                          challenges_array(chs), _p(rows), ctypes.c_int(len(rows)), _p(v))
    return v


def f162_mul(a, b):
    a = np.ascontiguousarray(a, np.uint64)
    b = np.ascontiguousarray(b, np.uint64)
    out = np.zeros(3, np.uint64)
    lib().orc_f162_mul(_p(a), _p(b), _p(out))
    return out


def eq_table(ps):
    ps = np.ascontiguousarray(ps, np.uint64).reshape(-1, 3)
    t = np.zeros((1 << len(ps), 3), np.uint64)
    lib().orc_eq_table(_p(ps), ctypes.c_int(len(ps)), _p(t))
    return t


def row_evaluate(w, p0):
    w = np.ascontiguousarray(w, np.uint64).reshape(-1, 3)
    p0 = np.ascontiguousarray(p0, np.uint64).reshape(-1, 3)
    cols = len(w) >> len(p0)
    u = np.zeros((cols, 3), np.uint64)
    lib().orc_row_evaluate(_p(w), ctypes.c_size_t(len(w)), _p(p0), ctypes.c_int(len(p0)), _p(u))
    return u


def claim(u, p1):
    u = np.ascontiguousarray(u, np.uint64).reshape(-1, 3)
    p1 = np.ascontiguousarray(p1, np.uint64).reshape(-1, 3)
    t = np.zeros(3, np.uint64)
    lib().orc_claim(_p(u), _p(p1), ctypes.c_int(len(p1)), _p(t))
    return t


def fold_binary(u, chs):
    u = np.ascontiguousarray(u, np.uint64).reshape(-1, 3)
    out = np.zeros(3, np.uint64)
    lib().orc_fold_binary(_p(u), challenges_array(chs), ctypes.c_size_t(len(chs)), _p(out))
    return out


class Transcript:
    """challenge.rs:66-160."""

    def __init__(self, domain):
        self.buf = ctypes.create_string_buffer(lib().orc_transcript_size())
        d = bytes(domain)
        lib().orc_transcript_new(self.buf, ctypes.c_char_p(d), ctypes.c_size_t(len(d)))

    def clone(self):
        t = Transcript.__new__(Transcript)
        t.buf = ctypes.create_string_buffer(self.buf.raw, len(self.buf.raw))
        return t

    def absorb_bytes(self, b):
        b = bytes(b)
        lib().orc_absorb_bytes(self.buf, ctypes.c_char_p(b), ctypes.c_size_t(len(b)))

    def absorb_u64(self, x):
        lib().orc_absorb_u64(self.buf, ctypes.c_uint64(x))

    def fill(self, label, n):
        out = (ctypes.c_uint8 * n)()
        lib().orc_fill(self.buf, ctypes.c_char_p(bytes(label)), out, ctypes.c_size_t(n))
        return bytes(out)

    def sample_short_challenge(self, weight=28, bound=12.0):
        c = Challenge()
        att = lib().orc_sample_short_challenge(self.buf, ctypes.c_uint32(weight), ctypes.c_double(bound), ctypes.byref(c))
        return c, att


def canonical_norm_sq(c):
    return lib().orc_canonical_norm_sq(ctypes.byref(c))


def make_params(wl, cl, primes, seed, dropped_bits=0):
    p = OParams()
    p.witness_log_len = wl
    p.column_log_len = cl
    p.n_primes = len(primes)
    for i, q in enumerate(primes):
        p.primes[i] = q
    p.dropped_bits = dropped_bits
    p.recursion = 0
    for i, b in enumerate(bytes(seed)):
        p.matrix_seed[i] = b
    return p


def derive_evaluation_point(t, params, comp):
    comp = np.ascontiguousarray(comp, np.int16)
    pt = np.zeros((params.witness_log_len, 3), np.uint64)
    lib().orc_derive_evaluation_point(t.buf, ctypes.byref(params), _p(comp), _p(pt))
    rows = params.witness_log_len - params.column_log_len
    return pt[:rows].copy(), pt[rows:].copy()


def derive_folding_challenges(t, u):
    u = np.ascontiguousarray(u, np.uint64).reshape(-1, 3)
    arr = (Challenge * len(u))()
    lib().orc_derive_folding_challenges(t.buf, _p(u), ctypes.c_size_t(len(u)), arr)
    return list(arr)


def fold_commitment(primes, cols, comp, chs):
    pr = np.array(primes, np.uint32)
    comp = np.ascontiguousarray(comp, np.int16)
    out = np.zeros((4, len(primes), N162), np.int16)
    lib().orc_fold_commitment(_p(pr), ctypes.c_int(len(primes)), ctypes.c_int(cols), _p(comp), challenges_array(chs),
                              _p(out))
    return out


def verify_folded_opening(params, a, folded_comm, v, p0, u_folded):
    a = np.ascontiguousarray(a, np.int16)
    fc = np.ascontiguousarray(folded_comm, np.int16)
    v = np.ascontiguousarray(v, np.int16)
    p0 = np.ascontiguousarray(p0, np.uint64)
    uf = np.ascontiguousarray(u_folded, np.uint64)
    return bool(lib().orc_verify_folded_opening(ctypes.byref(params), _p(a), _p(fc), _p(v), _p(p0), _p(uf)))


# ---- bit-dropped commitments (bd/, wire/mod.rs:386-397, verifier.rs:232-275) ----

def residue_bits(q):
    return int(lib().orc_residue_bits(ctypes.c_uint32(q)))


def top_bits(q0, d):
    return int(lib().orc_top_bits(ctypes.c_uint32(q0), ctypes.c_uint32(d)))


def bd_cap(cols, d):
    return int(lib().orc_bd_cap(ctypes.c_uint32(cols), ctypes.c_uint32(d)))


def drop_bits(primes, cols, raw, d):
    """raw [limb][c][648] -> top [c*648] u16, digits [limb-1][c*648] u16"""
    pr = np.array(primes, np.uint32)
    raw = np.ascontiguousarray(raw, np.uint32)
    top = np.zeros(cols * N, np.uint16)
    digits = np.zeros((len(primes) - 1, cols * N), np.uint16)
    lib().orc_drop_bits(_p(pr), ctypes.c_int(len(primes)), ctypes.c_int(cols), _p(raw), ctypes.c_uint32(d),
                        _p(top), _p(digits))
    return top, digits


def pack_dropped(primes, cols, d, top, digits):
    pr = np.array(primes, np.uint32)
    top = np.ascontiguousarray(top, np.uint16)
    digits = np.ascontiguousarray(digits, np.uint16)
    bits = top_bits(primes[0], d) + sum(residue_bits(q) for q in primes[1:])
    out = np.zeros((cols * N * bits + 7) // 8 + 8, np.uint8)
    n = lib().orc_pack_dropped(_p(pr), ctypes.c_int(len(primes)), ctypes.c_int(cols), ctypes.c_uint32(d), _p(top),
                               _p(digits), _p(out))
    return bytes(out[:n])


def bd_residual(primes, len_ring, cols, a, d, top, digits, chs, v):
    pr = np.array(primes, np.uint32)
    a = np.ascontiguousarray(a, np.int16)
    top = np.ascontiguousarray(top, np.uint16)
    digits = np.ascontiguousarray(digits, np.uint16)
    v = np.ascontiguousarray(v, np.int16)
    lo, hi = ctypes.c_uint64(), ctypes.c_uint64()
    lib().orc_bd_residual(_p(pr), ctypes.c_int(len(primes)), ctypes.c_int(len_ring), ctypes.c_int(cols), _p(a),
                          ctypes.c_uint32(d), _p(top), _p(digits), challenges_array(chs), _p(v), ctypes.byref(lo),
                          ctypes.byref(hi))
    return lo.value | (hi.value << 64)


def derive_evaluation_point_dropped(t, params, packed):
    pk = np.frombuffer(bytes(packed), np.uint8).copy()
    pt = np.zeros((params.witness_log_len, 3), np.uint64)
    lib().orc_derive_evaluation_point_dropped(t.buf, ctypes.byref(params), _p(pk), ctypes.c_size_t(len(pk)), _p(pt))
    rows = params.witness_log_len - params.column_log_len
    return pt[:rows].copy(), pt[rows:].copy()


def verify_opening_bd(params, a, top, digits, chs, v, p0, u_folded):
    a = np.ascontiguousarray(a, np.int16)
    top = np.ascontiguousarray(top, np.uint16)
    digits = np.ascontiguousarray(digits, np.uint16)
    v = np.ascontiguousarray(v, np.int16)
    p0 = np.ascontiguousarray(p0, np.uint64)
    uf = np.ascontiguousarray(u_folded, np.uint64)
    return bool(lib().orc_verify_opening_bd(ctypes.byref(params), _p(a), _p(top), _p(digits), challenges_array(chs),
                                            _p(v), _p(p0), _p(uf)))


# --------------------------------------------------------------------------------------------
# the cross-field switch (switch.rs:40-209, fields/crossfield.rs, fields/sumcheck.rs). B128 values
# are uint64 pairs (low word first), F162 values uint64 triples.
# --------------------------------------------------------------------------------------------

def _u64(a, shape):
    return np.ascontiguousarray(np.asarray(a, np.uint64).reshape(shape))


def b128_mul(a, b):
    a, b = _u64(a, (2,)), _u64(b, (2,))
    out = np.zeros(2, np.uint64)
    lib().orc_b128_mul(_p(a), _p(b), _p(out))
    return out


def eq_b128(r):
    r = _u64(r, (-1, 2))
    t = np.zeros((1 << len(r), 2), np.uint64)
    lib().orc_eq_b128(_p(r), ctypes.c_int(len(r)), _p(t))
    return t


def eq_f162_msb(r):
    r = _u64(r, (-1, 3))
    t = np.zeros((1 << len(r), 3), np.uint64)
    lib().orc_eq_f162_msb(_p(r), ctypes.c_int(len(r)), _p(t))
    return t


def b128_mle(trace, r):
    trace, r = _u64(trace, (-1, 2)), _u64(r, (-1, 2))
    out = np.zeros(2, np.uint64)
    lib().orc_b128_mle(_p(trace), ctypes.c_int(len(r)), _p(r), _p(out))
    return out


def bit_mle(trace, r_lo, r_hi):
    trace, r_lo, r_hi = _u64(trace, (-1, 2)), _u64(r_lo, (7, 2)), _u64(r_hi, (-1, 2))
    out = np.zeros(2, np.uint64)
    lib().orc_bit_mle(_p(trace), ctypes.c_int(len(r_hi)), _p(r_lo), _p(r_hi), _p(out))
    return out


def transparent_coeff(r_hi, r_pp, batch):
    r_hi, r_pp, batch = _u64(r_hi, (-1, 2)), _u64(r_pp, (-1, 3)), _u64(batch, (128, 3))
    out = np.zeros(3, np.uint64)
    lib().orc_transparent_coeff(_p(r_hi), _p(r_pp), ctypes.c_int(len(r_hi)), _p(batch), _p(out))
    return out


def _absorb_b128(self, xs):
    xs = _u64(xs, (-1, 2))
    lib().orc_absorb_b128(self.buf, _p(xs), ctypes.c_size_t(len(xs)))


def _absorb_f162(self, xs):
    xs = _u64(xs, (-1, 3))
    lib().orc_absorb_f162(self.buf, _p(xs), ctypes.c_size_t(len(xs)))


def _sample_f162(self, label, n):
    out = np.zeros((n, 3), np.uint64)
    lib().orc_sample_f162(self.buf, ctypes.c_char_p(bytes(label)), ctypes.c_size_t(n), _p(out))
    return out


def _sample_b128(self, label, n):
    out = np.zeros((n, 2), np.uint64)
    lib().orc_sample_b128(self.buf, ctypes.c_char_p(bytes(label)), ctypes.c_size_t(n), _p(out))
    return out


Transcript.absorb_b128 = _absorb_b128
Transcript.absorb_f162 = _absorb_f162
Transcript.sample_f162 = _sample_f162
Transcript.sample_b128 = _sample_b128


def switch_prove(trace, cl, r_lo, r_hi, claim, transcript):
    """switch.rs:137-164: (partials [128][2], rounds [l][2][3], opened [3], p0, p1)."""
    trace, r_hi = _u64(trace, (-1, 2)), _u64(r_hi, (-1, 2))
    r_lo = _u64(r_lo if r_lo is not None else np.zeros((0, 2)), (-1, 2))
    claim = _u64(claim, (2,))
    l = len(r_hi)
    partials = np.zeros((128, 2), np.uint64)
    rounds = np.zeros((l, 2, 3), np.uint64)
    opened = np.zeros(3, np.uint64)
    p0 = np.zeros((l - cl, 3), np.uint64)
    p1 = np.zeros((cl, 3), np.uint64)
    lib().orc_switch_prove(_p(trace), ctypes.c_int(l), ctypes.c_int(cl), _p(r_lo), ctypes.c_int(len(r_lo)), _p(r_hi),
                           _p(claim), transcript.buf, _p(partials), _p(rounds), _p(opened), _p(p0), _p(p1))
    return partials, rounds, opened, p0, p1


def switch_verify(l, cl, r_lo, r_hi, claim, partials, rounds, opened, transcript):
    """switch.rs:166-209: (accepted, p0, p1)."""
    r_hi = _u64(r_hi, (-1, 2))
    r_lo = _u64(r_lo if r_lo is not None else np.zeros((0, 2)), (-1, 2))
    claim, partials = _u64(claim, (2,)), _u64(partials, (128, 2))
    rounds, opened = _u64(rounds, (l, 2, 3)), _u64(opened, (3,))
    p0 = np.zeros((l - cl, 3), np.uint64)
    p1 = np.zeros((cl, 3), np.uint64)
    ok = lib().orc_switch_verify(ctypes.c_int(l), ctypes.c_int(cl), _p(r_lo), ctypes.c_int(len(r_lo)), _p(r_hi),
                                 _p(claim), _p(partials), _p(rounds), _p(opened), transcript.buf, _p(p0), _p(p1))
    return bool(ok), p0, p1


# =============================================================================================
# wire formats (wire/mod.rs, wire/rans.rs)
# =============================================================================================

WIRE_OK, WIRE_TRUNCATED, WIRE_MALFORMED = 0, 1, 2
LANES = 64


def commitment_bytes(primes, cols):
    pr = np.array(primes, np.uint32)
    return lib().orc_commitment_bytes(_p(pr), ctypes.c_int(len(primes)), ctypes.c_int(cols))


def pack_commitment(primes, cols, comp):
    """wire/mod.rs:226-243: comp [c][4][limb][162] int16 centered -> bytes"""
    pr = np.array(primes, np.uint32)
    comp = np.ascontiguousarray(comp, np.int16)
    assert comp.size == cols * 4 * len(primes) * N162
    out = np.zeros(commitment_bytes(primes, cols) + 8, np.uint8)
    n = lib().orc_pack_commitment(_p(pr), ctypes.c_int(len(primes)), ctypes.c_int(cols), _p(comp), _p(out))
    return bytes(out[:n])


def unpack_commitment(primes, cols, data):
    """wire/mod.rs:246-301: (error, comp [c][4][limb][162])"""
    pr = np.array(primes, np.uint32)
    buf = np.frombuffer(bytes(data) + b"\0", np.uint8)
    comp = np.zeros((cols, 4, len(primes), N162), np.int16)
    err = lib().orc_unpack_commitment(_p(pr), ctypes.c_int(len(primes)), ctypes.c_int(cols), _p(buf),
                                      ctypes.c_size_t(len(data)), _p(comp))
    return err, comp


def unpack_dropped(primes, cols, d, data):
    """wire/mod.rs:399-431: (error, top, digits)"""
    pr = np.array(primes, np.uint32)
    buf = np.frombuffer(bytes(data) + b"\0", np.uint8)
    top = np.zeros(cols * N, np.uint16)
    digits = np.zeros((len(primes) - 1, cols * N), np.uint16)
    err = lib().orc_unpack_dropped(_p(pr), ctypes.c_int(len(primes)), ctypes.c_int(cols), ctypes.c_uint32(d),
                                   _p(buf), ctypes.c_size_t(len(data)), _p(top), _p(digits))
    return err, top, digits


def pack_f162(values):
    """wire/mod.rs:445-453"""
    v = np.ascontiguousarray(np.asarray(values, np.uint64).reshape(-1, 3))
    out = np.zeros((len(v) * 162 + 7) // 8 + 8, np.uint8)
    n = lib().orc_pack_f162(_p(v), ctypes.c_size_t(len(v)), _p(out))
    return bytes(out[:n])


def unpack_f162(data, count):
    """wire/mod.rs:456-470: (error, values [count][3])"""
    buf = np.frombuffer(bytes(data) + b"\0", np.uint8)
    v = np.zeros((count, 3), np.uint64)
    err = lib().orc_unpack_f162(_p(buf), ctypes.c_size_t(len(data)), ctypes.c_size_t(count), _p(v))
    return err, v


def rans_quantize(counts):
    """rans.rs:236-271: the last entry is the escape symbol"""
    c = np.ascontiguousarray(counts, np.uint64)
    f = np.zeros(len(c), np.uint32)
    lib().orc_rans_quantize(_p(c), ctypes.c_size_t(len(c)), _p(f))
    return f


def rans_encode(v, base_q, ntt=False):
    """rans.rs:18-99: v [count][648] int16 -> bytes"""
    v = np.ascontiguousarray(np.asarray(v, np.int16).reshape(-1, N))
    out = np.zeros(lib().orc_rans_bound(ctypes.c_size_t(len(v))), np.uint8)
    n = lib().orc_rans_encode(_p(v), ctypes.c_size_t(len(v)), ctypes.c_int(int(ntt)), ctypes.c_uint16(base_q),
                              _p(out))
    return bytes(out[:n])


def rans_decode(data):
    """rans.rs:101-196: (error, v [count][648] int16, ntt)"""
    data = bytes(data)
    buf = np.frombuffer(data + b"\0", np.uint8)
    count = lib().orc_rans_count(_p(buf), ctypes.c_size_t(len(data)))
    cap = min(count, 1 << 16)
    v = np.zeros((max(cap, 1), N), np.int16)
    cnt, ntt = ctypes.c_uint32(), ctypes.c_int()
    err = lib().orc_rans_decode(_p(buf), ctypes.c_size_t(len(data)), _p(v), ctypes.c_size_t(cap), ctypes.byref(cnt),
                                ctypes.byref(ntt))
    return err, v[:cnt.value] if err == WIRE_OK else None, bool(ntt.value)


def entropy_bytes(v):
    """rans.rs:500-524"""
    v = np.ascontiguousarray(v, np.int16).ravel()
    return lib().orc_entropy_bytes(_p(v), ctypes.c_size_t(v.size))


def a_times_v(primes, len_ring, a, v):
    """verifier.rs:205-224: A v decomposed, [4][limb][162] int16 (a as key_expand returns it)."""
    pr = np.array(primes, np.uint32)
    a = np.ascontiguousarray(a, np.int16)
    v = np.ascontiguousarray(v, np.int16)
    assert v.size == len_ring * N
    out = np.zeros((4, len(primes), N162), np.int16)
    lib().orc_a_times_v(_p(pr), ctypes.c_int(len(primes)), ctypes.c_int(len_ring), _p(a), _p(v), _p(out))
    return out
