"""labinius on MI355X: the commit / evaluate / fold prover behind the reference's API.

A thin Python mirror of the Rust surface (crates/pcs/src/scheme/*.rs) over the C ABI of
``liblabinius.so`` (include/labinius.h). The arithmetic runs in HIP kernels on the GPU; the
transcript, the challenge sampler and the verifier run on the host in C++, as in the reference.
There is no CPU fallback: without the built library or a HIP device every entry point raises.
"""
import ctypes
import enum
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
# LABINIUS_LIB names another build of the same library (a sanitizer build of the host code, say)
LIB_PATH = os.environ.get("LABINIUS_LIB") or os.path.join(HERE, "liblabinius.so")

N = 648
N162 = 162

LBN_OK = 0
LBN_REJECTED = 13


class Modulus(enum.IntEnum):
    """ring/mod.rs:87-95, by prime."""

    Q2917_Q_S = 2917
    Q3889_FS_S = 3889
    Q4861_Q_S = 4861
    Q9721_FS_S = 9721
    Q12637_Q_S = 12637
    Q17497_FS_L = 17497
    Q19441_FS_L = 19441

    def is_quadratic(self):
        return (self.value - 1) % 972 == 0 and (self.value - 1) % 1944 != 0


ALL_MODULI = list(Modulus)


class ParamError(ValueError):
    """scheme/params.rs:22-33."""

    CODES = {1: "ColumnsExceedWitness", 2: "TooFewColumns", 3: "ColumnTooShort", 4: "DuplicateModulus",
             5: "BaseIsAlsoExtra", 6: "UnknownModulus"}

    def __init__(self, code, msg=""):
        self.code = code
        self.kind = self.CODES.get(code, str(code))
        super().__init__(f"{self.kind}: {msg}")


class WitnessError(ValueError):
    """scheme/witness.rs:12-15 (WrongLength)."""


class VerificationError(Exception):
    """scheme/commitment.rs:449-451 (Rejected)."""


class LabiniusError(RuntimeError):
    def __init__(self, code, msg):
        self.code = code
        super().__init__(f"labinius error {code}: {msg}")


class F162Arr(ctypes.Structure):
    _fields_ = [("w", ctypes.c_uint64 * 3)]


class ShortChallenge(ctypes.Structure):
    """challenge.rs:231-235: sorted positions, sign bits, weight."""

    _fields_ = [("positions", ctypes.c_uint8 * 32), ("signs", ctypes.c_uint32), ("weight", ctypes.c_uint32)]

    def coeffs(self):
        c = np.zeros(N162, np.int8)
        for i in range(self.weight):
            c[self.positions[i]] = 1 - 2 * ((self.signs >> i) & 1)
        return c

    def to_f162(self):
        x = np.zeros(3, np.uint64)
        for i in range(self.weight):
            p = self.positions[i]
            x[p >> 6] |= np.uint64(1 << (p & 63))
        return x

    def __eq__(self, other):
        return (bytes(self.positions) == bytes(other.positions) and self.signs == other.signs
                and self.weight == other.weight)


class CParams(ctypes.Structure):
    _fields_ = [("witness_log_len", ctypes.c_uint32), ("column_log_len", ctypes.c_uint32), ("base", ctypes.c_uint32),
                ("n_extra", ctypes.c_uint32), ("extra", ctypes.c_uint32 * 6), ("opening", ctypes.c_uint32),
                ("dropped_bits", ctypes.c_uint32)]


_lib = None


def lib():
    """The HIP library; raises if it has not been built (there is no fallback)."""
    global _lib
    if _lib is None:
        # When torch (ROCm build) is present, its HIP runtime must be the one in the process before
        # the library resolves libamdhip64.so.7: loading the system runtime first and torch's
        # second leaves the library without a device.
        if not os.environ.get("LABINIUS_NO_TORCH"):
            try:
                import torch  # noqa: F401
            except Exception:
                pass
        if not os.path.exists(LIB_PATH):
            raise LabiniusError(-1, f"{LIB_PATH} is not built (run __graft_entry__.build())")
        L = ctypes.CDLL(LIB_PATH)
        L.lbn_last_error.restype = ctypes.c_char_p
        L.lbn_transcript_new.restype = ctypes.c_void_p
        L.lbn_transcript_new.argtypes = [ctypes.c_char_p, ctypes.c_size_t]
        L.lbn_transcript_clone.restype = ctypes.c_void_p
        L.lbn_transcript_clone.argtypes = [ctypes.c_void_p]
        L.lbn_transcript_free.argtypes = [ctypes.c_void_p]
        L.lbn_transcript_absorb_bytes.argtypes = [ctypes.c_void_p, ctypes.c_char_p, ctypes.c_size_t]
        L.lbn_transcript_absorb_u64.argtypes = [ctypes.c_void_p, ctypes.c_uint64]
        L.lbn_transcript_fill.argtypes = [ctypes.c_void_p, ctypes.c_char_p, ctypes.c_size_t, ctypes.c_void_p,
                                          ctypes.c_size_t]
        L.lbn_sample_short_challenge.restype = ctypes.c_uint64
        L.lbn_sample_short_challenge.argtypes = [ctypes.c_void_p, ctypes.c_uint32, ctypes.c_double, ctypes.c_void_p]
        L.lbn_bd_cap.restype = ctypes.c_uint64
        L.lbn_dwitness_device_ptr.restype = ctypes.c_void_p
        L.lbn_dwitness_device_ptr.argtypes = [ctypes.c_void_p]
        for name in ("lbn_pp_release", "lbn_prover_release", "lbn_dwitness_release"):
            getattr(L, name).argtypes = [ctypes.c_void_p]
        _lib = L
    return _lib


def _check(rc):
    if rc != LBN_OK:
        raise LabiniusError(rc, lib().lbn_last_error().decode())


def _p(a):
    return a.ctypes.data_as(ctypes.c_void_p)


# =============================================================================================
# parameters (scheme/params.rs)
# =============================================================================================

class Opening:
    CLEAR = ("clear", 0)
    RECURSIVE = ("recursive", 0)

    @staticmethod
    def bit_dropped(bits):
        return ("bit_dropped", int(bits))


class Suite:
    def __init__(self, name, witness_log_len, column_log_len_clear, column_log_len_recursive, moduli, moduli_bd,
                 dropped_bits):
        self.name = name
        self.witness_log_len = witness_log_len
        self.column_log_len_clear = column_log_len_clear
        self.column_log_len_recursive = column_log_len_recursive
        self.moduli = moduli
        self.moduli_bd = moduli_bd
        self.dropped_bits = dropped_bits


M = Modulus
# scheme/params.rs:91-128
SUITES = [
    Suite("sizes", 18, 7, 8, [M.Q9721_FS_S, M.Q12637_Q_S], [M.Q3889_FS_S, M.Q2917_Q_S, M.Q4861_Q_S], 10),
    Suite("sizem", 20, 8, 9, [M.Q3889_FS_S, M.Q2917_Q_S, M.Q4861_Q_S], [M.Q3889_FS_S, M.Q2917_Q_S, M.Q4861_Q_S], 9),
    Suite("sizel", 22, 9, 10, [M.Q3889_FS_S, M.Q2917_Q_S, M.Q4861_Q_S], [M.Q3889_FS_S, M.Q2917_Q_S, M.Q4861_Q_S], 8),
    Suite("sizexl", 24, 10, 11, [M.Q3889_FS_S, M.Q2917_Q_S, M.Q9721_FS_S], [M.Q3889_FS_S, M.Q2917_Q_S, M.Q4861_Q_S],
          7),
]


class Params:
    """scheme/params.rs:62-288."""

    def __init__(self, witness_log_len, column_log_len, extra_moduli=(), opening=Opening.CLEAR, base=Modulus.Q3889_FS_S):
        self.witness_log_len = int(witness_log_len)
        self.column_log_len = int(column_log_len)
        self.base = Modulus(int(base))
        self.extra_moduli = [Modulus(int(m)) for m in extra_moduli]
        self.opening = opening
        rc = lib().lbn_params_check(ctypes.byref(self.c()))
        if rc != LBN_OK:
            raise ParamError(rc, lib().lbn_last_error().decode())

    @classmethod
    def new(cls, witness_log_len, column_log_len, extra_moduli=(), opening=Opening.CLEAR):
        return cls(witness_log_len, column_log_len, extra_moduli, opening)

    @classmethod
    def with_base(cls, witness_log_len, column_log_len, base, extra_moduli=(), opening=Opening.CLEAR):
        return cls(witness_log_len, column_log_len, extra_moduli, opening, base)

    @classmethod
    def basic(cls):
        """Params::basic(): 2^18 F162 in 128 columns, moduli 3889 and 9721, Clear."""
        return cls(18, 7, [Modulus.Q9721_FS_S])

    @classmethod
    def sized(cls, suite, opening=Opening.CLEAR):
        cl = suite.column_log_len_recursive if opening[0] == "recursive" else suite.column_log_len_clear
        lst = suite.moduli_bd if opening[0] == "bit_dropped" else suite.moduli
        return cls(suite.witness_log_len, cl, lst[1:], opening, lst[0])

    def c(self):
        p = CParams()
        p.witness_log_len = self.witness_log_len
        p.column_log_len = self.column_log_len
        p.base = int(self.base)
        p.n_extra = len(self.extra_moduli)
        for i, m in enumerate(self.extra_moduli[:6]):
            p.extra[i] = int(m)
        kind, bits = self.opening
        p.opening = {"clear": 0, "bit_dropped": 1, "recursive": 2}[kind]
        p.dropped_bits = bits
        return p

    def witness_len(self):
        return 1 << self.witness_log_len

    def columns(self):
        return 1 << self.column_log_len

    def row_log_len(self):
        return self.witness_log_len - self.column_log_len

    def column_len(self):
        return 1 << self.row_log_len()

    def len_ring(self):
        return self.column_len() // 4

    def primes(self):
        return [int(self.base)] + [int(m) for m in self.extra_moduli]

    def dropped_bits(self):
        return self.opening[1] if self.opening[0] == "bit_dropped" else 0

    def fold_cap(self):
        import math
        return math.ceil(108.0 * (self.witness_len() // 4 * N))

    def __eq__(self, o):
        return (self.witness_log_len, self.column_log_len, self.base, self.extra_moduli, self.opening) == (
            o.witness_log_len, o.column_log_len, o.base, o.extra_moduli, o.opening)


class PublicParameters:
    """PublicParameters::from_seed (scheme/params.rs:320-339): A expanded and resident on the GPU."""

    def __init__(self, params, matrix_seed, handle):
        self.params = params
        self.matrix_seed = bytes(matrix_seed)
        self._h = handle

    @classmethod
    def from_seed(cls, params, matrix_seed):
        h = ctypes.c_void_p()
        seed = (ctypes.c_uint8 * 32).from_buffer_copy(bytes(matrix_seed))
        _check(lib().lbn_pp_create(ctypes.byref(params.c()), seed, ctypes.byref(h)))
        return cls(params, matrix_seed, h.value)

    @classmethod
    def host_only(cls, params, matrix_seed):
        """The same parameters without device state (the key and the shape): enough for a Verifier
        on a host without a GPU; a Prover cannot be built on them."""
        h = ctypes.c_void_p()
        seed = (ctypes.c_uint8 * 32).from_buffer_copy(bytes(matrix_seed))
        _check(lib().lbn_pp_create_host(ctypes.byref(params.c()), seed, ctypes.byref(h)))
        return cls(params, matrix_seed, h.value)

    def key(self):
        """CommitmentKey::row for every limb: [limb][len_ring][648] int16, centered."""
        p = self.params
        a = np.zeros((len(p.primes()), p.len_ring(), N), np.int16)
        _check(lib().lbn_pp_key(ctypes.c_void_p(self._h), _p(a), ctypes.c_size_t(a.size)))
        return a

    def __del__(self):
        if getattr(self, "_h", None) and _lib is not None:
            _lib.lbn_pp_release(ctypes.c_void_p(self._h))
            self._h = None


# =============================================================================================
# witness and prover (scheme/witness.rs, scheme/prover.rs)
# =============================================================================================

class Witness:
    """scheme/witness.rs:27-117: witness_len() F162 as an (n, 3) uint64 array."""

    def __init__(self, params, elements):
        self.params = params
        self.elements = elements

    @classmethod
    def random(cls, params, seed):
        n = params.witness_len()
        w = np.zeros((n, 3), np.uint64)
        s = (ctypes.c_uint8 * 32).from_buffer_copy(bytes(seed))
        _check(lib().lbn_witness_random(ctypes.byref(params.c()), s, _p(w), ctypes.c_size_t(n)))
        return cls(params, w)

    @classmethod
    def from_elements(cls, params, elements):
        e = np.ascontiguousarray(elements, np.uint64).reshape(-1, 3)
        if len(e) != params.witness_len():
            raise WitnessError("the witness is not witness_len() elements")
        return cls(params, e)

    @classmethod
    def lifted(cls, params, trace):
        """Witness::lifted: a B128 trace (n, 2) uint64 zero-extended into F162 (fields/scalar.rs:72-74)."""
        t = np.ascontiguousarray(trace, np.uint64).reshape(-1, 2)
        if len(t) != params.witness_len():
            raise WitnessError("the witness is not witness_len() elements")
        e = np.zeros((len(t), 3), np.uint64)
        e[:, :2] = t
        return cls(params, e)

    def row_evaluate(self, point, prover):
        return prover.row_evaluate(self, point)

    def mle_evaluate(self, point, prover):
        return prover.mle_evaluate(self, point)


class ResidentWitness:
    """A witness uploaded once to HBM (the prover's lbn_dwitness)."""

    def __init__(self, prover, witness, handle):
        self.prover = prover
        self.params = witness.params
        self._h = handle

    def device_ptr(self):
        return lib().lbn_dwitness_device_ptr(ctypes.c_void_p(self._h))

    def __del__(self):
        if getattr(self, "_h", None) and _lib is not None:
            _lib.lbn_dwitness_release(ctypes.c_void_p(self._h))
            self._h = None


class Dropped:
    """bd/mod.rs:7-56: per coefficient of every column the base residue's top bits and the mixed-radix
    digits over the other primes."""

    def __init__(self, params, top, digits):
        self.params = params
        self.top = np.ascontiguousarray(top, np.uint16)        # [columns * 648]
        self.digits = np.ascontiguousarray(digits, np.uint16)  # [limbs - 1][columns * 648]

    @classmethod
    def of_raw(cls, params, raw):
        """bd/digits.rs:121-182 from the raw rows [limb][column][648]."""
        count = params.columns() * N
        top = np.zeros(count, np.uint16)
        digits = np.zeros((len(params.primes()) - 1, count), np.uint16)
        raw = np.ascontiguousarray(raw, np.uint32)
        _check(lib().lbn_drop_bits(ctypes.byref(params.c()), _p(raw), _p(top), _p(digits)))
        return cls(params, top, digits)

    @classmethod
    def _empty(cls, params):
        count = params.columns() * N
        return np.zeros(count, np.uint16), np.zeros((len(params.primes()) - 1, count), np.uint16)

    @classmethod
    def of_device(cls, prover, params, raw):
        """The same digits computed on the GPU by prover (whose moduli params shares)."""
        top, digits = cls._empty(params)
        raw = np.ascontiguousarray(raw, np.uint32)
        _check(lib().lbn_drop_bits_device(ctypes.c_void_p(prover._h), ctypes.byref(params.c()), _p(raw), _p(top),
                                          _p(digits)))
        return cls(params, top, digits)

    @classmethod
    def of_last_commit(cls, prover):
        """The digits of prover's last commitment, from its raw rows still on the GPU."""
        top, digits = cls._empty(prover.params)
        _check(lib().lbn_commit_drop_bits(ctypes.c_void_p(prover._h), _p(top), _p(digits)))
        return cls(prover.params, top, digits)

    def dropped_bits(self):
        return self.params.dropped_bits()

    def pack(self):
        """wire::pack_dropped (wire/mod.rs:386-397)."""
        n = ctypes.c_size_t()
        _check(lib().lbn_pack_dropped(ctypes.byref(self.params.c()), _p(self.top), _p(self.digits), None,
                                      ctypes.c_size_t(0), ctypes.byref(n)))
        out = np.zeros(max(1, n.value), np.uint8)
        _check(lib().lbn_pack_dropped(ctypes.byref(self.params.c()), _p(self.top), _p(self.digits), _p(out),
                                      ctypes.c_size_t(out.size), None))
        return bytes(out[:n.value])

    def wire_bytes(self):
        n = ctypes.c_size_t()
        _check(lib().lbn_pack_dropped(ctypes.byref(self.params.c()), _p(self.top), _p(self.digits), None,
                                      ctypes.c_size_t(0), ctypes.byref(n)))
        return n.value


class Commitment:
    """scheme/commitment.rs:24-118: the 4 x columns matrix per modulus, plus the raw rows; in the
    bit-dropped mode the value is the Dropped form (scheme/prover.rs:69-72)."""

    def __init__(self, primes, columns, components, raw, dropped=None):
        self.primes = primes
        self.columns_ = columns
        self.components = components  # [column][row][limb][162] int16
        self.raw = raw                # [limb][column][648] uint32
        self.dropped = dropped        # Dropped or None

    def rows(self):
        return 4

    def columns(self):
        return self.columns_

    def moduli(self):
        return self.primes

    def element(self, row, column, modulus_index):
        return self.components[column, row, modulus_index]


class CommitmentOpening:
    def __init__(self, prover, handle):
        self.prover = prover
        self._h = handle


class EvaluationPoint:
    """scheme/commitment.rs:141-183."""

    def __init__(self, p0, p1):
        self.p0 = np.ascontiguousarray(p0, np.uint64).reshape(-1, 3)
        self.p1 = np.ascontiguousarray(p1, np.uint64).reshape(-1, 3)

    @classmethod
    def msb_first(cls, params, coordinates):
        c = np.ascontiguousarray(coordinates, np.uint64).reshape(-1, 3)
        assert len(c) == params.witness_log_len, "one coordinate per variable of the witness"
        high, low = c[:params.column_log_len], c[params.column_log_len:]
        return cls(low[::-1].copy(), high[::-1].copy())


class RowEvaluation:
    def __init__(self, values):
        self.values = np.ascontiguousarray(values, np.uint64).reshape(-1, 3)


class FoldingChallenges:
    def __init__(self, challenges):
        self.challenges = challenges

    def __len__(self):
        return len(self.challenges)

    def array(self):
        arr = (ShortChallenge * len(self.challenges))()
        for i, c in enumerate(self.challenges):
            arr[i] = c
        return arr


class FoldedWitness:
    def __init__(self, elements):
        self.elements = elements  # [len_ring][648] int16


class FoldedCommitment:
    def __init__(self, primes, rows):
        self.primes = primes
        self.rows = rows  # [4][limb][162] int16

    def element(self, row, modulus_index):
        return self.rows[row, modulus_index]


class Prover:
    """scheme/prover.rs:21-133: the device workspace and one HIP stream; new() runs a warm-up round."""

    def __init__(self, pp):
        self.pp = pp
        self.params = pp.params
        h = ctypes.c_void_p()
        _check(lib().lbn_prover_create(ctypes.c_void_p(pp._h), ctypes.byref(h)))
        self._h = h.value

    def upload(self, witness):
        h = ctypes.c_void_p()
        e = np.ascontiguousarray(witness.elements, np.uint64)
        _check(lib().lbn_witness_upload(ctypes.c_void_p(self._h), _p(e), ctypes.c_size_t(len(e)), ctypes.byref(h)))
        return ResidentWitness(self, witness, h.value)

    def upload_params(self, witness):
        """A witness of other Params that share this prover's key (row_log_len and moduli)."""
        h = ctypes.c_void_p()
        e = np.ascontiguousarray(witness.elements, np.uint64)
        _check(lib().lbn_witness_upload_params(ctypes.c_void_p(self._h), ctypes.byref(witness.params.c()), _p(e),
                                               ctypes.c_size_t(len(e)), ctypes.byref(h)))
        return ResidentWitness(self, witness, h.value)

    def commit_batch(self, resident, want_outputs=True):
        """Every witness of the list committed in one launch; one Commitment per witness, each equal
        to that witness's own Prover::commit."""
        n = len(resident)
        hs = (ctypes.c_void_p * n)(*[r._h for r in resident])
        if not want_outputs:
            _check(lib().lbn_commit_batch(ctypes.c_void_p(self._h), hs, ctypes.c_size_t(n), None, None))
            return None
        comps, raws = [], []
        cptr = (ctypes.c_void_p * n)()
        rptr = (ctypes.c_void_p * n)()
        for i, r in enumerate(resident):
            p = r.params
            limbs = len(p.primes())
            comps.append(np.zeros((p.columns(), 4, limbs, N162), np.int16))
            raws.append(np.zeros((limbs, p.columns(), N), np.uint32))
            cptr[i] = comps[i].ctypes.data
            rptr[i] = raws[i].ctypes.data
        _check(lib().lbn_commit_batch(ctypes.c_void_p(self._h), hs, ctypes.c_size_t(n), cptr, rptr))
        out = []
        for r, comp, raw in zip(resident, comps, raws):
            p = r.params
            dropped = Dropped.of_device(self, p, raw) if p.dropped_bits() > 0 else None
            out.append(Commitment(p.primes(), p.columns(), comp, raw, dropped))
        return out

    def batch_info(self):
        t, sp, ms = ctypes.c_uint32(), ctypes.c_uint32(), ctypes.c_float()
        _check(lib().lbn_batch_info(ctypes.c_void_p(self._h), ctypes.byref(t), ctypes.byref(sp), ctypes.byref(ms)))
        return {"tiles_per_xcd_round": t.value, "k_ranges": sp.value, "ms": ms.value}

    def _buffers(self):
        p = self.params
        limbs = len(p.primes())
        comp = np.zeros((p.columns(), 4, limbs, N162), np.int16)
        raw = np.zeros((limbs, p.columns(), N), np.uint32)
        return comp, raw

    def commit(self, witness):
        if isinstance(witness, ResidentWitness):
            return self.commit_resident(witness)
        if witness.params != self.params:
            raise AssertionError("the witness was built for other parameters")
        comp, raw = self._buffers()
        op = ctypes.c_void_p()
        e = np.ascontiguousarray(witness.elements, np.uint64)
        _check(lib().lbn_commit(ctypes.c_void_p(self._h), _p(e), ctypes.c_size_t(len(e)), _p(comp), _p(raw),
                                ctypes.byref(op)))
        return self._commitment(comp, raw), CommitmentOpening(self, op.value)

    def _commitment(self, comp, raw):
        p = self.params
        dropped = Dropped.of_last_commit(self) if p.dropped_bits() > 0 else None
        return Commitment(p.primes(), p.columns(), comp, raw, dropped)

    def commit_resident(self, rw, want_outputs=True):
        comp, raw = self._buffers()
        op = ctypes.c_void_p()
        _check(lib().lbn_commit_resident(ctypes.c_void_p(self._h), ctypes.c_void_p(rw._h),
                                         _p(comp) if want_outputs else None, _p(raw) if want_outputs else None,
                                         ctypes.byref(op)))
        if not want_outputs:
            return Commitment(self.params.primes(), self.params.columns(), comp, raw), CommitmentOpening(self, op.value)
        return self._commitment(comp, raw), CommitmentOpening(self, op.value)

    def fold(self, opening, challenges, want_output=True):
        p = self.params
        v = np.zeros((p.len_ring(), N), np.int16)
        arr = challenges.array()
        h, opening._h = opening._h, None
        if h is None:
            raise AssertionError("the opening was already consumed")
        _check(lib().lbn_fold(ctypes.c_void_p(self._h), ctypes.c_void_p(h), arr, ctypes.c_size_t(len(arr)),
                              _p(v) if want_output else None))
        return FoldedWitness(v)

    def fold_resident(self, rw, challenges, want_output=True):
        """The fold of a witness resident since its commitment, without an opening (lbn_fold_resident)."""
        p = self.params
        v = np.zeros((p.len_ring(), N), np.int16)
        arr = challenges.array()
        _check(lib().lbn_fold_resident(ctypes.c_void_p(self._h), ctypes.c_void_p(rw._h), arr,
                                       ctypes.c_size_t(len(arr)), _p(v) if want_output else None))
        return FoldedWitness(v)

    def recycle(self, opening):
        h, opening._h = opening._h, None
        _check(lib().lbn_recycle(ctypes.c_void_p(self._h), ctypes.c_void_p(h)))

    def _resident(self, witness):
        return witness if isinstance(witness, ResidentWitness) else self.upload(witness)

    def row_evaluate(self, witness, point):
        rw = self._resident(witness)
        p = self.params
        u = np.zeros((p.columns(), 3), np.uint64)
        p0 = np.ascontiguousarray(point.p0, np.uint64)
        _check(lib().lbn_row_evaluate(ctypes.c_void_p(self._h), ctypes.c_void_p(rw._h), _p(p0),
                                      ctypes.c_size_t(len(p0)), _p(u)))
        return RowEvaluation(u)

    def mle_evaluate(self, witness, point):
        rw = self._resident(witness)
        t = np.zeros(3, np.uint64)
        p0 = np.ascontiguousarray(point.p0, np.uint64)
        p1 = np.ascontiguousarray(point.p1, np.uint64)
        _check(lib().lbn_mle_evaluate(ctypes.c_void_p(self._h), ctypes.c_void_p(rw._h), _p(p0),
                                      ctypes.c_size_t(len(p0)), _p(p1), ctypes.c_size_t(len(p1)), _p(t)))
        return t

    def timings(self):
        """Device ms of the last commit / fold / row evaluation stages (include/labinius.h)."""
        ms = (ctypes.c_float * 11)()
        _check(lib().lbn_prover_timings(ctypes.c_void_p(self._h), ms, ctypes.c_size_t(11)))
        keys = ["commit", "commit_layout", "commit_gemm", "commit_finish", "fold", "fold_challenges", "fold_gemm",
                "fold_intt", "eval", "eval_eq", "eval_reduce"]
        return dict(zip(keys, list(ms)))

    def __del__(self):
        if getattr(self, "_h", None) and _lib is not None:
            _lib.lbn_prover_release(ctypes.c_void_p(self._h))
            self._h = None


# =============================================================================================
# transcript and verifier (challenge.rs, scheme/verifier.rs)
# =============================================================================================

class Transcript:
    """challenge.rs:66-160."""

    def __init__(self, domain=None, _handle=None):
        if _handle is not None:
            self._h = _handle
        else:
            d = bytes(domain)
            self._h = lib().lbn_transcript_new(d, len(d))

    def clone(self):
        return Transcript(_handle=lib().lbn_transcript_clone(self._h))

    def absorb_bytes(self, b):
        b = bytes(b)
        lib().lbn_transcript_absorb_bytes(self._h, b, len(b))

    def absorb_u64(self, x):
        lib().lbn_transcript_absorb_u64(self._h, x)

    def fill(self, label, n):
        out = (ctypes.c_uint8 * n)()
        lb = bytes(label)
        lib().lbn_transcript_fill(self._h, lb, len(lb), out, n)
        return bytes(out)

    def sample_short_challenge(self, weight=28, bound=12.0):
        c = ShortChallenge()
        att = lib().lbn_sample_short_challenge(self._h, weight, bound, ctypes.byref(c))
        return c, att

    def __del__(self):
        if getattr(self, "_h", None) and _lib is not None:
            _lib.lbn_transcript_free(self._h)
            self._h = None


class Verifier:
    """scheme/verifier.rs:21-337 (the Clear and bit-dropped modes)."""

    def __init__(self, pp):
        self.pp = pp
        self.params = pp.params

    def derive_evaluation_point(self, transcript, commitment):
        p = self.params
        p0 = np.zeros((p.row_log_len(), 3), np.uint64)
        p1 = np.zeros((p.column_log_len, 3), np.uint64)
        if p.dropped_bits() > 0:
            d = commitment.dropped
            _check(lib().lbn_derive_evaluation_point_dropped(ctypes.c_void_p(self.pp._h),
                                                             ctypes.c_void_p(transcript._h), _p(d.top),
                                                             _p(d.digits), _p(p0), _p(p1)))
            return EvaluationPoint(p0, p1)
        comp = np.ascontiguousarray(commitment.components, np.int16)
        _check(lib().lbn_derive_evaluation_point(ctypes.c_void_p(self.pp._h), ctypes.c_void_p(transcript._h),
                                                 _p(comp), _p(p0), _p(p1)))
        return EvaluationPoint(p0, p1)

    def derive_folding_challenges(self, transcript, row_evaluation):
        u = np.ascontiguousarray(row_evaluation.values, np.uint64)
        arr = (ShortChallenge * len(u))()
        _check(lib().lbn_derive_folding_challenges(ctypes.c_void_p(self.pp._h), ctypes.c_void_p(transcript._h),
                                                   _p(u), ctypes.c_size_t(len(u)), arr))
        return FoldingChallenges(list(arr))

    def fold_commitment(self, commitment, challenges, device=None):
        """fold.rs:799-873; with device (a Prover of the same parameters) on its GPU."""
        p = self.params
        out = np.zeros((4, len(p.primes()), N162), np.int16)
        comp = np.ascontiguousarray(commitment.components, np.int16)
        arr = challenges.array()
        if device is not None:
            _check(lib().lbn_fold_commitment_device(ctypes.c_void_p(device._h), _p(comp), arr,
                                                    ctypes.c_size_t(len(arr)), _p(out)))
        else:
            _check(lib().lbn_fold_commitment(ctypes.c_void_p(self.pp._h), _p(comp), arr, ctypes.c_size_t(len(arr)),
                                             _p(out)))
        return FoldedCommitment(p.primes(), out)

    def fold_row_evaluation(self, row_evaluation, challenges):
        u = np.ascontiguousarray(row_evaluation.values, np.uint64)
        out = np.zeros(3, np.uint64)
        arr = challenges.array()
        _check(lib().lbn_fold_row_evaluation(_p(u), arr, ctypes.c_size_t(len(arr)), _p(out)))
        return out

    def verify_evaluation(self, point, claimed_value, row_evaluation):
        p1 = np.ascontiguousarray(point.p1, np.uint64)
        t = np.ascontiguousarray(claimed_value, np.uint64)
        u = np.ascontiguousarray(row_evaluation.values, np.uint64)
        rc = lib().lbn_verify_evaluation(ctypes.c_void_p(self.pp._h), _p(p1), ctypes.c_size_t(len(p1)), _p(t), _p(u),
                                         ctypes.c_size_t(len(u)))
        if rc == LBN_REJECTED:
            raise VerificationError("Rejected")
        _check(rc)

    def bd_residual(self, commitment, challenges, folded_witness):
        """bd/residual.rs:77-130 as a Python int."""
        d = commitment.dropped
        v = np.ascontiguousarray(folded_witness.elements, np.int16)
        arr = challenges.array()
        lo, hi = ctypes.c_uint64(), ctypes.c_uint64()
        _check(lib().lbn_bd_residual(ctypes.c_void_p(self.pp._h), _p(d.top), _p(d.digits), arr,
                                     ctypes.c_size_t(len(arr)), _p(v), ctypes.c_size_t(len(v)), ctypes.byref(lo),
                                     ctypes.byref(hi)))
        return lo.value | (hi.value << 64)

    def verify_opening(self, commitment, challenges, point, folded_commitment, folded_witness, folded_row_value,
                       device=None):
        """Verifier::verify_opening (verifier.rs:281-337): the Clear message carries the folded
        commitment; the bit-dropped one does not (pass None). With device (a Prover of the same
        parameters) the Clear checks run on its GPU."""
        if self.params.dropped_bits() > 0:
            return self._verify_opening_bd(commitment, challenges, point, folded_witness, folded_row_value)
        fc = np.ascontiguousarray(folded_commitment.rows, np.int16)
        v = np.ascontiguousarray(folded_witness.elements, np.int16)
        p0 = np.ascontiguousarray(point.p0, np.uint64)
        fr = np.ascontiguousarray(folded_row_value, np.uint64)
        if fc.shape != (4, len(self.params.primes()), N162):
            raise VerificationError("Rejected")
        if device is not None:
            rc = lib().lbn_verify_opening_clear_device(ctypes.c_void_p(device._h), _p(fc), _p(v),
                                                       ctypes.c_size_t(len(v)), _p(p0), ctypes.c_size_t(len(p0)),
                                                       _p(fr))
        else:
            rc = lib().lbn_verify_opening_clear(ctypes.c_void_p(self.pp._h), _p(fc), _p(v), ctypes.c_size_t(len(v)),
                                                _p(p0), ctypes.c_size_t(len(p0)), _p(fr))
        if rc == LBN_REJECTED:
            raise VerificationError("Rejected")
        _check(rc)


    def _verify_opening_bd(self, commitment, challenges, point, folded_witness, folded_row_value):
        d = commitment.dropped
        if d is None:
            raise VerificationError("Rejected")
        v = np.ascontiguousarray(folded_witness.elements, np.int16)
        p0 = np.ascontiguousarray(point.p0, np.uint64)
        fr = np.ascontiguousarray(folded_row_value, np.uint64)
        arr = challenges.array()
        rc = lib().lbn_verify_opening_bit_dropped(ctypes.c_void_p(self.pp._h), _p(d.top), _p(d.digits), arr,
                                                  ctypes.c_size_t(len(arr)), _p(v), ctypes.c_size_t(len(v)),
                                                  _p(p0), ctypes.c_size_t(len(p0)), _p(fr))
        if rc == LBN_REJECTED:
            raise VerificationError("Rejected")
        _check(rc)


def claim(u, p1):
    """eval.rs:185-193: sum_j eq(p1)_j u_j."""
    u = np.ascontiguousarray(u, np.uint64).reshape(-1, 3)
    p1 = np.ascontiguousarray(p1, np.uint64).reshape(-1, 3)
    t = np.zeros(3, np.uint64)
    _check(lib().lbn_claim(_p(u), ctypes.c_size_t(len(u)), _p(p1), ctypes.c_size_t(len(p1)), _p(t)))
    return t


def bd_cap(params):
    """Params::bd_cap (scheme/params.rs:250-252)."""
    return int(lib().lbn_bd_cap(ctypes.byref(params.c())))


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
    _check(lib().lbn_blake3(mode, k, buf, ctypes.c_size_t(len(data)), out, ctypes.c_size_t(out_len)))
    return bytes(out)


def f162_mul(a, b):
    a = np.ascontiguousarray(a, np.uint64)
    b = np.ascontiguousarray(b, np.uint64)
    out = np.zeros(3, np.uint64)
    lib().lbn_f162_mul(_p(a), _p(b), _p(out))
    return out


def device_available():
    try:
        import torch
        return torch.cuda.is_available()
    except Exception:
        return False
