"""The wire forms (wire/mod.rs, wire/rans.rs): what a verifier receives, and its exact inverse.

Uniform data is bit-packed: a commitment slot at ceil(log2 q) bits, an F162 at 162 bits. The
folded witness is entropy-coded with a 64-lane interleaved static rANS against its own quantised
histogram. The prover's side runs on the GPU over what the commit and the fold left in HBM
(pack_commitment, encode); the verifier's side runs on the host (unpack_commitment, decode).
"""
import ctypes

import numpy as np

from . import (Dropped, FoldedWitness, LabiniusError, RowEvaluation, _check, _p, lib, N)

N162 = 162
LANES = 64
F162_BITS = 162
LBN_WIRE_TRUNCATED = 15
LBN_WIRE_MALFORMED = 16


class WireError(Exception):
    """wire/mod.rs:67-74: kind is "Truncated" (the bytes ran out before the object did) or
    "Malformed" (not a coding of any object of the expected shape)."""

    def __init__(self, kind, msg=""):
        super().__init__(f"{kind}: {msg}" if msg else kind)
        self.kind = kind

    def __eq__(self, other):
        return isinstance(other, WireError) and other.kind == self.kind

    def __hash__(self):
        return hash(self.kind)


Truncated = WireError("Truncated")
Malformed = WireError("Malformed")


def _wire_check(rc):
    if rc == LBN_WIRE_TRUNCATED:
        raise WireError("Truncated", lib().lbn_last_error().decode())
    if rc == LBN_WIRE_MALFORMED:
        raise WireError("Malformed", lib().lbn_last_error().decode())
    _check(rc)


def _buf(data):
    data = bytes(data)
    return data, np.frombuffer(data + b"\0", np.uint8)


def residue_bits(q):
    """wire/mod.rs:77-79: ceil(log2 q)."""
    return (int(q) - 1).bit_length()


def commitment_bytes(params):
    """Commitment::wire_bytes: wire/mod.rs:473-476, or bd/mod.rs:70-72 for bit-dropped parameters."""
    n = ctypes.c_size_t()
    _check(lib().lbn_commitment_wire_bytes(ctypes.byref(params.c()), ctypes.byref(n)))
    return n.value


def _sized(call):
    n = ctypes.c_size_t()
    _check(call(None, 0, ctypes.byref(n)))
    out = np.zeros(max(1, n.value), np.uint8)
    _check(call(_p(out), out.size, None))
    return bytes(out[:n.value])


def pack_commitment(prover):
    """wire/mod.rs:226-243 (and 386-397 in the bit-dropped mode) of prover's last commitment, packed
    on the GPU from the matrix the commit left in HBM."""
    h = ctypes.c_void_p(prover._h)
    return _sized(lambda out, cap, n: lib().lbn_commit_pack(h, out, ctypes.c_size_t(cap), n))


def pack_commitment_host(params, components):
    """wire/mod.rs:226-243 of host components [columns][4][limb][162] int16."""
    comp = np.ascontiguousarray(components, np.int16)
    if comp.size != params.columns() * 4 * len(params.primes()) * N162:
        raise ValueError("components do not match the parameters")
    c = params.c()
    return _sized(lambda out, cap, n: lib().lbn_pack_commitment(ctypes.byref(c), _p(comp), out,
                                                                ctypes.c_size_t(cap), n))


def unpack_commitment(params, data):
    """wire/mod.rs:246-301: the components [columns][4][limb][162] int16, or Dropped for bit-dropped
    parameters. Raises WireError."""
    data, buf = _buf(data)
    if params.dropped_bits() > 0:
        top, digits = Dropped._empty(params)
        _wire_check(lib().lbn_unpack_dropped(ctypes.byref(params.c()), _p(buf), ctypes.c_size_t(len(data)),
                                             _p(top), _p(digits)))
        return Dropped(params, top, digits)
    comp = np.zeros((params.columns(), 4, len(params.primes()), N162), np.int16)
    _wire_check(lib().lbn_unpack_commitment(ctypes.byref(params.c()), _p(buf), ctypes.c_size_t(len(data)),
                                            _p(comp)))
    return comp


def f162_bytes(count):
    """wire/mod.rs:473-476 (f162_bytes)"""
    return (count * F162_BITS + 7) // 8


def pack_f162(values):
    """wire/mod.rs:445-453"""
    v = np.ascontiguousarray(np.asarray(values, np.uint64).reshape(-1, 3))
    return _sized(lambda out, cap, n: lib().lbn_pack_f162(_p(v), ctypes.c_size_t(len(v)), out,
                                                          ctypes.c_size_t(cap), n))


def unpack_f162(data, count):
    """wire/mod.rs:456-470"""
    data, buf = _buf(data)
    v = np.zeros((count, 3), np.uint64)
    _wire_check(lib().lbn_unpack_f162(_p(buf), ctypes.c_size_t(len(data)), ctypes.c_size_t(count), _p(v)))
    return v


def pack_row_evaluation(row):
    """wire/mod.rs:434-436"""
    return pack_f162(row.values)


def unpack_row_evaluation(data, count):
    """wire/mod.rs:439-442"""
    return RowEvaluation(unpack_f162(data, count))


def _bounded(count, call):
    """One call into a buffer of the encoder's worst-case size (the size is known only after
    encoding, so asking for it first would encode twice)."""
    lib().lbn_wire_encode_bound.restype = ctypes.c_size_t
    out = np.empty(lib().lbn_wire_encode_bound(ctypes.c_size_t(count)), np.uint8)
    n = ctypes.c_size_t()
    _check(call(_p(out), out.size, ctypes.byref(n)))
    return bytes(out[:n.value])


def encode(prover, base_q):
    """rans.rs:18-99 of prover's last folded witness, on the GPU over the copy the fold left in HBM."""
    h = ctypes.c_void_p(prover._h)
    return _bounded(prover.params.len_ring(),
                    lambda out, cap, n: lib().lbn_fold_encode(h, ctypes.c_uint16(base_q), out, ctypes.c_size_t(cap), n))


def encode_message(prover, folded, base_q, ntt=False):
    """rans.rs:18-99 of a host message ([count][648] int16, or a FoldedWitness) on prover's GPU."""
    v = folded.elements if isinstance(folded, FoldedWitness) else folded
    v = np.ascontiguousarray(np.asarray(v, np.int16).reshape(-1, N))
    h = ctypes.c_void_p(prover._h)
    return _bounded(len(v), lambda out, cap, n: lib().lbn_wire_encode_device(h, _p(v), ctypes.c_size_t(len(v)),
                                                                             ctypes.c_int(int(ntt)),
                                                                             ctypes.c_uint16(base_q), out,
                                                                             ctypes.c_size_t(cap), n))


def decode(data):
    """rans.rs:101-196, on the host: the FoldedWitness. Raises WireError."""
    data, buf = _buf(data)
    count = int.from_bytes(data[4:8], "little") if len(data) >= 16 else 0
    cap = count if count <= 1 << 24 else 0
    v = np.zeros((max(cap, 1), N), np.int16)
    cnt, ntt = ctypes.c_uint32(), ctypes.c_int()
    _wire_check(lib().lbn_wire_decode(_p(buf), ctypes.c_size_t(len(data)), _p(v), ctypes.c_size_t(cap),
                                      ctypes.byref(cnt), ctypes.byref(ntt)))
    return FoldedWitness(v[:cnt.value])


def timings(prover):
    """Device ms of prover's last encode: total, statistics, the rANS kernel."""
    ms = (ctypes.c_float * 3)()
    _check(lib().lbn_wire_timings(ctypes.c_void_p(prover._h), ms, ctypes.c_size_t(3)))
    return list(ms)


def entropy_bytes(v):
    """rans.rs:500-524: the zeroth-order entropy of a message in bytes."""
    v = np.asarray(v.elements if isinstance(v, FoldedWitness) else v, np.int16).ravel()
    if v.size == 0:
        return 0.0
    _, counts = np.unique(v, return_counts=True)
    p = counts / v.size
    return float(-(counts * np.log2(p)).sum() / 8.0)


__all__ = ["WireError", "Truncated", "Malformed", "LANES", "residue_bits", "commitment_bytes", "pack_commitment",
           "pack_commitment_host", "unpack_commitment", "pack_f162", "unpack_f162", "pack_row_evaluation",
           "unpack_row_evaluation", "f162_bytes", "encode", "encode_message", "decode", "timings", "entropy_bytes",
           "LabiniusError"]
