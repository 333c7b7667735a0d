"""The cross-field switch (switch.rs): a claim on the multilinear extension of a B128 trace (or of
its bits) turned into the F162 claim that the commitment to Witness::lifted(trace) opens, at the
point where the switch's sumcheck ends. The prover runs on the GPU over the lifted trace resident
on a prover (Prover.upload); the verifier runs on the host. B128 values are uint64 pairs, low word
first; F162 values uint64 triples.
"""
import ctypes

import numpy as np

from . import EvaluationPoint, LabiniusError, LBN_REJECTED, _p, lib

LOG_WORD = 7


class SwitchProof:
    """switch.rs:18-26: the 128 partial evaluations, the sumcheck's rounds [e0, einf], and the
    value of the lifted trace at the point the switch ends on."""

    def __init__(self, partials, rounds, opened):
        self.partials = np.ascontiguousarray(partials, np.uint64).reshape(128, 2)
        self.rounds = np.ascontiguousarray(rounds, np.uint64).reshape(-1, 2, 3)
        self.opened = np.ascontiguousarray(opened, np.uint64).reshape(3)


class SwitchError(Exception):
    """switch.rs:28-37."""


def _b128(a, n=None):
    a = np.ascontiguousarray(np.asarray(a, np.uint64).reshape(-1, 2))
    if n is not None and len(a) != n:
        raise ValueError(f"{n} B128 values expected")
    return a


def _prove(prover, witness, r_lo, r_hi, claim, transcript):
    r_hi = _b128(r_hi)
    r_lo = _b128(r_lo) if r_lo is not None else np.zeros((0, 2), np.uint64)
    claim = _b128(claim, 1)
    l = len(r_hi)
    cl = witness.params.column_log_len
    partials = np.zeros((128, 2), np.uint64)
    rounds = np.zeros((l, 2, 3), np.uint64)
    opened = np.zeros(3, np.uint64)
    p0 = np.zeros((l - cl, 3), np.uint64)
    p1 = np.zeros((cl, 3), np.uint64)
    rc = lib().lbn_switch_prove(ctypes.c_void_p(prover._h), ctypes.c_void_p(witness._h), _p(r_lo),
                                ctypes.c_size_t(len(r_lo)), _p(r_hi), ctypes.c_size_t(l), _p(claim),
                                ctypes.c_void_p(transcript._h), _p(partials), _p(rounds), _p(opened), _p(p0), _p(p1))
    if rc != 0:
        raise LabiniusError(rc, lib().lbn_last_error().decode())
    return SwitchProof(partials, rounds, opened), EvaluationPoint(p0, p1)


def prove(prover, witness, r, claim, transcript):
    """switch.rs:63-72: prove mle(trace, r) == claim for the lifted trace resident as witness."""
    return _prove(prover, witness, None, r, claim, transcript)


def prove_bits(prover, witness, r_lo, r_hi, claim, transcript):
    """switch.rs:88-100: prove bit_mle(trace, r_lo, r_hi) == claim."""
    if len(_b128(r_lo)) != LOG_WORD:
        raise AssertionError("r_lo picks a bit: 7 coordinates")
    return _prove(prover, witness, r_lo, r_hi, claim, transcript)


def _verify(params, r_lo, r_hi, claim, proof, transcript):
    r_hi = _b128(r_hi)
    r_lo = _b128(r_lo) if r_lo is not None else np.zeros((0, 2), np.uint64)
    claim = _b128(claim, 1)
    l, cl = len(r_hi), params.column_log_len
    if l != params.witness_log_len:
        raise SwitchError("the point does not match the parameters")
    if len(proof.rounds) != l:
        raise SwitchError("the proof does not match the parameters")
    p0 = np.zeros((l - cl, 3), np.uint64)
    p1 = np.zeros((cl, 3), np.uint64)
    rc = lib().lbn_switch_verify(ctypes.byref(params.c()), _p(r_lo), ctypes.c_size_t(len(r_lo)), _p(r_hi),
                                 ctypes.c_size_t(l), _p(claim), _p(proof.partials), _p(proof.rounds),
                                 _p(proof.opened), ctypes.c_void_p(transcript._h), _p(p0), _p(p1))
    if rc == LBN_REJECTED:
        raise SwitchError(lib().lbn_last_error().decode())
    if rc != 0:
        raise LabiniusError(rc, lib().lbn_last_error().decode())
    return EvaluationPoint(p0, p1), proof.opened.copy()


def verify(params, r, claim, proof, transcript):
    """switch.rs:76-86: the point at which the commitment must open, and the value it must open to."""
    return _verify(params, None, r, claim, proof, transcript)


def verify_bits(params, r_lo, r_hi, claim, proof, transcript):
    """switch.rs:103-121."""
    if len(_b128(r_lo)) != LOG_WORD:
        raise SwitchError("the point does not pick a bit")
    return _verify(params, r_lo, r_hi, claim, proof, transcript)
