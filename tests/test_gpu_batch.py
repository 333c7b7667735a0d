"""The batched commitment (lbn_commit_batch): every witness's result equals its own commitment,
checked against the oracle on small batches and against the single-witness path (itself
oracle-checked) on the full 64-witness configuration."""
import sys

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

MATRIX_SEED = bytes([0x5A]) * 32


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module", autouse=True)
def _device():
    from tests.conftest import gpu_available
    if not gpu_available():
        pytest.skip("no GPU")


def test_small_uneven_batch_matches_oracle(lbn, orc):
    from labinius_amd import workloads as wl
    say("case: small uneven batch")
    key = wl.key_params()
    pp = lbn.PublicParameters.from_seed(key, MATRIX_SEED)
    prover = lbn.Prover(pp)
    specs = [(12, False), (13, True), (14, False), (12, False), (15, True), (13, False)]
    ws = [wl.batch_witness(i, e, lifted) for i, (e, lifted) in enumerate(specs)]
    rws = [prover.upload_params(w) for w in ws]
    a = orc.key_expand(key.primes(), key.len_ring(), orc.matrix_seed(MATRIX_SEED))
    for rep in range(2):  # the second call replays the captured graph
        comms = prover.commit_batch(rws)
        for w, c in zip(ws, comms):
            raw, comp = orc.commit(key.primes(), key.len_ring(), w.params.columns(), a, w.elements)
            assert np.array_equal(c.raw, raw) and np.array_equal(c.components, comp)
    info = prover.batch_info()
    assert info["tiles_per_xcd_round"] in (4, 8, 16)
    # a narrower batch next: a new graph, the same answers
    sub = [rws[1], rws[4]]
    comms = prover.commit_batch(sub)
    for w, c in zip([ws[1], ws[4]], comms):
        raw, _ = orc.commit(key.primes(), key.len_ring(), w.params.columns(), a, w.elements)
        assert np.array_equal(c.raw, raw)
    # and each equals the witness's own prover
    for w in ws[:3]:
        own = lbn.Prover(lbn.PublicParameters.from_seed(w.params, MATRIX_SEED))
        c1, op = own.commit(w)
        own.recycle(op)
        raw, _ = orc.commit(key.primes(), key.len_ring(), w.params.columns(), a, w.elements)
        assert np.array_equal(c1.raw, raw)


def test_full_batch_matches_one_at_a_time(lbn):
    """64 witnesses of 2^16 .. 2^22 F162 (a quarter of them lifted B128 traces) under the basic key:
    the batch equals the witness-by-witness commitments bit for bit."""
    from labinius_amd import workloads as wl
    say("case: full batch")
    key = wl.key_params()
    pp = lbn.PublicParameters.from_seed(key, MATRIX_SEED)
    prover = lbn.Prover(pp)
    specs = wl.batch_specs()
    assert len(specs) == 64 and sum(lifted for _, lifted in specs) == 16
    ws = [wl.batch_witness(i, e, lifted) for i, (e, lifted) in enumerate(specs)]
    rws = [prover.upload_params(w) for w in ws]
    comms = prover.commit_batch(rws)
    say("batch done", prover.batch_info())
    singles = {}
    for w, c in zip(ws, comms):
        e = w.params.witness_log_len
        if e not in singles:
            singles[e] = lbn.Prover(lbn.PublicParameters.from_seed(w.params, MATRIX_SEED))
        own = singles[e]
        rw = own.upload(w)
        c1, op = own.commit_resident(rw)
        own.recycle(op)
        assert np.array_equal(c1.raw, c.raw), e
        assert np.array_equal(c1.components, c.components), e
