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


def _times_x(v, k):
    """v x^k in B128 = GF(2)[x] / (x^128 + x^7 + x^2 + x + 1) (fields/scalar.rs:41-63), as ints."""
    v <<= k
    for i in range(v.bit_length() - 1, 127, -1):
        if (v >> i) & 1:
            v ^= (1 << i) | (0x87 << (i - 128))
    return v


def test_batched_lifted_traces_open_through_the_switch(lbn, orc):
    """The batched configuration's B128 quarter (switch.rs:132-164, examples/two_fields.rs:72-221): every
    lifted trace of the 64-witness batch (2^16 .. 2^22 words) is committed in the batch, its word claim
    at a random point goes through the switch on the GPU at its own size, the host verifier accepts the
    switch, the batched commitment opens at the switch's point to the switch's value (the row
    evaluation's MLE, an independent path), and the whole opening verifies. The smallest trace's claim
    is also the oracle's B128 MLE, and its switch proof equals the oracle's."""
    from labinius_amd import switch
    from labinius_amd import workloads as wl
    key = wl.key_params()
    committer = lbn.Prover(lbn.PublicParameters.from_seed(key, MATRIX_SEED))
    specs = wl.batch_specs()
    lifted = [(i, e) for i, (e, lf) in enumerate(specs) if lf]
    ws = [wl.batch_witness(i, e, True) for i, e in lifted]
    say(f"case: batched lifted traces, sizes {sorted(e for _, e in lifted)}")
    comms = committer.commit_batch([committer.upload_params(w) for w in ws])
    rng = np.random.default_rng(0x5317)
    provers = {}
    smallest = min(range(len(ws)), key=lambda k: lifted[k][1])
    for k, ((i, e), w, comm) in enumerate(zip(lifted, ws, comms)):
        params = w.params
        if e not in provers:
            pp = lbn.PublicParameters.from_seed(params, MATRIX_SEED)
            provers[e] = (lbn.Prover(pp), lbn.Verifier(pp))
        prover, verifier = provers[e]
        rw = prover.upload(w)
        trace = np.ascontiguousarray(w.elements[:, :2])
        r = rng.integers(0, 2**63, (e, 2), dtype=np.uint64)
        first, _ = switch.prove(prover, rw, r, np.zeros(2, np.uint64), lbn.Transcript(b"labinius/test/bswitch"))
        c = 0
        for b, v in enumerate(first.partials):
            c ^= _times_x(int(v[0]) | (int(v[1]) << 64), b)
        claim = np.array([c & (2**64 - 1), c >> 64], np.uint64)
        if k == smallest:
            assert np.array_equal(claim, orc.b128_mle(trace, r))
        t = lbn.Transcript(b"labinius/test/bswitch")
        proof, point = switch.prove(prover, rw, r, claim, t)
        if k == smallest:
            ref = orc.switch_prove(trace, params.column_log_len, None, r, claim,
                                   orc.Transcript(b"labinius/test/bswitch"))
            assert np.array_equal(proof.partials, ref[0]) and np.array_equal(proof.rounds, ref[1])
            assert np.array_equal(proof.opened, ref[2])
        assert np.array_equal(prover.mle_evaluate(rw, point), proof.opened), e
        row = prover.row_evaluate(rw, point)
        challenges = verifier.derive_folding_challenges(t, row)
        folded = prover.fold_resident(rw, challenges)
        vt = lbn.Transcript(b"labinius/test/bswitch")
        vpoint, value = switch.verify(params, r, claim, proof, vt)
        assert np.array_equal(vpoint.p0, point.p0) and np.array_equal(vpoint.p1, point.p1)
        verifier.verify_evaluation(vpoint, value, row)
        vch = verifier.derive_folding_challenges(vt, row)
        verifier.verify_opening(comm, vch, vpoint, verifier.fold_commitment(comm, vch), folded,
                                verifier.fold_row_evaluation(row, vch))
    say("every lifted trace opened through the switch")
