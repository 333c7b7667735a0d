"""The serving loop (labinius_amd.serving): every request's result is the same in the mix as alone,
the verifier accepts each, and admission keeps the resident witnesses under the HBM budget."""
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


def _stream(lbn, n, sizes, seed=7):
    from labinius_amd import workloads as wl
    rng = wl.Rng(seed)
    out = []
    for i in range(n):
        e = sizes[rng.below(len(sizes))]
        d = 8 if rng.below(2) else 0
        w = lbn.Witness.random(wl.witness_params(e), wl.witness_seed(b"labinius/test/serving", i))
        # a burst of 8, then one every 0.5 ms: the burst's later requests arrive while the first
        # one's batch is on the GPU and join the next batch
        out.append((i, e, d, w, 0.0 if i < 8 else 0.0005 * i))
    return out


def _verify(lbn, r):
    from labinius_amd import serving
    p = r.params()
    v = lbn.Verifier(lbn.PublicParameters.host_only(p, MATRIX_SEED))
    res = r.result
    fr = v.fold_row_evaluation(res.row, res.challenges)
    v.verify_evaluation(res.point, res.claimed, res.row)
    fc = None if p.dropped_bits() else v.fold_commitment(res.commitment, res.challenges)
    v.verify_opening(res.commitment, res.challenges, res.point, fc, res.folded, fr)


def test_results_are_the_same_alone_and_in_the_mix(lbn, orc):
    from labinius_amd import serving
    say("case: serving mix")
    sizes = (12, 13, 14)
    reqs = _stream(lbn, 16, sizes)
    server = serving.Server(MATRIX_SEED, sizes=sizes)
    try:
        mixed = serving.run_stream(server, reqs)
    finally:
        server.close()
    assert all(r.error is None for r in mixed), [r.error for r in mixed if r.error]
    assert server.reserved == 0
    # more than one request per batched commitment: requests joined batches while others ran
    assert len(server.batches) < len(reqs)
    alone = {r.id: r for r in serving.serve_unbatched(MATRIX_SEED, reqs, sizes=sizes)}
    for r in mixed:
        assert r.result.same_as(alone[r.id].result), r.id
        _verify(lbn, r)
    # and one of them against the oracle end to end
    r = mixed[3]
    p = r.params()
    a = orc.key_expand(p.primes(), p.len_ring(), orc.matrix_seed(MATRIX_SEED))
    raw, comp = orc.commit(p.primes(), p.len_ring(), p.columns(), a, r.witness.elements)
    assert np.array_equal(r.result.commitment.raw, raw)
    assert np.array_equal(r.result.folded.elements,
                          orc.fold(p.primes()[0], p.len_ring(), p.columns(), r.witness.elements,
                                   r.result.challenges.challenges))


def test_admission_holds_the_budget(lbn):
    from labinius_amd import serving
    say("case: serving budget")
    sizes = (12, 13)
    reqs = _stream(lbn, 8, sizes, seed=11)
    budget = (24 << 13) + (24 << 12)  # one 2^13 and one 2^12 witness at most
    server = serving.Server(MATRIX_SEED, hbm_budget_bytes=budget, sizes=sizes)
    try:
        done = serving.run_stream(server, reqs)
    finally:
        server.close()
    assert all(r.error is None for r in done)
    assert 0 < server.peak_reserved <= budget
    assert server.reserved == 0
