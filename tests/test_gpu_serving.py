"""The serving loop (the native lbn_server through labinius_amd.serving) at its workload: sizes 2^16
to 2^24 under one key, Clear and bit-dropped, in bursts, so that requests join batches while others
are on the GPU. Every request's result is the same in the mix as alone (the serial server, one at a
time), the verifier accepts each, and sampled outputs equal the oracle; admission holds the budget;
an oversize request is refused at submit; a request made to fail from the host fails alone; the
device pool stops allocating once warm."""
import sys

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

MATRIX_SEED = bytes([0x5A]) * 32
SIZES = (16, 18, 20, 22, 24)


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module", autouse=True)
def _device():
    from tests.conftest import gpu_available
    if not gpu_available():
        pytest.skip("no GPU")


def _stream(lbn, exps, seed=7, burst=8, gap=0.0005, tag=b"labinius/test/serving"):
    from labinius_amd import workloads as wl
    rng = wl.Rng(seed)
    out = []
    for i, e in enumerate(exps):
        d = wl.SERVING_DROPPED[e] if rng.below(2) else 0
        w = lbn.Witness.random(wl.witness_params(e), wl.witness_seed(tag, i))
        out.append((i, e, d, w, 0.0 if i < burst else gap * i))
    return out


def _verify(lbn, r):
    p = r.params()
    v = lbn.Verifier(lbn.PublicParameters.host_only(p, MATRIX_SEED))
    res = r.result
    fr = v.fold_row_evaluation(res.row, res.challenges)
    v.verify_evaluation(res.point, res.claimed, res.row)
    fc = None if p.dropped_bits() else v.fold_commitment(res.commitment, res.challenges)
    v.verify_opening(res.commitment, res.challenges, res.point, fc, res.folded, fr)


def test_results_are_the_same_alone_and_in_the_mix(lbn, orc):
    from labinius_amd import serving
    exps = [16, 24, 18, 16, 22, 20, 16, 18, 16, 20, 18, 22, 16, 16, 18]
    reqs = _stream(lbn, exps)
    assert {d > 0 for _, _, d, _, _ in reqs} == {True, False}
    say(f"case: serving mix of {len(reqs)}: sizes {exps}")
    server = serving.Server(MATRIX_SEED, sizes=SIZES)
    try:
        mixed = serving.run_stream(server, reqs)
        stats = server.stats()
    finally:
        server.close()
    assert all(r.status == 0 for r in mixed), [(r.id, r.status) for r in mixed]
    # requests joined batches while others ran: fewer batched commitments than requests
    assert stats["batches"] < len(reqs), stats
    say(f"served: {stats['batches']} batches for {len(reqs)} requests")
    alone = {r.id: r for r in serving.serve_unbatched(MATRIX_SEED, reqs, sizes=SIZES)}
    assert all(r.status == 0 for r in alone.values())
    for r in mixed:
        assert r.result.same_as(alone[r.id].result), r.id
        _verify(lbn, r)
    say("alone = mix; the verifier accepts every request")
    # the oracle: a 2^16 request end to end, and sampled columns and fold rows of the 2^24 one
    for r in (next(x for x in mixed if x.e == 16), next(x for x in mixed if x.e == 24)):
        p = r.params()
        primes, L, cols, clen = p.primes(), p.len_ring(), p.columns(), p.column_len()
        a = orc.key_expand(primes, L, orc.matrix_seed(MATRIX_SEED))
        sample = list(range(cols)) if cols <= 32 else [0, 1, cols // 2, cols - 1]
        for c in sample:
            raw, comp = orc.commit(primes, L, 1, a, r.witness.elements[c * clen:(c + 1) * clen])
            assert np.array_equal(r.result.commitment.raw[:, c], raw[:, 0]), (r.id, c)
            assert np.array_equal(r.result.commitment.components[c], comp[0]), (r.id, c)
            u = orc.row_evaluate(r.witness.elements[c * clen:(c + 1) * clen], r.result.point.p0)
            assert np.array_equal(r.result.row.values[c], u[0]), (r.id, c)
        rows = np.array([0, 1, L // 2, L - 1])
        assert np.array_equal(r.result.folded.elements[rows],
                              orc.fold_rows(primes[0], L, cols, r.witness.elements, r.result.challenges.challenges,
                                            rows)), r.id
    say("sampled outputs equal the oracle")


def test_admission_holds_the_budget_and_the_pool_warms(lbn):
    from labinius_amd import serving
    from labinius_amd import workloads as wl
    say("case: serving budget")
    sizes = (16, 18)
    reqs = _stream(lbn, [16, 18, 16, 16, 18, 16, 18, 16], seed=11, tag=b"labinius/test/budget")
    server = serving.Server(MATRIX_SEED, sizes=sizes)
    try:
        fixed = server.stats()["fixed_bytes"]
        one = serving.request_bytes(wl.witness_params(18))
        server.set_budget(fixed + one + serving.request_bytes(wl.witness_params(16)))
        done = serving.run_stream(server, reqs)
        first = server.stats()
        again = serving.run_stream(server, reqs)
        second = server.stats()
    finally:
        server.close()
    assert all(r.status == 0 for r in done + again)
    assert 0 < first["reserved_peak"] <= server.budget
    # at most one 2^18 and one 2^16 were held at once: the largest sum of the bytes held at one time
    for rs in (done, again):
        events = sorted([(r.admitted, r.bytes) for r in rs] + [(r.done, -r.bytes) for r in rs],
                        key=lambda ev: (ev[0], ev[1]))
        held = peak = 0
        for _, b in events:
            held += b
            peak = max(peak, held)
        assert peak <= server.budget - fixed, peak
    # the second pass found every buffer in the pool
    assert second["pool_mallocs"] == first["pool_mallocs"], (first, second)
    assert second["pool_reuses"] > first["pool_reuses"]


def test_oversize_request_is_rejected_at_submit(lbn):
    from labinius_amd import serving
    from labinius_amd import workloads as wl
    say("case: serving oversize")
    server = serving.Server(MATRIX_SEED, sizes=(16,), workers=2)
    try:
        fixed = server.stats()["fixed_bytes"]
        server.set_budget(fixed + 4 * serving.request_bytes(wl.witness_params(16)))
        w20 = lbn.Witness.random(wl.witness_params(20), bytes([3]) * 32)
        with pytest.raises(serving.AdmissionError):
            server.submit(1, 20, 0, w20)
        w = lbn.Witness.random(wl.witness_params(16), bytes([4]) * 32)
        r = server.wait(server.submit(2, 16, 0, w))
        assert r.status == 0
        other = lbn.Params(16, 4, [lbn.Modulus.Q9721_FS_S])  # row length 2^12, not the key's 2^11
        with pytest.raises(lbn.LabiniusError) as e:
            server.submit(3, 16, 0, lbn.Witness.random(other, bytes([5]) * 32), params=other)
        assert e.value.code == 8  # LBN_ERR_SHAPE
    finally:
        server.close()


def test_a_failing_request_fails_alone(lbn):
    from labinius_amd import serving
    say("case: serving failure isolation")
    exps = [16, 18, 16, 20, 16, 18]
    reqs = _stream(lbn, exps, seed=13, tag=b"labinius/test/failure")
    reqs = [req + ((i == 2),) for i, req in enumerate(reqs)]  # request 2's fold is made to fail
    server = serving.Server(MATRIX_SEED, sizes=(16, 18, 20), workers=3)
    try:
        done = serving.run_stream(server, reqs)
        after = server.wait(server.submit(99, 16, 0, reqs[0][3]))
        stats = server.stats()
    finally:
        server.close()
    assert done[2].status == 12 and done[2].result is None  # LBN_ERR_ARG: the challenge's weight
    assert all(r.status == 0 for i, r in enumerate(done) if i != 2)
    assert after.status == 0 and stats["failed"] == 1
    for i, r in enumerate(done):
        if i != 2:
            _verify(lbn, r)
    alone = {r.id: r for r in serving.serve_unbatched(MATRIX_SEED, [q[:5] for q in reqs], sizes=(16, 18, 20))}
    for i, r in enumerate(done):
        if i != 2:
            assert r.result.same_as(alone[r.id].result), r.id
