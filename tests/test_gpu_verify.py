"""The verifier's heavy steps on the GPU (lbn_verify_opening_clear_device, lbn_fold_commitment_device)
against the host verifier and the oracle: the folded commitment equals both bit for bit, honest
openings are accepted, and every tampering the host verifier rejects (a coefficient, the row
value, a folded-commitment slot, a coefficient past the bound, a norm past the cap) is rejected.
A sparse v with coefficients at +-half, whose A v comes from the oracle, checks the bit-plane
decomposition at its extremes."""
import sys

import numpy as np
import pytest

from tests.test_verify_planes import lift_bits

pytestmark = pytest.mark.gpu

N = 648
MATRIX_SEED = bytes([0x33]) * 32
WITNESS_SEED = bytes([0x44]) * 32


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module", autouse=True)
def _device():
    from tests.conftest import gpu_available
    if not gpu_available():
        pytest.skip("no GPU")


def _accepts(verifier, prover, args, device):
    from labinius_amd import VerificationError
    try:
        verifier.verify_opening(*args, device=prover if device else None)
        return True
    except VerificationError:
        return False


@pytest.mark.parametrize("wl,cl,base,extra", [(11, 3, 3889, [9721]), (12, 2, 2917, [3889, 19441]),
                                              (10, 1, 19441, [4861]), (14, 5, 4861, [12637, 17497])])
def test_device_verifier_matches_host(lbn, orc, wl, cl, base, extra):
    say(f"case: device verifier wl={wl} cl={cl} base={base} extra={extra}")
    params = lbn.Params.with_base(wl, cl, lbn.Modulus(base), [lbn.Modulus(q) for q in extra])
    pp = lbn.PublicParameters.from_seed(params, MATRIX_SEED)
    prover, verifier = lbn.Prover(pp), lbn.Verifier(pp)
    w = lbn.Witness.random(params, WITNESS_SEED)
    commitment, opening = prover.commit(w)
    t = lbn.Transcript(b"labinius/test/verify")
    point = verifier.derive_evaluation_point(t, commitment)
    row = prover.row_evaluate(w, point)
    challenges = verifier.derive_folding_challenges(t, row)
    folded = prover.fold(opening, challenges)

    fc_host = verifier.fold_commitment(commitment, challenges)
    fc_dev = verifier.fold_commitment(commitment, challenges, device=prover)
    assert np.array_equal(fc_dev.rows, fc_host.rows)
    assert np.array_equal(fc_dev.rows, orc.fold_commitment(params.primes(), params.columns(), commitment.components,
                                                           challenges.challenges))
    fr = verifier.fold_row_evaluation(row, challenges)
    args = [commitment, challenges, point, fc_host, folded, fr]
    assert _accepts(verifier, prover, args, True) and _accepts(verifier, prover, args, False)

    bad = lbn.FoldedWitness(folded.elements.copy())
    bad.elements[3, 7] += 2
    for i, tampered in enumerate([
            [commitment, challenges, point, fc_host, bad, fr],
            [commitment, challenges, point, fc_host, folded, fr ^ np.array([1, 0, 0], np.uint64)],
            [commitment, challenges, point, lbn.FoldedCommitment(params.primes(), fc_host.rows + 0), folded, fr]]):
        if i == 2:
            tampered[3].rows[1, 0, 5] += 1
        assert not _accepts(verifier, prover, tampered, True), i
        assert not _accepts(verifier, prover, tampered, False), i
    q0 = params.primes()[0]
    far = lbn.FoldedWitness(folded.elements.copy())
    far.elements[0, 0] = (q0 - 1) // 2 + 1
    assert not _accepts(verifier, prover, [commitment, challenges, point, fc_host, far, fr], True)
    heavy = lbn.FoldedWitness(np.full_like(folded.elements, min(200, (q0 - 1) // 2)))
    assert not _accepts(verifier, prover, [commitment, challenges, point, fc_host, heavy, fr], True)


def test_device_verifier_at_the_coefficient_bound(lbn, orc):
    # v zero but for coefficients at +-half: A v from the oracle, the row value from v mod 2 (2^16
    # witness elements, so that four coefficients at half stay under the norm cap)
    say("case: device verifier, coefficients at the bound")
    params = lbn.Params.with_base(16, 4, lbn.Modulus.Q19441_FS_L, [lbn.Modulus.Q3889_FS_S])
    pp = lbn.PublicParameters.from_seed(params, MATRIX_SEED)
    prover, verifier = lbn.Prover(pp), lbn.Verifier(pp)
    primes, L = params.primes(), params.len_ring()
    half = (primes[0] - 1) // 2
    v = np.zeros((L, N), np.int16)
    v[0, 0], v[1, 77], v[L - 1, 647], v[5, 300], v[6, 301] = half, -half, half, -half, 1
    a = orc.key_expand(primes, L, orc.matrix_seed(MATRIX_SEED))
    fc = lbn.FoldedCommitment(primes, orc.a_times_v(primes, L, a, v))
    rng = np.random.default_rng(8)
    p0 = rng.integers(0, 2**63, (params.row_log_len(), 3), dtype=np.uint64)
    p0[:, 2] &= np.uint64((1 << 34) - 1)
    fr = orc.row_evaluate(lift_bits(v.astype(np.int64) & 1), p0)[0]
    point = lbn.EvaluationPoint(p0, np.zeros((params.column_log_len, 3), np.uint64))
    args = [None, None, point, fc, lbn.FoldedWitness(v), fr]
    assert _accepts(verifier, prover, args, True) and _accepts(verifier, prover, args, False)
    fc.rows[2, 1, 100] += 1
    assert not _accepts(verifier, prover, args, True) and not _accepts(verifier, prover, args, False)
