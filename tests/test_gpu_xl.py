"""The headline round at full size: Params::sized(SUITES[3], Clear) = 2^24 F162 in 1024 columns,
moduli 3889 / 2917 / 9721, with the reference bench's seeds and transcript domain
(crates/bench/src/bin/labinius.rs:12-14,104-254; scheme/params.rs:119-127).

Each column's commitment and row value, and each folded ring element, depends on its own slice of
the witness only, so sampled columns and rows are checked against the oracle bit for bit; the MLE
claim is checked from all of u, the point and the challenges from the transcript, and the host
verifier accepts the whole opening."""
import sys

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

MATRIX_SEED = bytes([0x5A]) * 32
WITNESS_SEED = bytes([0xC7]) * 32


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module", autouse=True)
def _device():
    from tests.conftest import gpu_available
    if not gpu_available():
        pytest.skip("no GPU")


def test_suite_xl_round_bit_exact(lbn, orc):
    say("case: suite xl clear (2^24, 1024 columns, 3889 / 2917 / 9721)")
    params = lbn.Params.sized(lbn.SUITES[3], lbn.Opening.CLEAR)
    assert params.witness_len() == 1 << 24 and params.columns() == 1024
    assert params.primes() == [3889, 2917, 9721]
    primes, L, cols, clen = params.primes(), params.len_ring(), params.columns(), params.column_len()
    pp = lbn.PublicParameters.from_seed(params, MATRIX_SEED)
    prover = lbn.Prover(pp)
    w = lbn.Witness.random(params, WITNESS_SEED)
    a = orc.key_expand(primes, L, orc.matrix_seed(MATRIX_SEED))
    rw = prover.upload(w)
    commitment, opening = prover.commit_resident(rw)
    say("committed")
    sample = [0, 1, 511, cols - 1]
    for c in sample:
        raw_c, comp_c = orc.commit(primes, L, 1, a, w.elements[c * clen:(c + 1) * clen])
        assert np.array_equal(commitment.raw[:, c], raw_c[:, 0]), c
        assert np.array_equal(commitment.components[c], comp_c[0]), c
    verifier = lbn.Verifier(pp)
    t = lbn.Transcript(b"labinius/reference")
    point = verifier.derive_evaluation_point(t, commitment)
    ot = orc.Transcript(b"labinius/reference")
    op0, op1 = orc.derive_evaluation_point(ot, orc.make_params(24, 10, primes, MATRIX_SEED), commitment.components)
    assert np.array_equal(point.p0, op0) and np.array_equal(point.p1, op1)
    assert len(point.p0) == 14 and len(point.p1) == 10
    row = prover.row_evaluate(rw, point)
    for c in sample:
        assert np.array_equal(row.values[c], orc.row_evaluate(w.elements[c * clen:(c + 1) * clen], point.p0)[0]), c
    claimed = prover.mle_evaluate(rw, point)
    assert np.array_equal(claimed, orc.claim(row.values, point.p1))
    say("evaluated")
    challenges = verifier.derive_folding_challenges(t, row)
    och = orc.derive_folding_challenges(ot, row.values)
    assert all(bytes(x.positions) == bytes(y.positions) and x.signs == y.signs
               for x, y in zip(challenges.challenges, och))
    folded = prover.fold(opening, challenges)
    rows = np.array([0, 1, 2048, L - 1])
    assert np.array_equal(folded.elements[rows], orc.fold_rows(primes[0], L, cols, w.elements, och, rows))
    say("folded")
    fc = verifier.fold_commitment(commitment, challenges)
    fr = verifier.fold_row_evaluation(row, challenges)
    verifier.verify_evaluation(point, claimed, row)
    verifier.verify_opening(commitment, challenges, point, fc, folded, fr)
    say("verified")
    # Prover::commit(&Witness): the host witness uploaded through the prover's staged copies (four
    # threads and streams, pinned chunks) and the outputs copied back through its pinned buffer
    c_host, op_host = prover.commit(w)
    assert np.array_equal(c_host.raw, commitment.raw)
    assert np.array_equal(c_host.components, commitment.components)
    prover.recycle(op_host)
    say("host-witness commit equal")
