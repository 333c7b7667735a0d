"""The cross-field switch on the GPU (switch.rs): the device prover against the oracle, bit-exact
(partial evaluations, sumcheck rounds, opened value and point), for claims on words and on bits;
the host verifier accepts; and the two-field round of crates/pcs/examples/two_fields.rs: the
commitment to the lifted trace opens at the switch's point to the switch's value."""
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


def rand_b128(rng, n):
    return rng.integers(0, 2**63, (n, 2), dtype=np.uint64) * np.uint64(2) + rng.integers(0, 2, (n, 2), dtype=np.uint64)


@pytest.mark.parametrize("wl,cl,bits", [(10, 3, False), (10, 3, True), (13, 5, False), (16, 4, True), (17, 9, False)])
def test_switch_device_matches_oracle(lbn, orc, wl, cl, bits):
    from labinius_amd import switch
    say(f"case: switch wl={wl} cl={cl} bits={bits}")
    params = lbn.Params(wl, cl, [lbn.Modulus(9721)])
    prover = lbn.Prover(lbn.PublicParameters.from_seed(params, MATRIX_SEED))
    rng = np.random.default_rng(wl + 100 * bits)
    trace = rand_b128(rng, 1 << wl)
    rw = prover.upload(lbn.Witness.lifted(params, trace))
    r_hi = rand_b128(rng, wl)
    r_lo = rand_b128(rng, 7) if bits else None
    claim = orc.bit_mle(trace, r_lo, r_hi) if bits else orc.b128_mle(trace, r_hi)
    if bits:
        proof, point = switch.prove_bits(prover, rw, r_lo, r_hi, claim, lbn.Transcript(b"labinius/test/switch"))
    else:
        proof, point = switch.prove(prover, rw, r_hi, claim, lbn.Transcript(b"labinius/test/switch"))
    partials, rounds, opened, p0, p1 = orc.switch_prove(trace, cl, r_lo, r_hi, claim,
                                                        orc.Transcript(b"labinius/test/switch"))
    assert np.array_equal(proof.partials, partials)
    assert np.array_equal(proof.rounds, rounds)
    assert np.array_equal(proof.opened, opened)
    assert np.array_equal(point.p0, p0) and np.array_equal(point.p1, p1)
    t = lbn.Transcript(b"labinius/test/switch")
    vp, value = (switch.verify_bits(params, r_lo, r_hi, claim, proof, t) if bits
                 else switch.verify(params, r_hi, claim, proof, t))
    assert np.array_equal(vp.p0, p0) and np.array_equal(value, opened)
    # the resident witness is left as it was: a second proof on it is the same
    again, _ = switch.prove(prover, rw, r_hi, orc.b128_mle(trace, r_hi), lbn.Transcript(b"labinius/test/again"))
    ref = orc.switch_prove(trace, cl, None, r_hi, orc.b128_mle(trace, r_hi), orc.Transcript(b"labinius/test/again"))
    assert np.array_equal(again.rounds, ref[1]) and np.array_equal(again.opened, ref[2])


def test_two_field_round(lbn, orc):
    """crates/pcs/examples/two_fields.rs: commit Witness::lifted(trace), run the switch on a claim on
    the words, then open the commitment at the switch's point: the MLE there is the switch's value,
    and the verifier accepts the whole opening."""
    from labinius_amd import switch
    say("case: two-field round")
    params = lbn.Params(13, 4, [lbn.Modulus(9721)])
    pp = lbn.PublicParameters.from_seed(params, MATRIX_SEED)
    prover, verifier = lbn.Prover(pp), lbn.Verifier(pp)
    rng = np.random.default_rng(11)
    trace = rand_b128(rng, 1 << 13)
    w = lbn.Witness.lifted(params, trace)
    commitment, opening = prover.commit(w)
    rw = prover.upload(w)
    t = lbn.Transcript(b"labinius/test/two-fields")
    r = rand_b128(rng, 13)
    claim = orc.b128_mle(trace, r)
    proof, point = switch.prove(prover, rw, r, claim, t)
    assert np.array_equal(prover.mle_evaluate(w, point), proof.opened)
    row = prover.row_evaluate(w, point)
    challenges = verifier.derive_folding_challenges(t, row)
    folded = prover.fold(opening, challenges)
    vt = lbn.Transcript(b"labinius/test/two-fields")
    vpoint, value = switch.verify(params, r, claim, proof, vt)
    assert np.array_equal(vpoint.p0, point.p0) and np.array_equal(vpoint.p1, point.p1)
    verifier.verify_evaluation(vpoint, value, row)
    vchallenges = verifier.derive_folding_challenges(vt, row)
    assert all(bytes(x.positions) == bytes(y.positions) for x, y in zip(challenges.challenges, vchallenges.challenges))
    verifier.verify_opening(commitment, vchallenges, vpoint, verifier.fold_commitment(commitment, vchallenges), folded,
                            verifier.fold_row_evaluation(row, vchallenges))
