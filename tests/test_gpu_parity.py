"""Parity of the HIP path (through the C ABI) with the oracle on the same seeded inputs: the
commitment's raw rows and components, the folded witness, the row and MLE evaluations, and the
verifier's acceptance. Integer and F162 outputs are compared bit for bit."""
import ctypes
import sys

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

N = 648
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


def test_mfma_lane_layout(lbn):
    say("case: mfma probe")
    rng = np.random.default_rng(1)
    a = rng.integers(-128, 128, (16, 64)).astype(np.int8)
    b = rng.integers(-128, 128, (64, 16)).astype(np.int8)
    d = np.zeros((16, 16), np.int32)
    rc = lbn.lib().lbn_debug_mfma_probe(a.ctypes.data_as(ctypes.c_void_p), b.ctypes.data_as(ctypes.c_void_p),
                                        d.ctypes.data_as(ctypes.c_void_p))
    assert rc == 0
    assert np.array_equal(d, a.astype(np.int32) @ b.astype(np.int32))


def _setup(lbn, orc, params):
    pp = lbn.PublicParameters.from_seed(params, MATRIX_SEED)
    prover = lbn.Prover(pp)
    w = lbn.Witness.random(params, WITNESS_SEED)
    a = orc.key_expand(params.primes(), params.len_ring(), orc.matrix_seed(MATRIX_SEED))
    return pp, prover, w, a


CASES = [
    (10, 2, 3889, [9721]),
    (11, 3, 3889, [2917, 4861]),
    (10, 1, 3889, [12637, 17497, 19441]),
    (12, 4, 2917, [3889]),
    (11, 2, 19441, [4861]),
    (9, 2, 4861, []),
]


@pytest.mark.parametrize("wl,cl,base,extra", CASES)
def test_commit_fold_eval_parity(lbn, orc, wl, cl, base, extra):
    say(f"case: parity wl={wl} cl={cl} base={base} extra={extra}")
    params = lbn.Params.with_base(wl, cl, lbn.Modulus(base), [lbn.Modulus(q) for q in extra])
    pp, prover, w, a = _setup(lbn, orc, params)
    commitment, opening = prover.commit(w)
    raw, comp = orc.commit(params.primes(), params.len_ring(), params.columns(), a, w.elements)
    assert np.array_equal(commitment.raw, raw)
    assert np.array_equal(commitment.components, comp)
    verifier = lbn.Verifier(pp)
    t = lbn.Transcript(b"labinius/test/parity")
    point = verifier.derive_evaluation_point(t, commitment)
    ot = orc.Transcript(b"labinius/test/parity")
    op0, op1 = orc.derive_evaluation_point(ot, orc.make_params(wl, cl, params.primes(), MATRIX_SEED), comp)
    assert np.array_equal(point.p0, op0) and np.array_equal(point.p1, op1)
    row = prover.row_evaluate(w, point)
    assert np.array_equal(row.values, orc.row_evaluate(w.elements, point.p0))
    claimed = prover.mle_evaluate(w, point)
    assert np.array_equal(claimed, orc.claim(row.values, point.p1))
    challenges = verifier.derive_folding_challenges(t, row)
    och = orc.derive_folding_challenges(ot, row.values)
    assert all(bytes(x.positions) == bytes(y.positions) and x.signs == y.signs for x, y in zip(challenges.challenges, och))
    folded = prover.fold(opening, challenges)
    assert np.array_equal(folded.elements, orc.fold(base, params.len_ring(), params.columns(), w.elements, och))
    fc = verifier.fold_commitment(commitment, challenges)
    assert np.array_equal(fc.rows, orc.fold_commitment(params.primes(), params.columns(), comp, och))
    fr = verifier.fold_row_evaluation(row, challenges)
    verifier.verify_evaluation(point, claimed, row)
    verifier.verify_opening(commitment, challenges, point, fc, folded, fr)
    assert orc.verify_folded_opening(orc.make_params(wl, cl, params.primes(), MATRIX_SEED), a, fc.rows,
                                     folded.elements, point.p0, fr)


@pytest.mark.parametrize("chunk,wl,cl,extra", [(1, 13, 3, [2917]), (3, 16, 8, [9721]), (8, 18, 2, [2917])])
def test_gemm_k_chunks(lbn, orc, chunk, wl, cl, extra):
    """K split into many ranges whose residues are added into the output (the path every K past
    the int32 accumulators' exact range takes), here ranges of at most 1, 3 and 8 steps:
    commitments and folds stay bit-exact. (1, 13, 3) runs every step as its own range on the plain
    LUT; (18, 2) the commit's 20 736 steps (16 384 ring elements) as 8-step ranges on the
    replicated-LUT path."""
    say(f"case: K chunk {chunk} wl={wl} cl={cl} extra={extra}")
    L = lbn.lib()
    assert L.lbn_debug_gemm_chunk(ctypes.c_uint32(chunk)) == 0
    try:
        params = lbn.Params(wl, cl, [lbn.Modulus(q) for q in extra])
        pp, prover, w, a = _setup(lbn, orc, params)
        commitment, opening = prover.commit(w)
        raw, comp = orc.commit(params.primes(), params.len_ring(), params.columns(), a, w.elements)
        assert np.array_equal(commitment.raw, raw) and np.array_equal(commitment.components, comp)
        t = lbn.Transcript(b"labinius/test/chunks")
        verifier = lbn.Verifier(pp)
        point = verifier.derive_evaluation_point(t, commitment)
        row = prover.row_evaluate(w, point)
        challenges = verifier.derive_folding_challenges(t, row)
        folded = prover.fold(opening, challenges)
        assert np.array_equal(folded.elements, orc.fold(params.primes()[0], params.len_ring(), params.columns(),
                                                        w.elements, challenges.challenges))
    finally:
        assert L.lbn_debug_gemm_chunk(ctypes.c_uint32(0)) == 0
    assert L.lbn_debug_gemm_chunk(ctypes.c_uint32(3450)) != 0  # past the exact range


def test_basic_round_bit_exact(lbn, orc):
    """Params::basic() (2^18 F162, 128 columns, {3889, 9721}) with the reference bench's seeds."""
    say("case: basic round")
    params = lbn.Params.basic()
    pp, prover, w, a = _setup(lbn, orc, params)
    rw = prover.upload(w)
    commitment, opening = prover.commit_resident(rw)
    raw, comp = orc.commit(params.primes(), params.len_ring(), params.columns(), a, w.elements)
    assert np.array_equal(commitment.raw, raw)
    assert np.array_equal(commitment.components, comp)
    verifier = lbn.Verifier(pp)
    t = lbn.Transcript(b"labinius/reference")
    point = verifier.derive_evaluation_point(t, commitment)
    row = prover.row_evaluate(rw, point)
    assert np.array_equal(row.values, orc.row_evaluate(w.elements, point.p0))
    claimed = prover.mle_evaluate(rw, point)
    challenges = verifier.derive_folding_challenges(t, row)
    folded = prover.fold(opening, challenges)
    assert np.array_equal(folded.elements,
                          orc.fold(3889, params.len_ring(), params.columns(), w.elements, challenges.challenges))
    fc = verifier.fold_commitment(commitment, challenges)
    fr = verifier.fold_row_evaluation(row, challenges)
    verifier.verify_evaluation(point, claimed, row)
    verifier.verify_opening(commitment, challenges, point, fc, folded, fr)


def test_tampering_is_rejected(lbn, orc):
    # tests/fold.rs:29-242: a corrupted v, u or claim rejects
    say("case: tamper")
    params = lbn.Params(10, 2, [lbn.Modulus.Q9721_FS_S])
    pp, prover, w, a = _setup(lbn, orc, params)
    verifier = lbn.Verifier(pp)
    commitment, opening = prover.commit(w)
    t = lbn.Transcript(b"labinius/test/tamper")
    point = verifier.derive_evaluation_point(t, commitment)
    row = prover.row_evaluate(w, point)
    claimed = prover.mle_evaluate(w, point)
    challenges = verifier.derive_folding_challenges(t, row)
    folded = prover.fold(opening, challenges)
    fc = verifier.fold_commitment(commitment, challenges)
    fr = verifier.fold_row_evaluation(row, challenges)
    verifier.verify_opening(commitment, challenges, point, fc, folded, fr)
    bad = lbn.FoldedWitness(folded.elements.copy())
    bad.elements[1, 7] += 2
    with pytest.raises(lbn.VerificationError):
        verifier.verify_opening(commitment, challenges, point, fc, bad, fr)
    bad_claim = claimed.copy()
    bad_claim[0] ^= np.uint64(1)
    with pytest.raises(lbn.VerificationError):
        verifier.verify_evaluation(point, bad_claim, row)
    bad_row = lbn.RowEvaluation(row.values.copy())
    bad_row.values[0, 1] ^= np.uint64(4)
    with pytest.raises(lbn.VerificationError):
        verifier.verify_evaluation(point, claimed, bad_row)


def test_prover_runs_several_rounds_and_recycles(lbn, orc):
    # tests/api.rs: one prover, several rounds; recycle hands the buffer back
    say("case: rounds")
    params = lbn.Params(10, 2, [lbn.Modulus.Q2917_Q_S])
    pp, prover, w, a = _setup(lbn, orc, params)
    c1, o1 = prover.commit(w)
    prover.recycle(o1)
    c2, o2 = prover.commit(w)
    assert np.array_equal(c1.components, c2.components)
    w2 = lbn.Witness.random(params, bytes([1]) * 32)
    c3, o3 = prover.commit(w2)
    raw, comp = orc.commit(params.primes(), params.len_ring(), params.columns(), a, w2.elements)
    assert np.array_equal(c3.raw, raw)
    prover.recycle(o2)
    prover.recycle(o3)


def test_zero_witness_commits_to_zero(lbn, orc):
    # tests/commit.rs:344-355
    say("case: zero")
    params = lbn.Params(10, 2, [lbn.Modulus.Q9721_FS_S])
    pp, prover, _, _ = _setup(lbn, orc, params)
    w = lbn.Witness.from_elements(params, np.zeros((params.witness_len(), 3), np.uint64))
    c, o = prover.commit(w)
    assert not c.raw.any() and not c.components.any()
    prover.recycle(o)


@pytest.mark.parametrize("wl,cl", [(18, 1), (12, 5)])
def test_row_evaluation_many_rows(lbn, orc, wl, cl):
    # eq tables of 17 and 7 row variables (several 8-coordinate sub-tables, and one)
    say(f"case: row eval wl={wl} cl={cl}")
    params = lbn.Params(wl, cl, [])
    pp = lbn.PublicParameters.from_seed(params, MATRIX_SEED)
    prover = lbn.Prover(pp)
    w = lbn.Witness.random(params, WITNESS_SEED)
    rng = np.random.default_rng(wl)
    pt = rng.integers(0, 2**63, (wl, 3), dtype=np.uint64)
    pt[:, 2] &= np.uint64((1 << 34) - 1)
    point = lbn.EvaluationPoint(pt[:wl - cl], pt[wl - cl:])
    row = prover.row_evaluate(w, point)
    assert np.array_equal(row.values, orc.row_evaluate(w.elements, point.p0))
    assert np.array_equal(prover.mle_evaluate(w, point), orc.claim(row.values, point.p1))


@pytest.mark.parametrize("base,extra,d", [(3889, [9721], 10), (3889, [2917, 4861], 8),
                                          (9721, [17497, 19441, 2917, 12637], 12), (2917, [3889], 11)])
def test_drop_bits_device_matches_host(lbn, base, extra, d):
    """bd/digits.rs:121-182 on the GPU (every limb's inverse NTT, then Garner) against the host
    restatement, which the oracle tests pin: random raw rows, all-zero and all-(q - 1) rows, and a
    quadratic base limb. Then the resident form over the prover's own commitment."""
    say(f"case: drop_bits device base={base} extra={extra} d={d}")
    params = lbn.Params(13, 4, [lbn.Modulus(q) for q in extra], opening=lbn.Opening.bit_dropped(d),
                        base=lbn.Modulus(base))
    prover = lbn.Prover(lbn.PublicParameters.from_seed(params, MATRIX_SEED))
    with pytest.raises(lbn.LabiniusError):
        lbn.Dropped.of_last_commit(prover)  # no commitment yet
    primes, cols = params.primes(), params.columns()
    rng = np.random.default_rng(d)
    cases = [np.stack([rng.integers(0, q, (cols, N), dtype=np.uint32) for q in primes]),
             np.zeros((len(primes), cols, N), np.uint32),
             np.stack([np.full((cols, N), q - 1, np.uint32) for q in primes])]
    for raw in cases:
        host = lbn.Dropped.of_raw(params, raw)
        dev = lbn.Dropped.of_device(prover, params, raw)
        assert np.array_equal(dev.top, host.top) and np.array_equal(dev.digits, host.digits)
    commitment, _ = prover.commit(lbn.Witness.random(params, WITNESS_SEED))
    host = lbn.Dropped.of_raw(params, commitment.raw)
    assert np.array_equal(commitment.dropped.top, host.top) and np.array_equal(commitment.dropped.digits, host.digits)


@pytest.mark.parametrize("wl,cl,extra,d", [(12, 3, [2917, 4861], 8), (11, 2, [9721], 10), (10, 2, [2917, 12637], 7)])
def test_bit_dropped_round_bit_exact(lbn, orc, wl, cl, extra, d):
    # scheme/prover.rs:69-72 (drop_bits after the GPU commitment), verifier.rs:61-100 (the Dropped
    # arm of the point), verifier.rs:232-275 (the residual check)
    say(f"case: bit-dropped wl={wl} cl={cl} extra={extra} d={d}")
    params = lbn.Params(wl, cl, [lbn.Modulus(q) for q in extra], opening=lbn.Opening.bit_dropped(d))
    pp, prover, w, a = _setup(lbn, orc, params)
    primes, L, cols = params.primes(), params.len_ring(), params.columns()
    commitment, opening = prover.commit(w)
    raw, comp = orc.commit(primes, L, cols, a, w.elements)
    assert np.array_equal(commitment.raw, raw)
    top, digits = orc.drop_bits(primes, cols, raw, d)
    assert np.array_equal(commitment.dropped.top, top) and np.array_equal(commitment.dropped.digits, digits)
    verifier = lbn.Verifier(pp)
    t = lbn.Transcript(b"labinius/test/bd")
    point = verifier.derive_evaluation_point(t, commitment)
    ot = orc.Transcript(b"labinius/test/bd")
    oparams = orc.make_params(wl, cl, primes, MATRIX_SEED, dropped_bits=d)
    op0, op1 = orc.derive_evaluation_point_dropped(ot, oparams, orc.pack_dropped(primes, cols, d, top, digits))
    assert np.array_equal(point.p0, op0) and np.array_equal(point.p1, op1)
    row = prover.row_evaluate(w, point)
    assert np.array_equal(row.values, orc.row_evaluate(w.elements, point.p0))
    challenges = verifier.derive_folding_challenges(t, row)
    och = orc.derive_folding_challenges(ot, row.values)
    folded = prover.fold(opening, challenges)
    assert np.array_equal(folded.elements, orc.fold(primes[0], L, cols, w.elements, och))
    fr = verifier.fold_row_evaluation(row, challenges)
    verifier.verify_opening(commitment, challenges, point, None, folded, fr)
    assert orc.verify_opening_bd(oparams, a, top, digits, och, folded.elements, point.p0, fr)
    assert verifier.bd_residual(commitment, challenges, folded) == orc.bd_residual(primes, L, cols, a, d, top, digits,
                                                                                   och, folded.elements)


@pytest.mark.parametrize("mode", ["clear", "bit_dropped"])
def test_suite_l_round(lbn, orc, mode):
    """Params::sized(SUITES[2], mode): 2^22 F162 in 512 columns, moduli 3889 / 2917 / 4861 (8 dropped
    bits). Bit-exact against the oracle on sampled columns and rows (each is independent of the
    others), and the verifier accepts the whole opening."""
    say(f"case: suite l {mode}")
    opening_mode = lbn.Opening.CLEAR if mode == "clear" else lbn.Opening.bit_dropped(8)
    params = lbn.Params.sized(lbn.SUITES[2], opening_mode)
    pp, prover, w, a = _setup(lbn, orc, params)
    primes, L, cols, clen = params.primes(), params.len_ring(), params.columns(), params.column_len()
    rw = prover.upload(w)
    commitment, opening = prover.commit_resident(rw)
    sample = [0, 1, 257, cols - 1]
    for c in sample:
        raw_c, comp_c = orc.commit(primes, L, 1, a, w.elements[c * clen:(c + 1) * clen])
        assert np.array_equal(commitment.raw[:, c], raw_c[:, 0]), c
        assert np.array_equal(commitment.components[c], comp_c[0]), c
        if mode == "bit_dropped":
            top, digits = orc.drop_bits(primes, 1, raw_c, 8)
            assert np.array_equal(commitment.dropped.top[c * N:(c + 1) * N], top)
            assert np.array_equal(commitment.dropped.digits[:, c * N:(c + 1) * N], digits)
    verifier = lbn.Verifier(pp)
    t = lbn.Transcript(b"labinius/reference")
    point = verifier.derive_evaluation_point(t, commitment)
    row = prover.row_evaluate(rw, point)
    for c in sample:
        assert np.array_equal(row.values[c], orc.row_evaluate(w.elements[c * clen:(c + 1) * clen], point.p0)[0])
    claimed = prover.mle_evaluate(rw, point)
    assert np.array_equal(claimed, orc.claim(row.values, point.p1))
    challenges = verifier.derive_folding_challenges(t, row)
    folded = prover.fold(opening, challenges)
    rows = np.array([0, 1, 1000, L - 1])
    assert np.array_equal(folded.elements[rows],
                          orc.fold_rows(primes[0], L, cols, w.elements, challenges.challenges, rows))
    fr = verifier.fold_row_evaluation(row, challenges)
    verifier.verify_evaluation(point, claimed, row)
    fc = None if mode == "bit_dropped" else verifier.fold_commitment(commitment, challenges)
    verifier.verify_opening(commitment, challenges, point, fc, folded, fr)
