"""Pins the oracle's cross-field switch (switch.rs, fields/crossfield.rs, fields/sumcheck.rs):
B128 arithmetic against an independent big-integer restatement, the eq tables and MLEs at Boolean
points (which fixes the coordinate order), and the reference's own round-trip and rejection tests
(switch.rs:211-282), with the opened value checked against the lifted witness's MLE at the
switch's point through the main protocol's row evaluation."""
import numpy as np
import pytest

MASK64 = (1 << 64) - 1


def as_int(x):
    x = np.asarray(x, np.uint64)
    return int(x[0]) | (int(x[1]) << 64)


def as_pair(v):
    return np.array([v & MASK64, v >> 64], np.uint64)


def py_b128_mul(a, b):
    # carry-less product, then reduction by x^128 + x^7 + x^2 + x + 1 (0x87 below x^128)
    p = 0
    while b:
        if b & 1:
            p ^= a
        a <<= 1
        b >>= 1
    for i in range(254, 127, -1):
        if (p >> i) & 1:
            p ^= (1 << i) | (0x87 << (i - 128))
    return p


def rand_b128(rng, n):
    return rng.integers(0, 2**63, (n, 2), dtype=np.uint64) * np.uint64(2) + rng.integers(0, 2, (n, 2), dtype=np.uint64)


def test_b128_mul_matches_bigint_and_the_modulus(orc):
    rng = np.random.default_rng(1)
    for a, b in zip(rand_b128(rng, 64), rand_b128(rng, 64)):
        assert as_int(orc.b128_mul(a, b)) == py_b128_mul(as_int(a), as_int(b))
    x = as_pair(2)
    x127 = as_pair(1 << 127)
    assert as_int(orc.b128_mul(x, x127)) == 0x87  # x^128 = x^7 + x^2 + x + 1
    one = as_pair(1)
    a = rand_b128(rng, 1)[0]
    assert np.array_equal(orc.b128_mul(a, one), a)


def test_eq_tables_sum_to_one_and_pick_boolean_points(orc):
    rng = np.random.default_rng(2)
    r = rand_b128(rng, 5)
    t = orc.eq_b128(r)
    acc = 0
    for x in t:
        acc ^= as_int(x)
    assert acc == 1
    # at a Boolean point eq is the indicator of j, coordinate 0 the most significant bit
    for j in (0, 1, 6, 19, 31):
        rb = np.array([as_pair((j >> (4 - i)) & 1) for i in range(5)])
        tb = orc.eq_b128(rb)
        assert [as_int(x) for x in tb] == [int(i == j) for i in range(32)]
    rf = rng.integers(0, 2**63, (4, 3), dtype=np.uint64)
    tf = orc.eq_f162_msb(rf)
    s = np.bitwise_xor.reduce(tf, axis=0)
    assert list(s) == [1, 0, 0]
    for j in (0, 5, 10):
        rb = np.zeros((4, 3), np.uint64)
        for i in range(4):
            rb[i, 0] = (j >> (3 - i)) & 1
        tb = orc.eq_f162_msb(rb)
        assert [int(x[0]) for x in tb] == [int(i == j) for i in range(16)] and not tb[:, 1:].any()


def test_mles_at_boolean_points(orc):
    rng = np.random.default_rng(3)
    trace = rand_b128(rng, 64)
    for j in (0, 7, 42, 63):
        r = np.array([as_pair((j >> (5 - i)) & 1) for i in range(6)])
        assert np.array_equal(orc.b128_mle(trace, r), trace[j])
        for k in (0, 5, 64, 127):
            r_lo = np.array([as_pair((k >> (6 - i)) & 1) for i in range(7)])
            assert as_int(orc.bit_mle(trace, r_lo, r)) == (as_int(trace[j]) >> k) & 1
    # linearity in the trace
    r = rand_b128(rng, 6)
    t2 = rand_b128(rng, 64)
    assert np.array_equal(orc.b128_mle(trace ^ t2, r), orc.b128_mle(trace, r) ^ orc.b128_mle(t2, r))


def _lifted(trace):
    w = np.zeros((len(trace), 3), np.uint64)
    w[:, :2] = trace
    return w


@pytest.mark.parametrize("bits", [False, True])
def test_switch_round_trip_and_rejection(orc, bits):
    # switch.rs:222-281 (words_round_trip_and_rejection, bits_round_trip_and_rejection)
    l, cl = 10, 3
    rng = np.random.default_rng(4 + bits)
    trace = rand_b128(rng, 1 << l)
    r_hi = rand_b128(rng, l)
    r_lo = rand_b128(rng, 7) if bits else None
    claim = orc.bit_mle(trace, r_lo, r_hi) if bits else orc.b128_mle(trace, r_hi)
    partials, rounds, opened, p0, p1 = orc.switch_prove(trace, cl, r_lo, r_hi, claim, orc.Transcript(b"t"))
    ok, q0, q1 = orc.switch_verify(l, cl, r_lo, r_hi, claim, partials, rounds, opened, orc.Transcript(b"t"))
    assert ok and np.array_equal(q0, p0) and np.array_equal(q1, p1)
    # the opened value is the lifted witness's MLE at the switch's point (eval.rs:185-193)
    u = orc.row_evaluate(_lifted(trace), p0)
    assert np.array_equal(orc.claim(u, p1), opened)
    wrong = claim.copy()
    wrong[0] ^= np.uint64(1)
    assert not orc.switch_verify(l, cl, r_lo, r_hi, wrong, partials, rounds, opened, orc.Transcript(b"t"))[0]
    bad = opened.copy()
    bad[0] ^= np.uint64(1)
    assert not orc.switch_verify(l, cl, r_lo, r_hi, claim, partials, rounds, bad, orc.Transcript(b"t"))[0]
    br = rounds.copy()
    br[3, 1, 0] ^= np.uint64(1)
    assert not orc.switch_verify(l, cl, r_lo, r_hi, claim, partials, br, opened, orc.Transcript(b"t"))[0]
    # another transcript domain derives other challenges and another point
    _, _, _, o0, _ = orc.switch_prove(trace, cl, r_lo, r_hi, claim, orc.Transcript(b"u"))
    assert not np.array_equal(o0, p0)


@pytest.mark.parametrize("bits", [False, True])
def test_host_verifier_accepts_oracle_proofs(orc, lbn, bits):
    """The library's host verifier (lbn_switch_verify) on proofs the oracle made: the same point,
    and the same rejections."""
    from labinius_amd import switch
    l, cl = 10, 3
    params = lbn.Params(l, cl, [])
    rng = np.random.default_rng(7 + bits)
    trace = rand_b128(rng, 1 << l)
    r_hi = rand_b128(rng, l)
    r_lo = rand_b128(rng, 7) if bits else None
    claim = orc.bit_mle(trace, r_lo, r_hi) if bits else orc.b128_mle(trace, r_hi)
    partials, rounds, opened, p0, p1 = orc.switch_prove(trace, cl, r_lo, r_hi, claim, orc.Transcript(b"t"))
    proof = switch.SwitchProof(partials, rounds, opened)
    verify = (lambda c, pr: switch.verify_bits(params, r_lo, r_hi, c, pr, lbn.Transcript(b"t"))) if bits else \
        (lambda c, pr: switch.verify(params, r_hi, c, pr, lbn.Transcript(b"t")))
    point, value = verify(claim, proof)
    assert np.array_equal(point.p0, p0) and np.array_equal(point.p1, p1) and np.array_equal(value, opened)
    wrong = claim.copy()
    wrong[1] ^= np.uint64(1 << 40)
    with pytest.raises(switch.SwitchError):
        verify(wrong, proof)
    bad = switch.SwitchProof(partials, rounds, opened ^ np.array([0, 1, 0], np.uint64))
    with pytest.raises(switch.SwitchError):
        verify(claim, bad)
