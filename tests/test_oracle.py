"""Pins the oracle (oracle/oracle.c) against closed forms, definitions and published vectors,
chosen so that a plausible slip (a dropped term, a wrong sign or index, a transposed operand)
fails one of them. The reference's own tests these follow are cited per test."""
import math

import numpy as np
import pytest

N = 648
SPLIT = [3889, 9721, 17497, 19441]
QUAD = [2917, 4861, 12637]
ALL = sorted(SPLIT + QUAD)


def _blake3_vectors():
    import json
    import os
    return json.load(open(os.path.join(os.path.dirname(__file__), "golden", "blake3_spec_vectors.json")))


def test_blake3_published_vectors(orc):
    # the BLAKE3 specification's test vectors (tests/golden/blake3_spec_vectors.json)
    v = _blake3_vectors()
    for case in v["cases"]:
        data = bytes(i % 251 for i in range(case["input_len"]))
        assert orc.blake3(data).hex() == case["hash"]
        if "keyed_hash" in case:
            assert orc.blake3(data, key=v["key"].encode()).hex() == case["keyed_hash"]
        if "derive_key" in case:
            assert orc.blake3(data, context=v["context"]).hex() == case["derive_key"]
    for case in v["literal"]:
        assert orc.blake3(case["input_ascii"].encode()).hex() == case["hash"]


def test_blake3_xof_prefix_and_tree(orc):
    # the XOF extends the 32-byte hash; every prefix of a long read is the shorter read
    data = bytes(i % 251 for i in range(5000))
    long = orc.blake3(data, 200)
    for n in (1, 31, 32, 64, 65, 199):
        assert orc.blake3(data, n) == long[:n]
    # changing one byte in any chunk of a multi-chunk input changes the hash
    base = orc.blake3(data)
    for pos in (0, 1023, 1024, 2047, 3071, 4999):
        d = bytearray(data)
        d[pos] ^= 1
        assert orc.blake3(bytes(d)) != base


def test_blake3_against_python_tree(orc):
    # an independent, recursive (non-incremental) restatement of the spec's tree in tests/
    from tests.blake3_ref import blake3_hash
    for n in (0, 1, 63, 64, 65, 1023, 1024, 1025, 2048, 2049, 3072, 4096, 5121, 8192 + 7):
        data = bytes(i % 251 for i in range(n))
        assert orc.blake3(data, 40) == blake3_hash(data, 40), n


def test_tables_slot_exponents_are_units(orc):
    # tests/scalar_reference.rs:18-26 and the compile-time asserts of params.rs:334-345
    t = orc.tables()
    se = t["SLOT_EXP"].astype(int)
    assert sorted(se) == [u for u in range(1944) if math.gcd(u, 1944) == 1]
    qse = t["QUAD_SLOT_EXP"].astype(int)
    assert sorted(qse) == [u for u in range(972) if math.gcd(u, 972) == 1]
    # POW3 classes: the 162 units mod 486, each owning 4 slots with exponents v + 486 t
    p3 = t["POW3_SLOT_EXP"].astype(int)
    assert sorted(p3) == [u for u in range(486) if math.gcd(u, 486) == 1]
    s648 = t["SLOT_648"].astype(int)
    for tt in range(4):
        for s in range(162):
            assert se[s648[tt, s]] == p3[s] + 486 * tt
    qcs = t["QUAD_CLASS_SLOT"].astype(int)
    for s in range(162):
        assert qse[qcs[0, s]] == p3[s] and qse[qcs[1, s]] == p3[s] + 486


@pytest.mark.parametrize("q", ALL)
def test_roots(orc, q):
    # tests/scalar_reference.rs:28-45
    psi = orc.psi(q)
    if orc.is_quad(q):
        assert pow(psi, 972, q) == 1 and pow(psi, 486, q) != 1 and pow(psi, 324, q) != 1
        assert all(pow(x, 972, q) != 1 or pow(x, 486, q) == 1 or pow(x, 324, q) == 1 for x in range(2, psi))
    else:
        assert pow(psi, 1944, q) == 1 and pow(psi, 972, q) != 1 and pow(psi, 648, q) != 1
        assert all(pow(x, 1944, q) != 1 or pow(x, 972, q) == 1 or pow(x, 648, q) == 1 for x in range(2, psi))
    assert orc.is_quad(q) == (q in QUAD)


@pytest.mark.parametrize("q", ALL)
def test_ntt_is_evaluation(orc, q):
    # NTT(a)[j] = a(psi^SLOT_EXP[j]) (scalar.rs:86-95) and the quadratic leaves (scalar.rs:237-248)
    rng = np.random.default_rng(q)
    t = orc.tables()
    for trial in range(2):
        a = rng.integers(0, q, N).astype(np.uint32)
        y = orc.ntt(q, a)
        if orc.is_quad(q):
            for j in range(0, 324, 7):
                assert (int(y[2 * j]), int(y[2 * j + 1])) == orc.eval_quad_at(q, a, int(t["QUAD_SLOT_EXP"][j]))
        else:
            for j in range(0, N, 5):
                assert int(y[j]) == orc.eval_at(q, a, int(t["SLOT_EXP"][j]))
        assert np.array_equal(orc.intt(q, y), a)


@pytest.mark.parametrize("q", ALL)
def test_ntt_is_a_ring_homomorphism(orc, q):
    # tests/scalar_reference.rs (NTT(ab) = NTT(a) o NTT(b)), tests/ntt.rs:1909-1911 for quad
    rng = np.random.default_rng(q + 1)
    a = rng.integers(0, q, N).astype(np.uint32)
    b = rng.integers(0, q, N).astype(np.uint32)
    ab = orc.mul_mod_phi(a, b, q)
    na, nb, nab = orc.ntt(q, a), orc.ntt(q, b), orc.ntt(q, ab)
    if orc.is_quad(q):
        psi = orc.psi(q)
        qse = orc.tables()["QUAD_SLOT_EXP"]
        for j in range(324):
            c = pow(psi, int(qse[j]), q)
            a0, a1, b0, b1 = (int(x) for x in (na[2 * j], na[2 * j + 1], nb[2 * j], nb[2 * j + 1]))
            assert int(nab[2 * j]) == (a0 * b0 + c * a1 * b1) % q
            assert int(nab[2 * j + 1]) == (a0 * b1 + a1 * b0) % q
    else:
        assert np.array_equal(nab, (na.astype(np.uint64) * nb % q).astype(np.uint32))


def test_mul_mod_phi_relation(orc):
    # X^648 = X^324 - 1 (tests/scalar_reference.rs)
    q = 3889
    x324 = np.zeros(N, np.uint32)
    x324[324] = 1
    x = orc.mul_mod_phi(x324, x324, q)
    want = np.zeros(N, np.uint32)
    want[324] = 1
    want[0] = q - 1
    assert np.array_equal(x, want)


@pytest.mark.parametrize("q", ALL)
def test_decomposition_is_the_definition(orc, q):
    # ring/mod.rs:22-38: out[k][s] = y_k(theta^{v_s}), y_k(Y) = sum_m y[4m + k] Y^m, theta = psi^4
    # (psi'^2 for a quadratic prime), from the coefficient form of y
    rng = np.random.default_rng(q + 2)
    coeffs = rng.integers(0, q, N).astype(np.uint32)
    y = orc.ntt(q, coeffs)
    d = orc.decompose(q, y)
    psi = orc.psi(q)
    theta = pow(psi, 2 if orc.is_quad(q) else 4, q)
    p3 = orc.tables()["POW3_SLOT_EXP"]
    for s in range(0, 162, 11):
        z = pow(theta, int(p3[s]), q)
        for k in range(4):
            val = 0
            for m in reversed(range(162)):
                val = (val * z + int(coeffs[4 * m + k])) % q
            c = val if val <= (q - 1) // 2 else val - q
            assert int(d[k, s]) == c


def test_witness_lift(orc):
    # f162.rs:32-41: coefficient 4m + k is bit m of element k
    rng = np.random.default_rng(3)
    e = rng.integers(0, 2**63, (4, 3), dtype=np.uint64)
    e[:, 2] &= np.uint64((1 << 34) - 1)
    c = orc.lift(e)
    for m in range(162):
        for k in range(4):
            assert c[4 * m + k] == (int(e[k, m // 64]) >> (m % 64)) & 1


def _sum_commit_coeff(orc, q, a_coeffs, lifts):
    # y = sum_i a_i w_i in R_648 (coefficient domain), then transformed
    acc = np.zeros(N, np.uint64)
    for ai, wi in zip(a_coeffs, lifts):
        acc = (acc + orc.mul_mod_phi(ai, wi, q)) % q
    return orc.ntt(q, acc.astype(np.uint32))


@pytest.mark.parametrize("primes", [[3889, 9721], [2917, 4861], [12637, 17497, 19441]])
def test_commit_is_the_ring_product(orc, primes):
    # tests/commit.rs:86-101 sums scalar NTTs; here, independently, the R_648 product sum_i a_i w_i
    # with a_i = INTT(A_i), transformed back
    L, cols = 32, 2
    w = orc.witness_random(bytes([9]) * 32, 4 * L * cols)
    a = orc.key_expand(primes, L, 12345)
    raw, comp = orc.commit(primes, L, cols, a, w)
    for k, q in enumerate(primes):
        for c in range(cols):
            lifts = [orc.lift(w[4 * (c * L + i):4 * (c * L + i) + 4]) for i in range(L)]
            a_coeffs = [orc.intt(q, (a[k, i].astype(np.int64) % q).astype(np.uint32)) for i in range(L)]
            want = _sum_commit_coeff(orc, q, a_coeffs, lifts)
            assert np.array_equal(raw[k, c], want)
            assert np.array_equal(comp[c, :, k], orc.decompose(q, raw[k, c]))


def test_key_layout_and_range(orc):
    # key.rs:63-81: centered in [-(q-1)/2, (q-1)/2]; a shorter key's A is a prefix
    a = orc.key_expand([3889, 2917], 64, 777)
    for k, q in enumerate([3889, 2917]):
        assert a[k].min() >= -(q - 1) // 2 and a[k].max() <= (q - 1) // 2
    b = orc.key_expand([3889], 32, 777)
    assert np.array_equal(a[0, :32], b[0])


def _gf2_mul(a, b):
    return (a.astype(np.int64) @ b.astype(np.int64)) & 1


def _gf2_pow(m, e):
    r = np.eye(64, dtype=np.int64)
    while e:
        if e & 1:
            r = _gf2_mul(r, m)
        m = _gf2_mul(m, m)
        e >>= 1
    return r


def test_xorshift_is_the_full_period_generator(orc):
    # rng.rs:10-14 is Marsaglia's xorshift over GF(2)^64: the oracle's step must be linear, and
    # its matrix must have order 2^64 - 1 (the full period of the published (12, 25, 27) triple),
    # which a wrong shift or direction breaks
    cols = [orc.xorshift_step(1 << i) for i in range(64)]
    m = np.array([[(c >> r) & 1 for c in cols] for r in range(64)], np.int64)  # column i = step(e_i)
    rng = np.random.default_rng(31)
    for _ in range(20):
        a, b = (int(x) for x in rng.integers(0, 2**63, 2, dtype=np.uint64))
        assert orc.xorshift_step(a ^ b) == orc.xorshift_step(a) ^ orc.xorshift_step(b)
        v = np.array([(a >> i) & 1 for i in range(64)], np.int64)
        out = (m @ v) & 1
        assert sum(int(out[i]) << i for i in range(64)) == orc.xorshift_step(a)
    period = (1 << 64) - 1
    eye = np.eye(64, dtype=np.int64)
    assert np.array_equal(_gf2_pow(m, period), eye)
    for p in (3, 5, 17, 257, 641, 65537, 6700417):  # the prime factors of 2^64 - 1
        assert not np.array_equal(_gf2_pow(m, period // p), eye), p


def test_key_draws_are_uniform_and_centred(orc):
    # key.rs:63-81: A_k[i][j] = below(q) - (q - 1) / 2 with below(q) = (hi32 * q) >> 32: every
    # value of the centred range is hit about equally often, and each limb has its own stream
    q = 97
    a = orc.key_expand([q, 3889], 2048, 12345)
    counts = np.bincount(a[0].ravel().astype(np.int64) + (q - 1) // 2, minlength=q)
    assert counts.size == q and counts.min() > 0
    expect = a[0].size / q
    chi2 = float(((counts - expect) ** 2 / expect).sum())
    assert chi2 < 160.0  # 96 degrees of freedom; the 0.999 quantile is about 145
    assert not np.array_equal(a[0][:8] % q, a[1][:8] % q)


def test_sampler_statistics(orc):
    # tests/challenge.rs:106-125 (positions uniform over 20 000 unrejected attempts, within 15% of
    # weight / 162 each), :169-191 (the signs are balanced), :37-60 (absorbed data separates the
    # streams) and :235-271 (at bound 15 under two attempts per challenge on average)
    t = orc.Transcript(b"oracle/uniform")
    hits = np.zeros(162, np.int64)
    plus = total = 0
    n = 20000
    for _ in range(n):
        c, att = t.sample_short_challenge(28, 1e9)  # no rejection: one attempt each
        assert att == 1
        pos = list(c.positions[:c.weight])
        hits[pos] += 1
        s = sum((c.signs >> i) & 1 for i in range(c.weight))
        plus += c.weight - s
        total += c.weight
    expect = n * 28 / 162
    assert hits.min() > 0.85 * expect and hits.max() < 1.15 * expect
    assert abs(plus / total - 0.5) < 0.05
    a, b = orc.Transcript(b"oracle/sep"), orc.Transcript(b"oracle/sep")
    a.absorb_bytes(b"x")
    b.absorb_bytes(b"y")
    ca, cb = a.sample_short_challenge(30, 1e9)[0], b.sample_short_challenge(30, 1e9)[0]
    assert bytes(ca.positions) != bytes(cb.positions) or ca.signs != cb.signs
    t = orc.Transcript(b"oracle/accept")
    att = [t.sample_short_challenge(28, 15.0)[1] for _ in range(2000)]
    assert np.mean(att) < 2.0


def _challenges(orc, n, label=b"oracle/test"):
    t = orc.Transcript(label)
    return [t.sample_short_challenge()[0] for _ in range(n)]


@pytest.mark.parametrize("base", [3889, 2917, 9721, 4861])
def test_fold_is_the_integer_sum(orc, base):
    # v = sum_j c_j(-X^4) W_j (fold.rs:6-19): the NTT-domain computation equals the direct sparse
    # convolution over Z, reduced and centered (tests/fold.rs:89-99)
    L, cols = 32, 4
    w = orc.witness_random(bytes([5]) * 32, 4 * L * cols)
    ch = _challenges(orc, cols)
    v = orc.fold(base, L, cols, w, ch)
    direct = orc.fold_direct(base, L, cols, w, ch, np.arange(L))
    assert np.array_equal(v, direct)
    assert np.abs(v).max() < 200


def test_embed(orc):
    # fold.rs:287-300
    c = _challenges(orc, 1)[0]
    e = orc.embed(c)
    nz = np.nonzero(e)[0]
    assert len(nz) == c.weight and all(x % 4 == 0 for x in nz)
    for i in range(c.weight):
        m = c.positions[i]
        co = 1 - 2 * ((c.signs >> i) & 1)
        assert e[4 * m] == (co if m % 2 == 0 else -co)


def _f(x):
    return np.array(x, np.uint64)


def test_f162_field_laws(orc):
    # tests/fields.rs:115-149: x has order 243; x^162 = x^81 + 1; commutative, distributive
    x = _f([2, 0, 0])
    p = _f([1, 0, 0])
    for _ in range(162):
        p = orc.f162_mul(p, x)
    want = _f([1, 0, 0])
    want[1] |= np.uint64(1 << 17)  # x^81
    assert np.array_equal(p, want)
    for _ in range(81):
        p = orc.f162_mul(p, x)
    assert np.array_equal(p, _f([1, 0, 0]))
    rng = np.random.default_rng(4)
    for _ in range(5):
        a, b, c = (rng.integers(0, 2**63, 3, dtype=np.uint64) for _ in range(3))
        for v in (a, b, c):
            v[2] &= np.uint64((1 << 34) - 1)
        assert np.array_equal(orc.f162_mul(a, b), orc.f162_mul(b, a))
        assert np.array_equal(orc.f162_mul(a, b ^ c), orc.f162_mul(a, b) ^ orc.f162_mul(a, c))


def test_eq_row_mle_naive(orc):
    # tests/eval.rs:15-46: eq_naive, row_naive, mle_naive
    rows, cols = 4, 2
    rng = np.random.default_rng(6)
    w = rng.integers(0, 2**63, ((1 << rows) * (1 << cols), 3), dtype=np.uint64)
    w[:, 2] &= np.uint64((1 << 34) - 1)
    p0 = rng.integers(0, 2**63, (rows, 3), dtype=np.uint64)
    p1 = rng.integers(0, 2**63, (cols, 3), dtype=np.uint64)
    p0[:, 2] &= np.uint64((1 << 34) - 1)
    p1[:, 2] &= np.uint64((1 << 34) - 1)

    def eq_naive(ps, b):
        acc = _f([1, 0, 0])
        for k, x in enumerate(ps):
            f = x if (b >> k) & 1 else x ^ _f([1, 0, 0])
            acc = orc.f162_mul(acc, f)
        return acc

    eq = orc.eq_table(p0)
    for b in range(1 << rows):
        assert np.array_equal(eq[b], eq_naive(p0, b))
    u = orc.row_evaluate(w, p0)
    wdim = 1 << rows
    for j in range(1 << cols):
        s = _f([0, 0, 0])
        for i in range(wdim):
            s = s ^ orc.f162_mul(eq_naive(p0, i), w[i + wdim * j])
        assert np.array_equal(u[j], s)
    t = orc.claim(u, p1)
    m = _f([0, 0, 0])
    for b in range(len(w)):
        m = m ^ orc.f162_mul(orc.f162_mul(eq_naive(p0, b % wdim), eq_naive(p1, b // wdim)), w[b])
    assert np.array_equal(t, m)


def test_challenges_are_short_and_deterministic(orc):
    # challenge.rs:517-533; tests/challenge.rs: determinism, weight, sorted positions, bound
    a = _challenges(orc, 20, b"oracle/det")
    b = _challenges(orc, 20, b"oracle/det")
    assert all(bytes(x.positions) == bytes(y.positions) and x.signs == y.signs for x, y in zip(a, b))
    for c in a:
        pos = list(c.positions[:c.weight])
        assert c.weight == 28 and pos == sorted(set(pos)) and max(pos) < 162
        assert c.signs < (1 << 28)
        assert orc.canonical_norm_sq(c) <= 144.0 + 1e-9


def test_canonical_norm_is_horner(orc):
    # challenge.rs:441-461: the naive Horner evaluation over all 162 roots
    for c in _challenges(orc, 5, b"oracle/norm"):
        coeffs = [0] * 162
        for i in range(c.weight):
            coeffs[c.positions[i]] = 1 - 2 * ((c.signs >> i) & 1)
        best = 0.0
        for u in [x for x in range(1, 243) if x % 3]:
            z = complex(math.cos(2 * math.pi * u / 243), math.sin(2 * math.pi * u / 243))
            acc = 0
            for p in reversed(range(162)):
                acc = acc * z + coeffs[p]
            best = max(best, abs(acc) ** 2)
        assert abs(best - orc.canonical_norm_sq(c)) < 1e-6


def test_honest_round_verifies_and_tamper_rejects(orc):
    # tests/fold.rs:29-242 in miniature: commit, point, row, challenges, fold, verify
    wl, cl, primes = 10, 2, [3889, 9721]
    seed = bytes([41]) * 32
    params = orc.make_params(wl, cl, primes, seed)
    L = (1 << (wl - cl)) // 4
    a = orc.key_expand(primes, L, orc.matrix_seed(seed))
    w = orc.witness_random(bytes([43]) * 32, 1 << wl)
    raw, comp = orc.commit(primes, L, 1 << cl, a, w)
    t = orc.Transcript(b"oracle/round")
    p0, p1 = orc.derive_evaluation_point(t, params, comp)
    u = orc.row_evaluate(w, p0)
    ch = orc.derive_folding_challenges(t, u)
    v = orc.fold(primes[0], L, 1 << cl, w, ch)
    fc = orc.fold_commitment(primes, 1 << cl, comp, ch)
    uf = orc.fold_binary(u, ch)
    assert orc.verify_folded_opening(params, a, fc, v, p0, uf)
    bad = v.copy()
    bad[3, 7] += 2
    assert not orc.verify_folded_opening(params, a, fc, bad, p0, uf)
    bad_u = uf.copy()
    bad_u[0] ^= np.uint64(1)
    assert not orc.verify_folded_opening(params, a, fc, v, p0, bad_u)


# ---- bit-dropped commitments --------------------------------------------------------------------

def _bd_setup(orc, primes, d, wl=9, cl=2, seed=61):
    L, cols = (1 << (wl - cl)) // 4, 1 << cl
    a = orc.key_expand(primes, L, 4242 + seed)
    w = orc.witness_random(bytes([seed]) * 32, 4 * L * cols)
    raw, comp = orc.commit(primes, L, cols, a, w)
    return L, cols, a, w, raw


def _mixed_radix(x, primes):
    out = []
    for q in primes:
        out.append(x % q)
        x //= q
    return out


@pytest.mark.parametrize("primes,d", [([3889, 2917, 4861], 8), ([3889, 9721], 10), ([2917, 4861, 12637], 7)])
def test_drop_bits_is_the_mixed_radix_of_the_crt_value(orc, primes, d):
    # bd/digits.rs:121-182: per coefficient, the CRT value of the column's coefficients over all
    # primes written in mixed radix (digit k < q_k); the base digit keeps its top bits, rounded
    L, cols, a, w, raw = _bd_setup(orc, primes, d)
    top, digits = orc.drop_bits(primes, cols, raw, d)
    M = int(np.prod([int(q) for q in primes], dtype=object))
    for c in range(cols):
        coeffs = [orc.intt(q, raw[k, c]).astype(np.int64) for k, q in enumerate(primes)]
        for i in range(0, N, 7):
            # CRT by search over the base residue class (independent of Garner's formula)
            x = 0
            mod = 1
            for k, q in enumerate(primes):
                r = int(coeffs[k][i]) % q
                while x % q != r:
                    x += mod
                mod *= q
            assert x < M
            digs = _mixed_radix(x, primes)
            assert top[c * N + i] == (digs[0] + (1 << (d - 1))) >> d
            for k in range(1, len(primes)):
                assert digits[k - 1, c * N + i] == digs[k]


def test_pack_dropped_bit_layout(orc):
    # wire/mod.rs:386-397: top (top_bits) then each digit (residue_bits), LSB first
    primes, d = [3889, 2917, 4861], 8
    L, cols, a, w, raw = _bd_setup(orc, primes, d)
    top, digits = orc.drop_bits(primes, cols, raw, d)
    tb = orc.top_bits(primes[0], d)
    # bd/mod.rs:58-64 and wire/mod.rs:77-79 by their definitions
    assert tb == ((primes[0] - 1 + (1 << (d - 1))) >> d).bit_length()
    assert [orc.residue_bits(q) for q in primes[1:]] == [(q - 1).bit_length() for q in primes[1:]]
    bits = []
    for i in range(cols * N):
        fields = [(int(top[i]), tb)] + [(int(digits[k - 1, i]), (primes[k] - 1).bit_length())
                                        for k in range(1, len(primes))]
        for v, n in fields:
            bits += [(v >> b) & 1 for b in range(n)]
    want = bytearray((len(bits) + 7) // 8)
    for j, b in enumerate(bits):
        want[j // 8] |= b << (j % 8)
    assert orc.pack_dropped(primes, cols, d, top, digits) == bytes(want)


def test_bd_cap_formula(orc):
    # bd/mod.rs:71-74 evaluated by hand for two shapes: BD_CAP * N * columns * 28 * (4^d - 1) / 12
    assert orc.bd_cap(4, 8) == 24192 * 65535        # 4 * 72576 / 12 = 24192, exact
    assert orc.bd_cap(512, 8) == 3096576 * 65535    # 4 * 9289728 / 12 = 3096576, exact


@pytest.mark.parametrize("primes,d", [([3889, 2917, 4861], 8), ([3889, 9721], 6)])
def test_bd_residual_is_the_coefficient_domain_norm(orc, primes, d):
    # bd/residual.rs:77-130: the columns rebuilt from (top, digits), folded with the challenges,
    # subtracted from A v per prime; computed here in the coefficient domain with mul_mod_phi and a
    # big-integer CRT, then centered and squared
    L, cols, a, w, raw = _bd_setup(orc, primes, d)
    top, digits = orc.drop_bits(primes, cols, raw, d)
    ch = _challenges(orc, cols, b"oracle/bd")
    v = orc.fold(primes[0], L, cols, w, ch)
    got = orc.bd_residual(primes, L, cols, a, d, top, digits, ch, v)
    M = int(np.prod([int(q) for q in primes], dtype=object))
    cnt = cols * N
    rebuilt = []
    for j in range(cols):
        col = []
        for i in range(N):
            inner = 0
            for k in range(len(primes) - 1, 0, -1):
                inner = inner * primes[k] + int(digits[k - 1, j * N + i])
            col.append((int(top[j * N + i]) << d) + primes[0] * inner)
        rebuilt.append(col)
    z = []
    for k, q in enumerate(primes):
        acc = np.zeros(N, np.int64)
        for i in range(L):
            ai = orc.intt(q, (a[k, i].astype(np.int64) % q).astype(np.uint32))
            acc = (acc + orc.mul_mod_phi(ai, (v[i].astype(np.int64) % q).astype(np.uint32), q)) % q
        for j in range(cols):
            e = orc.embed(ch[j])
            cj = (np.asarray(e, np.int64) % q).astype(np.uint32)
            colq = np.array([x % q for x in rebuilt[j]], np.uint32)
            acc = (acc - orc.mul_mod_phi(cj, colq, q).astype(np.int64)) % q
        z.append(acc)
    want = 0
    for i in range(N):
        x, mod = 0, 1
        for k, q in enumerate(primes):
            r = int(z[k][i]) % q
            while x % q != r:
                x += mod
            mod *= q
        mag = M - x if x > M // 2 else x
        want += mag * mag
    assert got == want
    # honest: the residual is the rounding noise sum_j c_j e_j, far under the cap
    assert got <= orc.bd_cap(cols, d)


def test_bd_round_verifies_and_tamper_rejects(orc):
    # scheme/prover.rs:69-72 and verifier.rs:232-275 in miniature
    wl, cl, primes, d = 10, 2, [3889, 2917, 4861], 8
    seed = bytes([71]) * 32
    params = orc.make_params(wl, cl, primes, seed, dropped_bits=d)
    L, cols = (1 << (wl - cl)) // 4, 1 << cl
    a = orc.key_expand(primes, L, orc.matrix_seed(seed))
    w = orc.witness_random(bytes([73]) * 32, 1 << wl)
    raw, comp = orc.commit(primes, L, cols, a, w)
    top, digits = orc.drop_bits(primes, cols, raw, d)
    t = orc.Transcript(b"oracle/bd-round")
    p0, p1 = orc.derive_evaluation_point_dropped(t, params, orc.pack_dropped(primes, cols, d, top, digits))
    # the dropped arm absorbs different bytes than the matrix arm
    t2 = orc.Transcript(b"oracle/bd-round")
    q0, _ = orc.derive_evaluation_point(t2, params, comp)
    assert not np.array_equal(p0, q0)
    u = orc.row_evaluate(w, p0)
    ch = orc.derive_folding_challenges(t, u)
    v = orc.fold(primes[0], L, cols, w, ch)
    uf = orc.fold_binary(u, ch)
    assert orc.verify_opening_bd(params, a, top, digits, ch, v, p0, uf)
    bad = v.copy()
    bad[1, 5] += 1
    assert not orc.verify_opening_bd(params, a, top, digits, ch, bad, p0, uf)
    # one coefficient off by q0 q1 (the last digit): far over the cap. (A single top bit is a change
    # of 2^d, inside the noise the cap tolerates by design.)
    bad_digits = digits.copy()
    bad_digits[1, 17] = (int(bad_digits[1, 17]) + 1) % primes[2]
    assert not orc.verify_opening_bd(params, a, top, bad_digits, ch, v, p0, uf)
    uf_bad = uf.copy()
    uf_bad[1] ^= np.uint64(4)
    assert not orc.verify_opening_bd(params, a, top, digits, ch, v, p0, uf_bad)
