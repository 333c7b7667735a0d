"""The product's host side (C++ in liblabinius.so, no device needed) against the oracle: BLAKE3,
the seed expansions, the transcript and its challenge sampler, F162, parameter checks and the
verifier's pieces. Everything here is bit-exact."""
import ctypes

import numpy as np
import pytest


def test_library_exports_every_declared_symbol(lbn):
    import os
    import re
    root = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
    names = set()
    for h in ("labinius.h", "labinius_debug.h"):
        src = open(os.path.join(root, "include", h)).read()
        names |= set(re.findall(r"\b(lbn_[a-z0-9_]+)\s*\(", src))
    assert len(names) > 30
    L = lbn.lib()
    for name in sorted(names):
        assert hasattr(L, name), name


def test_blake3_matches_oracle(lbn, orc):
    for n in list(range(0, 200, 7)) + [1023, 1024, 1025, 2047, 2048, 2049, 4096, 4097, 65536 + 3, 200000]:
        data = bytes((i * 7 + 3) % 256 for i in range(n))
        assert lbn.blake3(data, 77) == orc.blake3(data, 77), n
        key = bytes(range(32))
        assert lbn.blake3(data, 40, key=key) == orc.blake3(data, 40, key=key)
        assert lbn.blake3(data, 40, context="ctx") == orc.blake3(data, 40, context="ctx")


def test_blake3_matches_spec_vectors(lbn):
    import json
    import os
    v = json.load(open(os.path.join(os.path.dirname(__file__), "golden", "blake3_spec_vectors.json")))
    for case in v["cases"]:
        data = bytes(i % 251 for i in range(case["input_len"]))
        assert lbn.blake3(data).hex() == case["hash"]
        if "keyed_hash" in case:
            assert lbn.blake3(data, key=v["key"].encode()).hex() == case["keyed_hash"]
        if "derive_key" in case:
            assert lbn.blake3(data, context=v["context"]).hex() == case["derive_key"]


def test_blake3_subtree_path_matches_oracle(lbn, orc):
    # large inputs take whole aligned subtrees (threaded from 1 MiB); a prefix shifts the alignment
    for n in (16 * 1024, 16 * 1024 + 1, 17 * 1024, 3_000_000, (1 << 20) + 5):
        d = bytes((i * 131 + 7) % 256 for i in range(n))
        assert lbn.blake3(d, 40) == orc.blake3(d, 40), n
    for pre in (1, 1024, 2047, 5000):
        t, o = lbn.Transcript(b"x"), orc.Transcript(b"x")
        t.absorb_bytes(bytes(pre))
        o.absorb_bytes(bytes(pre))
        big = bytes((i * 7) % 256 for i in range(2_500_000 + pre))
        t.absorb_bytes(big)
        o.absorb_bytes(big)
        assert t.fill(b"label", 64) == o.fill(b"label", 64), pre


def test_params_errors(lbn):
    # scheme/params.rs:179-195 and tests/api.rs:46-80
    P, M = lbn.Params, lbn.Modulus
    cases = [((5, 6, []), "ColumnsExceedWitness"), ((10, 0, []), "TooFewColumns"), ((10, 4, []), "ColumnTooShort"),
             ((10, 2, [M.Q9721_FS_S, M.Q9721_FS_S]), "DuplicateModulus"), ((10, 2, [M.Q3889_FS_S]), "BaseIsAlsoExtra")]
    for args, kind in cases:
        with pytest.raises(lbn.ParamError) as e:
            P(*args)
        assert e.value.kind == kind
    p = P.basic()
    assert (p.witness_log_len, p.column_log_len, p.primes()) == (18, 7, [3889, 9721])
    pl = P.sized(lbn.SUITES[2])
    assert pl.primes() == [3889, 2917, 4861] and pl.column_log_len == 9
    pxl = P.sized(lbn.SUITES[3])
    assert pxl.primes() == [3889, 2917, 9721] and pxl.column_log_len == 10


def test_witness_random_matches_oracle(lbn, orc):
    p = lbn.Params(11, 3, [lbn.Modulus.Q9721_FS_S])
    w = lbn.Witness.random(p, bytes([0xC7]) * 32)
    assert np.array_equal(w.elements, orc.witness_random(bytes([0xC7]) * 32, p.witness_len()))
    assert (w.elements[:, 2] >> np.uint64(34)).max() == 0
    w2 = lbn.Witness.random(p, bytes([0xC7]) * 32)
    assert np.array_equal(w.elements, w2.elements)
    with pytest.raises(lbn.WitnessError):
        lbn.Witness.from_elements(p, np.zeros((5, 3), np.uint64))


def _debug(lbn, params, seed, which, dtype):
    L = lbn.lib()
    n = ctypes.c_size_t()
    s = (ctypes.c_uint8 * 32).from_buffer_copy(bytes(seed))
    assert L.lbn_debug_tables(ctypes.byref(params.c()), s, which, None, 0, ctypes.byref(n)) == 0
    buf = np.zeros(n.value // np.dtype(dtype).itemsize, dtype)
    assert L.lbn_debug_tables(ctypes.byref(params.c()), s, which, buf.ctypes.data_as(ctypes.c_void_p),
                              ctypes.c_size_t(n.value), None) == 0
    return buf


@pytest.mark.parametrize("extra", [[9721], [2917, 4861], [12637, 17497, 19441]])
def test_key_expansion_matches_oracle(lbn, orc, extra):
    p = lbn.Params(10, 2, [lbn.Modulus(q) for q in extra])
    seed = bytes([0x5A]) * 32
    a = _debug(lbn, p, seed, 14, np.int16).reshape(len(p.primes()), p.len_ring(), 648)
    want = orc.key_expand(p.primes(), p.len_ring(), orc.matrix_seed(seed))
    assert np.array_equal(a, want)


def test_transcript_and_challenges_match_oracle(lbn, orc):
    a, b = lbn.Transcript(b"labinius/test"), orc.Transcript(b"labinius/test")
    for t in (a, b):
        t.absorb_bytes(b"hello")
        t.absorb_u64(12345)
    assert a.fill(b"label", 100) == b.fill(b"label", 100)
    ac, bc = a.clone(), b.clone()
    for _ in range(40):
        ca, na = a.sample_short_challenge()
        cb, nb = b.sample_short_challenge()
        assert bytes(ca.positions) == bytes(cb.positions) and ca.signs == cb.signs and na == nb
    assert ac.fill(b"x", 33) == bc.fill(b"x", 33)


def test_many_columns_point_and_challenges_match_oracle(lbn, orc):
    # 1024 columns: the commitment (2.7 MB) is absorbed in one update that takes whole subtrees on
    # threads, and the 1024 challenges are sampled on threads; both must equal the oracle's
    M = lbn.Modulus
    params = lbn.Params(17, 10, [M(9721)])
    seed = bytes([0x21]) * 32
    pp = lbn.PublicParameters.host_only(params, seed)
    verifier = lbn.Verifier(pp)
    primes, cols = params.primes(), params.columns()
    assert cols == 1024
    rng = np.random.default_rng(12)
    comp = rng.integers(-1500, 1501, (cols, 4, len(primes), 162)).astype(np.int16)
    commitment = lbn.Commitment(primes, cols, comp, None)
    t, ot = lbn.Transcript(b"labinius/test/cols"), orc.Transcript(b"labinius/test/cols")
    point = verifier.derive_evaluation_point(t, commitment)
    op0, op1 = orc.derive_evaluation_point(ot, orc.make_params(17, 10, primes, seed), comp)
    assert np.array_equal(point.p0, op0) and np.array_equal(point.p1, op1)
    u = rng.integers(0, 2**63, (cols, 3), dtype=np.uint64)
    u[:, 2] &= np.uint64((1 << 34) - 1)
    ch = verifier.derive_folding_challenges(t, lbn.RowEvaluation(u))
    och = orc.derive_folding_challenges(ot, u)
    assert len(ch.challenges) == cols
    for a, b in zip(ch.challenges, och):
        assert bytes(a.positions) == bytes(b.positions) and a.signs == b.signs


def test_f162_matches_oracle(lbn, orc):
    rng = np.random.default_rng(8)
    for _ in range(50):
        x, y = (rng.integers(0, 2**63, 3, dtype=np.uint64) for _ in range(2))
        x[2] &= np.uint64((1 << 34) - 1)
        y[2] &= np.uint64((1 << 34) - 1)
        assert np.array_equal(lbn.f162_mul(x, y), orc.f162_mul(x, y))


@pytest.mark.parametrize("extra,d", [((2917, 4861), 8), ((9721,), 10), ((2917, 12637), 7)])
def test_bit_dropped_host_side_matches_oracle(lbn, orc, extra, d):
    # bd/digits.rs:121-182, wire/mod.rs:386-397, verifier.rs:61-100 and 232-275, bd/residual.rs,
    # on a commitment the oracle computes; the verifier needs no device (host-only parameters)
    M = lbn.Modulus
    params = lbn.Params(10, 2, [M(q) for q in extra], opening=lbn.Opening.bit_dropped(d))
    seed = bytes([0x5A]) * 32
    pp = lbn.PublicParameters.host_only(params, seed)
    with pytest.raises(lbn.LabiniusError):
        lbn.Prover(pp)
    primes, L, cols = params.primes(), params.len_ring(), params.columns()
    a = pp.key()
    assert np.array_equal(a, orc.key_expand(primes, L, orc.matrix_seed(seed)))
    w = orc.witness_random(bytes([0x11]) * 32, params.witness_len())
    raw, comp = orc.commit(primes, L, cols, a, w)
    dropped = lbn.Dropped.of_raw(params, raw)
    top, digits = orc.drop_bits(primes, cols, raw, d)
    assert np.array_equal(dropped.top, top) and np.array_equal(dropped.digits, digits)
    packed = orc.pack_dropped(primes, cols, d, top, digits)
    assert dropped.pack() == packed and dropped.wire_bytes() == len(packed)
    assert lbn.bd_cap(params) == orc.bd_cap(cols, d)

    commitment = lbn.Commitment(primes, cols, comp, raw, dropped)
    verifier = lbn.Verifier(pp)
    t = lbn.Transcript(b"host/bd")
    point = verifier.derive_evaluation_point(t, commitment)
    ot = orc.Transcript(b"host/bd")
    p0, p1 = orc.derive_evaluation_point_dropped(ot, orc.make_params(10, 2, primes, seed, dropped_bits=d), packed)
    assert np.array_equal(point.p0, p0) and np.array_equal(point.p1, p1)
    u = orc.row_evaluate(w, p0)
    challenges = verifier.derive_folding_challenges(t, lbn.RowEvaluation(u))
    och = orc.derive_folding_challenges(ot, u)
    for c, o in zip(challenges.challenges, och):
        assert list(c.positions) == list(o.positions) and c.signs == o.signs and c.weight == o.weight
    v = orc.fold(primes[0], L, cols, w, och)
    folded = lbn.FoldedWitness(v)
    assert verifier.bd_residual(commitment, challenges, folded) == orc.bd_residual(primes, L, cols, a, d, top,
                                                                                   digits, och, v)
    uf = orc.fold_binary(u, och)
    verifier.verify_opening(commitment, challenges, point, None, folded, uf)
    bad = lbn.Dropped(params, top, digits.copy())
    bad.digits[-1, 5] = (int(bad.digits[-1, 5]) + primes[-1] // 2) % primes[-1]  # far over the cap
    with pytest.raises(lbn.VerificationError):
        verifier.verify_opening(lbn.Commitment(primes, cols, comp, raw, bad), challenges, point, None, folded, uf)
    v2 = v.copy()
    v2[0, 0] += 1
    with pytest.raises(lbn.VerificationError):
        verifier.verify_opening(commitment, challenges, point, None, lbn.FoldedWitness(v2), uf)
