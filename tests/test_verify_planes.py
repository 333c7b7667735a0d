"""The identity the device verifier rests on, checked through the oracle on the CPU.

The device computes A v (verifier.rs:205-224) by committing bit planes: with u = v + half
(half = (q0 - 1) / 2, so 0 <= u < q0 for an admissible v) split into binary ring elements u_b,
A v = sum_b 2^b A u_b - half A J (J the all-ones ring element), and the decomposition into the
four R_162 components is linear, so it commutes with the sum. The planes travel as F162 columns in
the commitment's lift (f162.rs:32-41), and the verifier's row check (eval.rs:202-249) is the row
value of the column v mod 2. Here the oracle's commit of those columns, combined, is compared with
the oracle's own A v, and the oracle verifier accepts the parity column's row value.
"""
import numpy as np
import pytest

N = 648


def lift_bits(bits):
    """[len_ring][648] bits -> F162 elements [4 len_ring][3]: element 4 r + k holds coefficient
    4 m + k of ring element r at bit m (f162.rs:32-41)."""
    b = np.asarray(bits, np.uint64).reshape(-1, 162, 4).transpose(0, 2, 1).reshape(-1, 162)
    out = np.zeros((b.shape[0], 3), np.uint64)
    for w in range(3):
        seg = b[:, 64 * w:min(64 * w + 64, 162)]
        out[:, w] = (seg << np.arange(seg.shape[1], dtype=np.uint64)).sum(axis=1, dtype=np.uint64)
    return out


def planes(v, q0):
    """The device verifier's columns for v: bits of v + half, then J, then v mod 2."""
    half = (q0 - 1) // 2
    nb = (q0 - 1).bit_length()
    u = v.astype(np.int64) + half
    cols = [lift_bits((u >> b) & 1) for b in range(nb)]
    cols.append(lift_bits(np.ones_like(v)))
    cols.append(lift_bits(v.astype(np.int64) & 1))
    return np.concatenate(cols), nb


def combine(comp, nb, half, primes):
    """sum_b 2^b comp_b - half comp_J, per limb, centered."""
    acc = -half * comp[nb].astype(np.int64)
    for b in range(nb):
        acc += comp[b].astype(np.int64) << b
    out = np.zeros(acc.shape, np.int16)
    for k, q in enumerate(primes):
        x = np.mod(acc[:, k], q)
        out[:, k] = np.where(x > (q - 1) // 2, x - q, x)
    return out


def test_lift_bits_is_the_oracle_lift(orc):
    rng = np.random.default_rng(3)
    bits = rng.integers(0, 2, (3, N))
    e = lift_bits(bits)
    for r in range(3):
        assert np.array_equal(orc.lift(e[4 * r:4 * r + 4]), bits[r])


@pytest.mark.parametrize("primes", [[3889, 9721], [2917, 3889, 19441], [19441, 4861]])
def test_planes_commit_to_a_times_v(orc, primes):
    L = 8
    a = orc.key_expand(primes, L, 99)
    rng = np.random.default_rng(len(primes))
    q0 = primes[0]
    half = (q0 - 1) // 2
    cases = [np.rint(rng.normal(0, 60, (L, N))).astype(np.int16)]
    sparse = np.zeros((L, N), np.int16)
    sparse[0, 0], sparse[3, 100], sparse[7, 647], sparse[5, 5] = half, -half, half, -1
    cases.append(sparse)
    cases.append(rng.integers(-half, half + 1, (L, N)).astype(np.int16))
    for v in cases:
        w, nb = planes(v, q0)
        _, comp = orc.commit(primes, L, nb + 2, a, w)
        assert np.array_equal(combine(comp, nb, half, primes), orc.a_times_v(primes, L, a, v))


def test_parity_column_is_the_row_check(orc):
    # an honest round: the oracle verifier accepts the parity column's row value as u'
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
    cols, nb = planes(v, primes[0])
    par_row = orc.row_evaluate(cols, p0)[nb + 1]
    assert np.array_equal(par_row, orc.fold_binary(u, ch))
    assert orc.verify_folded_opening(params, a, fc, v, p0, par_row)
    assert np.array_equal(orc.a_times_v(primes, L, a, v), fc)
