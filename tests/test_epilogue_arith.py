"""The GEMM epilogue's reductions mod q (labinius_amd/csrc/kernels.hip: mod_q_classes,
mod_q_short, mod_q_unsigned) restated in exact rational arithmetic with IEEE-754 rounding (an fma
rounds the exact a*b + c once; rint rounds half to even), checked on the sums that decide them:
random class values at the largest sizes each form allows, and z placed next to every half-integer
quotient, where the rounded quotient can land on either side. Each must give r = z mod q with
|r| <= (q + 1) / 2, the bound the E.Z sums' int32 partials rely on."""
import random
from fractions import Fraction

import pytest

M = 6755399441055744.0  # 1.5 x 2^52
MODULI_UNSIGNED = [2917, 3889, 4861, 9721, 12637, 16381]
MODULI_CENTRED = [17497, 19441]


def fl(x: Fraction) -> float:
    """x rounded to the nearest double (ties to even), as an fma's single rounding does"""
    return float(x)


def fma(a: float, b: float, c: float) -> float:
    return fl(Fraction(a) * Fraction(b) + Fraction(c))


def rint(x: float) -> float:
    return float(round(x))  # Python rounds half to even, as v_rndne_f64


def classes(ll: int, mid: int, hh: int, q: int) -> int:
    """mod_q_classes: the exact f64 sum, its rounded quotient, the fma remainder"""
    qd, qinv = float(q), 1.0 / q
    z = fma(16384.0, float(hh), fma(128.0, float(mid), float(ll)))
    assert z == ll + 128 * mid + 16384 * hh  # exact below 2^53
    t = rint(z * qinv)
    return int(fma(-t, qd, z))


def offset(D: float, q: int) -> int:
    """mod_q_offset: D = M + z; quotient by one fma with k0 = fl(M / q); M + r, r its low word"""
    qd, qinv = float(q), 1.0 / q
    k0 = M * qinv
    t = rint(fma(D, qinv, -k0))
    rp = fma(-t, qd, D)
    assert 2.0**52 <= rp < 2.0**53 and rp == int(rp)  # unit 1: the low word of its bits is r
    return int(rp) - int(M)


def short(ll: int, mid: int, hh: int, q: int) -> int:
    w = (ll + (mid << 7)) & 0xFFFFFFFF
    assert w == ll + 128 * mid  # the u32 sum is exact for the short form's K
    D = M + w  # M's bits with w in the low word
    return offset(fma(16384.0, float(hh), D), q)


def unsigned(ll: int, mid: int, hh: int, q: int) -> int:
    D = fma(128.0, float(mid), M + ll)
    return offset(fma(16384.0, float(hh), D), q)


def check(r: int, z: int, q: int) -> None:
    assert (r - z) % q == 0
    assert abs(r) <= (q + 1) // 2


def near_half(q: int, zmax: int, rng: random.Random, n: int):
    """z with z / q within one unit of a half-integer, spread up to zmax"""
    for _ in range(n):
        k = rng.randrange(zmax // q)
        for d in (-1, 0, 1):
            z = k * q + (q - 1) // 2 + d
            if 0 <= z < zmax:
                yield z


def split(z: int, hh_max: int, mid_max: int, rng: random.Random):
    """non-negative (ll, mid, hh) with ll + 2^7 mid + 2^14 hh = z inside the bounds"""
    hh = min(z >> 14, hh_max)
    rest = z - (hh << 14)
    mid = min(rest >> 7, mid_max)
    if mid and rng.random() < 0.5:
        mid -= rng.randrange(min(mid, 4) + 1)
    return rest - (mid << 7), mid, hh


@pytest.mark.parametrize("q", MODULI_UNSIGNED)
def test_short_form(q):
    """K <= 16 steps: each class at most K x 127^2 (mid twice that), ll + 2^7 mid < 2^32"""
    rng = random.Random(q)
    K = 16 * 64
    lo, mi = K * 127 * 127, 2 * K * 127 * 127
    zmax = lo + 128 * mi + 16384 * lo
    for _ in range(300):
        ll, mid, hh = rng.randrange(lo + 1), rng.randrange(mi + 1), rng.randrange(lo + 1)
        check(short(ll, mid, hh, q), ll + 128 * mid + 16384 * hh, q)
    check(short(lo, mi, lo, q), zmax, q)
    for z in near_half(q, zmax, rng, 300):
        ll, mid, hh = split(z, lo, mi, rng)
        if ll <= lo:
            check(short(ll, mid, hh, q), z, q)


@pytest.mark.parametrize("q", MODULI_UNSIGNED)
def test_unsigned_form(q):
    """any K the launcher allows (the int32 classes' bound, 1040 steps)"""
    rng = random.Random(q + 1)
    cap = 2**31 - 1
    zmax = cap + 128 * cap + 16384 * cap
    for _ in range(300):
        ll, mid, hh = rng.randrange(cap), rng.randrange(cap), rng.randrange(cap)
        check(unsigned(ll, mid, hh, q), ll + 128 * mid + 16384 * hh, q)
    check(unsigned(cap, cap, cap, q), zmax, q)
    for z in near_half(q, zmax, rng, 300):
        ll, mid, hh = split(z, cap, cap, rng)
        if ll <= cap:
            check(unsigned(ll, mid, hh, q), z, q)


@pytest.mark.parametrize("q", MODULI_UNSIGNED + MODULI_CENTRED)
def test_classes_form(q):
    """the general form, signed classes (centred digits above 2^14)"""
    rng = random.Random(q + 2)
    cap = 2**31 - 1
    for _ in range(300):
        ll, mid, hh = (rng.randrange(-cap, cap) for _ in range(3))
        check(classes(ll, mid, hh, q), ll + 128 * mid + 16384 * hh, q)
    for z in near_half(q, cap * 16384, rng, 200):
        hh, rest = divmod(z, 16384)
        check(classes(rest, 0, hh, q), z, q)
        check(classes(-rest, 0, -hh, q), -z, q)


def test_a_plausible_slip_fails():
    """the quotient's fma without its -k0 term (a dropped term) leaves M / q in t: the remainder
    leaves [2^52, 2^53) and is no longer z mod q, which the checks above must catch"""
    q, z = 9721, 9721 * 123456 + 4860
    qd, qinv = float(q), 1.0 / q
    t = rint(fma(M + z, qinv, 0.0))
    rp = fma(-t, qd, M + z)
    assert not (2.0**52 <= rp < 2.0**53)
    with pytest.raises(AssertionError):
        check(int(rp) - int(M), z, q)
