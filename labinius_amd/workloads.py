"""The mixed-size workloads: the batched commitment and the serving request stream.

Both share one key: row_log_len = 11 and moduli {3889, 9721}, the key of Params::basic()
(scheme/params.rs:207-210). CommitmentKey::random (key.rs:48-88) depends only on the row length,
the seed and the moduli, so every witness Params::new(e, e - 11, [9721], mode) commits under the
same A, and each result is checkable against that witness's own round.
"""
from . import Modulus, Opening, Params, Witness, blake3

ROW_LOG_LEN = 11
BATCH_SIZE = 64
BATCH_EXPONENTS = range(16, 23)
SERVING_SIZES = [(16, 0.40), (18, 0.30), (20, 0.20), (22, 0.08), (24, 0.02)]
# dropped bits by size, following the SUITES pattern (scheme/params.rs:91-128), 10 for 2^16
SERVING_DROPPED = {16: 10, 18: 10, 20: 9, 22: 8, 24: 7}
MASK64 = (1 << 64) - 1


class Rng:
    """xorshift64* (rng.rs:1-20)."""

    def __init__(self, seed):
        self.s = (seed * 0x9E3779B97F4A7C15 | 1) & MASK64

    def next(self):
        x = self.s
        x ^= x >> 12
        x ^= (x << 25) & MASK64
        x ^= x >> 27
        self.s = x
        return (x * 0x2545F4914F6CDD1D) & MASK64

    def below(self, n):
        return ((self.next() >> 32) * n) >> 32

    def uniform(self):
        return (self.next() >> 11) / float(1 << 53)


def key_params():
    """The shared key's own shape (Params::basic())."""
    return Params.basic()


def witness_params(e, opening=Opening.CLEAR):
    return Params(e, e - ROW_LOG_LEN, [Modulus.Q9721_FS_S], opening=opening)


def witness_seed(tag, i):
    return blake3(tag + int(i).to_bytes(8, "little"))


def batch_specs(seed=0x5EEDBA7C, count=BATCH_SIZE):
    """(exponent, lifted) per witness: exponents uniform over {16..22}; every fourth a B128 trace."""
    rng = Rng(seed)
    return [(BATCH_EXPONENTS[0] + rng.below(len(BATCH_EXPONENTS)), i % 4 == 3) for i in range(count)]


def batch_witness(i, e, lifted):
    p = witness_params(e)
    w = Witness.random(p, witness_seed(b"labinius/batch", i))
    if lifted:
        return Witness.lifted(p, w.elements[:, :2])
    return w


def serving_sizes_stratified(count, seed=0x57A71F1):
    """Exponents in the serving mix's exact proportions (at least one of each size), in a seeded
    random order: 40 % 2^16, 30 % 2^18, 20 % 2^20, 8 % 2^22, 2 % 2^24."""
    rng = Rng(seed)
    counts = [max(1, int(round(p * count))) for _, p in SERVING_SIZES]
    counts[0] += count - sum(counts)
    out = [e for (e, _), c in zip(SERVING_SIZES, counts) for _ in range(c)]
    for i in range(len(out) - 1, 0, -1):  # Fisher-Yates
        j = rng.below(i + 1)
        out[i], out[j] = out[j], out[i]
    return out


def serving_requests_stratified(rate_per_s, count, seed=0x5E4F1CE):
    """Poisson arrivals at rate_per_s (None: all at once) with stratified sizes and Clear or
    bit-dropped at random: (arrival seconds, exponent, dropped bits or 0, index)."""
    import math
    rng = Rng(seed)
    sizes = serving_sizes_stratified(count, seed ^ 0x5A5A)
    t, out = 0.0, []
    for i, e in enumerate(sizes):
        if rate_per_s:
            t += -math.log(1.0 - rng.uniform()) / rate_per_s
        d = SERVING_DROPPED[e] if rng.below(2) else 0
        out.append((t, e, d, i))
    return out


def serving_requests(rate_per_s, count, seed=0x5E4F1CE):
    """Poisson arrivals at rate_per_s: (arrival seconds, exponent, dropped bits or 0, index)."""
    import math
    rng = Rng(seed)
    t, out = 0.0, []
    for i in range(count):
        t += -math.log(1.0 - rng.uniform()) / rate_per_s
        u, acc, e = rng.uniform(), 0.0, SERVING_SIZES[-1][0]
        for size, pr in SERVING_SIZES:
            acc += pr
            if u < acc:
                e = size
                break
        d = SERVING_DROPPED[e] if rng.below(2) else 0
        out.append((t, e, d, i))
    return out
