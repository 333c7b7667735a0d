"""CPU replay of the GEMM tiles (tests/tile_replay.py) against the oracle: the host-built tables,
fragment layouts and epilogue weights reproduce the commitment's raw rows and the fold's
transform bit for bit, for splitting and quadratic limbs alike."""
import numpy as np
import pytest

from tests import tile_replay as tr


@pytest.mark.parametrize("extra", [[9721], [2917], [4861, 12637], [17497, 19441]])
def test_commit_replay_matches_oracle(lbn, orc, extra):
    p = lbn.Params(9, 1, [lbn.Modulus(q) for q in extra])
    seed = bytes([7]) * 32
    w = orc.witness_random(bytes([8]) * 32, p.witness_len())
    raw = tr.replay_commit(lbn, p, seed, w)
    a = orc.key_expand(p.primes(), p.len_ring(), orc.matrix_seed(seed))
    want, _ = orc.commit(p.primes(), p.len_ring(), p.columns(), a, w)
    assert np.array_equal(raw, want)


@pytest.mark.parametrize("base", [3889, 2917, 4861, 19441])
def test_fold_replay_matches_oracle(lbn, orc, base):
    p = lbn.Params.with_base(9, 2, lbn.Modulus(base), [])
    seed = bytes([7]) * 32
    w = orc.witness_random(bytes([9]) * 32, p.witness_len())
    t = orc.Transcript(b"replay")
    ch = [t.sample_short_challenge()[0] for _ in range(p.columns())]
    vhat = tr.replay_fold(lbn, p, seed, w, ch)
    v = np.array([orc.intt(base, vhat[i].astype(np.uint32)) for i in range(p.len_ring())]).astype(np.int64)
    v = np.where(v > (base - 1) // 2, v - base, v)
    want = orc.fold(base, p.len_ring(), p.columns(), w, ch)
    assert np.array_equal(v, want)
