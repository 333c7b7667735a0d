"""Parity of the device wire forms (through the C ABI) with the oracle: the commitment packed on
the GPU from what the commit left in HBM (the matrix, or the bit-dropped digits) against the
oracle's pack of the same matrix / digits, and the rANS coding of the folded witness on the GPU
against the oracle's encoder, byte for byte, then decoded on the host. The messages cover honest
folds at several shapes and bases, the reference's adversarial cases (tests/wire.rs:200-440), any
i16 (the escape symbol, a symbol table too large for LDS), every tail-group length, and the size
of the largest single-GPU workload."""
import sys

import numpy as np
import pytest

from tests.test_wire import _adversarial_cases, _any_i16, _lcg_message, _one_lane_message

pytestmark = pytest.mark.gpu

N = 648
MATRIX_SEED = bytes([21]) * 32
WITNESS_SEED = bytes([23]) * 32


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module", autouse=True)
def _device():
    from tests.conftest import gpu_available
    if not gpu_available():
        pytest.skip("no GPU")


def _round(lbn, params):
    pp = lbn.PublicParameters.from_seed(params, MATRIX_SEED)
    prover = lbn.Prover(pp)
    w = lbn.Witness.random(params, WITNESS_SEED)
    commitment, opening = prover.commit(w)
    verifier = lbn.Verifier(pp)
    t = lbn.Transcript(b"labinius/test/wire")
    point = verifier.derive_evaluation_point(t, commitment)
    row = prover.row_evaluate(w, point)
    challenges = verifier.derive_folding_challenges(t, row)
    return prover, commitment, opening, challenges, row


@pytest.fixture(scope="module")
def prover(lbn):
    params = lbn.Params.with_base(10, 2, lbn.Modulus.Q3889_FS_S, [lbn.Modulus.Q9721_FS_S])
    return lbn.Prover(lbn.PublicParameters.from_seed(params, MATRIX_SEED))


@pytest.mark.parametrize("wl,cl,base,extra", [(11, 3, 3889, [9721]), (10, 1, 19441, [2917, 4861]),
                                              (12, 4, 2917, [17497, 12637, 3889])])
def test_commitment_packed_on_device(lbn, orc, wl, cl, base, extra):
    from labinius_amd import wire
    say(f"case: device pack wl={wl} cl={cl} base={base} extra={extra}")
    params = lbn.Params.with_base(wl, cl, lbn.Modulus(base), [lbn.Modulus(q) for q in extra])
    prover, commitment, opening, challenges, row = _round(lbn, params)
    b = wire.pack_commitment(prover)
    assert b == orc.pack_commitment(params.primes(), params.columns(), commitment.components)
    assert len(b) == wire.commitment_bytes(params)
    assert np.array_equal(wire.unpack_commitment(params, b), commitment.components)
    rb = wire.pack_row_evaluation(row)
    assert rb == orc.pack_f162(row.values)
    assert np.array_equal(wire.unpack_row_evaluation(rb, params.columns()).values, row.values)
    prover.recycle(opening)


@pytest.mark.parametrize("extra,d", [([9721], 3), ([2917, 4861], 5)])
def test_dropped_commitment_packed_on_device(lbn, orc, extra, d):
    from labinius_amd import wire
    say(f"case: device pack bit-dropped extra={extra} d={d}")
    params = lbn.Params.with_base(11, 3, lbn.Modulus.Q3889_FS_S, [lbn.Modulus(q) for q in extra],
                                  lbn.Opening.bit_dropped(d))
    pp = lbn.PublicParameters.from_seed(params, MATRIX_SEED)
    prover = lbn.Prover(pp)
    w = lbn.Witness.random(params, WITNESS_SEED)
    commitment, opening = prover.commit(w)  # computes the digits on the device (lbn_commit_drop_bits)
    prover.recycle(opening)
    expect = orc.pack_dropped(params.primes(), params.columns(), d, commitment.dropped.top,
                              commitment.dropped.digits)
    host = lbn.Dropped.of_raw(params, commitment.raw)
    assert np.array_equal(host.top, commitment.dropped.top) and np.array_equal(host.digits, commitment.dropped.digits)
    assert wire.pack_commitment(prover) == expect
    # a commitment with no outputs: the packer computes the digits itself
    rw = prover.upload(w)
    _, opening = prover.commit_resident(rw, want_outputs=False)
    prover.recycle(opening)
    assert wire.pack_commitment(prover) == expect
    got = wire.unpack_commitment(params, expect)
    assert np.array_equal(got.top, host.top) and np.array_equal(got.digits, host.digits)


@pytest.mark.parametrize("wl,cl,base,extra", [(11, 3, 3889, [9721]), (12, 2, 19441, [3889]),
                                              (14, 6, 2917, [3889])])
def test_fold_encoded_on_device(lbn, orc, wl, cl, base, extra):
    from labinius_amd import wire
    say(f"case: device encode of a fold wl={wl} cl={cl} base={base}")
    params = lbn.Params.with_base(wl, cl, lbn.Modulus(base), [lbn.Modulus(q) for q in extra])
    prover, commitment, opening, challenges, row = _round(lbn, params)
    folded = prover.fold(opening, challenges)
    b = wire.encode(prover, base)
    assert b == orc.rans_encode(folded.elements, base)
    assert np.array_equal(wire.decode(b).elements, folded.elements)
    assert len(b) < folded.elements.size * 2


def _messages():
    out = [("empty", np.zeros((0, N), np.int16)), ("any i16", _any_i16())]
    for q in [3889, 19441]:
        for i, v in enumerate(_adversarial_cases(q)):
            out.append((f"adversarial q={q} case={i}", v))
    state = 0x123456789ABCDEF1
    for count in [1, 2, 3, 5, 7, 8]:
        v, state = _lcg_message(count, state)
        out.append((f"lcg count={count}", v))
    for k in [0, 31, 63]:
        out.append((f"one lane k={k}", _one_lane_message(k)))
    rng = np.random.default_rng(5)
    # occupied ranges between the LDS symbol table's 8192 and the LDS histogram's 12288, and past it
    out.append(("range 10000", (rng.integers(-5000, 5000, (40, N))).astype(np.int16)))
    out.append(("range 30000", (rng.integers(-15000, 15000, (40, N))).astype(np.int16)))
    return out


def test_messages_encoded_on_device(lbn, orc, prover):
    from labinius_amd import wire
    for name, v in _messages():
        say(f"case: device encode {name} shape={v.shape}")
        b = wire.encode_message(prover, v, 3889)
        assert b == orc.rans_encode(v, 3889), name
        assert np.array_equal(wire.decode(b).elements, v), name
    # the header's representation flag
    v = _lcg_message(2, 9)[0]
    b = wire.encode_message(prover, v, 9721, ntt=True)
    assert b == orc.rans_encode(v, 9721, ntt=True)


def test_large_fold_encoded_on_device(lbn, orc, prover):
    # a folded witness the size of the largest single-GPU workload's (512 ring elements), a
    # discrete Gaussian of sigma 120 then 56, in a row (the scratch is reused); the size bound is
    # tests/wire.rs:322-328
    from labinius_amd import wire
    rng = np.random.default_rng(77)
    for sigma in [120, 56]:
        say(f"case: device encode 512 elements sigma={sigma}")
        v = np.rint(rng.normal(0, sigma, (512, N))).astype(np.int16)
        b = wire.encode_message(prover, v, 19441)
        assert b == orc.rans_encode(v, 19441)
        assert np.array_equal(wire.decode(b).elements, v)
        assert len(b) <= wire.entropy_bytes(v) * 1.02 + 4 * 64 + 200
