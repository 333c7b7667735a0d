"""lbn_row_evaluate_noalloc: the row evaluation without a prover and without device allocation,
its eq tensor in the library's static device scratch. Checked through the C ABI against
lbn_row_evaluate (which is pinned to the oracle) on two calls in a row on the same scratch, the
second with a smaller witness and a shorter row, and against the oracle; sizes past the scratch
are refused."""
import ctypes
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


def _point(rows, seed):
    rng = np.random.default_rng(seed)
    p = rng.integers(0, 2**63, (rows, 3), dtype=np.uint64) * np.uint64(2) + rng.integers(0, 2, (rows, 3),
                                                                                           dtype=np.uint64)
    p[:, 2] &= np.uint64((1 << 34) - 1)
    return p


def _noalloc(lbn, w_dev, n, p0, cols):
    import torch
    L = lbn.lib()
    d_p0 = torch.from_numpy(p0.view(np.int64).copy()).cuda()
    d_u = torch.full((cols * 3,), -1, dtype=torch.int64, device="cuda")  # stale contents are overwritten
    torch.cuda.synchronize()
    rc = L.lbn_row_evaluate_noalloc(ctypes.c_void_p(w_dev), ctypes.c_size_t(n), ctypes.c_void_p(d_p0.data_ptr()),
                                    ctypes.c_size_t(len(p0)), ctypes.c_void_p(d_u.data_ptr()), ctypes.c_void_p(None))
    assert rc == 0, L.lbn_last_error()
    torch.cuda.synchronize()
    return d_u.cpu().numpy().view(np.uint64).reshape(cols, 3)


@pytest.mark.parametrize("first,second", [((16, 4), (12, 5)), ((18, 3), (13, 6))])
def test_noalloc_row_evaluation_matches_prover(lbn, orc, first, second):
    L = lbn.lib()
    for k, (wl, cl) in enumerate((first, second)):
        say(f"case: noalloc row evaluation call {k + 1}: wl={wl} cl={cl}")
        params = lbn.Params(wl, cl, [lbn.Modulus.Q9721_FS_S])
        prover = lbn.Prover(lbn.PublicParameters.from_seed(params, MATRIX_SEED))
        w = lbn.Witness.random(params, bytes([wl, cl]) * 16)
        rw = prover.upload(w)
        p0 = _point(wl - cl, 100 * wl + cl)
        want = prover.row_evaluate(rw, lbn.EvaluationPoint(p0, _point(cl, 1)))
        got = _noalloc(lbn, L.lbn_dwitness_device_ptr(ctypes.c_void_p(rw._h)), params.witness_len(), p0,
                       params.columns())
        assert np.array_equal(got, want.values)
        sample = [0, params.columns() - 1]
        clen = params.column_len()
        for c in sample:
            assert np.array_equal(got[c], orc.row_evaluate(w.elements[c * clen:(c + 1) * clen], p0)[0])


def test_noalloc_rejects_rows_past_the_scratch(lbn):
    import torch
    say("case: noalloc limits")
    L = lbn.lib()
    buf = torch.zeros(3 << 16, dtype=torch.int64, device="cuda")
    p = torch.zeros(3 * 16, dtype=torch.int64, device="cuda")
    u = torch.zeros(6, dtype=torch.int64, device="cuda")
    ptr = lambda t: ctypes.c_void_p(t.data_ptr())
    # 2^16 elements in one row of 2^16: row_log_len 16 is past the scratch's 15
    assert L.lbn_row_evaluate_noalloc(ptr(buf), ctypes.c_size_t(1 << 16), ptr(p), ctypes.c_size_t(16), ptr(u),
                                      ctypes.c_void_p(None)) == 8
    # rows shorter than 2^7, and a witness that is not whole rows
    assert L.lbn_row_evaluate_noalloc(ptr(buf), ctypes.c_size_t(1 << 10), ptr(p), ctypes.c_size_t(6), ptr(u),
                                      ctypes.c_void_p(None)) == 8
    assert L.lbn_row_evaluate_noalloc(ptr(buf), ctypes.c_size_t(3 << 7), ptr(p), ctypes.c_size_t(8), ptr(u),
                                      ctypes.c_void_p(None)) == 8
    assert L.lbn_row_evaluate_noalloc(None, ctypes.c_size_t(1 << 10), ptr(p), ctypes.c_size_t(8), ptr(u),
                                      ctypes.c_void_p(None)) == 12
