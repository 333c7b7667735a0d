"""Serving commit -> evaluate -> fold requests of mixed sizes under one key, with continuous
batching, through the library's native server (lbn_server_*, labinius_amd/csrc/server.cpp).

The server's scheduler thread admits queued requests in arrival order while their device memory
(the resident witness and the fold's byte layout, from the server's device pool) fits the HBM
budget, uploads them and commits every admitted request in one batched launch: requests that
arrive while earlier ones are on the GPU join the next batch. Worker threads, each with its own
provers and HIP streams, run the rest of every round (bit-dropped digits, the evaluation point,
the row evaluation and its claim, the folding challenges, the fold of the still-resident
witness), overlapping on the device with each other and with the next batch. A request's
transcript starts from its own domain, so its result does not depend on what else is in flight.

The public matrix A stays in HBM for the server's lifetime; nothing is allocated or freed per
request once the pool has warmed up. `Server(..., serial=True)` runs the same stream one request at
a time with nothing beside it: the unbatched comparison.
"""
import ctypes
import time

import numpy as np

from . import (N, N162, Commitment, Dropped, EvaluationPoint, FoldedWitness, FoldingChallenges, LabiniusError,
               Opening, RowEvaluation, ShortChallenge, _check, _p, lib)
from .workloads import key_params, witness_params

LBN_ERR_ADMISSION = 17
LBN_SERVER_FAIL_FOLD = 1


class AdmissionError(LabiniusError):
    """The request alone needs more device memory than the server's budget (or more columns than one
    batched commitment takes)."""


class _Config(ctypes.Structure):
    _fields_ = [("hbm_budget", ctypes.c_uint64), ("workers", ctypes.c_uint32), ("serial", ctypes.c_uint32),
                ("max_batch_columns", ctypes.c_uint32), ("n_sizes", ctypes.c_uint32), ("sizes", ctypes.c_uint32 * 8)]


class _Outputs(ctypes.Structure):
    _fields_ = [("components", ctypes.c_void_p), ("raw", ctypes.c_void_p), ("top", ctypes.c_void_p),
                ("digits", ctypes.c_void_p), ("p0", ctypes.c_void_p), ("p1", ctypes.c_void_p), ("u", ctypes.c_void_p),
                ("claim", ctypes.c_void_p), ("challenges", ctypes.c_void_p), ("v", ctypes.c_void_p)]


class _Info(ctypes.Structure):
    _fields_ = [("status", ctypes.c_int), ("arrival_s", ctypes.c_double), ("admitted_s", ctypes.c_double),
                ("committed_s", ctypes.c_double), ("done_s", ctypes.c_double), ("bytes", ctypes.c_uint64),
                ("batch", ctypes.c_uint32)]


class _Stats(ctypes.Structure):
    _fields_ = [(k, ctypes.c_uint64) for k in ("requests", "failed", "batches", "batched_requests", "batched_columns",
                                                "reserved_peak", "fixed_bytes", "pool_obtained", "pool_mallocs",
                                                "pool_reuses")]


class Result:
    def __init__(self, commitment, point, row, claimed, challenges, folded):
        self.commitment = commitment
        self.point = point
        self.row = row
        self.claimed = claimed
        self.challenges = challenges
        self.folded = folded

    def same_as(self, other):
        a, b = self, other
        if not (np.array_equal(a.commitment.raw, b.commitment.raw)
                and np.array_equal(a.commitment.components, b.commitment.components)):
            return False
        if (a.commitment.dropped is None) != (b.commitment.dropped is None):
            return False
        if a.commitment.dropped is not None and not (
                np.array_equal(a.commitment.dropped.top, b.commitment.dropped.top)
                and np.array_equal(a.commitment.dropped.digits, b.commitment.dropped.digits)):
            return False
        return (np.array_equal(a.point.p0, b.point.p0) and np.array_equal(a.point.p1, b.point.p1)
                and np.array_equal(a.row.values, b.row.values) and np.array_equal(a.claimed, b.claimed)
                and all(x == y for x, y in zip(a.challenges.challenges, b.challenges.challenges))
                and np.array_equal(a.folded.elements, b.folded.elements))


class Request:
    """One submitted request: its witness and output buffers stay referenced here until it is done."""

    def __init__(self, rid, e, dropped_bits, witness, params=None):
        self.id = rid
        self.e = e
        self.dropped_bits = dropped_bits
        self.witness = witness
        self._params = params
        p = self.params()
        limbs, cols = len(p.primes()), p.columns()
        self.w = np.ascontiguousarray(witness.elements, np.uint64)
        self.comp = np.zeros((cols, 4, limbs, N162), np.int16)
        self.raw = np.zeros((limbs, cols, N), np.uint32)
        self.top = np.zeros(cols * N, np.uint16)
        self.digits = np.zeros((max(1, limbs - 1), cols * N), np.uint16)
        self.p0 = np.zeros((p.row_log_len(), 3), np.uint64)
        self.p1 = np.zeros((p.column_log_len, 3), np.uint64)
        self.u = np.zeros((cols, 3), np.uint64)
        self.claim = np.zeros(3, np.uint64)
        self.ch = (ShortChallenge * cols)()
        self.v = np.zeros((p.len_ring(), N), np.int16)
        self.ticket = None
        self.arrival = self.admitted = self.committed = self.done = None
        self.bytes = 0
        self.batch = None
        self.status = None
        self.error = None
        self.result = None

    def params(self):
        if self._params is not None:
            return self._params
        return witness_params(self.e, Opening.bit_dropped(self.dropped_bits) if self.dropped_bits else Opening.CLEAR)

    def outputs(self):
        o = _Outputs()
        o.components, o.raw, o.top, o.digits = (self.comp.ctypes.data, self.raw.ctypes.data, self.top.ctypes.data,
                                                self.digits.ctypes.data)
        o.p0, o.p1, o.u, o.claim = self.p0.ctypes.data, self.p1.ctypes.data, self.u.ctypes.data, self.claim.ctypes.data
        o.challenges = ctypes.addressof(self.ch)
        o.v = self.v.ctypes.data
        return o

    def _collect(self, info):
        self.status = info.status
        self.arrival, self.admitted, self.committed, self.done = (info.arrival_s, info.admitted_s, info.committed_s,
                                                                 info.done_s)
        self.bytes = info.bytes
        self.batch = info.batch
        if info.status != 0:
            self.error = LabiniusError(info.status, "the request failed in the server")
            return
        p = self.params()
        dropped = Dropped(p, self.top, self.digits) if self.dropped_bits else None
        self.result = Result(Commitment(p.primes(), p.columns(), self.comp, self.raw, dropped),
                             EvaluationPoint(self.p0, self.p1), RowEvaluation(self.u), self.claim,
                             FoldingChallenges(list(self.ch)), FoldedWitness(self.v))

    def latency_ms(self):
        return 1e3 * (self.done - self.arrival)


class Server:
    """lbn_server: sizes are the witness_log_len set up in advance (provers, autotuned batch widths)."""

    def __init__(self, matrix_seed, hbm_budget_bytes=200e9, sizes=(16, 18, 20, 22, 24), workers=4, serial=False,
                 max_batch_columns=16384):
        self.seed = bytes(matrix_seed)
        self.budget = int(hbm_budget_bytes)
        self.sizes = tuple(sizes)
        cfg = _Config()
        cfg.hbm_budget = self.budget
        cfg.workers = workers
        cfg.serial = 1 if serial else 0
        cfg.max_batch_columns = max_batch_columns
        cfg.n_sizes = len(self.sizes)
        for i, e in enumerate(self.sizes):
            cfg.sizes[i] = e
        seed = (ctypes.c_uint8 * 32).from_buffer_copy(self.seed)
        h = ctypes.c_void_p()
        rc = lib().lbn_server_create(ctypes.byref(key_params().c()), seed, ctypes.byref(cfg), ctypes.byref(h))
        _check(rc)
        self._h = h.value
        self.open = {}

    def submit(self, rid, e, dropped_bits, witness, fail_fold=False, params=None):
        """params: the request's own Params (by default the serving key's row length at 2^e)."""
        r = Request(rid, e, dropped_bits, witness, params)
        p = r.params()
        dom = b"labinius/serving/" + int(rid).to_bytes(8, "little")
        dbuf = (ctypes.c_uint8 * len(dom)).from_buffer_copy(dom)
        t = ctypes.c_uint64()
        o = r.outputs()
        rc = lib().lbn_server_submit(ctypes.c_void_p(self._h), ctypes.byref(p.c()), _p(r.w), ctypes.c_size_t(len(r.w)),
                                     dbuf, ctypes.c_size_t(len(dom)), ctypes.byref(o),
                                     ctypes.c_uint32(LBN_SERVER_FAIL_FOLD if fail_fold else 0), ctypes.byref(t))
        if rc == LBN_ERR_ADMISSION:
            raise AdmissionError(rc, lib().lbn_last_error().decode())
        _check(rc)
        r.ticket = t.value
        self.open[r.ticket] = r
        return r

    def wait(self, r):
        info = _Info()
        _check(lib().lbn_server_wait(ctypes.c_void_p(self._h), ctypes.c_uint64(r.ticket), ctypes.byref(info)))
        self.open.pop(r.ticket, None)
        r._collect(info)
        return r

    def set_budget(self, hbm_budget_bytes):
        """A new admission budget (device bytes: the server's own workspaces plus its requests')."""
        self.budget = int(hbm_budget_bytes)
        rc = lib().lbn_server_set_budget(ctypes.c_void_p(self._h), ctypes.c_uint64(self.budget))
        if rc == LBN_ERR_ADMISSION:
            raise AdmissionError(rc, lib().lbn_last_error().decode())
        _check(rc)

    def stats(self):
        s = _Stats()
        _check(lib().lbn_server_stats_get(ctypes.c_void_p(self._h), ctypes.byref(s)))
        return {k: getattr(s, k) for k, _ in _Stats._fields_}

    def close(self):
        if getattr(self, "_h", None):
            for r in list(self.open.values()):
                self.wait(r)
            lib().lbn_server_destroy(ctypes.c_void_p(self._h))
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def request_bytes(params):
    """Device bytes a request holds from admission to the end of its fold (lbn_server_request_bytes)."""
    L = lib()
    L.lbn_server_request_bytes.restype = ctypes.c_uint64
    return int(L.lbn_server_request_bytes(ctypes.byref(params.c())))


def run_stream(server, requests):
    """Submit (rid, e, dropped_bits, witness, arrival offset s[, fail_fold]) at their arrival times;
    wait for all. Returns the Requests, with the server's own timestamps."""
    start = time.perf_counter()
    subs = []
    for req in requests:
        rid, e, d, w, at = req[:5]
        fail_fold = bool(req[5]) if len(req) > 5 else False
        now = time.perf_counter() - start
        if now < at:
            time.sleep(at - now)
        subs.append(server.submit(rid, e, d, w, fail_fold=fail_fold))
    for r in subs:
        server.wait(r)
    return subs


def serve_unbatched(matrix_seed, requests, sizes=(16, 18, 20, 22, 24), server=None):
    """The same stream one request at a time: no batching, no overlap."""
    own = server is None
    if own:
        server = Server(matrix_seed, sizes=sizes, serial=True, workers=1)
    try:
        return run_stream(server, requests)
    finally:
        if own:
            server.close()


def summarize(reqs):
    ok = [r for r in reqs if r.status == 0]
    lat = np.array([r.latency_ms() for r in ok]) if ok else np.zeros(1)
    first = min(r.arrival for r in reqs)
    last = max(r.done for r in reqs)
    return {"requests": len(reqs), "errors": sum(r.status != 0 for r in reqs),
            "p50_ms": round(float(np.percentile(lat, 50)), 3), "p99_ms": round(float(np.percentile(lat, 99)), 3),
            "mean_ms": round(float(lat.mean()), 3), "commits_per_s": round(len(ok) / (last - first), 1),
            "makespan_s": round(last - first, 3),
            "mean_requests_per_batch": round(len(reqs) / max(1, len({r.batch for r in ok})), 2)}
