"""Serving commit -> evaluate -> fold requests of mixed sizes under one key, with continuous
batching.

A scheduler thread admits queued requests in arrival order while the HBM budget holds their
resident witnesses, uploads them, and commits everything it admitted in one batched launch
(lbn_commit_batch): requests that arrive while earlier ones are on the GPU join the next batch.
Each committed request then goes to the worker of its size, which owns a prover (its own HIP
stream) and runs the rest of the round: the evaluation point from the transcript, the row
evaluation and its claim, the folding challenges and the fold of the still-resident witness
(lbn_fold_resident). Workers of different sizes run at once, so their kernels overlap on the
device with each other and with the next batched commitment. A request's transcript starts from
its own domain, so its result does not depend on what else is in flight.

The public matrix A is expanded once per prover and stays in HBM (the committer's and each
worker's device tables); a request holds device memory for its witness from admission to the end
of its fold.
"""
import collections
import queue
import threading
import time

import numpy as np

from . import (Commitment, Dropped, Opening, Prover, PublicParameters, Transcript, Verifier, claim)
from .workloads import key_params, witness_params


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
                and all(bytes(x.positions) == bytes(y.positions) and x.signs == y.signs
                        for x, y in zip(a.challenges.challenges, b.challenges.challenges))
                and np.array_equal(a.folded.elements, b.folded.elements))


class Request:
    def __init__(self, rid, e, dropped_bits, witness, arrival):
        self.id = rid
        self.e = e
        self.dropped_bits = dropped_bits
        self.witness = witness
        self.arrival = arrival
        self.admitted = self.committed = self.done = None
        self.result = None
        self.error = None
        self.rw = None
        self.commitment = None
        self.event = threading.Event()

    def hbm_bytes(self):
        """Device memory the request holds from admission to the end of its fold: the witness."""
        return 24 << self.e

    def params(self):
        return witness_params(self.e, Opening.bit_dropped(self.dropped_bits) if self.dropped_bits else Opening.CLEAR)

    def latency_ms(self):
        return 1e3 * (self.done - self.arrival)


def round_for(prover, verifier, rw, params, commitment, rid, opening=None):
    """Everything after the commitment, as one request's round (scheme/verifier.rs:61-111 for the
    transcript steps, Witness::row_evaluate, eval.rs claim, Prover::fold: of the opening when there
    is one, else of the resident witness)."""
    if params.dropped_bits() > 0 and commitment.dropped is None:
        commitment = Commitment(params.primes(), params.columns(), commitment.components, commitment.raw,
                                Dropped.of_device(prover, params, commitment.raw))
    t = Transcript(b"labinius/serving/" + int(rid).to_bytes(8, "little"))
    point = verifier.derive_evaluation_point(t, commitment)
    row = prover.row_evaluate(rw, point)
    claimed = claim(row.values, point.p1)
    challenges = verifier.derive_folding_challenges(t, row)
    folded = prover.fold(opening, challenges) if opening is not None else prover.fold_resident(rw, challenges)
    return Result(commitment, point, row, claimed, challenges, folded)


class _HostKeys:
    """Host-only public parameters per (size, mode): the transcript's view of each request's Params."""

    def __init__(self, seed):
        self.seed = seed
        self.lock = threading.Lock()
        self.pp = {}

    def verifier(self, params):
        k = (params.witness_log_len, params.dropped_bits())
        with self.lock:
            if k not in self.pp:
                self.pp[k] = Verifier(PublicParameters.host_only(params, self.seed))
            return self.pp[k]


class _Worker:
    def __init__(self, server, e):
        self.server = server
        self.e = e
        self.pp = PublicParameters.from_seed(witness_params(e), server.seed)
        self.prover = Prover(self.pp)
        self.q = queue.Queue()
        self.thread = threading.Thread(target=self._run, name=f"labinius-worker-{e}", daemon=True)
        self.thread.start()

    def _run(self):
        while True:
            r = self.q.get()
            if r is None:
                return
            try:
                p = r.params()
                r.result = round_for(self.prover, self.server.keys.verifier(p), r.rw, p, r.commitment, r.id)
            except Exception as ex:  # the request fails, the server goes on
                r.error = ex
            finally:
                r.rw = None  # frees the resident witness
                r.done = time.perf_counter()
                self.server._release(r)
                r.event.set()


class Server:
    def __init__(self, matrix_seed, hbm_budget_bytes=200e9, sizes=(16, 18, 20, 22, 24)):
        self.seed = bytes(matrix_seed)
        self.budget = hbm_budget_bytes
        self.reserved = 0
        self.peak_reserved = 0
        self.keys = _HostKeys(self.seed)
        self.committer = Prover(PublicParameters.from_seed(key_params(), self.seed))
        self.workers = {e: _Worker(self, e) for e in sizes}
        self.pending = collections.deque()
        self.cv = threading.Condition()
        self.running = True
        self.batches = []  # (requests, columns, ms) per batched commitment
        self.thread = threading.Thread(target=self._schedule, name="labinius-scheduler", daemon=True)
        self.thread.start()

    def submit(self, rid, e, dropped_bits, witness, arrival=None):
        if e not in self.workers:
            raise ValueError(f"no worker for 2^{e}")
        if 24 << e > self.budget:
            raise ValueError("the request alone exceeds the HBM budget")
        r = Request(rid, e, dropped_bits, witness, time.perf_counter() if arrival is None else arrival)
        with self.cv:
            self.pending.append(r)
            self.cv.notify_all()
        return r

    def _release(self, r):
        with self.cv:
            self.reserved -= r.hbm_bytes()
            self.cv.notify_all()

    def _schedule(self):
        while True:
            with self.cv:
                while self.running and not (self.pending and
                                            self.reserved + self.pending[0].hbm_bytes() <= self.budget):
                    self.cv.wait()
                if not self.running and not self.pending:
                    return
                batch = []
                while self.pending and self.reserved + self.pending[0].hbm_bytes() <= self.budget:
                    r = self.pending.popleft()
                    self.reserved += r.hbm_bytes()
                    batch.append(r)
                self.peak_reserved = max(self.peak_reserved, self.reserved)
            if not batch:
                continue
            t0 = time.perf_counter()
            try:
                for r in batch:
                    r.rw = self.committer.upload_params(r.witness)
                    r.admitted = time.perf_counter()
                comms = self.committer.commit_batch([r.rw for r in batch])
            except Exception as ex:
                for r in batch:
                    r.error = ex
                    r.rw = None
                    r.done = time.perf_counter()
                    self._release(r)
                    r.event.set()
                continue
            now = time.perf_counter()
            self.batches.append((len(batch), sum(r.witness.params.columns() for r in batch), 1e3 * (now - t0)))
            for r, c in zip(batch, comms):
                r.commitment = c
                r.committed = now
                self.workers[r.e].q.put(r)

    def close(self):
        with self.cv:
            self.running = False
            self.cv.notify_all()
        self.thread.join()
        for w in self.workers.values():
            w.q.put(None)
            w.thread.join()


def serve_unbatched(matrix_seed, requests, sizes=(16, 18, 20, 22, 24), provers=None):
    """The same stream, one request at a time on one stream per size: no batching, no overlap.
    requests: (rid, e, dropped_bits, witness, arrival offset s); returns Requests with timestamps."""
    seed = bytes(matrix_seed)
    keys = _HostKeys(seed)
    if provers is None:
        provers = {e: Prover(PublicParameters.from_seed(witness_params(e), seed)) for e in sizes}
    out = []
    start = time.perf_counter()
    for rid, e, d, w, at in requests:
        arrival = start + at
        now = time.perf_counter()
        if now < arrival:
            time.sleep(arrival - now)
        r = Request(rid, e, d, w, arrival)
        pr = provers[e]
        r.rw = pr.upload(w)
        r.admitted = time.perf_counter()
        r.commitment, op = pr.commit_resident(r.rw)
        r.committed = time.perf_counter()
        p = r.params()
        r.result = round_for(pr, keys.verifier(p), r.rw, p, r.commitment, rid, op)
        r.rw = None
        r.done = time.perf_counter()
        out.append(r)
    return out


def run_stream(server, requests):
    """Submit (rid, e, dropped_bits, witness, arrival offset s) at their arrival times; wait for all."""
    start = time.perf_counter()
    subs = []
    for rid, e, d, w, at in requests:
        arrival = start + at
        now = time.perf_counter()
        if now < arrival:
            time.sleep(arrival - now)
        subs.append(server.submit(rid, e, d, w, arrival))
    for r in subs:
        r.event.wait()
    return subs


def summarize(reqs):
    lat = np.array([r.latency_ms() for r in reqs])
    first = min(r.arrival for r in reqs)
    last = max(r.done for r in reqs)
    return {"requests": len(reqs), "errors": sum(r.error is not None for r in reqs),
            "p50_ms": round(float(np.percentile(lat, 50)), 3), "p99_ms": round(float(np.percentile(lat, 99)), 3),
            "mean_ms": round(float(lat.mean()), 3), "commits_per_s": round(len(reqs) / (last - first), 1),
            "makespan_s": round(last - first, 3)}
