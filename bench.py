"""Benchmark of the commit + evaluate + fold path at 2^24 F162 on one MI355X.

One step = Prover::commit + Witness::row_evaluate (and the MLE claim) + Prover::fold of one
Params::sized(SUITES[3], Clear) witness (2^24 uniform F162 in 1024 columns, moduli 3889 / 2917 /
9721), the witness and A already resident in HBM. The evaluation point and the folding
challenges are derived once up front by the host transcript (they are the verifier's steps;
the reference bench times them on their own rows). Outputs stay in HBM inside the timed region;
the copy-back of the commitment and the folded witness is timed separately (``d2h_ms``).

Prints one JSON line. ``--gpus N`` with N > 1 exits: this build measures one GPU.
"""
import argparse
import ctypes
import json
import os
import platform
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

MATRIX_SEED = bytes([0x5A]) * 32
WITNESS_SEED = bytes([0xC7]) * 32
PEAK_INT8_TOPS = 5000.0   # MI355X dense int8 MFMA, TOP/s
PEAK_HBM_GBS = 8000.0     # MI355X HBM3E, GB/s


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def tile_ops(params):
    """Algorithmic int8 ops of one commit GEMM launch: per ring element and limb, every 81 x 81
    block (8 per splitting limb; 16 quarter-blocks of 81 x 81 rows for a quadratic one) times the
    four digit products, 2 ops per multiply-add."""
    per_ring = 0
    for q in params.primes():
        quad = (q - 1) % 972 == 0 and (q - 1) % 1944 != 0
        per_ring += (2 * 4 * 162 * 81) if quad else (8 * 81 * 81)
    return 2 * 4 * per_ring * (params.witness_len() // 4)


def kpad(length):
    """a K index length padded to a power of two of at least one step of 64"""
    p = 64
    while p < length:
        p *= 2
    return p


def executed_ops(params):
    """int8 MFMA ops the commit GEMM executes, padding included: per (tile, group of 256 columns)
    and K step of 64 (sub-ring coefficient, ring element) pairs, 81 x kpad(L) / 64 steps, 8 waves x
    44 v_mfma_i32_16x16x64_i8 of 2 x 16 x 16 x 64 ops (per wave and 16-column block: five 16-row
    blocks x 4 digit products for rows 0..79, and 2 for a block holding rows 80 and 81's
    digits)."""
    ntiles = sum(16 if (q - 1) % 972 == 0 and (q - 1) % 1944 != 0 else 8 for q in params.primes())
    groups = -(-params.columns() // 256)
    kbs = 81 * kpad(params.len_ring()) // 64
    return ntiles * groups * kbs * 8 * 44 * 2 * 16 * 16 * 64


def stage_rooflines(params, stage, ms):
    """Per kernel stage: its algorithmic work, the roofline that bounds it (HBM at 8 TB/s or int8
    MFMA at 5 POP/s), the achieved rate and the fraction; then the whole path's fraction, the sum
    of the stages' floors over the measured step. Bytes are what each kernel must move at least:
    W and the two byte layouts for the layout pass, yacc in and raw + components out for the
    finish, the challenges' fragments and the fold's A', v-hat in and v out, eq, and W again for
    the row values.
    The smallest kernels (challenges, inverse NTT, eq) are bound by their latency, far from either
    roofline; they are listed for completeness."""
    n, r, L, k = params.witness_len(), params.columns(), params.len_ring(), len(params.primes())
    layout = lambda batches, kdim: -(-batches // 256) * 8 * (81 * kpad(kdim) // 64) * 2 * 1024
    work = {
        "commit_layout": ("hbm", 24 * n + layout(r, L) + layout(L, r)),
        "commit_gemm": ("mfma", tile_ops(params)),
        "commit_finish": ("hbm", r * 2 * k * 648 * 4 + k * r * 648 * 4 + r * 4 * k * 162 * 2),
        # the challenges' transform written and read, and the fold's A' (E folded in) written
        "fold_challenges": ("hbm", 8 * (2 * -(-r // 64) + 81 * kpad(r) // 64) * 12288),
        "fold_gemm": ("mfma", 8 * 81 * 81 * 4 * 2 * L * r),
        "fold_intt": ("hbm", L * 648 * 4 + L * 648 * 2),
        "eval_eq": ("hbm", (n // r) * 24),
        "eval_reduce": ("hbm", 24 * n),
    }
    out, floor = {}, 0.0
    for name, (bound, w) in work.items():
        t = stage.get(name, 0.0)
        peak = PEAK_HBM_GBS * 1e9 if bound == "hbm" else PEAK_INT8_TOPS * 1e12
        f = w / peak * 1e3
        floor += f
        ach = w / (t * 1e-3) if t > 0 else 0.0
        out[name] = {"ms": round(t, 4), "bound": bound, "work": w, "work_unit": "bytes" if bound == "hbm" else "ops",
                     "achieved": round(ach / 1e9, 1) if bound == "hbm" else round(ach / 1e12, 1),
                     "unit": "GB/s" if bound == "hbm" else "TOP/s", "floor_ms": round(f, 4),
                     "frac": round(f / t, 4) if t > 0 else None}
    return out, round(floor, 4), round(floor / ms, 4)


def mfma_rate(L, operands):
    """The device's int8 MFMA rate (TOP/s) after a warm-up, measured by lbn_debug_mfma_rate:
    operands 0 constant, 1 random bytes switching every product, 2 like the tile GEMM's digits."""
    t = ctypes.c_double()
    rc = L.lbn_debug_mfma_rate(ctypes.c_int(operands), ctypes.byref(t))
    if rc != 0:
        raise RuntimeError(L.lbn_last_error().decode())
    return round(t.value, 1)


def algorithmic_bytes(params):
    """Algorithmic HBM bytes of one commit + fold + eval unit (the reference's kept transform
    counted at 2 B per slot): W read by commit and eval, T written by commit and read by fold, A."""
    n = params.witness_len()
    w = 24 * n
    t = 2 * 648 * (n // 4)
    a = 2 * 648 * params.len_ring() * len(params.primes())
    v = 2 * 648 * params.len_ring()
    chat = 2 * 648 * params.columns()
    return dict(commit=w + t + a, fold=t + chat + v, eval=w, total=(w + t + a) + (t + chat + v) + w)


def profiled_traffic(suite):
    """HBM bytes per launch of the commit GEMM, from the committed rocprofv3 FETCH_SIZE / WRITE_SIZE
    summary of this workload (scaled by a plain-copy calibration); None for other workloads."""
    if suite != "xl":
        return None
    path = os.path.join(ROOT, "profiles", "commit_fold_eval_xl_hbm_traffic.md")
    try:
        for line in open(path):
            if line.startswith("traffic_bytes_per_launch tile_gemm_kernel (commit) ="):
                return int(line.split("=")[1])
    except OSError:
        pass
    return None


def cpu_model():
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return platform.machine()


def cpu_baseline(params, w, commitment, row, claimed, point, challenges, folded, reps=3):
    """The oracle (a single-threaded, scalar plain-C restatement of the reference algorithm, not the
    reference's AVX-512 code) on a bounded sample of the same workload, scaled to the whole step:
    the commit and the row evaluation of the witness's first columns (each column is independent)
    and the fold of a few output ring elements (each depends on every column at its own index),
    each the median of `reps` runs on one pinned core. The sample's outputs are also compared with
    the GPU's, so that the step's result is checked bit for bit where it is measured: raw rows and
    components of the sampled columns, their row values, the MLE claim from all row values, and
    the sampled folded ring elements. The reference's own single-core figures (other hardware,
    from its source comments) are quoted beside it."""
    import numpy as np
    import oracle as orc
    affinity = sorted(os.sched_getaffinity(0))
    core = affinity[0]
    try:
        os.sched_setaffinity(0, {core})
    except Exception:
        core = None

    def med(f):
        ts, out = [], None
        for _ in range(reps):
            t0 = time.perf_counter()
            out = f()
            ts.append(time.perf_counter() - t0)
        return out, float(np.median(ts))

    primes, cols, L, wdim = params.primes(), params.columns(), params.len_ring(), params.column_len()
    ncols, nrows = 32, 128
    a = orc.key_expand(primes, L, orc.matrix_seed(MATRIX_SEED))
    ws = w.elements[:ncols * wdim]
    (raw, comp), t_c = med(lambda: orc.commit(primes, L, ncols, a, ws))
    u, t_e = med(lambda: orc.row_evaluate(ws, point.p0))
    rows = np.arange(0, L, L // nrows)[:nrows]
    fv, t_f = med(lambda: orc.fold_rows(primes[0], L, cols, w.elements, challenges.challenges, rows))
    checked = {
        "columns": f"0..{ncols - 1} of {cols}: raw rows, components and row values",
        "fold_rows": f"{nrows} of {L} folded ring elements (every {L // nrows}th)",
        "claim": "from all row values",
        "raw_and_components": bool(np.array_equal(commitment.raw[:, :ncols], raw)
                                   and np.array_equal(commitment.components[:ncols], comp)),
        "row_values": bool(np.array_equal(row.values[:ncols], u)),
        "claim_equal": bool(np.array_equal(claimed, orc.claim(row.values, point.p1))),
        "folded": bool(np.array_equal(folded.elements[rows], fv)),
    }
    if not all(checked[k] for k in ("raw_and_components", "row_values", "claim_equal", "folded")):
        raise AssertionError(f"the GPU's step differs from the oracle: {checked}")
    t_commit, t_eval, t_fold = t_c / ncols * cols, t_e / ncols * cols, t_f / nrows * L
    total_ms = 1e3 * (t_commit + t_eval + t_fold)
    t0 = time.perf_counter()
    orc.rans_encode(folded.elements, primes[0])
    encode_ms = 1e3 * (time.perf_counter() - t0)
    return {"value": round(total_ms, 1), "unit": "ms", "cores": 1, "kind": "port",
            "wire_encode_ms": round(encode_ms, 2),
            "stages_ms": {"commit": round(1e3 * t_commit, 1), "eval": round(1e3 * t_eval, 1),
                          "fold": round(1e3 * t_fold, 1)},
            "sample": f"the oracle (scalar plain C, not the reference's AVX-512 path) on one pinned core "
                      f"(core {core} of {len(affinity)} in this process's affinity, os.cpu_count() "
                      f"{os.cpu_count()}, {cpu_model()}): commit and row evaluation of {ncols} of {cols} "
                      f"columns and the fold of {nrows} of {L} output ring elements, each the median of "
                      f"{reps} runs, scaled linearly to the whole step",
            "parity_checked": checked,
            "reference_published": {
                "hardware": "Intel i7-11850H, one core (the reference's own source comments; other hardware)",
                "suite_l_commit_ms": 295.9, "suite_l_fold_ms": 68.1,
                "source": "crates/pcs/src/simd/commit.rs:100-101, crates/pcs/src/fold.rs:25-29",
                "note": "suite l is 2^22 F162 in 512 columns; xl has 4x the witness"}}


def wire_bench(lbn, prover, params, folded, reps=5):
    """The wire forms of the step's outputs (wire/mod.rs, wire/rans.rs), on the GPU over what the
    step left in HBM: the commitment bit-packed at ceil(log2 q) bits a slot, and the folded witness
    rANS-coded (the min / max and counts on the device, the table quantised on the host, one
    wavefront for the 64 interleaved coders). Times are wall clock with the bytes copied back."""
    import numpy as np
    from labinius_amd import wire

    def med(f):
        ts, out = [], None
        for _ in range(reps):
            t0 = time.perf_counter()
            out = f()
            ts.append(1e3 * (time.perf_counter() - t0))
        return out, round(float(np.median(ts)), 3)

    packed, pack_ms = med(lambda: wire.pack_commitment(prover))
    coded, encode_ms = med(lambda: wire.encode(prover, params.primes()[0]))
    total, stats, kernel = wire.timings(prover)
    assert np.array_equal(wire.decode(coded).elements, folded.elements)
    coefs = folded.elements.size
    entropy = wire.entropy_bytes(folded.elements)
    return {"pack_commitment_ms": pack_ms, "commitment_bytes": len(packed),
            "commitment_bytes_at_i16": 2 * 4 * params.columns() * len(params.primes()) * 162,
            "encode_ms": encode_ms, "encode_device_ms": {"total": round(total, 4), "min_max_counts": round(stats, 4),
                                                         "rans_kernel": round(kernel, 4)},
            "folded_coefficients": coefs, "encoded_bytes": len(coded),
            "bits_per_coefficient": round(8 * len(coded) / coefs, 4),
            "entropy_bits_per_coefficient": round(8 * entropy / coefs, 4),
            "sigma": round(float(np.sqrt(np.mean(folded.elements.astype(np.float64) ** 2))), 2)}


def verify_bench(lbn, prover, verifier, commitment, challenges, point, folded, row, reps=3):
    """The verify side of the step (verifier.rs:178-231, fold.rs:799-873) for the step's own
    outputs: the folded commitment and the Clear opening check, on the host and on the GPU (the
    folded witness and the commitment copied in, the bit planes of v through the commit GEMM).
    Wall clock; both must accept."""
    import numpy as np

    def timed(f, n):
        ts, out = [], None
        for _ in range(n):
            t0 = time.perf_counter()
            out = f()
            ts.append(1e3 * (time.perf_counter() - t0))
        return out, round(float(np.median(ts)), 3)

    fc_host, fold_host_ms = timed(lambda: verifier.fold_commitment(commitment, challenges), 1)
    fc_dev, fold_dev_ms = timed(lambda: verifier.fold_commitment(commitment, challenges, device=prover), reps)
    assert np.array_equal(fc_host.rows, fc_dev.rows)
    fr = verifier.fold_row_evaluation(row, challenges)
    args = (commitment, challenges, point, fc_host, folded, fr)
    _, verify_host_ms = timed(lambda: verifier.verify_opening(*args), 1)
    _, verify_dev_ms = timed(lambda: verifier.verify_opening(*args, device=prover), reps)
    return {"fold_commitment_ms": {"host": fold_host_ms, "device": fold_dev_ms},
            "verify_opening_ms": {"host": verify_host_ms, "device": verify_dev_ms},
            "note": "host = the library's C++ verifier on one thread; device = the same checks on the GPU, "
                    "inputs copied in over PCIe; both accepted"}


def batched(lbn, reps):
    """The batched commitment of 64 uneven witnesses (2^16 .. 2^22 F162, a quarter of them lifted B128
    traces, one key) against the same 64 commitments one at a time on the GPU. Both leave their
    outputs in HBM. Latency is from the start of the workload to each commitment's completion."""
    import ctypes
    import numpy as np
    import torch
    from labinius_amd import workloads as wl
    key = wl.key_params()
    pp = lbn.PublicParameters.from_seed(key, MATRIX_SEED)
    prover = lbn.Prover(pp)
    specs = wl.batch_specs()
    t0 = time.time()
    ws = [wl.batch_witness(i, e, lifted) for i, (e, lifted) in enumerate(specs)]
    rws = [prover.upload_params(w) for w in ws]
    log(f"batch: {len(ws)} witnesses, {sum(w.params.witness_len() for w in ws) / 1e6:.1f} M F162, "
        f"{sum(w.params.columns() for w in ws)} columns, set up in {time.time() - t0:.1f} s")
    prover.commit_batch(rws, want_outputs=False)  # autotune and capture
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(reps):
        prover.commit_batch(rws, want_outputs=False)
    torch.cuda.synchronize()
    batch_ms = 1e3 * (time.perf_counter() - t0) / reps
    info = prover.batch_info()
    # one at a time: one prover per size (the same key), each commitment on its own
    L = lbn.lib()
    singles, res = {}, []
    for w in ws:
        e = w.params.witness_log_len
        if e not in singles:
            singles[e] = lbn.Prover(lbn.PublicParameters.from_seed(w.params, MATRIX_SEED))
        res.append((singles[e], singles[e].upload(w)))
    for pr, rw in res[:4]:  # warm
        lbn._check(L.lbn_commit_resident(ctypes.c_void_p(pr._h), ctypes.c_void_p(rw._h), None, None, None))
    seq = []
    for _ in range(reps):
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        done = []
        for pr, rw in res:
            lbn._check(L.lbn_commit_resident(ctypes.c_void_p(pr._h), ctypes.c_void_p(rw._h), None, None, None))
            done.append(1e3 * (time.perf_counter() - t0))
        seq.append(done)
    one_ms = float(np.mean([d[-1] for d in seq]))
    lat = np.array(seq).mean(axis=0)
    return {"witnesses": len(ws), "f162": int(sum(w.params.witness_len() for w in ws)),
            "columns": int(sum(w.params.columns() for w in ws)), "lifted_b128": int(sum(l for _, l in specs)),
            "batched_ms": round(batch_ms, 3), "batched_device_ms": round(info["ms"], 3),
            "batched_commits_per_s": round(len(ws) / batch_ms * 1e3, 1),
            "batched_p50_ms": round(batch_ms, 3), "batched_p99_ms": round(batch_ms, 3),
            "one_at_a_time_ms": round(one_ms, 3),
            "one_at_a_time_commits_per_s": round(len(ws) / one_ms * 1e3, 1),
            "one_at_a_time_p50_ms": round(float(np.percentile(lat, 50)), 3),
            "one_at_a_time_p99_ms": round(float(np.percentile(lat, 99)), 3),
            "speedup": round(one_ms / batch_ms, 3),
            "autotuned": {"tiles_per_xcd_round": info["tiles_per_xcd_round"],
                          "k_ranges": info["k_ranges"]},
            "graph": "one hipGraph launch per batch (layout, GEMM, finish)"}


def switch_bench(lbn, reps=5):
    """The cross-field switch (switch.rs) on lifted B128 traces of 2^20 and 2^22 words, the batched
    workload's largest, under its key (row length 2^11, moduli 3889 / 9721): the prover on the GPU
    over the resident lifted trace, the verifier on the host. The word claim is the one the partial
    evaluations give (sum_k v_k x^k), so the timed proof verifies."""
    import numpy as np
    from labinius_amd import switch
    from labinius_amd import workloads as wl

    def times_x(v, k):  # v x^k in B128 = GF(2)[x] / (x^128 + x^7 + x^2 + x + 1)
        v <<= k
        for i in range(v.bit_length() - 1, 127, -1):
            if (v >> i) & 1:
                v ^= (1 << i) | (0x87 << (i - 128))
        return v

    out = {}
    rng = np.random.default_rng(0x5317)
    for e in (20, 22):
        params = wl.witness_params(e)
        prover = lbn.Prover(lbn.PublicParameters.from_seed(params, MATRIX_SEED))
        trace = rng.integers(0, 2**63, (1 << e, 2), dtype=np.uint64) * np.uint64(2)
        rw = prover.upload(lbn.Witness.lifted(params, trace))
        r = rng.integers(0, 2**63, (e, 2), dtype=np.uint64)
        first, _ = switch.prove(prover, rw, r, np.zeros(2, np.uint64), lbn.Transcript(b"labinius/bench/switch"))
        c = 0
        for k, v in enumerate(first.partials):
            c ^= times_x(int(v[0]) | (int(v[1]) << 64), k)
        claim = np.array([c & (2**64 - 1), c >> 64], np.uint64)
        ts = []
        for _ in range(reps):
            t0 = time.perf_counter()
            proof, _ = switch.prove(prover, rw, r, claim, lbn.Transcript(b"labinius/bench/switch"))
            ts.append(1e3 * (time.perf_counter() - t0))
        t0 = time.perf_counter()
        switch.verify(params, r, claim, proof, lbn.Transcript(b"labinius/bench/switch"))
        verify_ms = 1e3 * (time.perf_counter() - t0)
        out[f"2^{e}"] = {"prove_ms": round(float(np.median(ts)), 3), "verify_ms_host": round(verify_ms, 3),
                         "rounds": e}
    out["note"] = ("prove = GPU passes plus one host transcript step per sumcheck round (one device "
                   "synchronisation each); the trace is resident as Witness::lifted")
    return out


def serving_bench(lbn, rates, count):
    """The serving loop (the native lbn_server): a stream of `count` commit -> evaluate -> fold requests
    with sizes in the mix's exact proportions (2^16 40 %, 2^18 30 %, 2^20 20 %, 2^22 8 %, 2^24 2 %),
    Clear or bit-dropped at random, at each Poisson arrival rate (None: all at once, the saturating
    case), through the continuous-batching server and through the same server run serially (one
    request at a time, unbatched). Latency is arrival to folded witness on the host; commits/s is
    completed requests over the span from the first arrival to the last completion."""
    import numpy as np
    from labinius_amd import serving
    from labinius_amd import workloads as wl
    rng = np.random.default_rng(1)
    spec = wl.serving_requests_stratified(None, count)
    wits = []
    for _, e, d, i in spec:
        x = rng.integers(0, 2**63, (1 << e, 3), dtype=np.uint64) * np.uint64(2) + rng.integers(0, 2, (1 << e, 3),
                                                                                                  dtype=np.uint64)
        x[:, 2] &= np.uint64((1 << 34) - 1)
        wits.append(lbn.Witness.from_elements(wl.witness_params(e), x))
    t0 = time.time()
    served = serving.Server(MATRIX_SEED)
    serial = serving.Server(MATRIX_SEED, serial=True, workers=1)
    log(f"servers set up in {time.time() - t0:.1f} s")
    warm = [(10**6 + k, e, wl.SERVING_DROPPED[e] * (k % 2), wits[[s_[1] for s_ in spec].index(e)], 0.0)
            for k, e in enumerate([16, 18, 20, 22, 24] * 2)]
    serving.run_stream(served, warm)
    serving.run_stream(serial, warm)
    out = {"requests_per_rate": count, "sizes": {f"2^{e}": sum(1 for s_ in spec if s_[1] == e) for e in (16, 18, 20, 22, 24)},
           "bit_dropped": sum(1 for s_ in spec if s_[2]), "hbm_budget_gb": served.budget / 1e9,
           "fixed_gb": {"served": round(served.stats()["fixed_bytes"] / 1e9, 3),
                        "serial": round(serial.stats()["fixed_bytes"] / 1e9, 3)},
           "rates": []}
    for rate in rates:
        arr = wl.serving_requests_stratified(rate, count)
        reqs = [(i, e, d, wits[i], t) for (t, e, d, i) in arr]
        before = served.stats()
        mixed = serving.run_stream(served, reqs)
        after = served.stats()
        alone = serving.run_stream(serial, reqs)
        ms = serving.summarize(mixed)
        ms["batches"] = after["batches"] - before["batches"]
        ms["pool_mallocs"] = after["pool_mallocs"] - before["pool_mallocs"]
        su = serving.summarize(alone)
        stage = lambda rs, a, b: round(1e3 * float(np.mean([getattr(r, b) - getattr(r, a) for r in rs])), 3)
        for m, rs in ((ms, mixed), (su, alone)):
            m["queue_ms"] = stage(rs, "arrival", "admitted")
            m["commit_ms"] = stage(rs, "admitted", "committed")
            m["round_ms"] = stage(rs, "committed", "done")
        big = [r.latency_ms() for r in mixed if r.e == 24]
        ms["latency_2^24_ms"] = [round(x, 3) for x in big]
        out["rates"].append({"rate_per_s": rate if rate else "all at once", "served": ms, "unbatched": su,
                             "speedup_commits_per_s": round(ms["commits_per_s"] / max(su["commits_per_s"], 1e-9), 3)})
        log(f"serving at {rate}/s: served {ms['commits_per_s']}/s p99 {ms['p99_ms']} ms; "
            f"unbatched {su['commits_per_s']}/s p99 {su['p99_ms']} ms")
    st = served.stats()
    out["pool"] = {"obtained_gb": round(st["pool_obtained"] / 1e9, 3), "mallocs": st["pool_mallocs"],
                   "reuses": st["pool_reuses"], "peak_reserved_gb": round(st["reserved_peak"] / 1e9, 3)}
    served.close()
    serial.close()
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--suite", default="xl", choices=["s", "m", "l", "xl"])
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-batched", action="store_true", help="skip the batched-commit comparison")
    ap.add_argument("--no-serving", action="store_true", help="skip the serving comparison")
    ap.add_argument("--no-switch", action="store_true", help="skip the cross-field switch timing")
    ap.add_argument("--no-wire", action="store_true", help="skip the wire encoding timing")
    ap.add_argument("--no-verify", action="store_true", help="skip the verify-side timing")
    ap.add_argument("--serving-rates", default="100,300,1000,0",
                    help="Poisson arrival rates, requests per second (0: all at once)")
    ap.add_argument("--serving-requests", type=int, default=100)
    ap.add_argument("--calibrate-copy", action="store_true",
                    help="run only a plain 1 GiB device copy (read 1 GiB, write 1 GiB) five times, to scale the "
                         "FETCH_SIZE / WRITE_SIZE counters of a profiled run")
    args = ap.parse_args()
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) > 1:
        print("bench.py measures one GPU: the prover runs on a single MI355X (columns would shard across "
              "GPUs with one all-reduce in the fold; not part of this build)", file=sys.stderr)
        sys.exit(2)

    if args.calibrate_copy:
        import torch
        src = torch.ones(1 << 28, dtype=torch.int32, device="cuda")
        dst = torch.empty_like(src)
        for _ in range(5):
            dst.copy_(src)
        torch.cuda.synchronize()
        print(json.dumps({"calibration": "copy", "bytes_read": 1 << 30, "bytes_written": 1 << 30, "launches": 5}))
        return

    import numpy as np
    import labinius_amd as lbn

    if args.suite == "s":
        params = lbn.Params.basic()
    else:
        params = lbn.Params.sized(lbn.SUITES[{"m": 1, "l": 2, "xl": 3}[args.suite]])
    log(f"params: 2^{params.witness_log_len} F162, {params.columns()} columns, moduli {params.primes()}")
    t0 = time.time()
    pp = lbn.PublicParameters.from_seed(params, MATRIX_SEED)
    w = lbn.Witness.random(params, WITNESS_SEED)
    prover = lbn.Prover(pp)
    verifier = lbn.Verifier(pp)
    rw = prover.upload(w)
    log(f"setup {time.time() - t0:.1f} s")

    # the point and the challenges, once (host transcript)
    commitment, opening = prover.commit_resident(rw)
    t = lbn.Transcript(b"labinius/reference")
    point = verifier.derive_evaluation_point(t, commitment)
    row = prover.row_evaluate(rw, point)
    challenges = verifier.derive_folding_challenges(t, row)
    prover.recycle(opening)
    log("point and challenges derived")

    L = lbn.lib()
    p0 = np.ascontiguousarray(point.p0, np.uint64)
    p1 = np.ascontiguousarray(point.p1, np.uint64)
    u = np.zeros((params.columns(), 3), np.uint64)
    tval = np.zeros(3, np.uint64)
    arr = challenges.array()
    hp, hw = ctypes.c_void_p(prover._h), ctypes.c_void_p(rw._h)
    vp = ctypes.c_void_p

    def step():
        op = ctypes.c_void_p()
        lbn._check(L.lbn_commit_resident(hp, hw, None, None, ctypes.byref(op)))
        lbn._check(L.lbn_mle_evaluate(hp, hw, p0.ctypes.data_as(vp), ctypes.c_size_t(len(p0)),
                                      p1.ctypes.data_as(vp), ctypes.c_size_t(len(p1)), tval.ctypes.data_as(vp)))
        lbn._check(L.lbn_fold(hp, op, arr, ctypes.c_size_t(len(arr)), None))
        return prover.timings()

    for _ in range(args.warmup):
        step()
    import torch
    torch.cuda.synchronize()
    stage = {k: 0.0 for k in prover.timings()}
    t0 = time.perf_counter()
    for i in range(args.steps):
        tm = step()
        for k in stage:
            stage[k] += tm[k]
        log(f"step {i + 1}/{args.steps}")
    torch.cuda.synchronize()
    ms = 1e3 * (time.perf_counter() - t0) / args.steps
    stage = {k: v / args.steps for k, v in stage.items()}

    # the same step with the outputs copied back (PCIe included)
    t0 = time.perf_counter()
    c2, o2 = prover.commit_resident(rw)
    claimed = prover.mle_evaluate(rw, point)
    folded = prover.fold(o2, challenges)
    d2h_ms = 1e3 * (time.perf_counter() - t0)

    # what a drop-in Prover::commit(&Witness) pays on top: W over PCIe from pageable host memory
    # (scheme/prover.rs:55-91 takes a host witness); the upload alone, and lbn_commit of the host
    # witness (upload + commit, outputs left in HBM), each the median of 3
    e_host = np.ascontiguousarray(w.elements, np.uint64)
    h2d, host_commit = [], []
    for _ in range(3):
        t0 = time.perf_counter()
        tmp = prover.upload(w)
        h2d.append(1e3 * (time.perf_counter() - t0))
        del tmp
        op = ctypes.c_void_p()
        t0 = time.perf_counter()
        lbn._check(L.lbn_commit(hp, e_host.ctypes.data_as(vp), ctypes.c_size_t(len(e_host)), None, None,
                                ctypes.byref(op)))
        host_commit.append(1e3 * (time.perf_counter() - t0))
        lbn._check(L.lbn_recycle(hp, op))
    host_witness = {"h2d_ms": round(float(np.median(h2d)), 3), "bytes": int(e_host.nbytes),
                    "h2d_gb_per_s": round(e_host.nbytes / (float(np.median(h2d)) * 1e-3) / 1e9, 2),
                    "host_witness_commit_ms": round(float(np.median(host_commit)), 3),
                    "resident_commit_ms": round(stage["commit"], 3),
                    "note": "lbn_commit of a host witness (Prover::commit(&Witness)): the upload from pageable "
                            "host memory plus the commit; never the bench's value, which has W resident"}

    gemm_ms = stage["commit_gemm"]
    ops = tile_ops(params)
    achieved = ops / (gemm_ms * 1e-3) / 1e12
    xops = executed_ops(params)
    rate_c, rate_r, rate_g = mfma_rate(L, 0), mfma_rate(L, 1), mfma_rate(L, 2)
    ab = algorithmic_bytes(params)
    sr, sr_floor, sr_frac = stage_rooflines(params, stage, ms)
    line = {
        "metric": "commit+fold+eval ms @2^24 bit-exact",
        "value": round(ms, 3),
        "unit": "ms",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms, 3),
        "higher_is_better": False,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "int8-mfma/u32 mod q; gf(2^162)",
        "data": "synthetic: Witness::random([0xC7;32]), A from [0x5A;32]",
        "config": {"workload": f"Params::sized(SUITES[{ {'s': 0, 'm': 1, 'l': 2, 'xl': 3}[args.suite] }], Clear)"
                               if args.suite != "s" else "Params::basic()",
                   "witness_log_len": params.witness_log_len, "columns": params.columns(),
                   "moduli": params.primes()},
        "stages_ms": {k: round(v, 4) for k, v in stage.items()},
        "d2h_inclusive_ms": round(d2h_ms, 3),
        "host_witness": host_witness,
        "hbm_floor_ms": round(ab["total"] / (PEAK_HBM_GBS * 1e9) * 1e3, 4),
        "hbm_roofline_frac": round(ab["total"] / (PEAK_HBM_GBS * 1e9) * 1e3 / ms, 4),
        "stage_rooflines": sr,
        "whole_path_floor_ms": sr_floor,
        "whole_path_roofline_frac": sr_frac,
        "roofline": {"kernel": "tile_gemm_kernel (commit)", "bound": "mfma", "achieved": round(achieved, 1),
                     "peak": PEAK_INT8_TOPS, "unit": "TFLOP/s",
                     "unit_note": "int8 MFMA operations (multiply and add counted separately), 10^12 per second",
                     "frac": round(achieved / PEAK_INT8_TOPS, 4), "traffic": profiled_traffic(args.suite),
                     "ops_per_launch": ops, "launch_ms": round(gemm_ms, 4),
                     "executed_ops_per_launch": xops,
                     "measured_mfma_rate": {"constant_operands": rate_c, "random_operands": rate_r,
                                            "gemm_digit_operands": rate_g, "unit": "TOP/s",
                                            "note": "after a warm-up of about 10 ms; random = random bytes "
                                                    "switching every product (the worst case), gemm digit = "
                                                    "operands like the tile GEMM's",
                                            "executed_frac_of_gemm_digit_rate":
                                                round(xops / (gemm_ms * 1e-3) / 1e12 / rate_g, 4)}},
    }
    if not args.no_wire:
        log("wire...")
        line["wire"] = wire_bench(lbn, prover, params, folded)
    if not args.no_verify:
        log("verify...")
        line["verify"] = verify_bench(lbn, prover, verifier, c2, challenges, point, folded, row)
    if not args.no_batched:
        log("batched commit...")
        line["batched"] = batched(lbn, 3)
    if not args.no_serving:
        log("serving...")
        line["serving"] = serving_bench(lbn, [float(x) or None for x in args.serving_rates.split(",")],
                                        args.serving_requests)
    if not args.no_switch:
        log("switch...")
        line["switch"] = switch_bench(lbn)
    if not args.no_cpu_baseline:
        log("cpu baseline...")
        line["cpu_baseline"] = cpu_baseline(params, w, c2, row, claimed, point, challenges, folded)
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
