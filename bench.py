"""SpMV benchmark on one MI355X: y = alpha*A*x + beta*y, FP64 CSR.

Workload (default): the circuit-like skewed matrix, 5,558,326 x 5,558,326,
~59.5M nonzeros (200 dense rows of 1e4..1.3e6 nonzeros at uniform random
columns, the rest 1..12 nonzeros near the diagonal), FP64 N(0,1) values, x
U[-1,1), alpha = 1, beta = 0.  A, x and y are resident in HBM before the timed
region.  One step = one SpMV (spx_plan_execute) over the whole matrix.

    python bench.py [--steps K] [--warmup W] [--kernel 1|2|3|4] [--items I] [--parts P]

Prints ONE JSON line: value = algorithmic GB/s of the whole step (wall clock,
synchronize on both sides), roofline = the dominant SpMV kernel's achieved
bandwidth from HIP events on its stream vs 8 TB/s, cpu_baseline = the
reference's host CSR loop (the oracle) on this machine's CPU.
"""
from __future__ import annotations

import argparse
import json
import os
import re
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "sparse-blas_amd"))

import numpy as np  # noqa: E402

HBM_PEAK_GBS = 8000.0


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def profiled_traffic(kernel, items, window):
    """HBM bytes per launch of the dominant kernel for this configuration, as
    recorded from FETCH_SIZE + WRITE_SIZE runs in profiles/*.md, else None."""
    pat = re.compile(r"traffic_bytes_per_launch: (\d+) \(kernel=(\d+) items=(\d+) window=(\d+)\)")
    for name in sorted(os.listdir(os.path.join(ROOT, "profiles"))):
        if not name.endswith(".md"):
            continue
        with open(os.path.join(ROOT, "profiles", name)) as f:
            for mt in pat.finditer(f.read()):
                if (int(mt[2]), int(mt[3]), int(mt[4])) == (kernel, items, window):
                    return int(mt[1])
    return None


def run_precision(a):
    """The precision form on the rail-like (4,284 x 1,092,610, values 1.0),
    ASIC-like (682,862 square, lognormal values with random sign) or circuit-like
    (5,558,326 square, N(0,1) values; `prec-circuit`) matrix: plain FP64 and
    FP32-storage plans of the same schedule timed back to back, each with its
    normwise and worst-row relative error against the oracle."""
    import spx
    from spx import workloads as W

    A = {"rail": W.rail, "asic": W.asic, "prec-circuit": W.circuit}[a.workload]()
    m, n, nnz = A["m"], A["n"], A["nnz"]
    log(f"workload {a.workload}: m={m} n={n} nnz={nnz} mean_row={A['mean_row']:.2f} max_row={A['max_row']}")
    x = W.vector(n, 31)
    xd, yd = spx.DeviceBuffer.from_array(x), spx.DeviceBuffer(m * 8)
    # the CPU baseline leg: the reference's host loop (the oracle) on 1 thread,
    # timed; its y is also the checker's reference
    import oracle
    reps = []
    for _ in range(3):
        t = time.perf_counter()
        ref = oracle.spmv(A["rowptr"], A["col"], A["val"], x, 1.0, 0.0)
        reps.append((time.perf_counter() - t) * 1e3)
    cpu_ms = float(np.median(reps))
    _, mag = oracle.spmv_dot2(A["rowptr"], A["col"], A["val"], x, 1.0, 0.0)
    cpu = {"value": round(W.bytes_model(m, n, nnz, 8) / (cpu_ms * 1e-3) / 1e9, 3), "unit": "GB/s", "cores": 1,
           "kind": "port", "sample": f"whole matrix, FP64, median of 3 ({cpu_ms:.1f} ms/SpMV); main.cu:350-356 "
                                     "loop restated in C"}
    res = {}
    for prec, tag, sv in ((spx.PREC_F64, "fp64", 8), (spx.PREC_F32, "fp32", 4)):
        P = spx.Plan(A["rowptr"], A["col"], A["val"], n=n, kernel=a.kernel, items=a.items, parts=a.parts,
                     sigma=a.sigma, precision=prec, long_row=a.long_row)
        log(f"{tag} plan: {P.info()}")
        tot, ker, wall = P.time(xd, yd, 1.0, 0.0, warmup=a.warmup, reps=a.steps, wall=True)
        P.execute(xd, yd)
        spx.synchronize()
        y = yd.download(np.float64, m)
        B = W.bytes_model(m, n, nnz, sv)
        row = np.abs(y - ref) / np.maximum(mag, 1e-300)
        res[tag] = {"ms_per_step": round(wall, 5), "kernel_ms": round(ker, 5), "GBs": round(B / (wall * 1e-3) / 1e9, 2),
                    "kernel_GBs": round(B / (ker * 1e-3) / 1e9, 2),
                    "normwise_rel_err": float(np.linalg.norm(y - ref) / np.linalg.norm(ref)),
                    "max_row_err_over_abs_sum": float(row.max())}
        log(f"{tag}: wall {wall:.4f} ms, kernel {ker:.4f} ms, err {res[tag]['normwise_rel_err']:.3e}")
        P.close()
    f32 = res["fp32"]
    print(json.dumps({
        "metric": f"precision form on the {a.workload}-like matrix: FP32 storage vs plain FP64",
        "value": f32["GBs"], "unit": "GB/s", "n_gpus": 1, "steps": a.steps, "warmup": a.warmup,
        "ms_per_step": f32["ms_per_step"], "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "f32 storage, f64 accumulation", "data": "synthetic",
        "config": {"workload": f"{a.workload}-like CSR {m}x{n}, nnz={nnz}", "kernel": a.kernel, "items": a.items,
                   "long_row": a.long_row},
        "fp64": res["fp64"], "fp32": f32, "speedup_fp32_vs_fp64": round(res["fp64"]["ms_per_step"] / f32["ms_per_step"], 3),
        "kernel_ratio_fp32_over_fp64": round(f32["kernel_ms"] / res["fp64"]["kernel_ms"], 3), "cpu_baseline": cpu,
    }), flush=True)


def run_c1(a):
    """The plain method on the n = 8,192 Zipf(1.5) matrix (rows on [1, 512],
    realized mean about 17.4 nonzeros) and on the reference harness's `g`
    matrix at n = 10,000: every kernel version with its default schedule,
    timed warm (the matrices stay in L2 / MALL, so these figures are labelled
    cache-resident, not HBM), each checked against the oracle, beside the
    oracle's own time on one host thread."""
    import spx
    from spx import workloads as W
    import oracle

    res = {}
    for name, A in (("zipf8192", W.zipf(8192, 1.5, 512, seed=1)), ("g10000", W.g_matrix(10_000))):
        m, n, nnz = A["m"], A["n"], A["nnz"]
        x = W.vector(n, 3)
        xd, yd = spx.DeviceBuffer.from_array(x), spx.DeviceBuffer(m * 8)
        reps = []
        for i in range(6):
            t = time.perf_counter()
            ref = oracle.spmv(A["rowptr"], A["col"], A["val"], x, 1.0, 0.0)
            if i:
                reps.append(time.perf_counter() - t)
        cpu_ms = float(np.median(reps)) * 1e3
        B = W.bytes_model(m, n, nnz, 8)
        row = {"m": m, "n": n, "nnz": nnz, "mean_row": round(A["mean_row"], 2), "max_row": A["max_row"],
               "bytes": B, "cpu_1thread_ms": round(cpu_ms, 4), "cpu_1thread_GBs": round(B / (cpu_ms * 1e-3) / 1e9, 2)}
        for kern in (1, 2, 3, 4):
            P = spx.Plan(A["rowptr"], A["col"], A["val"], n=n, kernel=kern)
            ev, kk, wv = P.time_each(xd, yd, 1.0, 0.0, warmup=a.warmup, reps=a.steps)
            P.execute(xd, yd)
            spx.synchronize()
            y = yd.download(np.float64, m)
            err = float(np.linalg.norm(y - ref) / np.linalg.norm(ref))
            if not err <= 1e-13:
                raise AssertionError(f"{name} kernel {kern}: normwise relative error {err:.3e} > 1e-13")
            ex = float(np.median(ev))
            row[f"kernel{kern}"] = {"execute_us_median": round(ex * 1e3, 2), "GBs": round(B / (ex * 1e-3) / 1e9, 1),
                                    "err_vs_oracle": err}
            P.close()
        log(f"{name}: " + json.dumps(row))
        res[name] = row
    best = min(res["zipf8192"][f"kernel{k}"]["execute_us_median"] for k in (1, 2, 3, 4))
    print(json.dumps({
        "metric": "SpMV GB/s on the n = 8,192 Zipf(1.5) matrix (cache-resident, not an HBM figure)",
        "value": round(res["zipf8192"]["bytes"] / (best * 1e-6) / 1e9, 1), "unit": "GB/s", "n_gpus": 1,
        "steps": a.steps, "warmup": a.warmup, "ms_per_step": round(best * 1e-3, 6), "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "f64", "data": "synthetic",
        "config": {"workload": "n=8192 Zipf(1.5) rows on [1, 512] (seed 1); reference g matrix n=10000"},
        "matrices": res,
        "cpu_baseline": {"value": res["zipf8192"]["cpu_1thread_GBs"], "unit": "GB/s", "cores": 1, "kind": "port",
                         "sample": "whole matrix, median of 5 reps after 1 warm-up; main.cu:350-356 loop restated in C"},
    }), flush=True)


def run_spmm(a):
    """Multi-RHS SpMM on the circuit-like matrix: the fused passes against
    looping the single-vector execute, k columns, column-major X and Y."""
    import ctypes as C

    import spx
    from spx import workloads as W

    A = W.circuit(window=a.window)
    m, n, nnz, k = A["m"], A["n"], A["nnz"], a.k
    P = spx.Plan(A["rowptr"], A["col"], A["val"], n=n, kernel=4, items=a.items, flags=spx.FLAG_FUSED_SPMM)
    X = W.vector(n * k, 5)
    xd, yd = spx.DeviceBuffer.from_array(X), spx.DeviceBuffer(m * k * 8)
    L = spx.lib()

    def timeit(fn):
        for _ in range(a.warmup):
            fn()
        spx.synchronize()
        t = time.perf_counter()
        for _ in range(a.steps):
            fn()
        spx.synchronize()
        return (time.perf_counter() - t) * 1e3 / a.steps

    fused = timeit(lambda: P.execute_multi(k, xd, n, yd, m))
    yf = yd.download(np.float64, m * k)

    def loop():
        for j in range(k):
            spx.check(L.spx_plan_execute(P.h, 1.0, C.c_void_p(xd.ptr + j * n * 8), 0.0,
                                         C.c_void_p(yd.ptr + j * m * 8), None))

    looped = timeit(loop)
    yl = yd.download(np.float64, m * k)
    B = 4 * (m + 1) + 12 * nnz + 8 * n * k + 8 * m * k
    print(json.dumps({
        "metric": f"fused SpMM, k={k} right-hand sides, vs looping the SpMV", "value": round(B / (fused * 1e-3) / 1e9, 2),
        "unit": "GB/s", "n_gpus": 1, "steps": a.steps, "warmup": a.warmup, "ms_per_step": round(fused, 4),
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f64", "data": "synthetic",
        "config": {"workload": f"circuit-like CSR {m}x{n}, nnz={nnz}, X {n}x{k} and Y {m}x{k} column-major"},
        "looped_ms": round(looped, 4), "speedup_vs_loop": round(looped / fused, 3),
        "rel_fused_vs_loop": float(np.linalg.norm(yf - yl) / np.linalg.norm(yl)),
    }), flush=True)


def run_batch(a):
    """The batched form on 512 uneven matrices (n ~ lognormal(ln 14142, 1.32)
    clipped to [1e3, 2e5], Zipf(1.3) rows on [1, min(2048, n)], values
    U[-1,1)): the plain form (one FP64 plan per matrix, the drop-in default
    schedule: kernel 2, 8 items, one partition, one execute per matrix)
    against the learned-selector + FP32-storage batched form (all matrices
    in one plan, one hipGraph replay per batch; the schedule is the fastest of
    the selector's shortlist timed on a sample of the batch).
    Prints one JSON line: matrices/s of both, the ratio, the same against
    looping the drop-in itself (host arrays per call, on every 8th member),
    and the FP32 form's normwise error against the oracle on sampled members."""
    import ctypes as C

    import spx
    from spx import selector as S
    from spx import workloads as W

    t0 = time.time()
    mats = W.batch(a.batch_count, seed=5)
    M = sum(A["m"] for A in mats)
    NNZ = sum(A["nnz"] for A in mats)
    log(f"batch: {len(mats)} matrices, rows {M}, nnz {NNZ} (generated in {time.time() - t0:.1f}s)")
    x = W.vector(M, 23)
    off = np.zeros(len(mats) + 1, np.int64)
    np.cumsum([A["m"] for A in mats], out=off[1:])
    xd, yd = spx.DeviceBuffer.from_array(x), spx.DeviceBuffer(M * 8)
    L = spx.lib()

    # plain form: per-matrix plans, default schedule, FP64
    t0 = time.time()
    plans = [spx.Plan(A["rowptr"], A["col"], A["val"], n=A["n"], kernel=2) for A in mats]
    log(f"plain: {len(plans)} plans built in {time.time() - t0:.1f}s")

    def plain_round():
        for i, P in enumerate(plans):
            spx.check(L.spx_plan_execute(P.h, 1.0, C.c_void_p(xd.ptr + 8 * int(off[i])), 0.0,
                                         C.c_void_p(yd.ptr + 8 * int(off[i])), None))

    for _ in range(a.warmup):
        plain_round()
    spx.synchronize()
    t = time.perf_counter()
    for _ in range(a.steps):
        plain_round()
    spx.synchronize()
    plain_ms = (time.perf_counter() - t) * 1e3 / a.steps
    for P in plans:
        P.close()
    del plans
    log(f"plain: {plain_ms:.3f} ms per batch")
    # the literal baseline: looping the plain drop-in (spMV_mgpu_v1, kernel 2,
    # host arrays in and out per call) over every 8th member, as matrices/s
    t = time.perf_counter()
    nd = 0
    for A in mats[::8]:
        yh = np.zeros(A["m"])
        rc, _ = spx.spMV_mgpu_v1(A["m"], A["n"], A["nnz"], 1.0, A["val"], A["rowptr"], A["col"], x[:A["n"]], 0.0,
                                 yh, 1, kernel=2)
        if rc != 0:
            raise RuntimeError(f"spMV_mgpu_v1 returned {rc}")
        nd += 1
    dropin_mps = nd / (time.perf_counter() - t)
    log(f"drop-in loop: {dropin_mps:.1f} matrices/s over {nd} members")

    # learned selector on the packed block-diagonal matrix
    t0 = time.time()
    rp = np.zeros(M + 1, np.int64)
    nz_off = np.cumsum([0] + [A["nnz"] for A in mats])
    for i, A in enumerate(mats):
        rp[off[i] + 1:off[i + 1] + 1] = A["rowptr"][1:] + nz_off[i]
    col = np.concatenate([A["col"] + np.int32(off[i]) for i, A in enumerate(mats)])
    cands = S.shortlist(S.features(dict(rowptr=rp, col=col, m=M, n=M)), top=a.autotune_top)
    del col, rp
    log(f"selector: shortlist {cands} ({time.time() - t0:.1f}s)")

    def batch_kw(kw):
        bkw = dict(kw)
        if "parts" in bkw:  # kernels 1-3: the selector's partition count over nnz-balanced streams
            bkw["mode"] = spx.MODE_NNZ
        return bkw

    # autotuner: the shortlist timed on every 16th member, packed in FP32 as
    # the real batch will be
    t0 = time.time()
    sample = mats[::16]
    ms_ = sum(A["m"] for A in sample)
    xs, ys = spx.DeviceBuffer.from_array(x[:ms_]), spx.DeviceBuffer(ms_ * 8)
    sched, ttimes = S.autotune(lambda kw: spx.BatchPlan(sample, precision=spx.PREC_F32, **batch_kw(kw)),
                               xs, ys, cands)
    xs.free()
    ys.free()
    kw = S.schedule(sched)
    log(f"autotune on {len(sample)} members: " +
        ", ".join(f"{s} {S.schedule(s)} {t:.3f} ms" for s, t in ttimes.items()) +
        f" -> {sched} ({time.time() - t0:.1f}s)")

    out = {}
    for prec, tag in ((spx.PREC_F32, "fp32"), (spx.PREC_F64, "fp64")):
        t0 = time.time()
        bp = spx.BatchPlan(mats, precision=prec, **batch_kw(kw))
        g = spx.Graph(bp, xd, yd)
        ms = g.time(warmup=a.warmup, reps=a.steps)
        log(f"batched {tag}: {ms:.3f} ms per batch (plan {time.time() - t0:.1f}s)")
        y = yd.download(np.float64, M)
        out[tag] = (ms, y)
        g.close()
        bp.close()

    # the CPU baseline leg: the reference's host loop (the oracle) on 1 thread
    # over 16 sampled members, timed; its y is also the checker's reference
    import oracle
    rng = np.random.default_rng(0)
    errs = []
    cpu_s, cpu_n = 0.0, 0
    for i in sorted(rng.choice(len(mats), min(16, len(mats)), replace=False)):
        A = mats[i]
        t = time.perf_counter()
        ref = oracle.spmv(A["rowptr"], A["col"], A["val"], x[off[i]:off[i + 1]], 1.0, 0.0)
        cpu_s += time.perf_counter() - t
        cpu_n += 1
        yi = out["fp32"][1][off[i]:off[i + 1]]
        errs.append(float(np.linalg.norm(yi - ref) / np.linalg.norm(ref)))
    B32 = sum(W.bytes_model(A["m"], A["n"], A["nnz"], 4) for A in mats)
    B64 = sum(W.bytes_model(A["m"], A["n"], A["nnz"], 8) for A in mats)
    ms32, ms64 = out["fp32"][0], out["fp64"][0]
    res = {
        "metric": "batched matrices/s: learned selector + FP32 storage + one graph vs plain FP64 per-matrix",
        "value": round(len(mats) / (ms32 * 1e-3), 1), "unit": "matrices/s", "n_gpus": 1, "steps": a.steps,
        "warmup": a.warmup, "ms_per_step": round(ms32, 4), "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "f32 storage, f64 accumulation", "data": "synthetic",
        "config": {"workload": f"{len(mats)} CSR matrices, n lognormal(ln 14142, 1.32) in [1e3, 2e5], "
                               f"Zipf(1.3) rows on [1, 2048], {NNZ} nnz total", "schedule": kw},
        "autotune": {"shortlist": [int(c) for c in cands], "chosen": int(sched),
                     "sample_ms": {str(k): round(v, 4) for k, v in ttimes.items()}},
        "plain_fp64_matrices_per_s": round(len(mats) / (plain_ms * 1e-3), 1),
        "dropin_loop_matrices_per_s": round(dropin_mps, 1),
        "speedup_vs_dropin_loop": round(len(mats) / (ms32 * 1e-3) / dropin_mps, 1),
        "batched_fp64_matrices_per_s": round(len(mats) / (ms64 * 1e-3), 1),
        "speedup_vs_plain": round(plain_ms / ms32, 3),
        "fp32_GBs": round(B32 / (ms32 * 1e-3) / 1e9, 1), "fp64_batched_GBs": round(B64 / (ms64 * 1e-3) / 1e9, 1),
        "plain_GBs": round(B64 / (plain_ms * 1e-3) / 1e9, 1),
        "fp32_normwise_rel_err_max": max(errs), "fp32_normwise_rel_err_median": float(np.median(errs)),
        "cpu_baseline": {"value": round(cpu_n / cpu_s, 1), "unit": "matrices/s", "cores": 1, "kind": "port",
                         "sample": f"{cpu_n} sampled members, FP64, one pass each ({cpu_s * 1e3:.1f} ms in all); "
                                   "main.cu:350-356 loop restated in C"},
    }
    print(json.dumps(res), flush=True)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=100)
    ap.add_argument("--warmup", type=int, default=10)
    ap.add_argument("--kernel", type=int, default=4)
    ap.add_argument("--parts", type=int, default=1)
    ap.add_argument("--items", type=int, default=2)
    ap.add_argument("--sigma", type=int, default=0)
    ap.add_argument("--window", type=int, default=1024,
                    help="half-width of the regular rows' column band (0 = uniform columns)")
    ap.add_argument("--dense", type=int, default=200,
                    help="dense rows of the circuit matrix (0: its regular rows alone)")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--cpu-reps", type=int, default=100)
    ap.add_argument("--uniform-reps", type=int, default=20,
                    help="executes timed on the uniform-column variant of the matrix (0 = skip)")
    ap.add_argument("--dropin-reps", type=int, default=3,
                    help="calls of the host-pointer drop-in timed end to end (0 = skip)")
    ap.add_argument("--cold-reps", type=int, default=30,
                    help="executes timed with L2/MALL flushed before each (0 = skip)")
    ap.add_argument("--versions-reps", type=int, default=20,
                    help="executes timed per reference kernel version 1-3 on the same matrix (0 = skip)")
    ap.add_argument("--copy-mb", type=int, default=0,
                    help="after the timed region, one device-to-device copy of this many MiB: the "
                         "known-size copy that scales the FETCH_SIZE/WRITE_SIZE counters")
    ap.add_argument("--workload", choices=["circuit", "batch", "rail", "asic", "prec-circuit", "spmm", "c1"], default="circuit",
                    help="circuit: the 5.5M-row SpMV (default); batch: 512 matrices, batched form vs plain; "
                         "rail / asic: the precision form (FP32 storage) vs plain FP64; spmm: fused multi-RHS")
    ap.add_argument("--k", type=int, default=8, help="spmm: right-hand sides")
    ap.add_argument("--long-row", type=int, default=0,
                    help="rail / asic: long-row threshold (0: automatic, < 0: no long rows)")
    ap.add_argument("--batch-count", type=int, default=512)
    ap.add_argument("--autotune-top", type=int, default=4,
                    help="batch: schedules of the selector's shortlist the autotuner times")
    a = ap.parse_args()
    if a.gpus != 1:
        print(json.dumps({"error": "the library drives one MI355X; --gpus must be 1 (sharding over "
                                   "several GPUs is designed, not implemented)"}))
        sys.exit(2)

    if a.workload == "batch":
        run_batch(a)
        return
    if a.workload in ("rail", "asic", "prec-circuit"):
        run_precision(a)
        return
    if a.workload == "spmm":
        run_spmm(a)
        return
    if a.workload == "c1":
        run_c1(a)
        return

    import spx
    from spx import workloads as W

    t0 = time.time()
    A = W.circuit(window=a.window, n_dense=a.dense)
    x = W.vector(A["n"], 22)
    m, n, nnz = A["m"], A["n"], A["nnz"]
    log(f"workload: m={m} n={n} nnz={nnz} mean_row={A['mean_row']:.2f} max_row={A['max_row']} "
        f"(generated in {time.time() - t0:.1f}s)")

    t0 = time.time()
    plan = spx.Plan(A["rowptr"], A["col"], A["val"], n=n, kernel=a.kernel, parts=a.parts, items=a.items,
                    sigma=a.sigma)
    info = plan.info()
    log(f"plan: {info} (built in {time.time() - t0:.1f}s)")
    xd = spx.DeviceBuffer.from_array(x)
    yd = spx.DeviceBuffer(m * 8)
    alpha, beta = 1.0, 0.0

    for _ in range(a.warmup):
        plan.execute(xd, yd, alpha, beta)
    spx.synchronize()

    # timed region: K back-to-back executes on one stream, nothing between
    # them; the library reads the host clock from the first timed enqueue to
    # the end of the last execute (synchronized on both sides)
    spx.synchronize()
    wall_ms = plan.time_wall(xd, yd, alpha, beta, warmup=0, reps=a.steps)
    spx.synchronize()
    # then K more with HIP events around each execute and around its SpMV
    # kernel (each event pair costs the stream a few microseconds)
    ex_arr, ker_arr, wall_ev = plan.time_each(xd, yd, alpha, beta, warmup=0, reps=a.steps)
    spx.synchronize()
    tot_ms, ker_ms = float(ex_arr.mean()), float(ker_arr.mean())
    ker_med = float(np.median(ker_arr))
    log(f"timed {a.steps} steps: wall {wall_ms:.4f} ms/step; with events: wall {wall_ev:.4f} ms/step, "
        f"{tot_ms:.4f} ms/execute, kernel mean {ker_ms:.4f} / median {ker_med:.4f} ms")

    # the same kernel with L2 and MALL cold (the HBM-labelled figure): before
    # every execute a 1 GiB buffer is read, which evicts both caches and leaves
    # no dirty lines; for comparison, a 512 MiB buffer overwritten instead
    # (its dirty lines are written back while the kernel runs)
    cold = None
    if a.cold_reps > 0:
        _, kc, _ = plan.time_each(xd, yd, alpha, beta, warmup=1, reps=a.cold_reps, flush_bytes=1 << 30)
        _, kw, _ = plan.time_each(xd, yd, alpha, beta, warmup=1, reps=a.cold_reps, flush_bytes=-(512 << 20))
        cold = {"mean": float(kc.mean()), "median": float(np.median(kc)), "write_flush_median": float(np.median(kw))}
        log(f"cold caches ({a.cold_reps} executes each): kernel mean {cold['mean']:.4f} / median "
            f"{cold['median']:.4f} ms after a 1 GiB read; median {cold['write_flush_median']:.4f} ms after a "
            f"512 MiB overwrite")

    # the reference's kernel versions on the same matrix (the drop-in's
    # kernels), each timed warm: their SpMV kernels and the whole execute
    versions = None
    if a.versions_reps > 0 and a.kernel == 4:
        versions = {}
        for kv, kw_ in ((1, {}), (2, {"items": 8}), (3, {})):
            Pv = spx.Plan(A["rowptr"], A["col"], A["val"], n=n, kernel=kv, **kw_)
            wv = Pv.time_wall(xd, yd, alpha, beta, warmup=3, reps=a.versions_reps)
            ev, kk, _ = Pv.time_each(xd, yd, alpha, beta, warmup=0, reps=a.versions_reps)
            versions[f"kernel{kv}"] = {"ms_per_step": round(wv, 5), "execute_ms_median": round(float(np.median(ev)), 5),
                                       "kernels_ms_median": round(float(np.median(kk)), 5),
                                       "GBs": round(W.bytes_model(m, n, nnz, 8, beta != 0.0) / (wv * 1e-3) / 1e9, 1)}
            log(f"kernel version {kv}: wall {wv:.4f} ms/step, execute median {np.median(ev):.4f} ms")
            Pv.close()

    # the same sizes with every row's columns uniform at random (no locality
    # at all: the gathers, not the matrix stream, bound it)
    uniform = None
    if a.uniform_reps > 0:
        Au = W.circuit(window=0)
        Pu = spx.Plan(Au["rowptr"], Au["col"], Au["val"], n=n, kernel=a.kernel, parts=a.parts, items=a.items,
                      sigma=a.sigma)
        yu = spx.DeviceBuffer(m * 8)  # its own y: the headline's y is checked below
        tu, ku, wu = Pu.time(xd, yu, alpha, beta, warmup=3, reps=a.uniform_reps, wall=True)
        yu.free()
        Bu = W.bytes_model(Au["m"], Au["n"], Au["nnz"], 8, beta != 0.0)
        uniform = {"workload": f"same generator, every row's columns uniform at random (nnz={Au['nnz']})",
                   "ms_per_step": round(wu, 5), "kernel_ms": round(ku, 5), "GBs": round(Bu / (wu * 1e-3) / 1e9, 2),
                   "kernel_frac_of_hbm_peak": round(Bu / (ku * 1e-3) / 1e9 / HBM_PEAK_GBS, 4),
                   "column_bands": Pu.info()["bands"]}
        log(f"uniform columns: wall {wu:.4f} ms/step, kernel {ku:.4f} ms")
        Pu.close()
        del Au

    # the drop-in path end to end: host arrays in, host y out (upload, plan,
    # SpMV, download, free per call), as the reference's callers use it
    dropin_ms = None
    dropin_phases = None
    if a.dropin_reps > 0:
        yh = np.zeros(m)
        ts = []
        for _ in range(a.dropin_reps + 1):
            t = time.perf_counter()
            rc, _ = spx.spMV_mgpu_v1(m, n, nnz, alpha, A["val"], A["rowptr"], A["col"], x, beta, yh, 1, kernel=2)
            ts.append(time.perf_counter() - t)
            if rc != 0:
                raise RuntimeError(f"spMV_mgpu_v1 returned {rc}")
        dropin_ms = float(np.median(ts[1:])) * 1e3
        dropin_phases = spx.dropin_last_times()
        log(f"drop-in spMV_mgpu_v1 (kernel 2) end to end: {dropin_ms:.1f} ms per call; last call's phases: "
            + ", ".join(f"{k} {v:.2f}" for k, v in dropin_phases.items()))

    if a.copy_mb > 0:
        nb = a.copy_mb << 20
        src, dst = spx.DeviceBuffer(nb), spx.DeviceBuffer(nb)
        spx.check(spx.lib().spx_memcpy_d2d(dst.ptr, src.ptr, nb))
        spx.synchronize()
        log(f"calibration copy: {a.copy_mb} MiB device to device")
        src.free()
        dst.free()

    B = W.bytes_model(m, n, nnz, 8, beta != 0.0)
    y = yd.download(np.float64, m)
    value = B / (wall_ms * 1e-3) / 1e9
    # the roofline figure: the SpMV kernel's mean duration with cold caches
    # when measured, else the warm mean
    k_roof = cold["mean"] if cold else ker_ms
    achieved = B / (k_roof * 1e-3) / 1e9
    traffic = profiled_traffic(a.kernel, info["items"], a.window)
    out = {
        "metric": "SpMV GB/s & % HBM peak; learned+FP32-storage vs plain FP64 at stated rel. err",
        "value": round(value, 2),
        "unit": "GB/s",
        "n_gpus": 1,
        "steps": a.steps,
        "warmup": a.warmup,
        "ms_per_step": round(wall_ms, 5),
        "ms_per_step_with_events": round(wall_ev, 5),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic",
        "config": {
            "workload": (f"circuit-like skewed CSR {m}x{n}, nnz={nnz} (200 dense rows 1e4..1.3e6 nnz at "
                         f"uniform columns; other rows 1..12 nnz within +-{a.window} of the diagonal), "
                         "FP64 values, y = A x (alpha=1, beta=0)"),
            "kernel": a.kernel, "parts": info["parts"], "sigma": info["sigma"], "items": info["items"],
            "mode": "segmented units, one launch" if a.kernel == 4 else "nnz-balanced partitions (v1)",
        },
        "gflops": round(2 * nnz / (wall_ms * 1e-3) / 1e9, 2),
        "uniform_columns": uniform,
        "dropin_end_to_end": None if dropin_ms is None else {
            "entry": "spMV_mgpu_v1(kernel=2), host arrays", "ms_per_call": round(dropin_ms, 2),
            "GBs": round(B / (dropin_ms * 1e-3) / 1e9, 2),
            "last_call_phases_ms": {k: round(v, 3) for k, v in dropin_phases.items()}},
        "ref_getB_gbps": round(W.ref_getB(m, nnz) / (wall_ms * 1e-3) / 1e9, 2),
        "roofline": {
            "bound": "hbm", "achieved": round(achieved, 2), "peak": HBM_PEAK_GBS, "unit": "GB/s",
            "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": traffic,
            "bytes_per_launch": B,
            "basis": ("algorithmic bytes / mean SpMV-kernel time with L2 and MALL cold (a 1 GiB read before each "
                      "execute); the long-row fold after the kernel is outside it") if cold else
                     "algorithmic bytes / mean SpMV-kernel time, warm",
            "kernel_ms_cold_mean": None if cold is None else round(cold["mean"], 5),
            "kernel_ms_cold_median": None if cold is None else round(cold["median"], 5),
            "kernel_ms_cold_write_flush_median": None if cold is None else round(cold["write_flush_median"], 5),
            "kernel_ms_warm_mean": round(ker_ms, 5), "kernel_ms_warm_median": round(ker_med, 5),
            "frac_warm_median": round(B / (ker_med * 1e-3) / 1e9 / HBM_PEAK_GBS, 4),
            "traffic_frac_cold": None if (traffic is None) else round(traffic / (k_roof * 1e-3) / 1e9 / HBM_PEAK_GBS, 4),
        },
        "kernel_versions": versions,
    }

    if not a.no_cpu:
        import oracle  # the CPU baseline leg: the reference's host loop, restated in C

        reps = []
        ref = None
        for i in range(a.cpu_reps + 1):
            t = time.perf_counter()
            ref = oracle.spmv(A["rowptr"], A["col"], A["val"], x, alpha, beta)
            if i:
                reps.append(time.perf_counter() - t)
        t_cpu = float(np.median(reps))
        out["cpu_baseline"] = {
            "value": round(B / t_cpu / 1e9, 3), "unit": "GB/s", "cores": 1, "kind": "port",
            "sample": f"whole matrix, median of {a.cpu_reps} reps after 1 warm-up, 1 thread "
                      f"({t_cpu * 1e3:.1f} ms/SpMV); main.cu:350-356 loop restated in C",
        }
        # the same loop with rows split statically over the host's threads
        threads = int(os.environ.get("OMP_NUM_THREADS") or os.cpu_count() or 1)
        reps = []
        for i in range(a.cpu_reps + 1):
            t = time.perf_counter()
            y_omp = oracle.spmv_omp(A["rowptr"], A["col"], A["val"], x, alpha, beta, threads=threads)
            if i:
                reps.append(time.perf_counter() - t)
        t_omp = float(np.median(reps))
        if not np.array_equal(y_omp, ref):
            raise AssertionError("the OpenMP host loop differs from the sequential one")
        out["cpu_baseline_openmp"] = {
            "value": round(B / t_omp / 1e9, 3), "unit": "GB/s", "cores": threads, "kind": "port",
            "sample": f"whole matrix, median of {a.cpu_reps} reps after 1 warm-up, {threads} OpenMP threads, "
                      f"static row split ({t_omp * 1e3:.1f} ms/SpMV)",
        }
        err = float(np.linalg.norm(y - ref) / np.linalg.norm(ref))
        out["err_vs_oracle"] = err
        if not err <= 1e-13:
            raise AssertionError(f"SpMV differs from the oracle: normwise relative error {err:.3e} > 1e-13")
        log(f"cpu baseline {t_cpu * 1e3:.1f} ms/SpMV (1 thread), {t_omp * 1e3:.1f} ms/SpMV ({threads} threads); "
            f"normwise rel err vs oracle {err:.3e}")
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
