"""Learned schedule selector: picks kernel, tile width and partition count
for a CSR matrix from 32 device-computed features with a small MLP
(32 -> 64 -> 64 -> 96, FP32) whose training and inference run in the
library's own HIP kernels.  It replaces the reference harness's exhaustive
(d, c) search (SB/test/dspmv_test.cu:314-332); exhaustive search stays
available (`exhaustive`) as the fallback and as the yardstick of regret.

Schedule s in [0, 96): kernel 1 + s // 24, width index (s // 6) % 4, parts
PARTS[s % 6] (kernel 4 ignores parts, so its six copies share one timing).

Training data are the method's own timings: `collect` times every distinct
schedule on seeded matrices (laws: Zipf, lognormal, banded, block-diagonal
with dense rows) and stores features and times; `train` fits the MLP on the
GPU against soft targets that favour near-best schedules.

    python -m spx.selector collect OUT.npz --seeds 0:1000 --schedules single
    python -m spx.selector merge TRAIN.npz PART1.npz PART2.npz ...
    python -m spx.selector retime DATA.npz OUT.npz --kernels 1,2,3
    python -m spx.selector train TRAIN.npz --model MODEL.npz
    python -m spx.selector evaluate HELDOUT.npz --model MODEL.npz --train TRAIN.npz
"""
from __future__ import annotations

import argparse
import ctypes as C
import math
import os
import sys
import time

import numpy as np

from . import DeviceBuffer, Plan, SpxError, check, lib, synchronize
from . import workloads as W

NSCHED = 96
NFEAT = 32
LANES = (2, 8, 32, 64)
ITEMS = (2, 4, 8, 16)
SIGMAS = (4, 8, 16, 32)
CHUNKS = (1, 2, 4, 8)
PARTS = (1, 2, 3, 4, 6, 8)
DEFAULT = 2 * 6 + 0 + 24  # kernel 2, items 8, one partition: the plain drop-in's v1 form

HERE = os.path.dirname(os.path.abspath(__file__))
MODEL_PATH = os.path.join(HERE, "selector_model.npz")


def schedule(s):
    """Plan keyword arguments of schedule s."""
    if not 0 <= s < NSCHED:
        raise ValueError("schedule out of range")
    k, w, p = 1 + s // 24, (s // 6) % 4, PARTS[s % 6]
    if k == 1:
        return dict(kernel=1, lanes=LANES[w], parts=p)
    if k == 2:
        return dict(kernel=2, items=ITEMS[w], parts=p)
    if k == 3:
        return dict(kernel=3, sigma=SIGMAS[w], parts=p)
    return dict(kernel=4, items=CHUNKS[w])


def representative(s):
    """The schedule whose timing stands for s (kernel 4 drops the parts)."""
    return s - s % 6 if s >= 72 else s


def distinct():
    return [s for s in range(NSCHED) if representative(s) == s]


def single_partition():
    """The distinct schedules with one partition (kernels 1-3) and kernel 4:
    on one GPU more stream partitions only split the same CUs."""
    return [s for s in distinct() if s >= 72 or s % 6 == 0]


def long_row_threshold(n):
    return max(4096, 4 * ((n + 2047) // 2048))


def features(A):
    """The 32 features of CSR matrix A (dict with rowptr, col, m, n), computed
    on the device by spx_sel_features."""
    m, n = A["m"], A["n"]
    rp = np.ascontiguousarray(A["rowptr"], np.int32)
    col = np.ascontiguousarray(A["col"], np.int32)
    nnz = int(rp[-1]) if m else 0
    dr = DeviceBuffer.from_array(rp)
    dc = DeviceBuffer.from_array(col if nnz else np.zeros(1, np.int32))
    df = DeviceBuffer(NFEAT * 4)
    check(lib().spx_sel_features(m, n, nnz, dr.ptr, dc.ptr, long_row_threshold(n), df.ptr, None))
    return df.download(np.float32, NFEAT)


class MLP:
    """The selector network; parameters live on the device."""

    def __init__(self, params=None, seed=0):
        self.np = lib().spx_sel_param_count()
        if params is None:
            rng = np.random.default_rng(seed)
            parts = []
            for fan_in, fan_out in ((32, 64), (64, 64), (64, 96)):
                parts.append(rng.normal(0, math.sqrt(2.0 / fan_in), fan_out * fan_in))
                parts.append(np.zeros(fan_out))
            params = np.concatenate(parts)
        params = np.ascontiguousarray(params, np.float32)
        if params.size != self.np:
            raise ValueError(f"expected {self.np} parameters")
        self.p = DeviceBuffer.from_array(params)
        self.m1 = DeviceBuffer.from_array(np.zeros(self.np, np.float32))
        self.m2 = DeviceBuffer.from_array(np.zeros(self.np, np.float32))
        self.step = 0

    def params(self):
        return self.p.download(np.float32, self.np)

    def forward(self, X):
        X = np.ascontiguousarray(X, np.float32).reshape(-1, NFEAT)
        B = X.shape[0]
        dx, dz = DeviceBuffer.from_array(X), DeviceBuffer(max(B, 1) * NSCHED * 4)
        work = DeviceBuffer(max(1, lib().spx_mlp_work_floats(B)) * 4)
        check(lib().spx_mlp_forward(self.p.ptr, dx.ptr, B, dz.ptr, work.ptr, None))
        synchronize()
        return dz.download(np.float32, B * NSCHED).reshape(B, NSCHED)

    def grad(self, X, T):
        """(mean loss, gradient) at the current parameters, no update."""
        return self._step(X, T, 0.0, 0.0)

    def _step(self, X, T, lr, wd):
        X = np.ascontiguousarray(X, np.float32).reshape(-1, NFEAT)
        T = np.ascontiguousarray(T, np.float32).reshape(-1, NSCHED)
        B = X.shape[0]
        dx, dt = DeviceBuffer.from_array(X), DeviceBuffer.from_array(T)
        dl, dg = DeviceBuffer(B * 4), DeviceBuffer(self.np * 4)
        work = DeviceBuffer(lib().spx_mlp_work_floats(B) * 4)
        if lr > 0:
            self.step += 1
        check(lib().spx_mlp_train_step(self.p.ptr, self.m1.ptr, self.m2.ptr, max(self.step, 1), C.c_float(lr),
                                       C.c_float(wd), dx.ptr, dt.ptr, B, dl.ptr, dg.ptr, work.ptr, None))
        synchronize()
        return float(dl.download(np.float32, B).mean()), dg.download(np.float32, self.np)

    def fit(self, X, T, epochs=400, batch=64, lr=3e-3, wd=1e-4, seed=0, log=None):
        rng = np.random.default_rng(seed)
        N = X.shape[0]
        loss = float("nan")
        for ep in range(epochs):
            order = rng.permutation(N)
            tot = 0.0
            for i in range(0, N, batch):
                idx = order[i:i + batch]
                l, _ = self._step(X[idx], T[idx], lr, wd)
                tot += l * len(idx)
            loss = tot / N
            if log and (ep % 50 == 0 or ep == epochs - 1):
                log(f"epoch {ep}: loss {loss:.4f}")
        return loss

    def choose(self, X, mask=None):
        """Argmax schedule per feature row, on the device (spx_mlp_argmax)."""
        Z = self.forward(X)
        B = Z.shape[0]
        dz, do = DeviceBuffer.from_array(Z), DeviceBuffer(max(B, 1) * 4)
        dm = DeviceBuffer.from_array(np.asarray(mask, np.int32)) if mask is not None else None
        check(lib().spx_mlp_argmax(dz.ptr, B, dm.ptr if dm else None, do.ptr, None))
        synchronize()
        return do.download(np.int32, B)

    def save(self, path):
        np.savez(path, params=self.params())

    @classmethod
    def load(cls, path=MODEL_PATH):
        with np.load(path, allow_pickle=False) as f:
            return cls(f["params"])


def targets(times, tau=0.05):
    """Soft targets: weight exp(-(t/t_best - 1)/tau) over the timed schedules,
    normalised per matrix (untimed entries, NaN, get zero)."""
    T = np.zeros_like(times, dtype=np.float64)
    for i, t in enumerate(times):
        ok = np.isfinite(t)
        r = t[ok] / t[ok].min() - 1.0
        w = np.exp(-r / tau)
        T[i, ok] = w / w.sum()
    return T.astype(np.float32)


def regret(times, choice):
    """t(choice) / t(best) - 1 per matrix."""
    t = np.asarray(times)
    c = np.asarray(choice)
    return t[np.arange(len(c)), c] / np.nanmin(t, axis=1) - 1.0


def select(A, model=None):
    """(schedule index, Plan kwargs) the trained model picks for A."""
    model = model or MLP.load()
    s = int(model.choose(features(A)[None, :])[0])
    return s, schedule(s)


def shortlist(f, top=4, model=None):
    """The model's `top` highest-scoring distinct schedules for feature row f,
    best first (the first is what select() picks)."""
    model = model or MLP.load()
    z = model.forward(np.asarray(f, np.float32)[None, :])[0]
    out = []
    for s in np.argsort(-z, kind="stable"):
        r = representative(int(s))
        if r not in out:
            out.append(r)
        if len(out) == top:
            break
    return out


def autotune(make_plan, x_dev, y_dev, candidates, warmup=2, reps=5):
    """Time each candidate schedule on the plans make_plan(kwargs) builds (the
    real operand: a batch sample in its final precision, say) and return
    (fastest schedule, {schedule: ms per execute}).  A candidate whose plan
    cannot be built is skipped."""
    times = {}
    for s in candidates:
        try:
            P = make_plan(schedule(s))
        except SpxError as e:
            print(f"autotune: schedule {s} {schedule(s)}: {e}", file=sys.stderr, flush=True)
            continue
        times[s] = P.time(x_dev, y_dev, warmup=warmup, reps=reps)[0]
        P.close()
    if not times:
        raise SpxError(-1, "autotune: no candidate schedule could be built")
    return min(times, key=times.get), times


def exhaustive(A, x=None, warmup=2, reps=5, only=None, threads=6, batch=16):
    """Time every distinct schedule on A (or those in `only`; the rest stay
    NaN) in ms per execute from HIP events; the copies of a kernel-4 schedule
    share its time.  Large matrices get fewer repetitions.  Plans are built
    `batch` at a time on `threads` host threads (the kernel-4 layout is built
    on the host), then timed one after another with the GPU otherwise idle."""
    from concurrent.futures import ThreadPoolExecutor

    x = W.vector(A["n"], 17) if x is None else x
    xd, yd = DeviceBuffer.from_array(x), DeviceBuffer(max(A["m"], 1) * 8)
    if A["nnz"] > 2e7:
        warmup, reps = 1, 3
    t = np.full(NSCHED, np.nan)

    def build(s):
        try:
            return s, Plan(A["rowptr"], A["col"], A["val"], n=A["n"], **schedule(s))
        except Exception as e:  # a schedule that cannot be built stays untimed (NaN)
            print(f"schedule {s} {schedule(s)}: {e}", file=sys.stderr, flush=True)
            return s, None

    # kernel-4 plans first: their host builds take longest
    order = sorted(distinct() if only is None else only, key=lambda s: (s < 72, s))
    with ThreadPoolExecutor(max(1, threads)) as pool:
        for i in range(0, len(order), batch):
            for s, P in list(pool.map(build, order[i:i + batch])):
                if P is None:
                    continue
                t[s] = P.time(xd, yd, warmup=warmup, reps=reps)[0]
                P.close()
    for s in range(NSCHED):
        t[s] = t[representative(s)]
    return t


def training_matrix(seed, nmin=1e4, nmax=6e6, nnz_cap=4e7):
    """A seeded matrix for the selector's data: n log-uniform in [nmin, nmax],
    mean row length log-uniform in [2, 200] (capped by nnz_cap), law uniform
    over Zipf, lognormal, banded (bandwidth 3..101), block-diagonal with
    dense rows."""
    rng = np.random.default_rng(1_000_003 * seed + 7)
    n = int(math.exp(rng.uniform(math.log(nmin), math.log(nmax))))
    mean = math.exp(rng.uniform(math.log(2), math.log(200)))
    mean = max(1.5, min(mean, nnz_cap / n))
    law = int(rng.integers(4))
    s = 10_000 + seed
    if law == 0:
        z = rng.uniform(1.05, 2.5)
        kmax = _zipf_kmax(z, mean, n)
        return W.zipf(n, z, kmax, s), "zipf"
    if law == 1:
        sg = rng.uniform(0.5, 2.0)
        return W.lognormal_rows(n, n, math.log(mean) - sg * sg / 2, sg, s, kmax=n), "lognormal"
    if law == 2:
        bw = int(rng.integers(3, 102))
        return W.banded(int(min(n, nnz_cap / bw)), bw, s), "banded"
    block = int(rng.integers(16, 513))
    nd = int(rng.integers(1, 33))
    dlen = int(min(n, max(1024, n * rng.uniform(0.005, 0.2), nnz_cap * 0.3 / nd)))
    return W.blockdiag(n, block, min(mean, block), nd, dlen, s), "blockdiag"


def _zipf_kmax(z, mean, n):
    """kmax at which Zipf(z) truncated to [1, kmax] has about this mean."""
    def zmean(k):
        i = np.arange(1, k + 1, dtype=np.float64)
        w = i ** -z
        return float((i * w).sum() / w.sum())
    lo, hi = 1, min(n, 1 << 20)
    if zmean(hi) <= mean:
        return hi
    while hi - lo > 1:
        mid = (lo + hi) // 2
        if zmean(mid) < mean:
            lo = mid
        else:
            hi = mid
    return hi


GEN_DEFAULT = (1e4, 6e6, 4e7)
# the generator bounds of data files that do not record theirs
GEN_LEGACY = (1e4, 2e6, 8e6)


def gen_params(f):
    """(nmin, nmax, nnz_cap) a data file's matrices were drawn with."""
    return tuple(float(v) for v in f["gen"]) if "gen" in f.files else GEN_LEGACY


def collect(seeds, out, schedules="all", prefetch=3, budget=None, **kw):
    """Features and exhaustive timings of training matrices; saved as npz
    after every matrix, with the generator bounds in `gen`.  schedules: "all"
    distinct ones, or "single" (one partition and kernel 4).  The next
    `prefetch` matrices are generated on host threads while the GPU times
    the current one.  With a budget (seconds), no matrix is started after it
    is spent; the seeds done are in the file."""
    from concurrent.futures import ThreadPoolExecutor

    only = None if schedules == "all" else single_partition()
    nmin, nmax, cap = kw.get("nmin", GEN_DEFAULT[0]), kw.get("nmax", GEN_DEFAULT[1]), kw.get("nnz_cap", GEN_DEFAULT[2])
    F, T, meta = [], [], []
    t_start = time.time()
    with ThreadPoolExecutor(max(1, prefetch)) as pool:
        futs = {}

        def ask(j):
            if j < len(seeds) and j not in futs:
                futs[j] = pool.submit(training_matrix, seeds[j], nmin, nmax, cap)

        for j in range(prefetch):
            ask(j)
        for i, seed in enumerate(seeds):
            if budget is not None and time.time() - t_start > budget:
                print(f"budget spent after {i} matrices", flush=True)
                break
            t0 = time.time()
            ask(i)
            A, law = futs.pop(i).result()
            ask(i + prefetch)
            tg = time.time() - t0
            f = features(A)
            t = exhaustive(A, only=only)
            F.append(f)
            T.append(t)
            meta.append((seed, A["m"], A["nnz"], law))
            best = int(np.nanargmin(t))
            print(f"[{i + 1}/{len(seeds)}] seed {seed} {law} m={A['m']} nnz={A['nnz']} best={best} "
                  f"{schedule(best)} {np.nanmin(t) * 1e3:.1f}us (wait {tg:.1f}s, {time.time() - t0:.1f}s, "
                  f"total {time.time() - t_start:.0f}s)", flush=True)
            np.savez(out, features=np.array(F, np.float32), times=np.array(T), seeds=np.array([m[0] for m in meta]),
                     m=np.array([m[1] for m in meta]), nnz=np.array([m[2] for m in meta]),
                     law=np.array([m[3] for m in meta]), gen=np.array([nmin, nmax, cap]))
            del A
    return out


def retime(path, out, kernels, schedules="all", start=0, budget=None, prefetch=3, min_n=0, resume=False):
    """Re-time the schedules of `kernels` (after a change to those kernels) on
    the matrices of data file `path`, from matrix `start` on, keeping every
    other timing; written to `out` after every matrix (the matrices not yet
    re-timed keep their old times, and `retimed` counts the done ones;
    `resume` continues from there).  Matrices with fewer than `min_n` rows
    (square: columns) are skipped, for a change that only affects larger
    ones."""
    from concurrent.futures import ThreadPoolExecutor

    with np.load(path, allow_pickle=False) as f:
        cols = {k: f[k] for k in f.files}
        gen = gen_params(f)
    pool_s = distinct() if schedules == "all" else single_partition()
    only = [s for s in pool_s if 1 + min(s, 95) // 24 in kernels]
    T = cols["times"].copy()
    seeds = [int(v) for v in cols["seeds"]]
    todo = [i for i in range(len(seeds)) if int(cols["m"][i]) >= min_n]
    done = int(cols["retimed"]) if resume and "retimed" in cols else start
    t_start = time.time()
    with ThreadPoolExecutor(max(1, prefetch)) as pool:
        futs = {}

        order = [i for i in todo if i >= done]

        def ask(q):
            if q < len(order) and order[q] not in futs:
                futs[order[q]] = pool.submit(training_matrix, seeds[order[q]], *gen)

        for q in range(prefetch):
            ask(q)
        for q, i in enumerate(order):
            if budget is not None and time.time() - t_start > budget:
                print(f"budget spent at matrix {i}", flush=True)
                break
            t0 = time.time()
            ask(q)
            A, _ = futs.pop(i).result()
            ask(q + prefetch)
            if A["nnz"] != int(cols["nnz"][i]):
                raise RuntimeError(f"seed {seeds[i]}: regenerated nnz {A['nnz']} != recorded {cols['nnz'][i]}")
            t = exhaustive(A, only=only)
            for s in range(NSCHED):
                if 1 + min(s, 95) // 24 in kernels:
                    T[i, s] = t[s]
            best = int(np.nanargmin(T[i]))
            print(f"[{i + 1}/{len(seeds)}] seed {seeds[i]} nnz={A['nnz']} best={best} {schedule(best)} "
                  f"({time.time() - t0:.1f}s, total {time.time() - t_start:.0f}s)", flush=True)
            cols["times"] = T
            cols["retimed"] = np.array(i + 1)
            np.savez(out, **cols)
            del A
        else:
            cols["retimed"] = np.array(len(seeds))
            np.savez(out, **cols)
    return out


def merge(parts, out):
    """Concatenate collected npz files (same fields, same generator bounds)
    into one."""
    keys = ("features", "times", "seeds", "m", "nnz", "law")
    cols = {k: [] for k in keys}
    gens = set()
    for path in parts:
        with np.load(path, allow_pickle=False) as f:
            for k in keys:
                cols[k].append(f[k])
            gens.add(gen_params(f))
    if len(gens) != 1:
        raise ValueError(f"parts drawn with different generator bounds: {sorted(gens)}")
    np.savez(out, gen=np.array(gens.pop()), **{k: np.concatenate(v) for k, v in cols.items()})
    return out


def _seeds(spec):
    a, b = spec.split(":")
    return list(range(int(a), int(b)))


def main(argv=None):
    ap = argparse.ArgumentParser(prog="python -m spx.selector")
    sub = ap.add_subparsers(dest="cmd", required=True)
    c = sub.add_parser("collect")
    c.add_argument("out")
    c.add_argument("--seeds", default="0:8")
    c.add_argument("--nmax", type=float, default=6e6)
    c.add_argument("--nnz-cap", type=float, default=4e7)
    c.add_argument("--schedules", choices=["all", "single"], default="all")
    c.add_argument("--budget", type=float, default=None, help="seconds after which no new matrix is started")
    t = sub.add_parser("train")
    t.add_argument("data", nargs="+")
    t.add_argument("--model", default=MODEL_PATH)
    t.add_argument("--epochs", type=int, default=400)
    e = sub.add_parser("evaluate")
    e.add_argument("data")
    e.add_argument("--model", default=MODEL_PATH)
    e.add_argument("--train", nargs="*", default=[], help="training sets: the best fixed schedule is chosen on them")
    r = sub.add_parser("retime")
    r.add_argument("data")
    r.add_argument("out")
    r.add_argument("--kernels", default="1,2,3")
    r.add_argument("--schedules", choices=["all", "single"], default="all")
    r.add_argument("--start", type=int, default=0)
    r.add_argument("--budget", type=float, default=None)
    r.add_argument("--min-n", type=int, default=0, help="skip matrices with fewer rows")
    r.add_argument("--resume", action="store_true", help="continue an interrupted re-timing of the same file")
    m = sub.add_parser("merge")
    m.add_argument("out")
    m.add_argument("parts", nargs="+")
    a = ap.parse_args(argv)
    if a.cmd == "merge":
        merge(a.parts, a.out)
    elif a.cmd == "retime":
        retime(a.data, a.out, [int(k) for k in a.kernels.split(",")], schedules=a.schedules, start=a.start,
               budget=a.budget, min_n=a.min_n, resume=a.resume)
    elif a.cmd == "collect":
        collect(_seeds(a.seeds), a.out, schedules=a.schedules, budget=a.budget, nmax=a.nmax, nnz_cap=a.nnz_cap)
    elif a.cmd == "train":
        X, Tm = [], []
        for path in a.data:
            with np.load(path, allow_pickle=False) as f:
                X.append(f["features"])
                Tm.append(f["times"])
        X, Tm = np.concatenate(X), np.concatenate(Tm)
        net = MLP(seed=0)
        loss = net.fit(X, targets(Tm), epochs=a.epochs, log=lambda s: print(s, flush=True))
        net.save(a.model)
        ch = net.choose(X)
        print(f"trained on {len(X)} matrices: loss {loss:.4f}, train regret mean "
              f"{regret(Tm, ch).mean():.4f}", flush=True)
    else:
        with np.load(a.data, allow_pickle=False) as f:
            X, Tm = f["features"], f["times"]
            nnz = f["nnz"] if "nnz" in f.files else None
        train_times = None
        if a.train:
            train_times = np.concatenate([np.load(p, allow_pickle=False)["times"] for p in a.train])
        net = MLP.load(a.model)
        print(report(net, X, Tm, train_times=train_times, nnz=nnz), flush=True)


def report(net, X, Tm, train_times=None, nnz=None):
    """Regret of the selector against exhaustive search, beside the default
    schedule and the best fixed schedule (chosen on train_times, else on these
    times); how often a stream-partition variant (parts > 1, kernels 1-3) is
    the fastest schedule, and what leaving them all out costs; with nnz, the
    selector's mean regret per nnz decade."""
    ch = net.choose(X)
    ref = Tm if train_times is None else train_times
    rel = ref / np.nanmin(ref, axis=1, keepdims=True)
    timed = ~np.isnan(rel).all(axis=0)  # schedules the reference set timed
    fixed = int(np.flatnonzero(timed)[np.argmin(np.nanmean(rel[:, timed], axis=0))])
    r = regret(Tm, ch)
    best = np.nanargmin(Tm, axis=1)
    multi = np.array([s < 72 and s % 6 != 0 for s in best])
    single = [s for s in range(NSCHED) if s >= 72 or s % 6 == 0]
    t1 = np.nanmin(Tm[:, single], axis=1)
    out = {
        "matrices": int(len(X)),
        "selector_regret_mean": float(r.mean()),
        "selector_regret_median": float(np.median(r)),
        "selector_regret_p90": float(np.quantile(r, 0.9)),
        "exact_best_fraction": float(np.mean(r <= 1e-12)),
        "default_regret_mean": float(regret(Tm, np.full(len(X), DEFAULT)).mean()),
        "fixed_best": int(fixed),
        "fixed_best_schedule": schedule(fixed),
        "fixed_regret_mean": float(regret(Tm, np.full(len(X), fixed)).mean()),
        "multi_partition_best_fraction": float(multi.mean()),
        "single_partition_only_regret_mean": float(np.mean(t1 / np.nanmin(Tm, axis=1) - 1.0)),
        "best_kernel_share": {int(k): float(np.mean(1 + np.minimum(best, 95) // 24 == k)) for k in (1, 2, 3, 4)},
    }
    if nnz is not None:
        dec = np.floor(np.log10(np.maximum(np.asarray(nnz, np.float64), 1.0))).astype(int)
        out["selector_regret_by_nnz_decade"] = {int(d): round(float(r[dec == d].mean()), 4) for d in np.unique(dec)}
    return out


if __name__ == "__main__":
    sys.exit(main())
