"""Benchmark of the particle step at 2^24 particles on one MI355X.

Workload (one bench "step"): one filter time step of six constant-velocity
bootstrap particle filters, one per vSMC resampling scheme (multinomial,
residual, stratified, systematic, residual-stratified, residual-systematic),
run in turn on one HIP stream, each with N = 2^24 particles, FP32 state and
likelihood, FP64 log-sum-exp, exact FP64 weight fold, resampling when
ESS < N/2.  W warmup steps (the first is the init step) then K timed steps;
W + K = 200 = T by default.

value = 6 * N * K / (device time of the K steps), inputs resident in HBM.
The same workload is also timed for the fused FP64 pipeline and for the
un-fused FP64 pipeline (one kernel per vSMC operation), and, as an extra,
with the six filters on six concurrent streams.
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

import vsmc_amd as V  # noqa: E402

METRIC = "particle-updates/s at 2^24 particles, mixed vs FP64 plain; p50/p99 step latency"
HBM_PEAK_GBS = 8000.0
SCHEME_NAMES = ["multinomial", "residual", "stratified", "systematic", "residual_stratified",
                "residual_systematic"]


def move_bytes(prec, prev_resampled):
    """Algorithmic HBM bytes per particle of one move-kernel launch.
    mixed:  state 16 B read + 16 B write; after a resample 4 B ancestor read and
            the log-weight in its compact form (the 4-byte FP32 log2 term it is
            computed from), otherwise 8 B log-weight read and 8 B write;
    fp64:   states 32 B read + 32 B write, log-weight 8 B write, plus 4 B
            ancestor (after a resample) or 8 B log-weight read (otherwise)."""
    if prec == V.MIXED:
        return 32 + (4 + 4 if prev_resampled else 8 + 8)
    return 2 * 32 + 8 + (4 if prev_resampled else 8)


def step_bytes(dim, hist, first=0, compact=False):
    """Minimal algorithmic HBM bytes per particle of the steps first.. of one
    filter's history (FP32 state of dim values): the state write (4 dim) and,
    after step 0, its read (4 dim); the log-weight write (8, or 4 in compact
    form on step 0 and after a resample when `compact`: the CV model's FP32
    pipeline); after step 0 the ancestor index (4) when the previous step
    resampled, else the old log-weight (8); on a resampling step one weight
    read (8) and one ancestor write (4).  The extra passes of the exact fold
    are not counted."""
    res = hist["resampled"].astype(bool)
    t = np.arange(len(res))
    prev = np.concatenate([[False], res[:-1]])
    lw_write = np.where(compact & ((t == 0) | prev), 4, 8)
    b = 4 * dim + lw_write + np.where(t > 0, 4 * dim + np.where(prev, 4, 8), 0) + np.where(res, 12, 0)
    return float(b[first:].sum())


def c1_step_bytes(prec, hist, first):
    """Algorithmic bytes per particle-update of the fused / plain step over the
    steps first.. of a filter's history: 48 + 24 r (mixed), 80 + 24 r (fused
    FP64), 168 + 120 r (plain FP64), r = 1 on a resampling step."""
    base, per_res = {V.MIXED: (48, 24), V.FP64_FUSED: (80, 24), V.FP64_PLAIN: (168, 120)}[prec]
    res = hist["resampled"][first:].astype(np.float64)
    return float(np.sum(base + per_res * res))


def whole_step_roofline(total_bytes, seconds, scope):
    """HBM roofline of a whole step loop: minimal algorithmic bytes of every
    timed particle-update over the timed interval."""
    a = total_bytes / seconds / 1e9
    return {"bound": "hbm", "achieved": round(a, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
            "frac": round(a / HBM_PEAK_GBS, 4), "traffic": None, "algorithmic_bytes": total_bytes,
            "scope": scope}


def run_precision(ctx, prec, N, W, K, obs, schemes, profile=True, concurrent=True,
                  per_filter=True):
    """The six filters of config 1, each advanced one step per bench step.
    concurrent: every filter has its own context (its own HIP stream, default
    priority), so one filter's single-workgroup scans overlap another's
    full-grid passes; the total is host wall-clock between full
    synchronisations.  Otherwise the filters share ctx's stream in turn, with
    a HIP event before every filter's step (per_filter) or only before every
    bench step (each event on the stream costs a few microseconds of idle GPU,
    so the headline run records one per bench step)."""
    T = W + K
    ctxs = [V.Context(0) for _ in schemes] if concurrent else [ctx] * len(schemes)
    filters = [V.Filter(c, V.cv_desc(N, T, precision=prec, scheme=s, seed=2024 + s,
                                     resample_seed=4242 + s)) for c, s in zip(ctxs, schemes)]
    for f in filters:
        f.set_obs(obs)
    for _ in range(W):
        for f in filters:
            f.run(1)
    uniq = list({id(c): c for c in ctxs}.values())

    def sync_all():
        for c in uniq:
            V.lib().vsmcx_ctx_sync(c.h)

    sync_all()
    for f in filters:
        f.profile(profile)
    if concurrent:
        tms = [V.Timer(c, K + 1) for c in ctxs]
    else:
        tm = V.Timer(ctx, (K * len(filters) if per_filter else K) + 1)
    t0 = time.perf_counter()
    m = 0
    for k in range(K):
        for j, f in enumerate(filters):
            if concurrent:
                tms[j].record(k)
            elif per_filter or j == 0:
                tm.record(m)
                m += 1
            f.run(1)
    if concurrent:
        for t in tms:
            t.record(K)
    else:
        tm.record(m)
    sync_all()
    wall = time.perf_counter() - t0
    if concurrent:
        per = np.array([[tms[j].elapsed(k, k + 1) for j in range(len(filters))] for k in range(K)])
        total_ms = wall * 1e3
        for t in tms:
            t.close()
    elif per_filter:
        per = np.array([tm.elapsed(i, i + 1) for i in range(K * len(filters))]).reshape(
            K, len(filters))
        total_ms = tm.elapsed(0, K * len(filters))
        tm.close()
    else:
        per = None
        step_ms = np.array([tm.elapsed(k, k + 1) for k in range(K)])
        total_ms = tm.elapsed(0, K)
        tm.close()
    out = {"total_ms": total_ms, "wall_s": wall, "per_filter_step_ms": per,
           "concurrent": concurrent}
    if per is not None:
        step_ms = per.sum(axis=1) if not concurrent else per.max(axis=1)
    out["step_ms"] = step_ms
    moves = []
    hist = []
    for f in filters:
        mt = f.move_times() if profile else np.zeros(0)
        f.finish()
        h = f.history()
        moves.append(mt)
        hist.append(h)
        f.close()
    if concurrent:
        for c in uniq:
            c.close()
    out["move_ms"] = moves
    out["hist"] = hist
    return out


def roofline(prec, res, N, W, schemes):
    """Dominant kernel (the move): algorithmic bytes / measured duration, over
    the move launches that have the GPU to themselves.  A multinomial filter of
    the fused pipelines computes its step's sorted uniforms on a side stream
    while its move runs, so its move launches share the GPU and are left out."""
    tot_b = 0.0
    tot_s = 0.0
    shared = [s == 0 and prec != V.FP64_PLAIN for s in schemes]
    for mt, h, sh in zip(res["move_ms"], res["hist"], shared):
        if sh:
            continue
        for k, ms in enumerate(mt):
            t = W + k
            prev = bool(h["resampled"][t - 1]) if t > 0 else False
            tot_b += N * move_bytes(prec, prev)
            tot_s += ms * 1e-3
    if tot_s == 0:
        return None
    gbs = tot_b / tot_s / 1e9
    nl = sum(len(m) for m, sh in zip(res["move_ms"], shared) if not sh)
    return {"kernel": "k_cv_move<float>" if prec == V.MIXED else
            ("k_cv_move<double>" if prec == V.FP64_FUSED else "k_plain_move"),
            "bound": "hbm", "achieved": round(gbs, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
            "frac": round(gbs / HBM_PEAK_GBS, 4),
            "traffic": counter_traffic("k_cv_move<float>") if prec == V.MIXED else None,
            "traffic_source": ("FETCH_SIZE + WRITE_SIZE per launch from a separate rocprofv3 --pmc "
                               "run of this workload, scaled by a 1 GiB copy "
                               "(profiles/hbm_traffic_counters.json)") if prec == V.MIXED else None,
            "bytes_per_launch": round(tot_b / nl), "avg_launch_ms": round(tot_s * 1e3 / nl, 5),
            "launches": ("the move launches of the filters whose move runs alone" +
                         ("; the multinomial filter's move overlaps its side-stream uniforms"
                          if any(shared) else ""))}


def counter_traffic(kernel):
    """HBM bytes per launch of `kernel` from the committed FETCH_SIZE / WRITE_SIZE
    counter runs (profiles/hbm_traffic_counters.json, each counter scaled by a
    1 GiB device copy measured the same way), or None."""
    path = os.path.join(ROOT, "profiles", "hbm_traffic_counters.json")
    try:
        with open(path) as fh:
            d = json.load(fh)
    except (OSError, ValueError):
        return None
    for name, v in d.get("kernels", {}).items():
        if name.endswith(kernel):
            return v["traffic_bytes"]
    return None


def run_batched(nfilters, seed, tune=True):
    """Config-3 workload: nfilters independent CV filters, particle counts
    log-uniform in [2^8, 2^16], horizons uniform in [50, 500], each with its
    own observations; all of them advanced together (the tuner picks one
    persistent launch or one tick per step)."""
    rng = np.random.default_rng(seed)
    ns = np.exp(rng.uniform(np.log(2 ** 8), np.log(2 ** 16), nfilters)).astype(np.uint64)
    Ts = rng.integers(50, 501, nfilters).astype(np.uint64)
    obs = V.cv_simulate_many(Ts, seed + 1)
    ctx = V.Context(0)
    b = V.Batch(ctx, ns, Ts, 100000 + np.arange(nfilters, dtype=np.uint64),
                200000 + np.arange(nfilters, dtype=np.uint64), precision=V.MIXED,
                scheme=V.SYSTEMATIC)
    for f in range(nfilters):
        b.set_obs(f, obs[f])
    shape = b.tune(1) if tune else (0, 0)
    report = b.tune_report() if tune else []
    strategy = b.strategy()
    done = np.array([b.steps(f) for f in range(nfilters)], np.uint64)
    ms = b.run()
    updates = float(np.sum(ns * (Ts - done)))
    hists = [b.history(f) for f in range(nfilters)]
    rfrac = float(np.mean([h["resampled"].mean() for h in hists]))
    nbytes = sum(float(ns[f]) * step_bytes(4, hists[f], int(done[f]), compact=True)
                 for f in range(nfilters))
    b.close()
    ctx.close()
    return {"value": updates / (ms * 1e-3), "unit": "particle-updates/s", "ms": ms,
            "filters": nfilters, "particles": int(ns.sum()), "particle_updates": updates,
            "launch": {"strategy": "ticks" if strategy == V.BATCH_TICKS else "persistent",
                       "workgroups": shape[0], "order": "most work first" if shape[1] == 0
                       else "as given", "tuned": tune,
                       "candidates_ms": [{"strategy": "ticks" if c[0] == V.BATCH_TICKS
                                          else "persistent", "workgroups": c[1],
                                          "order": c[2], "ms": round(c[3], 3)}
                                         for c in report]},
            "resample_frac": rfrac,
            "roofline": whole_step_roofline(
                nbytes, ms * 1e-3, "all kernels of every timed step: minimal algorithmic bytes "
                "(state read/write, log-weight write (4 B compact after a resample, else 8 B), "
                "ancestor or log-weight read; on resampling steps one weight read and one "
                "ancestor write) over the timed run"),
            "config": "config-3: CV filters, N log-uniform [2^8, 2^16], T uniform [50, 500], "
                      "mixed precision, systematic, ESS<N/2; the tuner times one persistent "
                      "launch (several grids and orders) and one tick per step on the same first "
                      "step, restoring the state after each; the timed run covers every step"}


def run_serving(seconds, rate, seed, budget_gb=200.0, max_jobs=8192, pacing="closed",
                obs_interval_ms=8.0):
    """Config-4 workload: Poisson arrivals (rate per second) for `seconds`, a
    50/50 mix of the CV model and the 10-D linear-Gaussian model, N
    log-uniform in [2^10, 2^20], T uniform in [20, 1000].  Pacing "closed": a
    job's next observation arrives as soon as its previous step has completed
    (one step at a time).  Pacing "open": a job's k-th observation arrives at
    its submission time + k * obs_interval_ms whatever the device does, so the
    latency includes queueing.  Jobs are admitted under the HBM budget; the run
    ends when every submitted job has finished."""
    rng = np.random.default_rng(seed)
    t_arr = np.cumsum(rng.exponential(1.0 / rate, int(rate * seconds * 1.5) + 16))
    t_arr = t_arr[t_arr < seconds]
    k = t_arr.size
    models = rng.integers(0, 2, k)
    ns = np.exp(rng.uniform(np.log(2 ** 10), np.log(2 ** 20), k)).astype(np.uint64)
    Ts = rng.integers(20, 1001, k)
    A = V.lg_matrix(10, seed=seed)
    y_cv = V.cv_simulate_many(np.maximum(Ts, 1), seed + 1)
    y_lg = V.lg_simulate_many(A, np.maximum(Ts, 1), seed + 2)
    ctx = V.Context(0)
    s = V.Serve(ctx, budget_gb * 1e9, max_jobs, 1000, A, precision=V.MIXED)
    ids = {}
    fed = {}
    data = {}
    sub_t = {}
    fully_fed = set()
    finished = 0
    dt_obs = obs_interval_ms * 1e-3
    t0 = time.perf_counter()
    nxt = 0
    while True:
        now = time.perf_counter() - t0
        while nxt < k and t_arr[nxt] <= now:
            j = s.submit(int(models[nxt]), int(ns[nxt]), int(Ts[nxt]), 500000 + nxt, 600000 + nxt)
            ids[j] = nxt
            fed[j] = 0
            sub_t[j] = t_arr[nxt]
            data[j] = y_lg[nxt] if models[nxt] == V.MODEL_LG else y_cv[nxt]
            nxt += 1
        if pacing == "closed":
            ready = s.ready()
            if ready.size:
                ys = [data[int(j)][fed[int(j)]] for j in ready]
                s.observe_many(ready, ys)
                for j in ready:
                    j = int(j)
                    fed[j] += 1
                    if fed[j] == len(data[j]):
                        fully_fed.add(j)
        else:
            due, ys = [], []
            for j in list(fed):
                if j in fully_fed:
                    continue
                upto = min(len(data[j]), int((now - sub_t[j]) / dt_obs) + 1)
                while fed[j] < upto:
                    due.append(j)
                    ys.append(data[j][fed[j]])
                    fed[j] += 1
                if fed[j] == len(data[j]):
                    fully_fed.add(j)
            if due:
                s.observe_many(np.array(due, np.uint64), ys)
        launched = s.poll(wait=True)
        for j in [j for j in fully_fed if s.job(j)[0] == 2]:
            fully_fed.discard(j)
            del fed[j]
            finished += 1
        if nxt >= k and finished == k:
            break
        if launched == 0 and nxt < k:
            time.sleep(max(0.0, min(1e-3, t_arr[nxt] - (time.perf_counter() - t0))))
    wall = time.perf_counter() - t0
    st = s.stats()
    nbytes = 0.0
    for j, q in ids.items():
        dim = 10 if models[q] == V.MODEL_LG else 4
        nbytes += float(ns[q]) * step_bytes(dim, s.history(j, int(Ts[q])), compact=dim == 4)
    s.close()
    ctx.close()
    pace = ("closed loop (next observation when the previous step completes)"
            if pacing == "closed" else
            f"open loop (a job's observations every {obs_interval_ms:g} ms from its submission)")
    return {"value": st["particle_updates"] / (st["elapsed_ms"] * 1e-3), "unit": "particle-updates/s",
            "p50_step_ms": st["p50_us"] * 1e-3, "p99_step_ms": st["p99_us"] * 1e-3,
            "jobs": k, "steps": st["steps"], "particle_updates": st["particle_updates"],
            "device_busy_frac": st["busy_ms"] / st["elapsed_ms"] if st["elapsed_ms"] else None,
            "peak_particles": st["peak_particles"], "capacity_particles": st["capacity_particles"],
            "wall_s": wall, "pacing": pacing,
            "roofline": whole_step_roofline(
                nbytes, st["elapsed_ms"] * 1e-3, "all kernels of every tick: minimal algorithmic "
                "bytes of every job step (FP32 state of 4 or 10 values read/written, log-weight "
                "write (CV: 4 B compact after a resample), ancestor or log-weight read; on "
                "resampling steps one weight read and one ancestor write) over the serving "
                "interval"),
            "config": f"config-4: Poisson arrivals {rate}/s for {seconds}s, 50/50 CV and 10-D "
                      f"linear-Gaussian (A spectral radius 0.9, q 0.1, r 0.2, s0 0.25), N "
                      f"log-uniform [2^10, 2^20], T uniform [20, 1000], {budget_gb:g} GB budget, "
                      f"mixed precision, systematic; pacing: {pace}"}


GMM_FLOPS_PER_POINT = 6  # per component: x - mu, * A, FMA (2), exp2, accumulate


def run_gmm(N, sweeps, seed, precision=V.MIXED, K=4, ndata=1000):
    """Config-2 workload: K = 4 mixture, ndata points from means (-3, 0, 2, 5)
    with component sds drawn in [0.5, 1.5], N particles x (4 mu, 4 lambda,
    4 weights), adaptive tempering to CESS = 0.9, `sweeps` sweeps of the three
    RW-MH moves per iteration, stratified resampling at ESS < N/2; run from
    the prior to alpha = 1."""
    y = V.gmm_simulate(ndata, seed)
    ctx = V.Context(0)
    g = V.Gmm(ctx, y, N, components=K, sweeps=sweeps, precision=precision, seed=seed + 1,
              resample_seed=seed + 2)
    t0 = time.perf_counter()
    g.initialize()
    iters, alpha = g.iterate()
    wall = time.perf_counter() - t0
    h = g.history()
    t_temper, t_res, t_mh = g.times()
    g.close()
    ctx.close()
    evals = float(N) * 3 * sweeps * iters
    flops = evals * ndata * (GMM_FLOPS_PER_POINT * K + 2)
    out = {"iterations": iters, "alpha": alpha, "wall_s": wall,
           "ms_per_iteration": wall * 1e3 / max(iters, 1),
           "particle_iterations_per_s": N * iters / wall,
           "mh_evaluations_per_s": evals / wall,
           "device_ms": {"temper_reweight": t_temper, "resample": t_res, "mh_sweeps": t_mh},
           "log_z": float(h["log_z"][-1]), "path_log_z": float(h["path_log_z"][-1]),
           "resample_frac": float(h["resampled"][1:].mean()) if iters else 0.0,
           "accept_rate": (h["accept"][1:].sum(0) / (float(N) * sweeps * max(iters, 1))).tolist()}
    if t_mh > 0:
        achieved = flops / (t_mh * 1e-3) / 1e12
        out["roofline"] = {"bound": "valu", "achieved": achieved, "peak": 157.3,
                           "unit": "TFLOP/s", "frac": achieved / 157.3, "traffic": None,
                           "kernel": "k_gmm_mh",
                           "flops_per_point": GMM_FLOPS_PER_POINT * K + 2}
    return out


def host_cpu():
    """(threads available to this process -- what nproc reports --, CPU model)."""
    try:
        threads = len(os.sched_getaffinity(0))
    except AttributeError:
        threads = os.cpu_count() or 1
    model = "unknown"
    try:
        with open("/proc/cpuinfo") as fh:
            for line in fh:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    return threads, model


def cpu_gmm_baseline(N, sweeps, seed, threads, iters, K=4, ndata=1000):
    """The GMM oracle's MH-evaluation rate over the first `iters` iterations
    at N particles (the rate does not depend on the iteration count)."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    from oracle_ctypes import Oracle
    o = Oracle()
    y = V.gmm_simulate(ndata, seed)
    t0 = time.perf_counter()
    _, _, _, rec = o.gmm_run(N, y, K=K, sweeps=sweeps, seed=seed + 1, resample_seed=seed + 2,
                             threads=threads, want_state=False, max_steps=iters)
    dt = time.perf_counter() - t0
    done = len(rec) - 1
    return {"value": float(N) * 3 * sweeps * done / dt, "unit": "mh-evaluations/s",
            "cores": threads, "kind": "port", "cpu": host_cpu()[1],
            "sample": f"GMM oracle (oracle/gmm_oracle.c, FP64, OpenMP over particles), N={N}, "
                      f"{ndata} points, {sweeps} sweeps, the first {done} iterations from the "
                      f"prior (initialisation included)",
            "seconds": round(dt, 3)}


def cpu_baseline(N, steps, threads):
    """vSMC's CPU step on the host cores, once on each of its SMP backends:
    the move (and the gather) over all threads, normalisation and resampling
    sequential as in vSMC.  value: the faster backend."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    from oracle_ctypes import Oracle
    o = Oracle()
    obs = V.cv_simulate(steps, 7)
    out = {}
    for backend in ("omp", "threads"):
        t0 = time.perf_counter()
        rec, phase = o.cv_run_timed(N, obs, 42, 4242, V.SYSTEMATIC, 0.5, threads=threads,
                                    backend=backend)
        dt = time.perf_counter() - t0
        out[backend] = {"value": N * steps / dt, "seconds": round(dt, 3),
                        "phase_s": {"move": round(phase[0], 3), "normalise": round(phase[1], 3),
                                    "resample": round(phase[2], 3), "copy": round(phase[3], 3)}}
    best = max(out, key=lambda b: out[b]["value"])
    return {"value": out[best]["value"], "unit": "particle-updates/s", "cores": threads,
            "kind": "port", "cpu": host_cpu()[1], "backend": best,
            "sample": f"vSMC-semantics CPU path (oracle/vsmc_oracle.c: move and gather over "
                      f"{threads} threads, sequential normalise/inversion/cfrp_trans), N={N}, "
                      f"{steps} steps, systematic, on vSMC's OpenMP and std::thread backends",
            "backends": out}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=190)
    ap.add_argument("--warmup", type=int, default=10)
    ap.add_argument("--n", type=int, default=1 << 24)
    ap.add_argument("--schemes", default="all")
    ap.add_argument("--no-fp64", action="store_true")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--cpu-steps", type=int, default=12)
    ap.add_argument("--cpu-n", type=int, default=1 << 24)
    ap.add_argument("--no-batch", action="store_true")
    ap.add_argument("--no-serve", action="store_true")
    ap.add_argument("--serve-seconds", type=float, default=0.5)
    ap.add_argument("--serve-rate", type=float, default=2000.0)
    ap.add_argument("--serve-open-ms", type=float, default=8.0,
                    help="also serve with open-loop pacing, one observation per job every this "
                         "many ms (0: skip)")
    ap.add_argument("--batch-filters", type=int, default=4096)
    ap.add_argument("--no-gmm", action="store_true")
    ap.add_argument("--serial-filters", action="store_true",
                    help="config 1: only the six filters in turn on one stream (the headline "
                         "workload; skips the extra run on six concurrent streams)")
    ap.add_argument("--gmm-n", type=int, default=1 << 20)
    ap.add_argument("--gmm-sweeps", type=int, default=5)
    ap.add_argument("--gmm-fp64", action="store_true",
                    help="also run the GMM workload in FP64 at the same size")
    ap.add_argument("--gmm-cpu-n", type=int, default=1 << 16)
    ap.add_argument("--gmm-cpu-iters", type=int, default=3)
    ap.add_argument("--no-probe", action="store_true",
                    help="skip the streaming-rate probe (copy, read, write of 1 GiB)")
    ap.add_argument("--copy-probe", action="store_true",
                    help="only run a device-to-device copy of 1 GiB five times (calibrates the "
                         "FETCH_SIZE / WRITE_SIZE counters against known bytes) and exit")
    a = ap.parse_args()
    if a.copy_probe:
        import torch
        src = torch.empty(1 << 27, dtype=torch.float64, device="cuda:0").fill_(1.0)
        dst = torch.empty_like(src)
        for _ in range(5):
            dst.copy_(src)
        torch.cuda.synchronize()
        print(json.dumps({"copy_probe_bytes_per_launch": 2 * src.numel() * 8,
                          "read_bytes": src.numel() * 8, "write_bytes": src.numel() * 8}))
        return
    if a.gpus != 1:
        print("bench.py measures one GPU: this build has no multi-GPU path "
              "(it would shard particles with one all-reduce of three FP64 scalars per step); "
              f"--gpus {a.gpus} is not supported", file=sys.stderr)
        sys.exit(2)
    W, K, N = max(1, a.warmup), a.steps, a.n
    schemes = list(range(6)) if a.schemes == "all" else [SCHEME_NAMES.index(s) for s in
                                                         a.schemes.split(",")]
    ks = min(K, 8)  # steps of the short serial run that gives the move kernel's roofline
    obs = V.cv_simulate(max(W + K, 2 + ks), 11)
    ctx = V.Context(0)
    # the streaming rates this GPU reaches, next to the published peak
    probe = None if a.no_probe else ctx.stream_probe(1 << 30, 5)
    res = {}
    precs = [V.MIXED] + ([] if a.no_fp64 else [V.FP64_FUSED, V.FP64_PLAIN])
    serial = {}
    for prec in precs:
        print(f"[bench] precision {prec}: {len(schemes)} filters x N={N}, W={W} K={K}",
              file=sys.stderr, flush=True)
        # the six filters in turn on one stream: C1's workload, one HIP event
        # per bench step
        res[prec] = run_precision(ctx, prec, N, W, K, obs[:W + K], schemes, profile=False,
                                  concurrent=False, per_filter=False)
        # the move kernel's roofline from a short run with every move launch
        # bracketed by HIP events
        serial[prec] = run_precision(ctx, prec, N, 2, ks, obs[:2 + ks], schemes, profile=True,
                                     concurrent=False)
    # each scheme's step time: a shorter in-turn run with an event before every
    # filter's step
    kp = min(K, 24)
    per_run = run_precision(ctx, V.MIXED, N, W, kp, obs[:W + kp], schemes, profile=False,
                            concurrent=False, per_filter=True)
    conc = None
    if not a.serial_filters:
        print(f"[bench] concurrent streams: {len(schemes)} filters x N={N}", file=sys.stderr,
              flush=True)
        conc = run_precision(ctx, V.MIXED, N, W, K, obs[:W + K], schemes, profile=False,
                             concurrent=True)
    ctx.close()

    nf = len(schemes)
    mixed = res[V.MIXED]
    value = nf * N * K / (mixed["total_ms"] * 1e-3)
    # a bench step: every filter one step, in turn
    steps_ms = mixed["step_ms"]

    def step_roofline(prec, r):
        b = sum(N * c1_step_bytes(prec, h, W) for h in r["hist"])
        return whole_step_roofline(
            b, r["total_ms"] * 1e-3, "every kernel of the K timed steps of the six filters: "
            "algorithmic bytes per particle-update " +
            {V.MIXED: "48 + 24 r", V.FP64_FUSED: "80 + 24 r",
             V.FP64_PLAIN: "168 + 120 r"}[prec] + " (r = 1 on a resampling step)")

    out = {
        "metric": METRIC, "value": value, "unit": "particle-updates/s", "n_gpus": 1,
        "steps": K, "warmup": W, "ms_per_step": mixed["total_ms"] / K,
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "fp32+fp64",
        "data": "synthetic (observations simulated from the CV model, seed 11)",
        "config": {"workload": "config-1: CV bootstrap filter, N=2^24 per filter, six filters "
                               "(one per resampling scheme) advanced one step each per bench "
                               "step, in turn on one HIP stream, ESS<N/2 resampling, T=W+K",
                   "n": N, "filters_per_step": nf,
                   "schemes": [SCHEME_NAMES[s] for s in schemes], "horizon": W + K},
        "p50_step_ms": float(np.percentile(steps_ms, 50)),
        "p99_step_ms": float(np.percentile(steps_ms, 99)),
        "roofline": roofline(V.MIXED, serial[V.MIXED], N, 2, schemes),
        "whole_step_roofline": step_roofline(V.MIXED, mixed),
    }
    if probe is not None:
        probe.update({"bytes": 1 << 30, "kernels": "vsmcx_stream_probe: grid-stride 16-B "
                      "loads/stores over 1 GiB, best of 5 launches at 1024 and 4096 workgroups",
                      "move_achieved_vs_copy": out["roofline"]["achieved"] / probe["copy_GBps"]})
        out["hbm_probe"] = probe
    per_scheme = {}  # each scheme's step within an in-turn run (events per filter step)
    for j, s in enumerate(schemes):
        ms = per_run["per_filter_step_ms"][:, j]
        h = mixed["hist"][j]
        per_scheme[SCHEME_NAMES[s]] = {
            "updates_per_s": N * len(ms) / (ms.sum() * 1e-3),
            "p50_ms": float(np.percentile(ms, 50)), "p99_ms": float(np.percentile(ms, 99)),
            "resample_frac": float(h["resampled"][W:].mean()), "log_z": float(h["log_z"][-1])}
    out["per_scheme"] = per_scheme
    if conc is not None:
        cms = conc["per_filter_step_ms"].max(axis=1)
        out["concurrent_streams"] = {
            "value": nf * N * K / (conc["total_ms"] * 1e-3), "ms_per_step": conc["total_ms"] / K,
            "p50_step_ms": float(np.percentile(cms, 50)),
            "p99_step_ms": float(np.percentile(cms, 99)),
            "note": "the same six filters, each on its own HIP stream (host wall-clock between "
                    "full synchronisations); not C1's in-turn workload"}
    for prec, name in ((V.FP64_FUSED, "fp64_fused"), (V.FP64_PLAIN, "fp64_plain")):
        if prec in res:
            r = res[prec]
            v = nf * N * K / (r["total_ms"] * 1e-3)
            out[name] = {"value": v, "ms_per_step": r["total_ms"] / K,
                         "roofline": roofline(prec, serial[prec], N, 2, schemes),
                         "whole_step_roofline": step_roofline(prec, r),
                         "log_z": {SCHEME_NAMES[s]: float(r["hist"][j]["log_z"][-1])
                                   for j, s in enumerate(schemes)}}
    if V.FP64_PLAIN in res:
        out["speedup_mixed_vs_fp64_plain"] = value / out["fp64_plain"]["value"]
    if V.FP64_FUSED in res:
        out["speedup_mixed_vs_fp64_fused"] = value / out["fp64_fused"]["value"]
    if not a.no_batch:
        print(f"[bench] batched: {a.batch_filters} filters", file=sys.stderr, flush=True)
        out["batched"] = run_batched(a.batch_filters, 13)
    if not a.no_serve:
        print(f"[bench] serving: {a.serve_rate}/s for {a.serve_seconds}s", file=sys.stderr,
              flush=True)
        out["serving"] = run_serving(a.serve_seconds, a.serve_rate, 17)
        if a.serve_open_ms > 0:
            print(f"[bench] serving open loop: observations every {a.serve_open_ms} ms",
                  file=sys.stderr, flush=True)
            out["serving"]["open_loop"] = run_serving(a.serve_seconds, a.serve_rate, 17,
                                                      pacing="open",
                                                      obs_interval_ms=a.serve_open_ms)
    if not a.no_gmm:
        print(f"[bench] gmm: N={a.gmm_n}, {a.gmm_sweeps} sweeps", file=sys.stderr, flush=True)
        gm = run_gmm(a.gmm_n, a.gmm_sweeps, 23)
        if a.gmm_fp64:
            gm["fp64"] = run_gmm(a.gmm_n, a.gmm_sweeps, 23, precision=V.FP64_FUSED)
        if not a.no_cpu:
            gm["cpu_baseline"] = cpu_gmm_baseline(a.gmm_cpu_n, a.gmm_sweeps, 23, host_cpu()[0],
                                                  a.gmm_cpu_iters)
        gm["config"] = ("config-2: K=4 GMM, 1000 points from means (-3, 0, 2, 5), sds U[0.5, 1.5], "
                        f"N={a.gmm_n} x 12 FP32 parameters, CESS 0.9 tempering, "
                        f"{a.gmm_sweeps} RW-MH sweeps (mu, lambda, weight) per iteration, "
                        "stratified resampling at ESS < N/2, mixed precision")
        out["gmm"] = gm
    if not a.no_cpu:
        threads, model = host_cpu()
        print(f"[bench] cpu baseline: N={a.cpu_n}, {a.cpu_steps} steps, {threads} threads "
              f"({model})", file=sys.stderr, flush=True)
        out["cpu_baseline"] = cpu_baseline(a.cpu_n, a.cpu_steps, threads)
    print(json.dumps(out))


if __name__ == "__main__":
    main()
