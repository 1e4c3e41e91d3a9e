"""Benchmark: MGPCG solves/s and latency on one MI355X.

Workload (default): the adaptive 3-level octree, 256^3 effective -- leaf tiles
at h=1/256 where the tile box meets |d|<0.05 of a sphere of radius 0.15 at the
domain centre, h=1/128 where it meets |d|<0.15, h=1/64 elsewhere, 2:1 face
balanced; the sphere is solid (NEUMANN), the top face Dirichlet, the other walls
Neumann; the RHS is the divergence of a seeded N(0,1) face-velocity field; FP32;
red-black Gauss-Seidel, 2 pre + 2 post sweeps, 10 bottom iterations; solved to
1e-6 relative residual from x=0.  A step is one complete solve of that problem
(inputs resident in HBM).

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--no-cpu-baseline]

Prints one JSON line.  The CPU baseline is the oracle (the OpenMP CPU
transcription) timed on this host's cores on a bounded sample of the same solve.

HBM traffic of the dominant kernel comes from two hardware-counter runs of the
kernels-only mode (one counter block each), scaled by what the same counter
reads for a plain copy of known size:

    rocprofv3 --pmc FETCH_SIZE -d OUT/fetch -o pmc --output-format csv -- python3 bench.py --kernels-only
    rocprofv3 --pmc WRITE_SIZE -d OUT/write -o pmc --output-format csv -- python3 bench.py --kernels-only
    python bench.py --summarize-pmc OUT/fetch/pmc_counter_collection.csv OUT/write/pmc_counter_collection.csv

The last command writes profiles/k_gs_color_adaptive256_hbm_traffic.txt, whose
traffic figure the benchmark line reports.
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

from cirrus_amd.workloads import ADAPTIVE256, PLAIN64, VORTEX256, VORTEX512, batch_scene_params, build_scenes  # noqa: E402

WORKLOADS = {
    # name: (scene kwargs, solver kwargs, description)
    "adaptive256": (ADAPTIVE256, dict(level_iters=2, coarsest_iters=10, smoother=0),
                    "adaptive octree 256^3 effective (levels 3-5, sphere band r=0.15), FP32, RB-GS 2+2, tol 1e-6"),
    "plain64": (PLAIN64, dict(level_iters=2, coarsest_iters=10, smoother=1),
                "uniform 64^3, solid sphere r=0.2, FP32, damped Jacobi 2+2 (w=2/3), tol 1e-6"),
}
PEAK_HBM_GBS = 8000.0


def plain_drop_in(ca, items, solver_kw, reps=1):
    """The plain FP32 drop-in on the given problems: one handle per problem,
    solved one after another from x = 0 (as GMGSolver::solve starts,
    src/GMGSolver.cu:537) to 1e-6 of the recursive residual, after one warm-up
    pass.  Returns solves/s and the latency spread."""
    S = ca.GMGSolver(1.0, 1.0)
    probs = [ca.Problem(t) for t, _ in items]
    lat = []
    for rep_i in range(reps + 1):
        for P, (_, b) in zip(probs, items):
            P.set(ca.CH_B, b)
            t = time.perf_counter()
            S.solve(P, False, 1000, 1e-6, solver_kw["level_iters"], solver_kw["coarsest_iters"], 1, False,
                    smoother=solver_kw["smoother"])
            dt = time.perf_counter() - t
            assert S.last["converged"] == 1
            if rep_i > 0:
                lat.append(dt)
    for P in probs:
        P.close()
    lat_ms = np.array(lat) * 1e3
    return {"problems": len(items), "solves_per_s": len(lat) / float(np.sum(lat)),
            "p50_ms": float(np.percentile(lat_ms, 50)), "p99_ms": float(np.percentile(lat_ms, 99))}


def run_served(ca, solver_kw, scenes, duration=60.0, lam=2000.0, max_resident=256, budget=200 << 30, rho=0.95,
               steps=50, storage=1, refine=True, seed=7, drain_s=20.0):
    """Serving form: projection requests arrive as a Poisson process of rate
    `lam` for `duration` seconds.  They come from simulated instances, each a
    sequence of `steps` correlated solves (a simulation issues step k+1 only
    after step k's pressure is back): an arriving request is taken by an idle
    instance that has steps left, or else starts a new instance, whose octree is
    drawn at random from `scenes` (lognormal leaf-tile counts) and admitted into
    the server under the HBM budget (at most `max_resident` instances); an
    instance that has run its `steps` solves is released, freeing its memory.
    Requests that find neither wait in arrival order.  An instance's first solve
    is cold (its scene's right-hand side from the host); each later one updates
    the right-hand side on the device as b = rho b + sqrt(1 - rho^2) rms(b0) u
    (u seeded unit-variance noise) and starts from the instance's previous
    pressure.  Latency is arrival -> result.  storage: V-cycle storage (0 FP32,
    1 BF16); refine: every request is an FP64 iterative refinement reported
    once |b - A x| / |b| < 1e-6 in FP64; otherwise the FP32 solve to 1e-6 of
    its recursive residual, each result's FP64 residual reported beside it.
    Requests not finished `drain_s` seconds after the window closes are
    counted as unfinished."""
    import collections
    rng = np.random.default_rng(seed)
    srv = ca.Server(budget, max_resident, level_iters=solver_kw["level_iters"], coarsest_iters=solver_kw["coarsest_iters"],
                    smoother=solver_kw["smoother"], storage=int(storage), refine=bool(refine),
                    verify_f64=0 if refine else 1)
    pool = [(sc.tree, np.ascontiguousarray(sc.rhs, np.float32)) for sc in scenes]
    sig0 = [float(np.sqrt(1 - rho * rho) * np.sqrt(np.mean(b0[b0 != 0] ** 2))) for _, b0 in pool]
    arrivals = np.cumsum(rng.exponential(1.0 / lam, int(lam * duration * 1.2) + 16))
    arrivals = arrivals[arrivals < duration]
    waiting = collections.deque()
    idle = collections.deque()       # instances with steps left and no request running
    inst_scene, inst_done = {}, {}   # instance -> scene index, solves finished
    running = {}                     # request tag -> instance
    lat, arr_t, iters, act, f64, admit_s, sizes = [], [], [], [], [], [], []
    n_admit = n_release = n_busy = 0
    nxt = 0
    t0 = time.perf_counter()
    deadline = duration + drain_s
    t_last = 0.0
    while True:
        now = time.perf_counter() - t0
        while nxt < len(arrivals) and arrivals[nxt] <= now:
            waiting.append(nxt)
            nxt += 1
        while waiting:
            j = waiting[0]
            if idle:
                inst = idle.popleft()
                q = inst_scene[inst]
                srv.perturb_rhs(inst, rho, sig0[q], 1000003 * inst + 7919 * inst_done[inst] + j)
                srv.submit(inst, None, warm=True, tag=j, download=False)
            elif len(inst_scene) < max_resident:
                q = int(rng.integers(len(pool)))
                t_a = time.perf_counter()
                try:
                    inst = srv.admit(pool[q][0])
                except ca.CirrusBusy:
                    n_busy += 1
                    break
                admit_s.append(time.perf_counter() - t_a)
                n_admit += 1
                sizes.append(int((pool[q][0].type == 1).sum()))
                inst_scene[inst], inst_done[inst] = q, 0
                srv.submit(inst, pool[q][1], warm=False, tag=j, download=False)
            else:
                break
            waiting.popleft()
            running[j] = inst
        act.append(srv.step())
        t_done = time.perf_counter() - t0
        while True:
            r = srv.next_result(with_x=False)
            if r is None:
                break
            j, inst, res, _ = r
            running.pop(j, None)
            lat.append(t_done - arrivals[j])
            arr_t.append(arrivals[j])
            iters.append(res["iters"])
            f64.append(res["rel_res_f64"])
            t_last = t_done
            inst_done[inst] += 1
            if inst_done[inst] >= steps:
                srv.release(inst)
                n_release += 1
                del inst_scene[inst], inst_done[inst]
            else:
                idle.append(inst)
        if nxt >= len(arrivals) and not running and not waiting:
            break
        if now > deadline:
            break
    info = srv.info()
    srv.close()
    lat_ms = np.array(lat) * 1e3
    # the same percentiles over the requests that arrived after the first
    # quarter of the window (the start admits every instance cold, one after another)
    late = np.array(arr_t) >= duration / 4
    prec = ("BF16 V-cycle storage" if storage else "FP32") + \
        (", FP64 iterative refinement (FP64 pressure kept per instance, every result checked in FP64)" if refine
         else ", every result's FP64 residual checked")
    done_s = max(t_last, 1e-9)
    return {
        "workload": f"Poisson requests at {lam:g}/s for {duration:g} s from simulated instances of {steps} "
                    f"warm-started correlated steps; instance octrees drawn per instance from {len(pool)} lognormal "
                    f"scenes, admitted and released during the stream under a {budget / 2**30:g} GB HBM budget "
                    f"(at most {max_resident} resident); {prec}, RB-GS 2+2, tol 1e-6 of |b|",
        "offered_per_s": lam, "window_s": duration, "requests": int(len(arrivals)),
        "completed": len(lat), "unfinished": int(len(arrivals) - len(lat)),
        "solves_per_s": len(lat) / done_s, "wall_s": done_s,
        "p50_ms": float(np.percentile(lat_ms, 50)) if len(lat) else None,
        "p99_ms": float(np.percentile(lat_ms, 99)) if len(lat) else None,
        "p50_ms_after_first_quarter": float(np.percentile(lat_ms[late], 50)) if late.any() else None,
        "p99_ms_after_first_quarter": float(np.percentile(lat_ms[late], 99)) if late.any() else None,
        "iterations_median": float(np.median(iters)) if iters else None,
        "active_mean": float(np.mean(act)) if act else 0.0,
        "instances_admitted": n_admit, "instances_released": n_release, "admissions_refused_busy": n_busy,
        "instance_leaf_tiles_median": int(np.median(sizes)) if sizes else None,
        "instance_leaf_tiles_max": int(np.max(sizes)) if sizes else None,
        "distinct_scenes": len(pool), "resident_at_end": info["instances"],
        "fp64_checked": len(f64), "fp64_rel_res_max": float(np.max(f64)) if f64 else None,
        "fp64_over_tol": int(np.sum(np.array(f64) >= 1e-6)) if f64 else None,
        "admit_ms_median": float(np.median(admit_s)) * 1e3 if admit_s else None,
        "admit_ms_max": float(np.max(admit_s)) * 1e3 if admit_s else None,
    }


def served_light_sweep(ca, solver_kw, scenes, args, budget, capacity):
    """The served_mixed form below capacity: one request window per offered
    rate of --serve-rate-light (fractions of `capacity`, served_mixed's
    completion rate, when below 1.5).  Returns the record of the highest rate
    whose p99 after the first quarter of its window stays within
    --light-p99-ms (else the lowest rate), with every point under "sweep"."""
    rates = []
    for tok in str(args.serve_rate_light).split(","):
        v = float(tok)
        if v <= 0:
            continue
        if v < 1.5:
            if not capacity:
                continue
            v *= capacity
        rates.append(v)
    if not rates or args.serve_seconds_light <= 0:
        return None
    points = []
    for lam in sorted(rates):
        f = run_served(ca, solver_kw, scenes, duration=args.serve_seconds_light, lam=lam,
                       max_resident=args.serve_instances, budget=budget, storage=1, refine=True)
        points.append(f)
        note(f"served_mixed_light: {lam:.0f}/s offered, p50 {f['p50_ms_after_first_quarter']:.1f} ms, "
             f"p99 {f['p99_ms_after_first_quarter']:.1f} ms after the first quarter")
    ok = [f for f in points if f["p99_ms_after_first_quarter"] is not None
          and f["p99_ms_after_first_quarter"] <= args.light_p99_ms]
    best = dict(ok[-1] if ok else points[0])
    best["p99_bound_ms"] = args.light_p99_ms
    best["p99_bound_met"] = bool(ok)
    best["sweep"] = [{k: f[k] for k in ("offered_per_s", "solves_per_s", "completed", "unfinished", "p50_ms",
                                        "p99_ms", "p50_ms_after_first_quarter", "p99_ms_after_first_quarter",
                                        "fp64_over_tol")} for f in points]
    return best


def run_batched(ca, scenes, B, reps, solver_kw, use_graph=None, storage=0, refine=False, tune=False, t_build=0.0,
                oracle_sample=False):
    """Solve nprob independent problems together; returns the form's record.
    storage: V-cycle storage (0 FP32, 1 BF16).  refine: FP64 iterative
    refinement to 1e-6 of the FP64 residual (inner MGPCG to 5e-7); otherwise
    the MGPCG to 1e-6 of its recursive residual, each pressure then checked in
    FP64.  tune: the autotuner picks the schedule and the smoother settings
    (bfloat16 storage allowed) on the batch's own right-hand sides first.
    scenes: the built scenes; B: their batched handle.  oracle_sample: keep
    the FP64 pressures of six problems for the FP64-oracle comparison."""
    nprob = len(scenes)
    b_all = np.concatenate([np.ascontiguousarray(sc.rhs, np.float32) for sc in scenes])
    S = ca.GMGSolver(1.0, 1.0)
    kw = dict(level_iters=solver_kw["level_iters"], smoother=solver_kw["smoother"], storage=storage)
    tune_rep = None
    if tune:
        B.set(ca.CH_B, b_all)
        # candidates timed as the refined solve this form runs
        kw, tune_rep = S.autotune_refined(B, 1e-6, 5e-7, 20, 1000, solver_kw["level_iters"], solver_kw["coarsest_iters"],
                                          smoother=solver_kw["smoother"], storage=storage, allow_16bit=True)

    def once():
        B.set(ca.CH_B, b_all)
        t = time.perf_counter()
        if tune:
            res = S.solve_refined(B, 1e-6, 5e-7, 20, 1000, kw["level_iters"], solver_kw["coarsest_iters"],
                                  smoother=kw["smoother"], use_graph=use_graph, storage=kw["storage"])
        elif refine:
            res = S.solve_refined(B, 1e-6, 5e-7, 20, 1000, solver_kw["level_iters"], solver_kw["coarsest_iters"],
                                  smoother=solver_kw["smoother"], use_graph=use_graph, storage=storage)
        else:
            res = S.solve_batch(B, 1000, 1e-6, solver_kw["level_iters"], solver_kw["coarsest_iters"],
                                smoother=solver_kw["smoother"], use_graph=use_graph, storage=storage, verify_f64=True)
        return time.perf_counter() - t, res

    once()  # warm-up
    times, res = [], None
    for _ in range(reps):
        dt, res = once()
        times.append(dt)
        assert all(r["converged"] == 1 for r in res)
    leaf = [int((sc.tree.type == 1).sum()) for sc in scenes]
    iters = [r["iters"] for r in res]
    check = None
    if oracle_sample and (refine or tune):
        # FP64 pressures of a sample of the problems (largest, smallest and
        # quantiles between) for the comparison with the FP64 oracle
        order = np.argsort(leaf, kind="stable")
        picks = sorted({int(order[int(round(f * (nprob - 1)))]) for f in (1.0, 0.9, 0.75, 0.5, 0.25, 0.0)})
        x64 = B.split(B.get_f64())
        bq = B.split(b_all)
        check = [(q, batch_scene_params(q)[0], scenes[q].tree.level, scenes[q].tree.coord, scenes[q].tree.type,
                  bq[q].copy(), x64[q].copy()) for q in picks]
    f64 = [r["rel_res"] if (refine or tune) else r["rel_res_f64"] for r in res]
    total = float(np.sum(times))
    # the V-cycle's dominant kernel on the pool's most populated level
    lv = int(np.bincount(np.concatenate([sc.tree.level[sc.tree.type != 4] for sc in scenes])).argmax())
    kid = {0: ca.KERNEL_GS_SWEEP, 1: ca.KERNEL_GS_SWEEP_BF16, 2: ca.KERNEL_GS_SWEEP_F16}[kw["storage"]]
    gus, gtiles = B.benchmark_kernel(kid, lv, reps=20)
    gbytes = gs_alg_bytes(gtiles, 2 if kw["storage"] else 4)
    prec = {0: "FP32", 1: "BF16 V-cycle storage", 2: "F16 V-cycle storage (power-of-two rhs scale)"}[kw["storage"]]
    how = "FP64 iterative refinement to 1e-6 of the FP64 residual" if (refine or tune) else \
        "tol 1e-6 each (recursive residual), every pressure checked in FP64"
    if tune:
        sm = "RB-GS" if kw["smoother"] == 0 else "damped Jacobi"
        how = (f"schedule and smoother chosen by the autotuner ({sm} {kw['level_iters']}+{kw['level_iters']}, "
               f"coarse shift {tune_rep['coarse_shift']}, XCD order {tune_rep['swizzle']}, "
               f"tile block {tune_rep['tile_block']}), " + how)
    return {
        "workload": f"{nprob} independent adaptive problems (sphere band, lognormal leaf-tile counts), "
                    f"{prec}, {'' if tune else 'RB-GS 2+2, '}{how}, solved together",
        "problems": nprob, "leaf_tiles_total": int(np.sum(leaf)), "leaf_tiles_median": int(np.median(leaf)),
        "leaf_tiles_min": int(np.min(leaf)), "leaf_tiles_max": int(np.max(leaf)),
        "solves_per_s": nprob * reps / total, "ms_per_batch": total / reps * 1e3,
        "iterations_min": int(np.min(iters)), "iterations_median": float(np.median(iters)),
        "iterations_max": int(np.max(iters)), "build_s": t_build, "reps": reps,
        "fp64_rel_res_max": float(np.max(f64)),
        "fp64_over_tol": int(np.sum(np.array(f64) >= 1e-6)),
        "outer_max": int(max(r["outer"] for r in res)) if (refine or tune) else None,
        "autotune": tune_rep,
        "gs_sweep": gs_record(lv, gtiles, gus, gbytes, "k_gs_color<float>" if kw["storage"] == 0 else
                              ("k_gs_color<bf16>" if kw["storage"] == 1 else "k_gs_color<f16>")),
        "_check": check,
    }


def run_plain64(ca, steps=50, warmup=5):
    """The uniform 64^3 configuration through the plain drop-in (solid sphere
    r=0.2, damped Jacobi 2+2 with w=2/3, 10 bottom iterations, FP32, tol
    1e-6, from x = 0), timed like the headline solve."""
    scene_kw, skw, desc = WORKLOADS["plain64"]
    sc = ca.Scene(**scene_kw)
    P = ca.Problem(sc.tree)
    b = np.ascontiguousarray(sc.rhs, np.float32)
    S = ca.GMGSolver(1.0, 1.0)
    lat = []
    for i in range(warmup + steps):
        P.set(ca.CH_B, b)
        t = time.perf_counter()
        S.solve(P, False, 1000, 1e-6, skw["level_iters"], skw["coarsest_iters"], 1, False, smoother=skw["smoother"],
                verify_f64=i == warmup + steps - 1)
        if i >= warmup:
            lat.append(time.perf_counter() - t)
        assert S.last["converged"] == 1
    lat_ms = np.array(lat) * 1e3
    out = {"workload": desc, "tiles": P.n, "leaf_tiles": P.n_leaf, "solves_per_s": steps / float(np.sum(lat)),
           "p50_ms": float(np.percentile(lat_ms, 50)), "p99_ms": float(np.percentile(lat_ms, 99)),
           "iterations": S.last["iters"], "fp64_rel_res": S.last["rel_res_f64"]}
    P.close()
    return out


def gs_record(level, tiles, us, alg_bytes, kname):
    """Roofline record of a colour sweep at batch scale: the hipEvent timing of
    this run and, where the committed kernel-trace summary covers the same
    launch, the profiler's per-dispatch mean."""
    rp = read_trace_mean(BATCH_TRACE_FILE, kname, tiles)
    return {"level": level, "tiles": tiles, "us": us, "alg_bytes": alg_bytes, "GBps": alg_bytes / (us * 1e-6) / 1e9,
            "frac_hbm_peak": alg_bytes / (us * 1e-6) / 1e9 / PEAK_HBM_GBS, "rocprof_us": rp,
            "frac_hbm_peak_rocprof": alg_bytes / (rp * 1e-6) / 1e9 / PEAK_HBM_GBS if rp else None}


def run_amg(ca, P, b_host, solver_kw, reps=10):
    """AMGSolver on the adaptive 256^3 workload: prepareTypesAndCoeffs once,
    then AMG-preconditioned solves to 1e-6 (same settings as the plain GMG
    solve); the stored-coefficient colour sweep on the finest level timed
    alone (algorithmic bytes 29 per cell: x, rhs, the four coefficients and the
    cell type read, x written)."""
    S = ca.AMGSolver(1.0, 1.0)
    t = time.perf_counter()
    S.prepareTypesAndCoeffs(P)
    t_prep = time.perf_counter() - t
    nu, nc = solver_kw["level_iters"], solver_kw["coarsest_iters"]
    lat = []
    for rep_i in range(reps + 2):
        P.set(ca.CH_B, b_host)
        t = time.perf_counter()
        S.solve(P, False, 1000, 1e-6, nu, nc, 1, False)
        if rep_i >= 2:
            lat.append(time.perf_counter() - t)
        assert S.last_batch[0]["converged"] == 1
    L = P.max_level
    us, tiles = P.benchmark_kernel(ca.KERNEL_AMG_GS, L, reps=100)
    alg = tiles * 512 * 29
    return {"workload": "adaptive octree 256^3 effective, AMG stored-coefficient V-cycle (RB-GS 2+2 on the stored "
                        "rows), FP32, tol 1e-6",
            "prepare_ms": t_prep * 1e3, "solves_per_s": len(lat) / float(np.sum(lat)),
            "ms_per_solve": float(np.median(lat)) * 1e3, "iterations": S.last_batch[0]["iters"],
            "amg_gs_sweep": {"level": L, "tiles": tiles, "us": us, "alg_bytes": alg,
                             "GBps": alg / (us * 1e-6) / 1e9, "frac_hbm_peak": alg / (us * 1e-6) / 1e9 / PEAK_HBM_GBS}}


VORTEX_KW = VORTEX512


def run_vortex(ca, solver_kw, reps=3, kw=VORTEX_KW):
    """Mixed-precision form on the 512x512x1024-effective vortex-tube octree
    (domain [0,1]x[0,1]x[0,2], two root tiles, leaf level 6 where the vorticity
    of 8 seeded Gaussian tubes exceeds the threshold, level 4 elsewhere): BF16
    V-cycle storage, FP32 smoothing and stencils, FP64 residual and x,
    iterative refinement to 1e-8 of the FP64 residual.  The plain FP32 solve to
    1e-6 of the same problem is timed beside it."""
    t0 = time.perf_counter()
    sc = ca.Scene(**kw)
    t_build = time.perf_counter() - t0
    P = ca.Problem(sc.tree)
    b = np.ascontiguousarray(sc.rhs, np.float32)
    S = ca.GMGSolver(1.0, 1.0)
    nu, nc, sm = solver_kw["level_iters"], solver_kw["coarsest_iters"], solver_kw["smoother"]

    def refined():
        P.set(ca.CH_B, b)
        return S.solve_refined(P, 1e-8, 1e-4, 20, 1000, nu, nc, smoother=sm, storage=ca.STORE_BF16)[0]

    def plain():
        P.set(ca.CH_B, b)
        S.solve(P, False, 1000, 1e-6, nu, nc, 1, False, smoother=sm, verify_f64=True)
        return S.last

    refined()
    plain()
    rr = [refined() for _ in range(reps)]
    x64 = P.get_f64()  # the refined pressure (a verified plain solve overwrites the FP64 array)
    pp = [plain() for _ in range(reps)]
    r = rr[-1]
    assert r["converged"] == 1
    L = P.max_level
    k32, nt = P.benchmark_kernel(ca.KERNEL_GS_SWEEP, L, reps=50)
    k16, _ = P.benchmark_kernel(ca.KERNEL_GS_SWEEP_BF16, L, reps=50)
    return {
        "workload": "vortex-tube octree 512x512x1024 effective (8 seeded Gaussian tubes, leaf level 6 where "
                    "|w| > 0.5, level 4 elsewhere), BF16 V-cycle storage, FP32 smoothing/stencils, FP64 residual "
                    "and x, iterative refinement to 1e-8",
        "tiles": P.n, "leaf_tiles": P.n_leaf, "leaf_cells": P.n_leaf * 512, "build_s": t_build,
        "solves_per_s": reps / sum(q["seconds"] for q in rr), "ms_per_solve": float(np.median([q["seconds"] for q in rr])) * 1e3,
        "refinement_steps": r["outer"], "iterations": r["iters"], "fp64_rel_res": r["rel_res"],
        "plain_f32_1e-6": {"ms_per_solve": float(np.median([q["seconds"] for q in pp])) * 1e3,
                           "iterations": pp[-1]["iters"], "fp64_rel_res": pp[-1]["rel_res_f64"]},
        "gs_sweep_level": L, "gs_sweep_tiles": nt,
        "gs_sweep_f32": {"us": k32, "GBps": gs_alg_bytes(nt, 4) / (k32 * 1e-6) / 1e9},
        "gs_sweep_bf16": {"us": k16, "GBps": gs_alg_bytes(nt, 2) / (k16 * 1e-6) / 1e9},
        "_check": (kw, sc.tree.level, sc.tree.coord, sc.tree.type, b, x64),
    }
TRAFFIC_FILE = os.path.join(ROOT, "profiles", "k_gs_color_adaptive256_hbm_traffic.txt")
SINGLE_TRACE_FILE = os.path.join(ROOT, "profiles", "k_gs_color_adaptive256_kernel_stats.txt")
BATCH_TRACE_FILE = os.path.join(ROOT, "profiles", "k_gs_color_batch256_kernel_stats.txt")
FORMS = ["plain64", "projection", "amg", "plain_drop_in", "batched", "batched_mixed", "batched_mixed_f16", "batched_autotuned", "served",
         "served_mixed", "served_mixed_light", "vortex"]


def note(msg):
    """Progress line on stderr (the JSON result goes to stdout)."""
    print(f"[bench] {msg}", file=sys.stderr, flush=True)
KERNEL_REPS = 20


def build_if_missing():
    import subprocess
    if not os.path.exists(os.path.join(ROOT, "cirrus_amd", "libcirrus_mgpcg.so")):
        subprocess.run([sys.executable, "-m", "cirrus_amd.build", "-q"], check=True, cwd=ROOT)


def cpu_baseline(scene_kw, solver_kw, budget_s=20.0, max_solves=8):
    """Oracle (CPU transcription, FP32, OpenMP) solves of the same problem."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import subprocess
    if not os.path.exists(os.path.join(ROOT, "oracle", "liboracle_mgpcg.so")):
        subprocess.run(["make", "-s", "-C", os.path.join(ROOT, "oracle")], check=True)
    import pyoracle as po
    o = po.Oracle(**scene_kw)
    b = o.get(po.CH_B)
    kw = dict(max_iters=1000, tol=1e-6, level_iters=solver_kw["level_iters"],
              coarsest_iters=solver_kw["coarsest_iters"], smoother=solver_kw["smoother"])
    o.solve(**kw)  # warm-up
    times, iters = [], 0
    t_start = time.perf_counter()
    while len(times) < max_solves and time.perf_counter() - t_start < budget_s:
        o.set(po.CH_B, b)
        r = o.solve(**kw)
        times.append(r["seconds"])
        iters = r["iters"]
    return dict({
        "value": len(times) / sum(times), "unit": "solves/s", "cores": po.get_threads(), "kind": "port",
        "sample": f"{len(times)} full solves of the same problem ({iters} iterations each), "
                  f"median {np.median(times) * 1e3:.1f} ms",
    }, **host_cpu())


def host_cpu():
    """Model, sockets and logical CPUs of this host (from /proc/cpuinfo) and
    the CPUs this process may run on."""
    model, sockets, logical = None, set(), 0
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                k, _, v = line.partition(":")
                k, v = k.strip(), v.strip()
                if k == "model name" and model is None:
                    model = v
                elif k == "physical id":
                    sockets.add(v)
                elif k == "processor":
                    logical += 1
    except OSError:
        pass
    return {"cpu_model": model, "sockets": len(sockets) or None, "host_logical_cpus": logical or None,
            "cpus_available": len(os.sched_getaffinity(0))}


def fp64_oracle_error(kw, level, coord, ttype, b32, x64, solver_kw, tol=1e-11):
    """Relative L2 error, over LEAF INTERIOR cells, of a GPU pressure x64
    (FP64; the library scene's tile order given by level/coord/type; solved
    for the FP32 right-hand side b32) against the FP64 oracle's solution of the
    same system to |b - A x| < tol |b|.  The oracle starts from the GPU's x64:
    it forms r = b - A x64 with its FP64 operator, solves A d = r by its FP64
    MGPCG until |r - A d| < tol |b|, and x64 + d is the reference pressure, so
    the error is |d| / |x64 + d| (the oracle's own error is then about tol
    times the condition number, far below the one measured)."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import pyoracle as po
    o64 = po.Oracle(fp64=True, **kw)
    lev, crd, typ, _ = o64.topology()
    ko = {(int(l), *map(int, c)): i for i, (l, c) in enumerate(zip(lev, crd))}
    to_o = np.array([ko[(int(l), *map(int, c))] for l, c in zip(level, coord)], np.int64)
    assert np.array_equal(typ[to_o], ttype)
    n = len(lev)
    leaf = typ == po.LEAF
    m = leaf[:, None] & (o64.celltypes() == po.INTERIOR)
    b = np.zeros((n, 512))
    b[to_o] = np.asarray(b32, np.float64)
    b[~m] = 0.0
    x0 = np.zeros((n, 512))
    x0[to_o] = x64
    x0[~leaf] = 0.0
    t = time.perf_counter()
    o64.set(po.CH_X, x0)
    o64.full_negative_laplacian(po.CH_X, po.CH_AP)
    r = np.where(m, b - o64.get(po.CH_AP), 0.0)
    bn, rn = float(np.linalg.norm(b[m])), float(np.linalg.norm(r[m]))
    o64.set(po.CH_B, r)
    r64 = o64.solve(max_iters=1000, tol=min(1.0, tol * bn / rn) if rn > 0 else 1.0,
                    level_iters=solver_kw["level_iters"], coarsest_iters=solver_kw["coarsest_iters"],
                    smoother=solver_kw["smoother"])
    t64 = time.perf_counter() - t
    d = o64.get(po.CH_X)
    xr = x0 + d
    err = float(np.linalg.norm(d[m]) / np.linalg.norm(xr[m]))
    o64.close()
    return err, {"gpu_fp64_rel_res_by_oracle": rn / bn, "oracle_fp64_rel_res": r64["rel_res"] * rn / bn,
                 "oracle_iterations": r64["iters"], "oracle_seconds": t64, "oracle_threads": po.get_threads()}


def gs_alg_bytes(tiles, s=4):
    """Algorithmic bytes of one red-black colour sweep: x and rhs read, x written
    for half the cells, the cell type read (2s + t per cell; s = 4 for FP32,
    2 for BF16 storage; t = 1)."""
    return tiles * 512 * (2 * s + 1)


def kernels_only(P, ca, level=None, blocks=(1,)):
    """The dominant kernel (FP32 and BF16 storage; for each tile block given)
    and the copy reference, KERNEL_REPS launches each (the workload for
    hardware-counter runs)."""
    L = P.max_level if level is None else level
    out = {"gs_level": L}
    for blk in blocks:
        P.set_schedule(0, True, blk)
        us, tiles = P.benchmark_kernel(ca.KERNEL_GS_SWEEP, L, reps=KERNEL_REPS)
        bus, _ = P.benchmark_kernel(ca.KERNEL_GS_SWEEP_BF16, L, reps=KERNEL_REPS)
        out.update({"gs_tiles": tiles, f"gs_us_block{blk}": us, f"gs_bf16_us_block{blk}": bus,
                    f"gs_GBps_block{blk}": gs_alg_bytes(tiles) / (us * 1e-6) / 1e9,
                    f"gs_bf16_GBps_block{blk}": gs_alg_bytes(tiles, 2) / (bus * 1e-6) / 1e9})
    P.set_schedule(0, True, 1)
    cus, ctiles = P.benchmark_kernel(ca.KERNEL_COPY, 0, reps=KERNEL_REPS)
    out.update({"copy_tiles": ctiles, "copy_us": cus})
    print(json.dumps(out), flush=True)


def _kname(raw):
    """Kernel name without return type, namespace and arguments (template
    arguments kept: k_gs_color<float>)."""
    import re
    m = re.match(r"_ZN2cg(\d+)", raw)
    if m:  # a mangled name the profiler could not demangle (bfloat16 template arguments)
        k = m.end()
        name = raw[k:k + int(m.group(1))]
        rest = raw[k + int(m.group(1)):]
        targ = {"IDF16bE": "<bf16>", "IfE": "<float>", "IDF16bLb0EE": "<bf16>", "IfLb0EE": "<float>",
                "IDF16bLb1EE": "<bf16, true>", "IfLb1EE": "<float, true>"}
        return name + next((v for t, v in targ.items() if rest.startswith(t)), "")
    n = raw.split("(")[0].strip()
    if n.startswith("void "):
        n = n[5:]
    # the colour sweep's <V, false> instance is the plain pass; <V, true> also sums r.z
    return n.replace("cg::", "").replace("__bf16", "bf16").replace(", false>", ">")


def _pmc_per_dispatch(path, counter):
    """Mean counter value per dispatch, by kernel name and grid size."""
    import csv
    import collections
    acc = collections.defaultdict(list)
    with open(path) as f:
        for r in csv.DictReader(f):
            if r.get("Counter_Name") != counter:
                continue
            name = _kname(r["Kernel_Name"])
            grid = int(r.get("Grid_Size", r.get("Grid_Size_X", 0)) or 0)
            acc[(name, grid)].append(float(r["Counter_Value"]))
    return {k: sum(v) / len(v) for k, v in acc.items()}


def summarize_pmc(fetch_csv, write_csv, title=None, out=None):
    """Scale FETCH_SIZE / WRITE_SIZE (KiB per dispatch) of the colour sweeps by
    the copy kernel's reading of its known traffic; write the profile summary.
    Default: the adaptive 256^3 workload's summary, whose traffic line feeds
    the roofline record."""
    fetch = _pmc_per_dispatch(fetch_csv, "FETCH_SIZE")
    write = _pmc_per_dispatch(write_csv, "WRITE_SIZE")

    def pick(d, name, required=True):
        ks = [k for k in d if k[0] == name]
        if not ks:
            if required:
                raise SystemExit(f"no {name} dispatches in the counter file")
            return None
        k = max(ks, key=lambda k: k[1])
        return k, d[k]

    (ck, cf), (_, cw) = pick(fetch, "k_copy_tiles"), pick(write, "k_copy_tiles")
    copy_tiles = ck[1] // 64
    copy_bytes = copy_tiles * 512 * 4
    f_scale = copy_bytes / (cf * 1024.0)
    w_scale = copy_bytes / (cw * 1024.0)
    default = title is None
    title = title or "the adaptive 256^3 workload"
    extra = "" if default else " --kernels-batch N"
    lines = [
        f"HBM traffic of the red-black colour sweeps (most populated level) on {title}",
        "",
        "commands:",
        f"  rocprofv3 --pmc FETCH_SIZE -d OUT/fetch -o pmc --output-format csv -- python3 bench.py --kernels-only{extra}",
        f"  rocprofv3 --pmc WRITE_SIZE -d OUT/write -o pmc --output-format csv -- python3 bench.py --kernels-only{extra}",
        "  python bench.py --summarize-pmc OUT/fetch/pmc_counter_collection.csv OUT/write/pmc_counter_collection.csv"
        + ("" if default else " --pmc-title '...' --pmc-out FILE"),
        "",
        f"copy reference k_copy_tiles: {copy_tiles} tiles, {copy_bytes} B read and {copy_bytes} B written per dispatch",
        f"  FETCH_SIZE {cf:.1f} KiB/dispatch -> read scale {f_scale:.4f}",
        f"  WRITE_SIZE {cw:.1f} KiB/dispatch -> write scale {w_scale:.4f}",
    ]
    main_traffic = main_tiles = None
    for name, s_bytes, per_wave in [("k_gs_color<float>", 4, 1), ("k_gs_color<bf16>", 2, 1),
                                    ("k_gs_color2<float>", 4, 2), ("k_gs_color2<bf16>", 2, 2)]:
        f = pick(fetch, name, required=per_wave == 1)
        w = pick(write, name, required=per_wave == 1)
        if f is None or w is None:
            continue
        (k, fv), (_, wv) = f, w
        # the counter run launches each sweep over the same list; its tile count
        # comes from the one-tile-per-wavefront grid
        tiles = pick(fetch, "k_gs_color<float>")[0][1] // 64
        rd, wr = fv * 1024.0 * f_scale, wv * 1024.0 * w_scale
        alg = gs_alg_bytes(tiles, s_bytes)
        lines += [f"{name}: {tiles} tiles, {per_wave} per wavefront (grid {k[1]} threads)",
                  f"  FETCH_SIZE {fv:.1f} KiB/dispatch -> {rd:.0f} B read",
                  f"  WRITE_SIZE {wv:.1f} KiB/dispatch -> {wr:.0f} B written",
                  f"  algorithmic bytes per launch {alg} (2s+t per cell, s = {s_bytes})",
                  f"  traffic / algorithmic = {(rd + wr) / alg:.3f}"]
        if name == "k_gs_color<float>":
            main_traffic, main_tiles = rd + wr, tiles
    lines += ["", f"traffic_bytes_per_launch = {main_traffic:.0f}", f"traffic_tiles = {main_tiles}"]
    path = out or TRAFFIC_FILE
    os.makedirs(os.path.dirname(os.path.abspath(path)), exist_ok=True)
    with open(path, "w") as fh:
        fh.write("\n".join(lines) + "\n")
    print("\n".join(lines))


STATS_FILE = os.path.join(ROOT, "profiles", "mgpcg_adaptive256_kernel_stats.txt")


def summarize_trace(trace_csv, stats_csv, title=None, out=None, bench_args=None):
    """Per-kernel and per-(kernel, grid) durations of a --kernel-trace --stats run;
    writes the profile summary (default: the adaptive 256^3 workload's)."""
    import csv
    import collections
    bench_args = bench_args or "--steps 5 --warmup 1 --no-cpu-baseline --batch 0 --serve-seconds 0 --no-vortex"
    lines = [
        f"Kernel durations, {title or 'adaptive 256^3 MGPCG workload'} (one MI355X)",
        "",
        "command:",
        f"  rocprofv3 --kernel-trace --stats -d OUT -o run --output-format csv -- python3 bench.py {bench_args}",
        "  python bench.py --summarize-trace OUT/run_kernel_trace.csv OUT/run_kernel_stats.csv"
        + ("" if title is None else " --trace-title '...' --trace-args '...' --trace-out FILE"),
        "",
        "rocprofv3 --stats (all launches of each kernel):",
        f"  {'kernel':24s} {'calls':>7s} {'avg us':>9s} {'total ms':>9s} {'%':>6s}",
    ]
    with open(stats_csv) as f:
        for r in csv.DictReader(f):
            name = _kname(r["Name"])
            lines.append(f"  {name:24s} {int(r['Calls']):7d} {float(r['AverageNs']) / 1e3:9.3f} "
                         f"{float(r['TotalDurationNs']) / 1e6:9.3f} {float(r['Percentage']):6.2f}")
    acc = collections.defaultdict(list)
    seq = []
    with open(trace_csv) as f:
        for r in csv.DictReader(f):
            name = _kname(r["Kernel_Name"])
            wg = int(r.get("Workgroup_Size_X", r.get("Workgroup_Size", 1)) or 1)
            grid = int(r.get("Grid_Size_X", r.get("Grid_Size", 0)) or 0)
            t0, t1 = int(r["Start_Timestamp"]), int(r["End_Timestamp"])
            acc[(name, grid // max(1, wg))].append((t1 - t0) / 1e3)
            seq.append((t0, name, grid // max(1, wg), (t1 - t0) / 1e3))
    lines += ["", "kernel trace by launch size (workgroups = tiles for the tile kernels):",
              f"  {'kernel':24s} {'WGs':>6s} {'calls':>6s} {'mean us':>8s} {'median us':>9s}"]
    for (name, wgs), v in sorted(acc.items(), key=lambda kv: -sum(kv[1])):
        v = sorted(v)
        lines.append(f"  {name:24s} {wgs:6d} {len(v):6d} {sum(v) / len(v):8.3f} {v[len(v) // 2]:9.3f}")
    # the roofline record's launches: bench.py times the colour sweep in one run
    # of back-to-back dispatches of the same size (its longest such run here)
    seq.sort()
    best, cur = [], []
    for t0, name, wgs, d in seq:
        if name == "k_gs_color<float>" and cur and cur[-1][0] == wgs:
            cur.append((wgs, d))
        else:
            cur = [(wgs, d)] if name == "k_gs_color<float>" else []
        if len(cur) > len(best):
            best = list(cur)
    if len(best) >= 20:
        d = [x[1] for x in best]
        lines += ["", f"roofline launches (bench.py's back-to-back k_gs_color<float> run): {len(d)} dispatches of "
                  f"{best[0][0]} workgroups, mean {sum(d) / len(d):.3f} us, median {sorted(d)[len(d) // 2]:.3f} us"]
    path = out or STATS_FILE
    os.makedirs(os.path.dirname(os.path.abspath(path)), exist_ok=True)
    with open(path, "w") as f:
        f.write("\n".join(lines) + "\n")
    print("\n".join(lines))


def read_trace_mean(path, kernel, tiles):
    """Mean per-dispatch duration (us) of `kernel` at `tiles` workgroups from a
    committed kernel-trace summary (its per-launch-size table), if present."""
    try:
        with open(path) as f:
            for line in f:
                parts = line.split()
                if len(parts) == 5 and parts[0] == kernel and parts[1] == str(tiles):
                    return float(parts[3])
    except (OSError, ValueError):
        pass
    return None


def read_traffic(tiles):
    """Traffic per launch from the committed counter summary, if it covers this launch."""
    try:
        kv = {}
        with open(TRAFFIC_FILE) as f:
            for line in f:
                if " = " in line:
                    k, v = line.split(" = ", 1)
                    kv[k.strip()] = v.strip()
        if int(kv["traffic_tiles"]) == tiles:
            return float(kv["traffic_bytes_per_launch"])
    except (OSError, KeyError, ValueError):
        pass
    return None


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=50)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--workload", default="adaptive256", choices=list(WORKLOADS))
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--pair-sweeps", type=int, default=-1,
                    help="one-launch red-black iterations on the measured handles: 1 on, 0 off, -1 the default")
    ap.add_argument("--no-oracle-check", action="store_true",
                    help="skip the FP64-oracle errors of the mixed-precision results (host solves after the timing)")
    ap.add_argument("--graph", action="store_true",
                    help="always replay one CG iteration's hipGraph (default: the library chooses per handle; "
                         "results identical)")
    ap.add_argument("--eager-max", type=int, default=-1,
                    help="levels of at most this many tiles use the one-wait prolongation pass (-1: the library default)")
    ap.add_argument("--fused16-max", type=int, default=-1,
                    help="16-bit levels of at most this many tiles run the first down-sweep iteration in one launch "
                         "(-1: the library default)")
    ap.add_argument("--batch", type=int, default=256,
                    help="problems in the batched form measured beside the plain solve (0: skip)")
    ap.add_argument("--batch-reps", type=int, default=3)
    ap.add_argument("--serve-seconds", type=float, default=60.0,
                    help="length of the Poisson request window of the served_mixed form (0: skip the served forms)")
    ap.add_argument("--serve-seconds-fp32", type=float, default=10.0,
                    help="length of the request window of the FP32 served form")
    ap.add_argument("--serve-seconds-light", type=float, default=15.0,
                    help="length of each request window of the served_mixed_light sweep")
    ap.add_argument("--serve-rate", type=float, default=2000.0)
    ap.add_argument("--serve-rate-light", default="0.5,0.65,0.8",
                    help="offered rates of the served_mixed_light sweep: comma list, as fractions of what served_mixed "
                         "completed per s when below 1.5, else requests/s")
    ap.add_argument("--light-p99-ms", type=float, default=50.0,
                    help="served_mixed_light reports the highest swept rate whose p99 (after the first quarter) "
                         "stays within this bound")
    ap.add_argument("--serve-instances", type=int, default=256, help="most instances resident at once")
    ap.add_argument("--serve-scenes", type=int, default=256,
                    help="lognormal scenes the served instances draw their octrees from")
    ap.add_argument("--serve-budget-gb", type=float, default=200.0, help="HBM admission budget of the served forms")
    ap.add_argument("--forms", default="all",
                    help=f"comma list of the forms to run beside the plain solve ({FORMS}), or all, or none")
    ap.add_argument("--no-vortex", action="store_true",
                    help="skip the mixed-precision refinement form on the vortex-tube octree")
    ap.add_argument("--kernels-only", action="store_true",
                    help="run only the dominant kernel and the copy reference (for counter runs)")
    ap.add_argument("--kernels-batch", type=int, default=0,
                    help="with --kernels-only: run the kernels on the pool of this many batched-form problems")
    ap.add_argument("--summarize-pmc", nargs=2, metavar=("FETCH_CSV", "WRITE_CSV"))
    ap.add_argument("--pmc-title", default=None, help="with --summarize-pmc: the workload the counters cover")
    ap.add_argument("--trace-title", default=None, help="with --summarize-trace: the workload the trace covers")
    ap.add_argument("--trace-args", default=None, help="with --summarize-trace: the bench.py arguments of the traced run")
    ap.add_argument("--trace-out", default=None, help="with --summarize-trace: the summary file to write")
    ap.add_argument("--pmc-out", default=None, help="with --summarize-pmc: the summary file to write")
    ap.add_argument("--summarize-trace", nargs=2, metavar=("TRACE_CSV", "STATS_CSV"))
    args = ap.parse_args()
    if args.summarize_pmc:
        summarize_pmc(*args.summarize_pmc, title=args.pmc_title, out=args.pmc_out)
        return
    if args.summarize_trace:
        summarize_trace(*args.summarize_trace, title=args.trace_title, out=args.trace_out, bench_args=args.trace_args)
        return
    if args.gpus != 1:
        print("bench.py measures one GPU: every solve here runs on a single MI355X; "
              "run it with --gpus 1 (independent problems would be split across GPUs by a front end).",
              file=sys.stderr)
        sys.exit(2)

    build_if_missing()
    import cirrus_amd as ca
    if ca.device_count() < 1:
        raise SystemExit("no HIP device visible")
    scene_kw, solver_kw, desc = WORKLOADS[args.workload]
    t0 = time.perf_counter()
    scene = ca.Scene(**scene_kw)
    t_build = time.perf_counter() - t0
    t0 = time.perf_counter()
    P = ca.Problem(scene.tree)
    t_create = time.perf_counter() - t0
    if args.pair_sweeps >= 0 or args.eager_max >= 0 or args.fused16_max >= 0:
        P.set_level_forms(args.eager_max, args.fused16_max, args.pair_sweeps)
    if args.kernels_only:
        if args.kernels_batch > 0:
            # the same kernels at batch scale, past the last-level cache: the
            # pool of the batched form's problems, on its most populated level
            scenes = build_scenes(ca, [batch_scene_params(q)[0] for q in range(args.kernels_batch)])
            B = ca.Problem.batch([sc.tree for sc in scenes])
            lv = int(np.bincount(np.concatenate([sc.tree.level[sc.tree.type != 4] for sc in scenes])).argmax())
            del scenes
            kernels_only(B, ca, level=lv, blocks=(1, 2, 4, 8))
        else:
            kernels_only(P, ca, blocks=(1, 2, 4, 8))
        return
    S = ca.GMGSolver(1.0, 1.0)

    def step():
        return S.solve(P, False, 1000, 1e-6, solver_kw["level_iters"], solver_kw["coarsest_iters"], 1, False,
                       smoother=solver_kw["smoother"], use_graph=True if args.graph else None)

    # the solve overwrites b with the residual: re-upload b before each solve,
    # outside the timed call (a host->device copy is not part of the solve)
    b_host = np.ascontiguousarray(scene.rhs, np.float32)

    def reset_b():
        P.set(ca.CH_B, b_host)

    for _ in range(args.warmup):
        reset_b()
        step()
    lat = []
    iters = None
    for _ in range(args.steps):
        reset_b()
        t = time.perf_counter()
        step()
        lat.append(time.perf_counter() - t)
        iters = S.last["iters"]
        assert S.last["converged"] == 1
    total = float(np.sum(lat))
    lat_ms = np.array(lat) * 1e3

    # dominant kernel: red-black GS colour sweep on the finest level, timed with
    # hipEvents on the library's stream; algorithmic bytes 2s+t per cell
    # (x and rhs read, x written for half the cells; s=4 B, t=1 B)
    L = P.max_level
    us, tiles = P.benchmark_kernel(ca.KERNEL_GS_SWEEP, L, reps=200)
    alg_bytes = gs_alg_bytes(tiles)
    traffic = read_traffic(tiles)
    rp_us = read_trace_mean(SINGLE_TRACE_FILE, "k_gs_color<float>", tiles)
    achieved = alg_bytes / (us * 1e-6) / 1e9
    per_kernel = {}
    for name, kid, lvl in [("apply", ca.KERNEL_APPLY, 0), ("restrict_L", ca.KERNEL_RESTRICT, L),
                           ("prolong_L", ca.KERNEL_PROLONG, L), ("gs_prolong_L", ca.KERNEL_GS_PROLONG, L), ("update_xr", ca.KERNEL_UPDATE_XR, 0),
                           ("gs_level1", ca.KERNEL_GS_SWEEP, 1), ("coarse_lds", ca.KERNEL_COARSE, 10),
                           ("empty_1wg", ca.KERNEL_EMPTY, 1), ("empty_2240wg", ca.KERNEL_EMPTY, 2240),
                           ("copy_all", ca.KERNEL_COPY, 0)]:
        u, nt = P.benchmark_kernel(kid, lvl, reps=100)
        per_kernel[name] = {"us": round(u, 3), "tiles": nt}

    out = {
        "metric": "MGPCG solves/s + p99 latency at 1e-6 rel. residual vs plain; V-cycle HBM GB/s",
        "value": args.steps / total,
        "unit": "solves/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": total / args.steps * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {"workload": desc, "tiles": P.n, "leaf_tiles": P.n_leaf, "max_level": L,
                   "leaf_cells": P.n_leaf * 512, "iterations": iters,
                   "launch": "hipGraph replay" if args.graph else "chosen per handle (stream launches at this size)"},
        "p50_ms": float(np.percentile(lat_ms, 50)),
        "p99_ms": float(np.percentile(lat_ms, 99)),
        "build_s": t_build,
        "create_s": t_create,
        "roofline": {"bound": "hbm", "kernel": f"k_gs_color level {L} ({tiles} tiles)", "achieved": achieved,
                     "peak": PEAK_HBM_GBS, "unit": "GB/s", "frac": achieved / PEAK_HBM_GBS, "traffic": traffic,
                     "avg_launch_us": us, "alg_bytes_per_launch": alg_bytes,
                     "rocprof_avg_launch_us": rp_us, "frac_rocprof": (alg_bytes / (rp_us * 1e-6) / 1e9 / PEAK_HBM_GBS)
                     if rp_us else None, "rocprof_summary": os.path.relpath(SINGLE_TRACE_FILE, ROOT)},
        "kernels_us": per_kernel,
    }
    note(f"plain solve: {out['value']:.1f} solves/s")
    forms = {}
    pending_checks = []  # (form, data) compared with the FP64 oracle at the end
    wanted = set(FORMS if args.forms == "all" else ([] if args.forms == "none" else args.forms.split(",")))
    unknown = wanted - set(FORMS)
    if unknown:
        raise SystemExit(f"unknown forms {sorted(unknown)}; known: {FORMS}")
    if "plain64" in wanted and args.workload != "plain64":
        forms["plain64"] = run_plain64(ca)
        note(f"plain64: {forms['plain64']['solves_per_s']:.1f} solves/s")
    # the whole projection around the solve (FluidEuler::project without the
    # velocity BCs): divergence of the scene's face velocities, the solve, the
    # gradient added back; velocities re-uploaded before each, outside the timing
    if "projection" in wanted:
        uvw = [np.ascontiguousarray(scene.uvw[:, a, :]) for a in range(3)]
        pt = []
        for rep_i in range(12):
            for a in range(3):
                P.set(ca.CH_U + a, uvw[a])
            t = time.perf_counter()
            pr = S.project(P, 1000, 1e-6, solver_kw["level_iters"], solver_kw["coarsest_iters"],
                           smoother=solver_kw["smoother"])[0]
            if rep_i >= 2:
                pt.append(time.perf_counter() - t)
            assert pr["converged"] == 1
        forms["projection"] = {
            "workload": desc + "; divergence of the face velocities, MGPCG, pressure gradient added to the velocities",
            "ms_per_projection": float(np.median(pt)) * 1e3, "projections_per_s": len(pt) / float(np.sum(pt)),
            "solve_ms": out["ms_per_step"], "iterations": pr["iters"]}
        note("projection done")
    if "amg" in wanted:
        forms["amg"] = run_amg(ca, P, b_host, solver_kw)
        forms["amg"]["vs_gmg_solve"] = forms["amg"]["solves_per_s"] / out["value"]
        note(f"amg: {forms['amg']['solves_per_s']:.1f} solves/s")
    batched_forms = [f for f in ("batched", "batched_mixed", "batched_mixed_f16", "batched_autotuned") if f in wanted]
    if args.batch > 0 and batched_forms:
        t0 = time.perf_counter()
        scenes = build_scenes(ca, [batch_scene_params(q)[0] for q in range(args.batch)])
        t_build_b = time.perf_counter() - t0
        items = [(sc.tree, np.ascontiguousarray(sc.rhs, np.float32)) for sc in scenes]
        # the same problems through the plain drop-in, one at a time: the
        # denominator of the batched forms' speed-up
        note(f"{args.batch} batch problems built in {t_build_b:.1f} s")
        base = None
        if "plain_drop_in" in wanted:
            base = plain_drop_in(ca, items, solver_kw)
            forms["plain_drop_in_batch_problems"] = dict(base, workload="the batched forms' problems, FP32, one handle "
                                                         "each, solved one after another from x = 0")
            note(f"plain drop-in on the batch problems: {base['solves_per_s']:.1f} solves/s")
        t0 = time.perf_counter()
        B = ca.Problem.batch([sc.tree for sc in scenes])
        t_pool = time.perf_counter() - t0
        if args.pair_sweeps >= 0 or args.eager_max >= 0 or args.fused16_max >= 0:
            B.set_level_forms(args.eager_max, args.fused16_max, args.pair_sweeps)
        note(f"pool of {args.batch} problems ({B.n} tiles) created in {t_pool:.2f} s")
        # FP32 storage is also refined in FP64, so that every pressure meets
        # 1e-6 of its FP64 residual like the mixed forms'
        for name, storage, refine, tune in [("batched", 0, True, False), ("batched_mixed", 1, True, False),
                                            ("batched_mixed_f16", 2, True, False), ("batched_autotuned", 0, False, True)]:
            if name not in wanted:
                continue
            f = run_batched(ca, scenes, B, args.batch_reps, solver_kw, use_graph=True if args.graph else None, storage=storage,
                            refine=refine, tune=tune, t_build=t_build_b, oracle_sample=name == "batched_mixed")
            chk = f.pop("_check")
            if chk:
                pending_checks.append((name, chk))
            if base:
                f["vs_plain_solves_per_s"] = f["solves_per_s"] / base["solves_per_s"]
            f["vs_plain_adaptive256"] = f["solves_per_s"] / out["value"]
            f["pool_create_s"] = t_pool
            forms[name] = f
            note(f"{name}: {f['solves_per_s']:.1f} solves/s")
        B.close()
        del items, scenes
    served_forms = [f for f in ("served", "served_mixed", "served_mixed_light") if f in wanted]
    if args.serve_seconds > 0 and served_forms:
        t0 = time.perf_counter()
        sscenes = build_scenes(ca, [batch_scene_params(q, seed=1)[0] for q in range(args.serve_scenes)])
        note(f"{len(sscenes)} serving scenes built in {time.perf_counter() - t0:.1f} s")
        base = None
        if "plain_drop_in" in wanted:
            # the scenes the instances draw from, through the plain drop-in, each
            # solved cold as the reference's projection does
            items = [(sc.tree, np.ascontiguousarray(sc.rhs, np.float32)) for sc in sscenes]
            base = plain_drop_in(ca, items, solver_kw, reps=1)
            forms["plain_drop_in_served_problems"] = dict(base, workload="the serving scenes (the instances draw from "
                                                          "them uniformly), FP32, one handle each, solved one after "
                                                          "another from x = 0")
            del items
        capacity = None
        budget = int(args.serve_budget_gb * (1 << 30))
        for name in served_forms:
            if name == "served":  # FP32 storage, every result FP64-verified (refinement)
                storage, lam, dur = 0, args.serve_rate, args.serve_seconds_fp32
            elif name == "served_mixed":  # the serving configuration as stated
                storage, lam, dur = 1, args.serve_rate, args.serve_seconds
            else:  # below capacity: a sweep of offered rates
                f = served_light_sweep(ca, solver_kw, sscenes, args, budget, capacity)
                if f is None:
                    continue
                if base:
                    f["vs_plain_solves_per_s"] = f["solves_per_s"] / base["solves_per_s"]
                forms[name] = f
                note(f"{name}: p99 within {args.light_p99_ms:g} ms up to {f['offered_per_s']:.0f}/s offered "
                     f"(p50 {f['p50_ms_after_first_quarter']:.1f} ms, p99 {f['p99_ms_after_first_quarter']:.1f} ms)")
                continue
            if lam <= 0 or dur <= 0:
                continue
            try:
                f = run_served(ca, solver_kw, sscenes, duration=dur, lam=lam, max_resident=args.serve_instances,
                               budget=budget, storage=storage, refine=True)
            except ca.CirrusError as e:  # recorded in the line; the other forms still run
                forms[name] = {"error": str(e)}
                note(f"{name}: {e}")
                continue
            if name == "served_mixed":
                capacity = f["solves_per_s"]
            if base:
                f["vs_plain_solves_per_s"] = f["solves_per_s"] / base["solves_per_s"]
            f["vs_plain_adaptive256"] = f["solves_per_s"] / out["value"]
            forms[name] = f
            note(f"{name}: {f['solves_per_s']:.1f} solves/s at {lam:.0f}/s offered, p99 {f['p99_ms']} ms")
        del sscenes
    if not args.no_vortex and "vortex" in wanted:
        f = run_vortex(ca, solver_kw)
        pending_checks.append(("refined_vortex512", f.pop("_check")))
        forms["refined_vortex512"] = f
        note("vortex done")
    if forms:
        out["forms"] = forms
        out["mixed_vs_plain"] = {k: forms[k]["vs_plain_solves_per_s"]
                                 for k in ("batched_mixed", "batched_mixed_f16", "served_mixed", "batched_autotuned")
                                 if k in forms and "vs_plain_solves_per_s" in forms[k]}
    if not args.no_cpu_baseline:
        out["cpu_baseline"] = cpu_baseline(scene_kw, solver_kw)
        if "plain64" in forms:
            p64_kw, p64_skw, _ = WORKLOADS["plain64"]
            forms["plain64"]["cpu_baseline"] = cpu_baseline(p64_kw, p64_skw, budget_s=5.0)
            forms["plain64"]["vs_cpu_baseline"] = forms["plain64"]["solves_per_s"] / forms["plain64"]["cpu_baseline"]["value"]
    # errors of the mixed-precision results against the FP64 oracle (after
    # every timed part; the oracle solves the same scenes on the host)
    if not args.no_oracle_check:
        for name, chk in pending_checks:
            note(f"FP64 oracle check: {name}")
            if name == "refined_vortex512":
                kw, level, coord, ttype, b32, x64 = chk
                err, info = fp64_oracle_error(kw, level, coord, ttype, b32, x64, solver_kw)
                forms[name]["fp64_oracle_check"] = dict(info, rel_l2_error_vs_fp64_oracle=err)
            else:
                rows = []
                for q, kw, level, coord, ttype, b32, x64 in chk:
                    err, info = fp64_oracle_error(kw, level, coord, ttype, b32, x64, solver_kw)
                    rows.append(dict(info, problem=q, leaf_tiles=int((ttype == 1).sum()), rel_l2_error_vs_fp64_oracle=err))
                forms[name]["fp64_oracle_check"] = {
                    "problems": rows, "rel_l2_error_max": max(r["rel_l2_error_vs_fp64_oracle"] for r in rows)}
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
