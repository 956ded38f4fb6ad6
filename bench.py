#!/usr/bin/env python3
"""Throughput benchmark: batched polar Newton power flow on one MI355X.

Workloads (``--config``):

* ``n1_10k`` (default): the 10,000-bus grid (spanning tree + 3,500 chords, 800 PV buses),
  every single-branch outage (13,499 cases) with re-slacking of islands, DC start, tol 1e-8
  p.u., max 20 iterations. One step is the whole N-1 solve on the device: islands
  (union-find), then Newton on the base Ybus minus each case's stamps. ``value`` is the plain
  FP64 form; the mixed-precision forms are timed beside it (``compare``, ``compare_2``, ``mixed_over_fp64``).
* ``ts2000``: the 2,000-bus synthetic grid (spanning tree + 600 chords, 200 PV buses, one
  slack), 8,760 hourly load cases (daily sinusoid x N(1, 0.05) per load, PV output x
  U(0, 1)), DC-power-flow start.
* ``pf200``: the 200-bus grid, 96 load-scaled cases from a flat start.
* ``multi48``: 48 grids of 100..20,000 buses with 4,096 cases in all, submitted at once to
  the continuous-batching server (one slot pool per grid, pools overlapping on the device),
  both precisions, against the plain FP64 one-grid-at-a-time loop.
* ``serve64``: Poisson arrivals (``--rate``, ``--requests``) of time-series and N-1 jobs
  over 64 cached topologies of 100..50,000 buses; p50/p99 request latency (from each
  request's scheduled arrival) and throughput of the server scheduling shortest work first,
  beside the same server running every worker at once and the plain FP64
  one-request-at-a-time loop.

A step is one complete batched solve of all cases from the start voltage, with the
inputs already resident in device memory. The JSON line reports converged cases/s over
the timed steps, p50/p99 batch latency, the roofline position of the dominant kernel
(HIP events on the library's stream, algorithmic bytes per launch; HBM traffic per launch
from the committed rocprofv3 counter summary when one matches) and the CPU oracle timed on
the host cores over a bounded sample of the same workload, with the per-case error of the
GPU results against it.
"""

from __future__ import annotations

import argparse
import ctypes as C
import json
import os
import re
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
for p in (os.path.join(ROOT, "parallel-pandapower-solver_amd"), ROOT):
    if p not in sys.path:
        sys.path.insert(0, p)

import numpy as np  # noqa: E402

HBM_PEAK_GBS = 8000.0
METRIC = "converged PF cases/sec at 1e-8 p.u. mismatch; p50/p99 solve latency"
MAX_ITER, TOL = 20, 1e-8
TRAFFIC_FILE = os.path.join(ROOT, "profiles", "traffic_summary.md")


def log(msg: str) -> None:
    print(msg, file=sys.stderr, flush=True)


def measured_traffic(kernel: str, workload: str, precision: str):
    """HBM bytes per launch (FETCH_SIZE + WRITE_SIZE, calibrated) of a kernel on a workload, from
    the lines 'traffic kernel=K workload=W precision=P bytes=B' of the committed summary."""
    if not os.path.exists(TRAFFIC_FILE):
        return None
    pat = re.compile(r"traffic kernel=(\S+) workload=(\S+) precision=(\S+) bytes=([0-9.eE+]+)")
    with open(TRAFFIC_FILE) as fh:
        for line in fh:
            m = pat.search(line)
            if m and m.group(1) == kernel and m.group(2) == workload and m.group(3) == precision:
                return float(m.group(4))
    return None


def stats_acc():
    return {"ms_eval": 0.0, "ms_lu": 0.0, "ms_step": 0.0, "ms_decide": 0.0, "ms_other": 0.0, "ms_refine": 0.0,
            "n_eval": 0, "n_lu": 0, "n_step": 0, "n_decide": 0, "n_refine": 0, "double_factorizations": 0,
            "single_factorizations": 0, "refinement_steps": 0, "newton_passes": 0, "line_search_passes": 0,
            "ms_lu_batch": 0.0, "n_lu_batch": 0, "lu_batch_cases": 0}


def add_stats(acc, L, bh):
    from p3smi import _lib

    st = _lib.Stats()
    L.p3smi_batch_stats(bh, C.byref(st))
    for key in acc:
        acc[key] += getattr(st, key)


def roofline(acc, bpc, steps, workload, precision):
    """Dominant kernel's achieved algorithmic GB/s over its event-timed launches.

    In an FP64 solve that is the LU's factor + solve launch of the batch shape, timed apart by
    the engine (ms_lu_batch over n_lu_batch launches, lu_batch_cases factorizations in them):
    bytes per case x cases in those launches / their time. The launches of the few-cases shape
    (the last stragglers of a solve) are in kernel_ms_per_step only. In mixed precision the LU
    launches also carry the refinement solves, which the byte count leaves out."""
    batch = precision == "double" and acc["n_lu_batch"] > 0
    kern = "lu" if acc["ms_lu"] >= acc["ms_eval"] else "eval"
    if kern == "lu" and batch:
        bytes_total = acc["lu_batch_cases"] * bpc[1]
        ms, n_launch = acc["ms_lu_batch"], acc["n_lu_batch"]
    else:
        facts = acc["double_factorizations"] + acc["single_factorizations"]
        bytes_total = facts * (bpc[1] if kern == "lu" else bpc[0])
        ms, n_launch = acc[f"ms_{kern}"], acc[f"n_{kern}"]
    achieved = bytes_total / (ms / 1e3) / 1e9 if ms > 0 else 0.0
    name = "k_lu_batch" if kern == "lu" and batch else f"k_{kern}"
    traffic = measured_traffic(name, workload, precision)
    avg_ms = ms / max(n_launch, 1)
    out = {
        "kernel": name,
        "bound": "hbm",
        "achieved": round(achieved, 2),
        "peak": HBM_PEAK_GBS,
        "unit": "GB/s",
        "frac": round(achieved / HBM_PEAK_GBS, 4),
        "traffic": traffic,
        "avg_launch_ms": round(avg_ms, 4),
        "launches": int(n_launch),
        "cases_per_launch": round(bytes_total / bpc[1 if kern == "lu" else 0] / max(n_launch, 1), 1),
        "bytes_per_case": float(bpc[1] if kern == "lu" else bpc[0]),
        "kernel_ms_per_step": {k: round(acc[f"ms_{k}"] / steps, 3)
                               for k in ("eval", "lu", "refine", "step", "decide", "other")},
    }
    if traffic:
        out["traffic_GB_per_s"] = round(traffic / (avg_ms / 1e3) / 1e9, 1)
    return out


def base_out(args, value, times, t_total, precision, config):
    return {
        "metric": METRIC,
        "value": round(value, 1),
        "unit": "cases/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(1e3 * t_total / args.steps, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": {"double": "f64", "mixed": "f32-lu+f64", "mixed_switch": "f32/f64-lu"}[precision],
        "data": "synthetic",
        "config": config,
        "latency_ms": {"p50": round(1e3 * float(np.percentile(times, 50)), 3),
                       "p99": round(1e3 * float(np.percentile(times, 99)), 3)},
    }


def lu_shape(info):
    return [int(info["lu_g"]), int(info["lu_ns"]), int(info["lu_few_g"]), int(info["lu_few_ns"]),
            int(info["lu_few_max"])]


def dv_hist(Vg, Vr, ok):
    """Per-case error against the oracle: max over buses of |dVm| and |dVa| for every case of
    the sample where both converged, counted in decades (upper bounds 1e-15 .. 1e-9, then
    above 1e-9)."""
    edges = [1e-15, 1e-14, 1e-13, 1e-12, 1e-11, 1e-10, 1e-9]
    out = {}
    for name, f in (("dVm", np.abs), ("dVa", np.angle)):
        a, b = Vg[:, ok], Vr[:, ok]
        fin = np.isfinite(a) & np.isfinite(b)
        d = np.where(fin, np.abs(f(np.where(fin, a, 1.0)) - f(np.where(fin, b, 1.0))), 0.0).max(axis=0, initial=0.0)
        counts = np.histogram(d, bins=[0.0] + edges + [np.inf])[0]
        out[name] = {("<=%.0e" % e if k < len(edges) else ">1e-09"): int(c)
                     for k, (e, c) in enumerate(zip(edges + [np.inf], counts))}
    return out


def time_best_of(fn, reps=3):
    """fn() timed `reps` times after one untimed warm-up call; (best seconds, last result)."""
    out = fn()
    best = np.inf
    for _ in range(reps):
        t0 = time.perf_counter()
        out = fn()
        best = min(best, time.perf_counter() - t0)
    return best, out


def dv_max(Vg, Vr, ok):
    """max |dVm|, |dVa| over the finite entries of the columns ok."""
    a, b = Vg[:, ok], Vr[:, ok]
    fin = np.isfinite(a) & np.isfinite(b)
    if not fin.any():
        return 0.0, 0.0
    return (float(np.abs(np.abs(a[fin]) - np.abs(b[fin])).max()),
            float(np.abs(np.angle(a[fin]) - np.angle(b[fin])).max()))


# ====================================================================== time series
def timeseries_workload(name: str):
    from p3smi import grids
    from p3smi.timeseries import dc_initial_voltage

    if name == "ts2000":
        g = grids.config_grid("ts2000", 0)
        S = grids.timeseries_daily(g, 8760, seed=1)
        V0 = dc_initial_voltage(g)  # replaced by the device DC start once the solver exists
        desc = "2000-bus synthetic grid, 8760 hourly cases, DC start (on the device), tol 1e-8, max_iter 20"
    else:
        g = grids.config_grid("pf200", 0)
        S = grids.timeseries_scaled(g, 96, seed=1)
        V0 = g.flat_start()
        desc = "200-bus synthetic grid, 96 load-scaled cases, flat start, tol 1e-8, max_iter 20"
    return g, S, V0, desc


def cpu_timeseries(args, g, S_c, V0_c, V, it, conv, out):
    from oracle.oracle import OracleSolver
    from p3smi.timeseries import usable_cpus

    Yp, Yj, Yx = g.csr()
    T = S_c.shape[1]
    cores = usable_cpus()
    ora = OracleSolver(Yp, Yj, Yx, g.pv, g.pq)
    v0 = (lambda k: V0_c) if V0_c.ndim == 1 else (lambda k: V0_c[:, :k])
    probe = min(T, 256)
    t0 = time.perf_counter()
    ora.solve_batch(S_c[:, :probe], v0(probe), MAX_ITER, TOL, 0)
    rate = probe / max(time.perf_counter() - t0, 1e-6)
    ns = int(min(T, max(probe, rate * args.cpu_seconds)))
    idx = np.unique(np.linspace(0, T - 1, ns).round().astype(int))  # spread over the year
    ns = idx.size
    log(f"[bench] cpu oracle sample: {ns} cases on {cores} threads, best of 3 after a warm-up")
    Ss = np.ascontiguousarray(S_c[:, idx])
    V0s = V0_c if V0_c.ndim == 1 else np.ascontiguousarray(V0_c[:, idx])
    t_cpu, ref = time_best_of(lambda: ora.solve_batch(Ss, V0s, MAX_ITER, TOL, 0))
    # one thread, a smaller sample (a few seconds)
    n1 = int(min(ns, max(8, rate / max(cores, 1) * 2.0)))
    S1 = np.ascontiguousarray(Ss[:, :n1])
    V01 = V0s if V0s.ndim == 1 else np.ascontiguousarray(V0s[:, :n1])
    t_one, r1 = time_best_of(lambda: ora.solve_batch(S1, V01, MAX_ITER, TOL, 1), reps=1)
    out["cpu_baseline"] = {
        "value": round(int(ref["converged"].sum()) / t_cpu, 1), "unit": "cases/s", "cores": cores, "kind": "port",
        "sample": f"{ns} of the {T} cases spread evenly over the set, solve call only, best of 3 after a "
                  f"warm-up, {cores} host threads",
        "single_thread": {"value": round(int(r1["converged"].sum()) / t_one, 1), "cases": n1}}
    okm = ref["converged"] & conv[idx]
    dvm, dva = dv_max(V[:, idx], ref["V"], okm)
    out["parity"] = {"cases": ns, "max_dVm": dvm, "max_dVa": dva,
                     "converged_equal": bool(np.array_equal(ref["converged"], conv[idx])),
                     "iterations_equal": bool(np.array_equal(ref["iterations"], it[idx])),
                     "per_case_error_hist": dv_hist(V[:, idx], ref["V"], okm)}


def run_timeseries(args) -> dict:
    from p3smi import _lib, autotune, nr_klu

    g, S, V0, desc = timeseries_workload(args.config)
    Yp, Yj, Yx = g.csr()
    T = S.shape[1]
    log(f"[bench] workload {args.config}: n={g.n} cases={T}")
    t0 = time.perf_counter()
    solver = nr_klu.Solver(Yp, Yj, Yx, g.pv, g.pq)
    t_setup = time.perf_counter() - t0
    if args.config == "ts2000":
        from p3smi.timeseries import dc_initial_voltage

        V0 = dc_initial_voltage(g, solver=solver)
    if not autotune.apply_cached(solver, T):
        log("[bench] no cached LU shape for this size bucket: tuning (not timed)")
        autotune.tune(solver, S, V0, reps=1, log=log)
    info = solver.info
    log(f"[bench] plan: m={info['m']} nnzJ={info['nnz_j']} nnzLU={info['nnz_lu']} pairs={info['n_pairs']} "
        f"levels={info['n_levels']} ({t_setup:.2f}s)")

    L = _lib.lib()
    prec = {"double": 0, "mixed": 1, "mixed_switch": 2}[args.precision]
    bh = C.c_void_p()
    _lib.check(L.p3smi_batch_create(solver._h, T, C.byref(bh)))
    S_c = np.ascontiguousarray(S, dtype=np.complex128)
    V0_c = np.ascontiguousarray(V0, dtype=np.complex128)
    _lib.check(L.p3smi_batch_upload(bh, _lib.ptr(S_c), _lib.ptr(V0_c), int(V0_c.ndim == 2)))

    def solve():
        _lib.check(L.p3smi_batch_solve(bh, MAX_ITER, TOL, 1, prec))

    for w in range(args.warmup):
        solve()
        log(f"[bench] warmup {w + 1}/{args.warmup} done")
    _lib.check(L.p3smi_batch_set_timing(bh, 1))
    acc = stats_acc()
    times = []
    t_all0 = time.perf_counter()
    for k in range(args.steps):
        ts = time.perf_counter()
        solve()
        times.append(time.perf_counter() - ts)
        add_stats(acc, L, bh)
        log(f"[bench] step {k + 1}/{args.steps}: {1e3 * times[-1]:.1f} ms")
    t_total = time.perf_counter() - t_all0
    _lib.check(L.p3smi_batch_set_timing(bh, 0))
    if args.main_only:  # for profiling: the timed steps and nothing else
        bpc = np.empty(4, dtype=np.float64)
        L.p3smi_batch_bytes_per_case(bh, _lib.ptr(bpc))
        out = base_out(args, T * args.steps / t_total, times, t_total, args.precision,
                       {"workload": f"{args.config}: {desc}, timed steps only"})
        out["roofline"] = roofline(acc, bpc, args.steps, args.config, args.precision)
        L.p3smi_batch_free(bh)
        return out
    # the same steps with each Newton iteration replayed as a captured hipGraph (timers off)
    _lib.check(L.p3smi_batch_set_graphs(bh, 1))
    solve()
    tg0 = time.perf_counter()
    for k in range(args.steps):
        solve()
    t_graph = time.perf_counter() - tg0
    _lib.check(L.p3smi_batch_set_graphs(bh, 0))
    log(f"[bench] graph replay: {1e3 * t_graph / args.steps:.1f} ms per step")

    V = np.empty((g.n, T), dtype=np.complex128)
    it = np.empty(T, dtype=np.int32)
    cv = np.empty(T, dtype=np.uint8)
    mis = np.empty(T, dtype=np.float64)
    _lib.check(L.p3smi_batch_download(bh, _lib.ptr(V), _lib.ptr(it), _lib.ptr(cv), _lib.ptr(mis)))
    conv = cv.astype(bool)
    n_conv = int(conv.sum())
    bpc = np.empty(4, dtype=np.float64)
    L.p3smi_batch_bytes_per_case(bh, _lib.ptr(bpc))
    config = {"workload": f"{args.config}: {desc}", "n_bus": g.n, "cases": T, "precision": args.precision,
              "nnz_lu": int(info["nnz_lu"]), "n_levels": int(info["n_levels"]), "lu_shape": lu_shape(info)}
    out = base_out(args, n_conv * args.steps / t_total, times, t_total, args.precision, config)
    out.update({"converged": n_conv, "iterations_hist": np.bincount(it, minlength=1).tolist(),
                "newton_passes_per_step": acc["newton_passes"] / args.steps,
                "graph_replay": {"ms_per_step": round(1e3 * t_graph / args.steps, 3),
                                 "value": round(n_conv * args.steps / t_graph, 1)} if prec == 0 else None,
                "roofline": roofline(acc, bpc, args.steps, args.config, args.precision)})
    # the drop-in boundary takes host arrays: one solve_batch call with S and V0 from host
    # memory and V back to it (PCIe transfers and layout changes included), for reference
    solver.solve_batch(S_c, V0_c, max_iter=MAX_ITER, tol=TOL, precision=args.precision)
    th = time.perf_counter()
    rh = solver.solve_batch(S_c, V0_c, max_iter=MAX_ITER, tol=TOL, precision=args.precision)
    t_host = time.perf_counter() - th
    out["one_call_host_arrays"] = {"ms": round(1e3 * t_host, 3),
                                   "value": round(int(rh["converged"].sum()) / t_host, 1)}
    # the step before the path: injections of every case built on the device from element
    # series (loads p, q at PQ buses, generators p at PV buses), written into the batch
    from p3smi.timeseries import SbusBuilder

    rng = np.random.default_rng(9)
    pq, pv = g.pq, g.pv
    series = [g.p_load[pq][:, None] * rng.uniform(0.7, 1.3, (pq.size, T)),
              g.q_load[pq][:, None] * rng.uniform(0.7, 1.3, (pq.size, T)),
              g.p_gen[pv][:, None] * rng.uniform(0.0, 1.0, (pv.size, T))]
    sb = SbusBuilder(g.n, g.sbus(), [{"bus": pq, "coef": -np.ones(pq.size), "base": g.p_load[pq], "imag": False},
                                     {"bus": pq, "coef": -np.ones(pq.size), "base": g.q_load[pq], "imag": True},
                                     {"bus": pv, "coef": np.ones(pv.size), "base": g.p_gen[pv], "imag": False}])
    sb.build(bh, series)
    ts = time.perf_counter()
    sb.build(bh, series)
    out["sbus_build_device_ms"] = round(1e3 * (time.perf_counter() - ts), 3)
    if not args.no_cpu:
        cpu_timeseries(args, g, S_c, V0_c, V, it, conv, out)
    L.p3smi_batch_free(bh)
    return out


# ====================================================================== N-1
def run_n1(args) -> dict:
    from p3smi import _lib, autotune, contingency, grids, nr_klu
    from p3smi.timeseries import dc_initial_voltage

    g = grids.config_grid("n1_10k", 0)
    Yp, Yj, Yx = g.csr()
    t0 = time.perf_counter()
    solver = nr_klu.Solver(Yp, Yj, Yx, g.pv, g.pq)
    autotune.apply_cached(solver, g.nbr)
    cs = contingency.from_grid(g, reslack=True, solver=solver)
    t_setup = time.perf_counter() - t0
    info = solver.info
    Lc = cs.L
    log(f"[bench] workload n1_10k: n={g.n} outages={Lc} nnzLU={info['nnz_lu']} levels={info['n_levels']} "
        f"({t_setup:.2f}s)")
    S = np.ascontiguousarray(g.sbus(), dtype=np.complex128)
    Vs = np.ascontiguousarray(dc_initial_voltage(g, solver=solver), dtype=np.complex128)  # on the device
    td = time.perf_counter()  # a second solve: the first one also loads the kernels
    dc_initial_voltage(g, solver=solver)
    t_dc = time.perf_counter() - td
    L = _lib.lib()
    it = np.empty(Lc, dtype=np.int32)
    cv = np.empty(Lc, dtype=np.uint8)
    mis = np.empty(Lc, dtype=np.float64)

    def solve(prec, V=None, served=None):
        _lib.check(L.p3smi_outages_solve(cs._h, _lib.ptr(S), _lib.ptr(Vs), MAX_ITER, TOL, 1, prec,
                                         _lib.ptr(V) if V is not None else None, _lib.ptr(it), _lib.ptr(cv),
                                         _lib.ptr(mis), _lib.ptr(served) if served is not None else None))

    main = args.precision
    other = "double" if main != "double" else "mixed"
    code = {"double": 0, "mixed": 1, "mixed_switch": 2}
    # the first solve allocates the resident batch whose kernel timers the steps read
    args.warmup = max(1, args.warmup)
    for w in range(args.warmup):
        solve(code[main])
        if not args.main_only:
            solve(code[other])
        log(f"[bench] warmup {w + 1}/{args.warmup} done")
    bh = C.c_void_p()
    _lib.check(L.p3smi_topology_last_batch(solver._h, C.byref(bh)))
    _lib.check(L.p3smi_batch_set_timing(bh, 1))

    def timed(prec):
        acc = stats_acc()
        times = []
        t_all0 = time.perf_counter()
        for k in range(args.steps):
            ts = time.perf_counter()
            solve(code[prec])
            times.append(time.perf_counter() - ts)
            add_stats(acc, L, bh)
            log(f"[bench] {prec} step {k + 1}/{args.steps}: {1e3 * times[-1]:.1f} ms")
        return acc, times, time.perf_counter() - t_all0, int(cv.astype(bool).sum()), it.copy()

    acc, times, t_total, n_conv, it_main = timed(main)
    bpc = np.empty(4, dtype=np.float64)
    L.p3smi_batch_bytes_per_case(bh, _lib.ptr(bpc))
    if args.main_only:  # for profiling: the main precision's timed steps and nothing else
        _lib.check(L.p3smi_batch_set_timing(bh, 0))
        out = base_out(args, n_conv * args.steps / t_total, times, t_total, main,
                       {"workload": f"n1_10k: {Lc} single-branch outages, main precision only"})
        out["roofline"] = roofline(acc, bpc, args.steps, "n1_10k", main)
        return out
    acc2, times2, t_total2, n_conv2, it_other = timed(other)
    third = ({"double", "mixed", "mixed_switch"} - {main, other}).pop()
    solve(code[third])  # warm-up of the third precision
    acc3, _, t_total3, n_conv3, it_third = timed(third)
    _lib.check(L.p3smi_batch_set_timing(bh, 0))

    # ratings for the post-processing screen: 125 % of each branch's loading in the intact
    # grid (an outage set with one empty group, solved and post-processed on the device)
    vn = np.full(g.nbr, 20.0)
    base = contingency.from_grid(g, groups=[np.zeros(0, np.int32)], reslack=True, solver=solver)
    base.solve(S, Vs, max_iter=MAX_ITER, tol=TOL)
    bl = contingency.find_overloads(base, 1e300, vn, vn, np.ones(g.nbr), cap=0).max_loading_percent
    rating = np.where(np.isfinite(bl) & (bl > 0), bl / 100.0 * 1.25, 1.0)
    del base
    # untimed: the full result table of the main precision (V, served) for the checks
    V = np.empty((g.n, Lc), dtype=np.complex128)
    served = np.empty((g.n, Lc), dtype=np.uint8)
    solve(code[main], V, served)
    conv = cv.astype(bool)
    # N-1 post-processing on the resident results: loadings of every (branch, case), the
    # overload list (above 125 % of the intact loading) with each branch's worst case, and
    # the voltage-band violations
    tp = time.perf_counter()
    rep = contingency.find_overloads(cs, 100.0, vn, vn, rating, cap=1 << 22)
    t_over = time.perf_counter() - tp
    tp = time.perf_counter()
    vv = contingency.find_voltage_violations(cs, 0.95, 1.05, cap=1 << 22)
    t_vv = time.perf_counter() - tp
    log(f"[bench] post-processing: {len(rep)} overloads in {1e3 * t_over:.1f} ms, {len(vv)} voltage violations "
        f"in {1e3 * t_vv:.1f} ms")
    config = {"workload": f"n1_10k: 10000-bus synthetic grid, every single-branch outage ({Lc} cases), re-slack, "
                          f"DC start (on the device), tol 1e-8, max_iter 20",
              "n_bus": g.n, "cases": Lc, "precision": main, "nnz_lu": int(info["nnz_lu"]),
              "n_levels": int(info["n_levels"]), "lu_shape": lu_shape(info)}
    out = base_out(args, n_conv * args.steps / t_total, times, t_total, main, config)
    value2 = n_conv2 * args.steps / t_total2
    out.update({
        "converged": n_conv,
        "iterations_hist": np.bincount(it_main, minlength=1).tolist(),
        "cases_with_unserved_buses": int((~served.astype(bool)).any(axis=0).sum()),
        "newton_passes_per_step": acc["newton_passes"] / args.steps,
        "roofline": roofline(acc, bpc, args.steps, "n1_10k", main),
        "compare": {"precision": other, "value": round(value2, 1), "ms_per_step": round(1e3 * t_total2 / args.steps, 3),
                    "converged": n_conv2, "iterations_equal": bool(np.array_equal(it_main, it_other)),
                    "kernel_ms_per_step": {k: round(acc2[f"ms_{k}"] / args.steps, 3)
                                           for k in ("eval", "lu", "refine", "step", "decide", "other")}},
        "compare_2": {"precision": third, "value": round(n_conv3 * args.steps / t_total3, 1),
                      "ms_per_step": round(1e3 * t_total3 / args.steps, 3), "converged": n_conv3,
                      "iterations_equal": bool(np.array_equal(it_third, it_main)),
                      "kernel_ms_per_step": {k: round(acc3[f"ms_{k}"] / args.steps, 3)
                                             for k in ("eval", "lu", "refine", "step", "decide", "other")}},
        "dc_start_device_ms": round(1e3 * t_dc, 3),
        "mixed_over_fp64": round((out["value"] / value2) if main == "mixed" else (value2 / out["value"]), 3),
        "mixed_switch_over_fp64": (round((n_conv3 * args.steps / t_total3) / out["value"], 3)
                                   if main == "double" and third == "mixed_switch" else None),
        "post_processing": {"branch_case_pairs": int(g.nbr) * int(Lc), "overloads": len(rep),
                            "overload_screen_ms": round(1e3 * t_over, 3), "voltage_violations": len(vv),
                            "voltage_screen_ms": round(1e3 * t_vv, 3)},
    })
    out["one_outage_at_a_time_fp64"] = one_at_a_time_n1(cs, S, Vs)
    if not args.no_cpu:
        cpu_n1(args, g, cs, Vs, V, it_main, conv, served.astype(bool), out)
    return out


def one_at_a_time_n1(cs, S, Vs, k=32):
    """The plain FP64 GPU form without batching: the first k outages solved one per call
    (each a one-case outage set, built outside the timed region)."""
    from p3smi.contingency import ContingencySet

    subs = [ContingencySet(cs.solver, cs.f, cs.t, *cs.yval.T, [cs.groups[j]], cs.ref, cs.gen_z, cs.reslack)
            for j in range(k)]
    subs[0].solve(S, Vs, max_iter=MAX_ITER, tol=TOL)  # warm-up
    t0 = time.perf_counter()
    conv = sum(int(sub.solve(S, Vs, max_iter=MAX_ITER, tol=TOL).converged.sum()) for sub in subs)
    wall = time.perf_counter() - t0
    return {"sample": k, "value": round(conv / wall, 1), "ms_per_case": round(1e3 * wall / k, 3), "converged": conv}


def cpu_n1(args, g, cs, Vs, V, it, conv, served, out):
    """Oracle on the first k outage cases: islands and per-case Ybus restated on the host, then
    the FP64 restatement of solve_batch_contingency timed on all host threads."""
    from oracle import islands
    from oracle.oracle import OracleSolver
    from p3smi.timeseries import usable_cpus

    Yp, Yj, Yx = g.csr()
    cores = usable_cpus()
    ora = OracleSolver(Yp, Yj, Yx, g.pv, g.pq)
    gen_bus = np.nonzero(np.isfinite(g.gen_z))[0]

    def inputs(idx):
        groups = [cs.groups[i] for i in idx]
        srv, _, pinned = islands.cases(g.n, g.f, g.t, cs.live, groups, g.ref, gen_bus, g.gen_z[gen_bus], True)
        V0, pin = islands.solver_inputs(Vs, srv, pinned)
        Yc = np.stack([islands.case_ybus(cs.solver.Yx, cs.ypos, cs.yval, gr) for gr in groups], axis=1)
        return (np.ascontiguousarray(np.abs(Yc)), np.ascontiguousarray(np.angle(Yc)), np.ascontiguousarray(V0),
                np.ascontiguousarray(pin), srv)

    S = g.sbus()
    probe_idx = np.arange(min(cs.L, 2 * cores))
    mag, ang, V0, pin, _ = inputs(probe_idx)
    t0 = time.perf_counter()
    ora.solve_batch_contingency(mag, ang, S, V0, pin, MAX_ITER, TOL, 0)
    rate = probe_idx.size / max(time.perf_counter() - t0, 1e-6)
    ns = int(min(cs.L, 2048, max(probe_idx.size, rate * args.cpu_seconds)))
    idx = np.unique(np.linspace(0, cs.L - 1, ns).round().astype(int))  # spread over the outage set
    ns = idx.size
    log(f"[bench] cpu oracle sample: {ns} outage cases on {cores} threads, best of 3 after a warm-up")
    mag, ang, V0, pin, srv = inputs(idx)
    t_cpu, ref = time_best_of(lambda: ora.solve_batch_contingency(mag, ang, S, V0, pin, MAX_ITER, TOL, 0))
    n1 = int(min(ns, max(4, rate / max(cores, 1) * 2.0)))
    sl = slice(0, n1)
    t_one, r1 = time_best_of(lambda: ora.solve_batch_contingency(
        np.ascontiguousarray(mag[:, sl]), np.ascontiguousarray(ang[:, sl]), S, np.ascontiguousarray(V0[:, sl]),
        np.ascontiguousarray(pin[:, sl]), MAX_ITER, TOL, 1), reps=1)
    out["cpu_baseline"] = {"value": round(int(ref["converged"].sum()) / t_cpu, 1), "unit": "cases/s", "cores": cores,
                           "kind": "port",
                           "sample": f"{ns} of the {cs.L} outage cases spread evenly over the set, "
                                     f"solve_batch_contingency call only, best of 3 after a warm-up, "
                                     f"{cores} host threads",
                           "single_thread": {"value": round(int(r1["converged"].sum()) / t_one, 1), "cases": n1}}
    Vr = ref["V"].copy()
    Vr[~srv] = np.nan
    okm = ref["converged"] & conv[idx]
    dvm, dva = dv_max(V[:, idx], Vr, okm)
    out["parity"] = {"cases": ns, "max_dVm": dvm, "max_dVa": dva,
                     "served_equal": bool(np.array_equal(srv, served[:, idx])),
                     "converged_equal": bool(np.array_equal(ref["converged"], conv[idx])),
                     "iterations_equal": bool(np.array_equal(ref["iterations"], it[idx])),
                     "per_case_error_hist": dv_hist(V[:, idx], Vr, okm)}


# ====================================================================== many grids
def _counts(rng, k, total, lo, hi):
    """k batch sizes, log-uniform in [lo, hi], rescaled to sum to `total` (each kept in range)."""
    c = np.exp(rng.uniform(np.log(lo), np.log(hi), k))
    c = np.clip(np.round(c * total / c.sum()), lo, hi).astype(int)
    while c.sum() != total:
        d = total - c.sum()
        i = int(np.argmax(c)) if d < 0 else int(np.argmin(c))
        c[i] = int(np.clip(c[i] + np.sign(d) * min(abs(d), hi), lo, hi))
    return c


def _run_requests(srv, jobs, arrivals=None):
    """Submit jobs (at their arrival offsets, seconds, if given), wait for all. Returns the
    requests, the wall time from the first submission to the last completion. A request's
    latency runs from its scheduled arrival to its completion."""
    reqs = []
    t0 = time.perf_counter()
    for k, job in enumerate(jobs):
        if arrivals is not None:
            dt = t0 + arrivals[k] - time.perf_counter()
            if dt > 0:
                time.sleep(dt)
        if job["kind"] == "ts":
            reqs.append(srv.submit(*job["topo"], job["S"], job["V0"]))
        else:
            reqs.append(srv.submit_contingency(job["key"], job["cases"], job["S"], job["V0"]))
        if arrivals is not None:  # latency counts from the scheduled arrival: a late submit is the server's
            r = reqs[-1]
            r.submit_lag = time.perf_counter() - (t0 + arrivals[k])
            r.t_submit = min(r.t_submit, t0 + arrivals[k])
    for r in reqs:
        r.wait()
    return reqs, max(r.t_done for r in reqs) - t0


def _latency(reqs):
    lat = np.array([r.latency for r in reqs])
    return {"p50": round(1e3 * float(np.percentile(lat, 50)), 3), "p99": round(1e3 * float(np.percentile(lat, 99)), 3)}


def _latency_by_size(reqs, jobs, fam, edges=(0, 1000, 10000, 10 ** 9)):
    """p50 latency (ms) of the requests per grid-size band (buses)."""
    n = np.array([fam[j["gid"]].n for j in jobs])
    lat = np.array([r.latency for r in reqs])
    out = {}
    for lo, hi in zip(edges[:-1], edges[1:]):
        sel = (n >= lo) & (n < hi)
        if sel.any():
            out[(lo, hi)] = round(1e3 * float(np.percentile(lat[sel], 50)), 1)
    return out


def _sequential(jobs, solvers, n1sets, arrivals=None):
    """The plain FP64 loop: one request at a time on the same GPU, in arrival order. Latency
    = queueing (if arrivals are given) + service, with the measured service times."""
    from p3smi import autotune

    t_free = 0.0
    lat, conv, cases = [], 0, 0
    t0 = time.perf_counter()
    for k, job in enumerate(jobs):
        ts = time.perf_counter()
        if job["kind"] == "ts":
            sol = solvers[job["gid"]]
            autotune.apply_cached(sol, job["S"].shape[1])
            r = sol.solve_batch(job["S"], job["V0"], max_iter=MAX_ITER, tol=TOL)
            c = r["converged"]
        else:
            from p3smi.contingency import ContingencySet

            full = n1sets[job["gid"]]
            sub = ContingencySet(full.solver, full.f, full.t, *full.yval.T, [full.groups[i] for i in job["cases"]],
                                 full.ref, full.gen_z, full.reslack)
            c = sub.solve(job["S"], job["V0"], max_iter=MAX_ITER, tol=TOL).converged
        svc = time.perf_counter() - ts
        arr = arrivals[k] if arrivals is not None else 0.0
        start = max(arr, t_free)
        t_free = start + svc
        lat.append(t_free - arr)
        conv += int(c.sum())
        cases += int(c.size)
    wall = time.perf_counter() - t0 if arrivals is None else t_free - (arrivals[0] if len(arrivals) else 0.0)
    return {"value": round(conv / wall, 1), "wall_s": round(wall, 3), "converged": conv, "cases": cases,
            "latency_ms": {"p50": round(1e3 * float(np.percentile(lat, 50)), 3),
                           "p99": round(1e3 * float(np.percentile(lat, 99)), 3)}}


def run_multi48(args) -> dict:
    """48 grids, sizes log-uniform in 100..20,000 buses, 4,096 load-scaled cases in batches
    of 1..256 per grid, all submitted at once to the continuous-batching server: every grid
    gets a slot pool sized to its batch, the pools iterate concurrently on their own streams."""
    from p3smi import grids, nr_klu
    from p3smi.serving import Server
    from p3smi.timeseries import dc_initial_voltage

    rng = np.random.default_rng(4)
    fam = grids.grid_family(48, 100, 20000, seed=4)
    counts = _counts(rng, 48, 4096, 1, 256)
    jobs, solvers = [], []
    for k, (g, b) in enumerate(zip(fam, counts)):
        Yp, Yj, Yx = g.csr()
        jobs.append({"kind": "ts", "gid": k, "topo": (Yp, Yj, Yx, g.pv, g.pq),
                     "S": np.ascontiguousarray(grids.timeseries_scaled(g, int(b), seed=100 + k, sigma=0.1)),
                     "V0": np.ascontiguousarray(dc_initial_voltage(g))})
        solvers.append(nr_klu.Solver(Yp, Yj, Yx, g.pv, g.pq))
    log(f"[bench] multi48: sizes {min(g.n for g in fam)}..{max(g.n for g in fam)} buses, {int(counts.sum())} cases")
    grp = run_group(args, jobs, solvers)
    res = {}
    other_prec = "mixed" if args.precision == "double" else "double"
    for prec, prio in ((args.precision, True), (other_prec, True), (args.precision, False)):
        srv = Server(tol=TOL, max_iter=MAX_ITER, precision=prec, max_slots=256, priority=prio)
        for w in range(args.warmup):
            _run_requests(srv, jobs)
            log(f"[bench] {prec} warmup {w + 1}/{args.warmup} done")
        walls, lats, conv = [], [], 0
        for k in range(args.steps):
            reqs, wall = _run_requests(srv, jobs)
            walls.append(wall)
            lats += [r.latency for r in reqs]
            conv = sum(int(r.converged.sum()) for r in reqs)
            log(f"[bench] {prec} {'shortest work first' if prio else 'all at once'} step {k + 1}/{args.steps}: "
                f"{1e3 * wall:.1f} ms")
        srv.close()
        res[prec if prio else "all_at_once"] = {"value": conv * len(walls) / sum(walls), "walls": walls, "conv": conv,
                     "lat": {"p50": round(1e3 * float(np.percentile(lats, 50)), 3),
                             "p99": round(1e3 * float(np.percentile(lats, 99)), 3)}}
    seq = _sequential(jobs, solvers, None)
    main, other, fair = res[args.precision], res[other_prec], res["all_at_once"]
    config = {"workload": "multi48: 48 synthetic grids, sizes log-uniform 100-20000 buses, 4096 load-scaled cases "
                          "(batches of 1-256 per grid, N(1, 0.1) load scaling), DC start, all submitted at once to "
                          "the continuous-batching server", "grids": 48, "cases": int(counts.sum()),
              "precision": args.precision}
    config["engine"] = "batch group: every grid's cases in one launch per kernel per pass, the passes planned on " \
                       f"the device and captured {grp['passes_per_graph']} to a graph launch"
    out = base_out(args, grp["value"], grp["walls"], sum(grp["walls"]), "double", config)
    out.update({"converged": grp["conv"], "group": {k: v for k, v in grp.items() if k not in ("walls",)},
                "server": {"precision": args.precision, "value": round(main["value"], 1), "latency_ms": main["lat"],
                           "converged": main["conv"]},
                "server_compare": {"precision": other_prec, "value": round(other["value"], 1),
                                   "latency_ms": other["lat"]},
                "server_all_workers_at_once": {"precision": args.precision, "value": round(fair["value"], 1),
                                               "latency_ms": fair["lat"]},
                "sequential_fp64_one_at_a_time": seq,
                "group_over_sequential": round(grp["value"] / max(seq["value"], 1e-9), 3)})
    return out


def run_group(args, jobs, solvers) -> dict:
    """All jobs in one batch group (FP64): inputs resident, each timed step one group solve of
    every case. Every job's result is checked bit-identical to its batch solved alone."""
    from p3smi import autotune
    from p3smi.group import Group

    for job, s in zip(jobs, solvers):
        autotune.apply_cached(s, job["S"].shape[1])
    K = 8
    grp = Group([(s, job["S"], job["V0"]) for job, s in zip(jobs, solvers)])
    try:
        for w in range(args.warmup):
            grp.solve(max_iter=MAX_ITER, tol=TOL, passes_per_graph=K)
        walls, res = [], None
        for k in range(args.steps):
            t0 = time.perf_counter()
            res = grp.solve(max_iter=MAX_ITER, tol=TOL, passes_per_graph=K)
            walls.append(time.perf_counter() - t0)
            log(f"[bench] group step {k + 1}/{args.steps}: {1e3 * walls[-1]:.1f} ms")
        passes = grp.passes()
    finally:
        grp.close()
    same = True
    for job, s, r in zip(jobs, solvers, res):
        a = s.solve_batch(job["S"], job["V0"], max_iter=MAX_ITER, tol=TOL)
        same &= bool(np.array_equal(a["iterations"], r["iterations"]) and np.array_equal(a["converged"], r["converged"])
                     and np.array_equal(a["V"].view(np.float64), r["V"].view(np.float64)))
    conv = sum(int(r["converged"].sum()) for r in res)
    return {"value": round(conv * len(walls) / sum(walls), 1), "walls": walls, "conv": conv,
            "ms_per_solve": round(1e3 * sum(walls) / len(walls), 3), "passes_per_graph": K, "passes": passes,
            "bit_identical_to_alone": same}


def _serve64_runner(args, fam, topos, starts, jobs, arrivals):
    """serve(priority, **server options) -> (requests, wall s, outage sets): one Server over the
    64 topologies, its topology cache warmed, the jobs submitted at their arrival offsets."""
    from p3smi import contingency, grids
    from p3smi.serving import Server, topology_key

    for j in jobs:
        if j["kind"] == "n1":
            j["key"] = topology_key(*topos[j["gid"]])
    warm = [{"kind": "ts", "gid": i, "topo": topos[i], "V0": starts[i],
             "S": np.ascontiguousarray(grids.timeseries_scaled(fam[i], 8, seed=7, sigma=0.1))} for i in range(len(fam))]

    def serve(priority, **kw):
        srv = Server(tol=TOL, max_iter=MAX_ITER, precision=args.precision, max_slots=1024, priority=priority, **kw)
        n1sets = {}
        for gid, g in enumerate(fam):
            key = srv.register_outages(*topos[gid], g.f, g.t, g.branch_stamps(),
                                       contingency.single_branch_outages(g.nbr), g.ref, g.gen_z, True)
            n1sets[gid] = srv.outage_sets[key]
        _run_requests(srv, warm)  # warm the topology cache: one small request per grid
        reqs, wall = _run_requests(srv, jobs, arrivals)
        srv.close()
        lat = _latency(reqs)
        log(f"[bench] served ({'shortest work first' if priority else 'all workers at once'}): {1e3 * wall:.1f} ms, "
            f"p50 {lat['p50']:.1f} ms, p99 {lat['p99']:.1f} ms; p50 by grid size: "
            + ", ".join(f"{lo}-{hi} buses {v}" for (lo, hi), v in _latency_by_size(reqs, jobs, fam).items()))
        return reqs, wall, n1sets

    return serve


def _serve64_jobs(args):
    """The 64 topologies, their solvers and DC starts, and the jobs with their arrival offsets."""
    from p3smi import grids, nr_klu
    from p3smi.timeseries import dc_initial_voltage

    rng = np.random.default_rng(5)
    fam = grids.grid_family(64, 100, 50000, seed=5)
    solvers, starts, topos = [], [], []
    for g in fam:
        Yp, Yj, Yx = g.csr()
        topos.append((Yp, Yj, Yx, g.pv, g.pq))
        solvers.append(nr_klu.Solver(Yp, Yj, Yx, g.pv, g.pq))
        starts.append(np.ascontiguousarray(dc_initial_voltage(g)))
    nreq = args.requests
    arrivals = np.cumsum(rng.exponential(1.0 / args.rate, nreq))
    arrivals -= arrivals[0]
    jobs = []
    for k in range(nreq):
        gid = int(rng.integers(0, len(fam)))
        g = fam[gid]
        b = int(np.clip(np.round(np.exp(rng.uniform(0.0, np.log(1024)))), 1, 1024))
        if rng.random() < 0.3:
            b = min(b, g.nbr)
            jobs.append({"kind": "n1", "gid": gid, "cases": np.sort(rng.choice(g.nbr, b, replace=False)),
                         "S": np.ascontiguousarray(g.sbus()), "V0": starts[gid]})
        else:
            jobs.append({"kind": "ts", "gid": gid, "topo": topos[gid], "V0": starts[gid],
                         "S": np.ascontiguousarray(grids.timeseries_scaled(g, b, seed=1000 + k, sigma=0.1))})
    total = sum(j["S"].shape[1] if j["kind"] == "ts" else j["cases"].size for j in jobs)
    log(f"[bench] serve64: {nreq} requests at {args.rate} req/s, {total} cases, "
        f"{sum(j['kind'] == 'n1' for j in jobs)} N-1 subsets")
    return fam, topos, starts, solvers, jobs, arrivals, total


def run_serve64(args) -> dict:
    """Poisson arrivals of (grid, batch) jobs over 64 cached topologies (100..50,000 buses,
    log-uniform), batch sizes log-uniform 1..1,024, 70 % time-series and 30 % N-1 subsets, at
    `--rate` requests/s, served shortest work first; beside it the same server with every worker
    running at once, and the plain FP64 one-request-at-a-time loop."""
    fam, topos, starts, solvers, jobs, arrivals, total = _serve64_jobs(args)
    nreq = len(jobs)
    serve = _serve64_runner(args, fam, topos, starts, jobs, arrivals)
    reqs, wall, n1sets = serve(True)
    conv = sum(int(r.converged.sum()) for r in reqs)
    fair, fair_wall, _ = serve(False)
    fair_conv = sum(int(r.converged.sum()) for r in fair)
    same = all(np.array_equal(a.V.view(np.float64), b.V.view(np.float64), equal_nan=True)
               and np.array_equal(a.iterations, b.iterations) and np.array_equal(a.converged, b.converged)
               for a, b in zip(reqs, fair))
    seq = _sequential(jobs, solvers, n1sets, arrivals)
    config = {"workload": f"serve64: Poisson arrivals at {args.rate} req/s, {nreq} requests over 64 cached synthetic "
                          "topologies (100-50000 buses, log-uniform), batches log-uniform 1-1024 cases, 70% time "
                          "series / 30% N-1 subsets, DC start", "requests": nreq, "cases": int(total),
              "precision": args.precision}
    out = base_out(args, conv / wall, [r.latency for r in reqs], wall * args.steps, args.precision, config)
    out["latency_ms"] = _latency(reqs)
    out["steps"] = 1
    out["ms_per_step"] = round(1e3 * wall, 3)
    out["latency_ms_by_grid_size_p50"] = {f"{lo}-{hi}": v for (lo, hi), v in _latency_by_size(reqs, jobs, fam).items()}
    out["submit_lag_ms_max"] = round(1e3 * max(getattr(r, "submit_lag", 0.0) for r in reqs), 2)
    out.update({"converged": conv,
                "all_workers_at_once": {"value": round(fair_conv / fair_wall, 1), "wall_s": round(fair_wall, 3),
                                        "latency_ms": _latency(fair), "converged": fair_conv,
                                        "results_bit_identical": bool(same)},
                "sequential_fp64_one_at_a_time": seq})
    return out


def calibrate() -> int:
    """The HBM reference patterns of the library, 1 GiB read and 1 GiB written per launch,
    three timed launches each after a warm-up: a linear copy and a copy of 512-byte rows
    between permuted positions. Under rocprofv3 --pmc their FETCH_SIZE and WRITE_SIZE per
    dispatch against the known 1 GiB give the scale of the counters for each access pattern."""
    from p3smi import _lib

    n = 1 << 30
    for pattern, name in ((0, "linear copy"), (1, "512-byte rows, permuted")):
        ms = _lib.hbm_probe(pattern, n, 3)
        print(json.dumps({"calibration": name, "bytes_read": n, "bytes_written": n, "ms_per_launch": round(ms, 4),
                          "GB_per_s": round(2 * n / (ms / 1e3) / 1e9, 1)}), flush=True)
    return 0


def main() -> int:
    ap = argparse.ArgumentParser(description=__doc__)
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--config", default="n1_10k", choices=["n1_10k", "ts2000", "pf200", "multi48", "serve64"])
    ap.add_argument("--rate", type=float, default=200.0, help="serve64: mean arrival rate (requests/s)")
    ap.add_argument("--requests", type=int, default=300, help="serve64: number of requests")
    ap.add_argument("--precision", default="double", choices=["double", "mixed", "mixed_switch"],
                    help="precision of the line's value (n1_10k times the other two beside it)")
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU oracle baseline")
    ap.add_argument("--main-only", action="store_true",
                    help="time the main solve only (no compare legs, graph replay, screens or CPU baseline)")
    ap.add_argument("--cpu-seconds", type=float, default=4.0,
                    help="target duration of one timed CPU baseline pass (best of 3 after a warm-up)")
    ap.add_argument("--calibrate", action="store_true",
                    help="only run the HBM reference patterns (scale of the traffic counters)")
    args = ap.parse_args()
    if args.calibrate:
        return calibrate()
    if args.gpus != 1:
        print(f"bench.py: this engine runs on one GPU; --gpus {args.gpus} is not supported (use --gpus 1)",
              file=sys.stderr)
        return 2

    from p3smi import _lib

    if _lib.device_count() < 1:
        print("bench.py: no HIP device visible", file=sys.stderr)
        return 1
    run = {"n1_10k": run_n1, "multi48": run_multi48, "serve64": run_serve64}.get(args.config, run_timeseries)
    out = run(args)
    print(json.dumps(out), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
