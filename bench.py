#!/usr/bin/env python3
"""Throughput benchmark: batched polar Newton power flow on one MI355X.

Workload (default ``ts2000``): the 2,000-bus synthetic grid (spanning tree + 600 chords,
200 PV buses, one slack), 8,760 hourly load cases (daily sinusoid x N(1, 0.05) per load,
PV output x U(0, 1)), DC-power-flow start, tol 1e-8 p.u., max 20 iterations, FP64.

A step is one complete batched solve of all cases from the start voltage, with the
inputs already resident in device memory. The JSON line reports converged cases/s over
the timed steps, p50/p99 batch latency, the roofline position of the dominant kernel
(HIP events on the library's stream, algorithmic bytes per launch) and the CPU oracle
timed on the host cores over a bounded sample of the same workload.
"""

from __future__ import annotations

import argparse
import ctypes as C
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
for p in (os.path.join(ROOT, "parallel-pandapower-solver_amd"), ROOT):
    if p not in sys.path:
        sys.path.insert(0, p)

import numpy as np  # noqa: E402

HBM_PEAK_GBS = 8000.0


def log(msg: str) -> None:
    print(msg, file=sys.stderr, flush=True)


def workload(name: str):
    from p3smi import grids
    from p3smi.timeseries import dc_initial_voltage

    if name == "ts2000":
        g = grids.config_grid("ts2000", 0)
        S = grids.timeseries_daily(g, 8760, seed=1)
        V0 = dc_initial_voltage(g)
        desc = "2000-bus synthetic grid, 8760 hourly cases, DC start, tol 1e-8, max_iter 20"
    elif name == "pf200":
        g = grids.config_grid("pf200", 0)
        S = grids.timeseries_scaled(g, 96, seed=1)
        V0 = g.flat_start()
        desc = "200-bus synthetic grid, 96 load-scaled cases, flat start, tol 1e-8, max_iter 20"
    else:
        raise SystemExit(f"unknown workload {name!r}")
    return g, S, V0, desc


def main() -> int:
    ap = argparse.ArgumentParser(description=__doc__)
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--config", default="ts2000")
    ap.add_argument("--precision", default="double", choices=["double", "mixed"])
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU oracle baseline")
    ap.add_argument("--cpu-seconds", type=float, default=10.0, help="target CPU baseline sample duration")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py: this engine runs on one GPU; --gpus {args.gpus} is not supported (use --gpus 1)",
              file=sys.stderr)
        return 2

    from p3smi import _lib, nr_klu
    from p3smi.timeseries import usable_cpus

    if _lib.device_count() < 1:
        print("bench.py: no HIP device visible", file=sys.stderr)
        return 1
    g, S, V0, desc = workload(args.config)
    Yp, Yj, Yx = g.csr()
    T = S.shape[1]
    log(f"[bench] workload {args.config}: n={g.n} cases={T}")
    from p3smi import autotune

    t0 = time.perf_counter()
    solver = nr_klu.Solver(Yp, Yj, Yx, g.pv, g.pq)
    t_setup = time.perf_counter() - t0
    if not autotune.apply_cached(solver):
        log("[bench] no cached LU shape for this size bucket: tuning (not timed)")
        autotune.tune(solver, S, V0, reps=1, log=log)
    info = solver.info
    log(f"[bench] plan: m={info['m']} nnzJ={info['nnz_j']} nnzLU={info['nnz_lu']} pairs={info['n_pairs']} "
        f"levels={info['n_levels']} ({t_setup:.2f}s)")

    L = _lib.lib()
    prec = {"double": 0, "mixed": 1}[args.precision]
    bh = C.c_void_p()
    _lib.check(L.p3smi_batch_create(solver._h, T, C.byref(bh)))
    S_c = np.ascontiguousarray(S, dtype=np.complex128)
    V0_c = np.ascontiguousarray(V0, dtype=np.complex128)
    _lib.check(L.p3smi_batch_upload(bh, _lib.ptr(S_c), _lib.ptr(V0_c), int(V0_c.ndim == 2)))

    def solve():
        _lib.check(L.p3smi_batch_solve(bh, 20, 1e-8, 1, prec))

    for w in range(args.warmup):
        solve()
        log(f"[bench] warmup {w + 1}/{args.warmup} done")

    _lib.check(L.p3smi_batch_set_timing(bh, 1))
    st = _lib.Stats()
    acc = {"ms_eval": 0.0, "ms_lu": 0.0, "ms_step": 0.0, "ms_decide": 0.0, "ms_other": 0.0, "n_eval": 0, "n_lu": 0,
           "n_step": 0, "n_decide": 0, "double_factorizations": 0, "newton_passes": 0, "line_search_passes": 0}
    times = []
    t_all0 = time.perf_counter()
    for k in range(args.steps):
        ts = time.perf_counter()
        solve()
        times.append(time.perf_counter() - ts)
        L.p3smi_batch_stats(bh, C.byref(st))
        for key in acc:
            acc[key] += getattr(st, key)
        log(f"[bench] step {k + 1}/{args.steps}: {1e3 * times[-1]:.1f} ms")
    t_total = time.perf_counter() - t_all0

    V = np.empty((g.n, T), dtype=np.complex128)
    it = np.empty(T, dtype=np.int32)
    cv = np.empty(T, dtype=np.uint8)
    mis = np.empty(T, dtype=np.float64)
    _lib.check(L.p3smi_batch_download(bh, _lib.ptr(V), _lib.ptr(it), _lib.ptr(cv), _lib.ptr(mis)))
    conv = cv.astype(bool)
    n_conv = int(conv.sum())
    value = n_conv * args.steps / t_total

    bpc = np.empty(4, dtype=np.float64)
    L.p3smi_batch_bytes_per_case(bh, _lib.ptr(bpc))
    # dominant kernel: the one with the largest total time
    kern = max(("lu", "eval"), key=lambda k: acc[f"ms_{k}"])
    if kern == "lu":
        cases_processed = acc["double_factorizations"]
        bytes_total = cases_processed * bpc[1]
        n_launch = acc["n_lu"]
    else:
        # eval passes touch the cases of their status mask; approximate with active cases per pass
        cases_processed = acc["double_factorizations"] + T * args.steps
        bytes_total = cases_processed * bpc[0]
        n_launch = acc["n_eval"]
    ms = acc[f"ms_{kern}"]
    achieved = bytes_total / (ms / 1e3) / 1e9 if ms > 0 else 0.0
    roofline = {
        "kernel": f"k_{kern}",
        "bound": "hbm",
        "achieved": round(achieved, 2),
        "peak": HBM_PEAK_GBS,
        "unit": "GB/s",
        "frac": round(achieved / HBM_PEAK_GBS, 4),
        "traffic": None,
        "avg_launch_ms": round(ms / max(n_launch, 1), 4),
        "bytes_per_case": float(bpc[1] if kern == "lu" else bpc[0]),
        "kernel_ms_per_step": {k: round(acc[f"ms_{k}"] / args.steps, 3) for k in ("eval", "lu", "step", "decide", "other")},
    }

    out = {
        "metric": "converged PF cases/sec at 1e-8 p.u. mismatch; p50/p99 solve latency",
        "value": round(value, 1),
        "unit": "cases/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(1e3 * t_total / args.steps, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64" if args.precision == "double" else "f32-lu+f64",
        "data": "synthetic",
        "config": {"workload": f"{args.config}: {desc}", "n_bus": g.n, "cases": T, "precision": args.precision,
                   "nnz_lu": int(info["nnz_lu"]), "n_levels": int(info["n_levels"]),
                   "lu_shape": [int(info["lu_g"]), int(info["lu_ns"]), int(info["lu_few_g"]), int(info["lu_few_ns"]),
                                int(info["lu_few_max"])]},
        "latency_ms": {"p50": round(1e3 * float(np.percentile(times, 50)), 3),
                       "p99": round(1e3 * float(np.percentile(times, 99)), 3)},
        "converged": n_conv,
        "iterations_hist": np.bincount(it, minlength=1).tolist(),
        "newton_passes_per_step": acc["newton_passes"] / args.steps,
        "roofline": roofline,
    }

    if not args.no_cpu:
        from oracle.oracle import OracleSolver

        cores = usable_cpus()
        ora = OracleSolver(Yp, Yj, Yx, g.pv, g.pq)
        probe = min(T, 256)
        t0 = time.perf_counter()
        ora.solve_batch(S_c[:, :probe], V0_c if V0_c.ndim == 1 else V0_c[:, :probe], 20, 1e-8, 0)
        rate = probe / max(time.perf_counter() - t0, 1e-6)
        ns = int(min(T, max(probe, rate * args.cpu_seconds)))
        log(f"[bench] cpu oracle sample: {ns} cases on {cores} threads")
        Ss = np.ascontiguousarray(S_c[:, :ns])
        t0 = time.perf_counter()
        ref = ora.solve_batch(Ss, V0_c if V0_c.ndim == 1 else V0_c[:, :ns], 20, 1e-8, 0)
        t_cpu = time.perf_counter() - t0
        cpu_value = int(ref["converged"].sum()) / t_cpu
        out["cpu_baseline"] = {"value": round(cpu_value, 1), "unit": "cases/s", "cores": cores, "kind": "port",
                               "sample": f"first {ns} of the {T} cases, one pass, all {cores} host threads"}
        okm = ref["converged"] & conv[:ns]
        dvm = float(np.abs(np.abs(V[:, :ns][:, okm]) - np.abs(ref["V"][:, okm])).max(initial=0.0))
        dva = float(np.abs(np.angle(V[:, :ns][:, okm]) - np.angle(ref["V"][:, okm])).max(initial=0.0))
        out["parity"] = {"cases": ns, "max_dVm": dvm, "max_dVa": dva,
                         "converged_equal": bool(np.array_equal(ref["converged"], conv[:ns])),
                         "iterations_equal": bool(np.array_equal(ref["iterations"], it[:ns]))}
    L.p3smi_batch_free(bh)
    print(json.dumps(out), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
