"""Pick the LU launch shape and the evaluation's row tile per grid-size bucket and batch-size
class by timing candidates on the device.

The batched LU kernel has two shape parameters: g, the cases one workgroup factors
together, and ns, its wavefronts (so 64 * ns / g task lists per case). Large g keeps
memory accesses in long contiguous lines; small g spreads one factorization over more
task lists and finishes it sooner. The best pair depends on the factor pattern and on how
many cases there are: a batch of fewer than 8,192 cases gives g = 64 fewer than 128
workgroups for the 256 compute units, where more task lists per case pay. So the tuner
times every candidate on the real batch and caches the winner per size bucket (<= 500,
<= 5k, <= 20k, > 20k buses) and case class ("few" < 8,192 cases <= "many").
"""

from __future__ import annotations

import ctypes as C
import json
import os
import time

import numpy as np

from . import _lib

BATCH_CANDIDATES = [(64, 8), (64, 16), (32, 8), (32, 16), (16, 8), (16, 16), (1, 8), (1, 16)]
FEW_CANDIDATES = [(1, 4), (1, 8), (1, 16), (16, 8)]
# group 0 is one entry per task, except at g = 64 where it is the 2x2-block schedule;
# dealing runs of a pivot's entries to one list (group > 0) lost to single entries on every
# grid measured, so the tuner keeps to group 0 unless asked for more
GROUP_CANDIDATES = [0]
# rows per tile of the mismatch/Jacobian evaluation (a workgroup of 4 wavefronts per tile)
EVAL_TILE_CANDIDATES = [8, 16, 32, 64]
BUCKETS = [(500, "le500"), (5000, "le5k"), (20000, "le20k"), (10**9, "gt20k")]
CACHE = os.path.join(os.path.dirname(os.path.abspath(__file__)), "autotune.json")


MANY_CASES = 8192
FEW_PROBE = 1  # cases of the batch on which the few-case shapes are timed


def bucket(n_bus: int) -> str:
    for lim, name in BUCKETS:
        if n_bus <= lim:
            return name
    return BUCKETS[-1][1]


def key(n_bus: int, cases: int) -> str:
    return f"{bucket(n_bus)}:{'many' if cases >= MANY_CASES else 'few'}"


def _time_solve(solver, bh, reps: int) -> float:
    L = _lib.lib()
    best = float("inf")
    for _ in range(reps):
        t0 = time.perf_counter()
        _lib.check(L.p3smi_batch_solve(bh, 20, 1e-8, 1, 0))
        best = min(best, time.perf_counter() - t0)
    return best


def tune(solver, S, V0, reps: int = 2, log=None) -> dict:
    """Time the candidate shapes on (S, V0) and set the fastest on ``solver``."""
    L = _lib.lib()
    S = np.ascontiguousarray(S, dtype=np.complex128)
    V0 = np.ascontiguousarray(V0, dtype=np.complex128)
    bh = C.c_void_p()
    _lib.check(L.p3smi_batch_create(solver._h, S.shape[1], C.byref(bh)))
    try:
        _lib.check(L.p3smi_batch_upload(bh, _lib.ptr(S), _lib.ptr(V0), int(V0.ndim == 2)))
        results = {}
        for g, ns in BATCH_CANDIDATES:
            for grp in GROUP_CANDIDATES:
                solver.set_lu_shape(g, ns, 1, 16, 0, grp)  # batch shape for every pass
                _time_solve(solver, bh, 1)
                results[(g, ns, grp)] = _time_solve(solver, bh, reps)
                if log:
                    log(f"[autotune] batch g={g} ns={ns} group={grp}: {1e3 * results[(g, ns, grp)]:.2f} ms")
        bg, bns, bgrp = min(results, key=results.get)
        # the few-case shape runs single solves and the last iterating cases of a batch, both
        # latency bound: in a whole batch's time it is lost in the noise, so its (g, ns) is
        # timed on FEW_PROBE cases that it factors alone (one: a 16-case probe picked 16
        # wavefronts, which doubled a one-outage call on the 10k grid against 4), and only
        # the threshold on the whole batch
        k = min(FEW_PROBE, S.shape[1])
        probe = {}
        sh = C.c_void_p()
        _lib.check(L.p3smi_batch_create(solver._h, k, C.byref(sh)))
        try:
            Sk = np.ascontiguousarray(S[:, :k])
            V0k = np.ascontiguousarray(V0[:, :k]) if V0.ndim == 2 else V0
            _lib.check(L.p3smi_batch_upload(sh, _lib.ptr(Sk), _lib.ptr(V0k), int(V0k.ndim == 2)))
            for fg, fns in FEW_CANDIDATES:
                solver.set_lu_shape(bg, bns, fg, fns, 1 << 30, bgrp)
                _time_solve(solver, sh, 1)
                probe[(fg, fns)] = _time_solve(solver, sh, reps)
                if log:
                    log(f"[autotune] few g={fg} ns={fns} on {k} cases: {1e3 * probe[(fg, fns)]:.2f} ms")
        finally:
            L.p3smi_batch_free(sh)
        fg, fns = min(probe, key=probe.get)
        few = {}
        for fmax in (64, 256, 1024):
            solver.set_lu_shape(bg, bns, fg, fns, fmax, bgrp)
            _time_solve(solver, bh, 1)
            few[fmax] = _time_solve(solver, bh, reps)
            if log:
                log(f"[autotune] few g={fg} ns={fns} max={fmax}: {1e3 * few[fmax]:.2f} ms")
        # a threshold of 0 (no few-case launches) is not offered: the whole batch's time cannot
        # tell the thresholds apart, and without the few-case shape a single solve takes the
        # batch shape (a one-outage call on the 10k grid: 69 against 24 ms)
        fmax = min(few, key=few.get)
        solver.set_lu_shape(bg, bns, fg, fns, fmax, bgrp)
        tiles = {}
        for rows in EVAL_TILE_CANDIDATES:
            solver.set_eval_tile(rows)
            _time_solve(solver, bh, 1)
            tiles[rows] = _time_solve(solver, bh, reps)
            if log:
                log(f"[autotune] eval tile {rows} rows: {1e3 * tiles[rows]:.2f} ms")
        best_rows = min(tiles, key=tiles.get)
        solver.set_eval_tile(best_rows)
        return {"g": bg, "ns": bns, "group": bgrp, "few_g": fg, "few_ns": fns, "few_max": fmax,
                "eval_rows": best_rows, "ms": 1e3 * tiles[best_rows]}
    finally:
        L.p3smi_batch_free(bh)


def load_cache() -> dict:
    if os.path.exists(CACHE):
        with open(CACHE) as fh:
            return json.load(fh)
    return {}


def save_cache(entries: dict) -> None:
    with open(CACHE, "w") as fh:
        json.dump(entries, fh, indent=1, sort_keys=True)


def apply_cached(solver, cases: int = 1) -> bool:
    """Set the cached shape for the solver's size bucket and a batch of `cases` cases (the
    other case class of the bucket when only that one is cached), if there is one."""
    cache = load_cache()
    b = bucket(solver.n)
    e = cache.get(key(solver.n, cases)) or cache.get(f"{b}:few") or cache.get(f"{b}:many")
    if not e:
        return False
    solver.set_lu_shape(e["g"], e["ns"], e["few_g"], e["few_ns"], e["few_max"], e.get("group", 0))
    if "eval_rows" in e:
        solver.set_eval_tile(e["eval_rows"])
    return True


# Representative batches of each (size bucket, case class) for the synthetic grid family:
# (bucket key, grid size in buses, cases). The "many" class of the > 20k bucket is tuned at
# 2,048 cases (a 50k-bus case needs ~20 MB of workspace).
REPRESENTATIVE = [("le500:few", 200, 96), ("le5k:few", 2000, 512), ("le5k:many", 2000, 8760),
                  ("le20k:few", 10000, 256), ("le20k:many", 10000, 13499), ("gt20k:few", 40000, 64),
                  ("gt20k:many", 40000, 1024)]
CONFIG_OF = {200: "pf200", 2000: "ts2000", 10000: "n1_10k"}  # the benchmark grids of those sizes


def main(argv=None) -> int:
    """Re-tune every (size bucket, case class) on synthetic grids and write the cache:
    python -m p3smi.autotune [cache path]."""
    import sys

    from . import grids, nr_klu
    from .timeseries import dc_initial_voltage

    argv = sys.argv[1:] if argv is None else argv
    out = argv[0] if argv else CACHE
    entries = {}
    for name, n_bus, cases in REPRESENTATIVE:
        g = grids.config_grid(CONFIG_OF[n_bus], 0) if n_bus in CONFIG_OF else grids.sized_grid(n_bus, 0)
        Yp, Yj, Yx = g.csr()
        s = nr_klu.Solver(Yp, Yj, Yx, g.pv, g.pq)
        S = grids.timeseries_scaled(g, cases, seed=1, sigma=0.1)
        t0 = time.perf_counter()
        entries[name] = tune(s, S, dc_initial_voltage(g, solver=s), reps=2, log=lambda m: print(m, flush=True))
        print(f"[autotune] {name}: {g.n} buses, {cases} cases -> {entries[name]} "
              f"({time.perf_counter() - t0:.1f} s)", flush=True)
    with open(out, "w") as fh:
        json.dump(entries, fh, indent=1, sort_keys=True)
    return 0


if __name__ == "__main__":
    raise SystemExit(main())
