"""Uneven batching: batches of several topologies solved together on the device.

The reference solves one topology's batch per call (``nr_klu.Solver.solve_batch``,
p3s/cpp/nr_klu.cpp:1049-1166). ``solve_group`` takes many (solver, Sbus, V0) jobs -- grids of
any size, any number of cases each -- and runs all their Newton iterations at once: every pass
is one launch per kernel over all jobs (per LU launch shape and evaluation variant), the passes
are decided on the device and captured ``passes_per_graph`` to a graph launch. Each job's result
is bit-identical to ``solver.solve_batch`` of that job alone (FP64).
"""

from __future__ import annotations

import ctypes as C

import numpy as np

from . import _lib
from ._lib import check, lib, ptr
from .nr_klu import Solver, _c128


class Group:
    """Resident batches of several topologies, solved together (``p3smi_group_*``)."""

    def __init__(self, jobs):
        """jobs: sequence of (Solver, Sbus (n, T) complex, V0 (n,) or (n, T) complex)."""
        L = lib()
        self.jobs = []
        self._batches = []
        self._h = None
        try:
            for solver, S, V0 in jobs:
                if not isinstance(solver, Solver):
                    raise TypeError("group: every job needs an nr_klu.Solver")
                S = _c128(S)
                if S.ndim != 2 or S.shape[0] != solver.n or S.shape[1] == 0:
                    raise RuntimeError("group: Sbus must be (n_bus, T) with T > 0")
                V0 = _c128(V0)
                per_col = V0.ndim == 2
                if (per_col and V0.shape != S.shape) or (not per_col and V0.shape != (solver.n,)):
                    raise RuntimeError("group: V0 must be (n,) or (n, T)")
                bh = C.c_void_p()
                check(L.p3smi_batch_create(solver._h, int(S.shape[1]), C.byref(bh)), "group: ")
                self._batches.append(bh)
                check(L.p3smi_batch_upload(bh, ptr(S), ptr(V0), int(per_col)), "group: ")
                self.jobs.append((solver, int(S.shape[1])))
            arr = (C.c_void_p * len(self._batches))(*[b.value for b in self._batches])
            h = C.c_void_p()
            check(L.p3smi_group_create(arr, len(self._batches), C.byref(h)), "group: ")
            self._h = h
        except Exception:
            self.close()
            raise

    def solve(self, max_iter: int = 30, tol: float = 1e-8, line_search: bool = True, passes_per_graph: int = 8):
        """Solve every job; returns a list of {V (n, T), iterations, converged, mismatch}."""
        L = lib()
        check(L.p3smi_group_solve(self._h, int(max_iter), float(tol), int(bool(line_search)), 0,
                                  int(passes_per_graph)), "group solve: ")
        out = []
        for (solver, T), bh in zip(self.jobs, self._batches):
            V = np.empty((solver.n, T), dtype=np.complex128)
            it = np.empty(T, dtype=np.int32)
            cv = np.empty(T, dtype=np.uint8)
            mis = np.empty(T, dtype=np.float64)
            check(L.p3smi_batch_download(bh, ptr(V), ptr(it), ptr(cv), ptr(mis)), "group: ")
            out.append({"V": V, "iterations": it, "converged": cv.astype(bool), "mismatch": mis})
        return out

    def passes(self) -> dict:
        """Passes of the last solve: on the device, on the host side (pivoting), graph launches."""
        o = np.zeros(3, dtype=np.int64)
        check(lib().p3smi_group_stats(self._h, ptr(o)))
        return {"device": int(o[0]), "host": int(o[1]), "graph_launches": int(o[2])}

    def stats(self) -> list:
        """Per job: the batch counters of the last solve (as ``Solver.last_stats``)."""
        res = []
        for bh in self._batches:
            st = _lib.Stats()
            check(lib().p3smi_batch_stats(bh, C.byref(st)))
            res.append({f: getattr(st, f) for f, _ in st._fields_})
        return res

    def close(self):
        L = lib()
        if self._h:
            L.p3smi_group_free(self._h)
            self._h = None
        for bh in self._batches:
            L.p3smi_batch_free(bh)
        self._batches = []

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def solve_group(jobs, max_iter: int = 30, tol: float = 1e-8, line_search: bool = True, passes_per_graph: int = 8):
    """Solve (solver, Sbus, V0) jobs of any topologies together; one result dict per job."""
    g = Group(jobs)
    try:
        return g.solve(max_iter, tol, line_search, passes_per_graph)
    finally:
        g.close()
