"""Power-flow drivers above the Newton path.

* ``calculate``: one operating point with the seed sequence and voltage-band acceptance of
  ``NewtonPowerflowCpp.calculate`` (p3s/NewtonPowerflowCpp.py:553-584): try the primary
  start, then (if given) a flat start; accept the first solve that converged AND keeps every
  |V| in the band (default [0.5, 1.5] p.u.). A converged solve outside the band is a
  voltage-collapse branch, not an operating point.
* ``calculate_batch``: the same rule for many cases at once on the device. Every case runs
  from its primary start; the band is tested on the device (``p3smi_batch_vrange``); the
  cases that fail either test run again, together, from the flat start.
* ``calculate_timeseries``: ``calculate_timeseries_cpp`` (p3s/NewtonPowerflowCpp.py:586-676):
  validate the inputs, solve the batch, raise unless every case converged to finite V.

A ``PowerFlow`` object keeps the solver of its topology between calls and refreshes only
the admittance values when the pattern is unchanged, as the reference caches its
``nr_klu.Solver`` (p3s/NewtonPowerflowCpp.py:533-541, 650-656).
"""

from __future__ import annotations

import ctypes as C

import numpy as np

from . import _lib
from .nr_klu import Solver, _prec

DEFAULT_BAND = (0.5, 1.5)


class LoadflowNotConverged(RuntimeError):
    """No start produced a converged, in-band solution (pandapower's exception of that name)."""


def _csr(Yp, Yj, Yx):
    return (np.ascontiguousarray(Yp, dtype=np.int32), np.ascontiguousarray(Yj, dtype=np.int32),
            np.ascontiguousarray(Yx, dtype=np.complex128))


def _in_band(vmin, vmax, band) -> np.ndarray:
    if band is None:
        return np.ones(np.shape(vmin), dtype=bool)
    return (np.asarray(vmin) >= band[0]) & (np.asarray(vmax) <= band[1])


class PowerFlow:
    """Solver cache of one grid plus the drivers above (see the module docstring)."""

    def __init__(self, Yp, Yj, Yx, pv, pq, precision: str = "double"):
        self.solver = None
        self._nnz = None
        self.precision = precision
        self.set_grid(Yp, Yj, Yx, pv, pq)

    def set_grid(self, Yp, Yj, Yx, pv, pq):
        """Rebuild the solver if the pattern size changed, else refresh Ybus values only."""
        Yp, Yj, Yx = _csr(Yp, Yj, Yx)
        pv = np.ascontiguousarray(pv, dtype=np.int32)
        pq = np.ascontiguousarray(pq, dtype=np.int32)
        if self.solver is None or self._nnz != Yj.shape[0]:
            self.solver = Solver(Yp, Yj, Yx, pv, pq)
            self._nnz = Yj.shape[0]
            from . import autotune

            autotune.apply_cached(self.solver)
        else:
            self.solver.update_Y(Yx)
        self.n = int(Yp.shape[0] - 1)

    # ------------------------------------------------------------------ one case
    def calculate(self, Sbus, V0, V_flat=None, max_iterations: int = 30, tolerance: float = 1e-8,
                  line_search: bool = True, voltage_band=DEFAULT_BAND) -> np.ndarray:
        """V (n,) of the first start (V0, then V_flat if given) that converges in band."""
        Sbus = np.ascontiguousarray(Sbus, dtype=np.complex128)
        seeds = [np.ascontiguousarray(V0, dtype=np.complex128)]
        if V_flat is not None:
            seeds.append(np.ascontiguousarray(V_flat, dtype=np.complex128))
        res = None
        for seed in seeds:
            res = self.solver.solve(Sbus, seed, max_iter=max_iterations, tol=tolerance, line_search=line_search,
                                    precision=self.precision)
            vm = np.abs(res["V"])
            if res["converged"] and bool(_in_band(vm.min(), vm.max(), voltage_band)):
                return res["V"]
        if res is not None and res["converged"]:
            vm = np.abs(res["V"])
            raise LoadflowNotConverged(
                f"Loadflow converged to a non-physical solution: bus voltages in [{vm.min():.3f}, {vm.max():.3f}] pu "
                f"fall outside the accepted band {tuple(voltage_band)}. Likely a voltage-collapse / spurious branch.")
        raise LoadflowNotConverged(f"Loadflow did not converge in {max_iterations} iterations "
                                   f"(reached {res['iterations'] if res is not None else 0}).")

    # ------------------------------------------------------------------ many cases
    def _device_solve(self, S, V0, v0_per_case, max_iterations, tolerance, line_search):
        """Resident-batch solve: V (n, B), iterations, converged, mismatch, vmin, vmax."""
        from . import autotune

        L = _lib.lib()
        B = S.shape[1]
        autotune.apply_cached(self.solver, B)
        bh = C.c_void_p()
        _lib.check(L.p3smi_batch_create(self.solver._h, B, C.byref(bh)))
        try:
            _lib.check(L.p3smi_batch_upload(bh, _lib.ptr(S), _lib.ptr(V0), int(v0_per_case)))
            _lib.check(L.p3smi_batch_solve(bh, int(max_iterations), float(tolerance), int(bool(line_search)),
                                           _prec(self.precision)))
            V = np.empty((self.n, B), dtype=np.complex128)
            it = np.empty(B, dtype=np.int32)
            cv = np.empty(B, dtype=np.uint8)
            mis = np.empty(B, dtype=np.float64)
            vmin = np.empty(B, dtype=np.float64)
            vmax = np.empty(B, dtype=np.float64)
            _lib.check(L.p3smi_batch_vrange(bh, _lib.ptr(vmin), _lib.ptr(vmax)))
            _lib.check(L.p3smi_batch_download(bh, _lib.ptr(V), _lib.ptr(it), _lib.ptr(cv), _lib.ptr(mis)))
        finally:
            L.p3smi_batch_free(bh)
        return V, it, cv.astype(bool), mis, vmin, vmax

    def calculate_batch(self, Sbus, V0, V_flat=None, max_iterations: int = 30, tolerance: float = 1e-8,
                        line_search: bool = True, voltage_band=DEFAULT_BAND) -> dict:
        """Seed-and-band rule per case: dict(V (n, B), accepted (B,), seed (B,) 0 = V0, 1 = flat,
        iterations, converged, mismatch, vmin, vmax) of the solve each case ended with."""
        S = np.ascontiguousarray(Sbus, dtype=np.complex128)
        if S.ndim != 2 or S.shape[0] != self.n:
            raise ValueError(f"calculate_batch: Sbus must be ({self.n}, B)")
        V0 = np.ascontiguousarray(V0, dtype=np.complex128)
        per_case = V0.ndim == 2
        V, it, cv, mis, vmin, vmax = self._device_solve(S, V0, per_case, max_iterations, tolerance, line_search)
        ok = cv & _in_band(vmin, vmax, voltage_band)
        seed = np.zeros(S.shape[1], dtype=np.int8)
        retry = np.flatnonzero(~ok)
        if V_flat is not None and retry.size:
            Vf = np.ascontiguousarray(V_flat, dtype=np.complex128)
            Sr = np.ascontiguousarray(S[:, retry])
            V2, it2, cv2, mis2, lo2, hi2 = self._device_solve(Sr, Vf, False, max_iterations, tolerance, line_search)
            V[:, retry] = V2
            it[retry], cv[retry], mis[retry], vmin[retry], vmax[retry] = it2, cv2, mis2, lo2, hi2
            ok[retry] = cv2 & _in_band(lo2, hi2, voltage_band)
            seed[retry] = 1
        return {"V": V, "accepted": ok, "seed": seed, "iterations": it, "converged": cv, "mismatch": mis,
                "vmin": vmin, "vmax": vmax}

    # ------------------------------------------------------------------ time series
    def calculate_timeseries(self, Sbus, V_init, max_iterations: int = 30, tolerance: float = 1e-8,
                             n_threads: int = 0) -> np.ndarray:
        """V (n, T) of every time step, or an error as calculate_timeseries_cpp raises it."""
        S = np.asarray(Sbus)
        if S.ndim != 2:
            raise ValueError(f"Sbus shape mismatch: expected ({self.n}, T), got {S.shape}")
        n_bus, T = S.shape
        if n_bus == 0:
            raise ValueError("Network has no buses")
        if T == 0:
            raise ValueError("Time series has zero time steps")
        if np.any(np.isnan(S)) or np.any(np.isinf(S)):
            raise ValueError("Sbus matrix contains NaN or inf values")
        v = np.asarray(V_init)
        if np.any(np.isnan(v)) or np.any(np.isinf(v)):
            raise ValueError("DC initial voltage contains NaN or inf values")
        if n_bus != self.n:
            raise ValueError(f"Sbus shape mismatch: expected ({self.n}, {T}), got {S.shape}")
        if v.shape[0] != n_bus:
            raise ValueError(f"V0 shape mismatch: expected ({n_bus},), got {v.shape}")
        res = self.solver.solve_batch(np.ascontiguousarray(S, dtype=np.complex128),
                                      np.ascontiguousarray(v, dtype=np.complex128), max_iter=max_iterations,
                                      tol=tolerance, n_threads=n_threads, precision=self.precision)
        V = res["V"]
        if np.any(np.isnan(V)) or np.any(np.isinf(V)):
            raise ValueError("batch solver returned NaN or inf voltages")
        if not bool(np.all(res["converged"])):
            n_bad = int((~res["converged"]).sum())
            raise LoadflowNotConverged(
                f"batch did not converge for {n_bad} of {T} time steps in {max_iterations} iterations.")
        return V
