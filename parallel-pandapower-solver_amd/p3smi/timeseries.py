"""Start voltages and host helpers around the batched solve.

``dc_initial_voltage`` restates the reference's DC-power-flow start
(p3s/timeseries.py:200-244 with the reduced DC solve of
p3s/NewtonPowerflowCpp.py:422-440): solve B[pvpq, pvpq] theta = P[pvpq] -
B[pvpq, ref] theta_ref with the DC susceptance matrix of the branches (diagonal
+sum 1/x, off-diagonal -1/x, p3s/models/TransmissionLineModel.py:146-148), keep the start
magnitudes at PV and PQ buses and take the DC angles; fall back to the flat start when
the DC angles are not finite. Given a solver, the solve runs on the device
(``Solver.dc_start``: the engine's own LU on the DC matrix); otherwise with scipy on the host.
"""

from __future__ import annotations

import os

import numpy as np
import scipy.sparse as sp
from scipy.sparse.linalg import spsolve


def dc_susceptance(n: int, f, t, x) -> sp.csr_matrix:
    """DC susceptance matrix B' of the branch list (p.u.)."""
    f = np.asarray(f)
    t = np.asarray(t)
    b = 1.0 / np.asarray(x, dtype=np.float64)
    rows = np.concatenate([f, t, f, t])
    cols = np.concatenate([f, t, t, f])
    vals = np.concatenate([b, b, -b, -b])
    return sp.coo_matrix((vals, (rows, cols)), shape=(n, n)).tocsr()


def dc_initial_voltage_arrays(Bdc, P, V_flat, ref, pv, pq) -> np.ndarray:
    """DC-initialised start voltage from B', the active injections P and the flat start."""
    flat = np.array(V_flat, dtype=np.complex128)
    voltage = flat.copy()
    ref = np.asarray(ref, dtype=np.int64)
    pv = np.asarray(pv, dtype=np.int64)
    pq = np.asarray(pq, dtype=np.int64)
    pvpq = np.concatenate([pv, pq])
    pv_vm = np.abs(voltage[pv])
    pq_vm = np.abs(voltage[pq])
    Bdc = sp.csr_matrix(Bdc)
    theta_ref = np.angle(voltage[ref])
    rhs = np.asarray(P, dtype=np.float64)[pvpq] - Bdc[pvpq][:, ref] @ theta_ref
    va = np.real(spsolve(Bdc[pvpq][:, pvpq].tocsc(), rhs))
    np.nan_to_num(va, copy=False)
    voltage[pvpq] = np.exp(1j * va)
    if pv.size:
        voltage[pv] = pv_vm * np.exp(1j * np.angle(voltage[pv]))
    if pq.size:
        voltage[pq] = pq_vm * np.exp(1j * np.angle(voltage[pq]))
    if not np.isfinite(np.abs(np.degrees(np.angle(voltage))).max()):
        return flat
    return voltage


def dc_initial_voltage(grid, Sbus=None, solver=None) -> np.ndarray:
    """DC start voltage of a synthetic grid (base-case injections unless ``Sbus`` given): on
    the device through ``solver`` (an ``nr_klu.Solver`` of the grid), else on the host."""
    S = grid.sbus() if Sbus is None else np.asarray(Sbus)
    Bdc = dc_susceptance(grid.n, grid.f, grid.t, grid.x)
    if solver is not None:
        return solver.dc_start(Bdc, S.real, grid.flat_start(), grid.ref)
    return dc_initial_voltage_arrays(Bdc, S.real, grid.flat_start(), grid.ref, grid.pv, grid.pq)


def usable_cpus() -> int:
    """CPUs this process may run on (p3s/timeseries.py:144-153)."""
    aff = getattr(os, "sched_getaffinity", None)
    if aff is not None:
        return len(aff(0))
    return os.cpu_count() or 1


class SbusBuilder:
    """Per-bus, per-case injections from element time series, built on the device.

    The reference's ``build_sbus_matrix`` (p3s/timeseries.py:36-141) starts from the base
    injection s0 and adds, for every time-series row of a load / sgen / gen variable, the
    delta coef * (value - base) to the element's bus (loads -1/sn, generation +1/sn, times
    scaling and in-service; q rows to the imaginary part). Its contribution list is ordered
    bus-major with the element order kept within a bus; this class forms the same list from
    arrays and hands it, with the stacked series, to ``p3smi_batch_build_sbus``, which writes
    S (n, T) straight into a resident batch (no host copy of the (n, T) matrix).

    sources: list of dicts, one per series (element, variable): ``bus`` (rows,) int, ``coef``
    (rows,), ``base`` (rows,), ``imag`` bool; rows with bus < 0 (unsupplied) are dropped.
    """

    def __init__(self, n_bus: int, s0, sources):
        self.n = int(n_bus)
        self.s0 = np.ascontiguousarray(s0, dtype=np.complex128)
        if self.s0.shape != (self.n,):
            raise ValueError("SbusBuilder: s0 must be (n_bus,)")
        bus, row, coef, base, imag = [], [], [], [], []
        self.offsets = []
        off = 0
        for src in sources:
            b = np.asarray(src["bus"], dtype=np.int64)
            keep = np.nonzero(b >= 0)[0]
            if np.any(b[keep] >= self.n):
                raise ValueError("SbusBuilder: element bus out of range")
            bus.append(b[keep])
            row.append(off + keep)
            coef.append(np.asarray(src["coef"], dtype=np.float64)[keep])
            base.append(np.asarray(src["base"], dtype=np.float64)[keep])
            imag.append(np.full(keep.size, bool(src["imag"])))
            self.offsets.append(off)
            off += b.size
        self.nrows = off
        bus_a = np.concatenate(bus) if bus else np.zeros(0, np.int64)
        order = np.argsort(bus_a, kind="stable")  # bus-major, element order kept within a bus
        self.ptr = np.zeros(self.n + 1, dtype=np.int32)
        self.ptr[1:] = np.cumsum(np.bincount(bus_a, minlength=self.n))
        self.row = np.ascontiguousarray(np.concatenate(row)[order], dtype=np.int32) if bus else np.zeros(0, np.int32)
        self.coef = np.ascontiguousarray(np.concatenate(coef)[order]) if bus else np.zeros(0)
        self.base = np.ascontiguousarray(np.concatenate(base)[order]) if bus else np.zeros(0)
        self.imag = np.ascontiguousarray(np.concatenate(imag)[order], dtype=np.uint8) if bus else np.zeros(0, np.uint8)

    def stack(self, series) -> np.ndarray:
        """The series (one (rows, T) array per source, in source order) as one (rows, T) array."""
        X = np.ascontiguousarray(np.concatenate([np.asarray(x, dtype=np.float64) for x in series], axis=0))
        if X.shape[0] != self.nrows:
            raise ValueError("SbusBuilder: series rows do not match the sources")
        return X

    def blocks(self, series) -> list:
        """The series as float64 C-contiguous (rows, T) arrays (copies only where the caller's
        are not), rows adding up to the sources' rows."""
        X = [np.ascontiguousarray(x, dtype=np.float64) for x in series]
        if any(x.ndim != 2 for x in X) or sum(x.shape[0] for x in X) != self.nrows:
            raise ValueError("SbusBuilder: series rows do not match the sources")
        if len({x.shape[1] for x in X}) > 1:
            raise ValueError("SbusBuilder: series differ in length")
        return X

    def build(self, batch_handle, series) -> None:
        """S of the batch (n, cases) from the series, on the device. Every series holds as
        many columns as the batch has cases. The series go to the device one by one (no
        stacked host copy)."""
        import ctypes as C

        from . import _lib

        X = self.blocks(series)
        cols = X[0].shape[1] if X else 0
        rows = np.array([x.shape[0] for x in X], dtype=np.int32)
        blk = (C.c_void_p * max(len(X), 1))(*[x.ctypes.data for x in X])
        L = _lib.lib()
        _lib.check(L.p3smi_batch_build_sbus_blocks(batch_handle, _lib.ptr(self.s0), _lib.ptr(self.ptr),
                                                   int(self.row.size), _lib.ptr(self.row), _lib.ptr(self.coef),
                                                   _lib.ptr(self.base), _lib.ptr(self.imag), int(cols), len(X),
                                                   _lib.ptr(rows),
                                                   C.cast(blk, C.c_void_p)))
