"""Start voltages and host helpers around the batched solve.

``dc_initial_voltage`` restates the reference's DC-power-flow start
(p3s/timeseries.py:200-244 with the reduced DC solve of
p3s/NewtonPowerflowCpp.py:422-440): solve B[pvpq, pvpq] theta = P[pvpq] -
B[pvpq, ref] theta_ref with the DC susceptance matrix of the branches (diagonal
+sum 1/x, off-diagonal -1/x, p3s/models/TransmissionLineModel.py:146-148), keep the start
magnitudes at PV and PQ buses and take the DC angles; fall back to the flat start when
the DC angles are not finite. It is a one-off host step per topology, like the reference's.
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


def dc_initial_voltage(grid, Sbus=None) -> np.ndarray:
    """DC start voltage of a synthetic grid (base-case injections unless ``Sbus`` given)."""
    S = grid.sbus() if Sbus is None else np.asarray(Sbus)
    Bdc = dc_susceptance(grid.n, grid.f, grid.t, grid.x)
    return dc_initial_voltage_arrays(Bdc, S.real, grid.flat_start(), grid.ref, grid.pv, grid.pq)


def usable_cpus() -> int:
    """CPUs this process may run on (p3s/timeseries.py:144-153)."""
    aff = getattr(os, "sched_getaffinity", None)
    if aff is not None:
        return len(aff(0))
    return os.cpu_count() or 1
