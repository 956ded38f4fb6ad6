"""Generate the small golden Newton fixtures (tests/golden/newton_small.npz).

The expected voltages come from a textbook full-step polar Newton written with scipy
(dS/dVm, dS/dVa as in the reference's own scipy check, p3s/cpp/tests/test_batch.py:32-70,
solved with a dense LU), NOT from the reference code and not from this project's solvers.

Grids:
  four_bus  -- the hand-built 4-bus meshed grid of the reference's wheel smoke test
               (p3s/cpp/ci/smoke_test.py:29-45, 104-109): lines 0-1, 0-2, 1-2, 2-3 and a
               shunt at bus 3; bus 0 slack, pv = [1], pq = [2, 3].
  geo30     -- a 30-bus synthetic grid of p3smi.grids (seed 11), 8 load-scaled cases.

Run:  python tests/golden/make_golden.py
"""

from __future__ import annotations

import os
import sys

import numpy as np
import scipy.sparse as sp

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, os.path.join(ROOT, "parallel-pandapower-solver_amd"))


def four_bus():
    n = 4
    Y = np.zeros((n, n), dtype=complex)
    for i, j, z, b in [(0, 1, 0.01 + 0.08j, 0.02), (0, 2, 0.02 + 0.10j, 0.03), (1, 2, 0.015 + 0.09j, 0.02),
                       (2, 3, 0.01 + 0.05j, 0.01)]:
        y = 1 / z
        Y[i, i] += y + 0.5j * b
        Y[j, j] += y + 0.5j * b
        Y[i, j] -= y
        Y[j, i] -= y
    Y[3, 3] += 0.05j
    pv = np.array([1])
    pq = np.array([2, 3])
    S = np.array([0.0, 0.4 + 0.0j, -0.6 - 0.2j, -0.3 - 0.1j])
    V0 = np.array([1.02, 1.01, 1.0, 1.0], dtype=complex)
    return Y, pv, pq, S, V0


def newton_dense(Y, S, V0, pv, pq, tol=1e-10, max_iter=30):
    """Full-step polar Newton with a dense solve; returns (V, iterations, converged, norms)."""
    Y = np.asarray(Y)
    pvpq = np.r_[pv, pq]
    V = V0.astype(complex).copy()

    def F(V):
        mis = V * np.conj(Y @ V) - S
        return np.r_[mis[pvpq].real, mis[pq].imag]

    f = F(V)
    norms = [np.max(np.abs(f))]
    it = 0
    while it < max_iter:
        if norms[-1] < tol:
            return V, it, True, norms
        it += 1
        Ibus = Y @ V
        dV = np.diag(V)
        dVn = np.diag(V / np.abs(V))
        dS_dVm = dV @ np.conj(Y @ dVn) + np.conj(np.diag(Ibus)) @ dVn
        dS_dVa = 1j * dV @ np.conj(np.diag(Ibus) - Y @ dV)
        J = np.block([[dS_dVa[np.ix_(pvpq, pvpq)].real, dS_dVm[np.ix_(pvpq, pq)].real],
                      [dS_dVa[np.ix_(pq, pvpq)].imag, dS_dVm[np.ix_(pq, pq)].imag]])
        dx = -np.linalg.solve(J, f)
        Va = np.angle(V)
        Vm = np.abs(V)
        Va[pvpq] += dx[: len(pvpq)]
        Vm[pq] += dx[len(pvpq):]
        V = Vm * np.exp(1j * Va)
        f = F(V)
        norms.append(np.max(np.abs(f)))
    return V, it, norms[-1] < tol, norms


def main():
    from p3smi import grids

    out = {}
    Y, pv, pq, S, V0 = four_bus()
    V, it, ok, _ = newton_dense(Y, S, V0, pv, pq, tol=1e-10)
    assert ok
    out.update(four_Y=Y, four_pv=pv, four_pq=pq, four_S=S, four_V0=V0, four_V=V, four_it=it)
    # batch of 16 load scalings, PV generator fixed (p3s/cpp/ci/smoke_test.py:122-125)
    scale = np.linspace(0.5, 1.3, 16)
    Sb = np.outer(S, scale)
    Sb[1, :] = S[1]
    Vb, itb = [], []
    for k in range(16):
        v, i, ok, _ = newton_dense(Y, Sb[:, k], V0, pv, pq, tol=1e-10)
        assert ok
        Vb.append(v)
        itb.append(i)
    out.update(four_Sb=Sb, four_Vb=np.stack(Vb, axis=1), four_itb=np.array(itb))

    g = grids.make_grid(30, 9, 4, seed=11, power_scale=0.05)
    Yg = g.ybus().toarray()
    Sg = grids.timeseries_scaled(g, 8, seed=5)
    V0g = g.flat_start()
    Vg, itg = [], []
    for k in range(8):
        v, i, ok, _ = newton_dense(Yg, Sg[:, k], V0g, g.pv, g.pq, tol=1e-8, max_iter=20)
        assert ok
        Vg.append(v)
        itg.append(i)
    out.update(geo30_Y=Yg, geo30_pv=g.pv, geo30_pq=g.pq, geo30_S=Sg, geo30_V0=V0g, geo30_V=np.stack(Vg, axis=1),
               geo30_it=np.array(itg))
    np.savez_compressed(os.path.join(HERE, "newton_small.npz"), **out)
    print("wrote", os.path.join(HERE, "newton_small.npz"), {k: np.shape(v) for k, v in out.items()})


if __name__ == "__main__":
    main()
