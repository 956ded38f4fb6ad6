"""Generate the small golden Newton fixtures (tests/golden/newton_small.npz).

The expected voltages come from a textbook full-step polar Newton written with scipy
(dS/dVm, dS/dVa as in the reference's own scipy check, p3s/cpp/tests/test_batch.py:32-70,
solved with a dense LU), NOT from the reference code and not from this project's solvers.

Grids:
  four_bus  -- the hand-built 4-bus meshed grid of the reference's wheel smoke test
               (p3s/cpp/ci/smoke_test.py:29-45, 104-109): lines 0-1, 0-2, 1-2, 2-3 and a
               shunt at bus 3; bus 0 slack, pv = [1], pq = [2, 3].
  geo30     -- a 30-bus synthetic grid of p3smi.grids (seed 11), 8 load-scaled cases.
  damped    -- the geo30 grid from 48 perturbed starts, with and without the Armijo line
               search, max_iter 20 and 5 (newton_armijo_dense restates the reference's
               documented damping rules on the dense solve; backtracking, best-trial
               commits and the iteration limit all occur).

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


def newton_armijo_dense(Y, S, V0, pv, pq, tol=1e-8, max_iter=20, line_search=True):
    """Damped polar Newton with a dense solve, following the loop the reference documents
    (p3s/cpp/nr_klu.cpp:588-708): stop if ||F|| < tol (converged) or non-finite; count the
    iteration; full step accepted iff ||F1|| <= (1 - 1e-4) ||F0||; else alpha = 0.5^k,
    k = 1..10, accept the first with ||F|| <= (1 - 1e-4 alpha) ||F0||, or commit the
    smallest-residual trial (the full step included) if none passes. The step of iteration
    max_iter is taken without a further convergence test. Returns (V, iterations, converged,
    iterations that backtracked, iterations that committed the best trial, largest condition
    number of the Jacobians solved)."""
    Y = np.asarray(Y)
    pvpq = np.r_[pv, pq]
    npvpq = len(pvpq)

    def F(Vm, Va):
        V = Vm * np.exp(1j * Va)
        mis = V * np.conj(Y @ V) - S
        return np.r_[mis[pvpq].real, mis[pq].imag]

    def norm(f):
        return np.max(np.abs(f)) if np.isfinite(np.sum(np.abs(f))) else np.finfo(float).max

    Vm, Va = np.abs(V0).astype(float), np.angle(V0).astype(float)
    nrm = norm(F(Vm, Va))
    it, converged, n_bt, n_best, kappa = 0, False, 0, 0, 0.0
    while it < max_iter:
        if nrm < tol:
            converged = True
            break
        if nrm >= np.finfo(float).max:
            break
        it += 1
        V = Vm * np.exp(1j * Va)
        Ibus = Y @ V
        dV = np.diag(V)
        dVn = np.diag(V / np.abs(V))
        dS_dVm = dV @ np.conj(Y @ dVn) + np.conj(np.diag(Ibus)) @ dVn
        dS_dVa = 1j * dV @ np.conj(np.diag(Ibus) - Y @ dV)
        J = np.block([[dS_dVa[np.ix_(pvpq, pvpq)].real, dS_dVm[np.ix_(pvpq, pq)].real],
                      [dS_dVa[np.ix_(pq, pvpq)].imag, dS_dVm[np.ix_(pq, pq)].imag]])
        kappa = max(kappa, float(np.linalg.cond(J)))
        try:
            dx = -np.linalg.solve(J, F(Vm, Va))
        except np.linalg.LinAlgError:
            break

        def trial(a):
            m2, a2 = Vm.copy(), Va.copy()
            a2[pvpq] += a * dx[:npvpq]
            m2[pq] += a * dx[npvpq:]
            return m2, a2

        m1, a1 = trial(1.0)
        n1 = norm(F(m1, a1))
        if not line_search or n1 <= (1.0 - 1e-4) * nrm:
            Vm, Va, nrm = m1, a1, n1
            continue
        alpha, best_alpha, best_nrm = 1.0, 1.0, n1
        n_bt += 1
        for k in range(1, 11):
            alpha *= 0.5
            mk, ak = trial(alpha)
            nk = norm(F(mk, ak))
            if nk < best_nrm:
                best_nrm, best_alpha = nk, alpha
            if nk <= (1.0 - 1e-4 * alpha) * nrm:
                break
            if k == 10:
                alpha = best_alpha
                n_best += 1
        Vm, Va = trial(alpha)
        nrm = norm(F(Vm, Va))
    return Vm * np.exp(1j * Va), it, converged, n_bt, n_best, kappa


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
    # damped Newton from perturbed starts (heavier loads, PQ magnitudes x0.5..1.5): the
    # backtracking branch and the best-trial commit occur; with max_iter 5 some cases end
    # at the iteration limit (not converged, after the step of iteration 5)
    rng = np.random.default_rng(21)
    T = 48
    k = rng.uniform(1.0, 3.0, T)
    Sd = g.p_gen[:, None] * k[None, :] - (g.p_load + 1j * g.q_load)[:, None] * k[None, :]
    V0d = np.repeat(V0g[:, None], T, axis=1)
    V0d[g.pq, :] *= rng.uniform(0.5, 1.5, T)[None, :]
    stats = {}
    for ls in (True, False):
        for mi in (20, 5):
            res = [newton_armijo_dense(Yg, Sd[:, c], V0d[:, c], g.pv, g.pq, tol=1e-8, max_iter=mi, line_search=ls)
                   for c in range(T)]
            tag = ("ls" if ls else "nols") + ("" if mi == 20 else "_lim")
            out[f"damped_{tag}_V"] = np.stack([r[0] for r in res], axis=1)
            out[f"damped_{tag}_it"] = np.array([r[1] for r in res])
            out[f"damped_{tag}_cv"] = np.array([r[2] for r in res])
            stats[tag] = (sum(r[2] for r in res), sum(r[3] > 0 for r in res), sum(r[4] > 0 for r in res))
    out.update(damped_S=Sd, damped_V0=V0d)
    # loads past the grid's transfer limit (x20..60): no solution; the line search finds no
    # sufficient decrease and commits its best trial (first at iterations 4-9). Kept: the
    # first 16 draws whose Jacobians stay below condition number 1e5 over 6 iterations, so
    # that any LU with threshold or partial pivoting gives the same steps to roundoff (on
    # worse iterates a threshold-pivoting LU, as KLU's, and a dense partial-pivoting solve
    # take visibly different steps). After 6 iterations the state agrees between solvers to
    # ~1e-12 (roundoff grows about tenfold per iteration here); after 20 only flags and
    # counts are compared.
    ko = []
    for kk in rng.uniform(20.0, 60.0, 64):
        Sk = g.p_gen * kk - (g.p_load + 1j * g.q_load) * kk
        if newton_armijo_dense(Yg, Sk, V0g, g.pv, g.pq, tol=1e-8, max_iter=6, line_search=True)[5] < 1e5:
            ko.append(kk)
        if len(ko) == 16:
            break
    ko = np.array(ko)
    So = g.p_gen[:, None] * ko[None, :] - (g.p_load + 1j * g.q_load)[:, None] * ko[None, :]
    out.update(overload_S=So)
    for mi in (6, 20):
        res = [newton_armijo_dense(Yg, So[:, c], V0g, g.pv, g.pq, tol=1e-8, max_iter=mi, line_search=True)
               for c in range(So.shape[1])]
        out.update({f"overload{mi}_it": np.array([r[1] for r in res]), f"overload{mi}_cv": np.array([r[2] for r in res]),
                    f"overload{mi}_V": np.stack([r[0] for r in res], axis=1)})
        stats[f"overload{mi}"] = (sum(r[2] for r in res), sum(r[3] > 0 for r in res), sum(r[4] > 0 for r in res))
    print("damped sets (converged, cases that backtracked, cases with a best-trial commit):", stats)
    np.savez_compressed(os.path.join(HERE, "newton_small.npz"), **out)
    print("wrote", os.path.join(HERE, "newton_small.npz"), {k: np.shape(v) for k, v in out.items()})


if __name__ == "__main__":
    main()
