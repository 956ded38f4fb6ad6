"""The 10,000-bus N-1 set (every single-branch outage, 13,499 cases, re-slack, DC start) solved
whole on the GPU in FP64 and in mixed precision.

Against the CPU oracle on a spread sample of 1,024 outages: served masks and reference buses
bit-exact (union-find vs scipy connected components), convergence flags and iteration counts
exact, max |dVm| <= 1e-9 p.u. and max |dVa| <= 1e-9 rad (the reference's own GPU test demands
the same of its GPU path against nr_klu, tests/cuda/test_lean_lu.py:126-147,
tests/contingency/test_gpu_solve.py:80-111). On all 13,499 cases, properties that need no
oracle: mismatch below tol wherever converged, NaN exactly at the unserved buses, and the
mixed-precision form with the same flags and iteration counts as FP64 and voltages within
1e-9 of it."""

import numpy as np
import pytest

from oracle import islands
from oracle.oracle import OracleSolver
from p3smi import autotune, contingency, grids, nr_klu
from p3smi.timeseries import dc_initial_voltage

pytestmark = pytest.mark.gpu

TOL_V = 1e-9
N_SAMPLE = 1024


@pytest.fixture(scope="module")
def n1():
    g = grids.config_grid("n1_10k", 0)
    Yp, Yj, Yx = g.csr()
    solver = nr_klu.Solver(Yp, Yj, Yx, g.pv, g.pq)
    autotune.apply_cached(solver, g.nbr)
    cs = contingency.from_grid(g, reslack=True, solver=solver)
    return g, cs, np.ascontiguousarray(g.sbus()), np.ascontiguousarray(dc_initial_voltage(g))


def _dv(Va, Vb, cols):
    a, b = Va[:, cols], Vb[:, cols]
    fin = np.isfinite(a) & np.isfinite(b)
    assert np.array_equal(np.isfinite(a), np.isfinite(b))
    return (float(np.abs(np.abs(a[fin]) - np.abs(b[fin])).max(initial=0.0)),
            float(np.abs(np.angle(a[fin]) - np.angle(b[fin])).max(initial=0.0)))


def _oracle(g, cs, idx, V_start):
    """Islands, start voltages, pins and per-case Ybus of the outages idx restated on the host,
    then the FP64 restatement of solve_batch_contingency."""
    groups = [cs.groups[i] for i in idx]
    gen_bus = np.nonzero(np.isfinite(g.gen_z))[0]
    srv, ref_bus, pinned = islands.cases(g.n, g.f, g.t, cs.live, groups, g.ref, gen_bus, g.gen_z[gen_bus], True)
    V0, pin = islands.solver_inputs(V_start, srv, pinned)
    Yc = np.stack([islands.case_ybus(cs.solver.Yx, cs.ypos, cs.yval, gr) for gr in groups], axis=1)
    Yp, Yj, Yx = g.csr()
    r = OracleSolver(Yp, Yj, Yx, g.pv, g.pq).solve_batch_contingency(
        np.ascontiguousarray(np.abs(Yc)), np.ascontiguousarray(np.angle(Yc)), g.sbus(), V0, pin, 20, 1e-8)
    V = r["V"].copy()
    V[~srv] = np.nan
    return V, r["iterations"], r["converged"], srv, ref_bus


def test_n1_10k_every_outage_fp64_and_mixed(gpu, n1):
    g, cs, S, Vs = n1
    L = cs.L
    assert L == g.nbr == 13499
    r64 = cs.solve(S, Vs, max_iter=20, tol=1e-8, precision="double")
    rmx = cs.solve(S, Vs, max_iter=20, tol=1e-8, precision="mixed")
    print(f"n1_10k: {int(r64.converged.sum())}/{L} converged, iterations "
          f"{dict(zip(*np.unique(r64.iterations, return_counts=True)))}, "
          f"{int((~r64.served).any(axis=0).sum())} cases with unserved buses")
    # every case: converged cases meet tol; NaN exactly at the unserved buses
    for r in (r64, rmx):
        assert (r.mismatch[r.converged] < 1e-8).all()
        np.testing.assert_array_equal(np.isnan(r.V[:, r.converged]), ~r.served[:, r.converged])
        assert np.isnan(r.V[~r.served]).all()
    np.testing.assert_array_equal(rmx.served, r64.served)
    # mixed precision: the FP64 iteration, counts included
    np.testing.assert_array_equal(rmx.converged, r64.converged)
    np.testing.assert_array_equal(rmx.iterations, r64.iterations)
    dvm, dva = _dv(rmx.V, r64.V, r64.converged)
    assert dvm <= TOL_V and dva <= TOL_V, (dvm, dva)

    # a spread sample against the oracle: islands, flags, counts and voltages
    idx = np.unique(np.linspace(0, L - 1, N_SAMPLE).round().astype(int))
    assert idx.size >= 1000
    Vo, ito, cvo, srvo, refo = _oracle(g, cs, idx, Vs)
    sub = contingency.ContingencySet(cs.solver, cs.f, cs.t, *cs.yval.T, [cs.groups[i] for i in idx], cs.ref,
                                     cs.gen_z, cs.reslack)
    srv_s, ref_s, _ = sub.islands()
    np.testing.assert_array_equal(srv_s, srvo)
    np.testing.assert_array_equal(ref_s, refo)
    np.testing.assert_array_equal(r64.served[:, idx], srvo)
    for name, r in (("double", r64), ("mixed", rmx)):
        np.testing.assert_array_equal(r.converged[idx], cvo, err_msg=name)
        np.testing.assert_array_equal(r.iterations[idx], ito, err_msg=name)
        dvm, dva = _dv(r.V[:, idx], Vo, cvo)
        print(f"{name}: sample of {idx.size} outages, max|dVm| {dvm:.2e}, max|dVa| {dva:.2e}")
        assert dvm <= TOL_V and dva <= TOL_V, (name, dvm, dva)
