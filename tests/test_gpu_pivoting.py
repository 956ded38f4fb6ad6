"""The pivoting factorization behind the frozen plan, on the device, against the oracle.

nr_klu rejects the frozen pivot order when a refactorization's pivot ratio falls below
REUSE_RCOND_RATIO x the flat-start one and factors with KLU's partial pivoting for the rest of
the run (p3s/cpp/nr_klu.cpp:609-652); a column that does not converge on the frozen plan is run
again from its start with pivoting (:844-857). The reference's GPU test forces the fallback by
setting the ratio to 1e30 (tests/cuda/test_lean_lu.py:224-244). Here both the device engine and
the oracle take that setting, and their results must agree: convergence flags and iteration
counts exact, |dV| <= 1e-9; and the reruns of the non-converging cases happen on the same cases.

A fresh pivoting factorization runs in one of two ways, both checked here: the static-pivot LU
of the plan's order, kept when the threshold rule would take the diagonal at every column
(diagonal_first, the default), or the device Gilbert-Peierls LU choosing every pivot."""

import numpy as np
import pytest

from oracle import oracle as ora
from oracle.oracle import OracleSolver
from p3smi import grids, nr_klu
from p3smi.timeseries import dc_initial_voltage

pytestmark = pytest.mark.gpu

TOL_V = 1e-9


def _compare(r, ref, tag):
    np.testing.assert_array_equal(r["converged"], ref["converged"], err_msg=tag)
    np.testing.assert_array_equal(r["iterations"], ref["iterations"], err_msg=tag)
    ok = ref["converged"]
    dvm = np.abs(np.abs(r["V"][:, ok]) - np.abs(ref["V"][:, ok])).max(initial=0.0)
    dva = np.abs(np.angle(r["V"][:, ok]) - np.angle(ref["V"][:, ok])).max(initial=0.0)
    assert dvm <= TOL_V and dva <= TOL_V, (tag, dvm, dva)


@pytest.fixture
def forced_guard():
    ora.set_lean_guard(1e30)
    yield
    ora.set_lean_guard(1e-3)


@pytest.mark.parametrize("diagonal_first", [True, False])
@pytest.mark.parametrize("name", ["pf200", "ts2000"])
def test_every_factorization_pivots_when_the_frozen_plan_is_rejected(gpu, forced_guard, name, diagonal_first):
    g = grids.config_grid(name, 0)
    Yp, Yj, Yx = g.csr()
    if name == "pf200":
        S, V0 = grids.timeseries_scaled(g, 96, seed=1), g.flat_start()
    else:
        S, V0 = grids.timeseries_daily(g, 8760, seed=1)[:, ::73].copy(), dc_initial_voltage(g)
    s = nr_klu.Solver(Yp, Yj, Yx, g.pv, g.pq)
    s.set_guard(1e30)
    s.set_pivoting(diagonal_first)
    r = s.solve_batch(S, V0, max_iter=20, tol=1e-8)
    st = s.last_stats
    ora.reset_counters()
    ref = OracleSolver(Yp, Yj, Yx, g.pv, g.pq).solve_batch(S, V0, max_iter=20, tol=1e-8)
    oc = ora.counters()
    _compare(r, ref, name)
    assert r["converged"].all()
    # the first factorization of every case trips the guard and pivots; later ones refactor
    assert st["guard_trips"] == S.shape[1], st
    assert st["pivot_factorizations"] >= S.shape[1], st
    # the oracle's threshold pivoting took the diagonal at every column of these Jacobians, so the
    # device's first try stands for every case (and the Gilbert-Peierls path is not needed)
    assert oc["offdiagonal_pivots"] == 0, oc
    if diagonal_first:
        assert st["diagonal_factorizations"] == st["pivot_factorizations"] == S.shape[1], st
    else:
        assert st["diagonal_factorizations"] == 0, st
    # the same solve on the frozen plan: the same flags and iterations, voltages within the bar
    s.set_guard(1e-3)
    plain = s.solve_batch(S, V0, max_iter=20, tol=1e-8)
    assert s.last_stats["pivot_factorizations"] == 0
    _compare(r, plain, name + " frozen vs pivoting")


@pytest.mark.parametrize("diagonal_first", [True, False])
def test_reruns_of_non_converging_cases_match_the_oracle(gpu, diagonal_first):
    """Heavy loads from perturbed starts: some cases do not converge on the frozen plan; both the
    device and the oracle run each of them again with pivoting and report that run."""
    g = grids.config_grid("pf200", 0)
    Yp, Yj, Yx = g.csr()
    rng = np.random.default_rng(7)
    T = 256
    k = rng.uniform(1.0, 2.5, T)
    S = g.p_gen[:, None] * k[None, :] - (g.p_load + 1j * g.q_load)[:, None] * k[None, :]
    V0 = np.repeat(g.flat_start()[:, None], T, axis=1)
    V0[g.pq, :] *= rng.uniform(0.6, 1.4, T)[None, :]
    s = nr_klu.Solver(Yp, Yj, Yx, g.pv, g.pq)
    s.set_pivoting(diagonal_first)
    r = s.solve_batch(S, V0, max_iter=20, tol=1e-8)
    st = s.last_stats
    ora.reset_counters()
    ref = OracleSolver(Yp, Yj, Yx, g.pv, g.pq).solve_batch(S, V0, max_iter=20, tol=1e-8)
    oc = ora.counters()
    _compare(r, ref, "reruns")
    assert st["reruns"] == oc["column_reruns"] == int((~ref["converged"]).sum()), (st, oc)
    assert st["reruns"] > 0
    print(f"diagonal_first={diagonal_first}: {st['reruns']} reruns, {st['pivot_factorizations']} pivoting "
          f"factorizations ({st['diagonal_factorizations']} on the diagonal); oracle {oc}")
    if not diagonal_first:
        assert st["diagonal_factorizations"] == 0, st


@pytest.mark.parametrize("diagonal_first", [True, False])
def test_forced_pivoting_on_n1_outages(gpu, forced_guard, diagonal_first):
    """Per-case admittances and pinned buses on the pivoting path: every pf200 outage."""
    from p3smi import contingency
    from tests.test_gpu_contingency import _oracle_solve

    g = grids.config_grid("pf200", 0)
    cs = contingency.from_grid(g, reslack=True)
    cs.solver.set_guard(1e30)
    cs.solver.set_pivoting(diagonal_first)
    V_start = g.flat_start()
    r = cs.solve(g.sbus(), V_start, max_iter=20, tol=1e-8)
    V, it, cv, srv = _oracle_solve(cs, g, V_start)
    np.testing.assert_array_equal(r.served, srv)
    np.testing.assert_array_equal(r.converged, cv)
    np.testing.assert_array_equal(r.iterations, it)
    fin = np.isfinite(V[:, cv])
    dvm = np.abs(np.abs(r.V[:, cv])[fin] - np.abs(V[:, cv])[fin]).max()
    dva = np.abs(np.angle(r.V[:, cv])[fin] - np.angle(V[:, cv])[fin]).max()
    assert dvm <= TOL_V and dva <= TOL_V, (dvm, dva)
