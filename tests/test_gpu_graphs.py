"""Newton iterations replayed as captured hipGraphs give the same bits as the kernel-by-kernel
path, replay after replay, with eager solves and synchronisations between the replays (the
graphs record kernels and counter copies only; counters are cleared by a kernel)."""

import numpy as np
import pytest

from p3smi import contingency, grids, nr_klu
from p3smi.timeseries import dc_initial_voltage

pytestmark = pytest.mark.gpu


def _same(a, b, tag):
    assert np.array_equal(a["iterations"], b["iterations"]), tag
    assert np.array_equal(a["converged"], b["converged"]), tag
    assert np.array_equal(a["V"].view(np.float64), b["V"].view(np.float64)), tag
    assert np.array_equal(a["mismatch"], b["mismatch"]), tag


def _perturbed(g, T, seed):
    rng = np.random.default_rng(seed)
    k = rng.uniform(1.0, 2.5, T)
    S = g.p_gen[:, None] * k[None, :] - (g.p_load + 1j * g.q_load)[:, None] * k[None, :]
    V0 = np.repeat(g.flat_start()[:, None], T, axis=1)
    V0[g.pq, :] *= rng.uniform(0.6, 1.4, T)[None, :]
    return S, V0


@pytest.mark.parametrize("ls", [True, False])
def test_graph_replays_match_eager(gpu, ls):
    g = grids.config_grid("pf200", 0)
    Yp, Yj, Yx = g.csr()
    S, V0 = _perturbed(g, 256, 7)  # backtracking and Jacobian-rebuild phases both occur
    s = nr_klu.Solver(Yp, Yj, Yx, g.pv, g.pq)
    eager = s.solve_batch(S, V0, max_iter=20, tol=1e-8, line_search=ls)
    assert 0 < eager["converged"].sum() < S.shape[1]
    for rep in range(3):
        s.set_graphs(True)
        r = s.solve_batch(S, V0, max_iter=20, tol=1e-8, line_search=ls)
        _same(r, eager, f"graph solve {rep}")
        s.set_graphs(False)  # eager work between the replays
        _same(s.solve_batch(S, V0, max_iter=20, tol=1e-8, line_search=ls), eager, f"eager solve {rep}")


def test_graph_replays_ts2000_and_n1(gpu):
    g = grids.config_grid("ts2000", 0)
    Yp, Yj, Yx = g.csr()
    S = grids.timeseries_daily(g, 512, seed=2)
    V0 = dc_initial_voltage(g)
    s = nr_klu.Solver(Yp, Yj, Yx, g.pv, g.pq)
    eager = s.solve_batch(S, V0, max_iter=20, tol=1e-8)
    s.set_graphs(True)
    for rep in range(2):
        _same(s.solve_batch(S, V0, max_iter=20, tol=1e-8), eager, f"ts2000 graph solve {rep}")
    # N-1: per-case Ybus and pin mask select other kernels and buffers (another graph)
    p = grids.config_grid("pf200", 0)
    cs = contingency.from_grid(p, reslack=True)
    a = cs.solve(p.sbus(), p.flat_start(), max_iter=20, tol=1e-8)
    cs.solver.set_graphs(True)
    for rep in range(2):
        b = cs.solve(p.sbus(), p.flat_start(), max_iter=20, tol=1e-8)
        assert np.array_equal(a.iterations, b.iterations)
        assert np.array_equal(a.served, b.served)
        assert np.array_equal(a.V.view(np.float64), b.V.view(np.float64), equal_nan=True)
