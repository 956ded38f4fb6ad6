"""N-1 on the GPU == the CPU oracle: island masks and reference buses bit-exact, and the
contingency solves with identical convergence flags and iteration counts, |dV| <= 1e-9 and
NaN exactly at the unserved buses."""

import numpy as np
import pytest

from oracle import islands
from oracle.oracle import OracleSolver
from p3smi import contingency, grids, nr_klu
from tests.n1_fixtures import FIXTURES, stamps, ybus

pytestmark = pytest.mark.gpu


def _fixture_set(fx, reslack):
    Y = ybus(fx)
    s = nr_klu.Solver(Y.indptr, Y.indices, Y.data, fx["pv"], fx["pq"])
    yff, yft, ytf, ytt = stamps(fx)
    names = list(fx["groups"])
    return contingency.ContingencySet(s, fx["f"], fx["t"], yff, yft, ytf, ytt, [fx["groups"][k] for k in names],
                                      fx["ref"], fx["gen_z"], reslack, names)


@pytest.mark.parametrize("name", list(FIXTURES))
def test_islands_on_fixtures(gpu, name):
    fx = FIXTURES[name]()
    for reslack in (False, True):
        cs = _fixture_set(fx, reslack)
        served, ref_bus, pinned = cs.islands()
        for c, g in enumerate(cs.names):
            exp = fx["expect"][g][1 if reslack else 0]
            np.testing.assert_array_equal(served[:, c], exp, err_msg=f"{name}/{g} reslack={reslack}")
            if reslack and "ref_on" in fx:
                np.testing.assert_array_equal(ref_bus[:, c], fx["ref_on"][g])
                assert pinned[c] == 1


def _oracle_islands(cs, grid):
    gen_bus = np.nonzero(np.isfinite(grid.gen_z))[0]
    return islands.cases(grid.n, grid.f, grid.t, cs.live, cs.groups, grid.ref, gen_bus, grid.gen_z[gen_bus],
                         cs.reslack)


@pytest.mark.parametrize("reslack", [False, True])
def test_islands_every_outage_pf200_and_ts2000(gpu, reslack):
    for name in ("pf200", "ts2000"):
        g = grids.config_grid(name, 0)
        cs = contingency.from_grid(g, reslack=reslack)
        served, ref_bus, pinned = cs.islands()
        osrv, oref, opin = _oracle_islands(cs, g)
        np.testing.assert_array_equal(served, osrv)
        np.testing.assert_array_equal(ref_bus, oref)
        np.testing.assert_array_equal(pinned, [len(p) for p in opin])
        assert (~served).any(), "expected some islanding outages on a meshed tree grid"


def test_islands_sample_of_10k_outages(gpu):
    g = grids.config_grid("n1_10k", 0)
    rng = np.random.default_rng(1)
    groups = [np.array([k]) for k in rng.choice(g.nbr, 400, replace=False)]
    groups.append(np.array([0, 1, 2, 3]))  # a multi-branch group
    cs = contingency.from_grid(g, groups=groups, reslack=True)
    served, ref_bus, pinned = cs.islands()
    osrv, oref, opin = _oracle_islands(cs, g)
    np.testing.assert_array_equal(served, osrv)
    np.testing.assert_array_equal(ref_bus, oref)


def _oracle_solve(cs, grid, V_start, max_iter=20, tol=1e-8):
    osrv, oref, opin = _oracle_islands(cs, grid)
    V0, pin = islands.solver_inputs(V_start, osrv, opin)
    Yc = np.stack([islands.case_ybus(cs.solver.Yx, cs.ypos, cs.yval, gr) for gr in cs.groups], axis=1)
    Yp, Yj, Yx = grid.csr()
    o = OracleSolver(Yp, Yj, Yx, grid.pv, grid.pq)
    r = o.solve_batch_contingency(np.abs(Yc), np.angle(Yc), grid.sbus(), V0, pin, max_iter, tol)
    V = r["V"].copy()
    V[~osrv] = np.nan
    return V, r["iterations"], r["converged"], osrv


@pytest.mark.parametrize("reslack", [False, True])
def test_n1_solve_pf200_every_branch(gpu, reslack):
    g = grids.config_grid("pf200", 0)
    cs = contingency.from_grid(g, reslack=reslack)
    V_start = g.flat_start()
    r = cs.solve(g.sbus(), V_start, max_iter=20, tol=1e-8)
    V, it, cv, srv = _oracle_solve(cs, g, V_start)
    np.testing.assert_array_equal(r.served, srv)
    np.testing.assert_array_equal(np.isnan(r.V), np.isnan(V))
    np.testing.assert_array_equal(r.converged, cv)
    np.testing.assert_array_equal(r.iterations, it)
    ok = cv
    fin = ~np.isnan(V[:, ok])
    dvm = np.abs(np.abs(r.V[:, ok])[fin] - np.abs(V[:, ok])[fin]).max()
    dva = np.abs(np.angle(r.V[:, ok])[fin] - np.angle(V[:, ok])[fin]).max()
    assert dvm <= 1e-9 and dva <= 1e-9, (dvm, dva)
    assert cv.mean() > 0.9
