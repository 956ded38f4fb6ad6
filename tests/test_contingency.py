"""N-1 inputs on the host and the oracle's island rules (no GPU).

The island oracle (oracle/islands.py) is pinned against the expectations of the
reference's fixtures (restated in tests/n1_fixtures.py); the host-side stamp bookkeeping
is pinned against a Ybus rebuilt without the outaged branches
(tests/contingency/test_case_generator.py:53-62)."""

import numpy as np
import pytest

from oracle import islands
from p3smi import contingency, grids
from tests.n1_fixtures import FIXTURES, stamps, ybus


@pytest.mark.parametrize("name", list(FIXTURES))
def test_oracle_islands_match_fixture_expectations(name):
    fx = FIXTURES[name]()
    yff, yft, ytf, ytt = stamps(fx)
    live = (yft != 0) | (ytf != 0)
    gen_bus = np.nonzero(np.isfinite(fx["gen_z"]))[0]
    for gname, members in fx["groups"].items():
        for reslack, exp in ((False, fx["expect"][gname][0]), (True, fx["expect"][gname][1])):
            labels, nc = islands.components_after_outage(fx["n"], fx["f"], fx["t"], live, members)
            served, ref_bus, pinned = islands.resolve_references(labels, nc, fx["ref"], gen_bus,
                                                                 fx["gen_z"][gen_bus], reslack)
            np.testing.assert_array_equal(served, exp, err_msg=f"{name}/{gname} reslack={reslack}")
            if reslack and "ref_on" in fx:
                np.testing.assert_array_equal(ref_bus, fx["ref_on"][gname])
                assert pinned == [fx["ref_on"][gname][-1]]


def test_branch_positions_and_stamp_subtraction_match_rebuilt_ybus():
    g = grids.config_grid("pf200", 0)
    Yp, Yj, Yx = g.csr()
    pos = contingency.branch_positions(Yp, Yj, g.f, g.t)
    lookup = {}
    for r in range(g.n):
        for k in range(Yp[r], Yp[r + 1]):
            lookup[(r, Yj[k])] = k
    for k in range(g.nbr):
        f, t = g.f[k], g.t[k]
        assert tuple(pos[k]) == (lookup[(f, f)], lookup[(f, t)], lookup[(t, f)], lookup[(t, t)])
    yff, yft, ytf, ytt = g.branch_stamps()
    yval = np.stack([yff, yft, ytf, ytt], axis=1)
    rng = np.random.default_rng(0)
    for k in rng.choice(g.nbr, 25, replace=False):
        Yc = islands.case_ybus(Yx, pos, yval, [k])
        keep = np.ones(g.nbr, bool)
        keep[k] = False
        Yr = grids.build_ybus(g.n, g.f[keep], g.t[keep], yff[keep], yft[keep], ytf[keep], ytt[keep])
        import scipy.sparse as sp

        M = sp.csr_matrix((Yc, Yj, Yp), shape=(g.n, g.n))
        assert abs(M - Yr).max() < 1e-9


def test_parallel_branch_keeps_other_contribution():
    fx = FIXTURES["parallel_branch"]()
    Y = ybus(fx)
    yff, yft, ytf, ytt = stamps(fx)
    pos = contingency.branch_positions(Y.indptr, Y.indices, fx["f"], fx["t"])
    Yc = islands.case_ybus(Y.data, pos, np.stack([yff, yft, ytf, ytt], axis=1), [0])
    k01 = pos[1, 1]
    assert abs(Yc[k01]) > 1e-6
    assert abs(Yc[k01] - yft[1]) < 1e-12
