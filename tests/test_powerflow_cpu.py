"""Input validation of the power-flow drivers (no device needed: every check runs before a
solve). Messages follow calculate_timeseries_cpp (p3s/NewtonPowerflowCpp.py:634-672)."""

import numpy as np
import pytest

from p3smi import grids
from p3smi.powerflow import PowerFlow, _in_band


@pytest.fixture(scope="module")
def pf():
    g = grids.config_grid("pf200", 0)
    Yp, Yj, Yx = g.csr()
    return PowerFlow(Yp, Yj, Yx, g.pv, g.pq), g


def test_timeseries_validation(pf):
    p, g = pf
    V = g.flat_start()
    S = grids.timeseries_scaled(g, 4, seed=1)
    with pytest.raises(ValueError, match="zero time steps"):
        p.calculate_timeseries(np.zeros((g.n, 0), complex), V)
    bad = S.copy()
    bad[3, 1] = np.nan
    with pytest.raises(ValueError, match="Sbus matrix contains NaN or inf"):
        p.calculate_timeseries(bad, V)
    bad[3, 1] = np.inf
    with pytest.raises(ValueError, match="Sbus matrix contains NaN or inf"):
        p.calculate_timeseries(bad, V)
    Vb = V.copy()
    Vb[5] = np.nan
    with pytest.raises(ValueError, match="DC initial voltage contains NaN or inf"):
        p.calculate_timeseries(S, Vb)
    with pytest.raises(ValueError, match="V0 shape mismatch"):
        p.calculate_timeseries(S, V[:-1])
    with pytest.raises(ValueError, match="Sbus shape mismatch"):
        p.calculate_timeseries(S[:-1], V)


def test_band_rule():
    vmin = np.array([0.49, 0.5, 0.9, 1.0, np.nan])
    vmax = np.array([1.0, 1.5, 1.51, 1.2, 1.0])
    assert _in_band(vmin, vmax, (0.5, 1.5)).tolist() == [False, True, False, True, False]
    assert _in_band(vmin, vmax, None).all()


def test_solver_reused_on_same_pattern(pf):
    p, g = pf
    s0 = p.solver
    Yp, Yj, Yx = g.csr()
    p.set_grid(Yp, Yj, Yx * 1.01, g.pv, g.pq)
    assert p.solver is s0  # same pattern: values refreshed, plan kept
