"""The host DC start (p3s/timeseries.py:200-244, p3s/NewtonPowerflowCpp.py:422-440) pinned by
what the mathematics fixes: its angles solve B'(pvpq, pvpq) theta = P(pvpq) - B'(pvpq, ref)
theta_ref (residual at round-off), the reference angles are untouched, the start magnitudes are
kept, and a grid whose DC matrix is singular (an island without a reference) falls back to
angle 0 at the buses the solve cannot fix."""

import numpy as np
import pytest
import scipy.sparse as sp

from p3smi import grids
from p3smi.timeseries import dc_initial_voltage, dc_initial_voltage_arrays, dc_susceptance


def dc_residual(g, V):
    B = dc_susceptance(g.n, g.f, g.t, g.x).tocsr()
    th = np.angle(V)
    pvpq = np.concatenate([g.pv, g.pq])
    r = B[pvpq][:, pvpq] @ th[pvpq] + B[pvpq][:, g.ref] @ th[g.ref] - g.sbus().real[pvpq]
    return float(np.abs(r).max()), float(np.abs(g.sbus().real).max())


@pytest.mark.parametrize("name", ["pf200", "ts2000", "n1_10k"])
def test_host_dc_start_solves_the_dc_system(name):
    g = grids.config_grid(name, 0)
    V = dc_initial_voltage(g)
    res, scale = dc_residual(g, V)
    assert res <= 1e-12 * max(scale, 1.0), res
    flat = g.flat_start()
    np.testing.assert_allclose(np.abs(V), np.abs(flat), rtol=1e-15)
    np.testing.assert_array_equal(V[g.ref], flat[g.ref])
    assert np.abs(np.angle(V)).max() > 1e-3  # the start moved off the flat angles


def test_dc_start_of_a_singular_dc_matrix_keeps_finite_angles():
    # three buses, slack 0 connected to 1; bus 2 connected to nothing: B' singular
    B = sp.csr_matrix(np.array([[10.0, -10.0, 0.0], [-10.0, 10.0, 0.0], [0.0, 0.0, 0.0]]))
    V = dc_initial_voltage_arrays(B, np.array([0.0, -0.5, 0.1]), np.ones(3, complex), [0], [1], [2])
    assert np.isfinite(V).all()
