"""The DC start solved on the device (the engine's static-pivot LU on [B' 0; 0 I]) against the
host restatement: both solve B'(pvpq, pvpq) theta = P(pvpq) - B'(pvpq, ref) theta_ref to
round-off, agree with each other to round-off times the system's condition, and the Newton
solves started from either give the same flags, iteration counts and voltages."""

import numpy as np
import pytest

from oracle.oracle import OracleSolver
from p3smi import grids, nr_klu
from p3smi.timeseries import dc_initial_voltage
from tests.test_dc_start_cpu import dc_residual

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("name", ["pf200", "ts2000", "n1_10k"])
def test_device_dc_start_equals_host(gpu, name):
    g = grids.config_grid(name, 0)
    Yp, Yj, Yx = g.csr()
    s = nr_klu.Solver(Yp, Yj, Yx, g.pv, g.pq)
    Vd = dc_initial_voltage(g, solver=s)
    Vh = dc_initial_voltage(g)
    res, scale = dc_residual(g, Vd)
    assert res <= 1e-12 * max(scale, 1.0), res
    np.testing.assert_array_equal(Vd[g.ref], Vh[g.ref])
    np.testing.assert_allclose(np.abs(Vd), np.abs(Vh), rtol=1e-15, atol=0)
    assert np.abs(np.angle(Vd) - np.angle(Vh)).max() <= 1e-10
    if name == "ts2000":
        S = grids.timeseries_daily(g, 8760, seed=1)[:, ::97].copy()
        r = s.solve_batch(S, Vd, max_iter=20, tol=1e-8)
        ref = OracleSolver(Yp, Yj, Yx, g.pv, g.pq).solve_batch(S, Vh, max_iter=20, tol=1e-8)
        np.testing.assert_array_equal(r["converged"], ref["converged"])
        np.testing.assert_array_equal(r["iterations"], ref["iterations"])
        assert np.abs(np.abs(r["V"]) - np.abs(ref["V"])).max() <= 1e-9
        assert np.abs(np.angle(r["V"]) - np.angle(ref["V"])).max() <= 1e-9
