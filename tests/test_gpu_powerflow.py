"""Seed sequence and voltage-band acceptance (p3s/NewtonPowerflowCpp.py:553-584) and the
time-series driver (p3s/NewtonPowerflowCpp.py:586-676) on the device."""

import numpy as np
import pytest

from p3smi import grids
from p3smi.powerflow import LoadflowNotConverged, PowerFlow
from p3smi.timeseries import dc_initial_voltage

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def pf200():
    g = grids.config_grid("pf200", 0)
    Yp, Yj, Yx = g.csr()
    return PowerFlow(Yp, Yj, Yx, g.pv, g.pq), g


def test_calculate_accepts_in_band(gpu, pf200):
    p, g = pf200
    S = grids.timeseries_scaled(g, 1, seed=1)[:, 0]
    V = p.calculate(S, dc_initial_voltage(g), g.flat_start(), max_iterations=20)
    ref = p.solver.solve(S, dc_initial_voltage(g), max_iter=20)
    assert ref["converged"]
    assert np.array_equal(V, ref["V"])
    assert 0.5 <= np.abs(V).min() and np.abs(V).max() <= 1.5


def test_calculate_out_of_band_and_not_converged(gpu, pf200):
    p, g = pf200
    S = grids.timeseries_scaled(g, 1, seed=1)[:, 0]
    with pytest.raises(LoadflowNotConverged, match="non-physical solution"):
        p.calculate(S, g.flat_start(), g.flat_start(), max_iterations=20, voltage_band=(0.999, 1.001))
    with pytest.raises(LoadflowNotConverged, match="did not converge in 1 iterations"):
        p.calculate(S, g.flat_start(), g.flat_start(), max_iterations=1)


def test_calculate_batch_band_on_device(gpu, pf200):
    p, g = pf200
    S = grids.timeseries_scaled(g, 96, seed=1)
    V0 = dc_initial_voltage(g)
    r = p.calculate_batch(S, V0, None, max_iterations=20)
    vm = np.abs(r["V"])
    # device |V| range == host range of the returned voltages (|V| = Vm up to the last bit)
    assert np.allclose(r["vmin"], vm.min(axis=0), rtol=1e-15, atol=0)
    assert np.allclose(r["vmax"], vm.max(axis=0), rtol=1e-15, atol=0)
    assert r["accepted"].all() and (r["seed"] == 0).all()
    # a band between the cases' minima: the cases under it are retried from the flat start
    cut = float(np.median(r["vmin"]))
    band = (cut, 1.5)
    r2 = p.calculate_batch(S, V0, g.flat_start(), max_iterations=20, voltage_band=band)
    low = r["vmin"] < cut
    assert (r2["seed"] == low.astype(np.int8)).all()
    assert np.array_equal(r2["V"][:, ~low], r["V"][:, ~low])
    flat = p.solver.solve_batch(np.ascontiguousarray(S[:, low]), g.flat_start(), max_iter=20)
    assert np.array_equal(r2["iterations"][low], flat["iterations"])
    assert np.allclose(r2["V"][:, low], flat["V"], rtol=0, atol=1e-12)
    assert np.array_equal(r2["accepted"], r2["converged"] & (r2["vmin"] >= cut) & (r2["vmax"] <= 1.5))


def test_calculate_timeseries(gpu, pf200):
    p, g = pf200
    S = grids.timeseries_scaled(g, 32, seed=2)
    V0 = g.flat_start()
    V = p.calculate_timeseries(S, V0, max_iterations=20)
    ref = p.solver.solve_batch(S, V0, max_iter=20)
    assert np.array_equal(V, ref["V"])
    with pytest.raises(LoadflowNotConverged, match="did not converge for 32 of 32 time steps"):
        p.calculate_timeseries(S, V0, max_iterations=1)
