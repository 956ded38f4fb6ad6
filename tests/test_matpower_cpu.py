"""MATPOWER case ingestion (p3smi/matpower.py): Ybus against a dense incidence-matrix
construction, bus typing, injections and start voltage; the ingested 9-bus case solved by
the oracle."""

import os

import numpy as np
import pytest

from oracle.oracle import OracleSolver
from p3smi import matpower

HERE = os.path.dirname(os.path.abspath(__file__))

CASE4 = """
function mpc = case4t
mpc.version = '2';
mpc.baseMVA = 100;
% a bracket in a comment: [ 1 2 3 ];
mpc.bus = [
    10  3   0   0   0    0   1  1.00  0   110 1 1.1 0.9;
    20  2   20  5   0    0   1  1.00  0   110 1 1.1 0.9;
    30  2   40  10  2    15  1  1.00  0   110 1 1.1 0.9;
    40  1   30  12  0   -4   1  0.98  -2  20  1 1.1 0.9;
];
mpc.gen = [
    10  0   0   100 -100 1.03 100 1 200 0;
    20  50  0   100 -100 1.01 100 1 200 0;
    30  10  0   100 -100 1.02 100 0 200 0;
];
mpc.branch = [
    10  20  0.01  0.05  0.04  0 0 0 0     0   1 -360 360;
    20  30  0.02  0.08  0.06  0 0 0 0     0   1 -360 360;
    10  30  0.02  0.09  0.05  0 0 0 0     0   0 -360 360;
    30  40  0.002 0.06  0.0   0 0 0 0.95  5.0 1 -360 360;
    10  40  0.03  0.12  0.02  0 0 0 1.0   0   1 -360 360;
];
"""


def _dense_ybus(base, bus, branch):
    """Incidence-matrix construction: Ybus = Cf' Yf + Ct' Yt + diag(Ysh) with
    Yf = diag(Yff) Cf + diag(Yft) Ct and Yt = diag(Ytf) Cf + diag(Ytt) Ct."""
    n = bus.shape[0]
    idx = {int(b): i for i, b in enumerate(bus[:, 0])}
    on = branch[:, 10] > 0
    br = branch[on]
    nl = br.shape[0]
    Cf = np.zeros((nl, n))
    Ct = np.zeros((nl, n))
    for k in range(nl):
        Cf[k, idx[int(br[k, 0])]] = 1
        Ct[k, idx[int(br[k, 1])]] = 1
    Ys = 1 / (br[:, 2] + 1j * br[:, 3])
    Bc = br[:, 4]
    tap = np.ones(nl, complex)
    nz = br[:, 8] != 0
    tap[nz] = br[nz, 8]
    tap = tap * np.exp(1j * np.pi / 180 * br[:, 9])
    Ytt = Ys + 1j * Bc / 2
    Yff = Ytt / (tap * np.conj(tap))
    Yft = -Ys / np.conj(tap)
    Ytf = -Ys / tap
    Yf = np.diag(Yff) @ Cf + np.diag(Yft) @ Ct
    Yt = np.diag(Ytf) @ Cf + np.diag(Ytt) @ Ct
    Ysh = (bus[:, 4] + 1j * bus[:, 5]) / base
    return Cf.T @ Yf + Ct.T @ Yt + np.diag(Ysh)


def test_ybus_types_injections_start():
    c = matpower.parse(CASE4)
    assert c.n == 4 and c.base_mva == 100
    Y = c.ybus().toarray()
    assert np.allclose(Y, _dense_ybus(c.base_mva, c.bus, c.branch), rtol=1e-14, atol=1e-12)
    # the phase shifter makes Ybus non-symmetric, its pattern stays symmetric
    assert not np.allclose(Y, Y.T) and np.array_equal(Y != 0, (Y != 0).T)
    # the out-of-service branch 10-30 is absent
    assert Y[0, 2] == 0
    assert c.ref.tolist() == [0] and c.pv.tolist() == [1]
    assert c.pq.tolist() == [2, 3]  # bus 30 is typed PV but its generator is out of service
    assert np.allclose(c.sbus, np.array([0, (50 - 20) - 5j, -40 - 10j, -30 - 12j]) / 100)
    assert np.allclose(c.v0, [1.03, 1.01, 1.0, 0.98 * np.exp(-2j * np.pi / 180)])


def test_bad_cases_are_rejected():
    with pytest.raises(ValueError, match="mpc.branch is missing"):
        matpower.parse("mpc.baseMVA = 100;\nmpc.bus = [1 3 0 0 0 0 1 1 0 1 1 1 1];\nmpc.gen = [1 0 0 0 0 1 1 1];")
    bad = CASE4.replace("10  20  0.01", "99  20  0.01")
    with pytest.raises(ValueError, match="unknown bus 99"):
        matpower.parse(bad)


def test_wscc9_solves_on_the_oracle():
    c = matpower.load(os.path.join(HERE, "golden", "wscc9.m"))
    Yp, Yj, Yx = c.csr()
    r = OracleSolver(Yp, Yj, Yx, c.pv, c.pq).solve_batch(c.sbus[:, None], c.v0, 20, 1e-10, 0)
    assert r["converged"][0] and r["iterations"][0] <= 5
    V = r["V"][:, 0]
    S = V * np.conj(c.ybus() @ V)
    mis = np.r_[(S - c.sbus)[np.r_[c.pv, c.pq]].real, (S - c.sbus)[c.pq].imag]
    assert np.abs(mis).max() < 1e-10
    assert np.allclose(np.abs(V[c.ref]), 1.04) and np.allclose(np.abs(V[c.pv]), 1.025)
