"""Cases ingested from MATPOWER files solve on the device as on the oracle (flags and
iterations exact, |dV| <= 1e-9), including a phase-shifting transformer (non-symmetric Ybus)."""

import os

import numpy as np
import pytest

from oracle.oracle import OracleSolver
from p3smi import matpower, nr_klu

from .test_gpu_parity import _compare
from .test_matpower_cpu import CASE4

pytestmark = pytest.mark.gpu
HERE = os.path.dirname(os.path.abspath(__file__))


@pytest.mark.parametrize("name", ["wscc9", "case4"])
def test_ingested_cases_match_oracle(gpu, name):
    c = matpower.load(os.path.join(HERE, "golden", "wscc9.m")) if name == "wscc9" else matpower.parse(CASE4)
    Yp, Yj, Yx = c.csr()
    scale = np.linspace(0.6, 1.6, 64)
    S = np.outer(c.sbus, scale)
    gen_p = np.zeros(c.n, complex)
    np.add.at(gen_p, [c.bus_index[int(b)] for b in c.gen[c.gen[:, 7] > 0, 0]],
              (c.gen[c.gen[:, 7] > 0, 1] + 1j * c.gen[c.gen[:, 7] > 0, 2]) / c.base_mva)
    S = S + np.outer(gen_p, 1 - scale)  # scale the loads only
    r = nr_klu.Solver(Yp, Yj, Yx, c.pv, c.pq).solve_batch(S, c.v0, max_iter=20, tol=1e-8)
    ref = OracleSolver(Yp, Yj, Yx, c.pv, c.pq).solve_batch(S, c.v0, max_iter=20, tol=1e-8)
    _compare(r, ref, name)
    assert r["converged"].mean() > 0.9
