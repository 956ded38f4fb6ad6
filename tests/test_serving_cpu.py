import numpy as np

from p3smi import grids
from p3smi.serving import case_footprint, topology_key


def test_topology_key_is_content_hash():
    g = grids.config_grid("pf200", 0)
    Yp, Yj, Yx = g.csr()
    k1 = topology_key(Yp, Yj, Yx, g.pv, g.pq)
    assert k1 == topology_key(Yp.copy(), Yj.copy(), Yx.copy(), g.pv.copy(), g.pq.copy())
    Yx2 = Yx.copy()
    Yx2[0] += 1e-12
    assert topology_key(Yp, Yj, Yx2, g.pv, g.pq) != k1


def test_case_footprint_counts_factor_and_state():
    info = {"n": 2000, "m": 3798, "nnz_j": 26062, "nnz_lu": 38822}
    b = case_footprint(info)
    assert 8 * (26062 + 38822) < b < 2 * 8 * (26062 + 38822) + 200000
