"""Tiny p.u. grids restating the reference's N-1 fixtures (p3s/contingency/fixtures.py:66-314)
for topology, outage groups and island expectations. Impedances are arbitrary p.u. values
(the island rules do not depend on them); the expectations are the ones the fixtures'
docstrings and the reference's tests state (tests/contingency/test_case_generator.py:104-146).

Each fixture: dict(n, f, t, r, x, b, ref, pv, pq, gen_z, groups {name: [branch idx]},
expect {name: (served_off, served_on)}) with served masks for re-slacking off / on.
"""

import numpy as np


def _lines(pairs, r=0.01, x=0.05, b=0.02):
    f = np.array([p[0] for p in pairs])
    t = np.array([p[1] for p in pairs])
    k = len(pairs)
    return f, t, np.full(k, r), np.full(k, x), np.full(k, b)


def radial_spur():
    # meshed core 0-1, 1-2, 0-2; spur 2-3, 3-4; slack 0
    f, t, r, x, b = _lines([(0, 1), (1, 2), (0, 2), (2, 3), (3, 4)])
    n = 5
    allon = np.ones(n, bool)
    spur = np.array([1, 1, 1, 0, 0], bool)
    return dict(n=n, f=f, t=t, r=r, x=x, b=b, ref=[0], pv=[], pq=[1, 2, 3, 4], gen_z=np.full(n, np.inf),
                groups={"spur": [3], "core": [0]}, expect={"spur": (spur, spur), "core": (allon, allon)})


def parallel_branch():
    f, t, r, x, b = _lines([(0, 1), (0, 1), (1, 2), (0, 2)])
    n = 3
    allon = np.ones(n, bool)
    return dict(n=n, f=f, t=t, r=r, x=x, b=b, ref=[0], pv=[], pq=[1, 2], gen_z=np.full(n, np.inf),
                groups={"one_of_two": [0]}, expect={"one_of_two": (allon, allon)})


def parallel_branch_both():
    f, t, r, x, b = _lines([(0, 1), (0, 1), (1, 2), (0, 2)])
    n = 3
    allon = np.ones(n, bool)
    return dict(n=n, f=f, t=t, r=r, x=x, b=b, ref=[0], pv=[], pq=[1, 2], gen_z=np.full(n, np.inf),
                groups={"both": [0, 1]}, expect={"both": (allon, allon)})


def generator_island():
    # cluster A: a0=0, a1=1 (slack 0); cluster B: g0=2, g1=3 with a generator at g1; tie 1-2
    f, t, r, x, b = _lines([(0, 1), (1, 2), (2, 3)])
    n = 4
    gz = np.full(n, np.inf)
    gz[3] = 0.18
    off = np.array([1, 1, 0, 0], bool)
    on = np.ones(n, bool)
    return dict(n=n, f=f, t=t, r=r, x=x, b=b, ref=[0], pv=[3], pq=[1, 2], gen_z=gz, groups={"tie": [1]},
                expect={"tie": (off, on)}, ref_on={"tie": [0, 0, 3, 3]})


def multi_branch_group():
    # ring 0-1-2-3-0, spur chain 3-4-5; group "string" removes 3-4 and 4-5
    f, t, r, x, b = _lines([(0, 1), (1, 2), (2, 3), (3, 0), (3, 4), (4, 5)])
    n = 6
    exp = np.array([1, 1, 1, 1, 0, 0], bool)
    return dict(n=n, f=f, t=t, r=r, x=x, b=b, ref=[0], pv=[], pq=[1, 2, 3, 4, 5], gen_z=np.full(n, np.inf),
                groups={"string": [4, 5]}, expect={"string": (exp, exp)})


def open_tie_radial():
    # feeders 0-1-2 and 0-3-4; tie 2-4 open at bus 4: it keeps its charging at bus 2 but
    # has zero transfer admittance (Y_ft = Y_tf = 0) and is no edge
    f, t, r, x, b = _lines([(0, 1), (1, 2), (0, 3), (3, 4), (2, 4)])
    n = 5
    cut_a = np.array([1, 0, 0, 1, 1], bool)
    return dict(n=n, f=f, t=t, r=r, x=x, b=b, ref=[0], pv=[], pq=[1, 2, 3, 4], gen_z=np.full(n, np.inf),
                groups={"head_a": [0]}, expect={"head_a": (cut_a, cut_a)}, open_branch=4)


def lowest_z_tie_break():
    # island {2,3,4} with generators at 2 (z 0.2000001, within np.isclose of 0.2) and 4
    # (z 0.2): the lowest bus index among the best wins -> bus 2
    f, t, r, x, b = _lines([(0, 1), (1, 2), (2, 3), (3, 4)])
    n = 5
    gz = np.full(n, np.inf)
    gz[2] = 0.2000001
    gz[4] = 0.2
    gz[3] = 0.25
    return dict(n=n, f=f, t=t, r=r, x=x, b=b, ref=[0], pv=[2, 3, 4], pq=[1], gen_z=gz, groups={"cut": [1]},
                expect={"cut": (np.array([1, 1, 0, 0, 0], bool), np.ones(n, bool))}, ref_on={"cut": [0, 0, 2, 2, 2]})


FIXTURES = {f.__name__: f for f in (radial_spur, parallel_branch, parallel_branch_both, generator_island,
                                    multi_branch_group, open_tie_radial, lowest_z_tie_break)}


def stamps(fx):
    """(Yff, Yft, Ytf, Ytt) of the fixture's branches; the open tie couples nothing."""
    y = 1.0 / (fx["r"] + 1j * fx["x"])
    ysh = 0.5j * fx["b"]
    yff, yft, ytf, ytt = y + ysh, -y, -y, y + ysh
    if "open_branch" in fx:
        k = fx["open_branch"]
        yft = yft.copy()
        ytf = ytf.copy()
        ytt = ytt.copy()
        yff = yff.copy()
        yft[k] = 0.0
        ytf[k] = 0.0
        ytt[k] = 0.0  # the open end carries nothing; the closed end keeps its charging
        yff[k] = ysh[k]
    return yff, yft, ytf, ytt


def ybus(fx):
    import scipy.sparse as sp

    yff, yft, ytf, ytt = stamps(fx)
    f, t, n = fx["f"], fx["t"], fx["n"]
    rows = np.concatenate([f, f, t, t])
    cols = np.concatenate([f, t, f, t])
    vals = np.concatenate([yff, yft, ytf, ytt])
    Y = sp.coo_matrix((vals, (rows, cols)), shape=(n, n)).tocsr()
    Y.sum_duplicates()
    Y.sort_indices()
    # keep structural entries of open branches (explicit zeros) so the pattern is the full one
    return Y
