"""The N-1 post-processing restatement (oracle/postproc.py), pinned by physics and brute force."""

import numpy as np

from oracle import islands, postproc
from oracle.oracle import OracleSolver
from p3smi import contingency, grids


def _n1(subset, reslack=True):
    g = grids.config_grid("pf200", 0)
    Yp, Yj, Yx = g.csr()
    yval = np.stack(g.branch_stamps(), axis=1)
    live = (yval[:, 1] != 0) | (yval[:, 2] != 0)
    groups = [np.array([k]) for k in subset]
    gen_bus = np.nonzero(np.isfinite(g.gen_z))[0]
    srv, _, pinned = islands.cases(g.n, g.f, g.t, live, groups, g.ref, gen_bus, g.gen_z[gen_bus], reslack)
    V0, pin = islands.solver_inputs(g.flat_start(), srv, pinned)
    ypos = contingency.branch_positions(Yp, Yj, g.f, g.t)
    Yc = np.stack([islands.case_ybus(Yx, ypos, yval, gr) for gr in groups], axis=1)
    r = OracleSolver(Yp, Yj, Yx, g.pv, g.pq).solve_batch_contingency(np.abs(Yc), np.angle(Yc), g.sbus(), V0, pin,
                                                                      20, 1e-8)
    V = r["V"].copy()
    V[~srv] = np.nan
    return g, Yp, Yj, Yc, V, r["converged"], groups, yval, srv


def _kcl_residual(g, Yp, Yj, Yc, V, conv, tab, srv):
    worst = 0.0
    for c in np.nonzero(conv)[0]:
        rows = np.repeat(np.arange(g.n), np.diff(Yp))
        I = np.zeros(g.n, complex)
        np.add.at(I, rows, Yc[:, c] * V[Yj, c])
        Sbus = V[:, c] * np.conj(I)
        Sbr = np.zeros(g.n, complex)
        np.add.at(Sbr, g.f, np.nan_to_num(tab["S_from"][:, c]))
        np.add.at(Sbr, g.t, np.nan_to_num(tab["S_to"][:, c]))
        s = srv[:, c]
        worst = max(worst, float(np.abs(Sbr[s] - Sbus[s]).max()))
    return worst


def test_branch_flows_satisfy_kirchhoff_at_every_served_bus():
    subset = list(range(0, 259, 7))
    g, Yp, Yj, Yc, V, conv, groups, yval, srv = _n1(subset)
    assert conv.mean() > 0.9 and (~srv).any()
    ones = np.ones(g.nbr)
    tab = postproc.branch_table(V, conv, g.f, g.t, yval, groups, ones, ones, ones, np.zeros(g.nbr, bool))
    assert _kcl_residual(g, Yp, Yj, Yc, V, conv, tab, srv) < 1e-10
    # outaged branches carry nothing, non-converged cases are NaN
    for c, gr in enumerate(groups):
        assert (tab["loading"][gr, c] == 0).all() or not conv[c]
    # a planted mistake (the line charging dropped from the end stamps) breaks the balance
    bad_yval = yval.copy()
    bad_yval[:, 0] = -yval[:, 1]
    bad_yval[:, 3] = -yval[:, 2]
    bad = postproc.branch_table(V, conv, g.f, g.t, bad_yval, groups, ones, ones, ones, np.zeros(g.nbr, bool))
    assert _kcl_residual(g, Yp, Yj, Yc, V, conv, bad, srv) > 1e-3


def test_current_and_loading_formulas():
    """One branch by hand: I = |S| / (vn |V| sqrt 3), line and transformer loadings."""
    V = np.array([[1.02 + 0.01j], [0.98 - 0.03j]])
    y = 1 / (0.01 + 0.05j)
    yval = np.array([[y + 0.01j, -y, -y, y + 0.01j]])
    f, t = np.array([0]), np.array([1])
    for trafo, rating in ((False, 0.5), (True, 40.0)):
        tab = postproc.branch_table(V, [True], f, t, yval, [np.array([], int)], [110.0], [20.0], [rating], [trafo],
                                    sn_mva=100.0)
        Sf = np.conj((y + 0.01j) * V[0, 0] - y * V[1, 0]) * V[0, 0] * 100.0
        St = np.conj(-y * V[0, 0] + (y + 0.01j) * V[1, 0]) * V[1, 0] * 100.0
        i_f = abs(Sf) / (110.0 * abs(V[0, 0]) * np.sqrt(3))
        i_t = abs(St) / (20.0 * abs(V[1, 0]) * np.sqrt(3))
        # transformer: the apparent power at nominal voltage, I vn sqrt 3 = |S| / |V|, against sn_mva
        want = (max(abs(Sf) / abs(V[0, 0]), abs(St) / abs(V[1, 0])) / rating * 100 if trafo
                else max(i_f, i_t) / rating * 100)
        assert np.isclose(tab["loading"][0, 0], want, rtol=1e-12)
        assert np.isclose(tab["i_ka"][0, 0], max(i_f, i_t), rtol=1e-12)


def test_voltage_violations_match_brute_force():
    rng = np.random.default_rng(2)
    n, L = 9, 6
    V = (rng.uniform(0.9, 1.1, (n, L)) * np.exp(1j * rng.uniform(-0.2, 0.2, (n, L))))
    V[3, 2] = np.nan
    conv = np.array([1, 1, 0, 1, 1, 1], bool)
    vmin = np.full(n, 0.95)
    vmax = np.full(n, 1.05)
    vmin[4] = np.nan
    ins = np.ones(n, bool)
    ins[7] = False
    rows, cols, kind, vm, lim = postproc.voltage_violations(V.copy(), conv, vmin, vmax, ins, 1e-9)
    got = {(int(r), int(c), k) for r, c, k in zip(rows, cols, kind)}
    want = set()
    for i in range(n):
        for c in range(L):
            v = abs(V[i, c])
            if not conv[c] or not ins[i] or np.isnan(v):
                continue
            if v < vmin[i] - 1e-9:
                want.add((i, c, "low"))
            if v > vmax[i] + 1e-9:
                want.add((i, c, "high"))
    assert got == want and len(want) > 3
    dev = np.abs(vm - lim)
    assert (np.diff(dev) <= 0).all()
