"""N-1 post-processing on the device equals the numpy restatement of the reference
(oracle/postproc.py) on the same solve: branch tables to 1e-12 relative, the per-branch
worst cases and the violation lists exactly."""

import numpy as np
import pytest

from oracle import postproc
from p3smi import contingency, grids

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def solved():
    g = grids.config_grid("pf200", 0)
    cs = contingency.from_grid(g, reslack=True)
    res = cs.solve(g.sbus(), g.flat_start(), max_iter=20, tol=1e-8)
    assert res.converged.mean() > 0.9 and (~res.served).any()
    rng = np.random.default_rng(8)
    trafo = rng.random(g.nbr) < 0.15
    vn_from = np.where(trafo, 110.0, 20.0)
    vn_to = np.where(trafo, 20.0, 20.0)
    return g, cs, res, trafo, vn_from, vn_to


def _ref_table(g, cs, res, trafo, vn_from, vn_to, rating, sn):
    yval = np.stack(g.branch_stamps(), axis=1)
    return postproc.branch_table(res.V, res.converged, g.f, g.t, yval, cs.groups, vn_from, vn_to, rating, trafo,
                                 sn_mva=sn)


def _close(a, b, tag="", ctx=None):
    na, nb = np.isnan(a), np.isnan(b)
    if not np.array_equal(na, nb):
        r, c = np.nonzero(na != nb)
        info = ""
        if ctx is not None:
            g, res, cs = ctx
            info = "; ".join(f"(k={k}, c={j}: dev={a[k, j]!r} ref={b[k, j]!r} served f/t={res.served[g.f[k], j]}/"
                             f"{res.served[g.t[k], j]} conv={res.converged[j]} out={k in set(cs.groups[j].tolist())})"
                             for k, j in list(zip(r, c))[:6])
        raise AssertionError(f"{tag}: NaN pattern differs at {r.size} entries: {info}")
    m = ~nb
    assert np.allclose(a[m], b[m], rtol=1e-12, atol=1e-14), tag


def test_branch_table_matches_restatement(gpu, solved):
    g, cs, res, trafo, vn_from, vn_to = solved
    sn = 100.0
    ref0 = _ref_table(g, cs, res, trafo, vn_from, vn_to, np.ones(g.nbr), sn)
    rating = np.nanmax(np.where(trafo[:, None], ref0["loading"], ref0["i_ka"]), axis=1) * 0.9 + 1e-6
    tab = contingency.branch_loading(cs, vn_from, vn_to, rating, trafo, sn)
    ref = _ref_table(g, cs, res, trafo, vn_from, vn_to, rating, sn)
    for ours, theirs in (("p_from_mw", "p_from"), ("q_from_mvar", "q_from"), ("p_to_mw", "p_to"),
                         ("q_to_mvar", "q_to"), ("i_from_ka", "i_from"), ("i_to_ka", "i_to"), ("i_ka", "i_ka"),
                         ("loading_percent", "loading")):
        _close(getattr(tab, ours), ref[theirs], ours, (g, res, cs))
    best, arg = postproc.worst(ref["loading"])
    _close(tab.max_loading_percent, best)
    # the worst case of every branch; where two cases' loadings agree to ~1 ulp the argmax
    # may fall on either, and then both must be maxima within the table tolerance
    diff = np.nonzero(tab.worst_case != arg)[0]
    assert diff.size <= 2, diff
    for k in diff:
        assert tab.worst_case[k] >= 0 and arg[k] >= 0, k
        a_dev, a_ref = ref["loading"][k, tab.worst_case[k]], ref["loading"][k, arg[k]]
        assert np.isclose(a_dev, a_ref, rtol=1e-12, atol=0), (k, a_dev, a_ref)
    # outaged branches are zero in their own case
    for c, gr in enumerate(cs.groups):
        if res.converged[c]:
            assert (tab.loading_percent[gr, c] == 0).all()


def test_overloads_and_voltage_violations_match_restatement(gpu, solved):
    g, cs, res, trafo, vn_from, vn_to = solved
    sn = 100.0
    ref0 = _ref_table(g, cs, res, trafo, vn_from, vn_to, np.ones(g.nbr), sn)
    rating = np.nanmax(np.where(trafo[:, None], ref0["loading"], ref0["i_ka"]), axis=1) * 0.95 + 1e-6
    rep = contingency.find_overloads(cs, 100.0, vn_from, vn_to, rating, trafo, sn)
    ref = _ref_table(g, cs, res, trafo, vn_from, vn_to, rating, sn)
    r, c, ld, ik = postproc.overloads(ref["loading"], ref["i_ka"], 100.0)
    assert len(rep) == r.size > 10
    # the same entries (loadings agree to ~1 ulp, so near-ties may sort either way) ...
    a = np.lexsort((rep.case, rep.branch))
    b = np.lexsort((c, r))
    assert np.array_equal(rep.branch[a], r[b]) and np.array_equal(rep.case[a], c[b])
    assert np.allclose(rep.loading_percent[a], ld[b], rtol=1e-12) and np.allclose(rep.i_ka[a], ik[b], rtol=1e-12)
    # ... reported worst first
    assert (np.diff(rep.loading_percent) <= 0).all()
    vmin, vmax = np.full(g.n, 0.99), np.full(g.n, 1.03)
    ins = np.ones(g.n, bool)
    ins[5] = False
    vv = contingency.find_voltage_violations(cs, vmin, vmax, ins, 1e-9)
    rows, cols, kind, vm, lim = postproc.voltage_violations(res.V.copy(), res.converged, vmin, vmax, ins, 1e-9)
    assert len(vv) == rows.size > 10
    a = np.lexsort((vv.case, vv.bus))
    b = np.lexsort((cols, rows))
    assert np.array_equal(vv.bus[a], rows[b]) and np.array_equal(vv.case[a], cols[b])
    assert list(vv.kind[a]) == list(kind[b])
    assert np.allclose(vv.vm_pu[a], vm[b], rtol=1e-14)
    assert (np.diff(vv.deviation_pu) <= 0).all()
