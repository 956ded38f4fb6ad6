"""TEST INFRASTRUCTURE ONLY: numpy restatement of the reference's N-1 post-processing, the
checker of the device kernels.

* branch_table: compute_branch_loading (p3s/contingency/find_overloads.py:128-262) on a
  (n, L) voltage table with NaN at unserved buses: S_from = conj(Yf V) V_f sn (Yf row k =
  yff at f, yft at t), S_to likewise; I = |S| / (vn |V| sqrt 3); loading =
  max(I_from, I_to) / rating * 100 (line) or max(I vn sqrt 3) / rating * 100 (trafo);
  outaged branches zeroed, then non-converged cases NaN (assigned into the complex power
  arrays, which stores NaN + 0j: reactive powers read 0 there, as in the reference).
* overloads: find_overloads (:322-420) on that table.
* voltage_violations: find_voltage_violations (p3s/contingency/voltage_band.py:103-182).

Pinned by Kirchhoff's current law in tests/test_postproc_cpu.py: on a converged case the
branch end powers at a bus sum to V conj(Ybus V) there (the case's Ybus is the sum of the
stamps of the branches still in service). A dropped stamp term, a missing conjugate or a
swapped end fails it.
"""

import numpy as np
import scipy.sparse as sp


def branch_table(V, converged, f, t, yval, groups, vn_from, vn_to, rating, trafo, sn_mva=1.0):
    n, L = V.shape
    nb = len(f)
    rows = np.arange(nb)
    yff, yft, ytf, ytt = (yval[:, q] for q in range(4))
    Yf = sp.csr_matrix((np.r_[yff, yft], (np.r_[rows, rows], np.r_[f, t])), shape=(nb, n))
    Yt = sp.csr_matrix((np.r_[ytf, ytt], (np.r_[rows, rows], np.r_[f, t])), shape=(nb, n))
    S_from = np.conj(Yf @ V) * V[f, :] * sn_mva
    S_to = np.conj(Yt @ V) * V[t, :] * sn_mva
    vm = np.abs(V)
    sqrt3 = np.sqrt(3.0)
    vf = np.asarray(vn_from, dtype=float)[:, None]
    vt = np.asarray(vn_to, dtype=float)[:, None]
    tr = np.asarray(trafo, dtype=bool)
    with np.errstate(divide="ignore", invalid="ignore"):
        i_from = np.abs(S_from) / (vf * vm[f, :] * sqrt3)
        i_to = np.abs(S_to) / (vt * vm[t, :] * sqrt3)
        i_max = np.maximum(i_from, i_to)
        line_load = i_max / np.asarray(rating, dtype=float)[:, None] * 100.0
        s_max = np.maximum(i_from * vf * sqrt3, i_to * vt * sqrt3)
        trafo_load = s_max / np.asarray(rating, dtype=float)[:, None] * 100.0
    loading = np.where(tr[:, None], trafo_load, line_load)
    outaged = np.zeros((nb, L), dtype=bool)
    for c, g in enumerate(groups):
        outaged[np.asarray(g, dtype=np.int64), c] = True
    for arr in (S_from, S_to, i_from, i_to, i_max, loading):
        arr[outaged] = 0.0
    bad = ~np.asarray(converged, dtype=bool)
    for arr in (S_from, S_to, i_from, i_to, i_max, loading):
        arr[:, bad] = np.nan
    return {"p_from": S_from.real, "q_from": S_from.imag, "p_to": S_to.real, "q_to": S_to.imag, "i_from": i_from,
            "i_to": i_to, "i_ka": i_max, "loading": loading, "S_from": S_from, "S_to": S_to}


def worst(loading):
    """nanmax and first nanargmax per branch (NaN / -1 where every case is NaN)."""
    allnan = np.isnan(loading).all(axis=1)
    best = np.full(loading.shape[0], np.nan)
    arg = np.full(loading.shape[0], -1, dtype=np.int64)
    if (~allnan).any():
        best[~allnan] = np.nanmax(loading[~allnan], axis=1)
        arg[~allnan] = np.nanargmax(loading[~allnan], axis=1)
    return best, arg


def overloads(loading, i_ka, threshold):
    """(branch, case, loading, i_ka) above the threshold, worst first, ties by branch, case."""
    with np.errstate(invalid="ignore"):
        r, c = np.nonzero(loading > threshold)
    ld = loading[r, c]
    order = np.lexsort((c, r, -ld))
    return r[order], c[order], ld[order], i_ka[r, c][order]


def voltage_violations(V, converged, vmin, vmax, in_service, tol):
    vm = np.abs(V)
    vm[:, ~np.asarray(converged, dtype=bool)] = np.nan
    vm[~np.asarray(in_service, dtype=bool), :] = np.nan
    with np.errstate(invalid="ignore"):
        low = vm < (vmin - tol)[:, None]
        high = vm > (vmax + tol)[:, None]
    lr, lc = np.nonzero(low)
    hr, hc = np.nonzero(high)
    rows = np.concatenate([lr, hr])
    cols = np.concatenate([lc, hc])
    vm_v = vm[rows, cols]
    limit = np.concatenate([vmin[lr], vmax[hr]])
    kind = np.array(["low"] * lr.size + ["high"] * hr.size, dtype=object)
    order = np.argsort(np.abs(vm_v - limit), kind="stable")[::-1]
    return rows[order], cols[order], kind[order], vm_v[order], limit[order]
