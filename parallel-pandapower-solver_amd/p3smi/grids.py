"""Deterministic synthetic transmission grids and their load cases.

The grids are given directly in per-unit: every branch is a pi-model with series
impedance r + jx and total charging b, stamped into Ybus as

    Y_ff = Y_tt = y + j b/2,    Y_ft = Y_tf = -y,    y = 1 / (r + jx)

which is the line stamp of the reference's element model
(p3s/models/TransmissionLineModel.py:31-50) with the impedance already in p.u.
Parallel branches are summed (coo -> csr, duplicates added), columns sorted per row,
as the reference's case generator expects (p3s/contingency/case_generator.py:244-249).

Geometry. Buses are points in the unit square; the branch set is a random spanning tree
of the k-nearest-neighbour graph plus chords drawn from the same local graph. A planar-like
grid keeps the Jacobian's LU fill small (tens of entries per row), like real transmission
grids such as pegase; chords drawn uniformly at random would turn the grid into an
expander whose LU has a dense trailing front.

Bus typing follows the reference's rule (p3s/voltage_sources.py:80-97): ref = slack
buses, pv = buses holding a voltage-controlled generator (sorted), pq = the rest (sorted).
"""

from __future__ import annotations

from dataclasses import dataclass, field

import numpy as np
import scipy.sparse as sp
from scipy.sparse.csgraph import minimum_spanning_tree
from scipy.spatial import cKDTree


@dataclass
class Grid:
    """One synthetic grid in p.u. (all index arrays 0-based)."""

    n: int
    f: np.ndarray  # (nbr,) from bus
    t: np.ndarray  # (nbr,) to bus
    r: np.ndarray  # (nbr,) series resistance
    x: np.ndarray  # (nbr,) series reactance
    b: np.ndarray  # (nbr,) total line charging
    ref: np.ndarray  # slack bus indices
    pv: np.ndarray  # sorted PV bus indices
    pq: np.ndarray  # sorted PQ bus indices
    vm_set: np.ndarray  # (n,) voltage set-point (ref / PV), NaN elsewhere
    p_gen: np.ndarray  # (n,) scheduled generation at PV buses (p.u.)
    p_load: np.ndarray  # (n,) base active load (p.u.)
    q_load: np.ndarray  # (n,) base reactive load (p.u.)
    gen_z: np.ndarray = field(default=None)  # (n,) generator short-circuit |Z| (p.u.) at PV buses, inf elsewhere
    seed: int = 0

    # ---- derived network matrices -------------------------------------------------
    @property
    def nbr(self) -> int:
        return int(self.f.shape[0])

    def branch_stamps(self):
        """(Yff, Yft, Ytf, Ytt) complex arrays of every branch."""
        y = 1.0 / (self.r + 1j * self.x)
        ysh = 0.5j * self.b
        return y + ysh, -y, -y, y + ysh

    def ybus(self) -> sp.csr_matrix:
        """Ybus as a sorted CSR matrix (duplicates summed)."""
        return build_ybus(self.n, self.f, self.t, *self.branch_stamps())

    def csr(self):
        """(Yp int32, Yj int32, Yx complex128) of Ybus."""
        Y = self.ybus()
        return (
            np.ascontiguousarray(Y.indptr, dtype=np.int32),
            np.ascontiguousarray(Y.indices, dtype=np.int32),
            np.ascontiguousarray(Y.data, dtype=np.complex128),
        )

    def sbus(self) -> np.ndarray:
        """Base-case specified injections S = Pg - (Pd + jQd), (n,) complex."""
        return self.p_gen - self.p_load - 1j * self.q_load

    def flat_start(self) -> np.ndarray:
        """Start voltage: set-points at ref/PV, the mean set-point at PQ buses, angle 0.

        The mean of the generator set-points is the magnitude the reference seeds PQ
        buses with (p3s/timeseries.py:161-188, p3s/NewtonPowerflowCpp.py:214-217).
        """
        vm = np.full(self.n, np.nanmean(self.vm_set))
        has = np.isfinite(self.vm_set)
        vm[has] = self.vm_set[has]
        return vm.astype(np.complex128)


def build_ybus(n, f, t, yff, yft, ytf, ytt) -> sp.csr_matrix:
    """Sum the 2x2 branch stamps into a sorted CSR Ybus (coo -> csr adds duplicates)."""
    rows = np.concatenate([f, f, t, t])
    cols = np.concatenate([f, t, f, t])
    vals = np.concatenate([yff, yft, ytf, ytt]).astype(np.complex128)
    Y = sp.coo_matrix((vals, (rows, cols)), shape=(n, n)).tocsr()
    Y.sum_duplicates()
    Y.sort_indices()
    return Y


def _knn_edges(pts: np.ndarray, k: int) -> np.ndarray:
    """Undirected edge list (a < b) of the symmetrised k-nearest-neighbour graph."""
    tree = cKDTree(pts)
    _, idx = tree.query(pts, k=k + 1)
    a = np.repeat(np.arange(pts.shape[0]), k)
    b = idx[:, 1:].ravel()
    lo, hi = np.minimum(a, b), np.maximum(a, b)
    key = np.unique(lo.astype(np.int64) * pts.shape[0] + hi)
    return np.stack([key // pts.shape[0], key % pts.shape[0]], axis=1)


def make_grid(
    n_bus: int,
    n_chords: int,
    n_pv: int,
    seed: int,
    geometry: str = "geo",
    knn: int = 6,
    load_to_gen: float = 0.9,
    power_scale: float = 1.0,
) -> Grid:
    """Build one synthetic meshed grid.

    * Branches: a random spanning tree plus ``n_chords`` chords. ``geometry="geo"``
      draws both from the k-NN graph of random points in the unit square;
      ``geometry="uniform"`` uses a random tree and uniformly random chords.
      r ~ U(0.005, 0.05), x = r * U(2, 10), b ~ U(0, 0.05).
    * Buses: one slack (Vm = 1.02, Va = 0); ``n_pv`` PV buses with
      Vm ~ U(0.98, 1.05) and Pg ~ U(0.5, 3); every other bus is a PQ load with
      P ~ U(0.1, 1.0), Q = P tan(acos(pf)), pf ~ U(0.9, 0.99), scaled so that the total
      load is ``load_to_gen`` times the total scheduled generation.
    * ``power_scale`` multiplies every injection and every line charging b. Taken
      literally (scale 1) the distributions above push tens of p.u. through p.u.-sized
      impedances and no Newton method converges; the benchmark configurations use 0.05
      (200 buses) and 0.007 (2,000 and more), which keep the distributions' shapes and give
      realistic angle spreads.
    * The slack is the bus nearest the centre of the square (``geo``), else bus 0.
    """
    rng = np.random.default_rng(seed)
    n = int(n_bus)
    if geometry == "geo":
        pts = rng.random((n, 2))
        edges = _knn_edges(pts, min(knn, n - 1))
        w = rng.random(edges.shape[0]) + 1e-3
        G = sp.coo_matrix((w, (edges[:, 0], edges[:, 1])), shape=(n, n)).tocsr()
        T = minimum_spanning_tree(G).tocoo()
        tree = np.stack([T.row, T.col], axis=1)
        if tree.shape[0] != n - 1:  # k-NN graph disconnected: bridge the components
            ncomp, lab = sp.csgraph.connected_components(G + G.T, directed=False)
            extra = []
            for c in range(1, ncomp):
                a = np.nonzero(lab == 0)[0]
                bb = np.nonzero(lab == c)[0]
                d = cKDTree(pts[a]).query(pts[bb])
                j = int(np.argmin(d[0]))
                extra.append((a[d[1][j]], bb[j]))
                lab[lab == c] = 0
            tree = np.concatenate([tree, np.array(extra)], axis=0)
        in_tree = {(int(min(a, b)), int(max(a, b))) for a, b in tree}
        cand = np.array([e for e in map(tuple, edges) if (int(e[0]), int(e[1])) not in in_tree])
        pick = rng.choice(cand.shape[0], size=min(n_chords, cand.shape[0]), replace=False)
        chords = cand[np.sort(pick)]
    elif geometry == "uniform":
        perm = rng.permutation(n)
        parent = np.array([perm[rng.integers(0, i)] for i in range(1, n)])
        tree = np.stack([perm[1:], parent], axis=1)
        cf = rng.integers(0, n, size=4 * n_chords + 16)
        ct = rng.integers(0, n, size=4 * n_chords + 16)
        ok = cf != ct
        chords = np.stack([cf[ok], ct[ok]], axis=1)[:n_chords]
    else:
        raise ValueError("geometry must be 'geo' or 'uniform'")
    br = np.concatenate([tree, chords], axis=0).astype(np.int64)
    nbr = br.shape[0]
    r = rng.uniform(0.005, 0.05, nbr)
    x = r * rng.uniform(2.0, 10.0, nbr)
    b = rng.uniform(0.0, 0.05, nbr) * power_scale

    slack = int(np.argmin(((pts - 0.5) ** 2).sum(axis=1))) if geometry == "geo" else 0
    ref = np.array([slack], dtype=np.int64)
    others = rng.permutation(np.setdiff1d(np.arange(n), ref))
    pv = np.sort(others[:n_pv])
    pq = np.sort(others[n_pv:])
    vm_set = np.full(n, np.nan)
    vm_set[slack] = 1.02
    vm_set[pv] = rng.uniform(0.98, 1.05, pv.shape[0])
    p_gen = np.zeros(n)
    p_gen[pv] = rng.uniform(0.5, 3.0, pv.shape[0]) * power_scale
    p_load = np.zeros(n)
    q_load = np.zeros(n)
    p = rng.uniform(0.1, 1.0, pq.shape[0])
    pf = rng.uniform(0.9, 0.99, pq.shape[0])
    scale = load_to_gen * p_gen.sum() / max(p.sum(), 1e-12)
    p_load[pq] = p * scale
    q_load[pq] = p_load[pq] * np.tan(np.arccos(pf))
    gen_z = np.full(n, np.inf)
    gen_z[pv] = rng.uniform(0.1, 0.3, pv.shape[0])
    return Grid(
        n=n,
        f=br[:, 0].copy(),
        t=br[:, 1].copy(),
        r=r,
        x=x,
        b=b,
        ref=ref,
        pv=pv,
        pq=pq,
        vm_set=vm_set,
        p_gen=p_gen,
        p_load=p_load,
        q_load=q_load,
        gen_z=gen_z,
        seed=seed,
    )


# ---- the benchmark workloads ----------------------------------------------------------


def timeseries_scaled(grid: Grid, T: int, seed: int, sigma: float = 0.1, lo: float = 0.6, hi: float = 1.4):
    """(n, T) Sbus: every load scaled per case by N(1, sigma) clipped to [lo, hi]."""
    rng = np.random.default_rng(seed)
    k = np.clip(rng.normal(1.0, sigma, size=(grid.n, T)), lo, hi)
    S = grid.p_gen[:, None] - k * (grid.p_load + 1j * grid.q_load)[:, None]
    return np.ascontiguousarray(S, dtype=np.complex128)


def timeseries_daily(grid: Grid, T: int, seed: int, amplitude: float = 0.25, sigma: float = 0.05):
    """(n, T) hourly Sbus: a daily sinusoid (peak at 18:00) times N(1, sigma) per load,
    and every PV generator's Pg scaled by U(0, 1) per case."""
    rng = np.random.default_rng(seed)
    hours = np.arange(T)
    prof = 1.0 + amplitude * np.sin(2.0 * np.pi * (hours - 12.0) / 24.0)
    k = prof[None, :] * rng.normal(1.0, sigma, size=(grid.n, T))
    g = rng.uniform(0.0, 1.0, size=(grid.n, T))
    S = g * grid.p_gen[:, None] - k * (grid.p_load + 1j * grid.q_load)[:, None]
    return np.ascontiguousarray(S, dtype=np.complex128)


CONFIGS = {
    # name: (n_bus, chords, n_pv, cases, power_scale)
    "pf200": (200, 60, 30, 96, 0.05),
    "ts2000": (2000, 600, 200, 8760, 0.007),
    "n1_10k": (10000, 3500, 800, 13499, 0.007),
}


def config_grid(name: str, seed: int = 0) -> Grid:
    n, ch, npv, _, ps = CONFIGS[name]
    return make_grid(n, ch, npv, seed=seed, power_scale=ps)


def size_scale(n_bus: int) -> float:
    """Injection scale of a grid of any size: 0.007 up to 10,000 buses, then falling as
    1/sqrt(n) (the total load grows with n faster than the grid's transfer capacity; at 50,000
    buses a flat 0.007 lands some cases on a voltage-collapse branch)."""
    return 0.007 * min(1.0, (10000.0 / max(int(n_bus), 1)) ** 0.5)


def sized_grid(n_bus: int, seed: int, pv_frac: float = 0.09, chord_frac: float = 0.3) -> Grid:
    """A grid of the multi-grid workloads: tree + chord_frac * n chords, pv_frac * n PV buses."""
    n = int(n_bus)
    return make_grid(n, int(chord_frac * n), max(1, int(round(pv_frac * n))), seed=seed, power_scale=size_scale(n))


def grid_family(count: int, lo: int, hi: int, seed: int) -> list:
    """`count` grids with sizes log-uniform in [lo, hi] buses (seeded), PV share U(8%, 10%)."""
    rng = np.random.default_rng(seed)
    sizes = np.exp(rng.uniform(np.log(lo), np.log(hi), count)).round().astype(int)
    pv = rng.uniform(0.08, 0.10, count)
    return [sized_grid(int(n), seed=int(seed * 1000 + k), pv_frac=float(f)) for k, (n, f) in enumerate(zip(sizes, pv))]
