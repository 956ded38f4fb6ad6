"""Host side of the device Sbus builder (p3s/timeseries.py:36-141): the contribution list
and the numpy restatement used as the checker, pinned by a scalar brute force."""

import numpy as np
import pytest

from oracle.sbus import build_sbus
from p3smi.timeseries import SbusBuilder


def _case(rng, n=7, T=5):
    sources = [
        {"bus": np.array([3, 1, -1, 3, 6]), "coef": rng.uniform(-2, -0.5, 5), "base": rng.uniform(0, 1, 5),
         "imag": False},
        {"bus": np.array([3, 1, -1, 3, 6]), "coef": rng.uniform(-2, -0.5, 5), "base": rng.uniform(0, 1, 5),
         "imag": True},
        {"bus": np.array([0, 3]), "coef": rng.uniform(0.5, 2, 2), "base": rng.uniform(0, 1, 2), "imag": False},
    ]
    series = [rng.uniform(0, 2, (5, T)), rng.uniform(0, 1, (5, T)), rng.uniform(0, 3, (2, T))]
    s0 = rng.normal(size=n) + 1j * rng.normal(size=n)
    return n, T, s0, sources, series


def _brute(n, T, s0, sources, series):
    out = np.zeros((n, T), dtype=np.complex128)
    for i in range(n):
        for t in range(T):
            re, im = float(s0[i].real), float(s0[i].imag)
            for si, src in enumerate(sources):  # source order, then row order: stable bus-major
                for r, b in enumerate(src["bus"]):
                    if b != i:
                        continue
                    v = float(src["coef"][r]) * (float(series[si][r][t]) - float(src["base"][r]))
                    if src["imag"]:
                        im = im + v
                    else:
                        re = re + v
            out[i, t] = complex(re, im)
    return out


def test_restatement_equals_brute_force():
    rng = np.random.default_rng(3)
    n, T, s0, sources, series = _case(rng)
    a = build_sbus(n, s0, sources, series)
    b = _brute(n, T, s0, sources, series)
    assert np.array_equal(a.view(np.float64), b.view(np.float64))
    # a plausible mistake (one delta taken from the wrong side) must show
    bad = [dict(src) for src in sources]
    bad[2]["coef"] = -bad[2]["coef"]
    assert not np.array_equal(build_sbus(n, s0, bad, series), b)


def test_contribution_list_is_bus_major_and_stable():
    rng = np.random.default_rng(4)
    n, T, s0, sources, series = _case(rng)
    sb = SbusBuilder(n, s0, sources)
    assert sb.ptr[0] == 0 and sb.ptr[-1] == sb.row.size == 10  # 12 rows, two unsupplied
    X = sb.stack(series)
    # rebuild from the list on the host, in the list's order: equals the restatement
    re = np.repeat(s0.real[:, None], T, axis=1)
    im = np.repeat(s0.imag[:, None], T, axis=1)
    for i in range(n):
        for k in range(sb.ptr[i], sb.ptr[i + 1]):
            v = sb.coef[k] * (X[sb.row[k]] - sb.base[k])
            if sb.imag[k]:
                im[i] = im[i] + v
            else:
                re[i] = re[i] + v
    got = re + 1j * im
    assert np.array_equal(got, build_sbus(n, s0, sources, series))
    # within bus 3: source 0 rows 0, 3, then source 1 rows 0, 3, then source 2 row 1
    assert sb.row[sb.ptr[3]:sb.ptr[4]].tolist() == [0, 3, 5, 8, 11]


def test_builder_rejects_bad_inputs():
    with pytest.raises(ValueError):
        SbusBuilder(3, np.zeros(2, complex), [])
    with pytest.raises(ValueError):
        SbusBuilder(3, np.zeros(3, complex), [{"bus": [5], "coef": [1.0], "base": [0.0], "imag": False}])
    sb = SbusBuilder(3, np.zeros(3, complex), [{"bus": [1, 2], "coef": [1.0, 1.0], "base": [0.0, 0.0], "imag": False}])
    with pytest.raises(ValueError):
        sb.stack([np.zeros((3, 4))])
    with pytest.raises(ValueError):
        sb.blocks([np.zeros((1, 4)), np.zeros((2, 4))])  # 3 rows for 2 elements
    with pytest.raises(ValueError):
        sb.blocks([np.zeros((1, 4)), np.zeros((1, 5))])  # series of different lengths
    with pytest.raises(ValueError):
        sb.blocks([np.zeros(8)])
    X = sb.blocks([np.arange(4.0)[None, :], np.ones((1, 4), np.float32)])
    assert [x.dtype for x in X] == [np.float64, np.float64] and all(x.flags.c_contiguous for x in X)
