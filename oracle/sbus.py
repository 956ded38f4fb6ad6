"""TEST INFRASTRUCTURE ONLY: numpy restatement of the reference's Sbus time-series builder
(p3s/timeseries.py:36-141, the `_fill_sbus` loop at :118-141), used as the checker of the
device builder. out[i, t] = s0[i] + sum_k coef[k] * (x_k[t] - base[k]) over the
contributions of bus i in bus-major order (element order kept within a bus), imaginary
contributions on the imaginary part; every product and sum rounded one by one.
Pinned by a brute-force scalar loop over (bus, case) in tests/test_sbus_cpu.py."""

import numpy as np


def build_sbus(n, s0, sources, series):
    """sources: dicts (bus, coef, base, imag) per series; series: (rows, T) per source."""
    T = series[0].shape[1]
    contrib = []  # (bus, k order, source, row, coef, base, imag)
    for si, (src, x) in enumerate(zip(sources, series)):
        for r, b in enumerate(np.asarray(src["bus"])):
            if b < 0:
                continue
            contrib.append((int(b), si, r, float(src["coef"][r]), float(src["base"][r]), bool(src["imag"])))
    order = sorted(range(len(contrib)), key=lambda k: (contrib[k][0], k))  # stable bus-major
    out_re = np.repeat(np.real(s0)[:, None], T, axis=1).astype(np.float64)
    out_im = np.repeat(np.imag(s0)[:, None], T, axis=1).astype(np.float64)
    for k in order:
        b, si, r, c, base, im = contrib[k]
        v = c * (np.asarray(series[si][r], dtype=np.float64) - base)
        if im:
            out_im[b] = out_im[b] + v
        else:
            out_re[b] = out_re[b] + v
    out = np.empty((n, T), dtype=np.complex128)
    out.real = out_re
    out.imag = out_im
    return out
