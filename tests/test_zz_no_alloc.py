"""p3smi_outages_islands_noalloc: the islands of an outage set computed into caller-owned
device arrays with the library's static scratch (no device allocation by the call). Two calls
in a row, the second on a smaller set, each equal bit for bit to p3smi_outages_islands. The
test's own device buffers come from the HIP runtime through ctypes (plumbing only)."""

import ctypes as C

import numpy as np
import pytest

from p3smi import _lib, contingency, grids

pytestmark = pytest.mark.gpu


class DeviceArray:
    """A device buffer for a numpy shape/dtype, filled with a value, read back on demand."""

    _hip = None

    def __init__(self, shape, dtype, fill):
        if DeviceArray._hip is None:
            DeviceArray._hip = C.CDLL("libamdhip64.so")
        self.host = np.full(shape, fill, dtype=dtype)
        self.ptr = C.c_void_p()
        assert self._hip.hipMalloc(C.byref(self.ptr), C.c_size_t(self.host.nbytes)) == 0
        assert self._hip.hipMemcpy(self.ptr, self.host.ctypes.data_as(C.c_void_p), C.c_size_t(self.host.nbytes),
                                   1) == 0  # host to device

    def get(self):
        out = np.empty_like(self.host)
        assert self._hip.hipMemcpy(out.ctypes.data_as(C.c_void_p), self.ptr, C.c_size_t(out.nbytes), 2) == 0
        return out

    def free(self):
        self._hip.hipFree(self.ptr)


def _noalloc(cs, n, L, with_ref=True):
    served = DeviceArray((n, L), np.uint8, 9)
    ref = DeviceArray((n, L), np.int32, -7) if with_ref else None
    pinned = DeviceArray((L,), np.int32, 123)  # must be cleared by the call
    try:
        _lib.check(_lib.lib().p3smi_outages_islands_noalloc(cs._h, served.ptr, ref.ptr if with_ref else None,
                                                            pinned.ptr), "islands_noalloc: ")
        return served.get().astype(bool), (ref.get() if with_ref else None), pinned.get()
    finally:
        for a in (served, ref, pinned):
            if a is not None:
                a.free()


def test_islands_noalloc_equals_islands(gpu):
    # ts2000 with re-slacking: 2,599 cases of 2,000 buses, more than one scratch chunk
    # (64 MiB / 32 kB per case); then the 200-bus set, smaller, on the same scratch
    cs = None
    for name, reslack in (("ts2000", True), ("pf200", True), ("pf200", False)):
        g = grids.config_grid(name, 0)
        cs = contingency.from_grid(g, reslack=reslack)
        served, ref_bus, pinned = cs.islands()
        s2, r2, p2 = _noalloc(cs, g.n, cs.L)
        np.testing.assert_array_equal(s2, served, err_msg=name)
        np.testing.assert_array_equal(r2, ref_bus, err_msg=name)
        np.testing.assert_array_equal(p2, pinned, err_msg=name)
        assert (~served).any()
        s3, _, p3 = _noalloc(cs, g.n, cs.L, with_ref=False)
        np.testing.assert_array_equal(s3, served)
        np.testing.assert_array_equal(p3, pinned)
    with pytest.raises(_lib.P3smiError, match="null"):
        _lib.check(_lib.lib().p3smi_outages_islands_noalloc(cs._h, None, None, None))
