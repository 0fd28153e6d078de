"""edm_cell_pairs_noalloc, the allocation-free pair-list build, against
edm_cell_pairs (which allocates its scratch per call) through the C ABI:
the same pairs, count and found count, bit for bit, on two calls in a row
(the second smaller, so stale scratch from the first would show), and the
size limit refused with a clear error."""

import ctypes

import numpy as np
import pytest

from ed_mrdt import _hip
from ed_mrdt.free_kernels import DeviceArray

pytestmark = pytest.mark.gpu


def _call(name, pos, n, L, cutoff, cap):
    nx = ny = max(3, int(L // cutoff))
    out = DeviceArray(np.full((cap, 2), -7, dtype=np.int32))
    np_, nf = ctypes.c_int64(-1), ctypes.c_int64(-1)
    rc = getattr(_hip.lib(), name)(pos.ptr, n, nx, ny, L / nx, L / ny, L, L, 1, 1, cutoff, cap,
                                   out.ptr, ctypes.byref(np_), ctypes.byref(nf), None)
    _hip.check(rc, name)
    return out.numpy(), np_.value, nf.value


def test_noalloc_pairs_equal_the_allocating_entry_point():
    g = np.random.default_rng(12)
    for n, L, cap in ((60000, 20.0, 60000 * 30), (3000, 4.0, 1000)):
        pos = DeviceArray(g.uniform(0.0, L, (n, 2)))
        want, wn, wf = _call("edm_cell_pairs", pos, n, L, 0.165, cap)
        got, gn, gf = _call("edm_cell_pairs_noalloc", pos, n, L, 0.165, cap)
        assert (gn, gf) == (wn, wf) and wn > 0
        np.testing.assert_array_equal(got, want)
        print(f"n={n}: {gn} pairs stored of {gf}")


def test_noalloc_refuses_sizes_beyond_its_scratch():
    L = _hip.lib()
    pos = DeviceArray(np.zeros((4, 2)))
    out = DeviceArray(np.zeros((4, 2), dtype=np.int32))
    np_, nf = ctypes.c_int64(0), ctypes.c_int64(0)
    for n, nx, ny in (((1 << 20) + 1, 3, 3), (4, 2048, 1024)):
        rc = L.edm_cell_pairs_noalloc(pos.ptr, n, nx, ny, 1.0, 1.0, 3.0, 3.0, 1, 1, 0.5, 4,
                                      out.ptr, ctypes.byref(np_), ctypes.byref(nf), None)
        assert rc == -1 and b"static scratch" in L.edm_last_error(), (n, nx, ny)
