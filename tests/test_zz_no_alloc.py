"""fest_resnorm_device, the allocation-free norm entry point: through the C
ABI, its norms (written to a device buffer the caller owns, partial sums in
the library's static per-device scratch) equal fest_resnorm's bit for bit,
on two calls in a row, the second on a smaller grid (fewer partial sums than
the first call left in the shared scratch).  The caller's buffer comes from
the HIP runtime the library links: hipMalloc is looked up through the
library's handle, whose dependency tree holds that runtime (a process that
has imported torch also maps torch's bundled copy of the runtime under
another file name, and a lookup by file name may find that one).  The test
needs nothing but the library."""
import ctypes
import sys

import numpy as np
import pytest

from festgpu import cases as C
from festgpu.native import GpuBlock, lib

pytestmark = pytest.mark.gpu

H2D, D2H = 1, 2  # hipMemcpyHostToDevice, hipMemcpyDeviceToHost


def hip():
    rt = lib()  # dlsym through libfest_hip.so reaches the runtime it links
    rt.hipMalloc.argtypes = [ctypes.POINTER(ctypes.c_void_p), ctypes.c_size_t]
    rt.hipFree.argtypes = [ctypes.c_void_p]
    rt.hipMemcpy.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t, ctypes.c_int]
    for f in (rt.hipMalloc, rt.hipFree, rt.hipMemcpy):
        f.restype = ctypes.c_int
    return rt


def test_resnorm_device_matches_resnorm():
    L = lib()
    L.fest_resnorm_device.argtypes = [ctypes.c_void_p, ctypes.c_void_p]
    L.fest_resnorm_device.restype = ctypes.c_int
    rt = hip()
    for ni, nj, nk in ((40, 24, 16), (12, 8, 6)):
        print("resnorm_device", ni, nj, nk, flush=True)
        sys.stdout.flush()
        case, nodes, q = C.sst_bump(ni, nj, nk)
        g = GpuBlock(case)
        g.set_grid(nodes, wall_dist=True)
        g.upload_state(q)
        g.get_next_solution(1)
        ref = g.resnorm()
        buf = np.full(8, np.nan)
        dev = ctypes.c_void_p()
        assert rt.hipMalloc(ctypes.byref(dev), buf.nbytes) == 0
        try:
            assert rt.hipMemcpy(dev, buf.ctypes.data, buf.nbytes, H2D) == 0
            rc = L.fest_resnorm_device(g.h, dev)
            assert rc == 0, L.fest_last_error(g.h)
            g.sync()
            got = np.empty(8)
            assert rt.hipMemcpy(got.ctypes.data, dev, got.nbytes, D2H) == 0
        finally:
            rt.hipFree(dev)
        assert np.array_equal(got[: case.n_var + 1], ref), (got, ref)
        assert np.all(np.isnan(got[case.n_var + 1:]))  # nothing written past the norms
        g.close()
