"""fest_resnorm_device, the allocation-free norm entry point: through the C
ABI, its norms (written to a device buffer the caller owns, partial sums in
the library's static per-device scratch) equal fest_resnorm's bit for bit,
on two calls in a row, the second on a smaller grid (fewer partial sums than
the first call left in the shared scratch)."""
import ctypes
import sys

import numpy as np
import pytest

from festgpu import cases as C
from festgpu.native import GpuBlock, lib

pytestmark = pytest.mark.gpu


def test_resnorm_device_matches_resnorm():
    import torch
    L = lib()
    L.fest_resnorm_device.argtypes = [ctypes.c_void_p, ctypes.c_void_p]
    L.fest_resnorm_device.restype = ctypes.c_int
    for ni, nj, nk in ((40, 24, 16), (12, 8, 6)):
        print("resnorm_device", ni, nj, nk, flush=True)
        sys.stdout.flush()
        case, nodes, q = C.sst_bump(ni, nj, nk)
        g = GpuBlock(case)
        g.set_grid(nodes, wall_dist=True)
        g.upload_state(q)
        g.get_next_solution(1)
        ref = g.resnorm()
        out = torch.full((8,), float("nan"), dtype=torch.float64, device="cuda")
        rc = L.fest_resnorm_device(g.h, ctypes.c_void_p(out.data_ptr()))
        assert rc == 0, L.fest_last_error(g.h)
        g.sync()
        got = out.cpu().numpy()
        assert np.array_equal(got[: case.n_var + 1], ref), (got, ref)
        assert np.all(np.isnan(got[case.n_var + 1:]))  # nothing written past the norms
        g.close()
