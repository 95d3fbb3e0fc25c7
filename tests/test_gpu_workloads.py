"""The benchmark workloads beyond the SST bump, against the oracle: the
turbulent flat plate (PPM + LDFSS(0), SST, y+ about 1 clustering, a 1/7
power-law layer in the initial state) shrunk and at its full 512x256x128
size, and the 64-block chain's generator at a small size (the batched
tests run the chain itself).  Tolerance 1e-12 relative."""
import sys

import numpy as np
import pytest

from binding import Oracle
from festgpu import cases as C
from festgpu.native import GpuBlock

pytestmark = pytest.mark.gpu
TOL = 1e-12


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def relerr(a, b):
    s = np.abs(b).max()
    return np.abs(a - b).max() / s


def test_flat_plate_small_steps():
    case, nodes, q = C.flat_plate(40, 24, 8)
    say("flat plate 40x24x8, 3 RK4 steps")
    g = GpuBlock(case)
    g.set_grid(nodes, wall_dist=True)
    g.upload_state(q)
    o = Oracle(case)
    o.set_nodes(nodes)
    o.wall_dist()
    o.set_state(q)
    assert np.array_equal(g.download("dist"), o.arr("dist"))
    for it in range(1, 4):
        assert o.step(it) == 0
        g.get_next_solution(it)
    out = g.download("qp", "residue", "dt")
    inner = (slice(None), slice(3, -3), slice(3, -3), slice(3, -3))
    assert relerr(out["qp"][inner], o.arr("qp")[inner]) <= TOL
    assert relerr(out["residue"], o.arr("res")) <= TOL
    assert relerr(out["dt"], o.arr("dt")) <= TOL
    g.close()


def test_flat_plate_full_size_residual():
    """One residual evaluation of the 16.8M-cell plate (512x256x128) against
    the oracle; the wall distance is the device's (bit-exact with the
    oracle's at the sizes the oracle finishes, and O(cells x wall nodes)
    on one host core here)."""
    case, nodes, q = C.flat_plate()
    say("flat plate 512x256x128, one residual")
    g = GpuBlock(case)
    g.set_grid(nodes, wall_dist=True)
    g.upload_state(q)
    g.residual(1)
    out = g.download("residue", "dist", "F", "G", "H")
    g.close()
    o = Oracle(case)
    o.set_nodes(nodes)
    o.arr("dist")[...] = out["dist"]
    o.set_state(q)
    o.temp_refresh()
    assert o.residual(1) == 0
    assert relerr(out["residue"], o.arr("res")) <= TOL
    for n in ("F", "G", "H"):
        assert relerr(out[n], o.arr(n)) <= TOL, n
