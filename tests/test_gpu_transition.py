"""GPU parity of the lctm2015 transition model: the intermittency as
variable 8 (ghost rules, its gradient slot, face states and convective
flux, its viscous flux), Menter's 2015 F1 blending, the transition source
with the wall-normal sensor, and the RK update that leaves the
intermittency untouched (src/source.f90:273-463, viscous.f90:659-746,
viscosity.f90:265-279, 390-404, CC.f90, update.f90:456-484).  The HIP path
runs through the C ABI against the oracle, whose lctm2015 is pinned bit for
bit to the reference build (tests/test_oracle_vs_ref.py).  Tolerances: the
face fluxes, gradients, mu_t, F1 and the state 1e-12 relative; the residue
1e-10 per variable (the sources' exp and sqrt); the intermittency's ghost
cells bit for bit (copies and fixed values only)."""
import dataclasses
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
    a, b = np.asarray(a), np.asarray(b)
    s = np.abs(b).max()
    return np.abs(a - b).max() / s if s > 0 else np.abs(a - b).max()


def pair(case, nodes, q):
    o = Oracle(case)
    o.set_nodes(nodes)
    o.wall_dist()
    o.set_state(q)
    g = GpuBlock(case)
    g.set_grid(nodes, wall_dist=True)
    g.upload_state(q)
    return g, o


@pytest.mark.parametrize("turb", ["sst", "sst2003"])
@pytest.mark.parametrize("flux,recon", [("ausm", "muscl"), ("slau", "weno"), ("ldfss0", "ppm")])
def test_lctm2015_residual(turb, flux, recon):
    case, nodes, q = C.sst_bump(20, 12, 8, flux=flux, recon=recon, turb=turb, transition="lctm2015")
    say("lctm2015 residual", turb, flux, recon)
    g, o = pair(case, nodes, q)
    o.temp_refresh()
    assert o.residual(3) == 0
    g.residual(3)
    out = g.download("qp", "F", "G", "H", "residue", "mut", "f1", "grads")
    for n, on in (("F", "F"), ("G", "G"), ("H", "H")):
        assert relerr(out[n], o.arr(on)) <= TOL, n
    for l in range(8):
        assert relerr(out["residue"][l], o.arr("res")[l]) <= 1e-10, l
    inner = (slice(2, -2),) * 3
    assert relerr(out["mut"][inner], o.arr("mut")[inner]) <= TOL
    assert relerr(out["f1"][inner], o.arr("f1")[inner]) <= TOL
    gg = out["grads"]
    for d, n in enumerate(("gx", "gy", "gz")):
        assert relerr(gg[d], o.arr(n)) <= TOL, n
    # the intermittency's ghost cells: copies and fixed values, bit for bit
    assert np.array_equal(out["qp"][7], o.arr("qp")[7])
    g.close()


@pytest.mark.parametrize("bcs", [(-1, -2, -5, -6, -9, -9), (-3, -4, -7, -5, -8, -8), (-11, -4, -5, -8, -9, -9)])
def test_lctm2015_boundaries(bcs):
    case, nodes, q = C.sst_bump(16, 10, 6, flux="ausmUP", recon="muscl", transition="lctm2015")
    case = dataclasses.replace(case, bc_id=bcs, fix_tgm=(0.7, 0.6, 0.5, 0.4, 0.3, 0.2),
                               fix_tp=(1.02 * case.p_inf,) * 6)
    say("lctm2015 boundaries", bcs)
    g, o = pair(case, nodes, q)
    o.temp_refresh()
    assert o.residual(1) == 0
    g.residual(1)
    out = g.download("qp", "residue")
    assert np.array_equal(out["qp"][7], o.arr("qp")[7])
    assert relerr(out["qp"], o.arr("qp")) <= TOL
    for l in range(8):
        assert relerr(out["residue"][l], o.arr("res")[l]) <= 1e-10, l
    g.close()


@pytest.mark.parametrize("turb", ["sst", "sst2003"])
def test_lctm2015_rk4_steps(turb):
    """Three RK4 steps: the state within 1e-12, the intermittency unchanged
    in every interior cell, the norms (the intermittency's scaled by
    rho V gamma_inf) within 1e-10."""
    case, nodes, q = C.sst_bump(20, 12, 8, flux="slau", recon="weno", turb=turb, transition="lctm2015")
    say("lctm2015 RK4", turb)
    g, o = pair(case, nodes, q)
    for it in (1, 2, 3):
        assert o.step(it) == 0
        g.get_next_solution(it)
    got = g.download("qp")
    inner = (slice(None), slice(3, -3), slice(3, -3), slice(3, -3))
    assert relerr(got[inner], o.arr("qp")[inner]) <= TOL
    assert np.array_equal(got[7, 3:-3, 3:-3, 3:-3], q[7, 3:-3, 3:-3, 3:-3])
    np.testing.assert_allclose(g.resnorm(), o.resnorm(), rtol=1e-10)
    g.close()
