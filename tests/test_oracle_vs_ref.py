"""The oracle against the reference's own Fortran modules, built in place
with amdflang into oracle/_ref (the recipe is oracle/Makefile).  Every pass
of one residual evaluation -- ghost cells, face states, convective fluxes,
gradients, viscosity, viscous fluxes, residue and SST source -- must agree
bit for bit, as must the geometry and the wall distance.

Skipped when the reference build is absent (e.g. on a machine without the
reference sources); tests/golden holds fixtures from the same build.
"""
import dataclasses

import numpy as np
import pytest

from binding import Oracle, ref_available, run_reference
from festgpu import cases as C
from bc_cases import POLE, TOTAL_PRESSURE, pole_case, total_pressure_case

pytestmark = [pytest.mark.ref,
              pytest.mark.skipif(not ref_available(), reason="oracle/_ref/libfestref.so not built")]

NAMES = ["qp", "xl", "xr", "yl", "yr", "zl", "zr", "F", "G", "H", "gx", "gy", "gz", "mu", "res"]


def compare(case, nodes, q, it=1, wall_dist=False, extra=()):
    o = Oracle(case)
    o.set_nodes(nodes)
    if wall_dist:
        o.wall_dist()
    o.set_state(q)
    o.temp_refresh()
    temp, qin = o.arr("temp").copy(), o.arr("qp").copy()
    fn, dist = o.arr("nodes").copy(), o.arr("dist").copy()
    st = o.residual(it)
    assert st == 0
    names = NAMES + list(extra) + ["vol", "cx", "cy", "cz", "iA", "inx", "jny", "knz"]
    if wall_dist:
        names.append("dist")
    ref = run_reference(case, fn, qin, temp, 5, names, it=it, dist=None if wall_dist else dist,
                        wall_dist=wall_dist)
    bad = {}
    for n in names:
        a, b = o.arr(n), ref[n]
        if n in OUTER:
            # the left state of face 0 and the right state of face N+1 lie
            # outside every face the fluxes read; MUSCL leaves them unwritten
            a, b = a[OUTER[n]], b[OUTER[n]]
        if n in ("mut", "f1"):
            # computed over cells 0..N; the reference allocates mu_t without
            # a value (viscosity.f90:561) and never reads the outer layers
            a, b = a[2:-2, 2:-2, 2:-2], b[2:-2, 2:-2, 2:-2]
        if not np.array_equal(a, b, equal_nan=True):
            bad[n] = float(np.nanmax(np.abs(a - b)))
    assert not bad, bad


OUTER = {"xl": np.s_[..., 1:], "xr": np.s_[..., :-1], "yl": np.s_[..., 1:, :], "yr": np.s_[..., :-1, :],
         "zl": np.s_[:, 1:], "zr": np.s_[:, :-1]}


@pytest.mark.parametrize("flux", ["ausm", "ausmP", "ausmUP", "slau", "ldfss0", "van_leer"])
@pytest.mark.parametrize("recon", ["none", "muscl", "ppm", "weno"])
def test_sst_bump_all_schemes(flux, recon):
    case, nodes, q = C.sst_bump(14, 10, 6)
    case = dataclasses.replace(case, flux=flux, recon=recon)
    compare(case, nodes, q, it=3, wall_dist=True, extra=("mut", "f1"))


@pytest.mark.parametrize("bcs,tw,it", [
    ((-9, -9, -5, -5, -9, -9), (0, 0, 300, 300, 0, 0), 1),
    ((-3, -4, -6, -6, -8, -8), (0,) * 6, 1),
    ((-1, -2, -5, -5, -6, -6), (0, 0, -1, 0.5, 0, 0), 2),
    ((-3, -4, -5, -5, -9, -9), (0, 0, 350, -2, 0, 0), 3),
    ((-8, -8, -5, -8, -9, -9), (0,) * 6, 1),
])
def test_laminar_boundaries(bcs, tw, it):
    case, _, q = C.laminar_box(10, 8, 6)
    nodes = C.cartesian_nodes(case, warp=0.05)
    case = dataclasses.replace(case, bc_id=bcs, fix_twall=tw, fix_rho=(case.rho_inf * 1.01,) * 6,
                               fix_u=(case.u_inf,) * 6, fix_v=(3.0,) * 6, fix_w=(0.0,) * 6,
                               fix_p=(case.p_inf * 0.99,) * 6, mu_sutherland=1, accur=1)
    compare(case, nodes, q, it=it)


def test_laminar_box_warped():
    case, _, q = C.laminar_box(12, 10, 6)
    compare(case, C.cartesian_nodes(case, warp=0.03), q)


def compare_lusgs(case, nodes, q, it=1, wall_dist=False):
    """One 'implicit' time step: the oracle's residual, time step and LU-SGS
    update against the reference's residual and update_with_lusgs
    (src/lusgs.f90:134-183) fed the oracle's time step (time.f90 needs MPI
    and is not built)."""
    case = dataclasses.replace(case, time_accuracy="implicit")
    o = Oracle(case)
    o.set_nodes(nodes)
    if wall_dist:
        o.wall_dist()
    o.set_state(q)
    o.temp_refresh()
    temp, qin = o.arr("temp").copy(), o.arr("qp").copy()
    fn, dist = o.arr("nodes").copy(), o.arr("dist").copy()
    assert o.step(it) == 0
    ref = run_reference(case, fn, qin, temp, 5, ["qp", "res"], it=it, dist=None if wall_dist else dist,
                        wall_dist=wall_dist, lusgs_dt=o.arr("dt").copy())
    assert np.array_equal(o.arr("res"), ref["res"])
    got, want = o.arr("qp")[:, 3:-3, 3:-3, 3:-3], ref["qp"][:, 3:-3, 3:-3, 3:-3]
    assert not np.array_equal(got, qin[:, 3:-3, 3:-3, 3:-3])
    assert np.array_equal(got, want), float(np.abs(got - want).max())


@pytest.mark.parametrize("mu_mode", ["constant", "sutherland", "inviscid"])
@pytest.mark.parametrize("bcs", [(-9, -9, -5, -5, -9, -9), (-3, -4, -6, -5, -8, -8), (-1, -2, -5, -6, -6, -6)])
def test_lusgs_laminar(mu_mode, bcs):
    case, _, q = C.laminar_box(10, 8, 6)
    case = dataclasses.replace(case, bc_id=bcs, fix_rho=(case.rho_inf,) * 6, fix_u=(case.u_inf,) * 6,
                               fix_v=(2.0,) * 6, fix_w=(0.0,) * 6, fix_p=(case.p_inf,) * 6,
                               fix_twall=(0, 0, 320.0, 0, 0, 0), mu_sutherland=int(mu_mode == "sutherland"),
                               mu_ref=0.0 if mu_mode == "inviscid" else case.mu_ref)
    compare_lusgs(case, C.cartesian_nodes(case, warp=0.05), q)


@pytest.mark.parametrize("flux,recon", [("ausm", "muscl"), ("slau", "weno"), ("van_leer", "none")])
def test_lusgs_sst(flux, recon):
    case, nodes, q = C.sst_bump(14, 10, 6)
    case = dataclasses.replace(case, flux=flux, recon=recon)
    compare_lusgs(case, nodes, q, it=2, wall_dist=True)


# ------------------------------------------------------------------ 2-D
# One cell in k (kmx == 2): the reference zeroes the convective k flux
# (src/face/flux/convective/*.f90, e.g. ausm.f90:205-207), the z gradients
# (src/gradients.f90:328-336) and the SST viscous k flux
# (src/viscous.f90:101-105); ghost layers of a k face reach into the other
# face's.  The reference leaves part of its face-state arrays unwritten
# there, so the comparison covers everything downstream of them.
NAMES_2D = ["qp", "F", "G", "H", "res", "gx", "gy", "gz", "mu"]


def compare_2d(case, nodes, q, it=1, wall_dist=False, extra=()):
    o = Oracle(case)
    o.set_nodes(nodes)
    if wall_dist:
        o.wall_dist()
    o.set_state(q)
    o.temp_refresh()
    temp, qin = o.arr("temp").copy(), o.arr("qp").copy()
    fn, dist = o.arr("nodes").copy(), o.arr("dist").copy()
    assert o.residual(it) == 0
    names = NAMES_2D + list(extra)
    ref = run_reference(case, fn, qin, temp, 5, names, it=it, dist=None if wall_dist else dist,
                        wall_dist=wall_dist)
    cut = lambda n, a: a[2:-2, 2:-2, 2:-2] if n in ("mut", "f1") else a  # cells 0..N, as in compare()
    bad = {n: float(np.nanmax(np.abs(cut(n, o.arr(n)) - cut(n, ref[n])))) for n in names
           if not np.array_equal(cut(n, o.arr(n)), cut(n, ref[n]), equal_nan=True)}
    assert not bad, bad
    assert not o.arr("H")[0].any()  # no mass through the k faces


@pytest.mark.parametrize("flux,recon", [("ausm", "muscl"), ("slau", "weno"), ("ausmUP", "ppm"), ("van_leer", "none")])
@pytest.mark.parametrize("viscous", [True, False])
def test_two_dimensional(flux, recon, viscous):
    case, _, q = C.laminar_box(10, 8, 1)
    case = dataclasses.replace(case, flux=flux, recon=recon, bc_id=(-8, -4, -6, -6, -6, -6),
                               fix_p=(case.p_inf,) * 6, mu_ref=case.mu_ref if viscous else 0.0)
    nodes = C.cartesian_nodes(case, warp=0.04)
    compare_2d(case, nodes, q)
    compare_lusgs(case, nodes, q)


def test_two_dimensional_sst_periodic_k():
    """SST with a single-block periodic k direction of one cell: the
    periodic copy's source and destination layers overlap."""
    case, nodes, q = C.sst_bump(14, 10, 1)
    compare_2d(case, nodes, q, it=2, wall_dist=True, extra=("mut", "f1"))
    compare_lusgs(case, nodes, q, it=2, wall_dist=True)


# ------------------------------------------------------------ SST-2003
@pytest.mark.parametrize("flux,recon", [("slau", "weno"), ("ausm", "muscl")])
def test_sst2003(flux, recon):
    """The 2003 form of SST: strain-rate eddy viscosity with 1e-10 floors
    (src/viscosity.f90:215-275), production limited to 10 D_k and fixed
    gammas in the source (src/source.f90:205-211); everything else as SST."""
    case, nodes, q = C.sst_bump(14, 10, 6)
    case = dataclasses.replace(case, turb="sst2003", flux=flux, recon=recon)
    compare(case, nodes, q, it=3, wall_dist=True, extra=("mut", "f1"))
    compare_lusgs(case, nodes, q, it=2, wall_dist=True)


def test_sst2003_differs_from_sst():
    """Guard against the 2003 switch being a no-op: on the same state the
    eddy viscosity and the k and omega residues change."""
    case, nodes, q = C.sst_bump(14, 10, 6)
    outs = []
    for turb in ("sst", "sst2003"):
        o = Oracle(dataclasses.replace(case, turb=turb))
        o.set_nodes(nodes)
        o.wall_dist()
        o.set_state(q)
        o.temp_refresh()
        assert o.residual(1) == 0
        outs.append((o.arr("mut").copy(), o.arr("res").copy()))
    assert not np.array_equal(outs[0][0], outs[1][0])
    assert not np.array_equal(outs[0][1][5:], outs[1][1][5:])



@pytest.mark.parametrize("face,turb", TOTAL_PRESSURE)
def test_total_pressure_inlet(face, turb):
    case, nodes, q = total_pressure_case(face, turb)
    sst = turb != "none"
    compare(case, nodes, q, it=3, wall_dist=sst, extra=("mut", "f1") if sst else ())
    compare_lusgs(case, nodes, q, it=2, wall_dist=sst)


@pytest.mark.parametrize("bcs,recon", POLE)
def test_pole(bcs, recon):
    """Pole faces: zero area and the first interior face's normal on the
    pole face and its ghost faces, zero mass flux, the flat ghost copy."""
    case, nodes, q = pole_case(bcs, recon)
    compare(case, nodes, q, it=1)
    compare_lusgs(case, nodes, q, it=1)

def test_pole_sst():
    case, nodes, q = C.sst_bump(14, 10, 6, flux="ausm", recon="muscl")
    case = dataclasses.replace(case, bc_id=(-3, -4, -5, -7, -9, -9))
    compare(case, nodes, q, it=2, wall_dist=True, extra=("mut", "f1"))


@pytest.mark.parametrize("flux,recon", [("slau", "weno"), ("ausm", "muscl"), ("van_leer", "none")])
def test_spalart_allmaras(flux, recon):
    """SA on the bump channel: nu-tilde ghosts (anti-symmetric at the wall),
    mu_t = rho nu~ fv1, the nu-tilde viscous flux, the SA source, and the SA
    LU-SGS with its source Jacobian on the diagonal."""
    case, nodes, q = C.sst_bump(14, 10, 6, flux=flux, recon=recon, turb="sa")
    compare(case, nodes, q, it=3, wall_dist=True, extra=("mut",))
    compare_lusgs(case, nodes, q, it=2, wall_dist=True)


@pytest.mark.parametrize("bcs", [(-3, -4, -5, -8, -9, -9), (-1, -2, -5, -6, -6, -6), (-11, -4, -5, -5, -7, -7)])
def test_spalart_allmaras_boundaries(bcs):
    """Iteration 1, where the inlets impose the fixed nu-tilde, across the
    boundary kinds (total-pressure inlet and poles included)."""
    case, nodes, q = C.sst_bump(14, 10, 6, flux="ausmUP", recon="ppm", turb="sa")
    case = dataclasses.replace(case, bc_id=bcs, fix_tv=(1.3 * case.tv_inf,) * 6, fix_tp=(1.01 * case.p_inf,) * 6)
    compare(case, nodes, q, it=1, wall_dist=True, extra=("mut",))
    compare_lusgs(case, nodes, q, it=1, wall_dist=True)


def test_spalart_allmaras_two_dimensional():
    """One cell in k: the SA viscous flux still runs in k (viscous.f90:94-97)."""
    case, nodes, q = C.sst_bump(14, 10, 1, flux="slau", recon="muscl", turb="sa")
    compare_2d(case, nodes, q, it=2, wall_dist=True, extra=("mut",))


@pytest.mark.parametrize("flux,recon", [("slau", "weno"), ("ausm", "muscl"), ("van_leer", "none")])
def test_k_kl(flux, recon):
    """k-kL on the bump channel: k and kL ghosts (both anti-symmetric at the
    wall), mu_t = cmu^(1/4) rho kL / sqrt(k), the k-kL viscous flux, and the
    source with its von Karman length from second velocity derivatives
    built from the neighbours' gradients (src/source.f90:607-832); and the
    k-kL LU-SGS (its flux floors, sigma 1, the source terms on the
    diagonal, src/lusgs.f90:1198-1677)."""
    case, nodes, q = C.sst_bump(14, 10, 6, flux=flux, recon=recon, turb="kkl")
    compare(case, nodes, q, it=3, wall_dist=True, extra=("mut",))
    compare_lusgs(case, nodes, q, it=2, wall_dist=True)


@pytest.mark.parametrize("bcs", [(-3, -4, -5, -8, -9, -9), (-1, -2, -5, -6, -6, -6), (-11, -4, -5, -5, -7, -7)])
def test_k_kl_boundaries(bcs):
    """Iteration 1, where the inlets impose the fixed values: kL from the
    fixed kL values except at the subsonic inlet, which takes the omega
    values (0 under k-kL); the mu_t ghosts skip the poles."""
    case, nodes, q = C.sst_bump(14, 10, 6, flux="ausmUP", recon="ppm", turb="kkl")
    case = dataclasses.replace(case, bc_id=bcs, fix_tk=(1.2 * case.tk_inf,) * 6,
                               fix_tkl=(1.3 * case.tkl_inf,) * 6, fix_tp=(1.01 * case.p_inf,) * 6)
    compare(case, nodes, q, it=1, wall_dist=True, extra=("mut",))
    compare_lusgs(case, nodes, q, it=1, wall_dist=True)


def test_k_kl_two_dimensional():
    """One cell in k: the k-kL viscous flux still runs in k (viscous.f90:106-109)."""
    case, nodes, q = C.sst_bump(14, 10, 1, flux="slau", recon="muscl", turb="kkl")
    compare_2d(case, nodes, q, it=2, wall_dist=True, extra=("mut",))


def test_k_kl_subsonic_inlet_takes_the_omega_values():
    """Guard on the quirk above: with the subsonic inlet at imin, the kL
    ghosts there are the (zero) omega values, not the fixed kL values."""
    case, nodes, q = C.sst_bump(14, 10, 6, flux="ausm", recon="muscl", turb="kkl")
    o = Oracle(case)
    o.set_nodes(nodes)
    o.wall_dist()
    o.set_state(q)
    o.temp_refresh()
    assert o.residual(1) == 0
    kl = o.arr("qp")[6]
    assert np.all(kl[3:-3, 3:-3, 0:3] == 0.0)
    assert np.all(kl[3:-3, 3:-3, -3:] > 0.0)


@pytest.mark.parametrize("turb", ["sa", "sst", "sst2003"])
def test_transition_bc(turb):
    """The 'bc' transition: the intermittency gamma_BC in the SA and SST
    sources (src/source.f90:467-604, 985-1194), with the free-stream
    turbulence intensity and the Reynolds number of the free stream."""
    case, nodes, q = C.sst_bump(14, 10, 6, flux="ausm", recon="muscl", turb=turb)
    case = dataclasses.replace(case, transition="bc", tu_inf=0.8)
    extra = ("mut",) if turb == "sa" else ("mut", "f1")
    compare(case, nodes, q, it=3, wall_dist=True, extra=extra)
    compare_lusgs(case, nodes, q, it=2, wall_dist=True)


def test_transition_bc_changes_the_turbulence_residue():
    """Guard against the bc switch being a no-op on the cases above: the
    intermittency changes the first turbulence equation in every cell."""
    from binding import Oracle
    for turb in ("sa", "sst"):
        res = {}
        for tr in ("none", "bc"):
            case, nodes, q = C.sst_bump(14, 10, 6, flux="ausm", recon="muscl", turb=turb)
            case = dataclasses.replace(case, transition=tr, tu_inf=0.8)
            o = Oracle(case)
            o.set_nodes(nodes)
            o.wall_dist()
            o.set_state(q)
            o.temp_refresh()
            assert o.residual(3) == 0
            res[tr] = o.arr("res")[5].copy()
        assert (res["bc"] != res["none"]).all(), turb


@pytest.mark.parametrize("flux", ["ausmP", "slau", "ldfss0"])
def test_weno_nm(flux):
    """WENO for non-uniform grids (src/face/state/weno_NM.f90:21-147) on the
    stretched bump grid, SST with wall distance."""
    case, nodes, q = C.sst_bump(14, 10, 6)
    case = dataclasses.replace(case, flux=flux, recon="weno_NM")
    compare(case, nodes, q, it=3, wall_dist=True, extra=("mut", "f1"))


def test_weno_nm_laminar_boundaries():
    case, _, q = C.laminar_box(10, 8, 6)
    nodes = C.cartesian_nodes(case, warp=0.05)
    case = dataclasses.replace(case, recon="weno_NM", bc_id=(-3, -4, -6, -5, -8, -8),
                               fix_twall=(0, 0, 0, 300, 0, 0), fix_rho=(case.rho_inf * 1.01,) * 6,
                               fix_u=(case.u_inf,) * 6, fix_v=(3.0,) * 6, fix_w=(0.0,) * 6,
                               fix_p=(case.p_inf * 0.99,) * 6)
    compare(case, nodes, q, it=1)


@pytest.mark.parametrize("recon", ["muscl", "ppm"])
@pytest.mark.parametrize("pbs", [(1, 1, 1), (1, 0, 1), (0, 1, 0)])
def test_pressure_based_switch(recon, pbs):
    """The pressure-based switch of MUSCL and PPM (src/face/state/muscl.f90:
    37-112, ppm.f90:109-186) with the boundary overrides after it, SST."""
    case, nodes, q = C.sst_bump(14, 10, 6)
    case = dataclasses.replace(case, flux="ausmUP", recon=recon, pbs=pbs)
    compare(case, nodes, q, it=3, wall_dist=True, extra=("mut", "f1"))


@pytest.mark.parametrize("recon", ["muscl", "ppm", "weno"])
def test_pressure_based_switch_laminar(recon):
    """Laminar box with walls, far field and inlets; under WENO the switch
    is read and not applied, as in compute_face_interpolant."""
    case, _, q = C.laminar_box(10, 8, 6)
    nodes = C.cartesian_nodes(case, warp=0.05)
    case = dataclasses.replace(case, recon=recon, pbs=(1, 1, 1), bc_id=(-1, -2, -5, -6, -9, -9),
                               fix_twall=(0, 0, 310, 0, 0, 0), fix_rho=(case.rho_inf * 1.01,) * 6,
                               fix_u=(case.u_inf,) * 6, fix_v=(3.0,) * 6, fix_w=(0.0,) * 6,
                               fix_p=(case.p_inf * 0.99,) * 6, accur=1)
    compare(case, nodes, q, it=1)


def test_lusgs_weno_nm_pressure_switch():
    case, nodes, q = C.sst_bump(12, 8, 5)
    compare_lusgs(dataclasses.replace(case, recon="weno_NM"), nodes, q, it=2, wall_dist=True)
    compare_lusgs(dataclasses.replace(case, recon="muscl", pbs=(1, 1, 1)), nodes, q, it=2, wall_dist=True)


@pytest.mark.parametrize("turb", ["sst", "sst2003"])
@pytest.mark.parametrize("flux,recon", [("ausm", "muscl"), ("slau", "weno"), ("ldfss0", "ppm")])
def test_transition_lctm2015(turb, flux, recon):
    """lctm2015: the intermittency as variable 8 -- its ghost cells, face
    states and convective flux, its gradient slot, Menter's 2015 F1
    blending (src/viscosity.f90:265-279, 390-404), its viscous flux
    (src/viscous.f90:659-746) and the transition source with the CC.f90
    pressure-gradient sensor (src/source.f90:273-463) -- bit for bit."""
    case, nodes, q = C.sst_bump(14, 10, 6, flux=flux, recon=recon, turb=turb, transition="lctm2015")
    compare(case, nodes, q, it=3, wall_dist=True, extra=("mut", "f1"))


@pytest.mark.parametrize("bcs", [(-1, -2, -5, -6, -9, -9), (-3, -4, -7, -5, -8, -8), (-11, -4, -5, -8, -9, -9)])
def test_transition_lctm2015_boundaries(bcs):
    """The intermittency's ghost rules: fixed at inlets, a flat copy at
    outlets, walls, slip walls and the pole, the turbulence variables'
    inflow/outflow choice at far-field and total-pressure faces."""
    case, nodes, q = C.sst_bump(12, 9, 6, flux="ausmUP", recon="muscl", transition="lctm2015")
    case = dataclasses.replace(case, bc_id=bcs, fix_tgm=(0.7, 0.6, 0.5, 0.4, 0.3, 0.2),
                               fix_tp=(1.02 * case.p_inf,) * 6)
    compare(case, nodes, q, it=1, wall_dist=True, extra=("mut", "f1"))


def test_transition_lctm2015_changes_the_residue():
    """Guard against a silent no-op: against plain SST on the same state,
    the transition changes the k and omega residues (F1 blending, the
    intermittency-weighted production) and gives variable 8 a residue."""
    res = {}
    for tr in ("none", "lctm2015"):
        case, nodes, q = C.sst_bump(14, 10, 6, flux="ausm", recon="muscl", transition=tr)
        o = Oracle(case)
        o.set_nodes(nodes)
        o.wall_dist()
        o.set_state(q)
        o.temp_refresh()
        assert o.residual(3) == 0
        res[tr] = o.arr("res").copy()
    assert (res["lctm2015"][5] != res["none"][5]).all()
    assert (res["lctm2015"][6] != res["none"][6]).mean() > 0.5
    assert np.abs(res["lctm2015"][7]).max() > 0
