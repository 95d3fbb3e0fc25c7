"""Pins the oracle (the C restatement in oracle/) with the reference's own
known-answer tests and with closed forms.

Each KAT below re-expresses one of the reference's unit tests on the same
inputs and checks the same intervals or exact values:
tests/test_{ausm,ausmp,ausmUP,slau,ldfss,van_leer}.f90 (face fluxes),
tests/test_{muscl,weno,ppm,first_order}.f90 (face states),
tests/test_residue.f90, tests/test_gradient.f90, tests/test_time.f90,
tests/test_viscosity.f90.  The closed forms cover what those leave open:
flux consistency F(q,q) = Euler flux, freestream preservation on a warped
grid, the uniform-grid time step with its viscous correction, the RK4
composition of the step and the norm definition.
"""
import dataclasses
import math

import numpy as np
import pytest

from binding import Oracle
from festgpu import cases as C
from props import boundary_flux_mismatch

PROFILE = np.array([-2.0, -1.0, 0.0, 1.0, 2.0, 3.0, 5.0, 7.0, 7.0, 7.0])  # qp(-2..7,1,1,:)


def tiny_case(imx, jmx, kmx, **kw):
    base = dict(flux="ausm", recon="muscl", turb="none", bc_id=(0,) * 6, accur=1, lim=(1, 1, 1), tlim=(1, 1, 1),
                rho_inf=1.0, u_inf=0.0, v_inf=0.0, w_inf=0.0, p_inf=1.0, vel_mag=1.0, mu_ref=0.0, gm=1.4)
    base.update(kw)
    return C.Case(imx, jmx, kmx, **base)


# ------------------------------------------------------------------ fluxes
FLUX_KAT = {  # scheme: interval of F(2,1,1,2) (tests/test_<scheme>.f90)
    "ausm": (4.48, 4.5), "ausmP": (4.5, 4.6), "ausmUP": (5.0, 5.2),
    "slau": (4.5, 4.7), "ldfss0": (4.9, 5.1), "van_leer": (4.9, 5.1),
}


@pytest.mark.parametrize("scheme", list(FLUX_KAT))
def test_flux_kat(scheme):
    o = Oracle(tiny_case(2, 2, 2, flux=scheme, MInf=0.0))
    q1 = np.array([1.0, 2.0, 3.0, 4.0, 5.0])
    f1 = o.face_flux(q1, q1, (1.0, 0.0, 0.0))
    assert 8.9 < f1[1] < 9.1
    f2 = o.face_flux([1.0, 2.0, 0.0, 0.0, 1.0], [0.125, 0.0, 0.0, 0.0, 0.1], (1.0, 0.0, 0.0))
    lo, hi = FLUX_KAT[scheme]
    assert lo < f2[1] < hi, (scheme, f2[1])


def euler_flux(q, n, gm=1.4):
    r, u, v, w, p = q[:5]
    vn = u * n[0] + v * n[1] + w * n[2]
    H = 0.5 * (u * u + v * v + w * w) + gm / (gm - 1) * p / r
    return np.array([r * vn, r * vn * u + p * n[0], r * vn * v + p * n[1], r * vn * w + p * n[2], r * vn * H])


@pytest.mark.parametrize("scheme", list(FLUX_KAT))
@pytest.mark.parametrize("mach", [0.3, -0.7, 1.6, -2.2])
def test_flux_consistency(scheme, mach):
    """F(q, q) is the exact Euler flux; for |M| > 1 all schemes upwind."""
    o = Oracle(tiny_case(2, 2, 2, flux=scheme, MInf=0.5))
    n = np.array([0.6, 0.48, 0.64])
    c = math.sqrt(1.4 * 2.0 / 1.3)
    q = np.array([1.3, mach * c * n[0] + 0.1 * n[1], mach * c * n[1] - 0.1 * n[0], mach * c * n[2], 2.0])
    f = o.face_flux(q, q, n, A=1.7)
    np.testing.assert_allclose(f, 1.7 * euler_flux(q, n), rtol=1e-12, atol=1e-12)
    if abs(mach) > 1 and scheme != "ausmUP":
        q2 = q.copy()
        q2[0] *= 0.9
        q2[4] *= 1.1
        up, dn = (q, q2) if mach > 0 else (q2, q)
        np.testing.assert_allclose(o.face_flux(up, dn, n), euler_flux(q, n), rtol=1e-12, atol=1e-12)


def test_flux_zero_mass_flag():
    """Wall faces carry no mass flux; the pressure flux remains (src/boundary/bc.f90:53-66)."""
    o = Oracle(tiny_case(2, 2, 2, flux="ausmP"))
    qL = np.array([1.0, 0.3, 0.1, 0.0, 1.0])
    qR = np.array([1.1, 0.2, 0.0, 0.1, 1.2])
    f = o.face_flux(qL, qR, (1.0, 0.0, 0.0), zero=0)
    assert f[0] == 0.0 and f[4] == 0.0
    assert f[1] != 0.0


# ---------------------------------------------------------- face states
def states_1d(recon, lim):
    case = tiny_case(5, 2, 2, recon=recon, lim=(lim,) * 3, tlim=(lim,) * 3)
    o = Oracle(case)
    o.arr("vol")[...] = 1.0  # cells(:,:,:)%volume = 1 (test_weno_NM.f90:33)
    q = o.arr("qp")
    q[...] = 1.0
    for l in range(5):
        q[l, 3, 3, :] = PROFILE
    o.stage(1, 2)
    return o.arr("xl")[0, 0, 0, :].copy(), o.arr("xr")[0, 0, 0, :].copy()


def test_muscl_kat():
    L, R = states_1d("muscl", 0)  # tests/test_muscl.f90:27-49 (limiter off)
    assert L[1] == 0.5 and R[1] == 0.5
    assert L[2] == 1.5 and R[2] == 1.5
    assert L[3] == 2.5 and 2.3 < R[3] < 2.4
    assert 3.82 < L[4] < 3.85 and R[4] == 4.0
    assert L[5] == 6.0 and 6.32 < R[5] < 6.35
    # unlimited kappa=1/3 MUSCL is (-q[i-1] + 5 q[i] + 2 q[i+1])/6 at the upper face
    q = PROFILE  # index 0 is cell -2
    for f in range(1, 6):
        i = f - 1 + 2
        assert L[f] == pytest.approx((-q[i - 1] + 5 * q[i] + 2 * q[i + 1]) / 6, rel=1e-14)


def test_weno_kat():
    L, R = states_1d("weno", 1)  # tests/test_weno.f90:19-42
    assert 0.48 <= L[1] < 0.52 and 0.48 < R[1] < 0.52
    assert 1.48 <= L[2] < 1.52 and 1.48 < R[2] < 1.52
    assert 2.4 <= L[3] < 2.5 and 2.4 < R[3] < 2.5
    assert 3.6 <= L[4] < 3.7 and 3.8 < R[4] < 4.1
    assert 5.9 <= L[5] < 6.1 and 6.9 < R[5] < 7.1


def test_weno_linear_exact():
    """On linear data every smoothness indicator is equal, the weights are the
    linear ones and the reconstruction is exact."""
    case = tiny_case(8, 2, 2, recon="weno")
    o = Oracle(case)
    q = o.arr("qp")
    q[...] = 1.0
    for l in range(5):
        q[l, 3, 3, :] = 0.25 * np.arange(-2, 11) + 3.0
    o.stage(1, 2)
    L, R = o.arr("xl")[0, 0, 0, :], o.arr("xr")[0, 0, 0, :]
    faces = np.arange(0, 10)
    expect = 0.25 * (faces - 0.5) + 3.0  # face f between cells f-1 and f
    np.testing.assert_allclose(L[1:9], expect[1:9], rtol=1e-13)
    np.testing.assert_allclose(R[0:9], expect[0:9], rtol=1e-13)


def test_ppm_kat():
    L, R = states_1d("ppm", 1)  # tests/test_ppm.f90:26-49
    assert 0.48 < L[1] < 0.52 and 0.48 < R[1] < 0.52
    assert 1.48 < L[2] < 1.52 and 1.48 < R[2] < 1.52
    assert 2.3 < L[3] < 2.5 and 2.3 < R[3] < 2.5
    assert 3.8 < L[4] < 4.0 and 3.8 < R[4] < 4.0
    assert 6.0 < L[5] < 6.2 and 6.5 < R[5] < 6.7


def test_first_order_kat():
    L, R = states_1d("none", 1)  # tests/test_first_order.f90:24-47
    for f in range(0, 7):
        assert L[f] == PROFILE[f - 1 + 2] and R[f] == PROFILE[f + 2]


# ------------------------------------------------------- residue, gradient
def test_residue_kat():
    o = Oracle(tiny_case(2, 2, 2))  # tests/test_residue.f90:19-31
    F, G, H = o.arr("F"), o.arr("G"), o.arr("H")
    F[:, 0, 0, 0] = [1.0, 2.0, 3.0, 4.0, 5.0]
    F[:, 0, 0, 1] = [2.0, 2.0, 9.0, 0.0, 5.1]
    G[...] = 1.0
    H[...] = 1.0
    o.stage(1, 16)
    r = o.arr("res")[:, 0, 0, 0]
    assert r[0] == 1.0 and r[1] == 0.0 and r[2] == 6.0 and 0.09 < r[4] < 1.1


def test_residue_telescopes_to_boundary_flux():
    """The oracle's mean-flow residues on the SST bump sum to the net
    boundary flux (the property the GPU checks at full size); a residue
    scattered with a shifted face index breaks it."""
    case, nodes, q = C.sst_bump(24, 12, 8)
    o = Oracle(case)
    o.set_nodes(nodes)
    o.wall_dist()
    o.set_state(q)
    o.temp_refresh()
    assert o.residual(1) == 0
    F, G, H, res = (o.arr(n).copy() for n in ("F", "G", "H", "res"))
    for l in range(5):
        assert boundary_flux_mismatch(F, G, H, res, l) <= 1e-13, l
    # planted mistake: the residue built from G shifted by one j face
    bad = res.copy()
    bad[:, :, :-1, :] += G[:, :, 1:-1, :] - G[:, :, 2:, :]
    assert max(boundary_flux_mismatch(F, G, H, bad, l) for l in range(5)) > 1e-6


def test_gradient_kat():
    case = tiny_case(2, 2, 2, bc_id=(3,) * 6, mu_ref=1.0)  # tests/test_gradient.f90:35-62
    o = Oracle(case)
    o.arr("vol")[...] = 2.0
    o.arr("iA")[...] = 2.0
    o.arr("inx")[...] = 1.0
    o.arr("jA")[...] = 1.0
    o.arr("jny")[...] = 1.0
    o.arr("kA")[...] = 1.0
    o.arr("knz")[...] = 1.0
    q = o.arr("qp")
    q[...] = 1.0
    for i, v in ((0, 2.0), (1, 4.0), (2, 8.0), (3, 16.0)):
        q[:, :, :, i + 2] = v
    o.arr("temp")[...] = 1.0
    o.stage(1, 8)
    gx = o.arr("gx")
    assert gx[0, 1, 1, 0] == 1.5 and gx[0, 1, 1, 1] == 3.0 and gx[0, 1, 1, 2] == 6.0


def test_gradient_linear_exact():
    """Green-Gauss is exact for a linear field on a Cartesian grid."""
    case = tiny_case(7, 6, 5, bc_id=(-9,) * 6, mu_ref=1e-3, recon="muscl")
    o = Oracle(case)
    o.set_nodes(C.cartesian_nodes(case, spacing=(0.5, 0.7, 0.9)))
    cx, cy, cz = o.arr("cx"), o.arr("cy"), o.arr("cz")
    q = o.arr("qp")
    q[...] = 1.0
    q[1] = 2.0 + 0.3 * cx - 0.2 * cy + 0.1 * cz
    q[4] = 1.0e5
    o.arr("temp")[...] = 300.0
    o.stage(1, 8)
    g = o.arr("gx")[0, 1:-1, 1:-1, 1:-1], o.arr("gy")[0, 1:-1, 1:-1, 1:-1], o.arr("gz")[0, 1:-1, 1:-1, 1:-1]
    np.testing.assert_allclose(g[0], 0.3, rtol=1e-12)
    np.testing.assert_allclose(g[1], -0.2, rtol=1e-12)
    np.testing.assert_allclose(g[2], 0.1, rtol=1e-12)


# ------------------------------------------------------------- time step
def test_time_kat():
    case = tiny_case(2, 2, 2, time_local=1, CFL=1.0)  # tests/test_time.f90:32-57
    o = Oracle(case)
    for n in ("xl", "xr", "yl", "yr", "zl", "zr"):
        o.arr(n)[...] = 1.0
    for d in ("i", "j", "k"):
        o.arr(d + "A")[...] = 1.0
    o.arr("inx")[...] = 1.0
    o.arr("jny")[...] = 1.0
    o.arr("knz")[...] = 1.0
    o.arr("vol")[...] = 1.0
    o.arr("qp")[...] = 0.0
    o.time_step()
    dt = o.arr("dt")[0, 0, 0]
    assert 0.135 < dt < 0.145
    assert dt == pytest.approx(1.0 / (6 * math.sqrt(1.4)), rel=1e-15)


@pytest.mark.parametrize("viscous", [False, True])
def test_time_step_uniform_closed_form(viscous):
    """Uniform state on a uniform Cartesian grid: dt = CFL h / (sum |u_d| + 6 c)
    inviscid; with viscosity dt = 2 CFL V / (6 h gm mu / (rho Pr) + 2 CFL V / dt_inv)
    (src/time.f90:230-234, 434-444)."""
    case = tiny_case(6, 5, 5, bc_id=(-9,) * 6, time_local=1, CFL=0.7, flux="ausmP",
                     mu_ref=(2e-3 if viscous else 0.0), Pr=0.72)
    h = 0.4
    o = Oracle(case)
    o.set_nodes(C.cartesian_nodes(case, spacing=(h, h, h)))
    rho, u, v, w, p = 1.2, 30.0, -20.0, 10.0, 1.0e5
    q = o.arr("qp")
    for l, val in enumerate((rho, u, v, w, p)):
        q[l] = val
    o.temp_refresh()
    assert o.residual(1) == 0
    o.time_step()
    c = math.sqrt(1.4 * p / rho)
    dt_inv = 0.7 * h / (2 * (abs(u) + abs(v) + abs(w)) + 6 * c)
    if not viscous:
        np.testing.assert_allclose(o.arr("dt"), dt_inv, rtol=1e-13)
    else:
        V = h ** 3
        ls = 1.4 * (6 * h * h * 2e-3 / (rho * h)) / 0.72
        np.testing.assert_allclose(o.arr("dt"), 0.7 * (2 / (ls + 2 * 0.7 * V / dt_inv)) * V, rtol=1e-13)


def test_time_step_uniform_closed_form_sst():
    """add_turbulent_time (src/time.f90:450-531) on a uniform SST state of a
    uniform periodic grid: every gradient is zero, so the vorticity is zero
    and the eddy viscosity is rho a1 k / (a1 omega) = rho k / omega exactly
    (viscosity.f90:343-388); the turbulent pass repeats the viscous one
    with mu_t and Pr_t on the viscous time step.  Pr != Pr_t and
    mu != mu_t, so a swap of either fails."""
    mu, Pr, tPr, h = 2e-3, 0.72, 0.9, 0.01
    case = tiny_case(6, 5, 5, bc_id=(-9,) * 6, time_local=1, CFL=0.7, flux="ausmP", turb="sst",
                     mu_ref=mu, Pr=Pr, tPr=tPr, tk_inf=500.0, tw_inf=600.0, lim=(1, 1, 1), tlim=(1, 1, 1))
    o = Oracle(case)
    o.set_nodes(C.cartesian_nodes(case, spacing=(h, h, h)))
    o.arr("dist")[...] = 1.0
    rho, u, v, w, p, k, om = 1.2, 30.0, -20.0, 10.0, 1.0e5, 500.0, 600.0
    q = o.arr("qp")
    for l, val in enumerate((rho, u, v, w, p, k, om)):
        q[l] = val
    o.temp_refresh()
    assert o.residual(1) == 0
    inner = (slice(3, -3),) * 3
    mut = rho * k / om
    np.testing.assert_allclose(o.arr("mut")[inner], mut, rtol=1e-14)
    o.time_step()
    c = math.sqrt(1.4 * p / rho)
    dt_inv = 0.7 * h / (2 * (abs(u) + abs(v) + abs(w)) + 6 * c)
    V = h ** 3
    ls = 1.4 * (6 * h * h * mu / (rho * h)) / Pr
    dt_v = 0.7 * (2 / (ls + 2 * 0.7 * V / dt_inv)) * V
    lt = 1.4 * (6 * h * h * mut / (rho * h)) / tPr
    dt_t = 0.7 * (2 / (lt + 2 * 0.7 * V / dt_v)) * V
    np.testing.assert_allclose(o.arr("dt"), dt_t, rtol=1e-13)
    # the pass changes the step: mu_t is not negligible here
    assert abs(dt_t - dt_v) / dt_v > 1e-2


# ----------------------------------------------------------- invariants
@pytest.mark.parametrize("scheme", list(FLUX_KAT))
@pytest.mark.parametrize("recon", ["none", "muscl", "ppm", "weno"])
def test_freestream_preservation(scheme, recon):
    """A uniform state on a warped periodic grid has zero residue: the face
    area vectors of every cell close (src/geometry.f90:236-314)."""
    case = tiny_case(9, 8, 7, flux=scheme, recon=recon, bc_id=(-9,) * 6, MInf=0.3)
    o = Oracle(case)
    o.set_nodes(C.cartesian_nodes(case, warp=0.04))
    q = o.arr("qp")
    for l, val in enumerate((1.1, 80.0, 20.0, -10.0, 9.0e4)):
        q[l] = val
    o.temp_refresh()
    assert o.residual(1) == 0
    fscale = np.abs(o.arr("F")).max()
    assert np.abs(o.arr("res")).max() < 1e-10 * fscale


def test_sutherland_kat():
    """mu(T_ref) = mu_ref (tests/test_viscosity.f90)."""
    case = tiny_case(5, 5, 5, bc_id=(-9,) * 6, mu_ref=1.8e-5, mu_sutherland=1, T_ref=300.0, S_ref=110.4,
                     recon="muscl")
    o = Oracle(case)
    o.set_nodes(C.cartesian_nodes(case))
    q = o.arr("qp")
    for l, val in enumerate((1.0, 1.0, 0.0, 0.0, 287.0 * 300.0)):
        q[l] = val
    o.temp_refresh()
    o.stage(1, 8)
    np.testing.assert_allclose(o.arr("mu")[1:-1, 1:-1, 1:-1], 1.8e-5, rtol=1e-15)


def test_freestream_values():
    """k_inf, omega_inf and the Reynolds convention of src/state.f90:86-99
    (tests/test_state.f90 checks k_inf = 1.5 for |V| tu = 1)."""
    c = C.Case(5, 5, 5, turb="sst")
    C.freestream(c, mach=0.2, reynolds=1e6, tu_percent=0.1, mu_ratio=0.01)
    assert c.tk_inf == pytest.approx(1.5 * (c.vel_mag * 1e-3) ** 2, rel=1e-15)
    assert c.tw_inf == pytest.approx(c.rho_inf * c.tk_inf / (c.mu_ref * 0.01), rel=1e-15)
    assert c.rho_inf * c.vel_mag / c.mu_ref == pytest.approx(1e6, rel=1e-12)


def test_ghost_nodes_affine_exact():
    """Ghost nodes are linear extrapolations (src/grid.f90:137-236): on a grid
    whose coordinates are an integer-valued affine map of the node indices
    every operation is exact, so all three ghost layers of every face, edge
    and corner must equal the map itself bit for bit."""
    case = tiny_case(7, 6, 5)
    o = Oracle(case)
    n = np.zeros(case.node_shape())
    k, j, i = np.meshgrid(np.arange(-2, case.kmx + 4), np.arange(-2, case.jmx + 4), np.arange(-2, case.imx + 4),
                          indexing="ij")
    exact = np.stack([2.0 * i + j - k + 5.0, 1.0 * j + 3.0 * k, 1.0 * i + 2.0 * k + 0.5 * j])
    n[:, 3:-3, 3:-3, 3:-3] = exact[:, 3:-3, 3:-3, 3:-3]
    o.set_nodes(n)
    assert np.array_equal(o.arr("nodes"), exact)
    # unit-cube cells of the map have the map's determinant as volume
    det = np.linalg.det(np.array([[2.0, 1.0, -1.0], [0.0, 1.0, 3.0], [1.0, 0.5, 2.0]]))
    np.testing.assert_allclose(o.arr("vol")[3:-3, 3:-3, 3:-3], det, rtol=1e-14)


def test_rk4_composition():
    """The step equals the RK4 combination of four residual evaluations taken
    with the same stale Temp and the stage-1 time step (src/update.f90:176-187)."""
    case, nodes, q0 = C.laminar_box(10, 8, 6)
    case = dataclasses.replace(case, time_local=1)
    o = Oracle(case)
    o.set_nodes(C.cartesian_nodes(case, warp=0.03))
    gm = case.gm

    def to_cons(q):
        u = q[:5].copy()
        u[1:5] = q[1:5] * q[0]
        u[4] = (u[4] / (gm - 1) + 0.5 * (u[1] ** 2 + u[2] ** 2 + u[3] ** 2)) / u[0]
        return u

    def to_prim(u):
        q = u.copy()
        q[1:] = u[1:] / u[0]
        q[4] = (gm - 1) * q[0] * (q[4] - 0.5 * (q[1] ** 2 + q[2] ** 2 + q[3] ** 2))
        return q

    inner = (slice(None), slice(3, -3), slice(3, -3), slice(3, -3))
    o.set_state(q0)
    o.temp_refresh()
    temp = o.arr("temp").copy()
    U0 = to_cons(o.arr("qp").copy()[inner])
    vol = o.arr("vol")[3:-3, 3:-3, 3:-3]
    Rs = []
    coef = [0.5, 0.5, 1.0]
    for s in range(4):
        assert o.residual(1) == 0
        if s == 0:
            o.time_step()
            dt = o.arr("dt").copy()
        R = o.arr("res").copy()
        Rs.append(R)
        if s < 3:
            q = o.arr("qp").copy()
            q[inner] = to_prim(U0 - coef[s] * R * dt / vol)
            o.set_state(q)
            o.arr("temp")[...] = temp
    U4 = U0 - (Rs[0] + 2 * Rs[1] + 2 * Rs[2] + Rs[3]) * dt / vol / 6
    o2 = Oracle(case)
    o2.set_nodes(C.cartesian_nodes(case, warp=0.03))
    o2.set_state(q0)
    assert o2.step(1) == 0
    np.testing.assert_allclose(o2.arr("qp")[inner], to_prim(U4), rtol=1e-12)


def _to_cons(q, gm):
    u = q.copy()
    u[1:] = q[1:] * q[0]
    u[4] = (u[4] / (gm - 1) + 0.5 * (u[1] ** 2 + u[2] ** 2 + u[3] ** 2)) / u[0]
    return u


def _to_prim(u, gm):
    q = u.copy()
    q[1:] = u[1:] / u[0]
    q[4] = (gm - 1) * q[0] * (q[4] - 0.5 * (q[1] ** 2 + q[2] ** 2 + q[3] ** 2))
    return q


STEP_RULES = {  # src/update.f90:162-187: per stage (TF, SF); None = no R accumulator
    "RK4": [(0.5, 1.0), (0.5, 2.0), (1.0, 2.0), (1.0 / 6.0, 1.0)],
    "RK2": [(0.5, 1.0), (0.5, 1.0)],
    "none": [(1.0, None)],
}


@pytest.mark.parametrize("rule,turb,time_local", [
    ("RK4", "sst", 1), ("RK4", "sst", 0), ("RK2", "none", 1), ("RK2", "sst", 1), ("none", "none", 0),
    ("none", "sst", 1)])
def test_step_composition(rule, turb, time_local):
    """A step equals its stages composed from separately evaluated residuals
    (src/update.f90:162-195, 367-485): stale Temp, the stage-1 time step,
    R_store = sum SF R with R replaced by the sum at the last stage, the
    stage state U0 - TF R dt/V.  With SST the k and omega residues are first
    divided by 1 + beta omega dt and 1 + 2 beta omega dt, beta blended by the
    stage's F1 and omega the stage's current value."""
    if turb == "sst":
        case, nodes, q0 = C.sst_bump(14, 10, 6)
        wall = True
    else:
        case, _, q0 = C.laminar_box(10, 8, 6)
        nodes = C.cartesian_nodes(case, warp=0.03)
        wall = False
    case = dataclasses.replace(case, time_accuracy=rule, time_local=time_local)
    gm = case.gm
    inner = (slice(None), slice(3, -3), slice(3, -3), slice(3, -3))
    o = Oracle(case)
    o.set_nodes(nodes)
    if wall:
        o.wall_dist()
    o.set_state(q0)
    o.temp_refresh()
    temp = o.arr("temp").copy()
    vol = o.arr("vol")[3:-3, 3:-3, 3:-3].copy()
    U0 = _to_cons(o.arr("qp").copy()[inner], gm)
    acc = None
    stages = STEP_RULES[rule]
    for s, (TF, SF) in enumerate(stages):
        assert o.residual(1) == 0
        if s == 0:
            o.time_step()
            dt = o.arr("dt").copy()
        q = o.arr("qp").copy()
        R = o.arr("res").copy()
        if turb == "sst":
            F1 = o.arr("f1")[3:-3, 3:-3, 3:-3]
            w = q[6][3:-3, 3:-3, 3:-3]
            beta = 0.075 * F1 + (1.0 - F1) * 0.0828
            R[5] = R[5] / (1 + beta * w * dt)
            R[6] = R[6] / (1 + 2 * beta * w * dt)
        base = U0 if SF is not None else _to_cons(q[inner], gm)
        if SF is not None:
            acc = SF * R if acc is None else acc + SF * R
            if s == len(stages) - 1:
                R = acc
        new = _to_prim(base - R * (TF * dt / vol), gm)
        if turb == "sst":
            assert (new[5] >= 0).all() and (new[6] >= 0).all()
        q[inner] = new
        o.set_state(q)
        o.arr("temp")[...] = temp
    expect = o.arr("qp").copy()[inner]
    o2 = Oracle(case)
    o2.set_nodes(nodes)
    if wall:
        o2.wall_dist()
    o2.set_state(q0)
    assert o2.step(1) == 0
    np.testing.assert_allclose(o2.arr("qp")[inner], expect, rtol=1e-12, atol=0)


@pytest.mark.parametrize("rule", ["TVDRK2", "TVDRK3"])
@pytest.mark.parametrize("turb", ["none", "sst"])
def test_tvd_step_composition(rule, turb):
    """TVD-RK steps (src/update.f90:196-214): plain stage updates
    U = U(q) - R dt/V from the current state (the stage-1 time step, the SST
    point-implicit scaling), and the blends q = 1/2 U0 + 1/2 q (TVDRK2) or
    q = 3/4 U0 + 1/4 q and q = 1/3 U0 + 2/3 q (TVDRK3) of the primitive state
    with the stored base."""
    if turb == "sst":
        case, nodes, q0 = C.sst_bump(14, 10, 6)
        wall = True
    else:
        case, _, q0 = C.laminar_box(10, 8, 6)
        nodes = C.cartesian_nodes(case, warp=0.03)
        wall = False
    case = dataclasses.replace(case, time_accuracy=rule, time_local=1)
    gm = case.gm
    inner = (slice(None), slice(3, -3), slice(3, -3), slice(3, -3))
    o = Oracle(case)
    o.set_nodes(nodes)
    if wall:
        o.wall_dist()
    o.set_state(q0)
    o.temp_refresh()
    temp = o.arr("temp").copy()
    vol = o.arr("vol")[3:-3, 3:-3, 3:-3].copy()
    U0 = o.arr("qp").copy()
    dt = None

    def stage():
        nonlocal dt
        assert o.residual(1) == 0
        if dt is None:
            o.time_step()
            dt = o.arr("dt").copy()
        q = o.arr("qp").copy()
        R = o.arr("res").copy()
        if turb == "sst":
            F1 = o.arr("f1")[3:-3, 3:-3, 3:-3]
            w = q[6][3:-3, 3:-3, 3:-3]
            beta = 0.075 * F1 + (1.0 - F1) * 0.0828
            R[5] = R[5] / (1 + beta * w * dt)
            R[6] = R[6] / (1 + 2 * beta * w * dt)
        new = _to_prim(_to_cons(q[inner], gm) - R * (1.0 * dt / vol), gm)
        if turb == "sst":
            assert (new[5] >= 0).all() and (new[6] >= 0).all()
        q[inner] = new
        o.set_state(q)
        o.arr("temp")[...] = temp

    def blend(a, b):
        o.set_state(a * U0 + b * o.arr("qp").copy())
        o.arr("temp")[...] = temp

    stage()
    stage()
    if rule == "TVDRK2":
        blend(0.5, 0.5)
    else:
        blend(0.75, 0.25)
        stage()
        blend(1.0 / 3.0, 2.0 / 3.0)
    expect = o.arr("qp").copy()[inner]
    o2 = Oracle(case)
    o2.set_nodes(nodes)
    if wall:
        o2.wall_dist()
    o2.set_state(q0)
    assert o2.step(1) == 0
    np.testing.assert_allclose(o2.arr("qp")[inner], expect, rtol=1e-12, atol=0)


def test_resnorm_closed_form():
    case, nodes, q = C.laminar_box(6, 5, 4)
    o = Oracle(case)
    rng = np.random.default_rng(0)
    R = rng.standard_normal(o.arr("res").shape)
    o.arr("res")[...] = R
    for n in ("F", "G", "H"):
        o.arr(n)[...] = rng.standard_normal(o.arr(n).shape)
    out = o.resnorm()
    v, rho = case.vel_mag, case.rho_inf
    scale = [rho * v, rho * v * v, rho * v * v, rho * v * v, 0.5 * rho * v ** 3 + 3.5 * case.p_inf]
    for l in range(5):
        assert out[l + 1] == pytest.approx(math.sqrt((R[l] ** 2).sum()) / scale[l], rel=1e-12)
    F, G, H = o.arr("F")[0], o.arr("G")[0], o.arr("H")[0]
    m = F[:, :, 0].sum() - F[:, :, -1].sum() + G[:, 0, :].sum() - G[:, -1, :].sum() + H[0].sum() - H[-1].sum()
    assert out[0] == pytest.approx(abs(m), rel=1e-10)


# ------------------------------------------------ MUSCL limiters, switches
def _muscl_row(r, limiter, lim=1):
    """Upper- and lower-face states of cell 2 of a row with q(1) = 0,
    q(2) = 1, q(3) = 1 + r: backward difference 1, forward difference r."""
    case = tiny_case(6, 2, 2, recon="muscl", lim=(lim,) * 3, tlim=(lim,) * 3, limiter=limiter)
    o = Oracle(case)
    q = o.arr("qp")
    q[...] = 1.0
    row = np.full(11, 7.0)
    row[:3] = (-2.0, -1.0, 0.0)   # cells -2 .. 0
    row[3:6] = (0.0, 1.0, 1.0 + r)  # cells 1, 2, 3
    for l in range(5):
        q[l, 3, 3, :] = row
    o.stage(1, 2)
    return o.arr("xl")[0, 0, 0, 3], o.arr("xr")[0, 0, 0, 2]


@pytest.mark.parametrize("r", [-2.0, -0.5, 0.0, 0.5, 1.0, 3.0, 10.0])
def test_van_albada_closed_form(r):
    """psi(r) = (r^2 + r)/(r^2 + 1) has psi(1/r) = psi(r)/r, so both
    kappa = 1/3 face states reduce to q -+ psi(r) bd / 2."""
    va = (r * r + r) / (r * r + 1.0)
    L, R = _muscl_row(r, "van_albada")
    assert L == pytest.approx(1.0 + 0.5 * va, rel=1e-12, abs=1e-13)
    assert R == pytest.approx(1.0 - 0.5 * va, rel=1e-12, abs=1e-13)


@pytest.mark.parametrize("r,psi,psi_inv", [(-1.0, 0.0, 0.0), (0.25, 0.5, 2.0), (1.0, 1.0, 1.0), (3.0, 2.0, 5.0 / 9.0),
                                           (10.0, 2.0, 0.2)])
def test_koren_closed_form(r, psi, psi_inv):
    """Koren-alpha limiter max(0, min(2r, (2r+1)/3, 2)): 0 for r <= 0, 1 at
    r = 1, 2 from r = 2.5; face states q + (2/3 psi(r) + 4/3 r psi(1/r)) bd/4
    and q - (4/3 psi(r) + 2/3 r psi(1/r)) bd/4."""
    L, R = _muscl_row(r, "koren")
    assert L == pytest.approx(1.0 + 0.25 * (2 / 3 * psi + 4 / 3 * r * psi_inv), rel=1e-12, abs=1e-13)
    assert R == pytest.approx(1.0 - 0.25 * (4 / 3 * psi + 2 / 3 * r * psi_inv), rel=1e-12, abs=1e-13)
    # limiter switch off: psi = 1, the unlimited kappa = 1/3 states
    L0, R0 = _muscl_row(r, "koren", lim=0)
    assert L0 == pytest.approx(1.0 + 0.25 * (2 / 3 + 4 / 3 * r), rel=1e-13)
    assert R0 == pytest.approx(1.0 - 0.25 * (4 / 3 + 2 / 3 * r), rel=1e-13)


def test_weno_nm_kat():
    L, R = states_1d("weno_NM", 1)  # tests/test_weno_NM.f90:19-46, unit volumes
    assert 0.48 <= L[1] < 0.52 and 0.48 < R[1] < 0.52
    assert 1.48 <= L[2] < 1.52 and 1.48 < R[2] < 1.52
    assert 2.4 <= L[3] < 2.5 and 2.4 < R[3] < 2.5
    assert 3.6 <= L[4] < 3.7 and 3.8 < R[4] < 4.1
    assert 5.9 <= L[5] < 6.1 and 6.9 < R[5] < 7.1


def test_weno_nm_is_weno5_on_equal_volumes():
    """With equal volumes every blend weight is 1/2 and WENO-NM's
    polynomials and smoothness indicators are WENO5's (weno.f90:59-87)
    rewritten: the two agree to rounding, on a wiggly profile too."""
    case, nodes, q = C.laminar_box(10, 8, 6)
    out = {}
    for recon in ("weno", "weno_NM"):
        o = Oracle(dataclasses.replace(case, recon=recon))
        o.set_nodes(nodes)
        o.set_state(q)
        o.stage(1, 2)
        out[recon] = [o.arr(n).copy() for n in ("xl", "xr", "yl", "yr", "zl", "zr")]
    for a, b in zip(out["weno"], out["weno_NM"]):
        np.testing.assert_allclose(b, a, rtol=1e-13, atol=0)
    # on the stretched bump grid the volumes differ and so do the states
    c2, n2, q2 = C.sst_bump(10, 8, 4)
    yl = {}
    for recon in ("weno", "weno_NM"):
        o = Oracle(dataclasses.replace(c2, recon=recon))
        o.set_nodes(n2)
        o.set_state(q2)
        o.stage(1, 2)
        yl[recon] = o.arr("yl").copy()
    assert np.abs(yl["weno"] - yl["weno_NM"]).max() > 1e-6 * np.abs(yl["weno"]).max()


def _pb_states(case, nodes, q, pbs):
    o = Oracle(dataclasses.replace(case, pbs=pbs))
    o.set_nodes(nodes)
    o.set_state(q)
    o.stage(1, 2)
    return [o.arr(n).copy() for n in ("xl", "xr", "yl", "yr", "zl", "zr")]


@pytest.mark.parametrize("recon", ["muscl", "ppm"])
def test_pressure_switch_limits(recon):
    """pdif = 1 - |dp|/(|dp| + p_inf): with uniform pressure it is 1 and the
    switch changes nothing; with pressure jumps far above p_inf it tends to
    0 and the face states become the first-order cell values."""
    case, _, q = C.laminar_box(8, 6, 4)
    case = dataclasses.replace(case, recon=recon, bc_id=(0,) * 6)
    nodes = C.cartesian_nodes(case)
    q = q.copy()
    q[4] = case.p_inf
    a, b = _pb_states(case, nodes, q, (0, 0, 0)), _pb_states(case, nodes, q, (1, 1, 1))
    for x, y in zip(a, b):
        assert np.array_equal(x, y)
    # pressure jumps of 1e12 p_inf between every i-neighbour pair two
    # apart (period 4 in i): pdif ~ 1e-12
    kk, jj, ii = np.indices(q[4].shape)
    q[4] = case.p_inf * (1.0 + 1e12 * ((ii // 2) % 2))
    L, R = _pb_states(case, nodes, q, (1, 0, 0))[:2]
    u = q[1]
    # i faces 1..imx of interior rows: left = cell f-1, right = cell f
    for f in range(1, case.imx + 1):
        np.testing.assert_allclose(L[1, :, :, f], u[3:-3, 3:-3, f + 1], rtol=1e-10)
        np.testing.assert_allclose(R[1, :, :, f], u[3:-3, 3:-3, f + 2], rtol=1e-10)
    # the j states are untouched when only the i switch is on
    a0 = _pb_states(case, nodes, q, (0, 0, 0))
    assert np.array_equal(a0[2], _pb_states(case, nodes, q, (1, 0, 0))[2])
    assert not np.array_equal(a0[0], L)


def test_lctm2015_rk_update_leaves_the_intermittency():
    """The conservative RK update writes variables 1-5 and the two
    turbulence variables only (src/update.f90:456-484): under lctm2015 the
    intermittency keeps its value through RK steps, while its residue is
    computed (and enters the non-physical-state check)."""
    case, nodes, q = C.sst_bump(12, 8, 6, flux="ausm", recon="muscl", transition="lctm2015")
    o = Oracle(case)
    o.set_nodes(nodes)
    o.wall_dist()
    o.set_state(q)
    g0 = o.arr("qp")[7, 3:-3, 3:-3, 3:-3].copy()
    k0 = o.arr("qp")[5, 3:-3, 3:-3, 3:-3].copy()
    for it in (1, 2):
        assert o.step(it) == 0
    assert np.array_equal(o.arr("qp")[7, 3:-3, 3:-3, 3:-3], g0)
    assert not np.array_equal(o.arr("qp")[5, 3:-3, 3:-3, 3:-3], k0)
    assert np.abs(o.arr("res")[7]).max() > 0


def test_lctm2015_intermittency_source_closed_form():
    """With the intermittency at 1 everywhere its production vanishes
    (P_gm carries gamma (1 - gamma), src/source.f90:430) and so does the
    production limiter term of k; with k tiny and omega large the turbulent
    Reynolds number RT = rho k/(mu omega) is ~1e-20, so Fturb =
    exp(-(RT/2)**4) = 1.  The intermittency residue is then the face-flux
    divergence plus D_gm V = 0.06 rho Omega (50 - 1) V, Omega the vorticity
    magnitude of the cell's gradients."""
    case, nodes, q = C.sst_bump(10, 8, 5, flux="ausm", recon="none", transition="lctm2015")
    o = Oracle(case)
    o.set_nodes(nodes)
    o.wall_dist()
    qq = q.copy()
    qq[5] = 1e-12
    qq[6] = 1e8
    qq[7] = 1.0
    o.set_state(qq)
    o.temp_refresh()
    assert o.residual(3) == 0
    gx, gy, gz = o.arr("gx"), o.arr("gy"), o.arr("gz")
    sl = (slice(1, -1),) * 3
    vort = np.sqrt((gy[2][sl] - gz[1][sl]) ** 2 + (gz[0][sl] - gx[2][sl]) ** 2 + (gx[1][sl] - gy[0][sl]) ** 2)
    rho = qq[0, 3:-3, 3:-3, 3:-3]
    vol = o.arr("vol")[3:-3, 3:-3, 3:-3]
    F, G, H = o.arr("F")[7], o.arr("G")[7], o.arr("H")[7]
    div = (F[:, :, 1:] - F[:, :, :-1]) + (G[:, 1:, :] - G[:, :-1, :]) + (H[1:] - H[:-1])
    want = div + 0.06 * rho * vort * 49.0 * vol
    np.testing.assert_allclose(o.arr("res")[7], want, rtol=1e-10, atol=1e-12 * np.abs(want).max())
