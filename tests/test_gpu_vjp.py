"""The discrete adjoint of the inviscid residual (fest_vjp), through the C ABI.

fest_vjp returns bar = (dR/dq)^T w for the residual R that fest_residual
computes, the ghost cells following from the boundary rules.  The checks:

- Jacobian columns: bar[v, c] = d<w, R>/dq[v, c] by central differences
  (h = 1e-7 of the variable's free-stream scale) of the GPU's own residual,
  for every interior input of grids with each supported boundary rule, and
  for seeded samples of inputs under every flux scheme and reconstruction
  (2-D grids included);
- the same columns from the oracle's residual: the adjoint of the residual
  the oracle restates (and which is pinned to the reference);
- the transposition identity <bar, dq> = d/dh <w, R(q + h dq)> over 32
  seeded directions dq of all interior cells at once, for a scheme whose
  residual is smooth at this state (WENO5 + van Leer, subsonic: no limiter
  or switch changes branch).  With MUSCL, PPM or the AUSM family a
  direction that moves every cell at once crosses some limiter kink or
  switch inside [-h, h] (there are ~10^4 of them), and the difference
  quotient then carries that kink's jump (1e-4..1e-2 relative), while
  single-input columns stay at ~1e-8: that is why those schemes are
  checked column by column;
- torch.autograd.gradcheck of the residual as a differentiable function
  (festgpu.autograd) on a 2-D 6 x 5 grid: the full Jacobian, column by
  column from finite differences, row by row from fest_vjp;
- the unsteady adjoint of 12 RK4 steps with checkpoints (festgpu.adjoint)
  against central differences of the whole trajectory;
- ghost cells hold no adjoint without interface faces, the adjoint is
  linear in w and deterministic;
- the laminar viscous terms (viscous fluxes, Green-Gauss gradients and
  their ghost-cell rule, Sutherland's law) by the same columns;
- configurations it does not differentiate fail with FEST_E_CONFIG.

Tolerance: 1e-6 relative (per column: of the largest |bar| of that
variable; per direction: of the larger side).  The differences leave ~1e-8
of rounding at h = 1e-7 (measured), far below the 1e-3..1 that a wrong
sign, a dropped term or a missed ghost rule gives.
"""
import ctypes
import dataclasses
import sys

import numpy as np
import pytest

from binding import Oracle
from festgpu import cases as C
from festgpu import native
from festgpu.native import FestError, GpuBlock

pytestmark = pytest.mark.gpu
H = 1e-7
TOL = 1e-6
INNER = (slice(None),) + (slice(3, -3),) * 3


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def scales(case):
    return np.array([case.rho_inf, case.vel_mag, case.vel_mag, case.vel_mag, case.p_inf])


def directions(case, seed, n):
    """n seeded directions on the interior cells, scaled per variable."""
    rng = np.random.default_rng(seed)
    shape = (5, case.kmx - 1, case.jmx - 1, case.imx - 1)
    s = scales(case)[:, None, None, None]
    return [s * rng.uniform(-1.0, 1.0, shape) for _ in range(n)]


def weights(case, seed):
    rng = np.random.default_rng(seed + 99)
    return rng.uniform(-1.0, 1.0, (5, case.kmx - 1, case.jmx - 1, case.imx - 1))


def gpu_residual(g, q):
    g.upload_state(q)
    g.residual(1)
    return g.download("residue")


def fd_dot(resid, q, dq, w):
    qp, qm = q.copy(), q.copy()
    qp[INNER] += H * dq
    qm[INNER] -= H * dq
    return float(np.sum(w * (resid(qp) - resid(qm)))) / (2 * H)


def check_identity(g, case, q, resid, seed, ndir):
    w = weights(case, seed)
    g.upload_state(q)
    bar = g.vjp(w)
    worst = 0.0
    for dq in directions(case, seed, ndir):
        a = float(np.sum(bar[INNER] * dq))
        b = fd_dot(resid, q, dq, w)
        err = abs(a - b) / max(abs(a), abs(b), 1e-300)
        worst = max(worst, err)
    return bar, worst


def check_columns(g, case, q, resid, seed, sample=None):
    """max over the checked inputs (v, k, j, i) of |bar - FD| / max|bar_v|;
    sample: how many seeded inputs, None for every interior input."""
    w = weights(case, seed)
    g.upload_state(q)
    bar = g.vjp(w)
    inner = bar[INNER]
    shape = inner.shape
    n = int(np.prod(shape))
    idx = range(n) if sample is None else np.random.default_rng(seed + 7).choice(n, sample, replace=False)
    sc = scales(case)
    top = np.abs(inner).reshape(5, -1).max(axis=1)
    worst, where = 0.0, None
    for f in idx:
        v, k, j, i = np.unravel_index(int(f), shape)
        d = H * sc[v]
        qp, qm = q.copy(), q.copy()
        qp[v, 3 + k, 3 + j, 3 + i] += d
        qm[v, 3 + k, 3 + j, 3 + i] -= d
        fd = float(np.sum(w * (resid(qp) - resid(qm)))) / (2 * d)
        err = abs(inner[v, k, j, i] - fd) / top[v]
        if err > worst:
            worst, where = err, (v, k, j, i, inner[v, k, j, i], fd)
    if worst > TOL:
        say("  worst column (v, k, j, i, adjoint, difference):", where)
    return bar, worst


BC_MIXES = [
    (-3, -4, -6, -6, -9, -9),   # subsonic inlet/outlet, slip walls, periodic k
    (-8, -8, -6, -8, -9, -9),   # far field around a slip wall
    (-1, -2, -5, -7, -9, -9),   # supersonic inlet/outlet, no-slip wall, pole
    (-11, -4, -6, -8, -6, -6),  # total-pressure inlet, slip walls in k
    (-9, -9, -5, -6, -9, -9),   # periodic i, no-slip and slip walls
]


@pytest.mark.parametrize("bcs", BC_MIXES)
def test_jacobian_columns_boundary_rules(bcs):
    """Every interior input of a 10 x 7 x 5 grid, MUSCL + AUSM+."""
    say("vjp columns bcs", bcs)
    case, nodes, q = C.euler_bump(10, 7, 5, bc_id=bcs)
    g = GpuBlock(case)
    g.set_grid(nodes)
    bar, worst = check_columns(g, case, q, lambda x: gpu_residual(g, x), 11)
    say("  worst", worst)
    assert worst <= TOL
    # no interface faces: every ghost cell ends without adjoint
    ghost = np.ones(case.cell_shape(), bool)
    ghost[3:-3, 3:-3, 3:-3] = False
    assert np.all(bar[:, ghost] == 0.0)


@pytest.mark.parametrize("flux", ["ausm", "ausmP", "ausmUP", "slau", "ldfss0", "van_leer"])
@pytest.mark.parametrize("recon", ["none", "muscl", "ppm", "weno", "weno_NM"])
def test_jacobian_columns_schemes(flux, recon):
    """120 seeded inputs per scheme pair on a 12 x 9 x 5 grid."""
    say("vjp columns", flux, recon)
    case, nodes, q = C.euler_bump(12, 9, 5, flux=flux, recon=recon, seed=23)
    g = GpuBlock(case)
    g.set_grid(nodes)
    _, worst = check_columns(g, case, q, lambda x: gpu_residual(g, x), 5, sample=120)
    say("  worst", worst)
    assert worst <= TOL


@pytest.mark.parametrize("recon,pbs", [("muscl", (1, 1, 1)), ("ppm", (1, 1, 1)), ("muscl", (0, 1, 0))])
def test_jacobian_columns_pressure_based_switch(recon, pbs):
    """Every interior input of a 10 x 7 x 5 grid with the pressure-based
    switch: each face state then also reads the stencil's pressures."""
    say("vjp columns pbs", recon, pbs)
    case, nodes, q = C.euler_bump(10, 7, 5, recon=recon, bc_id=(-3, -4, -6, -8, -9, -9), seed=31)
    case = dataclasses.replace(case, pbs=pbs)
    g = GpuBlock(case)
    g.set_grid(nodes)
    _, worst = check_columns(g, case, q, lambda x: gpu_residual(g, x), 17)
    say("  worst", worst)
    assert worst <= TOL


VISC_MIXES = [
    ((-3, -4, -5, -6, -9, -9), 1, "muscl"),  # adiabatic no-slip wall, slip wall, periodic k; Sutherland
    ((-8, -8, -5, -8, -6, -6), 1, "weno"),   # far fields, slip walls in k
    ((-11, -4, -5, -5, -9, -9), 0, "muscl"),  # total-pressure inlet, two walls; constant viscosity
    ((-9, -9, -5, -6, -9, -9), 1, "none"),   # periodic i
]


@pytest.mark.parametrize("bcs,suth,recon", VISC_MIXES)
def test_jacobian_columns_viscous(bcs, suth, recon):
    """Every interior input of a 10 x 7 x 5 laminar grid: the viscous
    fluxes, the Green-Gauss gradients and their ghost-cell rule, and the
    viscosity's temperature dependence on top of the inviscid residual.
    (The state of euler_bump's seed 41 puts a MUSCL limiter kink on one
    pressure column: the inviscid adjoint shows the same 2e-3 there.)"""
    say("vjp columns viscous", bcs, suth, recon)
    case, nodes, q = C.euler_bump(10, 7, 5, bc_id=bcs, recon=recon)
    case = dataclasses.replace(case, mu_ref=0.02, mu_sutherland=suth, S_ref=0.4)
    g = GpuBlock(case)
    g.set_grid(nodes)
    bar, worst = check_columns(g, case, q, lambda x: gpu_residual(g, x), 29)
    say("  worst", worst)
    assert worst <= TOL
    # the viscous terms are there: the inviscid adjoint differs
    inv = GpuBlock(dataclasses.replace(case, mu_ref=0.0, mu_sutherland=0))
    inv.set_grid(nodes)
    inv.upload_state(q)
    assert np.max(np.abs(inv.vjp(weights(case, 29)) - bar)) > 1e-3 * np.max(np.abs(bar))


def test_viscous_adjoint_refusals():
    """2-D viscous grids and the node adjoint of a viscous block are not
    differentiated."""
    case, nodes, q = C.euler_bump(12, 8, 1, bc_id=(-3, -4, -5, -8, -9, -9))
    case = dataclasses.replace(case, mu_ref=0.02, mu_sutherland=1)
    g = GpuBlock(case)
    g.set_grid(nodes)
    g.upload_state(q)
    with pytest.raises(FestError) as e:
        g.vjp(weights(case, 0))
    assert e.value.code == -3
    case, nodes, q = C.euler_bump(10, 7, 5, bc_id=(-3, -4, -5, -6, -9, -9))
    case = dataclasses.replace(case, mu_ref=0.02, mu_sutherland=1)
    g = GpuBlock(case)
    g.set_grid(nodes)
    g.upload_state(q)
    with pytest.raises(FestError) as e:
        g.vjp(weights(case, 0), nodes=True)
    assert e.value.code == -3


@pytest.mark.parametrize("limiter", ["koren", "van_albada"])
def test_jacobian_columns_two_dimensional(limiter):
    say("vjp 2-D", limiter)
    case, nodes, q = C.euler_bump(12, 8, 1, bc_id=(-3, -4, -6, -8, -9, -9))
    case = dataclasses.replace(case, limiter=limiter)
    g = GpuBlock(case)
    g.set_grid(nodes)
    _, worst = check_columns(g, case, q, lambda x: gpu_residual(g, x), 3)
    say("  worst", worst)
    assert worst <= TOL


@pytest.mark.parametrize("bcs", [(-3, -4, -6, -6, -9, -9), (-8, -8, -5, -6, -6, -6)])
def test_transposition_identity_32_directions(bcs):
    """WENO5 + van Leer at M 0.5: a smooth residual, 32 seeded directions.
    A far field runs through it only where the flow crosses it (imin, imax
    of the second mix): where the normal velocity is near zero (a jmax far
    field above a channel) its inflow/outflow switch is a jump that
    directions moving every cell cross; the columns cover that face."""
    say("vjp directions", bcs)
    case, nodes, q = C.euler_bump(14, 10, 6, flux="van_leer", recon="weno", bc_id=bcs, seed=4)
    g = GpuBlock(case)
    g.set_grid(nodes)
    _, worst = check_identity(g, case, q, lambda x: gpu_residual(g, x), 4, 32)
    say("  worst", worst)
    assert worst <= TOL


@pytest.mark.parametrize("flux,recon,bcs", [
    ("ausmP", "muscl", (-3, -4, -6, -8, -9, -9)),
    ("slau", "weno", (-8, -8, -6, -8, -9, -9)),
    ("ldfss0", "ppm", (-1, -2, -5, -7, -9, -9)),
])
def test_adjoint_of_the_oracle_residual(flux, recon, bcs):
    """Columns of the oracle's residual: 150 seeded inputs."""
    say("vjp vs oracle", flux, recon, bcs)
    case, nodes, q = C.euler_bump(12, 9, 5, flux=flux, recon=recon, bc_id=bcs, seed=31)
    g = GpuBlock(case)
    g.set_grid(nodes)
    o = Oracle(case)
    o.set_nodes(nodes)

    def oracle_residual(x):
        o.set_state(x)
        assert o.residual(1) == 0
        return o.arr("res").copy()

    _, worst = check_columns(g, case, q, oracle_residual, 17, sample=150)
    say("  worst", worst)
    assert worst <= TOL


def test_adjoint_is_linear_in_the_weights():
    case, nodes, q = C.euler_bump(12, 9, 5)
    g = GpuBlock(case)
    g.set_grid(nodes)
    g.upload_state(q)
    w1, w2 = weights(case, 1), weights(case, 2)
    b1, b2 = g.vjp(w1), g.vjp(w2)
    b12 = g.vjp(2.0 * w1 - 3.0 * w2)
    scale = np.abs(b12).max()
    assert np.abs(b12 - (2.0 * b1 - 3.0 * b2)).max() <= 1e-13 * scale
    assert np.array_equal(g.vjp(w1), b1)  # deterministic: no atomics


def test_gradcheck_through_torch():
    torch = pytest.importorskip("torch")
    from festgpu.autograd import inviscid_residual
    case, nodes, q = C.euler_bump(6, 5, 1, bc_id=(-3, -4, -6, -8, -9, -9), seed=41)
    g = GpuBlock(case)
    g.set_grid(nodes)
    qi = torch.from_numpy(np.ascontiguousarray(q[INNER])).requires_grad_(True)
    say("gradcheck 6x5x1")
    assert torch.autograd.gradcheck(lambda x: inviscid_residual(x, g), (qi,), eps=1e-6, atol=1e-7, rtol=1e-5)


@pytest.mark.parametrize("change", [
    dict(mu_ref=1e-3, bc_id=(-3, -4, 0, -8, -9, -9)),  # viscous with an interface face
    dict(bc_id=(-3, -4, -5, -8, -9, -9), fix_twall=(0, 0, 2.0, 0, 0, 0)),  # isothermal density rule
])
def test_undifferentiated_configurations_fail(change):
    case, nodes, q = C.euler_bump(12, 9, 5)
    case = dataclasses.replace(case, **change)
    g = GpuBlock(case)
    g.set_grid(nodes)
    g.upload_state(q)
    with pytest.raises(FestError) as e:
        g.vjp(weights(case, 0))
    assert e.value.code == -3


def test_turbulent_handles_fail_fp32_handles_take_the_fp64_adjoint():
    """The SST bump has seven variables: the raw entry point refuses it.  A
    handle in the FP32 face mode gets the adjoint of the FP64 residual:
    the same as an FP64 handle's, bit for bit."""
    case, nodes, q = C.sst_bump(12, 9, 5)
    g = GpuBlock(case)
    g.set_grid(nodes, wall_dist=True)
    g.upload_state(q)
    w = np.zeros(case.n_var * case.ncells)
    out = np.zeros(5 * int(np.prod(case.cell_shape())))
    with pytest.raises(FestError) as e:
        g._check(native.lib().fest_vjp(g.h, native._ptr(w), native._ptr(out), None))
    assert e.value.code == -3
    ec, en, eq = C.euler_bump(12, 9, 5)
    bars = []
    for prec in (0, 1):
        gp = GpuBlock(ec, precision=prec)
        gp.set_grid(en)
        gp.upload_state(eq)
        bars.append(gp.vjp(weights(ec, 0)))
        gp.close()
    assert np.array_equal(bars[0], bars[1])


def warped(nodes):
    """The grid moved by a smooth warp of all three coordinates, so that no
    face normal has an exact zero component (see node_columns)."""
    nodes = nodes.copy()
    inn = (slice(3, -3),) * 3
    X, Y, Z = (nodes[c][inn].copy() for c in range(3))
    nodes[0][inn] = X + 0.01 * np.sin(2 * np.pi * (Y + Z))
    nodes[1][inn] = Y + 0.005 * np.sin(2 * np.pi * (X + Z)) * (1.0 - Y)
    nodes[2][inn] = Z + 0.01 * np.sin(2 * np.pi * (X + Y))
    return nodes


def node_columns(case, nodes, q, seed, sample=None):
    """max over real node coordinates of |bar_nodes - FD| / max|bar_nodes|,
    FD by central differences of <w, R> under node displacements of 1e-7,
    the ghost nodes and metrics recomputed by fest_set_grid.

    The grid is first warped in all three coordinates.  On the plain bump
    grid the wall's normal and the I faces beside it have exact zero
    components away from the bump (n_x = 0, r_y = 0, and n_z = r_z = 0
    everywhere), so the ghost cells of a slip wall or far field equal the
    interior cell exactly in some velocity components and the limiter of
    that component (MUSCL, and the boundary MUSCL of PPM and WENO) sits
    exactly on its kink: moving a node tilts the normal and steps across
    the kink, and a central difference there averages the two one-sided
    slopes while the adjoint takes the branch the residual took (measured:
    the difference is the same at h = 1e-7, 1e-8, 1e-9)."""
    nodes = warped(nodes)
    g = GpuBlock(case)
    g.set_grid(nodes)
    g.upload_state(q)
    w = weights(case, seed)
    _, bn = g.vjp(w, nodes=True)
    real = (slice(None), slice(3, case.kmx + 3), slice(3, case.jmx + 3), slice(3, case.imx + 3))
    ghost = np.ones(case.node_shape(), bool)
    ghost[real] = False
    assert np.all(bn[ghost] == 0.0)
    inner = bn[real]
    top = np.abs(inner).max()
    h = 1e-7  # the grid spans O(1); its smallest spacing is 0.02
    shape = inner.shape
    n = int(np.prod(shape))
    idx = range(n) if sample is None else np.random.default_rng(seed + 3).choice(n, sample, replace=False)

    def R(x):
        g.set_grid(x)
        g.upload_state(q)
        g.residual(1)
        return g.download("residue")

    worst = 0.0
    for f in idx:
        c, k, j, i = np.unravel_index(int(f), shape)
        xp, xm = nodes.copy(), nodes.copy()
        xp[c, 3 + k, 3 + j, 3 + i] += h
        xm[c, 3 + k, 3 + j, 3 + i] -= h
        fd = float(np.sum(w * (R(xp) - R(xm)))) / (2 * h)
        worst = max(worst, abs(inner[c, k, j, i] - fd) / top)
    return worst


@pytest.mark.parametrize("flux,recon,bcs,nk", [
    ("ausmP", "muscl", (-3, -4, -6, -6, -9, -9), 5),   # slip walls on jmin and jmax
    ("slau", "weno", (-8, -8, -6, -6, -6, -6), 5),     # far-field normals; slip walls on j and k max
    ("ldfss0", "ppm", (-11, -4, -5, -6, -9, -9), 5),   # total-pressure inlet
    ("van_leer", "none", (-1, -2, -6, -8, -9, -9), 5),
    ("ausmUP", "muscl", (-3, -4, -6, -6, -9, -9), 1),  # 2-D
])
def test_node_adjoint_columns(flux, recon, bcs, nk):
    """Every real node coordinate of an 8 x 6 x nk grid.  Far fields sit
    where the flow crosses them (or on a first-order scheme): above the
    channel the normal velocity is near zero and node moves flip the
    Riemann state's inflow/outflow branch, a jump."""
    say("vjp nodes", flux, recon, bcs, nk)
    case, nodes, q = C.euler_bump(8, 6, nk, flux=flux, recon=recon, bc_id=bcs, seed=13)
    worst = node_columns(case, nodes, q, 19)
    say("  worst", worst)
    assert worst <= TOL


def test_metric_adjoints_against_uploaded_geometry():
    """The adjoints of the stored face normals and areas (fest_device_ptr
    "vjp_metrics_*" after fest_vjp with bar_nodes) against central
    differences of <w, R> with one metric entry moved (fest_upload_geometry):
    every face of the first two layers next to each j face, 40 seeded faces
    of each direction elsewhere, all four entries."""
    case, nodes, q = C.euler_bump(8, 6, 5, flux="slau", recon="weno", bc_id=(-8, -8, -6, -6, -6, -6), seed=29)
    nodes = warped(nodes)
    g = GpuBlock(case)
    g.set_grid(nodes)
    g.upload_state(q)
    w = weights(case, 29)
    g.vjp(w, nodes=True)
    L = native.lib()
    L.hipMemcpy.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t, ctypes.c_int]
    L.hipMemcpy.restype = ctypes.c_int
    mb = []
    for name in ("vjp_metrics_i", "vjp_metrics_j", "vjp_metrics_k"):
        p, n = g.device_ptr(name)
        out = np.empty(n)
        assert L.hipMemcpy(out.ctypes.data, p, out.nbytes, 2) == 0
        mb.append(out.reshape(4, -1))
    cells = g.download("cells")
    faces = g.download("faces")

    def R(fc):
        g.upload_geometry(cells, *fc)
        g.upload_state(q)
        g.residual(1)
        return g.download("residue")

    ext = [(case.imx, case.jmx - 1, case.kmx - 1), (case.imx - 1, case.jmx, case.kmx - 1),
           (case.imx - 1, case.jmx - 1, case.kmx)]
    rng = np.random.default_rng(31)
    worst = 0.0
    for d in range(3):
        ni, nj, nk = ext[d]
        allf = [(i, j, k) for k in range(1, nk + 1) for j in range(1, nj + 1) for i in range(1, ni + 1)]
        near = [f for f in allf if f[1] <= 2 or f[1] >= nj - 1]
        far = [f for f in allf if f not in near]
        pick = near + [far[x] for x in rng.choice(len(far), min(40, len(far)), replace=False)]
        top = np.abs(mb[d]).max()
        for i, j, k in pick:
            f = (i - 1) + ni * ((j - 1) + nj * (k - 1))
            for c, comp in enumerate((1, 2, 3, 0)):  # (nx, ny, nz, A) in the download's (A, nx, ny, nz)
                fp = [x.copy() for x in faces.values()]
                fm = [x.copy() for x in faces.values()]
                fp[d][comp, k + 2, j + 2, i + 2] += H
                fm[d][comp, k + 2, j + 2, i + 2] -= H
                fd = float(np.sum(w * (R(fp) - R(fm)))) / (2 * H)
                worst = max(worst, abs(mb[d][c, f] - fd) / top)
    say("metric adjoints worst", worst)
    assert worst <= TOL


def test_node_adjoint_refused_for_pole_and_weno_nm():
    for change in (dict(bc_id=(-1, -2, -5, -7, -9, -9)), dict(recon="weno_NM")):
        case, nodes, q = C.euler_bump(12, 9, 5, **change)
        g = GpuBlock(case)
        g.set_grid(nodes)
        g.upload_state(q)
        g.vjp(weights(case, 0))  # the state adjoint alone is fine
        with pytest.raises(FestError) as e:
            g.vjp(weights(case, 0), nodes=True)
        assert e.value.code == -3


def test_unsteady_adjoint_with_checkpoints():
    """d J(q_N) / d q_0 through 12 RK4 steps (global time step, WENO5 + van
    Leer), J the pressure x-force on the jmin slip wall: the host-driven step equals the
    device's RK4 step, the reverse sweep with checkpoints every 4 steps
    matches central differences of the whole trajectory, and equals the
    sweep without intermediate checkpoints bit for bit."""
    from festgpu.adjoint import INNER as IN, RK4Adjoint, wall_force_x
    dt = 0.004
    case, nodes, q = C.euler_bump(12, 9, 5, flux="van_leer", recon="weno", bc_id=(-3, -4, -6, -8, -9, -9),
                                  seed=51)
    case = dataclasses.replace(case, time_local=0, global_time_step=dt, time_accuracy="RK4")
    g = GpuBlock(case)
    g.set_grid(nodes)
    adj = RK4Adjoint(g, dt)
    qi = np.ascontiguousarray(q[IN])
    g.upload_state(q)
    g.get_next_solution(1)
    qdev = g.download("qp")[IN]
    qhost = adj.step(qi)
    d = np.abs(qhost - qdev).max() / np.abs(qdev).max()
    say("host step vs device step", d)
    assert d <= 1e-13
    J, dJ = wall_force_x(g, 2)
    N, K = 12, 4
    qN, grad = adj.gradient(qi, N, K, dJ)
    assert adj.vjp_calls == 4 * N
    worst = 0.0
    for dq in directions(case, 61, 4):
        a = float(np.sum(grad * dq))
        jp = J(adj.forward(qi + H * dq, N, N)[0])
        jm = J(adj.forward(qi - H * dq, N, N)[0])
        b = (jp - jm) / (2 * H)
        worst = max(worst, abs(a - b) / max(abs(a), abs(b)))
    say("  unsteady worst", worst)
    assert worst <= TOL
    _, grad1 = adj.gradient(qi, N, N, dJ)
    assert np.array_equal(grad, grad1)


@pytest.mark.parametrize("recon", ["muscl", "weno"])
def test_adjoint_of_split_blocks_composes(recon):
    """A block cut in two along i (interface faces between the halves,
    festgpu.multiblock.split with consistent ghost layers): each half's
    fest_vjp leaves on its interface ghost cells the adjoint owed to the
    other half; adding every half's adjoint at the global cell indices (the
    transpose of the interface copy) gives the uncut block's adjoint."""
    from festgpu.multiblock import split
    case, nodes, q = C.euler_bump(16, 8, 5, recon=recon, bc_id=(-3, -4, -6, -6, -9, -9), seed=37)
    w = weights(case, 37)
    g = GpuBlock(case)
    g.set_grid(nodes)
    g.upload_state(q)
    whole = g.vjp(w)
    blocks, _ = split(case, nodes, q, 0, [9])
    total = np.zeros_like(whole)
    s = 1
    for cb, nb, qb in blocks:
        gb = GpuBlock(cb)
        gb.set_grid(nb)
        gb.upload_state(qb)
        n = cb.imx  # nodes s..s+n-1: cells s..s+n-2, padded local -2..n+2
        wb = np.ascontiguousarray(w[..., s - 1: s - 1 + n - 1])
        bar = gb.vjp(wb)
        # local padded i index a <-> global padded index a + s - 1
        total[..., s - 1: s - 1 + n + 5] += bar
        s += n - 1
    inner = (slice(None),) + (slice(3, -3),) * 3
    d = np.abs(total[inner] - whole[inner]).max() / np.abs(whole[inner]).max()
    say("split blocks", recon, d)
    assert d <= 1e-13


@pytest.mark.parametrize("axis,bcs", [
    (0, (-3, -4, -6, -6, -9, -9)),   # cut along i: two interface links
    (2, (-3, -4, -6, -6, -9, -9)),   # cut along k of a periodic k: interfaces and a cross-block periodic pair
])
def test_group_adjoint_equals_the_uncut_block(axis, bcs):
    """fest_group_vjp on a block cut in two: the interfaces exchanged, each
    half's adjoint taken, the interface ghost adjoints sent back to the
    cells they copied; every ghost cell ends at zero and the halves'
    interiors equal the uncut block's adjoint; the halves' node adjoints,
    summed on the nodes they share, equal the uncut block's.  The uncut block is a group
    too (split with no cut): a periodic direction then becomes the
    cross-block periodic pair (-10) on the block itself, whose faces take
    the full stencil where the single-block periodic rule (-9) takes the
    boundary face states, so the two residuals differ."""
    from festgpu.multiblock import split
    from festgpu.native import GpuGroup
    case, nodes, q = C.euler_bump(12, 8, 8, recon="weno", bc_id=bcs, seed=43)
    w = weights(case, 43)
    (uc, un, uq), = split(case, nodes, q, axis, [])[0]
    g = GpuBlock(uc)
    g.set_grid(un)
    g.upload_state(uq)
    g1 = GpuGroup([g], split(case, nodes, q, axis, [])[1])
    (whole,), (whole_nodes,) = g1.vjp([w], nodes=True)
    g1.close()
    cut = (case.imx + 1) // 2 if axis == 0 else (case.kmx + 1) // 2
    blocks, links = split(case, nodes, q, axis, [cut])
    gs = []
    for cb, nb, qb in blocks:
        gb = GpuBlock(cb)
        gb.set_grid(nb)
        gb.upload_state(qb)
        gs.append(gb)
    grp = GpuGroup(gs, links)
    ax = 3 - axis  # array axis of the cut direction in (k, j, i) + the variable axis
    ws, s = [], 1
    for cb, _, _ in blocks:
        n = (cb.imx, cb.jmx, cb.kmx)[axis]
        sl = [slice(None)] * 4
        sl[ax] = slice(s - 1, s - 1 + n - 1)
        ws.append(np.ascontiguousarray(w[tuple(sl)]))
        s += n - 1
    bars, bns = grp.vjp(ws, nodes=True)
    # node adjoints: each block's real nodes at their global node index,
    # a node on the cut shared by the two blocks
    total_nodes = np.zeros_like(whole_nodes)
    s = 1
    for (cb, _, _), bn in zip(blocks, bns):
        n = (cb.imx, cb.jmx, cb.kmx)[axis]
        src = [slice(None), slice(3, cb.kmx + 3), slice(3, cb.jmx + 3), slice(3, cb.imx + 3)]
        dst = list(src)
        dst[ax] = slice(s + 2, s + 2 + n)
        total_nodes[tuple(dst)] += bn[tuple(src)]
        s += n - 1
    dn = np.abs(total_nodes - whole_nodes).max() / np.abs(whole_nodes).max()
    say("group node adjoint", axis, dn)
    assert dn <= 1e-13
    worst, s = 0.0, 1
    scale = np.abs(whole).max()
    for (cb, _, _), bar in zip(blocks, bars):
        n = (cb.imx, cb.jmx, cb.kmx)[axis]
        ghost = np.ones(cb.cell_shape(), bool)
        ghost[3:-3, 3:-3, 3:-3] = False
        assert np.all(bar[:, ghost] == 0.0)
        inner = (slice(None),) + (slice(3, -3),) * 3
        sl = [slice(None), slice(3, -3), slice(3, -3), slice(3, -3)]
        sl[ax] = slice(s + 2, s + 2 + n - 1)
        worst = max(worst, np.abs(bar[inner] - whole[tuple(sl)]).max() / scale)
        s += n - 1
    grp.close()
    say("group adjoint", axis, worst)
    assert worst <= 1e-13
