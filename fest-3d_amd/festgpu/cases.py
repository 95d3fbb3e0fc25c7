"""Synthetic FEST-3D cases: grids, freestream, boundary setup and seeded
initial states for the workloads this package is measured on.

All arrays use the reference's Fortran layout (column-major with explicit
lower bounds, src/update.f90:143-165) stored in C-ordered numpy arrays whose
last axis is i:  a padded cell array (-2:imx+2,-2:jmx+2,-2:kmx+2,1:n_var) is
``np.empty((n_var, kmx+5, jmx+5, imx+5))`` and element (i,j,k,l) lives at
``[l-1, k+2, j+2, i+2]``.  Node arrays (-2:imx+3,...) are
``(3, kmx+6, jmx+6, imx+6)``.

Random numbers come from a counter-based SplitMix64 stream,
``u_n = (mix(seed + (n+1)*0x9E3779B97F4A7C15) >> 11) * 2**-53``, so every
array is reproducible element by element.

Freestream conventions follow src/state.f90:75-133: Reynolds number per unit
length L = 1 (``mu_ref = rho_inf*|V_inf|/Re``), ``tu_inf`` in percent,
``k_inf = 1.5 (|V| tu)^2``, ``omega_inf = rho k / (mu_ref * mu_ratio)``.
"""
from __future__ import annotations

import dataclasses
import math
from typing import List, Optional

import numpy as np

# flux / reconstruction / turbulence ids shared with the C-ABI
FLUX = {"ausm": 0, "ausmP": 1, "ausmUP": 2, "slau": 3, "ldfss0": 4, "van_leer": 5}
RECON = {"none": 0, "muscl": 1, "ppm": 2, "weno": 3, "weno_NM": 4}
LIMITER = {"koren": 0, "van_albada": 1}
TURB = {"none": 0, "sst": 1, "sst2003": 2, "sa": 3, "kkl": 4}
TRANSITION = {"none": 0, "bc": 1, "lctm2015": 2}
TIME_ACC = {"none": 0, "RK2": 2, "RK4": 4, "TVDRK2": 12, "TVDRK3": 13, "implicit": 20}

_GOLDEN = np.uint64(0x9E3779B97F4A7C15)
_M1 = np.uint64(0xBF58476D1CE4E5B9)
_M2 = np.uint64(0x94D049BB133111EB)


def splitmix_uniform(seed: int, n: int, offset: int = 0) -> np.ndarray:
    """n uniform doubles in [0,1) from draws offset..offset+n-1 of stream `seed`."""
    with np.errstate(over="ignore"):
        idx = np.arange(offset + 1, offset + n + 1, dtype=np.uint64)
        z = np.uint64(seed) + idx * _GOLDEN
        z = (z ^ (z >> np.uint64(30))) * _M1
        z = (z ^ (z >> np.uint64(27))) * _M2
        z = z ^ (z >> np.uint64(31))
    return (z >> np.uint64(11)).astype(np.float64) * (2.0 ** -53)


@dataclasses.dataclass
class Case:
    """Everything one block needs: extent, scheme, flow, boundary values."""

    imx: int
    jmx: int
    kmx: int
    flux: str = "ausmP"
    recon: str = "muscl"
    turb: str = "none"
    time_local: int = 0          # 1: 'l' local steps, 0: 'g' global step
    time_accuracy: str = "RK4"
    accur: int = 0
    lim: tuple = (1, 1, 1)
    tlim: tuple = (0, 0, 0)
    mu_sutherland: int = 0
    bc_id: tuple = (-9, -9, -5, -5, -9, -9)
    global_time_step: float = -1.0
    CFL: float = 0.5
    rho_inf: float = 1.0
    u_inf: float = 0.0
    v_inf: float = 0.0
    w_inf: float = 0.0
    p_inf: float = 101325.0
    tk_inf: float = 0.0
    tw_inf: float = 0.0
    vel_mag: float = 0.0
    MInf: float = 0.0
    gm: float = 1.4
    R_gas: float = 287.0
    mu_ref: float = 0.0
    T_ref: float = 300.0
    S_ref: float = 110.0
    Pr: float = 0.72
    tPr: float = 0.9
    fix_rho: tuple = (0.0,) * 6
    fix_u: tuple = (0.0,) * 6
    fix_v: tuple = (0.0,) * 6
    fix_w: tuple = (0.0,) * 6
    fix_p: tuple = (0.0,) * 6
    fix_tk: tuple = (0.0,) * 6
    fix_tw: tuple = (0.0,) * 6
    fix_twall: tuple = (0.0,) * 6
    fix_tp: tuple = (0.0,) * 6      # total pressure of a -11 face
    tv_inf: float = 0.0             # Spalart-Allmaras free-stream nu-tilde
    fix_tv: tuple = (0.0,) * 6      # fixed nu-tilde of each face
    transition: str = "none"       # 'bc': the intermittency of the SA and SST sources;
                                    # 'lctm2015': the intermittency equation (variable 8, SST only)
    tu_inf: float = 0.0             # free-stream turbulence intensity, percent
    tkl_inf: float = 0.0            # k-kL free-stream kL
    fix_tkl: tuple = (0.0,) * 6     # fixed kL of each face (k-kL)
    pbs: tuple = (0, 0, 0)          # i/j/kPB_switch: the pressure-based switch of MUSCL and PPM
    limiter: str = "koren"          # MUSCL limiter: the reference's Koren-alpha, or "van_albada"
    tgm_inf: float = 1.0            # free-stream intermittency (lctm2015; vartypes.f90:258)
    fix_tgm: tuple = (0.0,) * 6     # fixed intermittency of each face (lctm2015)
    name: str = "case"

    @property
    def sst_family(self) -> bool:
        """SST k-omega, in its original or its 2003 form: k and omega are
        variables 6 and 7."""
        return self.turb in ("sst", "sst2003")

    @property
    def two_equation(self) -> bool:
        """SST or k-kL: two turbulence variables, 6 and 7."""
        return self.sst_family or self.turb == "kkl"

    @property
    def n_var(self) -> int:
        """src/state.f90:281-320: lctm2015 adds the intermittency."""
        n = 7 if self.two_equation else 6 if self.turb == "sa" else 5
        return n + 1 if self.transition == "lctm2015" else n

    @property
    def n_grad(self) -> int:
        """Gradient slots u, v, w, T, then k, omega (SST), k, kL (k-kL) or
        nu-tilde (SA), then the intermittency under lctm2015
        (src/gradients.f90:236-275)."""
        n = 6 if self.two_equation else 5 if self.turb == "sa" else 4
        return n + 1 if self.transition == "lctm2015" else n

    @property
    def ncells(self) -> int:
        return (self.imx - 1) * (self.jmx - 1) * (self.kmx - 1)

    def cell_shape(self, nvar: Optional[int] = None):
        s = (self.kmx + 5, self.jmx + 5, self.imx + 5)
        return s if nvar is None else (nvar,) + s

    def node_shape(self):
        return (3, self.kmx + 6, self.jmx + 6, self.imx + 6)


def freestream(case: Case, T_inf=300.0, p_inf=101325.0, mach=0.5, reynolds=1000.0,
               tu_percent=0.1, mu_ratio=0.01, direction=(1.0, 0.0, 0.0)) -> Case:
    """Fill the dimensional freestream of `case` (src/state.f90:75-133)."""
    gm, R = case.gm, case.R_gas
    rho = p_inf / (R * T_inf)
    c = math.sqrt(gm * p_inf / rho)
    V = mach * c
    d = np.asarray(direction, dtype=float)
    d = d / np.linalg.norm(d)
    case.rho_inf, case.p_inf = rho, p_inf
    case.u_inf, case.v_inf, case.w_inf = (float(V * d[0]), float(V * d[1]), float(V * d[2]))
    case.vel_mag = math.sqrt(case.u_inf ** 2 + case.v_inf ** 2 + case.w_inf ** 2)
    case.MInf = case.vel_mag / math.sqrt(gm * p_inf / rho)
    case.mu_ref = rho * case.vel_mag * 1.0 / reynolds
    case.T_ref = T_inf
    if case.sst_family:
        tu = tu_percent / 100
        case.tk_inf = 1.5 * ((case.vel_mag * tu) ** 2)
        case.tw_inf = rho * case.tk_inf / (case.mu_ref * mu_ratio)
    if case.turb == "sa":
        case.tv_inf = mu_ratio * case.mu_ref / rho
    if case.turb == "kkl":
        # src/state.f90:101-103: both from the free-stream sound speed
        case.tk_inf = 9 * (1e-9) * (c ** 2)
        case.tkl_inf = 1.5589 * (1e-6) * (case.mu_ref * c) / rho
    case.tu_inf = tu_percent
    return case


def case_to_json(case: Case) -> str:
    """Lossless JSON of a case (floats round-trip through repr)."""
    import json

    def conv(v):
        if isinstance(v, (tuple, list)):
            return [conv(x) for x in v]
        if isinstance(v, (np.floating, float)):
            return float(v)
        if isinstance(v, (np.integer, int)):
            return int(v)
        return v

    return json.dumps({k: conv(v) for k, v in dataclasses.asdict(case).items()})


def case_from_json(text: str) -> Case:
    import json
    d = json.loads(text)
    return Case(**{k: tuple(v) if isinstance(v, list) else v for k, v in d.items()})


def freestream_state(case: Case) -> np.ndarray:
    q = np.empty(case.cell_shape(case.n_var))
    vals = [case.rho_inf, case.u_inf, case.v_inf, case.w_inf, case.p_inf]
    vals += ([case.tv_inf] if case.turb == "sa" else [case.tk_inf, case.tkl_inf] if case.turb == "kkl"
             else [case.tk_inf, case.tw_inf])
    if case.transition == "lctm2015":
        vals = vals[:7] + [case.tgm_inf]  # src/state.f90:255-268
    for l in range(case.n_var):
        q[l] = vals[l]
    return q


def cartesian_nodes(case: Case, spacing=(1.0, 1.0, 1.0), warp: float = 0.0, seed: int = 0) -> np.ndarray:
    """Uniform Cartesian nodes 0..(n-1)*h; optional smooth warp for tests."""
    x = np.zeros(case.node_shape())
    i = np.arange(1, case.imx + 1)
    j = np.arange(1, case.jmx + 1)
    k = np.arange(1, case.kmx + 1)
    K, J, I = np.meshgrid(k, j, i, indexing="ij")
    X = (I - 1) * spacing[0]
    Y = (J - 1) * spacing[1]
    Z = (K - 1) * spacing[2]
    if warp:
        Lx, Ly, Lz = X.max() or 1.0, Y.max() or 1.0, Z.max() or 1.0
        X = X + warp * Lx * np.sin(2 * np.pi * Y / Ly) * np.sin(2 * np.pi * Z / Lz)
        Y = Y + 0.5 * warp * Ly * np.sin(2 * np.pi * X / Lx) * np.cos(np.pi * Z / Lz)
        Z = Z + 0.5 * warp * Lz * np.sin(np.pi * X / Lx) * np.sin(2 * np.pi * Y / Ly)
    x[0, 3:-3, 3:-3, 3:-3] = X
    x[1, 3:-3, 3:-3, 3:-3] = Y
    x[2, 3:-3, 3:-3, 3:-3] = Z
    return x


def tanh_stretch(n_nodes: int, first: float, height: float = 1.0) -> np.ndarray:
    """Wall-clustered distribution s_0=0..s_{n-1}=height with s_1 = first.

    s(eta) = height*(1 + tanh(beta*(eta-1))/tanh(beta)); beta solved by
    bisection so that the first spacing equals `first`."""
    eta = np.linspace(0.0, 1.0, n_nodes)

    def first_of(beta):
        return height * (1 + math.tanh(beta * (eta[1] - 1)) / math.tanh(beta))

    lo, hi = 1e-6, 50.0
    for _ in range(200):
        mid = 0.5 * (lo + hi)
        if first_of(mid) > first:
            lo = mid
        else:
            hi = mid
    beta = 0.5 * (lo + hi)
    return height * (1 + np.tanh(beta * (eta - 1)) / math.tanh(beta))


def bump_nodes(case: Case, length=1.0, height=1.0, span=0.25, bump=0.05, first=1e-5) -> np.ndarray:
    """Channel over a sinusoidal bump y_w = bump*sin^2(pi(x-0.3)/0.4) on
    0.3L..0.7L; tanh wall-normal clustering with first spacing `first`,
    uniform in x and z."""
    x = np.zeros(case.node_shape())
    xs = np.linspace(0.0, length, case.imx)
    yw = np.where((xs >= 0.3 * length) & (xs <= 0.7 * length),
                  bump * length * np.sin(np.pi * (xs - 0.3 * length) / (0.4 * length)) ** 2, 0.0)
    s = tanh_stretch(case.jmx, first / height, 1.0)
    zs = np.linspace(0.0, span, case.kmx)
    X = np.broadcast_to(xs[None, None, :], (case.kmx, case.jmx, case.imx))
    Y = yw[None, None, :] + (height - yw[None, None, :]) * s[None, :, None]
    Z = np.broadcast_to(zs[:, None, None], (case.kmx, case.jmx, case.imx))
    x[0, 3:-3, 3:-3, 3:-3] = X
    x[1, 3:-3, 3:-3, 3:-3] = Y
    x[2, 3:-3, 3:-3, 3:-3] = Z
    return x


def gaussian_bumps(case: Case, q: np.ndarray, seed: int, nbumps=8, width=4.0, amp=0.01):
    """Adds seeded Gaussian bumps (width in cells) to rho,u,v,w,p of the
    interior cells of q, amplitude `amp` times rho_inf, |V_inf|, p_inf."""
    ni, nj, nk = case.imx - 1, case.jmx - 1, case.kmx - 1
    u = splitmix_uniform(seed, nbumps * 8)
    ic = np.arange(1, ni + 1) - 0.5
    jc = np.arange(1, nj + 1) - 0.5
    kc = np.arange(1, nk + 1) - 0.5
    K, J, I = np.meshgrid(kc, jc, ic, indexing="ij")
    scales = [case.rho_inf, case.vel_mag, case.vel_mag, case.vel_mag, case.p_inf]
    for b in range(nbumps):
        r = u[8 * b: 8 * b + 8]
        ci, cj, ck = r[0] * ni, r[1] * nj, r[2] * nk
        g = np.exp(-((I - ci) ** 2 + (J - cj) ** 2 + (K - ck) ** 2) / (2 * width * width))
        for l in range(5):
            q[l, 3:-3, 3:-3, 3:-3] += amp * scales[l] * (2 * r[3 + l] - 1) * g
    return q


def uniform_noise(case: Case, q: np.ndarray, seed: int, amp: float):
    """Seeded uniform noise of relative amplitude `amp` on every variable of
    the interior cells (velocities scaled by |V_inf|)."""
    n = case.ncells
    inner = (slice(3, -3),) * 3
    for l in range(case.n_var):
        r = splitmix_uniform(seed, n, offset=l * n).reshape(case.kmx - 1, case.jmx - 1, case.imx - 1)
        if l in (1, 2, 3):
            q[(l,) + inner] += amp * case.vel_mag * (2 * r - 1)
        else:
            q[(l,) + inner] *= 1 + amp * (2 * r - 1)
    return q


# ----------------------------------------------------------------- workloads
def laminar_box(ni=64, nj=64, nk=4, seed=1, flux="ausmP", recon="muscl", time_local=0):
    """Laminar box: periodic in i and k (BC -9), isothermal no-slip walls at
    j (Tw = 300 K), M=0.5, Re=1000, Pr=0.72, unit spacing, 8 Gaussian bumps
    of 1%.  At 64x64x4 cells this is the reference's CPU-runnable case."""
    c = Case(ni + 1, nj + 1, nk + 1, flux=flux, recon=recon, turb="none", time_local=time_local,
             accur=0, lim=(1, 1, 1), tlim=(0, 0, 0), bc_id=(-9, -9, -5, -5, -9, -9), CFL=0.5,
             name=f"laminar_box_{ni}x{nj}x{nk}")
    freestream(c, mach=0.5, reynolds=1000.0)
    c.Pr = 0.72
    c.fix_twall = (0.0, 0.0, 300.0, 300.0, 0.0, 0.0)
    nodes = cartesian_nodes(c)
    q = gaussian_bumps(c, freestream_state(c), seed)
    return c, nodes, q


def sst_bump(ni=256, nj=128, nk=128, seed=2, flux="slau", recon="weno", first=1e-5, turb="sst",
             transition="none"):
    """RANS SST channel over a 0.05L sinusoidal bump: WENO5 + SLAU, M=0.2,
    Re=1e6 per L, Tu=0.1%, mu_t/mu=0.01, 0.5% seeded noise.  Boundaries:
    subsonic inlet (-3), subsonic outlet (-4), adiabatic wall at jmin (-5),
    far field at jmax (-8), periodic in k (-9).  turb="sa" gives the same
    channel with Spalart-Allmaras (nu-tilde = mu_ratio mu / rho), turb="kkl"
    with k-kL (whose fixed omega values, never read, stay 0).
    transition="lctm2015" adds the intermittency (free stream and inlet 1,
    interior cells seeded uniformly in [0.05, 1] so that its production and
    destruction terms are both active)."""
    c = Case(ni + 1, nj + 1, nk + 1, flux=flux, recon=recon, turb=turb, time_local=1,
             accur=1, lim=(1, 1, 1), tlim=(1, 1, 1), mu_sutherland=1,
             bc_id=(-3, -4, -5, -8, -9, -9), CFL=0.5, name=f"sst_bump_{ni}x{nj}x{nk}",
             transition=transition)
    freestream(c, mach=0.2, reynolds=1.0e6, tu_percent=0.1, mu_ratio=0.01)
    c.Pr = 0.72
    c.tPr = 0.9
    inf = (c.rho_inf, c.u_inf, c.v_inf, c.w_inf, c.p_inf, c.tk_inf, c.tw_inf)
    c.fix_rho = (inf[0],) * 6
    c.fix_u = (inf[1],) * 6
    c.fix_v = (inf[2],) * 6
    c.fix_w = (inf[3],) * 6
    c.fix_p = (inf[4],) * 6
    c.fix_tk = (inf[5],) * 6
    c.fix_tw = (inf[6],) * 6
    c.fix_twall = (0.0,) * 6
    c.fix_tv = (c.tv_inf,) * 6
    c.fix_tkl = (c.tkl_inf,) * 6
    c.fix_tgm = (c.tgm_inf,) * 6
    nodes = bump_nodes(c, first=first)
    q = uniform_noise(c, freestream_state(c), seed, 0.005)
    if transition == "lctm2015":
        r = splitmix_uniform(seed + 101, c.ncells).reshape(c.kmx - 1, c.jmx - 1, c.imx - 1)
        q[7, 3:-3, 3:-3, 3:-3] = 0.05 + 0.95 * r
    return c, nodes, q


def flat_plate(ni=512, nj=256, nk=128, seed=5, flux="ldfss0", recon="ppm", turb="sst", first=5.5e-6,
               reynolds=5.0e6, length=2.0, height=0.5, span=0.25):
    """Turbulent flat plate: PPM + LDFSS(0), SST, M=0.2, Re=5e6 per L, plate
    length 2L along the whole lower wall (adiabatic, -5), subsonic inlet
    (-3) and outlet (-4), far field at jmax (-8), periodic in k.  Wall-normal
    tanh clustering with first spacing 5.5e-6 L (y+ about 1 at Re_x = 5e6).
    Initial state: a 1/7-power-law boundary layer u = U (y/delta)^(1/7) with
    delta(x) = 0.37 x Re_x^(-1/5) (x measured from 0.05L ahead of the inlet),
    plus 1% seeded noise; the inlet itself fixes uniform free-stream values
    (the reference's inlet imposes fixed values only)."""
    c = Case(ni + 1, nj + 1, nk + 1, flux=flux, recon=recon, turb=turb, time_local=1,
             accur=1, lim=(1, 1, 1), tlim=(1, 1, 1), mu_sutherland=1,
             bc_id=(-3, -4, -5, -8, -9, -9), CFL=0.5, name=f"flat_plate_{ni}x{nj}x{nk}")
    freestream(c, mach=0.2, reynolds=reynolds, tu_percent=0.1, mu_ratio=0.01)
    c.Pr = 0.72
    c.tPr = 0.9
    inf = (c.rho_inf, c.u_inf, c.v_inf, c.w_inf, c.p_inf, c.tk_inf, c.tw_inf)
    c.fix_rho, c.fix_u, c.fix_v, c.fix_w = (inf[0],) * 6, (inf[1],) * 6, (inf[2],) * 6, (inf[3],) * 6
    c.fix_p, c.fix_tk, c.fix_tw = (inf[4],) * 6, (inf[5],) * 6, (inf[6],) * 6
    c.fix_twall = (0.0,) * 6
    c.fix_tv = (c.tv_inf,) * 6
    c.fix_tkl = (c.tkl_inf,) * 6
    x = np.zeros(c.node_shape())
    xs = np.linspace(0.0, length, c.imx)
    ys = tanh_stretch(c.jmx, first / height, 1.0) * height
    zs = np.linspace(0.0, span, c.kmx)
    x[0, 3:-3, 3:-3, 3:-3] = xs[None, None, :]
    x[1, 3:-3, 3:-3, 3:-3] = ys[None, :, None]
    x[2, 3:-3, 3:-3, 3:-3] = zs[:, None, None]
    q = freestream_state(c)
    # cell centres of the interior cells
    xc = 0.5 * (xs[1:] + xs[:-1]) + 0.05 * length
    yc = 0.5 * (ys[1:] + ys[:-1])
    rex = c.rho_inf * c.vel_mag * xc / c.mu_ref
    delta = 0.37 * xc * rex ** -0.2
    prof = np.minimum(1.0, (yc[:, None] / delta[None, :]) ** (1.0 / 7.0))  # (j, i)
    q[1, 3:-3, 3:-3, 3:-3] = c.u_inf * prof[None, :, :]
    q = uniform_noise(c, q, seed, 0.01)
    return c, x, q


def euler_bump(ni=16, nj=12, nk=6, seed=7, flux="ausmP", recon="muscl", bc_id=(-3, -4, -6, -8, -9, -9),
               mach=0.5, amp=0.05, noise=0.01):
    """Inviscid (mu_ref 0) channel over the 0.05L bump in nondimensional
    units: free stream rho 1, p 1/gamma (sound speed 1), |V| = mach along
    x; every face's fixed values are the free stream's, the no-slip wall
    keeps the symmetric density (Tw 0) and a -11 face gets a total pressure
    0.2% above the free stream's.  Moderate wall clustering (first spacing
    0.02), 8 Gaussian bumps of amplitude `amp` and seeded noise of `noise`,
    so the limiters and the flux switches take varied branches."""
    c = Case(ni + 1, nj + 1, nk + 1, flux=flux, recon=recon, turb="none", time_local=1, accur=0,
             lim=(1, 1, 1), tlim=(0, 0, 0), bc_id=tuple(bc_id), CFL=0.5, name=f"euler_bump_{ni}x{nj}x{nk}")
    c.rho_inf, c.p_inf = 1.0, 1.0 / c.gm
    c.u_inf, c.v_inf, c.w_inf = float(mach), 0.0, 0.0
    c.vel_mag = c.MInf = float(mach)
    c.R_gas = 1.0 / c.gm  # T = p / (R rho) = 1 in the free stream
    c.T_ref = 1.0
    c.mu_ref = 0.0
    c.fix_rho, c.fix_u, c.fix_v = (c.rho_inf,) * 6, (c.u_inf,) * 6, (0.0,) * 6
    c.fix_w, c.fix_p, c.fix_twall = (0.0,) * 6, (c.p_inf,) * 6, (0.0,) * 6
    pt = 1.002 * c.p_inf * (1 + 0.5 * (c.gm - 1) * mach * mach) ** (c.gm / (c.gm - 1))
    c.fix_tp = (pt,) * 6
    nodes = bump_nodes(c, first=0.02, span=0.25 if nk > 1 else 0.05)
    q = gaussian_bumps(c, freestream_state(c), seed, amp=amp)
    q = uniform_noise(c, q, seed + 1, noise)
    return c, nodes, q


def chain_blocks(nblocks=64, lo=16, hi=96, seed=3, flux="ausmUP", recon="muscl", h=1.0 / 64, warp=0.02):
    """A seeded chain of `nblocks` laminar blocks (MUSCL + AUSM-UP, M=0.8):
    extents drawn from U{lo..hi}; block b+1 sits against block b across the
    faces of direction (b mod 3) and keeps block b's two extents along that
    face (the reference's interfaces join faces of equal extent), drawing
    the third anew.  Cells are cubes of side h, and every node is moved by
    the smooth warp x += warp sin(2 pi y) sin(2 pi z) of the global
    coordinates, so the nodes of joined faces coincide.  Faces not on the
    chain: subsonic inlet (-3) on i-min, outlet (-4) on i-max, adiabatic
    walls (-5) on j and k.  State: free stream plus 1% seeded noise.

    Returns ([(case, nodes, q)], links)."""
    from .multiblock import Link, aligned_link
    u = splitmix_uniform(seed, 3 + nblocks)
    ext = [int(lo + min(hi - lo, int(u[a] * (hi - lo + 1)))) for a in range(3)]
    sizes, origins = [list(ext)], [[0, 0, 0]]
    for b in range(1, nblocks):
        d = (b - 1) % 3
        e = list(sizes[-1])
        o = list(origins[-1])
        o[d] += e[d]
        e[d] = int(lo + min(hi - lo, int(u[3 + b] * (hi - lo + 1))))
        sizes.append(e)
        origins.append(o)
    blocks = []
    for b in range(nblocks):
        ni, nj, nk = sizes[b]
        ids = [-3, -4, -5, -5, -5, -5]
        if b > 0:
            d = (b - 1) % 3
            ids[2 * d] = b - 1
        if b < nblocks - 1:
            d = b % 3
            ids[2 * d + 1] = b + 1
        c = Case(ni + 1, nj + 1, nk + 1, flux=flux, recon=recon, turb="none", time_local=1, accur=0,
                 lim=(1, 1, 1), tlim=(0, 0, 0), bc_id=tuple(ids), CFL=0.5, name=f"chain{nblocks}_b{b}")
        freestream(c, mach=0.8, reynolds=1.0e4)
        c.Pr = 0.72
        c.fix_rho, c.fix_u, c.fix_v = (c.rho_inf,) * 6, (c.u_inf,) * 6, (c.v_inf,) * 6
        c.fix_w, c.fix_p, c.fix_twall = (c.w_inf,) * 6, (c.p_inf,) * 6, (0.0,) * 6
        nodes = np.zeros(c.node_shape())
        ii = origins[b][0] + np.arange(ni + 1)
        jj = origins[b][1] + np.arange(nj + 1)
        kk = origins[b][2] + np.arange(nk + 1)
        K, J, I = np.meshgrid(kk * h, jj * h, ii * h, indexing="ij")
        nodes[0, 3:-3, 3:-3, 3:-3] = I + warp * np.sin(2 * np.pi * J) * np.sin(2 * np.pi * K)
        nodes[1, 3:-3, 3:-3, 3:-3] = J
        nodes[2, 3:-3, 3:-3, 3:-3] = K
        q = uniform_noise(c, freestream_state(c), seed * 1000 + b, 0.01)
        blocks.append((c, nodes, q))
    cases = [c for c, _, _ in blocks]
    links = []
    for b in range(nblocks - 1):
        d = b % 3
        links.append(aligned_link(cases, b, 2 * d + 2, b + 1, 2 * d + 1))
        links.append(aligned_link(cases, b + 1, 2 * d + 1, b, 2 * d + 2))
    return blocks, links


WORKLOADS = {"laminar_box": laminar_box, "sst_bump": sst_bump, "flat_plate": flat_plate}
