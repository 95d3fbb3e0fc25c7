"""Readers for a FEST-3D case directory, so that the reference's own cases
run on the engine as they ship:

  system/control.md, fvscheme.md, flow.md   src/read_write/read/read.f90
  system/mesh/layout/layout.md             src/layout.f90 (one block per line)
  system/mesh/gridfiles/grid_XX.txt        src/grid.f90:60-140
  system/mesh/layout/mapping.txt           src/mapping.f90:64-258
  system/mesh/bc/bc_XX.md                  src/boundary/read_bc.f90:16-190
  time_directories/NNNN/process_XX.dat     src/read_write/write/write_output_tec.f90:244-305

The .md files are read by their "## title" lines, the value on the next
non-empty line; the reference reads them in a fixed order, and the titles
are what that order means.  Fixed boundary values default to the free
stream, as read_bc.f90's set_value does when a line carries no number.
"""
from __future__ import annotations

import dataclasses
import math
import os
from typing import Dict, List, Tuple

import numpy as np

from .cases import Case, freestream_state
from .multiblock import Link

_FACES = ("imn", "imx", "jmn", "jmx", "kmn", "kmx")


def read_md(path: str) -> Dict[str, List[str]]:
    """{title: [value lines]} of a system/*.md file."""
    out: Dict[str, List[str]] = {}
    key = None
    with open(path) as f:
        for line in f:
            s = line.strip()
            if s.startswith("##"):
                key = s[2:].strip()
                out[key] = []
            elif s and key is not None and not s.startswith("="):
                out[key].append(s)
    return out


def _first(md: Dict[str, List[str]], title: str) -> str:
    return md[title][0].split()[0]


def read_layout(path: str) -> List[Tuple[str, str, Tuple[int, ...]]]:
    """[(grid file, bc file, (imin, imax, jmin, jmax, kmin, kmax) ids)]"""
    blocks = []
    with open(path) as f:
        for line in f:
            s = line.split()
            if len(s) == 9 and not line.lstrip().startswith("#") and s[1].startswith("grid"):
                blocks.append((s[1], s[2], tuple(int(x) for x in s[3:9])))
    return blocks


def read_grid(path: str):
    """(imx, jmx, kmx, nodes (3, kmx, jmx, imx)); the file holds a header
    line "imx jmx kmx" and then x y z per node, i fastest."""
    with open(path) as f:
        imx, jmx, kmx = (int(x) for x in f.readline().split()[:3])
        a = np.loadtxt(f, dtype=np.float64, max_rows=imx * jmx * kmx)
    return imx, jmx, kmx, a.T.reshape(3, kmx, jmx, imx)


def padded_nodes(imx: int, jmx: int, kmx: int, nodes: np.ndarray) -> np.ndarray:
    """Interior nodes placed in the (-2:imx+3,...) node array; the ghost
    nodes are left for the engine's geometry setup (src/grid.f90:137-236)."""
    out = np.zeros((3, kmx + 6, jmx + 6, imx + 6))
    out[:, 3:kmx + 3, 3:jmx + 3, 3:imx + 3] = nodes
    return out


def read_mapping(path: str) -> List[Link]:
    """One link per map line: block b1's face f1 receives from block b2's face
    f2 over b2's node ranges with the direction switch (mapping.f90:121-171)."""
    links = []
    with open(path) as f:
        for line in f:
            s = line.split()
            if not s or s[0].startswith("#"):
                continue
            v = [int(x) for x in s[:14]]
            b1, f1, _, _, _, _, b2, f2, s21, e21, s22, e22, switch = v[:13]
            links.append(Link(b1, f1, b2, f2, s21, e21, s22, e22, switch))
    return links


def read_bc(path: str, case: Case) -> Case:
    """Fixed values of the six faces (read_bc.f90:16-190).  Under k-kL the
    omega values are neither filled nor read (read_bc.f90:98-108, 198-202):
    they stay 0.  lctm2015's intermittency: the free stream, or FIX_tgm
    (read_bc.f90:123-131, 213-218)."""
    vals = {k: list(getattr(case, k)) for k in ("fix_rho", "fix_u", "fix_v", "fix_w", "fix_p", "fix_tk",
                                                  "fix_tw", "fix_twall", "fix_tp", "fix_tv", "fix_tkl", "fix_tgm")}
    inf = dict(fix_rho=case.rho_inf, fix_u=case.u_inf, fix_v=case.v_inf, fix_w=case.w_inf, fix_p=case.p_inf,
               fix_tk=case.tk_inf, fix_tw=0.0 if case.turb == "kkl" else case.tw_inf, fix_twall=0.0, fix_tp=0.0,
               fix_tv=case.tv_inf, fix_tkl=case.tkl_inf,
               fix_tgm=case.tgm_inf if case.transition == "lctm2015" else 0.0)
    for k in vals:  # fill_fixed_values: the free stream everywhere
        vals[k] = [inf[k]] * 6
    names = {"FIX_DENSITY": "fix_rho", "FIX_X_SPEED": "fix_u", "FIX_Y_SPEED": "fix_v", "FIX_Z_SPEED": "fix_w",
             "FIX_PRESSURE": "fix_p", "WALL_TEMPERATURE": "fix_twall", "FIX_tk": "fix_tk", "FIX_tw": "fix_tw",
             "TOTAL_PRESSURE": "fix_tp", "FIX_tv": "fix_tv", "FIX_tkl": "fix_tkl"}
    if case.turb == "kkl":
        del names["FIX_tw"]
    if case.transition == "lctm2015":
        names["FIX_tgm"] = "fix_tgm"
    face = -1
    with open(path) as f:
        for line in f:
            s = line.strip()
            if s.startswith("#"):
                face += 1
            elif s.startswith("- ") and 0 <= face < 6:
                parts = s[2:].split()
                key = names.get(parts[0])
                if key and len(parts) > 1:
                    try:
                        vals[key][face] = float(parts[1])
                    except ValueError:
                        pass
    return dataclasses.replace(case, **{k: tuple(v) for k, v in vals.items()})


def read_case(root: str):
    """The blocks of a case directory: [(Case, padded nodes)], links, and the
    control values {CFL, max_iter, ...}."""
    sysd = os.path.join(root, "system")
    control = read_md(os.path.join(sysd, "control.md"))
    scheme = read_md(os.path.join(sysd, "fvscheme.md"))
    flow = read_md(os.path.join(sysd, "flow.md"))
    lim = [int(x) for x in " ".join(scheme["Switch: Limiter - Pressure based switch"]).split()]
    tlim = [int(x) for x in " ".join(scheme["Turbulence Limiter Switch"]).split()]
    turb = _first(scheme, "Turbulence model")
    transition = _first(scheme, "Transition model")
    if transition not in ("none", "bc", "lctm2015"):
        # src/read_write/read/read.f90:219-222
        raise ValueError(f"transition model {transition!r} is not supported (none, bc, lctm2015)")
    if transition == "lctm2015" and turb not in ("sst", "sst2003"):
        # src/source.f90:124-139 stops on lctm2015 outside SST
        raise ValueError("lctm2015 needs the sst or sst2003 turbulence model")
    if len(lim) < 6:
        raise ValueError("the limiter line needs three limiter and three pressure-based switches")
    pbs = tuple(lim[3:6])
    ts = scheme["Time Step"][0].split()
    rho, u, v, w, p = (float(_first(flow, t)) for t in ("Free Stream Density", "Free Stream X-Speed",
                                                         "Free Stream Y-Speed", "Free Stream Z-Speed",
                                                         "Free Stream Pressure"))
    gm, R = float(_first(flow, "Specific heat ratio")), float(_first(flow, "Gas Constant"))
    prs = flow["Prandtl numbers"][0].split()
    # src/state.f90:75-133
    vel = math.sqrt(u ** 2 + v ** 2 + w ** 2)
    mu_ref = float(_first(flow, "Reference Viscosity"))
    base = Case(imx=2, jmx=2, kmx=2, flux=_first(scheme, "Inviscid Flux Scheme"),
                recon=_first(scheme, "Higher Order Method"), turb=turb,
                time_local=1 if ts[0] == "l" else 0,
                global_time_step=float(ts[1]) if len(ts) > 1 else -1.0,
                time_accuracy=_first(scheme, "Time Integration Method"),
                accur=int(_first(scheme, "Higher Order Boundary Conditions")), lim=tuple(lim[:3]),
                tlim=tuple(tlim[:3]), mu_sutherland=int(_first(flow, "Viscosity law") == "sutherland_law"),
                CFL=float(_first(control, "CFL")), rho_inf=rho, u_inf=u, v_inf=v, w_inf=w, p_inf=p,
                vel_mag=vel, MInf=vel / math.sqrt(gm * p / rho), gm=gm, R_gas=R, mu_ref=mu_ref,
                T_ref=float(_first(flow, "Reference Temperature")),
                S_ref=float(_first(flow, "Sutherland Temperature")), Pr=float(prs[0]), tPr=float(prs[1]),
                transition=transition, pbs=pbs,
                # percent, as flow%tu_inf (src/read_write/read/read.f90)
                tu_inf=float(_first(flow, "Free Stream Turbulence Intensity")),
                # read.f90:306-310; vartypes.f90:258 defaults it to 1
                tgm_inf=float(_first(flow, "Free Stream Intermittency")) if "Free Stream Intermittency" in flow
                else 1.0)
    if turb in ("sst", "sst2003"):
        tu = float(_first(flow, "Free Stream Turbulence Intensity")) / 100
        ratio = float(_first(flow, "Free Stream Viscosity Ratio"))
        tk = 1.5 * ((vel * tu) ** 2)
        base = dataclasses.replace(base, tk_inf=tk, tw_inf=rho * tk / (mu_ref * ratio))
    if turb == "sa":  # src/state.f90:105-106
        ratio = float(_first(flow, "Free Stream Viscosity Ratio"))
        base = dataclasses.replace(base, tv_inf=ratio * mu_ref / rho)
    if turb == "kkl":  # src/state.f90:101-103, from the free-stream sound speed
        a = math.sqrt(gm * p / rho)
        base = dataclasses.replace(base, tk_inf=9 * (1e-9) * (a ** 2), tkl_inf=1.5589 * (1e-6) * (mu_ref * a) / rho)
    mesh = os.path.join(sysd, "mesh")
    blocks = []
    for b, (gfile, bcfile, ids) in enumerate(read_layout(os.path.join(mesh, "layout", "layout.md"))):
        imx, jmx, kmx, nodes = read_grid(os.path.join(mesh, "gridfiles", gfile))
        c = dataclasses.replace(base, imx=imx, jmx=jmx, kmx=kmx, bc_id=ids, name=f"{os.path.basename(root)}_{b}")
        c = read_bc(os.path.join(mesh, "bc", bcfile), c)
        blocks.append((c, padded_nodes(imx, jmx, kmx, nodes)))
    links = read_mapping(os.path.join(mesh, "layout", "mapping.txt"))
    ctl = dict(CFL=float(_first(control, "CFL")), max_iter=int(_first(control, "Maximum Iterations")),
               restart=int(_first(control, "Restart level")))
    return blocks, links, ctl


def read_tecplot(path: str, case: Case):
    """(nodes (3, kmx, jmx, imx), {name: cell array (kmx-1, jmx-1, imx-1)})
    of a block-packed Tecplot dump: x y z at the nodes, then the cell-centred
    variables named in the header (u v w Density Pressure ...)."""
    with open(path) as f:
        lines = f.read().split("\n")
    names, start = [], 0
    for n, line in enumerate(lines):
        tok = line.split()
        if not tok:
            continue
        try:
            float(tok[0])
            start = n
            break
        except ValueError:
            pass
        if tok[0].lower().startswith("zone") or "=" in tok[0] and not tok[0].startswith("variables"):
            continue
        if tok[0] == "variables":
            tok = tok[2:]
        if not any("=" in t for t in tok):
            names += tok
    a = np.array(" ".join(lines[start:]).split(), dtype=np.float64)
    imx, jmx, kmx = case.imx, case.jmx, case.kmx
    nn, ncl = imx * jmx * kmx, (imx - 1) * (jmx - 1) * (kmx - 1)
    nodes = a[:3 * nn].reshape(3, kmx, jmx, imx)
    cells = {}
    for v, name in enumerate(names[3:]):
        cells[name] = a[3 * nn + v * ncl:3 * nn + (v + 1) * ncl].reshape(kmx - 1, jmx - 1, imx - 1)
    return nodes, cells


def read_restart_list(path: str) -> List[str]:
    """The restart variables of system/output_control.md (its second
    {...} list)."""
    lists, cur = [], None
    with open(path) as f:
        for line in f:
            s = line.strip()
            if s.startswith("{"):
                cur = []
            elif s.startswith("}"):
                lists.append(cur)
                cur = None
            elif cur is not None and s:
                cur.append(s)
    return lists[1] if len(lists) > 1 else []


def restart_state(case: Case, path: str, restart_list: List[str]) -> np.ndarray:
    """The padded state a restart starts from: the free stream everywhere,
    then the interior cells of the variables in the restart list read from
    the dump (src/state.f90:178-188, src/read_write/read/read_output_tec.f90:
    100-136; other variables keep the free stream)."""
    qp = freestream_state(case)
    _, cells = read_tecplot(path, case)
    kmx, jmx, imx = case.kmx, case.jmx, case.imx
    inner = (slice(3, kmx + 2), slice(3, jmx + 2), slice(3, imx + 2))
    # the variable slot of each restart name (read_output_tec.f90:100-136
    # and the model's state pointers, :60-92): k and omega of SST, k and kL
    # of k-kL, nu-tilde of SA
    take = {"Velocity": [(1, "u"), (2, "v"), (3, "w")], "Density": [(0, "Density")], "Pressure": [(4, "Pressure")],
            "TKE": [(5, "TKE")], "Omega": [(6, "Omega")] if case.sst_family else [],
            "Kl": [(6, "Kl")] if case.turb == "kkl" else [], "tv": [(5, "tv")] if case.turb == "sa" else [],
            "tgm": [(7, "tgm")] if case.transition == "lctm2015" else []}  # read_output_tec.f90:130-131
    for var in restart_list:
        for l, name in take.get(var, []):
            if l < case.n_var:
                qp[(l,) + inner] = cells[name]
    return qp


def initial_states(root: str, blocks, ctl) -> List[np.ndarray]:
    """The state each block starts from: the free stream, or the restart
    dump time_directories/NNNN/process_XX.dat (control.md "Restart level")."""
    if ctl["restart"] == 0:
        return [freestream_state(c) for c, _ in blocks]
    rl = read_restart_list(os.path.join(root, "system", "output_control.md"))
    return [restart_state(c, os.path.join(root, "time_directories", f"{ctl['restart']:04d}", f"process_{b:02d}.dat"),
                          rl) for b, (c, _) in enumerate(blocks)]
