"""Solution output in the reference's formats: block-packed Tecplot dumps
(cell-centred and nodal), legacy VTK, the restart log, the residual-norm
file and the time-directory layout with its purge rule
(src/read_write/write/*.f90, src/resnorm.f90:1-377).

Numbers are written with Fortran edit descriptors restated here: ESw.dEe,
Ew.dEe, F0.d and Iw.m.  Each rounds the exact binary value of the double to
the digits shown, as Python's own %E/%f conversions do, so the text matches
the reference's byte for byte.
"""
import math
import os
import shutil
from typing import Dict, Iterable, List, Optional, Sequence

import numpy as np

from .cases import Case


# ------------------------------------------------------------ edit descriptors
def _nonfinite(x: float, w: int) -> str:
    """NaN, Inf, -Inf right-justified in w (w = 0: as they are)."""
    s = "NaN" if math.isnan(x) else ("Inf" if x > 0 else "-Inf")
    if w == 0:
        return s
    return s.rjust(w) if len(s) <= w else "*" * w


def f_es(x: float, w: int, d: int, e: int) -> str:
    """ESw.dEe: one nonzero digit before the point, d after, an exponent of
    e digits."""
    x = float(x)
    if not math.isfinite(x):
        return _nonfinite(x, w)
    m, ex = ("%.*E" % (d, x)).split("E")
    ex = int(ex)
    if abs(ex) >= 10 ** e:
        return "*" * w
    s = "%sE%s%0*d" % (m, "-" if ex < 0 else "+", e, abs(ex))
    return s.rjust(w) if len(s) <= w else "*" * w


def f_e(x: float, w: int, d: int, e: Optional[int] = None) -> str:
    """Ew.dEe: 0.ddd...E+xx with d significant digits; a leading zero when
    the width leaves room for it.  Without e (plain Ew.d) the exponent has
    two digits after E, or three digits and no E beyond 99."""
    x = float(x)
    if not math.isfinite(x):
        return _nonfinite(x, w)
    if x == 0.0:
        digits, ex = "0" * d, 0
    else:
        m, ex = ("%.*E" % (d - 1, abs(x))).split("E")
        digits, ex = m.replace(".", ""), int(ex) + 1
    sign = "-" if math.copysign(1.0, x) < 0 else ""
    es = "-" if ex < 0 else "+"
    if e is None:
        if abs(ex) > 999:
            return "*" * w
        body = ".%sE%s%02d" % (digits, es, abs(ex)) if abs(ex) <= 99 else ".%s%s%03d" % (digits, es, abs(ex))
    else:
        if abs(ex) >= 10 ** e:
            return "*" * w
        body = ".%sE%s%0*d" % (digits, es, e, abs(ex))
    s = sign + "0" + body
    if len(s) > w:
        s = sign + body
    return s.rjust(w) if len(s) <= w else "*" * w


def f_f0(x: float, d: int) -> str:
    """F0.d: the shortest field, d digits after the point, no leading zero
    before the point."""
    x = float(x)
    if not math.isfinite(x):
        return _nonfinite(x, 0)
    s = "%.*f" % (d, x)
    neg = s.startswith("-")
    if neg:
        s = s[1:]
    if s.startswith("0."):
        s = s[1:]
    return ("-" if neg else "") + s


def f_i(n: int, w: int, m: int = 0) -> str:
    """Iw.m (w = 0: I0)."""
    s = "%0*d" % (m, abs(n)) if m else str(abs(n))
    if n < 0:
        s = "-" + s
    if w == 0:
        return s
    return s.rjust(w) if len(s) <= w else "*" * w


# ------------------------------------------------------------ fields
GRAD_NAMES = {"Dudx": (0, 0), "Dudy": (1, 0), "Dudz": (2, 0), "Dvdx": (0, 1), "Dvdy": (1, 1), "Dvdz": (2, 1),
              "Dwdx": (0, 2), "Dwdy": (1, 2), "Dwdz": (2, 2), "DTdx": (0, 3), "DTdy": (1, 3), "DTdz": (2, 3),
              "Dtkdx": (0, 4), "Dtkdy": (1, 4), "Dtkdz": (2, 4), "Dtwdx": (0, 5), "Dtwdy": (1, 5),
              "Dtwdz": (2, 5)}


def block_fields(case: Case, qp: np.ndarray, mu=None, mut=None, dist=None, f1=None, grads=None) -> Dict[str, np.ndarray]:
    """The writable variables of one block as interior cell arrays
    (kmx-1, jmx-1, imx-1), named as in output_control.md, from the padded
    device/oracle arrays (cells -2..n+2; gradients 0..n)."""
    inner = (slice(3, -3),) * 3
    out = {"u": qp[1][inner], "v": qp[2][inner], "w": qp[3][inner],
           "Density": qp[0][inner], "Pressure": qp[4][inner]}
    if case.sst_family:
        out["TKE"], out["Omega"] = qp[5][inner], qp[6][inner]
    if case.turb == "kkl":
        out["TKE"], out["Kl"] = qp[5][inner], qp[6][inner]
    if case.turb == "sa":  # nu-tilde, write_output_tec.f90:163-164
        out["tv"] = qp[5][inner]
    if case.transition == "lctm2015":  # write_output_tec.f90:166-167
        out["tgm"] = qp[7][inner]
    for name, a in (("Mu", mu), ("Mu_t", mut), ("Wall_distance", dist), ("F1", f1)):
        if a is not None:
            out[name] = a[inner]
    if grads is not None:
        gin = (slice(1, -1),) * 3
        ng = grads[0].shape[0]
        for name, (d, l) in GRAD_NAMES.items():
            if l < ng:
                out[name] = grads[d][l][gin]
    return out


def _columns(w_list: Sequence[str], fields: Dict[str, np.ndarray]):
    """(header names, arrays) in w_list order; 'Velocity' is u, v, w."""
    names, arrays = [], []
    for w in w_list:
        if w == "do not write":
            continue
        if w == "Velocity":
            names.append("Velocity")
            arrays += [fields["u"], fields["v"], fields["w"]]
        else:
            names.append(w)
            arrays.append(fields[w])
    return names, arrays


def _nodes3(nodes: np.ndarray, shape) -> np.ndarray:
    """(3, kmx, jmx, imx) from either that or the padded node array."""
    kmx, jmx, imx = shape
    if nodes.shape[1:] == (kmx, jmx, imx):
        return nodes
    return nodes[:, 3:3 + kmx, 3:3 + jmx, 3:3 + imx]


def _tec_header(w_list, dims, process_id, solution_time, nodal):
    imx, jmx, kmx = dims
    lines = ["variables = x y z "]
    total = 3
    for w in w_list:
        if w == "Velocity":
            lines.append(" u v w ")
            total += 3
        elif w != "do not write":
            lines.append(w.strip() + " ")
            total += 1
    lines.append("zone T=block%s  i=%s j=%s k=%s Datapacking=Block" % (
        f_i(process_id, 4, 4), f_i(imx, 5, 5), f_i(jmx, 5, 5), f_i(kmx - 1 if nodal else kmx, 5, 5)))
    lines.append(" Varlocation=([1-3]=Nodal)")
    lines.append("Varlocation=([4-%s]=%s)" % (f_i(total, 2, 2), "Nodal" if nodal else "CELLCENTERED"))
    lines.append("STRANDID=%s" % f_i(1, 4, 4))
    lines.append("SOLUTIONTIME=%s" % f_i(solution_time, 4, 4))
    return lines


def write_tecplot(path: str, nodes: np.ndarray, w_list: Sequence[str], fields: Dict[str, np.ndarray],
                  process_id: int = 0, solution_time: int = 0) -> None:
    """Block-packed cell-centred Tecplot dump (write_output_tec.f90:60-306):
    the header, x y z at the nodes, then each listed variable over the
    interior cells, one ES28.15E4 value per line, i fastest."""
    kmx1, jmx1, imx1 = next(iter(fields.values())).shape
    dims = (imx1 + 1, jmx1 + 1, kmx1 + 1)
    n3 = _nodes3(nodes, dims[::-1])
    _, arrays = _columns(w_list, fields)
    with open(path, "w") as f:
        f.write("\n".join(_tec_header(w_list, dims, process_id, solution_time, False)) + "\n")
        for a in list(n3) + arrays:
            f.write("".join(f_es(v, 28, 15, 4) + "\n" for v in np.asarray(a, dtype=np.float64).ravel()))


def write_tecplot_nodal(path: str, nodes: np.ndarray, w_list: Sequence[str], fields_padded: Dict[str, np.ndarray],
                        process_id: int = 0, solution_time: int = 0) -> None:
    """Nodal Tecplot dump (write_output_tec_node.f90:200-280): node layers
    k = 1..kmx-1 and, per cell layer k, the average of the four cells
    around each (i, j) node column, 35 E25.15 values a record.  The fields
    are padded cell arrays (-2..n+2)."""
    a0 = next(iter(fields_padded.values()))
    dims = (a0.shape[2] - 5, a0.shape[1] - 5, a0.shape[0] - 5)
    imx, jmx, kmx = dims
    n3 = _nodes3(nodes, dims[::-1])[:, :kmx - 1]
    _, arrays = _columns(w_list, fields_padded)

    def record(vals):
        out = []
        for s in range(0, len(vals), 35):
            out.append("".join(f_e(v, 25, 15) for v in vals[s:s + 35]) + "\n")
        return "".join(out)

    with open(path, "w") as f:
        f.write("\n".join(_tec_header(w_list, dims, process_id, solution_time, True)) + "\n")
        for a in n3:
            f.write(record(a.ravel()))
        for a in arrays:
            o = 2  # padded index of cell 0
            c = a[o + 1:o + kmx, :, :]
            avg = 0.25 * (((c[:, o + 1:o + jmx + 1, o + 1:o + imx + 1] + c[:, o + 1:o + jmx + 1, o:o + imx])
                           + c[:, o:o + jmx, o + 1:o + imx + 1]) + c[:, o:o + jmx, o:o + imx])
            f.write(record(avg.ravel()))


VTK_NAMES = {"Wall_distance": "dist", "Dissipation": "Dissipation"}


def write_vtk(path: str, nodes: np.ndarray, w_list: Sequence[str], fields: Dict[str, np.ndarray],
              data_format: str = "ASCII") -> None:
    """Legacy-VTK structured grid (write_output_vtk.f90:51-296): F0.16 node
    coordinates, then the cell data; the CELL_DATA line comes with the
    velocity vectors, as in the reference."""
    kmx1, jmx1, imx1 = next(iter(fields.values())).shape
    imx, jmx, kmx = imx1 + 1, jmx1 + 1, kmx1 + 1
    n3 = _nodes3(nodes, (kmx, jmx, imx))
    with open(path, "w") as f:
        f.write("# vtk DataFile Version 3.1\ncfd-iitm output\n%s\nDATASET STRUCTURED_GRID\n" % data_format.strip())
        f.write("DIMENSIONS %d %d %d\n" % (imx, jmx, kmx))
        f.write("POINTS %d DOUBLE\n" % (imx * jmx * kmx))
        X, Y, Z = (a.ravel() for a in n3)
        f.write("".join("%s %s %s\n" % (f_f0(x, 16), f_f0(y, 16), f_f0(z, 16)) for x, y, z in zip(X, Y, Z)))
        f.write("\n")
        for w in w_list:
            if w == "do not write":
                continue
            if w == "Velocity":
                f.write("CELL_DATA %d\n" % (imx1 * jmx1 * kmx1))
                f.write("VECTORS Velocity FLOAT\n")
                U, V, W = (fields[c].ravel() for c in "uvw")
                f.write("".join("%s %s %s\n" % (f_es(u, 27, 16, 4), f_es(v, 27, 16, 4), f_es(x, 27, 16, 4))
                                for u, v, x in zip(U, V, W)))
            else:
                name = VTK_NAMES.get(w, w[0].lower() + w[1:] if w in GRAD_NAMES else w)
                f.write("SCALARS %s FLOAT\nLOOKUP_TABLE default\n" % name.strip())
                f.write("".join(f_es(v, 25, 16, 4) + "\n" for v in fields[w].ravel()))
            f.write("\n")


# ------------------------------------------------------------ restart log
def write_restart_log(path: str, turbulence: str, iteration: int, previous_res: Iterable[float]) -> None:
    """time_directories/NNNN/restart/process_XX (dump_solution.f90:120-152):
    the model ('viscous' for laminar), the iteration count, and the n_var+1
    reference residual norms, F0.16 each."""
    with open(path, "w") as f:
        f.write(("viscous" if turbulence == "none" else turbulence) + "\n")
        f.write(f_i(iteration, 0) + "\n")
        f.write("".join(f_f0(v, 16) + "\n" for v in previous_res))


def dump_dir(root: str, checkpoint_count: int) -> str:
    """time_directories/NNNN (dump_solution.f90:94-101)."""
    return os.path.join(root, "time_directories", f_i(checkpoint_count, 4, 4))


def checkpoint(root: str, checkpoint_count: int, purge_write: int, write_blocks) -> str:
    """One dump (dump_solution.f90:19-92): make NNNN and NNNN/restart, call
    write_blocks(directory) to write each block's files, then remove the
    directory purge_write dumps back when purge_write is set."""
    d = dump_dir(root, checkpoint_count)
    os.makedirs(os.path.join(d, "restart"), exist_ok=True)
    write_blocks(d)
    purge = checkpoint_count - purge_write
    if purge_write != 0 and purge > 0:
        shutil.rmtree(dump_dir(root, purge), ignore_errors=True)
    return d


def is_checkpoint_iteration(current_iter: int, checkpoint_iter: int, max_iters: int) -> bool:
    """dump_solution.f90:32-34."""
    return checkpoint_iter != 0 and (current_iter % checkpoint_iter == 0 or current_iter == max_iters)


# ------------------------------------------------------------ residual norms
def _norm(v) -> float:
    """sqrt(sum(v**2)) summed left to right, as the reference's SUM."""
    t = 0.0
    for x in v:
        t += float(x) * float(x)
    return math.sqrt(t)


RES_TURB = {"TKE": ("sst", "kkl", "sst2003"), "Tv": ("sa", "saBC"), "Dissipation": ("ke",),
            "Omega": ("sst", "sst2003"), "Kl": ("kkl",)}


class ResnormFile:
    """The resnorm file (src/resnorm.f90): absolute norms scaled by the free
    stream, summed over blocks, relative norms against those of the first
    Res_itr iterations (or of the restart log), written every
    res_write_interval iterations, at Res_itr and at iteration 1.

    update() takes the assembled Res_abs(0:n_var): a block's resnorm() for
    one block, the group's resnorm() for several.  assemble() restates that
    assembly from each block's (signed mass imbalance, sum of squared
    residues per variable)."""

    def __init__(self, path: str, case: Case, res_list: Sequence[str], write_interval: int = 10,
                 res_itr: int = 3, precision: int = 6, last_iter: int = 0,
                 previous_res: Optional[Sequence[float]] = None, append: bool = False):
        self.case, self.res_list = case, list(res_list)
        self.interval, self.res_itr, self.n = write_interval, res_itr, precision
        self.last_iter = last_iter
        nv = case.n_var
        sc = np.zeros(nv + 1)
        rho, vm, gm = case.rho_inf, case.vel_mag, case.gm
        sc[0] = 1.0
        sc[1] = rho * vm
        sc[2] = sc[3] = sc[4] = rho * vm * vm
        sc[5] = (0.5 * rho * vm ** 3 + ((gm / (gm - 1.0)) * case.p_inf))
        if case.two_equation:
            # src/resnorm.f90:147-153
            sc[6] = rho * vm * case.tk_inf
            sc[7] = rho * vm * (case.tkl_inf if case.turb == "kkl" else case.tw_inf)
        self.scale = sc
        self.restarted = previous_res is not None
        self.save = np.array(previous_res, dtype=np.float64) if previous_res is not None else np.zeros(nv + 1)
        self.abs = np.zeros(nv + 1)
        self.rel = np.zeros(nv + 1)
        self.f = open(path, "a" if append else "w")
        self.f.write("Iteration  " + "".join(r.strip() + "  " for r in self.res_list) + "\n")

    def close(self):
        self.f.close()

    def assemble(self, block_norms) -> np.ndarray:
        """Res_abs from [(merror, [sum residue_l^2 for l = 1..n_var])] per
        block (src/resnorm.f90:171-225)."""
        tot = np.zeros(self.case.n_var + 1)
        for merror, sq in block_norms:
            ra = np.empty(self.case.n_var + 1)
            ra[1:] = np.asarray(sq, dtype=np.float64) / self.scale[1:] ** 2
            ra[0] = merror / self.scale[0]
            tot += ra
        tot[1:] = np.sqrt(tot[1:])
        tot[0] = abs(tot[0])
        return tot

    def update(self, current_iter: int, res_abs) -> np.ndarray:
        """find_resnorm after the assembly, for one iteration: the relative
        norms (returned) and the file line."""
        self.abs = np.asarray(res_abs, dtype=np.float64).copy()
        if current_iter <= self.res_itr and not self.restarted:
            self.save = self.abs.copy()
        with np.errstate(divide="ignore", invalid="ignore"):
            self.rel = self.abs / self.save
        if current_iter % self.interval == 0 or current_iter == self.res_itr or current_iter == 1:
            self.f.write(self.line(current_iter))
        return self.rel

    def _value(self, name: str):
        a, r, t = self.abs, self.rel, self.case.turb
        table = {"Mass_abs": a[0], "Resnorm_abs": _norm(a[1:]),
                 "Viscous_abs": _norm(a[1:6]), "Continuity_abs": a[1], "X_mom_abs": a[2],
                 "Y_mom_abs": a[3], "Z_mom_abs": a[4], "Energy_abs": a[5],
                 "Mass_rel": r[0], "Resnorm_rel": _norm(r[1:]),
                 "Viscous_rel": _norm(r[1:6]), "Continuity_rel": r[1], "X-mom_rel": r[2],
                 "Y-mom_rel": r[3], "Z-mom_rel": r[4], "Energy_rel": r[5]}
        if name in table:
            return table[name]
        if name in ("Turbulent_abs", "Turbulent_rel"):
            if t == "none":
                return None
            v = a if name.endswith("abs") else r
            return _norm(v[6:])
        for key, models in RES_TURB.items():
            for kind, v in (("_abs", a), ("_rel", r)):
                if name == key + kind:
                    if t not in models:
                        return None
                    return v[7] if key in ("Dissipation", "Omega", "Kl") else v[6]
        return _norm(a[1:])  # unknown names: the absolute norm, with a warning there

    def line(self, current_iter: int) -> str:
        fmt = lambda v: f_e(v, self.n + 8, self.n, 2) + "    "
        out = f_i(current_iter + self.last_iter, 0) + "    "
        for name in self.res_list:
            v = self._value(name.strip())
            if v is not None:
                out += fmt(v)
        return out + "\n"
