"""Runs the Fortran host program fest-3d_amd/fest_gpu_drive (built from
fortran/fest_gpu_drive.f90 over the ISO_C_BINDING module fortran/fest_gpu.f90)
on a seeded case: writes its stream input, starts it as a child process and
reads back the state, residue and norms it wrote."""
import os
import subprocess
import tempfile

import numpy as np

from .cases import Case
from .native import c_structs

HERE = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
DRIVER = os.path.join(HERE, "fest_gpu_drive")


def write_input(path, case: Case, nodes: np.ndarray, qp: np.ndarray, nsteps: int, wall_dist: bool):
    e, s, f, b = c_structs(case)
    hdr = [e.imx, e.jmx, e.kmx, e.n_var, s.flux, s.recon, s.turb, s.time_local, s.time_accuracy, s.accur,
           *s.lim, *s.tlim, f.mu_sutherland, *b.id, nsteps, int(wall_dist), s.transition, *s.pbs,
           s.muscl_limiter]
    assert len(hdr) == 30
    flow = [f.density_inf, f.x_speed_inf, f.y_speed_inf, f.z_speed_inf, f.pressure_inf, f.tk_inf, f.tw_inf,
            f.vel_mag, f.MInf, f.gm, f.R_gas, f.mu_ref, f.T_ref, f.Sutherland_temp, f.Pr, f.tPr, f.tv_inf,
            f.tu_inf, f.tkl_inf, f.tgm_inf]
    fixed = [b.fixed_density, b.fixed_x_speed, b.fixed_y_speed, b.fixed_z_speed, b.fixed_pressure, b.fixed_tk,
             b.fixed_tw, b.fixed_wall_temperature, b.fixed_Tpressure, b.fixed_tv, b.fixed_tkl, b.fixed_tgm]
    with open(path, "wb") as fh:
        fh.write(np.asarray(hdr, dtype=np.int32).tobytes())
        fh.write(np.asarray([s.global_time_step, case.CFL] + flow, dtype=np.float64).tobytes())
        fh.write(np.asarray([list(v) for v in fixed], dtype=np.float64).tobytes())
        fh.write(np.ascontiguousarray(nodes, dtype=np.float64).tobytes())
        fh.write(np.ascontiguousarray(qp, dtype=np.float64).tobytes())


def read_output(path, case: Case):
    nv = case.n_var
    raw = np.fromfile(path, dtype=np.float64)
    nq = nv * int(np.prod(case.cell_shape()))
    nr = nv * (case.imx - 1) * (case.jmx - 1) * (case.kmx - 1)
    assert raw.size == nq + nr + nv + 1, (raw.size, nq, nr)
    qp = raw[:nq].reshape((nv,) + case.cell_shape())
    res = raw[nq:nq + nr].reshape((nv, case.kmx - 1, case.jmx - 1, case.imx - 1))
    return {"qp": qp, "residue": res, "resnorm": raw[nq + nr:]}


def run(case: Case, nodes, qp, nsteps: int, wall_dist: bool = False, timeout: float = 120.0):
    """Runs `nsteps` RK steps through the Fortran host; raises if it fails."""
    if not os.path.exists(DRIVER):
        raise FileNotFoundError(f"{DRIVER} is not built: run make -C fest-3d_amd")
    with tempfile.TemporaryDirectory() as d:
        fin, fout = os.path.join(d, "in.bin"), os.path.join(d, "out.bin")
        write_input(fin, case, nodes, qp, nsteps, wall_dist)
        p = subprocess.run([DRIVER, fin, fout], capture_output=True, text=True, timeout=timeout)
        if p.returncode != 0:
            raise RuntimeError(f"fest_gpu_drive failed ({p.returncode}): {p.stdout}{p.stderr}")
        return read_output(fout, case)
