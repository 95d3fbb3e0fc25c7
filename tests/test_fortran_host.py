"""The Fortran side of the boundary: the ISO_C_BINDING module
(fest-3d_amd/fortran/fest_gpu.f90), the drop-in `update` module
(fest-3d_amd/fortran/update_gpu.f90) and a Fortran host program
(fest_gpu_drive) that runs RK steps through them.

CPU: the host program is built and linked against libfest_hip.so; the
drop-in module compiles against the reference's own `vartypes` and
`wall_dist` modules (the amdflang build in oracle/_ref) and exports the
reference's public procedures.  GPU: the Fortran host's results equal the
Python host's bit for bit (same library, same inputs) and the oracle's
within 1e-12."""
import os
import shutil
import subprocess

import numpy as np
import pytest

from conftest import ROOT
from binding import Oracle, ref_available
from festgpu import cases as C
from festgpu import fortran_host as FH

PKG = os.path.join(ROOT, "fest-3d_amd")
FLANG = "/opt/rocm/bin/amdflang"


def test_driver_is_built_and_linked():
    assert os.path.exists(FH.DRIVER), "run make -C fest-3d_amd"
    p = subprocess.run([FH.DRIVER], capture_output=True, text=True, timeout=60)
    assert p.returncode == 2 and "usage" in p.stdout  # argument check, no device call
    out = subprocess.run(["ldd", FH.DRIVER], capture_output=True, text=True, timeout=60).stdout
    assert "libfest_hip.so" in out


def test_input_stream_layout(tmp_path):
    case, nodes, q = C.laminar_box(6, 5, 4)
    fin = tmp_path / "in.bin"
    FH.write_input(str(fin), case, nodes, q, 3, False)
    n = os.path.getsize(fin)
    assert n == 30 * 4 + (22 + 72) * 8 + 8 * (nodes.size + q.size)
    hdr = np.fromfile(fin, dtype=np.int32, count=30)
    assert list(hdr[:4]) == [case.imx, case.jmx, case.kmx, 5] and hdr[23] == 3


@pytest.mark.ref
@pytest.mark.skipif(not (ref_available() and os.path.exists(FLANG)), reason="reference build or amdflang absent")
def test_update_module_compiles_against_reference_types(tmp_path):
    ref = os.path.join(ROOT, "oracle", "_ref")
    for src, obj in (("fest_gpu.f90", "fest_gpu.o"), ("update_gpu.f90", "update.o"),
                     ("resnorm_gpu.f90", "resnorm.o")):
        p = subprocess.run([FLANG, "-cpp", "-fdefault-real-8", "-O2", "-fPIC", "-I", ref, "-I", str(tmp_path),
                            "-module-dir", str(tmp_path), "-c", os.path.join(PKG, "fortran", src),
                            "-o", str(tmp_path / obj)], capture_output=True, text=True, timeout=300)
        assert p.returncode == 0, p.stderr
    syms = subprocess.run(["nm", str(tmp_path / "update.o")], capture_output=True, text=True).stdout
    for name in ("setup_update", "get_next_solution", "gpu_state_to_host"):
        assert f"_QMupdateP{name}" in syms, name
    syms = subprocess.run(["nm", str(tmp_path / "resnorm.o")], capture_output=True, text=True).stdout
    for name in ("setup_resnorm", "find_resnorm"):
        assert f"_QMresnormP{name}" in syms, name


def _compare(case, nodes, q, nsteps, wall):
    from festgpu.native import GpuBlock
    fort = FH.run(case, nodes, q, nsteps, wall_dist=wall)
    g = GpuBlock(case)
    g.set_grid(nodes, wall_dist=wall)
    g.upload_state(q)
    for it in range(1, nsteps + 1):
        g.get_next_solution(it)
    py = g.download("qp", "residue")
    norms = g.resnorm()
    assert np.array_equal(fort["qp"], py["qp"])
    assert np.array_equal(fort["residue"], py["residue"])
    assert np.array_equal(fort["resnorm"], norms)
    o = Oracle(case)
    o.set_nodes(nodes)
    if wall:
        o.wall_dist()
    o.set_state(q)
    for it in range(1, nsteps + 1):
        assert o.step(it) == 0
    ref = o.arr("qp")
    err = np.abs(fort["qp"] - ref).max() / np.abs(ref).max()
    assert err <= 1e-12, err


@pytest.mark.gpu
def test_fortran_host_laminar_rk4():
    case, _, q = C.laminar_box(16, 12, 6)
    print("fortran host laminar 16x12x6, 3 RK4 steps", flush=True)
    _compare(case, C.cartesian_nodes(case, warp=0.03), q, 3, False)


@pytest.mark.gpu
def test_fortran_host_sst_bump():
    case, nodes, q = C.sst_bump(24, 16, 8)
    print("fortran host sst bump 24x16x8, 2 RK4 steps", flush=True)
    _compare(case, nodes, q, 2, True)
