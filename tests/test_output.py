"""Output in the reference's formats (festgpu.output).

Pinned by:
  - the Fortran runtime: every edit descriptor the reference's writers use
    (ES28.15E4, ES27.16E4, ES25.16E4, E25.15, E14.6E2, F0.16, I0, Iw.m) is
    formatted by amdflang from the same binary doubles -- zeros of both
    signs, denormals, rounding ties, exponents past the field, NaN and Inf,
    and seeded random values -- and the text must be identical;
  - the reference's shipped files: its Tecplot dumps (Tfp, 0002) are read
    and written back byte for byte, and its restart logs (Tfp and
    SmoothBump) line for line.  Those logs hold ten norms, where the
    current dump_solution.f90:149-152 writes n_var+1 (state.f90:66): they
    come from an older build, so the count is the writer's argument;
  - the reference's rules, restated in tests: the nodal average of four
    cells, the VTK layout, which iterations reach the resnorm file and the
    relative norms frozen after iteration 3, the purge of old dumps.
"""
import dataclasses
import os
import shutil
import subprocess

import numpy as np
import pytest

from festgpu import cases as C
from festgpu import case_io as IO
from festgpu import output as O

HERE = os.path.dirname(os.path.abspath(__file__))
GOLD = os.path.join(HERE, "golden")
FLANG = "/opt/rocm/bin/amdflang"

DESCRIPTORS = ["ES28.15E4", "ES27.16E4", "ES25.16E4", "E25.15", "E14.6E2", "F0.16"]


def py_format(desc, v):
    if desc.startswith("ES"):
        w, rest = desc[2:].split(".")
        d, e = rest.split("E")
        return O.f_es(v, int(w), int(d), int(e))
    if desc.startswith("E"):
        w, rest = desc[1:].split(".")
        if "E" in rest:
            d, e = rest.split("E")
            return O.f_e(v, int(w), int(d), int(e))
        return O.f_e(v, int(w), int(rest))
    return O.f_f0(v, int(desc.split(".")[1]))


def sample_values():
    rng = np.random.default_rng(7)
    v = [0.0, -0.0, 1.5, -0.33333, 1e-300, 1e300, 2.5e-310, -7.25e12, 0.5, 0.99999995, 9.9999999999999999,
         123456789.123456789, 1e99, 1e100, 9.995e99, -1e-100, 1.0 / 3.0, 2.0 / 3.0, 0.05, 0.25, 0.125,
         1.0000000000000002, 0.9999999999999999, 5e-324, 1.7976931348623157e308, 1e-5, 101325.0,
         float("nan"), float("inf"), float("-inf")]
    v += list(rng.standard_normal(40) * 10.0 ** rng.integers(-12, 12, 40))
    v += list(rng.random(20))
    return np.array(v, dtype=np.float64)


@pytest.mark.skipif(not os.path.exists(FLANG), reason="amdflang absent")
def test_edit_descriptors_match_the_fortran_runtime(tmp_path):
    v = sample_values()
    (tmp_path / "v.bin").write_bytes(v.tobytes())
    lines = "\n".join("    write(*, '(%s)') v(n)" % d for d in DESCRIPTORS)
    src = f"""program fmtcheck
  implicit none
  integer, parameter :: m = {v.size}
  real(8) :: v(m)
  integer :: n
  open(10, file='v.bin', access='stream', form='unformatted', status='old')
  read(10) v
  close(10)
  do n = 1, m
{lines}
    write(*, '(I0,a,I4.4,a,I5.5,a,I2.2)') n * 137 - 40, '|', n * 3, '|', n * 997, '|', n
  end do
end program
"""
    (tmp_path / "f.f90").write_text(src)
    r = subprocess.run([FLANG, "-O0", "f.f90", "-o", "f"], cwd=tmp_path, capture_output=True, text=True)
    assert r.returncode == 0, r.stderr
    out = subprocess.run(["./f"], cwd=tmp_path, capture_output=True, text=True).stdout.split("\n")
    k = 0
    for n, x in enumerate(v, start=1):
        for d in DESCRIPTORS:
            assert py_format(d, x) == out[k], (d, repr(x), py_format(d, x), out[k])
            k += 1
        ints = "%s|%s|%s|%s" % (O.f_i(n * 137 - 40, 0), O.f_i(n * 3, 4, 4), O.f_i(n * 997, 5, 5), O.f_i(n, 2, 2))
        assert ints == out[k], (ints, out[k])
        k += 1


def tfp_blocks():
    blocks, _, _ = IO.read_case(os.path.join(GOLD, "Tfp"))
    return blocks


@pytest.mark.parametrize("b", [0, 1])
def test_tecplot_dump_rewrites_byte_for_byte(tmp_path, b):
    """The reference's own dump, read and written back."""
    case = tfp_blocks()[b][0]
    src = os.path.join(GOLD, "Tfp", "time_directories", "0002", f"process_{b:02d}.dat")
    nodes, cells = IO.read_tecplot(src, case)
    w_list = ["Velocity", "Density", "Pressure", "Mu"]
    dst = tmp_path / "out.dat"
    O.write_tecplot(str(dst), nodes, w_list, cells, process_id=b, solution_time=2)
    assert dst.read_bytes() == open(src, "rb").read()


@pytest.mark.parametrize("case,turb,iteration", [("Tfp", "sst", 20000), ("SmoothBump", "none", 10000)])
@pytest.mark.parametrize("b", [0, 1])
def test_restart_log_lines(tmp_path, case, turb, iteration, b):
    d = "0002" if case == "Tfp" else "0010"
    src = os.path.join(GOLD, case, "time_directories", d, "restart", f"process_{b:02d}")
    ref = open(src, "rb").read()
    nres = ref.decode().count("\n") - 2
    dst = tmp_path / "log"
    O.write_restart_log(str(dst), turb, iteration, [0.0] * nres)
    assert dst.read_bytes() == ref


def padded_fields(case, q):
    return {"u": q[1], "v": q[2], "w": q[3], "Density": q[0], "Pressure": q[4]}


def test_nodal_tecplot_average_and_records(tmp_path):
    case, nodes, q = C.laminar_box(7, 5, 3)
    p = tmp_path / "n.dat"
    O.write_tecplot_nodal(str(p), nodes, ["Velocity", "Pressure"], padded_fields(case, q), process_id=3,
                          solution_time=12)
    text = p.read_text().split("\n")
    assert text[:3] == ["variables = x y z ", " u v w ", "Pressure "]
    assert text[3] == "zone T=block0003  i=00008 j=00006 k=00003 Datapacking=Block"
    assert text[5] == "Varlocation=([4-07]=Nodal)"
    assert text[7] == "SOLUTIONTIME=0012"
    body = [ln for ln in text[8:] if ln]
    assert all(len(ln) % 25 == 0 and len(ln) <= 35 * 25 for ln in body)
    imx, jmx, kmx = case.imx, case.jmx, case.kmx
    nn = imx * jmx * (kmx - 1)
    # each array starts a new record: 3 coordinates and 4 variables
    per = -(-nn // 35)
    assert len(body) == 7 * per
    vals = [[ln[s:s + 25] for ln in body[a * per:(a + 1) * per] for s in range(0, len(ln), 25)]
            for a in range(7)]
    # pressure at node (i, j) of cell layer k: 0.25 * the four cells around it
    P = q[4]
    for k, j, i in ((1, 2, 3), (kmx - 1, 1, 1), (2, jmx, imx)):
        c = P[k + 2]
        avg = 0.25 * (((c[j + 2, i + 2] + c[j + 2, i + 1]) + c[j + 1, i + 2]) + c[j + 1, i + 1])
        assert vals[6][((k - 1) * jmx + (j - 1)) * imx + (i - 1)] == O.f_e(avg, 25, 15)
    X = nodes[0, 3:3 + kmx - 1, 3:3 + jmx, 3:3 + imx].ravel()
    assert vals[0] == [O.f_e(x, 25, 15) for x in X]


def test_vtk_layout(tmp_path):
    case, nodes, q = C.laminar_box(5, 4, 3)
    f = O.block_fields(case, q, dist=np.ones_like(q[0]))
    p = tmp_path / "v.vtk"
    O.write_vtk(str(p), nodes, ["Velocity", "Density", "Wall_distance"], f)
    L = p.read_text().split("\n")
    assert L[:6] == ["# vtk DataFile Version 3.1", "cfd-iitm output", "ASCII", "DATASET STRUCTURED_GRID",
                     "DIMENSIONS 6 5 4", "POINTS 120 DOUBLE"]
    pts = np.array([[float(t) for t in ln.split()] for ln in L[6:126]])
    np.testing.assert_array_equal(pts[:, 0], nodes[0, 3:-3, 3:-3, 3:-3].ravel())
    assert L[126] == ""
    nc = 5 * 4 * 3
    assert L[127:129] == ["CELL_DATA %d" % nc, "VECTORS Velocity FLOAT"]
    uvw = np.array([[float(t) for t in ln.split()] for ln in L[129:129 + nc]])
    np.testing.assert_array_equal(uvw[:, 1], q[2][3:-3, 3:-3, 3:-3].ravel())  # 17 digits round-trip
    assert all(len(ln) == 27 * 3 + 2 for ln in L[129:129 + nc])
    s0 = 129 + nc + 1
    assert L[s0:s0 + 2] == ["SCALARS Density FLOAT", "LOOKUP_TABLE default"]
    assert L[s0 + 2 + nc + 1] == "SCALARS dist FLOAT"


def test_resnorm_file(tmp_path):
    case, _, _ = C.sst_bump(8, 6, 4)
    p = tmp_path / "resnorm"
    r = O.ResnormFile(str(p), case, ["Resnorm_abs", "Mass_abs", "TKE_rel", "Kl_abs", "Turbulent_abs"])
    rng = np.random.default_rng(3)
    rels = []
    for it in range(1, 13):
        blocks = [(rng.standard_normal(), rng.random(7) * 10.0 ** -it) for _ in range(2)]
        rels.append(r.update(it, r.assemble(blocks)))
    r.close()
    lines = p.read_text().split("\n")
    assert lines[0] == "Iteration  Resnorm_abs  Mass_abs  TKE_rel  Kl_abs  Turbulent_abs  "
    written = [int(ln.split()[0]) for ln in lines[1:] if ln]
    assert written == [1, 3, 10]          # iteration 1, Res_itr = 3, every 10
    # Kl_abs belongs to k-kL and is left out; four fields of 14 + 4 spaces
    assert len(lines[3]) == len("10    ") + 4 * 18
    # the relative norms use the iteration-3 absolutes from then on
    np.testing.assert_allclose(rels[11] * r.save, r.abs, rtol=1e-15)


def test_resnorm_absolute_assembles_blocks():
    case, _, _ = C.laminar_box(6, 5, 4)
    r = O.ResnormFile(os.devnull, case, ["Resnorm_abs"])
    a = r.assemble([(0.5, [1.0, 4.0, 9.0, 16.0, 25.0]), (-2.0, [3.0, 0.0, 0.0, 0.0, 0.0])])
    sc = r.scale
    assert a[0] == 1.5
    assert a[1] == pytest.approx(np.sqrt(1.0 / sc[1] ** 2 + 3.0 / sc[1] ** 2), rel=1e-15)
    assert a[5] == pytest.approx(5.0 / sc[5], rel=1e-15)
    r.close()


def test_checkpoint_directories_and_purge(tmp_path):
    root = str(tmp_path)
    made = []
    for count in range(1, 5):
        made.append(O.checkpoint(root, count, purge_write=2,
                                 write_blocks=lambda d: open(os.path.join(d, "process_00.dat"), "w").close()))
    left = sorted(os.listdir(os.path.join(root, "time_directories")))
    assert left == ["0003", "0004"]
    assert os.path.isdir(os.path.join(made[-1], "restart"))
    assert O.is_checkpoint_iteration(200, 100, 1000) and O.is_checkpoint_iteration(999, 100, 999)
    assert not O.is_checkpoint_iteration(150, 0, 150)


def oracle_pieces(o, case):
    """A block's (signed mass imbalance, sum R_l^2) from the oracle's
    residue and fluxes (src/resnorm.f90:171-191)."""
    res, F, G, H = o.arr("res"), o.arr("F"), o.arr("G"), o.arr("H")
    faces = (F[0, :, :, 0], F[0, :, :, -1], G[0, :, 0, :], G[0, :, -1, :], H[0, 0, :, :], H[0, -1, :, :])
    merror = faces[0].sum() - faces[1].sum() + faces[2].sum() - faces[3].sum() + faces[4].sum() - faces[5].sum()
    return merror, [float(np.sum(res[l] ** 2)) for l in range(case.n_var)], sum(np.abs(f).sum() for f in faces)


@pytest.mark.gpu
def test_group_resnorm_and_outputs_from_the_device(tmp_path):
    """Three blocks stepped on the GPU: the group's Res_abs against the
    oracle blocks' pieces assembled in block order, the resnorm file fed
    from it, and a Tecplot dump of a device block read back."""
    from binding import Oracle, group_step
    from festgpu import multiblock as M
    from festgpu.native import GpuBlock, GpuGroup
    case, _, q = C.laminar_box(24, 8, 8)
    case = dataclasses.replace(case, bc_id=(-8, -8) + tuple(case.bc_id[2:]))
    nodes = C.cartesian_nodes(case, warp=0.02)
    blocks, links = M.split_i(case, nodes, q, [8, 15])
    gs, os_ = [], []
    for c, n, qb in blocks:
        g = GpuBlock(c)
        g.set_grid(n)
        g.upload_state(qb)
        gs.append(g)
        o = Oracle(c)
        o.set_nodes(n)
        o.set_state(qb)
        os_.append(o)
    group = GpuGroup(gs, links)
    rf = O.ResnormFile(str(tmp_path / "resnorm"), case, ["Resnorm_abs", "Mass_abs", "Continuity_rel"])
    for it in (1, 2, 3):
        group.get_next_solution(it)
        assert group_step(os_, links, it) == 0
        ra = group.resnorm()
        pieces = [oracle_pieces(o, c) for o, (c, _, _) in zip(os_, blocks)]
        want = rf.assemble([p[:2] for p in pieces])
        np.testing.assert_allclose(ra[1:], want[1:], rtol=1e-10)
        # the mass imbalance cancels boundary fluxes: bound it by their size
        assert abs(ra[0] - want[0]) <= 1e-12 * sum(p[2] for p in pieces)
        rf.update(it, ra)
    rf.close()
    assert len(open(tmp_path / "resnorm").read().split("\n")) == 4  # header, iterations 1 and 3, end
    c0 = blocks[0][0]
    out = gs[0].download("qp", "mu")
    f = O.block_fields(c0, out["qp"], mu=out["mu"])
    p = tmp_path / "process_00.dat"
    O.write_tecplot(str(p), blocks[0][1], ["Velocity", "Density", "Pressure", "Mu"], f)
    _, cells = IO.read_tecplot(str(p), c0)
    np.testing.assert_allclose(cells["Pressure"], out["qp"][4][3:-3, 3:-3, 3:-3], rtol=1e-15)
    group.close()


@pytest.mark.parametrize("turb,names", [("sa", ["tv"]), ("kkl", ["TKE", "Kl"]), ("sst", ["TKE", "Omega"])])
def test_restart_round_trip_turbulence_variables(tmp_path, turb, names):
    """A dump written with the model's turbulence variables restarts them
    (read_output_tec.f90:100-136: 'TKE', 'Omega', 'Kl', 'tv'), not the free
    stream.  Values on a 2^-10 lattice survive the 16 digits of ES28.15E4
    exactly."""
    case, nodes, q = C.sst_bump(6, 5, 4, turb=turb)
    q = np.round(q * 1024.0) / 1024.0
    fields = O.block_fields(case, q)
    w_list = ["Velocity", "Density", "Pressure"] + names
    p = tmp_path / "process_00.dat"
    O.write_tecplot(str(p), nodes, w_list, fields)
    got = IO.restart_state(case, str(p), w_list)
    inner = (slice(None), slice(3, -3), slice(3, -3), slice(3, -3))
    assert np.array_equal(got[inner], q[inner])
    # variables left out of the list keep the free stream
    got2 = IO.restart_state(case, str(p), w_list[:3])
    fs = C.freestream_state(case)
    assert np.array_equal(got2[5:], fs[5:])
