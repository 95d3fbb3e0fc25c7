"""Three of the reference's own test cases run as they ship
(tests/golden/<case>, copied by tests/golden/copy_cases.py), against the
results of the reference's report (tests/Report.txt):

  SmoothBump  subsonic inviscid flow over a smooth bump, AUSM + MUSCL, two
              48x48x1 blocks (a 2-D grid) joined by an interface, far
              field, subsonic outlet, slip walls; LU-SGS at CFL 1000 from
              the free stream for 3000 iterations.  Relative change in
              entropy (tests/SmoothBump/pp/entropy.py:5-30): 7.883E-07.
  Lfp         laminar flat plate, two blocks 20x40x1 and 36x40x1, no-slip
              adiabatic plate; LU-SGS at CFL 2000 from the free stream for
              5000 iterations.  Drag coefficient (tests/Lfp/pp/Surface.py:
              85-160): 1.329E-03.
  Tfp         SST flat plate, two blocks 6x48x1 and 28x48x1; LU-SGS at CFL
              3000 for 10000 iterations, restarted from the state of its
              time_directories/0002 (velocity, density, pressure; k and
              omega at the free stream).  Reported drag coefficient
              2.872E-03; the oracle and the HIP path give 2.869E-03, 0.12%
              below it (within the report's 2% tolerance; the drag is still
              rising at 10000 iterations).

The first two match the report to its four printed digits on the oracle
and on the GPU, which also follows the oracle's trajectory over the first
steps within 1e-12 and ends at its converged states.  Iteration counts are
the cases' own (system/control.md).
"""
import os
import sys

import numpy as np
import pytest

from binding import Oracle, group_exchange, group_step, group_wall_dist
from festgpu import case_io as IO
import postproc as PP

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
# The report's own tolerance is 2% (tests/Report.txt); the checks below hold
# every case to 0.2% of the reported value, and to the four digits this
# path reaches as a regression value.  Tfp's 0.12% gap to the report is
# open (its cause is not found).
REPORT_RTOL = 2e-3
CASES = {
    # case: (measure, reported value, value this path reaches, plate length)
    "SmoothBump": ("entropy", "7.883E-07", "7.883E-07", None),
    "Lfp": ("drag", "1.329E-03", "1.329E-03", 0.05),
    "Tfp": ("drag", "2.872E-03", "2.869E-03", 2 * 0.04),
}


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def load(name):
    root = os.path.join(GOLDEN, name)
    blocks, links, ctl = IO.read_case(root)
    return blocks, links, ctl, IO.initial_states(root, blocks, ctl)


def measure(name, blocks, qps, vols, grads=None, mus=None):
    kind, _, _, length = CASES[name]
    cases = [c for c, _ in blocks]
    if kind == "entropy":
        return PP.entropy_error(qps, vols, cases[0])
    return PP.drag_coefficient(cases, [n for _, n in blocks], qps, grads, mus, length)


def oracle_run(name, n=None):
    blocks, links, ctl, q0 = load(name)
    os_ = []
    for (c, nd), q in zip(blocks, q0):
        o = Oracle(c)
        o.set_nodes(nd)
        o.set_state(q)
        os_.append(o)
    if blocks[0][0].turb == "sst":
        assert group_wall_dist(os_) == 0
    n = n or ctl["max_iter"]
    for it in range(1, n + 1):
        assert group_step(os_, links, it) == 0, it
    return blocks, links, ctl, os_


def oracle_measure(name, blocks, links, os_, it):
    if CASES[name][0] == "drag":
        # gradients and viscosity of the final state (the solution dump's)
        for o in os_:
            o.temp_refresh()
        assert group_exchange(os_, links) == 0
        for o in os_:
            assert o.residual(it) == 0
    vols = [o.arr("vol")[3:-3, 3:-3, 3:-3].copy() for o in os_]
    return measure(name, blocks, [o.arr("qp") for o in os_], vols,
                   [(o.arr("gx"), o.arr("gy"), o.arr("gz")) for o in os_], [o.arr("mu") for o in os_])


def test_case_reader():
    blocks, links, ctl, q0 = load("SmoothBump")
    assert ctl == dict(CFL=1000.0, max_iter=3000, restart=0)
    assert [(c.imx, c.jmx, c.kmx) for c, _ in blocks] == [(49, 49, 2), (49, 49, 2)]
    assert [c.bc_id for c, _ in blocks] == [(-8, 1, -6, -6, -6, -6), (0, -4, -6, -6, -6, -6)]
    c = blocks[0][0]
    assert (c.flux, c.recon, c.turb, c.time_accuracy, c.time_local, c.lim) == \
        ("ausm", "muscl", "none", "implicit", 1, (0, 0, 0))
    assert (c.rho_inf, c.u_inf, c.p_inf, c.mu_ref, c.gm, c.R_gas) == (1.225, 170.14, 101325.0, 0.0, 1.4, 287.0)
    # no numbers in the boundary files: every fixed value is the free stream
    assert c.fix_p == (101325.0,) * 6 and c.fix_rho == (1.225,) * 6
    assert [tuple(L) for L in links] == [(0, 2, 1, 1, 1, 49, 1, 2, 0), (1, 1, 0, 2, 1, 49, 1, 2, 0)]
    n0, n1 = blocks[0][1], blocks[1][1]
    assert np.array_equal(n0[:, 3:-3, 3:-3, 3 + 48], n1[:, 3:-3, 3:-3, 3])  # a common face
    assert np.all(q0[0][0] == 1.225)
    # the restart: velocity, density, pressure of the dump; k, omega free stream
    blocks, links, ctl, q0 = load("Tfp")
    assert ctl["restart"] == 2 and blocks[0][0].turb == "sst" and blocks[0][0].mu_sutherland == 1
    inner = q0[1][:, 3:-3, 3:-3, 3:-3]
    assert inner[1].max() > 60.0 and inner[1].min() < 5.0  # the dump's boundary layer
    assert np.all(q0[1][5] == blocks[1][0].tk_inf) and np.all(q0[1][6] == blocks[1][0].tw_inf)
    assert np.all(q0[1][0, :, :, 0] == blocks[1][0].rho_inf)  # ghosts keep the free stream


def test_group_wall_distance_is_decomposition_invariant():
    """A block cut in two sees the walls of both parts: the cells keep the
    distance of the uncut block (the surface nodes of all blocks form one
    set, src/wall/wall.f90:55-111)."""
    from festgpu import cases as C
    from festgpu import multiblock as M
    case, nodes, q = C.sst_bump(20, 10, 4)
    whole = Oracle(case)
    whole.set_nodes(nodes)
    whole.wall_dist()
    blocks, links = M.split_i(case, nodes, q, [9])
    os_ = []
    for c, n, qb in blocks:
        o = Oracle(c)
        o.set_nodes(n)
        os_.append(o)
    assert group_wall_dist(os_) == 0
    d = whole.arr("dist")[3:-3, 3:-3, 3:-3]
    got = np.concatenate([os_[0].arr("dist")[3:-3, 3:-3, 3:-3], os_[1].arr("dist")[3:-3, 3:-3, 3:-3]], axis=-1)
    assert np.array_equal(got, d)
    # block 1 alone misses the wall of block 0: its own distances differ
    alone = Oracle(blocks[1][0])
    alone.set_nodes(blocks[1][1])
    alone.wall_dist()
    assert not np.array_equal(alone.arr("dist"), os_[1].arr("dist"))


@pytest.mark.parametrize("name", list(CASES))
def test_oracle_reproduces_reference_report(name):
    blocks, links, ctl, os_ = oracle_run(name)
    value = oracle_measure(name, blocks, links, os_, ctl["max_iter"] + 1)
    reported = float(CASES[name][1])
    assert abs(value - reported) / reported < REPORT_RTOL, value
    assert "{:.3E}".format(value) == CASES[name][2], value


@pytest.mark.gpu
@pytest.mark.parametrize("name", list(CASES))
def test_gpu_reproduces_reference_report(name):
    from festgpu.native import GpuBlock, GpuGroup
    blocks, links, ctl, q0 = load(name)
    say(name, ": first steps against the oracle")
    gs = []
    for (c, nd), q in zip(blocks, q0):
        g = GpuBlock(c)
        g.set_grid(nd)
        g.upload_state(q)
        gs.append(g)
    group = GpuGroup(gs, links)
    blocks_o, links_o, ctl_o, os_ = oracle_run(name, 10)
    if blocks[0][0].turb == "sst":
        group.wall_dist()
        assert np.array_equal(gs[0].download("dist"), os_[0].arr("dist"))
    for it in range(1, 11):
        group.get_next_solution(it)
    for g, o in zip(gs, os_):
        ref = o.arr("qp")[:, 3:-3, 3:-3, 3:-3]
        err = np.abs(g.download("qp")[:, 3:-3, 3:-3, 3:-3] - ref).max() / np.abs(ref).max()
        assert err <= 1e-12, err
    n = ctl["max_iter"]
    say(name, ": iterations 11 ..", n)
    for it in range(11, n + 1):
        group.get_next_solution(it)
    if CASES[name][0] == "drag":
        group.exchange()
        for g in gs:
            g.residual(n + 1)
    out = [g.download("qp", "grads", "mu") for g in gs]
    vols = [o.arr("vol")[3:-3, 3:-3, 3:-3].copy() for o in os_]
    value = measure(name, blocks, [x["qp"] for x in out], vols, [x["grads"] for x in out], [x["mu"] for x in out])
    say(name, "measure", value)
    reported = float(CASES[name][1])
    assert abs(value - reported) / reported < REPORT_RTOL, value
    assert "{:.3E}".format(value) == CASES[name][2], value
    group.close()


def gpu_run_to_end(name, precision):
    """The case's full iteration count on a GPU group; its measure."""
    from festgpu.native import GpuBlock, GpuGroup
    blocks, links, ctl, q0 = load(name)
    gs = []
    for (c, nd), q in zip(blocks, q0):
        g = GpuBlock(c, precision=precision)
        g.set_grid(nd)
        g.upload_state(q)
        gs.append(g)
    group = GpuGroup(gs, links)
    if blocks[0][0].turb == "sst":
        group.wall_dist()
    n = ctl["max_iter"]
    for it in range(1, n + 1):
        group.get_next_solution(it)
    group.exchange()
    for g in gs:
        g.residual(n + 1)
    out = [g.download("qp", "grads", "mu", "cells") for g in gs]
    vols = [x["cells"][0][3:-3, 3:-3, 3:-3].copy() for x in out]
    value = measure(name, blocks, [x["qp"] for x in out], vols, [x["grads"] for x in out], [x["mu"] for x in out])
    group.close()
    return value


@pytest.mark.gpu
@pytest.mark.parametrize("name", ["Lfp", "Tfp"])
def test_gpu_fp32_faces_reproduce_reference_report(name):
    """The flat plates run to the end with FP32 face arithmetic
    (FEST_PREC_FP32_FLUX) and in FP64: the converged drag coefficients
    agree to 1e-6 relative (measured: 3e-8 on Lfp, 7e-8 on Tfp), and the
    FP32 one is within the report's tolerance of the reported value."""
    say(name, ": FP64 and FP32 faces to the end")
    v64 = gpu_run_to_end(name, 0)
    v32 = gpu_run_to_end(name, 1)
    say(name, "drag FP64", v64, "FP32", v32, "relative difference", (v32 - v64) / v64)
    reported = float(CASES[name][1])
    assert abs(v32 - reported) / reported < REPORT_RTOL, v32
    assert abs(v32 - v64) / abs(v64) < 1e-6, (v32, v64)
