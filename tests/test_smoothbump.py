"""The reference's SmoothBump case run as it ships (tests/golden/SmoothBump,
copied from the reference's tests/SmoothBump): subsonic inviscid flow over
a smooth bump, AUSM with MUSCL, two 48x48x1 blocks (a 2-D grid) joined by
an interface, far field, subsonic outlet and slip walls, integrated from
the free stream by LU-SGS at CFL 1000 for the case's 3000 iterations.

The reference's own report (tests/Report.txt) gives the result: a relative
change in entropy of 7.883E-07 (tests/SmoothBump/pp/entropy.py:5-30).  The
oracle must reproduce it on the CPU and the HIP path on the GPU, where the
first steps must also follow the oracle's trajectory within 1e-12 and the
converged states agree.
"""
import os
import sys

import numpy as np
import pytest

from binding import Oracle, group_step
from festgpu import case_io as IO
from festgpu import cases as C

ROOT = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "SmoothBump")
REPORTED = "7.883E-07"  # tests/Report.txt, "Calculated relative change in entropy"
ITERATIONS = 3000       # system/control.md, "Maximum Iterations"


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def oracle_blocks(blocks):
    out = []
    for c, n in blocks:
        o = Oracle(c)
        o.set_nodes(n)
        o.set_state(C.freestream_state(c))
        out.append(o)
    return out


def volumes(os_):
    return [o.arr("vol")[3:-3, 3:-3, 3:-3].copy() for o in os_]


def run_oracle(blocks, links, n):
    os_ = oracle_blocks(blocks)
    for it in range(1, n + 1):
        assert group_step(os_, links, it) == 0, it
    return os_


def test_case_reader():
    blocks, links, ctl = IO.read_case(ROOT)
    assert ctl == dict(CFL=1000.0, max_iter=ITERATIONS)
    assert [(c.imx, c.jmx, c.kmx) for c, _ in blocks] == [(49, 49, 2), (49, 49, 2)]
    assert [c.bc_id for c, _ in blocks] == [(-8, 1, -6, -6, -6, -6), (0, -4, -6, -6, -6, -6)]
    c = blocks[0][0]
    assert (c.flux, c.recon, c.turb, c.time_accuracy, c.time_local, c.lim) == \
        ("ausm", "muscl", "none", "implicit", 1, (0, 0, 0))
    assert (c.rho_inf, c.u_inf, c.p_inf, c.mu_ref, c.gm, c.R_gas) == (1.225, 170.14, 101325.0, 0.0, 1.4, 287.0)
    # no numbers in the boundary files: every fixed value is the free stream
    assert c.fix_p == (101325.0,) * 6 and c.fix_rho == (1.225,) * 6
    assert [tuple(L) for L in links] == [(0, 2, 1, 1, 1, 49, 1, 2, 0), (1, 1, 0, 2, 1, 49, 1, 2, 0)]
    # the blocks meet on a common face
    n0, n1 = blocks[0][1], blocks[1][1]
    assert np.array_equal(n0[:, 3:-3, 3:-3, 3 + 48], n1[:, 3:-3, 3:-3, 3])


def test_oracle_reproduces_reference_report():
    blocks, links, ctl = IO.read_case(ROOT)
    os_ = run_oracle(blocks, links, ITERATIONS)
    ds = IO.entropy_error([o.arr("qp") for o in os_], volumes(os_), blocks[0][0])
    assert "{:.3E}".format(ds) == REPORTED, ds
    # converged: the continuity residual fell by six orders from the start
    assert max(o.resnorm()[1] for o in os_) < 1e-8


@pytest.mark.gpu
def test_gpu_reproduces_reference_report():
    from festgpu.native import GpuBlock, GpuGroup
    blocks, links, ctl = IO.read_case(ROOT)
    say("smooth bump: 2 blocks 48x48x1, first steps against the oracle")
    gs = []
    for c, n in blocks:
        g = GpuBlock(c)
        g.set_grid(n)
        g.upload_state(C.freestream_state(c))
        gs.append(g)
    group = GpuGroup(gs, links)
    os_ = oracle_blocks(blocks)
    for it in range(1, 11):
        group.get_next_solution(it)
        assert group_step(os_, links, it) == 0
    for g, o in zip(gs, os_):
        ref = o.arr("qp")[:, 3:-3, 3:-3, 3:-3]
        err = np.abs(g.download("qp")[:, 3:-3, 3:-3, 3:-3] - ref).max() / np.abs(ref).max()
        assert err <= 1e-12, err
    say("smooth bump: iterations 11..3000")
    for it in range(11, ITERATIONS + 1):
        group.get_next_solution(it)
    qs = [g.download("qp") for g in gs]
    ds = IO.entropy_error(qs, volumes(os_), blocks[0][0])
    say("entropy change", ds)
    assert "{:.3E}".format(ds) == REPORTED, ds
    assert max(g.resnorm()[1] for g in gs) < 1e-8
    say("smooth bump: the oracle's 3000 iterations")
    for it in range(11, ITERATIONS + 1):
        assert group_step(os_, links, it) == 0
    for q, o in zip(qs, os_):
        ref = o.arr("qp")[:, 3:-3, 3:-3, 3:-3]
        err = np.abs(q[:, 3:-3, 3:-3, 3:-3] - ref).max() / np.abs(ref).max()
        assert err <= 1e-9, err
    group.close()
