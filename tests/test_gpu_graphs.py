"""The time step replayed as a captured HIP graph (fest_set_option "graphs")
against the same step launched kernel by kernel, bit for bit: the graph
holds the same kernels with the same arguments.  Every replay is checked,
with eager work of another handle, downloads and synchronisations between
the replays, across the inflow-fix classes (iterations <= 2 and later) and
a change of the CFL number (a re-capture).  Also the deferred error check
("check_interval")."""
import dataclasses
import sys

import numpy as np
import pytest

from festgpu import cases as C
from festgpu import multiblock as M
from festgpu.native import FestError, GpuBlock, GpuGroup

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def _block(case, nodes, q, graphs, wall):
    g = GpuBlock(case)
    g.set_option("graphs", graphs)
    g.set_grid(nodes, wall_dist=wall)
    g.upload_state(q)
    return g


@pytest.mark.parametrize("rule", ["RK4", "TVDRK3", "implicit"])
def test_graph_step_equals_eager_step(rule):
    case, nodes, q = C.sst_bump(20, 12, 8)
    case = dataclasses.replace(case, time_accuracy=rule)
    say("graph vs eager", rule)
    a = _block(case, nodes, q, 1, True)
    b = _block(case, nodes, q, 0, True)
    for it in range(1, 8):
        cfl = 0.5 if it < 5 else 0.4
        b.get_next_solution(it, cfl)
        b.download("qp")
        a.get_next_solution(it, cfl)
        qa, qb = a.download("qp"), b.download("qp")
        assert np.array_equal(qa, qb), it
        ra, rb = a.resnorm(), b.resnorm()
        assert np.array_equal(ra, rb), it
    a.close()
    b.close()


def test_graph_group_step_equals_eager():
    case, nodes, q = C.sst_bump(24, 10, 6)
    blocks, links = M.split_i(case, nodes, q, [9, 17])
    say("group graph vs eager")
    runs = []
    for graphs in (1, 0):
        gs = []
        for c, n, qb in blocks:
            g = GpuBlock(c)
            g.set_grid(n)
            g.upload_state(qb)
            gs.append(g)
        grp = GpuGroup(gs, links)
        grp.set_option("graphs", graphs)
        grp.wall_dist()
        runs.append((grp, gs))
    for it in range(1, 6):
        for grp, _ in runs[::-1]:
            grp.get_next_solution(it)
        for ga, gb in zip(runs[0][1], runs[1][1]):
            assert np.array_equal(ga.download("qp"), gb.download("qp")), it
        assert np.array_equal(runs[0][0].resnorm(), runs[1][0].resnorm())
    for grp, gs in runs:
        grp.close()
        for g in gs:
            g.close()


def test_deferred_error_check():
    """With check_interval 3 a non-physical state does not stop the step
    call itself; the next fest_sync (or norm, or third step) reports it."""
    case, nodes, q = C.laminar_box(12, 10, 6)
    q = q.copy()
    q[0, 6, 6, 6] = -1.0
    g = GpuBlock(case)
    g.set_grid(nodes)
    g.upload_state(q)
    g.set_option("check_interval", 3)
    g.get_next_solution(1)  # no check yet
    with pytest.raises(FestError) as e:
        g.sync()
    assert e.value.code in (-1, -2)
    with pytest.raises(FestError):
        g.set_option("check_interval", 0)
    g.close()
