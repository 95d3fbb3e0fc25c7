"""Batched group steps (fest_group_set_option "batched"): every phase of a
stage is one launch over the tile work list of all blocks (k_multi.h).  The
tiles run the per-block kernels' own bodies, so a batched step equals the
per-block step bit for bit, whatever the work-group count (a persistent
launch walks the list), the tile order or the j/k face-tile shape (64 x 4
or 32 x 8); both follow the oracle's group
step (the reference's interface exchange, src/interface1.f90) to 1e-12.
The autotuner leaves the state as it found it."""
import dataclasses
import sys

import numpy as np
import pytest

from binding import Oracle, group_step, group_wall_dist
from festgpu import cases as C
from festgpu import multiblock as M
from festgpu.native import FestError, GpuBlock, GpuGroup

pytestmark = pytest.mark.gpu
TOL = 1e-12


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def gpu_group(blocks, links, precision=0, wall=False, **opts):
    gs = []
    for c, n, q in blocks:
        g = GpuBlock(c, precision=precision)
        g.set_grid(n)
        g.upload_state(q)
        gs.append(g)
    grp = GpuGroup(gs, links)
    if wall:
        grp.wall_dist()
    for k, v in opts.items():
        grp.set_option(k, v)
    return grp, gs


def close(grp, gs):
    grp.close()
    for g in gs:
        g.close()


def states(gs):
    return [g.download("qp") for g in gs]


@pytest.mark.parametrize("opts", [dict(batched=1), dict(batched=1, grid_cap=7), dict(batched=1, order=1),
                                  dict(batched=1, grid_cap=3, order=1), dict(batched=1, tile_i=32),
                                  dict(batched=1, tile_i=32, grid_cap=5, order=1)])
def test_batched_chain_equals_per_block(opts):
    blocks, links = C.chain_blocks(nblocks=6, lo=8, hi=16, seed=3)
    say("batched chain", opts)
    a, ga = gpu_group(blocks, links, **opts)
    b, gb = gpu_group(blocks, links)
    for it in range(1, 5):
        a.get_next_solution(it)
        b.get_next_solution(it)
        for x, y in zip(states(ga), states(gb)):
            assert np.array_equal(x, y), it
    assert np.array_equal(a.resnorm(), b.resnorm())
    # and the oracle's group step
    os_ = []
    for c, n, q in blocks:
        o = Oracle(c)
        o.set_nodes(n)
        o.set_state(q)
        os_.append(o)
    for it in range(1, 5):
        assert group_step(os_, links, it) == 0
    for x, o in zip(states(ga), os_):
        ref = o.arr("qp")[:, 3:-3, 3:-3, 3:-3]
        err = np.abs(x[:, 3:-3, 3:-3, 3:-3] - ref).max() / np.abs(ref).max()
        assert err <= TOL, err
    # the group norms assemble the blocks' shares (src/resnorm.f90:193-213):
    # for l >= 1, sqrt of the sum of the blocks' squared norms
    want = np.sqrt(sum(o.resnorm()[1:] ** 2 for o in os_))
    np.testing.assert_allclose(a.resnorm()[1:], want, rtol=1e-11)
    close(a, ga)
    close(b, gb)


@pytest.mark.parametrize("precision", [0, 1])
def test_batched_sst_split_equals_per_block(precision):
    """The SST bump cut along k into three blocks (a cross-block periodic
    pair at the ends), wall distance over the group, FP64 and FP32 faces."""
    case, nodes, q = C.sst_bump(24, 12, 12)
    blocks, links = M.split(case, nodes, q, 2, [5, 9])
    say("batched sst split, precision", precision)
    a, ga = gpu_group(blocks, links, precision=precision, wall=True, batched=1, tile_i=32 if precision else 64)
    b, gb = gpu_group(blocks, links, precision=precision, wall=True)
    for it in range(1, 4):
        a.get_next_solution(it)
        b.get_next_solution(it)
    for x, y in zip(states(ga), states(gb)):
        assert np.array_equal(x, y)
    close(a, ga)
    close(b, gb)


def test_autotune_keeps_the_state():
    blocks, links = C.chain_blocks(nblocks=5, lo=8, hi=14, seed=4)
    say("autotune")
    a, ga = gpu_group(blocks, links)
    a.get_next_solution(1)
    before = states(ga)
    chosen, ms = a.autotune(4)
    say("autotune times (ms/step)", [round(t, 4) for t in ms], "chosen", chosen)
    assert len(ms) == len(GpuGroup.AUTOTUNE_CANDIDATES) and 0 <= chosen < len(ms)
    assert all(t > 0 for t in ms)
    for x, y in zip(before, states(ga)):
        assert np.array_equal(x, y)
    # the chosen schedule steps on like the per-block one
    b, gb = gpu_group(blocks, links)
    b.get_next_solution(1)
    for it in (2, 3):
        a.get_next_solution(it)
        b.get_next_solution(it)
    for x, y in zip(states(ga), states(gb)):
        assert np.array_equal(x, y)
    close(a, ga)
    close(b, gb)


def test_batched_rejects_mixed_blocks():
    case, nodes, q = C.laminar_box(16, 8, 6)
    blocks, links = M.split_i(case, nodes, q, [9])
    c1 = dataclasses.replace(blocks[1][0], recon="weno")
    blocks[1] = (c1, blocks[1][1], blocks[1][2])
    grp, gs = gpu_group(blocks, links)
    with pytest.raises(FestError):
        grp.set_option("batched", 1)
    close(grp, gs)
