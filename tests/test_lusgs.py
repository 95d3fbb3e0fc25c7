"""The 'implicit' integrator: one residual evaluation, the time step and the
LU-SGS update (src/update.f90:215-218, src/lusgs.f90:134-1196).

The oracle's LU-SGS is pinned bit for bit against the reference's own
lusgs.f90 (tests/test_oracle_vs_ref.py, and the lusgs_*.npz fixtures of
tests/test_golden.py where the reference is absent).  Here:
  - CPU: properties the mathematics fixes.  As the time step goes to zero
    the diagonal vol/dt dominates and the update tends to the explicit
    first-order update -R dt/vol ('none'); the relative gap shrinks in
    proportion to dt.  A uniform free stream stays uniform.
  - GPU: the hyperplane-ordered device sweeps against the oracle's loop
    order within 1e-12 (the residue feeding them differs from the oracle by
    FMA rounding only), on ragged extents, laminar, inviscid and SST, a
    multiblock group and the full 256x128x128 SST grid.
"""
import dataclasses
import sys

import numpy as np
import pytest

from binding import Oracle, group_step
from festgpu import cases as C
from festgpu import multiblock as M

TOL = 1e-12


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def relerr(a, b):
    a, b = np.asarray(a), np.asarray(b)
    assert np.isfinite(a).all() and np.isfinite(b).all()
    return np.abs(a - b).max() / np.abs(b).max()


def inner(a):
    return a[:, 3:-3, 3:-3, 3:-3]


def one_step(case, nodes, q, it=1):
    o = Oracle(case)
    o.set_nodes(nodes)
    o.set_state(q)
    assert o.step(it) == 0
    return o


# ------------------------------------------------------------------ CPU
def test_small_time_step_tends_to_explicit_update():
    case, _, q = C.laminar_box(12, 10, 6)
    nodes = C.cartesian_nodes(case, warp=0.04)
    gaps = []
    for dt in (1e-7, 1e-8):
        base = dataclasses.replace(case, time_local=0, global_time_step=dt)
        imp = one_step(dataclasses.replace(base, time_accuracy="implicit"), nodes, q)
        exp = one_step(dataclasses.replace(base, time_accuracy="none"), nodes, q)
        d_imp = inner(imp.arr("qp")) - inner(q)
        d_exp = inner(exp.arr("qp")) - inner(q)
        assert np.abs(d_exp).max() > 0
        gaps.append(np.abs(d_imp - d_exp).max() / np.abs(d_exp).max())
    # first order in dt: ten times smaller step, about ten times smaller gap
    assert gaps[0] < 1e-3, gaps
    assert 5.0 < gaps[0] / gaps[1] < 20.0, gaps


def test_uniform_free_stream_is_preserved():
    case, _, _ = C.laminar_box(10, 8, 6)
    case = dataclasses.replace(case, time_accuracy="implicit", bc_id=(-8,) * 6)
    q = C.freestream_state(case)
    o = one_step(case, C.cartesian_nodes(case, warp=0.05), q)
    err = relerr(inner(o.arr("qp")), inner(q))
    assert err < 1e-13, err


# ------------------------------------------------------------------ GPU
def gpu_pair(case, nodes, q, wall=False):
    from festgpu.native import GpuBlock
    o = Oracle(case)
    o.set_nodes(nodes)
    if wall:
        o.wall_dist()
    o.set_state(q)
    g = GpuBlock(case)
    g.set_grid(nodes, wall_dist=wall)
    g.upload_state(q)
    return g, o


@pytest.mark.gpu
@pytest.mark.parametrize("kind,ni,nj,nk", [
    ("laminar", 16, 12, 6),
    ("laminar", 9, 17, 5),      # ragged hyperplanes, j longest
    ("inviscid", 13, 7, 11),
    ("sutherland_walls", 12, 10, 8),
])
def test_gpu_lusgs_laminar(kind, ni, nj, nk):
    case, _, q = C.laminar_box(ni, nj, nk)
    case = dataclasses.replace(case, time_accuracy="implicit")
    if kind == "inviscid":
        case = dataclasses.replace(case, mu_ref=0.0, flux="slau", recon="weno")
    if kind == "sutherland_walls":
        case = dataclasses.replace(case, mu_sutherland=1, bc_id=(-3, -4, -5, -6, -8, -8),
                                   fix_rho=(case.rho_inf,) * 6, fix_u=(case.u_inf,) * 6, fix_p=(case.p_inf,) * 6,
                                   fix_twall=(0, 0, 320.0, 0, 0, 0))
    say("lusgs", kind, ni, nj, nk)
    g, o = gpu_pair(case, C.cartesian_nodes(case, warp=0.04), q)
    for it in (1, 2):
        assert o.step(it) == 0
        g.get_next_solution(it)
        out = g.download("qp", "residue", "dt")
        assert relerr(out["residue"], o.arr("res")) <= 1e-10
        assert relerr(out["dt"], o.arr("dt")) <= TOL
        assert relerr(inner(out["qp"]), inner(o.arr("qp"))) <= TOL, it


@pytest.mark.gpu
@pytest.mark.parametrize("ni,nj,nk", [(20, 12, 6), (11, 9, 14)])
def test_gpu_lusgs_sst(ni, nj, nk):
    case, nodes, q = C.sst_bump(ni, nj, nk)
    case = dataclasses.replace(case, time_accuracy="implicit")
    say("lusgs sst", ni, nj, nk)
    g, o = gpu_pair(case, nodes, q, wall=True)
    for it in (1, 2):
        assert o.step(it) == 0
        g.get_next_solution(it)
    out = g.download("qp")
    assert relerr(inner(out), inner(o.arr("qp"))) <= TOL


@pytest.mark.gpu
def test_gpu_lusgs_group():
    """A three-block viscous grid, one block re-indexed: each block's LU-SGS
    runs on its own after the exchange that opens the residual."""
    from festgpu.native import GpuBlock, GpuGroup
    kinds = {1: "jk_reverse"}
    case, _, q = C.laminar_box(24, 8, 8)
    case = dataclasses.replace(case, time_accuracy="implicit", bc_id=(-8, -8) + tuple(case.bc_id[2:]))
    nodes = C.cartesian_nodes(case, warp=0.02)
    blocks, links = M.split_i(case, nodes, q, [8, 15])
    for b, k_ in kinds.items():
        blocks[b] = M.reindex(*blocks[b], k_)
    links = M.reindex_links(links, [c for c, _, _ in blocks], kinds)
    say("lusgs group: 3 blocks, 2 steps")
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
    for it in (1, 2):
        group.get_next_solution(it)
        assert group_step(os_, links, it) == 0
    for g, o in zip(gs, os_):
        assert relerr(inner(g.download("qp")), inner(o.arr("qp"))) <= TOL
    group.close()


@pytest.mark.gpu
def test_gpu_lusgs_full_size_sst_bump():
    """One implicit step on the 256x128x128 SST bump grid (4.2M cells, 510
    hyperplanes each way) against the oracle; the wall distance comes from
    the device (its parity is pinned on smaller grids)."""
    from festgpu.native import GpuBlock
    case, nodes, q = C.sst_bump()
    case = dataclasses.replace(case, time_accuracy="implicit")
    say("lusgs sst_bump 256x128x128")
    g = GpuBlock(case)
    g.set_grid(nodes, wall_dist=True)
    g.upload_state(q)
    g.get_next_solution(1)
    out = g.download("qp", "dist")
    o = Oracle(case)
    o.set_nodes(nodes)
    o.arr("dist")[...] = out["dist"]
    o.set_state(q)
    assert o.step(1) == 0
    assert relerr(inner(out["qp"]), inner(o.arr("qp"))) <= TOL
