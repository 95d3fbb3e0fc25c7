"""Multiblock grids: the interface exchange (src/interface1.f90:96-705)
with the orientation maps (src/mapping.f90:64-258).

The reference's exchange needs MPI and does not build here, so the oracle's
restatement is pinned by domain decomposition: a uniform Cartesian block cut
along i into sub-blocks (interfaces on the cuts, a cross-block periodic pair
where the block was i-periodic) must reproduce the single block.  On an
integer grid the ghost-node extrapolation of each sub-block is exact, so
the sub-blocks see the same metrics, states and operation order as the
single block:
  - inviscid flow: bit for bit over several RK4 steps;
  - sub-blocks re-indexed with j and k reversed or swapped (the links then
    need reversed ranges and the direction switch): equal to rounding, with a
    fixed time step (the reference's local step takes the normal velocity of
    a cell's lower faces from the cell and of its upper faces from the
    neighbour, src/time.f90:180-240, so a mirrored block steps differently).
A wrong range, direction or switch moves ghost layers by whole cells and
fails these by orders of magnitude.  Viscous flow is not decomposition
invariant in the reference: the gradient of an interface ghost cell reads
the ghost cells of the block's edges, which the edge/corner averages of
src/boundary/bc_primitive.f90:209-224 fill at every block corner, and the
stale Temp of the interface ghosts is refreshed at the step start from the
values the last stage exchanged; viscous multiblock runs are checked GPU
against oracle only.  The GPU exchange is checked against the oracle's on
the same grids."""
import dataclasses

import numpy as np
import pytest

from binding import Oracle, group_exchange, group_step, group_wall_dist
from festgpu import cases as C
from festgpu import multiblock as M


def single_and_split(ni, nj, nk, cuts, viscous, kinds=None, seed=1, fixed_dt=None):
    case, _, q = C.laminar_box(ni, nj, nk, seed=seed)
    # far field on the outer i faces: a single-block periodic face (-9) takes
    # first-order ghost states while a cross-block periodic pair (-10) is
    # reconstructed like an interior face, so only non-periodic ends compare
    case = dataclasses.replace(case, time_local=1, mu_ref=(case.mu_ref if viscous else 0.0),
                               bc_id=(-8, -8) + tuple(case.bc_id[2:]))
    if fixed_dt:
        case = dataclasses.replace(case, time_local=0, global_time_step=fixed_dt)
    nodes = C.cartesian_nodes(case)  # unit spacing: integer coordinates
    blocks, links = M.split_i(case, nodes, q, cuts)
    kinds = kinds or {}
    for b, kind in kinds.items():
        blocks[b] = M.reindex(*blocks[b], kind)
    if kinds:
        links = M.reindex_links(links, [c for c, _, _ in blocks], kinds)
    return case, nodes, q, blocks, links


def oracles(blocks):
    out = []
    for c, n, q in blocks:
        o = Oracle(c)
        o.set_nodes(n)
        o.set_state(q)
        out.append(o)
    return out


def undo(kind, a):
    """Inverse re-indexing of a cell array (..., k, j, i)."""
    if kind is None:
        return a
    if kind == "jk_reverse":
        return a[..., ::-1, ::-1, :]
    # jk_swap is a quarter turn: three more make the identity
    for _ in range(3):
        a = np.swapaxes(a, -3, -2)[..., ::-1, :]
    return a


def assemble(case, blocks, cuts, arrays, kinds=None):
    """The interior cells of the sub-blocks, put back in the single block's
    index order: (n_var, k, j, i) over cells 1..imx-1 etc."""
    kinds = kinds or {}
    bounds = [1] + sorted(cuts) + [case.imx]
    parts = []
    for b, a in enumerate(arrays):
        a = undo(kinds.get(b), a)
        parts.append(a[:, 3:-3, 3:-3, 3:-3][..., : bounds[b + 1] - bounds[b]])
    return np.concatenate(parts, axis=-1)


@pytest.mark.parametrize("cuts", [[9], [7, 15]])
def test_exchange_fills_ghosts_from_neighbours(cuts):
    """After one exchange the ghost layers of every cut hold the neighbour's
    cells, i.e. exactly the single block's cells at those positions."""
    case, _, q = C.laminar_box(22, 6, 5)
    nodes = C.cartesian_nodes(case)
    blocks, links = M.split_i(case, nodes, q, cuts)  # i-periodic: -10 pair at the ends
    os_ = oracles(blocks)
    for o in os_:
        o.arr("qp")[...] = 0.0  # forget the consistent ghosts of the split
    for o, (_, _, qb) in zip(os_, blocks):
        o.arr("qp")[:, 3:-3, 3:-3, 3:-3] = qb[:, 3:-3, 3:-3, 3:-3]
    assert group_exchange(os_, links) == 0
    bounds = [1] + sorted(cuts) + [case.imx]
    for b, o in enumerate(os_):
        s = bounds[b]
        got = o.arr("qp")
        imx_b = bounds[b + 1] - s + 1
        for layer in (1, 2, 3):
            # imin ghost layer 1-l of block b = single cell s-l (wrapping periodically)
            ci = (s - layer - 1) % (case.imx - 1) + 1
            np.testing.assert_array_equal(got[:, 3:-3, 3:-3, 1 - layer + 2], q[:, 3:-3, 3:-3, ci + 2])
            ci = (s + imx_b - 1 + layer - 1 - 1) % (case.imx - 1) + 1
            np.testing.assert_array_equal(got[:, 3:-3, 3:-3, imx_b + layer - 1 + 2], q[:, 3:-3, 3:-3, ci + 2])


@pytest.mark.parametrize("cuts", [[11], [6, 13, 17]])
def test_split_equals_single_block_inviscid(cuts):
    case, nodes, q, blocks, links = single_and_split(24, 8, 5, cuts, False)
    single = Oracle(case)
    single.set_nodes(nodes)
    single.set_state(q)
    os_ = oracles(blocks)
    for it in range(1, 4):
        assert single.step(it) == 0
        assert group_step(os_, links, it) == 0
    got = assemble(case, blocks, cuts, [o.arr("qp").copy() for o in os_])
    assert np.array_equal(got, single.arr("qp")[:, 3:-3, 3:-3, 3:-3])


@pytest.mark.parametrize("kind", ["jk_reverse", "jk_swap"])
def test_reoriented_block_matches_single_block(kind):
    cuts = [12]
    kinds = {1: kind}
    case, nodes, q, blocks, links = single_and_split(24, 8, 8, cuts, False, kinds, fixed_dt=2e-4)
    single = Oracle(case)
    single.set_nodes(nodes)
    single.set_state(q)
    os_ = oracles(blocks)
    for it in range(1, 3):
        assert single.step(it) == 0
        assert group_step(os_, links, it) == 0
    got = assemble(case, blocks, cuts, [o.arr("qp").copy() for o in os_], kinds)
    ref = single.arr("qp")[:, 3:-3, 3:-3, 3:-3]
    err = np.abs(got - ref).max() / np.abs(ref).max()
    assert err < 1e-13, err


def test_bad_map_is_rejected():
    case, nodes, q, blocks, links = single_and_split(20, 6, 5, [10], False)
    os_ = oracles(blocks)
    bad = [links[0]._replace(hi1=links[0].hi1 - 1)]
    assert group_exchange(os_, bad) != 0


# ------------------------------------------------------------------ GPU
def gpu_blocks(blocks):
    from festgpu.native import GpuBlock
    out = []
    for c, n, q in blocks:
        g = GpuBlock(c)
        g.set_grid(n)
        g.upload_state(q)
        out.append(g)
    return out


@pytest.mark.gpu
@pytest.mark.parametrize("kind", [None, "jk_reverse", "jk_swap"])
def test_gpu_exchange_matches_oracle(kind):
    """The device exchange moves every element where the oracle's does, bit
    for bit, including reversed and swapped maps (ghosts zeroed first)."""
    from festgpu.native import GpuGroup
    kinds = {1: kind} if kind else {}
    case, _, q = C.laminar_box(20, 8, 8)
    nodes = C.cartesian_nodes(case)
    blocks, links = M.split_i(case, nodes, q, [7, 13])
    for b, k_ in kinds.items():
        blocks[b] = M.reindex(*blocks[b], k_)
    if kinds:
        links = M.reindex_links(links, [c for c, _, _ in blocks], kinds)
    print("exchange", kind, flush=True)
    zeroed = []
    for c, n, qb in blocks:
        z = np.zeros_like(qb)
        z[:, 3:-3, 3:-3, 3:-3] = qb[:, 3:-3, 3:-3, 3:-3]
        zeroed.append((c, n, z))
    gs = gpu_blocks(zeroed)
    group = GpuGroup(gs, links)
    group.exchange()
    os_ = oracles(zeroed)
    assert group_exchange(os_, links) == 0
    for g, o in zip(gs, os_):
        assert np.array_equal(g.download("qp"), o.arr("qp"))
    group.close()


@pytest.mark.gpu
def test_gpu_group_steps_match_oracle():
    """Three RK4 steps of a viscous three-block grid, one block re-indexed
    (swapped j/k), on the device against the oracle's multiblock step."""
    from festgpu.native import GpuGroup
    kinds = {2: "jk_swap"}
    case, _, q = C.laminar_box(24, 8, 8)
    case = dataclasses.replace(case, bc_id=(-8, -8) + tuple(case.bc_id[2:]))
    nodes = C.cartesian_nodes(case, warp=0.02)
    blocks, links = M.split_i(case, nodes, q, [8, 15])
    for b, k_ in kinds.items():
        blocks[b] = M.reindex(*blocks[b], k_)
    links = M.reindex_links(links, [c for c, _, _ in blocks], kinds)
    print("group steps: 3 blocks, 3 RK4 steps", flush=True)
    gs = gpu_blocks(blocks)
    group = GpuGroup(gs, links)
    os_ = oracles(blocks)
    for it in range(1, 4):
        group.get_next_solution(it)
        assert group_step(os_, links, it) == 0
    for g, o in zip(gs, os_):
        ref = o.arr("qp")[:, 3:-3, 3:-3, 3:-3]
        got = g.download("qp")[:, 3:-3, 3:-3, 3:-3]
        err = np.abs(got - ref).max() / np.abs(ref).max()
        assert err <= 1e-12, err
    group.close()


@pytest.mark.gpu
def test_gpu_group_rejects_bad_map():
    from festgpu.native import FestError, GpuGroup
    case, _, q = C.laminar_box(20, 6, 5)
    blocks, links = M.split_i(case, C.cartesian_nodes(case), q, [10])
    gs = gpu_blocks(blocks)
    with pytest.raises(FestError):
        GpuGroup(gs, [links[0]._replace(lo2=2)])


def test_threaded_group_step_equals_serial():
    """orc_group_step_mt (the CPU baseline's one-thread-per-block run) gives
    the serial group step's state bit for bit: the blocks share nothing."""
    from binding import group_step_mt
    case, nodes, q = C.sst_bump(24, 10, 6)
    blocks, links = M.split_i(case, nodes, q, [7, 13, 19])
    runs = []
    for mt in (False, True):
        os_ = []
        for c, n, qb in blocks:
            o = Oracle(c)
            o.set_nodes(n)
            o.set_state(qb)
            os_.append(o)
        assert group_wall_dist(os_) == 0
        for it in range(1, 4):
            st = group_step_mt(os_, links, it, 3) if mt else group_step(os_, links, it)
            assert st == 0
        runs.append([o.arr("qp").copy() for o in os_])
    for a, b in zip(*runs):
        assert np.array_equal(a, b)


@pytest.mark.parametrize("axis,cuts", [(1, [5]), (2, [4, 7])])
def test_split_along_j_and_k_equals_single_block_inviscid(axis, cuts):
    """split() along j or k: an integer Cartesian block with far-field faces
    on the cut axis, inviscid, three RK4 steps, bit for bit."""
    case, _, q = C.laminar_box(8, 10, 10, seed=3)
    ids = list(case.bc_id)
    ids[2 * axis], ids[2 * axis + 1] = -8, -8
    case = dataclasses.replace(case, time_local=1, mu_ref=0.0, bc_id=tuple(ids))
    nodes = C.cartesian_nodes(case)
    blocks, links = M.split(case, nodes, q, axis, cuts)
    single = Oracle(case)
    single.set_nodes(nodes)
    single.set_state(q)
    os_ = oracles(blocks)
    for it in range(1, 4):
        assert single.step(it) == 0
        assert group_step(os_, links, it) == 0
    bounds = [1] + sorted(cuts) + [(case.imx, case.jmx, case.kmx)[axis]]
    parts = [o.arr("qp")[:, 3:-3, 3:-3, 3:-3] for o in os_]
    parts = [np.take(p, np.arange(bounds[b + 1] - bounds[b]), axis=3 - axis) for b, p in enumerate(parts)]
    got = np.concatenate(parts, axis=3 - axis)
    assert np.array_equal(got, single.arr("qp")[:, 3:-3, 3:-3, 3:-3])


def test_chain_blocks_join_matching_faces():
    """The seeded chain of uneven blocks: every link joins faces of equal
    extent whose nodes coincide (after the global warp), the links pass the
    oracle's map check, and an oracle group step runs."""
    blocks, links = C.chain_blocks(nblocks=7, lo=6, hi=12, seed=3)
    for L in links:
        cr, nr, _ = blocks[L.block]
        cs, ns, _ = blocks[L.from_block]
        d = (L.face - 1) // 2
        ax = 3 - d  # node array axis of the face normal
        idx_r = 3 if L.face % 2 == 1 else nr.shape[ax] - 4
        idx_s = 3 if L.from_face % 2 == 1 else ns.shape[ax] - 4
        fr = np.take(nr[:, 3:-3, 3:-3, 3:-3], idx_r - 3, axis=ax)
        fs = np.take(ns[:, 3:-3, 3:-3, 3:-3], idx_s - 3, axis=ax)
        assert fr.shape == fs.shape
        np.testing.assert_allclose(fr, fs, rtol=0, atol=1e-12)
    os_ = oracles(blocks)
    assert group_exchange(os_, links) == 0
    for it in (1, 2):
        assert group_step(os_, links, it) == 0
    full = C.chain_blocks.__defaults__
    assert full[0] == 64  # the benchmark's chain has 64 blocks
