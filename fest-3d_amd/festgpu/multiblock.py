"""Multiblock grids for the interface exchange: cut a single block along i
into sub-blocks joined by interface links, and re-index a sub-block (j and
k reversed, or swapped) so that its links need the reference's
orientation maps (src/mapping.f90:64-258, src/interface1.f90:96-705).

A link (fields as in fest_link / orc_link) says that the three ghost layers
of face `face` (1..6 = imin..kmax) of block `block` receive the three
interior layers next to face `from_face` of block `from_block`.  The sender
packs them in its natural loop order (first tangential direction fastest:
j on i-faces, i on j- and k-faces); lo1/hi1 and lo2/hi2 are the node ranges
the receiver walks its first and second tangential direction through
(hi < lo walks backwards), dir_switch 1 makes the first direction the outer
loop."""
import dataclasses
from collections import namedtuple

import numpy as np

from .cases import Case

Link = namedtuple("Link", "block face from_block from_face lo1 hi1 lo2 hi2 dir_switch")


def _tangential_nodes(case: Case, face: int):
    if face <= 2:
        return case.jmx, case.kmx
    if face <= 4:
        return case.imx, case.kmx
    return case.imx, case.jmx


def aligned_link(blocks, b, face, b2, face2):
    """Link between faces whose tangential axes run the same way."""
    n1, n2 = _tangential_nodes(blocks[b], face)
    return Link(b, face, b2, face2, 1, n1, 1, n2, 0)


def split_i(case: Case, nodes: np.ndarray, q: np.ndarray, cuts):
    """Cut `case` at the i-nodes in `cuts` (1 < cut < imx); split() along i."""
    return split(case, nodes, q, 0, cuts)


def split(case: Case, nodes: np.ndarray, q: np.ndarray, axis: int, cuts):
    """Cut `case` at the nodes `cuts` (1 < cut < N) of direction `axis`
    (0 i, 1 j, 2 k).  Returns the blocks [(case, nodes, q)] and the links; a
    periodic original (-9 on both faces of the axis) becomes a cross-block
    periodic pair (-10, src/boundary/bc.f90:27-32).  Each sub-block's state
    includes three cells of its neighbours on each cut, i.e. consistent
    ghost layers."""
    dims = (case.imx, case.jmx, case.kmx)
    N = dims[axis]
    bounds = [1] + sorted(cuts) + [N]
    nb = len(bounds) - 1
    fmin, fmax = 2 * axis, 2 * axis + 1
    periodic = case.bc_id[fmin] == -9 and case.bc_id[fmax] == -9
    ax = 3 - axis  # array axis of direction `axis` in (component, k, j, i)
    blocks = []
    for b in range(nb):
        s, e = bounds[b], bounds[b + 1]  # node range of block b
        n = e - s + 1
        ids = list(case.bc_id)
        ids[fmin] = (b - 1) if b > 0 else (-10 if periodic else case.bc_id[fmin])
        ids[fmax] = (b + 1) if b < nb - 1 else (-10 if periodic else case.bc_id[fmax])
        ext = dict(zip(("imx", "jmx", "kmx"), dims))
        ext[("imx", "jmx", "kmx")[axis]] = n
        cb = dataclasses.replace(case, bc_id=tuple(ids), name=f"{case.name}_b{b}", **ext)
        # nodes: array index = node index + 2; cells: array index = cell index + 2
        nb_nodes = np.zeros(cb.node_shape())
        src_n = [slice(None)] * 4
        dst_n = [slice(None)] * 4
        src_n[ax] = slice(s + 2, e + 3)
        dst_n[ax] = slice(3, -3)
        nb_nodes[tuple(dst_n)] = nodes[tuple(src_n)]
        lo = s - 3 + 2  # cell index s-3 .. e+2
        src = np.clip(np.arange(lo, lo + n + 5), 0, q.shape[ax] - 1)
        qb = np.take(q, src, axis=ax)
        blocks.append((cb, nb_nodes, qb))
    cases = [c for c, _, _ in blocks]
    links = []
    for b in range(nb - 1):
        links.append(aligned_link(cases, b, fmax + 1, b + 1, fmin + 1))
        links.append(aligned_link(cases, b + 1, fmin + 1, b, fmax + 1))
    if periodic:
        links.append(aligned_link(cases, 0, fmin + 1, nb - 1, fmax + 1))
        links.append(aligned_link(cases, nb - 1, fmax + 1, 0, fmin + 1))
    return blocks, links


# face permutation under a re-indexing: new face -> old face
_FACES = {"jk_reverse": {1: 1, 2: 2, 3: 4, 4: 3, 5: 6, 6: 5},
          "jk_swap": {1: 1, 2: 2, 3: 6, 4: 5, 5: 3, 6: 4}}


def reindex(case: Case, nodes: np.ndarray, q: np.ndarray, kind: str):
    """Re-index a block without moving it: "jk_reverse" runs j and k
    backwards (a half turn about the i axis), "jk_swap" makes new j the old
    k reversed and new k the old j (a quarter turn; needs jmx == kmx).  Both
    keep the cells right-handed.  Physical coordinates and velocities are
    unchanged; the boundary faces follow their cells."""
    # a padded axis reversed end to end maps node n to N+1-n and cell c to
    # N-c, ghost layers onto ghost layers (array axes: component, k, j, i)
    if kind == "jk_reverse":
        nn = nodes[:, ::-1, ::-1, :].copy()
        nq = q[:, ::-1, ::-1, :].copy()
        nc = dataclasses.replace(case)
    elif kind == "jk_swap":
        assert case.jmx == case.kmx
        # new k = old j, new j = old k reversed
        nn = np.swapaxes(nodes, 1, 2)[:, :, ::-1, :].copy()
        nq = np.swapaxes(q, 1, 2)[:, :, ::-1, :].copy()
        nc = dataclasses.replace(case)
    else:
        raise ValueError(kind)
    perm = _FACES[kind]

    def per_face(v):
        return tuple(v[perm[f] - 1] for f in range(1, 7))

    nc = dataclasses.replace(nc, bc_id=per_face(case.bc_id), fix_rho=per_face(case.fix_rho),
                             fix_u=per_face(case.fix_u), fix_v=per_face(case.fix_v), fix_w=per_face(case.fix_w),
                             fix_p=per_face(case.fix_p), fix_tk=per_face(case.fix_tk), fix_tw=per_face(case.fix_tw),
                             fix_twall=per_face(case.fix_twall))
    return nc, nn, nq


# (j, k) of a block re-indexed by `kind` as a signed permutation of the
# original (j, k): new[a] = sign * old[col]
_TANG = {None: ((0, 1), (1, 1)), "jk_reverse": ((0, 1), (-1, -1)), "jk_swap": ((1, 0), (-1, 1))}


def _compose(rcv, snd):
    """receiver (j, k) as a signed permutation of the sender's (j, k)"""
    (rc, rs), (sc, ss) = _TANG[rcv], _TANG[snd]
    # old[sc[a]] = ss[a] * snd[a]  ->  snd[a] = ss[a] * old[sc[a]]
    inv_col = [0, 0]
    inv_sgn = [1, 1]
    for a_ in range(2):
        inv_col[sc[a_]] = a_
        inv_sgn[sc[a_]] = ss[a_]
    cols = tuple(inv_col[rc[a_]] for a_ in range(2))
    sgns = tuple(rs[a_] * inv_sgn[rc[a_]] for a_ in range(2))
    return cols, sgns


def reindex_links(links, cases, kinds):
    """Links of an i-cut grid after re-indexing blocks by `kinds` (block ->
    kind or None).  Only i-face links (tangential j, k) are handled."""
    out = []
    for L in links:
        assert L.face <= 2 and L.from_face <= 2
        cols, sgns = _compose(kinds.get(L.block), kinds.get(L.from_block))
        c = cases[L.block]
        ranges = [(1, c.jmx) if sgns[0] > 0 else (c.jmx, 1), (1, c.kmx) if sgns[1] > 0 else (c.kmx, 1)]
        # receiver direction a runs with sender direction cols[a]; the
        # sender's first direction is the inner loop
        switch = 0 if cols[0] == 0 else 1
        out.append(Link(L.block, L.face, L.from_block, L.from_face, ranges[0][0], ranges[0][1], ranges[1][0],
                        ranges[1][1], switch))
    return out
