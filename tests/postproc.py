"""The measures the reference's own case post-processing computes from a
solution dump, restated for the case tests: the smooth-bump entropy change
(tests/SmoothBump/pp/entropy.py:5-30 of the reference) and the flat-plate
drag coefficient (tests/Lfp/pp/Surface.py:85-160).  Test-side code: the
engine does not compute them.
"""
import math

import numpy as np


def entropy_error(states, volumes, case) -> float:
    """Relative change of entropy over the grid, the measure of the smooth
    bump test (tests/SmoothBump/pp/entropy.py:5-30):
    sqrt(sum_blocks |(s - s_inf) V / s_inf|^2 / sum V), s = p / rho^gamma."""
    s_inf = case.p_inf / case.rho_inf ** case.gm
    err, vol = [], []
    for q, v in zip(states, volumes):
        rho, p = q[0, 3:-3, 3:-3, 3:-3], q[4, 3:-3, 3:-3, 3:-3]
        s = p / np.power(rho, case.gm)
        err.append(np.sqrt(np.sum(np.power((s - s_inf) * v / s_inf, 2))))
        vol.append(np.sum(v))
    return float(np.sqrt(sum(np.power(np.array(err), 2)) / sum(vol)))


def drag_coefficient(cases, nodes, qps, grads, mus, length: float) -> float:
    """Drag coefficient of the no-slip j faces, as the flat-plate tests
    compute it from a solution dump (tests/Lfp/pp/Surface.py:85-160): the
    viscous stress of the first cell row, with the molecular viscosity and
    the cell gradients, projected on the face's unit normal, plus the
    pressure force, over 0.5 rho |V|^2 and the plate length.

    nodes: padded node arrays; qps, mus: padded cell arrays; grads: the
    gradient arrays (3 directions, n_grad, 0:kmx, 0:jmx, 0:imx), variables
    u, v, w first."""
    c0 = cases[0]
    vel = math.sqrt(c0.u_inf ** 2 + c0.v_inf ** 2 + c0.w_inf ** 2)
    dyn = 0.5 * c0.rho_inf * vel * vel
    cd = []
    for case, nd, qp, g, mu in zip(cases, nodes, qps, grads, mus):
        imx, jmx, kmx = case.imx, case.jmx, case.kmx
        X = nd[:, 3:kmx + 3, 3:jmx + 3, 3:imx + 3]  # (3, k, j, i)
        for face, bc in enumerate(case.bc_id):
            if bc != -5:
                continue
            if face not in (2, 3):
                raise ValueError("the flat-plate measure covers no-slip j faces only")
            jf = 0 if face == 2 else jmx - 1        # face node row (0-based)
            jc = 1 if face == 2 else jmx - 1        # first cell row (cell index)
            # j-face area vectors from the nodes, tests/geometry.f90:44-73
            d1 = X[:, 1:, jf, 1:] - X[:, :-1, jf, :-1]
            d2 = X[:, :-1, jf, 1:] - X[:, 1:, jf, :-1]
            yn = 0.5 * np.stack([d1[1] * d2[2] - d1[2] * d2[1], d1[2] * d2[0] - d1[0] * d2[2],
                                 d1[0] * d2[1] - d1[1] * d2[0]])
            area = np.linalg.norm(yn, axis=0)
            n = yn / area
            if face == 3:
                n = -n
            nx, ny, nz = n
            cells = (slice(1, kmx), jc, slice(1, imx))
            dudx, dudy, dudz = (g[d][0][cells] for d in range(3))
            dvdx, dvdy, dvdz = (g[d][1][cells] for d in range(3))
            dwdx, dwdy, dwdz = (g[d][2][cells] for d in range(3))
            pc = (slice(3, kmx + 2), jc + 2, slice(3, imx + 2))
            m, p = mu[pc], qp[4][pc]
            delv = dudx + dvdy + dwdz
            txx = m * ((dudx + dudx) - 2.0 * delv / 3.0)
            tyy = m * ((dvdy + dvdy) - 2.0 * delv / 3.0)
            tzz = m * ((dwdz + dwdz) - 2.0 * delv / 3.0)
            txy = m * (dudy + dvdx)
            tyz = m * (dwdy + dvdz)
            txz = m * (dudz + dwdx)
            Fx = txx * nx + txy * ny + txz * nz
            cp = (p - c0.p_inf) / dyn
            cd.append(np.sum(Fx * area / dyn))
            cd.append(np.sum(-cp * nx * area))
    return float(sum(cd) / length)
