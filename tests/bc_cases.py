"""Cases for the total-pressure inlet (-11) and the pole (-7), shared by the
oracle-vs-reference and the GPU parity tests."""
import dataclasses

from festgpu import cases as C


def total_pressure_case(face, turb, ni=14, nj=10, nk=6):
    """The SST bump with one face turned into a total-pressure inlet.  The
    total pressure is 0.2% above the free stream's isentropic one: inflow at
    imin, outflow at imax, and at jmax a normal velocity that is noise about
    zero, so the inflow/outflow branch changes along the face."""
    case, nodes, q = C.sst_bump(ni, nj, nk)
    ids = list(case.bc_id)
    ids[face] = -11
    m2 = (case.u_inf ** 2 + case.v_inf ** 2 + case.w_inf ** 2) / (case.gm * case.p_inf / case.rho_inf)
    pt = 1.002 * case.p_inf * (1 + 0.5 * (case.gm - 1) * m2) ** (case.gm / (case.gm - 1))
    fix_tp = [0.0] * 6
    fix_tp[face] = pt
    case = dataclasses.replace(case, bc_id=tuple(ids), fix_tp=tuple(fix_tp), turb=turb)
    if turb == "none":
        q = q[:5]
    return case, nodes, q


def pole_case(bcs, recon, ni=10, nj=8, nk=6):
    """A warped laminar box with pole faces.  A face that truly collapses to
    an axis gives negative ghost volumes, which the reference's
    compute_volumes rejects, so this grid keeps its ghost cells positive."""
    case, _, q = C.laminar_box(ni, nj, nk, recon=recon)
    nodes = C.cartesian_nodes(case, warp=0.04)
    case = dataclasses.replace(case, bc_id=bcs, mu_sutherland=1, fix_rho=(case.rho_inf,) * 6,
                               fix_u=(case.u_inf,) * 6, fix_v=(0.0,) * 6, fix_w=(0.0,) * 6,
                               fix_p=(case.p_inf,) * 6, fix_twall=(0, 0, 310.0, 0, 0, 0))
    return case, nodes, q


TOTAL_PRESSURE = [(0, "sst"), (1, "sst"), (3, "sst"), (0, "none"), (3, "sst2003")]
POLE = [
    ((-9, -9, -7, -5, -9, -9), "muscl"),   # the pole turns on the first-cell MUSCL states
    ((-7, -4, -5, -7, -6, -6), "weno"),
    ((-3, -7, -6, -6, -7, -8), "none"),
]
