"""Algorithmic HBM bytes of the hot kernels, per launch, for the roofline.

Each figure counts the bytes a kernel must move at least once per launch:
every array element it reads or writes, taken once (stencil neighbours come
from cache), FP64 throughout.  Face arrays are counted per face, cell arrays
per cell of the launch's index box.  The numbers follow the kernels in
fest-3d_amd/csrc:

  flux_i/j/k (k_flux<D>, one thread per face of direction D):
    read  q (n_var planes), the face's A,nx,ny,nz (4), both cell centres
          (3, one per cell), the cell gradients (3 x n_grad), mu (+ mu_t, F1
          with SST)
    write the face flux (n_var); the face sound speed at stage 1 (1, a
          quarter per launch over the four RK stages)
  grad (k_grad<NG>, one thread per cell 0..imx x 0..jmx x 0..kmx):
    read  the n_grad gradient variables, volume, the i/j/k face metrics
          (3 x 4), rho and p when mu follows Sutherland, rho, dist and mu
          (when constant) with SST
    write 3 x n_grad gradients, mu (Sutherland), mu_t, F1 and the two SST
          source terms
  update (k_update, one thread per interior cell), averaged over the four
    RK4 stages:
    read  the three face fluxes (3 x n_var, one face per direction per
          cell), the RK base state (n_var), dt, volume, the R accumulator
          (stages 2-4), the SST source (2), F1 and omega (SST)
    write q (n_var), the R accumulator (stages 1-3), the residue (stage 4)
"""

W = 8  # bytes per FP64 value


def _ngrad(case):
    return 6 if case.turb == "sst" else 4


def launch_items(case, group):
    """Number of faces or cells one launch of `group` processes."""
    I, J, K = case.imx, case.jmx, case.kmx
    return {
        "flux_i": I * (J - 1) * (K - 1),
        "flux_j": (I - 1) * J * (K - 1),
        "flux_k": (I - 1) * (J - 1) * K,
        "grad": (I + 1) * (J + 1) * (K + 1),
        "update": (I - 1) * (J - 1) * (K - 1),
    }.get(group)


def bytes_per_item(case, group):
    """Algorithmic bytes per face or cell of one launch of `group`."""
    nv, ng, sst = case.n_var, _ngrad(case), case.turb == "sst"
    if group in ("flux_i", "flux_j", "flux_k"):
        reads = nv + 4 + 3 + 3 * ng + (3 if sst else 1)
        writes = nv + 0.25
        return W * (reads + writes)
    if group == "grad":
        reads = ng + 1 + 12
        if case.mu_sutherland:
            reads += 2
        if sst:
            reads += 1 + (0 if case.mu_sutherland else 2)  # dist (+ rho, mu)
        writes = 3 * ng + (1 if case.mu_sutherland else 0) + (4 if sst else 0)
        return W * (reads + writes)
    if group == "update":
        base = 3 * nv + nv + 2 + nv  # fluxes, RK base, dt + vol, q write
        if sst:
            base += 2 + 2  # source, F1 + omega
        acc = (nv + 2 * nv + 2 * nv + 2 * nv) / 4.0  # R: write; read+write x2; read + residue write
        return W * (base + acc)
    return None


def bytes_per_launch(case, group):
    n, b = launch_items(case, group), bytes_per_item(case, group)
    if n is None or b is None:
        return None
    return n * b


# whole RK stage at the fused minimum: q read and write, the RK base, the R
# accumulator, three owned faces x {A,nx,ny,nz}, volume, time step, cell
# centres, stale temperature and wall distance (stage-averaged)
BYTES_PER_CELL_STAGE = {"sst": 406.0, "none": 326.0}
HBM_PEAK_GBPS = 8000.0
