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
  lusgs_sweep (k_lusgs_sweep, the LU-SGS update of the 'implicit' step: one
    launch per hyperplane i+j+k = s, 2 (ni+nj+nk-2) launches a step, which
    the engine's event timer brackets as one; priced per bracket, i.e. both
    sweeps over every interior cell), per interior cell and sweep, from the
    hyperplane-major storage:
    read  q (n_var), three faces x {A,nx,ny,nz}, volume, mu, mu_t, F1, three
          radii, three Old fluxes (3 x n_var), the diagonal (1, or 3 with
          SST), the right-hand side (residue or own value, n_var), the
          neighbours' sweep values (n_var, each value once)
    write the sweep value (n_var)
  lusgs_diag (k_to_skew + k_lusgs_prep, once a step), per interior cell:
    the copies of q, mu, mu_t, F1 and the residue into hyperplane-major
    storage (read + write, 2 x (2 n_var + 3)); the prep's stencil reads (q,
    six faces x 4 metrics counted once a face, volume, mu, mu_t, F1, centres,
    dt: n_var + 12 + 8) and writes (three radii, six Old fluxes, diagonal:
    3 + 6 n_var + 1 or 3)
  lusgs_apply (k_lusgs_apply): read q and the sweep value, write q (3 n_var)
"""

W = 8  # bytes per FP64 value


def _ngrad(case):
    return 6 if case.sst_family else 4


def lusgs_launches(case):
    """Hyperplane launches of the two LU-SGS sweeps of one step."""
    return 2 * ((case.imx - 1) + (case.jmx - 1) + (case.kmx - 1) - 2)


def launch_items(case, group):
    """Number of faces or cells one timed launch of `group` processes (for
    the sweep, cell-sweeps of both sweeps of a step)."""
    I, J, K = case.imx, case.jmx, case.kmx
    if group == "lusgs_sweep":
        return 2 * (I - 1) * (J - 1) * (K - 1)
    return {
        "flux_i": I * (J - 1) * (K - 1),
        "flux_j": (I - 1) * J * (K - 1),
        "flux_k": (I - 1) * (J - 1) * K,
        "grad": (I + 1) * (J + 1) * (K + 1),
        "update": (I - 1) * (J - 1) * (K - 1),
        "lusgs_diag": (I - 1) * (J - 1) * (K - 1),
        "lusgs_apply": (I - 1) * (J - 1) * (K - 1),
    }.get(group)


def bytes_per_item(case, group):
    """Algorithmic bytes per face or cell of one launch of `group`."""
    nv, ng, sst = case.n_var, _ngrad(case), case.sst_family
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
    ndg = 3 if sst else 1
    if group == "lusgs_sweep":
        reads = nv + 12 + 4 + 3 + 3 * nv + ndg + nv + nv
        return W * (reads + nv)
    if group == "lusgs_diag":
        copies = 2 * (2 * nv + 3)
        return W * (copies + (nv + 12 + 8) + (3 + 6 * nv + ndg))
    if group == "lusgs_apply":
        return W * 3 * nv
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
