"""Size-independent properties of the residual shared by the oracle and
GPU tests."""
import math

import numpy as np


def boundary_flux_mismatch(F, G, H, res, l):
    """Finite-volume telescoping: the residue of one variable summed over
    all cells equals the net flux through the block's six boundary faces,
    since every interior face flux enters two cells with opposite signs
    (residue assembly, reference src/face/flux/convective/scheme.f90:111-141).  Returns
    |sum(res) - boundary sum| / sum(|res|)."""
    tot = math.fsum(res[l].ravel())
    bnd = math.fsum(np.concatenate([
        F[l, :, :, -1].ravel(), -F[l, :, :, 0].ravel(),
        G[l, :, -1, :].ravel(), -G[l, :, 0, :].ravel(),
        H[l, -1].ravel(), -H[l, 0].ravel()]))
    return abs(tot - bnd) / np.abs(res[l]).sum()
