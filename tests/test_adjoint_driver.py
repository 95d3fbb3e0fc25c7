"""The host side of the unsteady adjoint (festgpu.adjoint.RK4Adjoint) on the
CPU, with a stand-in block whose residual is a fixed smooth nonlinear map
and whose vjp is that map's exact transpose: the primitive <-> conservative
maps invert each other, their transposes pass the dot-product test against
central differences, and the gradient through RK4 steps with checkpoints
matches central differences of the whole trajectory and does not depend on
the checkpoint interval.  The device residual and its transpose are checked
on the GPU (tests/test_gpu_vjp.py)."""
import numpy as np

from festgpu import cases as C
from festgpu.adjoint import INNER, RK4Adjoint


class StandIn:
    """R(q) = A tanh(q / s) per cell and variable (mixing the five
    variables through a fixed matrix A), the residue layout of a block."""

    def __init__(self, case, seed=0):
        self.case = case
        rng = np.random.default_rng(seed)
        self.A = rng.uniform(-1.0, 1.0, (5, 5)) * 0.3
        self.s = np.array([1.0, 0.5, 0.5, 0.5, 1.0 / 1.4])[:, None, None, None]
        self.q = None

    def download(self, name):
        if name == "cells":
            c = self.case
            cells = np.ones((4,) + c.cell_shape())
            cells[0] *= 0.01
            return cells
        assert name == "residue"
        qi = self.q[INNER]
        return np.einsum("lm,m...->l...", self.A, np.tanh(qi / self.s))

    def upload_state(self, q):
        self.q = q.copy()

    def residual(self, it=1):
        pass

    def vjp(self, w):
        qi = self.q[INNER]
        d = (1.0 - np.tanh(qi / self.s) ** 2) / self.s
        out = np.zeros_like(self.q)
        out[INNER] = np.einsum("lm,l...->m...", self.A, w) * d
        return out


def setup(seed=0):
    case, _, q = C.euler_bump(6, 5, 4, seed=seed)
    blk = StandIn(case, seed)
    return case, blk, RK4Adjoint(blk, dt=0.002), np.ascontiguousarray(q[INNER])


def test_primitive_conservative_maps_and_transposes():
    case, blk, adj, q = setup()
    assert np.allclose(adj.prim(adj.cons(q)), q, rtol=1e-14, atol=1e-14)
    rng = np.random.default_rng(5)
    h = 1e-6
    for _ in range(4):
        dq, bU = rng.uniform(-1, 1, q.shape), rng.uniform(-1, 1, q.shape)
        fd = np.sum(bU * (adj.cons(q + h * dq) - adj.cons(q - h * dq))) / (2 * h)
        assert abs(np.sum(adj.cons_T(q, bU) * dq) - fd) <= 1e-8 * abs(fd)
        U = adj.cons(q)
        dU = rng.uniform(-1, 1, q.shape)
        fd = np.sum(bU * (adj.prim(U + h * dU) - adj.prim(U - h * dU))) / (2 * h)
        assert abs(np.sum(adj.prim_T(U, bU) * dU) - fd) <= 1e-8 * abs(fd)


def test_rk4_gradient_with_checkpoints_matches_differences():
    case, blk, adj, q = setup(3)
    rng = np.random.default_rng(7)
    w = rng.uniform(-1, 1, q.shape)

    def J(qN):
        return float(np.sum(w * qN))

    N = 9
    _, g3 = adj.gradient(q, N, 3, lambda qN: w)
    _, g9 = adj.gradient(q, N, 9, lambda qN: w)
    _, g2 = adj.gradient(q, N, 2, lambda qN: w)
    assert np.array_equal(g3, g9) and np.array_equal(g2, g9)
    h = 1e-6
    for _ in range(4):
        dq = rng.uniform(-1, 1, q.shape) * np.array([1.0, 0.5, 0.5, 0.5, 0.7])[:, None, None, None]
        fd = (J(adj.forward(q + h * dq, N, N)[0]) - J(adj.forward(q - h * dq, N, N)[0])) / (2 * h)
        a = float(np.sum(g3 * dq))
        assert abs(a - fd) <= 1e-7 * max(abs(a), abs(fd)), (a, fd)


def test_autograd_wrapper_with_the_stand_in():
    """festgpu.autograd's Function around the stand-in block: gradcheck
    (finite-difference Jacobian against the vjp rows)."""
    import pytest
    torch = pytest.importorskip("torch")
    from festgpu.autograd import inviscid_residual
    case, _, q = C.euler_bump(4, 5, 4, seed=2)
    blk = StandIn(case, 2)
    qi = torch.from_numpy(np.ascontiguousarray(q[INNER])).requires_grad_(True)
    assert torch.autograd.gradcheck(lambda x: inviscid_residual(x, blk), (qi,), eps=1e-6, atol=1e-8, rtol=1e-6)
