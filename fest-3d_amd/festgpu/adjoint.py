"""The unsteady discrete adjoint of RK4 steps of the inviscid residual, with
checkpoints and recomputation.

The step is the reference's RK4 (src/update.f90:176-187) with a global time
step (dt fixed, `global_time_step` > 0, src/time.f90:291-334):

    U_s     = C(q_n) - a_s dt/V R(q_{s-1}),  q_s = P(U_s),  a = 1/2, 1/2, 1
    q_{n+1} = P(C(q_n) - dt/(6V) (R(q_0) + 2 R(q_1) + 2 R(q_2) + R(q_3)))

with C and P the primitive <-> conservative maps of src/update.f90:367-485
and R the residual of fest_residual.  The forward pass evaluates R on the
device (fest_residual) and the stage arithmetic here; the reverse pass
applies the transposes of P and C here and R's transpose on the device
(fest_vjp).  States are kept every `every` steps; the reverse sweep
recomputes the steps of one interval from its checkpoint, keeps their stage
states and walks them backwards, so memory is (nsteps/every + 4 every)
states instead of 4 nsteps.

Inviscid blocks without interface faces (what fest_vjp differentiates).
"""
import numpy as np

INNER = (slice(None),) + (slice(3, -3),) * 3
A_STAGE = (0.5, 0.5, 1.0)
W_FINAL = (1.0, 2.0, 2.0, 1.0)


class RK4Adjoint:
    """Forward RK4 steps and their adjoint on one GpuBlock (grid set)."""

    def __init__(self, block, dt: float):
        if block.case.mu_ref != 0.0:
            # the stage states are uploaded with zero ghost cells, whose
            # temperatures the viscous gradients would read
            raise ValueError("RK4Adjoint drives inviscid blocks (mu_ref 0)")
        self.g = block
        self.case = block.case
        self.gm = self.case.gm
        self.dt = float(dt)
        vol = block.download("cells")[0]
        self.dtv = self.dt / vol[3:-3, 3:-3, 3:-3]
        self.residual_calls = 0
        self.vjp_calls = 0

    # ---- the primitive <-> conservative maps and their transposes
    def cons(self, q):
        r, u, v, w, p = q
        return np.stack([r, r * u, r * v, r * w, p / (self.gm - 1.0) + 0.5 * r * (u * u + v * v + w * w)])

    def prim(self, U):
        r = U[0]
        u, v, w = U[1] / r, U[2] / r, U[3] / r
        p = (self.gm - 1.0) * (U[4] - 0.5 * r * (u * u + v * v + w * w))
        return np.stack([r, u, v, w, p])

    def cons_T(self, q, bU):
        r, u, v, w, p = q
        k = 0.5 * (u * u + v * v + w * w)
        return np.stack([bU[0] + u * bU[1] + v * bU[2] + w * bU[3] + k * bU[4],
                         r * bU[1] + r * u * bU[4], r * bU[2] + r * v * bU[4], r * bU[3] + r * w * bU[4],
                         bU[4] / (self.gm - 1.0)])

    def prim_T(self, U, bq):
        r = U[0]
        u, v, w = U[1] / r, U[2] / r, U[3] / r
        g1 = self.gm - 1.0
        k = 0.5 * (u * u + v * v + w * w)
        return np.stack([bq[0] - (u * bq[1] + v * bq[2] + w * bq[3]) / r + g1 * k * bq[4],
                         bq[1] / r - g1 * u * bq[4], bq[2] / r - g1 * v * bq[4], bq[3] / r - g1 * w * bq[4],
                         g1 * bq[4]])

    # ---- device calls
    def _padded(self, qi):
        q = np.zeros((5,) + self.case.cell_shape())
        q[INNER] = qi
        return q

    def residual(self, qi):
        self.g.upload_state(self._padded(qi))
        self.g.residual(1)
        self.residual_calls += 1
        return self.g.download("residue")

    def vjp(self, qi, bar_r):
        self.g.upload_state(self._padded(qi))
        self.vjp_calls += 1
        return self.g.vjp(bar_r)[INNER]

    # ---- one step
    def step(self, qn, keep=False):
        """q_{n+1} from q_n (interior arrays (5, nk, nj, ni)); with keep, also
        the stage states and conservative states the reverse pass needs."""
        Cn = self.cons(qn)
        qs, Us, R = [qn], [], []
        for s in range(4):
            R.append(self.residual(qs[-1]))
            if s < 3:
                U = Cn - A_STAGE[s] * self.dtv * R[-1]
                Us.append(U)
                qs.append(self.prim(U))
        U4 = Cn - (self.dtv / 6.0) * sum(w * r for w, r in zip(W_FINAL, R))
        q1 = self.prim(U4)
        return (q1, (qs, Us, U4)) if keep else q1

    def step_T(self, qn, tape, bar_q1):
        """bar q_n from bar q_{n+1}, given the step's stage states."""
        qs, Us, U4 = tape
        bU = self.prim_T(U4, bar_q1)
        bCn = bU.copy()
        bR = [-(w / 6.0) * self.dtv * bU for w in W_FINAL]
        for s in (3, 2, 1):
            bq = self.vjp(qs[s], bR[s])
            bUs = self.prim_T(Us[s - 1], bq)
            bCn += bUs
            bR[s - 1] = bR[s - 1] - A_STAGE[s - 1] * self.dtv * bUs
        return self.cons_T(qn, bCn) + self.vjp(qn, bR[0])

    # ---- trajectories
    def forward(self, q0, nsteps, every):
        """q_N and the checkpoints {n: q_n} every `every` steps."""
        cps, q = {0: q0.copy()}, q0.copy()
        for n in range(nsteps):
            q = self.step(q)
            if (n + 1) % every == 0 and n + 1 < nsteps:
                cps[n + 1] = q.copy()
        return q, cps

    def gradient(self, q0, nsteps, every, dJ_dqN):
        """dJ/dq_0 for an objective J(q_N) whose gradient is dJ_dqN(q_N)."""
        qN, cps = self.forward(q0, nsteps, every)
        bar = dJ_dqN(qN)
        for start in sorted(cps, reverse=True):
            stop = min(start + every, nsteps)
            q, tapes = cps[start], []
            for _ in range(start, stop):  # recompute the interval, keep its stages
                q1, tape = self.step(q, keep=True)
                tapes.append((q, tape))
                q = q1
            for qn, tape in reversed(tapes):
                bar = self.step_T(qn, tape, bar)
        return qN, bar


def wall_force_x(block, face=2):
    """J = sum over a wall face of (p - p_inf) n_x A, with p the pressure of
    the first interior cell and n the face normal as stored (pointing to
    increasing j): the pressure part of the wall-force measure of
    tests/Lfp/pp/Surface.py:95-136.  Returns the functions J(q) and dJ/dq
    of the interior state.  face: 2 (jmin) or 3 (jmax)."""
    c = block.case
    jf = block.download("faces")[1]  # (4, kmx+5, jmx+6, imx+5): A, nx, ny, nz
    jn = 3 if face == 2 else c.jmx + 2  # padded index of the face j = 1 or j = jmx
    jc = 0 if face == 2 else c.jmx - 2  # interior index of the cell next to it
    w = jf[0, 3:-3, jn, 3:-3] * jf[1, 3:-3, jn, 3:-3]

    def J(q):
        return float(np.sum((q[4, :, jc, :] - c.p_inf) * w))

    def dJ(q):
        out = np.zeros_like(q)
        out[4, :, jc, :] = w
        return out

    return J, dJ
