// k_vjp.h -- the discrete adjoint (vector-Jacobian product) of the inviscid
// residual R(q) = sum over directions of F(n+1) - F(n) (src/face/flux/
// convective/scheme.f90:111-141, the fluxes of src/face/flux/convective/*.f90
// on the face states of src/face/state/*.f90), with respect to the cell
// state, through the ghost-cell rules of src/boundary/bc_primitive.f90.
//
// Each face's flux depends on the stencil cells n-3..n+2 of its direction
// through the face states qL, qR.  The two factors of the chain rule are
// taken apart, both in dual numbers running the face's own arithmetic
// (FaceMath): k_vjp_flux differentiates the flux with tangents on qL, then
// on qR (the 5 x 10 flux Jacobian, contracted at once with the face's
// adjoint flux barF = barR(n-1) - barR(n) into bar qL, bar qR);
// k_vjp_recon, one thread per face and variable, differentiates that
// variable's reconstruction with tangents on its six stencil values.
// k_vjp_gather sums, for every cell, the contributions of the faces whose
// stencils hold it (in a fixed order: deterministic), and k_vjp_bc moves the
// adjoints of ghost cells onto the interior cells the ghost rules read.
#pragma once
#pragma clang fp contract(fast)
#include "blk.h"
#include "k_bc.h"
#include "k_dual.h"
#include "k_flux.h"

namespace fest {

constexpr int VJP_STENCIL = 6;  // stencil cells n-3..n+2 of a face
constexpr int VJP_PER_FACE = VJP_STENCIL * 5;

// face t of direction D (faces 1..N along D, 1..N-1 across, in the F/G/H
// order): its indices and the stencil's cell and stride
struct VjpFace {
  int i, j, k, n, N;
  long c, s;
};
template <int D>
__device__ __forceinline__ VjpFace vjp_face(const Blk &b, long t) {
  VjpFace f;
  const int ni = b.nfi[D], nj = b.nfj[D];
  f.i = (int)(t % ni) + 1;
  f.j = (int)((t / ni) % nj) + 1;
  f.k = (int)(t / ((long)ni * nj)) + 1;
  f.N = D == 0 ? b.imx : D == 1 ? b.jmx : b.kmx;
  f.n = D == 0 ? f.i : D == 1 ? f.j : f.k;
  f.c = ci(b, f.i, f.j, f.k);
  f.s = cstride(b, D);
  return f;
}

// one thread per face of direction D: fbar[l][face] = bar qL_l and
// fbar[5 + l][face] = bar qR_l, the flux's adjoint on the face states.  The
// ten tangents (qL then qR) go through the flux KT at a time: fewer
// tangents per pass keep fewer values live (the FP64 flux arithmetic is
// register-bound) at the price of more passes
constexpr int VJP_FLUX_TANGENTS = 5;
template <int D, int RECON>
__global__ void __launch_bounds__(256) k_vjp_flux(Blk b, const double *__restrict__ barR, double *fbar) {
  constexpr int KT = VJP_FLUX_TANGENTS;
  using DK = Dual<KT>;
  const long nf = b.nfl[D];
  const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= nf) return;
  const VjpFace f = vjp_face<D>(b, t);
  const int ii = D == 0, jj = D == 1, kk = D == 2;
  // barF = barR of the cell below (n-1) minus that of the cell above (n)
  double barF[5];
  const bool below = f.n >= 2, above = f.n <= f.N - 1;
  const long rb = below ? ri(b, f.i - ii, f.j - jj, f.k - kk) : 0, ra = above ? ri(b, f.i, f.j, f.k) : 0;
  bool any = false;
#pragma unroll
  for (int l = 0; l < 5; ++l) {
    barF[l] = (below ? barR[(long)l * b.nres + rb] : 0.0) - (above ? barR[(long)l * b.nres + ra] : 0.0);
    any |= barF[l] != 0.0;
  }
  // the k faces of a 2-D grid carry no convective flux
  if (D == 2 && b.kmx == 2) any = false;
  double bq[10] = {0, 0, 0, 0, 0, 0, 0, 0, 0, 0};  // bar qL, then bar qR
  if (any) {
    double q[10];  // qL, then qR
    typename FaceMath<double>::Aux ax;
    ax.pb = false;
    if constexpr (RECON == 4) FaceMath<double>::nm_blends(b.vol + f.c, f.s, ax);
#pragma unroll
    for (int l = 0; l < 5; ++l)
      FaceMath<double>::template face_states<RECON>(b, D, f.n, f.N, l, b.qp + (long)l * b.nc + f.c, f.s, ax, q[l],
                                                    q[5 + l]);
    const long fo = fi(b, D, f.i, f.j, f.k);
    const DK nx = b.fx[D][fo], ny = b.fy[D][fo], nz = b.fz[D][fo], A = b.fA[D][fo];
    const DK zero = (f.n == 1) ? (double)b.mz[D][0] : (f.n == f.N) ? (double)b.mz[D][1] : 1.0;
#pragma unroll
    for (int p0 = 0; p0 < 10; p0 += KT) {
      DK L[5], R[5], fl[5];
#pragma unroll
      for (int l = 0; l < 5; ++l) {
        L[l] = DK(q[l]);
        R[l] = DK(q[5 + l]);
      }
#pragma unroll
      for (int e = 0; e < KT; ++e) {
        const int x = p0 + e;  // tangent e of this pass: input x
        if (x < 5) L[x].d[e] = 1.0;
        else if (x < 10) R[x - 5].d[e] = 1.0;
      }
      FaceMath<DK>::template inviscid_flux<5>(b, L, R, nx, ny, nz, A, zero, fl);
#pragma unroll
      for (int e = 0; e < KT; ++e) {
        double acc = 0.0;
#pragma unroll
        for (int l = 0; l < 5; ++l) acc += fl[l].d[e] * barF[l];
        if (p0 + e < 10) bq[p0 + e] = acc;
      }
    }
  }
#pragma unroll
  for (int l = 0; l < 10; ++l) fbar[(long)l * nf + t] = bq[l];
}

// one thread per face of direction D and variable l (l-major):
// contrib[m * 5 + l][face] = bar qL_l dqL_l/du_m + bar qR_l dqR_l/du_m, the
// face's contribution to the adjoint of variable l of stencil cell n-3+m
template <int D, int RECON>
__global__ void __launch_bounds__(256) k_vjp_recon(Blk b, const double *__restrict__ fbar, double *contrib) {
  using D6 = Dual<VJP_STENCIL>;
  const long nf = b.nfl[D];
  const long tt = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (tt >= 5 * nf) return;
  const int l = (int)(tt / nf);
  const long t = tt - (long)l * nf;
  const double bl = fbar[(long)l * nf + t], br = fbar[(long)(5 + l) * nf + t];
  double o[VJP_STENCIL] = {0, 0, 0, 0, 0, 0};
  if (bl != 0.0 || br != 0.0) {
    const VjpFace f = vjp_face<D>(b, t);
    typename FaceMath<D6>::Aux ax;
    ax.pb = false;
    if constexpr (RECON == 4) FaceMath<D6>::nm_blends(b.vol + f.c, f.s, ax);
    D6 u[VJP_STENCIL], L, R;
#pragma unroll
    for (int m = 0; m < VJP_STENCIL; ++m) {
      u[m] = D6(b.qp[(long)l * b.nc + f.c + (long)(m - 3) * f.s]);
      u[m].d[m] = 1.0;
    }
    FaceMath<D6>::template face_states<RECON>(b, D, f.n, f.N, l, &u[3], 1, ax, L, R);
#pragma unroll
    for (int m = 0; m < VJP_STENCIL; ++m) o[m] = bl * L.d[m] + br * R.d[m];
  }
#pragma unroll
  for (int m = 0; m < VJP_STENCIL; ++m) contrib[(long)(m * 5 + l) * nf + t] = o[m];
}

// one thread per padded cell: the sum over directions and over the faces
// whose stencils hold the cell (face n = c + 3 - m of stencil position m);
// c0, c1, c2: the contributions of the three directions (k_vjp_recon)
__global__ void __launch_bounds__(256) k_vjp_gather(Blk b, const double *__restrict__ c0,
                                                     const double *__restrict__ c1, const double *__restrict__ c2,
                                                     double *bar) {
  const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= b.nc) return;
  const int i = (int)(t % b.pi) - 2, j = (int)((t / b.pi) % b.pj) - 2, k = (int)(t / ((long)b.pi * b.pj)) - 2;
  double acc[5] = {0, 0, 0, 0, 0};
  const double *cs[3] = {c0, c1, c2};
  const long nf[3] = {b.nfl[0], b.nfl[1], b.nfl[2]};
#pragma unroll
  for (int d = 0; d < 3; ++d) {
    const int N = d == 0 ? b.imx : d == 1 ? b.jmx : b.kmx;
    const int cn = d == 0 ? i : d == 1 ? j : k;
    // the cross coordinates must be those of faces of direction d
    const int a = d == 0 ? j : i, bb = d == 2 ? j : k;
    const int na = d == 0 ? b.jmx : b.imx, nb = d == 2 ? b.jmx : b.kmx;
    if (a < 1 || a > na - 1 || bb < 1 || bb > nb - 1) continue;
    for (int m = 0; m < VJP_STENCIL; ++m) {
      const int n = cn + 3 - m;
      if (n < 1 || n > N) continue;
      const long f = d == 0 ? fli(b, 0, n, j, k) : d == 1 ? fli(b, 1, i, n, k) : fli(b, 2, i, j, n);
#pragma unroll
      for (int v = 0; v < 5; ++v) acc[v] += cs[d][(long)(m * 5 + v) * nf[d] + f];
    }
  }
#pragma unroll
  for (int v = 0; v < 5; ++v) bar[(long)v * b.nc + t] = acc[v];
}

// the transpose of one face's ghost-cell rule on the adjoint: the ghost
// cells of the face interior hand their adjoints to the cells their values
// were read from, then hold none.  Supported rules (fest_vjp checks): -9
// periodic, -5 no-slip wall with the symmetric density (wall temperature
// in [0, 1]), -6 slip wall, -1..-4 inlets and outlets, -7 pole, -8 far
// field and -11 total-pressure inlet (their Riemann state differentiated in
// dual numbers); interfaces (>= 0, -10) keep their ghost adjoints, which
// belong to the neighbouring block.
__global__ void __launch_bounds__(256) k_vjp_bc(Blk b, int face, double *bar) {
  const int a = (int)(blockIdx.x * blockDim.x + threadIdx.x) + 1;
  const int bb = (int)blockIdx.y + 1;
  const int na = face_na(b, face), nb = face_nb(b, face);
  if (a > na - 1 || bb > nb - 1) return;
  const FaceBC p = face_params(b, face);
  const int id = p.id;
  if (id >= 0 || id == -10) return;
  const long nc = b.nc;
  auto G = [&](int v, int L) -> double & { return bar[(long)v * nc + fcell(b, face, L, a, bb)]; };
  double gb[3][5];
#pragma unroll
  for (int g = 0; g < 3; ++g)
#pragma unroll
    for (int v = 0; v < 5; ++v) {
      gb[g][v] = G(v, -g);
      G(v, -g) = 0.0;
    }
  switch (id) {
    case -9: {  // ghost -g copies the cell across the block (k_bc_face_b)
      const int d = face >> 1, hi = face & 1;
      const int Nd = d == 0 ? b.imx : d == 1 ? b.jmx : b.kmx;
      for (int g = 0; g < 3; ++g) {
        const int L = hi ? 1 + g : Nd - 1 - g;  // along d
        const long src = d == 0 ? ci(b, L, a, bb) : d == 1 ? ci(b, a, L, bb) : ci(b, a, bb, L);
#pragma unroll
        for (int v = 0; v < 5; ++v) bar[(long)v * nc + src] += gb[g][v];
      }
      break;
    }
    case -5:    // no-slip wall: rho, p symmetric copies, velocity anti
    case -6: {  // slip wall: rho, p symmetric copies, velocity reflected
      // ghost -g = (c2 q(g+1) - c3 q(g+2)) / c1 (copy_bc.f90's symmetric copy)
      const double cc2 = b.c2 / b.c1, cc3 = b.c3 / b.c1;
      double nx = 0, ny = 0, nz = 0, rx = 0, ry = 0, rz = 0;
      if (id == -6) {
        // u_g = u_c - 2 (n . u_c) r, n the boundary normal and r the
        // I-face normal at the same index (flow_tangency_col)
        bnormal(b, face, a, bb, nx, ny, nz);
        long f;
        switch (face) {
          case 0: f = fi(b, 0, 1, a, bb); break;
          case 1: f = fi(b, 0, b.imx, a, bb); break;
          case 2: f = fi(b, 0, a, 1, bb); break;
          case 3: f = fi(b, 0, a, b.jmx, bb); break;
          case 4: f = fi(b, 0, a, bb, 1); break;
          default: f = fi(b, 0, a, bb, b.kmx); break;
        }
        rx = b.fx[0][f]; ry = b.fy[0][f]; rz = b.fz[0][f];
      }
      for (int g = 0; g < 3; ++g) {
        const double *w = gb[g];
        G(0, g + 1) += cc2 * w[0];
        G(0, g + 2) -= cc3 * w[0];
        G(4, g + 1) += cc2 * w[4];
        G(4, g + 2) -= cc3 * w[4];
        if (id == -5) {
          G(1, g + 1) -= w[1];
          G(2, g + 1) -= w[2];
          G(3, g + 1) -= w[3];
        } else {
          const double rg = rx * w[1] + ry * w[2] + rz * w[3];
          G(1, g + 1) += w[1] - 2.0 * nx * rg;
          G(2, g + 1) += w[2] - 2.0 * ny * rg;
          G(3, g + 1) += w[3] - 2.0 * nz * rg;
        }
      }
      break;
    }
    case -8:      // far field and
    case -11: {   // total-pressure inlet: layer 0 from the Riemann state of
                  // the first interior cell, layers -1 and -2 its copies
      using DT = Dual<5>;
      DT c[5], g[5];
      const long c1 = fcell(b, face, 1, a, bb);
#pragma unroll
      for (int l = 0; l < 5; ++l) {
        c[l] = DT(b.qp[(long)l * nc + c1]);
        c[l].d[l] = 1.0;
      }
      riemann_state(b, face, a, bb, c, true, id == -11, p.fix[FX_TP], g);
#pragma unroll
      for (int v = 0; v < 5; ++v) {
        double acc = 0.0;
#pragma unroll
        for (int l = 0; l < 5; ++l) acc += g[l].d[v] * (gb[0][l] + gb[1][l] + gb[2][l]);
        G(v, 1) += acc;
      }
      break;
    }
    case -2:  // supersonic outlet: flat copies
    case -7:  // pole: flat copies
      for (int g = 0; g < 3; ++g)
#pragma unroll
        for (int v = 0; v < 5; ++v) G(v, 1) += gb[g][v];
      break;
    case -3:  // subsonic inlet: rho, u, v, w fixed, p a flat copy
      for (int g = 0; g < 3; ++g) G(4, 1) += gb[g][4];
      break;
    case -4:  // subsonic outlet: rho, u, v, w flat copies, p fixed
      for (int g = 0; g < 3; ++g)
#pragma unroll
        for (int v = 0; v < 4; ++v) G(v, 1) += gb[g][v];
      break;
    default:  // -1: fixed values
      break;
  }
}

}  // namespace fest
