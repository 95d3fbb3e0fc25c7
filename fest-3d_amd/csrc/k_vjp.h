// k_vjp.h -- the discrete adjoint (vector-Jacobian product) of the inviscid
// residual R(q) = sum over directions of F(n+1) - F(n) (src/face/flux/
// convective/scheme.f90:111-141, the fluxes of src/face/flux/convective/*.f90
// on the face states of src/face/state/*.f90), with respect to the cell
// state, through the ghost-cell rules of src/boundary/bc_primitive.f90.
//
// Each face's flux depends on the stencil cells n-3..n+2 of its direction.
// k_vjp_faces evaluates the face's own arithmetic (FaceMath) in dual numbers,
// once per stencil cell with the tangents seeded on that cell's five
// variables: the exact local Jacobian, one 5 x 5 block per stencil cell,
// contracted at once with the face's adjoint flux barF = barR(n-1) - barR(n).
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

// one thread per face of direction D (faces 1..N along D, 1..N-1 across);
// contrib[face][m][v]: d(barF . F)/dq(cell n-3+m, variable v)
template <int D, int RECON>
__global__ void __launch_bounds__(256) k_vjp_faces(Blk b, const double *__restrict__ barR, double *contrib) {
  using DT = Dual<5>;
  using M = FaceMath<DT>;
  const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= b.nfl[D]) return;
  const int ni = b.nfi[D], nj = b.nfj[D];
  const int i = (int)(t % ni) + 1, j = (int)((t / ni) % nj) + 1, k = (int)(t / ((long)ni * nj)) + 1;
  const int ii = D == 0, jj = D == 1, kk = D == 2;
  const int N = D == 0 ? b.imx : D == 1 ? b.jmx : b.kmx;
  const int n = D == 0 ? i : D == 1 ? j : k;
  double *out = contrib + t * VJP_PER_FACE;
#pragma unroll
  for (int q = 0; q < VJP_PER_FACE; ++q) out[q] = 0.0;
  // the k faces of a 2-D grid carry no convective flux
  if (D == 2 && b.kmx == 2) return;
  // barF = barR of the cell below (n-1) minus that of the cell above (n)
  double barF[5];
  const bool below = n >= 2, above = n <= N - 1;
  const long rb = below ? ri(b, i - ii, j - jj, k - kk) : 0, ra = above ? ri(b, i, j, k) : 0;
  bool any = false;
#pragma unroll
  for (int l = 0; l < 5; ++l) {
    barF[l] = (below ? barR[(long)l * b.nres + rb] : 0.0) - (above ? barR[(long)l * b.nres + ra] : 0.0);
    any |= barF[l] != 0.0;
  }
  if (!any) return;
  const long s = cstride(b, D);
  const long c = ci(b, i, j, k);
  const long fo = fi(b, D, i, j, k);
  const DT nx = b.fx[D][fo], ny = b.fy[D][fo], nz = b.fz[D][fo], A = b.fA[D][fo];
  const DT zero = (n == 1) ? (double)b.mz[D][0] : (n == N) ? (double)b.mz[D][1] : 1.0;
  typename M::Aux ax;
  ax.pb = false;
  if constexpr (RECON == 4) M::nm_blends(b.vol + c, s, ax);
  // the stencil cells each reconstruction reads (k_flux.h face_states)
  constexpr int m0 = RECON == 0 ? 2 : RECON == 1 ? 1 : 0;
  constexpr int m1 = RECON == 0 ? 3 : RECON == 1 ? 4 : 5;
  for (int m = m0; m <= m1; ++m) {
    DT qL[5], qR[5], fl[5];
#pragma unroll
    for (int l = 0; l < 5; ++l) {
      DT u[VJP_STENCIL];
#pragma unroll
      for (int mm = 0; mm < VJP_STENCIL; ++mm) {
        u[mm] = DT(b.qp[(long)l * b.nc + c + (long)(mm - 3) * s]);
        if (mm == m) u[mm].d[l] = 1.0;
      }
      M::template face_states<RECON>(b, D, n, N, l, &u[3], 1, ax, qL[l], qR[l]);
    }
    M::template inviscid_flux<5>(b, qL, qR, nx, ny, nz, A, zero, fl);
#pragma unroll
    for (int v = 0; v < 5; ++v) {
      double acc = 0.0;
#pragma unroll
      for (int l = 0; l < 5; ++l) acc += fl[l].d[v] * barF[l];
      out[m * 5 + v] = acc;
    }
  }
}

// one thread per padded cell: the sum over directions and over the faces
// whose stencils hold the cell (face n = c + 3 - m of stencil position m)
__global__ void __launch_bounds__(256) k_vjp_gather(Blk b, const double *__restrict__ c0,
                                                     const double *__restrict__ c1, const double *__restrict__ c2,
                                                     double *bar) {
  const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= b.nc) return;
  const int i = (int)(t % b.pi) - 2, j = (int)((t / b.pi) % b.pj) - 2, k = (int)(t / ((long)b.pi * b.pj)) - 2;
  double acc[5] = {0, 0, 0, 0, 0};
  const double *cs[3] = {c0, c1, c2};
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
      for (int v = 0; v < 5; ++v) acc[v] += cs[d][f * VJP_PER_FACE + m * 5 + v];
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
