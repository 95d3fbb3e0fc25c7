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
// fbar[5 + l][face] = bar qR_l, the flux's adjoint on the face states, and,
// when mbar is given, mbar[c][face] = the adjoint of the face's stored unit
// normal (c = 0..2) and area (c = 3).  The tangents (qL, qR, then the
// metrics) go through the flux KT at a time: fewer tangents per pass keep
// fewer values live (the FP64 flux arithmetic is register-bound) at the
// price of more passes
constexpr int VJP_FLUX_TANGENTS = 5;
template <int D, int RECON>
__global__ void __launch_bounds__(256) k_vjp_flux(Blk b, const double *__restrict__ barR, double *fbar,
                                                  double *mbar) {
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
  const int ntan = mbar ? 14 : 10;
  double bq[14] = {0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0};  // bar qL, bar qR, bar (n, A)
  if (any) {
    double q[10];  // qL, then qR
    typename FaceMath<double>::Aux ax;
    ax.pb = (RECON == 1 || RECON == 2) && b.pbs[D];
    if (ax.pb) FaceMath<double>::pb_factors(b, f.n, f.N, b.qp + 4 * b.nc + f.c, f.s, ax);
    if constexpr (RECON == 4) FaceMath<double>::nm_blends(b.vol + f.c, f.s, ax);
#pragma unroll
    for (int l = 0; l < 5; ++l)
      FaceMath<double>::template face_states<RECON>(b, D, f.n, f.N, l, b.qp + (long)l * b.nc + f.c, f.s, ax, q[l],
                                                    q[5 + l]);
    const long fo = fi(b, D, f.i, f.j, f.k);
    const double m0[4] = {b.fx[D][fo], b.fy[D][fo], b.fz[D][fo], b.fA[D][fo]};
    const DK zero = (f.n == 1) ? (double)b.mz[D][0] : (f.n == f.N) ? (double)b.mz[D][1] : 1.0;
#pragma unroll
    for (int p0 = 0; p0 < 14; p0 += KT) {
      if (p0 < ntan) {
        DK L[5], R[5], fl[5], m[4];
#pragma unroll
        for (int l = 0; l < 5; ++l) {
          L[l] = DK(q[l]);
          R[l] = DK(q[5 + l]);
        }
#pragma unroll
        for (int c = 0; c < 4; ++c) m[c] = DK(m0[c]);
#pragma unroll
        for (int e = 0; e < KT; ++e) {
          const int x = p0 + e;  // tangent e of this pass: input x
          if (x < 5) L[x].d[e] = 1.0;
          else if (x < 10) R[x - 5].d[e] = 1.0;
          else if (x < 14) m[x - 10].d[e] = 1.0;
        }
        FaceMath<DK>::template inviscid_flux<5>(b, L, R, m[0], m[1], m[2], m[3], zero, fl);
#pragma unroll
        for (int e = 0; e < KT; ++e) {
          double acc = 0.0;
#pragma unroll
          for (int l = 0; l < 5; ++l) acc += fl[l].d[e] * barF[l];
          if (p0 + e < 14) bq[p0 + e] = acc;
        }
      }
    }
  }
#pragma unroll
  for (int l = 0; l < 10; ++l) fbar[(long)l * nf + t] = bq[l];
  if (mbar) {
#pragma unroll
    for (int c = 0; c < 4; ++c) mbar[(long)c * nf + t] = bq[10 + c];
  }
}

// one thread per face of direction D and variable l (l-major):
// contrib[m * 5 + l][face] = bar qL_l dqL_l/du_m + bar qR_l dqR_l/du_m, the
// face's contribution to the adjoint of variable l of stencil cell n-3+m.
// Under the pressure-based switch (MUSCL, PPM) the face states of every
// variable also read the pressures of the stencil: the tangents 6..11 sit
// on them, and a variable l < 4 leaves its pressure contributions in
// pbc[m * 4 + l][face] (the pressure's own thread adds them to its slot).
template <int D, int RECON>
__global__ void __launch_bounds__(256) k_vjp_recon(Blk b, const double *__restrict__ fbar, double *contrib,
                                                   double *pbc) {
  using D6 = Dual<VJP_STENCIL>;
  using D12 = Dual<2 * VJP_STENCIL>;
  const long nf = b.nfl[D];
  const long tt = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (tt >= 5 * nf) return;
  const int l = (int)(tt / nf);
  const long t = tt - (long)l * nf;
  const double bl = fbar[(long)l * nf + t], br = fbar[(long)(5 + l) * nf + t];
  const bool pb = (RECON == 1 || RECON == 2) && b.pbs[D];
  double o[VJP_STENCIL] = {0, 0, 0, 0, 0, 0}, op[VJP_STENCIL] = {0, 0, 0, 0, 0, 0};
  if (bl != 0.0 || br != 0.0) {
    const VjpFace f = vjp_face<D>(b, t);
    if constexpr (RECON == 1 || RECON == 2) {
      if (pb) {
        typename FaceMath<D12>::Aux ax;
        ax.pb = true;
        D12 u[VJP_STENCIL], pu[VJP_STENCIL], L, R;
#pragma unroll
        for (int m = 0; m < VJP_STENCIL; ++m) {
          u[m] = D12(b.qp[(long)l * b.nc + f.c + (long)(m - 3) * f.s]);
          u[m].d[m] = 1.0;
          pu[m] = D12(b.qp[4L * b.nc + f.c + (long)(m - 3) * f.s]);
          pu[m].d[VJP_STENCIL + m] = 1.0;
        }
        FaceMath<D12>::pb_factors(b, f.n, f.N, &pu[3], 1, ax);
        FaceMath<D12>::template face_states<RECON>(b, D, f.n, f.N, l, &u[3], 1, ax, L, R);
#pragma unroll
        for (int m = 0; m < VJP_STENCIL; ++m) {
          o[m] = bl * L.d[m] + br * R.d[m];
          op[m] = bl * L.d[VJP_STENCIL + m] + br * R.d[VJP_STENCIL + m];
        }
      }
    }
    if (!pb) {
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
  }
  if (l == 4) {
#pragma unroll
    for (int m = 0; m < VJP_STENCIL; ++m) o[m] += op[m];
  } else if (pb) {
#pragma unroll
    for (int m = 0; m < VJP_STENCIL; ++m) pbc[(long)(m * 4 + l) * nf + t] = op[m];
  }
#pragma unroll
  for (int m = 0; m < VJP_STENCIL; ++m) contrib[(long)(m * 5 + l) * nf + t] = o[m];
}

// one thread per padded cell: the sum over directions and over the faces
// whose stencils hold the cell (face n = c + 3 - m of stencil position m);
// c0, c1, c2: the contributions of the three directions (k_vjp_recon), with
// the pressure-based switch's share after the face-state adjoints
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
      // the pressure-based switch's share of the pressure, from variables 0..3
      if ((b.recon == 1 || b.recon == 2) && b.pbs[d]) {
        const double *pbc = cs[d] + (long)(VJP_PER_FACE + 10) * nf[d];
#pragma unroll
        for (int v = 0; v < 4; ++v) acc[4] += pbc[(long)(m * 4 + v) * nf[d] + f];
      }
    }
  }
#pragma unroll
  for (int v = 0; v < 5; ++v) bar[(long)v * b.nc + t] = acc[v];
}

// the boundary face of `face` at (a, bb) (bnormal's face) in the compact
// face order of its direction, and the I face whose normal the slip wall
// reflects with (flow_tangency_col): compact I-face index, or -1 for the j
// and k max faces, whose I face lies in the first ghost row
__device__ __forceinline__ long vjp_bface(const Blk &b, int face, int a, int bb) {
  switch (face) {
    case 0: return fli(b, 0, 1, a, bb);
    case 1: return fli(b, 0, b.imx, a, bb);
    case 2: return fli(b, 1, a, 1, bb);
    case 3: return fli(b, 1, a, b.jmx, bb);
    case 4: return fli(b, 2, a, bb, 1);
    default: return fli(b, 2, a, bb, b.kmx);
  }
}
__device__ __forceinline__ long vjp_rface(const Blk &b, int face, int a, int bb) {
  switch (face) {
    case 0: return fli(b, 0, 1, a, bb);
    case 1: return fli(b, 0, b.imx, a, bb);
    case 2: return fli(b, 0, a, 1, bb);
    case 4: return fli(b, 0, a, bb, 1);
    default: return -1;
  }
}

// the transpose of one face's ghost-cell rule on the adjoint: the ghost
// cells of the face interior hand their adjoints to the cells their values
// were read from, then hold none.  Supported rules (fest_vjp checks): -9
// periodic, -5 no-slip wall with the symmetric density (wall temperature
// in [0, 1]), -6 slip wall, -1..-4 inlets and outlets, -7 pole, -8 far
// field and -11 total-pressure inlet (their Riemann state differentiated in
// dual numbers); interfaces (>= 0, -10) keep their ghost adjoints, which
// belong to the neighbouring block.  With mb0..mb2 (the metric adjoints of
// the three directions, k_vjp_flux), the rules that read face normals (-6,
// -8, -11) add their normals' adjoints there; the slip wall's ghost-row I
// face of the j and k max faces goes to gface[c][(a-1) + (na-1)(bb-1)].
__global__ void __launch_bounds__(256) k_vjp_bc(Blk b, int face, double *bar, double *mb0, double *mb1, double *mb2,
                                                double *gface) {
  const int a = (int)(blockIdx.x * blockDim.x + threadIdx.x) + 1;
  const int bb = (int)blockIdx.y + 1;
  const int na = face_na(b, face), nb = face_nb(b, face);
  if (a > na - 1 || bb > nb - 1) return;
  const FaceBC p = face_params(b, face);
  const int id = p.id;
  if (id >= 0 || id == -10) return;
  const long nc = b.nc;
  const int d = face >> 1;
  double *mb = d == 0 ? mb0 : d == 1 ? mb1 : mb2;  // the boundary face's direction
  const long nfd = d == 0 ? b.nfl[0] : d == 1 ? b.nfl[1] : b.nfl[2];
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
      const int hi = face & 1;
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
      long rf = -1;
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
        rf = vjp_rface(b, face, a, bb);
      }
      double bn[3] = {0, 0, 0}, br[3] = {0, 0, 0};
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
          const long cc = fcell(b, face, g + 1, a, bb);
          const double ux = b.qp[nc + cc], uy = b.qp[2 * nc + cc], uz = b.qp[3 * nc + cc];
          const double rg = rx * w[1] + ry * w[2] + rz * w[3];
          const double nu = nx * ux + ny * uy + nz * uz;
          G(1, g + 1) += w[1] - 2.0 * nx * rg;
          G(2, g + 1) += w[2] - 2.0 * ny * rg;
          G(3, g + 1) += w[3] - 2.0 * nz * rg;
          bn[0] -= 2.0 * rg * ux; bn[1] -= 2.0 * rg * uy; bn[2] -= 2.0 * rg * uz;
          br[0] -= 2.0 * nu * w[1]; br[1] -= 2.0 * nu * w[2]; br[2] -= 2.0 * nu * w[3];
        }
      }
      if (id == -6 && mb) {
        const long bf = vjp_bface(b, face, a, bb);
#pragma unroll
        for (int c = 0; c < 3; ++c) mb[(long)c * nfd + bf] += bn[c];
        if (rf >= 0) {
#pragma unroll
          for (int c = 0; c < 3; ++c) mb0[(long)c * b.nfl[0] + rf] += br[c];
        } else if (gface) {
          const long ng = (long)(na - 1) * (nb - 1), gi = (a - 1) + (long)(na - 1) * (bb - 1);
#pragma unroll
          for (int c = 0; c < 3; ++c) gface[(long)c * ng + gi] = br[c];
        }
      }
      break;
    }
    case -8:      // far field and
    case -11: {   // total-pressure inlet: layer 0 from the Riemann state of
                  // the first interior cell, layers -1 and -2 its copies
      using DT = Dual<8>;
      DT c[5], nrm[3], g[5];
      const long c1 = fcell(b, face, 1, a, bb);
#pragma unroll
      for (int l = 0; l < 5; ++l) {
        c[l] = DT(b.qp[(long)l * nc + c1]);
        c[l].d[l] = 1.0;
      }
      double n3[3];
      bnormal(b, face, a, bb, n3[0], n3[1], n3[2]);
#pragma unroll
      for (int x = 0; x < 3; ++x) {
        nrm[x] = DT(n3[x]);
        nrm[x].d[5 + x] = 1.0;
      }
      riemann_state(b, face, c, nrm, true, id == -11, p.fix[FX_TP], g);
      double gs[5];
#pragma unroll
      for (int l = 0; l < 5; ++l) gs[l] = gb[0][l] + gb[1][l] + gb[2][l];
#pragma unroll
      for (int v = 0; v < 8; ++v) {
        double acc = 0.0;
#pragma unroll
        for (int l = 0; l < 5; ++l) acc += g[l].d[v] * gs[l];
        if (v < 5) G(v, 1) += acc;
        else if (mb) mb[(long)(v - 5) * nfd + vjp_bface(b, face, a, bb)] += acc;
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

// ---- the node coordinates: the face metrics come from the nodes as
// S = (d1 x d2)/2, A = |S|, n = S/A (k_face_metrics), with d1 = P1 - P0 and
// d2 = P2 - P3 of the face's four nodes.  From the adjoints of the stored
// n and A, bar S = bar_A n + (bar_n - n (n . bar_n))/A, then
// bar d1 = (d2 x bar S)/2 and bar d2 = (bar S x d1)/2.

// the face's four nodes (k_face_metrics): p[0..3] = P0, P1, P2, P3
__device__ __forceinline__ void vjp_face_nodes(const Blk &b, int d, int i, int j, int k, long (&p)[4]) {
  p[0] = ndi(b, i, j, k);
  if (d == 0) { p[1] = ndi(b, i, j + 1, k + 1); p[2] = ndi(b, i, j, k + 1); p[3] = ndi(b, i, j + 1, k); }
  else if (d == 1) { p[1] = ndi(b, i + 1, j, k + 1); p[2] = ndi(b, i + 1, j, k); p[3] = ndi(b, i, j, k + 1); }
  else { p[1] = ndi(b, i + 1, j + 1, k); p[2] = ndi(b, i, j + 1, k); p[3] = ndi(b, i + 1, j, k); }
}
// bar d1, bar d2 of one face from the adjoints of its unit normal bn and area bA
__device__ __forceinline__ void vjp_face_diagonals(const Blk &b, int d, int i, int j, int k, const double (&bn)[3],
                                                   double bA, double (&out)[6]) {
  const long nn = (long)(b.imx + 6) * (b.jmx + 6) * (b.kmx + 6);
  const double *X = b.nodes, *Y = X + nn, *Z = Y + nn;
  long p[4];
  vjp_face_nodes(b, d, i, j, k, p);
  const double d1[3] = {X[p[1]] - X[p[0]], Y[p[1]] - Y[p[0]], Z[p[1]] - Z[p[0]]};
  const double d2[3] = {X[p[2]] - X[p[3]], Y[p[2]] - Y[p[3]], Z[p[2]] - Z[p[3]]};
  const double S[3] = {0.5 * (d1[1] * d2[2] - d1[2] * d2[1]), 0.5 * (d1[2] * d2[0] - d1[0] * d2[2]),
                       0.5 * (d1[0] * d2[1] - d1[1] * d2[0])};
  const double A = sqrt(S[0] * S[0] + S[1] * S[1] + S[2] * S[2]);
  double bS[3] = {0, 0, 0};
  if (A != 0.0) {
    const double n[3] = {S[0] / A, S[1] / A, S[2] / A};
    const double nb = n[0] * bn[0] + n[1] * bn[1] + n[2] * bn[2];
#pragma unroll
    for (int c = 0; c < 3; ++c) bS[c] = bA * n[c] + (bn[c] - n[c] * nb) / A;
  }
  out[0] = 0.5 * (d2[1] * bS[2] - d2[2] * bS[1]);
  out[1] = 0.5 * (d2[2] * bS[0] - d2[0] * bS[2]);
  out[2] = 0.5 * (d2[0] * bS[1] - d2[1] * bS[0]);
  out[3] = 0.5 * (bS[1] * d1[2] - bS[2] * d1[1]);
  out[4] = 0.5 * (bS[2] * d1[0] - bS[0] * d1[2]);
  out[5] = 0.5 * (bS[0] * d1[1] - bS[1] * d1[0]);
}

// one thread per face of direction D: dbar[c][face] = bar d1 (c = 0..2),
// bar d2 (c = 3..5) from mbar
template <int D>
__global__ void __launch_bounds__(256) k_vjp_metrics(Blk b, const double *__restrict__ mbar, double *dbar) {
  const long nf = b.nfl[D];
  const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= nf) return;
  const VjpFace f = vjp_face<D>(b, t);
  const double bn[3] = {mbar[t], mbar[nf + t], mbar[2 * nf + t]};
  double out[6];
  vjp_face_diagonals(b, D, f.i, f.j, f.k, bn, mbar[3 * nf + t], out);
#pragma unroll
  for (int c = 0; c < 6; ++c) dbar[(long)c * nf + t] = out[c];
}

// the ghost-row I faces of a slip wall on the j max (face 3) or k max
// (face 5) face: gface holds bar n (3 planes) on entry, bar d1 and bar d2
// (6 planes) on exit; one thread per face-interior column
__global__ void __launch_bounds__(256) k_vjp_metrics_ghost(Blk b, int face, double *gface) {
  const int a = (int)(blockIdx.x * blockDim.x + threadIdx.x) + 1;
  const int bb = (int)blockIdx.y + 1;
  const int na = face_na(b, face), nb = face_nb(b, face);
  if (a > na - 1 || bb > nb - 1) return;
  const long ng = (long)(na - 1) * (nb - 1), gi = (a - 1) + (long)(na - 1) * (bb - 1);
  const double bn[3] = {gface[gi], gface[ng + gi], gface[2 * ng + gi]};
  double out[6];
  if (face == 3) vjp_face_diagonals(b, 0, a, b.jmx, bb, bn, 0.0, out);
  else vjp_face_diagonals(b, 0, a, bb, b.kmx, bn, 0.0, out);
#pragma unroll
  for (int c = 0; c < 6; ++c) gface[(long)c * ng + gi] = out[c];
}

// one thread per padded node: bar of the real nodes (1..imx, 1..jmx,
// 1..kmx) as the sum over the faces that hold the node (P0: -bar d1, P1:
// +bar d1, P2: +bar d2, P3: -bar d2), in a fixed order; ghost nodes get 0.
// g3, g5 (or null): the ghost-row I faces of slip walls on faces 3 and 5,
// whose ghost nodes are the linear extrapolations 2 X(N) - X(N-1)
// (k_ghost_nodes)
__global__ void __launch_bounds__(256) k_vjp_nodes(Blk b, const double *__restrict__ d0, const double *__restrict__ d1,
                                                   const double *__restrict__ d2, const double *__restrict__ g3,
                                                   const double *__restrict__ g5, double *bar_nodes) {
  const long nn = (long)(b.imx + 6) * (b.jmx + 6) * (b.kmx + 6);
  const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= nn) return;
  const int I = (int)(t % (b.imx + 6)) - 2, J = (int)((t / (b.imx + 6)) % (b.jmx + 6)) - 2,
            K = (int)(t / ((long)(b.imx + 6) * (b.jmx + 6))) - 2;
  double acc[3] = {0, 0, 0};
  if (I >= 1 && I <= b.imx && J >= 1 && J <= b.jmx && K >= 1 && K <= b.kmx) {
    const double *db[3] = {d0, d1, d2};
#pragma unroll
    for (int d = 0; d < 3; ++d) {
      const long nf = b.nfl[d];
      const int ie = b.imx - 1 + (d == 0), je = b.jmx - 1 + (d == 1), ke = b.kmx - 1 + (d == 2);
      // (face offset, node role): the faces of direction d holding the node
      int oi[4], oj[4], ok[4];
      oi[0] = 0; oj[0] = 0; ok[0] = 0;  // P0
      if (d == 0) { oi[1] = 0; oj[1] = -1; ok[1] = -1; oi[2] = 0; oj[2] = 0; ok[2] = -1; oi[3] = 0; oj[3] = -1; ok[3] = 0; }
      else if (d == 1) { oi[1] = -1; oj[1] = 0; ok[1] = -1; oi[2] = -1; oj[2] = 0; ok[2] = 0; oi[3] = 0; oj[3] = 0; ok[3] = -1; }
      else { oi[1] = -1; oj[1] = -1; ok[1] = 0; oi[2] = 0; oj[2] = -1; ok[2] = 0; oi[3] = -1; oj[3] = 0; ok[3] = 0; }
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const int fi_ = I + oi[r], fj = J + oj[r], fk = K + ok[r];
        if (fi_ < 1 || fi_ > ie || fj < 1 || fj > je || fk < 1 || fk > ke) continue;
        const long f = fli(b, d, fi_, fj, fk);
        const int c0 = (r == 0 || r == 1) ? 0 : 3;
        const double sg = (r == 1 || r == 2) ? 1.0 : -1.0;
#pragma unroll
        for (int c = 0; c < 3; ++c) acc[c] += sg * db[d][(long)(c0 + c) * nf + f];
      }
    }
    // slip wall on j max: I faces (a, jmx, bb), a = 1..imx-1, bb = 1..kmx-1,
    // nodes P0 (a,jmx,bb), P1 (a,jmx+1,bb+1), P2 (a,jmx,bb+1), P3 (a,jmx+1,bb)
    if (g3 && (J == b.jmx || J == b.jmx - 1)) {
      const int na = b.imx, nb = b.kmx;
      const long ng = (long)(na - 1) * (nb - 1);
      const double w = J == b.jmx ? 2.0 : -1.0;  // weight of a ghost node's row
#pragma unroll
      for (int s = 0; s < 2; ++s) {  // faces bb = K (s 0) and bb = K - 1 (s 1)
        const int a = I, bb = K - s;
        if (a < 1 || a > na - 1 || bb < 1 || bb > nb - 1) continue;
        const long gi = (a - 1) + (long)(na - 1) * (bb - 1);
#pragma unroll
        for (int c = 0; c < 3; ++c) {
          const double bd1 = g3[(long)c * ng + gi], bd2 = g3[(long)(3 + c) * ng + gi];
          if (s == 0) acc[c] += (J == b.jmx ? -bd1 : 0.0) + w * (-bd2);  // P0, P3 (ghost)
          else acc[c] += (J == b.jmx ? bd2 : 0.0) + w * bd1;              // P2, P1 (ghost)
        }
      }
    }
    // slip wall on k max: I faces (a, bb, kmx), a = 1..imx-1, bb = 1..jmx-1,
    // nodes P0 (a,bb,kmx), P1 (a,bb+1,kmx+1), P2 (a,bb,kmx+1), P3 (a,bb+1,kmx)
    if (g5 && (K == b.kmx || K == b.kmx - 1)) {
      const int na = b.imx, nb = b.jmx;
      const long ng = (long)(na - 1) * (nb - 1);
      const double w = K == b.kmx ? 2.0 : -1.0;
#pragma unroll
      for (int s = 0; s < 2; ++s) {  // faces bb = J (s 0) and bb = J - 1 (s 1)
        const int a = I, bb = J - s;
        if (a < 1 || a > na - 1 || bb < 1 || bb > nb - 1) continue;
        const long gi = (a - 1) + (long)(na - 1) * (bb - 1);
#pragma unroll
        for (int c = 0; c < 3; ++c) {
          const double bd1 = g5[(long)c * ng + gi], bd2 = g5[(long)(3 + c) * ng + gi];
          if (s == 0) acc[c] += (K == b.kmx ? -bd1 : 0.0) + w * bd2;  // P0, P2 (ghost)
          else acc[c] += (K == b.kmx ? -bd2 : 0.0) + w * bd1;          // P3, P1 (ghost)
        }
      }
    }
  }
#pragma unroll
  for (int c = 0; c < 3; ++c) bar_nodes[(long)c * nn + t] = acc[c];
}

// ---- the laminar viscous terms (mu_ref != 0, five variables, 3-D)
//
// The viscous flux of a face (FaceMath<T>::visc_face) is a function of 22
// values: u, v, w, T of the cell below (cm) and above (c), the two cells'
// viscosities, and the average of their gradients of u, v, w, T.  T is
// p/(R rho) of the state; the gradients are the Green-Gauss sums of
// k_grad_b (of u, v, w and the temperature plane) and, in the ghost cells
// of boundary faces, the rule of k_grad_bc_b; the viscosities come from
// Sutherland's law of T when the case uses it (src/viscosity.f90), else they
// are constants.  The temperature plane is the state's p/(R rho) at upload
// (src/update.f90:170, taken before the ghost refill): its interior cells
// are functions of the state, its ghost cells constants of the call.
constexpr int VJP_VISC_IN = 22;

// visc_face's arithmetic on the 22 inputs x = (u, v, w, T of cm, u, v, w, T
// of c, mu of cm, mu of c, the averaged gradients [var][x, y, z]); dl the
// centre offset c - cm, id2 = 1/|dl|^2.  v[0..3]: the momentum and energy
// components of the viscous flux.
template <typename T>
__device__ __forceinline__ void vjp_visc_core(const Blk &b, const T (&x)[VJP_VISC_IN], const double (&dl)[3],
                                              double id2, double nx, double ny, double nz, double A, T (&v)[4]) {
  T gr[4][3];
#pragma unroll
  for (int l = 0; l < 4; ++l) {
    const T del = x[4 + l] - x[l];
    const T ncmp = (del - (x[10 + 3 * l] * T(dl[0]) + x[11 + 3 * l] * T(dl[1]) + x[12 + 3 * l] * T(dl[2]))) * T(id2);
#pragma unroll
    for (int c = 0; c < 3; ++c) gr[l][c] = x[10 + 3 * l + c] + ncmp * T(dl[c]);
  }
  const T tm = T(0.5) * (x[8] + x[9]);
  const T div3 = (gr[0][0] + gr[1][1] + gr[2][2]) * T(1.0 / 3.0);
  const T Txx = T(2.0) * tm * (gr[0][0] - div3), Tyy = T(2.0) * tm * (gr[1][1] - div3);
  const T Tzz = T(2.0) * tm * (gr[2][2] - div3);
  const T Txy = tm * (gr[1][0] + gr[0][1]), Txz = tm * (gr[2][0] + gr[0][2]), Tyz = tm * (gr[2][1] + gr[1][2]);
  const T K = tm * T(b.iPr * b.cpR);
  const T uf = T(0.5) * (x[0] + x[4]), vf = T(0.5) * (x[1] + x[5]), wf = T(0.5) * (x[2] + x[6]);
  const T Nx = T(nx), Ny = T(ny), Nz = T(nz), AA = T(A);
  v[0] = (Txx * Nx + Txy * Ny + Txz * Nz) * AA;
  v[1] = (Txy * Nx + Tyy * Ny + Tyz * Nz) * AA;
  v[2] = (Txz * Nx + Tyz * Ny + Tzz * Nz) * AA;
  v[3] = AA * ((Txx * uf + Txy * vf + Txz * wf + K * gr[3][0]) * Nx + (Txy * uf + Tyy * vf + Tyz * wf + K * gr[3][1]) * Ny +
               (Txz * uf + Tyz * vf + Tzz * wf + K * gr[3][2]) * Nz);
}

// one thread per face of direction D: vf[e][face] = d<barF, -v>/dx_e, the
// adjoint of the face's 22 inputs (the residue takes the flux minus v)
template <int D>
__global__ void __launch_bounds__(256) k_vjp_visc(Blk b, const double *__restrict__ barR, double *vf) {
  constexpr int KT = 6;
  using DK = Dual<KT>;
  const long nf = b.nfl[D];
  const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= nf) return;
  const VjpFace f = vjp_face<D>(b, t);
  const int ii = D == 0, jj = D == 1, kk = D == 2;
  const bool below = f.n >= 2, above = f.n <= f.N - 1;
  const long rb = below ? ri(b, f.i - ii, f.j - jj, f.k - kk) : 0, ra = above ? ri(b, f.i, f.j, f.k) : 0;
  double bv[4];
  bool any = false;
#pragma unroll
  for (int l = 0; l < 4; ++l) {
    const double bF = (below ? barR[(long)(l + 1) * b.nres + rb] : 0.0) -
                      (above ? barR[(long)(l + 1) * b.nres + ra] : 0.0);
    bv[l] = -bF;
    any |= bF != 0.0;
  }
  double out[VJP_VISC_IN];
#pragma unroll
  for (int e = 0; e < VJP_VISC_IN; ++e) out[e] = 0.0;
  if (any) {
    const long c = f.c, cm = f.c - f.s, nc = b.nc;
    const long gL = gri(b, f.i - ii, f.j - jj, f.k - kk), gR = gri(b, f.i, f.j, f.k);
    const long fo = fi(b, D, f.i, f.j, f.k);
    const double nx = b.fx[D][fo], ny = b.fy[D][fo], nz = b.fz[D][fo], A = b.fA[D][fo];
    const double dl[3] = {b.cx[c] - b.cx[cm], b.cy[c] - b.cy[cm], b.cz[c] - b.cz[cm]};
    const double id2 = 1.0 / (dl[0] * dl[0] + dl[1] * dl[1] + dl[2] * dl[2]);
    double x0[VJP_VISC_IN];
    const long cc[2] = {cm, c};
#pragma unroll
    for (int s = 0; s < 2; ++s) {
      x0[4 * s + 0] = b.qp[nc + cc[s]];
      x0[4 * s + 1] = b.qp[2 * nc + cc[s]];
      x0[4 * s + 2] = b.qp[3 * nc + cc[s]];
      x0[4 * s + 3] = b.qp[4 * nc + cc[s]] / (b.qp[cc[s]] * b.R);
      x0[8 + s] = b.mu[cc[s]];
    }
#pragma unroll
    for (int l = 0; l < 4; ++l)
#pragma unroll
      for (int q = 0; q < 3; ++q) x0[10 + 3 * l + q] = 0.5 * (b.g[q][l * b.ngr + gL] + b.g[q][l * b.ngr + gR]);
#pragma unroll
    for (int p0 = 0; p0 < VJP_VISC_IN; p0 += KT) {
      DK x[VJP_VISC_IN], v[4];
#pragma unroll
      for (int e = 0; e < VJP_VISC_IN; ++e) {
        x[e] = DK(x0[e]);
        if (e >= p0 && e < p0 + KT) x[e].d[e - p0] = 1.0;
      }
      vjp_visc_core<DK>(b, x, dl, id2, nx, ny, nz, A, v);
#pragma unroll
      for (int e = 0; e < KT; ++e)
        if (p0 + e < VJP_VISC_IN) out[p0 + e] = bv[0] * v[0].d[e] + bv[1] * v[1].d[e] + bv[2] * v[2].d[e] + bv[3] * v[3].d[e];
    }
  }
#pragma unroll
  for (int e = 0; e < VJP_VISC_IN; ++e) vf[(long)e * nf + t] = out[e];
}

// one thread per padded cell: the sum of the viscous faces' adjoints of the
// cell (as c of the face below it, as cm of the face above it, in each
// direction, in a fixed order).  Adds bar u, v, w to bar; writes bt, the
// adjoint of T = p/(R rho) of the state (with Sutherland's law folded in),
// and gb[(3 l + comp) ngr + g], the adjoint of the cell's gradient (half
// of each face's averaged-gradient adjoint), for the cells holding one
// (0..imx, 0..jmx, 0..kmx); zero elsewhere.
__global__ void __launch_bounds__(256) k_vjp_visc_gather(Blk b, const double *__restrict__ v0,
                                                         const double *__restrict__ v1,
                                                         const double *__restrict__ v2, double *bar, double *bt,
                                                         double *gb) {
  const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= b.nc) return;
  const int i = (int)(t % b.pi) - 2, j = (int)((t / b.pi) % b.pj) - 2, k = (int)(t / ((long)b.pi * b.pj)) - 2;
  if (i < 0 || i > b.imx || j < 0 || j > b.jmx || k < 0 || k > b.kmx) {
    bt[t] = 0.0;
    return;
  }
  const double *vs[3] = {v0, v1, v2};
  double acc[17] = {};  // u, v, w, T, mu, then the 12 gradient adjoints (twice the cell's share)
  const int p[3] = {i, j, k};
#pragma unroll
  for (int d = 0; d < 3; ++d) {
    const int N = d == 0 ? b.imx : d == 1 ? b.jmx : b.kmx;
    const int o1 = d == 0 ? 1 : 0, o2 = d == 2 ? 1 : 2;  // the two other directions
    const int a = p[o1], bb = p[o2];
    const int Na = o1 == 0 ? b.imx : b.jmx, Nb = o2 == 1 ? b.jmx : b.kmx;
    if (a < 1 || a > Na - 1 || bb < 1 || bb > Nb - 1) continue;
#pragma unroll
    for (int side = 0; side < 2; ++side) {  // 0: the face below (the cell is its c), 1: above (its cm)
      const int n = p[d] + side;
      if (n < 1 || n > N) continue;
      int q[3] = {i, j, k};
      q[d] = n;
      const long nf = b.nfl[d];
      const long fo = (long)(q[0] - 1) + (long)b.nfi[d] * ((long)(q[1] - 1) + (long)b.nfj[d] * (q[2] - 1));
      const double *v = vs[d];
      const int base = side == 0 ? 4 : 0;
#pragma unroll
      for (int e = 0; e < 4; ++e) acc[e] += v[(long)(base + e) * nf + fo];
      acc[4] += v[(long)(side == 0 ? 9 : 8) * nf + fo];
#pragma unroll
      for (int e = 0; e < 12; ++e) acc[5 + e] += v[(long)(10 + e) * nf + fo];
    }
  }
  bar[b.nc + t] += acc[0];
  bar[2 * b.nc + t] += acc[1];
  bar[3 * b.nc + t] += acc[2];
  double bT = acc[3];
  if (b.mu_suth && acc[4] != 0.0) {
    // mu = mu_ref (T/T_ref)^1.5 (T_ref + S)/(T + S): dmu/dT = mu (1.5/T - 1/(T + S))
    const double T = b.qp[4 * b.nc + t] / (b.qp[t] * b.R);
    const double x = T / b.T_ref;
    const double mu = b.mu_ref * (x * sqrt(x)) * ((b.T_ref + b.S_ref) / (T + b.S_ref));
    bT += acc[4] * mu * (1.5 / T - 1.0 / (T + b.S_ref));
  }
  bt[t] = bT;
  const long g = gri(b, i, j, k);
#pragma unroll
  for (int e = 0; e < 12; ++e) gb[(long)e * b.ngr + g] = 0.5 * acc[5 + e];
}

// the transpose of k_grad_bc_b on one face (boundary ids < 0), one thread
// per face-interior column: the ghost cell's gradient adjoint goes to the
// interior cell's gradient and to the differences of the two cells' values
// it reads, then is cleared (the Green-Gauss sum the rule overwrites is
// unused).  u, v, w of either cell into bar, T of the interior cell into bt
// (the ghost cell's temperature is a constant of the call)
__global__ void __launch_bounds__(256) k_vjp_grad_bc(Blk b, int face, double *bar, double *bt, double *gb) {
  const int a = blockIdx.x * blockDim.x + threadIdx.x + 1, bb = blockIdx.y + 1;
  const int na = face_na(b, face), nb = face_nb(b, face);
  if (a > na - 1 || bb > nb - 1) return;
  const int id = b.bc[face];
  if (id >= 0) return;
  const int d = face >> 1, hi = face & 1;
  const double Tw = b.fix[FX_TWALL][face];
  const int N = d == 0 ? b.imx : d == 1 ? b.jmx : b.kmx;
  const int fp = hi ? N : 1;
  int fi_, fj_, fk_;
  if (d == 0) { fi_ = fp; fj_ = a; fk_ = bb; } else if (d == 1) { fi_ = a; fj_ = fp; fk_ = bb; } else { fi_ = a; fj_ = bb; fk_ = fp; }
  int Ii = fi_, Ij = fj_, Ik = fk_, Gi = fi_, Gj = fj_, Gk = fk_;
  if (hi) { if (d == 0) Ii--; else if (d == 1) Ij--; else Ik--; }
  else { if (d == 0) Gi--; else if (d == 1) Gj--; else Gk--; }
  const long fo = fi(b, 0, fi_, fj_, fk_);  // the rule reads every direction's metrics at the I-face offset
  const double nx = b.fx[d][fo], ny = b.fy[d][fo], nz = b.fz[d][fo], A = b.fA[d][fo];
  const long cI = ci(b, Ii, Ij, Ik), cG = ci(b, Gi, Gj, Gk);
  const double vI = b.vol[cI];
  const double cx = A * nx / vI, cy = A * ny / vI, cz = A * nz / vI;
  const double sig = hi ? -1.0 : 1.0;
  const long gI = gri(b, Ii, Ij, Ik), gG = gri(b, Gi, Gj, Gk);
#pragma unroll
  for (int l = 0; l < 4; ++l) {
    double *gx = gb + (long)(3 * l) * b.ngr, *gy = gx + b.ngr, *gz = gy + b.ngr;
    const double rx = gx[gG], ry = gy[gG], rz = gz[gG];
    gx[gG] = 0.0;
    gy[gG] = 0.0;
    gz[gG] = 0.0;
    const double dot = rx * nx + ry * ny + rz * nz;
    double ix = rx - dot * nx, iy = ry - dot * ny, iz = rz - dot * nz;
    if (l == 3 && id == -5 && (Tw < 1. && Tw >= 0.)) {
      ix -= rx;
      iy -= ry;
      iz -= rz;
    } else {
      const double bdq = sig * (cx * rx + cy * ry + cz * rz);
      if (l < 3) {
        bar[(long)(l + 1) * b.nc + cI] += bdq;
        bar[(long)(l + 1) * b.nc + cG] -= bdq;
      } else {
        bt[cI] += bdq;
      }
    }
    gx[gI] += ix;
    gy[gI] += iy;
    gz[gI] += iz;
  }
}

// one thread per padded cell x: the transpose of the Green-Gauss gradients
// (k_grad_b) onto u, v, w and the temperature of x, from the gradient
// adjoints of x and its six neighbours; then the temperature adjoint of x
// to its density and pressure (T = p/(R rho)).  The temperature plane's
// ghost cells are constants of the call: their share is dropped.
__global__ void __launch_bounds__(256) k_vjp_grad(Blk b, const double *__restrict__ gb, const double *__restrict__ bt,
                                                  double *bar) {
  const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= b.nc) return;
  const int i = (int)(t % b.pi) - 2, j = (int)((t / b.pi) % b.pj) - 2, k = (int)(t / ((long)b.pi * b.pj)) - 2;
  double acc[4] = {0, 0, 0, 0};
  const int p[3] = {i, j, k};
  // (the cell c holding a gradient, sign, the face of c between c and x);
  // c = x: all six faces of x
#pragma unroll
  for (int m = 0; m < 7; ++m) {
    int q[3] = {i, j, k};
    const int d = m == 0 ? 0 : (m - 1) >> 1;
    const int up = m == 0 ? 0 : ((m - 1) & 1) ? 1 : -1;  // c = x + up e_d
    q[d] += up;
    if (q[0] < 0 || q[0] > b.imx || q[1] < 0 || q[1] > b.jmx || q[2] < 0 || q[2] > b.kmx) continue;
    const long c = ci(b, q[0], q[1], q[2]);
    const double ivol2 = 1.0 / (2 * b.vol[c]);
    double co[3] = {0, 0, 0};
    if (m == 0) {
#pragma unroll
      for (int e = 0; e < 3; ++e) {
        const int e0 = e == 0, e1 = e == 1, e2 = e == 2;
        const long lo = fi(b, e, q[0], q[1], q[2]), hi = fi(b, e, q[0] + e0, q[1] + e1, q[2] + e2);
        co[0] += b.fx[e][hi] * b.fA[e][hi] - b.fx[e][lo] * b.fA[e][lo];
        co[1] += b.fy[e][hi] * b.fA[e][hi] - b.fy[e][lo] * b.fA[e][lo];
        co[2] += b.fz[e][hi] * b.fA[e][hi] - b.fz[e][lo] * b.fA[e][lo];
      }
    } else {
      // x = c - e_d: x is c's lower neighbour, through c's lower face (up = +1);
      // x = c + e_d: through c's upper face (up = -1)
      const long f = up > 0 ? fi(b, d, q[0], q[1], q[2]) : fi(b, d, p[0], p[1], p[2]);
      const double sg = up > 0 ? -1.0 : 1.0;
      co[0] = sg * b.fx[d][f] * b.fA[d][f];
      co[1] = sg * b.fy[d][f] * b.fA[d][f];
      co[2] = sg * b.fz[d][f] * b.fA[d][f];
    }
    const long g = gri(b, q[0], q[1], q[2]);
#pragma unroll
    for (int l = 0; l < 4; ++l)
      acc[l] += ivol2 * (co[0] * gb[(long)(3 * l) * b.ngr + g] + co[1] * gb[(long)(3 * l + 1) * b.ngr + g] +
                         co[2] * gb[(long)(3 * l + 2) * b.ngr + g]);
  }
  bar[b.nc + t] += acc[0];
  bar[2 * b.nc + t] += acc[1];
  bar[3 * b.nc + t] += acc[2];
  const bool inner = i >= 1 && i <= b.imx - 1 && j >= 1 && j <= b.jmx - 1 && k >= 1 && k <= b.kmx - 1;
  const double bT = bt[t] + (inner ? acc[3] : 0.0);
  if (bT != 0.0) {
    const double rho = b.qp[t], T = b.qp[4 * b.nc + t] / (rho * b.R);
    bar[t] -= bT * T / rho;
    bar[4 * b.nc + t] += bT / (rho * b.R);
  }
}

}  // namespace fest
