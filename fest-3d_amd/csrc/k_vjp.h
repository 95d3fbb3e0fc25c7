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

}  // namespace fest
