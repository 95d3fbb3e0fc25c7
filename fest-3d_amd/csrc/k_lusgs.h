// k_lusgs.h -- the LU-SGS implicit update of the 'implicit' integrator
// (src/lusgs.f90:134-1196, laminar and SST variants; dispatch
// src/update.f90:215-218).
//
// The reference sweeps the cells in loop order: forward k, j, i ascending,
// backward i, j, k descending.  A cell's forward value needs only its i-, j-
// and k- neighbours' forward values, its backward value only the i+, j+, k+
// neighbours' backward values and its own forward value.  All cells of one
// hyperplane i + j + k = s therefore depend only on hyperplane s - 1 (s + 1
// backward): one launch per hyperplane reproduces the loop order's results
// exactly, with every cell's arithmetic in the reference's operation order.
//
// delQstar and delQ share one array: a cell's backward value overwrites its
// own forward value, which nothing else reads afterwards; the layer of cells
// next to the interior (index 0 and imx, jmx, kmx) stays zero, as the
// reference's arrays are zero there.
#pragma once
// unfused, as the reference: the sweep arithmetic matches it bit for bit
#pragma clang fp contract(off)
#include "blk.h"

namespace fest {

struct LuFace {
  double A, nx, ny, nz, vol, mm, tm, F1;
};

// Flux / SSTFlux (src/lusgs.f90:490-609, 1027-1196): the flux through face
// f of the neighbour state ql moved by du (conservative), against the cell's
// own primitive state qr
template <int NV>
__device__ __forceinline__ void lu_flux(const Blk &b, const double (&ql)[NV], const double (&qr)[NV],
                                        const double (&du)[NV], const LuFace &f, double (&out)[NV]) {
  const double gm = b.gm, R_gas = b.R;
  double U[NV], W[NV];
  U[0] = ql[0];
  U[1] = ql[0] * ql[1];
  U[2] = ql[0] * ql[2];
  U[3] = ql[0] * ql[3];
  U[4] = (ql[4] / (gm - 1.0)) + (0.5 * ql[0] * (ql[1] * ql[1] + ql[2] * ql[2] + ql[3] * ql[3]));
  if constexpr (NV == 7) {
    U[5] = ql[0] * ql[5];
    U[6] = ql[0] * ql[6];
  }
#pragma unroll
  for (int l = 0; l < NV; ++l) U[l] = U[l] + du[l];
  W[0] = U[0];
  W[1] = U[1] / U[0];
  W[2] = U[2] / U[0];
  W[3] = U[3] / U[0];
  W[4] = (gm - 1.0) * (U[4] - (0.5 * (U[1] * U[1] + U[2] * U[2] + U[3] * U[3]) / U[0]));
  if constexpr (NV == 7) {
    W[5] = U[5] / U[0];
    W[6] = U[6] / U[0];
    // a negative k or omega falls back to the neighbour's old value
    W[5] = W[5] + 0.5 * (1. - copysign(1., W[5])) * (ql[5] - W[5]);
    W[6] = W[6] + 0.5 * (1. - copysign(1., W[6])) * (ql[6] - W[6]);
  }
  const double *P = qr;
  const double nx = f.nx, ny = f.ny, nz = f.nz, Area = f.A, Volume = f.vol, mmu = f.mm, tmu = f.tm;
  const double Vn = (W[1] * nx) + (W[2] * ny) + (W[3] * nz);
  const double uface = 0.5 * (W[1] + P[1]), vface = 0.5 * (W[2] + P[2]), wface = 0.5 * (W[3] + P[3]);
  double Fx[NV];
  Fx[0] = W[0] * Vn;
  Fx[1] = (W[1] * Fx[0]) + (W[4] * nx);
  Fx[2] = (W[2] * Fx[0]) + (W[4] * ny);
  Fx[3] = (W[3] * Fx[0]) + (W[4] * nz);
  const double HalfRhoUsquare = 0.5 * W[0] * (W[1] * W[1] + W[2] * W[2] + W[3] * W[3]);
  const double RhoHt = ((gm / (gm - 1.0)) * W[4]) + HalfRhoUsquare;
  Fx[4] = RhoHt * Vn;
  if constexpr (NV == 7) {
    Fx[5] = W[5] * Fx[0];
    Fx[6] = W[6] * Fx[0];
  }
  const double mu = mmu + tmu;
  const double T1 = W[4] / (W[0] * R_gas), T2 = P[4] / (P[0] * R_gas);
  const double dTdx = (T2 - T1) * nx * Area / Volume, dTdy = (T2 - T1) * ny * Area / Volume,
               dTdz = (T2 - T1) * nz * Area / Volume;
  const double dudx = (P[1] - W[1]) * nx * Area / Volume, dudy = (P[1] - W[1]) * ny * Area / Volume,
               dudz = (P[1] - W[1]) * nz * Area / Volume;
  const double dvdx = (P[2] - W[2]) * nx * Area / Volume, dvdy = (P[2] - W[2]) * ny * Area / Volume,
               dvdz = (P[2] - W[2]) * nz * Area / Volume;
  const double dwdx = (P[3] - W[3]) * nx * Area / Volume, dwdy = (P[3] - W[3]) * ny * Area / Volume,
               dwdz = (P[3] - W[3]) * nz * Area / Volume;
  const double trace = dudx + dvdy + dwdz;
  const double Tauxx = 2. * mu * (dudx - trace / 3.0), Tauyy = 2. * mu * (dvdy - trace / 3.0),
               Tauzz = 2. * mu * (dwdz - trace / 3.0);
  const double Tauxy = mu * (dvdx + dudy), Tauxz = mu * (dwdx + dudz), Tauyz = mu * (dwdy + dvdz);
  const double K_heat = (mmu / b.Pr + tmu / b.tPr) * gm * R_gas / (gm - 1.0);
  const double Qx = K_heat * dTdx, Qy = K_heat * dTdy, Qz = K_heat * dTdz;
  Fx[1] = Fx[1] - (Tauxx * nx + Tauxy * ny + Tauxz * nz);
  Fx[2] = Fx[2] - (Tauxy * nx + Tauyy * ny + Tauyz * nz);
  Fx[3] = Fx[3] - (Tauxz * nx + Tauyz * ny + Tauzz * nz);
  Fx[4] = Fx[4] - (Tauxx * uface + Tauxy * vface + Tauxz * wface + Qx) * nx;
  Fx[4] = Fx[4] - (Tauxy * uface + Tauyy * vface + Tauyz * wface + Qy) * ny;
  Fx[4] = Fx[4] - (Tauxz * uface + Tauyz * vface + Tauzz * wface + Qz) * nz;
  if constexpr (NV == 7) {
    const double F1 = f.F1;
    const double dtkdx = (P[5] - W[5]) * nx * Area / Volume, dtkdy = (P[5] - W[5]) * ny * Area / Volume,
                 dtkdz = (P[5] - W[5]) * nz * Area / Volume;
    const double dtwdx = (P[6] - W[6]) * nx * Area / Volume, dtwdy = (P[6] - W[6]) * ny * Area / Volume,
                 dtwdz = (P[6] - W[6]) * nz * Area / Volume;
    const double sk = 0.85 * F1 + 1.0 * (1.0 - F1), sw = 0.5 * F1 + 0.856 * (1.0 - F1);
    Fx[5] = Fx[5] + (mmu + sk * tmu) * (dtkdx * nx + dtkdy * ny + dtkdz * nz);
    Fx[6] = Fx[6] + (mmu + sw * tmu) * (dtwdx * nx + dtwdy * ny + dtwdz * nz);
  }
#pragma unroll
  for (int l = 0; l < NV; ++l) out[l] = Fx[l] * Area;
}

// face n (0..5 = i-, j-, k-, i+, j+, k+) of cell c0 = (i, j, k) with
// neighbour cn: the Flist arrays of src/lusgs.f90:260-312
__device__ __forceinline__ LuFace lu_face(const Blk &b, int n, int i, int j, int k, long c0, long cn) {
  const int d = n % 3;
  const bool up = n >= 3;
  const long g = fi(b, d, i + (d == 0 && up), j + (d == 1 && up), k + (d == 2 && up));
  LuFace f;
  f.A = b.fA[d][g];
  f.nx = up ? b.fx[d][g] : -b.fx[d][g];
  f.ny = up ? b.fy[d][g] : -b.fy[d][g];
  f.nz = up ? b.fz[d][g] : -b.fz[d][g];
  f.vol = 0.5 * (b.vol[cn] + b.vol[c0]);
  // mu holds zeros for inviscid flow and mu_t for laminar flow, the dummy
  // arrays of src/lusgs.f90:115-130
  f.mm = 0.5 * (b.mu[cn] + b.mu[c0]);
  f.tm = 0.5 * (b.mut[cn] + b.mut[c0]);
  f.F1 = 0.5 * (b.f1[cn] + b.f1[c0]);
  return f;
}

// SpectralRadius (src/lusgs.f90:612-660) of face f times its area
template <int NV>
__device__ __forceinline__ double lu_radius(const Blk &b, const double (&ql)[NV], const double (&qr)[NV],
                                            const LuFace &f, long c1, long c2) {
  const double gm = b.gm;
  double Vn = 0.5 * (((ql[1] + qr[1]) * f.nx) + ((ql[2] + qr[2]) * f.ny) + ((ql[3] + qr[3]) * f.nz));
  Vn = fabs(Vn);
  const double a = 0.5 * (sqrt(gm * ql[4] / ql[0]) + sqrt(gm * qr[4] / qr[0]));
  const double rho = 0.5 * (ql[0] + qr[0]);
  const double dx = b.cx[c1] - b.cx[c2], dy = b.cy[c1] - b.cy[c2], dz = b.cz[c1] - b.cz[c2];
  const double dist = sqrt(dx * dx + dy * dy + dz * dz);
  const double vis = gm * (f.mm / b.Pr + f.tm / b.tPr) / (rho * dist);
  return (Vn + a + vis) * f.A;
}

template <int NV>
__device__ __forceinline__ void lu_load(const Blk &b, long c, double (&q)[NV]) {
#pragma unroll
  for (int l = 0; l < NV; ++l) q[l] = b.qp[(long)l * b.nc + c];
}

// Diagonal of every interior cell: the six spectral radii times area
// (lam, 6 planes) and vol/dt + 0.5 sum(lam) (ldg); src/lusgs.f90:321-344.
// Both sweeps use the same values.
template <int NV>
__global__ void __launch_bounds__(256) k_lusgs_diag(Blk b) {
  int i, j;
  if (!flat_ij(b.imx - 1, b.jmx - 1, i, j, (int)blockIdx.x)) return;
  i += 1;
  j += 1;
  const int k = blockIdx.y + 1;
  const long c0 = ci(b, i, j, k), r = ri(b, i, j, k);
  double Q0[NV], Qn[NV];
  lu_load<NV>(b, c0, Q0);
  double L[6];
#pragma unroll
  for (int n = 0; n < 6; ++n) {
    const long cn = c0 + (n < 3 ? -cstride(b, n) : cstride(b, n - 3));
    lu_load<NV>(b, cn, Qn);
    const LuFace f = lu_face(b, n, i, j, k, c0, cn);
    L[n] = lu_radius<NV>(b, Qn, Q0, f, cn, c0);
    b.lam[(long)n * b.nres + r] = L[n];
  }
  const double sumL = L[0] + L[1] + L[2] + L[3] + L[4] + L[5];
  b.ldg[r] = (b.vol[c0] / b.dt[r]) + 0.5 * sumL;
}

// One hyperplane i + j + k = s of the forward (BACK = 0, delQstar) or
// backward (BACK = 1, delQ) sweep.  Threads run over the (j, k) pairs;
// a pair whose i falls outside 1..imx-1 has no cell on this hyperplane.
template <int NV, int BACK>
__global__ void __launch_bounds__(256) k_lusgs_sweep(Blk b, int s) {
  const int nj = b.jmx - 1, nk = b.kmx - 1;
  const int t = (int)(blockIdx.x * blockDim.x + threadIdx.x);
  if (t >= nj * nk) return;
  const int k = t / nj + 1, j = t - (k - 1) * nj + 1;
  const int i = s - j - k;
  if (i < 1 || i > b.imx - 1) return;
  const long c0 = ci(b, i, j, k), r = ri(b, i, j, k);
  const long nc = b.nc;
  double Q0[NV], Qn[NV], dn[NV], zero[NV], New[NV], Old[NV], S[NV];
  lu_load<NV>(b, c0, Q0);
#pragma unroll
  for (int l = 0; l < NV; ++l) zero[l] = 0.0;
#pragma unroll
  for (int d = 0; d < 3; ++d) {
    const int n = BACK ? d + 3 : d;
    const long cn = BACK ? c0 + cstride(b, d) : c0 - cstride(b, d);
    lu_load<NV>(b, cn, Qn);
#pragma unroll
    for (int l = 0; l < NV; ++l) dn[l] = b.dq[(long)l * nc + cn];
    const LuFace f = lu_face(b, n, i, j, k, c0, cn);
    lu_flux<NV>(b, Qn, Q0, dn, f, New);
    lu_flux<NV>(b, Qn, Q0, zero, f, Old);
    const double lam = b.lam[(long)n * b.nres + r];
#pragma unroll
    for (int l = 0; l < NV; ++l) {
      const double term = (New[l] - Old[l]) - lam * dn[l];
      S[l] = d == 0 ? term : S[l] + term;
    }
  }
  double D[NV];
  const double dg = b.ldg[r];
#pragma unroll
  for (int l = 0; l < NV; ++l) D[l] = dg;
  if constexpr (NV == 7) {
    const double F1 = b.f1[c0], vol = b.vol[c0];
    const double beta = F1 * 0.075 + (1.0 - F1) * 0.0828;
    D[5] = (D[5] + (0.09 * Q0[6]) * vol);
    D[6] = (D[6] + 2.0 * beta * Q0[6] * vol);
  }
#pragma unroll
  for (int l = 0; l < NV; ++l) {
    double *p = b.dq + (long)l * nc + c0;
    if (!BACK) *p = (-b.res[(long)l * b.nres + r] - 0.5 * S[l]) / D[l];
    else *p = *p - 0.5 * S[l] / D[l];
  }
}

// qp += delQ in conservative variables (src/lusgs.f90:456-486, 996-1022)
template <int NV>
__global__ void __launch_bounds__(256) k_lusgs_apply(Blk b) {
  int i, j;
  if (!flat_ij(b.imx - 1, b.jmx - 1, i, j, (int)blockIdx.x)) return;
  i += 1;
  j += 1;
  const int k = blockIdx.y + 1;
  const long c = ci(b, i, j, k), nc = b.nc;
  const double gm = b.gm;
  double q[NV], U[NV];
  lu_load<NV>(b, c, q);
  U[0] = q[0];
  U[1] = q[0] * q[1];
  U[2] = q[0] * q[2];
  U[3] = q[0] * q[3];
  U[4] = (q[4] / (gm - 1.0)) + (0.5 * q[0] * (q[1] * q[1] + q[2] * q[2] + q[3] * q[3]));
  if constexpr (NV == 7) {
    U[5] = q[0] * q[5];
    U[6] = q[0] * q[6];
  }
#pragma unroll
  for (int l = 0; l < NV; ++l) U[l] = U[l] + b.dq[(long)l * nc + c];
  b.qp[c] = U[0];
  b.qp[nc + c] = U[1] / U[0];
  b.qp[2 * nc + c] = U[2] / U[0];
  b.qp[3 * nc + c] = U[3] / U[0];
  b.qp[4 * nc + c] = (gm - 1.0) * (U[4] - (0.5 * (U[1] * U[1] + U[2] * U[2] + U[3] * U[3]) / U[0]));
  if constexpr (NV == 7) {
    if (U[5] > 0) b.qp[5 * nc + c] = U[5] / U[0];
    if (U[6] > 0) b.qp[6 * nc + c] = U[6] / U[0];
  }
}

}  // namespace fest
