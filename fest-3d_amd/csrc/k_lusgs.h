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
// reference's arrays are zero there.  The sweeps read only hyperplane-major
// copies (Sk), so qp is updated after the backward sweep.
#pragma once
// unfused, as the reference: the sweep arithmetic matches it bit for bit
#pragma clang fp contract(off)
#include "blk.h"

namespace fest {

struct LuFace {
  double A, nx, ny, nz, vol, mm, tm, F1;
};

// Flux / SAFlux / SSTFlux (src/lusgs.f90:490-609, 2104-2259, 1027-1196): the flux through face
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
  if constexpr (NV == 6) U[5] = ql[0] * ql[5];
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
    if (b.kkl) {
      // KKLFlux floors k and kL (src/lusgs.f90:1605-1606)
      W[5] = fmx(W[5], 1e-8);
      W[6] = fmx(W[6], 1e-8);
    } else {
      // a negative k or omega falls back to the neighbour's old value
      W[5] = W[5] + 0.5 * (1. - copysign(1., W[5])) * (ql[5] - W[5]);
      W[6] = W[6] + 0.5 * (1. - copysign(1., W[6])) * (ql[6] - W[6]);
    }
  }
  if constexpr (NV == 6) W[5] = fmx(U[5] / U[0], 1e-8);
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
  if constexpr (NV == 6) Fx[5] = W[5] * Fx[0];
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
    // k-kL: F1 stays 0 and sigma_w2 is 1, so sk = sw = 1 (src/lusgs.f90:1671-1672)
    const double sk = 0.85 * F1 + 1.0 * (1.0 - F1), sw = 0.5 * F1 + b.sigma_w2 * (1.0 - F1);
    Fx[5] = Fx[5] + (mmu + sk * tmu) * (dtkdx * nx + dtkdy * ny + dtkdz * nz);
    Fx[6] = Fx[6] + (mmu + sw * tmu) * (dtwdx * nx + dtwdy * ny + dtwdz * nz);
  }
  if constexpr (NV == 6) {
    // nu-tilde diffusion with 0.25 (rho_L + rho_R)(nu_L + nu_R), added
    const double muCap = 0.25 * (P[0] + W[0]) * (P[5] + W[5]);
    const double dtvdx = (P[5] - W[5]) * nx * Area / Volume, dtvdy = (P[5] - W[5]) * ny * Area / Volume,
                 dtvdz = (P[5] - W[5]) * nz * Area / Volume;
    Fx[5] = Fx[5] + (mmu + muCap) * (dtvdx * nx + dtvdy * ny + dtvdz * nz) / Sa::sigma;
  }
#pragma unroll
  for (int l = 0; l < NV; ++l) out[l] = Fx[l] * Area;
}

// face n (0..5 = i-, j-, k-, i+, j+, k+) of cell c0 = (i, j, k) with
// neighbour cn, from the natural-layout arrays: the Flist arrays of
// src/lusgs.f90:260-312
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

// SpectralRadius (src/lusgs.f90:612-660) of face f times its area; c1 the
// neighbour's centre, c2 the cell's
template <int NV>
__device__ __forceinline__ double lu_radius(const Blk &b, const double (&ql)[NV], const double (&qr)[NV],
                                            const LuFace &f, double c1x, double c1y, double c1z, double c2x,
                                            double c2y, double c2z) {
  const double gm = b.gm;
  double Vn = 0.5 * (((ql[1] + qr[1]) * f.nx) + ((ql[2] + qr[2]) * f.ny) + ((ql[3] + qr[3]) * f.nz));
  Vn = fabs(Vn);
  const double a = 0.5 * (sqrt(gm * ql[4] / ql[0]) + sqrt(gm * qr[4] / qr[0]));
  const double rho = 0.5 * (ql[0] + qr[0]);
  const double dx = c1x - c2x, dy = c1y - c2y, dz = c1z - c2z;
  const double dist = sqrt(dx * dx + dy * dy + dz * dz);
  const double vis = gm * (f.mm / b.Pr + f.tm / b.tPr) / (rho * dist);
  return (Vn + a + vis) * f.A;
}

template <int NV>
__device__ __forceinline__ void lu_load(const Blk &b, long c, double (&q)[NV]) {
#pragma unroll
  for (int l = 0; l < NV; ++l) q[l] = b.qp[(long)l * b.nc + c];
}

// The SA diagonal (src/lusgs.f90:1882-1914, the same backward): D plus the
// linearised SA production and destruction, one value for all six
// equations; gi the cell's gradient index
__device__ __forceinline__ double sa_diagonal(const Blk &b, long c0, long gi, double rho, double q6, double vol,
                                              double D) {
  const double *gx = b.g[0], *gy = b.g[1], *gz = b.g[2];
  const long ngr = b.ngr;
  const double Omega = sqrt(sqd(gy[2 * ngr + gi] - gz[ngr + gi]) + sqd(gz[gi] - gx[2 * ngr + gi]) +
                            sqd(gx[ngr + gi] - gy[gi]));
  const double dist_i = b.dist[c0], dist_i_2 = dist_i * dist_i, k2 = Sa::kappa * Sa::kappa;
  const double nu = b.mu[c0] / rho;
  const double Ji = q6 / nu, Ji_2 = Ji * Ji, Ji_3 = Ji_2 * Ji;
  const double fv1 = Ji_3 / (Ji_3 + Sa::cv1_3);
  const double fv2 = 1.0 - Ji / (1.0 + (Ji * fv1));
  const double inv_k2_d2 = 1.0 / (k2 * dist_i_2);
  const double Shat = fmx(Omega + q6 * fv2 * inv_k2_d2, 1.0e-10);
  const double inv_Shat = 1.0 / Shat;
  const double jc = Ji_3 + Sa::cv1_3, jf = 1.0 + Ji * fv1;
  const double dfv1 = 3.0 * Ji_2 * Sa::cv1_3 / (nu * (jc * jc));
  const double dfv2 = -((1.0 / nu) - Ji_2 * dfv1) / (jf * jf);
  const double dShat = (fv2 + q6 * dfv2) * inv_k2_d2;
  D = D - Sa::cb1 * (q6 * dShat + Shat) * vol;
  const double r = fmn(q6 * inv_Shat * inv_k2_d2, 10.0);
  const double r5 = (((r * r) * r) * r) * r;
  const double g = r + Sa::cw2 * ((r5 * r) - r);
  const double g_6 = ((((g * g) * g) * g) * g) * g;
  const double glim = pow((1.0 + Sa::cw3_6) / (g_6 + Sa::cw3_6), 1.0 / 6.0);
  const double fw = g * glim;
  const double dr = (Shat - q6 * dShat) * inv_Shat * inv_Shat * inv_k2_d2;
  const double dg = dr * (1.0 + Sa::cw2 * (6.0 * r5 - 1.0));
  const double dfw = dg * glim * (1.0 - g_6 / (g_6 + Sa::cw3_6));
  return D + Sa::cw1 * (dfw * q6 + 2 * fw) * q6 / dist_i_2 * vol;
}

// Hyperplane-major ("skewed") storage of everything the sweeps read.  Cell
// (i, j, k), 0 <= i <= imx, 0 <= j <= jmx, 0 <= k <= kmx, lies on hyperplane
// s = i + j + k at position r = k sk_j + j; component c of it is stored at
// (s NC + c) sk_p + r.  One hyperplane's components are one contiguous block
// (few TLB pages per launch), its cells are contiguous per component, and
// the i-, j-, k- neighbours of a cell sit on hyperplane s - 1 at r, r - 1,
// r - sk_j: every sweep load reads whole cache lines.
template <int NV>
struct Sk {
  static constexpr int Q = 0;              // state (cells 0..n+1)
  static constexpr int MU = NV;            // mu, mu_t, F1 (cells 0..n+1)
  static constexpr int MUT = NV + 1;
  static constexpr int F1 = NV + 2;
  static constexpr int RES = NV + 3;       // residue (interior)
  static constexpr int VOL = 2 * NV + 3;   // volume (cells 0..n+1)
  static constexpr int FACE = 2 * NV + 4;  // lower face d: A, nx, ny, nz at FACE + 4 d (cells 1..n+1)
  static constexpr int DQ = 2 * NV + 16;   // delQstar / delQ (zero outside the interior)
  static constexpr int DG = 3 * NV + 16;   // diagonal (interior; SST: 3 values)
  static constexpr int LAM = 3 * NV + 19;  // radius x area of the lower face d (cells 1..n+1)
  static constexpr int OLD = 3 * NV + 22;  // Old flux of face n = 0..5, NV each (interior)
  static constexpr int NC = 9 * NV + 22;
};
__host__ __device__ inline int sk_ncomp(int nv) { return 9 * nv + 22; }

__device__ __forceinline__ const double *sk_at(const Blk &b, int NC, int s, int c) {
  return b.sk + ((long)s * NC + c) * b.sk_p;
}

// Copies n natural-layout planes into skewed components c0.. for the cells
// lo <= i, j, k <= imx-1+hi (layout 0: padded cell arrays, 1: interior
// arrays of the residue's extent, 2 + d: face arrays of direction d, the
// lower face of each cell).  32 x 32 (i, j) tiles of one k go through LDS:
// rows are read whole, and each diagonal i + j = const of the tile, a run
// of consecutive positions on one hyperplane, is written whole.
struct SkewIn {
  const double *p[8];
  int n, c0, layout, lo, hi;
};
__global__ void __launch_bounds__(256) k_to_skew(Blk b, SkewIn a) {
  __shared__ double tile[32][33];
  const int NC = sk_ncomp(b.nv);
  const int ilo = a.lo, ihi = b.imx - 1 + a.hi, jhi = b.jmx - 1 + a.hi;
  const int ntx = (ihi - ilo + 32) / 32;
  const int i0 = ilo + ((int)blockIdx.x % ntx) * 32, j0 = ilo + ((int)blockIdx.x / ntx) * 32;
  const int k = ilo + (int)blockIdx.y;
  const int tid = threadIdx.x;
  for (int c = 0; c < a.n; ++c) {
    const double *src = a.p[c];
#pragma unroll
    for (int m = 0; m < 4; ++m) {
      const int ii = tid & 31, jj = (tid >> 5) + 8 * m, i = i0 + ii, j = j0 + jj;
      if (i <= ihi && j <= jhi) {
        long o = a.layout == 0 ? ci(b, i, j, k) : a.layout == 1 ? ri(b, i, j, k) : fi(b, a.layout - 2, i, j, k);
        tile[jj][ii] = src[o];
      }
    }
    __syncthreads();
#pragma unroll
    for (int m = 0; m < 4; ++m) {
      // position e of the tile in diagonal order: diagonals 0..31 first,
      // then 32..62 as the mirror image of 30..0
      const int e = tid + 256 * m, em = e < 528 ? e : 1023 - e;
      int sg = (int)((sqrtf((float)(8 * em + 1)) - 1.0f) * 0.5f);
      // exact whatever sqrtf rounds to: sg (sg + 1) / 2 <= em < (sg + 1) (sg + 2) / 2
      if ((sg + 1) * (sg + 2) / 2 <= em) ++sg;
      if (sg * (sg + 1) / 2 > em) --sg;
      const int pos = em - sg * (sg + 1) / 2;
      int jj = pos, ii = sg - pos;
      if (e >= 528) { jj = 31 - jj; ii = 31 - ii; }
      const int i = i0 + ii, j = j0 + jj;
      if (i <= ihi && j <= jhi) {
        double *dst = b.sk + ((long)(i + j + k) * NC + a.c0 + c) * b.sk_p + (long)k * b.sk_j + j;
        *dst = tile[jj][ii];
      }
    }
    __syncthreads();
  }
}

// Everything of a cell the sweeps use that does not depend on them, written
// into hyperplane-major storage: the radius x area of the lower faces
// (src/lusgs.f90:612-660; a face's radius is the same seen from either of
// its cells -- the normal flips sign, every other operand pairs up
// commutatively -- so the backward sweep reads a cell's upper faces from the
// neighbours), the six "Old" fluxes of the neighbours' unmoved states
// (src/lusgs.f90:314-319, 424-429) and the diagonal vol/dt + 0.5 sum of the
// six radii (src/lusgs.f90:329-345).  16 x 16 (i, j) tiles of one k over
// the cells 1..n+1 read the stencil whole; each value goes through LDS and
// out along the tile's diagonals, runs of consecutive positions of one
// hyperplane.  Values of cells outside the interior are never read.
template <int NV>
__global__ void __launch_bounds__(256) k_lusgs_prep(Blk b) {
  using C = Sk<NV>;
  constexpr int NT = NV + 1;  // values staged per round: a face's Old flux and radius
  __shared__ double tile[NT][16][17];
  const int ntx = (b.imx + 15) / 16;
  const int i0 = 1 + ((int)blockIdx.x % ntx) * 16, j0 = 1 + ((int)blockIdx.x / ntx) * 16;
  const int k = 1 + (int)blockIdx.y;
  const int tid = threadIdx.x;
  const int ii = tid & 15, jj = tid >> 4;
  const int i = min(i0 + ii, b.imx), j = min(j0 + jj, b.jmx);  // clamped lanes compute, never write
  const bool inner = i <= b.imx - 1 && j <= b.jmx - 1 && k <= b.kmx - 1;
  // where the tile's diagonal order puts this thread on the way out
  int oi, oj;
  {
    const int e = tid, em = e < 136 ? e : 255 - e;
    int sg = (int)((sqrtf((float)(8 * em + 1)) - 1.0f) * 0.5f);
    if ((sg + 1) * (sg + 2) / 2 <= em) ++sg;
    if (sg * (sg + 1) / 2 > em) --sg;
    oj = em - sg * (sg + 1) / 2;
    oi = sg - oj;
    if (e >= 136) {
      oj = 15 - oj;
      oi = 15 - oi;
    }
  }
  const bool out_ok = i0 + oi <= b.imx && j0 + oj <= b.jmx;
  double *const out_base =
      b.sk + (long)(i0 + oi + j0 + oj + k) * C::NC * b.sk_p + (long)k * b.sk_j + (j0 + oj);
  const long c0 = ci(b, i, j, k);
  double Q0[NV], Qn[NV], zero[NV], Old[NV], L[6];
  lu_load<NV>(b, c0, Q0);
#pragma unroll
  for (int l = 0; l < NV; ++l) zero[l] = 0.0;
#pragma unroll 1
  for (int n = 0; n < 6; ++n) {
    const int d = n % 3;
    const long cn = n < 3 ? c0 - cstride(b, d) : c0 + cstride(b, d);
    lu_load<NV>(b, cn, Qn);
    const LuFace f = lu_face(b, n, i, j, k, c0, cn);
    L[n] = lu_radius<NV>(b, Qn, Q0, f, b.cx[cn], b.cy[cn], b.cz[cn], b.cx[c0], b.cy[c0], b.cz[c0]);
    lu_flux<NV>(b, Qn, Q0, zero, f, Old);
#pragma unroll
    for (int l = 0; l < NV; ++l) tile[l][jj][ii] = Old[l];
    tile[NV][jj][ii] = L[n];
    __syncthreads();
    if (out_ok) {
#pragma unroll
      for (int l = 0; l < NV; ++l) out_base[(long)(C::OLD + n * NV + l) * b.sk_p] = tile[l][oj][oi];
      if (n < 3) out_base[(long)(C::LAM + n) * b.sk_p] = tile[NV][oj][oi];
    }
    __syncthreads();
  }
  double dg = 0.0, d5 = 0.0, d6 = 0.0;
  if (inner) {
    const double sumL = L[0] + L[1] + L[2] + L[3] + L[4] + L[5];
    const double vol = b.vol[c0];
    dg = (vol / b.dt[ri(b, i, j, k)]) + 0.5 * sumL;
    if constexpr (NV == 6) dg = sa_diagonal(b, c0, gri(b, i, j, k), Q0[0], Q0[5], vol, dg);
    if constexpr (NV == 7) {
      if (b.kkl) {
        // the k-kL source terms (src/lusgs.f90:1359-1361)
        const double mu = b.mu[c0], d2 = b.dist[c0] * b.dist[c0];
        d5 = dg + (2.5 * (pow(0.09, 0.75)) * Q0[0] * (pow(Q0[5], 1.5)) * vol / Q0[6]);
        d5 = d5 + (2 * mu * vol / d2);
        d6 = dg + (6 * mu * vol / d2);
      } else {
        const double F1 = b.f1[c0], w = Q0[6];
        const double beta = F1 * 0.075 + (1.0 - F1) * 0.0828;
        d5 = (dg + (0.09 * w) * vol);
        d6 = (dg + 2.0 * beta * w * vol);
      }
    }
  }
  tile[0][jj][ii] = dg;
  tile[1][jj][ii] = d5;
  tile[2][jj][ii] = d6;
  __syncthreads();
  if (out_ok) {
    out_base[(long)C::DG * b.sk_p] = tile[0][oj][oi];
    if constexpr (NV == 7) {
      out_base[(long)(C::DG + 1) * b.sk_p] = tile[1][oj][oi];
      out_base[(long)(C::DG + 2) * b.sk_p] = tile[2][oj][oi];
    }
  }
}

// One hyperplane i + j + k = s of the forward (BACK = 0, delQstar) or
// backward (BACK = 1, delQ) sweep, four lanes per cell: lane d = 0, 1, 2
// takes face d (the cell's lower face forward, its upper face backward):
// the flux of the neighbour's state moved by its sweep value (the "New"
// flux of src/lusgs.f90:308-310, 418-420) minus the Old flux and the radius
// term; lane 3 takes the diagonal and the right-hand side, sums the three
// face terms in the reference's order and writes the cell.  Cells run over
// the (j, k) pairs; a pair whose i falls outside 1..imx-1 has no cell on
// this hyperplane.
template <int NV, int BACK>
__global__ void __launch_bounds__(256) k_lusgs_sweep(Blk b, int s) {
  using C = Sk<NV>;
  constexpr int NC = C::NC;
  const int nj = b.jmx - 1, nk = b.kmx - 1;
  const int t = (int)(blockIdx.x * blockDim.x + threadIdx.x);
  const int cell = t >> 2, q = t & 3;
  // the four lanes of a cell leave together
  if (cell >= nj * nk) return;
  const int k = cell / nj + 1, j = cell - (k - 1) * nj + 1;
  const int i = s - j - k;
  if (i < 1 || i > b.imx - 1) return;
  const long P = b.sk_p;
  const int r = k * b.sk_j + j;
  double *const W0 = b.sk + (long)s * NC * P;  // this hyperplane
  const double *const S0 = W0;
  double v[NV];  // lanes 0..2: face term; lane 3: right-hand side
  double D[NV];
  if (q < 3) {
    const int d = q;
    const int dr = d == 0 ? 0 : d == 1 ? 1 : b.sk_j;
    const int rn = BACK ? r + dr : r - dr;
    const double *const S1 = sk_at(b, NC, BACK ? s + 1 : s - 1, 0);
    // the lower face d of the cell forward, of the neighbour backward
    const double *const SF = BACK ? S1 : S0;
    const int rf = BACK ? rn : r;
    double Q0[NV], Qn[NV], dn[NV], New[NV];
#pragma unroll
    for (int l = 0; l < NV; ++l) {
      Q0[l] = S0[(C::Q + l) * P + r];
      Qn[l] = S1[(C::Q + l) * P + rn];
      dn[l] = S1[(C::DQ + l) * P + rn];
    }
    LuFace f;
    f.A = SF[(C::FACE + 4 * d) * P + rf];
    // outward normal of this cell: the lower face's own normal flipped forward
    const double nx = SF[(C::FACE + 4 * d + 1) * P + rf], ny = SF[(C::FACE + 4 * d + 2) * P + rf],
                 nz = SF[(C::FACE + 4 * d + 3) * P + rf];
    f.nx = BACK ? nx : -nx;
    f.ny = BACK ? ny : -ny;
    f.nz = BACK ? nz : -nz;
    f.vol = 0.5 * (S1[C::VOL * P + rn] + S0[C::VOL * P + r]);
    f.mm = 0.5 * (S1[C::MU * P + rn] + S0[C::MU * P + r]);
    f.tm = 0.5 * (S1[C::MUT * P + rn] + S0[C::MUT * P + r]);
    f.F1 = 0.5 * (S1[C::F1 * P + rn] + S0[C::F1 * P + r]);
    lu_flux<NV>(b, Qn, Q0, dn, f, New);
    const double lam = SF[(C::LAM + d) * P + rf];
#pragma unroll
    for (int l = 0; l < NV; ++l) v[l] = (New[l] - S0[(C::OLD + (BACK * 3 + d) * NV + l) * P + r]) - lam * dn[l];
  } else {
    const double dg = S0[C::DG * P + r];
#pragma unroll
    for (int l = 0; l < NV; ++l) D[l] = dg;
    if constexpr (NV == 7) {
      D[5] = S0[(C::DG + 1) * P + r];
      D[6] = S0[(C::DG + 2) * P + r];
    }
#pragma unroll
    for (int l = 0; l < NV; ++l) v[l] = BACK ? S0[(C::DQ + l) * P + r] : -S0[(C::RES + l) * P + r];
  }
  double T0[NV], T1[NV], T2[NV];
#pragma unroll
  for (int l = 0; l < NV; ++l) {
    T0[l] = __shfl(v[l], 0, 4);
    T1[l] = __shfl(v[l], 1, 4);
    T2[l] = __shfl(v[l], 2, 4);
  }
  if (q != 3) return;
#pragma unroll
  for (int l = 0; l < NV; ++l) {
    const double S = (T0[l] + T1[l]) + T2[l];
    W0[(C::DQ + l) * P + r] = BACK ? v[l] - 0.5 * S / D[l] : (v[l] - 0.5 * S) / D[l];
  }
}

// qp += delQ in conservative variables (src/lusgs.f90:456-486, 996-1022)
template <int NV>
__global__ void __launch_bounds__(256) k_lusgs_apply(Blk b) {
  using C = Sk<NV>;
  int i, j;
  if (!flat_ij(b.imx - 1, b.jmx - 1, i, j, (int)blockIdx.x)) return;
  i += 1;
  j += 1;
  const int k = blockIdx.y + 1;
  const long c = ci(b, i, j, k), nc = b.nc;
  const double *dq = sk_at(b, C::NC, i + j + k, C::DQ) + (long)k * b.sk_j + j;
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
  if constexpr (NV == 6) U[5] = q[0] * q[5];
#pragma unroll
  for (int l = 0; l < NV; ++l) U[l] = U[l] + dq[(long)l * b.sk_p];
  b.qp[c] = U[0];
  b.qp[nc + c] = U[1] / U[0];
  b.qp[2 * nc + c] = U[2] / U[0];
  b.qp[3 * nc + c] = U[3] / U[0];
  b.qp[4 * nc + c] = (gm - 1.0) * (U[4] - (0.5 * (U[1] * U[1] + U[2] * U[2] + U[3] * U[3]) / U[0]));
  if constexpr (NV == 7) {
    if (b.kkl) {  // src/lusgs.f90:1502-1505
      b.qp[5 * nc + c] = fmx(U[5] / U[0], 1.e-8);
      b.qp[6 * nc + c] = fmx(U[6] / U[0], 1.e-8);
    } else {
      if (U[5] > 0) b.qp[5 * nc + c] = U[5] / U[0];
      if (U[6] > 0) b.qp[6 * nc + c] = U[6] / U[0];
    }
  }
  if constexpr (NV == 6) b.qp[5 * nc + c] = fmx(U[5] / U[0], 1.e-8);  // src/lusgs.f90:2096-2097
}

}  // namespace fest
