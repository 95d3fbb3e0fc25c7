// k_flux.h -- one thread per face of one sweep direction: reconstruct the
// left/right states of every variable from the cell stencil along the sweep
// (MUSCL, WENO5, PPM or first order), apply the boundary-face rules, then the
// inviscid flux and the laminar/SST viscous flux, and store F, G or H.
//
// Face n lies between cells n-1 and n of the sweep.  Its left state is the
// upper-face state of cell n-1 and its right state the lower-face state of
// cell n (src/face/state/muscl.f90:190-193, weno.f90:74-87, ppm.f90:47-100);
// the boundary rules follow src/boundary/boundary_state_reconstruction.f90
// (first/last interior cell MUSCL when ppm_flag, ghost or averaged states on
// the boundary face).
#pragma once
// the residual kernels let multiply-adds fuse into FMA: results differ from
// the unfused reference order by rounding only, far inside the 1e-12 parity
// bound, and the FP64 instruction count, which bounds these kernels, drops
#pragma clang fp contract(fast)
#include "blk.h"

namespace fest {

// ---- reconstruction of one cell from its stencil u[-2..2] (u points at the cell)
__device__ __forceinline__ void muscl_cell(const double *u, long s, int sw, double &up, double &lo) {
  const double alpha = 2. / 3., phi = 1.0, kappa = 1. / 3., eps = 1e-14;
  double fd = u[s] - u[0];
  double bd = u[0] - u[-s];
  double r = fd / (bd + fsgn(eps, bd));
  double psi1 = fmx(0., fmn(fmn(2 * r, alpha * (r - 1.0) + 1.0), 2.));
  r = bd / (fd + fsgn(eps, fd));
  double psi2 = fmx(0., fmn(fmn(2 * r, alpha * (r - 1.0) + 1.0), 2.));
  psi1 = (1 - (1 - psi1) * sw);
  psi2 = (1 - (1 - psi2) * sw);
  up = u[0] + 0.25 * phi * (((1. - kappa) * psi1 * bd) + ((1. + kappa) * psi2 * fd));
  lo = u[0] - 0.25 * phi * (((1. + kappa) * psi1 * bd) + ((1. - kappa) * psi2 * fd));
}

// boundary-cell MUSCL of boundary_state_reconstruction.f90:100-117 (no eps)
__device__ __forceinline__ void bmuscl_cell(const double *u, long s, int sw, double &up, double &lo) {
  const double phi = 1.0, kappa = 1. / 3.;
  double fd = u[s] - u[0];
  double bd = u[0] - u[-s];
  double r = fd / bd;
  double psi1 = fmx(0., fmn(fmn(2 * r, (2 + r) / 3.), 2.));
  psi1 = (1 - (1 - psi1) * sw);
  r = bd / fd;
  double psi2 = fmx(0., fmn(fmn(2 * r, (2 + r) / 3.), 2.));
  psi2 = (1 - (1 - psi2) * sw);
  up = u[0] + 0.25 * phi * (((1. - kappa) * psi1 * bd) + ((1. + kappa) * psi2 * fd));
  lo = u[0] - 0.25 * phi * (((1. + kappa) * psi1 * bd) + ((1. - kappa) * psi2 * fd));
}

// WENO5 upper- and lower-face states of a cell (weno.f90:59-87), Jiang-Shu
// smoothness indicators and linear weights (0.1, 0.6, 0.3), eps = 1e-6.
// The normalised weights w_i/sum(w) with w_i = g_i/D_i, D_i = (eps+B_i)^2 are
// formed as g_i E_i / sum(g_j E_j) with E_i the product of the other two
// D's, and the 1/6 of the candidate polynomials is folded into the same
// single division; this is the reference's quotient up to rounding.
__device__ __forceinline__ double weno_combine(double D1, double D2, double D3, double Q1, double Q2, double Q3) {
  const double g1 = 1.0 / 10.0, g2 = 6.0 / 10.0, g3 = 3.0 / 10.0;
  double e1 = g1 * (D2 * D3), e2 = g2 * (D1 * D3), e3 = g3 * (D1 * D2);
  return (e1 * Q1 + e2 * Q2 + e3 * Q3) / (6.0 * (e1 + e2 + e3));
}
__device__ __forceinline__ void weno_betas(double Um2, double Um1, double U0, double Up1, double Up2, double &D1,
                                           double &D2, double &D3) {
  const double eps = 1e-6;
  double B1 = (13.0 / 12.0) * sqd(Um2 - 2.0 * Um1 + U0) + (1.0 / 4.0) * sqd(Um2 - 4.0 * Um1 + 3.0 * U0);
  double B2 = (13.0 / 12.0) * sqd(Um1 - 2.0 * U0 + Up1) + (1.0 / 4.0) * sqd(Um1 - Up1);
  double B3 = (13.0 / 12.0) * sqd(U0 - 2.0 * Up1 + Up2) + (1.0 / 4.0) * sqd(3.0 * U0 - 4.0 * Up1 + Up2);
  D1 = sqd(eps + B1); D2 = sqd(eps + B2); D3 = sqd(eps + B3);
}
__device__ __forceinline__ double weno_up(double Um2, double Um1, double U0, double Up1, double Up2) {
  double D1, D2, D3;
  weno_betas(Um2, Um1, U0, Up1, Up2, D1, D2, D3);
  return weno_combine(D1, D2, D3, 2.0 * Um2 - 7.0 * Um1 + 11.0 * U0, -1.0 * Um1 + 5.0 * U0 + 2.0 * Up1,
                      2.0 * U0 + 5.0 * Up1 - 1.0 * Up2);
}
__device__ __forceinline__ double weno_lo(double Um2, double Um1, double U0, double Up1, double Up2) {
  double D1, D2, D3;
  weno_betas(Um2, Um1, U0, Up1, Up2, D1, D2, D3);
  return weno_combine(D3, D2, D1, 2.0 * Up2 - 7.0 * Up1 + 11.0 * U0, -1.0 * Up1 + 5.0 * U0 + 2.0 * Um1,
                      2.0 * U0 + 5.0 * Um1 - 1.0 * Um2);
}

// PPM face estimate at the face between cells m-1 and m; u points at cell m
__device__ __forceinline__ double ppm_face(const double *u, long s) {
  return (7. * (u[0] + u[-s]) - (u[s] + u[-2 * s])) / 12.;
}
// remove_extrema on one cell (ppm.f90:86-101): Lc is the cell's upper-face
// value, Rc its lower-face value
__device__ __forceinline__ void ppm_extrema(double qc, double &Lc, double &Rc) {
  if ((Lc - qc) * (qc - Rc) <= 0) {
    Lc = qc; Rc = qc;
  } else {
    double dqrl = Lc - Rc;
    double dq6 = 6. * (qc - 0.5 * (Lc + Rc));
    if (dqrl * dq6 > dqrl * dqrl) Rc = 3. * qc - 2. * Lc;
    else if (-dqrl * dqrl > dqrl * dq6) Lc = 3. * qc - 2. * Rc;
  }
}

// left (L) and right (R) states of face n (u points at cell n, stride s);
// RECON is the reconstruction (0 none, 1 MUSCL, 2 PPM, 3 WENO5)
template <int RECON>
__device__ __forceinline__ void face_states(const Blk &b, int d, int n, int N, int l, const double *u, long s,
                                            double &L, double &R) {
  constexpr int recon = RECON;
  const int sw = (l >= 5) ? b.tlim[d] : b.lim[d];
  const int minbc = b.bc[2 * d], maxbc = b.bc[2 * d + 1];
  const bool on = recon != 0;
  const bool bmin = on && minbc < 0 && minbc != -10;
  const bool bmax = on && maxbc < 0 && maxbc != -10;
  const bool gmin = (minbc == -8 || minbc == -9), gmax = (maxbc == -8 || maxbc == -9);
  const bool pf = b.ppm_flag;
  double dummy;
  // left state: upper face of cell n-1
  if (bmin && n == 1) {
    L = gmin ? u[-s] : 0.5 * (u[-s] + u[0]);
  } else if (bmax && n == N && gmax) {
    L = u[0];
  } else if ((bmin && pf && n == 2) || (bmax && pf && n == N)) {
    bmuscl_cell(u - s, s, sw, L, dummy);
  } else {
    const double *c = u - s;
    if constexpr (recon == 0) {
      L = c[0];
    } else if constexpr (recon == 1) {
      muscl_cell(c, s, sw, L, dummy);
    } else if constexpr (recon == 2) {
      double Lc = ppm_face(u, s), Rc = ppm_face(c, s);
      if (b.lim[d] == 1) ppm_extrema(c[0], Lc, Rc);
      L = Lc;
    } else {
      L = weno_up(c[-2 * s], c[-s], c[0], c[s], c[2 * s]);
    }
  }
  // right state: lower face of cell n
  if (bmax && n == N) {
    R = gmax ? u[0] : 0.5 * (u[-s] + u[0]);
  } else if (bmin && n == 1 && gmin) {
    R = u[-s];
  } else if ((bmin && pf && n == 1) || (bmax && pf && n == N - 1)) {
    bmuscl_cell(u, s, sw, dummy, R);
  } else {
    if constexpr (recon == 0) {
      R = u[0];
    } else if constexpr (recon == 1) {
      muscl_cell(u, s, sw, dummy, R);
    } else if constexpr (recon == 2) {
      double Lc = ppm_face(u + s, s), Rc = ppm_face(u, s);
      if (b.lim[d] == 1) ppm_extrema(u[0], Lc, Rc);
      R = Rc;
    } else {
      R = weno_lo(u[-2 * s], u[-s], u[0], u[s], u[2 * s]);
    }
  }
}

// ---- inviscid flux of one face (src/face/flux/convective/*.f90)
__device__ __forceinline__ double alpha_sub(double M) { return fabs(M) < 1.0 ? 1.0 : 0.0; }

template <int NV>
__device__ __forceinline__ void inviscid_flux(const Blk &b, const double (&qL)[NV], const double (&qR)[NV], double nx,
                                              double ny, double nz, double A, double zero, double (&out)[NV]) {
  const double gm = b.gm;
  constexpr int nv = NV;
  double rL = qL[0], uL = qL[1], vL = qL[2], wL = qL[3], pL = qL[4];
  double rR = qR[0], uR = qR[1], vR = qR[2], wR = qR[3], pR = qR[4];
  const int sch = b.flux_scheme;
  if (sch == 1 || sch == 2 || sch == 3) {
    double HL = (0.5 * (uL * uL + vL * vL + wL * wL)) + ((gm / (gm - 1.)) * pL / rL);
    double HR = (0.5 * (uR * uR + vR * vR + wR * wR)) + ((gm / (gm - 1.)) * pR / rR);
    double VnL = uL * nx + vL * ny + wL * nz;
    double VnR = uR * nx + vR * ny + wR * nz;
    double mass, pbar;
    if (sch == 1) {  // AUSM+, ausmP.f90:105-158
      double cs = sqrt(2.0 * (gm - 1.0) * (0.5 * (HL + HR)) / (gm + 1.0));
      double cL = cs * cs / (fmx(cs, fabs(VnL)));
      double cR = cs * cs / (fmx(cs, fabs(VnR)));
      double C = fmn(cL, cR);
      const double iC = 1.0 / C;  // one reciprocal for the quotients by C
      double ML = VnL * iC, MR = VnR * iC;
      double aL = alpha_sub(ML), aR = alpha_sub(MR);
      double FmL = (0.5 * (1.0 + fsgn(1.0, ML)) * (1.0 - aL) * ML) + aL * 0.25 * sqd(1.0 + ML);
      double betaL = (0.5 * (1.0 + fsgn(1.0, ML)) * (1.0 - aL)) + aL * 0.25 * sqd(1.0 + ML) * (2.0 - ML);
      double FmR = (0.5 * (1.0 - fsgn(1.0, MR)) * (1.0 - aR) * MR) - aR * 0.25 * sqd(1.0 - MR);
      double betaR = (0.5 * (1.0 - fsgn(1.0, MR)) * (1.0 - aR)) + aR * 0.25 * sqd(1.0 - MR) * (2.0 + MR);
      FmL = FmL + aL * 0.1250 * sqd(ML * ML - 1.0);
      betaL = betaL + aL * 0.1875 * sqd(ML * ML - 1.0) * ML;
      FmR = FmR - aR * 0.1250 * sqd(MR * MR - 1.0);
      betaR = betaR - aR * 0.1875 * sqd(MR * MR - 1.0) * MR;
      double Mface = FmL + FmR;
      pbar = betaL * pL + betaR * pR;
      mass = (Mface > 0.0) ? Mface * C * rL : Mface * C * rR;
    } else if (sch == 2) {  // AUSM+-up, ausmUP.f90:113-189 (uses the freestream Mach)
      const double Kp = 0.25, Ku = 0.75, sigma = 1.0;
      double cs = sqrt(2.0 * (gm - 1.0) * (0.5 * (HL + HR)) / (gm + 1.0));
      double cL = cs * cs / (fmx(cs, VnL));
      double cR = cs * cs / (fmx(cs, -VnR));
      double C = fmn(cL, cR);
      const double iC = 1.0 / C;  // one reciprocal for the quotients by C
      double ML = VnL * iC, MR = VnR * iC;
      double Mb = sqrt(0.5 * ((VnL * VnL) + (VnR * VnR)) / (C * C));
      double Mo = sqrt(fmn(1.0, fmx(Mb * Mb, b.MInf * b.MInf)));
      double fna = Mo * (2.0 - Mo);
      double alfa = 3.0 * (-4.0 + (5.0 * fna * fna)) / 16.0;
      double aL = alpha_sub(ML), aR = alpha_sub(MR);
      double FmL = (0.5 * (1.0 + fsgn(1.0, ML)) * (1.0 - aL) * ML) + aL * 0.25 * sqd(1.0 + ML);
      double betaL = (0.5 * (1.0 + fsgn(1.0, ML)) * (1.0 - aL)) + aL * 0.25 * sqd(1.0 + ML) * (2.0 - ML);
      double FmR = (0.5 * (1.0 - fsgn(1.0, MR)) * (1.0 - aR) * MR) - aR * 0.25 * sqd(1.0 - MR);
      double betaR = (0.5 * (1.0 - fsgn(1.0, MR)) * (1.0 - aR)) + aR * 0.25 * sqd(1.0 - MR) * (2.0 + MR);
      FmL = FmL + aL * 0.125 * sqd(ML * ML - 1.0);
      betaL = betaL + aL * alfa * sqd(ML * ML - 1.0) * ML;
      FmR = FmR - aR * 0.125 * sqd(MR * MR - 1.0);
      betaR = betaR - aR * alfa * sqd(MR * MR - 1.0) * MR;
      double Pu = -Ku * betaL * betaR * (rL + rR) * fna * C * (VnR - VnL);
      double Mp = -2.0 * Kp * fmx(1.0 - (sigma * Mb * Mb), 0.0) * (pR - pL) / (fna * (rL + rR) * C * C);
      double Mface = FmL + FmR + Mp;
      pbar = betaL * pL + betaR * pR + Pu;
      mass = (Mface > 0.0) ? Mface * C * rL : Mface * C * rR;
    } else {  // SLAU, slau.f90:108-161
      double cL = sqrt(gm * pL / rL), cR = sqrt(gm * pR / rR);
      double C = 0.5 * (cL + cR);
      double delp = pR - pL;
      const double iC = 1.0 / C;  // one reciprocal for the quotients by C
      double ML = VnL * iC, MR = VnR * iC;
      double aL = alpha_sub(ML), aR = alpha_sub(MR);
      double betaL = (1.0 - aL) * 0.5 * (1.0 + fsgn(1.0, ML)) + (aL) * 0.25 * (2.0 - ML) * sqd(ML + 1.0);
      double betaR = (1.0 - aR) * 0.5 * (1.0 - fsgn(1.0, MR)) + (aR) * 0.25 * (2.0 + MR) * sqd(MR - 1.0);
      double vt = sqrt(0.5 * ((uL * uL) + (vL * vL) + (wL * wL) + (uR * uR) + (vR * vR) + (wR * wR)));
      double Mcap = fmn(1.0, vt * iC);
      double Xi = sqd(1.0 - Mcap);
      double Vnabs = (rL * fabs(VnL) + rR * fabs(VnR)) / (rL + rR);
      double fnG = -1.0 * fmx(fmn(ML, 0.0), -1.0) * fmn(fmx(MR, 0.0), 1.0);
      pbar = 0.5 * ((pL + pR) + (betaL - betaR) * (pL - pR) + (1.0 - Xi) * (betaL + betaR - 1.0) * (pL + pR));
      double VnabsL = (1.0 - fnG) * Vnabs + fnG * fabs(VnL);
      double VnabsR = (1.0 - fnG) * Vnabs + fnG * fabs(VnR);
      mass = 0.5 * ((rL * (VnL + VnabsL) + rR * (VnR - VnabsR)) - (Xi * delp * iC));
    }
    mass = mass * zero;
    double Fp0 = 0.5 * (mass + fabs(mass)), Fm0 = 0.5 * (mass - fabs(mass));
    out[0] = (Fp0 + Fm0) * A;
    out[1] = ((Fp0 * uL + Fm0 * uR) + (pbar * nx)) * A;
    out[2] = ((Fp0 * vL + Fm0 * vR) + (pbar * ny)) * A;
    out[3] = ((Fp0 * wL + Fm0 * wR) + (pbar * nz)) * A;
    out[4] = (Fp0 * HL + Fm0 * HR) * A;
#pragma unroll
    for (int l = 5; l < nv; ++l) out[l] = (Fp0 * qL[l] + Fm0 * qR[l]) * A;
    return;
  }
  // AUSM (0), LDFSS(0) (4), van Leer (5): ausm.f90, ldfss0.f90, van_leer.f90
  double ssa = 0.5 * (sqrt(gm * pL / rL) + sqrt(gm * pR / rR));
  double MpL = (uL * nx + vL * ny + wL * nz) / ssa;
  double ap = 0.5 * (1.0 + fsgn(1.0, MpL));
  double bl = fabs(MpL) < 1.0 ? -1.0 : 0.0;
  double Mplus = 0.25 * sqd(1. + MpL);
  double Dplus = 0.25 * sqd(1. + MpL) * (2. - MpL);
  double cplus = (ap * (1.0 + bl) * MpL) - bl * Mplus;
  double sDp = (ap * (1. + bl)) - (bl * Dplus);
  double MpR = (uR * nx + vR * ny + wR * nz) / ssa;
  double am = 0.5 * (1.0 - fsgn(1.0, MpR));
  double br = fabs(MpR) < 1.0 ? -1.0 : 0.0;
  double Mminus = -0.25 * sqd(1. - MpR);
  double Dminus = 0.25 * sqd(1. - MpR) * (2. + MpR);
  double cminus = (am * (1.0 + br) * MpR) - br * Mminus;
  double sDm = (am * (1. + br)) - (br * Dminus);
  if (sch == 0) {
    double tc = cplus + cminus;
    cplus = fmx(0., tc);
    cminus = fmn(0., tc);
  } else if (sch == 4) {
    double Ml = 0.25 * bl * br * sqd(sqrt((MpL * MpL + MpR * MpR) * 0.5) - 1);
    double Mpl = Ml * (1 - (pL - pR) / (2 * rL * (ssa * ssa)));
    double Mml = Ml * (1 - (pL - pR) / (2 * rR * (ssa * ssa)));
    cplus = cplus - Mpl;
    cminus = cminus + Mml;
  }
  double Fp0 = rL * ssa * cplus, Fm0 = rR * ssa * cminus;
  Fp0 = Fp0 * zero;
  Fm0 = Fm0 * zero;
  double Fp[5], Fm[5];
  Fp[0] = Fp0; Fm[0] = Fm0;
  Fp[1] = (Fp0 * uL) + (sDp * pL * nx);
  Fp[2] = (Fp0 * vL) + (sDp * pL * ny);
  Fp[3] = (Fp0 * wL) + (sDp * pL * nz);
  Fp[4] = Fp0 * ((0.5 * (uL * uL + vL * vL + wL * wL)) + ((gm / (gm - 1.)) * pL / rL));
  Fm[1] = (Fm0 * uR) + (sDm * pR * nx);
  Fm[2] = (Fm0 * vR) + (sDm * pR * ny);
  Fm[3] = (Fm0 * wR) + (sDm * pR * nz);
  Fm[4] = Fm0 * ((0.5 * (uR * uR + vR * vR + wR * wR)) + ((gm / (gm - 1.)) * pR / rR));
#pragma unroll
  for (int l = 0; l < 5; ++l) out[l] = Fp[l] * A + Fm[l] * A;
#pragma unroll
  for (int l = 5; l < nv; ++l) out[l] = (Fp0 * qL[l]) * A + (Fm0 * qR[l]) * A;
}

// ---- viscous face terms (src/viscous.f90:144-447): v[l], l = 1..NV-1, is
// what the reference subtracts from the inviscid flux of a face between
// cells cm and c (gradient indices gL, gR) with unit normal n and area A;
// the laminar and SST parts of one component are summed before the
// subtraction.  laminar_only keeps the laminar part (with mu_t) alone: the
// k faces of a 2-D grid, where the reference skips the SST viscous flux
// (src/viscous.f90:101-105).  NV 6 is Spalart-Allmaras: its nu-tilde flux
// runs on every face (src/viscous.f90:94-97, 570-656)
template <int NV>
__device__ __forceinline__ void visc_face(const Blk &b, long c, long cm, long gL, long gR, double nx, double ny,
                                          double nz, double A, bool laminar_only, double (&v)[NV]) {
  const double *qp = b.qp;
  const long nc = b.nc;
  double delx = b.cx[c] - b.cx[cm], dely = b.cy[c] - b.cy[cm], delz = b.cz[c] - b.cz[cm];
  double d2 = delx * delx + dely * dely + delz * delz;
  double id2 = 1.0 / d2;  // normal_comp*del/d_LR = (del_phi - grad.del)*del/d_LR^2
  const double *gx = b.g[0], *gy = b.g[1], *gz = b.g[2];
  const long ngr = b.ngr;
  double gr[4][3];
#pragma unroll
  for (int l = 0; l < 4; ++l) {
    gr[l][0] = 0.5 * (gx[l * ngr + gL] + gx[l * ngr + gR]);
    gr[l][1] = 0.5 * (gy[l * ngr + gL] + gy[l * ngr + gR]);
    gr[l][2] = 0.5 * (gz[l * ngr + gL] + gz[l * ngr + gR]);
  }
  double T_LE = qp[4 * nc + cm] / (qp[cm] * b.R);
  double T_RE = qp[4 * nc + c] / (qp[c] * b.R);
  double del[4] = {qp[nc + c] - qp[nc + cm], qp[2 * nc + c] - qp[2 * nc + cm], qp[3 * nc + c] - qp[3 * nc + cm],
                   T_RE - T_LE};
#pragma unroll
  for (int l = 0; l < 4; ++l) {
    double ncmp = (del[l] - (gr[l][0] * delx + gr[l][1] * dely + gr[l][2] * delz)) * id2;
    gr[l][0] = gr[l][0] + ncmp * delx;
    gr[l][1] = gr[l][1] + ncmp * dely;
    gr[l][2] = gr[l][2] + ncmp * delz;
  }
  double mu_f = 0.5 * (b.mu[cm] + b.mu[c]);
  // compute_viscous_fluxes_laminar takes mu_t whenever the run is turbulent
  // (src/viscous.f90:270-277)
  double mut_f = (NV >= 6) ? 0.5 * (b.mut[cm] + b.mut[c]) : 0.0;
  double tm = mu_f + mut_f;
  double dudx = gr[0][0], dudy = gr[0][1], dudz = gr[0][2];
  double dvdx = gr[1][0], dvdy = gr[1][1], dvdz = gr[1][2];
  double dwdx = gr[2][0], dwdy = gr[2][1], dwdz = gr[2][2];
  double div3 = (dudx + dvdy + dwdz) * (1.0 / 3.0);
  double Txx = 2. * tm * (dudx - div3);
  double Tyy = 2. * tm * (dvdy - div3);
  double Tzz = 2. * tm * (dwdz - div3);
  double Txy = tm * (dvdx + dudy);
  double Txz = tm * (dwdx + dudz);
  double Tyz = tm * (dwdy + dvdz);
  double K_heat = (mu_f * b.iPr + mut_f * b.itPr) * b.cpR;
  double Qx = K_heat * gr[3][0], Qy = K_heat * gr[3][1], Qz = K_heat * gr[3][2];
  double uf = 0.5 * (qp[nc + cm] + qp[nc + c]);
  double vf = 0.5 * (qp[2 * nc + cm] + qp[2 * nc + c]);
  double wf = 0.5 * (qp[3 * nc + cm] + qp[3 * nc + c]);
  v[0] = 0.0;
  v[1] = (Txx * nx + Txy * ny + Txz * nz) * A;
  v[2] = (Txy * nx + Tyy * ny + Tyz * nz) * A;
  v[3] = (Txz * nx + Tyz * ny + Tzz * nz) * A;
  v[4] = A * (((Txx * uf + Txy * vf + Txz * wf + Qx) * nx) + ((Txy * uf + Tyy * vf + Tyz * wf + Qy) * ny) +
              ((Txz * uf + Tyz * vf + Tzz * wf + Qz) * nz));
  if constexpr (NV == 7) {
    v[5] = 0.0;
    v[6] = 0.0;
  }
  if constexpr (NV == 6) {
    // nu-tilde diffusion with (mu + rho_f nu~_f) / sigma, sigma = 2/3; the
    // quotient taken as a product with 1.5 (within an ulp of it)
    double g0 = 0.5 * (gx[4 * ngr + gL] + gx[4 * ngr + gR]);
    double g1 = 0.5 * (gy[4 * ngr + gL] + gy[4 * ngr + gR]);
    double g2 = 0.5 * (gz[4 * ngr + gL] + gz[4 * ngr + gR]);
    double ncmp = ((qp[5 * nc + c] - qp[5 * nc + cm]) - (g0 * delx + g1 * dely + g2 * delz)) * id2;
    g0 = g0 + ncmp * delx;
    g1 = g1 + ncmp * dely;
    g2 = g2 + ncmp * delz;
    double rhof = 0.5 * (qp[cm] + qp[c]);
    double mtv = 0.5 * (qp[5 * nc + cm] + qp[5 * nc + c]) * rhof;
    v[NV == 6 ? 5 : 0] = (A * ((mu_f + mtv) * (g0 * nx + g1 * ny + g2 * nz))) * 1.5;
  }
  if (NV == 7 && !laminar_only) {
    // SST k / omega diffusion and the -2/3 rho k stress, src/viscous.f90:328-447
    double gt[2][3];
#pragma unroll
    for (int l = 0; l < 2; ++l) {
      gt[l][0] = 0.5 * (gx[(4 + l) * ngr + gL] + gx[(4 + l) * ngr + gR]);
      gt[l][1] = 0.5 * (gy[(4 + l) * ngr + gL] + gy[(4 + l) * ngr + gR]);
      gt[l][2] = 0.5 * (gz[(4 + l) * ngr + gL] + gz[(4 + l) * ngr + gR]);
    }
    double dt2[2] = {qp[5 * nc + c] - qp[5 * nc + cm], qp[6 * nc + c] - qp[6 * nc + cm]};
#pragma unroll
    for (int l = 0; l < 2; ++l) {
      double ncmp = (dt2[l] - (gt[l][0] * delx + gt[l][1] * dely + gt[l][2] * delz)) * id2;
      gt[l][0] = gt[l][0] + ncmp * delx;
      gt[l][1] = gt[l][1] + ncmp * dely;
      gt[l][2] = gt[l][2] + ncmp * delz;
    }
    double F1 = 0.5 * (b.f1[cm] + b.f1[c]);
    double sigma_kf = 0.85 * F1 + 1.0 * (1.0 - F1);
    double sigma_wf = 0.5 * F1 + 0.856 * (1.0 - F1);
    double rhof = 0.5 * (qp[cm] + qp[c]);
    double tkf = 0.5 * (qp[5 * nc + cm] + qp[5 * nc + c]);
    double Tkk = -(2.0 * rhof * tkf * (1.0 / 3.0));
    v[1] = v[1] + Tkk * nx * A;
    v[2] = v[2] + Tkk * ny * A;
    v[3] = v[3] + Tkk * nz * A;
    double kfl = A * ((mu_f + sigma_kf * mut_f) * (gt[0][0] * nx + gt[0][1] * ny + gt[0][2] * nz));
    v[4] = v[4] + kfl;
    v[NV == 7 ? 5 : 0] = kfl;
    v[NV == 7 ? 6 : 0] = A * ((mu_f + sigma_wf * mut_f) * (gt[1][0] * nx + gt[1][1] * ny + gt[1][2] * nz));
  }
}

// ---- the face kernel// ---- the face kernel
// D: sweep direction.  Faces: n = 1..N along D, 1..N-1 across.  NV: 5
// (laminar), 6 (Spalart-Allmaras) or 7 (SST) variables; RECON: the reconstruction.  Compile-time
// counts keep the per-variable arrays in registers.
// Registers are fitted to 4 waves per SIMD: 5 spills the SST/WENO variants
// and measured slower even where it does not spill.
template <int D, int NV, int RECON, bool FLAT = false>
__global__ void __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(4))) k_flux(Blk b, int store_csnd) {
  const int ii = D == 0, jj = D == 1, kk = D == 2;
  int i, j;
  // i sweep: flat rows (imx faces per row); j, k sweeps: 64 x 4 tiles
  const int blk = plane_block(b.xcd_remap);
  const bool in = (D == 0) ? flat_ij(b.imx - 1 + ii, b.jmx - 1 + jj, i, j, blk)
                           : tile_ij(b.imx - 1 + ii, b.jmx - 1 + jj, i, j, blk);
  if (!in) return;
  i += 1;
  j += 1;
  const int k = blockIdx.y + 1;
  if (k > b.kmx - 1 + kk) return;
  const int N = D == 0 ? b.imx : D == 1 ? b.jmx : b.kmx;
  const int n = D == 0 ? i : D == 1 ? j : k;
  const long s = cstride(b, D);
  const long c = ci(b, i, j, k);
  const long cm = c - s;
  constexpr int nv = NV;
  const long o = fli(b, D, i, j, k);
  // FLAT: the k faces of a 2-D grid (one cell in k), no convective flux
  // (src/face/flux/convective/ausm.f90:205-207 and alike in every scheme)
  constexpr bool flat = FLAT;
  double out[NV];
  double csf = 0.0;
  const long fo = fi(b, D, i, j, k);
  const double nx = b.fx[D][fo], ny = b.fy[D][fo], nz = b.fz[D][fo], A = b.fA[D][fo];
  {
    double qL[NV], qR[NV];
#pragma unroll
    for (int l = 0; l < nv; ++l) {
      const double *u = b.qp + (long)l * b.nc + c;
      face_states<RECON>(b, D, n, N, l, u, s, qL[l], qR[l]);
    }
    if (store_csnd) csf = 0.5 * (sqrt(b.gm * qL[4] / qL[0]) + sqrt(b.gm * qR[4] / qR[0]));
    const double zero = (n == 1) ? (double)b.mz[D][0] : (n == N) ? (double)b.mz[D][1] : 1.0;
    if constexpr (flat) {
#pragma unroll
      for (int l = 0; l < nv; ++l) out[l] = 0.0;
      (void)zero;
    } else {
      inviscid_flux(b, qL, qR, nx, ny, nz, A, zero, out);
    }
  }
  if (b.mu_ref != 0.0) {
    double v[NV];
    visc_face<NV>(b, c, cm, gri(b, i - ii, j - jj, k - kk), gri(b, i, j, k), nx, ny, nz, A, flat, v);
#pragma unroll
    for (int l = 1; l < nv; ++l) out[l] = out[l] - v[l];
  }

  bool bad = false;
#pragma unroll
  for (int l = 0; l < nv; ++l) {
    bad |= isnan(out[l]);
    b.flux[D][(long)l * b.nfl[D] + o] = out[l];
  }
  if (bad) atomicOr(b.err, ERR_NAN_FLUX);
  if (store_csnd) b.csnd[D][o] = csf;
}

}  // namespace fest
