// k_grad.h -- Green-Gauss cell gradients (src/gradients.f90:276-482) and the
// ghost-cell gradient rule (src/gradients.f90:486-676), molecular, SST (with
// the F1 blending function) and Spalart-Allmaras eddy viscosity
// (src/viscosity.f90:52-546), and the SST and SA sources.
//
// One thread per cell computes all 3*n_grad gradients (the reference makes
// 3*n_grad separate passes): the 6 face metrics and the volume are read once
// and reused for every variable.
#pragma once
// the residual kernels let multiply-adds fuse into FMA: results differ from
// the unfused reference order by rounding only, far inside the 1e-12 parity
// bound, and the FP64 instruction count, which bounds these kernels, drops
#pragma clang fp contract(fast)
#include "blk.h"
#include "k_bc.h"

namespace fest {

// gradient slots: u, v, w, T, then k, omega (SST) or nu-tilde (SA), then
// lctm2015's intermittency
__device__ __forceinline__ const double *grad_var(const Blk &b, int l) {
  return l < 3 ? b.qp + (long)(l + 1) * b.nc : l == 3 ? b.temp : b.qp + (long)(l + 1) * b.nc;
}

// SST eddy viscosity and F1 of one cell from its gradients
// (src/viscosity.f90:343-388); S2003: the 2003 form, with the strain rate in
// place of the vorticity and 1e-10 floors (src/viscosity.f90:215-275)
template <bool S2003, int G>
__device__ __forceinline__ void sst_mut_f1(double rho, double tk, double tw, double dist, double mu,
                                           const double (&g)[G][3], double &mut, double &f1) {
  const double bstar = 0.09, a1 = 0.31, sigma_w2 = 0.856;
  double var1 = sqrt(tk) / (bstar * tw * dist);
  double var2 = 500 * (mu / rho) / ((dist * dist) * tw);
  double arg2 = fmx(2 * var1, var2);
  double F = tanh(arg2 * arg2);
  double NUM = rho * a1 * tk, DENOM;
  if constexpr (!S2003) {
    double wx = (g[2][1] - g[1][2]);
    double wy = (g[0][2] - g[2][0]);
    double wz = (g[1][0] - g[0][1]);
    double vort = sqrt(wx * wx + wy * wy + wz * wz);
    DENOM = fmx(fmx((a1 * tw), vort * F), 1.e-20);
  } else {
    double sxx = g[0][0], syy = g[1][1], szz = g[2][2];
    double syz = g[2][1] + g[1][2], szx = g[0][2] + g[2][0], sxy = g[1][0] + g[0][1];
    double SijSij = (2.0 * (sxx * sxx)) + (2.0 * (syy * syy)) + (2.0 * (szz * szz)) + syz * syz + szx * szx + sxy * sxy;
    DENOM = fmx(fmx((a1 * tw), sqrt(SijSij) * F), 1.0e-10);
  }
  mut = NUM / DENOM;
  double CD = fmx(2 * rho * sigma_w2 * (g[4][0] * g[5][0] + g[4][1] * g[5][1] + g[4][2] * g[5][2]) / tw,
                  S2003 ? 1.0e-10 : 1e-20);
  double right = 4 * (rho * sigma_w2 * tk) / (CD * (dist * dist));
  double left = fmx(var1, var2);
  double arg1 = fmn(left, right);
  f1 = tanh(arg1 * arg1 * arg1 * arg1);
}

// SA eddy viscosity rho nu~ fv1 (src/viscosity.f90:149-163)
__device__ __forceinline__ double sa_mut(double rho, double tv, double mu) {
  double xi = tv * rho / mu;
  double x3 = (xi * xi) * xi;
  return rho * tv * (x3 / (x3 + Sa::cv1_3));
}

// SA source of an interior cell, times its volume (src/source.f90:835-983).
// grho is the cell's own Green-Gauss density gradient, whose k-face normal
// is (nx, nx, nx) as in the reference (source.f90:887).
__device__ __forceinline__ double sa_source(double rho, double tv, double mu, double dist, double vol,
                                            const double (&g)[5][3], const double (&grho)[3]) {
  double vort = sqrt(sqd(g[2][1] - g[1][2]) + sqd(g[0][2] - g[2][0]) + sqd(g[1][0] - g[0][1]));
  double CD1 = Sa::cb2 * ((g[4][0] * g[4][0]) + (g[4][1] * g[4][1]) + (g[4][2] * g[4][2]));
  double CD2 = (grho[0] * g[4][0]) + (grho[1] * g[4][1]) + (grho[2] * g[4][2]);
  double kd = Sa::kappa * dist, kd2 = kd * kd;
  double nu = mu / rho;
  double xi = tv / nu;
  double x3 = (xi * xi) * xi;
  double fv1 = x3 / (x3 + Sa::cv1_3);
  double fv2 = 1.0 - xi / (1.0 + (xi * fv1));
  double scap = fmx(vort + (tv * fv2 / kd2), 0.3 * vort);
  double r = fmn(tv / (scap * kd2), 10.0);
  double r6 = ((((r * r) * r) * r) * r) * r;
  double gg = r + Sa::cw2 * (r6 - r);
  double g6 = ((((gg * gg) * gg) * gg) * gg) * gg;
  double fw = gg * pow((1.0 + Sa::cw3_6) / (g6 + Sa::cw3_6), 1.0 / 6.0);
  double td = tv / dist;
  double D_v = rho * Sa::cw1 * fw * (td * td);
  double P_v = rho * Sa::cb1 * scap * tv;
  double lamda = rho * CD1 / Sa::sigma - CD2 * (nu + tv) / Sa::sigma;
  return (P_v - D_v + lamda) * vol;
}

// The 'bc' intermittency 1 - exp(-(term1 + term2)) (src/source.f90:566-580,
// 1128-1142)
__device__ __forceinline__ double gamma_bc(const Blk &b, double nu_t, double vmag, double dist, double re_v) {
  const double chi_1 = 0.002, chi_2 = 5.0;
  double nu_cr = chi_2 / b.Re;
  double nu_bc = nu_t / (vmag * dist);
  double re_theta = re_v / 2.193;
  double re_theta_t = (803.73 * pow(b.tu_inf + 0.6067, -1.027));
  double term1 = sqrt(fmx(re_theta - re_theta_t, 0.) / (chi_1 * re_theta_t));
  double term2 = sqrt(fmx(nu_bc - nu_cr, 0.0) / nu_cr);
  return 1.0 - exp(-(term1 + term2));
}

// SA source with the 'bc' transition (src/source.f90:985-1194): production
// times gamma_BC and destruction in kinematic form, the cross diffusion as
// in sa_source
__device__ __forceinline__ double sa_bc_source(const Blk &b, double rho, double vmag, double tv, double mu,
                                               double dist, double vol, const double (&g)[5][3],
                                               const double (&grho)[3]) {
  double Omega = sqrt(sqd(g[2][1] - g[1][2]) + sqd(g[0][2] - g[2][0]) + sqd(g[1][0] - g[0][1]));
  double CD1 = Sa::cb2 * ((g[4][0] * g[4][0]) + (g[4][1] * g[4][1]) + (g[4][2] * g[4][2]));
  double CD2 = (grho[0] * g[4][0]) + (grho[1] * g[4][1]) + (grho[2] * g[4][2]);
  double d2 = dist * dist, k2 = Sa::kappa * Sa::kappa;
  double nu = mu / rho;
  double Ji = tv / nu, Ji_3 = (Ji * Ji) * Ji;
  double fv1 = Ji_3 / (Ji_3 + Sa::cv1_3);
  double fv2 = 1.0 - Ji / (1.0 + (Ji * fv1));
  double inv_k2_d2 = 1.0 / (k2 * d2);
  double Shat = fmx(Omega + tv * fv2 * inv_k2_d2, 1.0e-10);
  double inv_Shat = 1.0 / Shat;
  double gBC = gamma_bc(b, tv * fv1, vmag, dist, d2 * Omega / nu);
  double Production = gBC * Sa::cb1 * Shat * tv * vol;
  double r = fmn(tv * inv_Shat * inv_k2_d2, 10.0);
  double r6 = ((((r * r) * r) * r) * r) * r;
  double gg = r + Sa::cw2 * (r6 - r);
  double g6 = ((((gg * gg) * gg) * gg) * gg) * gg;
  double fw = gg * pow((1.0 + Sa::cw3_6) / (g6 + Sa::cw3_6), 1.0 / 6.0);
  double Destruction = (Sa::cw1 * fw * tv * tv / d2) * vol;
  double lamda = (rho * CD1 / Sa::sigma - CD2 * (nu + tv) / Sa::sigma) * vol;
  return Production - Destruction + lamda;
}

// SST source with the 'bc' transition (src/source.f90:467-604): production
// mu_t Omega^2 without the divergence term, limited at 20 D_k, the original
// gama1/gama2 for either SST form, P_k scaled by gamma_BC after P_omega
template <int G>
__device__ __forceinline__ void sst_bc_source(const Blk &b, double rho, double vmag, double tk, double tw,
                                              double mut, double F1, double mu, double dist, double vol,
                                              const double (&g)[G][3], double &S_k, double &S_w) {
  const double sigma_w1 = 0.5, sigma_w2 = 0.856, beta1 = 0.075, beta2 = 0.0828, bstar = 0.09, kap = 0.41;
  const double gama1 = (beta1 / bstar) - ((sigma_w1 * (kap * kap)) / sqrt(bstar));
  const double gama2 = (beta2 / bstar) - ((sigma_w2 * (kap * kap)) / sqrt(bstar));
  double vort = sqrt(sqd(g[2][1] - g[1][2]) + sqd(g[0][2] - g[2][0]) + sqd(g[1][0] - g[0][1]));
  double CD = 2 * rho * sigma_w2 * (g[4][0] * g[5][0] + g[4][1] * g[5][1] + g[4][2] * g[5][2]) / tw;
  double gama = gama1 * F1 + gama2 * (1. - F1);
  double beta = beta1 * F1 + beta2 * (1. - F1);
  double D_k = bstar * rho * tw * tk;
  double D_w = beta * rho * (tw * tw);
  double P_k = fmn(mut * (vort * vort), 20.0 * D_k);
  double P_w = (rho * gama / mut) * P_k;
  double lamda = (1. - F1) * CD;
  P_k = gamma_bc(b, mut / rho, vmag, dist, rho * dist * dist * vort / mu) * P_k;
  S_k = (P_k - D_k) * vol;
  S_w = (P_w - D_w + lamda) * vol;
}

// SST source of an interior cell, times its volume (src/source.f90:158-270);
// S2003: production limited to 10 D_k, CD floored at 1e-10, gama1 = 5/9,
// gama2 = 0.44 (src/source.f90:205-211)
template <bool S2003, int G>
__device__ __forceinline__ void sst_source(double rho, double tk, double tw, double mut, double F1, double vol,
                                           const double (&g)[G][3], double &S_k, double &S_w) {
  const double sigma_w1 = 0.5, sigma_w2 = 0.856, beta1 = 0.075, beta2 = 0.0828, bstar = 0.09, kap = 0.41;
  const double gama1 = S2003 ? 5.0 / 9.0 : (beta1 / bstar) - ((sigma_w1 * (kap * kap)) / sqrt(bstar));
  const double gama2 = S2003 ? 0.44 : (beta2 / bstar) - ((sigma_w2 * (kap * kap)) / sqrt(bstar));
  double vort = sqrt(sqd(g[2][1] - g[1][2]) + sqd(g[0][2] - g[2][0]) + sqd(g[1][0] - g[0][1]));
  double CD = 2 * rho * sigma_w2 * (g[4][0] * g[5][0] + g[4][1] * g[5][1] + g[4][2] * g[5][2]) / tw;
  CD = fmx(CD, S2003 ? 1e-10 : 1e-20);
  double gama = gama1 * F1 + gama2 * (1. - F1);
  double beta = beta1 * F1 + beta2 * (1. - F1);
  double D_k = bstar * rho * tw * tk;
  double D_w = beta * rho * (tw * tw);
  double divergence = g[0][0] + g[1][1] + g[2][2];
  double P_k = mut * (vort * vort) - ((2.0 / 3.0) * rho * tk * divergence);
  P_k = fmn(P_k, (S2003 ? 10 : 20) * D_k);
  double P_w = (rho * gama / mut) * P_k;
  double lamda = (1. - F1) * CD;
  S_k = (P_k - D_k) * vol;
  S_w = (P_w - D_w + lamda) * vol;
}

// lctm2015's blending of F1 (src/viscosity.f90:265-279, 390-404):
// F1 = max(F1, exp(-(Ry/120)**8)).  Ry = rho d sqrt(k)/mu reads the
// reference's scalars density and tk, which still hold the last cell of its
// F1 loop, (imx, jmx, kmx), not the current cell's: they are passed in.
__device__ __forceinline__ double lctm_f1(double F1, double rho_last, double tk_last, double dist, double mu) {
  const double var1 = rho_last * dist * sqrt(tk_last) / mu;
  const double x = var1 / 120;
  const double x8 = x * x * x * x * x * x * x * x;
  return fmx(F1, exp(-x8));
}

// lctm2015 source of an interior cell, times its volume (src/source.f90:
// 273-463): the SST terms with Kato-Launder production rho mu_t |Omega| S,
// weighted by the intermittency, plus the onset limiter term Pk_lim; the
// intermittency's production and destruction.  dvdy is the wall-normal
// sensor of the cell (k_lctm_sensor); g holds slots u, v, w, T, k, omega.
template <bool S2003, int G>
__device__ __forceinline__ void lctm_source(double rho, double tk, double tw, double gmi, double mut, double F1,
                                            double mu, double dist, double dvdy, double vol,
                                            const double (&g)[G][3], double &S_k, double &S_w, double &S_gm) {
  const double sigma_w1 = 0.5, sigma_w2 = 0.856, beta1 = 0.075, beta2 = 0.0828, bstar = 0.09, kap = 0.41;
  const double gama1 = S2003 ? 5.0 / 9.0 : (beta1 / bstar) - ((sigma_w1 * (kap * kap)) / sqrt(bstar));
  const double gama2 = S2003 ? 0.44 : (beta2 / bstar) - ((sigma_w2 * (kap * kap)) / sqrt(bstar));
  const int limiter = S2003 ? 10 : 20;
  double vort = sqrt(sqd(g[2][1] - g[1][2]) + sqd(g[0][2] - g[2][0]) + sqd(g[1][0] - g[0][1]));
  double strain = sqrt(sqd(g[2][1] + g[1][2]) + sqd(g[0][2] + g[2][0]) + sqd(g[1][0] + g[0][1]) +
                       2 * sqd(g[0][0]) + 2 * sqd(g[1][1]) + 2 * sqd(g[2][2]));
  double CD = 2 * rho * sigma_w2 * (g[4][0] * g[5][0] + g[4][1] * g[5][1] + g[4][2] * g[5][2]) / tw;
  CD = fmx(CD, S2003 ? 1e-10 : 1e-20);
  double gama = gama1 * F1 + gama2 * (1. - F1);
  double beta = beta1 * F1 + beta2 * (1. - F1);
  double D_k = bstar * rho * tw * tk;
  double D_w = beta * rho * (tw * tw);
  double divergence = g[0][0] + g[1][1] + g[2][2];
  double P_k = mut * (vort * strain) - ((2.0 / 3.0) * rho * tk * divergence);
  P_k = fmn(P_k, limiter * D_k);
  double P_w = (rho * gama / mut) * P_k;
  double lamda = (1. - F1) * CD;
  double lamd = (-7.57e-3) * (dvdy * dist * dist * rho / mu) + 0.0128;
  lamd = fmn(fmx(lamd, -1.0), 1.0);
  double Fpg = (lamd >= 0.0) ? fmn(1.0 + 14.68 * lamd, 1.5) : fmn(1.0 - 7.34 * lamd, 3.0);
  Fpg = fmx(Fpg, 0.0);
  double TuL = fmn(100.0 * sqrt(2.0 * tk / 3.0) / (tw * dist), 100.0);
  double Re_theta = 100.0 + 1000.0 * exp(-TuL * Fpg);
  double Rev = rho * dist * dist * strain / mu;
  double RT = rho * tk / (mu * tw);
  const double rt2 = 0.5 * RT, rt35 = RT / 3.5;
  double Fturb = exp(-(rt2 * rt2 * rt2 * rt2));
  double Fonset2 = fmn(Rev / (2.2 * Re_theta), 2.0);
  double Fonset3 = fmx(1.0 - rt35 * rt35 * rt35, 0.0);
  double Fonset = fmx(Fonset2 - Fonset3, 0.0);
  double P_gm = 100 * rho * strain * gmi * (1.0 - gmi) * Fonset;
  double D_gm = 0.06 * rho * vort * gmi * Fturb * ((50.0 * gmi) - 1.0);
  double Fon_lim = fmn(fmx((Rev / (2.2 * 1100.0)) - 1.0, 0.0), 3.0);
  double Pk_lim = 5 * fmx(gmi - 0.2, 0.0) * (1.0 - gmi) * Fon_lim * fmx(3 * mu - mut, 0.0) * strain * vort;
  S_k = (gmi * P_k - fmx(gmi, 0.1) * D_k + Pk_lim) * vol;
  S_w = (P_w - D_w + lamda) * vol;
  S_gm = (P_gm - D_gm) * vol;
}

// One thread per cell 0..imx x 0..jmx x 0..kmx: all Green-Gauss gradients,
// then (same cell, gradients in registers) the molecular viscosity, the SST
// mu_t and F1 (src/viscosity.f90:52-388) and, for interior cells, the SST
// source.  Interior gradients are final here: the ghost-gradient rule only
// writes ghost cells.  NG is the number of gradient slots (4 laminar, 6
// SST); a compile-time count keeps all 3*NG gradients in registers.
// F32: the FP32 mode when only FP32 gradients are kept (g32only) -- the
// Green-Gauss sums run in FP32 over the FP32 state copy (Temp from its FP64
// array), and mu, mu_t, F1 and the source take the FP32 gradients widened.
template <int NG, bool S2003, bool BC, bool KKL, bool F32 = false>
__device__ __forceinline__ void k_grad_b(const Blk &b, const Wg &wg) {
  int i, j;
  if (!flat_ij(b.imx + 1, b.jmx + 1, i, j, plane_block(wg, b.xcd_remap))) return;
  const int k = wg.y;
  if (k > b.kmx) return;
  long c = ci(b, i, j, k);
  long f0 = fi(b, 0, i, j, k), f1 = fi(b, 0, i + 1, j, k);
  long g0 = fi(b, 1, i, j, k), g1 = fi(b, 1, i, j + 1, k);
  long h0 = fi(b, 2, i, j, k), h1 = fi(b, 2, i, j, k + 1);
  double Ai0 = b.fA[0][f0], Ai1 = b.fA[0][f1], Aj0 = b.fA[1][g0], Aj1 = b.fA[1][g1];
  double Ak0 = b.fA[2][h0], Ak1 = b.fA[2][h1];
  double n[3][6];  // [component][face]
  n[0][0] = b.fx[0][f0]; n[1][0] = b.fy[0][f0]; n[2][0] = b.fz[0][f0];
  n[0][1] = b.fx[1][g0]; n[1][1] = b.fy[1][g0]; n[2][1] = b.fz[1][g0];
  n[0][2] = b.fx[2][h0]; n[1][2] = b.fy[2][h0]; n[2][2] = b.fz[2][h0];
  n[0][3] = b.fx[0][f1]; n[1][3] = b.fy[0][f1]; n[2][3] = b.fz[0][f1];
  n[0][4] = b.fx[1][g1]; n[1][4] = b.fy[1][g1]; n[2][4] = b.fz[1][g1];
  n[0][5] = b.fx[2][h1]; n[1][5] = b.fy[2][h1]; n[2][5] = b.fz[2][h1];
  double vol = b.vol[c];
  // one reciprocal for all 3*NG quotients by 2V (the reference divides each;
  // the product differs from the quotient by at most one ulp)
  const double ivol2 = 1.0 / (2 * vol);
  const long si = 1, sj = b.pi, sk = (long)b.pi * b.pj;
  long gidx = gri(b, i, j, k);
  bool bad = false;
  int ndir = (b.kmx > 2) ? 3 : 2;
  double gg[NG][3];
  if constexpr (F32) {
    float nf[3][6], Af[6] = {(float)Ai0, (float)Aj0, (float)Ak0, (float)Ai1, (float)Aj1, (float)Ak1};
#pragma unroll
    for (int d = 0; d < 3; ++d)
#pragma unroll
      for (int f = 0; f < 6; ++f) nf[d][f] = (float)n[d][f];
    const float iv2 = (float)ivol2;
#pragma unroll
    for (int l = 0; l < NG; ++l) {
      float vc, vmi, vmj, vmk, vpi, vpj, vpk;
      if (l == 3) {
        const double *v = b.temp;
        vc = (float)v[c]; vmi = (float)v[c - si]; vmj = (float)v[c - sj]; vmk = (float)v[c - sk];
        vpi = (float)v[c + si]; vpj = (float)v[c + sj]; vpk = (float)v[c + sk];
      } else {
        const float *v = b.q32 + (long)(l + 1) * b.nc;
        vc = v[c]; vmi = v[c - si]; vmj = v[c - sj]; vmk = v[c - sk];
        vpi = v[c + si]; vpj = v[c + sj]; vpk = v[c + sk];
      }
      const float sm_i = vmi + vc, sm_j = vmj + vc, sm_k = vmk + vc;
      const float sp_i = vpi + vc, sp_j = vpj + vc, sp_k = vpk + vc;
#pragma unroll
      for (int d = 0; d < 3; ++d) {
        float g;
        if (d < ndir) {
          g = -(sm_i * nf[d][0] * Af[0]);
          g = g - sm_j * nf[d][1] * Af[1];
          g = g - sm_k * nf[d][2] * Af[2];
          g = g + sp_i * nf[d][3] * Af[3];
          g = g + sp_j * nf[d][4] * Af[4];
          g = g + sp_k * nf[d][5] * Af[5];
          g = g * iv2;
        } else {
          g = 0.0f;
        }
        bad |= isnan(g);
        gg[l][d] = (double)g;
        b.g32[d][(long)l * b.ngr + gidx] = g;
      }
    }
  } else {
#pragma unroll
  for (int l = 0; l < NG; ++l) {
    const double *v = grad_var(b, l);
    double vc = v[c];
    double sm_i = v[c - si] + vc, sm_j = v[c - sj] + vc, sm_k = v[c - sk] + vc;
    double sp_i = v[c + si] + vc, sp_j = v[c + sj] + vc, sp_k = v[c + sk] + vc;
#pragma unroll
    for (int d = 0; d < 3; ++d) {
      double g;
      if (d < ndir) {
        g = -(sm_i * n[d][0] * Ai0);
        g = g - sm_j * n[d][1] * Aj0;
        g = g - sm_k * n[d][2] * Ak0;
        g = g + sp_i * n[d][3] * Ai1;
        g = g + sp_j * n[d][4] * Aj1;
        g = g + sp_k * n[d][5] * Ak1;
        g = g * ivol2;
      } else {
        g = 0.0;
      }
      bad |= isnan(g);
      gg[l][d] = g;
      if (!b.g32only) b.g[d][(long)l * b.ngr + gidx] = g;
      if (b.fp32) b.g32[d][(long)l * b.ngr + gidx] = (float)g;
    }
  }
  }
  if (bad) atomicOr(b.err, ERR_NAN_GRAD);
  double rho = b.qp[c];
  double mu;
  if (b.mu_suth) {
    double T = b.qp[4 * b.nc + c] / (rho * b.R);
    const double x = T / b.T_ref;  // x**1.5 as x*sqrt(x): within an ulp of pow, without its log/exp
    mu = b.mu_ref * (x * sqrt(x)) * ((b.T_ref + b.S_ref) / (T + b.S_ref));
    b.mu[c] = mu;
    if (isnan(mu)) atomicOr(b.err, ERR_NAN_MU);
  } else {
    mu = b.mu[c];
  }
  if constexpr (KKL) {
    // k-kL mu_t (src/viscosity.f90:469-483); its source reads neighbouring
    // gradients and runs after the ghost-cell gradients (k_kkl_source)
    const double tk = b.qp[5 * b.nc + c], tkl = b.qp[6 * b.nc + c];
    double mut = pow(0.09, 0.25) * rho * tkl / (fmx(sqrt(tk), 1.e-20));
    if (tkl < 1.e-14 || tk < 1.e-14) mut = 0.0;
    b.mut[c] = mut;
    return;
  }
  if constexpr (NG == 6 || NG == 7) {
  double tk = b.qp[5 * b.nc + c], tw = b.qp[6 * b.nc + c];
  double mut, F1;
  sst_mut_f1<S2003>(rho, tk, tw, b.dist[c], mu, gg, mut, F1);
  if constexpr (NG == 7) {
    const long cl = ci(b, b.imx, b.jmx, b.kmx);
    F1 = lctm_f1(F1, b.qp[cl], b.qp[5 * b.nc + cl], b.dist[c], mu);
  }
  b.mut[c] = mut;
  b.f1[c] = F1;
  if (i >= 1 && i <= b.imx - 1 && j >= 1 && j <= b.jmx - 1 && k >= 1 && k <= b.kmx - 1) {
    double S_k, S_w;
    if constexpr (NG == 7) {
      double S_gm;
      lctm_source<S2003>(rho, tk, tw, b.qp[7 * b.nc + c], mut, F1, mu, b.dist[c], b.dvdy[c], vol, gg, S_k, S_w,
                         S_gm);
      b.src[2 * b.nres + ri(b, i, j, k)] = S_gm;
    } else if constexpr (BC) {
      const double u = b.qp[b.nc + c], v = b.qp[2 * b.nc + c], w = b.qp[3 * b.nc + c];
      sst_bc_source(b, rho, sqrt(u * u + v * v + w * w), tk, tw, mut, F1, mu, b.dist[c], vol, gg, S_k, S_w);
    } else {
      sst_source<S2003>(rho, tk, tw, mut, F1, vol, gg, S_k, S_w);
    }
    long r = ri(b, i, j, k);
    b.src[r] = S_k;
    b.src[b.nres + r] = S_w;
  }
  }
  if constexpr (NG == 5) {
  double tv = b.qp[5 * b.nc + c];
  b.mut[c] = sa_mut(rho, tv, mu);
  if (i >= 1 && i <= b.imx - 1 && j >= 1 && j <= b.jmx - 1 && k >= 1 && k <= b.kmx - 1) {
    // the density gradient over the face sums rho_neighbour + rho
    const double *q0 = b.qp;
    double rf[6] = {q0[c - si] + rho, q0[c - sj] + rho, q0[c - sk] + rho,
                    q0[c + si] + rho, q0[c + sj] + rho, q0[c + sk] + rho};
    double grho[3];
#pragma unroll
    for (int d = 0; d < 3; ++d) {
      // the k-face normal of the lower k face is (nx, nx, nx)
      double nk0 = n[0][2];
      double g = -(rf[0] * n[d][0] * Ai0);
      g = g - rf[1] * n[d][1] * Aj0;
      g = g - rf[2] * nk0 * Ak0;
      g = g + rf[3] * n[d][3] * Ai1;
      g = g + rf[4] * n[d][4] * Aj1;
      g = g + rf[5] * n[d][5] * Ak1;
      grho[d] = g / (2.0 * vol);
    }
    if constexpr (BC) {
      const double u = b.qp[b.nc + c], v = b.qp[2 * b.nc + c], w = b.qp[3 * b.nc + c];
      b.src[ri(b, i, j, k)] = sa_bc_source(b, rho, sqrt(u * u + v * v + w * w), tv, mu, b.dist[c], vol, gg, grho);
    } else {
      b.src[ri(b, i, j, k)] = sa_source(rho, tv, mu, b.dist[c], vol, gg, grho);
    }
  }
  }
}
template <int NG, bool S2003, bool BC = false, bool KKL = false, bool F32 = false>
__global__ void __launch_bounds__(256) k_grad(Blk b) {
  k_grad_b<NG, S2003, BC, KKL, F32>(b, FEST_WG_HW);
}

// mu_t / F1 ghost rules (src/viscosity.f90:407-465), face interiors
__device__ __forceinline__ void visc_ghost(const Blk &b, int face, int a, int bb) {
  int id = b.bc[face];
  if (!(id == -5 || (id <= -1 && id >= -9))) return;
  if (b.kkl && id == -7) return;  // k-kL: -4..-1, -6, -8, -9 and -5 (src/viscosity.f90:485-531)
  long cg = fcell(b, face, 0, a, bb), cn = fcell(b, face, 1, a, bb);
  b.mut[cg] = (id == -5) ? -b.mut[cn] : b.mut[cn];
  if (b.ng >= 6) b.f1[cg] = b.f1[cn];
}


// ghost-cell gradients on the face interiors of every face with a
// boundary id < 0.  The reference passes Jfaces/Kfaces to a dummy declared
// with the I-face extents, so the J/K metrics are read at the I-face linear
// offset (src/gradients.f90:523-588 calling :591); that is reproduced.
__device__ __forceinline__ void k_grad_bc_b(const Blk &b, const Wg &wg) {
  int face = wg.z;
  int a = wg.x * blockDim.x + threadIdx.x + 1;
  int bb = wg.y * blockDim.y + threadIdx.y + 1;
  int na = face_na(b, face), nb = face_nb(b, face);
  if (a > na - 1 || bb > nb - 1) return;
  int id = b.bc[face];
  if (id >= 0) return;
  if (b.turb) visc_ghost(b, face, a, bb);
  int d = face >> 1, hi = face & 1;
  double Tw = b.fix[FX_TWALL][face];
  // face position and the interior (I) / ghost (G) cells
  int fi_, fj_, fk_;
  int N = d == 0 ? b.imx : d == 1 ? b.jmx : b.kmx;
  int fp = hi ? N : 1;
  if (d == 0) { fi_ = fp; fj_ = a; fk_ = bb; } else if (d == 1) { fi_ = a; fj_ = fp; fk_ = bb; } else { fi_ = a; fj_ = bb; fk_ = fp; }
  int Ii = fi_, Ij = fj_, Ik = fk_, Gi = fi_, Gj = fj_, Gk = fk_;
  if (hi) { if (d == 0) Ii--; else if (d == 1) Ij--; else Ik--; }
  else { if (d == 0) Gi--; else if (d == 1) Gj--; else Gk--; }
  long fo = fi(b, 0, fi_, fj_, fk_);  // I-face offset for every direction
  double nx = b.fx[d][fo], ny = b.fy[d][fo], nz = b.fz[d][fo], A = b.fA[d][fo];
  long cI = ci(b, Ii, Ij, Ik), cG = ci(b, Gi, Gj, Gk);
  double vI = b.vol[cI];
  double c_x = A * nx / vI, c_y = A * ny / vI, c_z = A * nz / vI;
  double sig = hi ? -1.0 : 1.0;
  long gI = gri(b, Ii, Ij, Ik), gG = gri(b, Gi, Gj, Gk);
  double *gx = b.g[0], *gy = b.g[1], *gz = b.g[2];
  float *hx = b.g32[0], *hy = b.g32[1], *hz = b.g32[2];
  // the interior gradient: FP64, or its FP32 copy when that is all there is
  auto gin = [&](int c, long o) -> double {
    if (b.g32only) return (double)(c == 0 ? hx : c == 1 ? hy : hz)[o];
    return (c == 0 ? gx : c == 1 ? gy : gz)[o];
  };
  for (int l = 0; l < b.ng; ++l) {
    double dq = b.qp[(long)(l + 1) * b.nc + cI] - b.qp[(long)(l + 1) * b.nc + cG];
    if (l == 3) dq = b.temp[cI] - b.temp[cG];
    double ox = sig * dq * c_x, oy = sig * dq * c_y, oz = sig * dq * c_z;
    long o = (long)l * b.ngr;
    double ix = gin(0, o + gI), iy = gin(1, o + gI), iz = gin(2, o + gI);
    if (l == 3 && id == -5 && (Tw < 1. && Tw >= 0.)) {
      ox = -ix; oy = -iy; oz = -iz;
    }
    double dot = (ix * nx) + (iy * ny) + (iz * nz);
    const double rx = ox + (ix - dot * nx), ry = oy + (iy - dot * ny), rz = oz + (iz - dot * nz);
    if (!b.g32only) {
      gx[o + gG] = rx;
      gy[o + gG] = ry;
      gz[o + gG] = rz;
    }
    if (b.fp32) {
      hx[o + gG] = (float)rx;
      hy[o + gG] = (float)ry;
      hz[o + gG] = (float)rz;
    }
  }
}
__global__ void k_grad_bc(Blk b) {
  k_grad_bc_b(b, FEST_WG_HW);
}

// FP32 copy of the padded state for the FP32 face sweeps
__device__ __forceinline__ void k_to_f32_b(const Blk &b, const Wg &wg) {
  long t = (long)wg.x * blockDim.x + threadIdx.x;
  if (t >= b.nc * b.nv) return;
  b.q32[t] = (float)b.qp[t];
}
__global__ void k_to_f32(Blk b) {
  k_to_f32_b(b, FEST_WG_HW);
}

// FP64 gradients widened from the FP32 copy (downloads in FP32 mode)
__global__ void k_g32_widen(Blk b) {
  long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= b.ngr * b.ng) return;
  for (int d = 0; d < 3; ++d) b.g[d][t] = (double)b.g32[d][t];
}

// A -11 face has no mu_t / F1 ghost rule (src/viscosity.f90:284-339,
// 409-465 list -1..-9): its ghost cells keep the SST values
// calculate_viscosity computes there, from the ghost gradients k_grad_bc has
// written.  Launched after k_grad_bc, only when an SST block has a -11 face.
__device__ __forceinline__ void k_mut_ghost_tp_b(const Blk &b, const Wg &wg) {
  int face = wg.z;
  int a = wg.x * blockDim.x + threadIdx.x + 1;
  int bb = wg.y * blockDim.y + threadIdx.y + 1;
  int na = face_na(b, face), nb = face_nb(b, face);
  if (a > na - 1 || bb > nb - 1) return;
  if (b.bc[face] != -11 || b.ng < 6) return;
  int d = face >> 1, hi = face & 1;
  int N = d == 0 ? b.imx : d == 1 ? b.jmx : b.kmx;
  int g = hi ? N : 0;  // the ghost layer next to the face
  int Gi, Gj, Gk;
  if (d == 0) { Gi = g; Gj = a; Gk = bb; } else if (d == 1) { Gi = a; Gj = g; Gk = bb; } else { Gi = a; Gj = bb; Gk = g; }
  long cG = ci(b, Gi, Gj, Gk), gG = gri(b, Gi, Gj, Gk);
  double gg[6][3];
#pragma unroll
  for (int l = 0; l < 6; ++l) {
    gg[l][0] = b.g[0][(long)l * b.ngr + gG];
    gg[l][1] = b.g[1][(long)l * b.ngr + gG];
    gg[l][2] = b.g[2][(long)l * b.ngr + gG];
  }
  double mut, F1;
  double rho = b.qp[cG], tk = b.qp[5 * b.nc + cG], tw = b.qp[6 * b.nc + cG];
  if (b.turb == FEST_TURB_SST2003) sst_mut_f1<true>(rho, tk, tw, b.dist[cG], b.mu[cG], gg, mut, F1);
  else sst_mut_f1<false>(rho, tk, tw, b.dist[cG], b.mu[cG], gg, mut, F1);
  if (b.lctm) {
    const long cl = ci(b, b.imx, b.jmx, b.kmx);
    F1 = lctm_f1(F1, b.qp[cl], b.qp[5 * b.nc + cl], b.dist[cG], b.mu[cG]);
  }
  b.mut[cG] = mut;
  b.f1[cG] = F1;
}
__global__ void k_mut_ghost_tp(Blk b) {
  k_mut_ghost_tp_b(b, FEST_WG_HW);
}

// k-kL source (src/source.f90:607-832), after the ghost-cell gradients:
// production from the full stress with -2/3 rho k, limited to 20 D_k; the
// von Karman length from the Laplacian of the velocity, each second
// derivative a Green-Gauss sum over the six faces of the face-averaged
// first derivatives (ghost cells' included); its limiters, and the
// near-wall damping of both equations.  Writes S_k, S_kL (times the volume)
// where the SST source goes.
__device__ __forceinline__ void k_kkl_source_b(const Blk &b, const Wg &wg) {
  int i, j;
  if (!flat_ij(b.imx - 1, b.jmx - 1, i, j, plane_block(wg, b.xcd_remap))) return;
  i += 1;
  j += 1;
  const int k = wg.y + 1;
  if (k > b.kmx - 1) return;
  const long c = ci(b, i, j, k), r = ri(b, i, j, k), gi = gri(b, i, j, k), nc = b.nc, ngr = b.ngr;
  const double rho = b.qp[c], tk = b.qp[5 * nc + c], tkl = b.qp[6 * nc + c];
  const double mut = b.mut[c], mu = b.mu[c], dist = b.dist[c], vol = b.vol[c];
  double g[3][3];  // g[v][d]: derivative of velocity v along d
#pragma unroll
  for (int v = 0; v < 3; ++v)
#pragma unroll
    for (int d = 0; d < 3; ++d) g[v][d] = b.g[d][(long)v * ngr + gi];
  double S[3][3];
#pragma unroll
  for (int a = 0; a < 3; ++a)
#pragma unroll
    for (int e = 0; e < 3; ++e) S[a][e] = 0.5 * (g[a][e] + g[e][a]);
  const double delv = g[0][0] + g[1][1] + g[2][2];
  double P_k = 0.;
#pragma unroll
  for (int a = 0; a < 3; ++a)
#pragma unroll
    for (int e = 0; e < 3; ++e) {
      double tau = mut * (2 * S[a][e]);
      if (a == e) tau = mut * (2 * S[a][e] - (2.0 / 3.0) * delv) - (2.0 / 3.0) * rho * tk;
      P_k = P_k + tau * g[a][e];
    }
  const double D_k = (pow(0.09, 0.75)) * rho * (pow(tk, 2.5)) / fmx(tkl, 1.e-20);
  P_k = fmn(P_k, 20 * D_k);
  const long f0 = fi(b, 0, i, j, k), f1 = fi(b, 0, i + 1, j, k), g0 = fi(b, 1, i, j, k), g1 = fi(b, 1, i, j + 1, k);
  const long h0 = fi(b, 2, i, j, k), h1 = fi(b, 2, i, j, k + 1);
  const long nb[6] = {gri(b, i - 1, j, k), gri(b, i, j - 1, k), gri(b, i, j, k - 1),
                      gri(b, i + 1, j, k), gri(b, i, j + 1, k), gri(b, i, j, k + 1)};
  double lap[3];
#pragma unroll
  for (int v = 0; v < 3; ++v) {
    lap[v] = 0.;
#pragma unroll
    for (int d = 0; d < 3; ++d) {
      const double *G = b.g[d] + (long)v * ngr;
      const double *const *nn = d == 0 ? b.fx : d == 1 ? b.fy : b.fz;
      const double c0 = G[gi];
      double t = (-(G[nb[0]] + c0) * nn[0][f0] * b.fA[0][f0] - (G[nb[1]] + c0) * nn[1][g0] * b.fA[1][g0] -
                  (G[nb[2]] + c0) * nn[2][h0] * b.fA[2][h0] + (G[nb[3]] + c0) * nn[0][f1] * b.fA[0][f1] +
                  (G[nb[4]] + c0) * nn[1][g1] * b.fA[1][g1] + (G[nb[5]] + c0) * nn[2][h1] * b.fA[2][h1]) /
                 (2 * vol);
      lap[v] = lap[v] + t;
    }
  }
  const double udd = sqrt(sqd(lap[0]) + sqd(lap[1]) + sqd(lap[2]));
  double ss = 0.;
#pragma unroll
  for (int a = 0; a < 3; ++a)
#pragma unroll
    for (int e = 0; e < 3; ++e) ss = ss + sqd(S[a][e]);
  const double ud = sqrt(2 * ss);
  const double kappa = 0.41;
  double Lvk = kappa * fabs(ud / fmx(udd, 1.e-20));
  const double fp = fmn(fmx(P_k / D_k, 0.5), 1.0);
  Lvk = fmx(Lvk, tkl / fmx((tk * 10.0), 1.e-20));
  Lvk = fmn(Lvk, 1.3 * kappa * dist * fp);
  const double eta = rho * dist * sqrt(0.3 * tk) / (20 * mu);
  const double e2 = eta * eta;
  const double fphi = (1 + 4.7 * eta) / (1 + (e2 * eta) * eta);
  const double cphi1 = (1.2 - 0.97 * (sqd(tkl / fmx(tk * Lvk, 1.e-20))));
  const double P_kl = cphi1 * tkl * P_k / fmx(tk, 1.e-20);
  const double D_kl = 0.13 * rho * (pow(tk, 1.5));
  const double d2 = sqd(dist);
  const double S_k = P_k - D_k - 2 * mu * tk / d2;
  const double S_kl = P_kl - D_kl - 6 * mu * tkl * fphi / d2;
  b.src[r] = S_k * vol;
  b.src[b.nres + r] = S_kl * vol;
}
__global__ void __launch_bounds__(256) k_kkl_source(Blk b) {
  k_kkl_source_b(b, FEST_WG_HW);
}

}  // namespace fest
