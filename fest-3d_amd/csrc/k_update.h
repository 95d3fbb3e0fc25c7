// k_update.h -- residue (flux divergence, src/face/flux/convective/
// scheme.f90:111-141), SST source (src/source.f90:158-270), the
// conservative RK stage update (src/update.f90:367-485), the local/global
// time step (src/time.f90:122-531) and the residual norms
// (src/resnorm.f90:171-225).
#pragma once
// the residual kernels let multiply-adds fuse into FMA: results differ from
// the unfused reference order by rounding only, far inside the 1e-12 parity
// bound, and the FP64 instruction count, which bounds these kernels, drops
#pragma clang fp contract(fast)
#include "blk.h"

namespace fest {

// Temp = p/(R rho) on the whole padded array (src/update.f90:170) and the RK
// base U_store = qp (src/update.f90:178)
__device__ __forceinline__ void k_step_begin_b(const Blk &b, const Wg &wg, int store_u) {
  long t = (long)wg.x * blockDim.x + threadIdx.x;
  if (t >= b.nc) return;
  double rho = b.qp[t];
  b.temp[t] = b.qp[4 * b.nc + t] / (b.R * rho);
  if (store_u)
    for (int l = 0; l < b.nv; ++l) b.ust[(long)l * b.nc + t] = b.qp[(long)l * b.nc + t];
}
__global__ void k_step_begin(Blk b, int store_u) {
  k_step_begin_b(b, FEST_WG_HW, store_u);
}

// q = a U_store + c q over the whole padded primitive array, the TVD-RK
// blends of src/update.f90:204,206,213
__global__ void k_blend(Blk b, double a, double c) {
  long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= b.nc * b.nv) return;
  b.qp[t] = a * b.ust[t] + c * b.qp[t];
}

__device__ __forceinline__ double sspd(const Blk &b, int d, long o) { return b.csnd[d][o]; }

// local time step of one cell (src/time.f90:122-246, 366-531)
__device__ __forceinline__ void k_dt_b(const Blk &b, const Wg &wg) {
  int i, j;
  bool active = flat_ij(b.imx - 1, b.jmx - 1, i, j, (int)wg.x);
  i += 1;
  j += 1;
  const int k = wg.y + 1;
  active = active && k <= b.kmx - 1;
  double dtv = 1e300;
  if (active) {
    const long nc = b.nc;
    const double *qp = b.qp;
    long c = ci(b, i, j, k);
    long c4 = c + 1, c5 = c + b.pi, c6 = c + (long)b.pi * b.pj;
    long f1 = fi(b, 0, i, j, k), f2 = fi(b, 1, i, j, k), f3 = fi(b, 2, i, j, k);
    long f4 = fi(b, 0, i + 1, j, k), f5 = fi(b, 1, i, j + 1, k), f6 = fi(b, 2, i, j, k + 1);
    double u = qp[nc + c], v = qp[2 * nc + c], w = qp[3 * nc + c];
    double l1 = fabs((u * b.fx[0][f1]) + (v * b.fy[0][f1]) + (w * b.fz[0][f1])) + sspd(b, 0, fli(b, 0, i, j, k));
    double l2 = fabs((u * b.fx[1][f2]) + (v * b.fy[1][f2]) + (w * b.fz[1][f2])) + sspd(b, 1, fli(b, 1, i, j, k));
    double l3 = fabs((u * b.fx[2][f3]) + (v * b.fy[2][f3]) + (w * b.fz[2][f3])) + sspd(b, 2, fli(b, 2, i, j, k));
    double l4 = fabs((qp[nc + c4] * b.fx[0][f4]) + (qp[2 * nc + c4] * b.fy[0][f4]) + (qp[3 * nc + c4] * b.fz[0][f4])) +
                sspd(b, 0, fli(b, 0, i + 1, j, k));
    double l5 = fabs((qp[nc + c5] * b.fx[1][f5]) + (qp[2 * nc + c5] * b.fy[1][f5]) + (qp[3 * nc + c5] * b.fz[1][f5])) +
                sspd(b, 1, fli(b, 1, i, j + 1, k));
    double l6 = fabs((qp[nc + c6] * b.fx[2][f6]) + (qp[2 * nc + c6] * b.fy[2][f6]) + (qp[3 * nc + c6] * b.fz[2][f6])) +
                sspd(b, 2, fli(b, 2, i, j, k + 1));
    double A1 = b.fA[0][f1], A2 = b.fA[1][f2], A3 = b.fA[2][f3], A4 = b.fA[0][f4], A5 = b.fA[1][f5], A6 = b.fA[2][f6];
    double lsum = (A1 * l1) + (A2 * l2) + (A3 * l3) + (A4 * l4) + (A5 * l5) + (A6 * l6);
    double vol = b.vol[c];
    dtv = 1. / lsum;
    dtv = dtv * vol * b.CFLv;
    if (b.mu_ref != 0.0) {
      long c1 = c - 1, c2 = c - b.pi, c3 = c - (long)b.pi * b.pj;
      const double *cx = b.cx, *cy = b.cy, *cz = b.cz;
      double rho = qp[c];
      for (int pass = 0; pass < 2; ++pass) {
        if (pass == 1 && !b.turb) break;
        const double *m = pass == 0 ? b.mu : b.mut;
        double prn = pass == 0 ? b.Pr : b.tPr;
        double m1 = m[c] / (rho * fabs(((cx[c1] - cx[c]) * b.fx[0][f1]) + ((cy[c1] - cy[c]) * b.fy[0][f1]) + ((cz[c1] - cz[c]) * b.fz[0][f1])));
        double m2 = m[c] / (rho * fabs(((cx[c2] - cx[c]) * b.fx[1][f2]) + ((cy[c2] - cy[c]) * b.fy[1][f2]) + ((cz[c2] - cz[c]) * b.fz[1][f2])));
        double m3 = m[c] / (rho * fabs(((cx[c3] - cx[c]) * b.fx[2][f3]) + ((cy[c3] - cy[c]) * b.fy[2][f3]) + ((cz[c3] - cz[c]) * b.fz[2][f3])));
        double m4 = m[c4] / (qp[c4] * fabs(((cx[c] - cx[c4]) * b.fx[0][f4]) + ((cy[c] - cy[c4]) * b.fy[0][f4]) + ((cz[c] - cz[c4]) * b.fz[0][f4])));
        double m5 = m[c5] / (qp[c5] * fabs(((cx[c] - cx[c5]) * b.fx[1][f5]) + ((cy[c] - cy[c5]) * b.fy[1][f5]) + ((cz[c] - cz[c5]) * b.fz[1][f5])));
        double m6 = m[c6] / (qp[c6] * fabs(((cx[c] - cx[c6]) * b.fx[2][f6]) + ((cy[c] - cy[c6]) * b.fy[2][f6]) + ((cz[c] - cz[c6]) * b.fz[2][f6])));
        double ls = (A1 * m1) + (A2 * m2) + (A3 * m3) + (A4 * m4) + (A5 * m5) + (A6 * m6);
        ls = b.gm * ls / prn;
        ls = 2. / (ls + (2. * b.CFLv * vol / dtv));
        dtv = b.CFLv * (ls * vol);
      }
    }
    b.dt[ri(b, i, j, k)] = dtv;
  }
  if (!b.time_local) {
    // block minimum for the global step (src/time.f90:248-289)
    __shared__ double smin[256];
    int tid = threadIdx.x;
    smin[tid] = dtv;
    __syncthreads();
    for (int st = 128; st > 0; st >>= 1) {
      if (tid < st) smin[tid] = fmn(smin[tid], smin[tid + st]);
      __syncthreads();
    }
    if (tid == 0) {
      long blk = wg.x + (long)wg.nx * wg.y;
      b.red[blk] = smin[0];
    }
  }
}
__global__ void __launch_bounds__(256) k_dt(Blk b) {
  k_dt_b(b, FEST_WG_HW);
}

// global step: minimum of the block minima, broadcast into delta_t
__global__ void k_dt_min(Blk b, long nblk) {
  __shared__ double s[256];
  double m = 1e300;
  for (long t = threadIdx.x; t < nblk; t += blockDim.x) m = fmn(m, b.red[t]);
  s[threadIdx.x] = m;
  __syncthreads();
  for (int st = 128; st > 0; st >>= 1) {
    if (threadIdx.x < st) s[threadIdx.x] = fmn(s[threadIdx.x], s[threadIdx.x + st]);
    __syncthreads();
  }
  if (threadIdx.x == 0) b.dtmin[0] = s[0];
}
__global__ void k_dt_fill(Blk b, int use_fixed) {
  long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= b.nres) return;
  b.dt[t] = use_fixed ? b.global_dt : b.dtmin[0];
}

enum { UPD_FIRST = 1, UPD_LAST = 2, UPD_WRITE_RES = 4, UPD_APPLY = 8, UPD_STORE = 16 };

// the SA point-implicit destruction factor of the update (src/update.f90:
// 405-420): xi takes the conservative nu-tilde times the current density
__device__ __forceinline__ double sa_implicit(const Blk &b, long c, long gi, double u1_0, double u1_5, double dt) {
  const double *gx = b.g[0], *gy = b.g[1], *gz = b.g[2];
  const long ngr = b.ngr;
  double vort = sqrt(sqd(gy[2 * ngr + gi] - gz[ngr + gi]) + sqd(gz[gi] - gx[2 * ngr + gi]) +
                     sqd(gx[ngr + gi] - gy[gi]));
  double dist = b.dist[c];
  double kd = Sa::kappa * dist, kd2 = kd * kd;
  double xi = u1_5 * b.qp[c] / b.mu[c];
  double x3 = (xi * xi) * xi;
  double fv1 = x3 / (x3 + Sa::cv1_3);
  double fv2 = 1.0 - xi / (1 + xi * fv1);
  double scap = vort + u1_5 * fv2 / kd2;
  double rsa = fmn(u1_5 / (scap * kd2), 10.0);
  double r6 = ((((rsa * rsa) * rsa) * rsa) * rsa) * rsa;
  double g = rsa + Sa::cw2 * (r6 - rsa);
  double g6 = ((((g * g) * g) * g) * g) * g;
  double fw = g * pow((1.0 + Sa::cw3_6) / (g6 + Sa::cw3_6), 1.0 / 6.0);
  return 1. + ((-1.0 * u1_0 * Sa::cb1 * scap) + (2.0 * u1_0 * Sa::cw1 * fw * u1_5 / (dist * dist))) * dt;
}

// residue + SST/SA source, then (UPD_APPLY) the conservative update of the cell
__device__ __forceinline__ void k_update_b(const Blk &b, const Wg &wg, double TF, double SF, int flags) {
  int i, j;
  if (!flat_ij(b.imx - 1, b.jmx - 1, i, j, plane_block(wg, b.xcd_remap))) return;
  i += 1;
  j += 1;
  const int k = wg.y + 1;
  if (k > b.kmx - 1) return;
  const int nv = b.nv;
  const long nc = b.nc;
  long c = ci(b, i, j, k);
  long r = ri(b, i, j, k);
  long F0 = fli(b, 0, i, j, k), G0 = fli(b, 1, i, j, k), H0 = fli(b, 2, i, j, k);
  long sG = b.nfi[1], sH = (long)b.nfi[2] * b.nfj[2];
  double R[8];
  if (b.fp32) {
    // FP32 mode: the face fluxes are stored in FP32, the sums taken in FP64
    for (int l = 0; l < nv; ++l) {
      const float *F = (const float *)b.flux[0] + (long)l * b.nfl[0];
      const float *G = (const float *)b.flux[1] + (long)l * b.nfl[1];
      const float *H = (const float *)b.flux[2] + (long)l * b.nfl[2];
      R[l] = ((double)F[F0 + 1] - (double)F[F0]) + ((double)G[G0 + sG] - (double)G[G0]) +
             ((double)H[H0 + sH] - (double)H[H0]);
    }
  } else {
    for (int l = 0; l < nv; ++l) {
      const double *F = b.flux[0] + (long)l * b.nfl[0];
      const double *G = b.flux[1] + (long)l * b.nfl[1];
      const double *H = b.flux[2] + (long)l * b.nfl[2];
      R[l] = (F[F0 + 1] - F[F0]) + (G[G0 + sG] - G[G0]) + (H[H0 + sH] - H[H0]);
    }
  }
  double rho = b.qp[c];
  const bool sst = b.ng >= 6, sa = b.ng == 5;
  if (sst) {
    // SST source (src/source.f90:158-270), computed with the gradients
    R[5] = R[5] - b.src[r];
    R[6] = R[6] - b.src[b.nres + r];
  }
  if (b.lctm) R[7] = R[7] - b.src[2 * b.nres + r];  // src/source.f90:454-456
  if (sa) R[5] = R[5] - b.src[r];  // SA source (src/source.f90:835-983)
  if (flags & UPD_WRITE_RES)
    for (int l = 0; l < nv; ++l) b.res[(long)l * b.nres + r] = R[l];
  if (!(flags & UPD_APPLY)) return;
  // src/update.f90:367-485
  const double gm = b.gm;
  const double *Qu = (flags & UPD_STORE) ? b.ust : b.qp;
  double u1[8], u2[8];
  u1[0] = Qu[c];
  for (int l = 1; l < nv; ++l) u1[l] = Qu[(long)l * nc + c] * u1[0];
  u1[4] = (u1[4] / (gm - 1.) + 0.5 * (u1[1] * u1[1] + u1[2] * u1[2] + u1[3] * u1[3])) / u1[0] + 0.0;
  double dt = b.dt[r];
  if (sa) R[5] = R[5] / sa_implicit(b, c, gri(b, i, j, k), u1[0], u1[5], dt);
  if (sst && b.kkl) {
    // k-kL point-implicit destruction, sqrt(rho) in the k factor
    // (src/update.f90:399-404)
    const double mu = b.mu[c], dist = b.dist[c];
    const double eta = u1[0] * dist * (sqrt(0.3 * u1[5]) / (20 * mu));
    const double fphi = (1 + 4.7 * eta) / (1 + ((eta * eta) * eta) * eta);
    const double d2 = sqd(dist);
    R[5] = R[5] / (1. + ((2.5 * ((pow(0.09, 0.75)) * sqrt(u1[0]) * (pow(u1[5], 1.5)) / fmx(u1[6], 1.e-20)) +
                          (2 * mu / d2)) * dt));
    R[6] = R[6] / (1. + (6 * mu * fphi / d2) * dt);
  } else if (sst) {
    double F1 = b.f1[c];
    double beta = 0.075 * F1 + (1. - F1) * 0.0828;
    double w7 = b.qp[6 * nc + c];
    R[5] = R[5] / (1 + (beta * w7 * dt));
    R[6] = R[6] / (1 + (2 * beta * w7 * dt));
  }
  if (flags & UPD_STORE) {
    for (int l = 0; l < nv; ++l) {
      double rn = (flags & UPD_FIRST) ? 0.0 + SF * R[l] : b.rst[(long)l * b.nres + r] + SF * R[l];
      if (flags & UPD_LAST) R[l] = rn;
      else b.rst[(long)l * b.nres + r] = rn;
    }
  }
  double fac = TF * dt / b.vol[c];
  for (int l = 0; l < nv; ++l) u2[l] = u1[l] - R[l] * fac;
  for (int l = 1; l < nv; ++l) u2[l] = u2[l] / u2[0];
  u2[4] = (gm - 1.) * u2[0] * (u2[4] - (0.5 * (u2[1] * u2[1] + u2[2] * u2[2] + u2[3] * u2[3])) - 0.0);
  bool bad = (u2[0] < 0.) || (u2[4] < 0.);
  for (int l = 0; l < nv; ++l) bad |= isnan(u2[l]);
  if (bad) {
    atomicOr(b.err, ERR_NONPHYS);
    return;
  }
  // variables 1-5 and the two turbulence variables; lctm2015's
  // intermittency is not written back (src/update.f90:456-484)
  for (int l = 0; l < 5; ++l) b.qp[(long)l * nc + c] = u2[l];
  if (sst) {
    if (u2[5] >= 0.) b.qp[5 * nc + c] = u2[5];
    if (u2[6] >= 0.) b.qp[6 * nc + c] = u2[6];
  }
  if (sa) b.qp[5 * nc + c] = fmx(u2[5], 1.e-12);
}
__global__ void __launch_bounds__(256) k_update(Blk b, double TF, double SF, int flags) {
  k_update_b(b, FEST_WG_HW, TF, SF, flags);
}

// per-block sums of residue^2 and of the boundary mass fluxes: blocks take
// whole i rows (j, k) of the residue in turn, threads the cells of a row.
// NORM_ACC sums per block: the mass imbalance and up to 8 variables.
constexpr int NORM_ACC = 9;
__global__ void __launch_bounds__(256) k_norm_partial(Blk b, double *red) {
  __shared__ double s[NORM_ACC][256];
  const int nv = b.nv;
  double acc[NORM_ACC] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
  const int ni = b.imx - 1, nj = b.jmx - 1;
  const long rows = (long)nj * (b.kmx - 1);
  for (long r = blockIdx.x; r < rows; r += gridDim.x) {
    const int j = (int)(r % nj) + 1, k = (int)(r / nj) + 1;
    const long row = (r) * (long)ni;
    for (int i = 1 + (int)threadIdx.x; i <= ni; i += blockDim.x) {
      const long t = row + (i - 1);
#pragma unroll
      for (int l = 0; l < NORM_ACC - 1; ++l)
        if (l < nv) {
          double v = b.res[(long)l * b.nres + t];
          acc[1 + l] += v * v;
        }
      // the mass fluxes (FP32 in the FP32 mode)
      auto fl = [&](int d, long o) -> double {
        return b.fp32 ? (double)((const float *)b.flux[d])[o] : b.flux[d][o];
      };
      if (i == 1) acc[0] += fl(0, fli(b, 0, 1, j, k));
      if (i == ni) acc[0] -= fl(0, fli(b, 0, b.imx, j, k));
      if (j == 1) acc[0] += fl(1, fli(b, 1, i, 1, k));
      if (j == nj) acc[0] -= fl(1, fli(b, 1, i, b.jmx, k));
      if (k == 1) acc[0] += fl(2, fli(b, 2, i, j, 1));
      if (k == b.kmx - 1) acc[0] -= fl(2, fli(b, 2, i, j, b.kmx));
    }
  }
#pragma unroll
  for (int l = 0; l < NORM_ACC; ++l) s[l][threadIdx.x] = acc[l];
  __syncthreads();
  for (int st = 128; st > 0; st >>= 1) {
    if (threadIdx.x < st)
#pragma unroll
      for (int l = 0; l < NORM_ACC; ++l) s[l][threadIdx.x] += s[l][threadIdx.x + st];
    __syncthreads();
  }
  if (threadIdx.x < NORM_ACC) red[(long)blockIdx.x * NORM_ACC + threadIdx.x] = s[threadIdx.x][0];
}

// final norms, src/resnorm.f90:128-225: 256 threads sum the block
// partials in a fixed order (strided, then a tree).  out receives the n_var+1
// norms and, with shares, from out[NORM_SHARE] the block's unassembled share.
__global__ void __launch_bounds__(256) k_norm_final(Blk b, int nblk, const double *red, double *out, int shares) {
  __shared__ double s[NORM_ACC][256];
  double part[NORM_ACC] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
  for (int t = threadIdx.x; t < nblk; t += blockDim.x)
#pragma unroll
    for (int l = 0; l < NORM_ACC; ++l) part[l] += red[(long)t * NORM_ACC + l];
#pragma unroll
  for (int l = 0; l < NORM_ACC; ++l) s[l][threadIdx.x] = part[l];
  __syncthreads();
  for (int st = 128; st > 0; st >>= 1) {
    if (threadIdx.x < st)
#pragma unroll
      for (int l = 0; l < NORM_ACC; ++l) s[l][threadIdx.x] += s[l][threadIdx.x + st];
    __syncthreads();
  }
  if (threadIdx.x != 0) return;
  double acc[NORM_ACC];
#pragma unroll
  for (int l = 0; l < NORM_ACC; ++l) acc[l] = s[l][0];
  double v = b.vel_mag;
  double scale[NORM_ACC];
  scale[0] = 1.;
  scale[1] = b.rho_inf * v;
  scale[2] = scale[3] = scale[4] = b.rho_inf * v * v;
  scale[5] = (0.5 * b.rho_inf * (v * v * v) + ((b.gm / (b.gm - 1.)) * b.p_inf));
  scale[6] = b.rho_inf * v * (b.ng == 5 ? b.tv_inf : b.tk_inf);  // nu-tilde for SA (resnorm.f90:158)
  scale[7] = b.rho_inf * v * (b.kkl ? b.tkl_inf : b.tw_inf);  // kL for k-kL (resnorm.f90:153)
  // lctm2015's intermittency: resnorm.f90:128-168 leaves Res_scale(8)
  // unset, so the reference's value is undefined; scaled here like the
  // other turbulence variables
  scale[8] = b.rho_inf * v * b.tgm_inf;
  out[0] = fabs(acc[0] / scale[0]);
  for (int l = 1; l <= b.nv; ++l) out[l] = sqrt(acc[l] / (scale[l] * scale[l]));
  if (!shares) return;
  // the block's share before the cross-block assembly (src/resnorm.f90:
  // 171-225): the signed mass imbalance and sum(R_l^2)/scale_l^2
  out[NORM_SHARE] = acc[0] / scale[0];
  for (int l = 1; l <= b.nv; ++l) out[NORM_SHARE + l] = acc[l] / (scale[l] * scale[l]);
}

}  // namespace fest
