// k_bc.h -- ghost-cell boundary conditions (one launch per block face, in
// the reference's face order imin, imax, jmin, jmax, kmin, kmax), then the
// edge/corner averages.  src/boundary/bc_primitive.f90:55-227 and the
// routines it calls; every face kernel gives one thread a column of three
// ghost cells, which only reads its own column, so the per-face passes keep
// the reference's sequential semantics.  The bodies and helpers are here,
// their per-block kernels in k_bc_kernels.h (one translation unit).
#pragma once
// ghost cells and geometry keep the reference's unfused operation order
#pragma clang fp contract(off)
#include "blk.h"

namespace fest {

// cell of layer L (L >= 1 interior counted from the face, L <= 0 ghost) at
// tangential position (a, bb) of a face
__device__ __forceinline__ long fcell(const Blk &b, int face, int L, int a, int bb) {
  switch (face) {
    case 0: return ci(b, L, a, bb);
    case 1: return ci(b, b.imx - L, a, bb);
    case 2: return ci(b, a, L, bb);
    case 3: return ci(b, a, b.jmx - L, bb);
    case 4: return ci(b, a, bb, L);
    default: return ci(b, a, bb, b.kmx - L);
  }
}
__host__ __device__ inline int face_na(const Blk &b, int face) { return face < 2 ? b.jmx : b.imx; }
__host__ __device__ inline int face_nb(const Blk &b, int face) { return face < 4 ? b.kmx : b.jmx; }

enum { CP_ANTI = 0, CP_FLAT = 1, CP_SYMM = 2 };

// boundary id and fixed values of one face, gathered with compile-time
// indices: indexing the by-value kernel argument with a run-time face would
// make the compiler copy the whole argument block to private memory
struct FaceBC {
  int id;
  double fix[FX_N];
};
__device__ __forceinline__ FaceBC face_params(const Blk &b, int face) {
  FaceBC p;
  p.id = b.bc[0];
#pragma unroll
  for (int l = 0; l < FX_N; ++l) p.fix[l] = b.fix[l][0];
#pragma unroll
  for (int f = 1; f < 6; ++f)
    if (face == f) {
      p.id = b.bc[f];
#pragma unroll
      for (int l = 0; l < FX_N; ++l) p.fix[l] = b.fix[l][f];
    }
  return p;
}

// src/boundary/copy_bc.f90:58-131 on one column
__device__ __forceinline__ void copy3col(const Blk &b, double *v, int type, int face, int a, int bb) {
  double a1, a2, a3;
  if (type == CP_ANTI) { a1 = 1.; a2 = -1.; a3 = 0.; }
  else if (type == CP_FLAT) { a1 = 1.; a2 = 1.; a3 = 0.; }
  else { a1 = b.c1; a2 = b.c2; a3 = b.c3; }
  // ghost layer g takes interior layer g+1 (anti, symm) or layer 1 (flat)
#pragma unroll
  for (int g = 0; g < 3; ++g) {
    const int src = type == CP_FLAT ? 1 : g + 1;
    v[fcell(b, face, -g, a, bb)] = (a2 * v[fcell(b, face, src, a, bb)] - a3 * v[fcell(b, face, src + 1, a, bb)]) / a1;
  }
}

// src/boundary/bc_primitive.f90:484-525 on one column
__device__ __forceinline__ void fixcol(const Blk &b, double *v, double val, int face, int a, int bb) {
#pragma unroll
  for (int g = 0; g < 3; ++g) v[fcell(b, face, -g, a, bb)] = val;
}

// the turbulence variables: k and omega (SST), k and kL (k-kL), nu-tilde
// (SA); lctm2015's intermittency takes a flat copy wherever they copy
// (bc_primitive.f90:302, 385, 437, 475, 555) and its own fixed value
// wherever they are fixed (:262, :343).  kl_tw: the subsonic inlet fixes
// k-kL's kL from the omega values (src/boundary/bc_primitive.f90:333-336),
// which k-kL never reads from the boundary file (read_bc.f90:98-108)
__device__ __forceinline__ bool two_eq(const Blk &b) { return b.ng >= 6; }
__device__ __forceinline__ void turb_copycol(const Blk &b, int type, int face, int a, int bb) {
  if (b.ng >= 5) copy3col(b, b.qp + 5 * b.nc, type, face, a, bb);
  if (two_eq(b)) copy3col(b, b.qp + 6 * b.nc, type, face, a, bb);
  if (b.lctm) copy3col(b, b.qp + 7 * b.nc, CP_FLAT, face, a, bb);
}
__device__ __forceinline__ void turb_fixcol(const Blk &b, const FaceBC &p, int face, int a, int bb, bool kl_tw = false) {
  if (two_eq(b)) {
    fixcol(b, b.qp + 5 * b.nc, p.fix[FX_TK], face, a, bb);
    fixcol(b, b.qp + 6 * b.nc, (b.kkl && !kl_tw) ? p.fix[FX_TKL] : p.fix[FX_TW], face, a, bb);
  }
  if (b.ng == 5) fixcol(b, b.qp + 5 * b.nc, p.fix[FX_TV], face, a, bb);
  if (b.lctm) fixcol(b, b.qp + 7 * b.nc, p.fix[FX_TGM], face, a, bb);
}

// unit normal of the boundary face at (a, bb); the metric planes are picked
// per case with constant member indices (a run-time index into the
// by-value argument's pointer arrays would put them in private memory)
__device__ __forceinline__ void bnormal(const Blk &b, int face, int a, int bb, double &nx, double &ny, double &nz) {
  long f;
  const double *px, *py, *pz;
  switch (face) {
    case 0: f = fi(b, 0, 1, a, bb); px = b.fx[0]; py = b.fy[0]; pz = b.fz[0]; break;
    case 1: f = fi(b, 0, b.imx, a, bb); px = b.fx[0]; py = b.fy[0]; pz = b.fz[0]; break;
    case 2: f = fi(b, 1, a, 1, bb); px = b.fx[1]; py = b.fy[1]; pz = b.fz[1]; break;
    case 3: f = fi(b, 1, a, b.jmx, bb); px = b.fx[1]; py = b.fy[1]; pz = b.fz[1]; break;
    case 4: f = fi(b, 2, a, bb, 1); px = b.fx[2]; py = b.fy[2]; pz = b.fz[2]; break;
    default: f = fi(b, 2, a, bb, b.kmx); px = b.fx[2]; py = b.fy[2]; pz = b.fz[2]; break;
  }
  nx = px[f]; ny = py[f]; nz = pz[f];
}

// src/boundary/bc_primitive.f90:1778-1945
__device__ __forceinline__ void temp_based_density_col(const Blk &b, const FaceBC &p_, int face, int a, int bb) {
  double Tw = p_.fix[FX_TWALL], R = b.R, gm = b.gm;
  double *rho = b.qp, *p = b.qp + 4 * b.nc;
  if (Tw < 0.0) {
    long c1 = fcell(b, face, 1, a, bb);
    double stag = (p[c1] / (R * rho[c1])) * (1 + (0.5 * (gm - 1.) * gm * p[c1] / rho[c1]));
    for (int g = 0; g < 3; ++g) {
      long cg = fcell(b, face, -g, a, bb);
      rho[cg] = p[cg] / (R * stag);
    }
  } else if (Tw > 1.0) {
    for (int g = 0; g < 3; ++g) {
      long cg = fcell(b, face, -g, a, bb), cn = fcell(b, face, 1 + g, a, bb);
      rho[cg] = p[cg] / (R * (2 * Tw - (p[cn] / (R * rho[cn]))));
    }
  } else {
    copy3col(b, rho, CP_SYMM, face, a, bb);
  }
}

// src/boundary/bc_primitive.f90:564-643 (Sutherland mu whatever mu_variation)
__device__ __forceinline__ void omega_wall_col(const Blk &b, int face, int a, int bb) {
  double *rho = b.qp, *p = b.qp + 4 * b.nc, *tw = b.qp + 6 * b.nc;
  long c0 = fcell(b, face, 0, a, bb), c1 = fcell(b, face, 1, a, bb);
  for (int l = 1; l <= 3; ++l) {
    double T_face = 0.5 * ((p[c0] / rho[c0]) + (p[c1] / rho[c1])) / b.R;
    double mu = b.mu_ref * pow(T_face / b.T_ref, 1.5) * ((b.T_ref + b.S_ref) / (T_face + b.S_ref));
    double r = 0.5 * (rho[c0] + rho[c1]);
    double d2 = 2 * b.dist[c1];
    tw[fcell(b, face, 1 - l, a, bb)] = 120 * mu / (r * 0.075 * (d2 * d2)) - tw[fcell(b, face, l, a, bb)];
  }
}

// far-field Riemann state of one column (src/boundary/bc_primitive.f90:645-1234)
// or, when total, the total-pressure inlet (bc_primitive.f90:1237-1776):
// the same Riemann-invariant velocity, then the pressure from the total
// pressure tp and the boundary Mach number and the density from the
// boundary speed of sound.  c is the first interior cell's rho, u, v, w, p
// and nrm the boundary face's stored unit normal (bnormal); returns Unb and,
// when full, the ghost layer 0 state g.  Written for a type T: double, or
// the dual numbers of the adjoint (k_dual.h)
template <typename T>
__device__ __forceinline__ T riemann_state(const Blk &b, int face, const T (&c)[5], const T (&nrm)[3], bool full,
                                           bool total, double tp, T (&g)[5]) {
  const double gm = b.gm;
  T nx = nrm[0], ny = nrm[1], nz = nrm[2];
  if ((face & 1) == 0) { nx = -nx; ny = -ny; nz = -nz; }
  T cexp = sqrt(gm * c[4] / c[0]);
  double cinf = sqrt(gm * b.p_inf / b.rho_inf);
  T Unexp = c[1] * nx + c[2] * ny + c[3] * nz;
  T Uninf = b.u_inf * nx + b.v_inf * ny + b.w_inf * nz;
  T Rinf = Uninf - 2 * cinf / (gm - 1.);
  T Rexp = Unexp + 2 * cexp / (gm - 1.);
  T Unb = 0.5 * (Rexp + Rinf);
  T Cb = 0.25 * (gm - 1.) * (Rexp - Rinf);
  if (!full) return Unb;
  if (Unb > 0.) {
    T vd = Unb - Unexp;
    g[1] = c[1] + vd * nx;
    g[2] = c[2] + vd * ny;
    g[3] = c[3] + vd * nz;
    if (!total) {
      T sen = c[4] / pow(c[0], gm);
      g[0] = pow(Cb * Cb / (gm * sen), 1. / (gm - 1.));
      g[4] = (g[0] * Cb * Cb / gm);
    }
  } else {
    T vd = Unb - Uninf;
    g[1] = b.u_inf + vd * nx;
    g[2] = b.v_inf + vd * ny;
    g[3] = b.w_inf + vd * nz;
    if (!total) {
      double sen = b.p_inf / pow(b.rho_inf, gm);
      g[0] = pow(Cb * Cb / (gm * sen), 1. / (gm - 1.));
      g[4] = (g[0] * Cb * Cb / gm);
    }
  }
  if (total) {
    T Mb = sqrt(g[1] * g[1] + g[2] * g[2] + g[3] * g[3]) / Cb;
    g[4] = tp / pow((1 + 0.5 * (gm - 1.) * Mb * Mb), gm / (gm - 1.));
    g[0] = gm * g[4] / (Cb * Cb);
  }
  return Unb;
}
// riemann_state on the resident state; when write != 0 it stores layer 0
// of rho, u, v, w, p
__device__ __forceinline__ double riemann_col(const Blk &b, int face, int a, int bb, int write, bool total, double tp) {
  const long c1 = fcell(b, face, 1, a, bb), cg = fcell(b, face, 0, a, bb);
  double c[5], g[5], nrm[3];
#pragma unroll
  for (int l = 0; l < 5; ++l) c[l] = b.qp[(long)l * b.nc + c1];
  bnormal(b, face, a, bb, nrm[0], nrm[1], nrm[2]);
  const double Unb = riemann_state(b, face, c, nrm, write != 0, total, tp, g);
  if (write) {
#pragma unroll
    for (int l = 0; l < 5; ++l) b.qp[(long)l * b.nc + cg] = g[l];
  }
  return Unb;
}

// src/boundary/FT_bc.f90:36-106; j and k faces reflect with the I-face
// normal stored at the same index, as the reference does
__device__ __forceinline__ void flow_tangency_col(const Blk &b, int face, int a, int bb) {
  double *u = b.qp + b.nc, *v = b.qp + 2 * b.nc, *w = b.qp + 3 * b.nc;
  double nx, ny, nz;
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
  double rx = b.fx[0][f], ry = b.fy[0][f], rz = b.fz[0][f];
  for (int l = 1; l <= 3; ++l) {
    long c = fcell(b, face, l, a, bb), g = fcell(b, face, 1 - l, a, bb);
    double dot = u[c] * nx + v[c] * ny + w[c] * nz;
    u[g] = u[c] - (2.0 * dot * rx);
    v[g] = v[c] - (2.0 * dot * ry);
    w[g] = w[c] - (2.0 * dot * rz);
  }
}

// One face of populate_ghost_primitive.  The grid spans the padded
// tangential plane (-2..na+2) x (-2..nb+2); most conditions act on the face
// interior 1..na-1 x 1..nb-1, the periodic and far-field plane copies on the
// whole plane.
// blockIdx.z selects face0 + z: the min and max faces of one direction run
// in one launch when they cannot touch each other's cells (see the caller).
__device__ __forceinline__ void k_bc_face_b(const Blk &b, const Wg &wg, int face0, int current_iter) {
  const int face = face0 + (int)wg.z;
  int a = (int)(wg.x * blockDim.x + threadIdx.x) - 2;
  int bb = (int)(wg.y * blockDim.y + threadIdx.y) - 2;
  int na = face_na(b, face), nb = face_nb(b, face);
  if (a > na + 2 || bb > nb + 2) return;
  const FaceBC p = face_params(b, face);
  const int id = p.id;
  bool inner = (a >= 1 && a <= na - 1 && bb >= 1 && bb <= nb - 1);
  double *q = b.qp;
  const long nc = b.nc;
  switch (id) {
    case -1:  // supersonic inlet
      if (inner && current_iter <= 2) {
        for (int l = 0; l < 5; ++l) fixcol(b, q + l * nc, p.fix[l], face, a, bb);
        turb_fixcol(b, p, face, a, bb);
      }
      break;
    case -2:  // supersonic outlet
      if (inner) {
        for (int l = 0; l < 5; ++l) copy3col(b, q + l * nc, CP_FLAT, face, a, bb);
        turb_copycol(b, CP_FLAT, face, a, bb);
      }
      break;
    case -3:  // subsonic inlet
      if (inner) {
        if (current_iter <= 2) {
          for (int l = 0; l < 4; ++l) fixcol(b, q + l * nc, p.fix[l], face, a, bb);
          turb_fixcol(b, p, face, a, bb, true);
        }
        copy3col(b, q + 4 * nc, CP_FLAT, face, a, bb);
      }
      break;
    case -4:  // subsonic outlet
      if (inner) {
        for (int l = 0; l < 4; ++l) copy3col(b, q + l * nc, CP_FLAT, face, a, bb);
        if (current_iter <= 2) fixcol(b, q + 4 * nc, p.fix[FX_P], face, a, bb);
        turb_copycol(b, CP_FLAT, face, a, bb);
      }
      break;
    case -5:  // no-slip wall
      if (inner) {
        copy3col(b, q + 4 * nc, CP_SYMM, face, a, bb);
        temp_based_density_col(b, p, face, a, bb);
        copy3col(b, q + 1 * nc, CP_ANTI, face, a, bb);
        copy3col(b, q + 2 * nc, CP_ANTI, face, a, bb);
        copy3col(b, q + 3 * nc, CP_ANTI, face, a, bb);
        if (b.ng >= 5) copy3col(b, q + 5 * nc, CP_ANTI, face, a, bb);  // k or nu-tilde
        if (two_eq(b) && b.kkl) copy3col(b, q + 6 * nc, CP_ANTI, face, a, bb);  // kL, :548-550
        else if (two_eq(b)) omega_wall_col(b, face, a, bb);
        if (b.lctm) copy3col(b, q + 7 * nc, CP_FLAT, face, a, bb);  // :553-558
      }
      break;
    case -6:  // slip wall
      if (inner) {
        copy3col(b, q, CP_SYMM, face, a, bb);
        copy3col(b, q + 4 * nc, CP_SYMM, face, a, bb);
        turb_copycol(b, CP_SYMM, face, a, bb);
        flow_tangency_col(b, face, a, bb);
      }
      break;
    case -7:  // pole
      if (inner) {
        for (int l = 0; l < 5; ++l) copy3col(b, q + l * nc, CP_FLAT, face, a, bb);
        turb_copycol(b, CP_FLAT, face, a, bb);
      }
      break;
    case -8:     // far field
    case -11: {  // total-pressure inlet
      const bool total = id == -11;
      const double tp = p.fix[FX_TP];
      if (inner) {
        riemann_col(b, face, a, bb, 1, total, tp);
        if (b.turb) {
          // the whole-face turbulent copy/fix of the reference ends as set by
          // the last face cell in its loop order
          double last = riemann_col(b, face, na - 1, nb - 1, 0, total, tp);
          if (last > 0.) turb_copycol(b, CP_FLAT, face, a, bb);
          else turb_fixcol(b, p, face, a, bb);
        }
      }
      for (int l = 0; l < b.nv; ++l) {
        double v0 = q[l * nc + fcell(b, face, 0, a, bb)];
        q[l * nc + fcell(b, face, -1, a, bb)] = v0;
        q[l * nc + fcell(b, face, -2, a, bb)] = v0;
      }
      break;
    }
    case -9: {  // single-block periodic: whole plane, three layers
      // the reference's array assignment reads its source layers whole
      // before writing; with one or two cells in the direction they overlap
      // the destination layers.  Layers g = 0..2 sit at d0 + g*st (ghosts)
      // and s0 + g*st (their sources) along the face normal's stride st.
      long d0, s0, st;
      switch (face) {
        case 0: d0 = ci(b, -2, a, bb); s0 = ci(b, b.imx - 3, a, bb); st = 1; break;
        case 1: d0 = ci(b, b.imx, a, bb); s0 = ci(b, 1, a, bb); st = 1; break;
        case 2: d0 = ci(b, a, -2, bb); s0 = ci(b, a, b.jmx - 3, bb); st = b.pi; break;
        case 3: d0 = ci(b, a, b.jmx, bb); s0 = ci(b, a, 1, bb); st = b.pi; break;
        case 4: d0 = ci(b, a, bb, -2); s0 = ci(b, a, bb, b.kmx - 3); st = (long)b.pi * b.pj; break;
        default: d0 = ci(b, a, bb, b.kmx); s0 = ci(b, a, bb, 1); st = (long)b.pi * b.pj; break;
      }
      for (int l = 0; l < b.nv; ++l) {
        double *v = q + l * nc;
        const double v0 = v[s0], v1 = v[s0 + st], v2 = v[s0 + 2 * st];
        v[d0] = v0;
        v[d0 + st] = v1;
        v[d0 + 2 * st] = v2;
      }
      break;
    }
    default:
      break;  // interfaces (>=0) and -10 are filled by the interface exchange
  }
}

// edges and corners, src/boundary/bc_primitive.f90:209-224: phase 0 the
// four lines along i, phase 1 the four lines along k, phase 2 the corners
__device__ __forceinline__ void k_bc_edges_b(const Blk &b, const Wg &wg, int phase) {
  int t = wg.x * blockDim.x + threadIdx.x;
  const int I = b.imx, J = b.jmx, K = b.kmx;
#define Q(i, j, k) v[ci(b, i, j, k)]
  if (phase == 0) {
    int n = I + 5;
    if (t >= 4 * n * b.nv) return;
    int l = t / (4 * n), r = t % (4 * n), s = r / n, i = r % n - 2;
    double *v = b.qp + (long)l * b.nc;
    if (s == 0) Q(i, 0, 0) = 0.33 * (Q(i, 1, 1) + Q(i, 0, 1) + Q(i, 1, 0));
    else if (s == 1) Q(i, 0, K) = 0.33 * (Q(i, 1, K - 1) + Q(i, 0, K - 1) + Q(i, 1, K));
    else if (s == 2) Q(i, J, 0) = 0.33 * (Q(i, J - 1, 1) + Q(i, J, 1) + Q(i, J - 1, 0));
    else Q(i, J, K) = 0.33 * (Q(i, J - 1, K - 1) + Q(i, J, K - 1) + Q(i, J - 1, K));
  } else if (phase == 1) {
    int n = K + 5;
    if (t >= 4 * n * b.nv) return;
    int l = t / (4 * n), r = t % (4 * n), s = r / n, k = r % n - 2;
    double *v = b.qp + (long)l * b.nc;
    if (s == 0) Q(I, 0, k) = 0.33 * (Q(I - 1, 1, k) + Q(I - 1, 0, k) + Q(I, 1, k));
    else if (s == 1) Q(0, 0, k) = 0.33 * (Q(1, 1, k) + Q(1, 0, k) + Q(0, 1, k));
    else if (s == 2) Q(0, J, k) = 0.33 * (Q(1, J - 1, k) + Q(1, J, k) + Q(0, J - 1, k));
    else Q(I, J, k) = 0.33 * (Q(I - 1, J - 1, k) + Q(I - 1, J, k) + Q(I, J - 1, k));
  } else {
    if (t >= 8 * b.nv) return;
    int l = t / 8, s = t % 8;
    double *v = b.qp + (long)l * b.nc;
    switch (s) {
      case 0: Q(0, 0, 0) = 0.33 * (Q(1, 0, 0) + Q(0, 1, 0) + Q(0, 0, 1)); break;
      case 1: Q(I, 0, 0) = 0.33 * (Q(I - 1, 0, 0) + Q(I, 1, 0) + Q(I, 0, 1)); break;
      case 2: Q(0, J, 0) = 0.33 * (Q(1, J, 0) + Q(0, J - 1, 0) + Q(0, J, 1)); break;
      case 3: Q(0, 0, K) = 0.33 * (Q(1, 0, K) + Q(0, 1, K) + Q(0, 0, K - 1)); break;
      case 4: Q(I, J, 0) = 0.33 * (Q(I - 1, J, 0) + Q(I, J - 1, 0) + Q(I, J, 1)); break;
      case 5: Q(I, 0, K) = 0.33 * (Q(I - 1, 0, K) + Q(I, 1, K) + Q(I, 0, K - 1)); break;
      case 6: Q(0, J, K) = 0.33 * (Q(1, J, K) + Q(0, J - 1, K) + Q(0, J, K - 1)); break;
      default: Q(I, J, K) = 0.33 * (Q(I - 1, J, K) + Q(I, J - 1, K) + Q(I, J, K - 1)); break;
    }
  }
#undef Q
}

}  // namespace fest
