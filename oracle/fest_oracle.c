/*
 * fest_oracle.c -- ORACLE (test infrastructure only).
 *
 * A plain-C, single-threaded restatement of FEST-3D's explicit residual
 * evaluation and RK stage update.  It follows the reference pass by pass and
 * in the reference's operation order, so that the GPU path can be checked
 * against it.  Only tests/, __graft_entry__.smoke() and bench.py's
 * cpu_baseline leg may load it, and only as the checker.
 *
 * Parity pins (see tests/): each pass is compared against the amdflang build
 * of the reference modules (oracle/_ref, when /root/reference is present),
 * against the reference's own known-answer tests (tests/test_*.f90), and
 * against closed forms for the parts the reference build cannot run
 * (src/time.f90 and src/update.f90 use MPI and do not link here): the RK4
 * amplification polynomial, the uniform-grid time step and the norms.
 * Parity unpinned: the point-implicit turbulence factors of the RK update
 * (SA, src/update.f90:405-420; k-kL, :399-404), which only update.f90 holds.
 *
 * Semantics choices that the reference leaves to the compiler:
 *   - MAX/MIN with a NaN argument follow amdflang's lowering, `a>b ? a : b`
 *     and `a<b ? a : b`, folded left to right.
 *   - SIGN(a,b) with b = -0.0 returns -|a| (copysign).
 *   - No floating-point contraction (build with -ffp-contract=off).
 */
#include "fest_oracle.h"
#include <math.h>
#include <stdio.h>
#include <pthread.h>
#include <stdlib.h>
#include <string.h>

struct orc {
  orc_case c;
  int imx, jmx, kmx, nv, ngrad;
  int pi, pj, pk;
  size_t ncell, nif, njf, nkf, nnode;
  size_t nxs, nys, nzs, nF, nG, nH, nR, ngr;
  double *a[ORC_NARRAYS];
  size_t n[ORC_NARRAYS];
  int *mfz, *mgz, *mhz;
  int ppm_flag;
  double c1, c2, c3;
};

/* ---------------------------------------------------------------- indexing */
#define IMX (s->imx)
#define JMX (s->jmx)
#define KMX (s->kmx)
static inline size_t CI(const orc *s, int i, int j, int k) {
  return (size_t)(i + 2) + (size_t)s->pi * ((size_t)(j + 2) + (size_t)s->pj * (size_t)(k + 2));
}
static inline size_t IFI(const orc *s, int i, int j, int k) {
  return (size_t)(i + 2) + (size_t)(s->imx + 6) * ((size_t)(j + 2) + (size_t)(s->jmx + 5) * (size_t)(k + 2));
}
static inline size_t JFI(const orc *s, int i, int j, int k) {
  return (size_t)(i + 2) + (size_t)(s->imx + 5) * ((size_t)(j + 2) + (size_t)(s->jmx + 6) * (size_t)(k + 2));
}
static inline size_t KFI(const orc *s, int i, int j, int k) {
  return (size_t)(i + 2) + (size_t)(s->imx + 5) * ((size_t)(j + 2) + (size_t)(s->jmx + 5) * (size_t)(k + 2));
}
static inline size_t NDI(const orc *s, int i, int j, int k) {
  return (size_t)(i + 2) + (size_t)(s->imx + 6) * ((size_t)(j + 2) + (size_t)(s->jmx + 6) * (size_t)(k + 2));
}
/* face-state arrays, direction d = 0 (x), 1 (y), 2 (z) */
static inline size_t XSI(const orc *s, int i, int j, int k) {
  return (size_t)i + (size_t)(s->imx + 2) * ((size_t)(j - 1) + (size_t)(s->jmx - 1) * (size_t)(k - 1));
}
static inline size_t YSI(const orc *s, int i, int j, int k) {
  return (size_t)(i - 1) + (size_t)(s->imx - 1) * ((size_t)j + (size_t)(s->jmx + 2) * (size_t)(k - 1));
}
static inline size_t ZSI(const orc *s, int i, int j, int k) {
  return (size_t)(i - 1) + (size_t)(s->imx - 1) * ((size_t)(j - 1) + (size_t)(s->jmx - 1) * (size_t)k);
}
static inline size_t FI(const orc *s, int i, int j, int k) {
  return (size_t)(i - 1) + (size_t)s->imx * ((size_t)(j - 1) + (size_t)(s->jmx - 1) * (size_t)(k - 1));
}
static inline size_t GI(const orc *s, int i, int j, int k) {
  return (size_t)(i - 1) + (size_t)(s->imx - 1) * ((size_t)(j - 1) + (size_t)s->jmx * (size_t)(k - 1));
}
static inline size_t HI(const orc *s, int i, int j, int k) {
  return (size_t)(i - 1) + (size_t)(s->imx - 1) * ((size_t)(j - 1) + (size_t)(s->jmx - 1) * (size_t)(k - 1));
}
#define RI HI
static inline size_t GRI(const orc *s, int i, int j, int k) {
  return (size_t)i + (size_t)(s->imx + 1) * ((size_t)j + (size_t)(s->jmx + 1) * (size_t)k);
}

#define QP(l, i, j, k) s->a[ORC_QP][(size_t)(l) * s->ncell + CI(s, i, j, k)]
#define QPL(l) (s->a[ORC_QP] + (size_t)(l) * s->ncell)

/* flang's lowering of MAX/MIN (see header comment) */
static inline double fmx(double a, double b) { return a > b ? a : b; }
static inline double fmn(double a, double b) { return a < b ? a : b; }
static inline double fsign(double a, double b) { return copysign(fabs(a), b); }
static inline double sq(double x) { return x * x; }

/* SST constants, src/global/global_sst.f90:6-14 */
static const double sigma_k1 = 0.85, sigma_k2 = 1.0, sigma_w1 = 0.5, sigma_w2 = 0.856;
static const double beta1 = 0.075, beta2 = 0.0828, bstar = 0.09, kappa_sst = 0.41, a1_sst = 0.31;
/* Spalart-Allmaras, src/global/global_sa.f90; cw1, cv1^3 and cw3^6 are the
 * reference's parameter expressions evaluated in double */
static const double cb1 = 0.1355, cb2 = 0.6220, cw2 = 0.3, cw3 = 2.0, cv1 = 7.1, sigma_sa = 2. / 3.,
                    kappa_sa = 0.41;
#define CW1_SA ((cb1 / (kappa_sa * kappa_sa)) + ((1 + cb2) / sigma_sa))
#define CV1_3 ((cv1 * cv1) * cv1)
#define CW3_6 (((((cw3 * cw3) * cw3) * cw3) * cw3) * cw3)
/* x**n with an integer constant n, multiplied left to right as amdflang
 * evaluates it */
static inline double ipow(double x, int n) {
  double r = x;
  for (int m = 1; m < n; ++m) r = r * x;
  return r;
}
#define SST(s) ((s)->c.turb == ORC_TURB_SST || (s)->c.turb == ORC_TURB_SST2003)
#define SA(s) ((s)->c.turb == ORC_TURB_SA)
#define KKL(s) ((s)->c.turb == ORC_TURB_KKL)
/* the lctm2015 transition model: intermittency as variable 8 (SST only) */
#define LCTM(s) ((s)->c.transition == 2)
/* src/global/global_kkl.f90:5-14; sigma_k = sigma_phi = 1 */
static const double zeta1_kkl = 1.2, zeta2_kkl = 0.97, zeta3_kkl = 0.13, cmu_kkl = 0.09, kappa_kkl = 0.41,
                    c11_kkl = 10.0, c12_kkl = 1.3, cd1_kkl = 4.7;

/* ----------------------------------------------------------- allocation */
orc *orc_create(const orc_case *c) {
  orc *s = (orc *)calloc(1, sizeof(orc));
  if (!s) return NULL;
  s->c = *c;
  s->imx = c->imx; s->jmx = c->jmx; s->kmx = c->kmx; s->nv = c->n_var;
  s->ngrad = (c->turb == ORC_TURB_SA) ? 5 : (c->turb != ORC_TURB_NONE) ? 6 : 4;
  if (c->transition == 2) s->ngrad += 1; /* the intermittency slot (src/gradients.f90:262-273) */
  s->pi = s->imx + 5; s->pj = s->jmx + 5; s->pk = s->kmx + 5;
  s->ncell = (size_t)s->pi * s->pj * s->pk;
  s->nif = (size_t)(s->imx + 6) * (s->jmx + 5) * (s->kmx + 5);
  s->njf = (size_t)(s->imx + 5) * (s->jmx + 6) * (s->kmx + 5);
  s->nkf = (size_t)(s->imx + 5) * (s->jmx + 5) * (s->kmx + 6);
  s->nnode = (size_t)(s->imx + 6) * (s->jmx + 6) * (s->kmx + 6);
  s->nxs = (size_t)(s->imx + 2) * (s->jmx - 1) * (s->kmx - 1);
  s->nys = (size_t)(s->imx - 1) * (s->jmx + 2) * (s->kmx - 1);
  s->nzs = (size_t)(s->imx - 1) * (s->jmx - 1) * (s->kmx + 2);
  s->nF = (size_t)s->imx * (s->jmx - 1) * (s->kmx - 1);
  s->nG = (size_t)(s->imx - 1) * s->jmx * (s->kmx - 1);
  s->nH = (size_t)(s->imx - 1) * (s->jmx - 1) * s->kmx;
  s->nR = (size_t)(s->imx - 1) * (s->jmx - 1) * (s->kmx - 1);
  s->ngr = (size_t)(s->imx + 1) * (s->jmx + 1) * (s->kmx + 1);
  size_t nv = (size_t)s->nv;
  size_t sz[ORC_NARRAYS] = {0};
  sz[ORC_QP] = s->ncell * nv; sz[ORC_TEMP] = s->ncell; sz[ORC_MU] = s->ncell;
  sz[ORC_MUT] = s->ncell; sz[ORC_F1] = s->ncell; sz[ORC_DIST] = s->ncell;
  sz[ORC_VOL] = sz[ORC_CX] = sz[ORC_CY] = sz[ORC_CZ] = s->ncell;
  sz[ORC_IA] = sz[ORC_INX] = sz[ORC_INY] = sz[ORC_INZ] = s->nif;
  sz[ORC_JA] = sz[ORC_JNX] = sz[ORC_JNY] = sz[ORC_JNZ] = s->njf;
  sz[ORC_KA] = sz[ORC_KNX] = sz[ORC_KNY] = sz[ORC_KNZ] = s->nkf;
  sz[ORC_XL] = sz[ORC_XR] = s->nxs * nv;
  sz[ORC_YL] = sz[ORC_YR] = s->nys * nv;
  sz[ORC_ZL] = sz[ORC_ZR] = s->nzs * nv;
  sz[ORC_F] = s->nF * nv; sz[ORC_G] = s->nG * nv; sz[ORC_H] = s->nH * nv;
  sz[ORC_RES] = s->nR * nv; sz[ORC_DT] = s->nR;
  sz[ORC_USTORE] = s->ncell * nv; sz[ORC_RSTORE] = s->nR * nv;
  sz[ORC_GX] = sz[ORC_GY] = sz[ORC_GZ] = s->ngr * (size_t)s->ngrad;
  sz[ORC_NODES] = s->nnode * 3;
  for (int id = 0; id < ORC_NARRAYS; ++id) {
    s->n[id] = sz[id];
    s->a[id] = (double *)calloc(sz[id] ? sz[id] : 1, sizeof(double));
    if (!s->a[id]) { orc_destroy(s); return NULL; }
  }
  /* mu = mu_ref initially (src/viscosity.f90:565-568) */
  for (size_t n = 0; n < s->ncell; ++n) s->a[ORC_MU][n] = c->mu_ref;
  /* src/boundary/bc.f90:45-67 */
  s->c2 = 1 + c->accur;
  s->c3 = 0.5 * c->accur;
  s->c1 = s->c2 - s->c3;
  s->mfz = (int *)malloc(sizeof(int) * (s->imx + 1));
  s->mgz = (int *)malloc(sizeof(int) * (s->jmx + 1));
  s->mhz = (int *)malloc(sizeof(int) * (s->kmx + 1));
  for (int i = 0; i <= s->imx; ++i) s->mfz[i] = 1;
  for (int j = 0; j <= s->jmx; ++j) s->mgz[j] = 1;
  for (int k = 0; k <= s->kmx; ++k) s->mhz[k] = 1;
#define WALLISH(id) ((id) == -5 || (id) == -6 || (id) == -7)
  if (WALLISH(c->bc_id[0])) s->mfz[1] = 0;
  if (WALLISH(c->bc_id[2])) s->mgz[1] = 0;
  if (WALLISH(c->bc_id[4])) s->mhz[1] = 0;
  if (WALLISH(c->bc_id[1])) s->mfz[s->imx] = 0;
  if (WALLISH(c->bc_id[3])) s->mgz[s->jmx] = 0;
  if (WALLISH(c->bc_id[5])) s->mhz[s->kmx] = 0;
#undef WALLISH
  /* src/boundary/boundary_state_reconstruction.f90:45-47 */
  s->ppm_flag = (c->recon == ORC_RECON_PPM || c->recon == ORC_RECON_WENO || c->recon == ORC_RECON_WENO_NM) ? 1 : 0;
  for (int f = 0; f < 6; ++f) if (c->bc_id[f] == -7) s->ppm_flag = 1;
  return s;
}

void orc_destroy(orc *s) {
  if (!s) return;
  for (int id = 0; id < ORC_NARRAYS; ++id) free(s->a[id]);
  free(s->mfz); free(s->mgz); free(s->mhz);
  free(s);
}

double *orc_array(orc *s, int id, size_t *count) {
  if (id < 0 || id >= ORC_NARRAYS) return NULL;
  if (count) *count = s->n[id];
  return s->a[id];
}

/* ------------------------------------------------------------- geometry */
/* src/grid.f90:137-236: linear extrapolation of ghost nodes, i then j then k */
static void ghost_grid(orc *s) {
  double *X[3] = {s->a[ORC_NODES], s->a[ORC_NODES] + s->nnode, s->a[ORC_NODES] + 2 * s->nnode};
  for (int c = 0; c < 3; ++c) {
    double *x = X[c];
    for (int k = -2; k <= KMX + 3; ++k)
      for (int j = -2; j <= JMX + 3; ++j) {
        x[NDI(s, 0, j, k)] = 2 * x[NDI(s, 1, j, k)] - x[NDI(s, 2, j, k)];
        x[NDI(s, -1, j, k)] = 2 * x[NDI(s, 0, j, k)] - x[NDI(s, 1, j, k)];
        x[NDI(s, -2, j, k)] = 2 * x[NDI(s, -1, j, k)] - x[NDI(s, 0, j, k)];
        x[NDI(s, IMX + 1, j, k)] = 2 * x[NDI(s, IMX, j, k)] - x[NDI(s, IMX - 1, j, k)];
        x[NDI(s, IMX + 2, j, k)] = 2 * x[NDI(s, IMX + 1, j, k)] - x[NDI(s, IMX, j, k)];
        x[NDI(s, IMX + 3, j, k)] = 2 * x[NDI(s, IMX + 2, j, k)] - x[NDI(s, IMX + 1, j, k)];
      }
  }
  for (int c = 0; c < 3; ++c) {
    double *x = X[c];
    for (int k = -2; k <= KMX + 3; ++k)
      for (int i = -2; i <= IMX + 3; ++i) {
        x[NDI(s, i, 0, k)] = 2 * x[NDI(s, i, 1, k)] - x[NDI(s, i, 2, k)];
        x[NDI(s, i, -1, k)] = 2 * x[NDI(s, i, 0, k)] - x[NDI(s, i, 1, k)];
        x[NDI(s, i, -2, k)] = 2 * x[NDI(s, i, -1, k)] - x[NDI(s, i, 0, k)];
        x[NDI(s, i, JMX + 1, k)] = 2 * x[NDI(s, i, JMX, k)] - x[NDI(s, i, JMX - 1, k)];
        x[NDI(s, i, JMX + 2, k)] = 2 * x[NDI(s, i, JMX + 1, k)] - x[NDI(s, i, JMX, k)];
        x[NDI(s, i, JMX + 3, k)] = 2 * x[NDI(s, i, JMX + 2, k)] - x[NDI(s, i, JMX + 1, k)];
      }
  }
  for (int c = 0; c < 3; ++c) {
    double *x = X[c];
    for (int j = -2; j <= JMX + 3; ++j)
      for (int i = -2; i <= IMX + 3; ++i) {
        x[NDI(s, i, j, 0)] = 2 * x[NDI(s, i, j, 1)] - x[NDI(s, i, j, 2)];
        x[NDI(s, i, j, -1)] = 2 * x[NDI(s, i, j, 0)] - x[NDI(s, i, j, 1)];
        x[NDI(s, i, j, -2)] = 2 * x[NDI(s, i, j, -1)] - x[NDI(s, i, j, 0)];
        x[NDI(s, i, j, KMX + 1)] = 2 * x[NDI(s, i, j, KMX)] - x[NDI(s, i, j, KMX - 1)];
        x[NDI(s, i, j, KMX + 2)] = 2 * x[NDI(s, i, j, KMX + 1)] - x[NDI(s, i, j, KMX)];
        x[NDI(s, i, j, KMX + 3)] = 2 * x[NDI(s, i, j, KMX + 2)] - x[NDI(s, i, j, KMX + 1)];
      }
  }
}

/* src/geometry.f90:342-356 */
static double vol_tet(const double *p1, const double *p2, const double *p3, const double *p4) {
  double v = (p4[0] - p1[0]) * ((p2[1] - p1[1]) * (p3[2] - p1[2]) - (p2[2] - p1[2]) * (p3[1] - p1[1]))
           + (p4[1] - p1[1]) * ((p2[2] - p1[2]) * (p3[0] - p1[0]) - (p2[0] - p1[0]) * (p3[2] - p1[2]))
           + (p4[2] - p1[2]) * ((p2[0] - p1[0]) * (p3[1] - p1[1]) - (p2[1] - p1[1]) * (p3[0] - p1[0]));
  return -v / 6.;
}

/* src/geometry.f90:367-446: max of two five-tetrahedron splits */
static double vol_hex(double p[8][3]) {
  double v1 = 0.;
  v1 = v1 + vol_tet(p[0], p[4], p[7], p[5]);
  v1 = v1 + vol_tet(p[6], p[7], p[5], p[2]);
  v1 = v1 + vol_tet(p[7], p[3], p[0], p[2]);
  v1 = v1 + vol_tet(p[5], p[0], p[2], p[7]);
  v1 = v1 + vol_tet(p[0], p[1], p[5], p[2]);
  double v = 0.;
  v = v + vol_tet(p[1], p[5], p[4], p[6]);
  v = v + vol_tet(p[7], p[4], p[6], p[3]);
  v = v + vol_tet(p[4], p[0], p[1], p[3]);
  v = v + vol_tet(p[6], p[1], p[3], p[4]);
  v = v + vol_tet(p[1], p[2], p[6], p[3]);
  return fmx(v, v1);
}

int orc_geometry(orc *s) {
  ghost_grid(s);
  const double *X = s->a[ORC_NODES], *Y = X + s->nnode, *Z = Y + s->nnode;
#define NX(i, j, k) X[NDI(s, i, j, k)]
#define NY(i, j, k) Y[NDI(s, i, j, k)]
#define NZ(i, j, k) Z[NDI(s, i, j, k)]
  /* area vectors, src/geometry.f90:236-314 */
  double *A, *nx, *ny, *nz;
  A = s->a[ORC_IA]; nx = s->a[ORC_INX]; ny = s->a[ORC_INY]; nz = s->a[ORC_INZ];
  for (int k = -2; k <= KMX + 2; ++k)
    for (int j = -2; j <= JMX + 2; ++j)
      for (int i = -2; i <= IMX + 3; ++i) {
        double d1x = NX(i, j + 1, k + 1) - NX(i, j, k), d1y = NY(i, j + 1, k + 1) - NY(i, j, k),
               d1z = NZ(i, j + 1, k + 1) - NZ(i, j, k);
        double d2x = NX(i, j, k + 1) - NX(i, j + 1, k), d2y = NY(i, j, k + 1) - NY(i, j + 1, k),
               d2z = NZ(i, j, k + 1) - NZ(i, j + 1, k);
        size_t f = IFI(s, i, j, k);
        nx[f] = 0.5 * (d1y * d2z - d1z * d2y);
        ny[f] = 0.5 * (d1z * d2x - d1x * d2z);
        nz[f] = 0.5 * (d1x * d2y - d1y * d2x);
      }
  A = s->a[ORC_JA]; nx = s->a[ORC_JNX]; ny = s->a[ORC_JNY]; nz = s->a[ORC_JNZ];
  for (int k = -2; k <= KMX + 2; ++k)
    for (int j = -2; j <= JMX + 3; ++j)
      for (int i = -2; i <= IMX + 2; ++i) {
        double d1x = NX(i + 1, j, k + 1) - NX(i, j, k), d1y = NY(i + 1, j, k + 1) - NY(i, j, k),
               d1z = NZ(i + 1, j, k + 1) - NZ(i, j, k);
        double d2x = NX(i + 1, j, k) - NX(i, j, k + 1), d2y = NY(i + 1, j, k) - NY(i, j, k + 1),
               d2z = NZ(i + 1, j, k) - NZ(i, j, k + 1);
        size_t f = JFI(s, i, j, k);
        nx[f] = 0.5 * (d1y * d2z - d1z * d2y);
        ny[f] = 0.5 * (d1z * d2x - d1x * d2z);
        nz[f] = 0.5 * (d1x * d2y - d1y * d2x);
      }
  A = s->a[ORC_KA]; nx = s->a[ORC_KNX]; ny = s->a[ORC_KNY]; nz = s->a[ORC_KNZ];
  for (int k = -2; k <= KMX + 3; ++k)
    for (int j = -2; j <= JMX + 2; ++j)
      for (int i = -2; i <= IMX + 2; ++i) {
        double d1x = NX(i + 1, j + 1, k) - NX(i, j, k), d1y = NY(i + 1, j + 1, k) - NY(i, j, k),
               d1z = NZ(i + 1, j + 1, k) - NZ(i, j, k);
        double d2x = NX(i, j + 1, k) - NX(i + 1, j, k), d2y = NY(i, j + 1, k) - NY(i + 1, j, k),
               d2z = NZ(i, j + 1, k) - NZ(i + 1, j, k);
        size_t f = KFI(s, i, j, k);
        nx[f] = 0.5 * (d1y * d2z - d1z * d2y);
        ny[f] = 0.5 * (d1z * d2x - d1x * d2z);
        nz[f] = 0.5 * (d1x * d2y - d1y * d2x);
      }
  /* areas and unit normals, src/geometry.f90:50-234 (pole branches: -7) */
  const int fa[3][4] = {{ORC_IA, ORC_INX, ORC_INY, ORC_INZ},
                        {ORC_JA, ORC_JNX, ORC_JNY, ORC_JNZ},
                        {ORC_KA, ORC_KNX, ORC_KNY, ORC_KNZ}};
  const size_t fn[3] = {s->nif, s->njf, s->nkf};
  for (int d = 0; d < 3; ++d) {
    A = s->a[fa[d][0]]; nx = s->a[fa[d][1]]; ny = s->a[fa[d][2]]; nz = s->a[fa[d][3]];
    for (size_t f = 0; f < fn[d]; ++f) A[f] = sqrt(sq(nx[f]) + sq(ny[f]) + sq(nz[f]));
    for (size_t f = 0; f < fn[d]; ++f)
      if (A[f] != 0.) { nx[f] = nx[f] / A[f]; ny[f] = ny[f] / A[f]; nz[f] = nz[f] / A[f]; }
  }
  /* pole faces, src/geometry.f90:101-186 and :227-232: the pole face and
   * its three ghost faces get zero area and the unit normal of the first
   * interior face (index 2, or N-1 at a max face), over the whole plane */
  for (int f = 0; f < 6; ++f) {
    if (s->c.bc_id[f] != -7) continue;
    /* on a one-cell k extent the two k faces' layers overlap; not restated */
    if (f >= 4 && KMX == 2) return ORC_E_CONFIG;
    int d = f / 2, hi = f & 1;
    int N = d == 0 ? IMX : d == 1 ? JMX : KMX;
    int ae = (d == 0 ? JMX : IMX) + 2, be = (d == 2 ? JMX : KMX) + 2;
    int src = hi ? N - 1 : 2;
    A = s->a[fa[d][0]]; nx = s->a[fa[d][1]]; ny = s->a[fa[d][2]]; nz = s->a[fa[d][3]];
    for (int l = 0; l < 4; ++l) {
      int n = hi ? N + l : 1 - l;
      for (int b = -2; b <= be; ++b)
        for (int a = -2; a <= ae; ++a) {
          size_t t, o;
          if (d == 0) { t = IFI(s, n, a, b); o = IFI(s, src, a, b); }
          else if (d == 1) { t = JFI(s, a, n, b); o = JFI(s, a, src, b); }
          else { t = KFI(s, a, b, n); o = KFI(s, a, b, src); }
          A[t] = 0.;
          nx[t] = nx[o]; ny[t] = ny[o]; nz[t] = nz[o];
        }
    }
  }
  /* volumes, src/geometry.f90:448-496 */
  double *vol = s->a[ORC_VOL];
  for (size_t n = 0; n < s->ncell; ++n) vol[n] = 1.;
  for (int k = 0; k <= KMX; ++k)
    for (int j = 0; j <= JMX; ++j)
      for (int i = 0; i <= IMX; ++i) {
        double p[8][3];
        const int o[8][3] = {{0, 0, 0}, {1, 0, 0}, {1, 1, 0}, {0, 1, 0},
                             {0, 0, 1}, {1, 0, 1}, {1, 1, 1}, {0, 1, 1}};
        for (int q = 0; q < 8; ++q) {
          p[q][0] = NX(i + o[q][0], j + o[q][1], k + o[q][2]);
          p[q][1] = NY(i + o[q][0], j + o[q][1], k + o[q][2]);
          p[q][2] = NZ(i + o[q][0], j + o[q][1], k + o[q][2]);
        }
        vol[CI(s, i, j, k)] = vol_hex(p);
        if (vol[CI(s, i, j, k)] <= 0.0) {
          int ghost = (i == 0 || i == IMX || j == 0 || j == JMX || k == 0 || k == KMX);
          if (!ghost) return ORC_E_NONPHYSICAL;
        }
      }
  for (size_t n = 0; n < s->ncell; ++n)
    if (vol[n] <= 0.0) return ORC_E_NONPHYSICAL;
  /* centroids, src/geometry.f90:500-545 */
  double *cc[3] = {s->a[ORC_CX], s->a[ORC_CY], s->a[ORC_CZ]};
  const double *XX[3] = {X, Y, Z};
  for (int c = 0; c < 3; ++c) {
    const double *x = XX[c];
    for (int k = -2; k <= KMX + 2; ++k)
      for (int j = -2; j <= JMX + 2; ++j)
        for (int i = -2; i <= IMX + 2; ++i)
          cc[c][CI(s, i, j, k)] = 0.125 * (x[NDI(s, i, j, k)] + x[NDI(s, i + 1, j, k)] +
                                           x[NDI(s, i + 1, j + 1, k)] + x[NDI(s, i + 1, j + 1, k + 1)] +
                                           x[NDI(s, i + 1, j, k + 1)] + x[NDI(s, i, j + 1, k)] +
                                           x[NDI(s, i, j + 1, k + 1)] + x[NDI(s, i, j, k + 1)]);
  }
#undef NX
#undef NY
#undef NZ
  return ORC_OK;
}

/* rounds to the 11 significant digits of an ES18.10E3 edit descriptor */
static double es18_10(double x) {
  char buf[64];
  snprintf(buf, sizeof buf, "%.10E", x);
  return strtod(buf, NULL);
}

typedef struct { size_t n, cap; double *x, *y, *z; } wall_set;

/* src/wall/wall.f90:198-288: the nodes of every no-slip face in face and
 * loop order, rounded like the ES18.10E3 surface-node file
 * (src/wall/wall.f90:101) that every process reads back */
static void wall_nodes(orc *s, wall_set *w) {
  const double *X = s->a[ORC_NODES], *Y = X + s->nnode, *Z = Y + s->nnode;
  for (int f = 0; f < 6; ++f) {
    if (s->c.bc_id[f] != -5) continue;
    int im = 1, ix = IMX, jm = 1, jx = JMX, km = 1, kx = KMX;
    switch (f) {
      case 0: ix = 1; break;
      case 1: im = IMX; break;
      case 2: jx = 1; break;
      case 3: jm = JMX; break;
      case 4: kx = 1; break;
      case 5: km = KMX; break;
    }
    for (int k = km; k <= kx; ++k)
      for (int j = jm; j <= jx; ++j)
        for (int i = im; i <= ix; ++i) {
          if (w->n == w->cap) {
            w->cap = w->cap ? 2 * w->cap : 1024;
            w->x = (double *)realloc(w->x, w->cap * sizeof(double));
            w->y = (double *)realloc(w->y, w->cap * sizeof(double));
            w->z = (double *)realloc(w->z, w->cap * sizeof(double));
          }
          w->x[w->n] = es18_10(X[NDI(s, i, j, k)]);
          w->y[w->n] = es18_10(Y[NDI(s, i, j, k)]);
          w->z[w->n] = es18_10(Z[NDI(s, i, j, k)]);
          ++w->n;
        }
  }
}

/* src/wall/wall_dist.f90:84-131: node distance to the nearest surface node,
 * cell distance the mean of its eight nodes' */
static void dist_from_walls(orc *s, const wall_set *w) {
  const double *X = s->a[ORC_NODES], *Y = X + s->nnode, *Z = Y + s->nnode;
  double *nd = (double *)malloc(s->nnode * sizeof(double));
  for (int k = -2; k <= KMX + 3; ++k)
    for (int j = -2; j <= JMX + 3; ++j)
      for (int i = -2; i <= IMX + 3; ++i) {
        size_t q = NDI(s, i, j, k);
        double best = 1.e+20;
        for (size_t n = 0; n < w->n; ++n) {
          double cd = sqrt(sq(w->x[n] - X[q]) + sq(w->y[n] - Y[q]) + sq(w->z[n] - Z[q]));
          best = fmn(best, cd);
        }
        nd[q] = best;
      }
  double *dist = s->a[ORC_DIST];
  for (int k = -2; k <= KMX + 2; ++k)
    for (int j = -2; j <= JMX + 2; ++j)
      for (int i = -2; i <= IMX + 2; ++i)
        dist[CI(s, i, j, k)] = 0.125 * (nd[NDI(s, i, j, k)] + nd[NDI(s, i, j + 1, k)] +
                                        nd[NDI(s, i, j + 1, k + 1)] + nd[NDI(s, i, j, k + 1)] +
                                        nd[NDI(s, i + 1, j, k + 1)] + nd[NDI(s, i + 1, j, k)] +
                                        nd[NDI(s, i + 1, j + 1, k)] + nd[NDI(s, i + 1, j + 1, k + 1)]);
  free(nd);
}

int orc_wall_dist(orc *s) {
  wall_set w = {0, 0, NULL, NULL, NULL};
  wall_nodes(s, &w);
  dist_from_walls(s, &w);
  free(w.x); free(w.y); free(w.z);
  return ORC_OK;
}

int orc_group_wall_dist(orc **blocks, int nblocks) {
  wall_set w = {0, 0, NULL, NULL, NULL};
  for (int b = 0; b < nblocks; ++b) wall_nodes(blocks[b], &w);
  for (int b = 0; b < nblocks; ++b) dist_from_walls(blocks[b], &w);
  free(w.x); free(w.y); free(w.z);
  return ORC_OK;
}

/* ------------------------------------------------------- ghost cells (C2) */
enum { CP_ANTI, CP_FLAT, CP_SYMM };

/* src/boundary/copy_bc.f90:58-131 */
static void copy3(orc *s, double *v, int type, int face) {
  double a1 = 1, a2 = 1, a3 = 0;
  int i1 = 1, i2 = 2, i3 = 3;
  if (type == CP_ANTI) { a1 = 1.; i1 = 1; a2 = -1.; i2 = 2; a3 = 0.; i3 = 3; }
  else if (type == CP_FLAT) { a1 = 1.; i1 = 1; a2 = 1.; i2 = 1; a3 = 0.; i3 = 1; }
  else { a1 = s->c1; i1 = 1; a2 = s->c2; i2 = 2; a3 = s->c3; i3 = 3; }
  const int src[3] = {i1, i2, i3};
  for (int g = 0; g < 3; ++g) {
    int is = src[g];
    switch (face) {
      case 0:
        for (int k = 1; k <= KMX - 1; ++k) for (int j = 1; j <= JMX - 1; ++j)
          v[CI(s, -g, j, k)] = (a2 * v[CI(s, is, j, k)] - a3 * v[CI(s, is + 1, j, k)]) / a1;
        break;
      case 1:
        for (int k = 1; k <= KMX - 1; ++k) for (int j = 1; j <= JMX - 1; ++j)
          v[CI(s, IMX + g, j, k)] = (a2 * v[CI(s, IMX - is, j, k)] - a3 * v[CI(s, IMX - is - 1, j, k)]) / a1;
        break;
      case 2:
        for (int k = 1; k <= KMX - 1; ++k) for (int i = 1; i <= IMX - 1; ++i)
          v[CI(s, i, -g, k)] = (a2 * v[CI(s, i, is, k)] - a3 * v[CI(s, i, is + 1, k)]) / a1;
        break;
      case 3:
        for (int k = 1; k <= KMX - 1; ++k) for (int i = 1; i <= IMX - 1; ++i)
          v[CI(s, i, JMX + g, k)] = (a2 * v[CI(s, i, JMX - is, k)] - a3 * v[CI(s, i, JMX - is - 1, k)]) / a1;
        break;
      case 4:
        for (int j = 1; j <= JMX - 1; ++j) for (int i = 1; i <= IMX - 1; ++i)
          v[CI(s, i, j, -g)] = (a2 * v[CI(s, i, j, is)] - a3 * v[CI(s, i, j, is + 1)]) / a1;
        break;
      case 5:
        for (int j = 1; j <= JMX - 1; ++j) for (int i = 1; i <= IMX - 1; ++i)
          v[CI(s, i, j, KMX + g)] = (a2 * v[CI(s, i, j, KMX - is)] - a3 * v[CI(s, i, j, KMX - is - 1)]) / a1;
        break;
    }
  }
}

/* src/boundary/bc_primitive.f90:484-525 */
static void fixv(orc *s, double *v, double val, int face) {
  for (int g = 0; g < 3; ++g) {
    switch (face) {
      case 0: for (int k = 1; k <= KMX - 1; ++k) for (int j = 1; j <= JMX - 1; ++j) v[CI(s, -g, j, k)] = val; break;
      case 1: for (int k = 1; k <= KMX - 1; ++k) for (int j = 1; j <= JMX - 1; ++j) v[CI(s, IMX + g, j, k)] = val; break;
      case 2: for (int k = 1; k <= KMX - 1; ++k) for (int i = 1; i <= IMX - 1; ++i) v[CI(s, i, -g, k)] = val; break;
      case 3: for (int k = 1; k <= KMX - 1; ++k) for (int i = 1; i <= IMX - 1; ++i) v[CI(s, i, JMX + g, k)] = val; break;
      case 4: for (int j = 1; j <= JMX - 1; ++j) for (int i = 1; i <= IMX - 1; ++i) v[CI(s, i, j, -g)] = val; break;
      case 5: for (int j = 1; j <= JMX - 1; ++j) for (int i = 1; i <= IMX - 1; ++i) v[CI(s, i, j, KMX + g)] = val; break;
    }
  }
}

/* Face geometry helper: ghost layer g (0,1,2 -> 0,-1,-2 or imx,imx+1,...),
 * interior layer m (1,2,3 -> 1,2,3 or imx-1,...), tangential (a,b). */
static size_t face_cell(const orc *s, int face, int layer_ghost, int a, int b) {
  /* layer_ghost: >0 interior index counted from the face (1..), <=0 ghost (0,-1,-2) */
  switch (face) {
    case 0: return CI(s, layer_ghost, a, b);
    case 1: return CI(s, IMX - layer_ghost, a, b);
    case 2: return CI(s, a, layer_ghost, b);
    case 3: return CI(s, a, JMX - layer_ghost, b);
    case 4: return CI(s, a, b, layer_ghost);
    default: return CI(s, a, b, KMX - layer_ghost);
  }
}
static void face_range(const orc *s, int face, int *na, int *nb) {
  if (face < 2) { *na = JMX - 1; *nb = KMX - 1; }
  else if (face < 4) { *na = IMX - 1; *nb = KMX - 1; }
  else { *na = IMX - 1; *nb = JMX - 1; }
}

/* src/boundary/bc_primitive.f90:1778-1945 */
static void temp_based_density(orc *s, int face) {
  double Tw = s->c.fix_twall[face], R = s->c.R_gas, gm = s->c.gm;
  double *rho = QPL(0), *p = QPL(4);
  int na, nb; face_range(s, face, &na, &nb);
  if (Tw < 0.0) {
    for (int b = 1; b <= nb; ++b) for (int a = 1; a <= na; ++a) {
      size_t ci = face_cell(s, face, 1, a, b);
      double stag = (p[ci] / (R * rho[ci])) * (1 + (0.5 * (gm - 1.) * gm * p[ci] / rho[ci]));
      for (int g = 0; g < 3; ++g) {
        size_t cg = face_cell(s, face, -g, a, b);
        rho[cg] = p[cg] / (R * stag);
      }
    }
  } else if (Tw > 1.0) {
    for (int b = 1; b <= nb; ++b) for (int a = 1; a <= na; ++a)
      for (int g = 0; g < 3; ++g) {
        size_t cg = face_cell(s, face, -g, a, b), ci = face_cell(s, face, 1 + g, a, b);
        rho[cg] = p[cg] / (R * (2 * Tw - (p[ci] / (R * rho[ci]))));
      }
  } else {
    copy3(s, rho, CP_SYMM, face);
  }
}

/* src/boundary/bc_primitive.f90:564-643 (Sutherland mu always, Q13) */
static void set_omega_at_wall(orc *s, int face) {
  double *rho = QPL(0), *p = QPL(4), *tw = QPL(6), *dist = s->a[ORC_DIST];
  int na, nb; face_range(s, face, &na, &nb);
  const orc_case *c = &s->c;
  for (int l = 1; l <= 3; ++l)
    for (int b = 1; b <= nb; ++b) for (int a = 1; a <= na; ++a) {
      size_t c0 = face_cell(s, face, 0, a, b), c1 = face_cell(s, face, 1, a, b);
      double T_face = 0.5 * ((p[c0] / rho[c0]) + (p[c1] / rho[c1])) / c->R_gas;
      double mu = c->mu_ref * pow(T_face / c->T_ref, 1.5) * ((c->T_ref + c->S_ref) / (T_face + c->S_ref));
      double r = 0.5 * (rho[c0] + rho[c1]);
      double d2 = 2 * dist[c1];
      tw[face_cell(s, face, 1 - l, a, b)] = 120 * mu / (r * beta1 * (d2 * d2)) - tw[face_cell(s, face, l, a, b)];
    }
}

/* lctm2015's intermittency takes a flat copy at every face that copies,
 * whatever the turbulence variables take (src/boundary/bc_primitive.f90:
 * 302, 385, 437, 475, 555) */
static void tgm_copy(orc *s, int face) {
  if (LCTM(s)) copy3(s, QPL(7), CP_FLAT, face);
}
static void turb_copy(orc *s, int type, int face) {
  if (SST(s) || KKL(s)) { copy3(s, QPL(5), type, face); copy3(s, QPL(6), type, face); }
  if (SA(s)) copy3(s, QPL(5), type, face);
  tgm_copy(s, face);
}
/* kl_tw: the subsonic inlet fixes k-kL's kL from the omega values
 * (src/boundary/bc_primitive.f90:333-336), which k-kL never reads from
 * the boundary file (read_bc.f90:98-108), so they stay 0 */
static void turb_fix_kl(orc *s, int face, int kl_tw) {
  if (SST(s)) {
    fixv(s, QPL(5), s->c.fix_tk[face], face);
    fixv(s, QPL(6), s->c.fix_tw[face], face);
  }
  if (KKL(s)) {
    fixv(s, QPL(5), s->c.fix_tk[face], face);
    fixv(s, QPL(6), kl_tw ? s->c.fix_tw[face] : s->c.fix_tkl[face], face);
  }
  if (SA(s)) fixv(s, QPL(5), s->c.fix_tv[face], face);
  if (LCTM(s)) fixv(s, QPL(7), s->c.fix_tgm[face], face); /* :262, :343 */
}
static void turb_fix(orc *s, int face) { turb_fix_kl(s, face, 0); }

/* face-normal unit vector stored at the boundary face (outward sign applied
 * by the caller) */
static void bface_normal(const orc *s, int face, int a, int b, double *nx, double *ny, double *nz) {
  size_t f;
  switch (face) {
    case 0: f = IFI(s, 1, a, b); *nx = s->a[ORC_INX][f]; *ny = s->a[ORC_INY][f]; *nz = s->a[ORC_INZ][f]; break;
    case 1: f = IFI(s, IMX, a, b); *nx = s->a[ORC_INX][f]; *ny = s->a[ORC_INY][f]; *nz = s->a[ORC_INZ][f]; break;
    case 2: f = JFI(s, a, 1, b); *nx = s->a[ORC_JNX][f]; *ny = s->a[ORC_JNY][f]; *nz = s->a[ORC_JNZ][f]; break;
    case 3: f = JFI(s, a, JMX, b); *nx = s->a[ORC_JNX][f]; *ny = s->a[ORC_JNY][f]; *nz = s->a[ORC_JNZ][f]; break;
    case 4: f = KFI(s, a, b, 1); *nx = s->a[ORC_KNX][f]; *ny = s->a[ORC_KNY][f]; *nz = s->a[ORC_KNZ][f]; break;
    default: f = KFI(s, a, b, KMX); *nx = s->a[ORC_KNX][f]; *ny = s->a[ORC_KNY][f]; *nz = s->a[ORC_KNZ][f]; break;
  }
}

/* src/boundary/bc_primitive.f90:645-1234 (far field) and, with total != 0,
 * :1237-1776 (total-pressure inlet): the same Riemann-invariant velocity;
 * the total-pressure inlet then takes the ghost pressure from the face's
 * total pressure and the boundary Mach number, and the density from the
 * boundary speed of sound */
static void riemann_face(orc *s, int face, int total) {
  const orc_case *c = &s->c;
  double gm = c->gm;
  double *rho = QPL(0), *u = QPL(1), *v = QPL(2), *w = QPL(3), *p = QPL(4);
  int na, nb; face_range(s, face, &na, &nb);
  double sg = (face % 2 == 0) ? -1.0 : 1.0; /* min faces use -n */
  int last_positive = 0;
  for (int b = 1; b <= nb; ++b) for (int a = 1; a <= na; ++a) {
    size_t ci = face_cell(s, face, 1, a, b), cg = face_cell(s, face, 0, a, b);
    double fnx, fny, fnz;
    bface_normal(s, face, a, b, &fnx, &fny, &fnz);
    double mx = (sg < 0) ? -fnx : fnx, my = (sg < 0) ? -fny : fny, mz = (sg < 0) ? -fnz : fnz;
    double uu = u[ci], vv = v[ci], ww = w[ci];
    double uf = c->u_inf, vf = c->v_inf, wf = c->w_inf;
    double cexp = sqrt(gm * p[ci] / rho[ci]);
    double cinf = sqrt(gm * c->p_inf / c->rho_inf);
    double Unexp = uu * mx + vv * my + ww * mz;
    double Uninf = uf * mx + vf * my + wf * mz;
    double Rinf = Uninf - 2 * cinf / (gm - 1.);
    double Rexp = Unexp + 2 * cexp / (gm - 1.);
    double Unb = 0.5 * (Rexp + Rinf);
    double Cb = 0.25 * (gm - 1.) * (Rexp - Rinf);
    if (Unb > 0.) {
      double vel_diff = Unb - Unexp;
      u[cg] = u[ci] + vel_diff * mx;
      v[cg] = v[ci] + vel_diff * my;
      w[cg] = w[ci] + vel_diff * mz;
      if (!total) {
        double sen = p[ci] / pow(rho[ci], gm);
        rho[cg] = pow(Cb * Cb / (gm * sen), 1. / (gm - 1.));
        p[cg] = (rho[cg] * Cb * Cb / gm);
      }
      last_positive = 1;
    } else {
      double vel_diff = Unb - Uninf;
      u[cg] = c->u_inf + vel_diff * mx;
      v[cg] = c->v_inf + vel_diff * my;
      w[cg] = c->w_inf + vel_diff * mz;
      if (!total) {
        double sen = c->p_inf / pow(c->rho_inf, gm);
        rho[cg] = pow(Cb * Cb / (gm * sen), 1. / (gm - 1.));
        p[cg] = (rho[cg] * Cb * Cb / gm);
      }
      last_positive = 0;
    }
    if (total) {
      double Mb = sqrt(u[cg] * u[cg] + v[cg] * v[cg] + w[cg] * w[cg]) / Cb;
      p[cg] = c->fix_tp[face] / pow((1 + 0.5 * (gm - 1.) * Mb * Mb), gm / (gm - 1.));
      rho[cg] = gm * p[cg] / (Cb * Cb);
    }
  }
  /* the turbulent ghosts end as set by the last face cell in loop order:
   * a whole-face flat copy if it was outflow, the fixed values otherwise */
  if (s->c.turb != ORC_TURB_NONE) {
    if (last_positive) turb_copy(s, CP_FLAT, face); else turb_fix(s, face);
  }
  /* qp(-1,:,:,:) = qp(0,:,:,:); qp(-2,:,:,:) = qp(0,:,:,:) on the full plane */
  for (int l = 0; l < s->nv; ++l) {
    double *q = QPL(l);
    for (int g = 1; g <= 2; ++g) {
      switch (face) {
        case 0: for (int k = -2; k <= KMX + 2; ++k) for (int j = -2; j <= JMX + 2; ++j) q[CI(s, -g, j, k)] = q[CI(s, 0, j, k)]; break;
        case 1: for (int k = -2; k <= KMX + 2; ++k) for (int j = -2; j <= JMX + 2; ++j) q[CI(s, IMX + g, j, k)] = q[CI(s, IMX, j, k)]; break;
        case 2: for (int k = -2; k <= KMX + 2; ++k) for (int i = -2; i <= IMX + 2; ++i) q[CI(s, i, -g, k)] = q[CI(s, i, 0, k)]; break;
        case 3: for (int k = -2; k <= KMX + 2; ++k) for (int i = -2; i <= IMX + 2; ++i) q[CI(s, i, JMX + g, k)] = q[CI(s, i, JMX, k)]; break;
        case 4: for (int j = -2; j <= JMX + 2; ++j) for (int i = -2; i <= IMX + 2; ++i) q[CI(s, i, j, -g)] = q[CI(s, i, j, 0)]; break;
        case 5: for (int j = -2; j <= JMX + 2; ++j) for (int i = -2; i <= IMX + 2; ++i) q[CI(s, i, j, KMX + g)] = q[CI(s, i, j, KMX)]; break;
      }
    }
  }
}

/* src/boundary/FT_bc.f90:36-106; the j and k faces reflect with the I-face
 * normal stored at the same index (reference behaviour, kept) */
static void flow_tangency(orc *s, int face) {
  double *u = QPL(1), *v = QPL(2), *w = QPL(3);
  const double *inx = s->a[ORC_INX], *iny = s->a[ORC_INY], *inz = s->a[ORC_INZ];
  int na, nb; face_range(s, face, &na, &nb);
  for (int b = 1; b <= nb; ++b) for (int a = 1; a <= na; ++a)
    for (int l = 1; l <= 3; ++l) {
      double fnx, fny, fnz;
      bface_normal(s, face, a, b, &fnx, &fny, &fnz);
      size_t fi;
      switch (face) {
        case 0: fi = IFI(s, 1, a, b); break;
        case 1: fi = IFI(s, IMX, a, b); break;
        case 2: fi = IFI(s, a, 1, b); break;
        case 3: fi = IFI(s, a, JMX, b); break;
        case 4: fi = IFI(s, a, b, 1); break;
        default: fi = IFI(s, a, b, KMX); break;
      }
      size_t ci = face_cell(s, face, l, a, b), cg = face_cell(s, face, 1 - l, a, b);
      double dot = u[ci] * fnx + v[ci] * fny + w[ci] * fnz;
      u[cg] = u[ci] - (2.0 * dot * inx[fi]);
      v[cg] = v[ci] - (2.0 * dot * iny[fi]);
      w[cg] = w[ci] - (2.0 * dot * inz[fi]);
    }
}

/* src/boundary/bc_primitive.f90:1948-1979.  The array assignments read
 * their right-hand side whole before writing; with one or two cells in a
 * direction the ranges overlap, so the min faces copy layers in ascending
 * order and the max faces in descending order, which reads every source
 * layer before it is overwritten. */
static void periodic_bc(orc *s, int face) {
  for (int l = 0; l < s->nv; ++l) {
    double *q = QPL(l);
    for (int gg = 0; gg < 3; ++gg) {
      int g = (face & 1) ? 2 - gg : gg;
      switch (face) {
        case 0: for (int k = -2; k <= KMX + 2; ++k) for (int j = -2; j <= JMX + 2; ++j) q[CI(s, -2 + g, j, k)] = q[CI(s, IMX - 3 + g, j, k)]; break;
        case 1: for (int k = -2; k <= KMX + 2; ++k) for (int j = -2; j <= JMX + 2; ++j) q[CI(s, IMX + g, j, k)] = q[CI(s, 1 + g, j, k)]; break;
        case 2: for (int k = -2; k <= KMX + 2; ++k) for (int i = -2; i <= IMX + 2; ++i) q[CI(s, i, -2 + g, k)] = q[CI(s, i, JMX - 3 + g, k)]; break;
        case 3: for (int k = -2; k <= KMX + 2; ++k) for (int i = -2; i <= IMX + 2; ++i) q[CI(s, i, JMX + g, k)] = q[CI(s, i, 1 + g, k)]; break;
        case 4: for (int j = -2; j <= JMX + 2; ++j) for (int i = -2; i <= IMX + 2; ++i) q[CI(s, i, j, -2 + g)] = q[CI(s, i, j, KMX - 3 + g)]; break;
        case 5: for (int j = -2; j <= JMX + 2; ++j) for (int i = -2; i <= IMX + 2; ++i) q[CI(s, i, j, KMX + g)] = q[CI(s, i, j, 1 + g)]; break;
      }
    }
  }
}

/* src/boundary/bc_primitive.f90:55-226 */
static int populate_ghost_primitive(orc *s, int current_iter) {
  const orc_case *c = &s->c;
  for (int f = 0; f < 6; ++f) {
    switch (c->bc_id[f]) {
      case -1: /* supersonic inlet */
        if (current_iter <= 2) {
          fixv(s, QPL(0), c->fix_rho[f], f); fixv(s, QPL(1), c->fix_u[f], f);
          fixv(s, QPL(2), c->fix_v[f], f); fixv(s, QPL(3), c->fix_w[f], f);
          fixv(s, QPL(4), c->fix_p[f], f); turb_fix(s, f);
        }
        break;
      case -2: /* supersonic outlet */
        for (int l = 0; l < 5; ++l) copy3(s, QPL(l), CP_FLAT, f);
        turb_copy(s, CP_FLAT, f);
        break;
      case -3: /* subsonic inlet */
        if (current_iter <= 2) {
          fixv(s, QPL(0), c->fix_rho[f], f); fixv(s, QPL(1), c->fix_u[f], f);
          fixv(s, QPL(2), c->fix_v[f], f); fixv(s, QPL(3), c->fix_w[f], f);
          turb_fix_kl(s, f, 1);
        }
        copy3(s, QPL(4), CP_FLAT, f);
        break;
      case -4: /* subsonic outlet */
        for (int l = 0; l < 4; ++l) copy3(s, QPL(l), CP_FLAT, f);
        if (current_iter <= 2) fixv(s, QPL(4), c->fix_p[f], f);
        turb_copy(s, CP_FLAT, f);
        break;
      case -5: /* wall */
        copy3(s, QPL(4), CP_SYMM, f);
        temp_based_density(s, f);
        copy3(s, QPL(1), CP_ANTI, f); copy3(s, QPL(2), CP_ANTI, f); copy3(s, QPL(3), CP_ANTI, f);
        if (SST(s)) { copy3(s, QPL(5), CP_ANTI, f); set_omega_at_wall(s, f); }
        if (SA(s)) copy3(s, QPL(5), CP_ANTI, f); /* no_slip, bc_primitive.f90:540-541 */
        if (KKL(s)) { copy3(s, QPL(5), CP_ANTI, f); copy3(s, QPL(6), CP_ANTI, f); } /* :548-550 */
        tgm_copy(s, f); /* :553-558 */
        break;
      case -6: /* slip wall */
        copy3(s, QPL(0), CP_SYMM, f); copy3(s, QPL(4), CP_SYMM, f);
        turb_copy(s, CP_SYMM, f);
        flow_tangency(s, f);
        break;
      case -7: /* pole */
        for (int l = 0; l < 5; ++l) copy3(s, QPL(l), CP_FLAT, f);
        turb_copy(s, CP_FLAT, f);
        break;
      case -8: riemann_face(s, f, 0); break;
      case -9: periodic_bc(s, f); break;
      case -11: riemann_face(s, f, 1); break;
      default: break; /* interfaces (>=0) and -10 untouched */
    }
  }
  /* edges and corners, src/boundary/bc_primitive.f90:209-224 */
  for (int l = 0; l < s->nv; ++l) {
    double *q = QPL(l);
#define Q(i, j, k) q[CI(s, i, j, k)]
    for (int i = -2; i <= IMX + 2; ++i) {
      Q(i, 0, 0) = 0.33 * (Q(i, 1, 1) + Q(i, 0, 1) + Q(i, 1, 0));
      Q(i, 0, KMX) = 0.33 * (Q(i, 1, KMX - 1) + Q(i, 0, KMX - 1) + Q(i, 1, KMX));
      Q(i, JMX, 0) = 0.33 * (Q(i, JMX - 1, 1) + Q(i, JMX, 1) + Q(i, JMX - 1, 0));
      Q(i, JMX, KMX) = 0.33 * (Q(i, JMX - 1, KMX - 1) + Q(i, JMX, KMX - 1) + Q(i, JMX - 1, KMX));
    }
    for (int k = -2; k <= KMX + 2; ++k) Q(IMX, 0, k) = 0.33 * (Q(IMX - 1, 1, k) + Q(IMX - 1, 0, k) + Q(IMX, 1, k));
    for (int k = -2; k <= KMX + 2; ++k) Q(0, 0, k) = 0.33 * (Q(1, 1, k) + Q(1, 0, k) + Q(0, 1, k));
    for (int k = -2; k <= KMX + 2; ++k) Q(0, JMX, k) = 0.33 * (Q(1, JMX - 1, k) + Q(1, JMX, k) + Q(0, JMX - 1, k));
    for (int k = -2; k <= KMX + 2; ++k) Q(IMX, JMX, k) = 0.33 * (Q(IMX - 1, JMX - 1, k) + Q(IMX - 1, JMX, k) + Q(IMX, JMX - 1, k));
    Q(0, 0, 0) = 0.33 * (Q(1, 0, 0) + Q(0, 1, 0) + Q(0, 0, 1));
    Q(IMX, 0, 0) = 0.33 * (Q(IMX - 1, 0, 0) + Q(IMX, 1, 0) + Q(IMX, 0, 1));
    Q(0, JMX, 0) = 0.33 * (Q(1, JMX, 0) + Q(0, JMX - 1, 0) + Q(0, JMX, 1));
    Q(0, 0, KMX) = 0.33 * (Q(1, 0, KMX) + Q(0, 1, KMX) + Q(0, 0, KMX - 1));
    Q(IMX, JMX, 0) = 0.33 * (Q(IMX - 1, JMX, 0) + Q(IMX, JMX - 1, 0) + Q(IMX, JMX, 1));
    Q(IMX, 0, KMX) = 0.33 * (Q(IMX - 1, 0, KMX) + Q(IMX, 1, KMX) + Q(IMX, 0, KMX - 1));
    Q(0, JMX, KMX) = 0.33 * (Q(1, JMX, KMX) + Q(0, JMX - 1, KMX) + Q(0, JMX, KMX - 1));
    Q(IMX, JMX, KMX) = 0.33 * (Q(IMX - 1, JMX, KMX) + Q(IMX, JMX - 1, KMX) + Q(IMX, JMX, KMX - 1));
#undef Q
  }
  return ORC_OK;
}

/* ------------------------------------------------ reconstruction (C3, C4) */
typedef struct { double *L, *R; size_t n; size_t (*idx)(const orc *, int, int, int); } fstates;

static fstates states_of(orc *s, int d) {
  fstates f;
  if (d == 0) { f.L = s->a[ORC_XL]; f.R = s->a[ORC_XR]; f.n = s->nxs; f.idx = XSI; }
  else if (d == 1) { f.L = s->a[ORC_YL]; f.R = s->a[ORC_YR]; f.n = s->nys; f.idx = YSI; }
  else { f.L = s->a[ORC_ZL]; f.R = s->a[ORC_ZR]; f.n = s->nzs; f.idx = ZSI; }
  return f;
}

/* MUSCL, src/face/state/muscl.f90:115-200.  With the van Albada option the
 * limiter is (r^2 + r)/(r^2 + 1), the form muscl.f90:180,186 carries as a
 * comment; the Koren-alpha limiter is the reference's. */
static void muscl_dir(orc *s, int d) {
  fstates fs = states_of(s, d);
  int ii = d == 0, jj = d == 1, kk = d == 2;
  const int valb = s->c.muscl_limiter == ORC_LIM_VAN_ALBADA;
  double alpha = 2. / 3., phi = 1.0, kappa = 1. / 3., eps = 1e-14;
  int sw = s->c.lim[d];
  for (int l = 0; l < s->nv; ++l) {
    if (l >= 5) sw = s->c.tlim[d];
    double *q = QPL(l);
    double *Lp = fs.L + (size_t)l * fs.n, *Rp = fs.R + (size_t)l * fs.n;
    for (int k = 1 - kk; k <= KMX - 1 + kk; ++k)
      for (int j = 1 - jj; j <= JMX - 1 + jj; ++j)
        for (int i = 1 - ii; i <= IMX - 1 + ii; ++i) {
          double fd = q[CI(s, i + ii, j + jj, k + kk)] - q[CI(s, i, j, k)];
          double bd = q[CI(s, i, j, k)] - q[CI(s, i - ii, j - jj, k - kk)];
          double r = fd / (bd + fsign(eps, bd));
          double psi1 = valb ? ((r * r) + r) / ((r * r) + 1.0) : fmx(0., fmn(fmn(2 * r, alpha * (r - 1.0) + 1.0), 2.));
          r = bd / (fd + fsign(eps, fd));
          double psi2 = valb ? ((r * r) + r) / ((r * r) + 1.0) : fmx(0., fmn(fmn(2 * r, alpha * (r - 1.0) + 1.0), 2.));
          psi1 = (1 - (1 - psi1) * sw);
          psi2 = (1 - (1 - psi2) * sw);
          Lp[fs.idx(s, i + ii, j + jj, k + kk)] = q[CI(s, i, j, k)] + 0.25 * phi *
              (((1. - kappa) * psi1 * bd) + ((1. + kappa) * psi2 * fd));
          Rp[fs.idx(s, i, j, k)] = q[CI(s, i, j, k)] - 0.25 * phi *
              (((1. + kappa) * psi1 * bd) + ((1. - kappa) * psi2 * fd));
        }
  }
}

/* WENO5, src/face/state/weno.f90:24-93 */
static void weno_dir(orc *s, int d) {
  fstates fs = states_of(s, d);
  int ii = d == 0, jj = d == 1, kk = d == 2;
  double g1 = 1.0 / 10.0, g2 = 6.0 / 10.0, g3 = 3.0 / 10.0, eps = 1e-6;
  for (int l = 0; l < s->nv; ++l) {
    double *q = QPL(l);
    double *Lp = fs.L + (size_t)l * fs.n, *Rp = fs.R + (size_t)l * fs.n;
    for (int k = 1 - kk; k <= KMX - 1 + kk; ++k)
      for (int j = 1 - jj; j <= JMX - 1 + jj; ++j)
        for (int i = 1 - ii; i <= IMX - 1 + ii; ++i) {
          double Um2 = q[CI(s, i - 2 * ii, j - 2 * jj, k - 2 * kk)];
          double Um1 = q[CI(s, i - ii, j - jj, k - kk)];
          double U0 = q[CI(s, i, j, k)];
          double Up1 = q[CI(s, i + ii, j + jj, k + kk)];
          double Up2 = q[CI(s, i + 2 * ii, j + 2 * jj, k + 2 * kk)];
          double P1 = (2.0 * Um2 - 7.0 * Um1 + 11.0 * U0) / 6.0;
          double P2 = (-1.0 * Um1 + 5.0 * U0 + 2.0 * Up1) / 6.0;
          double P3 = (2.0 * U0 + 5.0 * Up1 - 1.0 * Up2) / 6.0;
          double B1 = (13.0 / 12.0) * sq(Um2 - 2.0 * Um1 + U0) + (1.0 / 4.0) * sq(Um2 - 4.0 * Um1 + 3.0 * U0);
          double B2 = (13.0 / 12.0) * sq(Um1 - 2.0 * U0 + Up1) + (1.0 / 4.0) * sq(Um1 - Up1);
          double B3 = (13.0 / 12.0) * sq(U0 - 2.0 * Up1 + Up2) + (1.0 / 4.0) * sq(3.0 * U0 - 4.0 * Up1 + Up2);
          double w1 = g1 / sq(eps + B1), w2 = g2 / sq(eps + B2), w3 = g3 / sq(eps + B3);
          Lp[fs.idx(s, i + ii, j + jj, k + kk)] = (w1 * P1 + w2 * P2 + w3 * P3) / (w1 + w2 + w3);
          P1 = (2.0 * Up2 - 7.0 * Up1 + 11.0 * U0) / 6.0;
          P2 = (-1.0 * Up1 + 5.0 * U0 + 2.0 * Um1) / 6.0;
          P3 = (2.0 * U0 + 5.0 * Um1 - 1.0 * Um2) / 6.0;
          w1 = g1 / sq(eps + B3); w2 = g2 / sq(eps + B2); w3 = g3 / sq(eps + B1);
          Rp[fs.idx(s, i, j, k)] = (w1 * P1 + w2 * P2 + w3 * P3) / (w1 + w2 + w3);
        }
  }
}

/* PPM, src/face/state/ppm.f90:21-107 */
static void ppm_dir(orc *s, int d) {
  fstates fs = states_of(s, d);
  int ii = d == 0, jj = d == 1, kk = d == 2;
  for (int l = 0; l < s->nv; ++l) {
    double *q = QPL(l);
    double *Lp = fs.L + (size_t)l * fs.n, *Rp = fs.R + (size_t)l * fs.n;
    for (int k = 1 - kk; k <= KMX - 1 + 2 * kk; ++k)
      for (int j = 1 - jj; j <= JMX - 1 + 2 * jj; ++j)
        for (int i = 1 - ii; i <= IMX - 1 + 2 * ii; ++i) {
          double v = (7. * (q[CI(s, i, j, k)] + q[CI(s, i - ii, j - jj, k - kk)]) -
                      (q[CI(s, i + ii, j + jj, k + kk)] + q[CI(s, i - 2 * ii, j - 2 * jj, k - 2 * kk)])) / 12.;
          Lp[fs.idx(s, i, j, k)] = v;
          Rp[fs.idx(s, i, j, k)] = v;
        }
    if (s->c.lim[d] == 1) {
      for (int k = 1 - kk; k <= KMX - 1 + kk; ++k)
        for (int j = 1 - jj; j <= JMX - 1 + jj; ++j)
          for (int i = 1 - ii; i <= IMX - 1 + ii; ++i) {
            size_t iL = fs.idx(s, i + ii, j + jj, k + kk), iR = fs.idx(s, i, j, k);
            double qc = q[CI(s, i, j, k)];
            if ((Lp[iL] - qc) * (qc - Rp[iR]) <= 0) {
              Lp[iL] = qc; Rp[iR] = qc;
            } else {
              double dqrl = Lp[iL] - Rp[iR];
              double dq6 = 6. * (qc - 0.5 * (Lp[iL] + Rp[iR]));
              if (dqrl * dq6 > dqrl * dqrl) Rp[iR] = 3. * qc - 2. * Lp[iL];
              else if (-dqrl * dqrl > dqrl * dq6) Lp[iL] = 3. * qc - 2. * Rp[iR];
            }
          }
    }
  }
}

/* WENO for non-uniform grids, src/face/state/weno_NM.f90:21-128: the
 * stencil values are blended with volume weights before the Jiang-Shu
 * polynomials and smoothness indicators; sums in element order. */
static void weno_nm_dir(orc *s, int d) {
  fstates fs = states_of(s, d);
  int ii = d == 0, jj = d == 1, kk = d == 2;
  const double g[3] = {1.0 / 10.0, 6.0 / 10.0, 3.0 / 10.0}, eps = 1e-6;
  const double *vol = s->a[ORC_VOL];
  for (int l = 0; l < s->nv; ++l) {
    double *q = QPL(l);
    double *Lp = fs.L + (size_t)l * fs.n, *Rp = fs.R + (size_t)l * fs.n;
    for (int k = 1 - kk; k <= KMX - 1 + kk; ++k)
      for (int j = 1 - jj; j <= JMX - 1 + jj; ++j)
        for (int i = 1 - ii; i <= IMX - 1 + ii; ++i) {
          double U[5], V[5];
          for (int m = -2; m <= 2; ++m) {
            size_t c = CI(s, i + m * ii, j + m * jj, k + m * kk);
            U[m + 2] = q[c];
            V[m + 2] = vol[c];
          }
          double alpha12 = V[4] / (V[3] + V[4]);
          double alpha01 = V[3] / (V[2] + V[3]);
          double alpha10 = V[2] / (V[1] + V[2]);
          double alpha21 = V[1] / (V[0] + V[1]);
          double U01 = (1.0 - alpha01) * U[2] + alpha01 * U[3];
          double U12 = (1.0 - alpha12) * U[3] + alpha12 * U[4];
          double U10 = (1.0 - alpha10) * U[1] + alpha10 * U[2];
          double U21 = (1.0 - alpha21) * U[0] + alpha21 * U[1];
          double U00 = U[1] + (1.0 - alpha21) * (U[1] - U[0]);
          double U11 = U[3] + alpha12 * (U[3] - U[4]);
          double P[3], B[3], w[3];
          P[0] = (6.0 * U[2] - 1.0 * U10 - 2.0 * U00) / 3.0;
          P[1] = (-1.0 * U10 + 2.0 * U[2] + 2.0 * U01) / 3.0;
          P[2] = (2.0 * U01 + 2.0 * U[3] - 1.0 * U12) / 3.0;
          B[0] = (13.0 / 12.0) * sq(2 * U10 - 2.0 * U00) + (1.0 / 4.0) * sq(4 * U[2] - 2.0 * U10 - 2.0 * U00);
          B[1] = (13.0 / 12.0) * sq(2 * U10 - 4.0 * U[2] + 2 * U01) + (1.0 / 4.0) * sq(-2 * U10 + 2.0 * U01);
          B[2] = (13.0 / 12.0) * sq(2 * U01 - 4.0 * U[3] + 2 * U12) + (1.0 / 4.0) * sq(-6 * U01 + 8.0 * U[3] - 2.0 * U12);
          for (int m = 0; m < 3; ++m) w[m] = g[m] / sq(eps + B[m]);
          Lp[fs.idx(s, i + ii, j + jj, k + kk)] = ((w[0] * P[0] + w[1] * P[1]) + w[2] * P[2]) / ((w[0] + w[1]) + w[2]);
          P[0] = (6.0 * U[2] - 1.0 * U01 - 2.0 * U11) / 3.0;
          P[1] = (-1.0 * U01 + 2.0 * U[2] + 2.0 * U10) / 3.0;
          P[2] = (2.0 * U10 + 2.0 * U[1] - 1.0 * U21) / 3.0;
          B[0] = (13.0 / 12.0) * sq(2 * U01 - 2.0 * U11) + (1.0 / 4.0) * sq(4 * U[2] - 2.0 * U01 - 2.0 * U11);
          B[1] = (13.0 / 12.0) * sq(2 * U01 - 4.0 * U[2] + 2 * U10) + (1.0 / 4.0) * sq(-2 * U01 + 2.0 * U10);
          B[2] = (13.0 / 12.0) * sq(2 * U10 - 4.0 * U[1] + 2 * U21) + (1.0 / 4.0) * sq(-6 * U10 + 8.0 * U[1] - 2.0 * U21);
          for (int m = 0; m < 3; ++m) w[m] = g[m] / sq(eps + B[m]);
          Rp[fs.idx(s, i, j, k)] = ((w[0] * P[0] + w[1] * P[1]) + w[2] * P[2]) / ((w[0] + w[1]) + w[2]);
        }
  }
}

/* Pressure-based switch, src/face/state/muscl.f90:37-112 (and its copy in
 * ppm.f90:109-186): pdif = 1 - |p(c+1) - p(c-1)| / (|...| + p_inf) over the
 * interior cells along d, copied into the two ghost cells along d from the
 * first and last interior cell; then every face state of faces 1..N is
 * pulled towards its cell value by that cell's pdif. */
static void pb_switch_dir(orc *s, int d) {
  fstates fs = states_of(s, d);
  int ii = d == 0, jj = d == 1, kk = d == 2;
  const int N = d == 0 ? IMX : d == 1 ? JMX : KMX;
  const double *p = QPL(4);
  /* pdif over (0:imx,0:jmx,0:kmx) */
  const int ni = IMX + 1, nj = JMX + 1;
  double *pdif = (double *)malloc(sizeof(double) * (size_t)ni * nj * (KMX + 1));
#define PD(i, j, k) pdif[(size_t)(i) + (size_t)ni * ((size_t)(j) + (size_t)nj * (size_t)(k))]
  for (int k = 1; k <= KMX - 1; ++k)
    for (int j = 1; j <= JMX - 1; ++j)
      for (int i = 1; i <= IMX - 1; ++i) {
        double pd2 = fabs(p[CI(s, i + ii, j + jj, k + kk)] - p[CI(s, i - ii, j - jj, k - kk)]);
        PD(i, j, k) = 1 - (pd2 / (pd2 + s->c.p_inf));
      }
  for (int k = 1; k <= KMX - 1; ++k)
    for (int j = 1; j <= JMX - 1; ++j)
      for (int i = 1; i <= IMX - 1; ++i) {
        int n = d == 0 ? i : d == 1 ? j : k;
        if (n == 1) PD(i - ii, j - jj, k - kk) = PD(i, j, k);
        if (n == N - 1) PD(i + ii, j + jj, k + kk) = PD(i, j, k);
      }
  for (int l = 0; l < s->nv; ++l) {
    const double *q = QPL(l);
    double *Lp = fs.L + (size_t)l * fs.n, *Rp = fs.R + (size_t)l * fs.n;
    for (int k = 1; k <= KMX - 1 + kk; ++k)
      for (int j = 1; j <= JMX - 1 + jj; ++j)
        for (int i = 1; i <= IMX - 1 + ii; ++i) {
          size_t f = fs.idx(s, i, j, k);
          size_t cm = CI(s, i - ii, j - jj, k - kk), c = CI(s, i, j, k);
          Lp[f] = q[cm] + (PD(i - ii, j - jj, k - kk) * (Lp[f] - q[cm]));
          Rp[f] = q[c] - (PD(i, j, k) * (q[c] - Rp[f]));
        }
  }
#undef PD
  free(pdif);
}

/* first order, src/face/state/face_interpolant.f90:61-76 */
static void none_dir(orc *s, int d) {
  fstates fs = states_of(s, d);
  int ii = d == 0, jj = d == 1, kk = d == 2;
  for (int l = 0; l < s->nv; ++l) {
    double *q = QPL(l);
    double *Lp = fs.L + (size_t)l * fs.n, *Rp = fs.R + (size_t)l * fs.n;
    for (int k = 1 - kk; k <= KMX - 1 + 2 * kk; ++k)
      for (int j = 1 - jj; j <= JMX - 1 + 2 * jj; ++j)
        for (int i = 1 - ii; i <= IMX - 1 + 2 * ii; ++i) {
          Lp[fs.idx(s, i, j, k)] = q[CI(s, i - ii, j - jj, k - kk)];
          Rp[fs.idx(s, i, j, k)] = q[CI(s, i, j, k)];
        }
  }
}

/* src/boundary/boundary_state_reconstruction.f90:78-411 (one face) */
static void reconstruct_boundary_face(orc *s, int face) {
  int d = face / 2, hi = face % 2;
  fstates fs = states_of(s, d);
  int ii = d == 0, jj = d == 1, kk = d == 2;
  int n = (d == 0) ? IMX : (d == 1) ? JMX : KMX;
  double phi = 1.0, kappa = 1. / 3.;
  int id = s->c.bc_id[face];
  int na, nb; face_range(s, face, &na, &nb);
  /* cell index along the sweep: 1 for min, n-1 for max */
  int cidx = hi ? n - 1 : 1;
  if (s->ppm_flag) {
    int sw = s->c.lim[d];
    for (int l = 0; l < s->nv; ++l) {
      if (d < 2) { if (l >= 5) sw = s->c.tlim[d]; }
      else sw = (l >= 5) ? s->c.tlim[d] : s->c.lim[d];
      double *q = QPL(l);
      double *Lp = fs.L + (size_t)l * fs.n, *Rp = fs.R + (size_t)l * fs.n;
      for (int b = 1; b <= nb; ++b) for (int a = 1; a <= na; ++a) {
        int i, j, k;
        if (d == 0) { i = cidx; j = a; k = b; } else if (d == 1) { i = a; j = cidx; k = b; } else { i = a; j = b; k = cidx; }
        if (d == 2 && hi) Rp[fs.idx(s, i, j, k + 1)] = 0.5 * (q[CI(s, i, j, k)] + q[CI(s, i, j, k + 1)]);
        double fd = q[CI(s, i + ii, j + jj, k + kk)] - q[CI(s, i, j, k)];
        double bd = q[CI(s, i, j, k)] - q[CI(s, i - ii, j - jj, k - kk)];
        double r = fd / bd;
        double psi1 = fmx(0., fmn(fmn(2 * r, (2 + r) / 3.), 2.));
        psi1 = (1 - (1 - psi1) * sw);
        r = bd / fd;
        double psi2 = fmx(0., fmn(fmn(2 * r, (2 + r) / 3.), 2.));
        psi2 = (1 - (1 - psi2) * sw);
        Lp[fs.idx(s, i + ii, j + jj, k + kk)] = q[CI(s, i, j, k)] + 0.25 * phi *
            (((1. - kappa) * psi1 * bd) + ((1. + kappa) * psi2 * fd));
        Rp[fs.idx(s, i, j, k)] = q[CI(s, i, j, k)] - 0.25 * phi *
            (((1. + kappa) * psi1 * bd) + ((1. - kappa) * psi2 * fd));
      }
    }
  } else if (d == 2 && hi) {
    for (int l = 0; l < s->nv; ++l) {
      double *q = QPL(l);
      double *Rp = fs.R + (size_t)l * fs.n;
      for (int b = 1; b <= nb; ++b) for (int a = 1; a <= na; ++a)
        Rp[fs.idx(s, a, b, KMX)] = 0.5 * (q[CI(s, a, b, KMX - 1)] + q[CI(s, a, b, KMX)]);
    }
  }
  /* face-state overrides at the boundary face */
  int fidx = hi ? n : 1;     /* face index */
  int gidx = hi ? n : 0;     /* ghost cell adjacent to the face */
  int iidx = hi ? n - 1 : 1; /* interior cell adjacent to the face */
  for (int l = 0; l < s->nv; ++l) {
    double *q = QPL(l);
    double *Lp = fs.L + (size_t)l * fs.n, *Rp = fs.R + (size_t)l * fs.n;
    for (int b = 1; b <= nb; ++b) for (int a = 1; a <= na; ++a) {
      size_t fi, cg, cin;
      if (d == 0) { fi = fs.idx(s, fidx, a, b); cg = CI(s, gidx, a, b); cin = CI(s, iidx, a, b); }
      else if (d == 1) { fi = fs.idx(s, a, fidx, b); cg = CI(s, a, gidx, b); cin = CI(s, a, iidx, b); }
      else { fi = fs.idx(s, a, b, fidx); cg = CI(s, a, b, gidx); cin = CI(s, a, b, iidx); }
      if (id == -8 || id == -9) {
        Lp[fi] = q[cg]; Rp[fi] = q[cg];
      } else if (!hi) {
        Lp[fi] = 0.5 * (q[cg] + q[cin]);
      } else {
        Rp[fi] = 0.5 * (q[cin] + q[cg]);
      }
    }
  }
}

static int compute_face_interpolant(orc *s) {
  for (int d = 0; d < 3; ++d) {
    switch (s->c.recon) {
      case ORC_RECON_NONE: none_dir(s, d); break;
      case ORC_RECON_MUSCL: muscl_dir(s, d); break;
      case ORC_RECON_PPM: ppm_dir(s, d); break;
      case ORC_RECON_WENO: weno_dir(s, d); break;
      case ORC_RECON_WENO_NM: weno_nm_dir(s, d); break;
      default: return ORC_E_CONFIG;
    }
    /* compute_muscl_states / compute_ppm_states only */
    if (s->c.pbs[d] == 1 && (s->c.recon == ORC_RECON_MUSCL || s->c.recon == ORC_RECON_PPM)) pb_switch_dir(s, d);
  }
  /* src/boundary/boundary_state_reconstruction.f90:28-75 */
  if (s->c.recon != ORC_RECON_NONE)
    for (int f = 0; f < 6; ++f) {
      int id = s->c.bc_id[f];
      if (id < 0 && id != -10) reconstruct_boundary_face(s, f);
    }
  return ORC_OK;
}

/* ----------------------------------------------------------- fluxes (C5) */
typedef struct { double r, u, v, w, p; } prim;

/* convective face flux of one face for the selected scheme; the
 * turbulent scalars are carried by the split mass flux.  Sources:
 * src/face/flux/convective/{ausm,ausmP,ausmUP,slau,ldfss0,van_leer}.f90 */
static void face_flux(const orc *s, const double *qL, const double *qR, double nx, double ny, double nz,
                      double A, int zero, double *out) {
  const orc_case *c = &s->c;
  double gm = c->gm;
  int nv = s->nv;
  double rL = qL[0], uL = qL[1], vL = qL[2], wL = qL[3], pL = qL[4];
  double rR = qR[0], uR = qR[1], vR = qR[2], wR = qR[3], pR = qR[4];
  double Fp[8], Fm[8];
  if (c->flux == ORC_FLUX_AUSMP || c->flux == ORC_FLUX_AUSMUP || c->flux == ORC_FLUX_SLAU) {
    double HL = (0.5 * (uL * uL + vL * vL + wL * wL)) + ((gm / (gm - 1.)) * pL / rL);
    double HR = (0.5 * (uR * uR + vR * vR + wR * wR)) + ((gm / (gm - 1.)) * pR / rR);
    double VnL = uL * nx + vL * ny + wL * nz;
    double VnR = uR * nx + vR * ny + wR * nz;
    double mass, pbar;
    if (c->flux == ORC_FLUX_AUSMP) {
      /* ausmP.f90:105-158 */
      double cs = sqrt(2.0 * (gm - 1.0) * (0.5 * (HL + HR)) / (gm + 1.0));
      double cL = cs * cs / (fmx(cs, fabs(VnL)));
      double cR = cs * cs / (fmx(cs, fabs(VnR)));
      double C = fmn(cL, cR);
      double ML = VnL / C, MR = VnR / C;
      double alphaL = (double)((0 > 1 - (int)floor(fabs(ML))) ? 0 : 1 - (int)floor(fabs(ML)));
      double alphaR = (double)((0 > 1 - (int)floor(fabs(MR))) ? 0 : 1 - (int)floor(fabs(MR)));
      double FmL = (0.5 * (1.0 + fsign(1.0, ML)) * (1.0 - alphaL) * ML) + alphaL * 0.25 * sq(1.0 + ML);
      double betaL = (0.5 * (1.0 + fsign(1.0, ML)) * (1.0 - alphaL)) + alphaL * 0.25 * sq(1.0 + ML) * (2.0 - ML);
      double FmR = (0.5 * (1.0 - fsign(1.0, MR)) * (1.0 - alphaR) * MR) - alphaR * 0.25 * sq(1.0 - MR);
      double betaR = (0.5 * (1.0 - fsign(1.0, MR)) * (1.0 - alphaR)) + alphaR * 0.25 * sq(1.0 - MR) * (2.0 + MR);
      FmL = FmL + alphaL * 0.1250 * sq(ML * ML - 1.0);
      betaL = betaL + alphaL * 0.1875 * sq(ML * ML - 1.0) * ML;
      FmR = FmR - alphaR * 0.1250 * sq(MR * MR - 1.0);
      betaR = betaR - alphaR * 0.1875 * sq(MR * MR - 1.0) * MR;
      double Mface = FmL + FmR;
      pbar = betaL * pL + betaR * pR;
      if (Mface > 0.0) mass = Mface * C * rL; else mass = Mface * C * rR;
    } else if (c->flux == ORC_FLUX_AUSMUP) {
      /* ausmUP.f90:113-189 */
      const double Kp = 0.25, Ku = 0.75, sigma = 1.0;
      double cs = sqrt(2.0 * (gm - 1.0) * (0.5 * (HL + HR)) / (gm + 1.0));
      double cL = cs * cs / (fmx(cs, VnL));
      double cR = cs * cs / (fmx(cs, -VnR));
      double C = fmn(cL, cR);
      double ML = VnL / C, MR = VnR / C;
      double Mb = sqrt(0.5 * ((VnL * VnL) + (VnR * VnR)) / (C * C));
      double Mo = sqrt(fmn(1.0, fmx(Mb * Mb, c->MInf * c->MInf)));
      double fna = Mo * (2.0 - Mo);
      double alfa = 3.0 * (-4.0 + (5.0 * fna * fna)) / 16.0;
      double alphaL = (double)((0 > 1 - (int)floor(fabs(ML))) ? 0 : 1 - (int)floor(fabs(ML)));
      double alphaR = (double)((0 > 1 - (int)floor(fabs(MR))) ? 0 : 1 - (int)floor(fabs(MR)));
      double FmL = (0.5 * (1.0 + fsign(1.0, ML)) * (1.0 - alphaL) * ML) + alphaL * 0.25 * sq(1.0 + ML);
      double betaL = (0.5 * (1.0 + fsign(1.0, ML)) * (1.0 - alphaL)) + alphaL * 0.25 * sq(1.0 + ML) * (2.0 - ML);
      double FmR = (0.5 * (1.0 - fsign(1.0, MR)) * (1.0 - alphaR) * MR) - alphaR * 0.25 * sq(1.0 - MR);
      double betaR = (0.5 * (1.0 - fsign(1.0, MR)) * (1.0 - alphaR)) + alphaR * 0.25 * sq(1.0 - MR) * (2.0 + MR);
      FmL = FmL + alphaL * 0.125 * sq(ML * ML - 1.0);
      betaL = betaL + alphaL * alfa * sq(ML * ML - 1.0) * ML;
      FmR = FmR - alphaR * 0.125 * sq(MR * MR - 1.0);
      betaR = betaR - alphaR * alfa * sq(MR * MR - 1.0) * MR;
      double Pu = -Ku * betaL * betaR * (rL + rR) * fna * C * (VnR - VnL);
      double Mp = -2.0 * Kp * fmx(1.0 - (sigma * Mb * Mb), 0.0) * (pR - pL) / (fna * (rL + rR) * C * C);
      double Mface = FmL + FmR + Mp;
      pbar = betaL * pL + betaR * pR + Pu;
      if (Mface > 0.0) mass = Mface * C * rL; else mass = Mface * C * rR;
    } else {
      /* slau.f90:108-161 */
      double cL = sqrt(gm * pL / rL), cR = sqrt(gm * pR / rR);
      double C = 0.5 * (cL + cR);
      double delp = pR - pL;
      double ML = VnL / C, MR = VnR / C;
      double alphaL = fmx(0.0, 1.0 - floor(fabs(ML)));
      double alphaR = fmx(0.0, 1.0 - floor(fabs(MR)));
      double betaL = (1.0 - alphaL) * 0.5 * (1.0 + fsign(1.0, ML)) + (alphaL) * 0.25 * (2.0 - ML) * sq(ML + 1.0);
      double betaR = (1.0 - alphaR) * 0.5 * (1.0 - fsign(1.0, MR)) + (alphaR) * 0.25 * (2.0 + MR) * sq(MR - 1.0);
      double vtface = sqrt(0.5 * ((uL * uL) + (vL * vL) + (wL * wL) + (uR * uR) + (vR * vR) + (wR * wR)));
      double Mcap = fmn(1.0, vtface / C);
      double Xi = sq(1.0 - Mcap);
      double Vnabs = (rL * fabs(VnL) + rR * fabs(VnR)) / (rL + rR);
      double fnG = -1.0 * fmx(fmn(ML, 0.0), -1.0) * fmn(fmx(MR, 0.0), 1.0);
      pbar = 0.5 * ((pL + pR) + (betaL - betaR) * (pL - pR) + (1.0 - Xi) * (betaL + betaR - 1.0) * (pL + pR));
      double VnabsL = (1.0 - fnG) * Vnabs + fnG * fabs(VnL);
      double VnabsR = (1.0 - fnG) * Vnabs + fnG * fabs(VnR);
      mass = 0.5 * ((rL * (VnL + VnabsL) + rR * (VnR - VnabsR)) - (Xi * delp / C));
    }
    mass = mass * zero;
    Fp[0] = 0.5 * (mass + fabs(mass));
    Fp[1] = Fp[0] * uL; Fp[2] = Fp[0] * vL; Fp[3] = Fp[0] * wL; Fp[4] = Fp[0] * HL;
    Fm[0] = 0.5 * (mass - fabs(mass));
    Fm[1] = Fm[0] * uR; Fm[2] = Fm[0] * vR; Fm[3] = Fm[0] * wR; Fm[4] = Fm[0] * HR;
    for (int l = 5; l < nv; ++l) { Fp[l] = Fp[0] * qL[l]; Fm[l] = Fm[0] * qR[l]; }
    for (int l = 0; l < nv; ++l) out[l] = Fp[l] + Fm[l];
    out[1] = out[1] + (pbar * nx);
    out[2] = out[2] + (pbar * ny);
    out[3] = out[3] + (pbar * nz);
    for (int l = 0; l < nv; ++l) out[l] = out[l] * A;
    return;
  }
  /* AUSM, LDFSS(0), van Leer: ausm.f90:58-150, ldfss0.f90:60-164, van_leer.f90:56-145 */
  double ssa = 0.5 * (sqrt(gm * pL / rL) + sqrt(gm * pR / rR));
  double fns = uL * nx + vL * ny + wL * nz;
  double MpL = fns / ssa;
  double alpha_plus = 0.5 * (1.0 + fsign(1.0, MpL));
  int tb = 1 - (int)floor(fabs(MpL));
  double beta_left = -(double)(0 > tb ? 0 : tb);
  double M_plus = 0.25 * sq(1. + MpL);
  double D_plus = 0.25 * sq(1. + MpL) * (2. - MpL);
  double c_plus = (alpha_plus * (1.0 + beta_left) * MpL) - beta_left * M_plus;
  double scrD_plus = (alpha_plus * (1. + beta_left)) - (beta_left * D_plus);
  fns = uR * nx + vR * ny + wR * nz;
  double MpR = fns / ssa;
  double alpha_minus = 0.5 * (1.0 - fsign(1.0, MpR));
  tb = 1 - (int)floor(fabs(MpR));
  double beta_right = -(double)(0 > tb ? 0 : tb);
  double M_minus = -0.25 * sq(1. - MpR);
  double D_minus = 0.25 * sq(1. - MpR) * (2. + MpR);
  double c_minus = (alpha_minus * (1.0 + beta_right) * MpR) - beta_right * M_minus;
  double scrD_minus = (alpha_minus * (1. + beta_right)) - (beta_right * D_minus);
  if (c->flux == ORC_FLUX_AUSM) {
    double temp_c = c_plus + c_minus;
    c_plus = fmx(0., temp_c);
    c_minus = fmn(0., temp_c);
  } else if (c->flux == ORC_FLUX_LDFSS0) {
    double M_ldfss = 0.25 * beta_left * beta_right * sq(sqrt((MpL * MpL + MpR * MpR) * 0.5) - 1);
    double M_plus_ldfss = M_ldfss * (1 - (pL - pR) / (2 * rL * (ssa * ssa)));
    double M_minus_ldfss = M_ldfss * (1 - (pL - pR) / (2 * rR * (ssa * ssa)));
    c_plus = c_plus - M_plus_ldfss;
    c_minus = c_minus + M_minus_ldfss;
  }
  Fp[0] = rL * ssa * c_plus;
  Fm[0] = rR * ssa * c_minus;
  Fp[0] = Fp[0] * zero;
  Fm[0] = Fm[0] * zero;
  Fp[1] = (Fp[0] * uL) + (scrD_plus * pL * nx);
  Fp[2] = (Fp[0] * vL) + (scrD_plus * pL * ny);
  Fp[3] = (Fp[0] * wL) + (scrD_plus * pL * nz);
  Fp[4] = Fp[0] * ((0.5 * (uL * uL + vL * vL + wL * wL)) + ((gm / (gm - 1.)) * pL / rL));
  Fm[1] = (Fm[0] * uR) + (scrD_minus * pR * nx);
  Fm[2] = (Fm[0] * vR) + (scrD_minus * pR * ny);
  Fm[3] = (Fm[0] * wR) + (scrD_minus * pR * nz);
  Fm[4] = Fm[0] * ((0.5 * (uR * uR + vR * vR + wR * wR)) + ((gm / (gm - 1.)) * pR / rR));
  for (int l = 5; l < nv; ++l) { Fp[l] = Fp[0] * qL[l]; Fm[l] = Fm[0] * qR[l]; }
  for (int l = 0; l < nv; ++l) out[l] = Fp[l] * A + Fm[l] * A;
}

static int compute_fluxes(orc *s) {
  int nv = s->nv;
  double qL[8], qR[8], out[8];
  for (int d = 0; d < 3; ++d) {
    fstates fs = states_of(s, d);
    int ii = d == 0, jj = d == 1, kk = d == 2;
    double *Fl = d == 0 ? s->a[ORC_F] : d == 1 ? s->a[ORC_G] : s->a[ORC_H];
    size_t nfl = d == 0 ? s->nF : d == 1 ? s->nG : s->nH;
    size_t (*fidx)(const orc *, int, int, int) = d == 0 ? FI : d == 1 ? GI : HI;
    const double *fA, *fx, *fy, *fz;
    size_t (*gidx)(const orc *, int, int, int) = d == 0 ? IFI : d == 1 ? JFI : KFI;
    if (d == 0) { fA = s->a[ORC_IA]; fx = s->a[ORC_INX]; fy = s->a[ORC_INY]; fz = s->a[ORC_INZ]; }
    else if (d == 1) { fA = s->a[ORC_JA]; fx = s->a[ORC_JNX]; fy = s->a[ORC_JNY]; fz = s->a[ORC_JNZ]; }
    else { fA = s->a[ORC_KA]; fx = s->a[ORC_KNX]; fy = s->a[ORC_KNY]; fz = s->a[ORC_KNZ]; }
    if (d == 2 && KMX == 2) {
      for (size_t n = 0; n < nfl * (size_t)nv; ++n) Fl[n] = 0.;
    } else {
      for (int k = 1; k <= KMX - 1 + kk; ++k)
        for (int j = 1; j <= JMX - 1 + jj; ++j)
          for (int i = 1; i <= IMX - 1 + ii; ++i) {
            size_t si = fs.idx(s, i, j, k);
            for (int l = 0; l < nv; ++l) { qL[l] = fs.L[(size_t)l * fs.n + si]; qR[l] = fs.R[(size_t)l * fs.n + si]; }
            size_t gi = gidx(s, i, j, k);
            int zero = ii * s->mfz[i] + jj * s->mgz[j] + kk * s->mhz[k];
            face_flux(s, qL, qR, fx[gi], fy[gi], fz[gi], fA[gi], zero, out);
            size_t fo = fidx(s, i, j, k);
            for (int l = 0; l < nv; ++l) Fl[(size_t)l * nfl + fo] = out[l];
          }
    }
    for (size_t n = 0; n < nfl * (size_t)nv; ++n)
      if (isnan(Fl[n])) return ORC_E_NAN;
  }
  return ORC_OK;
}

/* ----------------------------------------------------- gradients (C6) */
#define GX(l, i, j, k) s->a[ORC_GX][(size_t)(l) * s->ngr + GRI(s, i, j, k)]
#define GY(l, i, j, k) s->a[ORC_GY][(size_t)(l) * s->ngr + GRI(s, i, j, k)]
#define GZ(l, i, j, k) s->a[ORC_GZ][(size_t)(l) * s->ngr + GRI(s, i, j, k)]

/* src/gradients.f90:405-482 */
static int gradient_G(orc *s, double *grad, const double *var, int dir) {
  const double *iA = s->a[ORC_IA], *jA = s->a[ORC_JA], *kA = s->a[ORC_KA];
  const double *in = s->a[ORC_INX + dir], *jn = s->a[ORC_JNX + dir], *kn = s->a[ORC_KNX + dir];
  const double *vol = s->a[ORC_VOL];
  for (int k = 0; k <= KMX; ++k)
    for (int j = 0; j <= JMX; ++j)
      for (int i = 0; i <= IMX; ++i) {
        double vc = var[CI(s, i, j, k)];
        double g = -(var[CI(s, i - 1, j, k)] + vc) * in[IFI(s, i, j, k)] * iA[IFI(s, i, j, k)]
                   - (var[CI(s, i, j - 1, k)] + vc) * jn[JFI(s, i, j, k)] * jA[JFI(s, i, j, k)]
                   - (var[CI(s, i, j, k - 1)] + vc) * kn[KFI(s, i, j, k)] * kA[KFI(s, i, j, k)]
                   + (var[CI(s, i + 1, j, k)] + vc) * in[IFI(s, i + 1, j, k)] * iA[IFI(s, i + 1, j, k)]
                   + (var[CI(s, i, j + 1, k)] + vc) * jn[JFI(s, i, j + 1, k)] * jA[JFI(s, i, j + 1, k)]
                   + (var[CI(s, i, j, k + 1)] + vc) * kn[KFI(s, i, j, k + 1)] * kA[KFI(s, i, j, k + 1)];
        grad[GRI(s, i, j, k)] = g / (2 * vol[CI(s, i, j, k)]);
      }
  for (size_t n = 0; n < s->ngr; ++n) if (isnan(grad[n])) return ORC_E_NAN;
  return ORC_OK;
}

/* src/gradients.f90:486-676.  apply_gradient_bc_face declares its `faces`
 * dummy with the I-face extents (-2:imx+3,-2:jmx+2,-2:kmx+2) and is called
 * with Jfaces and Kfaces too (gradients.f90:559,571,583,595); by sequence
 * association the J/K metrics are then read at the I-face linear offset.
 * That is reproduced here: every direction indexes its face arrays with IFI. */
static void gradient_bc_face(orc *s, int face) {
  int d = face / 2, hi = face % 2;
  int id = s->c.bc_id[face];
  double Tw = s->c.fix_twall[face];
  int ng = s->ngrad;
  const double *fA = s->a[d == 0 ? ORC_IA : d == 1 ? ORC_JA : ORC_KA];
  const double *fx = s->a[d == 0 ? ORC_INX : d == 1 ? ORC_JNX : ORC_KNX];
  const double *fy = s->a[d == 0 ? ORC_INY : d == 1 ? ORC_JNY : ORC_KNY];
  const double *fz = s->a[d == 0 ? ORC_INZ : d == 1 ? ORC_JNZ : ORC_KNZ];
  size_t (*gidx)(const orc *, int, int, int) = IFI;
  int n = d == 0 ? IMX : d == 1 ? JMX : KMX;
  int fpos = hi ? n : 1;
  int sig = hi ? -1 : 1;
  int na, nb; face_range(s, face, &na, &nb);
  const double *T = s->a[ORC_TEMP], *vol = s->a[ORC_VOL];
  for (int b = 1; b <= nb; ++b) for (int a = 1; a <= na; ++a) {
    int fi_, fj_, fk_;
    if (d == 0) { fi_ = fpos; fj_ = a; fk_ = b; } else if (d == 1) { fi_ = a; fj_ = fpos; fk_ = b; } else { fi_ = a; fj_ = b; fk_ = fpos; }
    /* interior I = face - iu, ghost G = face - il */
    int Ii = fi_, Ij = fj_, Ik = fk_, Gi = fi_, Gj = fj_, Gk = fk_;
    if (hi) { if (d == 0) Ii--; else if (d == 1) Ij--; else Ik--; }
    else { if (d == 0) Gi--; else if (d == 1) Gj--; else Gk--; }
    size_t gf = gidx(s, fi_, fj_, fk_);
    double nx = fx[gf], ny = fy[gf], nz = fz[gf];
    double vI = vol[CI(s, Ii, Ij, Ik)];
    double c_x = fA[gf] * nx / vI, c_y = fA[gf] * ny / vI, c_z = fA[gf] * nz / vI;
    double T_I = T[CI(s, Ii, Ij, Ik)], T_G = T[CI(s, Gi, Gj, Gk)];
    for (int l = 0; l < ng; ++l) {
      double dq = QP(l + 1, Ii, Ij, Ik) - QP(l + 1, Gi, Gj, Gk);
      GX(l, Gi, Gj, Gk) = sig * dq * c_x;
      GY(l, Gi, Gj, Gk) = sig * dq * c_y;
      GZ(l, Gi, Gj, Gk) = sig * dq * c_z;
    }
    GX(3, Gi, Gj, Gk) = sig * (T_I - T_G) * c_x;
    GY(3, Gi, Gj, Gk) = sig * (T_I - T_G) * c_y;
    GZ(3, Gi, Gj, Gk) = sig * (T_I - T_G) * c_z;
    if (id == -5 && (Tw < 1. && Tw >= 0.)) {
      GX(3, Gi, Gj, Gk) = -GX(3, Ii, Ij, Ik);
      GY(3, Gi, Gj, Gk) = -GY(3, Ii, Ij, Ik);
      GZ(3, Gi, Gj, Gk) = -GZ(3, Ii, Ij, Ik);
    }
    for (int l = 0; l < ng; ++l) {
      double dot = (GX(l, Ii, Ij, Ik) * nx) + (GY(l, Ii, Ij, Ik) * ny) + (GZ(l, Ii, Ij, Ik) * nz);
      GX(l, Gi, Gj, Gk) = GX(l, Gi, Gj, Gk) + (GX(l, Ii, Ij, Ik) - dot * nx);
      GY(l, Gi, Gj, Gk) = GY(l, Gi, Gj, Gk) + (GY(l, Ii, Ij, Ik) - dot * ny);
      GZ(l, Gi, Gj, Gk) = GZ(l, Gi, Gj, Gk) + (GZ(l, Ii, Ij, Ik) - dot * nz);
    }
  }
}

/* src/gradients.f90:276-402 */
static int evaluate_all_gradients(orc *s) {
  double *gxs[3] = {s->a[ORC_GX], s->a[ORC_GY], s->a[ORC_GZ]};
  const double *vars[7] = {QPL(1), QPL(2), QPL(3), s->a[ORC_TEMP], QPL(5), QPL(6), LCTM(s) ? QPL(7) : NULL};
  int ng = s->ngrad;
  for (int dir = 0; dir < 3; ++dir) {
    if (dir == 2 && KMX <= 2) {
      for (size_t n = 0; n < s->ngr * (size_t)ng; ++n) gxs[2][n] = 0.0;
      continue;
    }
    for (int l = 0; l < ng; ++l) {
      int st = gradient_G(s, gxs[dir] + (size_t)l * s->ngr, vars[l], dir);
      if (st) return st;
    }
  }
  for (int f = 0; f < 6; ++f)
    if (s->c.bc_id[f] < 0) gradient_bc_face(s, f);
  return ORC_OK;
}

/* ---------------------------------------------------- viscosity (C7) */
/* src/viscosity.f90:52-546 (constant / Sutherland mu, SST mu_t and F1) */
static int calculate_viscosity(orc *s) {
  const orc_case *c = &s->c;
  double *mu = s->a[ORC_MU], *mut = s->a[ORC_MUT], *F1 = s->a[ORC_F1], *dist = s->a[ORC_DIST];
  int s2003 = c->turb == ORC_TURB_SST2003;
  if (c->mu_sutherland) {
    for (int k = 0; k <= KMX; ++k) for (int j = 0; j <= JMX; ++j) for (int i = 0; i <= IMX; ++i) {
      size_t ci = CI(s, i, j, k);
      double T = QP(4, i, j, k) / (QP(0, i, j, k) * c->R_gas);
      mu[ci] = c->mu_ref * (pow(T / c->T_ref, 1.5)) * ((c->T_ref + c->S_ref) / (T + c->S_ref));
    }
  }
  if (SA(s)) {
    /* src/viscosity.f90:149-163: mu_t = rho nu~ fv1 */
    for (int k = 0; k <= KMX; ++k) for (int j = 0; j <= JMX; ++j) for (int i = 0; i <= IMX; ++i) {
      size_t ci = CI(s, i, j, k);
      double tv = QP(5, i, j, k), density = QP(0, i, j, k);
      double xi = tv * density / mu[ci];
      double fv1 = ipow(xi, 3) / (ipow(xi, 3) + CV1_3);
      mut[ci] = density * tv * fv1;
    }
  }
  if (SST(s)) {
    for (int k = 0; k <= KMX; ++k) for (int j = 0; j <= JMX; ++j) for (int i = 0; i <= IMX; ++i) {
      size_t ci = CI(s, i, j, k);
      double density = QP(0, i, j, k), tk = QP(5, i, j, k), tw = QP(6, i, j, k);
      double var1 = sqrt(tk) / (bstar * tw * dist[ci]);
      double var2 = 500 * (mu[ci] / density) / ((dist[ci] * dist[ci]) * tw);
      double arg2 = fmx(2 * var1, var2);
      double F = tanh(arg2 * arg2);
      double NUM = density * a1_sst * tk, DENOM;
      if (!s2003) {
        double wx = (GY(2, i, j, k) - GZ(1, i, j, k));
        double wy = (GZ(0, i, j, k) - GX(2, i, j, k));
        double wz = (GX(1, i, j, k) - GY(0, i, j, k));
        double wijwij = wx * wx + wy * wy + wz * wz;
        double vort = sqrt(wijwij);
        DENOM = fmx(fmx((a1_sst * tw), vort * F), 1.e-20);
      } else {
        /* src/viscosity.f90:215-275: the strain rate, limits 1e-10 */
        double sxx = GX(0, i, j, k), syy = GY(1, i, j, k), szz = GZ(2, i, j, k);
        double syz = GY(2, i, j, k) + GZ(1, i, j, k), szx = GZ(0, i, j, k) + GX(2, i, j, k);
        double sxy = GX(1, i, j, k) + GY(0, i, j, k);
        double SijSij = (2.0 * (sxx * sxx)) + (2.0 * (syy * syy)) + (2.0 * (szz * szz)) + syz * syz + szx * szx +
                        sxy * sxy;
        double strain = sqrt(SijSij);
        DENOM = fmx(fmx((a1_sst * tw), strain * F), 1.0e-10);
      }
      mut[ci] = NUM / DENOM;
      double CD = fmx(2 * density * sigma_w2 * (GX(4, i, j, k) * GX(5, i, j, k) + GY(4, i, j, k) * GY(5, i, j, k) +
                                                 GZ(4, i, j, k) * GZ(5, i, j, k)) / tw,
                      s2003 ? 1.0e-10 : 1e-20);
      double right = 4 * (density * sigma_w2 * tk) / (CD * (dist[ci] * dist[ci]));
      double left = fmx(var1, var2);
      double arg1 = fmn(left, right);
      F1[ci] = tanh(arg1 * arg1 * arg1 * arg1);
    }
    if (LCTM(s)) {
      /* src/viscosity.f90:265-279 (sst) and 390-404 (sst2003): Menter's
       * 2015 blending, F1 = max(F1, exp(-(Ry/120)**8)).  Ry reads the
       * scalars density and tk, which still hold the F1 loop's last cell
       * (imx, jmx, kmx): the reference does not reload them per cell */
      const double density = QP(0, IMX, JMX, KMX), tk = QP(5, IMX, JMX, KMX);
      for (int k = 0; k <= KMX; ++k) for (int j = 0; j <= JMX; ++j) for (int i = 0; i <= IMX; ++i) {
        size_t ci = CI(s, i, j, k);
        double var1 = density * dist[ci] * sqrt(tk) / mu[ci];
        double var2 = exp(-ipow(var1 / 120, 8));
        F1[ci] = fmx(F1[ci], var2);
      }
    }
  }
  if (KKL(s)) {
    /* src/viscosity.f90:469-483: mu_t = cmu^(1/4) rho kL / sqrt(k), zero
     * where k or kL is below 1e-14 */
    double cc = pow(cmu_kkl, 0.25);
    for (int k = 0; k <= KMX; ++k) for (int j = 0; j <= JMX; ++j) for (int i = 0; i <= IMX; ++i) {
      size_t ci = CI(s, i, j, k);
      double density = QP(0, i, j, k), tk = QP(5, i, j, k), tkl = QP(6, i, j, k);
      mut[ci] = cc * density * tkl / (fmx(sqrt(tk), 1.e-20));
      if (tkl < 1.e-14 || tk < 1.e-14) mut[ci] = 0.0;
    }
  }
  if (c->turb != ORC_TURB_NONE) {
    for (int f = 0; f < 6; ++f) {
      int id = c->bc_id[f];
      if (id >= 0 || id == -10) continue;
      int anti = (id == -5);
      if (!(id == -5 || (id <= -1 && id >= -9))) continue;
      if (KKL(s) && id == -7) continue; /* k-kL: -4..-1, -6, -8, -9 symm and -5 anti (:485-531) */
      int na, nb; face_range(s, f, &na, &nb);
      for (int b = 1; b <= nb; ++b) for (int a = 1; a <= na; ++a) {
        size_t cg = face_cell(s, f, 0, a, b), ci = face_cell(s, f, 1, a, b);
        mut[cg] = anti ? -mut[ci] : mut[ci];
        if (SST(s)) F1[cg] = F1[ci];
      }
    }
  }
  for (size_t n = 0; n < s->ncell; ++n) if (isnan(mu[n])) return ORC_E_NAN;
  return ORC_OK;
}

/* ----------------------------------------------- viscous fluxes (C8) */
/* src/viscous.f90:144-325 (laminar) and 328-447 (SST) */
static void viscous_dir(orc *s, int d, int sst) {
  const orc_case *c = &s->c;
  int ii = d == 0, jj = d == 1, kk = d == 2;
  double *Fl = d == 0 ? s->a[ORC_F] : d == 1 ? s->a[ORC_G] : s->a[ORC_H];
  size_t nfl = d == 0 ? s->nF : d == 1 ? s->nG : s->nH;
  size_t (*fidx)(const orc *, int, int, int) = d == 0 ? FI : d == 1 ? GI : HI;
  size_t (*gidx)(const orc *, int, int, int) = d == 0 ? IFI : d == 1 ? JFI : KFI;
  const double *fA = s->a[d == 0 ? ORC_IA : d == 1 ? ORC_JA : ORC_KA];
  const double *fx = s->a[d == 0 ? ORC_INX : d == 1 ? ORC_JNX : ORC_KNX];
  const double *fy = s->a[d == 0 ? ORC_INY : d == 1 ? ORC_JNY : ORC_KNY];
  const double *fz = s->a[d == 0 ? ORC_INZ : d == 1 ? ORC_JNZ : ORC_KNZ];
  const double *cx = s->a[ORC_CX], *cy = s->a[ORC_CY], *cz = s->a[ORC_CZ];
  const double *mu = s->a[ORC_MU], *mut = s->a[ORC_MUT], *F1a = s->a[ORC_F1];
  int turbulent = (c->turb != ORC_TURB_NONE);
  for (int k = 1; k <= KMX - 1 + kk; ++k)
    for (int j = 1; j <= JMX - 1 + jj; ++j)
      for (int i = 1; i <= IMX - 1 + ii; ++i) {
        int im = i - ii, jm = j - jj, km = k - kk;
        size_t cR = CI(s, i, j, k), cL = CI(s, im, jm, km);
        size_t gf = gidx(s, i, j, k), fo = fidx(s, i, j, k);
        double delx = cx[cR] - cx[cL], dely = cy[cR] - cy[cL], delz = cz[cR] - cz[cL];
        double d_LR = sqrt(delx * delx + dely * dely + delz * delz);
        double nx = fx[gf], ny = fy[gf], nz = fz[gf], area = fA[gf];
        double mu_f = 0.5 * (mu[cL] + mu[cR]);
        if (!sst) {
          double g[4][3];
          for (int l = 0; l < 4; ++l) {
            g[l][0] = 0.5 * (GX(l, im, jm, km) + GX(l, i, j, k));
            g[l][1] = 0.5 * (GY(l, im, jm, km) + GY(l, i, j, k));
            g[l][2] = 0.5 * (GZ(l, im, jm, km) + GZ(l, i, j, k));
          }
          double T_LE = QP(4, im, jm, km) / (QP(0, im, jm, km) * c->R_gas);
          double T_RE = QP(4, i, j, k) / (QP(0, i, j, k) * c->R_gas);
          double del[4];
          del[0] = QP(1, i, j, k) - QP(1, im, jm, km);
          del[1] = QP(2, i, j, k) - QP(2, im, jm, km);
          del[2] = QP(3, i, j, k) - QP(3, im, jm, km);
          del[3] = T_RE - T_LE;
          for (int l = 0; l < 4; ++l) {
            double nc = (del[l] - (g[l][0] * delx + g[l][1] * dely + g[l][2] * delz)) / d_LR;
            g[l][0] = g[l][0] + (nc * delx / d_LR);
            g[l][1] = g[l][1] + (nc * dely / d_LR);
            g[l][2] = g[l][2] + (nc * delz / d_LR);
          }
          double dudx = g[0][0], dudy = g[0][1], dudz = g[0][2];
          double dvdx = g[1][0], dvdy = g[1][1], dvdz = g[1][2];
          double dwdx = g[2][0], dwdy = g[2][1], dwdz = g[2][2];
          double mut_f = turbulent ? 0.5 * (mut[cL] + mut[cR]) : 0.0;
          double total_mu = mu_f + mut_f;
          double Txx = 2. * total_mu * (dudx - ((dudx + dvdy + dwdz) / 3.));
          double Tyy = 2. * total_mu * (dvdy - ((dudx + dvdy + dwdz) / 3.));
          double Tzz = 2. * total_mu * (dwdz - ((dudx + dvdy + dwdz) / 3.));
          double Txy = total_mu * (dvdx + dudy);
          double Txz = total_mu * (dwdx + dudz);
          double Tyz = total_mu * (dwdy + dvdz);
          double K_heat = (mu_f / c->Pr + mut_f / c->tPr) * c->gm * c->R_gas / (c->gm - 1);
          double Qx = K_heat * g[3][0], Qy = K_heat * g[3][1], Qz = K_heat * g[3][2];
          double uface = 0.5 * (QP(1, im, jm, km) + QP(1, i, j, k));
          double vface = 0.5 * (QP(2, im, jm, km) + QP(2, i, j, k));
          double wface = 0.5 * (QP(3, im, jm, km) + QP(3, i, j, k));
          Fl[1 * nfl + fo] = Fl[1 * nfl + fo] - ((Txx * nx + Txy * ny + Txz * nz) * area);
          Fl[2 * nfl + fo] = Fl[2 * nfl + fo] - ((Txy * nx + Tyy * ny + Tyz * nz) * area);
          Fl[3 * nfl + fo] = Fl[3 * nfl + fo] - ((Txz * nx + Tyz * ny + Tzz * nz) * area);
          Fl[4 * nfl + fo] = Fl[4 * nfl + fo] - (area * (((Txx * uface + Txy * vface + Txz * wface + Qx) * nx) +
                                                        ((Txy * uface + Tyy * vface + Tyz * wface + Qy) * ny) +
                                                        ((Txz * uface + Tyz * vface + Tzz * wface + Qz) * nz)));
        } else if (sst == 4) {
          /* compute_viscous_fluxes_lctm2015, src/viscous.f90:659-746:
           * (mu + mu_t) grad(gamma), no sigma */
          double g0 = 0.5 * (GX(6, im, jm, km) + GX(6, i, j, k));
          double g1 = 0.5 * (GY(6, im, jm, km) + GY(6, i, j, k));
          double g2 = 0.5 * (GZ(6, im, jm, km) + GZ(6, i, j, k));
          double deltgm = QP(7, i, j, k) - QP(7, im, jm, km);
          double nc = (deltgm - (g0 * delx + g1 * dely + g2 * delz)) / d_LR;
          g0 = g0 + (nc * delx / d_LR);
          g1 = g1 + (nc * dely / d_LR);
          g2 = g2 + (nc * delz / d_LR);
          double mut_f = 0.5 * (mut[cL] + mut[cR]);
          Fl[7 * nfl + fo] = Fl[7 * nfl + fo] - (area * ((mu_f + mut_f) * (g0 * nx + g1 * ny + g2 * nz)));
        } else if (sst == 2) {
          /* compute_viscous_fluxes_sa, src/viscous.f90:570-656 */
          double g0 = 0.5 * (GX(4, im, jm, km) + GX(4, i, j, k));
          double g1 = 0.5 * (GY(4, im, jm, km) + GY(4, i, j, k));
          double g2 = 0.5 * (GZ(4, im, jm, km) + GZ(4, i, j, k));
          double deltv = QP(5, i, j, k) - QP(5, im, jm, km);
          double nc = (deltv - (g0 * delx + g1 * dely + g2 * delz)) / d_LR;
          g0 = g0 + (nc * delx / d_LR);
          g1 = g1 + (nc * dely / d_LR);
          g2 = g2 + (nc * delz / d_LR);
          double rhoface = 0.5 * (QP(0, im, jm, km) + QP(0, i, j, k));
          double mut_f = 0.5 * (QP(5, im, jm, km) + QP(5, i, j, k)) * rhoface;
          Fl[5 * nfl + fo] =
              Fl[5 * nfl + fo] - (area * ((mu_f + mut_f) * (g0 * nx + g1 * ny + g2 * nz))) / sigma_sa;
        } else {
          double g[2][3];
          for (int l = 0; l < 2; ++l) {
            g[l][0] = 0.5 * (GX(4 + l, im, jm, km) + GX(4 + l, i, j, k));
            g[l][1] = 0.5 * (GY(4 + l, im, jm, km) + GY(4 + l, i, j, k));
            g[l][2] = 0.5 * (GZ(4 + l, im, jm, km) + GZ(4 + l, i, j, k));
          }
          double del[2];
          del[0] = QP(5, i, j, k) - QP(5, im, jm, km);
          del[1] = QP(6, i, j, k) - QP(6, im, jm, km);
          for (int l = 0; l < 2; ++l) {
            double nc = (del[l] - (g[l][0] * delx + g[l][1] * dely + g[l][2] * delz)) / d_LR;
            g[l][0] = g[l][0] + (nc * delx / d_LR);
            g[l][1] = g[l][1] + (nc * dely / d_LR);
            g[l][2] = g[l][2] + (nc * delz / d_LR);
          }
          double mut_f = 0.5 * (mut[cL] + mut[cR]);
          double F1 = 0.5 * (F1a[cL] + F1a[cR]);
          double sigma_kf = sigma_k1 * F1 + sigma_k2 * (1.0 - F1);
          double sigma_wf = sigma_w1 * F1 + sigma_w2 * (1.0 - F1);
          if (sst == 3) sigma_kf = sigma_wf = 1.0; /* compute_viscous_fluxes_kkl, src/viscous.f90:450-567 */
          double rhoface = 0.5 * (QP(0, im, jm, km) + QP(0, i, j, k));
          double tkface = 0.5 * (QP(5, im, jm, km) + QP(5, i, j, k));
          double Txx = -(2.0 * rhoface * tkface / 3.0);
          Fl[1 * nfl + fo] = Fl[1 * nfl + fo] - (Txx * nx * area);
          Fl[2 * nfl + fo] = Fl[2 * nfl + fo] - (Txx * ny * area);
          Fl[3 * nfl + fo] = Fl[3 * nfl + fo] - (Txx * nz * area);
          double kflux = area * ((mu_f + sigma_kf * mut_f) * (g[0][0] * nx + g[0][1] * ny + g[0][2] * nz));
          Fl[4 * nfl + fo] = Fl[4 * nfl + fo] - kflux;
          Fl[5 * nfl + fo] = Fl[5 * nfl + fo] - kflux;
          Fl[6 * nfl + fo] = Fl[6 * nfl + fo] - (area * ((mu_f + sigma_wf * mut_f) * (g[1][0] * nx + g[1][1] * ny + g[1][2] * nz)));
        }
      }
}

static int compute_viscous_fluxes(orc *s) {
  for (int d = 0; d < 3; ++d) viscous_dir(s, d, 0);
  if (SST(s)) {
    viscous_dir(s, 0, 1);
    viscous_dir(s, 1, 1);
    if (KMX != 2) viscous_dir(s, 2, 1);
  }
  /* the SA flux runs in all three directions, 2-D grids too
   * (src/viscous.f90:94-97) */
  if (SA(s))
    for (int d = 0; d < 3; ++d) viscous_dir(s, d, 2);
  /* so does the k-kL flux (src/viscous.f90:106-109) */
  if (KKL(s))
    for (int d = 0; d < 3; ++d) viscous_dir(s, d, 3);
  /* lctm2015: the intermittency flux, none on the k faces of a 2-D grid
   * (src/viscous.f90:114-128) */
  if (LCTM(s)) {
    viscous_dir(s, 0, 4);
    viscous_dir(s, 1, 4);
    if (KMX != 2) viscous_dir(s, 2, 4);
  }
  size_t nn[3] = {s->nF, s->nG, s->nH};
  double *arr[3] = {s->a[ORC_F], s->a[ORC_G], s->a[ORC_H]};
  for (int d = 0; d < 3; ++d)
    for (size_t n = 0; n < nn[d] * (size_t)s->nv; ++n) if (isnan(arr[d][n])) return ORC_E_NAN;
  return ORC_OK;
}

/* ---------------------------------------------- residue + source (C9, C10) */
/* src/face/flux/convective/scheme.f90:111-141 */
static void compute_residue(orc *s) {
  double *F = s->a[ORC_F], *G = s->a[ORC_G], *H = s->a[ORC_H], *R = s->a[ORC_RES];
  for (int l = 0; l < s->nv; ++l)
    for (int k = 1; k <= KMX - 1; ++k) for (int j = 1; j <= JMX - 1; ++j) for (int i = 1; i <= IMX - 1; ++i)
      R[(size_t)l * s->nR + RI(s, i, j, k)] =
          (F[(size_t)l * s->nF + FI(s, i + 1, j, k)] - F[(size_t)l * s->nF + FI(s, i, j, k)]) +
          (G[(size_t)l * s->nG + GI(s, i, j + 1, k)] - G[(size_t)l * s->nG + GI(s, i, j, k)]) +
          (H[(size_t)l * s->nH + HI(s, i, j, k + 1)] - H[(size_t)l * s->nH + HI(s, i, j, k)]);
}

/* add_sa_source, src/source.f90:835-983.  The density gradient is a
 * Green-Gauss sum of its own over the face averages; its k-face normal is
 * (nx, nx, nx), as the reference writes it (source.f90:887). */
static void add_sa_source(orc *s) {
  double *R = s->a[ORC_RES], *vol = s->a[ORC_VOL], *mu = s->a[ORC_MU], *dist = s->a[ORC_DIST];
  const double cw1 = CW1_SA;
  for (int k = 1; k <= KMX - 1; ++k) for (int j = 1; j <= JMX - 1; ++j) for (int i = 1; i <= IMX - 1; ++i) {
    size_t ci = CI(s, i, j, k);
    double density = QP(0, i, j, k), tv = QP(5, i, j, k);
    double rf[6] = {QP(0, i - 1, j, k) + density, QP(0, i, j - 1, k) + density, QP(0, i, j, k - 1) + density,
                    QP(0, i + 1, j, k) + density, QP(0, i, j + 1, k) + density, QP(0, i, j, k + 1) + density};
    size_t f[6] = {IFI(s, i, j, k), JFI(s, i, j, k), KFI(s, i, j, k),
                   IFI(s, i + 1, j, k), JFI(s, i, j + 1, k), KFI(s, i, j, k + 1)};
    const int arr[6][4] = {{ORC_IA, ORC_INX, ORC_INY, ORC_INZ}, {ORC_JA, ORC_JNX, ORC_JNY, ORC_JNZ},
                           {ORC_KA, ORC_KNX, ORC_KNX, ORC_KNX}, {ORC_IA, ORC_INX, ORC_INY, ORC_INZ},
                           {ORC_JA, ORC_JNX, ORC_JNY, ORC_JNZ}, {ORC_KA, ORC_KNX, ORC_KNY, ORC_KNZ}};
    double gr[3];
    for (int c3 = 0; c3 < 3; ++c3) {
      double g = -(rf[0] * s->a[arr[0][1 + c3]][f[0]] * s->a[arr[0][0]][f[0]]);
      g = g - rf[1] * s->a[arr[1][1 + c3]][f[1]] * s->a[arr[1][0]][f[1]];
      g = g - rf[2] * s->a[arr[2][1 + c3]][f[2]] * s->a[arr[2][0]][f[2]];
      g = g + rf[3] * s->a[arr[3][1 + c3]][f[3]] * s->a[arr[3][0]][f[3]];
      g = g + rf[4] * s->a[arr[4][1 + c3]][f[4]] * s->a[arr[4][0]][f[4]];
      g = g + rf[5] * s->a[arr[5][1 + c3]][f[5]] * s->a[arr[5][0]][f[5]];
      gr[c3] = g / (2.0 * vol[ci]);
    }
    double vort = sqrt(sq(GY(2, i, j, k) - GZ(1, i, j, k)) + sq(GZ(0, i, j, k) - GX(2, i, j, k)) +
                       sq(GX(1, i, j, k) - GY(0, i, j, k)));
    double CD1 = cb2 * ((GX(4, i, j, k) * GX(4, i, j, k)) + (GY(4, i, j, k) * GY(4, i, j, k)) +
                        (GZ(4, i, j, k) * GZ(4, i, j, k)));
    double CD2 = ((gr[0] * GX(4, i, j, k)) + (gr[1] * GY(4, i, j, k)) + (gr[2] * GZ(4, i, j, k)));
    double kd = kappa_sa * dist[ci], kd2 = kd * kd;
    double nu = mu[ci] / density;
    double xi = tv / nu;
    double fv1 = ipow(xi, 3) / (ipow(xi, 3) + CV1_3);
    double fv2 = 1.0 - xi / (1.0 + (xi * fv1));
    double scap = fmx(vort + (tv * fv2 / (kd2)), 0.3 * vort);
    double r = fmn(tv / (scap * kd2), 10.0);
    double g = r + cw2 * (ipow(r, 6) - r);
    double fw = g * pow((1.0 + CW3_6) / (ipow(g, 6) + CW3_6), 1.0 / 6.0);
    double td = tv / dist[ci];
    double D_v = density * cw1 * fw * (td * td);
    double P_v = density * cb1 * scap * tv;
    double lamda = density * CD1 / sigma_sa - CD2 * (nu + tv) / sigma_sa;
    double S_v = (P_v - D_v + lamda) * vol[ci];
    R[(size_t)5 * s->nR + RI(s, i, j, k)] = R[(size_t)5 * s->nR + RI(s, i, j, k)] - S_v;
  }
}


/* The density gradient of the SA sources: a Green-Gauss sum over the face
 * averages rho_neighbour + rho, whose k-face normal is (nx, nx, nx)
 * (source.f90:887, 1039) */
static void sa_rho_gradient(const orc *s, int i, int j, int k, double gr[3]) {
  double density = QP(0, i, j, k);
  double rf[6] = {QP(0, i - 1, j, k) + density, QP(0, i, j - 1, k) + density, QP(0, i, j, k - 1) + density,
                  QP(0, i + 1, j, k) + density, QP(0, i, j + 1, k) + density, QP(0, i, j, k + 1) + density};
  size_t f[6] = {IFI(s, i, j, k), JFI(s, i, j, k), KFI(s, i, j, k),
                 IFI(s, i + 1, j, k), JFI(s, i, j + 1, k), KFI(s, i, j, k + 1)};
  const int arr[6][4] = {{ORC_IA, ORC_INX, ORC_INY, ORC_INZ}, {ORC_JA, ORC_JNX, ORC_JNY, ORC_JNZ},
                         {ORC_KA, ORC_KNX, ORC_KNX, ORC_KNX}, {ORC_IA, ORC_INX, ORC_INY, ORC_INZ},
                         {ORC_JA, ORC_JNX, ORC_JNY, ORC_JNZ}, {ORC_KA, ORC_KNX, ORC_KNY, ORC_KNZ}};
  double vol = s->a[ORC_VOL][CI(s, i, j, k)];
  for (int c3 = 0; c3 < 3; ++c3) {
    double g = -(rf[0] * s->a[arr[0][1 + c3]][f[0]] * s->a[arr[0][0]][f[0]]);
    g = g - rf[1] * s->a[arr[1][1 + c3]][f[1]] * s->a[arr[1][0]][f[1]];
    g = g - rf[2] * s->a[arr[2][1 + c3]][f[2]] * s->a[arr[2][0]][f[2]];
    g = g + rf[3] * s->a[arr[3][1 + c3]][f[3]] * s->a[arr[3][0]][f[3]];
    g = g + rf[4] * s->a[arr[4][1 + c3]][f[4]] * s->a[arr[4][0]][f[4]];
    g = g + rf[5] * s->a[arr[5][1 + c3]][f[5]] * s->a[arr[5][0]][f[5]];
    gr[c3] = g / (2 * vol);
  }
}

/* The 'bc' intermittency gamma_BC = 1 - exp(-(term1 + term2)) of
 * source.f90:1128-1142 and 566-580; nu_t, re_v from the caller */
static double gamma_bc(double nu_t, double vmag, double dist, double re_v, double Re, double tu) {
  double chi_1 = 0.002, chi_2 = 5.0;
  double nu_cr = chi_2 / Re;
  double nu_bc = nu_t / (vmag * dist);
  double re_theta = re_v / 2.193;
  double re_theta_t = (803.73 * pow(tu + 0.6067, -1.027));
  double term1 = sqrt(fmx(re_theta - re_theta_t, 0.) / (chi_1 * re_theta_t));
  double term2 = sqrt(fmx(nu_bc - nu_cr, 0.0) / nu_cr);
  return 1.0 - exp(-(term1 + term2));
}

static double reynolds(const orc *s) {
  /* state.f90:89 */
  return s->c.rho_inf * s->c.vel_mag * 1.0 / s->c.mu_ref;
}

/* add_saBC_source, src/source.f90:985-1194: production (kinematic, times
 * gamma_BC) and destruction without density, the cross diffusion with */
static void add_sa_bc_source(orc *s) {
  double *R = s->a[ORC_RES], *vol = s->a[ORC_VOL], *mu = s->a[ORC_MU], *dist = s->a[ORC_DIST];
  const double cw1 = CW1_SA, Re = reynolds(s), tu = s->c.tu_inf;
  for (int k = 1; k <= KMX - 1; ++k) for (int j = 1; j <= JMX - 1; ++j) for (int i = 1; i <= IMX - 1; ++i) {
    size_t ci = CI(s, i, j, k);
    double density = QP(0, i, j, k), u = QP(1, i, j, k), v = QP(2, i, j, k), w = QP(3, i, j, k);
    double tv = QP(5, i, j, k);
    double vmag = sqrt(u * u + v * v + w * w);
    double gr[3];
    sa_rho_gradient(s, i, j, k, gr);
    double Omega = sqrt(sq(GY(2, i, j, k) - GZ(1, i, j, k)) + sq(GZ(0, i, j, k) - GX(2, i, j, k)) +
                        sq(GX(1, i, j, k) - GY(0, i, j, k)));
    double CD1 = cb2 * ((GX(4, i, j, k) * GX(4, i, j, k)) + (GY(4, i, j, k) * GY(4, i, j, k)) +
                        (GZ(4, i, j, k) * GZ(4, i, j, k)));
    double CD2 = ((gr[0] * GX(4, i, j, k)) + (gr[1] * GY(4, i, j, k)) + (gr[2] * GZ(4, i, j, k)));
    double dist_i = dist[ci], dist_i_2 = dist_i * dist_i, k2 = kappa_sa * kappa_sa;
    double nu = mu[ci] / density;
    double Ji = tv / nu, Ji_2 = Ji * Ji, Ji_3 = Ji_2 * Ji;
    double fv1 = (Ji_3) / ((Ji_3) + (CV1_3));
    double fv2 = 1.0 - Ji / (1.0 + (Ji * fv1));
    double S = Omega;
    double inv_k2_d2 = 1.0 / (k2 * dist_i_2);
    double Shat = S + tv * fv2 * inv_k2_d2;
    Shat = fmx(Shat, 1.0e-10);
    double inv_Shat = 1.0 / Shat;
    double nu_t = tv * fv1;
    double re_v = dist_i_2 * Omega / nu;
    double gBC = gamma_bc(nu_t, vmag, dist_i, re_v, Re, tu);
    double Production = gBC * cb1 * Shat * tv * vol[ci];
    double r = fmn(tv * inv_Shat * inv_k2_d2, 10.0);
    double g = r + cw2 * ((ipow(r, 6)) - r);
    double g_6 = ipow(g, 6);
    double glim = pow((1.0 + CW3_6) / (g_6 + CW3_6), 1.0 / 6.0);
    double fw = g * glim;
    double Destruction = (cw1 * fw * tv * tv / dist_i_2) * (vol[ci]);
    double lamda = (density * CD1 / sigma_sa - CD2 * (nu + tv) / sigma_sa) * vol[ci];
    double S_v = (Production - Destruction + lamda);
    R[(size_t)5 * s->nR + RI(s, i, j, k)] = R[(size_t)5 * s->nR + RI(s, i, j, k)] - S_v;
  }
}

/* add_sst_bc_source, src/source.f90:467-604: the SST source with the
 * production mu_t Omega^2 (no divergence term), limited at 20 D_k, the
 * original gama1/gama2 whatever the SST form, and P_k scaled by gamma_BC
 * after P_omega is formed */
static void add_sst_bc_source(orc *s) {
  double *R = s->a[ORC_RES], *mut = s->a[ORC_MUT], *F1a = s->a[ORC_F1], *vol = s->a[ORC_VOL];
  double *mu = s->a[ORC_MU], *dist = s->a[ORC_DIST];
  const double gama1 = (beta1 / bstar) - ((sigma_w1 * (kappa_sst * kappa_sst)) / sqrt(bstar));
  const double gama2 = (beta2 / bstar) - ((sigma_w2 * (kappa_sst * kappa_sst)) / sqrt(bstar));
  const double Re = reynolds(s), tu = s->c.tu_inf;
  for (int k = 1; k <= KMX - 1; ++k) for (int j = 1; j <= JMX - 1; ++j) for (int i = 1; i <= IMX - 1; ++i) {
    size_t ci = CI(s, i, j, k);
    double density = QP(0, i, j, k), tk = QP(5, i, j, k), tw = QP(6, i, j, k);
    double vort = sqrt(sq(GY(2, i, j, k) - GZ(1, i, j, k)) + sq(GZ(0, i, j, k) - GX(2, i, j, k)) +
                       sq(GX(1, i, j, k) - GY(0, i, j, k)));
    double CD = 2 * density * sigma_w2 * (GX(4, i, j, k) * GX(5, i, j, k) + GY(4, i, j, k) * GY(5, i, j, k) +
                                          GZ(4, i, j, k) * GZ(5, i, j, k)) / tw;
    double F1 = F1a[ci];
    double gama = gama1 * F1 + gama2 * (1. - F1);
    double beta = beta1 * F1 + beta2 * (1. - F1);
    double D_k = bstar * density * tw * tk;
    double D_w = beta * density * (tw * tw);
    double P_k = mut[ci] * (vort * vort);
    P_k = fmn(P_k, 20.0 * D_k);
    double P_w = (density * gama / mut[ci]) * P_k;
    double lamda = (1. - F1) * CD;
    double u = QP(1, i, j, k), v = QP(2, i, j, k), w = QP(3, i, j, k);
    double vmag = sqrt(u * u + v * v + w * w);
    double nu_t = mut[ci] / density;
    double re_v = density * dist[ci] * dist[ci] * vort / mu[ci];
    double gBC = gamma_bc(nu_t, vmag, dist[ci], re_v, Re, tu);
    P_k = gBC * P_k;
    double S_k = P_k - D_k;
    double S_w = P_w - D_w + lamda;
    S_k = S_k * vol[ci];
    S_w = S_w * vol[ci];
    R[(size_t)5 * s->nR + RI(s, i, j, k)] = R[(size_t)5 * s->nR + RI(s, i, j, k)] - S_k;
    R[(size_t)6 * s->nR + RI(s, i, j, k)] = R[(size_t)6 * s->nR + RI(s, i, j, k)] - S_w;
  }
}

static void add_sst_source(orc *s);

/* src/source.f90:607-832 (add_kkl_source): production limited to 20 D_k,
 * the von Karman length from the second velocity derivatives (each a
 * Green-Gauss sum over the six faces of the face-averaged first
 * derivatives, the ghost-cell gradients included), its limiters, and the
 * near-wall damping of both equations */
static void add_kkl_source(orc *s) {
  const double *mu = s->a[ORC_MU], *mut = s->a[ORC_MUT], *dist = s->a[ORC_DIST], *vol = s->a[ORC_VOL];
  const double *iA = s->a[ORC_IA], *jA = s->a[ORC_JA], *kA = s->a[ORC_KA];
  const double *inx[3] = {s->a[ORC_INX], s->a[ORC_INY], s->a[ORC_INZ]};
  const double *jnx[3] = {s->a[ORC_JNX], s->a[ORC_JNY], s->a[ORC_JNZ]};
  const double *knx[3] = {s->a[ORC_KNX], s->a[ORC_KNY], s->a[ORC_KNZ]};
  double *res = s->a[ORC_RES];
  for (int k = 1; k <= KMX - 1; ++k) for (int j = 1; j <= JMX - 1; ++j) for (int i = 1; i <= IMX - 1; ++i) {
    size_t ci = CI(s, i, j, k), ri = RI(s, i, j, k);
    double density = QP(0, i, j, k), tk = QP(5, i, j, k), tkl = QP(6, i, j, k);
    /* g[v][d]: derivative of velocity v along d */
    double g[3][3];
    for (int v = 0; v < 3; ++v) { g[v][0] = GX(v, i, j, k); g[v][1] = GY(v, i, j, k); g[v][2] = GZ(v, i, j, k); }
    double S[3][3];
    for (int a = 0; a < 3; ++a) for (int b = 0; b < 3; ++b) S[a][b] = 0.5 * (g[a][b] + g[b][a]);
    double delv = g[0][0] + g[1][1] + g[2][2];
    double P_k = 0.;
    for (int a = 0; a < 3; ++a) for (int b = 0; b < 3; ++b) {
      double tau = mut[ci] * (2 * S[a][b]);
      if (a == b) tau = mut[ci] * (2 * S[a][b] - (2.0 / 3.0) * delv) - (2.0 / 3.0) * density * tk;
      P_k = P_k + tau * g[a][b];
    }
    double D_k = (pow(cmu_kkl, 0.75)) * density * (pow(tk, 2.5)) / fmx(tkl, 1.e-20);
    P_k = fmn(P_k, 20 * D_k);
    /* d2[v][d]: second derivative of velocity v along d, twice */
    double lap[3];
    for (int v = 0; v < 3; ++v) {
      lap[v] = 0.;
      for (int d = 0; d < 3; ++d) {
#define GD(v_, d_, i_, j_, k_) ((d_) == 0 ? GX(v_, i_, j_, k_) : (d_) == 1 ? GY(v_, i_, j_, k_) : GZ(v_, i_, j_, k_))
        double c0 = GD(v, d, i, j, k);
        double t = (-(GD(v, d, i - 1, j, k) + c0) * inx[d][IFI(s, i, j, k)] * iA[IFI(s, i, j, k)]
                    - (GD(v, d, i, j - 1, k) + c0) * jnx[d][JFI(s, i, j, k)] * jA[JFI(s, i, j, k)]
                    - (GD(v, d, i, j, k - 1) + c0) * knx[d][KFI(s, i, j, k)] * kA[KFI(s, i, j, k)]
                    + (GD(v, d, i + 1, j, k) + c0) * inx[d][IFI(s, i + 1, j, k)] * iA[IFI(s, i + 1, j, k)]
                    + (GD(v, d, i, j + 1, k) + c0) * jnx[d][JFI(s, i, j + 1, k)] * jA[JFI(s, i, j + 1, k)]
                    + (GD(v, d, i, j, k + 1) + c0) * knx[d][KFI(s, i, j, k + 1)] * kA[KFI(s, i, j, k + 1)]) /
                   (2 * vol[ci]);
#undef GD
        lap[v] = lap[v] + t;
      }
    }
    double udd = sqrt(sq(lap[0]) + sq(lap[1]) + sq(lap[2]));
    double ss = 0.;
    for (int a = 0; a < 3; ++a) for (int b = 0; b < 3; ++b) ss = ss + sq(S[a][b]);
    double ud = sqrt(2 * ss);
    double Lvk = kappa_kkl * fabs(ud / fmx(udd, 1.e-20));
    double fp = fmn(fmx(P_k / D_k, 0.5), 1.0);
    Lvk = fmx(Lvk, tkl / fmx((tk * c11_kkl), 1.e-20));
    Lvk = fmn(Lvk, c12_kkl * kappa_kkl * dist[ci] * fp);
    double eta = density * dist[ci] * sqrt(0.3 * tk) / (20 * mu[ci]);
    double fphi = (1 + cd1_kkl * eta) / (1 + ipow(eta, 4));
    double cphi2 = zeta3_kkl;
    double cphi1 = (zeta1_kkl - zeta2_kkl * (sq(tkl / fmx(tk * Lvk, 1.e-20))));
    double P_kl = cphi1 * tkl * P_k / fmx(tk, 1.e-20);
    double D_kl = cphi2 * density * (pow(tk, 1.5));
    double d2 = sq(dist[ci]);
    double S_k = P_k - D_k - 2 * mu[ci] * tk / d2;
    double S_kl = P_kl - D_kl - 6 * mu[ci] * tkl * fphi / d2;
    S_k = S_k * vol[ci];
    S_kl = S_kl * vol[ci];
    res[5 * s->nR + ri] = res[5 * s->nR + ri] - S_k;
    res[6 * s->nR + ri] = res[6 * s->nR + ri] - S_kl;
  }
}
/* compute_gradient of src/CC.f90:130-195: the Green-Gauss gradient of a
 * cell field over cells 0..imx x 0..jmx x 0..kmx into a padded cell array
 * that is zero elsewhere */
static void cc_gradient(orc *s, double *grad, const double *var, int dir) {
  const double *iA = s->a[ORC_IA], *jA = s->a[ORC_JA], *kA = s->a[ORC_KA];
  const double *in = s->a[ORC_INX + dir], *jn = s->a[ORC_JNX + dir], *kn = s->a[ORC_KNX + dir];
  const double *vol = s->a[ORC_VOL];
  for (size_t n = 0; n < s->ncell; ++n) grad[n] = 0.0;
  for (int k = 0; k <= KMX; ++k)
    for (int j = 0; j <= JMX; ++j)
      for (int i = 0; i <= IMX; ++i) {
        double vc = var[CI(s, i, j, k)];
        grad[CI(s, i, j, k)] =
            (-(var[CI(s, i - 1, j, k)] + vc) * in[IFI(s, i, j, k)] * iA[IFI(s, i, j, k)]
             - (var[CI(s, i, j - 1, k)] + vc) * jn[JFI(s, i, j, k)] * jA[JFI(s, i, j, k)]
             - (var[CI(s, i, j, k - 1)] + vc) * kn[KFI(s, i, j, k)] * kA[KFI(s, i, j, k)]
             + (var[CI(s, i + 1, j, k)] + vc) * in[IFI(s, i + 1, j, k)] * iA[IFI(s, i + 1, j, k)]
             + (var[CI(s, i, j + 1, k)] + vc) * jn[JFI(s, i, j + 1, k)] * jA[JFI(s, i, j + 1, k)]
             + (var[CI(s, i, j, k + 1)] + vc) * kn[KFI(s, i, j, k + 1)] * kA[KFI(s, i, j, k + 1)]) /
            (2 * vol[CI(s, i, j, k)]);
      }
}

/* add_sst_source_lctm2015, src/source.f90:273-463, with the wall-normal
 * velocity-gradient sensor of src/CC.f90.  find_DCCVn takes the gradient
 * of the wall distance (CC.f90:117-120 pass `dist`, not CCVn), and the
 * cell-centre normal is that same gradient normalised at set-up
 * (CC.f90:83-98), so dvdy = |grad d|^2 / (|grad d| + 1e-12). */
static void add_sst_source_lctm2015(orc *s) {
  double *R = s->a[ORC_RES], *mut = s->a[ORC_MUT], *F1a = s->a[ORC_F1], *vol = s->a[ORC_VOL];
  const double *mu = s->a[ORC_MU], *dist = s->a[ORC_DIST];
  int s2003 = s->c.turb == ORC_TURB_SST2003;
  double gama1 = s2003 ? 5.0 / 9.0 : (beta1 / bstar) - ((sigma_w1 * (kappa_sst * kappa_sst)) / sqrt(bstar));
  double gama2 = s2003 ? 0.44 : (beta2 / bstar) - ((sigma_w2 * (kappa_sst * kappa_sst)) / sqrt(bstar));
  int limiter = s2003 ? 10 : 20;
  double cd_floor = s2003 ? 1e-10 : 1e-20;
  size_t nc = s->ncell;
  double *D = (double *)malloc(sizeof(double) * 3 * nc), *N = (double *)malloc(sizeof(double) * 3 * nc);
  if (!D || !N) { free(D); free(N); return; }
  for (int d = 0; d < 3; ++d) cc_gradient(s, N + d * nc, dist, d);
  for (size_t n = 0; n < nc; ++n) {
    double mag = sqrt(N[n] * N[n] + N[nc + n] * N[nc + n] + N[2 * nc + n] * N[2 * nc + n]);
    N[n] = N[n] / (mag + 1e-12);
    N[nc + n] = N[nc + n] / (mag + 1e-12);
    N[2 * nc + n] = N[2 * nc + n] / (mag + 1e-12);
  }
  for (int d = 0; d < 3; ++d) cc_gradient(s, D + d * nc, dist, d);
  for (int k = 1; k <= KMX - 1; ++k) for (int j = 1; j <= JMX - 1; ++j) for (int i = 1; i <= IMX - 1; ++i) {
    size_t ci = CI(s, i, j, k), ri = RI(s, i, j, k);
    double density = QP(0, i, j, k), tk = QP(5, i, j, k), tw = QP(6, i, j, k), gmi = QP(7, i, j, k);
    double vort = sqrt(sq(GY(2, i, j, k) - GZ(1, i, j, k)) + sq(GZ(0, i, j, k) - GX(2, i, j, k)) +
                       sq(GX(1, i, j, k) - GY(0, i, j, k)));
    double strain = sqrt(sq(GY(2, i, j, k) + GZ(1, i, j, k)) + sq(GZ(0, i, j, k) + GX(2, i, j, k)) +
                         sq(GX(1, i, j, k) + GY(0, i, j, k)) + 2 * sq(GX(0, i, j, k)) + 2 * sq(GY(1, i, j, k)) +
                         2 * sq(GZ(2, i, j, k)));
    double CD = 2 * density * sigma_w2 * (GX(4, i, j, k) * GX(5, i, j, k) + GY(4, i, j, k) * GY(5, i, j, k) +
                                          GZ(4, i, j, k) * GZ(5, i, j, k)) / tw;
    CD = fmx(CD, cd_floor);
    double F1 = F1a[ci];
    double gama = gama1 * F1 + gama2 * (1. - F1);
    double beta = beta1 * F1 + beta2 * (1. - F1);
    double D_k = bstar * density * tw * tk;
    double D_w = beta * density * (tw * tw);
    double divergence = GX(0, i, j, k) + GY(1, i, j, k) + GZ(2, i, j, k);
    double P_k = mut[ci] * (vort * strain) - ((2.0 / 3.0) * density * tk * divergence);
    P_k = fmn(P_k, limiter * D_k);
    double P_w = (density * gama / mut[ci]) * P_k;
    double lamda = (1. - F1) * CD;
    double dvdy = D[ci] * N[ci] + D[nc + ci] * N[nc + ci] + D[2 * nc + ci] * N[2 * nc + ci];
    double dd = dist[ci];
    double lamd = (-7.57e-3) * (dvdy * dd * dd * density / mu[ci]) + 0.0128;
    lamd = fmn(fmx(lamd, -1.0), 1.0);
    double Fpg = (lamd >= 0.0) ? fmn(1.0 + 14.68 * lamd, 1.5) : fmn(1.0 - 7.34 * lamd, 3.0);
    Fpg = fmx(Fpg, 0.0);
    double TuL = fmn(100.0 * sqrt(2.0 * tk / 3.0) / (tw * dd), 100.0);
    double Re_theta = 100.0 + 1000.0 * exp(-TuL * Fpg);
    double Rev = density * dd * dd * strain / mu[ci];
    double RT = density * tk / (mu[ci] * tw);
    double Fturb = exp(-ipow(0.5 * RT, 4));
    double Fonset1 = Rev / (2.2 * Re_theta);
    double Fonset2 = fmn(Fonset1, 2.0);
    double Fonset3 = fmx(1.0 - ipow(RT / 3.5, 3), 0.0);
    double Fonset = fmx(Fonset2 - Fonset3, 0.0);
    double P_gm = 100 * density * strain * gmi * (1.0 - gmi) * Fonset;
    double D_gm = 0.06 * density * vort * gmi * Fturb * ((50.0 * gmi) - 1.0);
    double Fon_lim = fmn(fmx((Rev / (2.2 * 1100.0)) - 1.0, 0.0), 3.0);
    double Pk_lim = 5 * fmx(gmi - 0.2, 0.0) * (1.0 - gmi) * Fon_lim * fmx(3 * mu[ci] - mut[ci], 0.0) * strain * vort;
    double S_k = gmi * P_k - fmx(gmi, 0.1) * D_k + Pk_lim;
    double S_w = P_w - D_w + lamda;
    double S_gm = P_gm - D_gm;
    S_k = S_k * vol[ci];
    S_w = S_w * vol[ci];
    S_gm = S_gm * vol[ci];
    R[5 * s->nR + ri] = R[5 * s->nR + ri] - S_k;
    R[6 * s->nR + ri] = R[6 * s->nR + ri] - S_w;
    R[7 * s->nR + ri] = R[7 * s->nR + ri] - S_gm;
  }
  free(D);
  free(N);
}

/* add_source_term_residue, src/source.f90:94-155 */
static void add_source_term_residue(orc *s) {
  int bc = s->c.transition == 1;
  if (SST(s) && LCTM(s)) { add_sst_source_lctm2015(s); return; }
  if (SST(s)) { if (bc) add_sst_bc_source(s); else add_sst_source(s); }
  if (SA(s)) { if (bc) add_sa_bc_source(s); else add_sa_source(s); }
  if (KKL(s)) add_kkl_source(s); /* no transition model (src/source.f90:147-150) */
}

/* src/source.f90:158-270; sst2003 limits the production to 10 D_k, floors
 * CD at 1e-10 and sets gama1 = 5/9, gama2 = 0.44 (source.f90:205-211) */
static void add_sst_source(orc *s) {
  double *R = s->a[ORC_RES], *mut = s->a[ORC_MUT], *F1a = s->a[ORC_F1], *vol = s->a[ORC_VOL];
  int s2003 = s->c.turb == ORC_TURB_SST2003;
  double gama1 = s2003 ? 5.0 / 9.0 : (beta1 / bstar) - ((sigma_w1 * (kappa_sst * kappa_sst)) / sqrt(bstar));
  double gama2 = s2003 ? 0.44 : (beta2 / bstar) - ((sigma_w2 * (kappa_sst * kappa_sst)) / sqrt(bstar));
  int limiter = s2003 ? 10 : 20;
  double cd_floor = s2003 ? 1e-10 : 1e-20;
  for (int k = 1; k <= KMX - 1; ++k) for (int j = 1; j <= JMX - 1; ++j) for (int i = 1; i <= IMX - 1; ++i) {
    size_t ci = CI(s, i, j, k);
    double density = QP(0, i, j, k), tk = QP(5, i, j, k), tw = QP(6, i, j, k);
    double vort = sqrt(sq(GY(2, i, j, k) - GZ(1, i, j, k)) + sq(GZ(0, i, j, k) - GX(2, i, j, k)) +
                       sq(GX(1, i, j, k) - GY(0, i, j, k)));
    double CD = 2 * density * sigma_w2 * (GX(4, i, j, k) * GX(5, i, j, k) + GY(4, i, j, k) * GY(5, i, j, k) +
                                          GZ(4, i, j, k) * GZ(5, i, j, k)) / tw;
    CD = fmx(CD, cd_floor);
    double F1 = F1a[ci];
    double gama = gama1 * F1 + gama2 * (1. - F1);
    double beta = beta1 * F1 + beta2 * (1. - F1);
    double D_k = bstar * density * tw * tk;
    double D_w = beta * density * (tw * tw);
    double divergence = GX(0, i, j, k) + GY(1, i, j, k) + GZ(2, i, j, k);
    double P_k = mut[ci] * (vort * vort) - ((2.0 / 3.0) * density * tk * divergence);
    P_k = fmn(P_k, limiter * D_k);
    double P_w = (density * gama / mut[ci]) * P_k;
    double lamda = (1. - F1) * CD;
    double S_k = P_k - D_k;
    double S_w = P_w - D_w + lamda;
    S_k = S_k * vol[ci];
    S_w = S_w * vol[ci];
    R[5 * s->nR + RI(s, i, j, k)] = R[5 * s->nR + RI(s, i, j, k)] - S_k;
    R[6 * s->nR + RI(s, i, j, k)] = R[6 * s->nR + RI(s, i, j, k)] - S_w;
  }
}

void orc_temp_refresh(orc *s) {
  double *T = s->a[ORC_TEMP];
  for (size_t n = 0; n < s->ncell; ++n) T[n] = QPL(4)[n] / (s->c.R_gas * QPL(0)[n]);
}

/* lctm2015 runs with SST and SST-2003 only: src/source.f90:124-131 stops
 * on it under SA, and k-kL's source never reads it (:147-150) */
static int transition_ok(const orc *s) { return !LCTM(s) || SST(s); }

int orc_stage(orc *s, int current_iter, int mask) {
  int st;
  if (!transition_ok(s)) return ORC_E_CONFIG;
  if ((mask & 1) && (st = populate_ghost_primitive(s, current_iter))) return st;
  if ((mask & 2) && (st = compute_face_interpolant(s))) return st;
  if ((mask & 4) && (st = compute_fluxes(s))) return st;
  if ((mask & 8) && s->c.mu_ref != 0.0) {
    if ((st = evaluate_all_gradients(s))) return st;
    if ((st = calculate_viscosity(s))) return st;
    if ((st = compute_viscous_fluxes(s))) return st;
  }
  if (mask & 16) {
    compute_residue(s);
    add_source_term_residue(s);
  }
  return ORC_OK;
}

void orc_face_flux(orc *s, const double *qL, const double *qR, double nx, double ny, double nz, double A,
                   int zero, double *out) {
  face_flux(s, qL, qR, nx, ny, nz, A, zero, out);
}

int orc_residual(orc *s, int current_iter) {
  int st;
  if (!transition_ok(s)) return ORC_E_CONFIG;
  if ((st = populate_ghost_primitive(s, current_iter))) return st;
  if ((st = compute_face_interpolant(s))) return st;
  if ((st = compute_fluxes(s))) return st;
  if (s->c.mu_ref != 0.0) {
    if ((st = evaluate_all_gradients(s))) return st;
    if ((st = calculate_viscosity(s))) return st;
    if ((st = compute_viscous_fluxes(s))) return st;
  }
  compute_residue(s);
  add_source_term_residue(s);
  return ORC_OK;
}

/* ------------------------------------------------------- time step (D) */
static double sspeed(const orc *s, const fstates *fs, size_t si) {
  double gm = s->c.gm;
  return 0.5 * (sqrt(gm * fs->L[4 * fs->n + si] / fs->L[si]) + sqrt(gm * fs->R[4 * fs->n + si] / fs->R[si]));
}

/* src/time.f90:122-246 and 366-531 */
static void local_time_step(orc *s) {
  const orc_case *c = &s->c;
  fstates xs = states_of(s, 0), ys = states_of(s, 1), zs = states_of(s, 2);
  double *dt = s->a[ORC_DT], *vol = s->a[ORC_VOL];
  const double *iA = s->a[ORC_IA], *ix = s->a[ORC_INX], *iy = s->a[ORC_INY], *iz = s->a[ORC_INZ];
  const double *jA = s->a[ORC_JA], *jx = s->a[ORC_JNX], *jy = s->a[ORC_JNY], *jz = s->a[ORC_JNZ];
  const double *kA = s->a[ORC_KA], *kx = s->a[ORC_KNX], *ky = s->a[ORC_KNY], *kz = s->a[ORC_KNZ];
  for (int k = 1; k <= KMX - 1; ++k) for (int j = 1; j <= JMX - 1; ++j) for (int i = 1; i <= IMX - 1; ++i) {
    double xsa = sspeed(s, &xs, XSI(s, i, j, k)), ysa = sspeed(s, &ys, YSI(s, i, j, k)), zsa = sspeed(s, &zs, ZSI(s, i, j, k));
    size_t f1 = IFI(s, i, j, k), f2 = JFI(s, i, j, k), f3 = KFI(s, i, j, k);
    size_t f4 = IFI(s, i + 1, j, k), f5 = JFI(s, i, j + 1, k), f6 = KFI(s, i, j, k + 1);
    double l1 = fabs((QP(1, i, j, k) * ix[f1]) + (QP(2, i, j, k) * iy[f1]) + (QP(3, i, j, k) * iz[f1])) + xsa;
    double l2 = fabs((QP(1, i, j, k) * jx[f2]) + (QP(2, i, j, k) * jy[f2]) + (QP(3, i, j, k) * jz[f2])) + ysa;
    double l3 = fabs((QP(1, i, j, k) * kx[f3]) + (QP(2, i, j, k) * ky[f3]) + (QP(3, i, j, k) * kz[f3])) + zsa;
    xsa = sspeed(s, &xs, XSI(s, i + 1, j, k)); ysa = sspeed(s, &ys, YSI(s, i, j + 1, k)); zsa = sspeed(s, &zs, ZSI(s, i, j, k + 1));
    double l4 = fabs((QP(1, i + 1, j, k) * ix[f4]) + (QP(2, i + 1, j, k) * iy[f4]) + (QP(3, i + 1, j, k) * iz[f4])) + xsa;
    double l5 = fabs((QP(1, i, j + 1, k) * jx[f5]) + (QP(2, i, j + 1, k) * jy[f5]) + (QP(3, i, j + 1, k) * jz[f5])) + ysa;
    double l6 = fabs((QP(1, i, j, k + 1) * kx[f6]) + (QP(2, i, j, k + 1) * ky[f6]) + (QP(3, i, j, k + 1) * kz[f6])) + zsa;
    double lsum = (iA[f1] * l1) + (jA[f2] * l2) + (kA[f3] * l3) + (iA[f4] * l4) + (jA[f5] * l5) + (kA[f6] * l6);
    size_t r = RI(s, i, j, k);
    dt[r] = 1. / lsum;
    dt[r] = dt[r] * vol[CI(s, i, j, k)] * c->CFL;
  }
  if (c->mu_ref == 0.0) return;
  const double *cx = s->a[ORC_CX], *cy = s->a[ORC_CY], *cz = s->a[ORC_CZ];
  for (int pass = 0; pass < 2; ++pass) {
    if (pass == 1 && c->turb == ORC_TURB_NONE) break;
    const double *m = pass == 0 ? s->a[ORC_MU] : s->a[ORC_MUT];
    double prn = pass == 0 ? c->Pr : c->tPr;
    for (int k = 1; k <= KMX - 1; ++k) for (int j = 1; j <= JMX - 1; ++j) for (int i = 1; i <= IMX - 1; ++i) {
      size_t cc = CI(s, i, j, k);
      size_t c1 = CI(s, i - 1, j, k), c2 = CI(s, i, j - 1, k), c3 = CI(s, i, j, k - 1);
      size_t c4 = CI(s, i + 1, j, k), c5 = CI(s, i, j + 1, k), c6 = CI(s, i, j, k + 1);
      size_t f1 = IFI(s, i, j, k), f2 = JFI(s, i, j, k), f3 = KFI(s, i, j, k);
      size_t f4 = IFI(s, i + 1, j, k), f5 = JFI(s, i, j + 1, k), f6 = KFI(s, i, j, k + 1);
      double rho = QPL(0)[cc];
      double l1 = m[cc] / (rho * fabs(((cx[c1] - cx[cc]) * ix[f1]) + ((cy[c1] - cy[cc]) * iy[f1]) + ((cz[c1] - cz[cc]) * iz[f1])));
      double l2 = m[cc] / (rho * fabs(((cx[c2] - cx[cc]) * jx[f2]) + ((cy[c2] - cy[cc]) * jy[f2]) + ((cz[c2] - cz[cc]) * jz[f2])));
      double l3 = m[cc] / (rho * fabs(((cx[c3] - cx[cc]) * kx[f3]) + ((cy[c3] - cy[cc]) * ky[f3]) + ((cz[c3] - cz[cc]) * kz[f3])));
      double l4 = m[c4] / (QPL(0)[c4] * fabs(((cx[cc] - cx[c4]) * ix[f4]) + ((cy[cc] - cy[c4]) * iy[f4]) + ((cz[cc] - cz[c4]) * iz[f4])));
      double l5 = m[c5] / (QPL(0)[c5] * fabs(((cx[cc] - cx[c5]) * jx[f5]) + ((cy[cc] - cy[c5]) * jy[f5]) + ((cz[cc] - cz[c5]) * jz[f5])));
      double l6 = m[c6] / (QPL(0)[c6] * fabs(((cx[cc] - cx[c6]) * kx[f6]) + ((cy[cc] - cy[c6]) * ky[f6]) + ((cz[cc] - cz[c6]) * kz[f6])));
      double lsum = (iA[f1] * l1) + (jA[f2] * l2) + (kA[f3] * l3) + (iA[f4] * l4) + (jA[f5] * l5) + (kA[f6] * l6);
      lsum = c->gm * lsum / prn;
      size_t r = RI(s, i, j, k);
      lsum = 2. / (lsum + (2. * c->CFL * vol[cc] / dt[r]));
      dt[r] = c->CFL * (lsum * vol[cc]);
    }
  }
}

int orc_time_step(orc *s) {
  double *dt = s->a[ORC_DT];
  if (s->c.time_local) { local_time_step(s); return ORC_OK; }
  if (s->c.global_time_step > 0) {
    for (size_t n = 0; n < s->nR; ++n) dt[n] = s->c.global_time_step;
    return ORC_OK;
  }
  local_time_step(s);
  double m = dt[0];
  for (size_t n = 1; n < s->nR; ++n) m = fmn(m, dt[n]) ; /* minval */
  for (size_t n = 0; n < s->nR; ++n) dt[n] = m;
  return ORC_OK;
}

/* ------------------------------------------------------ update (E, a17) */
/* src/update.f90:367-485 with the conservative branch */
static int update_with(orc *s, double TF, double SF, int TU, int use_store) {
  const orc_case *c = &s->c;
  int nv = s->nv;
  double gm = c->gm;
  double *dt = s->a[ORC_DT], *vol = s->a[ORC_VOL], *F1a = s->a[ORC_F1];
  double *Rs = s->a[ORC_RSTORE], *res = s->a[ORC_RES];
  const double *Qu = use_store ? s->a[ORC_USTORE] : s->a[ORC_QP];
  int status = ORC_OK;
  for (int k = 1; k <= KMX - 1; ++k) for (int j = 1; j <= JMX - 1; ++j) for (int i = 1; i <= IMX - 1; ++i) {
    size_t ci = CI(s, i, j, k), ri = RI(s, i, j, k);
    double u1[8] = {0}, u2[8] = {0}, R[8] = {0};
    u1[0] = Qu[ci];
    for (int l = 1; l < nv; ++l) u1[l] = Qu[(size_t)l * s->ncell + ci] * u1[0];
    double KE = 0.;
    u1[4] = (u1[4] / (gm - 1.) + 0.5 * (u1[1] * u1[1] + u1[2] * u1[2] + u1[3] * u1[3])) / u1[0] + KE;
    for (int l = 0; l < nv; ++l) R[l] = res[(size_t)l * s->nR + ri];
    if (SST(s)) {
      double beta = beta1 * F1a[ci] + (1. - F1a[ci]) * beta2;
      R[5] = R[5] / (1 + (beta * QP(6, i, j, k) * dt[ri]));
      R[6] = R[6] / (1 + (2 * beta * QP(6, i, j, k) * dt[ri]));
    }
    if (SA(s)) {
      /* src/update.f90:405-420: the point-implicit destruction, with
       * xi = U1(6) * rho / mu from the conservative nu-tilde */
      const double *mu = s->a[ORC_MU], *dist = s->a[ORC_DIST];
      double vort = sqrt(sq(GY(2, i, j, k) - GZ(1, i, j, k)) + sq(GZ(0, i, j, k) - GX(2, i, j, k)) +
                         sq(GX(1, i, j, k) - GY(0, i, j, k)));
      double kd = kappa_sa * dist[ci], kd2 = kd * kd;
      double xi = u1[5] * QP(0, i, j, k) / mu[ci];
      double fv1 = ipow(xi, 3) / (ipow(xi, 3) + CV1_3);
      double fv2 = 1.0 - xi / (1 + xi * fv1);
      double scap = vort + u1[5] * fv2 / (kd2);
      double rsa = fmn(u1[5] / (scap * kd2), 10.0);
      double g = rsa + cw2 * (ipow(rsa, 6) - rsa);
      double fw = g * pow((1.0 + CW3_6) / (ipow(g, 6) + CW3_6), 1.0 / 6.0);
      double d2 = dist[ci] * dist[ci];
      R[5] = R[5] / (1. + ((-1.0 * u1[0] * cb1 * scap) + (2.0 * u1[0] * CW1_SA * fw * u1[5] / d2)) * dt[ri]);
    }
    if (KKL(s)) {
      /* src/update.f90:399-404: the point-implicit destruction, with
       * sqrt(rho) in the k factor */
      const double *mu = s->a[ORC_MU], *dist = s->a[ORC_DIST];
      double eta = u1[0] * dist[ci] * (sqrt(0.3 * u1[5]) / (20 * mu[ci]));
      double fphi = (1 + cd1_kkl * eta) / (1 + ipow(eta, 4));
      double d2 = sq(dist[ci]);
      R[5] = R[5] / (1. + ((2.5 * ((pow(cmu_kkl, 0.75)) * sqrt(u1[0]) * (pow(u1[5], 1.5)) / fmx(u1[6], 1.e-20)) +
                            (2 * mu[ci] / d2)) * dt[ri]));
      R[6] = R[6] / (1. + (6 * mu[ci] * fphi / d2) * dt[ri]);
    }
    if (use_store) {
      for (int l = 0; l < nv; ++l) Rs[(size_t)l * s->nR + ri] = Rs[(size_t)l * s->nR + ri] + SF * R[l];
      if (TU) for (int l = 0; l < nv; ++l) R[l] = Rs[(size_t)l * s->nR + ri];
    }
    for (int l = 0; l < nv; ++l) u2[l] = u1[l] - R[l] * (TF * dt[ri] / vol[ci]);
    for (int l = 1; l < nv; ++l) u2[l] = u2[l] / u2[0];
    u2[4] = (gm - 1.) * u2[0] * (u2[4] - (0.5 * (u2[1] * u2[1] + u2[2] * u2[2] + u2[3] * u2[3])) - KE);
    int bad = (u2[0] < 0.) || (u2[4] < 0.);
    for (int l = 0; l < nv; ++l) if (isnan(u2[l])) bad = 1;
    if (bad) { status = ORC_E_NONPHYSICAL; continue; }
    for (int l = 0; l < 5; ++l) QP(l, i, j, k) = u2[l];
    if (SST(s) || KKL(s)) {
      if (u2[5] >= 0.) QP(5, i, j, k) = u2[5];
      if (u2[6] >= 0.) QP(6, i, j, k) = u2[6];
    }
    if (SA(s)) QP(5, i, j, k) = fmx(u2[5], 1.e-12);
  }
  return status;
}

/* ------------------------------------------------------- LU-SGS (f1) */
/* One face of a cell as the sweeps see it: the Flist arrays of
 * src/lusgs.f90:260-312 (area, outward normal of the cell, the averages of
 * volume, mu, mu_t and, for SST, F1 over the two cells). */
typedef struct { double A, nx, ny, nz, vol, mm, tm, F1; } lu_face;

/* Flux / SSTFlux, src/lusgs.f90:490-609 and 1027-1196: the face flux with
 * the state of neighbour ql moved by du (conservative), against the cell's
 * own state qr for the face averages and differences. */
static void lu_flux(const orc *s, int nv, const double *ql, const double *qr, const double *du, const lu_face *f,
                    double *out) {
  const orc_case *c = &s->c;
  double gm = c->gm, R_gas = c->R_gas;
  double U[7], W[7];
  const double *P = qr;
  U[0] = ql[0];
  U[1] = ql[0] * ql[1];
  U[2] = ql[0] * ql[2];
  U[3] = ql[0] * ql[3];
  U[4] = (ql[4] / (gm - 1.0)) + (0.5 * ql[0] * (ql[1] * ql[1] + ql[2] * ql[2] + ql[3] * ql[3]));
  if (nv == 7) { U[5] = ql[0] * ql[5]; U[6] = ql[0] * ql[6]; }
  if (nv == 6) U[5] = ql[0] * ql[5];
  for (int l = 0; l < nv; ++l) U[l] = U[l] + du[l];
  W[0] = U[0];
  W[1] = U[1] / U[0];
  W[2] = U[2] / U[0];
  W[3] = U[3] / U[0];
  W[4] = (gm - 1.0) * (U[4] - (0.5 * (U[1] * U[1] + U[2] * U[2] + U[3] * U[3]) / U[0]));
  if (nv == 7) {
    W[5] = U[5] / U[0];
    W[6] = U[6] / U[0];
    if (KKL(s)) {
      /* KKLFlux floors k and kL (src/lusgs.f90:1605-1606) */
      W[5] = fmx(W[5], 1e-8);
      W[6] = fmx(W[6], 1e-8);
    } else {
      /* a negative k or omega falls back to the neighbour's old value */
      W[5] = W[5] + 0.5 * (1. - copysign(1., W[5])) * (ql[5] - W[5]);
      W[6] = W[6] + 0.5 * (1. - copysign(1., W[6])) * (ql[6] - W[6]);
    }
  }
  if (nv == 6) W[5] = fmx(U[5] / U[0], 1e-8); /* SAFlux, src/lusgs.f90:2145-2146 */
  double nx = f->nx, ny = f->ny, nz = f->nz, Area = f->A, Volume = f->vol, mmu = f->mm, tmu = f->tm;
  double Vn = (W[1] * nx) + (W[2] * ny) + (W[3] * nz);
  double uface = 0.5 * (W[1] + P[1]), vface = 0.5 * (W[2] + P[2]), wface = 0.5 * (W[3] + P[3]);
  double Fx[7];
  Fx[0] = W[0] * Vn;
  Fx[1] = (W[1] * Fx[0]) + (W[4] * nx);
  Fx[2] = (W[2] * Fx[0]) + (W[4] * ny);
  Fx[3] = (W[3] * Fx[0]) + (W[4] * nz);
  double HalfRhoUsquare = 0.5 * W[0] * (W[1] * W[1] + W[2] * W[2] + W[3] * W[3]);
  double RhoHt = ((gm / (gm - 1.0)) * W[4]) + HalfRhoUsquare;
  Fx[4] = RhoHt * Vn;
  if (nv == 7) { Fx[5] = W[5] * Fx[0]; Fx[6] = W[6] * Fx[0]; }
  if (nv == 6) Fx[5] = (W[5] * Fx[0]);
  double mu = mmu + tmu;
  double T1 = W[4] / (W[0] * R_gas), T2 = P[4] / (P[0] * R_gas);
  double dTdx = (T2 - T1) * nx * Area / Volume, dTdy = (T2 - T1) * ny * Area / Volume,
         dTdz = (T2 - T1) * nz * Area / Volume;
  double dudx = (P[1] - W[1]) * nx * Area / Volume, dudy = (P[1] - W[1]) * ny * Area / Volume,
         dudz = (P[1] - W[1]) * nz * Area / Volume;
  double dvdx = (P[2] - W[2]) * nx * Area / Volume, dvdy = (P[2] - W[2]) * ny * Area / Volume,
         dvdz = (P[2] - W[2]) * nz * Area / Volume;
  double dwdx = (P[3] - W[3]) * nx * Area / Volume, dwdy = (P[3] - W[3]) * ny * Area / Volume,
         dwdz = (P[3] - W[3]) * nz * Area / Volume;
  double trace = dudx + dvdy + dwdz;
  double Tauxx = 2. * mu * (dudx - trace / 3.0), Tauyy = 2. * mu * (dvdy - trace / 3.0),
         Tauzz = 2. * mu * (dwdz - trace / 3.0);
  double Tauxy = mu * (dvdx + dudy), Tauxz = mu * (dwdx + dudz), Tauyz = mu * (dwdy + dvdz);
  double K_heat = (mmu / c->Pr + tmu / c->tPr) * gm * R_gas / (gm - 1.0);
  double Qx = K_heat * dTdx, Qy = K_heat * dTdy, Qz = K_heat * dTdz;
  Fx[1] = Fx[1] - (Tauxx * nx + Tauxy * ny + Tauxz * nz);
  Fx[2] = Fx[2] - (Tauxy * nx + Tauyy * ny + Tauyz * nz);
  Fx[3] = Fx[3] - (Tauxz * nx + Tauyz * ny + Tauzz * nz);
  Fx[4] = Fx[4] - (Tauxx * uface + Tauxy * vface + Tauxz * wface + Qx) * nx;
  Fx[4] = Fx[4] - (Tauxy * uface + Tauyy * vface + Tauyz * wface + Qy) * ny;
  Fx[4] = Fx[4] - (Tauxz * uface + Tauyz * vface + Tauzz * wface + Qz) * nz;
  if (nv == 7) {
    double F1 = f->F1;
    double dtkdx = (P[5] - W[5]) * nx * Area / Volume, dtkdy = (P[5] - W[5]) * ny * Area / Volume,
           dtkdz = (P[5] - W[5]) * nz * Area / Volume;
    double dtwdx = (P[6] - W[6]) * nx * Area / Volume, dtwdy = (P[6] - W[6]) * ny * Area / Volume,
           dtwdz = (P[6] - W[6]) * nz * Area / Volume;
    double sk = sigma_k1 * F1 + sigma_k2 * (1.0 - F1), sw = sigma_w1 * F1 + sigma_w2 * (1.0 - F1);
    if (KKL(s)) sk = sw = 1.0; /* sigma_k, sigma_phi (src/lusgs.f90:1671-1672) */
    Fx[5] = Fx[5] + (mmu + sk * tmu) * (dtkdx * nx + dtkdy * ny + dtkdz * nz);
    Fx[6] = Fx[6] + (mmu + sw * tmu) * (dtwdx * nx + dtwdy * ny + dtwdz * nz);
  }
  if (nv == 6) {
    /* nu-tilde diffusion with the face's 0.25 (rho_L + rho_R)(nu_L + nu_R),
     * added (src/lusgs.f90:2219, 2254) */
    double muCap = 0.25 * (P[0] + W[0]) * (P[5] + W[5]);
    double dtvdx = (P[5] - W[5]) * nx * Area / Volume, dtvdy = (P[5] - W[5]) * ny * Area / Volume,
           dtvdz = (P[5] - W[5]) * nz * Area / Volume;
    Fx[5] = Fx[5] + (mmu + muCap) * (dtvdx * nx + dtvdy * ny + dtvdz * nz) / sigma_sa;
  }
  for (int l = 0; l < nv; ++l) out[l] = Fx[l] * Area;
}

/* SpectralRadius, src/lusgs.f90:612-660: convective plus viscous spectral
 * radius of a face times its area; c1, c2 the two cell centres */
static double lu_radius(const orc *s, const double *ql, const double *qr, const lu_face *f, const double *c1,
                        const double *c2) {
  const orc_case *c = &s->c;
  double gm = c->gm;
  double Vn = 0.5 * (((ql[1] + qr[1]) * f->nx) + ((ql[2] + qr[2]) * f->ny) + ((ql[3] + qr[3]) * f->nz));
  Vn = fabs(Vn);
  double a = 0.5 * (sqrt(gm * ql[4] / ql[0]) + sqrt(gm * qr[4] / qr[0]));
  double rho = 0.5 * (ql[0] + qr[0]);
  double dist = sqrt((c1[0] - c2[0]) * (c1[0] - c2[0]) + (c1[1] - c2[1]) * (c1[1] - c2[1]) +
                     (c1[2] - c2[2]) * (c1[2] - c2[2]));
  double vis = gm * (f->mm / c->Pr + f->tm / c->tPr) / (rho * dist);
  return (Vn + a + vis) * f->A;
}

/* index into the sweep arrays delQstar/delQ over (0:imx,0:jmx,0:kmx) */
static inline size_t DQI(const orc *s, int i, int j, int k) {
  return (size_t)i + (size_t)(s->imx + 1) * ((size_t)j + (size_t)(s->jmx + 1) * (size_t)k);
}

/* The six faces, neighbour states and centres of cell (i,j,k), in the
 * reference's numbering 1..6 = i-, j-, k-, i+, j+, k+ (src/lusgs.f90:246-312). */
static void lu_cell(const orc *s, int i, int j, int k, double Q[7][7], double C[7][3], lu_face F[6]) {
  int nv = s->nv, turb = s->c.turb != ORC_TURB_NONE, sst = SST(s);
  const int di[7] = {0, -1, 0, 0, 1, 0, 0}, dj[7] = {0, 0, -1, 0, 0, 1, 0}, dk[7] = {0, 0, 0, -1, 0, 0, 1};
  const double *vol = s->a[ORC_VOL], *mu = s->a[ORC_MU], *mut = s->a[ORC_MUT], *F1 = s->a[ORC_F1];
  int lam_mu = s->c.mu_ref != 0.0;
  size_t c0 = CI(s, i, j, k);
  for (int n = 0; n < 7; ++n) {
    size_t cn = CI(s, i + di[n], j + dj[n], k + dk[n]);
    for (int l = 0; l < nv; ++l) Q[n][l] = s->a[ORC_QP][(size_t)l * s->ncell + cn];
    C[n][0] = s->a[ORC_CX][cn]; C[n][1] = s->a[ORC_CY][cn]; C[n][2] = s->a[ORC_CZ][cn];
  }
  for (int n = 1; n <= 6; ++n) {
    int d = (n - 1) % 3, up = n > 3;
    int fi_ = i + (d == 0 && up), fj = j + (d == 1 && up), fk = k + (d == 2 && up);
    size_t g = d == 0 ? IFI(s, fi_, fj, fk) : d == 1 ? JFI(s, fi_, fj, fk) : KFI(s, fi_, fj, fk);
    const double *A = s->a[d == 0 ? ORC_IA : d == 1 ? ORC_JA : ORC_KA];
    const double *nx = s->a[d == 0 ? ORC_INX : d == 1 ? ORC_JNX : ORC_KNX];
    const double *ny = s->a[d == 0 ? ORC_INY : d == 1 ? ORC_JNY : ORC_KNY];
    const double *nz = s->a[d == 0 ? ORC_INZ : d == 1 ? ORC_JNZ : ORC_KNZ];
    size_t cn = CI(s, i + di[n], j + dj[n], k + dk[n]);
    lu_face *f = &F[n - 1];
    f->A = A[g];
    f->nx = up ? nx[g] : -nx[g];
    f->ny = up ? ny[g] : -ny[g];
    f->nz = up ? nz[g] : -nz[g];
    f->vol = 0.5 * (vol[cn] + vol[c0]);
    /* mmu points at zeros for inviscid flow and tmu for laminar flow
     * (src/lusgs.f90:115-130) */
    f->mm = lam_mu ? 0.5 * (mu[cn] + mu[c0]) : 0.0;
    f->tm = turb ? 0.5 * (mut[cn] + mut[c0]) : 0.0;
    f->F1 = sst ? 0.5 * (F1[cn] + F1[c0]) : 0.0;
  }
}

/* The SA diagonal, src/lusgs.f90:1882-1914 (and 2034-2066 backward): the
 * convective/viscous D plus the linearised SA production and destruction,
 * one value for all six equations. */
static double sa_diagonal(const orc *s, int i, int j, int k, double D) {
  const double *mu = s->a[ORC_MU], *dist = s->a[ORC_DIST], *vol = s->a[ORC_VOL];
  size_t c0 = CI(s, i, j, k);
  double density = QP(0, i, j, k), q6 = QP(5, i, j, k);
  double Omega = sqrt(sq(GY(2, i, j, k) - GZ(1, i, j, k)) + sq(GZ(0, i, j, k) - GX(2, i, j, k)) +
                      sq(GX(1, i, j, k) - GY(0, i, j, k)));
  double dist_i = dist[c0], dist_i_2 = dist_i * dist_i, k2 = kappa_sa * kappa_sa;
  double nu = mu[c0] / density;
  double Ji = q6 / nu, Ji_2 = Ji * Ji, Ji_3 = Ji_2 * Ji;
  double fv1 = (Ji_3) / ((Ji_3) + (CV1_3));
  double fv2 = 1.0 - Ji / (1.0 + (Ji * fv1));
  double inv_k2_d2 = 1.0 / (k2 * dist_i_2);
  double Shat = Omega + q6 * fv2 * inv_k2_d2;
  Shat = fmx(Shat, 1.0e-10);
  double inv_Shat = 1.0 / Shat;
  double jc = Ji_3 + CV1_3;
  double dfv1 = 3.0 * Ji_2 * CV1_3 / (nu * (jc * jc));
  double jf = 1.0 + Ji * fv1;
  double dfv2 = -((1.0 / nu) - Ji_2 * dfv1) / (jf * jf);
  double dShat = (fv2 + q6 * dfv2) * inv_k2_d2;
  D = D - cb1 * (q6 * dShat + Shat) * vol[c0];
  double r = fmn(q6 * inv_Shat * inv_k2_d2, 10.0);
  double g = r + cw2 * ((ipow(r, 6)) - r);
  double g_6 = ipow(g, 6);
  double glim = pow((1.0 + CW3_6) / (g_6 + CW3_6), 1.0 / 6.0);
  double fw = g * glim;
  double dr = (Shat - q6 * dShat) * inv_Shat * inv_Shat * inv_k2_d2;
  double dg = dr * (1.0 + cw2 * (6.0 * (ipow(r, 5)) - 1.0));
  double dfw = dg * glim * (1.0 - g_6 / (g_6 + CW3_6));
  D = D + CW1_SA * (dfw * q6 + 2 * fw) * q6 / dist_i_2 * vol[c0];
  return D;
}

/* update_laminar_variables / update_SST_variables, src/lusgs.f90:186-488 and
 * 686-1024: forward sweep for delQstar, backward sweep for delQ, then the
 * conservative update of the interior cells. */
static int lusgs(orc *s) {
  const orc_case *c = &s->c;
  /* update_lctm2015 (src/lusgs.f90:2262-2669) is not restated */
  if (LCTM(s)) return ORC_E_CONFIG;
  int nv = s->nv, sst = SST(s), sa = SA(s), kkl = KKL(s);
  double gm = c->gm;
  size_t ndq = (size_t)(IMX + 1) * (JMX + 1) * (KMX + 1);
  double *dqs = calloc(ndq * (size_t)nv, sizeof(double));
  double *dq = calloc(ndq * (size_t)nv, sizeof(double));
  if (!dqs || !dq) { free(dqs); free(dq); return ORC_E_CONFIG; }
  const double *res = s->a[ORC_RES], *dt = s->a[ORC_DT], *vol = s->a[ORC_VOL], *F1a = s->a[ORC_F1];
  const double zero[7] = {0, 0, 0, 0, 0, 0, 0};
  double Q[7][7], C[7][3], L[6], D[7], New[3][7], Old[3][7];
  lu_face F[6];
#define DQS(l, i, j, k) dqs[(size_t)(l) * ndq + DQI(s, i, j, k)]
#define DQ(l, i, j, k) dq[(size_t)(l) * ndq + DQI(s, i, j, k)]
  for (int pass = 0; pass < 2; ++pass) {
    int ni = IMX - 1, nj = JMX - 1, nk = KMX - 1;
    /* forward: k, j, i ascending; backward: i, j, k descending */
    for (int a = 0; a < (pass ? ni : nk); ++a)
      for (int b = 0; b < nj; ++b)
        for (int e = 0; e < (pass ? nk : ni); ++e) {
          int i = pass ? ni - a : e + 1, j = pass ? nj - b : b + 1, k = pass ? nk - e : a + 1;
          lu_cell(s, i, j, k, Q, C, F);
          const int di[3] = {1, 0, 0}, dj[3] = {0, 1, 0}, dk[3] = {0, 0, 1};
          for (int d = 0; d < 3; ++d) {
            int n = pass ? d + 4 : d + 1; /* neighbour 1..6 */
            int si = pass ? 1 : -1;
            double dn[7];
            for (int l = 0; l < nv; ++l)
              dn[l] = pass ? DQ(l, i + si * di[d], j + si * dj[d], k + si * dk[d])
                           : DQS(l, i + si * di[d], j + si * dj[d], k + si * dk[d]);
            lu_flux(s, nv, Q[n], Q[0], dn, &F[n - 1], New[d]);
            lu_flux(s, nv, Q[n], Q[0], zero, &F[n - 1], Old[d]);
          }
          for (int n = 1; n <= 6; ++n) L[n - 1] = lu_radius(s, Q[n], Q[0], &F[n - 1], C[n], C[0]);
          size_t c0 = CI(s, i, j, k), r = RI(s, i, j, k);
          double sumL = L[0] + L[1] + L[2] + L[3] + L[4] + L[5];
          for (int l = 0; l < nv; ++l) D[l] = (vol[c0] / dt[r]) + 0.5 * sumL;
          if (sst) {
            double beta = F1a[c0] * beta1 + (1.0 - F1a[c0]) * beta2;
            D[5] = (D[5] + (bstar * QP(6, i, j, k)) * vol[c0]);
            D[6] = (D[6] + 2.0 * beta * QP(6, i, j, k) * vol[c0]);
          }
          if (sa) {
            double d0 = sa_diagonal(s, i, j, k, D[0]);
            for (int l = 0; l < nv; ++l) D[l] = d0;
          }
          if (kkl) {
            /* src/lusgs.f90:1359-1361 and 1468-1470 */
            const double *mu = s->a[ORC_MU], *dist = s->a[ORC_DIST];
            double d2 = dist[c0] * dist[c0];
            D[5] = D[5] + (2.5 * (pow(cmu_kkl, 0.75)) * QP(0, i, j, k) * (pow(QP(5, i, j, k), 1.5)) * vol[c0] /
                           QP(6, i, j, k));
            D[5] = D[5] + (2 * mu[c0] * vol[c0] / d2);
            D[6] = D[6] + (6 * mu[c0] * vol[c0] / d2);
          }
          for (int l = 0; l < nv; ++l) {
            double dI = New[0][l] - Old[0][l], dJ = New[1][l] - Old[1][l], dK = New[2][l] - Old[2][l];
            if (!pass) {
              double du = -res[(size_t)l * s->nR + r] -
                          0.5 * ((dI - L[0] * DQS(l, i - 1, j, k)) + (dJ - L[1] * DQS(l, i, j - 1, k)) +
                                 (dK - L[2] * DQS(l, i, j, k - 1)));
              DQS(l, i, j, k) = du / D[l];
            } else {
              DQ(l, i, j, k) = DQS(l, i, j, k) - 0.5 *
                                                     ((dI - L[3] * DQ(l, i + 1, j, k)) + (dJ - L[4] * DQ(l, i, j + 1, k)) +
                                                      (dK - L[5] * DQ(l, i, j, k + 1))) /
                                                     D[l];
            }
          }
        }
  }
  for (int k = 1; k <= KMX - 1; ++k) for (int j = 1; j <= JMX - 1; ++j) for (int i = 1; i <= IMX - 1; ++i) {
    double U[7];
    U[0] = QP(0, i, j, k);
    U[1] = QP(0, i, j, k) * QP(1, i, j, k);
    U[2] = QP(0, i, j, k) * QP(2, i, j, k);
    U[3] = QP(0, i, j, k) * QP(3, i, j, k);
    U[4] = (QP(4, i, j, k) / (gm - 1.0)) +
           (0.5 * QP(0, i, j, k) * (QP(1, i, j, k) * QP(1, i, j, k) + QP(2, i, j, k) * QP(2, i, j, k) +
                                    QP(3, i, j, k) * QP(3, i, j, k)));
    if (sst || kkl) { U[5] = QP(0, i, j, k) * QP(5, i, j, k); U[6] = QP(0, i, j, k) * QP(6, i, j, k); }
    if (sa) U[5] = QP(0, i, j, k) * QP(5, i, j, k);
    for (int l = 0; l < nv; ++l) U[l] = U[l] + DQ(l, i, j, k);
    QP(0, i, j, k) = U[0];
    QP(1, i, j, k) = U[1] / U[0];
    QP(2, i, j, k) = U[2] / U[0];
    QP(3, i, j, k) = U[3] / U[0];
    QP(4, i, j, k) = (gm - 1.0) * (U[4] - (0.5 * (U[1] * U[1] + U[2] * U[2] + U[3] * U[3]) / U[0]));
    if (sst && !KKL(s)) {
      if (U[5] > 0) QP(5, i, j, k) = U[5] / U[0];
      if (U[6] > 0) QP(6, i, j, k) = U[6] / U[0];
    }
    if (KKL(s)) { /* src/lusgs.f90:1502-1505 */
      QP(5, i, j, k) = fmx(U[5] / U[0], 1.e-8);
      QP(6, i, j, k) = fmx(U[6] / U[0], 1.e-8);
    }
    if (sa) QP(5, i, j, k) = fmx(U[5] / U[0], 1.e-8);
  }
#undef DQS
#undef DQ
  free(dqs);
  free(dq);
  return ORC_OK;
}

/* src/update.f90:129-226 (RK4, RK2, TVD-RK, first-order 'none' and the
 * LU-SGS 'implicit' branches) */
int orc_step(orc *s, int current_iter) {
  int st;
  orc_temp_refresh(s);
  int ta = s->c.time_accuracy;
  if (ta == ORC_TA_IMPLICIT) {
    if ((st = orc_residual(s, current_iter))) return st;
    if ((st = orc_time_step(s))) return st;
    return lusgs(s);
  }
  if (ta == ORC_TA_NONE) {
    if ((st = orc_residual(s, current_iter))) return st;
    if ((st = orc_time_step(s))) return st;
    return update_with(s, 1., 1., 0, 0);
  }
  size_t nq = s->ncell * (size_t)s->nv;
  if (ta == ORC_TA_TVDRK2 || ta == ORC_TA_TVDRK3) {
    /* src/update.f90:196-214: plain stage updates, then blends of the
     * primitive state with U_store over the whole padded array */
    double *q = s->a[ORC_QP], *u = s->a[ORC_USTORE];
    memcpy(u, q, nq * sizeof(double));
    if ((st = orc_residual(s, current_iter))) return st;
    if ((st = orc_time_step(s))) return st;
    if ((st = update_with(s, 1., 1., 0, 0))) return st;
    if ((st = orc_residual(s, current_iter))) return st;
    if ((st = update_with(s, 1., 1., 0, 0))) return st;
    if (ta == ORC_TA_TVDRK2) {
      for (size_t n = 0; n < nq; ++n) q[n] = 0.5 * u[n] + 0.5 * q[n];
      return ORC_OK;
    }
    for (size_t n = 0; n < nq; ++n) q[n] = 0.75 * u[n] + 0.25 * q[n];
    if ((st = orc_residual(s, current_iter))) return st;
    if ((st = update_with(s, 1., 1., 0, 0))) return st;
    for (size_t n = 0; n < nq; ++n) q[n] = (1. / 3.) * u[n] + (2. / 3.) * q[n];
    return ORC_OK;
  }
  for (size_t n = 0; n < s->nR * (size_t)s->nv; ++n) s->a[ORC_RSTORE][n] = 0.;
  memcpy(s->a[ORC_USTORE], s->a[ORC_QP], nq * sizeof(double));
  if (ta == ORC_TA_RK4) {
    const double TFs[4] = {0.5, 0.5, 1.0, 1. / 6.};
    const double SFs[4] = {1., 2., 2., 1.};
    for (int stage = 0; stage < 4; ++stage) {
      if ((st = orc_residual(s, current_iter))) return st;
      if (stage == 0 && (st = orc_time_step(s))) return st;
      if ((st = update_with(s, TFs[stage], SFs[stage], stage == 3, 1))) return st;
    }
    return ORC_OK;
  }
  if (ta == ORC_TA_RK2) {
    if ((st = orc_residual(s, current_iter))) return st;
    if ((st = orc_time_step(s))) return st;
    if ((st = update_with(s, 0.5, 1., 0, 1))) return st;
    if ((st = orc_residual(s, current_iter))) return st;
    return update_with(s, 0.5, 1., 1, 1);
  }
  return ORC_E_CONFIG;
}

/* src/resnorm.f90:128-225 */
void orc_resnorm(orc *s, double *out) {
  const orc_case *c = &s->c;
  double scale[9];
  scale[0] = 1.;
  scale[1] = c->rho_inf * c->vel_mag;
  scale[2] = scale[3] = scale[4] = c->rho_inf * c->vel_mag * c->vel_mag;
  scale[5] = (0.5 * c->rho_inf * (c->vel_mag * c->vel_mag * c->vel_mag) + ((c->gm / (c->gm - 1.)) * c->p_inf));
  /* nu-tilde for SA, k for SST (src/resnorm.f90:150-158) */
  scale[6] = c->rho_inf * c->vel_mag * (c->turb == ORC_TURB_SA ? c->tv_inf : c->tk_inf);
  scale[7] = c->rho_inf * c->vel_mag * (c->turb == ORC_TURB_KKL ? c->tkl_inf : c->tw_inf);
  /* lctm2015's intermittency: setup_scale (resnorm.f90:128-168) leaves
   * Res_scale(8) unset, so that norm has no defined value there; it is
   * scaled here like the other turbulence variables, by rho V gamma_inf */
  scale[8] = c->rho_inf * c->vel_mag * c->tgm_inf;
  double *R = s->a[ORC_RES];
  for (int l = 0; l < s->nv; ++l) {
    double sum = 0.;
    for (size_t n = 0; n < s->nR; ++n) sum += R[(size_t)l * s->nR + n] * R[(size_t)l * s->nR + n];
    out[l + 1] = sqrt(sum / (scale[l + 1] * scale[l + 1]));
  }
  double *F = s->a[ORC_F], *G = s->a[ORC_G], *H = s->a[ORC_H];
  double sF1 = 0, sFn = 0, sG1 = 0, sGn = 0, sH1 = 0, sHn = 0;
  for (int k = 1; k <= KMX - 1; ++k) for (int j = 1; j <= JMX - 1; ++j) { sF1 += F[FI(s, 1, j, k)]; sFn += F[FI(s, IMX, j, k)]; }
  for (int k = 1; k <= KMX - 1; ++k) for (int i = 1; i <= IMX - 1; ++i) { sG1 += G[GI(s, i, 1, k)]; sGn += G[GI(s, i, JMX, k)]; }
  for (int j = 1; j <= JMX - 1; ++j) for (int i = 1; i <= IMX - 1; ++i) { sH1 += H[HI(s, i, j, 1)]; sHn += H[HI(s, i, j, KMX)]; }
  double merror = (sF1 - sFn + sG1 - sGn + sH1 - sHn);
  out[0] = fabs(merror / scale[0]);
}

/* ------------------------------------------------------------ multiblock */
/* ghost cell of layer l (1..3) at tangential cell (a, b) of face f (1..6),
 * src/interface1.f90 receive loops: qp(1-l,..), qp(imx+l-1,..), ... */
static size_t ghost_cell(const orc *s, int f, int l, int a, int b) {
  switch (f) {
    case 1: return CI(s, 1 - l, a, b);
    case 2: return CI(s, IMX + l - 1, a, b);
    case 3: return CI(s, a, 1 - l, b);
    case 4: return CI(s, a, JMX + l - 1, b);
    case 5: return CI(s, a, b, 1 - l);
    default: return CI(s, a, b, KMX + l - 1);
  }
}
/* interior cell of layer l (1..3) next to face f, src/interface1.f90 pack
 * loops: qp(l,..), qp(imx-l,..), ... */
static size_t layer_cell(const orc *s, int f, int l, int a, int b) {
  switch (f) {
    case 1: return CI(s, l, a, b);
    case 2: return CI(s, IMX - l, a, b);
    case 3: return CI(s, a, l, b);
    case 4: return CI(s, a, JMX - l, b);
    case 5: return CI(s, a, b, l);
    default: return CI(s, a, b, KMX - l);
  }
}
/* cell extents (node count - 1) of the first and second tangential
 * direction of face f */
static void face_extents(const orc *s, int f, int *n1, int *n2) {
  if (f <= 2) { *n1 = JMX - 1; *n2 = KMX - 1; }
  else if (f <= 4) { *n1 = IMX - 1; *n2 = KMX - 1; }
  else { *n1 = IMX - 1; *n2 = JMX - 1; }
}
/* node range of the map -> cell loop start/end/step,
 * change_map_to_particular_range (src/mapping.f90:186-258) */
static void cell_range(int lo, int hi, int *P_lo, int *P_hi, int *dir) {
  *P_lo = lo; *P_hi = hi; *dir = 1;
  if (lo == 1) *P_lo = 1;
  if (hi == 1) { *P_hi = 1; *dir = -1; }
  if (lo > 1) { *P_lo = lo - 1; *dir = -1; }
  if (hi > 1) *P_hi = hi - 1;
}

int orc_exchange(orc **blocks, int nblocks, const orc_link *links, int nlinks) {
  /* pack every sender first (the reference's SENDRECV reads interior layers
   * only and writes ghost layers only, so the order of links is free) */
  double **buf = (double **)calloc((size_t)(nlinks > 0 ? nlinks : 1), sizeof(double *));
  if (!buf) return ORC_E_CONFIG;
  int st = ORC_OK;
  for (int e = 0; e < nlinks && st == ORC_OK; ++e) {
    const orc_link *L = &links[e];
    if (L->block < 0 || L->block >= nblocks || L->from_block < 0 || L->from_block >= nblocks ||
        L->face < 1 || L->face > 6 || L->from_face < 1 || L->from_face > 6) { st = ORC_E_CONFIG; break; }
    const orc *src = blocks[L->from_block];
    int n1, n2;
    face_extents(src, L->from_face, &n1, &n2);
    buf[e] = (double *)malloc(sizeof(double) * (size_t)src->nv * 3 * (size_t)n1 * (size_t)n2);
    if (!buf[e]) { st = ORC_E_CONFIG; break; }
    size_t count = 0;
    for (int n = 0; n < src->nv; ++n)
      for (int l = 1; l <= 3; ++l)
        for (int b = 1; b <= n2; ++b)
          for (int a = 1; a <= n1; ++a)
            buf[e][count++] = src->a[ORC_QP][(size_t)n * src->ncell + layer_cell(src, L->from_face, l, a, b)];
  }
  for (int e = 0; e < nlinks && st == ORC_OK; ++e) {
    const orc_link *L = &links[e];
    orc *dst = blocks[L->block];
    const orc *src = blocks[L->from_block];
    int sn1, sn2, dn1, dn2;
    face_extents(src, L->from_face, &sn1, &sn2);
    face_extents(dst, L->face, &dn1, &dn2);
    if ((long)sn1 * sn2 != (long)dn1 * dn2 || src->nv != dst->nv) { st = ORC_E_CONFIG; break; }
    int lo1, hi1, d1, lo2, hi2, d2;
    cell_range(L->lo1, L->hi1, &lo1, &hi1, &d1);
    cell_range(L->lo2, L->hi2, &lo2, &hi2, &d2);
    if ((hi1 - lo1) * d1 + 1 != dn1 || (hi2 - lo2) * d2 + 1 != dn2) { st = ORC_E_CONFIG; break; }
    size_t count = 0;
    for (int n = 0; n < dst->nv; ++n) {
      double *q = dst->a[ORC_QP] + (size_t)n * dst->ncell;
      for (int l = 1; l <= 3; ++l) {
        if (L->dir_switch == 0) {
          for (int b = lo2; d2 > 0 ? b <= hi2 : b >= hi2; b += d2)
            for (int a = lo1; d1 > 0 ? a <= hi1 : a >= hi1; a += d1) q[ghost_cell(dst, L->face, l, a, b)] = buf[e][count++];
        } else {
          for (int a = lo1; d1 > 0 ? a <= hi1 : a >= hi1; a += d1)
            for (int b = lo2; d2 > 0 ? b <= hi2 : b >= hi2; b += d2) q[ghost_cell(dst, L->face, l, a, b)] = buf[e][count++];
        }
      }
    }
  }
  for (int e = 0; e < nlinks; ++e) free(buf[e]);
  free(buf);
  return st;
}

int orc_group_step(orc **blocks, int nb, const orc_link *links, int nl, int current_iter) {
  int st;
  for (int b = 0; b < nb; ++b) {
    if (blocks[b]->c.time_accuracy != blocks[0]->c.time_accuracy) return ORC_E_CONFIG;
    orc_temp_refresh(blocks[b]);
  }
  int ta = blocks[0]->c.time_accuracy;
  if (ta == ORC_TA_NONE || ta == ORC_TA_IMPLICIT) {
    /* LU-SGS runs on each block by itself: its sweep arrays are zero in
     * the ghost layers, interfaces included (src/lusgs.f90:251,361) */
    if ((st = orc_exchange(blocks, nb, links, nl))) return st;
    for (int b = 0; b < nb; ++b) {
      if ((st = orc_residual(blocks[b], current_iter))) return st;
      if ((st = orc_time_step(blocks[b]))) return st;
      if ((st = ta == ORC_TA_NONE ? update_with(blocks[b], 1., 1., 0, 0) : lusgs(blocks[b]))) return st;
    }
    return ORC_OK;
  }
  for (int b = 0; b < nb; ++b) {
    orc *s = blocks[b];
    size_t nq = s->ncell * (size_t)s->nv;
    for (size_t n = 0; n < s->nR * (size_t)s->nv; ++n) s->a[ORC_RSTORE][n] = 0.;
    memcpy(s->a[ORC_USTORE], s->a[ORC_QP], nq * sizeof(double));
  }
  const double TF4[4] = {0.5, 0.5, 1.0, 1. / 6.}, SF4[4] = {1., 2., 2., 1.};
  const double TF2[2] = {0.5, 0.5}, SF2[2] = {1., 1.};
  int nst = ta == ORC_TA_RK4 ? 4 : ta == ORC_TA_RK2 ? 2 : 0;
  if (!nst) return ORC_E_CONFIG;
  for (int stage = 0; stage < nst; ++stage) {
    if ((st = orc_exchange(blocks, nb, links, nl))) return st;
    for (int b = 0; b < nb; ++b) {
      orc *s = blocks[b];
      if ((st = orc_residual(s, current_iter))) return st;
      if (stage == 0 && (st = orc_time_step(s))) return st;
      double TF = nst == 4 ? TF4[stage] : TF2[stage], SF = nst == 4 ? SF4[stage] : SF2[stage];
      if ((st = update_with(s, TF, SF, stage == nst - 1, 1))) return st;
    }
  }
  return ORC_OK;
}

/* ------------------------------------------------ threaded group step */
typedef struct {
  orc **blocks;
  int nb, t, nt, stage, nst, current_iter;
  int st;
} mt_job;

static void *mt_stage(void *arg) {
  mt_job *j = (mt_job *)arg;
  const double TF4[4] = {0.5, 0.5, 1.0, 1. / 6.}, SF4[4] = {1., 2., 2., 1.};
  const double TF2[2] = {0.5, 0.5}, SF2[2] = {1., 1.};
  j->st = ORC_OK;
  for (int b = j->t; b < j->nb; b += j->nt) {
    orc *s = j->blocks[b];
    int st;
    if (j->stage < 0) {
      /* step start: Temp refresh, R_store = 0, U_store = qp */
      orc_temp_refresh(s);
      for (size_t n = 0; n < s->nR * (size_t)s->nv; ++n) s->a[ORC_RSTORE][n] = 0.;
      memcpy(s->a[ORC_USTORE], s->a[ORC_QP], s->ncell * (size_t)s->nv * sizeof(double));
      continue;
    }
    if ((st = orc_residual(s, j->current_iter))) { j->st = st; return NULL; }
    if (j->stage == 0 && (st = orc_time_step(s))) { j->st = st; return NULL; }
    double TF = j->nst == 4 ? TF4[j->stage] : TF2[j->stage], SF = j->nst == 4 ? SF4[j->stage] : SF2[j->stage];
    if ((st = update_with(s, TF, SF, j->stage == j->nst - 1, 1))) { j->st = st; return NULL; }
  }
  return NULL;
}

static int mt_run(orc **blocks, int nb, int nt, int stage, int nst, int current_iter) {
  pthread_t th[256];
  mt_job jobs[256];
  int created[256];
  if (nt > 256) nt = 256;
  if (nt > nb) nt = nb;
  int st = ORC_OK;
  for (int t = 0; t < nt; ++t) {
    jobs[t] = (mt_job){blocks, nb, t, nt, stage, nst, current_iter, ORC_OK};
    created[t] = pthread_create(&th[t], NULL, mt_stage, &jobs[t]) == 0;
    if (!created[t]) mt_stage(&jobs[t]); /* no thread: run its share here */
  }
  for (int t = 0; t < nt; ++t) {
    if (created[t]) pthread_join(th[t], NULL);
    if (jobs[t].st && !st) st = jobs[t].st;
  }
  return st;
}

int orc_group_step_mt(orc **blocks, int nb, const orc_link *links, int nl, int current_iter, int nthreads) {
  int ta = blocks[0]->c.time_accuracy;
  for (int b = 0; b < nb; ++b)
    if (blocks[b]->c.time_accuracy != ta) return ORC_E_CONFIG;
  int nst = ta == ORC_TA_RK4 ? 4 : ta == ORC_TA_RK2 ? 2 : 0;
  if (!nst || nthreads < 1) return ORC_E_CONFIG;
  int st;
  if ((st = mt_run(blocks, nb, nthreads, -1, nst, current_iter))) return st;
  for (int stage = 0; stage < nst; ++stage) {
    if ((st = orc_exchange(blocks, nb, links, nl))) return st;
    if ((st = mt_run(blocks, nb, nthreads, stage, nst, current_iter))) return st;
  }
  return ORC_OK;
}
