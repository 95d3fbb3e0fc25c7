// blk.h -- the device-resident block: extents, pitches, array pointers and
// the scheme/flow/boundary constants, passed by value to every kernel.
//
// HBM layout (structure of arrays, i fastest, one plane per variable):
//   padded cell arrays  (-2:imx+2,-2:jmx+2,-2:kmx+2)   pitch pi=imx+5, pj=jmx+5
//   I / J / K face arrays (-2:imx+3,..), (..,-2:jmx+3,..), (..,-2:kmx+3)
//   F / G / H            (1:imx,1:jmx-1,1:kmx-1), (1:imx-1,1:jmx,..), (..,1:kmx)
//   residue, R_store, delta_t (1:imx-1,1:jmx-1,1:kmx-1)
//   LU-SGS arrays      hyperplane-major slots (i+j+k, k, j), see k_lusgs.h
//   gradients            (0:imx,0:jmx,0:kmx) x n_grad per direction
// These are the reference's own extents (src/update.f90:143-165), so the
// host<->device copies are plain memcpy and index mapping is exact.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "fest_hip.h"

namespace fest {

// the fixed values of each face: rho, u, v, w, p, tk, tw, wall
// temperature, total pressure, nu-tilde, kL, intermittency
enum { FX_RHO = 0, FX_U, FX_V, FX_W, FX_P, FX_TK, FX_TW, FX_TWALL, FX_TP, FX_TV, FX_TKL, FX_TGM, FX_N };

// residual norms per block: n_var + 1 absolute norms, then from
// NORM_SHARE the block's unassembled share (n_var + 1 values)
constexpr int NORM_SHARE = 9, NORM_SLOTS = 18;

struct Blk {
  int imx, jmx, kmx, nv, ng;
  int pi, pj, pk;
  long nc;                    // cells per padded plane
  int fpi[3], fpj[3], fpk[3]; // face array pitches per direction
  long nfc[3];                // face array sizes
  int nfi[3], nfj[3], nfk[3]; // F/G/H extents
  long nfl[3];                // F/G/H plane sizes
  long ngr;                   // gradient plane size
  long nres;                  // residue plane size
  double *qp, *temp, *mu, *mut, *f1, *dist;
  double *vol, *cx, *cy, *cz;
  double *fA[3], *fx[3], *fy[3], *fz[3];
  double *flux[3];
  double *csnd[3];            // stage-1 face sound speed (time step)
  double *res, *dt, *ust, *rst;
  double *src;                // SST source S_k, S_w (and lctm2015's S_gamma) of the interior cells (x volume)
  double *dvdy;               // lctm2015: the wall-normal sensor of each cell (k_lctm_sensor)
  // LU-SGS inputs and sweep values in hyperplane-major storage (k_lusgs.h)
  int sk_j;                   // jmx + 1: positions per k row of a hyperplane
  long sk_p, sk_n;            // positions per hyperplane; hyperplanes x sk_p
  double *sk;                 // 3 nv + 23 components x sk_n
  double *g[3];
  // FP32 mode: FP32 copies of the padded state (refreshed after the ghost
  // cells of every residual evaluation) and of the gradients, read by the
  // face sweeps; g32only: the FP64 gradients are not written
  float *q32;
  float *g32[3];
  int g32only;
  double *nodes;              // grid nodes (3 planes) when set on device
  double *red;                // reduction partials
  double *norms;              // n_var+1 absolute norms; from 8, the block's unassembled share
  double *dtmin;              // global time step
  unsigned *err;              // error bits
  // scheme
  int flux_scheme, recon, turb, ppm_flag, time_local, time_acc, accur, mu_suth;
  int xcd_remap;  // XCD-aware block order of the sweep kernels (plane_block)
  int lim[3], tlim[3];
  int pbs[3];                 // pressure-based switch per direction (MUSCL, PPM)
  int valb;                   // 1: van Albada MUSCL limiter
  int bc[6];
  int mz[3][2];               // make_*_flux_zero at the first and last face
  // flow
  double gm, R, mu_ref, T_ref, S_ref, Pr, tPr, MInf, vel_mag;
  double iPr, itPr, cpR;  // 1/Pr, 1/Pr_t, gamma R/(gamma-1): the heat-flux factors
  double rho_inf, u_inf, v_inf, w_inf, p_inf, tk_inf, tw_inf, tv_inf;
  int trans;                  // 1: the 'bc' intermittency in the turbulence source
  int lctm;                   // 1: lctm2015, the intermittency as variable 8
  double tgm_inf;             // its free-stream value
  int kkl;                    // k-kL: variables 6, 7 are k and kL
  double tkl_inf;
  double sigma_w2;            // SST's 0.856; 1 under k-kL, whose F1 stays 0
  int fp32;                   // 1: face reconstruction and flux arithmetic in FP32
  double tu_inf, Re;          // free-stream turbulence intensity (percent), rho V / mu_ref
  double c1, c2, c3;
  double fix[FX_N][6];
  double global_dt;
  double CFLv;                // CFL of the current call
};

enum : unsigned { ERR_NONPHYS = 1u, ERR_NAN_FLUX = 2u, ERR_NAN_GRAD = 4u, ERR_NAN_MU = 8u, ERR_CONFIG = 16u };

__host__ __device__ inline long ci(const Blk &b, int i, int j, int k) {
  return (long)(i + 2) + (long)b.pi * ((long)(j + 2) + (long)b.pj * (long)(k + 2));
}
// face array index of direction d
__host__ __device__ inline long fi(const Blk &b, int d, int i, int j, int k) {
  return (long)(i + 2) + (long)b.fpi[d] * ((long)(j + 2) + (long)b.fpj[d] * (long)(k + 2));
}
// F/G/H index of direction d
__host__ __device__ inline long fli(const Blk &b, int d, int i, int j, int k) {
  return (long)(i - 1) + (long)b.nfi[d] * ((long)(j - 1) + (long)b.nfj[d] * (long)(k - 1));
}
__host__ __device__ inline long ri(const Blk &b, int i, int j, int k) {
  return (long)(i - 1) + (long)(b.imx - 1) * ((long)(j - 1) + (long)(b.jmx - 1) * (long)(k - 1));
}
__host__ __device__ inline long gri(const Blk &b, int i, int j, int k) {
  return (long)i + (long)(b.imx + 1) * ((long)j + (long)(b.jmx + 1) * (long)k);
}
__host__ __device__ inline long ndi(const Blk &b, int i, int j, int k) {
  return (long)(i + 2) + (long)(b.imx + 6) * ((long)(j + 2) + (long)(b.jmx + 6) * (long)(k + 2));
}
// Launch geometry of the cell and face sweeps: 256-thread blocks cover the
// i-fastest (i, j) plane of one k as a flat index and blockIdx.y is k, so an
// i extent that is not a multiple of 64 leaves no idle lanes per row.
constexpr int FLAT_BLOCK = 256;

// Block of the (i, j) plane this work-group handles.  Work-groups go to the
// 8 XCDs round-robin by linear id.  With remap the grid carries a multiple of
// 8 slots per plane (blockIdx.x) and XCD e takes the contiguous range of
// blocks e*S/8 .. (e+1)*S/8-1 of every plane, so stencil neighbours along j
// (adjacent rows) and along k (the same range of the next plane) are read
// through one L2.  Blocks past the plane's end find no cells.
// The work-group's place in its launch: the hardware block index, or an
// entry of a group's tile work list (nx = 0 there: no XCD reordering).
struct Wg {
  int x, y, z, nx;
};
#define FEST_WG_HW (::fest::Wg{(int)blockIdx.x, (int)blockIdx.y, (int)blockIdx.z, (int)gridDim.x})

__device__ __forceinline__ int plane_block(const Wg &wg, int remap) {
  const int s = wg.x;
  if (!remap || wg.nx == 0) return s;
  const int S = wg.nx;
  return (s & 7) * (S >> 3) + (s >> 3);
}

__device__ __forceinline__ bool flat_ij(int ni, int nj, int &i, int &j, int blk) {
  int t = blk * FLAT_BLOCK + (int)threadIdx.x;
  j = t / ni;
  i = t - j * ni;
  return j < nj;
}

// Tiled variant: each 256-thread block is a TI (i) x 256/TI (j) tile --
// 64 x 4 by default, 32 x 8 for narrow blocks of a batched group -- tiles
// numbered i-fastest in blockIdx.x, blockIdx.y = k.  Sweeps whose stencil
// runs along j or k keep whole rows of one i-range in one block.
constexpr int TILE_I = 64, TILE_J = 4;
template <int TI = TILE_I>
__device__ __forceinline__ bool tile_ij(int ni, int nj, int &i, int &j, int blk) {
  constexpr int TJ = 256 / TI;
  const int ntx = (ni + TI - 1) / TI;
  const int tx = blk % ntx, ty = blk / ntx;
  i = tx * TI + ((int)threadIdx.x & (TI - 1));
  j = ty * TJ + ((int)threadIdx.x / TI);
  return i < ni && j < nj;
}

// stride of the padded cell arrays along direction d
__host__ __device__ inline long cstride(const Blk &b, int d) {
  return d == 0 ? 1L : d == 1 ? (long)b.pi : (long)b.pi * b.pj;
}

// amdflang's lowering of MAX/MIN (NaN in the first operand yields the
// second); kept so that limiter branches agree with the reference build
__device__ __forceinline__ double fmx(double a, double b) { return a > b ? a : b; }
__device__ __forceinline__ double fmn(double a, double b) { return a < b ? a : b; }
__device__ __forceinline__ double fsgn(double a, double b) { return copysign(fabs(a), b); }
__device__ __forceinline__ double sqd(double x) { return x * x; }
// FP32 forms for the single-precision face arithmetic
__device__ __forceinline__ float fmx(float a, float b) { return a > b ? a : b; }
__device__ __forceinline__ float fmn(float a, float b) { return a < b ? a : b; }
__device__ __forceinline__ float fsgn(float a, float b) { return __builtin_copysignf(__builtin_fabsf(a), b); }
__device__ __forceinline__ float sqd(float x) { return x * x; }

// Spalart-Allmaras constants (src/global/global_sa.f90); cw1 and cv1^3 as
// the reference's parameter expressions evaluate in double
struct Sa {
  static constexpr double cb1 = 0.1355, cb2 = 0.6220, cw2 = 0.3, cv1 = 7.1, sigma = 2. / 3., kappa = 0.41;
  static constexpr double cw1 = (cb1 / (kappa * kappa)) + ((1 + cb2) / sigma);
  static constexpr double cv1_3 = (cv1 * cv1) * cv1;
  static constexpr double cw3_6 = 64.0;
};

}  // namespace fest
