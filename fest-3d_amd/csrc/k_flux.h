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
//
// The face arithmetic is written once for a type T: double (the reference's
// precision) or float (the single-precision flux mode).  The stencil, the
// face metrics and the gradients are stored in FP64 and converted on load;
// differences of cell centres, of velocities and of temperatures between
// the two cells of a face are formed in FP64 before the conversion, since
// the near-wall cells of a stretched grid differ in the fifth digit.  The
// flux is stored in FP64, and the residue, the update and the norms stay
// FP64.
#pragma once
// the residual kernels let multiply-adds fuse into FMA: results differ from
// the unfused reference order by rounding only, far inside the 1e-12 parity
// bound, and the FP64 instruction count, which bounds these kernels, drops
#pragma clang fp contract(fast)
#include <type_traits>

#include "blk.h"

namespace fest {

// FP64: the IEEE operations.  FP32 (the mixed-precision mode): the
// hardware reciprocal and square root (v_rcp_f32, v_sqrt_f32, 1 ulp), whose
// error is far below the FP32 face arithmetic's own.
__device__ __forceinline__ double tsqrt(double x) { return sqrt(x); }
__device__ __forceinline__ float tsqrt(float x) { return __builtin_amdgcn_sqrtf(x); }
__device__ __forceinline__ double tdiv(double a, double b) { return a / b; }
__device__ __forceinline__ float tdiv(float a, float b) { return a * __builtin_amdgcn_rcpf(b); }
__device__ __forceinline__ double tabs(double x) { return fabs(x); }
__device__ __forceinline__ float tabs(float x) { return __builtin_fabsf(x); }

template <typename T>
struct FaceMath {
  // ---- reconstruction of one cell from its neighbours um (below) and up
  // (above); upv is the cell's upper-face value, lo its lower-face value
  // valb: van Albada (r^2+r)/(r^2+1), the limiter muscl.f90:180,186 leaves
  // commented out, in place of the Koren-alpha form
  static __device__ __forceinline__ void muscl_cell(T um, T u0, T up, int sw, T &upv, T &lo, bool valb = false) {
    const T alpha = T(2.) / T(3.), phi = T(1.0), kappa = T(1.) / T(3.), eps = T(1e-14);
    T fd = up - u0;
    T bd = u0 - um;
    T r = tdiv(fd, bd + fsgn(eps, bd));
    T psi1 = valb ? tdiv((r * r) + r, (r * r) + T(1.0)) : fmx(T(0.), fmn(fmn(2 * r, alpha * (r - T(1.0)) + T(1.0)), T(2.)));
    r = tdiv(bd, fd + fsgn(eps, fd));
    T psi2 = valb ? tdiv((r * r) + r, (r * r) + T(1.0)) : fmx(T(0.), fmn(fmn(2 * r, alpha * (r - T(1.0)) + T(1.0)), T(2.)));
    psi1 = (1 - (1 - psi1) * sw);
    psi2 = (1 - (1 - psi2) * sw);
    upv = u0 + T(0.25) * phi * (((T(1.) - kappa) * psi1 * bd) + ((T(1.) + kappa) * psi2 * fd));
    lo = u0 - T(0.25) * phi * (((T(1.) + kappa) * psi1 * bd) + ((T(1.) - kappa) * psi2 * fd));
  }

  // boundary-cell MUSCL of boundary_state_reconstruction.f90:100-117 (no eps)
  static __device__ __forceinline__ void bmuscl_cell(T um, T u0, T up, int sw, T &upv, T &lo) {
    const T phi = T(1.0), kappa = T(1.) / T(3.);
    T fd = up - u0;
    T bd = u0 - um;
    T r = tdiv(fd, bd);
    T psi1 = fmx(T(0.), fmn(fmn(2 * r, (2 + r) / T(3.)), T(2.)));
    psi1 = (1 - (1 - psi1) * sw);
    r = tdiv(bd, fd);
    T psi2 = fmx(T(0.), fmn(fmn(2 * r, (2 + r) / T(3.)), T(2.)));
    psi2 = (1 - (1 - psi2) * sw);
    upv = u0 + T(0.25) * phi * (((T(1.) - kappa) * psi1 * bd) + ((T(1.) + kappa) * psi2 * fd));
    lo = u0 - T(0.25) * phi * (((T(1.) + kappa) * psi1 * bd) + ((T(1.) - kappa) * psi2 * fd));
  }

  // WENO5 upper- and lower-face states of a cell (weno.f90:59-87), Jiang-Shu
  // smoothness indicators and linear weights (0.1, 0.6, 0.3), eps = 1e-6.
  // In FP64 the normalised weights w_i/sum(w) with w_i = g_i/D_i,
  // D_i = (eps+B_i)^2, are formed as g_i E_i / sum(g_j E_j) with E_i the
  // product of the other two D's, and the 1/6 of the candidate polynomials
  // is folded into the same single division: the reference's quotient up to
  // rounding.  In FP32 that product overflows at strong pressure jumps, so
  // the weights are the quotients g_i/D_i themselves.
  static __device__ __forceinline__ T weno_combine(T D1, T D2, T D3, T Q1, T Q2, T Q3) {
    const T g1 = T(1.0) / T(10.0), g2 = T(6.0) / T(10.0), g3 = T(3.0) / T(10.0);
    if constexpr (!std::is_same<T, float>::value) {
      T e1 = g1 * (D2 * D3), e2 = g2 * (D1 * D3), e3 = g3 * (D1 * D2);
      return tdiv(e1 * Q1 + e2 * Q2 + e3 * Q3, T(6.0) * (e1 + e2 + e3));
    } else {
      T w1 = tdiv(g1, D1), w2 = tdiv(g2, D2), w3 = tdiv(g3, D3);
      return tdiv(w1 * Q1 + w2 * Q2 + w3 * Q3, T(6.0) * (w1 + w2 + w3));
    }
  }
  static __device__ __forceinline__ void weno_betas(T Um2, T Um1, T U0, T Up1, T Up2, T &D1, T &D2, T &D3) {
    const T eps = T(1e-6);
    T B1 = (T(13.0) / T(12.0)) * sqd(Um2 - T(2.0) * Um1 + U0) + (T(1.0) / T(4.0)) * sqd(Um2 - T(4.0) * Um1 + T(3.0) * U0);
    T B2 = (T(13.0) / T(12.0)) * sqd(Um1 - T(2.0) * U0 + Up1) + (T(1.0) / T(4.0)) * sqd(Um1 - Up1);
    T B3 = (T(13.0) / T(12.0)) * sqd(U0 - T(2.0) * Up1 + Up2) + (T(1.0) / T(4.0)) * sqd(T(3.0) * U0 - T(4.0) * Up1 + Up2);
    D1 = sqd(eps + B1); D2 = sqd(eps + B2); D3 = sqd(eps + B3);
  }
  static __device__ __forceinline__ T weno_up(T Um2, T Um1, T U0, T Up1, T Up2) {
    T D1, D2, D3;
    weno_betas(Um2, Um1, U0, Up1, Up2, D1, D2, D3);
    return weno_combine(D1, D2, D3, T(2.0) * Um2 - T(7.0) * Um1 + T(11.0) * U0,
                        T(-1.0) * Um1 + T(5.0) * U0 + T(2.0) * Up1, T(2.0) * U0 + T(5.0) * Up1 - T(1.0) * Up2);
  }
  static __device__ __forceinline__ T weno_lo(T Um2, T Um1, T U0, T Up1, T Up2) {
    T D1, D2, D3;
    weno_betas(Um2, Um1, U0, Up1, Up2, D1, D2, D3);
    return weno_combine(D3, D2, D1, T(2.0) * Up2 - T(7.0) * Up1 + T(11.0) * U0,
                        T(-1.0) * Up1 + T(5.0) * U0 + T(2.0) * Um1, T(2.0) * U0 + T(5.0) * Um1 - T(1.0) * Um2);
  }

  // WENO for non-uniform grids (weno_NM.f90:59-121): a21, a10, a01, a12
  // are the volume blends V(m)/(V(m-1)+V(m)) of the cell's stencil, m = -1..2
  // relative to the cell.  Polynomials and indicators as the reference, the
  // weights combined with one division (weno_combine; 2Q/6 = P/3 exactly).
  static __device__ __forceinline__ T weno_nm(T Um2, T Um1, T U0, T Up1, T Up2, T a21, T a10, T a01, T a12,
                                              bool upper) {
    const T eps = T(1e-6);
    T U01 = (T(1.0) - a01) * U0 + a01 * Up1;
    T U12 = (T(1.0) - a12) * Up1 + a12 * Up2;
    T U10 = (T(1.0) - a10) * Um1 + a10 * U0;
    T U21 = (T(1.0) - a21) * Um2 + a21 * Um1;
    T Q1, Q2, Q3, B1, B2, B3;
    if (upper) {
      T U00 = Um1 + (T(1.0) - a21) * (Um1 - Um2);
      Q1 = T(6.0) * U0 - T(1.0) * U10 - T(2.0) * U00;
      Q2 = T(-1.0) * U10 + T(2.0) * U0 + T(2.0) * U01;
      Q3 = T(2.0) * U01 + T(2.0) * Up1 - T(1.0) * U12;
      B1 = (T(13.0) / T(12.0)) * sqd(T(2) * U10 - T(2.0) * U00) + (T(1.0) / T(4.0)) * sqd(T(4) * U0 - T(2.0) * U10 - T(2.0) * U00);
      B2 = (T(13.0) / T(12.0)) * sqd(T(2) * U10 - T(4.0) * U0 + T(2) * U01) + (T(1.0) / T(4.0)) * sqd(T(-2) * U10 + T(2.0) * U01);
      B3 = (T(13.0) / T(12.0)) * sqd(T(2) * U01 - T(4.0) * Up1 + T(2) * U12) + (T(1.0) / T(4.0)) * sqd(T(-6) * U01 + T(8.0) * Up1 - T(2.0) * U12);
    } else {
      T U11 = Up1 + a12 * (Up1 - Up2);
      Q1 = T(6.0) * U0 - T(1.0) * U01 - T(2.0) * U11;
      Q2 = T(-1.0) * U01 + T(2.0) * U0 + T(2.0) * U10;
      Q3 = T(2.0) * U10 + T(2.0) * Um1 - T(1.0) * U21;
      B1 = (T(13.0) / T(12.0)) * sqd(T(2) * U01 - T(2.0) * U11) + (T(1.0) / T(4.0)) * sqd(T(4) * U0 - T(2.0) * U01 - T(2.0) * U11);
      B2 = (T(13.0) / T(12.0)) * sqd(T(2) * U01 - T(4.0) * U0 + T(2) * U10) + (T(1.0) / T(4.0)) * sqd(T(-2) * U01 + T(2.0) * U10);
      B3 = (T(13.0) / T(12.0)) * sqd(T(2) * U10 - T(4.0) * Um1 + T(2) * U21) + (T(1.0) / T(4.0)) * sqd(T(-6) * U10 + T(8.0) * Um1 - T(2.0) * U21);
    }
    return weno_combine(sqd(eps + B1), sqd(eps + B2), sqd(eps + B3), T(2) * Q1, T(2) * Q2, T(2) * Q3);
  }

  // PPM face estimate at the face between cells m-1 and m
  static __device__ __forceinline__ T ppm_face(T qm2, T qm1, T qm, T qp1) {
    return (T(7.) * (qm + qm1) - (qp1 + qm2)) / T(12.);
  }
  // remove_extrema on one cell (ppm.f90:86-101): Lc is the cell's upper-face
  // value, Rc its lower-face value
  static __device__ __forceinline__ void ppm_extrema(T qc, T &Lc, T &Rc) {
    if ((Lc - qc) * (qc - Rc) <= 0) {
      Lc = qc; Rc = qc;
    } else {
      T dqrl = Lc - Rc;
      T dq6 = T(6.) * (qc - T(0.5) * (Lc + Rc));
      if (dqrl * dq6 > dqrl * dqrl) Rc = T(3.) * qc - T(2.) * Lc;
      else if (-dqrl * dqrl > dqrl * dq6) Lc = T(3.) * qc - T(2.) * Rc;
    }
  }

  // left (L) and right (R) states of face n (u points at cell n, stride s);
  // RECON is the reconstruction (0 none, 1 MUSCL, 2 PPM, 3 WENO5).  w(m)
  // is cell n-3+m: the reconstructions read cells n-1..n (first order),
  // n-2..n+1 (MUSCL) or n-3..n+2 (PPM, WENO5), all inside the padded array
  // for faces 1..N.
  // per-face values shared by the variables: the pressure-based switch
  // factors of the two cells (muscl.f90:37-112, ppm.f90:109-186) and the
  // WENO-NM volume blends A(m) = V(m)/(V(m-1)+V(m)), m = n-2..n+2
  struct Aux {
    bool pb;
    T pdL, pdR;
    T A[5];
  };
  // pdif of the cells n-1 and n from the pressure stencil p(m) = cell
  // n-3+m; the ghost cells 0 and N take the first and last interior
  // cell's value
  template <typename Q>
  static __device__ __forceinline__ void pb_factors(const Blk &b, int n, int N, const Q *p, long s, Aux &ax) {
    auto w = [p, s](int m) { return T(p[(long)(m - 3) * s]); };
    const T pin = T(b.p_inf);
    T pd2 = (n == 1) ? tabs(w(4) - w(2)) : tabs(w(3) - w(1));
    ax.pdL = T(1) - tdiv(pd2, pd2 + pin);
    pd2 = (n == N) ? tabs(w(3) - w(1)) : tabs(w(4) - w(2));
    ax.pdR = T(1) - tdiv(pd2, pd2 + pin);
  }
  static __device__ __forceinline__ void nm_blends(const double *vol, long s, Aux &ax) {
#pragma unroll
    for (int m = 0; m < 5; ++m) {
      const T v1 = T(vol[(long)(m - 2) * s]), v0 = T(vol[(long)(m - 3) * s]);
      ax.A[m] = tdiv(v1, v0 + v1);
    }
  }

  template <int RECON, typename Q>
  static __device__ __forceinline__ void face_states(const Blk &b, int d, int n, int N, int l, const Q *u,
                                                     long s, const Aux &ax, T &L, T &R) {
    constexpr int recon = RECON;
    // cell n-3+m, read where it is used
    auto w = [u, s](int m) { return T(u[(long)(m - 3) * s]); };
    const int sw = (l >= 5) ? b.tlim[d] : b.lim[d];
    const int minbc = b.bc[2 * d], maxbc = b.bc[2 * d + 1];
    const bool on = recon != 0;
    const bool bmin = on && minbc < 0 && minbc != -10;
    const bool bmax = on && maxbc < 0 && maxbc != -10;
    const bool gmin = (minbc == -8 || minbc == -9), gmax = (maxbc == -8 || maxbc == -9);
    const bool pf = b.ppm_flag;
    T dummy;
    // left state: upper face of cell n-1
    if (bmin && n == 1) {
      L = gmin ? w(2) : T(0.5) * (w(2) + w(3));
    } else if (bmax && n == N && gmax) {
      L = w(3);
    } else if ((bmin && pf && n == 2) || (bmax && pf && n == N)) {
      bmuscl_cell(w(1), w(2), w(3), sw, L, dummy);
    } else {
      if constexpr (recon == 0) {
        L = w(2);
      } else if constexpr (recon == 1) {
        muscl_cell(w(1), w(2), w(3), sw, L, dummy, b.valb);
      } else if constexpr (recon == 2) {
        T Lc = ppm_face(w(1), w(2), w(3), w(4)), Rc = ppm_face(w(0), w(1), w(2), w(3));
        if (b.lim[d] == 1) ppm_extrema(w(2), Lc, Rc);
        L = Lc;
      } else if constexpr (recon == 3) {
        L = weno_up(w(0), w(1), w(2), w(3), w(4));
      } else {
        L = weno_nm(w(0), w(1), w(2), w(3), w(4), ax.A[0], ax.A[1], ax.A[2], ax.A[3], true);
      }
      if constexpr (recon == 1 || recon == 2) {
        if (ax.pb) L = w(2) + (ax.pdL * (L - w(2)));
      }
    }
    // right state: lower face of cell n
    if (bmax && n == N) {
      R = gmax ? w(3) : T(0.5) * (w(2) + w(3));
    } else if (bmin && n == 1 && gmin) {
      R = w(2);
    } else if ((bmin && pf && n == 1) || (bmax && pf && n == N - 1)) {
      bmuscl_cell(w(2), w(3), w(4), sw, dummy, R);
    } else {
      if constexpr (recon == 0) {
        R = w(3);
      } else if constexpr (recon == 1) {
        muscl_cell(w(2), w(3), w(4), sw, dummy, R, b.valb);
      } else if constexpr (recon == 2) {
        T Lc = ppm_face(w(2), w(3), w(4), w(5)), Rc = ppm_face(w(1), w(2), w(3), w(4));
        if (b.lim[d] == 1) ppm_extrema(w(3), Lc, Rc);
        R = Rc;
      } else if constexpr (recon == 3) {
        R = weno_lo(w(1), w(2), w(3), w(4), w(5));
      } else {
        R = weno_nm(w(1), w(2), w(3), w(4), w(5), ax.A[1], ax.A[2], ax.A[3], ax.A[4], false);
      }
      if constexpr (recon == 1 || recon == 2) {
        if (ax.pb) R = w(3) - (ax.pdR * (w(3) - R));
      }
    }
  }

  // ---- inviscid flux of one face (src/face/flux/convective/*.f90)
  static __device__ __forceinline__ T alpha_sub(T M) { return tabs(M) < T(1.0) ? T(1.0) : T(0.0); }

  template <int NV>
  static __device__ __forceinline__ void inviscid_flux(const Blk &b, const T (&qL)[NV], const T (&qR)[NV], T nx, T ny,
                                                       T nz, T A, T zero, T (&out)[NV]) {
    const T gm = T(b.gm);
    // gamma/(gamma-1): the reference forms (gm/(gm-1))*p/rho, i.e. the
    // product first, then the quotient by rho
    const T gmr = gm / (gm - T(1.));
    constexpr int nv = NV;
    T rL = qL[0], uL = qL[1], vL = qL[2], wL = qL[3], pL = qL[4];
    T rR = qR[0], uR = qR[1], vR = qR[2], wR = qR[3], pR = qR[4];
    const int sch = b.flux_scheme;
    if (sch == 1 || sch == 2 || sch == 3) {
      T HL = (T(0.5) * (uL * uL + vL * vL + wL * wL)) + tdiv(gmr * pL, rL);
      T HR = (T(0.5) * (uR * uR + vR * vR + wR * wR)) + tdiv(gmr * pR, rR);
      T VnL = uL * nx + vL * ny + wL * nz;
      T VnR = uR * nx + vR * ny + wR * nz;
      T mass, pbar;
      if (sch == 1) {  // AUSM+, ausmP.f90:105-158
        T cs = tsqrt(tdiv(T(2.0) * (gm - T(1.0)) * (T(0.5) * (HL + HR)), gm + T(1.0)));
        T cL = tdiv(cs * cs, fmx(cs, tabs(VnL)));
        T cR = tdiv(cs * cs, fmx(cs, tabs(VnR)));
        T C = fmn(cL, cR);
        const T iC = tdiv(T(1.0), C);  // one reciprocal for the quotients by C
        T ML = VnL * iC, MR = VnR * iC;
        T aL = alpha_sub(ML), aR = alpha_sub(MR);
        T FmL = (T(0.5) * (T(1.0) + fsgn(T(1.0), ML)) * (T(1.0) - aL) * ML) + aL * T(0.25) * sqd(T(1.0) + ML);
        T betaL = (T(0.5) * (T(1.0) + fsgn(T(1.0), ML)) * (T(1.0) - aL)) + aL * T(0.25) * sqd(T(1.0) + ML) * (T(2.0) - ML);
        T FmR = (T(0.5) * (T(1.0) - fsgn(T(1.0), MR)) * (T(1.0) - aR) * MR) - aR * T(0.25) * sqd(T(1.0) - MR);
        T betaR = (T(0.5) * (T(1.0) - fsgn(T(1.0), MR)) * (T(1.0) - aR)) + aR * T(0.25) * sqd(T(1.0) - MR) * (T(2.0) + MR);
        FmL = FmL + aL * T(0.1250) * sqd(ML * ML - T(1.0));
        betaL = betaL + aL * T(0.1875) * sqd(ML * ML - T(1.0)) * ML;
        FmR = FmR - aR * T(0.1250) * sqd(MR * MR - T(1.0));
        betaR = betaR - aR * T(0.1875) * sqd(MR * MR - T(1.0)) * MR;
        T Mface = FmL + FmR;
        pbar = betaL * pL + betaR * pR;
        mass = (Mface > T(0.0)) ? Mface * C * rL : Mface * C * rR;
      } else if (sch == 2) {  // AUSM+-up, ausmUP.f90:113-189 (uses the freestream Mach)
        const T Kp = T(0.25), Ku = T(0.75), sigma = T(1.0);
        const T MInf = T(b.MInf);
        T cs = tsqrt(tdiv(T(2.0) * (gm - T(1.0)) * (T(0.5) * (HL + HR)), gm + T(1.0)));
        T cL = tdiv(cs * cs, fmx(cs, VnL));
        T cR = tdiv(cs * cs, fmx(cs, -VnR));
        T C = fmn(cL, cR);
        const T iC = tdiv(T(1.0), C);  // one reciprocal for the quotients by C
        T ML = VnL * iC, MR = VnR * iC;
        T Mb = tsqrt(tdiv(T(0.5) * ((VnL * VnL) + (VnR * VnR)), C * C));
        T Mo = tsqrt(fmn(T(1.0), fmx(Mb * Mb, MInf * MInf)));
        T fna = Mo * (T(2.0) - Mo);
        T alfa = T(3.0) * (T(-4.0) + (T(5.0) * fna * fna)) / T(16.0);
        T aL = alpha_sub(ML), aR = alpha_sub(MR);
        T FmL = (T(0.5) * (T(1.0) + fsgn(T(1.0), ML)) * (T(1.0) - aL) * ML) + aL * T(0.25) * sqd(T(1.0) + ML);
        T betaL = (T(0.5) * (T(1.0) + fsgn(T(1.0), ML)) * (T(1.0) - aL)) + aL * T(0.25) * sqd(T(1.0) + ML) * (T(2.0) - ML);
        T FmR = (T(0.5) * (T(1.0) - fsgn(T(1.0), MR)) * (T(1.0) - aR) * MR) - aR * T(0.25) * sqd(T(1.0) - MR);
        T betaR = (T(0.5) * (T(1.0) - fsgn(T(1.0), MR)) * (T(1.0) - aR)) + aR * T(0.25) * sqd(T(1.0) - MR) * (T(2.0) + MR);
        FmL = FmL + aL * T(0.125) * sqd(ML * ML - T(1.0));
        betaL = betaL + aL * alfa * sqd(ML * ML - T(1.0)) * ML;
        FmR = FmR - aR * T(0.125) * sqd(MR * MR - T(1.0));
        betaR = betaR - aR * alfa * sqd(MR * MR - T(1.0)) * MR;
        T Pu = -Ku * betaL * betaR * (rL + rR) * fna * C * (VnR - VnL);
        T Mp = tdiv(T(-2.0) * Kp * fmx(T(1.0) - (sigma * Mb * Mb), T(0.0)) * (pR - pL), fna * (rL + rR) * C * C);
        T Mface = FmL + FmR + Mp;
        pbar = betaL * pL + betaR * pR + Pu;
        mass = (Mface > T(0.0)) ? Mface * C * rL : Mface * C * rR;
      } else {  // SLAU, slau.f90:108-161
        T cL = tsqrt(tdiv(gm * pL, rL)), cR = tsqrt(tdiv(gm * pR, rR));
        T C = T(0.5) * (cL + cR);
        T delp = pR - pL;
        const T iC = tdiv(T(1.0), C);  // one reciprocal for the quotients by C
        T ML = VnL * iC, MR = VnR * iC;
        T aL = alpha_sub(ML), aR = alpha_sub(MR);
        T betaL = (T(1.0) - aL) * T(0.5) * (T(1.0) + fsgn(T(1.0), ML)) + (aL) * T(0.25) * (T(2.0) - ML) * sqd(ML + T(1.0));
        T betaR = (T(1.0) - aR) * T(0.5) * (T(1.0) - fsgn(T(1.0), MR)) + (aR) * T(0.25) * (T(2.0) + MR) * sqd(MR - T(1.0));
        T vt = tsqrt(T(0.5) * ((uL * uL) + (vL * vL) + (wL * wL) + (uR * uR) + (vR * vR) + (wR * wR)));
        T Mcap = fmn(T(1.0), vt * iC);
        T Xi = sqd(T(1.0) - Mcap);
        T Vnabs = tdiv(rL * tabs(VnL) + rR * tabs(VnR), rL + rR);
        T fnG = T(-1.0) * fmx(fmn(ML, T(0.0)), T(-1.0)) * fmn(fmx(MR, T(0.0)), T(1.0));
        pbar = T(0.5) * ((pL + pR) + (betaL - betaR) * (pL - pR) + (T(1.0) - Xi) * (betaL + betaR - T(1.0)) * (pL + pR));
        T VnabsL = (T(1.0) - fnG) * Vnabs + fnG * tabs(VnL);
        T VnabsR = (T(1.0) - fnG) * Vnabs + fnG * tabs(VnR);
        mass = T(0.5) * ((rL * (VnL + VnabsL) + rR * (VnR - VnabsR)) - (Xi * delp * iC));
      }
      mass = mass * zero;
      T Fp0 = T(0.5) * (mass + tabs(mass)), Fm0 = T(0.5) * (mass - tabs(mass));
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
    T ssa = T(0.5) * (tsqrt(tdiv(gm * pL, rL)) + tsqrt(tdiv(gm * pR, rR)));
    T MpL = tdiv(uL * nx + vL * ny + wL * nz, ssa);
    T ap = T(0.5) * (T(1.0) + fsgn(T(1.0), MpL));
    T bl = tabs(MpL) < T(1.0) ? T(-1.0) : T(0.0);
    T Mplus = T(0.25) * sqd(T(1.) + MpL);
    T Dplus = T(0.25) * sqd(T(1.) + MpL) * (T(2.) - MpL);
    T cplus = (ap * (T(1.0) + bl) * MpL) - bl * Mplus;
    T sDp = (ap * (T(1.) + bl)) - (bl * Dplus);
    T MpR = tdiv(uR * nx + vR * ny + wR * nz, ssa);
    T am = T(0.5) * (T(1.0) - fsgn(T(1.0), MpR));
    T br = tabs(MpR) < T(1.0) ? T(-1.0) : T(0.0);
    T Mminus = T(-0.25) * sqd(T(1.) - MpR);
    T Dminus = T(0.25) * sqd(T(1.) - MpR) * (T(2.) + MpR);
    T cminus = (am * (T(1.0) + br) * MpR) - br * Mminus;
    T sDm = (am * (T(1.) + br)) - (br * Dminus);
    if (sch == 0) {
      T tc = cplus + cminus;
      cplus = fmx(T(0.), tc);
      cminus = fmn(T(0.), tc);
    } else if (sch == 4) {
      T Ml = T(0.25) * bl * br * sqd(tsqrt((MpL * MpL + MpR * MpR) * T(0.5)) - 1);
      T Mpl = Ml * (1 - tdiv(pL - pR, 2 * rL * (ssa * ssa)));
      T Mml = Ml * (1 - tdiv(pL - pR, 2 * rR * (ssa * ssa)));
      cplus = cplus - Mpl;
      cminus = cminus + Mml;
    }
    T Fp0 = rL * ssa * cplus, Fm0 = rR * ssa * cminus;
    Fp0 = Fp0 * zero;
    Fm0 = Fm0 * zero;
    T Fp[5], Fm[5];
    Fp[0] = Fp0; Fm[0] = Fm0;
    Fp[1] = (Fp0 * uL) + (sDp * pL * nx);
    Fp[2] = (Fp0 * vL) + (sDp * pL * ny);
    Fp[3] = (Fp0 * wL) + (sDp * pL * nz);
    Fp[4] = Fp0 * ((T(0.5) * (uL * uL + vL * vL + wL * wL)) + tdiv(gmr * pL, rL));
    Fm[1] = (Fm0 * uR) + (sDm * pR * nx);
    Fm[2] = (Fm0 * vR) + (sDm * pR * ny);
    Fm[3] = (Fm0 * wR) + (sDm * pR * nz);
    Fm[4] = Fm0 * ((T(0.5) * (uR * uR + vR * vR + wR * wR)) + tdiv(gmr * pR, rR));
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
  static __device__ __forceinline__ void visc_face(const Blk &b, long c, long cm, long gL, long gR, T nx, T ny, T nz,
                                                   T A, bool laminar_only, T (&v)[NV]) {
    const double *qp = b.qp;
    const long nc = b.nc;
    const T delx = T(b.cx[c] - b.cx[cm]), dely = T(b.cy[c] - b.cy[cm]), delz = T(b.cz[c] - b.cz[cm]);
    T d2 = delx * delx + dely * dely + delz * delz;
    T id2 = tdiv(T(1.0), d2);  // normal_comp*del/d_LR = (del_phi - grad.del)*del/d_LR^2
    // FP32 mode reads the FP32 copy of the gradients
    using GS = typename std::conditional<sizeof(T) == sizeof(float), const float, const double>::type;
    GS *gx, *gy, *gz;
    if constexpr (sizeof(T) == sizeof(float)) {
      gx = b.g32[0]; gy = b.g32[1]; gz = b.g32[2];
    } else {
      gx = b.g[0]; gy = b.g[1]; gz = b.g[2];
    }
    const long ngr = b.ngr;
    T gr[4][3];
#pragma unroll
    for (int l = 0; l < 4; ++l) {
      gr[l][0] = T(0.5) * (T(gx[l * ngr + gL]) + T(gx[l * ngr + gR]));
      gr[l][1] = T(0.5) * (T(gy[l * ngr + gL]) + T(gy[l * ngr + gR]));
      gr[l][2] = T(0.5) * (T(gz[l * ngr + gL]) + T(gz[l * ngr + gR]));
    }
    const double T_LE = qp[4 * nc + cm] / (qp[cm] * b.R);
    const double T_RE = qp[4 * nc + c] / (qp[c] * b.R);
    T del[4] = {T(qp[nc + c] - qp[nc + cm]), T(qp[2 * nc + c] - qp[2 * nc + cm]), T(qp[3 * nc + c] - qp[3 * nc + cm]),
                T(T_RE - T_LE)};
#pragma unroll
    for (int l = 0; l < 4; ++l) {
      T ncmp = (del[l] - (gr[l][0] * delx + gr[l][1] * dely + gr[l][2] * delz)) * id2;
      gr[l][0] = gr[l][0] + ncmp * delx;
      gr[l][1] = gr[l][1] + ncmp * dely;
      gr[l][2] = gr[l][2] + ncmp * delz;
    }
    T mu_f = T(0.5) * (T(b.mu[cm]) + T(b.mu[c]));
    // compute_viscous_fluxes_laminar takes mu_t whenever the run is turbulent
    // (src/viscous.f90:270-277)
    T mut_f = (NV >= 6) ? T(0.5) * (T(b.mut[cm]) + T(b.mut[c])) : T(0.0);
    T tm = mu_f + mut_f;
    T dudx = gr[0][0], dudy = gr[0][1], dudz = gr[0][2];
    T dvdx = gr[1][0], dvdy = gr[1][1], dvdz = gr[1][2];
    T dwdx = gr[2][0], dwdy = gr[2][1], dwdz = gr[2][2];
    T div3 = (dudx + dvdy + dwdz) * (T(1.0) / T(3.0));
    T Txx = T(2.) * tm * (dudx - div3);
    T Tyy = T(2.) * tm * (dvdy - div3);
    T Tzz = T(2.) * tm * (dwdz - div3);
    T Txy = tm * (dvdx + dudy);
    T Txz = tm * (dwdx + dudz);
    T Tyz = tm * (dwdy + dvdz);
    T K_heat = (mu_f * T(b.iPr) + mut_f * T(b.itPr)) * T(b.cpR);
    T Qx = K_heat * gr[3][0], Qy = K_heat * gr[3][1], Qz = K_heat * gr[3][2];
    T uf = T(0.5) * (T(qp[nc + cm]) + T(qp[nc + c]));
    T vf = T(0.5) * (T(qp[2 * nc + cm]) + T(qp[2 * nc + c]));
    T wf = T(0.5) * (T(qp[3 * nc + cm]) + T(qp[3 * nc + c]));
    v[0] = T(0.0);
    v[1] = (Txx * nx + Txy * ny + Txz * nz) * A;
    v[2] = (Txy * nx + Tyy * ny + Tyz * nz) * A;
    v[3] = (Txz * nx + Tyz * ny + Tzz * nz) * A;
    v[4] = A * (((Txx * uf + Txy * vf + Txz * wf + Qx) * nx) + ((Txy * uf + Tyy * vf + Tyz * wf + Qy) * ny) +
                ((Txz * uf + Tyz * vf + Tzz * wf + Qz) * nz));
    if constexpr (NV >= 7) {
      v[5] = T(0.0);
      v[6] = T(0.0);
    }
    if constexpr (NV == 8) v[NV - 1] = T(0.0);
    if constexpr (NV == 6) {
      // nu-tilde diffusion with (mu + rho_f nu~_f) / sigma, sigma = 2/3; the
      // quotient taken as a product with 1.5 (within an ulp of it)
      T g0 = T(0.5) * (T(gx[4 * ngr + gL]) + T(gx[4 * ngr + gR]));
      T g1 = T(0.5) * (T(gy[4 * ngr + gL]) + T(gy[4 * ngr + gR]));
      T g2 = T(0.5) * (T(gz[4 * ngr + gL]) + T(gz[4 * ngr + gR]));
      T ncmp = (T(qp[5 * nc + c] - qp[5 * nc + cm]) - (g0 * delx + g1 * dely + g2 * delz)) * id2;
      g0 = g0 + ncmp * delx;
      g1 = g1 + ncmp * dely;
      g2 = g2 + ncmp * delz;
      T rhof = T(0.5) * (T(qp[cm]) + T(qp[c]));
      T mtv = T(0.5) * (T(qp[5 * nc + cm]) + T(qp[5 * nc + c])) * rhof;
      v[NV == 6 ? 5 : 0] = (A * ((mu_f + mtv) * (g0 * nx + g1 * ny + g2 * nz))) * T(1.5);
    }
    if (NV >= 7 && (!laminar_only || b.kkl)) {
      // SST k / omega diffusion and the -2/3 rho k stress, src/viscous.f90:328-447;
      // k-kL's has sigma_k = sigma_phi = 1 and runs on the k faces of a 2-D
      // grid too (:450-567, :106-109)
      T gt[2][3];
#pragma unroll
      for (int l = 0; l < 2; ++l) {
        gt[l][0] = T(0.5) * (T(gx[(4 + l) * ngr + gL]) + T(gx[(4 + l) * ngr + gR]));
        gt[l][1] = T(0.5) * (T(gy[(4 + l) * ngr + gL]) + T(gy[(4 + l) * ngr + gR]));
        gt[l][2] = T(0.5) * (T(gz[(4 + l) * ngr + gL]) + T(gz[(4 + l) * ngr + gR]));
      }
      T dt2[2] = {T(qp[5 * nc + c] - qp[5 * nc + cm]), T(qp[6 * nc + c] - qp[6 * nc + cm])};
#pragma unroll
      for (int l = 0; l < 2; ++l) {
        T ncmp = (dt2[l] - (gt[l][0] * delx + gt[l][1] * dely + gt[l][2] * delz)) * id2;
        gt[l][0] = gt[l][0] + ncmp * delx;
        gt[l][1] = gt[l][1] + ncmp * dely;
        gt[l][2] = gt[l][2] + ncmp * delz;
      }
      // k-kL leaves F1 at 0 and sets sigma_w2 to 1, so both blends are 1
      T F1 = T(0.5) * (T(b.f1[cm]) + T(b.f1[c]));
      T sigma_kf = T(0.85) * F1 + T(1.0) * (T(1.0) - F1);
      T sigma_wf = T(0.5) * F1 + T(b.sigma_w2) * (T(1.0) - F1);
      T rhof = T(0.5) * (T(qp[cm]) + T(qp[c]));
      T tkf = T(0.5) * (T(qp[5 * nc + cm]) + T(qp[5 * nc + c]));
      T Tkk = -(T(2.0) * rhof * tkf * (T(1.0) / T(3.0)));
      v[1] = v[1] + Tkk * nx * A;
      v[2] = v[2] + Tkk * ny * A;
      v[3] = v[3] + Tkk * nz * A;
      T kfl = A * ((mu_f + sigma_kf * mut_f) * (gt[0][0] * nx + gt[0][1] * ny + gt[0][2] * nz));
      v[4] = v[4] + kfl;
      v[NV >= 7 ? 5 : 0] = kfl;
      v[NV >= 7 ? 6 : 0] = A * ((mu_f + sigma_wf * mut_f) * (gt[1][0] * nx + gt[1][1] * ny + gt[1][2] * nz));
    }
    if constexpr (NV == 8) {
      // lctm2015's intermittency: (mu + mu_t) grad(gamma), no sigma, none on
      // the k faces of a 2-D grid (src/viscous.f90:114-128, 659-746)
      if (!laminar_only) {
        T g0 = T(0.5) * (T(gx[6 * ngr + gL]) + T(gx[6 * ngr + gR]));
        T g1 = T(0.5) * (T(gy[6 * ngr + gL]) + T(gy[6 * ngr + gR]));
        T g2 = T(0.5) * (T(gz[6 * ngr + gL]) + T(gz[6 * ngr + gR]));
        T ncmp = (T(qp[7 * nc + c] - qp[7 * nc + cm]) - (g0 * delx + g1 * dely + g2 * delz)) * id2;
        g0 = g0 + ncmp * delx;
        g1 = g1 + ncmp * dely;
        g2 = g2 + ncmp * delz;
        v[NV - 1] = A * ((mu_f + mut_f) * (g0 * nx + g1 * ny + g2 * nz));
      }
    }
  }
};

// ---- the face kernel
// D: sweep direction.  Faces: n = 1..N along D, 1..N-1 across.  NV: 5
// (laminar), 6 (Spalart-Allmaras), 7 (SST) or 8 (SST with lctm2015)
// variables; RECON: the
// reconstruction; T: the face arithmetic, double or float.  Compile-time
// counts keep the per-variable arrays in registers.
// Registers are fitted to 4 waves per SIMD: 5 spills the SST/WENO variants
// and measured slower even where it does not spill.
// The i sweep stages its stencils in LDS: a block takes 256 consecutive
// offsets p of the padded (i, j) plane of one k, i.e. the cells whose lower
// i face it computes.  The stencil of every face, cells p-3..p+2, is then a
// contiguous run of that plane, rows included, so one coalesced load per
// variable fills sq[l][0..260] (offsets p0-3..p0+257) and each face reads its
// six cells from LDS.  Offsets that are not faces (ghost columns and rows,
// the tail) take part in the load and the barrier, then leave.
constexpr int FLUX_I_SPAN = 256;
__host__ __device__ inline long flux_i_first(const Blk &b) { return 3 + 3L * b.pi; }
__host__ __device__ inline long flux_i_span(const Blk &b) {
  return (b.imx + 2) + (long)(b.jmx + 1) * b.pi - flux_i_first(b) + 1;
}

template <int D, int NV, int RECON, bool FLAT, typename T, int TI = TILE_I>
__device__ __forceinline__ void k_flux_b(const Blk &b, const Wg &wg, int store_csnd) {
  using M = FaceMath<T>;
  const int ii = D == 0, jj = D == 1, kk = D == 2;
  int i, j;
  const int blk = plane_block(wg, b.xcd_remap);
  const int k = wg.y + 1;
  // the stencil's storage: the FP64 state, or its FP32 copy in FP32 mode
  using QS = typename std::conditional<sizeof(T) == sizeof(float), float, double>::type;
  const QS *qsrc;
  if constexpr (sizeof(T) == sizeof(float)) qsrc = b.q32;
  else qsrc = b.qp;
  const QS *lds_row = nullptr;  // the i sweep's stencil row in LDS
  if constexpr (D == 0) {
    __shared__ QS sq[NV][FLUX_I_SPAN + 5];
    const int t = (int)threadIdx.x;
    const long p0 = flux_i_first(b) + (long)blk * FLUX_I_SPAN;
    const long p = p0 + t;
    const long plane = (long)b.pi * b.pj;
    i = (int)(p % b.pi) - 2;
    j = (int)(p / b.pi) - 2;
    const bool active = p < flux_i_first(b) + flux_i_span(b) && i >= 1 && i <= b.imx && j >= 1 &&
                        j <= b.jmx - 1 && k <= b.kmx - 1;
    const long kbase = plane * (long)(k + 2);
#pragma unroll
    for (int l = 0; l < NV; ++l) {
      const QS *ql = qsrc + (long)l * b.nc + kbase;
      for (int e = t; e < FLUX_I_SPAN + 5; e += FLUX_I_SPAN) {
        const long pe = p0 - 3 + e;
        sq[l][e] = (pe < plane && k <= b.kmx - 1) ? ql[pe] : QS(0);
      }
    }
    __syncthreads();
    if (!active) return;
    lds_row = &sq[0][t + 3];
  } else {
    // j, k sweeps: 64 x 4 tiles
    if (!tile_ij<TI>(b.imx - 1 + ii, b.jmx - 1 + jj, i, j, blk)) return;
    i += 1;
    j += 1;
    if (k > b.kmx - 1 + kk) return;
  }
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
  T out[NV];
  T csf = T(0.0);
  const long fo = fi(b, D, i, j, k);
  const T nx = T(b.fx[D][fo]), ny = T(b.fy[D][fo]), nz = T(b.fz[D][fo]), A = T(b.fA[D][fo]);
  {
    T qL[NV], qR[NV];
    typename M::Aux ax;
    ax.pb = (RECON == 1 || RECON == 2) && b.pbs[D];
    if (ax.pb) {
      if constexpr (D == 0) M::pb_factors(b, n, N, lds_row + 4 * (FLUX_I_SPAN + 5), 1, ax);
      else M::pb_factors(b, n, N, qsrc + 4 * b.nc + c, s, ax);
    }
    if constexpr (RECON == 4) M::nm_blends(b.vol + c, s, ax);
#pragma unroll
    for (int l = 0; l < nv; ++l) {
      if constexpr (D == 0) {
        M::template face_states<RECON>(b, D, n, N, l, lds_row + l * (FLUX_I_SPAN + 5), 1, ax, qL[l], qR[l]);
      } else {
        const QS *u = qsrc + (long)l * b.nc + c;
        M::template face_states<RECON>(b, D, n, N, l, u, s, ax, qL[l], qR[l]);
      }
    }
    if (store_csnd) csf = T(0.5) * (tsqrt(T(b.gm) * qL[4] / qL[0]) + tsqrt(T(b.gm) * qR[4] / qR[0]));
    const T zero = T((n == 1) ? (double)b.mz[D][0] : (n == N) ? (double)b.mz[D][1] : 1.0);
    if constexpr (flat) {
#pragma unroll
      for (int l = 0; l < nv; ++l) out[l] = T(0.0);
      (void)zero;
    } else {
      M::template inviscid_flux<NV>(b, qL, qR, nx, ny, nz, A, zero, out);
    }
  }
  if (b.mu_ref != 0.0) {
    T v[NV];
    M::template visc_face<NV>(b, c, cm, gri(b, i - ii, j - jj, k - kk), gri(b, i, j, k), nx, ny, nz, A, flat, v);
#pragma unroll
    for (int l = 1; l < nv; ++l) out[l] = out[l] - v[l];
  }

  bool bad = false;
#pragma unroll
  for (int l = 0; l < nv; ++l) {
    bad |= isnan(out[l]);
    // FP32 mode keeps the face fluxes in FP32 (the update widens them)
    if constexpr (sizeof(T) == sizeof(float)) ((float *)b.flux[D])[(long)l * b.nfl[D] + o] = out[l];
    else b.flux[D][(long)l * b.nfl[D] + o] = (double)out[l];
  }
  if (bad) atomicOr(b.err, ERR_NAN_FLUX);
  if (store_csnd) b.csnd[D][o] = (double)csf;
}
template <int D, int NV, int RECON, bool FLAT = false, typename T = double>
__global__ void __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(4))) k_flux(Blk b, int store_csnd) {
  k_flux_b<D, NV, RECON, FLAT, T>(b, FEST_WG_HW, store_csnd);
}

}  // namespace fest
