/*
 * fest_oracle.h -- CPU restatement of FEST-3D's explicit residual and RK
 * update, used ONLY as the checker by tests/, __graft_entry__.smoke() and the
 * cpu_baseline leg of bench.py.  It is test infrastructure, never part of the
 * product path.
 *
 * Every array uses the reference's Fortran column-major layout with its
 * explicit lower bounds (src/update.f90:143-165, src/face/state/
 * face_interpolant.f90:43-56, src/gradients.f90:124-126):
 *   padded cells  (-2:imx+2,-2:jmx+2,-2:kmx+2[,n_var])
 *   I/J/K faces   (-2:imx+3,...), (...,-2:jmx+3,...), (...,-2:kmx+3)
 *   nodes         (-2:imx+3,-2:jmx+3,-2:kmx+3)
 *   F (1:imx,1:jmx-1,1:kmx-1,n_var), G (1:imx-1,1:jmx,..), H (..,1:kmx)
 *   residue, R_store, delta_t (1:imx-1,1:jmx-1,1:kmx-1[,n_var])
 *   face states x (0:imx+1,1:jmx-1,1:kmx-1,n_var), y, z analogues
 *   gradients (0:imx,0:jmx,0:kmx,n_grad), slots [u,v,w,T(,k,w)(,gamma)]
 */
#ifndef FEST_ORACLE_H
#define FEST_ORACLE_H
#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

enum { ORC_FLUX_AUSM = 0, ORC_FLUX_AUSMP = 1, ORC_FLUX_AUSMUP = 2,
       ORC_FLUX_SLAU = 3, ORC_FLUX_LDFSS0 = 4, ORC_FLUX_VANLEER = 5 };
enum { ORC_RECON_NONE = 0, ORC_RECON_MUSCL = 1, ORC_RECON_PPM = 2,
       ORC_RECON_WENO = 3, ORC_RECON_WENO_NM = 4 };
/* the MUSCL limiter: the reference's Koren-alpha form, or van Albada
 * (the form muscl.f90:180,186 leaves commented out) */
enum { ORC_LIM_KOREN = 0, ORC_LIM_VAN_ALBADA = 1 };
enum { ORC_TURB_NONE = 0, ORC_TURB_SST = 1, ORC_TURB_SST2003 = 2, ORC_TURB_SA = 3, ORC_TURB_KKL = 4 };
enum { ORC_TA_NONE = 0, ORC_TA_RK2 = 2, ORC_TA_RK4 = 4, ORC_TA_TVDRK2 = 12, ORC_TA_TVDRK3 = 13,
       ORC_TA_IMPLICIT = 20 };

/* Status codes: same meaning as the product library's. */
enum { ORC_OK = 0, ORC_E_NONPHYSICAL = -1, ORC_E_NAN = -2, ORC_E_CONFIG = -3 };

typedef struct {
  int32_t imx, jmx, kmx, n_var;
  int32_t flux, recon, turb;
  int32_t time_local;          /* 1: 'l' local time step, 0: 'g' global    */
  int32_t time_accuracy;       /* ORC_TA_*                                  */
  int32_t accur;
  int32_t lim[3], tlim[3];     /* i/j/k limiter switches, turbulent ones   */
  int32_t mu_sutherland;       /* 0 constant, 1 sutherland_law             */
  int32_t bc_id[6];            /* imin,imax,jmin,jmax,kmin,kmax            */
  double global_time_step;
  double CFL;
  double rho_inf, u_inf, v_inf, w_inf, p_inf, tk_inf, tw_inf;
  double vel_mag, MInf;
  double gm, R_gas, mu_ref, T_ref, S_ref, Pr, tPr;
  double fix_rho[6], fix_u[6], fix_v[6], fix_w[6], fix_p[6];
  double fix_tk[6], fix_tw[6], fix_twall[6];
  double fix_tp[6]; /* total pressure of a -11 face */
  double tv_inf;     /* Spalart-Allmaras free-stream nu-tilde */
  double fix_tv[6];  /* fixed nu-tilde of each face */
  int32_t transition; /* 0 none, 1 'bc', 2 'lctm2015' (SST and SST-2003 only) */
  double tu_inf;      /* free-stream turbulence intensity, percent */
  double tkl_inf;     /* free-stream kL (k-kL) */
  double fix_tkl[6];  /* fixed kL of each face (k-kL) */
  int32_t pbs[3];     /* i/j/kPB_switch: pressure-based switch of MUSCL and PPM */
  int32_t muscl_limiter; /* ORC_LIM_* */
  double tgm_inf;     /* free-stream intermittency (lctm2015, variable 8) */
  double fix_tgm[6];  /* fixed intermittency of each face (lctm2015) */
} orc_case;

typedef struct orc orc;

/* Array ids for orc_array(). */
enum { ORC_QP = 0, ORC_TEMP, ORC_MU, ORC_MUT, ORC_F1, ORC_DIST,
       ORC_VOL, ORC_CX, ORC_CY, ORC_CZ,
       ORC_IA, ORC_INX, ORC_INY, ORC_INZ,
       ORC_JA, ORC_JNX, ORC_JNY, ORC_JNZ,
       ORC_KA, ORC_KNX, ORC_KNY, ORC_KNZ,
       ORC_XL, ORC_XR, ORC_YL, ORC_YR, ORC_ZL, ORC_ZR,
       ORC_F, ORC_G, ORC_H, ORC_RES, ORC_DT, ORC_USTORE, ORC_RSTORE,
       ORC_GX, ORC_GY, ORC_GZ, ORC_NODES, ORC_NARRAYS };

orc *orc_create(const orc_case *c);
void orc_destroy(orc *s);
double *orc_array(orc *s, int id, size_t *count);

/* ghost nodes + face/cell metrics from the interior nodes stored in
 * ORC_NODES (src/grid.f90:137-236, src/geometry.f90:50-545). */
int orc_geometry(orc *s);
/* brute-force wall distance from the no-slip faces (src/wall/wall.f90:198-288,
 * src/wall/wall_dist.f90:84-131), surface nodes rounded like the
 * ES18.10E3 text round trip of src/wall/wall.f90:101. */
int orc_wall_dist(orc *s);
/* The same over a multiblock grid: every block's distance to the no-slip
 * surface nodes of all blocks, the one surface-node file every process
 * reads (src/wall/wall.f90:55-111). */
int orc_group_wall_dist(orc **blocks, int nblocks);

/* Temp = p/(R rho) over the padded array (src/update.f90:170). */
void orc_temp_refresh(orc *s);
/* Selected passes of one residual evaluation: mask bit 0 ghost cells,
 * 1 face states (+ boundary states), 2 convective fluxes, 3 gradients,
 * viscosity and viscous fluxes, 4 residue and turbulence source. */
int orc_stage(orc *s, int current_iter, int mask);
/* One convective face flux of the selected scheme (n_var values). */
void orc_face_flux(orc *s, const double *qL, const double *qR, double nx, double ny, double nz, double A,
                   int zero, double *out);
/* One residual evaluation C2..C10 (src/update.f90:495-547). */
int orc_residual(orc *s, int current_iter);
/* Local / global time step (src/time.f90:122-334). */
int orc_time_step(orc *s);
/* One full time step of get_next_solution (src/update.f90:129-226). */
int orc_step(orc *s, int current_iter);
/* Absolute residual norms Res_abs(0:n_var) (src/resnorm.f90:171-225). */
void orc_resnorm(orc *s, double *out);

/* One interface link of a multiblock grid: the three ghost layers of face
 * `face` (1..6: imin..kmax) of block `block` receive the three interior
 * layers next to face `from_face` of block `from_block`.  lo1/hi1, lo2/hi2
 * are the map's node ranges for the receiving face's first and second
 * tangential direction (j,k on i-faces; i,k on j-faces; i,j on k-faces),
 * dir_switch the map's direction swap (src/mapping.f90:64-258). */
typedef struct {
  int32_t block, face, from_block, from_face;
  int32_t lo1, hi1, lo2, hi2, dir_switch;
} orc_link;
/* The interface exchange of src/interface1.f90:96-705 over all links: every
 * sender packs its layers in its own loop order, every receiver unpacks them
 * in the order its map ranges give. */
int orc_exchange(orc **blocks, int nblocks, const orc_link *links, int nlinks);
/* One time step of get_next_solution on every block of a multiblock grid,
 * the interface exchange opening every residual evaluation
 * (src/update.f90:129-226, 534). */
int orc_group_step(orc **blocks, int nblocks, const orc_link *links, int nlinks, int current_iter);
/* The same RK step with the blocks of each stage spread over `nthreads`
 * threads (one block per thread at a time), as the reference spreads its
 * blocks over MPI processes; the exchange stays serial.  RK2/RK4 only.
 * Results equal orc_group_step's bit for bit (blocks share no state). */
int orc_group_step_mt(orc **blocks, int nblocks, const orc_link *links, int nlinks, int current_iter,
                      int nthreads);

#ifdef __cplusplus
}
#endif
#endif
