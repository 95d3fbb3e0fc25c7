/*
 * fest_hip.h -- C ABI of the MI355X explicit-residual engine for FEST-3D.
 *
 * The entry points replace the body of the reference's time-integration
 * module and the residual passes it calls; a Fortran host binds them with
 * ISO_C_BINDING and keeps the reference's call surface:
 *
 *   setup_update(control, scheme, flow, dims)          src/update.f90:85-126
 *   get_next_solution(qp, Temp, residue, delta_t, cells, F, G, H,
 *                     Ifaces, Jfaces, Kfaces, control, scheme, flow, bc, dims)
 *                                                      src/update.f90:129-226
 *   get_total_conservative_Residue(...)                src/update.f90:495-547
 *   compute_time_step(...)                             src/time.f90:291-334
 *   find_resnorm(...) (absolute norms)                 src/resnorm.f90:62-91,171-225
 *
 * Conventions
 *   - Host arrays are plain double arrays in the reference's Fortran layout
 *     (column-major, explicit lower bounds); derived-type arrays are passed
 *     component by component (the Fortran side extracts cells%volume,
 *     Ifaces%nx, ...).  Nothing is retained after a call returns.
 *   - State and geometry stay resident on the device between calls.
 *   - Every function returns 0 on success or a negative FEST_E_* code;
 *     fest_last_error() gives the message.  The Fortran wrapper turns a
 *     nonzero code into the print + STOP of the reference's Fatal_error
 *     (src/error.h:17).
 *   - One host thread per handle; calls on one handle are not re-entrant.
 */
#ifndef FEST_HIP_H
#define FEST_HIP_H
#include <stdint.h>
#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

/* error codes */
#define FEST_OK 0
#define FEST_E_NONPHYSICAL (-1) /* rho<0, p<0 or NaN in the update, src/update.f90:448-452 */
#define FEST_E_NAN (-2)         /* NaN in fluxes, gradients or mu (the reference's isnan passes) */
#define FEST_E_CONFIG (-3)      /* unknown/unsupported scheme, BC or extent */
#define FEST_E_HIP (-4)         /* HIP runtime error or no device */
#define FEST_E_ARG (-5)         /* bad argument (null pointer, bad mask) */

/* scheme ids; the Fortran side maps schemetype strings onto them */
enum { FEST_FLUX_AUSM = 0, FEST_FLUX_AUSMP = 1, FEST_FLUX_AUSMUP = 2,
       FEST_FLUX_SLAU = 3, FEST_FLUX_LDFSS0 = 4, FEST_FLUX_VANLEER = 5 };
enum { FEST_RECON_NONE = 0, FEST_RECON_MUSCL = 1, FEST_RECON_PPM = 2,
       FEST_RECON_WENO = 3, FEST_RECON_WENO_NM = 4 /* 'weno_NM', src/face/state/weno_NM.f90 */ };
/* MUSCL limiter: the reference's Koren-alpha form (src/face/state/
 * muscl.f90:178,184), or van Albada (r^2+r)/(r^2+1), the form the reference
 * leaves commented out at muscl.f90:180,186 */
enum { FEST_LIM_KOREN = 0, FEST_LIM_VAN_ALBADA = 1 };
enum { FEST_TURB_NONE = 0, FEST_TURB_SST = 1, FEST_TURB_SST2003 = 2 /* src/viscosity.f90:215, source.f90:205 */,
       FEST_TURB_SA = 3 /* 'sa', src/viscosity.f90:149, source.f90:835, lusgs.f90:1680 */,
       FEST_TURB_KKL = 4 /* 'kkl', src/viscosity.f90:469, source.f90:607, viscous.f90:450, lusgs.f90:1198 */ };
enum { FEST_TRANS_NONE = 0, FEST_TRANS_BC = 1 /* the 'bc' intermittency of the SA and SST sources */,
       FEST_TRANS_LCTM2015 = 2 /* 'lctm2015': the intermittency as variable 8 (n_var 8), SST and
                                  SST-2003, FP64, RK (src/source.f90:273-463, viscous.f90:659-746,
                                  viscosity.f90:265-279, CC.f90) */ };
enum { FEST_TA_NONE = 0, FEST_TA_RK2 = 2, FEST_TA_RK4 = 4, FEST_TA_TVDRK2 = 12, FEST_TA_TVDRK3 = 13,
       FEST_TA_IMPLICIT = 20 /* LU-SGS, src/lusgs.f90 */ };

/* type(extent), src/vartypes.f90:21-26 */
typedef struct { int32_t imx, jmx, kmx, n_var; } fest_extent;

/* the hot-path fields of type(schemetype), src/vartypes.f90:195-236 */
typedef struct {
  int32_t flux;            /* scheme_name  -> FEST_FLUX_*                */
  int32_t recon;           /* interpolant  -> FEST_RECON_*               */
  int32_t turb;            /* turbulence   -> FEST_TURB_*                */
  int32_t time_local;      /* time_stepping_method: 1 'l', 0 'g'         */
  int32_t time_accuracy;   /* time_step_accuracy -> FEST_TA_*            */
  int32_t accur;           /* accur (higher-order symmetry copy)         */
  int32_t lim[3];          /* i/j/klimiter_switch                        */
  int32_t tlim[3];         /* i/j/ktlimiter_switch                       */
  double global_time_step; /* global_time_step (used when > 0, 'g')      */
  int32_t transition;      /* transition -> FEST_TRANS_* (source.f90:125-145) */
  int32_t pbs[3];          /* i/j/kPB_switch: pressure-based switch of MUSCL
                              and PPM (muscl.f90:37-112, ppm.f90:109-186)  */
  int32_t muscl_limiter;   /* FEST_LIM_*                                 */
} fest_scheme;

/* the hot-path fields of type(flowtype), src/vartypes.f90:239-290 */
typedef struct {
  double density_inf, x_speed_inf, y_speed_inf, z_speed_inf, pressure_inf;
  double tk_inf, tw_inf, vel_mag, MInf;
  double gm, R_gas, mu_ref, T_ref, Sutherland_temp, Pr, tPr;
  int32_t mu_sutherland;   /* mu_variation: 0 'constant', 1 'sutherland_law' */
  double tv_inf;           /* Spalart-Allmaras free-stream nu-tilde (state.f90:106) */
  double tu_inf;           /* free-stream turbulence intensity, percent (vartypes.f90) */
  double tkl_inf;          /* k-kL free-stream kL (state.f90:103) */
  double tgm_inf;          /* lctm2015 free-stream intermittency (vartypes.f90:258, read.f90:306-310) */
} fest_flow;

/* the hot-path fields of type(boundarytype), src/vartypes.f90:293-365 */
typedef struct {
  int32_t id[6];           /* imin,imax,jmin,jmax,kmin,kmax ids          */
  double fixed_density[6], fixed_x_speed[6], fixed_y_speed[6], fixed_z_speed[6];
  double fixed_pressure[6], fixed_tk[6], fixed_tw[6], fixed_wall_temperature[6];
  double fixed_Tpressure[6];  /* total pressure of a -11 face (read_bc.f90:79-80) */
  double fixed_tv[6];         /* fixed nu-tilde of each face (read_bc.f90:106-113) */
  double fixed_tkl[6];        /* fixed kL of each face (read_bc.f90:103-106, 198-202) */
  double fixed_tgm[6];        /* fixed intermittency of each face (read_bc.f90:123-131, 213-218) */
} fest_bc;

typedef struct fest_handle fest_handle;

/* precision of fest_create: the plain FP64 path, or face reconstruction and
 * flux arithmetic in FP32 -- reading FP32 copies of the state and the
 * gradients and storing the face fluxes in FP32 -- with metrics, the
 * residue sums, the RK update, the sources and the norms in FP64 (no
 * counterpart in the reference, which is FP64 throughout) */
enum { FEST_PREC_FP64 = 0, FEST_PREC_FP32_FLUX = 1 };

/* Allocates the device-resident block.  precision: FEST_PREC_*.
 * Replaces the allocations of setup_update/setup_scheme/setup_gradients/
 * setup_viscosity (src/update.f90:85-126, src/face/flux/convective/
 * scheme.f90:29-65, src/gradients.f90:102-200, src/viscosity.f90:548-592). */
int fest_create(fest_handle **h, const fest_extent *dims, const fest_scheme *scheme,
                const fest_flow *flow, const fest_bc *bc, int precision);
void fest_destroy(fest_handle *h);
const char *fest_last_error(const fest_handle *h);

/* Geometry, option 1: upload the reference's metrics as the solver holds
 * them after setup_geometry (src/geometry.f90:548-580): cells4 = four
 * padded cell planes [volume, centerx, centery, centerz]; if4/jf4/kf4 = four
 * face planes [A, nx, ny, nz] each; dist = wall distance (nullable, padded
 * cell layout, src/wall/wall_dist.f90:22). */
int fest_upload_geometry(fest_handle *h, const double *cells4, const double *if4,
                         const double *jf4, const double *kf4, const double *dist);

/* Geometry, option 2: compute it on the device from the grid nodes, three
 * planes over (-2:imx+3,-2:jmx+3,-2:kmx+3) with the interior 1..imx filled:
 * ghost nodes (src/grid.f90:137-236), face areas/normals, volumes and
 * centroids (src/geometry.f90:50-545) and, when with_wall_dist != 0, the
 * brute-force wall distance to the no-slip nodes (src/wall/wall.f90:198-288,
 * src/wall/wall_dist.f90:84-131). */
int fest_set_grid(fest_handle *h, const double *nodes, int with_wall_dist);

/* Upload the primitive state qp (-2:imx+2,-2:jmx+2,-2:kmx+2,1:n_var) and,
 * optionally, Temp (nullable). */
int fest_upload_state(fest_handle *h, const double *qp, const double *temp);

/* One full time step of get_next_solution (src/update.f90:129-226),
 * including the Temp refresh at the start (src/update.f90:170) and the
 * stage-1 time step.  current_iter feeds the inflow 'fix' rule
 * (src/boundary/bc_primitive.f90:236,318,366). */
int fest_get_next_solution(fest_handle *h, int current_iter, double cfl);

/* One residual evaluation, get_total_conservative_Residue
 * (src/update.f90:495-547), on the resident state and Temp. */
int fest_residual(fest_handle *h, int current_iter);

/* The discrete adjoint of the inviscid residual: bar_qp = (dR/dqp)^T
 * bar_residue at the resident state, R being what fest_residual computes
 * (get_total_conservative_Residue, src/update.f90:495-547: the fluxes of
 * src/face/flux/convective/ on the face states of src/face/state/,
 * differenced per cell as in src/face/flux/convective/scheme.f90:111-141),
 * with the ghost cells of the block's own boundary rules
 * (src/boundary/bc_primitive.f90) taken as functions of the interior cells.
 * No counterpart in the reference, which has no adjoint: it is for
 * optimisation and sensitivity callers.
 * bar_residue has the residue's layout (1:imx-1,1:jmx-1,1:kmx-1,1:5),
 * bar_qp the state's (-2:imx+2,-2:jmx+2,-2:kmx+2,1:5): interior cells hold
 * the adjoint, the ghost cells of interface faces (ids >= 0 and -10) the
 * share owed to the neighbouring block, every other ghost cell zero.  The
 * ghost cells are refilled first, as fest_residual does at current_iter 1.
 * Handles with five variables and at least 4 cells in every 3-D direction,
 * inviscid, or laminar (the viscous fluxes, the Green-Gauss gradients and
 * their ghost-cell rule, Sutherland's law; src/viscous.f90,
 * src/gradients.f90) on 3-D FP64 handles without interface faces, where
 * the temperature plane's ghost cells (the state's p/(R rho) at upload,
 * src/update.f90:170) are constants; else FEST_E_CONFIG; on a
 * FEST_PREC_FP32_FLUX handle it is the adjoint of the FP64 residual (the
 * face arithmetic in FP64) at the resident FP64 state.  The result
 * is the exact derivative of the branches (limiters, switches) the
 * residual takes at this state.  bar_nodes (nullable): also the adjoint
 * with respect to the grid nodes, (dR/dnodes)^T bar_residue in the layout
 * of fest_set_grid's nodes (-2:imx+3,-2:jmx+3,-2:kmx+3,1:3), through the
 * face metrics (src/geometry.f90:236-314) and the ghost-node extrapolation
 * (src/grid.f90:137-236) to the real nodes; ghost nodes hold zero.  Not
 * for the pole (-7), WENO-NM or viscous handles (FEST_E_CONFIG). */
int fest_vjp(fest_handle *h, const double *bar_residue, double *bar_qp, double *bar_nodes);

/* compute_time_step (src/time.f90:291-334) from the face states of the
 * last residual evaluation. */
int fest_time_step(fest_handle *h, double cfl);

/* Download selected arrays in the reference layout.  Each pointer may be
 * NULL when its bit is not set in `what`. */
#define FEST_DL_QP      (1u << 0)
#define FEST_DL_TEMP    (1u << 1)
#define FEST_DL_RESIDUE (1u << 2)
#define FEST_DL_F       (1u << 3)
#define FEST_DL_G       (1u << 4)
#define FEST_DL_H       (1u << 5)
#define FEST_DL_DT      (1u << 6)
#define FEST_DL_MU      (1u << 7)
#define FEST_DL_MUT     (1u << 8)
#define FEST_DL_GRADS   (1u << 9)  /* gradqp_x,y,z: 3*n_grad planes (0:imx,0:jmx,0:kmx) */
#define FEST_DL_F1      (1u << 10)
#define FEST_DL_DIST    (1u << 11)
#define FEST_DL_CELLS   (1u << 12) /* cells4 as in fest_upload_geometry */
#define FEST_DL_FACES   (1u << 13) /* if4, jf4, kf4 concatenated */
int fest_download(fest_handle *h, uint32_t what, double *qp, double *temp, double *residue,
                  double *F, double *G, double *H, double *delta_t, double *mu, double *mu_t,
                  double *grads, double *sst_f1, double *dist, double *cells4, double *faces);

/* Absolute residual norms Res_abs(0:n_var) of the last residue
 * (src/resnorm.f90:171-225): [|mass imbalance|, sqrt(sum R_l^2)/scale_l]. */
int fest_resnorm(fest_handle *h, double *res_abs);

/* fest_resnorm without allocation, waiting or host copies: the n_var+1
 * norms of the last residue go to res_abs_dev, device memory the caller
 * owns, in stream order (nothing waits for them).  No counterpart in the
 * reference: it is for callers that cannot allocate -- code under HIP graph
 * capture, serving loops, harnesses that forbid hipMalloc.  Its partial
 * sums live in a static per-device scratch of 72 KiB (at most 1024 partial
 * blocks; larger reductions are rejected with FEST_E_ARG), set up when a
 * handle is created on the device; calls that overlap on one device race
 * for it.  The handle's device must be current. */
int fest_resnorm_device(fest_handle *h, double *res_abs_dev);

/* Device pointers of the resident arrays for callers that keep the state
 * on the device.  name: "qp","temp","residue","F","G","H","dt"; count is
 * in doubles.  A handle created with FEST_PREC_FP32_FLUX holds F, G and H
 * as floats: their count is then in floats.  After a fest_vjp call with
 * bar_nodes, "vjp_metrics_i", "_j", "_k" are the adjoints of the stored
 * face normals and areas (4 planes: nx, ny, nz, A, in the F/G/H face
 * order); FEST_E_ARG before. */
int fest_device_ptr(fest_handle *h, const char *name, void **ptr, size_t *count);

/* Kernel timing (no counterpart in the reference; it is for benchmarks
 * and profiling): fest_timing resets the totals and selects the kernel
 * groups whose launches are bracketed with HIP events on the handle's
 * stream -- NULL or "" none, "*" all, else a comma-separated list of
 * "step_begin", "bc_face", "bc_edges", "grad", "grad_bc", "flux_i",
 * "flux_j", "flux_k", "time_step", "update", "lusgs_diag", "lusgs_sweep",
 * "lusgs_apply", "kkl_source" (the k-kL source pass, k-kL blocks only),
 * "vjp" (the adjoint kernels of fest_vjp, after its ghost-cell pass).
 * fest_kernel_time waits for
 * the stream and returns one group's summed event time and the number of
 * bracketed launches since the last reset. */
int fest_timing(fest_handle *h, const char *groups);
int fest_kernel_time(fest_handle *h, const char *group, double *total_ms, long *launches);

/* Synchronise the handle's stream and report a deferred device error
 * (FEST_E_NONPHYSICAL / FEST_E_NAN raised since the last check). */
int fest_sync(fest_handle *h);

/* Run-time options (no counterpart in the reference):
 *   "graphs"          1 (default): fest_get_next_solution replays the time
 *                     step as a captured HIP graph, one launch per step
 *                     (captured on the second call of each inflow-fix class,
 *                     iterations <= 2 and later; re-captured when the CFL
 *                     number changes); 0: launch kernel by kernel.  Steps
 *                     with kernel timing on (fest_timing) run kernel by
 *                     kernel.
 *   "check_interval"  n >= 1 (default 1): read the device's error word after
 *                     every n-th fest_get_next_solution, so that the host
 *                     waits for the device only then.  With n > 1 an error
 *                     is reported by the next check, fest_resnorm or
 *                     fest_sync; the reference stops at once (n = 1). */
int fest_set_option(fest_handle *h, const char *name, long value);

/* ---- multiblock grids
 * One interface link (the reference's interface map, src/mapping.f90:64-258):
 * the three ghost layers of face `face` (1..6 = imin, imax, jmin, jmax, kmin,
 * kmax) of block `block` receive the three interior layers next to face
 * `from_face` of block `from_block` (block = index into the group's array).
 * The sender packs them in its natural loop order, first tangential
 * direction fastest (j on i-faces, i on j- and k-faces); the receiver walks
 * its first / second tangential direction through the node ranges
 * lo1..hi1 / lo2..hi2 (the map's S1 E1 S2 E2 of the receiving side; a range
 * with hi < lo runs backwards) with the second direction as the outer loop,
 * or the first when dir_switch is set (src/interface1.f90:96-493).  A
 * cross-block periodic pair (bc id -10, src/boundary/bc.f90:27-32,
 * src/interface1.f90:496-705) is a link with aligned ranges. */
typedef struct {
  int32_t block, face, from_block, from_face;
  int32_t lo1, hi1, lo2, hi2;
  int32_t dir_switch;
} fest_link;

typedef struct fest_group fest_group;

/* A multiblock grid over existing blocks (each with its grid and state set
 * or to be set): every block's boundary ids >= 0 and -10 are filled by the
 * links.  The blocks' launches move to the group's stream while the group
 * exists; destroy the group before its blocks.  Replaces the per-rank
 * setup_interface (src/interface1.f90:60-92) of the reference's MPI run,
 * with all blocks resident on one device. */
int fest_group_create(fest_group **g, fest_handle *const *blocks, int nblocks, const fest_link *links,
                      int nlinks);
void fest_group_destroy(fest_group *g);
const char *fest_group_last_error(const fest_group *g);
/* apply_interface (src/interface1.f90:96-493) on every link, device to device. */
int fest_group_exchange(fest_group *g);
/* Wall distance of every block to the no-slip surface nodes of all blocks
 * of the group, as the reference's processes all read one shared
 * surface-node file: write_surfnode + find_wall_dist
 * (src/wall/wall.f90:55-111, src/wall/wall_dist.f90:84-131).  Call after
 * fest_set_grid on every block; replaces their single-block distances. */
int fest_group_wall_dist(fest_group *g);

/* fest_vjp over a group: bar_qps[n] = the adjoint of block n's cells for
 * the residual of every block, bar_residues[n] having block n's residue
 * layout.  The interfaces are exchanged first (as the group's residual
 * opens), each block's adjoint is taken as by fest_vjp, and the adjoints
 * the interface ghost cells hold are added to the cells they were copied
 * from (the transpose of apply_interface, src/interface1.f90:96-705), so
 * every ghost cell ends at zero.  No counterpart in the reference.  The
 * conditions of fest_vjp hold for every block.  bar_nodes (nullable, or
 * per block nullable): each block's node adjoint as fest_vjp's -- a
 * block's residual reads only its own nodes, so nothing crosses the
 * interfaces; a node two blocks share gets a share in each. */
int fest_group_vjp(fest_group *g, const double *const *bar_residues, double *const *bar_qps,
                   double *const *bar_nodes);
/* Absolute residual norms Res_abs(0:n_var) of the whole group, assembled
 * from every block's last residue as find_resnorm does across processes
 * (src/resnorm.f90:62-91,171-225): the blocks' signed mass imbalances and
 * sum(R_l^2)/scale_l^2 summed in block order, then |mass| and square roots. */
int fest_group_resnorm(fest_group *g, double *res_abs);
/* get_next_solution (src/update.f90:129-226) on every block, stage by stage,
 * the interface exchange opening every residual evaluation
 * (src/update.f90:534). */
int fest_group_get_next_solution(fest_group *g, int current_iter, double cfl);
/* fest_set_option's "graphs" and "check_interval" for a group's step (the
 * exchange and every block's kernels of one step form one graph), and its
 * schedule (no counterpart in the reference, whose blocks are MPI ranks):
 *   "batched"   0 (default): kernel by kernel per block; 1: every phase of a
 *               stage (ghost cells, gradients, each face sweep, the time
 *               step, the update) is one launch over the tile work list of
 *               all blocks.  Needs blocks of one kernel variant (same
 *               variables, reconstruction, model and precision; 3-D; local
 *               time steps; RK4, RK2 or none), else FEST_E_CONFIG.
 *   "grid_cap"  work-groups per batched launch, each walking the work list
 *               (a persistent launch); 0 (default): one per tile.
 *   "order"     0: each block's tiles together; 1: the blocks' tiles
 *               interleaved.
 *   "tile_i"    64 (default) or 32: the j and k face sweeps of a batched
 *               step in 64 x 4 or 32 x 8 face tiles (narrow blocks leave
 *               fewer lanes idle in the narrower tile).
 * The results do not depend on the schedule. */
int fest_group_set_option(fest_group *g, const char *name, long value);
/* Autotuner of the group step's schedule (no counterpart in the reference):
 * times `steps` (>= 3; the first two capture the graph) steps of every
 * candidate -- per-block launches, and batched with one work-group per
 * tile or a persistent 4, 8 or 16 work-groups per compute unit, tiles in
 * either order, 64 x 4 or 32 x 8 face tiles (17 in all) -- on a snapshot
 * of the state, restores the state and keeps
 * the fastest.  ms[c] receives candidate c's time per step (up to ms_len),
 * *ncand the number of candidates and *chosen the pick (nullable). */
int fest_group_autotune(fest_group *g, int steps, double *ms, int ms_len, int *ncand, int *chosen);

#ifdef __cplusplus
}
#endif
#endif
