// fest_api.hip -- host runtime and C ABI of the MI355X explicit-residual
// engine (declared in include/fest_hip.h).  One handle owns one block's
// device-resident arrays and a HIP stream; every kernel of a stage is
// launched on that stream.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <initializer_list>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

#include "fest_hip.h"
#include "blk.h"
#include "k_bc.h"
#include "k_bc_kernels.h"
#include "k_flux.h"
#include "k_geom.h"
#include "k_grad.h"
#include "k_iface.h"
#include "k_lusgs.h"
#include "k_multi.h"
#include "k_update.h"
#include "flux_launch.h"
#include "vjp_launch.h"

using namespace fest;

// kernel groups timed by fest_timing (one or a few launches each)
enum { KG_STEP_BEGIN = 0, KG_BC_FACE, KG_BC_EDGES, KG_GRAD, KG_GRAD_BC, KG_FLUX_I, KG_FLUX_J, KG_FLUX_K,
       KG_TIME_STEP, KG_UPDATE, KG_LU_DIAG, KG_LU_SWEEP, KG_LU_APPLY, KG_KKL_SOURCE, KG_VJP, KG_COUNT };
static const char *const kg_names[KG_COUNT] = {"step_begin", "bc_face", "bc_edges", "grad", "grad_bc",
                                               "flux_i", "flux_j", "flux_k", "time_step", "update",
                                               "lusgs_diag", "lusgs_sweep", "lusgs_apply", "kkl_source", "vjp"};

struct KgTimer {
  std::vector<std::pair<hipEvent_t, hipEvent_t>> pending;
  double ms = 0.0;
  long launches = 0;
};

// A time step captured from the stream into a HIP graph and replayed with
// one launch.  Kernel arguments are baked in at capture, so a graph is kept
// per inflow-fix class and re-captured when the CFL number changes or the
// arrays it was captured on move.  The first call of a class runs eagerly
// (one-time set-up such as the LU-SGS geometry copy stays out of the graph);
// the second captures.
struct StepGraphs {
  bool enabled = true;
  hipGraphExec_t exec[2] = {nullptr, nullptr};
  double cfl[2] = {0.0, 0.0};
  const void *key[2] = {nullptr, nullptr};
  int seen[2] = {0, 0};
  void reset() {
    for (int k = 0; k < 2; ++k) {
      if (exec[k]) (void)hipGraphExecDestroy(exec[k]);
      exec[k] = nullptr;
      seen[k] = 0;
      key[k] = nullptr;
    }
  }
};

struct fest_handle {
  Blk b;
  hipStream_t stream = nullptr;      // the stream every launch goes to
  hipStream_t own_stream = nullptr;  // the handle's own; a group may lend another
  struct fest_group *group = nullptr;
  std::string err;
  fest_extent dims;
  fest_scheme scheme;
  fest_flow flow;
  fest_bc bc;
  std::vector<void *> allocs;
  long red_cap = 0;        // doubles in b.red
  int norm_blocks = 0;
  bool have_grid = false, have_state = false;
  bool sk_geometry = false;  // the LU-SGS storage holds the current grid's geometry
  // opt-in event timing: groups whose bit is set record a start and a stop
  // event on the stream around their launches
  unsigned timed = 0;
  KgTimer timers[KG_COUNT];
  std::vector<hipEvent_t> ev_free;
  // the time step as captured HIP graphs, one per inflow-fix class
  // (iterations <= 2 and later, the only way current_iter enters a kernel)
  StepGraphs graphs;
  int check_every = 1;  // read the device error word after every n-th step
  int unchecked = 0;
  // fest_vjp's buffers, allocated at its first call: the residue adjoint,
  // the per-face contributions of each direction and the cell adjoint
  double *vjp_barR = nullptr, *vjp_c[3] = {}, *vjp_bar = nullptr;
  // and, for the node-coordinate adjoint: the metric adjoints, the slip
  // walls' ghost-row faces on the j and k max faces, the node adjoint
  double *vjp_m[3] = {}, *vjp_g3 = nullptr, *vjp_g5 = nullptr, *vjp_nodes = nullptr;
  // and, for a viscous block: the temperature and gradient adjoints
  double *vjp_bt = nullptr, *vjp_gb = nullptr;
};

// a group's view of its blocks; a block destroyed while in a group detaches
// itself and leaves the group unusable (its links point into the block)
// the phases of a batched group step (k_multi.h), each one launch over the
// tiles of every block
enum { PH_BEGIN = 0, PH_BC_A0, PH_BC_A1, PH_BC_A2, PH_BC_B0, PH_BC_B1, PH_BC_B2, PH_EDGE0, PH_EDGE1, PH_EDGE2,
       PH_F32, PH_GRAD, PH_GRAD_BC, PH_FLUX_I, PH_FLUX_J, PH_FLUX_K, PH_DT, PH_UPDATE, PH_COUNT };

struct fest_group {
  std::vector<fest_handle *> blocks;
  std::vector<fest_link> links;
  LinkDev *d_links = nullptr;
  LinkDev *d_vjp_links = nullptr;  // the link table on the blocks' cell adjoints (fest_group_vjp)
  int max_layer = 0;  // largest cells-per-layer of any link
  hipStream_t stream = nullptr;
  bool broken = false;
  std::string err;
  StepGraphs graphs;
  int check_every = 1;
  int unchecked = 0;
  // batched execution: the blocks' records and each phase's tile work list
  // on the device; the schedule (batched or not, work-groups per launch,
  // tile order) is set by fest_group_set_option or fest_group_autotune
  std::string batch_why;  // empty when one kernel variant fits every block
  int batched = 0;
  int grid_cap = 0;       // work-groups per batched launch; 0: one per tile
  int order = 0;          // 0: tiles block by block; 1: the blocks' tiles interleaved
  int lists_order = -1;   // order the device lists were built in
  int tile_i = TILE_I;    // i width of the j/k sweeps' face tiles: 64 (x 4) or 32 (x 8)
  int lists_tile = -1;    // tile width the device lists were built with
  Blk *d_blks = nullptr;
  double *d_norms = nullptr;  // NORM_SLOTS doubles per block: norms, then the block's share
  double blk_cfl = -1.0;  // CFL of the uploaded records
  WorkItem *d_items = nullptr;
  int ph_off[PH_COUNT] = {}, ph_n[PH_COUNT] = {};
};

// Scratch of fest_resnorm_device, the allocation-free norm entry point: the
// per-block partial sums of the norm reduction, NORM_ACC (9) doubles for
// each of at most NORM_BLOCKS_MAX blocks (72 KiB).  Each device holds its own copy until the
// process ends; host code reaches it only through the address that
// hipGetSymbolAddress returns for the current device (scratch_dev).  Every
// call on a device shares that device's copy: calls that overlap on one
// device race.
constexpr int NORM_BLOCKS_MAX = 1024;
constexpr int SCRATCH_DEVICES = 64;
static __device__ __align__(256) double scratch[NORM_ACC * NORM_BLOCKS_MAX];
static double *scratch_dev[SCRATCH_DEVICES];

namespace {

int fail(fest_handle *h, int code, const char *fmt, ...) {
  char buf[512];
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(buf, sizeof buf, fmt, ap);
  va_end(ap);
  if (h) h->err = buf;
  return code;
}

#define HIPCHK(h, call)                                                                 \
  do {                                                                                  \
    hipError_t e_ = (call);                                                             \
    if (e_ != hipSuccess) return fail(h, FEST_E_HIP, "%s: %s", #call, hipGetErrorString(e_)); \
  } while (0)

#define LAUNCHCHK(h)                                                                    \
  do {                                                                                  \
    hipError_t e_ = hipGetLastError();                                                  \
    if (e_ != hipSuccess) return fail(h, FEST_E_HIP, "launch: %s", hipGetErrorString(e_)); \
  } while (0)

int dalloc(fest_handle *h, double **p, long n) {
  void *q = nullptr;
  if (n <= 0) n = 1;
  hipError_t e = hipMalloc(&q, sizeof(double) * (size_t)n);
  if (e != hipSuccess) return fail(h, FEST_E_HIP, "hipMalloc(%ld doubles): %s", n, hipGetErrorString(e));
  h->allocs.push_back(q);
  *p = (double *)q;
  return 0;
}

hipEvent_t take_event(fest_handle *h) {
  if (!h->ev_free.empty()) {
    hipEvent_t e = h->ev_free.back();
    h->ev_free.pop_back();
    return e;
  }
  hipEvent_t e = nullptr;
  if (hipEventCreate(&e) != hipSuccess) return nullptr;
  return e;
}

// brackets the launches of one kernel group with events when it is timed
struct KgScope {
  fest_handle *h;
  int g;
  hipEvent_t e0 = nullptr;
  KgScope(fest_handle *h_, int g_) : h(h_), g(g_) {
    if (h->timed & (1u << g)) {
      e0 = take_event(h);
      if (e0 && hipEventRecord(e0, h->stream) != hipSuccess) {
        h->ev_free.push_back(e0);
        e0 = nullptr;
      }
    }
  }
  ~KgScope() {
    if (!e0) return;
    hipEvent_t e1 = take_event(h);
    if (!e1 || hipEventRecord(e1, h->stream) != hipSuccess) {
      h->ev_free.push_back(e0);
      if (e1) h->ev_free.push_back(e1);
      return;
    }
    h->timers[g].pending.emplace_back(e0, e1);
  }
};

// waits for the stream and folds the recorded event pairs into the totals
int settle_timers(fest_handle *h) {
  HIPCHK(h, hipStreamSynchronize(h->stream));
  for (auto &t : h->timers) {
    for (auto &pr : t.pending) {
      float ms = 0.f;
      HIPCHK(h, hipEventElapsedTime(&ms, pr.first, pr.second));
      t.ms += ms;
      t.launches += 1;
      h->ev_free.push_back(pr.first);
      h->ev_free.push_back(pr.second);
    }
    t.pending.clear();
  }
  return 0;
}

inline unsigned cdiv(long a, long b) { return (unsigned)((a + b - 1) / b); }

// blocks per plane, rounded up to a multiple of 8 for the XCD-aware order
inline unsigned plane_slots(long nblk, int pad8) { return pad8 ? (unsigned)((nblk + 7) / 8 * 8) : (unsigned)nblk; }

// grid over an (ni x nj x nk) index box: flat 256-thread blocks over the
// (i, j) plane, one grid row per k (the kernels use flat_ij)
inline void box(long ni, long nj, long nk, dim3 &g, dim3 &t, int pad8 = 0) {
  t = dim3(FLAT_BLOCK, 1, 1);
  g = dim3(plane_slots(cdiv(ni * nj, FLAT_BLOCK), pad8), (unsigned)nk, 1);
}

int check_err(fest_handle *h) {
  unsigned e = 0;
  HIPCHK(h, hipMemcpyAsync(&e, h->b.err, sizeof(unsigned), hipMemcpyDeviceToHost, h->stream));
  HIPCHK(h, hipStreamSynchronize(h->stream));
  if (!e) return 0;
  HIPCHK(h, hipMemsetAsync(h->b.err, 0, sizeof(unsigned), h->stream));
  if (e & ERR_NONPHYS) return fail(h, FEST_E_NONPHYSICAL, "non-physical state (rho<0, p<0 or NaN) in the update");
  if (e & ERR_NAN_FLUX) return fail(h, FEST_E_NAN, "NaN in the face fluxes");
  if (e & ERR_NAN_GRAD) return fail(h, FEST_E_NAN, "NaN in the gradients");
  if (e & ERR_NAN_MU) return fail(h, FEST_E_NAN, "NaN in the molecular viscosity");
  return fail(h, FEST_E_CONFIG, "device error bits 0x%x", e);
}

// grid of 64 x 4 tiles over an (ni x nj x nk) box (the kernels use tile_ij)
inline void tiles(long ni, long nj, long nk, dim3 &g, dim3 &t, int pad8) {
  t = dim3(TILE_I * TILE_J, 1, 1);
  g = dim3(plane_slots((long)cdiv(ni, TILE_I) * cdiv(nj, TILE_J), pad8), (unsigned)nk, 1);
}

// C2: the ghost cells of the block's own boundary rules, face by face, then
// the edges and corners
int launch_ghosts(fest_handle *h, int current_iter) {
  Blk &b = h->b;
  hipStream_t st = h->stream;
  // faces in the reference's order imin, imax, jmin, ...; the two faces of
  // one direction share a launch when the direction has at least 8 nodes:
  // each side then reads only its own interior layers 1..4 and ghost layers,
  // so neither sees the other's writes
  {
  KgScope ts(h, KG_BC_FACE);
  for (int d = 0; d < 3; ++d) {
    int N = d == 0 ? b.imx : d == 1 ? b.jmx : b.kmx;
    int na = face_na(b, 2 * d), nb = face_nb(b, 2 * d);
    dim3 t(64, 4, 1);
    if (N >= 8) {
      dim3 g(cdiv(na + 5, 64), cdiv(nb + 5, 4), 2);
      hipLaunchKernelGGL(k_bc_face, g, t, 0, st, b, 2 * d, current_iter);
    } else {
      dim3 g(cdiv(na + 5, 64), cdiv(nb + 5, 4), 1);
      hipLaunchKernelGGL(k_bc_face, g, t, 0, st, b, 2 * d, current_iter);
      hipLaunchKernelGGL(k_bc_face, g, t, 0, st, b, 2 * d + 1, current_iter);
    }
  }
  }
  {
    KgScope ts(h, KG_BC_EDGES);
    hipLaunchKernelGGL(k_bc_edges, dim3(cdiv(4L * (b.imx + 5) * b.nv, 256)), dim3(256), 0, st, b, 0);
    hipLaunchKernelGGL(k_bc_edges, dim3(cdiv(4L * (b.kmx + 5) * b.nv, 256)), dim3(256), 0, st, b, 1);
    hipLaunchKernelGGL(k_bc_edges, dim3(1), dim3(64), 0, st, b, 2);
  }
  LAUNCHCHK(h);
  return 0;
}

// C4..C6: gradients, viscosities (and turbulence sources) and the ghost
// cells' gradients of a viscous block; nothing for an inviscid one
int launch_grads(fest_handle *h) {
  Blk &b = h->b;
  hipStream_t st = h->stream;
  dim3 g, t;
  if (b.mu_ref != 0.0) {
    box(b.imx + 1, b.jmx + 1, b.kmx + 1, g, t, b.xcd_remap);
    {
      KgScope ts(h, KG_GRAD);
      if (b.lctm) {
        if (b.turb == FEST_TURB_SST2003) hipLaunchKernelGGL((k_grad<7, true>), g, t, 0, st, b);
        else hipLaunchKernelGGL((k_grad<7, false>), g, t, 0, st, b);
      } else if (b.trans) {
        if (b.turb == FEST_TURB_SST2003) hipLaunchKernelGGL((k_grad<6, true, true>), g, t, 0, st, b);
        else if (b.ng == 6) hipLaunchKernelGGL((k_grad<6, false, true>), g, t, 0, st, b);
        else hipLaunchKernelGGL((k_grad<5, false, true>), g, t, 0, st, b);
      } else if (b.g32only) {
        // FP32 mode: the gradient sums in FP32 (k_grad_b's F32 variant)
        if (b.turb == FEST_TURB_SST2003) hipLaunchKernelGGL((k_grad<6, true, false, false, true>), g, t, 0, st, b);
        else if (b.ng == 6) hipLaunchKernelGGL((k_grad<6, false, false, false, true>), g, t, 0, st, b);
        else hipLaunchKernelGGL((k_grad<4, false, false, false, true>), g, t, 0, st, b);
      } else if (b.turb == FEST_TURB_SST2003) {
        hipLaunchKernelGGL((k_grad<6, true>), g, t, 0, st, b);
      } else if (b.kkl) {
        hipLaunchKernelGGL((k_grad<6, false, false, true>), g, t, 0, st, b);
      } else if (b.ng == 6) {
        hipLaunchKernelGGL((k_grad<6, false>), g, t, 0, st, b);
      } else if (b.ng == 5) {
        hipLaunchKernelGGL((k_grad<5, false>), g, t, 0, st, b);
      } else {
        hipLaunchKernelGGL((k_grad<4, false>), g, t, 0, st, b);
      }
    }
    int mx = std::max(std::max(b.imx, b.jmx), b.kmx);
    dim3 gb(cdiv(mx, 64), cdiv(mx, 4), 6);
    {
      KgScope ts(h, KG_GRAD_BC);
      hipLaunchKernelGGL(k_grad_bc, gb, dim3(64, 4, 1), 0, st, b);
      bool tp = false;
      for (int f = 0; f < 6; ++f) tp |= b.bc[f] == -11;
      // k-kL's mu_t needs no gradients: the ghosts keep the value from the
      // ghost state
      if (tp && b.ng >= 6 && !b.kkl) hipLaunchKernelGGL(k_mut_ghost_tp, gb, dim3(64, 4, 1), 0, st, b);
    }
    if (b.kkl) {
      box(b.imx - 1, b.jmx - 1, b.kmx - 1, g, t, b.xcd_remap);
      KgScope ts(h, KG_KKL_SOURCE);
      hipLaunchKernelGGL(k_kkl_source, g, t, 0, st, b);
    }
    LAUNCHCHK(h);
  }
  return 0;
}

// C2..C9(+C10 in k_update): ghosts, gradients/viscosity, face fluxes
int launch_residual_front(fest_handle *h, int current_iter, int store_csnd) {
  Blk &b = h->b;
  hipStream_t st = h->stream;
  if (int rc = launch_ghosts(h, current_iter)) return rc;
  if (b.fp32) {
    KgScope ts(h, KG_BC_EDGES);
    hipLaunchKernelGGL(k_to_f32, dim3(cdiv(b.nc * b.nv, 256)), dim3(256), 0, st, b);
  }
  LAUNCHCHK(h);
  if (int rc = launch_grads(h)) return rc;
  dim3 g, t;
  // the i sweep: blocks of FLUX_I_SPAN consecutive padded-plane offsets
  g = dim3(plane_slots(cdiv(flux_i_span(b), FLUX_I_SPAN), b.xcd_remap), (unsigned)(b.kmx - 1), 1);
  t = dim3(FLUX_I_SPAN, 1, 1);
  {
    KgScope ts(h, KG_FLUX_I);
    launch_flux<0>(b, g, t, st, store_csnd);
  }
  tiles(b.imx - 1, b.jmx, b.kmx - 1, g, t, b.xcd_remap);
  {
    KgScope ts(h, KG_FLUX_J);
    launch_flux<1>(b, g, t, st, store_csnd);
  }
  tiles(b.imx - 1, b.jmx - 1, b.kmx, g, t, b.xcd_remap);
  {
    KgScope ts(h, KG_FLUX_K);
    launch_flux<2>(b, g, t, st, store_csnd);
  }
  LAUNCHCHK(h);
  return 0;
}

int launch_time_step(fest_handle *h) {
  Blk &b = h->b;
  dim3 g, t;
  KgScope ts(h, KG_TIME_STEP);
  if (!b.time_local && b.global_dt > 0) {
    hipLaunchKernelGGL(k_dt_fill, dim3(cdiv(b.nres, 256)), dim3(256), 0, h->stream, b, 1);
    LAUNCHCHK(h);
    return 0;
  }
  box(b.imx - 1, b.jmx - 1, b.kmx - 1, g, t);
  hipLaunchKernelGGL(k_dt, g, t, 0, h->stream, b);
  if (!b.time_local) {
    long nblk = (long)g.x * g.y * g.z;
    hipLaunchKernelGGL(k_dt_min, dim3(1), dim3(256), 0, h->stream, b, nblk);
    hipLaunchKernelGGL(k_dt_fill, dim3(cdiv(b.nres, 256)), dim3(256), 0, h->stream, b, 0);
  }
  LAUNCHCHK(h);
  return 0;
}

int launch_update(fest_handle *h, double TF, double SF, int flags) {
  Blk &b = h->b;
  dim3 g, t;
  box(b.imx - 1, b.jmx - 1, b.kmx - 1, g, t, b.xcd_remap);
  KgScope ts(h, KG_UPDATE);
  hipLaunchKernelGGL(k_update, g, t, 0, h->stream, b, TF, SF, flags);
  LAUNCHCHK(h);
  return 0;
}

// LU-SGS update (src/lusgs.f90:134-183): the hyperplane-major copies and
// the diagonal, one launch per hyperplane i + j + k = s forward and
// backward, then qp += delQ
// copies up to 8 planes into hyperplane-major components c0.. (k_to_skew)
void to_skew(fest_handle *h, std::initializer_list<const double *> planes, int c0, int layout, int lo, int hi) {
  Blk &b = h->b;
  SkewIn a{};
  a.n = 0;
  for (const double *p : planes) a.p[a.n++] = p;
  a.c0 = c0;
  a.layout = layout;
  a.lo = lo;
  a.hi = hi;
  const int ni = b.imx - 1 + hi - lo + 1, nj = b.jmx - 1 + hi - lo + 1, nk = b.kmx - 1 + hi - lo + 1;
  const dim3 g((unsigned)(cdiv(ni, 32) * cdiv(nj, 32)), (unsigned)nk);
  hipLaunchKernelGGL(k_to_skew, g, dim3(256), 0, h->stream, b, a);
}

template <int NV>
int launch_lusgs_nv(fest_handle *h) {
  using C = Sk<NV>;
  Blk &b = h->b;
  hipStream_t st = h->stream;
  dim3 g, t;
  box(b.imx - 1, b.jmx - 1, b.kmx - 1, g, t);
  {
    KgScope ts(h, KG_LU_DIAG);
    if (!h->sk_geometry) {
      // volumes and faces: once per grid
      to_skew(h, {b.vol}, C::VOL, 0, 0, 1);
      for (int d = 0; d < 3; ++d) to_skew(h, {b.fA[d], b.fx[d], b.fy[d], b.fz[d]}, C::FACE + 4 * d, 2 + d, 1, 1);
      h->sk_geometry = true;
    }
    const long nc = b.nc, nr = b.nres;
    // the NV state planes and residue planes (at most 7, one transpose each)
    if (NV == 7)
      to_skew(h, {b.qp, b.qp + nc, b.qp + 2 * nc, b.qp + 3 * nc, b.qp + 4 * nc, b.qp + 5 * nc, b.qp + 6 * nc},
              C::Q, 0, 0, 1);
    else if (NV == 6)
      to_skew(h, {b.qp, b.qp + nc, b.qp + 2 * nc, b.qp + 3 * nc, b.qp + 4 * nc, b.qp + 5 * nc}, C::Q, 0, 0, 1);
    else
      to_skew(h, {b.qp, b.qp + nc, b.qp + 2 * nc, b.qp + 3 * nc, b.qp + 4 * nc}, C::Q, 0, 0, 1);
    to_skew(h, {b.mu, b.mut, b.f1}, C::MU, 0, 0, 1);
    if (NV == 7)
      to_skew(h, {b.res, b.res + nr, b.res + 2 * nr, b.res + 3 * nr, b.res + 4 * nr, b.res + 5 * nr, b.res + 6 * nr},
              C::RES, 1, 1, 0);
    else if (NV == 6)
      to_skew(h, {b.res, b.res + nr, b.res + 2 * nr, b.res + 3 * nr, b.res + 4 * nr, b.res + 5 * nr}, C::RES, 1, 1, 0);
    else
      to_skew(h, {b.res, b.res + nr, b.res + 2 * nr, b.res + 3 * nr, b.res + 4 * nr}, C::RES, 1, 1, 0);
    // radii, Old fluxes and diagonal over the cells 1..n+1
    const dim3 gp((unsigned)(cdiv(b.imx, 16) * cdiv(b.jmx, 16)), (unsigned)b.kmx);
    hipLaunchKernelGGL(k_lusgs_prep<NV>, gp, dim3(256), 0, st, b);
  }
  LAUNCHCHK(h);
  const int smax = (b.imx - 1) + (b.jmx - 1) + (b.kmx - 1);
  // four lanes per (j, k) pair
  const dim3 gs(cdiv(4L * (b.jmx - 1) * (b.kmx - 1), 256)), ts256(256);
  {
    KgScope ts(h, KG_LU_SWEEP);
    for (int s = 3; s <= smax; ++s) hipLaunchKernelGGL((k_lusgs_sweep<NV, 0>), gs, ts256, 0, st, b, s);
    for (int s = smax; s >= 3; --s) hipLaunchKernelGGL((k_lusgs_sweep<NV, 1>), gs, ts256, 0, st, b, s);
  }
  LAUNCHCHK(h);
  {
    KgScope ts(h, KG_LU_APPLY);
    hipLaunchKernelGGL(k_lusgs_apply<NV>, g, t, 0, st, b);
  }
  LAUNCHCHK(h);
  return 0;
}
int launch_lusgs(fest_handle *h) {
  return h->b.nv == 7 ? launch_lusgs_nv<7>(h) : h->b.nv == 6 ? launch_lusgs_nv<6>(h) : launch_lusgs_nv<5>(h);
}

// One time step of get_next_solution (src/update.f90:129-226) as a sequence
// of operations: the step start (Temp refresh, U_store), residual
// evaluations, the stage-1 time step, stage updates, the TVD blends and the
// LU-SGS update.
enum { OP_BEGIN, OP_RESID, OP_DT, OP_UPDATE, OP_BLEND, OP_LUSGS };
struct StepOp {
  int kind;
  double a, c;  // TF, SF of an update; the two weights of a blend
  int flags;    // update flags; for OP_BEGIN: store U_store; for OP_RESID: store the face sound speed
};

std::vector<StepOp> step_plan(int ta) {
  std::vector<StepOp> p;
  const int S = UPD_APPLY | UPD_STORE;
  switch (ta) {
    case FEST_TA_RK4:  // :176-187
      p = {{OP_BEGIN, 0, 0, 1}, {OP_RESID, 0, 0, 1}, {OP_DT, 0, 0, 0}, {OP_UPDATE, 0.5, 1., S | UPD_FIRST},
           {OP_RESID, 0, 0, 0}, {OP_UPDATE, 0.5, 2., S}, {OP_RESID, 0, 0, 0}, {OP_UPDATE, 1.0, 2., S},
           {OP_RESID, 0, 0, 0}, {OP_UPDATE, 1. / 6., 1., S | UPD_LAST | UPD_WRITE_RES}};
      break;
    case FEST_TA_RK2:  // :188-195
      p = {{OP_BEGIN, 0, 0, 1}, {OP_RESID, 0, 0, 1}, {OP_DT, 0, 0, 0}, {OP_UPDATE, 0.5, 1., S | UPD_FIRST},
           {OP_RESID, 0, 0, 0}, {OP_UPDATE, 0.5, 1., S | UPD_LAST | UPD_WRITE_RES}};
      break;
    case FEST_TA_TVDRK3:  // :196-206
      p = {{OP_BEGIN, 0, 0, 1}, {OP_RESID, 0, 0, 1}, {OP_DT, 0, 0, 0}, {OP_UPDATE, 1., 1., UPD_APPLY},
           {OP_RESID, 0, 0, 0}, {OP_UPDATE, 1., 1., UPD_APPLY}, {OP_BLEND, 0.75, 0.25, 0},
           {OP_RESID, 0, 0, 0}, {OP_UPDATE, 1., 1., UPD_APPLY | UPD_WRITE_RES}, {OP_BLEND, 1. / 3., 2. / 3., 0}};
      break;
    case FEST_TA_TVDRK2:  // :207-214
      p = {{OP_BEGIN, 0, 0, 1}, {OP_RESID, 0, 0, 1}, {OP_DT, 0, 0, 0}, {OP_UPDATE, 1., 1., UPD_APPLY},
           {OP_RESID, 0, 0, 0}, {OP_UPDATE, 1., 1., UPD_APPLY | UPD_WRITE_RES}, {OP_BLEND, 0.5, 0.5, 0}};
      break;
    case FEST_TA_IMPLICIT:  // :215-218; the update pass writes the residue only
      p = {{OP_BEGIN, 0, 0, 0}, {OP_RESID, 0, 0, 1}, {OP_DT, 0, 0, 0}, {OP_UPDATE, 1., 1., UPD_WRITE_RES},
           {OP_LUSGS, 0, 0, 0}};
      break;
    default:  // "none", :171-174
      p = {{OP_BEGIN, 0, 0, 0}, {OP_RESID, 0, 0, 1}, {OP_DT, 0, 0, 0},
           {OP_UPDATE, 1., 1., UPD_APPLY | UPD_WRITE_RES}};
      break;
  }
  return p;
}

// one operation on one block; residual evaluations of a group are opened by
// the group's exchange (run_group_op)
int run_op(fest_handle *h, const StepOp &op, int current_iter) {
  Blk &b = h->b;
  switch (op.kind) {
    case OP_BEGIN: {
      KgScope ts(h, KG_STEP_BEGIN);
      hipLaunchKernelGGL(k_step_begin, dim3(cdiv(b.nc, 256)), dim3(256), 0, h->stream, b, op.flags);
      LAUNCHCHK(h);
      return 0;
    }
    case OP_RESID: return launch_residual_front(h, current_iter, op.flags);
    case OP_DT: return launch_time_step(h);
    case OP_UPDATE: return launch_update(h, op.a, op.c, op.flags);
    case OP_LUSGS: return launch_lusgs(h);
    default: {
      KgScope ts(h, KG_UPDATE);
      hipLaunchKernelGGL(k_blend, dim3(cdiv(b.nc * b.nv, 256)), dim3(256), 0, h->stream, b, op.a, op.c);
      LAUNCHCHK(h);
      return 0;
    }
  }
}

// Runs `step` (which only launches kernels on `st`) through the graph of
// the inflow-fix class of current_iter, capturing it on the class's second
// call; eagerly when graphs are off.  key identifies the arrays and
// constants the kernels were captured with.
template <typename F>
int run_graphed(StepGraphs &G, hipStream_t st, int current_iter, double cfl, const void *key, bool eager, F step,
                std::string &err) {
  const int k = current_iter <= 2 ? 0 : 1;
  if (eager || !G.enabled) return step();
  if (G.exec[k] && (G.cfl[k] != cfl || G.key[k] != key)) {
    (void)hipGraphExecDestroy(G.exec[k]);
    G.exec[k] = nullptr;
    G.seen[k] = 0;
  }
  if (!G.exec[k]) {
    if (G.seen[k]++ == 0) return step();
    hipError_t e = hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal);
    if (e != hipSuccess) {
      err = std::string("hipStreamBeginCapture: ") + hipGetErrorString(e);
      return FEST_E_HIP;
    }
    int rc = step();
    hipGraph_t graph = nullptr;
    e = hipStreamEndCapture(st, &graph);
    if (rc) {
      if (graph) (void)hipGraphDestroy(graph);
      return rc;
    }
    if (e != hipSuccess) {
      err = std::string("hipStreamEndCapture: ") + hipGetErrorString(e);
      return FEST_E_HIP;
    }
    e = hipGraphInstantiate(&G.exec[k], graph, nullptr, nullptr, 0);
    (void)hipGraphDestroy(graph);
    if (e != hipSuccess) {
      G.exec[k] = nullptr;
      err = std::string("hipGraphInstantiate: ") + hipGetErrorString(e);
      return FEST_E_HIP;
    }
    G.cfl[k] = cfl;
    G.key[k] = key;
  }
  hipError_t e = hipGraphLaunch(G.exec[k], st);
  if (e != hipSuccess) {
    err = std::string("hipGraphLaunch: ") + hipGetErrorString(e);
    return FEST_E_HIP;
  }
  return 0;
}

// surface nodes of the no-slip faces of one block, in the order of
// src/wall/wall.f90:198-288, rounded like the ES18.10E3 surface-node file
// the reference writes and reads back (src/wall/wall.f90:101)
int wall_nodes(fest_handle *h, std::vector<double> &wall) {
  Blk &b = h->b;
  long nn = (long)(b.imx + 6) * (b.jmx + 6) * (b.kmx + 6);
  std::vector<double> hn(3 * nn);
  HIPCHK(h, hipMemcpyAsync(hn.data(), b.nodes, sizeof(double) * 3 * nn, hipMemcpyDeviceToHost, h->stream));
  HIPCHK(h, hipStreamSynchronize(h->stream));
  for (int f = 0; f < 6; ++f) {
    if (b.bc[f] != -5) continue;
    int im = 1, ix = b.imx, jm = 1, jx = b.jmx, km = 1, kx = b.kmx;
    switch (f) {
      case 0: ix = 1; break;
      case 1: im = b.imx; break;
      case 2: jx = 1; break;
      case 3: jm = b.jmx; break;
      case 4: kx = 1; break;
      default: km = b.kmx; break;
    }
    for (int k = km; k <= kx; ++k)
      for (int j = jm; j <= jx; ++j)
        for (int i = im; i <= ix; ++i)
          for (int c = 0; c < 3; ++c) {
            char buf[64];
            snprintf(buf, sizeof buf, "%.10E", hn[c * nn + ndi(b, i, j, k)]);
            wall.push_back(strtod(buf, nullptr));
          }
  }
  return 0;
}

// lctm2015's wall-normal sensor from the current wall distance and metrics
// (src/CC.f90 setupCC after find_wall_dist, solver.f90:131-135)
int lctm_sensor(fest_handle *h) {
  Blk &b = h->b;
  if (!b.lctm) return 0;
  dim3 g(cdiv(b.imx - 1, 64), cdiv(b.jmx - 1, 4), (unsigned)(b.kmx - 1));
  hipLaunchKernelGGL(k_lctm_sensor, g, dim3(64, 4, 1), 0, h->stream, b);
  LAUNCHCHK(h);
  return 0;
}

// cell distances of one block to the surface nodes `wall`
// (src/wall/wall_dist.f90:84-131)
int dist_from_walls(fest_handle *h, const std::vector<double> &wall) {
  Blk &b = h->b;
  long nn = (long)(b.imx + 6) * (b.jmx + 6) * (b.kmx + 6);
  int nw = (int)(wall.size() / 3);
  double *dwall = nullptr, *nd = nullptr;
  hipError_t le = hipMalloc(&dwall, sizeof(double) * std::max<size_t>(3, wall.size()));
  if (le == hipSuccess) le = hipMalloc(&nd, sizeof(double) * nn);
  if (le == hipSuccess && nw)
    le = hipMemcpyAsync(dwall, wall.data(), sizeof(double) * wall.size(), hipMemcpyHostToDevice, h->stream);
  if (le == hipSuccess) {
    hipLaunchKernelGGL(k_node_dist, dim3(cdiv(nn, 256)), dim3(256), 0, h->stream, b, (const double *)dwall, nw, nd);
    dim3 g(cdiv(b.imx + 5, 64), cdiv(b.jmx + 5, 4), (unsigned)(b.kmx + 5));
    hipLaunchKernelGGL(k_cell_dist, g, dim3(64, 4, 1), 0, h->stream, b, (const double *)nd);
    if (b.lctm) {
      dim3 gs(cdiv(b.imx - 1, 64), cdiv(b.jmx - 1, 4), (unsigned)(b.kmx - 1));
      hipLaunchKernelGGL(k_lctm_sensor, gs, dim3(64, 4, 1), 0, h->stream, b);
    }
    le = hipGetLastError();
  }
  hipError_t se = hipStreamSynchronize(h->stream);
  if (le == hipSuccess) le = se;
  if (dwall) (void)hipFree(dwall);
  if (nd) (void)hipFree(nd);
  if (le != hipSuccess) return fail(h, FEST_E_HIP, "wall distance: %s", hipGetErrorString(le));
  return 0;
}

}  // namespace

extern "C" {

int fest_create(fest_handle **out, const fest_extent *dims, const fest_scheme *scheme, const fest_flow *flow,
                const fest_bc *bc, int precision) {
  if (!out || !dims || !scheme || !flow || !bc) return FEST_E_ARG;
  *out = nullptr;
  fest_handle *h = new fest_handle();
  h->dims = *dims; h->scheme = *scheme; h->flow = *flow; h->bc = *bc;
  auto bad = [&](const char *msg) { h->err = msg; *out = h; return FEST_E_CONFIG; };
  if (precision != FEST_PREC_FP64 && precision != FEST_PREC_FP32_FLUX)
    return bad("precision must be 0 (FP64) or 1 (FP32 face arithmetic)");
  // a 2-D grid has one cell in k (kmx == 2), src/face/flux/convective/*.f90
  if (dims->imx < 4 || dims->jmx < 4 || (dims->kmx < 4 && dims->kmx != 2))
    return bad("i and j need at least 3 cells, k at least 3 or exactly 1");
  if (scheme->turb != FEST_TURB_NONE && scheme->turb != FEST_TURB_SST && scheme->turb != FEST_TURB_SST2003 &&
      scheme->turb != FEST_TURB_SA && scheme->turb != FEST_TURB_KKL)
    return bad("turbulence model not supported");
  if (scheme->turb == FEST_TURB_KKL && scheme->transition != FEST_TRANS_NONE)
    return bad("k-kL has no transition model (src/source.f90:147-150)");
  int nv_expect = scheme->turb == FEST_TURB_SA ? 6 : scheme->turb != FEST_TURB_NONE ? 7 : 5;
  if (scheme->transition != FEST_TRANS_NONE && scheme->transition != FEST_TRANS_BC &&
      scheme->transition != FEST_TRANS_LCTM2015)
    return bad("transition model not supported (none, bc, lctm2015)");
  if (scheme->transition == FEST_TRANS_BC && (scheme->turb == FEST_TURB_NONE || flow->mu_ref == 0.0))
    return bad("the bc transition needs a turbulence model and mu_ref != 0");
  if (scheme->transition == FEST_TRANS_LCTM2015) {
    // src/source.f90:124-139: the lctm2015 source exists for SST only
    if (scheme->turb != FEST_TURB_SST && scheme->turb != FEST_TURB_SST2003)
      return bad("lctm2015 needs the SST or SST-2003 model (src/source.f90:124-139)");
    if (scheme->time_accuracy == FEST_TA_IMPLICIT)
      return bad("lctm2015 with LU-SGS (update_lctm2015, src/lusgs.f90:2262) is not built");
    if (precision != FEST_PREC_FP64) return bad("lctm2015 runs in FP64 only");
    nv_expect = 8;  // src/state.f90:309-317
  }
  if (dims->n_var != nv_expect) return bad("n_var does not match the turbulence model");
  if (scheme->flux < 0 || scheme->flux > 5) return bad("unknown flux scheme");
  if (scheme->recon < 0 || scheme->recon > 4) return bad("unknown reconstruction");
  if (scheme->muscl_limiter != FEST_LIM_KOREN && scheme->muscl_limiter != FEST_LIM_VAN_ALBADA)
    return bad("unknown MUSCL limiter");
  for (int d = 0; d < 3; ++d)
    if (scheme->pbs[d] != 0 && scheme->pbs[d] != 1) return bad("pressure-based switch must be 0 or 1");
  if (scheme->time_accuracy != FEST_TA_RK4 && scheme->time_accuracy != FEST_TA_RK2 &&
      scheme->time_accuracy != FEST_TA_NONE && scheme->time_accuracy != FEST_TA_TVDRK2 &&
      scheme->time_accuracy != FEST_TA_TVDRK3 && scheme->time_accuracy != FEST_TA_IMPLICIT)
    return bad("time integration not supported (RK4, RK2, TVDRK2, TVDRK3, none, implicit)");
  for (int f = 0; f < 6; ++f) {
    int id = bc->id[f];
    // ids >= 0 (interfaces) and -10 (cross-block periodic) take their ghost
    // layers from a group's interface exchange (fest_group_create)
    if (id < -11) return bad("unknown boundary id");
    // a k pole on a one-cell k extent would take its normal from the other
    // k face's pole layers
    if (id == -7 && f >= 4 && dims->kmx == 2) return bad("a pole needs at least two cells across it");
  }
  if (flow->mu_ref == 0.0 && scheme->turb) return bad("turbulence needs mu_ref != 0");
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || ndev < 1) {
    h->err = "no HIP device";
    *out = h;
    return FEST_E_HIP;
  }

  {
    // the device address of the norm scratch, once per device
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= SCRATCH_DEVICES) {
      h->err = "hipGetDevice failed or device index out of range";
      *out = h;
      return FEST_E_HIP;
    }
    if (!scratch_dev[dev]) {
      void *p = nullptr;
      hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
      if (e != hipSuccess || !p) {
        h->err = std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(e);
        *out = h;
        return FEST_E_HIP;
      }
      scratch_dev[dev] = (double *)p;
    }
  }

  Blk &b = h->b;
  memset(&b, 0, sizeof b);
  b.imx = dims->imx; b.jmx = dims->jmx; b.kmx = dims->kmx; b.nv = dims->n_var;
  b.ng = scheme->turb == FEST_TURB_SA ? 5 : scheme->turb ? 6 : 4;
  b.lctm = scheme->transition == FEST_TRANS_LCTM2015;
  if (b.lctm) b.ng = 7;  // the intermittency slot (src/gradients.f90:262-273)
  b.pi = b.imx + 5; b.pj = b.jmx + 5; b.pk = b.kmx + 5;
  b.nc = (long)b.pi * b.pj * b.pk;
  b.fpi[0] = b.imx + 6; b.fpj[0] = b.jmx + 5; b.fpk[0] = b.kmx + 5;
  b.fpi[1] = b.imx + 5; b.fpj[1] = b.jmx + 6; b.fpk[1] = b.kmx + 5;
  b.fpi[2] = b.imx + 5; b.fpj[2] = b.jmx + 5; b.fpk[2] = b.kmx + 6;
  for (int d = 0; d < 3; ++d) b.nfc[d] = (long)b.fpi[d] * b.fpj[d] * b.fpk[d];
  b.nfi[0] = b.imx; b.nfj[0] = b.jmx - 1; b.nfk[0] = b.kmx - 1;
  b.nfi[1] = b.imx - 1; b.nfj[1] = b.jmx; b.nfk[1] = b.kmx - 1;
  b.nfi[2] = b.imx - 1; b.nfj[2] = b.jmx - 1; b.nfk[2] = b.kmx;
  for (int d = 0; d < 3; ++d) b.nfl[d] = (long)b.nfi[d] * b.nfj[d] * b.nfk[d];
  b.ngr = (long)(b.imx + 1) * (b.jmx + 1) * (b.kmx + 1);
  b.nres = (long)(b.imx - 1) * (b.jmx - 1) * (b.kmx - 1);
  b.flux_scheme = scheme->flux; b.recon = scheme->recon; b.turb = scheme->turb;
  b.time_local = scheme->time_local; b.time_acc = scheme->time_accuracy; b.accur = scheme->accur;
  for (int d = 0; d < 3; ++d) { b.lim[d] = scheme->lim[d]; b.tlim[d] = scheme->tlim[d]; }
  // the switch acts under MUSCL and PPM only (compute_face_interpolant,
  // src/face/state/face_interpolant.f90:91-104)
  for (int d = 0; d < 3; ++d)
    b.pbs[d] = (scheme->recon == FEST_RECON_MUSCL || scheme->recon == FEST_RECON_PPM) ? scheme->pbs[d] : 0;
  b.valb = scheme->muscl_limiter == FEST_LIM_VAN_ALBADA;
  // src/boundary/boundary_state_reconstruction.f90:45-47
  b.ppm_flag = (scheme->recon == FEST_RECON_PPM || scheme->recon == FEST_RECON_WENO ||
                scheme->recon == FEST_RECON_WENO_NM) ? 1 : 0;
  for (int f = 0; f < 6; ++f)
    if (bc->id[f] == -7) b.ppm_flag = 1;
  {
    // XCD-aware block order of the sweep kernels; FEST_XCD_REMAP=0 switches
    // it off (a tuning switch: results do not depend on it)
    const char *env = getenv("FEST_XCD_REMAP");
    b.xcd_remap = (env && env[0] == '0') ? 0 : 1;
  }
  for (int f = 0; f < 6; ++f) b.bc[f] = bc->id[f];
  // make_*_flux_zero at the first/last face, src/boundary/bc.f90:60-66
  for (int d = 0; d < 3; ++d)
    for (int s = 0; s < 2; ++s) {
      int id = bc->id[2 * d + s];
      b.mz[d][s] = (id == -5 || id == -6 || id == -7) ? 0 : 1;
    }
  b.mu_suth = flow->mu_sutherland;
  b.gm = flow->gm; b.R = flow->R_gas; b.mu_ref = flow->mu_ref; b.T_ref = flow->T_ref;
  b.S_ref = flow->Sutherland_temp; b.Pr = flow->Pr; b.tPr = flow->tPr; b.MInf = flow->MInf;
  b.iPr = 1.0 / flow->Pr; b.itPr = 1.0 / flow->tPr; b.cpR = flow->gm * flow->R_gas / (flow->gm - 1.0);
  b.vel_mag = flow->vel_mag;
  b.rho_inf = flow->density_inf; b.u_inf = flow->x_speed_inf; b.v_inf = flow->y_speed_inf;
  b.w_inf = flow->z_speed_inf; b.p_inf = flow->pressure_inf; b.tk_inf = flow->tk_inf; b.tw_inf = flow->tw_inf;
  b.tv_inf = flow->tv_inf;
  b.kkl = scheme->turb == FEST_TURB_KKL;
  b.sigma_w2 = b.kkl ? 1.0 : 0.856;
  b.tkl_inf = flow->tkl_inf;
  b.trans = scheme->transition == FEST_TRANS_BC;
  b.fp32 = precision == FEST_PREC_FP32_FLUX;
  // FP32 mode keeps only FP32 gradients where nothing but the ghost rule
  // and the face sweeps read them (no SA/k-kL source or LU-SGS diagonal)
  b.g32only = b.fp32 && scheme->time_accuracy != FEST_TA_IMPLICIT &&
              (scheme->turb == FEST_TURB_NONE || scheme->turb == FEST_TURB_SST ||
               scheme->turb == FEST_TURB_SST2003);
  b.tu_inf = flow->tu_inf;
  b.tgm_inf = flow->tgm_inf;
  b.Re = flow->density_inf * flow->vel_mag * 1.0 / flow->mu_ref;  // state.f90:89
  b.c2 = 1 + scheme->accur;
  b.c3 = 0.5 * scheme->accur;
  b.c1 = b.c2 - b.c3;
  for (int f = 0; f < 6; ++f) {
    b.fix[FX_RHO][f] = bc->fixed_density[f]; b.fix[FX_U][f] = bc->fixed_x_speed[f];
    b.fix[FX_V][f] = bc->fixed_y_speed[f]; b.fix[FX_W][f] = bc->fixed_z_speed[f];
    b.fix[FX_P][f] = bc->fixed_pressure[f]; b.fix[FX_TK][f] = bc->fixed_tk[f];
    b.fix[FX_TW][f] = bc->fixed_tw[f]; b.fix[FX_TWALL][f] = bc->fixed_wall_temperature[f];
    b.fix[FX_TP][f] = bc->fixed_Tpressure[f];
    b.fix[FX_TV][f] = bc->fixed_tv[f];
    b.fix[FX_TKL][f] = bc->fixed_tkl[f];
    b.fix[FX_TGM][f] = bc->fixed_tgm[f];
  }
  b.global_dt = scheme->global_time_step;
  b.CFLv = 1.0;

  *out = h;
  int rc = 0;
  if (hipStreamCreate(&h->own_stream) != hipSuccess) return fail(h, FEST_E_HIP, "hipStreamCreate failed");
  h->stream = h->own_stream;
#define AL(p, n) if ((rc = dalloc(h, &(p), (n)))) return rc
  AL(b.qp, b.nc * b.nv); AL(b.temp, b.nc); AL(b.mu, b.nc); AL(b.mut, b.nc); AL(b.f1, b.nc); AL(b.dist, b.nc);
  AL(b.vol, b.nc); AL(b.cx, b.nc); AL(b.cy, b.nc); AL(b.cz, b.nc);
  for (int d = 0; d < 3; ++d) {
    AL(b.fA[d], b.nfc[d]); AL(b.fx[d], b.nfc[d]); AL(b.fy[d], b.nfc[d]); AL(b.fz[d], b.nfc[d]);
    AL(b.flux[d], b.nfl[d] * b.nv); AL(b.csnd[d], b.nfl[d]);
    AL(b.g[d], b.ngr * b.ng);
  }
  if (b.fp32) {
    // FP32 copies: (nc * nv) state values and 3 x (ngr * ng) gradients
    double *q32 = nullptr, *g32 = nullptr;
    AL(q32, (b.nc * b.nv + 1) / 2);
    AL(g32, (3 * b.ngr * b.ng + 1) / 2);
    b.q32 = (float *)q32;
    for (int d = 0; d < 3; ++d) b.g32[d] = (float *)g32 + (long)d * b.ngr * b.ng;
  }
  AL(b.res, b.nres * b.nv); AL(b.dt, b.nres); AL(b.src, (b.lctm ? 3 : 2) * b.nres);
  if (b.lctm) AL(b.dvdy, b.nc); AL(b.ust, b.nc * b.nv); AL(b.rst, b.nres * b.nv);
  AL(b.nodes, 3L * (b.imx + 6) * (b.jmx + 6) * (b.kmx + 6));
  // reduction partials: the larger of the time-step block count and 8 per norm block
  long nblk_dt = (long)cdiv((long)(b.imx - 1) * (b.jmx - 1), FLAT_BLOCK) * (b.kmx - 1);
  h->norm_blocks = (int)std::min<long>(1024, std::max<long>(1, cdiv(b.nres, 256)));
  h->red_cap = std::max<long>(nblk_dt, (long)NORM_ACC * h->norm_blocks);
  AL(b.red, h->red_cap);
  AL(b.norms, NORM_SLOTS); AL(b.dtmin, 1);
  if (b.time_acc == FEST_TA_IMPLICIT) {
    // LU-SGS storage over the hyperplanes of (0:imx, 0:jmx, 0:kmx); zero
    // where nothing is written (the sweep values outside the interior)
    b.sk_j = b.jmx + 1;
    b.sk_p = (long)b.sk_j * (b.kmx + 1);
    b.sk_n = (long)(b.imx + b.jmx + b.kmx + 1) * b.sk_p;
    const long nsk = sk_ncomp(b.nv) * b.sk_n;
    AL(b.sk, nsk);
    HIPCHK(h, hipMemsetAsync(b.sk, 0, sizeof(double) * nsk, h->stream));
  }
#undef AL
  void *e = nullptr;
  if (hipMalloc(&e, sizeof(unsigned)) != hipSuccess) return fail(h, FEST_E_HIP, "hipMalloc(err)");
  h->allocs.push_back(e);
  b.err = (unsigned *)e;
  HIPCHK(h, hipMemsetAsync(b.err, 0, sizeof(unsigned), h->stream));
  // molecular viscosity starts at mu_ref (src/viscosity.f90:565-568); F1 at 0
  hipLaunchKernelGGL(k_fill, dim3(cdiv(b.nc, 256)), dim3(256), 0, h->stream, b.mu, b.nc, b.mu_ref);
  hipLaunchKernelGGL(k_fill, dim3(cdiv(b.nc, 256)), dim3(256), 0, h->stream, b.f1, b.nc, 0.0);
  hipLaunchKernelGGL(k_fill, dim3(cdiv(b.nc, 256)), dim3(256), 0, h->stream, b.mut, b.nc, 0.0);
  hipLaunchKernelGGL(k_fill, dim3(cdiv(b.nc, 256)), dim3(256), 0, h->stream, b.dist, b.nc, 1.0);
  LAUNCHCHK(h);
  HIPCHK(h, hipStreamSynchronize(h->stream));
  return 0;
}

void fest_destroy(fest_handle *h) {
  if (!h) return;
  if (h->stream) (void)hipStreamSynchronize(h->stream);
  if (h->group) {
    auto &v = h->group->blocks;
    v.erase(std::remove(v.begin(), v.end(), h), v.end());
    h->group->broken = true;
    h->group->err = "a block of the group was destroyed";
  }
  for (auto &t : h->timers)
    for (auto &pr : t.pending) {
      (void)hipEventDestroy(pr.first);
      (void)hipEventDestroy(pr.second);
    }
  for (hipEvent_t e : h->ev_free) (void)hipEventDestroy(e);
  h->graphs.reset();
  for (void *p : h->allocs) (void)hipFree(p);
  if (h->own_stream) (void)hipStreamDestroy(h->own_stream);
  delete h;
}

const char *fest_last_error(const fest_handle *h) { return h ? h->err.c_str() : "null handle"; }

int fest_upload_geometry(fest_handle *h, const double *cells4, const double *if4, const double *jf4,
                         const double *kf4, const double *dist) {
  if (!h || !cells4 || !if4 || !jf4 || !kf4) return FEST_E_ARG;
  Blk &b = h->b;
  size_t nc = sizeof(double) * b.nc;
  double *cd[4] = {b.vol, b.cx, b.cy, b.cz};
  for (int a = 0; a < 4; ++a) HIPCHK(h, hipMemcpyAsync(cd[a], cells4 + a * b.nc, nc, hipMemcpyHostToDevice, h->stream));
  const double *src[3] = {if4, jf4, kf4};
  for (int d = 0; d < 3; ++d) {
    size_t n = sizeof(double) * b.nfc[d];
    double *fd[4] = {b.fA[d], b.fx[d], b.fy[d], b.fz[d]};
    for (int a = 0; a < 4; ++a)
      HIPCHK(h, hipMemcpyAsync(fd[a], src[d] + a * b.nfc[d], n, hipMemcpyHostToDevice, h->stream));
  }
  if (dist) HIPCHK(h, hipMemcpyAsync(b.dist, dist, nc, hipMemcpyHostToDevice, h->stream));
  if (int rc = lctm_sensor(h)) return rc;
  HIPCHK(h, hipStreamSynchronize(h->stream));
  h->have_grid = true;
  h->sk_geometry = false;
  h->graphs.reset();
  return 0;
}

int fest_set_grid(fest_handle *h, const double *nodes, int with_wall_dist) {
  if (!h || !nodes) return FEST_E_ARG;
  Blk &b = h->b;
  long nn = (long)(b.imx + 6) * (b.jmx + 6) * (b.kmx + 6);
  HIPCHK(h, hipMemcpyAsync(b.nodes, nodes, sizeof(double) * 3 * nn, hipMemcpyHostToDevice, h->stream));
  for (int d = 0; d < 3; ++d) {
    int na = d == 0 ? b.jmx : b.imx, nb = d == 2 ? b.jmx : b.kmx;
    dim3 g(cdiv(na + 6, 64), cdiv(nb + 6, 4), 3);
    hipLaunchKernelGGL(k_ghost_nodes, g, dim3(64, 4, 1), 0, h->stream, b, d);
  }
  for (int d = 0; d < 3; ++d) {
    dim3 g(cdiv(b.imx + 5 + (d == 0), 64), cdiv(b.jmx + 5 + (d == 1), 4), (unsigned)(b.kmx + 5 + (d == 2)));
    hipLaunchKernelGGL(k_face_metrics, g, dim3(64, 4, 1), 0, h->stream, b, d);
  }
  for (int f = 0; f < 6; ++f) {
    if (b.bc[f] != -7) continue;
    int d = f >> 1;
    int na = (d == 0 ? b.jmx : b.imx) + 5, nb = (d == 2 ? b.jmx : b.kmx) + 5;
    hipLaunchKernelGGL(k_pole_faces, dim3(cdiv(na, 64), cdiv(nb, 4), 4), dim3(64, 4, 1), 0, h->stream, b, f);
  }
  {
    dim3 g(cdiv(b.imx + 5, 64), cdiv(b.jmx + 5, 4), (unsigned)(b.kmx + 5));
    hipLaunchKernelGGL(k_cells, g, dim3(64, 4, 1), 0, h->stream, b);
  }
  LAUNCHCHK(h);
  if (int rc = lctm_sensor(h)) return rc;  // with the distance as it stands; again below with the walls'
  if (with_wall_dist) {
    std::vector<double> wall;
    int rc = wall_nodes(h, wall);
    if (rc || (rc = dist_from_walls(h, wall))) return rc;
  }
  int rc = check_err(h);
  if (rc) return fail(h, FEST_E_NONPHYSICAL, "non-positive cell volume in the grid");
  h->have_grid = true;
  h->sk_geometry = false;
  h->graphs.reset();
  return 0;
}

int fest_upload_state(fest_handle *h, const double *qp, const double *temp) {
  if (!h || !qp) return FEST_E_ARG;
  Blk &b = h->b;
  HIPCHK(h, hipMemcpyAsync(b.qp, qp, sizeof(double) * b.nc * b.nv, hipMemcpyHostToDevice, h->stream));
  if (temp) HIPCHK(h, hipMemcpyAsync(b.temp, temp, sizeof(double) * b.nc, hipMemcpyHostToDevice, h->stream));
  else {
    hipLaunchKernelGGL(k_step_begin, dim3(cdiv(b.nc, 256)), dim3(256), 0, h->stream, b, 0);
    LAUNCHCHK(h);
  }
  HIPCHK(h, hipStreamSynchronize(h->stream));
  h->have_state = true;
  return 0;
}

int fest_residual(fest_handle *h, int current_iter) {
  if (!h) return FEST_E_ARG;
  if (!h->have_grid || !h->have_state) return fail(h, FEST_E_ARG, "grid and state must be set first");
  int rc = launch_residual_front(h, current_iter, 1);
  if (rc) return rc;
  if ((rc = launch_update(h, 0.0, 0.0, UPD_WRITE_RES))) return rc;
  return check_err(h);
}

// why fest_vjp cannot differentiate the block's residual (empty: it can)
extern "C++" std::string vjp_reason(const Blk &b) {
  if (b.nv != 5) return "only the five flow variables are differentiated (no turbulence model)";
  if (b.mu_ref != 0.0) {
    if (b.kmx == 2) return "the viscous terms are differentiated on 3-D grids only";
    if (b.fp32) return "the viscous terms are differentiated on FP64 handles only";
    for (int f = 0; f < 6; ++f)
      if (b.bc[f] >= 0 || b.bc[f] == -10) return "the viscous terms are differentiated without interface faces";
  }
  // the ghost rules read interior layers 1..4 only with 4 interior cells
  if (b.imx < 5 || b.jmx < 5 || (b.kmx != 2 && b.kmx < 5))
    return "every 3-D direction needs at least 4 cells";
  for (int f = 0; f < 6; ++f)
    if (b.bc[f] == -5 && !(b.fix[FX_TWALL][f] >= 0.0 && b.fix[FX_TWALL][f] <= 1.0))
      return "no-slip walls with an adiabatic or isothermal density rule are not differentiated";
  return "";
}

// fest_vjp's state buffers, allocated at the first call
int vjp_buffers(fest_handle *h) {
  const Blk &b = h->b;
  int rc;
  if (h->vjp_bar) return 0;
  if ((rc = dalloc(h, &h->vjp_barR, 5 * b.nres))) return rc;
  for (int d = 0; d < 3; ++d)
    if ((rc = dalloc(h, &h->vjp_c[d], vjp_contrib_len(b, d)))) return rc;
  if (b.mu_ref != 0.0) {
    if ((rc = dalloc(h, &h->vjp_bt, b.nc))) return rc;
    if ((rc = dalloc(h, &h->vjp_gb, 12 * b.ngr))) return rc;
  }
  return dalloc(h, &h->vjp_bar, 5 * b.nc);
}

// why the node adjoint is not available on top of the state adjoint
extern "C++" std::string vjp_nodes_reason(const Blk &b) {
  // the pole rewrites its faces' metrics, WENO-NM reads cell volumes
  for (int f = 0; f < 6; ++f)
    if (b.bc[f] == -7) return "the node-coordinate adjoint does not cover the pole (-7)";
  if (b.recon == FEST_RECON_WENO_NM) return "the node-coordinate adjoint does not cover WENO-NM";
  if (b.mu_ref != 0.0) return "the node-coordinate adjoint covers the inviscid residual";
  return "";
}

// the node adjoint's buffers (allocated at the first use) in `nodes`
int vjp_node_buffers(fest_handle *h, VjpNodes &nodes) {
  const Blk &b = h->b;
  const long nn = (long)(b.imx + 6) * (b.jmx + 6) * (b.kmx + 6);
  int rc;
  if (!h->vjp_nodes) {
    for (int d = 0; d < 3; ++d)
      if ((rc = dalloc(h, &h->vjp_m[d], 4 * b.nfl[d]))) return rc;
    if ((rc = dalloc(h, &h->vjp_g3, 6L * (b.imx - 1) * (b.kmx - 1)))) return rc;
    if ((rc = dalloc(h, &h->vjp_g5, 6L * (b.imx - 1) * (b.jmx - 1)))) return rc;
    if ((rc = dalloc(h, &h->vjp_nodes, 3 * nn))) return rc;
  }
  for (int d = 0; d < 3; ++d) nodes.mbar[d] = h->vjp_m[d];
  // slip walls on the j and k max faces reflect with ghost-row I faces
  nodes.g3 = b.bc[3] == -6 ? h->vjp_g3 : nullptr;
  nodes.g5 = b.bc[5] == -6 && b.kmx != 2 ? h->vjp_g5 : nullptr;
  nodes.bar_nodes = h->vjp_nodes;
  return 0;
}

int fest_vjp(fest_handle *h, const double *bar_residue, double *bar_qp, double *bar_nodes) {
  if (!h || !bar_residue || !bar_qp) return FEST_E_ARG;
  if (!h->have_grid || !h->have_state) return fail(h, FEST_E_ARG, "grid and state must be set first");
  Blk &b = h->b;
  std::string why = vjp_reason(b);
  if (why.empty() && bar_nodes) why = vjp_nodes_reason(b);
  if (!why.empty()) return fail(h, FEST_E_CONFIG, "fest_vjp: %s", why.c_str());
  hipStream_t st = h->stream;
  int rc;
  if ((rc = vjp_buffers(h))) return rc;
  const long nn = (long)(b.imx + 6) * (b.jmx + 6) * (b.kmx + 6);
  VjpNodes nodes;
  if (bar_nodes && (rc = vjp_node_buffers(h, nodes))) return rc;
  HIPCHK(h, hipMemcpyAsync(h->vjp_barR, bar_residue, sizeof(double) * 5 * b.nres, hipMemcpyHostToDevice, st));
  // the ghost cells as fest_residual at its first iteration (fixed faces refixed)
  if ((rc = launch_ghosts(h, 1))) return rc;
  // the viscous faces' primal inputs: gradients and viscosities
  if ((rc = launch_grads(h))) return rc;
  {
    KgScope ts(h, KG_VJP);
    VjpVisc visc;
    if (b.mu_ref != 0.0) {
      visc.bt = h->vjp_bt;
      visc.gb = h->vjp_gb;
    }
    launch_vjp(b, st, h->vjp_barR, h->vjp_c, h->vjp_bar, nodes, visc);
  }
  LAUNCHCHK(h);
  HIPCHK(h, hipMemcpyAsync(bar_qp, h->vjp_bar, sizeof(double) * 5 * b.nc, hipMemcpyDeviceToHost, st));
  if (bar_nodes)
    HIPCHK(h, hipMemcpyAsync(bar_nodes, h->vjp_nodes, sizeof(double) * 3 * nn, hipMemcpyDeviceToHost, st));
  return check_err(h);
}

int fest_time_step(fest_handle *h, double cfl) {
  if (!h) return FEST_E_ARG;
  h->b.CFLv = cfl;
  int rc = launch_time_step(h);
  if (rc) return rc;
  return check_err(h);
}

int fest_get_next_solution(fest_handle *h, int current_iter, double cfl) {
  if (!h) return FEST_E_ARG;
  if (!h->have_grid || !h->have_state) return fail(h, FEST_E_ARG, "grid and state must be set first");
  h->b.CFLv = cfl;
  // kernels see current_iter only through current_iter <= 2 (the inflow
  // fix, src/boundary/bc_primitive.f90:236), so one value per class
  const int itn = current_iter <= 2 ? 1 : 3;
  auto step = [h, itn]() -> int {
    int rc;
    for (const StepOp &op : step_plan(h->b.time_acc))
      if ((rc = run_op(h, op, itn))) return rc;
    return 0;
  };
  int rc = run_graphed(h->graphs, h->stream, current_iter, cfl, h->b.qp, h->timed != 0, step, h->err);
  if (rc) return rc;
  if (++h->unchecked < h->check_every) return 0;
  h->unchecked = 0;
  return check_err(h);
}

int fest_set_option(fest_handle *h, const char *name, long value) {
  if (!h || !name) return FEST_E_ARG;
  if (!strcmp(name, "graphs")) {
    h->graphs.reset();
    h->graphs.enabled = value != 0;
    return 0;
  }
  if (!strcmp(name, "check_interval")) {
    if (value < 1) return fail(h, FEST_E_ARG, "check_interval must be >= 1");
    h->check_every = (int)value;
    h->unchecked = 0;
    return 0;
  }
  return fail(h, FEST_E_ARG, "unknown option '%s'", name);
}

int fest_download(fest_handle *h, uint32_t what, double *qp, double *temp, double *residue, double *F, double *G,
                  double *H, double *delta_t, double *mu, double *mu_t, double *grads, double *sst_f1, double *dist,
                  double *cells4, double *faces) {
  if (!h) return FEST_E_ARG;
  Blk &b = h->b;
  hipStream_t st = h->stream;
  auto cp = [&](void *dst, const void *src, size_t n) -> int {
    if (!dst) return FEST_E_ARG;
    hipError_t e = hipMemcpyAsync(dst, src, n * sizeof(double), hipMemcpyDeviceToHost, st);
    return e == hipSuccess ? 0 : fail(h, FEST_E_HIP, "download: %s", hipGetErrorString(e));
  };
  int rc = 0;
#define DL(bit, ptr, src, n) if ((what & (bit)) && (rc = cp(ptr, src, (size_t)(n)))) return rc
  DL(FEST_DL_QP, qp, b.qp, b.nc * b.nv);
  DL(FEST_DL_TEMP, temp, b.temp, b.nc);
  DL(FEST_DL_RESIDUE, residue, b.res, b.nres * b.nv);
  if (b.fp32) {
    // FP32 mode: the face fluxes are held in FP32; widened on the host
    double *outs[3] = {F, G, H};
    const uint32_t bits[3] = {FEST_DL_F, FEST_DL_G, FEST_DL_H};
    for (int d = 0; d < 3; ++d) {
      if (!(what & bits[d])) continue;
      if (!outs[d]) return FEST_E_ARG;
      const size_t n = (size_t)(b.nfl[d] * b.nv);
      std::vector<float> tmp(n);
      HIPCHK(h, hipMemcpyAsync(tmp.data(), b.flux[d], n * sizeof(float), hipMemcpyDeviceToHost, st));
      HIPCHK(h, hipStreamSynchronize(st));
      for (size_t m = 0; m < n; ++m) outs[d][m] = (double)tmp[m];
    }
  } else {
    DL(FEST_DL_F, F, b.flux[0], b.nfl[0] * b.nv);
    DL(FEST_DL_G, G, b.flux[1], b.nfl[1] * b.nv);
    DL(FEST_DL_H, H, b.flux[2], b.nfl[2] * b.nv);
  }
  DL(FEST_DL_DT, delta_t, b.dt, b.nres);
  DL(FEST_DL_MU, mu, b.mu, b.nc);
  DL(FEST_DL_MUT, mu_t, b.mut, b.nc);
  DL(FEST_DL_F1, sst_f1, b.f1, b.nc);
  DL(FEST_DL_DIST, dist, b.dist, b.nc);
#undef DL
  if (what & FEST_DL_GRADS) {
    if (!grads) return FEST_E_ARG;
    if (b.g32only) {
      hipLaunchKernelGGL(k_g32_widen, dim3(cdiv(b.ngr * b.ng, 256)), dim3(256), 0, st, b);
      LAUNCHCHK(h);
    }
    for (int d = 0; d < 3; ++d)
      if ((rc = cp(grads + (size_t)d * b.ngr * b.ng, b.g[d], (size_t)(b.ngr * b.ng)))) return rc;
  }
  if (what & FEST_DL_CELLS) {
    if (!cells4) return FEST_E_ARG;
    double *src[4] = {b.vol, b.cx, b.cy, b.cz};
    for (int a = 0; a < 4; ++a)
      if ((rc = cp(cells4 + (size_t)a * b.nc, src[a], (size_t)b.nc))) return rc;
  }
  if (what & FEST_DL_FACES) {
    if (!faces) return FEST_E_ARG;
    size_t off = 0;
    for (int d = 0; d < 3; ++d) {
      double *src[4] = {b.fA[d], b.fx[d], b.fy[d], b.fz[d]};
      for (int a = 0; a < 4; ++a) {
        if ((rc = cp(faces + off, src[a], (size_t)b.nfc[d]))) return rc;
        off += (size_t)b.nfc[d];
      }
    }
  }
  HIPCHK(h, hipStreamSynchronize(st));
  return 0;
}

int fest_resnorm(fest_handle *h, double *res_abs) {
  if (!h || !res_abs) return FEST_E_ARG;
  Blk &b = h->b;
  hipLaunchKernelGGL(k_norm_partial, dim3(h->norm_blocks), dim3(256), 0, h->stream, b, b.red);
  hipLaunchKernelGGL(k_norm_final, dim3(1), dim3(256), 0, h->stream, b, h->norm_blocks, (const double *)b.red,
                     b.norms, 1);
  LAUNCHCHK(h);
  HIPCHK(h, hipMemcpyAsync(res_abs, b.norms, sizeof(double) * (b.nv + 1), hipMemcpyDeviceToHost, h->stream));
  // the norms need a wait anyway: deferred device errors are reported here
  h->unchecked = 0;
  return check_err(h);
}

int fest_resnorm_device(fest_handle *h, double *res_abs_dev) {
  if (!h || !res_abs_dev) return FEST_E_ARG;
  Blk &b = h->b;
  int dev = 0;
  HIPCHK(h, hipGetDevice(&dev));
  if (dev < 0 || dev >= SCRATCH_DEVICES || !scratch_dev[dev])
    return fail(h, FEST_E_ARG, "no norm scratch on device %d: create the handle with that device current", dev);
  if (h->norm_blocks > NORM_BLOCKS_MAX)
    return fail(h, FEST_E_ARG, "%d norm blocks exceed the scratch's %d", h->norm_blocks, NORM_BLOCKS_MAX);
  double *red = scratch_dev[dev];
  hipLaunchKernelGGL(k_norm_partial, dim3(h->norm_blocks), dim3(256), 0, h->stream, b, red);
  hipLaunchKernelGGL(k_norm_final, dim3(1), dim3(256), 0, h->stream, b, h->norm_blocks, (const double *)red,
                     res_abs_dev, 0);
  LAUNCHCHK(h);
  return 0;
}

int fest_device_ptr(fest_handle *h, const char *name, void **ptr, size_t *count) {
  if (!h || !name || !ptr) return FEST_E_ARG;
  Blk &b = h->b;
  struct { const char *n; void *p; long c; } t[] = {
      {"qp", b.qp, b.nc * b.nv}, {"temp", b.temp, b.nc}, {"residue", b.res, b.nres * b.nv},
      {"F", b.flux[0], b.nfl[0] * b.nv}, {"G", b.flux[1], b.nfl[1] * b.nv}, {"H", b.flux[2], b.nfl[2] * b.nv},
      {"dt", b.dt, b.nres},
      // the metric adjoints of the last fest_vjp with bar_nodes: bar (nx, ny, nz, A) per face, 4 planes
      {"vjp_metrics_i", h->vjp_m[0], 4 * b.nfl[0]}, {"vjp_metrics_j", h->vjp_m[1], 4 * b.nfl[1]},
      {"vjp_metrics_k", h->vjp_m[2], 4 * b.nfl[2]}};
  for (auto &e : t)
    if (!strcmp(e.n, name)) {
      if (!e.p) return fail(h, FEST_E_ARG, "array '%s' is not allocated", name);
      *ptr = e.p;
      if (count) *count = (size_t)e.c;  // floats for F/G/H in FP32 mode
      return 0;
    }
  return fail(h, FEST_E_ARG, "unknown array '%s'", name);
}

int fest_timing(fest_handle *h, const char *groups) {
  if (!h) return FEST_E_ARG;
  int rc = settle_timers(h);
  if (rc) return rc;
  for (auto &t : h->timers) {
    t.ms = 0.0;
    t.launches = 0;
  }
  h->timed = 0;
  if (!groups || !*groups) return 0;
  if (!strcmp(groups, "*")) {
    h->timed = (1u << KG_COUNT) - 1;
    return 0;
  }
  std::string all(groups);
  size_t pos = 0;
  while (pos <= all.size()) {
    size_t e = all.find(',', pos);
    if (e == std::string::npos) e = all.size();
    std::string name = all.substr(pos, e - pos);
    int g = 0;
    while (g < KG_COUNT && name != kg_names[g]) ++g;
    if (g == KG_COUNT) {
      h->timed = 0;
      return fail(h, FEST_E_ARG, "unknown kernel group '%s'", name.c_str());
    }
    h->timed |= 1u << g;
    pos = e + 1;
  }
  return 0;
}

int fest_kernel_time(fest_handle *h, const char *group, double *total_ms, long *launches) {
  if (!h || !group || !total_ms) return FEST_E_ARG;
  int rc = settle_timers(h);
  if (rc) return rc;
  for (int g = 0; g < KG_COUNT; ++g)
    if (!strcmp(group, kg_names[g])) {
      *total_ms = h->timers[g].ms;
      if (launches) *launches = h->timers[g].launches;
      return 0;
    }
  return fail(h, FEST_E_ARG, "unknown kernel group '%s'", group);
}

// ---------------------------------------------------------------- groups
namespace {

int run_group_op(fest_group *g, const StepOp &op, int current_iter);

int gfail(fest_group *g, int code, const char *fmt, ...) {
  char buf[512];
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(buf, sizeof buf, fmt, ap);
  va_end(ap);
  if (g) g->err = buf;
  return code;
}

// cells of the first and second tangential direction of face f (1..6)
void face_cells(const Blk &b, int f, int &n1, int &n2) {
  if (f <= 2) { n1 = b.jmx - 1; n2 = b.kmx - 1; }
  else if (f <= 4) { n1 = b.imx - 1; n2 = b.kmx - 1; }
  else { n1 = b.imx - 1; n2 = b.jmx - 1; }
}

// map node range -> receiving loop start and step, as
// change_map_to_particular_range (src/mapping.f90:186-258)
void loop_range(int lo, int hi, int &start, int &end, int &dir) {
  start = lo; end = hi; dir = 1;
  if (lo == 1) start = 1;
  if (hi == 1) { end = 1; dir = -1; }
  if (lo > 1) { start = lo - 1; dir = -1; }
  if (hi > 1) end = hi - 1;
}

int group_exchange(fest_group *g) {
  if (g->links.empty()) return 0;
  dim3 grid(cdiv(3L * g->max_layer, 256), (unsigned)g->links.size(), 1);
  hipLaunchKernelGGL(k_exchange, grid, dim3(256), 0, g->stream, (const LinkDev *)g->d_links);
  hipError_t e = hipGetLastError();
  return e == hipSuccess ? 0 : gfail(g, FEST_E_HIP, "exchange launch: %s", hipGetErrorString(e));
}

// the per-block launch grid of a phase (as launch_residual_front and the
// step operations lay it out); false when the block has no tiles in it
bool phase_grid(const Blk &b, int ph, dim3 &g, int ti = TILE_I) {
  const int tj = 256 / ti;
  auto N = [&](int d) { return d == 0 ? b.imx : d == 1 ? b.jmx : b.kmx; };
  switch (ph) {
    case PH_BEGIN: g = dim3(cdiv(b.nc, 256), 1, 1); return true;
    case PH_BC_A0: case PH_BC_A1: case PH_BC_A2: {
      const int d = ph - PH_BC_A0;
      g = dim3(cdiv(face_na(b, 2 * d) + 5, 64), cdiv(face_nb(b, 2 * d) + 5, 4), N(d) >= 8 ? 2 : 1);
      return true;
    }
    case PH_BC_B0: case PH_BC_B1: case PH_BC_B2: {
      const int d = ph - PH_BC_B0;
      g = dim3(cdiv(face_na(b, 2 * d) + 5, 64), cdiv(face_nb(b, 2 * d) + 5, 4), 1);
      return N(d) < 8;
    }
    case PH_EDGE0: g = dim3(cdiv(4L * (b.imx + 5) * b.nv, 256), 1, 1); return true;
    case PH_EDGE1: g = dim3(cdiv(4L * (b.kmx + 5) * b.nv, 256), 1, 1); return true;
    case PH_EDGE2: g = dim3(1, 1, 1); return true;
    case PH_F32: g = dim3(cdiv(b.nc * b.nv, 256), 1, 1); return b.fp32 != 0;
    case PH_GRAD: g = dim3(cdiv((long)(b.imx + 1) * (b.jmx + 1), FLAT_BLOCK), b.kmx + 1, 1); return b.mu_ref != 0.0;
    case PH_GRAD_BC: {
      const int mx = std::max(std::max(b.imx, b.jmx), b.kmx);
      g = dim3(cdiv(mx, 64), cdiv(mx, 4), 6);
      return b.mu_ref != 0.0;
    }
    case PH_FLUX_I: g = dim3(cdiv(flux_i_span(b), FLUX_I_SPAN), b.kmx - 1, 1); return true;
    case PH_FLUX_J: g = dim3(cdiv(b.imx - 1, ti) * cdiv(b.jmx, tj), b.kmx - 1, 1); return true;
    case PH_FLUX_K: g = dim3(cdiv(b.imx - 1, ti) * cdiv(b.jmx - 1, tj), b.kmx, 1); return true;
    case PH_DT: case PH_UPDATE:
      g = dim3(cdiv((long)(b.imx - 1) * (b.jmx - 1), FLAT_BLOCK), b.kmx - 1, 1);
      return true;
    default: return false;
  }
}

// why one batched kernel variant cannot serve every block ("" when it can)
extern "C++" std::string batch_reason(const fest_group *g) {
  const Blk &a = g->blocks[0]->b;
  for (fest_handle *h : g->blocks) {
    const Blk &b = h->b;
    if (b.nv != a.nv || b.recon != a.recon || b.turb != a.turb || b.fp32 != a.fp32 ||
        (b.mu_ref != 0.0) != (a.mu_ref != 0.0) || b.time_acc != a.time_acc)
      return "blocks differ in variables, reconstruction, model, precision or integrator";
    if (b.turb != FEST_TURB_NONE && b.turb != FEST_TURB_SST && b.turb != FEST_TURB_SST2003)
      return "batched steps cover laminar and SST blocks";
    if (b.trans || b.lctm) return "batched steps carry no transition model";
    if (b.kmx == 2) return "batched steps need three-dimensional blocks";
    if (!b.time_local) return "batched steps use local time steps";
    if (b.time_acc != FEST_TA_RK4 && b.time_acc != FEST_TA_RK2 && b.time_acc != FEST_TA_NONE)
      return "batched steps cover RK4, RK2 and the single-stage update";
    for (int f = 0; f < 6; ++f)
      if (b.bc[f] == -11) return "the total-pressure inlet is not batched";
  }
  return "";
}

// every phase's work list, in `order`, uploaded as one device array
int build_lists(fest_group *g, int order) {
  std::vector<WorkItem> all;
  for (int ph = 0; ph < PH_COUNT; ++ph) {
    std::vector<std::vector<WorkItem>> per(g->blocks.size());
    for (size_t n = 0; n < g->blocks.size(); ++n) {
      const Blk &b = g->blocks[n]->b;
      dim3 gr;
      if (!phase_grid(b, ph, gr, g->tile_i)) continue;
      if (ph == PH_GRAD_BC) {
        // the ghost-gradient rule per face, exact tiles of the face
        // interior, none on interface faces (k_grad_bc_b leaves those):
        // the per-block launch's square grid over the largest extent
        // is mostly idle tiles on uneven blocks
        for (int z = 0; z < 6; ++z) {
          if (b.bc[z] >= 0) continue;
          const unsigned nx = cdiv(face_na(b, z) - 1, 64), ny = cdiv(face_nb(b, z) - 1, 4);
          for (unsigned y = 0; y < ny; ++y)
            for (unsigned x = 0; x < nx; ++x) per[n].push_back(WorkItem{(int)n, (int)x, (int)y, z});
        }
        continue;
      }
      for (unsigned z = 0; z < gr.z; ++z)
        for (unsigned y = 0; y < gr.y; ++y)
          for (unsigned x = 0; x < gr.x; ++x) per[n].push_back(WorkItem{(int)n, (int)x, (int)y, (int)z});
    }
    g->ph_off[ph] = (int)all.size();
    if (order == 0) {
      for (auto &v : per) all.insert(all.end(), v.begin(), v.end());
    } else {
      // round robin over the blocks, one tile of each in turn
      size_t mx = 0;
      for (auto &v : per) mx = std::max(mx, v.size());
      for (size_t t = 0; t < mx; ++t)
        for (auto &v : per)
          if (t < v.size()) all.push_back(v[t]);
    }
    g->ph_n[ph] = (int)all.size() - g->ph_off[ph];
  }
  if (g->d_items) (void)hipFree(g->d_items);
  g->d_items = nullptr;
  if (hipMalloc(&g->d_items, sizeof(WorkItem) * std::max<size_t>(1, all.size())) != hipSuccess)
    return gfail(g, FEST_E_HIP, "hipMalloc of the work lists failed");
  if (!all.empty() && hipMemcpy(g->d_items, all.data(), sizeof(WorkItem) * all.size(), hipMemcpyHostToDevice) !=
                          hipSuccess)
    return gfail(g, FEST_E_HIP, "work list upload failed");
  g->lists_order = order;
  g->lists_tile = g->tile_i;
  return 0;
}

// the blocks' records on the device (they carry the CFL number of the call)
int upload_blks(fest_group *g) {
  std::vector<Blk> v;
  for (fest_handle *h : g->blocks) v.push_back(h->b);
  if (!g->d_blks && hipMalloc(&g->d_blks, sizeof(Blk) * v.size()) != hipSuccess)
    return gfail(g, FEST_E_HIP, "hipMalloc of the block records failed");
  if (hipMemcpy(g->d_blks, v.data(), sizeof(Blk) * v.size(), hipMemcpyHostToDevice) != hipSuccess)
    return gfail(g, FEST_E_HIP, "block record upload failed");
  g->blk_cfl = v[0].CFLv;
  return 0;
}



// one operation on all blocks through the phases' work lists
int run_group_op_batched(fest_group *g, const StepOp &op, int current_iter) {
  const Blk &b0 = g->blocks[0]->b;
  hipStream_t st = g->stream;
  auto grid = [g](int ph) -> unsigned {
    return (unsigned)(g->grid_cap > 0 ? std::min(g->ph_n[ph], g->grid_cap) : g->ph_n[ph]);
  };
  // one statement, so that it nests under if / else like a function call
#define KM(ph, kern, bd, ...)                                                                        \
  do {                                                                                               \
    if (g->ph_n[ph])                                                                                 \
      hipLaunchKernelGGL(kern, dim3(grid(ph)), bd, 0, st, (const Blk *)g->d_blks,                    \
                         (const WorkItem *)(g->d_items + g->ph_off[ph]), g->ph_n[ph], ##__VA_ARGS__); \
  } while (0)
  int rc;
  switch (op.kind) {
    case OP_BEGIN: KM(PH_BEGIN, km_step_begin, dim3(256), op.flags); break;
    case OP_RESID:
      if ((rc = group_exchange(g))) return rc;
      for (int d = 0; d < 3; ++d) {
        KM(PH_BC_A0 + d, km_bc_face, dim3(64, 4, 1), 2 * d, current_iter);
        KM(PH_BC_B0 + d, km_bc_face, dim3(64, 4, 1), 2 * d + 1, current_iter);
      }
      KM(PH_EDGE0, km_bc_edges, dim3(256), 0);
      KM(PH_EDGE1, km_bc_edges, dim3(256), 1);
      KM(PH_EDGE2, km_bc_edges, dim3(256), 2);
      if (b0.fp32) KM(PH_F32, km_to_f32, dim3(256));
      if (b0.mu_ref != 0.0) {
        if (b0.g32only) {
          if (b0.turb == FEST_TURB_SST2003) KM(PH_GRAD, (km_grad<6, true, true>), dim3(256));
          else if (b0.ng == 6) KM(PH_GRAD, (km_grad<6, false, true>), dim3(256));
          else KM(PH_GRAD, (km_grad<4, false, true>), dim3(256));
        } else if (b0.turb == FEST_TURB_SST2003) {
          KM(PH_GRAD, (km_grad<6, true>), dim3(256));
        } else if (b0.ng == 6) {
          KM(PH_GRAD, (km_grad<6, false>), dim3(256));
        } else {
          KM(PH_GRAD, (km_grad<4, false>), dim3(256));
        }
        KM(PH_GRAD_BC, km_grad_bc, dim3(64, 4, 1));
      }
      for (int d = 0; d < 3; ++d) {
        const int ph = PH_FLUX_I + d;
        if (!g->ph_n[ph]) continue;
        const WorkItem *items = g->d_items + g->ph_off[ph];
        if (d == 0) launch_flux_batched<0>(b0, st, g->d_blks, items, g->ph_n[ph], grid(ph), g->tile_i, op.flags);
        else if (d == 1) launch_flux_batched<1>(b0, st, g->d_blks, items, g->ph_n[ph], grid(ph), g->tile_i, op.flags);
        else launch_flux_batched<2>(b0, st, g->d_blks, items, g->ph_n[ph], grid(ph), g->tile_i, op.flags);
      }
      break;
    case OP_DT: KM(PH_DT, km_dt, dim3(256)); break;
    case OP_UPDATE: KM(PH_UPDATE, km_update, dim3(256), op.a, op.c, op.flags); break;
    default:
      for (fest_handle *h : g->blocks)
        if ((rc = run_op(h, op, current_iter))) return gfail(g, rc, "%s", h->err.c_str());
      return 0;
  }
#undef KM
  hipError_t e = hipGetLastError();
  return e == hipSuccess ? 0 : gfail(g, FEST_E_HIP, "batched launch: %s", hipGetErrorString(e));
}

int group_check(fest_group *g) {
  for (size_t n = 0; n < g->blocks.size(); ++n) {
    int rc = check_err(g->blocks[n]);
    if (rc) return gfail(g, rc, "block %zu: %s", n, g->blocks[n]->err.c_str());
  }
  return 0;
}

}  // namespace

int fest_group_create(fest_group **out, fest_handle *const *blocks, int nblocks, const fest_link *links,
                      int nlinks) {
  if (!out || !blocks || nblocks < 1 || (nlinks > 0 && !links) || nlinks < 0) return FEST_E_ARG;
  fest_group *g = new fest_group();
  *out = g;
  for (int n = 0; n < nblocks; ++n) {
    if (!blocks[n]) return gfail(g, FEST_E_ARG, "block %d is null", n);
    if (blocks[n]->group) return gfail(g, FEST_E_ARG, "block %d is already in a group", n);
    if (blocks[n]->b.nv != blocks[0]->b.nv || blocks[n]->b.time_acc != blocks[0]->b.time_acc)
      return gfail(g, FEST_E_CONFIG, "blocks of a group need the same n_var and time integration");
    for (int m = 0; m < n; ++m)
      if (blocks[m] == blocks[n]) return gfail(g, FEST_E_ARG, "block %d appears twice", n);
  }
  std::vector<LinkDev> dl;
  for (int e = 0; e < nlinks; ++e) {
    const fest_link &L = links[e];
    if (L.block < 0 || L.block >= nblocks || L.from_block < 0 || L.from_block >= nblocks || L.face < 1 ||
        L.face > 6 || L.from_face < 1 || L.from_face > 6)
      return gfail(g, FEST_E_ARG, "link %d: block or face out of range", e);
    const Blk &r = blocks[L.block]->b, &s = blocks[L.from_block]->b;
    int rn1, rn2, sn1, sn2;
    face_cells(r, L.face, rn1, rn2);
    face_cells(s, L.from_face, sn1, sn2);
    int lo1, hi1, d1, lo2, hi2, d2;
    loop_range(L.lo1, L.hi1, lo1, hi1, d1);
    loop_range(L.lo2, L.hi2, lo2, hi2, d2);
    // the receiving loops must cover the face cells exactly, and the sender
    // must supply as many cells per layer as the receiver takes
    if ((hi1 - lo1) * d1 + 1 != rn1 || (hi2 - lo2) * d2 + 1 != rn2 || std::min(lo1, hi1) != 1 ||
        std::min(lo2, hi2) != 1 || (long)sn1 * sn2 != (long)rn1 * rn2)
      return gfail(g, FEST_E_CONFIG, "link %d: map ranges do not match the faces (%d x %d <- %d x %d)", e, rn1, rn2,
                   sn1, sn2);
    LinkDev d;
    memset(&d, 0, sizeof d);
    d.src = s.qp; d.dst = r.qp;
    d.src_nc = s.nc; d.dst_nc = r.nc;
    d.src_pi = s.pi; d.src_pj = s.pj; d.dst_pi = r.pi; d.dst_pj = r.pj;
    d.src_N[0] = s.imx; d.src_N[1] = s.jmx; d.src_N[2] = s.kmx;
    d.dst_N[0] = r.imx; d.dst_N[1] = r.jmx; d.dst_N[2] = r.kmx;
    d.src_face = L.from_face; d.dst_face = L.face;
    d.s_n1 = sn1; d.per_layer = sn1 * sn2;
    d.lo1 = lo1; d.d1 = d1; d.lo2 = lo2; d.d2 = d2;
    d.dir_switch = L.dir_switch ? 1 : 0;
    d.r_inner = d.dir_switch ? rn2 : rn1;
    d.nv = r.nv;
    dl.push_back(d);
    g->max_layer = std::max(g->max_layer, d.per_layer);
  }
  g->links.assign(links, links + nlinks);
  if (!dl.empty()) {
    if (hipMalloc(&g->d_links, sizeof(LinkDev) * dl.size()) != hipSuccess)
      return gfail(g, FEST_E_HIP, "hipMalloc of the link table failed");
    if (hipMemcpy(g->d_links, dl.data(), sizeof(LinkDev) * dl.size(), hipMemcpyHostToDevice) != hipSuccess)
      return gfail(g, FEST_E_HIP, "link table upload failed");
  }
  if (hipStreamCreate(&g->stream) != hipSuccess) return gfail(g, FEST_E_HIP, "hipStreamCreate failed");
  for (int n = 0; n < nblocks; ++n) {
    fest_handle *h = blocks[n];
    (void)hipStreamSynchronize(h->own_stream);
    h->stream = g->stream;
    h->group = g;
    g->blocks.push_back(h);
  }
  g->batch_why = batch_reason(g);
  return 0;
}

int fest_group_autotune(fest_group *g, int steps, double *ms, int ms_len, int *ncand, int *chosen) {
  if (!g || steps < 3) return FEST_E_ARG;
  if (g->broken) return FEST_E_ARG;
  for (fest_handle *h : g->blocks)
    if (!h->have_grid || !h->have_state) return gfail(g, FEST_E_ARG, "grid and state must be set on every block");
  hipDeviceProp_t prop;
  int dev = 0;
  if (hipGetDevice(&dev) != hipSuccess || hipGetDeviceProperties(&prop, dev) != hipSuccess)
    return gfail(g, FEST_E_HIP, "device properties unavailable");
  const int cu = prop.multiProcessorCount;
  struct Cand { int batched, cap, order, tile; };
  std::vector<Cand> cands = {{0, 0, 0, TILE_I}};
  if (g->batch_why.empty())
    for (int tile : {64, 32})
      for (int order = 0; order < 2; ++order)
        for (int cap : {0, 16 * cu, 8 * cu, 4 * cu}) cands.push_back({1, cap, order, tile});
  // snapshot of the state: a step reads only qp (Temp is refreshed from it)
  std::vector<double *> saved;
  for (fest_handle *h : g->blocks) {
    double *p = nullptr;
    if (hipMalloc(&p, sizeof(double) * h->b.nc * h->b.nv) != hipSuccess) {
      for (double *q : saved) (void)hipFree(q);
      return gfail(g, FEST_E_HIP, "hipMalloc of the state snapshot failed");
    }
    (void)hipMemcpyAsync(p, h->b.qp, sizeof(double) * h->b.nc * h->b.nv, hipMemcpyDeviceToDevice, g->stream);
    saved.push_back(p);
  }
  auto restore = [&]() {
    for (size_t n = 0; n < g->blocks.size(); ++n)
      (void)hipMemcpyAsync(g->blocks[n]->b.qp, saved[n], sizeof(double) * g->blocks[n]->b.nc * g->blocks[n]->b.nv,
                           hipMemcpyDeviceToDevice, g->stream);
  };
  const double cfl = g->blocks[0]->b.CFLv;
  const int keep_b = g->batched, keep_c = g->grid_cap, keep_o = g->order, keep_t = g->tile_i;
  hipEvent_t e0, e1;
  (void)hipEventCreate(&e0);
  (void)hipEventCreate(&e1);
  int rc = 0, best = 0;
  std::vector<double> t(cands.size(), 0.0);
  for (size_t c = 0; c < cands.size() && !rc; ++c) {
    g->graphs.reset();
    g->batched = cands[c].batched;
    g->grid_cap = cands[c].cap;
    g->order = cands[c].order;
    g->tile_i = cands[c].tile;
    g->blk_cfl = -1.0;
    restore();
    // the first two steps capture the graph; the rest are timed
    for (int s = 0; s < steps && !rc; ++s) {
      if (s == 2) (void)hipEventRecord(e0, g->stream);
      rc = fest_group_get_next_solution(g, 3 + s, cfl);
    }
    (void)hipEventRecord(e1, g->stream);
    if (hipEventSynchronize(e1) != hipSuccess) rc = gfail(g, FEST_E_HIP, "autotune timing failed");
    float el = 0.f;
    (void)hipEventElapsedTime(&el, e0, e1);
    t[c] = el / (steps - 2);
    if (t[c] < t[best]) best = (int)c;
  }
  restore();
  (void)hipStreamSynchronize(g->stream);
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  for (double *q : saved) (void)hipFree(q);
  g->graphs.reset();
  g->blk_cfl = -1.0;
  if (rc) {
    g->batched = keep_b;
    g->grid_cap = keep_c;
    g->order = keep_o;
    g->tile_i = keep_t;
    return rc;
  }
  g->batched = cands[best].batched;
  g->grid_cap = cands[best].cap;
  g->order = cands[best].order;
  g->tile_i = cands[best].tile;
  if (ncand) *ncand = (int)cands.size();
  if (chosen) *chosen = best;
  for (int c = 0; ms && c < (int)cands.size() && c < ms_len; ++c) ms[c] = t[c];
  return 0;
}

void fest_group_destroy(fest_group *g) {
  if (!g) return;
  if (g->stream) (void)hipStreamSynchronize(g->stream);
  g->graphs.reset();
  if (g->d_blks) (void)hipFree(g->d_blks);
  if (g->d_items) (void)hipFree(g->d_items);
  if (g->d_norms) (void)hipFree(g->d_norms);
  for (fest_handle *h : g->blocks) {
    h->stream = h->own_stream;
    h->group = nullptr;
  }
  if (g->d_links) (void)hipFree(g->d_links);
  if (g->d_vjp_links) (void)hipFree(g->d_vjp_links);
  if (g->stream) (void)hipStreamDestroy(g->stream);
  delete g;
}

const char *fest_group_last_error(const fest_group *g) { return g ? g->err.c_str() : "null group"; }

int fest_group_exchange(fest_group *g) {
  if (!g) return FEST_E_ARG;
  if (g->broken) return FEST_E_ARG;
  int rc = group_exchange(g);
  if (rc) return rc;
  if (hipStreamSynchronize(g->stream) != hipSuccess) return gfail(g, FEST_E_HIP, "exchange failed");
  return 0;
}

int fest_group_vjp(fest_group *g, const double *const *bar_residues, double *const *bar_qps,
                   double *const *bar_nodes) {
  if (!g || !bar_residues || !bar_qps) return FEST_E_ARG;
  if (g->broken) return FEST_E_ARG;
  const int nb = (int)g->blocks.size();
  for (int n = 0; n < nb; ++n) {
    fest_handle *h = g->blocks[n];
    if (!bar_residues[n] || !bar_qps[n]) return gfail(g, FEST_E_ARG, "block %d: null adjoint array", n);
    if (!h->have_grid || !h->have_state) return gfail(g, FEST_E_ARG, "block %d: grid and state must be set first", n);
    std::string why = vjp_reason(h->b);
    if (why.empty() && h->b.mu_ref != 0.0) why = "the group adjoint covers inviscid blocks";
    if (why.empty() && bar_nodes && bar_nodes[n]) why = vjp_nodes_reason(h->b);
    if (!why.empty()) return gfail(g, FEST_E_CONFIG, "block %d: fest_vjp: %s", n, why.c_str());
  }
  int rc;
  // the interface ghost cells as the group's residual sees them
  if ((rc = group_exchange(g))) return rc;
  for (int n = 0; n < nb; ++n) {
    fest_handle *h = g->blocks[n];
    if ((rc = vjp_buffers(h))) return gfail(g, rc, "block %d: %s", n, h->err.c_str());
    if (hipMemcpyAsync(h->vjp_barR, bar_residues[n], sizeof(double) * 5 * h->b.nres, hipMemcpyHostToDevice,
                       g->stream) != hipSuccess)
      return gfail(g, FEST_E_HIP, "block %d: adjoint upload failed", n);
    VjpNodes nodes;
    if (bar_nodes && bar_nodes[n] && (rc = vjp_node_buffers(h, nodes)))
      return gfail(g, rc, "block %d: %s", n, h->err.c_str());
    if ((rc = launch_ghosts(h, 1))) return gfail(g, rc, "block %d: %s", n, h->err.c_str());
    KgScope ts(h, KG_VJP);
    launch_vjp(h->b, g->stream, h->vjp_barR, h->vjp_c, h->vjp_bar, nodes, VjpVisc());
  }
  // the interface ghost cells' adjoints go back to the cells they copied
  if (!g->links.empty()) {
    if (!g->d_vjp_links) {
      std::vector<LinkDev> dl(g->links.size());
      if (hipMemcpy(dl.data(), g->d_links, sizeof(LinkDev) * dl.size(), hipMemcpyDeviceToHost) != hipSuccess)
        return gfail(g, FEST_E_HIP, "link table download failed");
      for (LinkDev &L : dl)
        for (fest_handle *h : g->blocks) {
          if (L.src == h->b.qp) L.src = h->vjp_bar;
          if (L.dst == h->b.qp) L.dst = h->vjp_bar;
        }
      if (hipMalloc(&g->d_vjp_links, sizeof(LinkDev) * dl.size()) != hipSuccess)
        return gfail(g, FEST_E_HIP, "hipMalloc of the adjoint link table failed");
      if (hipMemcpy(g->d_vjp_links, dl.data(), sizeof(LinkDev) * dl.size(), hipMemcpyHostToDevice) != hipSuccess)
        return gfail(g, FEST_E_HIP, "adjoint link table upload failed");
    }
    for (int e = (int)g->links.size() - 1; e >= 0; --e)
      hipLaunchKernelGGL(k_exchange_adjoint, dim3(cdiv(3L * g->max_layer, 256)), dim3(256), 0, g->stream,
                         (const LinkDev *)g->d_vjp_links, e);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return gfail(g, FEST_E_HIP, "adjoint exchange launch: %s", hipGetErrorString(e));
  }
  for (int n = 0; n < nb; ++n) {
    fest_handle *h = g->blocks[n];
    if (hipMemcpyAsync(bar_qps[n], h->vjp_bar, sizeof(double) * 5 * h->b.nc, hipMemcpyDeviceToHost, g->stream) !=
        hipSuccess)
      return gfail(g, FEST_E_HIP, "block %d: adjoint download failed", n);
    const long nn = (long)(h->b.imx + 6) * (h->b.jmx + 6) * (h->b.kmx + 6);
    if (bar_nodes && bar_nodes[n] &&
        hipMemcpyAsync(bar_nodes[n], h->vjp_nodes, sizeof(double) * 3 * nn, hipMemcpyDeviceToHost, g->stream) !=
            hipSuccess)
      return gfail(g, FEST_E_HIP, "block %d: node adjoint download failed", n);
  }
  for (int n = 0; n < nb; ++n)
    if ((rc = check_err(g->blocks[n]))) return gfail(g, rc, "block %d: %s", n, g->blocks[n]->err.c_str());
  return 0;
}

int fest_group_wall_dist(fest_group *g) {
  if (!g) return FEST_E_ARG;
  if (g->broken) return FEST_E_ARG;
  for (fest_handle *h : g->blocks)
    if (!h->have_grid) return gfail(g, FEST_E_ARG, "the grid must be set on every block");
  std::vector<double> wall;
  int rc;
  for (fest_handle *h : g->blocks)
    if ((rc = wall_nodes(h, wall))) return gfail(g, rc, "%s", h->err.c_str());
  for (fest_handle *h : g->blocks)
    if ((rc = dist_from_walls(h, wall))) return gfail(g, rc, "%s", h->err.c_str());
  return 0;
}

int fest_group_resnorm(fest_group *g, double *res_abs) {
  if (!g || !res_abs) return FEST_E_ARG;
  if (g->broken) return FEST_E_ARG;
  const int nv = g->blocks[0]->b.nv;
  const size_t nb = g->blocks.size();
  // every block's share into one device array, then one copy and one wait
  if (!g->d_norms && hipMalloc(&g->d_norms, sizeof(double) * NORM_SLOTS * nb) != hipSuccess)
    return gfail(g, FEST_E_HIP, "hipMalloc of the group norms failed");
  for (size_t n = 0; n < nb; ++n) {
    fest_handle *h = g->blocks[n];
    Blk &b = h->b;
    hipLaunchKernelGGL(k_norm_partial, dim3(h->norm_blocks), dim3(256), 0, g->stream, b, b.red);
    hipLaunchKernelGGL(k_norm_final, dim3(1), dim3(256), 0, g->stream, b, h->norm_blocks, (const double *)b.red,
                       g->d_norms + NORM_SLOTS * n, 1);
  }
  if (hipGetLastError() != hipSuccess) return gfail(g, FEST_E_HIP, "norm launch failed");
  std::vector<double> part(NORM_SLOTS * nb);
  if (hipMemcpyAsync(part.data(), g->d_norms, sizeof(double) * NORM_SLOTS * nb, hipMemcpyDeviceToHost, g->stream) !=
          hipSuccess ||
      hipStreamSynchronize(g->stream) != hipSuccess)
    return gfail(g, FEST_E_HIP, "norm download failed");
  // MPI_ALLGATHER then the sum in process order (src/resnorm.f90:193-213)
  std::vector<double> tot(nv + 1, 0.0);
  for (size_t n = 0; n < nb; ++n)
    for (int l = 0; l <= nv; ++l) tot[l] += part[NORM_SLOTS * n + NORM_SHARE + l];
  res_abs[0] = fabs(tot[0]);
  for (int l = 1; l <= nv; ++l) res_abs[l] = sqrt(tot[l]);
  return 0;
}

int fest_group_get_next_solution(fest_group *g, int current_iter, double cfl) {
  if (!g) return FEST_E_ARG;
  if (g->broken) return FEST_E_ARG;
  for (fest_handle *h : g->blocks)
    if (!h->have_grid || !h->have_state) return gfail(g, FEST_E_ARG, "grid and state must be set on every block");
  for (fest_handle *h : g->blocks) h->b.CFLv = cfl;
  int rc0;
  if (g->batched) {
    if ((g->lists_order != g->order || g->lists_tile != g->tile_i) && (rc0 = build_lists(g, g->order))) return rc0;
    if (g->blk_cfl != cfl) {
      // the records carry the CFL number; nothing may still read them
      if (hipStreamSynchronize(g->stream) != hipSuccess) return gfail(g, FEST_E_HIP, "stream wait failed");
      if ((rc0 = upload_blks(g))) return rc0;
    }
  }
  const int itn = current_iter <= 2 ? 1 : 3;
  bool timed = false;
  for (fest_handle *h : g->blocks) timed |= h->timed != 0;
  auto step = [g, itn]() -> int {
    int rc;
    for (const StepOp &op : step_plan(g->blocks[0]->b.time_acc))
      if ((rc = run_group_op(g, op, itn))) return rc;
    return 0;
  };
  int rc = run_graphed(g->graphs, g->stream, current_iter, cfl, g->blocks[0]->b.qp, timed, step, g->err);
  if (rc) return rc;
  if (++g->unchecked < g->check_every) return 0;
  g->unchecked = 0;
  return group_check(g);
}

int fest_group_set_option(fest_group *g, const char *name, long value) {
  if (!g || !name) return FEST_E_ARG;
  if (!strcmp(name, "batched")) {
    if (value && !g->batch_why.empty()) return gfail(g, FEST_E_CONFIG, "not batched: %s", g->batch_why.c_str());
    g->graphs.reset();
    g->batched = value != 0;
    g->blk_cfl = -1.0;
    return 0;
  }
  if (!strcmp(name, "grid_cap")) {
    if (value < 0) return gfail(g, FEST_E_ARG, "grid_cap must be >= 0");
    g->graphs.reset();
    g->grid_cap = (int)value;
    return 0;
  }
  if (!strcmp(name, "tile_i")) {
    if (value != 64 && value != 32) return gfail(g, FEST_E_ARG, "tile_i must be 64 or 32");
    g->graphs.reset();
    g->tile_i = (int)value;
    return 0;
  }
  if (!strcmp(name, "order")) {
    if (value != 0 && value != 1) return gfail(g, FEST_E_ARG, "order must be 0 or 1");
    g->graphs.reset();
    g->order = (int)value;
    return 0;
  }
  if (!strcmp(name, "graphs")) {
    g->graphs.reset();
    g->graphs.enabled = value != 0;
    return 0;
  }
  if (!strcmp(name, "check_interval")) {
    if (value < 1) return gfail(g, FEST_E_ARG, "check_interval must be >= 1");
    g->check_every = (int)value;
    g->unchecked = 0;
    return 0;
  }
  return gfail(g, FEST_E_ARG, "unknown option '%s'", name);
}

namespace {
// an operation on every block of a group; each residual evaluation opens
// with the interface exchange (src/update.f90:534)
int run_group_op(fest_group *g, const StepOp &op, int current_iter) {
  int rc;
  if (g->batched) return run_group_op_batched(g, op, current_iter);
  if (op.kind == OP_RESID && (rc = group_exchange(g))) return rc;
  for (fest_handle *h : g->blocks)
    if ((rc = run_op(h, op, current_iter))) return gfail(g, rc, "%s", h->err.c_str());
  return 0;
}
}  // namespace

int fest_sync(fest_handle *h) {
  if (!h) return FEST_E_ARG;
  HIPCHK(h, hipStreamSynchronize(h->stream));
  h->unchecked = 0;
  return check_err(h);
}

}  // extern "C"
