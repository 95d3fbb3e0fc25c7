// vjp.hip -- launches of the adjoint kernels (k_vjp.h) for every
// reconstruction.
#include <hip/hip_runtime.h>

#include "k_vjp.h"
#include "vjp_launch.h"

namespace fest {
namespace {

inline unsigned blocks(long n) { return (unsigned)((n + 255) / 256); }

template <int D, int RECON>
void vjp_dir_t(const Blk &b, hipStream_t st, const double *barR, double *c, double *mbar) {
  const long nf = b.nfl[D];
  double *fbar = c + (long)VJP_PER_FACE * nf;  // bar qL, bar qR after the contributions
  hipLaunchKernelGGL((k_vjp_flux<D, RECON>), dim3(blocks(nf)), dim3(256), 0, st, b, barR, fbar, mbar);
  double *pbc = fbar + 10 * nf;  // the pressure-based switch's pressure contributions, when it is on
  hipLaunchKernelGGL((k_vjp_recon<D, RECON>), dim3(blocks(5 * nf)), dim3(256), 0, st, b, fbar, c, pbc);
}
template <int D>
void vjp_dir(const Blk &b, hipStream_t st, const double *barR, double *c, double *mbar) {
  switch (b.recon) {
    case 0: vjp_dir_t<D, 0>(b, st, barR, c, mbar); break;
    case 1: vjp_dir_t<D, 1>(b, st, barR, c, mbar); break;
    case 2: vjp_dir_t<D, 2>(b, st, barR, c, mbar); break;
    case 3: vjp_dir_t<D, 3>(b, st, barR, c, mbar); break;
    default: vjp_dir_t<D, 4>(b, st, barR, c, mbar); break;
  }
}

}  // namespace

long vjp_contrib_len(const Blk &b, int d) {
  const bool pb = (b.recon == 1 || b.recon == 2) && b.pbs[d];
  return (long)b.nfl[d] * (VJP_PER_FACE + 10 + (pb ? 4 * VJP_STENCIL : 0));
}

void launch_vjp(const Blk &b, hipStream_t st, const double *barR, double *const contrib[3], double *bar,
                const VjpNodes &nodes, const VjpVisc &visc) {
  vjp_dir<0>(b, st, barR, contrib[0], nodes.mbar[0]);
  vjp_dir<1>(b, st, barR, contrib[1], nodes.mbar[1]);
  vjp_dir<2>(b, st, barR, contrib[2], nodes.mbar[2]);
  hipLaunchKernelGGL(k_vjp_gather, dim3(blocks(b.nc)), dim3(256), 0, st, b, contrib[0], contrib[1], contrib[2], bar);
  if (visc.bt) {
    hipLaunchKernelGGL(k_vjp_visc<0>, dim3(blocks(b.nfl[0])), dim3(256), 0, st, b, barR, contrib[0]);
    hipLaunchKernelGGL(k_vjp_visc<1>, dim3(blocks(b.nfl[1])), dim3(256), 0, st, b, barR, contrib[1]);
    hipLaunchKernelGGL(k_vjp_visc<2>, dim3(blocks(b.nfl[2])), dim3(256), 0, st, b, barR, contrib[2]);
    hipLaunchKernelGGL(k_vjp_visc_gather, dim3(blocks(b.nc)), dim3(256), 0, st, b, contrib[0], contrib[1], contrib[2],
                       bar, visc.bt, visc.gb);
    for (int face = 5; face >= 0; --face) {
      const int na = face_na(b, face), nb = face_nb(b, face);
      hipLaunchKernelGGL(k_vjp_grad_bc, dim3(blocks(na - 1), (unsigned)(nb - 1)), dim3(256), 0, st, b, face, bar,
                         visc.bt, visc.gb);
    }
    hipLaunchKernelGGL(k_vjp_grad, dim3(blocks(b.nc)), dim3(256), 0, st, b, visc.gb, visc.bt, bar);
  }
  // the k faces of a 2-D grid (one cell in k) feed no flux
  for (int face = b.kmx == 2 ? 3 : 5; face >= 0; --face) {
    const int na = face_na(b, face), nb = face_nb(b, face);
    double *gface = face == 3 ? nodes.g3 : face == 5 ? nodes.g5 : nullptr;
    hipLaunchKernelGGL(k_vjp_bc, dim3(blocks(na - 1), (unsigned)(nb - 1)), dim3(256), 0, st, b, face, bar,
                       nodes.mbar[0], nodes.mbar[1], nodes.mbar[2], gface);
  }
  if (!nodes.bar_nodes) return;
  // the metric adjoints to the face diagonals (into the contribution
  // buffers, which the gather has consumed), then to the nodes
  hipLaunchKernelGGL(k_vjp_metrics<0>, dim3(blocks(b.nfl[0])), dim3(256), 0, st, b, nodes.mbar[0], contrib[0]);
  hipLaunchKernelGGL(k_vjp_metrics<1>, dim3(blocks(b.nfl[1])), dim3(256), 0, st, b, nodes.mbar[1], contrib[1]);
  hipLaunchKernelGGL(k_vjp_metrics<2>, dim3(blocks(b.nfl[2])), dim3(256), 0, st, b, nodes.mbar[2], contrib[2]);
  if (nodes.g3)
    hipLaunchKernelGGL(k_vjp_metrics_ghost, dim3(blocks(b.imx - 1), (unsigned)(b.kmx - 1)), dim3(256), 0, st, b, 3,
                       nodes.g3);
  if (nodes.g5)
    hipLaunchKernelGGL(k_vjp_metrics_ghost, dim3(blocks(b.imx - 1), (unsigned)(b.jmx - 1)), dim3(256), 0, st, b, 5,
                       nodes.g5);
  const long nn = (long)(b.imx + 6) * (b.jmx + 6) * (b.kmx + 6);
  hipLaunchKernelGGL(k_vjp_nodes, dim3(blocks(nn)), dim3(256), 0, st, b, contrib[0], contrib[1], contrib[2], nodes.g3,
                     nodes.g5, nodes.bar_nodes);
}

}  // namespace fest
