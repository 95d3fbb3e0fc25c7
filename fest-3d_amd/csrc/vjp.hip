// vjp.hip -- launches of the adjoint kernels (k_vjp.h) for every
// reconstruction.
#include <hip/hip_runtime.h>

#include "k_vjp.h"
#include "vjp_launch.h"

namespace fest {
namespace {

template <int D, int RECON>
void vjp_dir_t(const Blk &b, hipStream_t st, const double *barR, double *c) {
  const long nf = b.nfl[D];
  double *fbar = c + (long)VJP_PER_FACE * nf;  // bar qL, bar qR after the contributions
  hipLaunchKernelGGL((k_vjp_flux<D, RECON>), dim3((unsigned)((nf + 255) / 256)), dim3(256), 0, st, b, barR, fbar);
  hipLaunchKernelGGL((k_vjp_recon<D, RECON>), dim3((unsigned)((5 * nf + 255) / 256)), dim3(256), 0, st, b, fbar, c);
}
template <int D>
void vjp_dir(const Blk &b, hipStream_t st, const double *barR, double *c) {
  switch (b.recon) {
    case 0: vjp_dir_t<D, 0>(b, st, barR, c); break;
    case 1: vjp_dir_t<D, 1>(b, st, barR, c); break;
    case 2: vjp_dir_t<D, 2>(b, st, barR, c); break;
    case 3: vjp_dir_t<D, 3>(b, st, barR, c); break;
    default: vjp_dir_t<D, 4>(b, st, barR, c); break;
  }
}

}  // namespace

long vjp_contrib_len(const Blk &b, int d) { return (long)b.nfl[d] * (VJP_PER_FACE + 10); }

void launch_vjp(const Blk &b, hipStream_t st, const double *barR, double *const contrib[3], double *bar) {
  vjp_dir<0>(b, st, barR, contrib[0]);
  vjp_dir<1>(b, st, barR, contrib[1]);
  vjp_dir<2>(b, st, barR, contrib[2]);
  hipLaunchKernelGGL(k_vjp_gather, dim3((unsigned)((b.nc + 255) / 256)), dim3(256), 0, st, b, contrib[0],
                     contrib[1], contrib[2], bar);
  // the k faces of a 2-D grid (one cell in k) feed no flux
  for (int face = b.kmx == 2 ? 3 : 5; face >= 0; --face) {
    const int na = face_na(b, face), nb = face_nb(b, face);
    hipLaunchKernelGGL(k_vjp_bc, dim3((unsigned)((na - 1 + 255) / 256), (unsigned)(nb - 1)), dim3(256), 0, st, b,
                       face, bar);
  }
}

}  // namespace fest
