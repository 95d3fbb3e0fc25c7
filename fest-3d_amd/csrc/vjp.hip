// vjp.hip -- launches of the adjoint kernels (k_vjp.h) for every
// reconstruction.
#include <hip/hip_runtime.h>

#include "k_vjp.h"
#include "vjp_launch.h"

namespace fest {
namespace {

template <int D>
void vjp_faces(const Blk &b, hipStream_t st, const double *barR, double *c) {
  const dim3 g((unsigned)((b.nfl[D] + 255) / 256)), t(256);
  switch (b.recon) {
    case 0: hipLaunchKernelGGL((k_vjp_faces<D, 0>), g, t, 0, st, b, barR, c); break;
    case 1: hipLaunchKernelGGL((k_vjp_faces<D, 1>), g, t, 0, st, b, barR, c); break;
    case 2: hipLaunchKernelGGL((k_vjp_faces<D, 2>), g, t, 0, st, b, barR, c); break;
    case 3: hipLaunchKernelGGL((k_vjp_faces<D, 3>), g, t, 0, st, b, barR, c); break;
    default: hipLaunchKernelGGL((k_vjp_faces<D, 4>), g, t, 0, st, b, barR, c); break;
  }
}

}  // namespace

long vjp_contrib_len(const Blk &b, int d) { return (long)b.nfl[d] * VJP_PER_FACE; }

void launch_vjp(const Blk &b, hipStream_t st, const double *barR, double *const contrib[3], double *bar) {
  vjp_faces<0>(b, st, barR, contrib[0]);
  vjp_faces<1>(b, st, barR, contrib[1]);
  vjp_faces<2>(b, st, barR, contrib[2]);
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
