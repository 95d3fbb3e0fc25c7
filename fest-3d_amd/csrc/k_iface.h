// k_iface.h -- multiblock interface exchange (src/interface1.f90:96-705):
// the three ghost layers of a receiving face take the three interior layers
// next to the sending face, element for element in the reference's order:
// the sender packs var, layer, then its face cells with the first tangential
// direction fastest; the receiver unpacks them walking its map ranges
// (src/mapping.f90:186-258), second direction outer unless the direction
// switch is set.  Sender and receiver are resident blocks, so every element
// is one device copy; the links read interior cells and write ghost cells
// only, so all links run in one launch.
#pragma once
#include "blk.h"

namespace fest {

struct LinkDev {
  const double *src;
  double *dst;
  long src_nc, dst_nc;        // cells per plane
  int src_pi, src_pj;         // padded pitches (imx+5, jmx+5)
  int dst_pi, dst_pj;
  int src_N[3], dst_N[3];     // imx, jmx, kmx
  int src_face, dst_face;     // 1..6
  int s_n1, per_layer;        // sender first tangential cells, cells per layer
  int lo1, d1, lo2, d2;       // receiver loop starts and steps
  int r_inner;                // receiver inner-loop length
  int dir_switch;
  int nv;
};

__device__ __forceinline__ long cell_at(int pi, int pj, int i, int j, int k) {
  return (long)(i + 2) + (long)pi * ((long)(j + 2) + (long)pj * (long)(k + 2));
}

// interior cell of layer l next to face f at tangential cell (a, b)
__device__ __forceinline__ long layer_cell(const LinkDev &L, int l, int a, int b) {
  const int *N = L.src_N;
  switch (L.src_face) {
    case 1: return cell_at(L.src_pi, L.src_pj, l, a, b);
    case 2: return cell_at(L.src_pi, L.src_pj, N[0] - l, a, b);
    case 3: return cell_at(L.src_pi, L.src_pj, a, l, b);
    case 4: return cell_at(L.src_pi, L.src_pj, a, N[1] - l, b);
    case 5: return cell_at(L.src_pi, L.src_pj, a, b, l);
    default: return cell_at(L.src_pi, L.src_pj, a, b, N[2] - l);
  }
}

// ghost cell of layer l (1..3) of face f at tangential cell (a, b)
__device__ __forceinline__ long ghost_cell(const LinkDev &L, int l, int a, int b) {
  const int *N = L.dst_N;
  switch (L.dst_face) {
    case 1: return cell_at(L.dst_pi, L.dst_pj, 1 - l, a, b);
    case 2: return cell_at(L.dst_pi, L.dst_pj, N[0] + l - 1, a, b);
    case 3: return cell_at(L.dst_pi, L.dst_pj, a, 1 - l, b);
    case 4: return cell_at(L.dst_pi, L.dst_pj, a, N[1] + l - 1, b);
    case 5: return cell_at(L.dst_pi, L.dst_pj, a, b, 1 - l);
    default: return cell_at(L.dst_pi, L.dst_pj, a, b, N[2] + l - 1);
  }
}

// one thread per (layer, face cell) of one link (blockIdx.y); all variables
__global__ void __launch_bounds__(256) k_exchange(const LinkDev *links) {
  const LinkDev L = links[blockIdx.y];
  const int e = (int)(blockIdx.x * blockDim.x + threadIdx.x);
  if (e >= 3 * L.per_layer) return;
  const int l = e / L.per_layer + 1;
  const int t = e - (l - 1) * L.per_layer;
  // sender: natural order, first tangential fastest
  const int as = t % L.s_n1 + 1, bs = t / L.s_n1 + 1;
  // receiver: the t-th position of its loops
  const int inner = t % L.r_inner, outer = t / L.r_inner;
  int ar, br;
  if (L.dir_switch == 0) {
    ar = L.lo1 + inner * L.d1;
    br = L.lo2 + outer * L.d2;
  } else {
    ar = L.lo1 + outer * L.d1;
    br = L.lo2 + inner * L.d2;
  }
  const long cs = layer_cell(L, l, as, bs), cd = ghost_cell(L, l, ar, br);
  for (int n = 0; n < L.nv; ++n) L.dst[(long)n * L.dst_nc + cd] = L.src[(long)n * L.src_nc + cs];
}

// the transpose of one link's copy on the adjoint arrays (links[li] built
// like the exchange's, with src and dst at the sender's and the receiver's
// cell adjoints): each receiving ghost cell's adjoint is added to the
// sender's interior cell it was copied from, then cleared.  One link per
// launch: two links may read the same sender cell.
__global__ void __launch_bounds__(256) k_exchange_adjoint(const LinkDev *links, int li) {
  const LinkDev L = links[li];
  const int e = (int)(blockIdx.x * blockDim.x + threadIdx.x);
  if (e >= 3 * L.per_layer) return;
  const int l = e / L.per_layer + 1;
  const int t = e - (l - 1) * L.per_layer;
  const int as = t % L.s_n1 + 1, bs = t / L.s_n1 + 1;
  const int inner = t % L.r_inner, outer = t / L.r_inner;
  int ar, br;
  if (L.dir_switch == 0) {
    ar = L.lo1 + inner * L.d1;
    br = L.lo2 + outer * L.d2;
  } else {
    ar = L.lo1 + outer * L.d1;
    br = L.lo2 + inner * L.d2;
  }
  const long cs = layer_cell(L, l, as, bs), cd = ghost_cell(L, l, ar, br);
  double *src = const_cast<double *>(L.src);
  for (int n = 0; n < L.nv; ++n) {
    const double v = L.dst[(long)n * L.dst_nc + cd];
    L.dst[(long)n * L.dst_nc + cd] = 0.0;
    src[(long)n * L.src_nc + cs] += v;
  }
}

}  // namespace fest
