// flux_launch.hip -- launchers of the face sweeps (declared in
// flux_launch.h): per-block k_flux and batched km_flux, specialised on the
// variable count, reconstruction, precision and, for batched j/k sweeps,
// the face-tile shape.
#include <hip/hip_runtime.h>

#include "flux_launch.h"
#include "k_flux.h"
#include "k_flux_walk.h"

namespace fest {
namespace {

// the flux kernel of sweep D, specialised on the variable count and the
// reconstruction
template <int D, int NV, bool FLAT, typename T>
void launch_flux_t(const Blk &b, dim3 g, dim3 t, hipStream_t st, int store_csnd) {
  switch (b.recon) {
    case 0: hipLaunchKernelGGL((k_flux<D, NV, 0, FLAT, T>), g, t, 0, st, b, store_csnd); break;
    case 1: hipLaunchKernelGGL((k_flux<D, NV, 1, FLAT, T>), g, t, 0, st, b, store_csnd); break;
    case 2: hipLaunchKernelGGL((k_flux<D, NV, 2, FLAT, T>), g, t, 0, st, b, store_csnd); break;
    case 3: hipLaunchKernelGGL((k_flux<D, NV, 3, FLAT, T>), g, t, 0, st, b, store_csnd); break;
    default: hipLaunchKernelGGL((k_flux<D, NV, 4, FLAT, T>), g, t, 0, st, b, store_csnd); break;
  }
}
// FP64 face arithmetic, or FP32 when the handle was created with
// FEST_PREC_FP32_FLUX
template <int D, int NV, bool FLAT>
void launch_flux_nv(const Blk &b, dim3 g, dim3 t, hipStream_t st, int store_csnd) {
  if constexpr (NV == 8) {
    launch_flux_t<D, NV, FLAT, double>(b, g, t, st, store_csnd);  // lctm2015 runs in FP64 only
  } else {
    if (b.fp32) launch_flux_t<D, NV, FLAT, float>(b, g, t, st, store_csnd);
    else launch_flux_t<D, NV, FLAT, double>(b, g, t, st, store_csnd);
  }
}
// the k faces of a 2-D grid (one cell in k) carry no convective flux
template <int D, int NV>
void launch_flux_v(const Blk &b, dim3 g, dim3 t, hipStream_t st, int store_csnd) {
  if constexpr (D == 2) {
    if (b.kmx == 2) {
      launch_flux_nv<D, NV, true>(b, g, t, st, store_csnd);
      return;
    }
  }
  launch_flux_nv<D, NV, false>(b, g, t, st, store_csnd);
}
}  // namespace

template <int D>
void launch_flux(const Blk &b, dim3 g, dim3 t, hipStream_t st, int store_csnd) {
  if (b.nv == 8) launch_flux_v<D, 8>(b, g, t, st, store_csnd);
  else if (b.nv == 7) launch_flux_v<D, 7>(b, g, t, st, store_csnd);
  else if (b.nv == 6) launch_flux_v<D, 6>(b, g, t, st, store_csnd);
  else launch_flux_v<D, 5>(b, g, t, st, store_csnd);
}

namespace {

template <int D, int NV, int RECON, typename T>
void km_flux_launch(hipStream_t st, const Blk *blks, const WorkItem *items, int n, unsigned grid, int tile_i,
                    int store_csnd) {
  // the j and k sweeps in 64 x 4 or 32 x 8 face tiles
  if constexpr (D != 0) {
    if (tile_i == 32) {
      hipLaunchKernelGGL((km_flux<D, NV, RECON, false, T, 32>), dim3(grid), dim3(256), 0, st, blks, items, n,
                         store_csnd);
      return;
    }
  }
  hipLaunchKernelGGL((km_flux<D, NV, RECON, false, T>), dim3(grid), dim3(256), 0, st, blks, items, n, store_csnd);
}
template <int D, int NV, typename T>
void km_flux_recon(int recon, hipStream_t st, const Blk *blks, const WorkItem *items, int n, unsigned grid,
                   int tile_i, int store_csnd) {
  switch (recon) {
    case 0: km_flux_launch<D, NV, 0, T>(st, blks, items, n, grid, tile_i, store_csnd); break;
    case 1: km_flux_launch<D, NV, 1, T>(st, blks, items, n, grid, tile_i, store_csnd); break;
    case 2: km_flux_launch<D, NV, 2, T>(st, blks, items, n, grid, tile_i, store_csnd); break;
    case 3: km_flux_launch<D, NV, 3, T>(st, blks, items, n, grid, tile_i, store_csnd); break;
    default: km_flux_launch<D, NV, 4, T>(st, blks, items, n, grid, tile_i, store_csnd); break;
  }
}

}  // namespace

// batched groups hold laminar or SST blocks (5 or 7 variables)
template <int D>
void launch_flux_batched(const Blk &b0, hipStream_t st, const Blk *blks, const WorkItem *items, int n,
                         unsigned grid, int tile_i, int store_csnd) {
  if (b0.nv == 7) {
    if (b0.fp32) km_flux_recon<D, 7, float>(b0.recon, st, blks, items, n, grid, tile_i, store_csnd);
    else km_flux_recon<D, 7, double>(b0.recon, st, blks, items, n, grid, tile_i, store_csnd);
  } else {
    if (b0.fp32) km_flux_recon<D, 5, float>(b0.recon, st, blks, items, n, grid, tile_i, store_csnd);
    else km_flux_recon<D, 5, double>(b0.recon, st, blks, items, n, grid, tile_i, store_csnd);
  }
}

template void launch_flux<0>(const Blk &, dim3, dim3, hipStream_t, int);
template void launch_flux<1>(const Blk &, dim3, dim3, hipStream_t, int);
template void launch_flux<2>(const Blk &, dim3, dim3, hipStream_t, int);
template void launch_flux_batched<0>(const Blk &, hipStream_t, const Blk *, const WorkItem *, int, unsigned, int,
                                     int);
template void launch_flux_batched<1>(const Blk &, hipStream_t, const Blk *, const WorkItem *, int, unsigned, int,
                                     int);
template void launch_flux_batched<2>(const Blk &, hipStream_t, const Blk *, const WorkItem *, int, unsigned, int,
                                     int);

}  // namespace fest
