// k_flux_walk.h -- the face sweeps of a batched group step: one launch of
// sweep D over the work list of all blocks' tiles, running k_flux_b on
// each tile (TI: the j/k sweeps' face-tile width).
#pragma once
#include "k_flux.h"
#include "k_walk.h"

namespace fest {

template <int D, int NV, int RECON, bool FLAT, typename T, int TI = TILE_I>
__global__ void __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(4)))
km_flux(const Blk *__restrict__ blks, const WorkItem *__restrict__ items, int n, int store_csnd) {
  FEST_WALK((k_flux_b<D, NV, RECON, FLAT, T, TI>(b, wg, store_csnd)))
}

}  // namespace fest
