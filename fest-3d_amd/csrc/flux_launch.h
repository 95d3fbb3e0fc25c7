// flux_launch.h -- launchers of the face-sweep kernels (k_flux, km_flux),
// compiled in their own translation unit (flux_launch.hip): the sweeps'
// many specialisations (direction x variables x reconstruction x
// precision x tile) then build beside the rest of the library.
#pragma once
#include <hip/hip_runtime.h>

#include "blk.h"

namespace fest {

struct WorkItem;

// sweep D of one block: k_flux<D, NV, RECON, FLAT, T> for the block's
// variable count, reconstruction and precision (FLAT on the k faces of a
// 2-D grid), grid g of blocks t, on stream st
template <int D>
void launch_flux(const Blk &b, dim3 g, dim3 t, hipStream_t st, int store_csnd);

// sweep D over the work list of a batched group step (n items, `grid`
// work-groups; the j and k sweeps in tile_i x 256/tile_i face tiles); b0 is
// any block of the group (they share one kernel variant)
template <int D>
void launch_flux_batched(const Blk &b0, hipStream_t st, const Blk *blks, const WorkItem *items, int n,
                         unsigned grid, int tile_i, int store_csnd);

extern template void launch_flux<0>(const Blk &, dim3, dim3, hipStream_t, int);
extern template void launch_flux<1>(const Blk &, dim3, dim3, hipStream_t, int);
extern template void launch_flux<2>(const Blk &, dim3, dim3, hipStream_t, int);
extern template void launch_flux_batched<0>(const Blk &, hipStream_t, const Blk *, const WorkItem *, int, unsigned,
                                            int, int);
extern template void launch_flux_batched<1>(const Blk &, hipStream_t, const Blk *, const WorkItem *, int, unsigned,
                                            int, int);
extern template void launch_flux_batched<2>(const Blk &, hipStream_t, const Blk *, const WorkItem *, int, unsigned,
                                            int, int);

}  // namespace fest
