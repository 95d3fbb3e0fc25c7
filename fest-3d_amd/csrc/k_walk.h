// k_walk.h -- the work-list walk of the batched group kernels: a work
// item names a block and one of its per-block launch's tiles; a batched
// launch walks the list, each work-group taking items t, t + gridDim.x, ...
// (one item per work-group when the grid covers the list).
#pragma once
#include "blk.h"

namespace fest {

struct WorkItem {
  int blk, x, y, z;
};

// the tile walk; the barrier after each tile keeps a body's LDS tiles
// (the i sweep's stencil rows) from being overwritten while read
#define FEST_WALK(BODY_CALL)                                          \
  for (int t = (int)blockIdx.x; t < n; t += (int)gridDim.x) {          \
    const WorkItem it = items[t];                                      \
    const Blk &b = blks[it.blk];                                       \
    const Wg wg{it.x, it.y, it.z, 0};                                  \
    BODY_CALL;                                                         \
    __syncthreads();                                                   \
  }

}  // namespace fest
