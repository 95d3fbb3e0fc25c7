// k_multi.h -- the stage kernels of every block of a multiblock group in one
// launch per phase.  A group keeps a device array of its blocks' Blk
// records and, per phase, a work list of tiles (block, x, y, z): the
// per-block launch's block indices, concatenated over the blocks.  One
// launch walks the list; with fewer work-groups than tiles each work-group
// takes tiles t, t + gridDim.x, ... in turn (a persistent launch).  The
// tile bodies are the per-block kernels' own (k_*_b), so a tile computes
// exactly what the per-block launch computes for it.  The face sweeps'
// batched kernel, km_flux, is in k_flux_walk.h.
#pragma once
#include "blk.h"
#include "k_bc.h"
#include "k_flux.h"
#include "k_grad.h"
#include "k_update.h"
#include "k_walk.h"

namespace fest {

__global__ void __launch_bounds__(256) km_step_begin(const Blk *__restrict__ blks,
                                                     const WorkItem *__restrict__ items, int n, int store_u) {
  FEST_WALK(k_step_begin_b(b, wg, store_u))
}
__global__ void __launch_bounds__(256) km_bc_face(const Blk *__restrict__ blks, const WorkItem *__restrict__ items,
                                                  int n, int face0, int current_iter) {
  FEST_WALK(k_bc_face_b(b, wg, face0, current_iter))
}
__global__ void __launch_bounds__(256) km_bc_edges(const Blk *__restrict__ blks, const WorkItem *__restrict__ items,
                                                   int n, int phase) {
  FEST_WALK(k_bc_edges_b(b, wg, phase))
}
template <int NG, bool S2003, bool F32 = false>
__global__ void __launch_bounds__(256) km_grad(const Blk *__restrict__ blks, const WorkItem *__restrict__ items,
                                               int n) {
  FEST_WALK((k_grad_b<NG, S2003, false, false, F32>(b, wg)))
}
__global__ void __launch_bounds__(256) km_grad_bc(const Blk *__restrict__ blks, const WorkItem *__restrict__ items,
                                                  int n) {
  FEST_WALK(k_grad_bc_b(b, wg))
}
__global__ void __launch_bounds__(256) km_dt(const Blk *__restrict__ blks, const WorkItem *__restrict__ items,
                                             int n) {
  FEST_WALK(k_dt_b(b, wg))
}
__global__ void __launch_bounds__(256) km_update(const Blk *__restrict__ blks, const WorkItem *__restrict__ items,
                                                 int n, double TF, double SF, int flags) {
  FEST_WALK(k_update_b(b, wg, TF, SF, flags))
}
__global__ void __launch_bounds__(256) km_to_f32(const Blk *__restrict__ blks, const WorkItem *__restrict__ items,
                                                 int n) {
  FEST_WALK(k_to_f32_b(b, wg))
}

}  // namespace fest
