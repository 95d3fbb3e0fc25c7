// k_bc_kernels.h -- the per-block kernels of the ghost-cell passes, whose
// bodies are in k_bc.h.  Not templates: one translation unit includes this.
#pragma once
#include "k_bc.h"

namespace fest {

__global__ void k_bc_face(Blk b, int face0, int current_iter) {
  k_bc_face_b(b, FEST_WG_HW, face0, current_iter);
}
__global__ void k_bc_edges(Blk b, int phase) {
  k_bc_edges_b(b, FEST_WG_HW, phase);
}

}  // namespace fest
