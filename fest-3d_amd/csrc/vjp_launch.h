// vjp_launch.h -- launcher of the adjoint of the inviscid residual
// (k_vjp.h), built in its own translation unit (vjp.hip).
#pragma once
#include <hip/hip_runtime.h>

#include "blk.h"

namespace fest {

// doubles of the face buffer of sweep direction d: the 30 contributions of
// each face, then its 10 face-state adjoints
long vjp_contrib_len(const Blk &b, int d);

// bar (5 planes of the padded cell array) = d(barR . R)/dq on the resident
// state, whose ghost cells the block's own boundary rules filled: the face
// Jacobians contracted with barR (two launches per direction), the per-cell
// gather, then
// the transposes of the ghost rules, face by face in reverse order.
// contrib[d] holds vjp_contrib_len(b, d) doubles.
void launch_vjp(const Blk &b, hipStream_t st, const double *barR, double *const contrib[3], double *bar);

}  // namespace fest
