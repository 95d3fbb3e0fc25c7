// vjp_launch.h -- launcher of the adjoint of the inviscid residual
// (k_vjp.h), built in its own translation unit (vjp.hip).
#pragma once
#include <hip/hip_runtime.h>

#include "blk.h"

namespace fest {

// doubles of the face buffer of sweep direction d: the 30 contributions of
// each face, its 10 face-state adjoints, and under the pressure-based
// switch its 24 pressure contributions from the other variables
long vjp_contrib_len(const Blk &b, int d);

// the node-coordinate part of the adjoint (all null when it is not wanted):
// the metric adjoints of each direction (4 planes of nfl[d]), the ghost-row
// I faces of slip walls on the j and k max faces (6 planes of the face
// interior each, null when that face is no slip wall) and the result
// (3 planes of the padded node array)
struct VjpNodes {
  double *mbar[3] = {nullptr, nullptr, nullptr};
  double *g3 = nullptr, *g5 = nullptr;
  double *bar_nodes = nullptr;
};

// the laminar viscous part (null when the block is inviscid): bt, the
// temperature adjoint (one plane of the padded cell array), gb, the gradient
// adjoints (12 planes of the gradient array, ngr each)
struct VjpVisc {
  double *bt = nullptr, *gb = nullptr;
};

// bar (5 planes of the padded cell array) = d(barR . R)/dq on the resident
// state, whose ghost cells the block's own boundary rules filled: the face
// Jacobians contracted with barR (two launches per direction), the per-cell
// gather, then the transposes of the ghost rules, face by face in reverse
// order; with nodes.bar_nodes, also d(barR . R)/d(node coordinates).  With
// visc.bt, the viscous faces' adjoints are added before the ghost rules'
// transposes (the viscous faces reuse contrib after the gather: 22 planes).
// contrib[d] holds vjp_contrib_len(b, d) doubles.
void launch_vjp(const Blk &b, hipStream_t st, const double *barR, double *const contrib[3], double *bar,
                const VjpNodes &nodes, const VjpVisc &visc);

}  // namespace fest
