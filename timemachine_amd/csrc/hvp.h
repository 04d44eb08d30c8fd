// Hessian-vector products of the potentials (hvp.hip): tangents of du/dx,
// du/dp and du/dlambda along a direction (vx, vp, vl) of coordinates,
// parameters and lambda, accumulated into caller-zeroed fixed-point buffers
// hx [N, 3] (2^36), hp [P] (the potential's du_dp exponents) and hl [1]
// (2^36). vx or vp may be null (a zero tangent); a null output is not
// requested.
#pragma once

#include "bonded.h"
#include "nb_state.h"

namespace tmamd {

void launch_bonded_hvp(BondedKind kind, const BondedTerms &t, const double *x, const double *vx, const double *vp,
                       double lambda, double vl, uint64_t *hx, uint64_t *hp, uint64_t *hl, hipStream_t stream);

// One system, plain parameters (a.params1 and a.pair_sys null).
void launch_pair_list_hvp(const PairListArgs &a, const double *vx, const double *vp, double vl, uint64_t *hx,
                          uint64_t *hp, uint64_t *hl, hipStream_t stream);

// Pairs of the tile list of the last prepare() of a one-system state without
// interpolated parameters; params [N, 3] and offset [N] (lambda offsets, may be
// null) in original atom order.
void launch_tiles_hvp(const NbState &st, const double *params, const int *offset, const double *boxes,
                      const double *vx, const double *vp, double vl, uint64_t *hx, uint64_t *hp, uint64_t *hl,
                      hipStream_t stream);

} // namespace tmamd
