// Hessian-vector products of the potentials (hvp.hip): tangents of du/dx,
// du/dp and du/dlambda along a direction (vx, vp, vl) of coordinates,
// parameters and lambda, accumulated into caller-zeroed fixed-point buffers
// hx [A, 3] (2^36), hp (the potential's du_dp layout and exponents) and hl
// (2^36, one word per reduction index). Every launcher takes one system or
// the batched layout of many (per-term / per-pair / per-tile system tables,
// as the force kernels).
#pragma once

#include "bonded.h"
#include "nb_state.h"

namespace tmamd {

struct HvpTangent {
    const double *vx = nullptr;  // [A, 3] or null (zero)
    const double *vp = nullptr;  // parameter tangent in the parameters' layout, or null (zero)
    double vl = 0.0;             // lambda tangent of every system when vls is null
    const double *vls = nullptr; // [S] per-system lambda tangents, or null
};

struct HvpOut {
    uint64_t *hx = nullptr; // [A, 3] or null (not requested)
    uint64_t *hp = nullptr; // parameters' layout or null
    uint64_t *hl = nullptr; // [reduction index] or null
};

// t.term_sys, lambdas and red_idx as launch_bonded (null: one system at lambda0, index 0).
void launch_bonded_hvp(BondedKind kind, const BondedTerms &t, const double *x, const double *lambdas, double lambda0,
                       const int *red_idx, const HvpTangent &v, const HvpOut &o, hipStream_t stream);

// Pair list with plain parameters (a.params1 and a.slots null); one or many systems.
void launch_pair_list_hvp(const PairListArgs &a, const HvpTangent &v, const HvpOut &o, hipStream_t stream);

// Pairs of the tile lists of the last prepare() (no interpolated parameters).
// params [A, 3] and offset [A] (lambda offsets, may be null) in global atom
// order; boxes [S, 9].
void launch_tiles_hvp(const NbState &st, const double *params, const int *offset, const double *boxes,
                      const HvpTangent &v, const HvpOut &o, hipStream_t stream);

} // namespace tmamd
