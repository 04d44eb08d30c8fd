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
    // nonbonded kernels only: du/dp itself [A, 3] (the values the tangents
    // differentiate, exact FP64 derivatives in the du_dp exponents), or null
    uint64_t *gp = nullptr;
};

// t.term_sys, lambdas and red_idx as launch_bonded (null: one system at lambda0, index 0).
void launch_bonded_hvp(BondedKind kind, const BondedTerms &t, const double *x, const double *lambdas, double lambda0,
                       const int *red_idx, const HvpTangent &v, const HvpOut &o, hipStream_t stream);

// Pair list with plain parameters (a.params1 and a.slots null); one or many systems.
void launch_pair_list_hvp(const PairListArgs &a, const HvpTangent &v, const HvpOut &o, hipStream_t stream);

// Pairs of the tile lists of the last prepare(). params [A, 3] are the values
// the pairs use (interpolated ones evaluated at lambda by the caller, see
// below) and offset [A] the lambda offsets (may be null), in global atom
// order; boxes [S, 9].
void launch_tiles_hvp(const NbState &st, const double *params, const int *offset, const double *boxes,
                      const HvpTangent &v, const HvpOut &o, hipStream_t stream);

// Interpolated nonbonded parameters (params [2N, 3]: p0 then p1, used as
// p(lambda) = (1 - lambda) p0 + lambda p1). The product is the one of the
// plain-parameter kernels at p(lambda) along the parameter tangent
// vp(lambda) = (1 - lambda) vp0 + lambda vp1 + (p1 - p0) vl, followed by the
// chain rule through p(lambda): with g the du/dp at p(lambda) and dg its
// tangent,
//   hp0 = (1 - lambda) dg - vl g,   hp1 = lambda dg + vl g,
//   hl += sum over atoms and columns of dg (p1 - p0) + g (vp1 - vp0).
// Device scratch of such a product: p(lambda) and its tangent (par [2, N, 3]),
// the du/dp at p(lambda) and its tangent (grad [2, N, 3]).
struct InterpHvpScratch {
    DeviceArray<double> par;
    DeviceArray<uint64_t> grad;
    void reserve(int N) {
        const size_t n = static_cast<size_t>(N) * 6;
        if (par.n != n) {
            par.alloc(n);
            grad.alloc(n);
        }
    }
};
// First step: p_lam [N, 3] and vp_lam [N, 3] (vp [2N, 3] may be null: zero).
void launch_interp_hvp_params(int N, double lambda, const double *params, const double *vp, double vl, double *p_lam,
                              double *vp_lam, hipStream_t stream);
// Last step: g (HvpOut::gp of the kernels) and dg [N, 3] in the nonbonded
// du_dp exponents; adds to hp
// [2N, 3] and hl [1] (either may be null).
void launch_interp_hvp_chain(int N, double lambda, const double *params, const double *vp, double vl, const uint64_t *g,
                             const uint64_t *dg, uint64_t *hp, uint64_t *hl, hipStream_t stream);

} // namespace tmamd
