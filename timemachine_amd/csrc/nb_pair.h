// Per-pair nonbonded arithmetic shared by the tile kernel and the pair-list
// (exclusion) kernel.
//
// Physics restated from the reference's FP64 JAX potential
// (timemachine/potentials/nonbonded.py:133-264) and its FP32 kernel split
// (timemachine/cpp/src/kernels/nonbonded_common.cuh:31-140):
//   u_es = s_q q_i q_j erfc(beta d) / d
//   u_lj = s_lj 4 eps_ij ((sig_ij/d)^12 - (sig_ij/d)^6),  sig_ij = sig_i + sig_j, eps_ij = eps_i eps_j
// with the per-atom parameter convention [q*sqrt(ONE_4PI_EPS0), sig/2, sqrt(eps)].
// LJ contributes only when both eps_i and eps_j are nonzero
// (nonbonded.py:214-225 masks eps_ij == 0 pairs, gradients included).
//
// Exclusions cancel exactly only if both kernels produce identical bits for
// the same pair, so every kernel that evaluates a pair calls these functions,
// and all translation units are compiled with -ffp-contract=off.
#pragma once

#include <cmath>

#include "fixed_point.h"
#include "hd.h"

namespace tmamd {

// erfc(x) = exp(-x^2) * P(1 / (1 + ERFC_P x)); coefficients from erfc_fit.py
// (approximation error 5.4e-9 relative on [0, 3], below FP32 rounding).
constexpr float ERFC_P = 0.3400000035762787f;
constexpr float ERFC_C0 = 0.0032670979853719473f;
constexpr float ERFC_C1 = 0.15022893249988556f;
constexpr float ERFC_C2 = 0.4216855764389038f;
constexpr float ERFC_C3 = -0.5369299054145813f;
constexpr float ERFC_C4 = 1.532942295074463f;
constexpr float ERFC_C5 = -1.497012972831726f;
constexpr float ERFC_C6 = 1.1914864778518677f;
constexpr float ERFC_C7 = -0.26566749811172485f;
constexpr float TWO_OVER_SQRT_PI_F = 1.1283791670955126f;

TM_HD inline float min_image_f(float d, float box, float inv_box) { return d - box * std::rint(d * inv_box); }

struct PairTerms {
    float prefactor; // (dU/dd) / d : du/dx_i = prefactor * (x_i - x_j)
    float u;         // pair energy (valid when requested)
    float erfc_inv_d; // erfc(beta d) / d, the charge-derivative factor
    float sig_grad;  // dU/dsig_ij (LJ pairs only)
    float eps_grad;  // dU/deps_ij (LJ pairs only)
    bool lj;
};

template <bool COMPUTE_U, bool COMPUTE_DP>
TM_HD inline PairTerms nb_pair_terms(
    float d2,
    float qi,
    float qj,
    float sig_i,
    float sig_j,
    float eps_i,
    float eps_j,
    float beta,
    float scale_q,
    float scale_lj) {
    PairTerms o;
    const float inv_d = TM_RSQF(d2);
    const float d = d2 * inv_d;
    const float inv_d2 = inv_d * inv_d;

    // erfc(beta d) and exp(-(beta d)^2), the latter with the rounding error of
    // x^2 folded back in (x2lo is the exact low part of x*x).
    const float x = beta * d;
    const float x2 = x * x;
    const float x2lo = __builtin_fmaf(x, x, -x2);
    float e = std::exp(-x2);
    e = e - e * x2lo;
    const float t = TM_RCPF(1.0f + ERFC_P * x);
    float poly = ERFC_C7;
    poly = poly * t + ERFC_C6;
    poly = poly * t + ERFC_C5;
    poly = poly * t + ERFC_C4;
    poly = poly * t + ERFC_C3;
    poly = poly * t + ERFC_C2;
    poly = poly * t + ERFC_C1;
    poly = poly * t + ERFC_C0;
    const float erfc_x = e * poly;

    const float sq = scale_q * (qi * qj);
    const float es_factor = -inv_d2 * (TWO_OVER_SQRT_PI_F * x * e + erfc_x);
    o.erfc_inv_d = inv_d * erfc_x;
    o.prefactor = sq * inv_d * es_factor;
    o.u = 0.0f;
    if (COMPUTE_U) {
        o.u = sq * o.erfc_inv_d;
    }
    o.sig_grad = 0.0f;
    o.eps_grad = 0.0f;
    o.lj = (eps_i != 0.0f) && (eps_j != 0.0f);
    if (o.lj) {
        const float eps_ij = eps_i * eps_j;
        const float sig_ij = sig_i + sig_j;
        const float s1 = sig_ij * inv_d;
        const float s2 = s1 * s1;
        const float s4 = s2 * s2;
        const float s6 = s4 * s2;
        const float se = scale_lj * eps_ij;
        o.prefactor = o.prefactor - se * (s6 * inv_d2) * (s6 * 48.0f - 24.0f);
        if (COMPUTE_U || COMPUTE_DP) {
            const float lj_shape = (s6 - 1.0f) * s6;
            if (COMPUTE_U) {
                o.u = o.u + 4.0f * se * lj_shape;
            }
            if (COMPUTE_DP) {
                o.sig_grad = 24.0f * se * (sig_ij * s4 * inv_d2) * (2.0f * s6 - 1.0f);
                o.eps_grad = 4.0f * scale_lj * lj_shape;
            }
        }
    }
    return o;
}

} // namespace tmamd
