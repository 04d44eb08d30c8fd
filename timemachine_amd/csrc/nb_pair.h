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
// the same pair, so every kernel that evaluates a pair calls these functions
// with the same inputs (FP64 coordinates and w, FP32 parameters), and all
// translation units are compiled with -ffp-contract=off.
#pragma once

#include <cmath>

#include "fixed_point.h"
#include "hd.h"

namespace tmamd {

// erfc(x) = exp(-x^2) * P(1 / (1 + ERFC_P x)); coefficients from erfc_fit.py
// (approximation error 5.4e-9 relative on [0, 3]).
constexpr double ERFC_P = 0.3400000035762787;
constexpr double ERFC_C0 = 0.0032670979853719473;
constexpr double ERFC_C1 = 0.15022893249988556;
constexpr double ERFC_C2 = 0.4216855764389038;
constexpr double ERFC_C3 = -0.5369299054145813;
constexpr double ERFC_C4 = 1.532942295074463;
constexpr double ERFC_C5 = -1.497012972831726;
constexpr double ERFC_C6 = 1.1914864778518677;
constexpr double ERFC_C7 = -0.26566749811172485;
constexpr double TWO_OVER_SQRT_PI = 1.1283791670955126;

TM_HD inline float min_image_f(float d, float box, float inv_box) { return d - box * std::rint(d * inv_box); }
// d - box * rint(d / box) with one rounding for the product and difference.
TM_HD inline double min_image_d(double d, double box, double inv_box) {
    return __builtin_fma(-box, std::rint(d * inv_box), d);
}

// Squared 3D distance, dz^2 + dy^2 + dx^2 with fused multiply-adds.
TM_HD inline double pair_d2(double dx, double dy, double dz) {
    return __builtin_fma(dx, dx, __builtin_fma(dy, dy, dz * dz));
}

// 1/sqrt(a): FP32 hardware seed, one Newton step in FP64 (relative error ~1e-14).
TM_HD inline double rsqrt_nr(double a) {
    const double y = static_cast<double>(TM_RSQF(static_cast<float>(a)));
    return y * (1.5 - (0.5 * a) * (y * y));
}

// 1/a: FP32 hardware seed, one Newton step in FP64.
TM_HD inline double rcp_nr(double a) {
    const double y = static_cast<double>(TM_RCPF(static_cast<float>(a)));
    return y * (2.0 - a * y);
}

constexpr double LOG2_E = 1.4426950408889634;
constexpr double LN_2 = 0.6931471805599453;

// exp(-a), a >= 0: 2^(-t) with t = a log2(e) in FP64; the FP32 hardware 2^x of
// t rounded to FP32, times 2^(-r) ~ 1 - r ln2 for the rounding remainder r
// (|r ln2| < 4e-7 for a < 9, so the dropped second-order term is < 1e-13).
// Relative error ~1 FP32 ulp, from the hardware 2^x.
TM_HD inline double exp_neg(double a) {
    const double t = a * LOG2_E;
    const float tf = static_cast<float>(t);
    const double e = static_cast<double>(TM_EXP2F(-tf));
    return e - e * ((t - static_cast<double>(tf)) * LN_2);
}

// P(t) = sum_k ERFC_Ck t^k in Estrin's form: four independent first-degree
// pieces, then t^2 and t^4 (9 operations, depth 4 instead of Horner's 7).
TM_HD inline double erfc_poly(double t) {
    const double t2 = t * t;
    const double t4 = t2 * t2;
    const double p01 = __builtin_fma(ERFC_C1, t, ERFC_C0);
    const double p23 = __builtin_fma(ERFC_C3, t, ERFC_C2);
    const double p45 = __builtin_fma(ERFC_C5, t, ERFC_C4);
    const double p67 = __builtin_fma(ERFC_C7, t, ERFC_C6);
    return __builtin_fma(__builtin_fma(p67, t2, p45), t4, __builtin_fma(p23, t2, p01));
}

struct PairTerms {
    double prefactor;  // (dU/dd) / d : du/dx_i = prefactor * (x_i - x_j)
    double u;          // pair energy (valid when requested)
    double erfc_inv_d; // erfc(beta d) / d, the charge-derivative factor
    double sig_grad;   // dU/dsig_ij (LJ pairs only)
    double eps_grad;   // dU/deps_ij (LJ pairs only)
    bool lj;
};

// Pair terms from the squared (4D) distance. Mixed precision chosen for the
// 1e-5 per-atom force target: FP64 everywhere except exp(-x^2) (exp_neg,
// relative error ~1 FP32 ulp) and the FP32 rsq/rcp seeds that one FP64
// Newton step refines. Parameters arrive as FP32 values (the
// reference's per-atom parameter precision).
template <bool COMPUTE_U, bool COMPUTE_DP>
TM_HD inline PairTerms nb_pair_terms(
    double d2,
    double qi,
    double qj,
    double sig_i,
    double sig_j,
    double eps_i,
    double eps_j,
    double beta,
    double scale_q,
    double scale_lj) {
    PairTerms o;
    const double inv_d = rsqrt_nr(d2);
    const double d = d2 * inv_d;
    const double inv_d2 = inv_d * inv_d;

    const double x = beta * d;
    const double e = exp_neg(x * x);
    const double poly = erfc_poly(rcp_nr(1.0 + ERFC_P * x));
    const double erfc_x = e * poly;

    const double sq = scale_q * (qi * qj);
    o.erfc_inv_d = inv_d * erfc_x;
    o.prefactor = -(sq * inv_d * inv_d2) * ((TWO_OVER_SQRT_PI * x + poly) * e);
    o.u = 0.0;
    if (COMPUTE_U) {
        o.u = sq * o.erfc_inv_d;
    }
    o.sig_grad = 0.0;
    o.eps_grad = 0.0;
    o.lj = (eps_i != 0.0) && (eps_j != 0.0);
    if (o.lj) {
        const double eps_ij = eps_i * eps_j;
        const double sig_ij = sig_i + sig_j;
        const double s1 = sig_ij * inv_d;
        const double s2 = s1 * s1;
        const double s4 = s2 * s2;
        const double s6 = s4 * s2;
        const double se = scale_lj * eps_ij;
        o.prefactor = o.prefactor - se * (s6 * inv_d2) * (s6 * 48.0 - 24.0);
        if (COMPUTE_U || COMPUTE_DP) {
            const double lj_shape = (s6 - 1.0) * s6;
            if (COMPUTE_U) {
                o.u = o.u + 4.0 * se * lj_shape;
            }
            if (COMPUTE_DP) {
                o.sig_grad = 24.0 * se * (sig_ij * s4 * inv_d2) * (2.0 * s6 - 1.0);
                o.eps_grad = 4.0 * scale_lj * lj_shape;
            }
        }
    }
    return o;
}

// The same pair terms with FP32 arithmetic after the FP64 squared distance
// (the reference's FP32 kernel split, nonbonded_common.cuh:31-140, with this
// file's erfc form evaluated in FP32): one FP32 rsq / rcp / 2^x each, no
// Newton steps, FP32 products. Coordinates stay FP64 up to d^2, which keeps
// the periodic minimum image exact. Results are returned widened to FP64.
template <bool COMPUTE_U, bool COMPUTE_DP>
TM_HD inline PairTerms nb_pair_terms_f32(
    double d2_64,
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
    const float d2 = static_cast<float>(d2_64);
    const float inv_d = TM_RSQF(d2);
    const float d = d2 * inv_d;
    const float inv_d2 = inv_d * inv_d;
    const float x = beta * d;
    const float e = TM_EXP2F(-(x * x) * static_cast<float>(LOG2_E));
    const float t = TM_RCPF(__builtin_fmaf(static_cast<float>(ERFC_P), x, 1.0f));
    float poly = static_cast<float>(ERFC_C7);
    poly = __builtin_fmaf(poly, t, static_cast<float>(ERFC_C6));
    poly = __builtin_fmaf(poly, t, static_cast<float>(ERFC_C5));
    poly = __builtin_fmaf(poly, t, static_cast<float>(ERFC_C4));
    poly = __builtin_fmaf(poly, t, static_cast<float>(ERFC_C3));
    poly = __builtin_fmaf(poly, t, static_cast<float>(ERFC_C2));
    poly = __builtin_fmaf(poly, t, static_cast<float>(ERFC_C1));
    poly = __builtin_fmaf(poly, t, static_cast<float>(ERFC_C0));
    const float erfc_x = e * poly;
    const float sq = scale_q * (qi * qj);
    const float erfc_inv_d = inv_d * erfc_x;
    float pre = -(sq * inv_d * inv_d2) * (__builtin_fmaf(static_cast<float>(TWO_OVER_SQRT_PI), x, poly) * e);
    float u = 0.0f;
    if (COMPUTE_U) {
        u = sq * erfc_inv_d;
    }
    o.sig_grad = 0.0;
    o.eps_grad = 0.0;
    o.lj = (eps_i != 0.0f) && (eps_j != 0.0f);
    if (o.lj) {
        const float eps_ij = eps_i * eps_j;
        const float sig_ij = sig_i + sig_j;
        const float s1 = sig_ij * inv_d;
        const float s2 = s1 * s1;
        const float s4 = s2 * s2;
        const float s6 = s4 * s2;
        const float se = scale_lj * eps_ij;
        pre = pre - se * (s6 * inv_d2) * __builtin_fmaf(s6, 48.0f, -24.0f);
        if (COMPUTE_U || COMPUTE_DP) {
            const float lj_shape = (s6 - 1.0f) * s6;
            if (COMPUTE_U) {
                u = u + 4.0f * se * lj_shape;
            }
            if (COMPUTE_DP) {
                o.sig_grad = 24.0f * se * (sig_ij * s4 * inv_d2) * (2.0f * s6 - 1.0f);
                o.eps_grad = 4.0f * scale_lj * lj_shape;
            }
        }
    }
    o.prefactor = pre;
    o.u = u;
    o.erfc_inv_d = erfc_inv_d;
    return o;
}

// du/dlambda of one pair: the 4D term dw * prefactor * (dw/dl_i - dw/dl_j)
// and, with interpolated parameters (dp/dl = p1 - p0), the parameter terms
// (k_nonbonded.cuh:464-507): du/dq_i dq_i/dl + ... for q, sig and eps.
// T must carry sig_grad / eps_grad (COMPUTE_DP) when interp is true.
TM_HD inline double pair_du_dl(const PairTerms &T, double dw, double dwdl_i, double dwdl_j, bool interp,
                               double scale_q, double qi, double qj, double ei, double ej, double dqi, double dqj,
                               double dsi, double dsj, double dei, double dej) {
    double dl = dw * T.prefactor * (dwdl_i - dwdl_j);
    if (interp) {
        dl = dl + T.erfc_inv_d * ((scale_q * qj) * dqi + (scale_q * qi) * dqj);
        if (T.lj) {
            dl = dl + T.sig_grad * (dsi + dsj) + T.eps_grad * (ej * dei + ei * dej);
        }
    }
    return dl;
}

// Parameter derivatives of the pair's force prefactor, for d(du/dlambda)/dp.
// A pair contributes prefactor * dw * (dw/dl_i - dw/dl_j) to du/dlambda, and
// the prefactor is linear in q_i q_j (Coulomb) and in eps_i eps_j (LJ):
//   es_unit = prefactor_es / (scale_q q_i q_j)
//   lj_unit = prefactor_lj / (eps_i eps_j)       (LJ pairs only; includes scale_lj)
//   lj_dsig = d prefactor_lj / d sig_ij          (LJ pairs only)
// Same arithmetic as nb_pair_terms for the shared parts.
struct PairDlDp {
    double es_unit;
    double lj_unit;
    double lj_dsig;
    bool lj;
};

TM_HD inline PairDlDp nb_pair_dl_dp(double d2, double sig_i, double sig_j, double eps_i, double eps_j, double beta,
                                    double scale_lj) {
    PairDlDp o;
    const double inv_d = rsqrt_nr(d2);
    const double d = d2 * inv_d;
    const double inv_d2 = inv_d * inv_d;

    const double x = beta * d;
    const double e = exp_neg(x * x);
    const double poly = erfc_poly(rcp_nr(1.0 + ERFC_P * x));
    o.es_unit = -(inv_d * inv_d2) * ((TWO_OVER_SQRT_PI * x + poly) * e);
    o.lj_unit = 0.0;
    o.lj_dsig = 0.0;
    o.lj = (eps_i != 0.0) && (eps_j != 0.0);
    if (o.lj) {
        const double sig_ij = sig_i + sig_j;
        const double s1 = sig_ij * inv_d;
        const double s2 = s1 * s1;
        const double s4 = s2 * s2;
        const double s6 = s4 * s2;
        // prefactor_lj = -se s^6 / d^2 (48 s^6 - 24), s = sig_ij / d, se = scale_lj eps_i eps_j
        o.lj_unit = -scale_lj * (s6 * inv_d2) * (s6 * 48.0 - 24.0);
        // d/dsig_ij: s^6 -> 6 s^5 / d, s^12 -> 12 s^11 / d
        o.lj_dsig = -(scale_lj * (eps_i * eps_j)) * (inv_d2 * inv_d) * (s4 * s1) * (s6 * 576.0 - 144.0);
    }
    return o;
}

} // namespace tmamd
