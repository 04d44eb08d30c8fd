// Host build of the arithmetic headers of timemachine_amd/csrc, used by the
// CPU tests to check the formulas (and Philox / fixed point bit patterns)
// before anything runs on a GPU. Built by tests/test_host_math.py with g++.
#include <cstdint>

#include <cmath>
#include <vector>

#include "bonded_terms.h"
#include "fire.h"
#include "fixed_point.h"
#include "nb_pair.h"
#include "philox.h"

using namespace tmamd;

extern "C" {

void hc_philox(const uint32_t *ctr, uint32_t k0, uint32_t k1, uint32_t *out) {
    Philox4 c{{ctr[0], ctr[1], ctr[2], ctr[3]}};
    Philox4 r = philox4x32_10(c, k0, k1);
    for (int i = 0; i < 4; i++) {
        out[i] = r.v[i];
    }
}

void hc_normal2(const uint32_t *bits, double *z) {
    Philox4 r{{bits[0], bits[1], bits[2], bits[3]}};
    normal2(r, z[0], z[1]);
}

uint64_t hc_round_to_fixed(double v) { return round_to_fixed(v); }

void hc_round_to_fixed3(double a, double b, double c, uint64_t *out) { round_to_fixed3(a, b, c, out[0], out[1], out[2]); }

uint64_t hc_f32_to_fixed(float v, float scale) { return f32_to_fixed(v, scale); }

// out: prefactor, u, erfc_inv_d, sig_grad, eps_grad, lj
void hc_pair_terms(double d2, double qi, double qj, double si, double sj, double ei, double ej, double beta, double sq,
                   double slj, double *out) {
    PairTerms t = nb_pair_terms<true, true>(d2, qi, qj, si, sj, ei, ej, beta, sq, slj);
    out[0] = t.prefactor;
    out[1] = t.u;
    out[2] = t.erfc_inv_d;
    out[3] = t.sig_grad;
    out[4] = t.eps_grad;
    out[5] = t.lj ? 1.0 : 0.0;
}

float hc_min_image(float d, float b) { return min_image_f(d, b, 1.0f / b); }
double hc_min_image_d(double d, double b) { return min_image_d(d, b, 1.0 / b); }

// out: g[2][3], dkb, db0, u, dl (FP64, as the kernels compute them)
void hc_bond(const double *xi, const double *xj, double kb, double b0, int mult, int offset, double lam, double *out) {
    BondOut<double> o = bond_term<double>(xi, xj, kb, b0, mult, lam_prefactor<double>(mult, offset, lam));
    for (int a = 0; a < 2; a++)
        for (int d = 0; d < 3; d++) out[a * 3 + d] = o.g[a][d];
    out[6] = o.dkb;
    out[7] = o.db0;
    out[8] = o.u;
    out[9] = o.dl;
}

// out: g[3][3], dka, da0, u, dl
void hc_angle(const double *xi, const double *xj, const double *xk, double ka, double a0, int mult, int offset,
              double lam, double *out) {
    AngleOut<double> o = angle_term<double>(xi, xj, xk, ka, a0, mult, lam_prefactor<double>(mult, offset, lam));
    for (int a = 0; a < 3; a++)
        for (int d = 0; d < 3; d++) out[a * 3 + d] = o.g[a][d];
    out[9] = o.dka;
    out[10] = o.da0;
    out[11] = o.u;
    out[12] = o.dl;
}

// out: g[4][3], dk, dphase, dperiod, u, dl
void hc_torsion(const double *xi, const double *xj, const double *xk, const double *xl, double k, double phase,
                double period, int mult, int offset, double lam, double *out) {
    TorsionOut<double> o =
        torsion_term<double>(xi, xj, xk, xl, k, phase, period, mult, lam_prefactor<double>(mult, offset, lam));
    for (int a = 0; a < 4; a++)
        for (int d = 0; d < 3; d++) out[a * 3 + d] = o.g[a][d];
    out[12] = o.dk;
    out[13] = o.dphase;
    out[14] = o.dperiod;
    out[15] = o.u;
    out[16] = o.dl;
}
}

// CPU model of the tile kernel's accumulation: every pair of `pairs` inside the
// cutoff adds fixed(prefactor * d) to i and subtracts it from j, exactly as the
// kernel does (min-image and d2 in FP32 on FP32-rounded coordinates).
extern "C" void hc_nonbonded_du_dx(int N, const double *x, const double *params, const double *box, double beta,
                                   double cutoff, int M, const int64_t *pairs, uint64_t *du_dx, uint64_t *u) {
    const double bx = box[0], by = box[4], bz = box[8];
    const double ibx = 1.0 / bx, iby = 1.0 / by, ibz = 1.0 / bz;
    for (int m = 0; m < M; m++) {
        const int i = (int)pairs[2 * m], j = (int)pairs[2 * m + 1];
        const double dx = min_image_d(x[3 * i] - x[3 * j], bx, ibx);
        const double dy = min_image_d(x[3 * i + 1] - x[3 * j + 1], by, iby);
        const double dz = min_image_d(x[3 * i + 2] - x[3 * j + 2], bz, ibz);
        const double d2 = pair_d2(dx, dy, dz);
        if (!(d2 < cutoff * cutoff)) continue;
        auto f32 = [](double v) { return (double)(float)v; };
        PairTerms t = nb_pair_terms<true, false>(d2, f32(params[3 * i]), f32(params[3 * j]), f32(params[3 * i + 1]),
                                                 f32(params[3 * j + 1]), f32(params[3 * i + 2]), f32(params[3 * j + 2]),
                                                 (double)(float)beta, 1.0, 1.0);
        const uint64_t f[3] = {f64_to_fixed(t.prefactor * dx, FIXED_SCALE), f64_to_fixed(t.prefactor * dy, FIXED_SCALE),
                               f64_to_fixed(t.prefactor * dz, FIXED_SCALE)};
        for (int d = 0; d < 3; d++) {
            du_dx[3 * i + d] += f[d];
            du_dx[3 * j + d] -= f[d];
        }
        u[0] += f64_to_fixed(t.u, FIXED_SCALE);
    }
}

// The same model with the FP32 pair terms (nb_pair_terms_f32): FP64
// displacement, minimum image and d^2 (4D with w), FP32 terms, FP64
// projection of the FP32 prefactor onto the FP64 displacement.
// Also writes du/dlambda.
// Pairs m < M_all are added in full (the tile kernel); pairs m >= M_all are
// exclusions, subtracted with their scales (the negated pair-list kernel).
extern "C" void hc_nonbonded_du_dx_f32(int N, const double *x, const double *w, const double *dwdl,
                                       const double *params, const double *box, double beta, double cutoff, int M,
                                       const int64_t *pairs, const double *scales, int M_all, int f32_terms,
                                       uint64_t *du_dx, uint64_t *u, uint64_t *du_dl) {
    const double bx = box[0], by = box[4], bz = box[8];
    const double ibx = 1.0 / bx, iby = 1.0 / by, ibz = 1.0 / bz;
    for (int m = 0; m < M; m++) {
        const bool neg = m >= M_all;
        const double sq = neg ? scales[2 * (m - M_all)] : 1.0, slj = neg ? scales[2 * (m - M_all) + 1] : 1.0;
        const int i = (int)pairs[2 * m], j = (int)pairs[2 * m + 1];
        const double dx = min_image_d(x[3 * i] - x[3 * j], bx, ibx);
        const double dy = min_image_d(x[3 * i + 1] - x[3 * j + 1], by, iby);
        const double dz = min_image_d(x[3 * i + 2] - x[3 * j + 2], bz, ibz);
        const double dw = w[i] - w[j];
        const double d2 = __builtin_fma(dw, dw, pair_d2(dx, dy, dz));
        if (!(d2 < cutoff * cutoff)) continue;
        const float qi = (float)params[3 * i], qj = (float)params[3 * j];
        const float si = (float)params[3 * i + 1], sj = (float)params[3 * j + 1];
        const float ei = (float)params[3 * i + 2], ej = (float)params[3 * j + 2];
        PairTerms t = f32_terms ? nb_pair_terms_f32<true, false>(d2, qi, qj, si, sj, ei, ej, (float)beta, (float)sq,
                                                                 (float)slj)
                                : nb_pair_terms<true, false>(d2, qi, qj, si, sj, ei, ej, (double)(float)beta, sq, slj);
        const uint64_t f[3] = {f64_to_fixed(t.prefactor * dx, FIXED_SCALE), f64_to_fixed(t.prefactor * dy, FIXED_SCALE),
                               f64_to_fixed(t.prefactor * dz, FIXED_SCALE)};
        const uint64_t fu = f64_to_fixed(t.u, FIXED_SCALE);
        const uint64_t fdl =
            f64_to_fixed(dw * t.prefactor * ((double)(float)dwdl[i] - (double)(float)dwdl[j]), FIXED_SCALE);
        for (int d = 0; d < 3; d++) {
            du_dx[3 * i + d] += neg ? 0 - f[d] : f[d];
            du_dx[3 * j + d] -= neg ? 0 - f[d] : f[d];
        }
        u[0] += neg ? 0 - fu : fu;
        du_dl[0] += neg ? 0 - fdl : fdl;
    }
}

extern "C" float hc_half_round(float x) { return half_round(x); }

// FIRE on F = -k (x - c) with the step functions of fire.h and sequential
// sums; n_sched - 1 steps. Writes the final positions and, per step, dt,
// alpha and n_pos.
extern "C" void hc_fire_harmonic(int n3, const double *x0, const double *c, double k, int n_sched, double *x_out,
                                 float *dt_out, float *alpha_out, int *npos_out) {
    std::vector<double> x(x0, x0 + n3), v(n3, 0.0), f(n3);
    for (int i = 0; i < n3; i++) {
        f[i] = -k * (x[i] - c[i]);
    }
    float dt = FIRE_DT_START, alpha = FIRE_ALPHA_START;
    int npos = 0;
    for (int step = 1; step < n_sched; step++) {
        for (int i = 0; i < n3; i++) {
            x[i] = fire_shift(x[i], v[i], f[i], dt);
        }
        double sf = 0.0, sv = 0.0, sp = 0.0;
        for (int i = 0; i < n3; i++) {
            const double fn = -k * (x[i] - c[i]);
            v[i] = fire_velocity(v[i], f[i], fn, dt);
            f[i] = fn;
            sf += fn * fn;
            sv += v[i] * v[i];
            sp += fn * v[i];
        }
        const double f_norm = std::sqrt(sf + static_cast<double>(FIRE_F_EPS));
        const double v_norm = std::sqrt(sv);
        for (int i = 0; i < n3; i++) {
            v[i] = fire_mix(v[i], f[i], v_norm, f_norm, alpha, sp);
        }
        fire_adapt(sp, dt, alpha, npos);
        dt_out[step - 1] = dt;
        alpha_out[step - 1] = alpha;
        npos_out[step - 1] = npos;
    }
    for (int i = 0; i < n3; i++) {
        x_out[i] = x[i];
    }
}
