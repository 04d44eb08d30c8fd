// Host build of the arithmetic headers of timemachine_amd/csrc, used by the
// CPU tests to check the formulas (and Philox / fixed point bit patterns)
// before anything runs on a GPU. Built by tests/test_host_math.py with g++.
#include <cstdint>

#include "bonded_terms.h"
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

uint64_t hc_f32_to_fixed(float v, float scale) { return f32_to_fixed(v, scale); }

// out: prefactor, u, erfc_inv_d, sig_grad, eps_grad, lj
void hc_pair_terms(float d2, float qi, float qj, float si, float sj, float ei, float ej, float beta, float sq,
                   float slj, float *out) {
    PairTerms t = nb_pair_terms<true, true>(d2, qi, qj, si, sj, ei, ej, beta, sq, slj);
    out[0] = t.prefactor;
    out[1] = t.u;
    out[2] = t.erfc_inv_d;
    out[3] = t.sig_grad;
    out[4] = t.eps_grad;
    out[5] = t.lj ? 1.0f : 0.0f;
}

float hc_min_image(float d, float b) { return min_image_f(d, b, 1.0f / b); }

// out: g[2][3], dkb, db0, u, dl
void hc_bond(const double *xi, const double *xj, float kb, float b0, int mult, int offset, double lam, float *out) {
    BondOut o = bond_term(xi, xj, kb, b0, mult, lam_prefactor(mult, offset, lam));
    for (int a = 0; a < 2; a++)
        for (int d = 0; d < 3; d++) out[a * 3 + d] = o.g[a][d];
    out[6] = o.dkb;
    out[7] = o.db0;
    out[8] = o.u;
    out[9] = o.dl;
}

// out: g[3][3], dka, da0, u, dl
void hc_angle(const double *xi, const double *xj, const double *xk, float ka, float a0, int mult, int offset, double lam,
              float *out) {
    AngleOut o = angle_term(xi, xj, xk, ka, a0, mult, lam_prefactor(mult, offset, lam));
    for (int a = 0; a < 3; a++)
        for (int d = 0; d < 3; d++) out[a * 3 + d] = o.g[a][d];
    out[9] = o.dka;
    out[10] = o.da0;
    out[11] = o.u;
    out[12] = o.dl;
}

// out: g[4][3], dk, dphase, dperiod, u, dl
void hc_torsion(const double *xi, const double *xj, const double *xk, const double *xl, float k, float phase,
                float period, int mult, int offset, double lam, float *out) {
    TorsionOut o = torsion_term(xi, xj, xk, xl, k, phase, period, mult, lam_prefactor(mult, offset, lam));
    for (int a = 0; a < 4; a++)
        for (int d = 0; d < 3; d++) out[a * 3 + d] = o.g[a][d];
    out[12] = o.dk;
    out[13] = o.dphase;
    out[14] = o.dperiod;
    out[15] = o.u;
    out[16] = o.dl;
}
}
