// Per-term bonded arithmetic (FP32 after an f64 coordinate difference), shared
// by the HIP kernels and a host build used to check derivatives on the CPU.
//
// Energies follow the reference's FP64 JAX definitions
// (timemachine/potentials/bonded.py:255-332 bonds and torsions, :150-229 cosine
// angles) and its FP32 kernels (kernels/k_harmonic_bond.cuh:4-71,
// k_harmonic_angle.cuh:4-92, k_periodic_torsion.cuh:21-146):
//   bond    u = pf kb/2 (|r_ij| - b0)^2               (b0 == 0: u = pf kb/2 |r_ij|^2)
//   angle   u = pf ka/2 (cos theta_ijk - cos a0)^2
//   torsion u = pf k (1 + cos(n phi - phase)),  phi = atan2((n1 x n2).r_kj/|r_kj|, n1.n2)
// with pf = lambda_offset + lambda_mult * lambda and du/dlambda = lambda_mult * (u / pf).
// Cross products are plain products and differences (no fused multiply-add),
// so cross(a, b) == -cross(b, a) bitwise and reversed index tuples give
// bitwise identical forces.
#pragma once

#include <cmath>

#include "hd.h"

namespace tmamd {

struct BondOut {
    float g[2][3]; // du/dx of atoms i, j
    float dkb, db0; // du/dp
    float u;
    float dl;
};

struct AngleOut {
    float g[3][3];
    float dka, da0;
    float u;
    float dl;
};

struct TorsionOut {
    float g[4][3];
    float dk, dphase, dperiod;
    float u;
    float dl;
};

TM_HD inline float lam_prefactor(int mult, int offset, double lambda) {
    return static_cast<float>(offset + mult * lambda);
}

TM_HD inline BondOut bond_term(const double *xi, const double *xj, float kb, float b0, int mult, float pf) {
    BondOut o;
    float dx[3];
    float d2 = 0.0f;
    for (int d = 0; d < 3; d++) {
        dx[d] = static_cast<float>(xi[d] - xj[d]);
        d2 = d2 + dx[d] * dx[d];
    }
    const float dij = sqrtf(d2);
    const float db = dij - b0;
    const float half_db2 = 0.5f * db * db;
    for (int d = 0; d < 3; d++) {
        const float gd = (b0 != 0.0f) ? kb * db * (dx[d] / dij) : kb * dx[d];
        o.g[0][d] = gd * pf;
        o.g[1][d] = -gd * pf;
    }
    o.dkb = half_db2 * pf;
    o.db0 = -kb * db * pf;
    o.u = kb * half_db2 * pf;
    o.dl = static_cast<float>(mult) * (kb * half_db2);
    return o;
}

TM_HD inline AngleOut angle_term(const double *xi, const double *xj, const double *xk, float ka, float a0, int mult,
                                 float pf) {
    AngleOut o;
    float a[3], b[3];
    float aa = 0.0f, bb = 0.0f, ab = 0.0f;
    for (int d = 0; d < 3; d++) {
        a[d] = static_cast<float>(xi[d] - xj[d]);
        b[d] = static_cast<float>(xk[d] - xj[d]);
        aa = aa + a[d] * a[d];
        bb = bb + b[d] * b[d];
        ab = ab + a[d] * b[d];
    }
    const float na = sqrtf(aa);
    const float nb = sqrtf(bb);
    const float nab = na * nb;
    const float cos_t = ab / nab;
    const float delta = cos_t - cosf(a0);
    const float dudc = pf * ka * delta;
    for (int d = 0; d < 3; d++) {
        const float dca = b[d] / nab - cos_t * (a[d] / aa);
        const float dcb = a[d] / nab - cos_t * (b[d] / bb);
        o.g[0][d] = dudc * dca;
        o.g[2][d] = dudc * dcb;
        o.g[1][d] = -(o.g[0][d] + o.g[2][d]);
    }
    const float half_d2 = 0.5f * delta * delta;
    o.dka = half_d2 * pf;
    o.da0 = pf * ka * delta * sinf(a0);
    o.u = ka * half_d2 * pf;
    o.dl = static_cast<float>(mult) * (ka * half_d2);
    return o;
}

TM_HD inline void cross3(const float *a, const float *b, float *c) {
    c[0] = a[1] * b[2] - a[2] * b[1];
    c[1] = a[2] * b[0] - a[0] * b[2];
    c[2] = a[0] * b[1] - a[1] * b[0];
}

TM_HD inline float dot3(const float *a, const float *b) { return a[0] * b[0] + a[1] * b[1] + a[2] * b[2]; }

TM_HD inline TorsionOut torsion_term(const double *xi, const double *xj, const double *xk, const double *xl, float kt,
                                     float phase, float period, int mult, float pf) {
    TorsionOut o;
    float rij[3], rkj[3], rkl[3];
    for (int d = 0; d < 3; d++) {
        rij[d] = static_cast<float>(xj[d] - xi[d]);
        rkj[d] = static_cast<float>(xj[d] - xk[d]);
        rkl[d] = static_cast<float>(xl[d] - xk[d]);
    }
    float n1[3], n2[3], n3[3];
    cross3(rij, rkj, n1);
    cross3(rkj, rkl, n2);
    cross3(n1, n2, n3);
    const float n1n1 = dot3(n1, n1);
    const float n2n2 = dot3(n2, n2);
    const float rkj2 = dot3(rkj, rkj);
    const float rkj_n = sqrtf(rkj2);
    const float y = dot3(n3, rkj) / rkj_n;
    const float x = dot3(n1, n2);
    const float phi = atan2f(y, x);
    // dphi/dx for the four atoms (Blondel & Karplus 1996, J. Comput. Chem. 17, 1132)
    const float fij = dot3(rij, rkj) / rkj2;
    const float fkl = dot3(rkl, rkj) / rkj2;
    float dphi[4][3];
    for (int d = 0; d < 3; d++) {
        dphi[0][d] = -(rkj_n / n1n1) * n1[d];
        dphi[3][d] = (rkj_n / n2n2) * n2[d];
        dphi[1][d] = (fij - 1.0f) * dphi[0][d] - fkl * dphi[3][d];
        dphi[2][d] = (fkl - 1.0f) * dphi[3][d] - fij * dphi[0][d];
    }
    const float arg = period * phi - phase;
    const float s = sinf(arg);
    const float c = cosf(arg);
    const float dudphi = -pf * kt * period * s;
    for (int a = 0; a < 4; a++) {
        for (int d = 0; d < 3; d++) {
            o.g[a][d] = dudphi * dphi[a][d];
        }
    }
    o.dk = (1.0f + c) * pf;
    o.dphase = kt * s * pf;
    o.dperiod = -kt * s * phi * pf;
    o.u = kt * (1.0f + c) * pf;
    o.dl = static_cast<float>(mult) * (kt * (1.0f + c));
    return o;
}

} // namespace tmamd
