// Per-term bonded arithmetic, shared by the HIP kernels and a host build used
// to check derivatives on the CPU. Templated on the working precision; the
// kernels instantiate it in FP64: bonded terms are a few thousand per system,
// FP64 runs at the FP32 rate on CDNA4, and stiff terms (O-H bonds at
// 4.6e5 kJ/mol/nm^2) lose ~1e-4 of the per-atom force to FP32 cancellation.
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

#include "dual.h"
#include "hd.h"

namespace tmamd {

// x_a - x_b as T. FP64 positions subtract in FP64 and convert (the kernels'
// expression); positions that carry a tangent (Dual) subtract as duals, which
// gives the terms' directional derivatives (the Hessian-vector products).
template <typename T> TM_HD inline T pos_diff(double a, double b) { return static_cast<T>(a - b); }
template <typename T> TM_HD inline T pos_diff(const Dual &a, const Dual &b) { return a - b; }

template <typename T> struct BondOut {
    T g[2][3]; // du/dx of atoms i, j
    T dkb, db0; // du/dp
    T u;
    T dl;
};

template <typename T> struct AngleOut {
    T g[3][3];
    T dka, da0;
    T u;
    T dl;
};

template <typename T> struct TorsionOut {
    T g[4][3];
    T dk, dphase, dperiod;
    T u;
    T dl;
};

template <typename T> TM_HD inline T lam_prefactor(int mult, int offset, double lambda) {
    return static_cast<T>(offset + mult * lambda);
}

template <typename T, typename P = double>
TM_HD inline BondOut<T> bond_term(const P *xi, const P *xj, T kb, T b0, int mult, T pf) {
    using std::sqrt;
    BondOut<T> o;
    T dx[3];
    T d2 = 0;
    for (int d = 0; d < 3; d++) {
        dx[d] = pos_diff<T>(xi[d], xj[d]);
        d2 = d2 + dx[d] * dx[d];
    }
    const T dij = sqrt(d2);
    const T db = dij - b0;
    const T half_db2 = T(0.5) * db * db;
    for (int d = 0; d < 3; d++) {
        const T gd = (b0 != T(0)) ? kb * db * (dx[d] / dij) : kb * dx[d];
        o.g[0][d] = gd * pf;
        o.g[1][d] = -gd * pf;
    }
    o.dkb = half_db2 * pf;
    o.db0 = -kb * db * pf;
    o.u = kb * half_db2 * pf;
    o.dl = static_cast<T>(mult) * (kb * half_db2);
    return o;
}

template <typename T, typename P = double>
TM_HD inline AngleOut<T> angle_term(const P *xi, const P *xj, const P *xk, T ka, T a0, int mult, T pf) {
    using std::cos;
    using std::sin;
    using std::sqrt;
    AngleOut<T> o;
    T a[3], b[3];
    T aa = 0, bb = 0, ab = 0;
    for (int d = 0; d < 3; d++) {
        a[d] = pos_diff<T>(xi[d], xj[d]);
        b[d] = pos_diff<T>(xk[d], xj[d]);
        aa = aa + a[d] * a[d];
        bb = bb + b[d] * b[d];
        ab = ab + a[d] * b[d];
    }
    const T na = sqrt(aa);
    const T nb = sqrt(bb);
    const T nab = na * nb;
    const T cos_t = ab / nab;
    const T delta = cos_t - cos(a0);
    const T dudc = pf * ka * delta;
    for (int d = 0; d < 3; d++) {
        const T dca = b[d] / nab - cos_t * (a[d] / aa);
        const T dcb = a[d] / nab - cos_t * (b[d] / bb);
        o.g[0][d] = dudc * dca;
        o.g[2][d] = dudc * dcb;
        o.g[1][d] = -(o.g[0][d] + o.g[2][d]);
    }
    const T half_d2 = T(0.5) * delta * delta;
    o.dka = half_d2 * pf;
    o.da0 = pf * ka * delta * sin(a0);
    o.u = ka * half_d2 * pf;
    o.dl = static_cast<T>(mult) * (ka * half_d2);
    return o;
}

template <typename T> TM_HD inline void cross3(const T *a, const T *b, T *c) {
    c[0] = a[1] * b[2] - a[2] * b[1];
    c[1] = a[2] * b[0] - a[0] * b[2];
    c[2] = a[0] * b[1] - a[1] * b[0];
}

template <typename T> TM_HD inline T dot3(const T *a, const T *b) { return a[0] * b[0] + a[1] * b[1] + a[2] * b[2]; }

template <typename T, typename P = double>
TM_HD inline TorsionOut<T> torsion_term(const P *xi, const P *xj, const P *xk, const P *xl, T kt, T phase, T period,
                                        int mult, T pf) {
    using std::atan2;
    using std::cos;
    using std::sin;
    using std::sqrt;
    TorsionOut<T> o;
    T rij[3], rkj[3], rkl[3];
    for (int d = 0; d < 3; d++) {
        rij[d] = pos_diff<T>(xj[d], xi[d]);
        rkj[d] = pos_diff<T>(xj[d], xk[d]);
        rkl[d] = pos_diff<T>(xl[d], xk[d]);
    }
    T n1[3], n2[3], n3[3];
    cross3(rij, rkj, n1);
    cross3(rkj, rkl, n2);
    cross3(n1, n2, n3);
    const T n1n1 = dot3(n1, n1);
    const T n2n2 = dot3(n2, n2);
    const T rkj2 = dot3(rkj, rkj);
    const T rkj_n = sqrt(rkj2);
    const T y = dot3(n3, rkj) / rkj_n;
    const T x = dot3(n1, n2);
    const T phi = atan2(y, x);
    // dphi/dx for the four atoms (Blondel & Karplus 1996, J. Comput. Chem. 17, 1132)
    const T fij = dot3(rij, rkj) / rkj2;
    const T fkl = dot3(rkl, rkj) / rkj2;
    T dphi[4][3];
    for (int d = 0; d < 3; d++) {
        dphi[0][d] = -(rkj_n / n1n1) * n1[d];
        dphi[3][d] = (rkj_n / n2n2) * n2[d];
        dphi[1][d] = (fij - T(1)) * dphi[0][d] - fkl * dphi[3][d];
        dphi[2][d] = (fkl - T(1)) * dphi[3][d] - fij * dphi[0][d];
    }
    const T arg = period * phi - phase;
    const T s = sin(arg);
    const T c = cos(arg);
    const T dudphi = -pf * kt * period * s;
    for (int a = 0; a < 4; a++) {
        for (int d = 0; d < 3; d++) {
            o.g[a][d] = dudphi * dphi[a][d];
        }
    }
    o.dk = (T(1) + c) * pf;
    o.dphase = kt * s * pf;
    o.dperiod = -kt * s * phi * pf;
    o.u = kt * (T(1) + c) * pf;
    o.dl = static_cast<T>(mult) * (kt * (T(1) + c));
    return o;
}

} // namespace tmamd
