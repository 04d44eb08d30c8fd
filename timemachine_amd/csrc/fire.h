// FIRE minimisation arithmetic (Bitzek, Koskinen, Gahler, Moseler, Gumbsch,
// "Structural relaxation made simple", PRL 97, 170201 (2006)), with the
// numerics of the reference's minimiser (_vendored/fire.py:60-140, called by
// md/minimizer.py:35-73):
//
// * the step size dt and the mixing factor alpha are binary16 numbers: the
//   constants go through np.min_scalar_type (all land in float16) and every
//   update of dt / alpha is a float16 product; dt**2 is a float16 square, so it
//   is 0 while dt < 1.7e-4 and the force term of the position update vanishes;
// * dt * 0.5 of the velocity update is a float32 product; positions,
//   velocities and forces are float64.
//
// One step of a system (R position, V velocity, F_old force of R):
//   R <- R + (dt V + dt2 F_old)                   dt2 = f16(dt * dt)
//   F <- force(R)
//   V <- V + (dt * 0.5f) (F_old + F)
//   |F| = sqrt(sum F^2 + 1e-6f), |V| = sqrt(sum V^2), P = F . V
//   V <- V + alpha (F |V| / |F| - V)
//   n_pos <- P >= 0 ? n_pos + 1 : 0
//   P > 0 and n_pos > n_min: dt <- min(f16(dt f_inc), dt_max), alpha <- f16(alpha f_alpha)
//   P < 0: dt <- f16(dt f_dec), alpha <- alpha_start, V <- 0
// Plain C++ arithmetic (every translation unit is built with
// -ffp-contract=off, so nothing is fused).
#pragma once

#include <cmath>
#include <cstdint>
#include <cstring>

#include "hd.h"

namespace tmamd {

// Round a float to the nearest binary16 value (ties to even, subnormals kept,
// overflow to infinity) and return it as a float.
TM_HD inline float half_round(float x) {
    const float a = std::fabs(x);
    if (!(a == a)) {
        return x; // NaN
    }
    if (a >= 65520.0f) {
        return x > 0 ? INFINITY : -INFINITY;
    }
    if (a < 6.103515625e-05f) {
        // subnormal range: multiples of 2^-24 (scaling by powers of two is exact)
        const float q = std::rint(a * 16777216.0f);
        const float r = q * 5.9604644775390625e-08f;
        return x < 0 ? -r : r;
    }
    uint32_t u;
    __builtin_memcpy(&u, &x, sizeof(u));
    // keep 10 of the 23 mantissa bits, ties to even
    u = u + 0xFFFu + ((u >> 13) & 1u);
    u &= ~0x1FFFu;
    float r;
    __builtin_memcpy(&r, &u, sizeof(r));
    return r;
}

// Constants of _vendored/fire.py:66-74 after np.min_scalar_type (binary16 values).
constexpr float FIRE_DT_START = 1.0013580322265625e-05f; // 1e-5
constexpr float FIRE_DT_MAX = 0.0010004043579101562f;    // 1e-3
constexpr int FIRE_N_MIN = 5;
constexpr float FIRE_F_INC = 1.099609375f;          // 1.1
constexpr float FIRE_F_DEC = 0.5f;                  // 0.5
constexpr float FIRE_ALPHA_START = 0.0999755859375f; // 0.1
constexpr float FIRE_F_ALPHA = 0.990234375f;        // 0.99
constexpr float FIRE_F_EPS = 1e-6f;                 // added to sum F^2

// float16 square of the step size (0 below dt = 1.7e-4).
TM_HD inline double fire_dt2(float dt) { return static_cast<double>(half_round(dt * dt)); }

// New position component.
TM_HD inline double fire_shift(double r, double v, double f_old, float dt) {
    const double dr = static_cast<double>(dt) * v + fire_dt2(dt) * f_old;
    return r + dr;
}

// Velocity after the force at the new position is known.
TM_HD inline double fire_velocity(double v, double f_old, double f, float dt) {
    return v + static_cast<double>(dt * 0.5f) * (f_old + f);
}

// Velocity mixing; f_norm includes the 1e-6 term.
TM_HD inline double fire_mix(double v, double f, double v_norm, double f_norm, float alpha, double P) {
    const double m = v + static_cast<double>(alpha) * (f * v_norm / f_norm - v);
    return P < 0 ? 0.0 : m;
}

// dt / alpha / n_pos update from P = F . V.
TM_HD inline void fire_adapt(double P, float &dt, float &alpha, int &n_pos) {
    n_pos = P >= 0 ? n_pos + 1 : 0;
    if (P > 0 && n_pos > FIRE_N_MIN) {
        const float up = half_round(dt * FIRE_F_INC);
        dt = up < FIRE_DT_MAX ? up : FIRE_DT_MAX;
        alpha = half_round(alpha * FIRE_F_ALPHA);
    }
    if (P < 0) {
        dt = half_round(dt * FIRE_F_DEC);
        alpha = FIRE_ALPHA_START;
    }
}

} // namespace tmamd
