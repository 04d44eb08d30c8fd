// Counter-based random numbers: Philox4x32-10 (Salmon, Moraes, Dror, Shaw,
// "Parallel random numbers: as easy as 1, 2, 3", SC'11) and Box-Muller.
//
// The reference draws its Langevin noise and barostat uniforms from curand's
// stateful generator (timemachine/cpp/src/integrator.cu:25-27, barostat.cu:39-41).
// Here every draw is a pure function of (seed, stream, step, atom), so a
// system's trajectory is bitwise identical whether it runs alone or inside a
// batch, and (x, v, box, step) is a complete checkpoint.
#pragma once

#include <cmath>
#include <cstdint>

#include "hd.h"

namespace tmamd {

struct Philox4 {
    uint32_t v[4];
};

TM_HD inline void mulhilo32(uint32_t a, uint32_t b, uint32_t &hi, uint32_t &lo) {
    const uint64_t p = static_cast<uint64_t>(a) * static_cast<uint64_t>(b);
    hi = static_cast<uint32_t>(p >> 32);
    lo = static_cast<uint32_t>(p);
}

TM_HD inline Philox4 philox4x32_10(Philox4 ctr, uint32_t k0, uint32_t k1) {
    const uint32_t M0 = 0xD2511F53u, M1 = 0xCD9E8D57u;
    const uint32_t W0 = 0x9E3779B9u, W1 = 0xBB67AE85u;
    for (int r = 0; r < 10; r++) {
        uint32_t hi0, lo0, hi1, lo1;
        mulhilo32(M0, ctr.v[0], hi0, lo0);
        mulhilo32(M1, ctr.v[2], hi1, lo1);
        Philox4 n;
        n.v[0] = hi1 ^ ctr.v[1] ^ k0;
        n.v[1] = lo1;
        n.v[2] = hi0 ^ ctr.v[3] ^ k1;
        n.v[3] = lo0;
        ctr = n;
        k0 += W0;
        k1 += W1;
    }
    return ctr;
}

// Uniform double in (0, 1] from two 32-bit words (53 random bits).
TM_HD inline double u01_open0(uint32_t a, uint32_t b) {
    const uint64_t k = (static_cast<uint64_t>(a >> 5) << 26) | static_cast<uint64_t>(b >> 6);
    return (static_cast<double>(k) + 1.0) * 0x1p-53;
}

// Two standard normals from one Philox block (Box-Muller).
TM_HD inline void normal2(const Philox4 &r, double &z0, double &z1) {
    const double u1 = u01_open0(r.v[0], r.v[1]);
    const double u2 = u01_open0(r.v[2], r.v[3]);
    const double rad = sqrt(-2.0 * log(u1));
    const double ang = 6.283185307179586476925286766559 * u2;
    z0 = rad * cos(ang);
    z1 = rad * sin(ang);
}

// Stream identifiers (fourth counter word).
constexpr uint32_t RNG_STREAM_LANGEVIN = 0x4C414E47u; // Langevin noise, two blocks per atom
constexpr uint32_t RNG_STREAM_BAROSTAT = 0x4241524Fu; // barostat move

} // namespace tmamd
