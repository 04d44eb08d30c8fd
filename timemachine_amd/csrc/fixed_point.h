// 64-bit fixed-point encoding of every accumulated quantity.
//
// Follows the reference's convention (timemachine/cpp/src/fixed_point.hpp:3-17,
// kernels/k_fixed_point.cuh:51-69): a value v is scaled by 2^k in the working
// precision, rounded to the nearest integer (ties to even, as llrint does) and
// summed as two's-complement uint64 modulo 2^64. Summation order therefore
// cannot change a result: totals are bitwise reproducible run to run and
// independent of how pairs are grouped into tiles, waves or batches.
//
// gfx950 has no f32 -> i64 convert instruction, so the rounding is done in f64
// (every f32 is exactly representable in f64, and the f32 product v*2^k is exact
// because 2^k is a power of two):
//   |r| < 2^51 : add 1.5*2^52; the ulp there is 1, so the f64 add rounds v to
//               an integer with ties to even, and the integer sits in the low
//               mantissa bits; subtracting the bit pattern of 1.5*2^52 gives
//               the two's-complement result.
//   otherwise  : rint, then split into a high word (floor(r/2^32)) and an exact
//               low word r - hi*2^32 in [0, 2^32).
// Values outside the int64 range and NaN map to 0x8000000000000000, which is
// what x86 llrint returns, so the CPU model in oracle/ matches bit for bit.
#pragma once

#include <cstdint>

#include "hd.h"

namespace tmamd {

constexpr float FIXED_SCALE_F = 68719476736.0f;       // 2^36: du_dx, u, du_dl, charge and bonded du_dp
constexpr float FIXED_SCALE_DSIG_F = 137438953472.0f; // 2^37: d u / d sigma
constexpr float FIXED_SCALE_DEPS_F = 274877906944.0f; // 2^38: d u / d epsilon
constexpr double FIXED_SCALE = 68719476736.0;
constexpr double FIXED_SCALE_DSIG = 137438953472.0;
constexpr double FIXED_SCALE_DEPS = 274877906944.0;

constexpr uint64_t FIXED_MAGIC_BITS = 0x4338000000000000ull; // bit pattern of 1.5 * 2^52
constexpr uint64_t FIXED_INVALID = 0x8000000000000000ull;

// Round an (already scaled) double to the nearest int64, ties to even.
TM_HD inline uint64_t round_to_fixed(double d) {
    if (d > -0x1p51 && d < 0x1p51) {
        double t = d + 0x1.8p52;
#ifdef __HIP_DEVICE_COMPILE__
        uint64_t bits = static_cast<uint64_t>(__double_as_longlong(t));
#else
        uint64_t bits;
        __builtin_memcpy(&bits, &t, sizeof(bits));
#endif
        return bits - FIXED_MAGIC_BITS;
    }
    if (!(d > -0x1p63 && d < 0x1p63)) {
        return FIXED_INVALID; // out of range or NaN
    }
    double r = __builtin_rint(d);
    double hi = __builtin_floor(r * 0x1p-32);
    double lo = __builtin_fma(hi, -0x1p32, r);
    uint32_t hi32 = static_cast<uint32_t>(static_cast<int32_t>(hi));
    uint32_t lo32 = static_cast<uint32_t>(lo);
    return (static_cast<uint64_t>(hi32) << 32) | lo32;
}

// Three values at once with the same result as three round_to_fixed calls:
// one range test for the common case that all three take the add-magic path.
TM_HD inline void round_to_fixed3(double a, double b, double c, uint64_t &fa, uint64_t &fb, uint64_t &fc) {
    // comparisons are false for NaN, which the general path handles
    const bool fast = (__builtin_fabs(a) < 0x1p51) & (__builtin_fabs(b) < 0x1p51) & (__builtin_fabs(c) < 0x1p51);
    if (fast) {
        const double ta = a + 0x1.8p52, tb = b + 0x1.8p52, tc = c + 0x1.8p52;
        uint64_t ba, bb, bc;
        __builtin_memcpy(&ba, &ta, sizeof(ba));
        __builtin_memcpy(&bb, &tb, sizeof(bb));
        __builtin_memcpy(&bc, &tc, sizeof(bc));
        fa = ba - FIXED_MAGIC_BITS;
        fb = bb - FIXED_MAGIC_BITS;
        fc = bc - FIXED_MAGIC_BITS;
        return;
    }
    fa = round_to_fixed(a);
    fb = round_to_fixed(b);
    fc = round_to_fixed(c);
}

// f32 value -> fixed point with scale 2^k (scale must be a power of two).
TM_HD inline uint64_t f32_to_fixed(float v, float scale) {
    return round_to_fixed(static_cast<double>(v * scale));
}

// f64 value -> fixed point (used by f64 accumulations such as centroids).
TM_HD inline uint64_t f64_to_fixed(double v, double scale) { return round_to_fixed(v * scale); }

TM_HD inline double fixed_to_double(uint64_t v, double scale) {
    return static_cast<double>(static_cast<int64_t>(v)) / scale;
}

} // namespace tmamd
