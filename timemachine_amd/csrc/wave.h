// Wave64 helpers for gfx950.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

namespace tmamd {

__device__ __forceinline__ int lane_id() { return __builtin_amdgcn_mbcnt_hi(~0u, __builtin_amdgcn_mbcnt_lo(~0u, 0u)); }

// Number of set bits of mask below this lane.
__device__ __forceinline__ int lanes_below(uint64_t mask) {
    return __builtin_amdgcn_mbcnt_hi(static_cast<uint32_t>(mask >> 32),
                                      __builtin_amdgcn_mbcnt_lo(static_cast<uint32_t>(mask), 0u));
}

__device__ __forceinline__ uint64_t ballot(bool p) { return __ballot(p); }

__device__ __forceinline__ int uniform(int v) { return __builtin_amdgcn_readfirstlane(v); }
__device__ __forceinline__ float uniform_f(float v) {
    return __int_as_float(__builtin_amdgcn_readfirstlane(__float_as_int(v)));
}

// Orders LDS traffic between lanes of one wave (no workgroup barrier).
__device__ __forceinline__ void wave_lds_sync() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// Rotate a value by one lane across the whole wave: lane l receives the value
// held by lane (l + 1) mod 64. GFX9-family DPP "wave_rol:1" (ctrl 0x134); all
// lanes must be active.
__device__ __forceinline__ uint32_t rot1_u32(uint32_t v) {
    // every lane has a source lane, so no "old" value is needed (bound_ctrl set)
    return static_cast<uint32_t>(__builtin_amdgcn_mov_dpp(static_cast<int>(v), 0x134, 0xF, 0xF, true));
}
__device__ __forceinline__ float rot1(float v) { return __uint_as_float(rot1_u32(__float_as_uint(v))); }
__device__ __forceinline__ int rot1(int v) { return static_cast<int>(rot1_u32(static_cast<uint32_t>(v))); }
__device__ __forceinline__ uint32_t rot1(uint32_t v) { return rot1_u32(v); }
__device__ __forceinline__ uint64_t rot1(uint64_t v) {
    uint32_t lo = rot1_u32(static_cast<uint32_t>(v));
    uint32_t hi = rot1_u32(static_cast<uint32_t>(v >> 32));
    return (static_cast<uint64_t>(hi) << 32) | lo;
}

__device__ __forceinline__ double rot1(double v) {
    return __longlong_as_double(static_cast<long long>(rot1(static_cast<uint64_t>(__double_as_longlong(v)))));
}

__device__ __forceinline__ uint64_t shfl_xor_u64(uint64_t v, int m) {
    uint32_t lo = static_cast<uint32_t>(__shfl_xor(static_cast<int>(static_cast<uint32_t>(v)), m, 64));
    uint32_t hi = static_cast<uint32_t>(__shfl_xor(static_cast<int>(static_cast<uint32_t>(v >> 32)), m, 64));
    return (static_cast<uint64_t>(hi) << 32) | lo;
}

// Sum of a uint64 over the wave (modulo 2^64, exact); result in every lane.
__device__ __forceinline__ uint64_t wave_sum_u64(uint64_t v) {
    for (int m = 32; m >= 1; m >>= 1) {
        v += shfl_xor_u64(v, m);
    }
    return v;
}

// base[key] += v summed over the lanes of each distinct key, one global atomic
// per key present in the wave (terms of one system are contiguous, so a wave
// holds one or two keys). Lanes with key < 0 or v == 0 contribute nothing.
// Must be reached by every lane of the wave.
__device__ __forceinline__ void wave_add_by_key(uint64_t *base, int key, uint64_t v) {
    const bool has = key >= 0 && v != 0;
    uint64_t pending = ballot(has);
    while (pending != 0) {
        const int leader = __ffsll(static_cast<long long>(pending)) - 1;
        const int k = __builtin_amdgcn_readlane(key, leader);
        const bool mine = has && key == k;
        const uint64_t tot = wave_sum_u64(mine ? v : 0);
        if (lane_id() == leader) {
            atomicAdd(reinterpret_cast<unsigned long long *>(base + k), static_cast<unsigned long long>(tot));
        }
        pending &= ~ballot(mine);
    }
}

__device__ __forceinline__ float wave_min(float v) {
    for (int m = 32; m >= 1; m >>= 1) {
        v = fminf(v, __shfl_xor(v, m, 64));
    }
    return v;
}

__device__ __forceinline__ float wave_max(float v) {
    for (int m = 32; m >= 1; m >>= 1) {
        v = fmaxf(v, __shfl_xor(v, m, 64));
    }
    return v;
}

} // namespace tmamd
