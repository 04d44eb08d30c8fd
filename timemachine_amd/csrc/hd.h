// Host/device qualifiers so arithmetic headers also compile in plain C++
// (host builds used to check the formulas on the CPU). The FP32 hardware
// seeds (rsq, rcp, 2^x; about 1 ulp) map to their libm counterparts on the host.
#pragma once

#if defined(__HIPCC__)
#include <hip/hip_runtime.h>
#define TM_HD __host__ __device__
#else
#define TM_HD
#endif

#if defined(__HIP_DEVICE_COMPILE__)
#define TM_RSQF(x) __builtin_amdgcn_rsqf(x)
#define TM_RCPF(x) __builtin_amdgcn_rcpf(x)
#define TM_EXP2F(x) __builtin_amdgcn_exp2f(x)
#else
#include <cmath>
#define TM_RSQF(x) (1.0f / std::sqrt(x))
#define TM_RCPF(x) (1.0f / (x))
#define TM_EXP2F(x) std::exp2(static_cast<float>(x))
#endif
