// Host stand-in for <hip/hip_runtime.h>, used only by tests/cxx/ukernel_host.cpp: lets g++ compile the
// library's per-pair micro-kernel headers (stkfmm_amd/csrc/ukernels*.hpp) for the CPU so that their
// algebra is checked against the oracle without a GPU.  The two rsqrt builtins become exact
// reciprocal square roots (the device refines its hardware seed to the same value within ~1e-15).
#pragma once
#include <cmath>
#define __device__
#define __host__
#define __forceinline__ inline
static inline double __builtin_amdgcn_rsq(double x) { return 1.0 / std::sqrt(x); }
static inline float __builtin_amdgcn_rsqf(float x) { return 1.0f / std::sqrt(x); }
