// p2p_clear.hpp -- the coincidence-mask shortcut shared by the item kernels (LDS and scalar-broadcast).
//
// The reference zeroes pairs with r2 = 0.  Each work item computes the bounding box of its targets once
// (a wave reduction over the lanes' targets); a group of sources (an LDS tile, or a 64-source chunk of
// a range in the scalar-broadcast kernel) is checked one source per lane with one ballot: when every
// source lies outside the box by ClearGap on some axis, every pair of the group has r2 > 0 and the
// group runs the micro-kernel's unmasked form (eval<R, false>), which gives the same results without
// the compare and select of each pair.  A NaN coordinate makes r2 NaN, which the mask turns into a
// zero contribution, so it must never pass: a NaN in any target's coordinates makes the box the whole
// space, and a source with a NaN coordinate counts as near however far its other coordinates are
// (a separation on one axis says nothing about a NaN on another).
#pragma once

#include <hip/hip_runtime.h>

#include "p2p_internal.hpp"
#include "ukernels.hpp"

namespace stkp2p {

// one pair through the micro-kernel, without the coincidence mask when MASK is false and the
// micro-kernel offers that form
template <class K, bool MASK, class R>
__device__ __forceinline__ void eval_pair(R *u, R dx, R dy, R dz, const R *f) {
    if constexpr (HasUnmaskedEval<K>::value)
        K::template eval<R, MASK>(u, dx, dy, dz, f);
    else
        K::template eval<R>(u, dx, dy, dz, f);
}

// a wave-uniform double held in scalar registers (every lane has the same value)
__device__ __forceinline__ double uniform_f64(double v) {
    const unsigned long long b = __builtin_bit_cast(unsigned long long, v);
    const unsigned lo = __builtin_amdgcn_readfirstlane((unsigned)b);
    const unsigned hi = __builtin_amdgcn_readfirstlane((unsigned)(b >> 32));
    return __builtin_bit_cast(double, ((unsigned long long)hi << 32) | lo);
}

// Minimum separation, per axis, between a source and the bounding box of an item's targets for
// which every pair of the source is known to have r2 > 0 in the pair arithmetic: FP64, |dx| >=
// 1e-150 keeps dx^2 above the smallest normal double; mixed, |dx| >= 4e-18 (after the rounding of
// the box edge, at least 2e-18) keeps the FP32 dx^2 above the smallest normal float.
template <int PREC> struct ClearGap;
template <> struct ClearGap<0> { static constexpr double value = 1e-150; };
template <> struct ClearGap<1> { static constexpr double value = 4e-18; };

// min / max that return NaN when either operand is NaN
__device__ __forceinline__ double min_nan(double a, double b) { return (a < b || a != a) ? a : b; }
__device__ __forceinline__ double max_nan(double a, double b) { return (a > b || a != a) ? a : b; }

// bounding box of the item's targets, widened by ClearGap, in scalar registers
template <int PREC>
struct TargetBox {
    double lo[3], hi[3];
    template <int T>
    __device__ void build(const double (&tx)[T], const double (&ty)[T], const double (&tz)[T]) {
        lo[0] = hi[0] = tx[0];
        lo[1] = hi[1] = ty[0];
        lo[2] = hi[2] = tz[0];
#pragma unroll
        for (int t = 1; t < T; ++t) {
            lo[0] = min_nan(lo[0], tx[t]);
            hi[0] = max_nan(hi[0], tx[t]);
            lo[1] = min_nan(lo[1], ty[t]);
            hi[1] = max_nan(hi[1], ty[t]);
            lo[2] = min_nan(lo[2], tz[t]);
            hi[2] = max_nan(hi[2], tz[t]);
        }
        for (int off = 1; off < kWave; off <<= 1) {
#pragma unroll
            for (int d = 0; d < 3; ++d) {
                lo[d] = min_nan(lo[d], __shfl_xor(lo[d], off));
                hi[d] = max_nan(hi[d], __shfl_xor(hi[d], off));
            }
        }
        // NaN propagates through the reductions: a NaN target makes some bound NaN, and then every
        // source is near
        const bool nan_box = lo[0] != lo[0] || lo[1] != lo[1] || lo[2] != lo[2] || hi[0] != hi[0] ||
                             hi[1] != hi[1] || hi[2] != hi[2];
#pragma unroll
        for (int d = 0; d < 3; ++d) {
            lo[d] = uniform_f64(nan_box ? -__builtin_inf() : lo[d] - ClearGap<PREC>::value);
            hi[d] = uniform_f64(nan_box ? __builtin_inf() : hi[d] + ClearGap<PREC>::value);
        }
    }
    // the source may coincide with a target: not separated on any axis, or a NaN coordinate
    __device__ bool near(double x, double y, double z) const {
        const bool sep = x < lo[0] || x > hi[0] || y < lo[1] || y > hi[1] || z < lo[2] || z > hi[2];
        return !sep || x != x || y != y || z != z;
    }
};

} // namespace stkp2p
