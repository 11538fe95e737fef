// p2p_redo.hpp -- the FP64 path of a mixed-mode work item whose FP32 sums overflowed.
//
// Mixed mode evaluates each pair in FP32.  A pair closer than FP32 can represent (the 1/r^3 of a
// scaled separation below ~1.4e-13 overflows; so does a source value near the FP32 range) leaves
// its FP32 partial sum, and with it the target's FP64 accumulator, non-finite where the FP64 result
// is finite.  At the end of each work item, one wave-uniform test on the accumulators finds such
// targets (none, almost always), and each of them has the item's pairs summed again in FP64 from the
// source arrays in HBM, replacing its accumulators.  Non-finite inputs take the same path and give
// the FP64 mode's result.
#pragma once

#include <hip/hip_runtime.h>

#include "p2p_clear.hpp"
#include "p2p_internal.hpp"
#include "ukernels.hpp"

namespace stkp2p {

// acc[KT] = the FP64 sum of the item's pairs for one target (coordinates tx, ty, tz, already scaled
// by lam): sources rg.x + jf, rg.x + jf + js, ... of each source range (the sources this lane takes
// in the item kernels), coordinates and length-valued source values scaled by lam as the staging does
template <class K>
__device__ __forceinline__ void fp64_item_sums(const ItemArgs &a, const Item &item, double *acc, int jf, int js,
                                               double tx, double ty, double tz, double lam) {
    constexpr int KS = K::KS, KS1 = K::KS1, KL = StagedCount<K, 0>::value, KT = K::KT;
#pragma unroll
    for (int k = 0; k < KT; ++k)
        acc[k] = 0.0;
    for (int r = item.rng_begin; r < item.rng_end; ++r) {
        const int2 rg = a.ranges[r];
        for (long long s = (long long)rg.x + jf; s < rg.y; s += js) {
            double vin[KS];
#pragma unroll
            for (int k = 0; k < KS; ++k) {
                vin[k] = (k < KS1) ? a.v1.p[s * a.v1.s + k * a.v1.fs] : a.v2.p[s * a.v2.s + (k - KS1) * a.v2.fs];
                if ((K::SRC_SCALE_MASK >> k) & 1u)
                    vin[k] *= lam;
            }
            double vst[KL];
            stage_for<K, 0>(vin, vst);
            const double sx = a.sc.p[s * a.sc.s] * lam;
            const double sy = a.sc.p[s * a.sc.s + a.sc.fs] * lam;
            const double sz = a.sc.p[s * a.sc.s + 2 * a.sc.fs] * lam;
            eval_pair<K, true>(acc, tx - sx, ty - sy, tz - sz, vst);
        }
    }
    // the caller scales by the mixed form's post(); where the FP64 form accumulates an output at
    // another power of 2, bring it to the mixed form's (exact)
#pragma unroll
    for (int k = 0; k < KT; ++k)
        if constexpr (PostsByPrecision<K>::value)
            acc[k] *= post_for<K, 0>(k) / post_for<K, 1>(k);
}

// the end-of-item test and the FP64 path for the targets that need it
template <class K, int T>
__device__ __forceinline__ void redo_nonfinite_fp64(const ItemArgs &a, const Item &item, double (&acc)[T][K::KT],
                                                    int jf, int js, const double (&tx)[T], const double (&ty)[T],
                                                    const double (&tz)[T], double lam) {
    bool fin[T];
    bool all = true;
#pragma unroll
    for (int t = 0; t < T; ++t) {
        fin[t] = true;
#pragma unroll
        for (int k = 0; k < K::KT; ++k)
            fin[t] = fin[t] && __builtin_isfinite(acc[t][k]);
        all = all && fin[t];
    }
    if (__builtin_expect(__all(all), 1))
        return;
#pragma unroll
    for (int t = 0; t < T; ++t)
        if (!fin[t])
            fp64_item_sums<K>(a, item, acc[t], jf, js, tx[t], ty[t], tz[t], lam);
}

} // namespace stkp2p
