// p2p_sgpr.hpp -- scalar-broadcast variant of the item kernel.
//
// The LDS variant stages 64 sources per tile and every lane reads each staged source back from LDS
// (one broadcast read per value, into VGPRs).  Here the per-source records are written once per
// execute into HBM by stage_records_kernel (coordinates in FP64, the staged values in the compute
// precision, one copy per octree level that occurs among the work items, scaled by 2^level in mixed
// mode), and the item kernel reads record j of a work item's source range with scalar loads
// (the record address is uniform across the wave): source data live in SGPRs and are consumed
// directly as VALU operands -- no LDS traffic, no barrier, and no VGPRs for the source values, which
// leaves room for more targets per lane.
//
// Tiles that spread a target over several lanes (lsplit > 0) read records with per-lane vector
// loads instead, since their lanes walk different sources.
//
// Coincidence mask: the uniform path walks each range in chunks of 64 records; a chunk whose
// sources are all clear of the item's target box (checked one source per lane, one ballot) runs the
// unmasked pair form (p2p_clear.hpp).
#pragma once

#include <hip/hip_runtime.h>

#include <type_traits>

#include "p2p_clear.hpp"
#include "p2p_internal.hpp"
#include "p2p_queue.hpp"
#include "p2p_redo.hpp"
#include "ukernels.hpp"

namespace stkp2p {

template <class K, int PREC>
struct RecLayout {
    // 8-byte words per record: x, y, z, then KL values (FP64: one word each; mixed: two FP32 per word)
    static constexpr int KL = StagedCount<K, PREC>::value;
    static constexpr int VW = (PREC == 0) ? KL : (KL + 1) / 2;
    static constexpr int WORDS = 3 + VW;
};

typedef const __attribute__((address_space(4))) double cdouble_t;
typedef const __attribute__((address_space(4))) float cfloat_t;

// one record per (level copy, source); grid.x over sources, grid.y over level copies
template <int UK, int PREC>
__global__ void __launch_bounds__(256) stage_records_kernel(StageArgs a) {
    using K = typename UkerOf<UK>::type;
    using L = RecLayout<K, PREC>;
    constexpr int KS = K::KS, KS1 = K::KS1, KL = L::KL;
    const long long s = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= a.n_src)
        return;
    const int copy = blockIdx.y;
    const double lam = (PREC == 0) ? 1.0 : ldexp(1.0, a.level_min + copy);
    double vin[KS];
#pragma unroll
    for (int k = 0; k < KS; ++k) {
        vin[k] = (k < KS1) ? a.v1.p[s * a.v1.s + k * a.v1.fs] : a.v2.p[s * a.v2.s + (k - KS1) * a.v2.fs];
        if (PREC != 0 && ((K::SRC_SCALE_MASK >> k) & 1u))
            vin[k] *= lam;
    }
    double vst[KL];
    stage_for<K, PREC>(vin, vst);
    double *rec = a.rec + ((long long)copy * a.n_src + s) * L::WORDS;
    rec[0] = a.sc.p[s * a.sc.s] * lam;
    rec[1] = a.sc.p[s * a.sc.s + a.sc.fs] * lam;
    rec[2] = a.sc.p[s * a.sc.s + 2 * a.sc.fs] * lam;
    if (PREC == 0) {
#pragma unroll
        for (int k = 0; k < KL; ++k)
            rec[3 + k] = vst[k];
    } else {
        float *fv = reinterpret_cast<float *>(rec + 3);
#pragma unroll
        for (int k = 0; k < 2 * L::VW; ++k)
            fv[k] = (k < KL) ? (float)vst[k] : 0.0f;
    }
}

template <int UK, int PREC, int T>
__global__ void __launch_bounds__(kWave) p2p_items_sgpr_kernel(ItemArgs a, const double *recs, int level_min) {
    using K = typename UkerOf<UK>::type;
    using L = RecLayout<K, PREC>;
    constexpr int KL = L::KL, KT = K::KT;
    constexpr int TT = kWave * T;
    const int lane = threadIdx.x;
    ItemCursor cur(a);
    for (;;) {
        // uniform by construction; readfirstlane lets the compiler keep the item in SGPRs (and the
        // scalar-broadcast variant's record reads as scalar loads)
        const int it = __builtin_amdgcn_readfirstlane(next_item(a, cur, lane));
        if (it < 0)
            break;

        const Item item = a.items[it];
        const double lam = (PREC == 0) ? 1.0 : ldexp(1.0, item.level);
        const int ls = item.lsplit;
        const int split = 1 << ls;
        const int sub = lane & (split - 1);
        const int slot0 = lane >> ls;
        const int tstride = kWave >> ls;
        const long long copy = (PREC == 0) ? 0 : item.level - level_min;
        const double *rbase = recs + copy * (long long)a.n_src_rec * L::WORDS;

        double tx[T], ty[T], tz[T];
#pragma unroll
        for (int t = 0; t < T; ++t) {
            const int i = slot0 + tstride * t;
            const long long idx = item.trg_begin + ((i < item.trg_count) ? i : item.trg_count - 1);
            tx[t] = a.tc.p[idx * a.tc.s] * lam;
            ty[t] = a.tc.p[idx * a.tc.s + a.tc.fs] * lam;
            tz[t] = a.tc.p[idx * a.tc.s + 2 * a.tc.fs] * lam;
        }
        constexpr bool CLEAR = HasUnmaskedEval<K>::value;
        TargetBox<PREC> box;
        if constexpr (CLEAR)
            box.build(tx, ty, tz);
        double acc[T][KT];
#pragma unroll
        for (int t = 0; t < T; ++t)
#pragma unroll
            for (int k = 0; k < KT; ++k)
                acc[t][k] = 0.0;

        // one source record against the lane's T targets; P is the record pointer type (scalar
        // address space for the uniform path, generic for the per-lane path)
        auto pair_fp64 = [&](auto p, auto unmasked) {
            constexpr bool MASK = !decltype(unmasked)::value;
            const double sx = p[0], sy = p[1], sz = p[2];
            double f[KL];
#pragma unroll
            for (int k = 0; k < KL; ++k)
                f[k] = p[3 + k];
#pragma unroll
            for (int t = 0; t < T; ++t)
                eval_pair<K, MASK>(acc[t], tx[t] - sx, ty[t] - sy, tz[t] - sz, f);
        };

        for (int r = item.rng_begin; r < item.rng_end; ++r) {
            const int2 rg = a.ranges[r];
            if (ls == 0) {
                cdouble_t *R = (cdouble_t *)(unsigned long long)(rbase);
                // records [c0, c1) of the range, masked or not
                auto chunk = [&](int c0, int c1, auto unmasked) {
                    constexpr bool MASK = !decltype(unmasked)::value;
                    if (PREC == 0) {
#pragma unroll 4
                        for (int j = c0; j < c1; ++j)
                            pair_fp64(R + (long long)j * L::WORDS, unmasked);
                    } else {
                        for (int j0 = c0; j0 < c1; j0 += kMixedAccBlock) {
                            const int j1 = (c1 - j0 < kMixedAccBlock) ? c1 : j0 + kMixedAccBlock;
                            float part[T][KT];
#pragma unroll
                            for (int t = 0; t < T; ++t)
#pragma unroll
                                for (int k = 0; k < KT; ++k)
                                    part[t][k] = 0.0f;
#pragma unroll 4
                            for (int j = j0; j < j1; ++j) {
                                cdouble_t *p = R + (long long)j * L::WORDS;
                                const double sx = p[0], sy = p[1], sz = p[2];
                                cfloat_t *pf = (cfloat_t *)(p + 3);
                                float f[KL];
#pragma unroll
                                for (int k = 0; k < KL; ++k)
                                    f[k] = pf[k];
#pragma unroll
                                for (int t = 0; t < T; ++t)
                                    eval_pair<K, MASK>(part[t], (float)(tx[t] - sx), (float)(ty[t] - sy),
                                                       (float)(tz[t] - sz), f);
                            }
#pragma unroll
                            for (int t = 0; t < T; ++t)
#pragma unroll
                                for (int k = 0; k < KT; ++k)
                                    acc[t][k] += (double)part[t][k];
                        }
                    }
                };
                // chunks of 64 records (a multiple of kMixedAccBlock, so the FP32 partial sums group
                // the same sources as a walk of the whole range)
                for (int c0 = rg.x; c0 < rg.y; c0 += kWave) {
                    const int c1 = (rg.y - c0 < kWave) ? rg.y : c0 + kWave;
                    bool clear = false;
                    if constexpr (CLEAR) {
                        bool near = false;
                        if (c0 + lane < c1) {
                            const double *p = rbase + (long long)(c0 + lane) * L::WORDS;
                            near = box.near(p[0], p[1], p[2]);
                        }
                        clear = !__any(near);
                    }
                    if (clear)
                        chunk(c0, c1, std::true_type{});
                    else
                        chunk(c0, c1, std::false_type{});
                }
            } else {
                // per-lane sources: lane sub takes every split-th source of the range
                if (PREC == 0) {
                    for (int j = rg.x + sub; j < rg.y; j += split)
                        pair_fp64(rbase + (long long)j * L::WORDS, std::false_type{});
                } else {
                    for (int j0 = rg.x; j0 < rg.y; j0 += kMixedAccBlock) {
                        const int j1 = (rg.y - j0 < kMixedAccBlock) ? rg.y : j0 + kMixedAccBlock;
                        float part[T][KT];
#pragma unroll
                        for (int t = 0; t < T; ++t)
#pragma unroll
                            for (int k = 0; k < KT; ++k)
                                part[t][k] = 0.0f;
                        for (int j = j0 + sub; j < j1; j += split) {
                            const double *p = rbase + (long long)j * L::WORDS;
                            const double sx = p[0], sy = p[1], sz = p[2];
                            const float *pf = reinterpret_cast<const float *>(p + 3);
                            float f[KL];
#pragma unroll
                            for (int k = 0; k < KL; ++k)
                                f[k] = pf[k];
#pragma unroll
                            for (int t = 0; t < T; ++t)
                                K::template eval<float>(part[t], (float)(tx[t] - sx), (float)(ty[t] - sy),
                                                        (float)(tz[t] - sz), f);
                        }
#pragma unroll
                        for (int t = 0; t < T; ++t)
#pragma unroll
                            for (int k = 0; k < KT; ++k)
                                acc[t][k] += (double)part[t][k];
                    }
                }
            }
        }

        if constexpr (PREC != 0)
            redo_nonfinite_fp64<K>(a, item, acc, sub, split, tx, ty, tz, lam);
        for (int off = 1; off < split; off <<= 1) {
#pragma unroll
            for (int t = 0; t < T; ++t)
#pragma unroll
                for (int k = 0; k < KT; ++k)
                    acc[t][k] += __shfl_xor(acc[t][k], off);
        }
        double *dst = a.partial + (long long)item.slot * KT * TT;
#pragma unroll
        for (int k = 0; k < KT; ++k) {
            const double osc = ((PREC == 0) ? 1.0 : ldexp(1.0, K::deg(k) * item.level)) * post_for<K, PREC>(k);
#pragma unroll
            for (int t = 0; t < T; ++t) {
                const int i = slot0 + tstride * t;
                if (sub == 0 && i < item.trg_count)
                    dst[(long long)k * TT + i] = acc[t][k] * osc;
            }
        }
    }
}

} // namespace stkp2p
