// p2p_items.hpp -- the item kernel template and its per-micro-kernel launch helpers.
//
//   p2p_items_kernel<UK, PREC, T>  the hot kernel.  One wavefront per workgroup; each lane holds T
//       targets in registers; the wave walks the source ranges of its work item, staging 64 sources
//       at a time into LDS with coalesced loads (one source per lane), then every lane reads each
//       staged source by broadcast and evaluates the micro-kernel against its T targets.  Sums go
//       to per-item partial slots ([slot][KT][64*T], lane-contiguous, coalesced).
//       Persistent mode: a grid sized to the device's occupancy; each wave takes the next item of
//       the cost-sorted queue with one atomic on a counter, until the queue is empty (every wave
//       reaches that exit).  Otherwise one workgroup per item, in queue order.
//
// PREC 0 = FP64 throughout.  PREC 1 = mixed: coordinates in FP64 scaled by 2^level, displacement
// in FP64, micro-kernel (rsqrt + 1 Newton step, polynomial) in FP32, FP32 partial sums over
// kMixedAccBlock sources added into FP64 per-target accumulators.
//
// Coincidence mask: source tiles clear of the item's target box run the unmasked pair form
// (p2p_clear.hpp); in a U-list only the tiles of the target leaf and of nearby parts of touching
// neighbours keep the mask.
//
// make_uker_ops<UK>() bundles every launch entry of one micro-kernel (LDS and scalar-broadcast
// item kernels for PREC x T, their occupancy queries, record staging, dimensions) into a UkerOps
// table row; each translation unit instantiates the rows of its micro-kernels, so the kernel
// family compiles in parallel.
#pragma once

#include <hip/hip_runtime.h>

#include <type_traits>

#include "p2p_clear.hpp"
#include "p2p_internal.hpp"
#include "p2p_queue.hpp"
#include "p2p_redo.hpp"
#include "p2p_sgpr.hpp"
#include "ukernels.hpp"

namespace stkp2p {

// STKP2P_MIN_WAVES_PER_EU (build option): ask the compiler to keep at least this many waves of the
// item kernel resident per SIMD (register budget 512 / waves).  0 leaves the choice to it.
#ifndef STKP2P_MIN_WAVES_PER_EU
#define STKP2P_MIN_WAVES_PER_EU 0
#endif
// STKP2P_MIXED_HILO (build option, default 0): in mixed mode each level-scaled coordinate is held
// as an FP32 pair hi + lo (hi = fl32(x), lo = fl32(x - hi), ~48 bits together) and the displacement
// is formed in FP32 as (t_hi - s_hi) + (t_lo - s_lo): the hi difference is exact whenever the two
// coordinates are within a factor 2 of each other (Sterbenz), the lo difference is tiny, so the
// result is the FP64 displacement rounded once to FP32 -- as with an FP64 subtraction and a
// conversion, but in three FP32 operations per axis instead of two FP64-rate ones.  0: FP64
// coordinates, FP64 subtraction then conversion.  Measured on MI355X (1M PVel U-list, T = 2): the
// hi/lo form is ~2% slower (more instructions and registers; the FP64 subtraction and conversion
// cost less than assumed), so it is off by default.
#ifndef STKP2P_MIXED_HILO
#define STKP2P_MIXED_HILO 0
#endif
// STKP2P_FAR_F32 (build option, default 1): in mixed mode, a source tile clear of the item's target
// box by a margin forms its displacements in FP32 from coordinates taken relative to the box centre
// c (rounded to FP32 once per target per item and once per source per tile) -- one FP32 subtraction
// per axis instead of an FP64 subtraction and a conversion.  Each component of t - c and s - c is
// rounded with relative error <= 2^-24, so a displacement component errs by at most
// 2^-24 (|t - c| + |s - c| + r) (the last term: the FP32 subtraction's own rounding).  A tile is
// taken only when every source s satisfies |h| + |s - c| <= STKP2P_FAR_RATIO * dist(s, box), with
// |h| the box's half-diagonal (so |t - c| <= |h| and dist(s, box) <= r for every target t): then
// the error is at most (STKP2P_FAR_RATIO + 1) 2^-24 r per component, 3.0e-7 r at the default 4.
// Cancellation between nearby coordinates -- the reason the displacement is FP64 -- is confined
// to the tiles that fail the test, which keep the FP64 displacement.  The far-tile loop raises the
// kernel's register count, which costs occupancy where no tile is far (direct sums of unsorted
// points: 512 Stokeslet instances measured 42.5 -> 44.8 ms), so it is a separate instantiation
// (FARB = true, mixed, T = 2) that plans use only where target and source ranges are spatially
// compact: tree-derived work lists (groups with octree levels) and plans that order their points
// (morton.hip).
// STKP2P_FP64_UNROLL_T1 (build option, default 4): unroll of the FP64 source loop at one target per
// lane (small direct sums, whose few waves per SIMD leave the FP64 dependency chains exposed; more
// sources in flight per wave hide them).  8192^2 LapPGrad item kernel, same box, twice each: unroll 2
// 71.5 / 71.2 us, 4 67.6 / 68.4 us, 8 69.7 / 69.6 us.  The sums are added in the same order whatever
// the unroll.
#ifndef STKP2P_FP64_UNROLL_T1
#define STKP2P_FP64_UNROLL_T1 4
#endif

#ifndef STKP2P_FAR_F32
#define STKP2P_FAR_F32 1
#endif
#ifndef STKP2P_FAR_RATIO
#define STKP2P_FAR_RATIO 4.0
#endif
#if STKP2P_MIN_WAVES_PER_EU > 0
#define STKP2P_ITEM_ATTR __attribute__((amdgpu_waves_per_eu(STKP2P_MIN_WAVES_PER_EU, 8)))
#else
#define STKP2P_ITEM_ATTR
#endif

// unroll factor of the mixed-mode source loop: MIXED_UNROLL of the micro-kernel, else 2
template <class K, class = void> struct MixedUnroll { static constexpr int value = 2; };
template <class K> struct MixedUnroll<K, decltype((void)K::MIXED_UNROLL)> {
    static constexpr int value = K::MIXED_UNROLL;
};

// the mixed-mode source loop takes each block's first source on its own (MIXED_PEEL of the
// micro-kernel, else not): measured on MI355X, same box, twice each: the fused PVel U-list 7.870 ->
// 7.830 ms, the RPY U-list 39.98 -> 39.78 ms, the single-layer PVel FMM lists 6.84 -> 6.80 ms (X
// list -1.3 %, W list +0.4 %), but the ordered Stokeslet instances 34.54 -> 34.86 ms
template <class K, class = void> struct MixedPeel { static constexpr bool value = false; };
// the mixed-mode source loop at one lane per target fully unrolled over a block, each source at an
// immediate LDS offset from one base and a scheduling barrier between sources (MIXED_BLOCK_UNROLL of
// the micro-kernel, else not): measured on MI355X, same box, twice each: the RPY U-list 41.67 ->
// 40.90 ms (126 VGPRs and no scratch, against 128 and an 8-byte spill), but the fused PVel U-list
// 8.10 -> 8.58 ms and the ordered Stokeslet instances 36.17 -> 39.3 ms (the fused PVel with two
// sources per iteration and a barrier between them: 7.64 -> 7.92 ms; it keeps the rolled loop)
template <class K, class = void> struct MixedBlockUnroll { static constexpr bool value = false; };
template <class K> struct MixedBlockUnroll<K, decltype((void)K::MIXED_BLOCK_UNROLL)> {
    static constexpr bool value = K::MIXED_BLOCK_UNROLL;
};
template <class K> struct MixedPeel<K, decltype((void)K::MIXED_PEEL)> {
    static constexpr bool value = K::MIXED_PEEL;
};

// minimum resident waves per SIMD asked of the compiler for the mixed-mode item kernel:
// MIXED_MIN_WAVES of the micro-kernel, else 1 (no constraint).  The RPY U-list kernel asks for 4: the
// compiler then fits 128 VGPRs (instead of 158, 3 waves) by spilling a few item-level values to
// scratch outside the pair loops -- measured 42.0 -> 41.4 ms on the 4M RPY U-list (same box), same
// results.  For the fused PVel the same request measured no gain.
template <class K, class = void> struct MixedMinWaves { static constexpr int value = 1; };
template <class K> struct MixedMinWaves<K, decltype((void)K::MIXED_MIN_WAVES)> {
    static constexpr int value = K::MIXED_MIN_WAVES;
};
// sources per FP32 partial sum in mixed mode (a multiple of kMixedAccBlock that divides the 64-source
// tile): MIXED_ACC_SOURCES of the micro-kernel, else kMixedAccBlock.  The path (masked / clear / far)
// is still chosen per kMixedAccBlock sources.
template <class K, class = void> struct MixedFlush { static constexpr int value = kMixedAccBlock; };
template <class K> struct MixedFlush<K, decltype((void)K::MIXED_ACC_SOURCES)> {
    static constexpr int value = K::MIXED_ACC_SOURCES;
    static_assert(value % kMixedAccBlock == 0 && kSrcTile % value == 0, "MIXED_ACC_SOURCES");
};

template <int UK, int PREC> struct ItemMinWaves {
    static constexpr int value = (PREC != 0) ? MixedMinWaves<typename UkerOf<UK>::type>::value : 1;
};

template <int UK, int PREC, int T, bool FARB = false>
__global__ void __launch_bounds__(kWave) STKP2P_ITEM_ATTR
__attribute__((amdgpu_waves_per_eu(ItemMinWaves<UK, PREC>::value))) p2p_items_kernel(ItemArgs a) {
    using K = typename UkerOf<UK>::type;
    constexpr int KL = StagedCount<K, PREC>::value, KT = K::KT, KS = K::KS, KS1 = K::KS1;
    constexpr int TT = kWave * T;

    // FP64: values as [KL][tile] (two consecutive sources per 16-byte read);
    // mixed: FP32 values as 16-byte aligned rows [tile][KLP] (one source per few 16-byte reads,
    // all at immediate offsets from one base address), led by the hi/lo coordinates with HILO
    constexpr bool HILO = (PREC != 0) && (STKP2P_MIXED_HILO != 0);
    constexpr int KOFF = HILO ? 6 : 0;
    // relative FP32 coordinates of the far-tile form after the values
    constexpr bool FAR = FARB && (PREC != 0) && !HILO && HasUnmaskedEval<K>::value && (STKP2P_FAR_F32 != 0);
    constexpr int FOFF = KOFF + KL;
    // row pitch in floats: whole 16-byte quads.  A pitch of a multiple of 4 quads (16 floats) puts the
    // 64 lanes' row writes of a staging pass on 4 of the 64 LDS banks (16-way conflicts), so it gets
    // one more quad; other pitches conflict at most 2-way and stay as they are (padding the 2-quad
    // rows of the PVel single layer measured 8 % slower on its W/X lists; the reads are wave-uniform
    // broadcasts either way)
    constexpr int KLQ = (KL + KOFF + (FAR ? 3 : 0) + 3) / 4;
    constexpr int KLP = 4 * ((KLQ % 4 == 0) ? KLQ + 1 : KLQ);
    __shared__ double s_x[HILO ? 1 : kSrcTile], s_y[HILO ? 1 : kSrcTile], s_z[HILO ? 1 : kSrcTile];
    __shared__ double s_vd[PREC == 0 ? KL : 1][kSrcTile];
    __shared__ __attribute__((aligned(16))) float s_vf[PREC == 0 ? 1 : kSrcTile][KLP];

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
        // lanes per target (uniform): lane = target slot * split + sub
        const int ls = item.lsplit;
        const int split = 1 << ls;
        const int sub = lane & (split - 1);
        const int slot0 = lane >> ls;
        const int tstride = kWave >> ls;

        double tx[T], ty[T], tz[T];
#pragma unroll
        for (int t = 0; t < T; ++t) {
            const int i = slot0 + tstride * t;
            const long long idx = item.trg_begin + ((i < item.trg_count) ? i : item.trg_count - 1);
            tx[t] = a.tc.p[idx * a.tc.s] * lam;
            ty[t] = a.tc.p[idx * a.tc.s + a.tc.fs] * lam;
            tz[t] = a.tc.p[idx * a.tc.s + 2 * a.tc.fs] * lam;
        }
        float thx[T], thy[T], thz[T], tlx[T], tly[T], tlz[T];
        if constexpr (HILO) {
#pragma unroll
            for (int t = 0; t < T; ++t) {
                thx[t] = (float)tx[t];
                thy[t] = (float)ty[t];
                thz[t] = (float)tz[t];
                tlx[t] = (float)(tx[t] - (double)thx[t]);
                tly[t] = (float)(ty[t] - (double)thy[t]);
                tlz[t] = (float)(tz[t] - (double)thz[t]);
            }
        }
        // target box for the unmasked shortcut on source tiles clear of it (p2p_clear.hpp)
        constexpr bool CLEAR = HasUnmaskedEval<K>::value && !HILO;
        TargetBox<PREC> box;
        if constexpr (CLEAR)
            box.build(tx, ty, tz);
        // far-tile form: box centre, half-extent and half-diagonal (wave-uniform), targets relative
        // to the centre in FP32
        double fc[3] = {0.0, 0.0, 0.0}, fh[3] = {0.0, 0.0, 0.0}, fhn = 0.0;
        float tfx[T], tfy[T], tfz[T];
        if constexpr (FAR) {
            // computed by the vector unit (no scalar FP64), then held in scalar registers
#pragma unroll
            for (int d = 0; d < 3; ++d) {
                fc[d] = uniform_f64(0.5 * (box.lo[d] + box.hi[d]));
                fh[d] = uniform_f64(0.5 * (box.hi[d] - box.lo[d]));
            }
            fhn = uniform_f64(__builtin_sqrt(fh[0] * fh[0] + fh[1] * fh[1] + fh[2] * fh[2]));
#pragma unroll
            for (int t = 0; t < T; ++t) {
                tfx[t] = (float)(tx[t] - fc[0]);
                tfy[t] = (float)(ty[t] - fc[1]);
                tfz[t] = (float)(tz[t] - fc[2]);
            }
        }
        double acc[T][KT];
#pragma unroll
        for (int t = 0; t < T; ++t)
#pragma unroll
            for (int k = 0; k < KT; ++k)
                acc[t][k] = 0.0;

        for (int r = item.rng_begin; r < item.rng_end; ++r) {
            const int2 rg = a.ranges[r];
            for (int base = rg.x; base < rg.y; base += kSrcTile) {
                const int n = (rg.y - base < kSrcTile) ? rg.y - base : kSrcTile;
                bool near = false; // this lane's source may coincide with a target of the item
                bool far = true;   // this lane's source (if any) passes the far-tile test
                __syncthreads(); // the previous tile has been read by every lane
                if (lane < n) {
                    const long long s = base + lane;
                    double x = a.sc.p[s * a.sc.s];
                    double y = a.sc.p[s * a.sc.s + a.sc.fs];
                    double z = a.sc.p[s * a.sc.s + 2 * a.sc.fs];
                    double vin[KS];
#pragma unroll
                    for (int k = 0; k < KS; ++k)
                        vin[k] = (k < KS1) ? a.v1.p[s * a.v1.s + k * a.v1.fs]
                                           : a.v2.p[s * a.v2.s + (k - KS1) * a.v2.fs];
                    if (PREC != 0) {
                        x *= lam;
                        y *= lam;
                        z *= lam;
#pragma unroll
                        for (int k = 0; k < KS; ++k)
                            if ((K::SRC_SCALE_MASK >> k) & 1u)
                                vin[k] *= lam;
                    }
                    if constexpr (CLEAR)
                        near = box.near(x, y, z);
                    // mixed: this source's LDS row, assembled in registers and written as 16-byte quads
                    float row[PREC == 0 ? 1 : KLP];
                    if constexpr (PREC != 0)
#pragma unroll
                        for (int k = 0; k < KLP; ++k)
                            row[k] = 0.0f;
                    if constexpr (FAR) {
                        const double rx = x - fc[0], ry = y - fc[1], rz = z - fc[2];
                        const double gx = __builtin_fmax(__builtin_fabs(rx) - fh[0], 0.0);
                        const double gy = __builtin_fmax(__builtin_fabs(ry) - fh[1], 0.0);
                        const double gz = __builtin_fmax(__builtin_fabs(rz) - fh[2], 0.0);
                        const double lhs = fhn + __builtin_sqrt(rx * rx + ry * ry + rz * rz);
                        // false for NaN coordinates (near is then true as well)
                        constexpr double R2 = STKP2P_FAR_RATIO * STKP2P_FAR_RATIO;
                        far = !near && lhs * lhs <= R2 * (gx * gx + gy * gy + gz * gz);
                        row[FOFF] = (float)rx;
                        row[FOFF + 1] = (float)ry;
                        row[FOFF + 2] = (float)rz;
                    }
                    double vst[KL];
                    stage_for<K, PREC>(vin, vst);
                    if constexpr (HILO) {
                        const float hx = (float)x, hy = (float)y, hz = (float)z;
                        row[0] = hx;
                        row[1] = hy;
                        row[2] = hz;
                        row[3] = (float)(x - (double)hx);
                        row[4] = (float)(y - (double)hy);
                        row[5] = (float)(z - (double)hz);
                    } else {
                        s_x[lane] = x;
                        s_y[lane] = y;
                        s_z[lane] = z;
                    }
#pragma unroll
                    for (int k = 0; k < KL; ++k) {
                        if constexpr (PREC == 0)
                            s_vd[k][lane] = vst[k];
                        else
                            row[KOFF + k] = (float)vst[k];
                    }
                    if constexpr (PREC != 0) {
                        float4 *dst = reinterpret_cast<float4 *>(&s_vf[lane][0]);
#pragma unroll
                        for (int q = 0; q < KLP / 4; ++q)
                            dst[q] = make_float4(row[4 * q], row[4 * q + 1], row[4 * q + 2], row[4 * q + 3]);
                    }
                }
                __syncthreads();
                // sources of this tile: every lane takes all of them (split == 1, uniform loop with
                // immediate LDS offsets) or every split-th one starting at its sub-lane
                // FP64: the whole tile, sources [jb, je) = [0, n); mixed: one block of kMixedAccBlock
                // sources [jb, je), whose FP32 partial sums then join the FP64 accumulators
                // mixed: FP32 partial sums of the current group of MixedFlush<K> sources, added into the
                // FP64 accumulators when the group ends
                // (-0 is the exact additive identity: -0 + x == x for every x, +0 included)
                float part[T][KT];
#pragma unroll
                for (int t = 0; t < T; ++t)
#pragma unroll
                    for (int k = 0; k < KT; ++k)
                        part[t][k] = -0.0f;
                auto tile_pairs = [&](auto strided, auto clear, auto farc, int jb, int je) {
                    constexpr bool STR = decltype(strided)::value;
                    constexpr bool MASK = !decltype(clear)::value;
                    constexpr bool FARM = decltype(farc)::value;
                    const int j_first = STR ? sub : 0;
                    const int j_step = STR ? split : 1;
                    if (PREC == 0) {
                        constexpr int U64 = (T == 1) ? STKP2P_FP64_UNROLL_T1 : 2;
                        auto src = [&](int j) {
                            const double sx = s_x[j], sy = s_y[j], sz = s_z[j];
                            double f[KL];
#pragma unroll
                            for (int k = 0; k < KL; ++k)
                                f[k] = s_vd[k][j];
#pragma unroll
                            for (int t = 0; t < T; ++t)
                                eval_pair<K, MASK>(acc[t], tx[t] - sx, ty[t] - sy, tz[t] - sz, f);
                        };
                        if (T == 1 && !STR && je - jb == kSrcTile) {
                            // a whole tile at one target per lane: a fixed trip count, so the
                            // unrolled body has no exit tests between its sources and their LDS reads
                            // are issued together (8192^2 LapPGrad item kernel 66.8 -> 65.9 us; at
                            // T = 2 the FMM surface lists measured 0.6-2 % slower, so they keep the
                            // loop below)
#pragma unroll U64
                            for (int j = 0; j < kSrcTile; ++j)
                                src(jb + j);
                        } else {
#pragma unroll U64
                            for (int j = jb + j_first; j < je; j += j_step)
                                src(j);
                        }
                    } else {
                        auto src = [&](int j) {
                            // coordinates first: LDS reads return in issue order, and the
                            // displacement is needed before the values
                            float hx, hy, hz, lx, ly, lz;
                            double sx, sy, sz;
                            if constexpr (FARM) {
                                hx = s_vf[j][FOFF];
                                hy = s_vf[j][FOFF + 1];
                                hz = s_vf[j][FOFF + 2];
                            } else if constexpr (HILO) {
                                hx = s_vf[j][0];
                                hy = s_vf[j][1];
                                hz = s_vf[j][2];
                                lx = s_vf[j][3];
                                ly = s_vf[j][4];
                                lz = s_vf[j][5];
                            } else {
                                sx = s_x[j];
                                sy = s_y[j];
                                sz = s_z[j];
                            }
                            float f[KL];
#pragma unroll
                            for (int k = 0; k < KL; ++k)
                                f[k] = s_vf[j][KOFF + k];
#pragma unroll
                            for (int t = 0; t < T; ++t) {
                                if constexpr (FARM)
                                    eval_pair<K, false>(part[t], tfx[t] - hx, tfy[t] - hy, tfz[t] - hz, f);
                                else if constexpr (HILO)
                                    eval_pair<K, MASK>(part[t], (thx[t] - hx) + (tlx[t] - lx),
                                                            (thy[t] - hy) + (tly[t] - ly),
                                                            (thz[t] - hz) + (tlz[t] - lz), f);
                                else
                                    eval_pair<K, MASK>(part[t], (float)(tx[t] - sx), (float)(ty[t] - sy),
                                                       (float)(tz[t] - sz), f);
                            }
                        };
                        // MixedPeel: the block's first source on its own -- where the sums start the
                        // block at -0 (groups of one block), its accumulations fold into plain
                        // products and the sums need no clearing.  Then the rest, MixedUnroll<K>
                        // sources per iteration (measured on MI355X, T = 2: the fused PVel U-list
                        // 8.37 ms at 1, 8.62 at 2, 8.91 at 4; RPY 1 % faster at 2 than at 1)
                        int j = jb + j_first;
                        if constexpr (!STR && MixedBlockUnroll<K>::value) {
                            // one lane per target: the block's sources fully unrolled, each at an
                            // immediate LDS offset from one base (no per-source address or loop
                            // counter), with a scheduling barrier between sources so that registers
                            // stay as in the rolled loop
#pragma unroll
                            for (int u = 0; u < kMixedAccBlock; ++u) {
                                if (jb + u < je)
                                    src(jb + u);
                                __builtin_amdgcn_sched_barrier(0);
                            }
                        } else {
                            if constexpr (MixedPeel<K>::value) {
                                if (j < je)
                                    src(j);
                                j += j_step;
                            }
#pragma unroll MixedUnroll<K>::value
                            for (; j < je; j += j_step)
                                src(j);
                        }
                    }
                };
                using F = std::false_type;
                using Tr = std::true_type;
                // sources [jb, je) in the form their flags allow: uniform decisions, one per call
                auto run = [&](int jb, int je, bool clear, bool far_tile) {
                    if (ls == 0) {
                        if (far_tile)
                            tile_pairs(F{}, Tr{}, std::integral_constant<bool, FAR>{}, jb, je);
                        else if (clear)
                            tile_pairs(F{}, Tr{}, F{}, jb, je);
                        else
                            tile_pairs(F{}, F{}, F{}, jb, je);
                    } else {
                        if (far_tile)
                            tile_pairs(Tr{}, Tr{}, std::integral_constant<bool, FAR>{}, jb, je);
                        else if (clear)
                            tile_pairs(Tr{}, Tr{}, F{}, jb, je);
                        else
                            tile_pairs(Tr{}, F{}, F{}, jb, je);
                    }
                };
                if constexpr (PREC == 0) {
                    // uniform: no lane holds a source near the targets
                    run(0, n, CLEAR && !__any(near), false);
                } else {
                    // per block of kMixedAccBlock sources (lane j holds source j's flags; lanes past n
                    // hold no source and count as clear and far): a block with no source near the
                    // targets drops the mask, and one whose every source passes the far test takes
                    // the FP32 displacement -- a finer choice than the whole tile's, which on the 1M
                    // PVel U-list moves ~10 % of the pairs from clear to far tiles
                    const unsigned long long nm = __ballot(near), fm = __ballot(far);
                    constexpr unsigned long long BM =
                        (kMixedAccBlock >= 64) ? ~0ull : ((1ull << kMixedAccBlock) - 1ull);
                    constexpr int FL = MixedFlush<K>::value;
                    for (int j0 = 0; j0 < n; j0 += kMixedAccBlock) {
                        const unsigned long long bm = BM << j0;
                        const bool bclear = CLEAR && (nm & bm) == 0ull;
                        const bool bfar = FAR && bclear && (fm & bm) == bm;
                        const int j1 = (n - j0 < kMixedAccBlock) ? n : j0 + kMixedAccBlock;
                        run(j0, j1, bclear, bfar);
                        // (a group of one block flushes unconditionally: j1 is then always a multiple
                        // of the block or n, and saying so lets the compiler start every block's
                        // sums from zero instead of carrying them between the forms' loops)
                        if (FL == kMixedAccBlock || j1 % FL == 0 || j1 == n) {
#pragma unroll
                            for (int t = 0; t < T; ++t)
#pragma unroll
                                for (int k = 0; k < KT; ++k) {
                                    acc[t][k] += (double)part[t][k];
                                    part[t][k] = -0.0f;
                                }
                        }
                    }
                }
            }
        }

        if constexpr (PREC != 0)
            redo_nonfinite_fp64<K>(a, item, acc, ls ? sub : 0, split, tx, ty, tz, lam);
        // lanes that share a target add their partial sums (butterfly over the low ls lane bits)
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

// ---- per-micro-kernel launch helpers -----------------------------------------------------------

template <int UK, int PREC, int T, bool FARB = false>
hipError_t launch_items_one(int grid, const ItemArgs &a, hipStream_t s) {
    hipLaunchKernelGGL((p2p_items_kernel<UK, PREC, T, FARB>), dim3(grid), dim3(kWave), 0, s, a);
    return hipGetLastError();
}

template <int UK, int PREC, int T, bool FARB = false>
hipError_t occ_items_one(int *blocks) {
    return hipOccupancyMaxActiveBlocksPerMultiprocessor(blocks, p2p_items_kernel<UK, PREC, T, FARB>, kWave, 0);
}

template <int UK, int PREC, int T>
hipError_t launch_sgpr_one(int grid, const ItemArgs &a, const double *recs, int level_min, hipStream_t s) {
    hipLaunchKernelGGL((p2p_items_sgpr_kernel<UK, PREC, T>), dim3(grid), dim3(kWave), 0, s, a, recs, level_min);
    return hipGetLastError();
}

template <int UK, int PREC, int T>
hipError_t occ_sgpr_one(int *blocks) {
    return hipOccupancyMaxActiveBlocksPerMultiprocessor(blocks, p2p_items_sgpr_kernel<UK, PREC, T>, kWave, 0);
}

// PREC in {0, 1} x T in {1, 2, 4}
#define STKP2P_PT_SWITCH(FN, ...)                                                                  \
    switch (prec * 8 + T) {                                                                        \
    case 1: return FN<UK, 0, 1>(__VA_ARGS__);                                                      \
    case 2: return FN<UK, 0, 2>(__VA_ARGS__);                                                      \
    case 4: return FN<UK, 0, 4>(__VA_ARGS__);                                                      \
    case 9: return FN<UK, 1, 1>(__VA_ARGS__);                                                      \
    case 10: return FN<UK, 1, 2>(__VA_ARGS__);                                                     \
    case 12: return FN<UK, 1, 4>(__VA_ARGS__);                                                     \
    default: return hipErrorInvalidValue;                                                          \
    }

// the LDS item kernel also takes prec = kPrecMixedFar (mixed with the far-tile form, T = 2 only)
#define STKP2P_ITEMS_SWITCH(FN, ...)                                                               \
    if (prec == kPrecMixedFar)                                                                     \
        return (T == 2) ? FN<UK, 1, 2, true>(__VA_ARGS__) : hipErrorInvalidValue;                 \
    STKP2P_PT_SWITCH(FN, __VA_ARGS__)

template <int UK>
hipError_t launch_items_uk(int prec, int T, int grid, const ItemArgs &a, hipStream_t s) {
    STKP2P_ITEMS_SWITCH(launch_items_one, grid, a, s)
}

template <int UK>
hipError_t occ_items_uk(int prec, int T, int *blocks) {
    STKP2P_ITEMS_SWITCH(occ_items_one, blocks)
}

#undef STKP2P_ITEMS_SWITCH

template <int UK>
hipError_t launch_sgpr_uk(int prec, int T, int grid, const ItemArgs &a, const double *recs, int lmin,
                          hipStream_t s) {
    STKP2P_PT_SWITCH(launch_sgpr_one, grid, a, recs, lmin, s)
}

template <int UK>
hipError_t occ_sgpr_uk(int prec, int T, int *blocks) {
    STKP2P_PT_SWITCH(occ_sgpr_one, blocks)
}

#undef STKP2P_PT_SWITCH

template <int UK>
hipError_t stage_records_uk(int prec, const StageArgs &a, int n_copies, hipStream_t s) {
    const dim3 grid((unsigned)((a.n_src + 255) / 256), (unsigned)n_copies);
    if (prec == 0)
        hipLaunchKernelGGL((stage_records_kernel<UK, 0>), grid, dim3(256), 0, s, a);
    else
        hipLaunchKernelGGL((stage_records_kernel<UK, 1>), grid, dim3(256), 0, s, a);
    return hipGetLastError();
}

template <int UK>
UkerOps make_uker_ops() {
    using K = typename UkerOf<UK>::type;
    UkerOps o;
    o.launch = launch_items_uk<UK>;
    o.occ = occ_items_uk<UK>;
    o.launch_sgpr = launch_sgpr_uk<UK>;
    o.occ_sgpr = occ_sgpr_uk<UK>;
    o.stage = stage_records_uk<UK>;
    o.words[0] = RecLayout<K, 0>::WORDS;
    o.words[1] = RecLayout<K, 1>::WORDS;
    o.ks = K::KS;
    o.ks1 = K::KS1;
    o.kt = K::KT;
    return o;
}

} // namespace stkp2p
