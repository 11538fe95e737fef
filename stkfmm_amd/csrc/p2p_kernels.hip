// p2p_kernels.hip -- gfx950 kernels of the STKFMM P2P near-field path.
//
//   p2p_items_kernel<UK, PREC, T>  the hot kernel.  One wavefront per workgroup; each lane holds T
//       targets in registers; the wave walks the source ranges of its work item, staging 64 sources
//       at a time into LDS with coalesced loads (one source per lane), then every lane reads each
//       staged source by broadcast and evaluates the micro-kernel against its T targets.  Sums go
//       to per-item partial slots ([slot][KT][64*T], lane-contiguous, coalesced).
//       Persistent mode: a grid sized to the device's occupancy; each wave takes the next item of
//       the cost-sorted queue with one atomic on a counter, until the queue is empty (every wave
//       reaches that exit).  Otherwise one workgroup per item, in queue order.
//   p2p_reduce_kernel  per target: sum of its tile's partial slots in slot order (deterministic),
//       times the scale factor, added (+=) into the caller's output (optionally permuted).
//   vjp kernels        adjoint of a direct sum (target-centric and source-centric passes).
//
// PREC 0 = FP64 throughout.  PREC 1 = mixed: coordinates in FP64 scaled by 2^level, displacement
// in FP64, micro-kernel (rsqrt + 1 Newton step, polynomial) in FP32, FP32 partial sums over
// kMixedAccBlock sources added into FP64 per-target accumulators.
#include <hip/hip_runtime.h>

#include <type_traits>

#include "p2p_internal.hpp"
#include "ukernels.hpp"
#include "p2p_sgpr.hpp"
#include "vjp_kernels.hpp"

namespace stkp2p {

// STKP2P_MIN_WAVES_PER_EU (build option): ask the compiler to keep at least this many waves of the
// item kernel resident per SIMD (register budget 512 / waves).  0 leaves the choice to it.
#ifndef STKP2P_MIN_WAVES_PER_EU
#define STKP2P_MIN_WAVES_PER_EU 0
#endif
#if STKP2P_MIN_WAVES_PER_EU > 0
#define STKP2P_ITEM_ATTR __attribute__((amdgpu_waves_per_eu(STKP2P_MIN_WAVES_PER_EU, 8)))
#else
#define STKP2P_ITEM_ATTR
#endif

template <int UK, int PREC, int T>
__global__ void __launch_bounds__(kWave) STKP2P_ITEM_ATTR p2p_items_kernel(ItemArgs a) {
    using K = typename UkerOf<UK>::type;
    constexpr int KL = K::KL, KT = K::KT, KS = K::KS, KS1 = K::KS1;
    constexpr int TT = kWave * T;

    // FP64: values as [KL][tile] (two consecutive sources per 16-byte read);
    // mixed: FP32 values as 16-byte aligned rows [tile][KLP] (one source per few 16-byte reads,
    // all at immediate offsets from one base address)
    constexpr int KLP = ((KL + 3) / 4) * 4;
    __shared__ double s_x[kSrcTile], s_y[kSrcTile], s_z[kSrcTile];
    __shared__ double s_vd[PREC == 0 ? KL : 1][kSrcTile];
    __shared__ __attribute__((aligned(16))) float s_vf[PREC == 0 ? 1 : kSrcTile][KLP];

    const int lane = threadIdx.x;
    bool first = true;
    for (;;) {
        int it;
        if (a.counter) {
            int v = 0;
            if (lane == 0)
                v = atomicAdd(a.counter, 1);
            it = __builtin_amdgcn_readfirstlane(v);
        } else {
            if (!first)
                break;
            it = blockIdx.x;
        }
        first = false;
        if (it >= a.n_items)
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
                    double vst[KL];
                    K::stage(vin, vst);
                    s_x[lane] = x;
                    s_y[lane] = y;
                    s_z[lane] = z;
#pragma unroll
                    for (int k = 0; k < KL; ++k) {
                        if (PREC == 0)
                            s_vd[k][lane] = vst[k];
                        else
                            s_vf[lane][k] = (float)vst[k];
                    }
                }
                __syncthreads();
                // sources of this tile: every lane takes all of them (split == 1, uniform loop with
                // immediate LDS offsets) or every split-th one starting at its sub-lane
                auto tile_pairs = [&](auto strided) {
                    constexpr bool STR = decltype(strided)::value;
                    const int j_first = STR ? sub : 0;
                    const int j_step = STR ? split : 1;
                    if (PREC == 0) {
#pragma unroll 2
                        for (int j = j_first; j < n; j += j_step) {
                            const double sx = s_x[j], sy = s_y[j], sz = s_z[j];
                            double f[KL];
#pragma unroll
                            for (int k = 0; k < KL; ++k)
                                f[k] = s_vd[k][j];
#pragma unroll
                            for (int t = 0; t < T; ++t)
                                K::template eval<double>(acc[t], tx[t] - sx, ty[t] - sy, tz[t] - sz, f);
                        }
                    } else {
                        for (int j0 = 0; j0 < n; j0 += kMixedAccBlock) {
                            const int j1 = (n - j0 < kMixedAccBlock) ? n : j0 + kMixedAccBlock;
                            float part[T][KT];
#pragma unroll
                            for (int t = 0; t < T; ++t)
#pragma unroll
                                for (int k = 0; k < KT; ++k)
                                    part[t][k] = 0.0f;
#pragma unroll 2
                            for (int j = j0 + j_first; j < j1; j += j_step) {
                                const double sx = s_x[j], sy = s_y[j], sz = s_z[j];
                                float f[KL];
#pragma unroll
                                for (int k = 0; k < KL; ++k)
                                    f[k] = s_vf[j][k];
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
                };
                if (ls == 0)
                    tile_pairs(std::false_type{});
                else
                    tile_pairs(std::true_type{});
            }
        }

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
            const double osc = ((PREC == 0) ? 1.0 : ldexp(1.0, K::deg(k) * item.level)) * K::post(k);
#pragma unroll
            for (int t = 0; t < T; ++t) {
                const int i = slot0 + tstride * t;
                if (sub == 0 && i < item.trg_count)
                    dst[(long long)k * TT + i] = acc[t][k] * osc;
            }
        }
    }
}

// out(perm[i], k) += scale * sum over the tile's slots of partial[slot][k][i]; one workgroup per
// tile, blockDim = 64*T.
__global__ void __launch_bounds__(256) p2p_reduce_kernel(ReduceArgs a) {
    const Tile tl = a.tiles[blockIdx.x];
    const int i = threadIdx.x;
    if (i >= tl.trg_count)
        return;
    const long long row = tl.trg_begin + i;
    const long long o = a.perm ? (long long)a.perm[row] : row;
    for (int k = 0; k < a.kt; ++k) {
        double s = 0.0;
        for (int slot = tl.slot_begin; slot < tl.slot_end; ++slot)
            s += a.partial[((long long)slot * a.kt + k) * a.tt + i];
        a.out.p[o * a.out.s + k * a.out.fs] += a.scale * s;
    }
}

__global__ void fill_int_kernel(int *p, int value) { *p = value; }

__global__ void fill_double_kernel(double *p, long long n, double value) {
    const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n)
        p[i] = value;
}

// zero-fill without hipMemsetAsync, so that the enqueued work stays valid inside a captured graph
static void zero_fill(double *p, long long n, hipStream_t s) {
    if (p && n > 0)
        hipLaunchKernelGGL(fill_double_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, p, n, 0.0);
}

// ---- dispatch ---------------------------------------------------------------------------------

template <int UK, int PREC, int T>
static hipError_t launch_one(int grid, const ItemArgs &a, hipStream_t s) {
    hipLaunchKernelGGL((p2p_items_kernel<UK, PREC, T>), dim3(grid), dim3(kWave), 0, s, a);
    return hipGetLastError();
}

template <int UK, int PREC, int T>
static hipError_t occ_one(int *blocks) {
    return hipOccupancyMaxActiveBlocksPerMultiprocessor(blocks, p2p_items_kernel<UK, PREC, T>, kWave, 0);
}

template <int UK, int PREC>
static hipError_t launch_t(int T, int grid, const ItemArgs &a, hipStream_t s) {
    switch (T) {
    case 1: return launch_one<UK, PREC, 1>(grid, a, s);
    case 2: return launch_one<UK, PREC, 2>(grid, a, s);
    case 4: return launch_one<UK, PREC, 4>(grid, a, s);
    default: return hipErrorInvalidValue;
    }
}

template <int UK, int PREC>
static hipError_t occ_t(int T, int *blocks) {
    switch (T) {
    case 1: return occ_one<UK, PREC, 1>(blocks);
    case 2: return occ_one<UK, PREC, 2>(blocks);
    case 4: return occ_one<UK, PREC, 4>(blocks);
    default: return hipErrorInvalidValue;
    }
}

template <int UK>
static hipError_t launch_p(int prec, int T, int grid, const ItemArgs &a, hipStream_t s) {
    return prec == 0 ? launch_t<UK, 0>(T, grid, a, s) : launch_t<UK, 1>(T, grid, a, s);
}

template <int UK>
static hipError_t occ_p(int prec, int T, int *blocks) {
    return prec == 0 ? occ_t<UK, 0>(T, blocks) : occ_t<UK, 1>(T, blocks);
}

#define STKP2P_UK_SWITCH(CALL)                                                                     \
    switch (uk) {                                                                                  \
    case 0: return CALL(0);                                                                        \
    case 1: return CALL(1);                                                                        \
    case 2: return CALL(2);                                                                        \
    case 3: return CALL(3);                                                                        \
    case 4: return CALL(4);                                                                        \
    case 5: return CALL(5);                                                                        \
    case 6: return CALL(6);                                                                        \
    case 7: return CALL(7);                                                                        \
    case 8: return CALL(8);                                                                        \
    case 9: return CALL(9);                                                                        \
    case 10: return CALL(10);                                                                      \
    case 11: return CALL(11);                                                                      \
    case 12: return CALL(12);                                                                      \
    case 13: return CALL(13);                                                                      \
    default: return hipErrorInvalidValue;                                                          \
    }

hipError_t launch_items(int uk, int prec, int T, int grid, const ItemArgs &a, hipStream_t s) {
    if (grid <= 0)
        return hipSuccess;
#define CALL(U) launch_p<U>(prec, T, grid, a, s)
    STKP2P_UK_SWITCH(CALL)
#undef CALL
}

hipError_t items_occupancy(int uk, int prec, int T, int *blocks_per_cu) {
#define CALL(U) occ_p<U>(prec, T, blocks_per_cu)
    STKP2P_UK_SWITCH(CALL)
#undef CALL
}

// ---- scalar-broadcast variant dispatch ----------------------------------------------------------

template <int UK, int PREC, int T>
static hipError_t launch_sgpr_one(int grid, const ItemArgs &a, const double *recs, int level_min, hipStream_t s) {
    hipLaunchKernelGGL((p2p_items_sgpr_kernel<UK, PREC, T>), dim3(grid), dim3(kWave), 0, s, a, recs, level_min);
    return hipGetLastError();
}

template <int UK, int PREC, int T>
static hipError_t occ_sgpr_one(int *blocks) {
    return hipOccupancyMaxActiveBlocksPerMultiprocessor(blocks, p2p_items_sgpr_kernel<UK, PREC, T>, kWave, 0);
}

template <int UK, int PREC>
static hipError_t launch_sgpr_t(int T, int grid, const ItemArgs &a, const double *recs, int lmin, hipStream_t s) {
    switch (T) {
    case 1: return launch_sgpr_one<UK, PREC, 1>(grid, a, recs, lmin, s);
    case 2: return launch_sgpr_one<UK, PREC, 2>(grid, a, recs, lmin, s);
    case 4: return launch_sgpr_one<UK, PREC, 4>(grid, a, recs, lmin, s);
    default: return hipErrorInvalidValue;
    }
}

template <int UK, int PREC>
static hipError_t occ_sgpr_t(int T, int *blocks) {
    switch (T) {
    case 1: return occ_sgpr_one<UK, PREC, 1>(blocks);
    case 2: return occ_sgpr_one<UK, PREC, 2>(blocks);
    case 4: return occ_sgpr_one<UK, PREC, 4>(blocks);
    default: return hipErrorInvalidValue;
    }
}

template <int UK>
static hipError_t stage_p(int prec, const StageArgs &a, int n_copies, hipStream_t s) {
    const dim3 grid((unsigned)((a.n_src + 255) / 256), (unsigned)n_copies);
    if (prec == 0)
        hipLaunchKernelGGL((stage_records_kernel<UK, 0>), grid, dim3(256), 0, s, a);
    else
        hipLaunchKernelGGL((stage_records_kernel<UK, 1>), grid, dim3(256), 0, s, a);
    return hipGetLastError();
}

template <int UK>
static int words_p(int prec) {
    using K = typename UkerOf<UK>::type;
    return prec == 0 ? RecLayout<K, 0>::WORDS : RecLayout<K, 1>::WORDS;
}

hipError_t launch_items_sgpr(int uk, int prec, int T, int grid, const ItemArgs &a, const double *recs,
                             int level_min, hipStream_t s) {
    if (grid <= 0)
        return hipSuccess;
#define CALL(U) (prec == 0 ? launch_sgpr_t<U, 0>(T, grid, a, recs, level_min, s) : launch_sgpr_t<U, 1>(T, grid, a, recs, level_min, s))
    STKP2P_UK_SWITCH(CALL)
#undef CALL
}

hipError_t items_sgpr_occupancy(int uk, int prec, int T, int *blocks_per_cu) {
#define CALL(U) (prec == 0 ? occ_sgpr_t<U, 0>(T, blocks_per_cu) : occ_sgpr_t<U, 1>(T, blocks_per_cu))
    STKP2P_UK_SWITCH(CALL)
#undef CALL
}

hipError_t launch_stage_records(int uk, int prec, const StageArgs &a, int n_copies, hipStream_t s) {
    if (a.n_src <= 0 || n_copies <= 0)
        return hipSuccess;
#define CALL(U) stage_p<U>(prec, a, n_copies, s)
    STKP2P_UK_SWITCH(CALL)
#undef CALL
}

int record_words(int uk, int prec) {
#define CALL(U) words_p<U>(prec)
    switch (uk) {
    case 0: return CALL(0);
    case 1: return CALL(1);
    case 2: return CALL(2);
    case 3: return CALL(3);
    case 4: return CALL(4);
    case 5: return CALL(5);
    case 6: return CALL(6);
    case 7: return CALL(7);
    case 8: return CALL(8);
    case 9: return CALL(9);
    case 10: return CALL(10);
    case 11: return CALL(11);
    case 12: return CALL(12);
    case 13: return CALL(13);
    default: return -1;
    }
#undef CALL
}

hipError_t launch_reduce(const ReduceArgs &a, hipStream_t s) {
    if (a.n_tiles <= 0)
        return hipSuccess;
    hipLaunchKernelGGL(p2p_reduce_kernel, dim3(a.n_tiles), dim3(a.tt), 0, s, a);
    return hipGetLastError();
}

hipError_t launch_fill_int(int *p, int value, hipStream_t s) {
    hipLaunchKernelGGL(fill_int_kernel, dim3(1), dim3(1), 0, s, p, value);
    return hipGetLastError();
}

// ---- kernel table -----------------------------------------------------------------------------

template <int UK> struct Dims {
    using K = typename UkerOf<UK>::type;
    static constexpr int ks = K::KS, ks1 = K::KS1, kt = K::KT;
};

#define STKP2P_DIMS_SWITCH(FIELD)                                                                  \
    switch (uk) {                                                                                  \
    case 0: return Dims<0>::FIELD;                                                                 \
    case 1: return Dims<1>::FIELD;                                                                 \
    case 2: return Dims<2>::FIELD;                                                                 \
    case 3: return Dims<3>::FIELD;                                                                 \
    case 4: return Dims<4>::FIELD;                                                                 \
    case 5: return Dims<5>::FIELD;                                                                 \
    case 6: return Dims<6>::FIELD;                                                                 \
    case 7: return Dims<7>::FIELD;                                                                 \
    case 8: return Dims<8>::FIELD;                                                                 \
    case 9: return Dims<9>::FIELD;                                                                 \
    case 10: return Dims<10>::FIELD;                                                               \
    case 11: return Dims<11>::FIELD;                                                               \
    case 12: return Dims<12>::FIELD;                                                               \
    case 13: return Dims<13>::FIELD;                                                               \
    default: return -1;                                                                            \
    }

int uker_ks(int uk) { STKP2P_DIMS_SWITCH(ks) }
int uker_ks1(int uk) { STKP2P_DIMS_SWITCH(ks1) }
int uker_kt(int uk) { STKP2P_DIMS_SWITCH(kt) }

// ScaleFactor<Real>(): 1/(4 pi) Laplace family (LaplaceLayerKernel.hpp:44-46), 1/(8 pi) Stokes
// family (StokesSingleLayerKernel.hpp:30-32, RPYKernel.hpp:20-22, StokesRegSingleLayerKernel.hpp:20-22)
double uker_scale(int uk) {
    const double pi = 3.141592653589793238462643383279;
    return (uk >= 0 && uk <= 5) ? 1.0 / (4.0 * pi) : 1.0 / (8.0 * pi);
}

// ---- adjoint dispatch -------------------------------------------------------------------------

template <int UK, int PREC>
static hipError_t launch_vjp_t(const VjpArgs &a, double *partial, int nchunk_t, int nchunk_s, hipStream_t s) {
    using V = typename VjpOf<UK>::type;
    constexpr int KS = V::KS;
    VjpPass p{a.sc, a.sv, a.tc, a.w, a.n_src, a.n_trg, 0, partial};
    if (a.g_tc && a.n_trg > 0 && a.n_src > 0) {
        p.chunk = (a.n_src + nchunk_t - 1) / nchunk_t;
        const int nc = (a.n_src + p.chunk - 1) / p.chunk;
        hipLaunchKernelGGL((vjp_target_kernel<V, PREC>), dim3((a.n_trg + kWave - 1) / kWave, nc), dim3(kWave), 0,
                           s, p);
        hipError_t e = hipGetLastError();
        if (e != hipSuccess)
            return e;
        hipLaunchKernelGGL(vjp_reduce_kernel, dim3((a.n_trg + 255) / 256), dim3(256), 0, s, (const double *)partial,
                           a.n_trg, 3, nc, 0, 3, a.g_tc, 3, 0, a.scale);
        e = hipGetLastError();
        if (e != hipSuccess)
            return e;
    } else if (a.g_tc && a.n_trg > 0) {
        zero_fill(a.g_tc, 3LL * a.n_trg, s);
    }
    if ((a.g_sc || a.g_sv) && a.n_src > 0 && a.n_trg > 0) {
        p.chunk = (a.n_trg + nchunk_s - 1) / nchunk_s;
        const int nc = (a.n_trg + p.chunk - 1) / p.chunk;
        hipLaunchKernelGGL((vjp_source_kernel<V, PREC>), dim3((a.n_src + kWave - 1) / kWave, nc), dim3(kWave), 0,
                           s, p);
        hipError_t e = hipGetLastError();
        if (e != hipSuccess)
            return e;
        const dim3 g((a.n_src + 255) / 256);
        hipLaunchKernelGGL(vjp_reduce_kernel, g, dim3(256), 0, s, (const double *)partial, a.n_src, 3 + KS, nc, 0, 3,
                           a.g_sc, 3, 0, a.scale);
        hipLaunchKernelGGL(vjp_reduce_kernel, g, dim3(256), 0, s, (const double *)partial, a.n_src, 3 + KS, nc, 3,
                           3 + KS, a.g_sv, KS, 0, a.scale);
        e = hipGetLastError();
        if (e != hipSuccess)
            return e;
    } else if (a.n_src > 0) {
        zero_fill(a.g_sc, 3LL * a.n_src, s);
        zero_fill(a.g_sv, (long long)KS * a.n_src, s);
    }
    return hipGetLastError();
}

bool uker_has_vjp(int uk) { return uk == 7 || uk == 8 || uk == 11 || uk == 13; }

hipError_t launch_vjp(int uk, int prec, const VjpArgs &a, double *partial, int nchunk_t, int nchunk_s,
                      hipStream_t s) {
    if (nchunk_t < 1 || nchunk_s < 1)
        return hipErrorInvalidValue;
    switch (uk) {
    case 7: return prec ? launch_vjp_t<7, 1>(a, partial, nchunk_t, nchunk_s, s) : launch_vjp_t<7, 0>(a, partial, nchunk_t, nchunk_s, s);
    case 8: return prec ? launch_vjp_t<8, 1>(a, partial, nchunk_t, nchunk_s, s) : launch_vjp_t<8, 0>(a, partial, nchunk_t, nchunk_s, s);
    case 11: return prec ? launch_vjp_t<11, 1>(a, partial, nchunk_t, nchunk_s, s) : launch_vjp_t<11, 0>(a, partial, nchunk_t, nchunk_s, s);
    case 13: return prec ? launch_vjp_t<13, 1>(a, partial, nchunk_t, nchunk_s, s) : launch_vjp_t<13, 0>(a, partial, nchunk_t, nchunk_s, s);
    default: return hipErrorInvalidValue;
    }
}

} // namespace stkp2p
