// p2p_kernels.hip -- gfx950 kernels of the STKFMM P2P near-field path (core micro-kernels).
//
//   p2p_items_kernel (p2p_items.hpp) instantiated for micro-kernels 0..13: Laplace P/PGrad/
//       PGradGrad single and double layer, Stokeslet, PVel single/double/fused, regularised
//       Stokeslet, RPY; the rest of the family (14..24) is instantiated in p2p_family.hip.
//   p2p_reduce_kernel  per target: sum of its tile's partial slots in slot order (deterministic),
//       times the scale factor, added (+=) into the caller's output (optionally permuted); four
//       output components per thread for kernels with more than four.
//   vjp kernels        adjoint of a direct sum: one pass over the pairs gives the target-position,
//                      source-position and source-value gradients (vjp_kernels.hpp).
//   The launch entries below dispatch through the per-micro-kernel UkerOps table.
#include <hip/hip_runtime.h>

#include <type_traits>

#include "p2p_internal.hpp"
#include "p2p_items.hpp"
#include "ukernels.hpp"
#include "vjp_kernels.hpp"

namespace stkp2p {


// out(perm[i], k) += scale * sum over the tile's slots of partial[slot][k][i]; one workgroup per
// (tile, group of KG output components k = KG*blockIdx.y + c), blockDim = 64*T.  Each (target, k) sums
// its slots in slot order, so the result does not depend on the grid shape.  KG = 4 for kernels with
// more than 4 outputs over at least 2048 tiles: the components of a group are summed side by side, so
// that their loads are in flight together (one workgroup per component left the 16-output kernels
// with ~10^5 one-wave workgroups, latency-bound: reduce 0.178 -> 0.141 ms on a 262,144-point
// U-list); KG = 1 otherwise, where grouping left too few workgroups (the 4-output PVel U-list reduce
// 0.106 -> 0.119 ms; 65,536-target direct sums of 512 tiles and ~64 slots each slower).
template <int KG>
__global__ void __launch_bounds__(256) p2p_reduce_kernel(ReduceArgs a) {
    const Tile tl = a.tiles[blockIdx.x];
    const int i = threadIdx.x;
    if (i >= tl.trg_count)
        return;
    const int k0 = blockIdx.y * KG;
    const int nk = min(KG, a.kt - k0);
    const long long row = tl.trg_begin + i;
    const long long o = a.perm ? (long long)a.perm[row] : row;
    const long long step = (long long)a.kt * a.tt;
    const double *__restrict__ part = a.partial + (long long)k0 * a.tt + i;
    double sum[KG];
#pragma unroll
    for (int c = 0; c < KG; ++c)
        sum[c] = 0.0;
    // unrolled so that the loads of several slots are in flight together (the adds keep slot order)
    constexpr int kSlotUnroll = KG == 1 ? 4 : 2;
#pragma unroll kSlotUnroll
    for (int slot = tl.slot_begin; slot < tl.slot_end; ++slot) {
        const double *__restrict__ q = part + slot * step;
#pragma unroll
        for (int c = 0; c < KG; ++c)
            if (c < nk)
                sum[c] += q[(long long)c * a.tt];
    }
#pragma unroll
    for (int c = 0; c < KG; ++c)
        if (c < nk)
            a.out.p[o * a.out.s + (k0 + c) * a.out.fs] += a.scale * sum[c];
}

// out(perm[i], k) += scale * (slot sums) for every k of target i in one thread: the same sums in the
// same order as p2p_reduce_kernel (bitwise the same result), one workgroup per tile, tiles dealt to
// the XCDs in contiguous runs (the scattered rows of neighbouring tiles then share one L2)
__global__ void __launch_bounds__(256) p2p_reduce_rows_kernel(ReduceArgs a) {
    const int ti = xcd_unit(blockIdx.x, a.n_tiles);
    if (ti < 0)
        return;
    const Tile tl = a.tiles[ti];
    const int i = threadIdx.x;
    if (i >= tl.trg_count)
        return;
    const long long row = tl.trg_begin + i;
    const long long o = a.perm ? (long long)a.perm[row] : row;
    for (int k = 0; k < a.kt; ++k) {
        double s = 0.0;
#pragma unroll 4
        for (int slot = tl.slot_begin; slot < tl.slot_end; ++slot)
            s += a.partial[((long long)slot * a.kt + k) * a.tt + i];
        a.out.p[o * a.out.s + k * a.out.fs] += a.scale * s;
    }
}

__global__ void fill_int_kernel(int *p, int n, int value) {
    if ((int)threadIdx.x < n)
        p[threadIdx.x] = value;
}

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

// ---- dispatch through the micro-kernel table ------------------------------------------------------

namespace {
struct CoreTable {
    UkerOps row[kNumCoreUker];
    CoreTable() {
        row[0] = make_uker_ops<0>();
        row[1] = make_uker_ops<1>();
        row[2] = make_uker_ops<2>();
        row[3] = make_uker_ops<3>();
        row[4] = make_uker_ops<4>();
        row[5] = make_uker_ops<5>();
        row[6] = make_uker_ops<6>();
        row[7] = make_uker_ops<7>();
        row[8] = make_uker_ops<8>();
        row[9] = make_uker_ops<9>();
        row[10] = make_uker_ops<10>();
        row[11] = make_uker_ops<11>();
        row[12] = make_uker_ops<12>();
        row[13] = make_uker_ops<13>();
    }
};
} // namespace

const UkerOps *uker_ops(int uk) {
    static const CoreTable core;
    if (uk >= 0 && uk < kNumCoreUker)
        return &core.row[uk];
    return uker_ops_family(uk);
}

hipError_t launch_items(int uk, int prec, int T, int grid, const ItemArgs &a, hipStream_t s) {
    if (grid <= 0)
        return hipSuccess;
    const UkerOps *o = uker_ops(uk);
    return o ? o->launch(prec, T, grid, a, s) : hipErrorInvalidValue;
}

hipError_t items_occupancy(int uk, int prec, int T, int *blocks_per_cu) {
    const UkerOps *o = uker_ops(uk);
    return o ? o->occ(prec, T, blocks_per_cu) : hipErrorInvalidValue;
}

hipError_t launch_items_sgpr(int uk, int prec, int T, int grid, const ItemArgs &a, const double *recs,
                             int level_min, hipStream_t s) {
    if (grid <= 0)
        return hipSuccess;
    const UkerOps *o = uker_ops(uk);
    return o ? o->launch_sgpr(prec, T, grid, a, recs, level_min, s) : hipErrorInvalidValue;
}

hipError_t items_sgpr_occupancy(int uk, int prec, int T, int *blocks_per_cu) {
    const UkerOps *o = uker_ops(uk);
    return o ? o->occ_sgpr(prec, T, blocks_per_cu) : hipErrorInvalidValue;
}

hipError_t launch_stage_records(int uk, int prec, const StageArgs &a, int n_copies, hipStream_t s) {
    if (a.n_src <= 0 || n_copies <= 0)
        return hipSuccess;
    const UkerOps *o = uker_ops(uk);
    return o ? o->stage(prec, a, n_copies, s) : hipErrorInvalidValue;
}

int record_words(int uk, int prec) {
    const UkerOps *o = uker_ops(uk);
    return (o && (prec == 0 || prec == 1)) ? o->words[prec] : -1;
}

hipError_t launch_reduce(const ReduceArgs &a, hipStream_t s) {
    if (a.n_tiles <= 0)
        return hipSuccess;
    if (a.kt > 4 && a.n_tiles >= 2048)
        hipLaunchKernelGGL(p2p_reduce_kernel<4>, dim3(a.n_tiles, (a.kt + 3) / 4), dim3(a.tt), 0, s, a);
    else
        hipLaunchKernelGGL(p2p_reduce_kernel<1>, dim3(a.n_tiles, a.kt), dim3(a.tt), 0, s, a);
    return hipGetLastError();
}

hipError_t launch_reduce_rows(const ReduceArgs &a, hipStream_t s) {
    if (a.n_tiles <= 0)
        return hipSuccess;
    hipLaunchKernelGGL(p2p_reduce_rows_kernel, dim3((unsigned)xcd_grid(a.n_tiles)), dim3(a.tt), 0, s, a);
    return hipGetLastError();
}

hipError_t launch_fill_int(int *p, int n, int value, hipStream_t s) {
    if (n < 1 || n > 256)
        return hipErrorInvalidValue;
    hipLaunchKernelGGL(fill_int_kernel, dim3(1), dim3(256), 0, s, p, n, value);
    return hipGetLastError();
}

// ---- kernel table -----------------------------------------------------------------------------

int uker_ks(int uk) {
    const UkerOps *o = uker_ops(uk);
    return o ? o->ks : -1;
}
int uker_ks1(int uk) {
    const UkerOps *o = uker_ops(uk);
    return o ? o->ks1 : -1;
}
int uker_kt(int uk) {
    const UkerOps *o = uker_ops(uk);
    return o ? o->kt : -1;
}

// ScaleFactor<Real>(): 1/(4 pi) for the Laplace family (LaplaceLayerKernel.hpp:45,206,238),
// -3/(4 pi) for stokes_traction (StokesSingleLayerKernel.hpp:166), -3/(8 pi) for
// stokes_doubletraction (StokesDoubleLayerKernel.hpp:217), 1/(8 pi) for the rest of the Stokes
// family (StokesSingleLayerKernel.hpp:30-32, RPYKernel.hpp:20-22, StokesRegSingleLayerKernel.hpp:20-22)
double uker_scale(int uk) {
    const double pi = 3.141592653589793238462643383279;
    if ((uk >= 0 && uk <= 5) || uk == 22 || uk == 23)
        return 1.0 / (4.0 * pi);
    if (uk == 18)
        return -3.0 / (4.0 * pi);
    if (uk == 19)
        return -3.0 / (8.0 * pi);
    return 1.0 / (8.0 * pi);
}

// ---- adjoint dispatch -------------------------------------------------------------------------

template <int UK, int PREC, int T>
static hipError_t launch_vjp_t(const VjpArgs &a, const VjpLayout &l, double *partial, hipStream_t s) {
    using V = typename VjpOf<UK>::type;
    constexpr int KS = V::KS, ND = 3 + KS;
    if (a.n_trg == 0 || a.n_src == 0) {
        if (a.g_tc)
            zero_fill(a.g_tc, 3LL * a.n_trg, s);
        if (a.g_sc)
            zero_fill(a.g_sc, 3LL * a.n_src, s);
        if (a.g_sv)
            zero_fill(a.g_sv, (long long)KS * a.n_src, s);
        return hipGetLastError();
    }
    double *tpart = partial, *spart = partial + (long long)l.n_chunks * 3 * a.n_trg;
    VjpPass p{a.sc, a.sv, a.tc, a.w, a.n_src, a.n_trg, l.trg_tiles_per_block, l.src_tiles_per_chunk, tpart, spart};
    hipLaunchKernelGGL((vjp_pair_kernel<V, PREC, T>), dim3(l.n_blocks, l.n_chunks), dim3(kWave), 0, s, p);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess)
        return e;
    const unsigned bt = (unsigned)((a.n_trg + 255) / 256), bs = (unsigned)((a.n_src + 255) / 256);
    if (a.g_tc)
        hipLaunchKernelGGL(vjp_reduce_kernel, dim3(bt, 3), dim3(256), 0, s, (const double *)tpart, a.n_trg, 3,
                           l.n_chunks, 0, 3, a.g_tc, 3, 0, a.scale);
    if (a.g_sc)
        hipLaunchKernelGGL(vjp_reduce_kernel, dim3(bs, 3), dim3(256), 0, s, (const double *)spart, a.n_src, ND,
                           l.n_blocks, 0, 3, a.g_sc, 3, 0, a.scale);
    if (a.g_sv)
        hipLaunchKernelGGL(vjp_reduce_kernel, dim3(bs, ND - 3), dim3(256), 0, s, (const double *)spart, a.n_src, ND,
                           l.n_blocks, 3, ND, a.g_sv, KS, 0, a.scale);
    return hipGetLastError();
}

template <int UK, int PREC>
static hipError_t launch_vjp_p(const VjpArgs &a, const VjpLayout &l, double *partial, hipStream_t s) {
    if constexpr (PREC != 0 || vjp_fp64_four_targets(UK)) // four targets per lane
        if (l.T == 4)
            return launch_vjp_t<UK, PREC, 4>(a, l, partial, s);
    if constexpr (PREC != 0) // eight: mixed mode only
        if (l.T == 8)
            return launch_vjp_t<UK, PREC, 8>(a, l, partial, s);
    return (l.T == 1) ? launch_vjp_t<UK, PREC, 1>(a, l, partial, s) : launch_vjp_t<UK, PREC, 2>(a, l, partial, s);
}

bool uker_has_vjp(int uk) {
    return uk == 0 || uk == 2 || uk == 6 || uk == 7 || uk == 8 || uk == 9 || uk == 10 || uk == 11 || uk == 12 ||
           uk == 13;
}

hipError_t launch_vjp(int uk, int prec, const VjpArgs &a, const VjpLayout &l, double *partial, hipStream_t s) {
    if (l.T != 1 && l.T != 2 && !(l.T == 4 && (prec != 0 || vjp_fp64_four_targets(uk))) &&
        !(l.T == 8 && prec != 0))
        return hipErrorInvalidValue;
    switch (uk) {
    case 0: return prec ? launch_vjp_p<0, 1>(a, l, partial, s) : launch_vjp_p<0, 0>(a, l, partial, s);
    case 2: return prec ? launch_vjp_p<2, 1>(a, l, partial, s) : launch_vjp_p<2, 0>(a, l, partial, s);
    case 6: return prec ? launch_vjp_p<6, 1>(a, l, partial, s) : launch_vjp_p<6, 0>(a, l, partial, s);
    case 9: return prec ? launch_vjp_p<9, 1>(a, l, partial, s) : launch_vjp_p<9, 0>(a, l, partial, s);
    case 10: return prec ? launch_vjp_p<10, 1>(a, l, partial, s) : launch_vjp_p<10, 0>(a, l, partial, s);
    case 12: return prec ? launch_vjp_p<12, 1>(a, l, partial, s) : launch_vjp_p<12, 0>(a, l, partial, s);
    case 7: return prec ? launch_vjp_p<7, 1>(a, l, partial, s) : launch_vjp_p<7, 0>(a, l, partial, s);
    case 8: return prec ? launch_vjp_p<8, 1>(a, l, partial, s) : launch_vjp_p<8, 0>(a, l, partial, s);
    case 11: return prec ? launch_vjp_p<11, 1>(a, l, partial, s) : launch_vjp_p<11, 0>(a, l, partial, s);
    case 13: return prec ? launch_vjp_p<13, 1>(a, l, partial, s) : launch_vjp_p<13, 0>(a, l, partial, s);
    default: return hipErrorInvalidValue;
    }
}

} // namespace stkp2p
