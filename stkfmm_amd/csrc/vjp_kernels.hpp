// vjp_kernels.hpp -- adjoint (vector-Jacobian product) kernels of the direct-sum P2P.
//
// For L = sum_t w_t . u_t with u_t = scale * sum_s K(x_t - x_s, f_s), and a pair with displacement
// r, source value f and cotangent w:
//   g_r = sum_k w_k dK_k/dr,   g_f = sum_k w_k dK_k/df
// so that dL/dx_t = scale * sum_s g_r,  dL/dx_s = -scale * sum_t g_r,  dL/df_s = scale * sum_t g_f.
// The per-pair formulas are differentiated by hand from the forward micro-kernels (ukernels.hpp).
// Coincident pairs contribute zero (the forward is masked there).  The reference has no adjoint.
//
// One pass evaluates each pair once (vjp_pair_kernel): a wave holds T targets per lane and walks a
// chunk of source tiles staged in LDS.  Within a tile, at step k lane i pairs its targets with
// source slot (i + k) mod 64 (a conflict-free lane-indexed LDS read); the target gradient stays in
// the lane, and the source-side sums (position gradient, value gradient) travel with their source:
// after every step they move one lane down the wave (a DPP rotation), so after 64 steps each has
// met every target of the tile and is back in the lane of its slot.  The grid is (target blocks) x
// (source chunks); every (source chunk, target) and (target block, source) partial is written by
// one wave in a fixed order and the partials are summed in a fixed order -- deterministic.
#pragma once

#include <hip/hip_runtime.h>

#include <type_traits>

#include "p2p_internal.hpp"
#include "vjp_pairs.hpp"

namespace stkp2p {

// lane i receives the value of lane i + 1 (lane 63 that of lane 0): DPP wave_rol:1.  Called with all
// 64 lanes active, so every lane reads an active lane and the "old" operand (the value a lane keeps
// when its source lane is inactive) is never used; passing v itself lets the move rotate in place
// instead of first zeroing a fresh register.
__device__ __forceinline__ float rotate_down(float v) {
    const int b = __builtin_bit_cast(int, v);
    return __builtin_bit_cast(float, __builtin_amdgcn_update_dpp(b, b, 0x134, 0xF, 0xF, false));
}
__device__ __forceinline__ double rotate_down(double v) {
    const unsigned long long b = __builtin_bit_cast(unsigned long long, v);
    const int lo0 = (int)(unsigned)b, hi0 = (int)(unsigned)(b >> 32);
    const int lo = __builtin_amdgcn_update_dpp(lo0, lo0, 0x134, 0xF, 0xF, false);
    const int hi = __builtin_amdgcn_update_dpp(hi0, hi0, 0x134, 0xF, 0xF, false);
    return __builtin_bit_cast(double, ((unsigned long long)(unsigned)hi << 32) | (unsigned)lo);
}

// tpart: [n_chunks][3][n_trg] target position gradient over the sources of one chunk
// spart: [n_blocks][3 + KS][n_src] source position and value gradients over the targets of one block
struct VjpPass {
    const double *sc, *sv, *tc, *w;  // AoS
    int n_src, n_trg;
    int trg_tiles_per_block;         // target tiles (64 T targets) per block
    int src_tiles_per_chunk;         // source tiles (64 sources) per chunk
    double *tpart;
    double *spart;
};

// eight targets per lane: at least 3 waves per SIMD asked of the compiler (168 VGPRs; left to itself
// it schedules the 8-target step into ~240, 2 waves per SIMD)
template <class V, int PREC, int T>
__global__ void __launch_bounds__(kWave) __attribute__((amdgpu_waves_per_eu((T >= 8 || (PREC == 0 && T >= 4)) ? 3 : 1)))
vjp_pair_kernel(VjpPass a) {
    using CT = typename std::conditional<PREC == 0, double, float>::type;
    constexpr int KS = V::KS, KT = V::KT, ND = 3 + KS, TT = kWave * T;
    __shared__ double s_x[kSrcTile], s_y[kSrcTile], s_z[kSrcTile];
    __shared__ CT s_v[KS][kSrcTile];
    const int lane = threadIdx.x;
    const int n_trg_tiles = (a.n_trg + TT - 1) / TT;
    const int n_src_tiles = (a.n_src + kSrcTile - 1) / kSrcTile;
    const int tt0 = blockIdx.x * a.trg_tiles_per_block;
    const int tt1 = min(tt0 + a.trg_tiles_per_block, n_trg_tiles);
    const int st0 = blockIdx.y * a.src_tiles_per_chunk;
    const int st1 = min(st0 + a.src_tiles_per_chunk, n_src_tiles);
    double *tpart = a.tpart + (long long)blockIdx.y * 3 * a.n_trg;
    double *spart = a.spart + (long long)blockIdx.x * ND * a.n_src;

    for (int st = st0; st < st1; ++st) {
        const int sbase = st * kSrcTile;
        const int ns = min(kSrcTile, a.n_src - sbase);
        __syncthreads(); // every lane has read the previous tile
        {
            // slot lane: its source, or (past the end) a copy of the last one with zero values, which
            // contributes nothing to the targets
            const long long s = sbase + min(lane, ns - 1);
            s_x[lane] = a.sc[3 * s];
            s_y[lane] = a.sc[3 * s + 1];
            s_z[lane] = a.sc[3 * s + 2];
#pragma unroll
            for (int k = 0; k < KS; ++k)
                s_v[k][lane] = (lane < ns) ? (CT)a.sv[KS * s + k] : CT(0);
        }
        __syncthreads();
        double sacc[ND];
#pragma unroll
        for (int d = 0; d < ND; ++d)
            sacc[d] = 0.0;

        for (int tt = tt0; tt < tt1; ++tt) {
            // targets tt*TT + lane + 64 t; past the end: a copy of the last one with a zero cotangent
            double tx[T], ty[T], tz[T];
            CT w[T][KT];
#pragma unroll
            for (int t = 0; t < T; ++t) {
                const int i = tt * TT + lane + kWave * t;
                const long long ic = (i < a.n_trg) ? i : a.n_trg - 1;
                tx[t] = a.tc[3 * ic];
                ty[t] = a.tc[3 * ic + 1];
                tz[t] = a.tc[3 * ic + 2];
#pragma unroll
                for (int k = 0; k < KT; ++k)
                    w[t][k] = (i < a.n_trg) ? (CT)a.w[KT * ic + k] : CT(0);
            }
            CT tg[T][3];
            CT rs[ND]; // source-side sums of the source currently in this lane
#pragma unroll
            for (int t = 0; t < T; ++t)
                tg[t][0] = tg[t][1] = tg[t][2] = CT(0);
#pragma unroll
            for (int d = 0; d < ND; ++d)
                rs[d] = CT(0);
            for (int k = 0; k < kSrcTile; ++k) {
                const int j = (lane + k) & (kSrcTile - 1);
                const double sx = s_x[j], sy = s_y[j], sz = s_z[j];
                CT f[KS];
#pragma unroll
                for (int q = 0; q < KS; ++q)
                    f[q] = s_v[q][j];
                // mixed mode: this step's source-value terms over the lane's T targets are summed
                // apart and added to the rotating FP32 sums once, so those take 64 additions per target
                // tile whatever T is.  Added straight in (2 FMAs per component per target), the value
                // gradients of the 65,536^2 fused PVel adjoint missed the 1e-6 bar at T = 4 (errorMaxRel
                // 1.3e-6; 2.6e-7 with the sums apart), PVel's at T = 8 (1.4e-6; 4.0e-7).  FP64 adds them
                // straight in.  -0 is the exact identity of addition (+0 is not: +0 + -0 = +0), so the
                // first "+= x" into each sum folds to "= x".
                constexpr int KG = (PREC != 0) ? KS : 1;
                CT gs[KG];
#pragma unroll
                for (int q = 0; q < KG; ++q)
                    gs[q] = CT(-0.0);
#pragma unroll
                for (int t = 0; t < T; ++t) {
                    CT gr[3] = {CT(-0.0), CT(-0.0), CT(-0.0)};
                    V::template vjp<CT>((CT)(tx[t] - sx), (CT)(ty[t] - sy), (CT)(tz[t] - sz), f, w[t], gr,
                                        (PREC != 0) ? gs : rs + 3);
#pragma unroll
                    for (int d = 0; d < 3; ++d) {
                        tg[t][d] += gr[d];
                        rs[d] -= gr[d];
                    }
                }
                if constexpr (PREC != 0)
#pragma unroll
                    for (int q = 0; q < KS; ++q)
                        rs[3 + q] += gs[q];
#pragma unroll
                for (int d = 0; d < ND; ++d)
                    rs[d] = rotate_down(rs[d]);
            }
            // 64 rotations: rs is back in the lane of its slot
#pragma unroll
            for (int d = 0; d < ND; ++d)
                sacc[d] += (double)rs[d];
#pragma unroll
            for (int t = 0; t < T; ++t) {
                const int i = tt * TT + lane + kWave * t;
                if (i < a.n_trg) {
#pragma unroll
                    for (int d = 0; d < 3; ++d) {
                        double *p = tpart + (long long)d * a.n_trg + i;
                        *p = ((st == st0) ? 0.0 : *p) + (double)tg[t][d];
                    }
                }
            }
        }
        if (lane < ns) {
#pragma unroll
            for (int d = 0; d < ND; ++d)
                spart[(long long)d * a.n_src + sbase + lane] = sacc[d];
        }
    }
}

// out[i*ld + off + d] = scale * sum_c partial[c][d][i] for d = d0 + blockIdx.y (grid.y = d1 - d0:
// one block row per component, so small point counts still give enough blocks); chunks are summed
// in chunk order, so the result does not depend on the grid shape
__global__ void vjp_reduce_kernel(const double *partial, int n, int nd, int nchunk, int d0, int d1,
                                  double *out, int ld, int off, double scale) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    const int d = d0 + (int)blockIdx.y;
    if (i >= n || out == nullptr || d >= d1)
        return;
    double s = 0.0;
    // unrolled so that the loads of several chunks are in flight together (the adds keep chunk order)
#pragma unroll 8
    for (int c = 0; c < nchunk; ++c)
        s += partial[((long long)c * nd + d) * n + i];
    out[(long long)i * ld + off + (d - d0)] = scale * s;
}

} // namespace stkp2p
