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
#include "ukernels.hpp"

namespace stkp2p {

// stokes_pvel: K0 = 2a/r^3, K_i = f_i/r + r_i (a - T)/r^3, a = f.r, T = TrD
struct VjpStokesPVel {
    static constexpr int KS = 4, KT = 4;
    template <class R>
    __device__ static void vjp(R dx, R dy, R dz, const R *f, const R *w, R *gr, R *gf) {
        const R r[3] = {dx, dy, dz};
        R r2 = dx * dx + dy * dy + dz * dz;
        R rinv = rsqrt_masked(r2);
        R rinv2 = rinv * rinv, rinv3 = rinv2 * rinv, rinv5 = rinv3 * rinv2;
        R a = f[0] * dx + f[1] * dy + f[2] * dz;
        R am = a - f[3];
        R wr = w[1] * dx + w[2] * dy + w[3] * dz;
        R wf = w[1] * f[0] + w[2] * f[1] + w[3] * f[2];
        R cf = (R(2) * w[0] + wr) * rinv3;
        R cr = wf * rinv3 + (R(6) * w[0] * a + R(3) * wr * am) * rinv5;
        R amr3 = am * rinv3;
#pragma unroll
        for (int i = 0; i < 3; ++i) {
            gr[i] += cf * f[i] + amr3 * w[1 + i] - cr * r[i];
            gf[i] += cf * r[i] + w[1 + i] * rinv;
        }
        gf[3] += -wr * rinv3;
    }
};

// stokes_doublepvel: Q = r.D.r, K0 = 2(-3Q/r^5 + tr/r^3), K_i = -3 r_i Q/r^5
struct VjpStokesDoublePVel {
    static constexpr int KS = 9, KT = 4;
    template <class R>
    __device__ static void vjp(R dx, R dy, R dz, const R *D, const R *w, R *gr, R *gf) {
        const R r[3] = {dx, dy, dz};
        R r2 = dx * dx + dy * dy + dz * dz;
        R rinv = rsqrt_masked(r2);
        R rinv2 = rinv * rinv, rinv3 = rinv2 * rinv, rinv5 = rinv3 * rinv2, rinv7 = rinv5 * rinv2;
        R Dr[3], DTr[3];
#pragma unroll
        for (int i = 0; i < 3; ++i) {
            Dr[i] = D[3 * i] * dx + D[3 * i + 1] * dy + D[3 * i + 2] * dz;
            DTr[i] = D[i] * dx + D[3 + i] * dy + D[6 + i] * dz;
        }
        R Q = dx * Dr[0] + dy * Dr[1] + dz * Dr[2];
        R tr = D[0] + D[4] + D[8];
        R wr = w[1] * dx + w[2] * dy + w[3] * dz;
        R c = R(6) * w[0] + R(3) * wr;
        R q5 = Q * rinv5;
        R cq7 = R(5) * c * Q * rinv7 - R(6) * w[0] * tr * rinv5;
        R crr5 = c * rinv5;
#pragma unroll
        for (int i = 0; i < 3; ++i) {
            gr[i] += -crr5 * (Dr[i] + DTr[i]) + cq7 * r[i] - R(3) * w[1 + i] * q5;
#pragma unroll
            for (int j = 0; j < 3; ++j)
                gf[3 * i + j] += -crr5 * r[i] * r[j] + ((i == j) ? R(2) * w[0] * rinv3 : R(0));
        }
    }
};

// rpy_ulapu: u = f/r + r a/r^3 + beta c, lap u = 2c, c = f/r^3 - 3 a r/r^5, beta = b^2/3
struct VjpRpyULapU {
    static constexpr int KS = 4, KT = 6;
    template <class R>
    __device__ static void vjp(R dx, R dy, R dz, const R *f, const R *w, R *gr, R *gf) {
        const R r[3] = {dx, dy, dz};
        R r2 = dx * dx + dy * dy + dz * dz;
        R rinv = rsqrt_masked(r2);
        R rinv2 = rinv * rinv, rinv3 = rinv2 * rinv, rinv5 = rinv3 * rinv2, rinv7 = rinv5 * rinv2;
        R a = f[0] * dx + f[1] * dy + f[2] * dz;
        const R third = R(STKP2P_RPY_ONE_OVER_THREE);
        R beta = third * f[3] * f[3];
        R v[3], c[3];
#pragma unroll
        for (int i = 0; i < 3; ++i) {
            v[i] = beta * w[i] + R(2) * w[3 + i];
            c[i] = f[i] * rinv3 - R(3) * a * r[i] * rinv5;
        }
        R wuf = w[0] * f[0] + w[1] * f[1] + w[2] * f[2];
        R wur = w[0] * dx + w[1] * dy + w[2] * dz;
        R vf = v[0] * f[0] + v[1] * f[1] + v[2] * f[2];
        R vr = v[0] * dx + v[1] * dy + v[2] * dz;
        R wuc = w[0] * c[0] + w[1] * c[1] + w[2] * c[2];
        R s_r = -wuf * rinv3 - R(3) * (wur * a + vf) * rinv5 + R(15) * a * vr * rinv7;
#pragma unroll
        for (int i = 0; i < 3; ++i) {
            gf[i] += w[i] * rinv + wur * r[i] * rinv3 + v[i] * rinv3 - R(3) * vr * r[i] * rinv5;
            gr[i] += s_r * r[i] + (w[i] * a + wur * f[i]) * rinv3 - R(3) * (f[i] * vr + a * v[i]) * rinv5;
        }
        gf[3] += R(2) * third * f[3] * wuc;
    }
};

// single + double layer at one point: f = [fx, fy, fz, TrD, D00..D22]
struct VjpStokesPVelFused {
    static constexpr int KS = 13, KT = 4;
    template <class R>
    __device__ static void vjp(R dx, R dy, R dz, const R *f, const R *w, R *gr, R *gf) {
        VjpStokesPVel::vjp(dx, dy, dz, f, w, gr, gf);
        VjpStokesDoublePVel::vjp(dx, dy, dz, f + 4, w, gr, gf + 4);
    }
};

// laplace_p: K = q/r
struct VjpLaplaceP {
    static constexpr int KS = 1, KT = 1;
    template <class R>
    __device__ static void vjp(R dx, R dy, R dz, const R *f, const R *w, R *gr, R *gf) {
        R r2 = dx * dx + dy * dy + dz * dz;
        R rinv = rsqrt_masked(r2);
        R c = -w[0] * f[0] * (rinv * rinv * rinv);
        gr[0] += c * dx;
        gr[1] += c * dy;
        gr[2] += c * dz;
        gf[0] += w[0] * rinv;
    }
};

// laplace_pgrad: K = [q/r, -q r/r^3]
struct VjpLaplacePGrad {
    static constexpr int KS = 1, KT = 4;
    template <class R>
    __device__ static void vjp(R dx, R dy, R dz, const R *f, const R *w, R *gr, R *gf) {
        const R r[3] = {dx, dy, dz};
        R r2 = dx * dx + dy * dy + dz * dz;
        R rinv = rsqrt_masked(r2);
        R rinv3 = rinv * (rinv * rinv), rinv5 = rinv3 * (rinv * rinv);
        R wg = w[1] * dx + w[2] * dy + w[3] * dz;
        R q = f[0];
        R cr = q * (R(3) * wg * rinv5 - w[0] * rinv3);
        R cw = -q * rinv3;
#pragma unroll
        for (int i = 0; i < 3; ++i)
            gr[i] += cr * r[i] + cw * w[1 + i];
        gf[0] += w[0] * rinv - wg * rinv3;
    }
};

// stokes_vel (Stokeslet): K = f/r + r (f.r)/r^3
struct VjpStokesVel {
    static constexpr int KS = 3, KT = 3;
    template <class R>
    __device__ static void vjp(R dx, R dy, R dz, const R *f, const R *w, R *gr, R *gf) {
        const R r[3] = {dx, dy, dz};
        R r2 = dx * dx + dy * dy + dz * dz;
        R rinv = rsqrt_masked(r2);
        R rinv3 = rinv * (rinv * rinv), rinv5 = rinv3 * (rinv * rinv);
        R a = f[0] * dx + f[1] * dy + f[2] * dz;
        R b = w[0] * dx + w[1] * dy + w[2] * dz;
        R wf = w[0] * f[0] + w[1] * f[1] + w[2] * f[2];
        R ar3 = a * rinv3, br3 = b * rinv3;
        R cr = wf * rinv3 + R(3) * a * b * rinv5;
#pragma unroll
        for (int i = 0; i < 3; ++i) {
            gr[i] += ar3 * w[i] + br3 * f[i] - cr * r[i];
            gf[i] += w[i] * rinv + br3 * r[i];
        }
    }
};

// stokes_regvel: rho^2 = r^2 + eps^2, K = ((rho^2 + eps^2) f + r (f.r))/rho^3; source [f(3), eps]
struct VjpStokesRegVel {
    static constexpr int KS = 4, KT = 3;
    template <class R>
    __device__ static void vjp(R dx, R dy, R dz, const R *f, const R *w, R *gr, R *gf) {
        const R r[3] = {dx, dy, dz};
        const R eps = f[3];
        R rr = dx * dx + dy * dy + dz * dz;
        R s = rsqrt_masked(rr + eps * eps);
        R s3 = s * (s * s), s5 = s3 * (s * s);
        R h = rr + R(2) * eps * eps;
        R a = f[0] * dx + f[1] * dy + f[2] * dz;
        R b = w[0] * dx + w[1] * dy + w[2] * dz;
        R wf = w[0] * f[0] + w[1] * f[1] + w[2] * f[2];
        R as3 = a * s3, bs3 = b * s3, hs3 = h * s3;
        R ab5 = R(3) * a * b * s5;
        R cr = wf * (R(2) * s3 - R(3) * h * s5) - ab5;
#pragma unroll
        for (int i = 0; i < 3; ++i) {
            gr[i] += cr * r[i] + as3 * w[i] + bs3 * f[i];
            gf[i] += w[i] * hs3 + bs3 * r[i];
        }
        gf[3] += eps * (wf * (R(4) * s3 - R(3) * h * s5) - ab5);
    }
};

// rpy_u: the velocity of rpy_ulapu (cotangent of the Laplacian zero)
struct VjpRpyU {
    static constexpr int KS = 4, KT = 3;
    template <class R>
    __device__ static void vjp(R dx, R dy, R dz, const R *f, const R *w, R *gr, R *gf);
};

// stk_ulapu: rpy_ulapu with b = 0
struct VjpStkULapU {
    static constexpr int KS = 3, KT = 6;
    template <class R>
    __device__ static void vjp(R dx, R dy, R dz, const R *f, const R *w, R *gr, R *gf);
};

template <int UK> struct VjpOf { using type = void; };
template <class R>
__device__ void VjpRpyU::vjp(R dx, R dy, R dz, const R *f, const R *w, R *gr, R *gf) {
    const R w6[6] = {w[0], w[1], w[2], R(0), R(0), R(0)};
    VjpRpyULapU::vjp(dx, dy, dz, f, w6, gr, gf);
}
template <class R>
__device__ void VjpStkULapU::vjp(R dx, R dy, R dz, const R *f, const R *w, R *gr, R *gf) {
    const R f4[4] = {f[0], f[1], f[2], R(0)};
    R g4[4] = {gf[0], gf[1], gf[2], R(0)};
    VjpRpyULapU::vjp(dx, dy, dz, f4, w, gr, g4);
    gf[0] = g4[0];
    gf[1] = g4[1];
    gf[2] = g4[2];
}

template <> struct VjpOf<0> { using type = VjpLaplaceP; };
template <> struct VjpOf<2> { using type = VjpLaplacePGrad; };
template <> struct VjpOf<6> { using type = VjpStokesVel; };
template <> struct VjpOf<9> { using type = VjpStokesRegVel; };
template <> struct VjpOf<10> { using type = VjpRpyU; };
template <> struct VjpOf<12> { using type = VjpStkULapU; };
template <> struct VjpOf<7> { using type = VjpStokesPVel; };
template <> struct VjpOf<8> { using type = VjpStokesDoublePVel; };
template <> struct VjpOf<11> { using type = VjpRpyULapU; };
template <> struct VjpOf<13> { using type = VjpStokesPVelFused; };

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

template <class V, int PREC, int T>
__global__ void __launch_bounds__(kWave) vjp_pair_kernel(VjpPass a) {
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
#pragma unroll
                for (int t = 0; t < T; ++t) {
                    CT gr[3] = {CT(0), CT(0), CT(0)};
                    V::template vjp<CT>((CT)(tx[t] - sx), (CT)(ty[t] - sy), (CT)(tz[t] - sz), f, w[t], gr, rs + 3);
#pragma unroll
                    for (int d = 0; d < 3; ++d) {
                        tg[t][d] += gr[d];
                        rs[d] -= gr[d];
                    }
                }
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
    for (int c = 0; c < nchunk; ++c)
        s += partial[((long long)c * nd + d) * n + i];
    out[(long long)i * ld + off + (d - d0)] = scale * s;
}

} // namespace stkp2p
