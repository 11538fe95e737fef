// vjp_kernels.hpp -- adjoint (vector-Jacobian product) kernels of the direct-sum P2P.
//
// For L = sum_t w_t . u_t with u_t = scale * sum_s K(x_t - x_s, f_s), and a pair with displacement
// r, source value f and cotangent w:
//   g_r = sum_k w_k dK_k/dr,   g_f = sum_k w_k dK_k/df
// so that dL/dx_t = scale * sum_s g_r,  dL/dx_s = -scale * sum_t g_r,  dL/df_s = scale * sum_t g_f.
// The per-pair formulas are differentiated by hand from the forward micro-kernels (ukernels.hpp).
// Coincident pairs contribute zero (the forward is masked there).  The reference has no adjoint.
//
// Two passes, each a chunked direct sum with a lane per entity:
//   target pass: lane = target, sources staged in LDS -> partial g_trg_coord
//   source pass: lane = source, targets (coords + w) staged in LDS -> partial g_src_coord, g_src_value
// then a reduction over chunks (deterministic order) writes scale * sum (overwrite).
#pragma once

#include <hip/hip_runtime.h>

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

template <int UK> struct VjpOf { using type = void; };
template <> struct VjpOf<7> { using type = VjpStokesPVel; };
template <> struct VjpOf<8> { using type = VjpStokesDoublePVel; };
template <> struct VjpOf<11> { using type = VjpRpyULapU; };
template <> struct VjpOf<13> { using type = VjpStokesPVelFused; };

// partial layout: [chunk][dim][n_lane]
struct VjpPass {
    const double *sc, *sv, *tc, *w;
    int n_src, n_trg;
    int chunk;          // entities of the other side per chunk
    double *partial;
};

// lane = target; sources [c*chunk, (c+1)*chunk) staged in LDS; partial dims 3
template <class V, int PREC>
__global__ void __launch_bounds__(kWave) vjp_target_kernel(VjpPass a) {
    using CT = typename std::conditional<PREC == 0, double, float>::type;
    constexpr int KS = V::KS, KT = V::KT;
    __shared__ double s_x[kSrcTile], s_y[kSrcTile], s_z[kSrcTile];
    __shared__ CT s_v[KS][kSrcTile];
    const int lane = threadIdx.x;
    const int t = blockIdx.x * kWave + lane;
    const int tt = (t < a.n_trg) ? t : a.n_trg - 1;
    const double tx = a.tc[3LL * tt], ty = a.tc[3LL * tt + 1], tz = a.tc[3LL * tt + 2];
    CT w[KT];
#pragma unroll
    for (int k = 0; k < KT; ++k)
        w[k] = (CT)a.w[(long long)KT * tt + k];
    double g[3] = {0.0, 0.0, 0.0};
    const int s0 = blockIdx.y * a.chunk;
    const int s1 = (s0 + a.chunk < a.n_src) ? s0 + a.chunk : a.n_src;
    for (int base = s0; base < s1; base += kSrcTile) {
        const int n = (s1 - base < kSrcTile) ? s1 - base : kSrcTile;
        __syncthreads();
        if (lane < n) {
            const long long s = base + lane;
            s_x[lane] = a.sc[3 * s];
            s_y[lane] = a.sc[3 * s + 1];
            s_z[lane] = a.sc[3 * s + 2];
#pragma unroll
            for (int k = 0; k < KS; ++k)
                s_v[k][lane] = (CT)a.sv[KS * s + k];
        }
        __syncthreads();
        CT gp[3] = {CT(0), CT(0), CT(0)};
        for (int j = 0; j < n; ++j) {
            CT f[KS], gf[KS];
#pragma unroll
            for (int k = 0; k < KS; ++k) {
                f[k] = s_v[k][j];
                gf[k] = CT(0);
            }
            V::template vjp<CT>((CT)(tx - s_x[j]), (CT)(ty - s_y[j]), (CT)(tz - s_z[j]), f, w, gp, gf);
        }
        g[0] += (double)gp[0];
        g[1] += (double)gp[1];
        g[2] += (double)gp[2];
    }
    if (t < a.n_trg) {
#pragma unroll
        for (int d = 0; d < 3; ++d)
            a.partial[((long long)blockIdx.y * 3 + d) * a.n_trg + t] = g[d];
    }
}

// lane = source; targets [c*chunk, (c+1)*chunk) staged in LDS; partial dims 3 + KS
template <class V, int PREC>
__global__ void __launch_bounds__(kWave) vjp_source_kernel(VjpPass a) {
    using CT = typename std::conditional<PREC == 0, double, float>::type;
    constexpr int KS = V::KS, KT = V::KT, ND = 3 + KS;
    __shared__ double s_x[kSrcTile], s_y[kSrcTile], s_z[kSrcTile];
    __shared__ CT s_w[KT][kSrcTile];
    const int lane = threadIdx.x;
    const int s = blockIdx.x * kWave + lane;
    const int ss = (s < a.n_src) ? s : a.n_src - 1;
    const double sx = a.sc[3LL * ss], sy = a.sc[3LL * ss + 1], sz = a.sc[3LL * ss + 2];
    CT f[KS];
#pragma unroll
    for (int k = 0; k < KS; ++k)
        f[k] = (CT)a.sv[(long long)KS * ss + k];
    double g[ND];
#pragma unroll
    for (int d = 0; d < ND; ++d)
        g[d] = 0.0;
    const int t0 = blockIdx.y * a.chunk;
    const int t1 = (t0 + a.chunk < a.n_trg) ? t0 + a.chunk : a.n_trg;
    for (int base = t0; base < t1; base += kSrcTile) {
        const int n = (t1 - base < kSrcTile) ? t1 - base : kSrcTile;
        __syncthreads();
        if (lane < n) {
            const long long t = base + lane;
            s_x[lane] = a.tc[3 * t];
            s_y[lane] = a.tc[3 * t + 1];
            s_z[lane] = a.tc[3 * t + 2];
#pragma unroll
            for (int k = 0; k < KT; ++k)
                s_w[k][lane] = (CT)a.w[KT * t + k];
        }
        __syncthreads();
        CT gr[3] = {CT(0), CT(0), CT(0)};
        CT gf[KS];
#pragma unroll
        for (int k = 0; k < KS; ++k)
            gf[k] = CT(0);
        for (int j = 0; j < n; ++j) {
            CT w[KT];
#pragma unroll
            for (int k = 0; k < KT; ++k)
                w[k] = s_w[k][j];
            V::template vjp<CT>((CT)(s_x[j] - sx), (CT)(s_y[j] - sy), (CT)(s_z[j] - sz), f, w, gr, gf);
        }
#pragma unroll
        for (int d = 0; d < 3; ++d)
            g[d] -= (double)gr[d];
#pragma unroll
        for (int k = 0; k < KS; ++k)
            g[3 + k] += (double)gf[k];
    }
    if (s < a.n_src) {
#pragma unroll
        for (int d = 0; d < ND; ++d)
            a.partial[((long long)blockIdx.y * ND + d) * a.n_src + s] = g[d];
    }
}

// out[i*ld + off + d] = scale * sum_c partial[c][d][i] for d in [d0, d1)
__global__ void vjp_reduce_kernel(const double *partial, int n, int nd, int nchunk, int d0, int d1,
                                  double *out, int ld, int off, double scale) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n || out == nullptr)
        return;
    for (int d = d0; d < d1; ++d) {
        double s = 0.0;
        for (int c = 0; c < nchunk; ++c)
            s += partial[((long long)c * nd + d) * n + i];
        out[(long long)i * ld + off + (d - d0)] = scale * s;
    }
}

} // namespace stkp2p
