// vjp_pairs.hpp -- the per-pair adjoints (vector-Jacobian products) of the P2P micro-kernels.
//
// For a pair with displacement r = x_t - x_s, source value f and target cotangent w, V::vjp adds
//   g_r = sum_k w_k dK_k/dr   to gr[0..2]   and   g_f = sum_k w_k dK_k/df   to gf[0..KS-1],
// where K is the micro-kernel's per-pair output before its ScaleFactor.  The formulas are
// differentiated by hand from the forward micro-kernels (ukernels.hpp); coincident pairs contribute
// zero (the forward is masked there).  The reference has no adjoint.  Kept apart from the kernels
// (vjp_kernels.hpp) so that the host test harness compiles them for the CPU.  gf accumulates across
// pairs (the source-side sums), so each product is added by its own statement: "gf += a b" is one FMA,
// while "gf += a b + c d" costs an extra add (the sum is rounded before it joins gf).
#pragma once

#include <hip/hip_runtime.h>

#include "ukernels.hpp"

namespace stkp2p {

// stokes_pvel: K0 = 2a/r^3, K_i = f_i/r + r_i (a - T)/r^3, a = f.r, T = TrD.  With wr = w.r, wf = w.f:
//   g_r = cf f + (a - T)/r^3 w - cr r,  cf = (2 w0 + wr)/r^3,
//   cr = wf/r^3 + (6 w0 a + 3 wr (a - T))/r^5 = (wf + (6 w0 a + 3 wr (a - T))/r^2)/r^3  (no 1/r^5),
//   g_f = cf r + w/r,  g_T = -wr/r^3.
struct VjpStokesPVel {
    static constexpr int KS = 4, KT = 4;
    template <class R>
    __device__ static void vjp(R dx, R dy, R dz, const R *f, const R *w, R *gr, R *gf) {
        const R r[3] = {dx, dy, dz};
        R r2 = dx * dx + dy * dy + dz * dz;
        R rinv = rsqrt_masked(r2);
        R rinv2 = rinv * rinv, rinv3 = rinv2 * rinv;
        R a = f[0] * dx + f[1] * dy + f[2] * dz;
        R am = a - f[3];
        R wr = w[1] * dx + w[2] * dy + w[3] * dz;
        R wf = w[1] * f[0] + w[2] * f[1] + w[3] * f[2];
        R cf = (R(2) * w[0] + wr) * rinv3;
        R cr = (wf + (R(6) * w[0] * a + R(3) * wr * am) * rinv2) * rinv3;
        R amr3 = am * rinv3;
#pragma unroll
        for (int i = 0; i < 3; ++i) {
            gr[i] += cf * f[i] + amr3 * w[1 + i] - cr * r[i];
            gf[i] += cf * r[i];
            gf[i] += w[1 + i] * rinv;
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
        const R d = R(2) * w[0] * rinv3;
#pragma unroll
        for (int i = 0; i < 3; ++i) {
            gr[i] += -crr5 * (Dr[i] + DTr[i]) + cq7 * r[i] - R(3) * w[1 + i] * q5;
            const R ci = -crr5 * r[i];
#pragma unroll
            for (int j = 0; j < 3; ++j)
                gf[3 * i + j] += ci * r[j];
            gf[4 * i] += d;
        }
    }
};

// rpy_ulapu: u = f/r + r a/r^3 + beta c, lap u = 2c, c = f/r^3 - 3 a r/r^5, beta = b^2/3, a = f.r.
// Cotangents w_u = w[0..2], w_l = w[3..5]; v = beta w_u + 2 w_l; wuf = w_u.f, wur = w_u.r, vf = v.f,
// vr = v.r.  Differentiating and collecting the powers of 1/r (p3 = -3/r^2):
//   g_f,i = w_u,i/r + (v_i + r_i q)/r^3,                 q = wur + p3 vr
//   g_r,i = s r_i + (f_i q + a m_i)/r^3,                 m_i = w_u,i + p3 v_i
//   s = (-wuf + (-3 (wur a + vf) + 15 a vr/r^2)/r^2)/r^3
//   g_b = (2/3) b w_u.c = (2/3) b (wuf + p3 a wur)/r^3
struct VjpRpyULapU {
    static constexpr int KS = 4, KT = 6;
    template <class R>
    __device__ static void vjp(R dx, R dy, R dz, const R *f, const R *w, R *gr, R *gf) {
        const R r[3] = {dx, dy, dz};
        R r2 = dx * dx + dy * dy + dz * dz;
        R rinv = rsqrt_masked(r2);
        R rinv2 = rinv * rinv, rinv3 = rinv2 * rinv;
        R p3 = R(-3) * rinv2;
        R a = f[0] * dx + f[1] * dy + f[2] * dz;
        const R third = R(STKP2P_RPY_ONE_OVER_THREE);
        R beta = third * f[3] * f[3];
        R v[3];
#pragma unroll
        for (int i = 0; i < 3; ++i)
            v[i] = beta * w[i] + R(2) * w[3 + i];
        R wuf = w[0] * f[0] + w[1] * f[1] + w[2] * f[2];
        R wur = w[0] * dx + w[1] * dy + w[2] * dz;
        R vf = v[0] * f[0] + v[1] * f[1] + v[2] * f[2];
        R vr = v[0] * dx + v[1] * dy + v[2] * dz;
        R q = wur + p3 * vr;
        R s_r = (-wuf + (R(-3) * (wur * a + vf) + R(15) * rinv2 * (a * vr)) * rinv2) * rinv3;
#pragma unroll
        for (int i = 0; i < 3; ++i) {
            gf[i] += (v[i] + r[i] * q) * rinv3;
            gf[i] += w[i] * rinv;
            const R m = w[i] + p3 * v[i];
            gr[i] += s_r * r[i] + (f[i] * q + a * m) * rinv3;
        }
        gf[3] += (R(2) * third * f[3]) * ((wuf + p3 * (a * wur)) * rinv3);
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
        gf[0] += w[0] * rinv;
        gf[0] += -wg * rinv3;
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
            gf[i] += w[i] * rinv;
            gf[i] += br3 * r[i];
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
            gf[i] += w[i] * hs3;
            gf[i] += bs3 * r[i];
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


} // namespace stkp2p
