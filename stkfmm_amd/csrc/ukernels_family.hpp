// ukernels_family.hpp -- the rest of the STKFMM kernel family, for gfx950: Stokes PVelGrad,
// PVelLaplacian and Traction (single and double layer), the regularised force/torque kernels of
// StokesRegVelOmega, and the Laplace quadrupole kernels of LapQPGradGrad.
//
// Same contract as ukernels.hpp (KS, KS1, KL, KT, SRC_SCALE_MASK, deg, post, stage, eval).
//
// These kernels carry powers of 1/r up to r^-9 (r^-7 for the double layers) in the reference's
// algebra.  Formed literally, rinv^9 overflows FP32 for a close pair (r < ~1e-4 in the work item's
// level-scaled frame), so every kernel here is evaluated in the unit-direction form: with
// s = 1/|r| (masked) and n = r s, a term  poly_k(r) * rinv^(k+m)  with poly_k homogeneous of degree
// k equals  poly_k(n) * s^m, and each output becomes a polynomial in n (entries bounded by the
// source strengths) times one power of s -- the power the output's own magnitude needs.  The same
// form serves FP64, where it agrees with the reference's order to rounding (~1e-15 relative).
// Coincident pairs: s = 0 gives n = 0 and every output carries at least one factor of s, so they
// contribute exactly 0, as the reference's masked approx_rsqrt does.
//
// Each struct cites the reference lines it restates and shows the rewrite from the reference's
// expression to the unit-direction form.
#pragma once

#include "ukernels.hpp"

namespace stkp2p {

// single-layer Stokes source [fx, fy, fz, TrD].  The reference keeps one running coefficient,
// c = f.r, and updates it in place: c - TrD for u, (c - TrD) + TrD for grad p / lap u, and
// ((c - TrD) + TrD) - TrD for grad u (StokesSingleLayerKernel.hpp:76-98, 139-145).  When
// |f.r| << |TrD| (close pairs) each round trip through TrD costs f.r relative accuracy, and the
// reference's results carry that; FP64 reproduces the sequence so that it agrees with the reference to
// rounding.  FP32 uses the exact expressions (there the round trip would cost ~1e-7 |TrD| / |f.r|).
// Unit-direction forms: with s = 1/r, n = r s,  c s^k r-powers become  (c s) s^(k-1).
template <class R>
struct SlFrame {
    R s, nx, ny, nz;
    R c0, c1, c2, c3; // (f.r) s, (f.r - TrD) s, and the two later states of the coefficient, times s
    __device__ SlFrame(R dx, R dy, R dz, const R *f) {
        const R r2 = dx * dx + dy * dy + dz * dz;
        s = rsqrt_masked(r2);
        nx = dx * s;
        ny = dy * s;
        nz = dz * s;
        if constexpr (sizeof(R) == 8) {
            const R cc = f[0] * dx + f[1] * dy + f[2] * dz;
            const R a1 = cc - f[3];
            const R a2 = a1 + f[3];
            const R a3 = a2 - f[3];
            c0 = cc * s;
            c1 = a1 * s;
            c2 = a2 * s;
            c3 = a3 * s;
        } else {
            c0 = f[0] * nx + f[1] * ny + f[2] * nz;
            c1 = c0 - f[3] * s;
            c2 = c0;
            c3 = c1;
        }
    }
};

// stokes_pvelgrad, StokesSingleLayerKernel.hpp:61-110; source [fx, fy, fz, TrD] ->
// [p, u(3), grad p(3), grad u(9) row-major du_i/dx_j].  With r^2 s^2 = 1 (c0..c3 of SlFrame):
//   p = 2 (f.r) s^3 = 2 c0 s^2
//   u_i = s^3 (r^2 f_i + (f.r - TrD) r_i) = s f_i + (c1 s) n_i
//   dp/dx_i = 2 s^5 (r^2 f_i - 3 r_i (f.r)) = 2 s^3 (f_i - 3 n_i c2)
//   du_i/dx_j = s^5 ((r^2 d_ij - 3 r_i r_j)(f.r - TrD) + r^2 (r_i f_j - r_j f_i))
//             = s^2 (d_ij c3 - 3 n_i n_j c3 + n_i f_j - n_j f_i)
// post() applies the 2 of p and grad p.
struct UkStokesPVelGrad {
    static constexpr int KS = 4, KS1 = 4, KL = 4, KT = 16;
    static constexpr unsigned SRC_SCALE_MASK = 0x8u;
    __host__ __device__ static constexpr int deg(int k) { return k == 0 ? 2 : (k < 4 ? 1 : (k < 7 ? 3 : 2)); }
    __host__ __device__ static constexpr double post(int k) { return (k == 0 || (k >= 4 && k < 7)) ? 2.0 : 1.0; }
    __device__ static void stage(const double *in, double *o) {
        o[0] = in[0]; o[1] = in[1]; o[2] = in[2]; o[3] = in[3];
    }
    template <class R>
    __device__ static void eval(R *u, R dx, R dy, R dz, const R *f) {
        const SlFrame<R> a(dx, dy, dz, f);
        const R n[3] = {a.nx, a.ny, a.nz};
        const R s2 = a.s * a.s;
        const R s3 = s2 * a.s;
        const R ms = a.c1 * a.s;
        const R g = R(3.0) * a.c2;
        u[0] += a.c0 * s2;
#pragma unroll
        for (int i = 0; i < 3; ++i) {
            u[1 + i] += f[i] * a.s;
            u[1 + i] += n[i] * ms;
            u[4 + i] += s3 * (f[i] - n[i] * g);
        }
        const R M = a.c3 * s2;
        R p3[3], ns[3];
#pragma unroll
        for (int i = 0; i < 3; ++i) {
            p3[i] = R(3.0) * n[i] * M;
            ns[i] = n[i] * s2;
        }
#pragma unroll
        for (int i = 0; i < 3; ++i)
#pragma unroll
            for (int j = 0; j < 3; ++j) {
                if (i == j) {
                    u[7 + 3 * i + j] += M - p3[i] * n[j];
                } else {
                    u[7 + 3 * i + j] += ns[i] * f[j] - p3[i] * n[j];
                    u[7 + 3 * i + j] -= ns[j] * f[i];
                }
            }
    }
};

// stokes_pvellaplacian, StokesSingleLayerKernel.hpp:125-154: [p, u(3), lap u(3)];
//   lap u_i = 2 s^5 (f_i r^2 - 3 r_i (f.r)) = 2 s^3 (f_i - 3 n_i c2)   (= grad p of PVelGrad)
struct UkStokesPVelLaplacian {
    static constexpr int KS = 4, KS1 = 4, KL = 4, KT = 7;
    static constexpr unsigned SRC_SCALE_MASK = 0x8u;
    __host__ __device__ static constexpr int deg(int k) { return k == 0 ? 2 : (k < 4 ? 1 : 3); }
    __host__ __device__ static constexpr double post(int k) { return (k == 0 || k >= 4) ? 2.0 : 1.0; }
    __device__ static void stage(const double *in, double *o) {
        o[0] = in[0]; o[1] = in[1]; o[2] = in[2]; o[3] = in[3];
    }
    template <class R>
    __device__ static void eval(R *u, R dx, R dy, R dz, const R *f) {
        const SlFrame<R> a(dx, dy, dz, f);
        const R n[3] = {a.nx, a.ny, a.nz};
        const R s2 = a.s * a.s;
        const R s3 = s2 * a.s;
        const R ms = a.c1 * a.s;
        const R g = R(3.0) * a.c2;
        u[0] += a.c0 * s2;
#pragma unroll
        for (int i = 0; i < 3; ++i) {
            u[1 + i] += f[i] * a.s;
            u[1 + i] += n[i] * ms;
            u[4 + i] += s3 * (f[i] - n[i] * g);
        }
    }
};

// stokes_traction, StokesSingleLayerKernel.hpp:169-194 (ScaleFactor -3/(4 pi), :166): stress row-major
//   sigma_ij = r_i r_j (f.r - TrD) s^5 + d_ij TrD r^2 s^5 / 3 = n_i n_j c1 s^2 + d_ij (TrD/3) s^3
// Staged: fx, fy, fz, TrD, TrD * invthree (the reference's invthree = 0.3333333333333333333).
struct UkStokesTraction {
    static constexpr int KS = 4, KS1 = 4, KL = 5, KT = 9;
    static constexpr unsigned SRC_SCALE_MASK = 0x8u;
    __host__ __device__ static constexpr int deg(int) { return 2; }
    __host__ __device__ static constexpr double post(int k) { return ((void)k, 1.0); }
    __device__ static void stage(const double *in, double *o) {
        o[0] = in[0]; o[1] = in[1]; o[2] = in[2]; o[3] = in[3];
        o[4] = in[3] * 0.3333333333333333333;
    }
    template <class R>
    __device__ static void eval(R *u, R dx, R dy, R dz, const R *f) {
        const SlFrame<R> a(dx, dy, dz, f);
        const R n[3] = {a.nx, a.ny, a.nz};
        const R s2 = a.s * a.s;
        const R mm = a.c1 * s2;
        const R diag = f[4] * (s2 * a.s);
#pragma unroll
        for (int i = 0; i < 3; ++i) {
            const R ai = n[i] * mm;
#pragma unroll
            for (int j = 0; j < 3; ++j) {
                u[3 * i + j] += ai * n[j];
                if (i == j)
                    u[3 * i + j] += diag;
            }
        }
    }
};

// ---- double layer (StokesDoubleLayerKernel.hpp) --------------------------------------------------
// Every double-layer kernel depends on D through S = D + D^T and trD only.  Staged (FP64, the
// reference's source-only sums, :88-103): S = [2 Dxx, 2 Dyy, 2 Dzz, Dxy + Dyx, Dxz + Dzx,
// Dyz + Dzy], trD = Dxx + Dyy + Dzz.  Per pair: c = S n (the reference's (D r + D^T r) / r),
// Qh = n.D.n = (n.c)/2 (the reference's r.D.r / r^2).
__device__ inline void stage_sym(const double *in, double *o) {
    o[0] = in[0] + in[0];
    o[1] = in[4] + in[4];
    o[2] = in[8] + in[8];
    o[3] = in[1] + in[3];
    o[4] = in[2] + in[6];
    o[5] = in[5] + in[7];
    o[6] = in[0] + in[4] + in[8];
}

template <class R>
struct DlFrame {
    R s, n[3], c[3], qh;
    __device__ DlFrame(R dx, R dy, R dz, const R *g) {
        const R r2 = dx * dx + dy * dy + dz * dz;
        s = rsqrt_masked(r2);
        n[0] = dx * s;
        n[1] = dy * s;
        n[2] = dz * s;
        c[0] = g[0] * n[0] + g[3] * n[1] + g[4] * n[2];
        c[1] = g[3] * n[0] + g[1] * n[1] + g[5] * n[2];
        c[2] = g[4] * n[0] + g[5] * n[1] + g[2] * n[2];
        qh = R(0.5) * (n[0] * c[0] + n[1] * c[1] + n[2] * c[2]);
    }
};

// stokes_doublepvel part of the double-layer gradient kernels (:106-111, :192-197):
//   p = 2 s^5 (r^2 trD - 3 r.D.r) = 2 s^3 (trD - 3 Qh),   u_i = -3 r_i (r.D.r) s^5 = -3 Qh s^2 n_i
//   dp/dx_i (:115-118, :201-204, with its rinv7 *= -3):
//     = 6 s^7 (5 r_i (r.D.r) - r^2 ((D r + D^T r)_i + r_i trD)) = 6 s^4 ((5 Qh - trD) n_i - c_i)
// post(): 2 for p, 6 for grad p.
template <class R>
__device__ inline void dl_pvel_gradp(R *u, const DlFrame<R> &a, R trd) {
    const R s2 = a.s * a.s;
    const R s3 = s2 * a.s;
    const R s4 = s2 * s2;
    const R h = R(-3.0) * a.qh;
    u[0] += s3 * (trd + h);
    const R v = h * s2;
    const R w = R(5.0) * a.qh - trd;
#pragma unroll
    for (int i = 0; i < 3; ++i) {
        u[1 + i] += a.n[i] * v;
        u[4 + i] += s4 * (w * a.n[i] - a.c[i]);
    }
}

// stokes_doublepvelgrad, StokesDoubleLayerKernel.hpp:71-142; D row-major (9) -> 16.
//   du_i/dx_j (:131-141, rinv7 already times -3; the bracket is of degree 4 in r):
//     = -3 s^7 (-5 (r.D.r) r_i r_j + r^2 r_i (D r + D^T r)_j + r^2 (r.D.r) d_ij)
//     = 3 s^3 (5 Qh n_i n_j - n_i c_j - Qh d_ij);   post() 3
struct UkStokesDoublePVelGrad {
    static constexpr int KS = 9, KS1 = 9, KL = 7, KT = 16;
    static constexpr unsigned SRC_SCALE_MASK = 0x1FFu;
    __host__ __device__ static constexpr int deg(int k) { return k == 0 ? 2 : (k < 4 ? 1 : (k < 7 ? 3 : 2)); }
    __host__ __device__ static constexpr double post(int k) { return k == 0 ? 2.0 : (k < 4 ? 1.0 : (k < 7 ? 6.0 : 3.0)); }
    __device__ static void stage(const double *in, double *o) { stage_sym(in, o); }
    template <class R>
    __device__ static void eval(R *u, R dx, R dy, R dz, const R *g) {
        const DlFrame<R> a(dx, dy, dz, g);
        dl_pvel_gradp(u, a, g[6]);
        const R s3 = a.s * a.s * a.s;
        const R q5 = R(5.0) * a.qh;
        const R dq = s3 * a.qh;
        R e[3];
#pragma unroll
        for (int j = 0; j < 3; ++j)
            e[j] = s3 * (q5 * a.n[j] - a.c[j]);
#pragma unroll
        for (int i = 0; i < 3; ++i)
#pragma unroll
            for (int j = 0; j < 3; ++j) {
                u[7 + 3 * i + j] += a.n[i] * e[j];
                if (i == j)
                    u[7 + 3 * i + j] -= dq;
            }
    }
};

// stokes_doublelaplacian, StokesDoubleLayerKernel.hpp:157-205; D (9) -> [p, u(3), lap u(3)]
// (lap u has the expression of the double-layer grad p, :202-204)
struct UkStokesDoubleLaplacian {
    static constexpr int KS = 9, KS1 = 9, KL = 7, KT = 7;
    static constexpr unsigned SRC_SCALE_MASK = 0x1FFu;
    __host__ __device__ static constexpr int deg(int k) { return k == 0 ? 2 : (k < 4 ? 1 : 3); }
    __host__ __device__ static constexpr double post(int k) { return k == 0 ? 2.0 : (k < 4 ? 1.0 : 6.0); }
    __device__ static void stage(const double *in, double *o) { stage_sym(in, o); }
    template <class R>
    __device__ static void eval(R *u, R dx, R dy, R dz, const R *g) {
        const DlFrame<R> a(dx, dy, dz, g);
        dl_pvel_gradp(u, a, g[6]);
    }
};

// stokes_doubletraction, StokesDoubleLayerKernel.hpp:220-272 (ScaleFactor -3/(8 pi), :217):
//   np = facp r^2 s^7 (r^2 trD - 3 r.D.r) = facp s^3 (trD - 3 Qh)    (facp = 0.66666666666666)
//   tv_ij = s^7 (-5 (r.D.r) r_i r_j + r^2 r_i (D r + D^T r)_j + r^2 (r.D.r) d_ij)   (degree 4)
//         = s^3 (-5 Qh n_i n_j + n_i c_j + Qh d_ij)
//   sigma_ij = d_ij np + tv_ij + tv_ji = d_ij (np + 2 s^3 Qh) + n_i h_j + n_j h_i,
//   h_j = s^3 (c_j - 5 Qh n_j)
struct UkStokesDoubleTraction {
    static constexpr int KS = 9, KS1 = 9, KL = 7, KT = 9;
    static constexpr unsigned SRC_SCALE_MASK = 0x1FFu;
    __host__ __device__ static constexpr int deg(int) { return 2; }
    __host__ __device__ static constexpr double post(int k) { return ((void)k, 1.0); }
    __device__ static void stage(const double *in, double *o) { stage_sym(in, o); }
    template <class R>
    __device__ static void eval(R *u, R dx, R dy, R dz, const R *g) {
        const DlFrame<R> a(dx, dy, dz, g);
        const R s3 = a.s * a.s * a.s;
        const R np = R(0.66666666666666) * (s3 * (g[6] - R(3.0) * a.qh));
        const R diag = np + R(2.0) * (s3 * a.qh);
        const R q5 = R(5.0) * a.qh;
        R h[3];
#pragma unroll
        for (int j = 0; j < 3; ++j)
            h[j] = s3 * (a.c[j] - q5 * a.n[j]);
#pragma unroll
        for (int i = 0; i < 3; ++i)
#pragma unroll
            for (int j = 0; j < 3; ++j) {
                u[3 * i + j] += a.n[i] * h[j];
                u[3 * i + j] += a.n[j] * h[i];
                if (i == j)
                    u[3 * i + j] += diag;
            }
    }
};

// ---- regularised force + torque (StokesRegSingleLayerKernel.hpp) ----------------------------------
// Source [fx, fy, fz, tx, ty, tz, eps]; rho^2 = r^2 + eps^2, s = 1/rho (masked at rho = 0, where the
// reference divides by zero), n = r s, q = r^2 s^2, e = eps^2 s^2 (q + e = 1).  Staged: f, t, eps^2.
template <class R>
struct RegFrame {
    R s, n[3], q, e;
    __device__ RegFrame(R dx, R dy, R dz, R eps2) {
        const R r2 = dx * dx + dy * dy + dz * dz;
        s = rsqrt_masked(r2 + eps2);
        n[0] = dx * s;
        n[1] = dy * s;
        n[2] = dz * s;
        const R ss = s * s;
        q = r2 * ss;
        e = eps2 * ss;
    }
};

// velocity of stokes_regftvel (:85-87) and stokes_regftvelomega (:143-145):
//   H1 f + H2 (f.r) r + rc (t x r),  H1 = (r^2 + 2 eps^2) s^3 = (q + 2e) s,  H2 r_i r_j = s n_i n_j,
//   rc (t x r) = (2 r^2 + 5 eps^2) s^5 / 2 (t x r) = (2q + 5e)/2 s^2 (t x n)
template <class R>
__device__ inline void reg_velocity(R *u, const RegFrame<R> &a, const R *f, R rc2) {
    const R h1 = (a.q + R(2.0) * a.e) * a.s;
    const R fn = f[0] * a.n[0] + f[1] * a.n[1] + f[2] * a.n[2];
    const R b = fn * a.s;
    const R tx = f[4] * a.n[2] - f[5] * a.n[1];
    const R ty = f[5] * a.n[0] - f[3] * a.n[2];
    const R tz = f[3] * a.n[1] - f[4] * a.n[0];
    u[0] += f[0] * h1;
    u[0] += a.n[0] * b;
    u[0] += rc2 * tx;
    u[1] += f[1] * h1;
    u[1] += a.n[1] * b;
    u[1] += rc2 * ty;
    u[2] += f[2] * h1;
    u[2] += a.n[2] * b;
    u[2] += rc2 * tz;
}

// stokes_regftvel, StokesRegSingleLayerKernel.hpp:55-88: [f, t, eps] -> u(3)
struct UkStokesRegFTVel {
    static constexpr int KS = 7, KS1 = 7, KL = 7, KT = 3;
    static constexpr unsigned SRC_SCALE_MASK = 0x78u;
    __host__ __device__ static constexpr int deg(int) { return 1; }
    __host__ __device__ static constexpr double post(int k) { return ((void)k, 1.0); }
    __device__ static void stage(const double *in, double *o) {
        for (int k = 0; k < 6; ++k)
            o[k] = in[k];
        o[6] = in[6] * in[6];
    }
    template <class R>
    __device__ static void eval(R *u, R dx, R dy, R dz, const R *f) {
        const RegFrame<R> a(dx, dy, dz, f[6]);
        const R rc2 = (R(2.0) * a.q + R(5.0) * a.e) * (R(0.5) * (a.s * a.s));
        reg_velocity(u, a, f, rc2);
    }
};

// stokes_regftvelomega, StokesRegSingleLayerKernel.hpp:103-150: [f, t, eps] -> [u(3), omega(3)]
//   omega = D1 t + D2 (t.r) r + rc (f x r),
//   D1 = (10 eps^4 - 7 eps^2 r^2 - 2 r^4) s^7 / 4 = (10 e^2 - 7 e q - 2 q^2)/4 s^3,
//   D2 r_i r_j = (21 eps^2 + 6 r^2) s^7 / 4 r_i r_j = (21 e + 6 q)/4 s^3 n_i n_j
struct UkStokesRegFTVelOmega {
    static constexpr int KS = 7, KS1 = 7, KL = 7, KT = 6;
    static constexpr unsigned SRC_SCALE_MASK = 0x78u;
    __host__ __device__ static constexpr int deg(int k) { return k < 3 ? 1 : 2; }
    __host__ __device__ static constexpr double post(int k) { return ((void)k, 1.0); }
    __device__ static void stage(const double *in, double *o) { UkStokesRegFTVel::stage(in, o); }
    template <class R>
    __device__ static void eval(R *u, R dx, R dy, R dz, const R *f) {
        const RegFrame<R> a(dx, dy, dz, f[6]);
        const R ss = a.s * a.s;
        const R rc2 = (R(2.0) * a.q + R(5.0) * a.e) * (R(0.5) * ss);
        reg_velocity(u, a, f, rc2);
        const R s34 = R(0.25) * (ss * a.s);
        const R d1 = (a.e * (R(10.0) * a.e - R(7.0) * a.q) - R(2.0) * (a.q * a.q)) * s34;
        const R d2 = (R(21.0) * a.e + R(6.0) * a.q) * s34;
        const R tn = f[3] * a.n[0] + f[4] * a.n[1] + f[5] * a.n[2];
        const R b = d2 * tn;
        const R fx = f[1] * a.n[2] - f[2] * a.n[1];
        const R fy = f[2] * a.n[0] - f[0] * a.n[2];
        const R fz = f[0] * a.n[1] - f[1] * a.n[0];
        u[3] += f[3] * d1;
        u[3] += a.n[0] * b;
        u[3] += rc2 * fx;
        u[4] += f[4] * d1;
        u[4] += a.n[1] * b;
        u[4] += rc2 * fy;
        u[5] += f[5] * d1;
        u[5] += a.n[2] * b;
        u[5] += rc2 * fz;
    }
};

// stokes_velomega, StokesRegSingleLayerKernel.hpp:165-201 (L2T kernel of StokesRegVelOmega, :223-229):
// f(3) -> [u(3), omega(3)];  u = f s + (f.n) s n  (the Stokeslet),  omega = r^2 s^5 (f x r) = s^2 (f x n)
struct UkStokesVelOmega {
    static constexpr int KS = 3, KS1 = 3, KL = 3, KT = 6;
    static constexpr unsigned SRC_SCALE_MASK = 0u;
    __host__ __device__ static constexpr int deg(int k) { return k < 3 ? 1 : 2; }
    __host__ __device__ static constexpr double post(int k) { return ((void)k, 1.0); }
    __device__ static void stage(const double *in, double *o) { o[0] = in[0]; o[1] = in[1]; o[2] = in[2]; }
    template <class R>
    __device__ static void eval(R *u, R dx, R dy, R dz, const R *f) {
        const R r2 = dx * dx + dy * dy + dz * dz;
        const R s = rsqrt_masked(r2);
        const R nx = dx * s, ny = dy * s, nz = dz * s;
        const R b = (f[0] * nx + f[1] * ny + f[2] * nz) * s;
        const R ss = s * s;
        u[0] += f[0] * s;
        u[0] += nx * b;
        u[1] += f[1] * s;
        u[1] += ny * b;
        u[2] += f[2] * s;
        u[2] += nz * b;
        u[3] += ss * (f[1] * nz - f[2] * ny);
        u[4] += ss * (f[2] * nx - f[0] * nz);
        u[5] += ss * (f[0] * ny - f[1] * nx);
    }
};

// ---- Laplace quadrupole (LaplaceLayerKernel.hpp) ----------------------------------------------------
// Q row-major (9), staged like the Stokes double layer (S = Q + Q^T, trQ).

// laplace_quadp, LaplaceLayerKernel.hpp:209-231:  p = (3 r.Q.r - r^2 trQ) s^5 = s^3 (3 Qh - trQ)
struct UkLaplaceQuadP {
    static constexpr int KS = 9, KS1 = 9, KL = 7, KT = 1;
    static constexpr unsigned SRC_SCALE_MASK = 0u;
    __host__ __device__ static constexpr int deg(int) { return 3; }
    __host__ __device__ static constexpr double post(int k) { return ((void)k, 1.0); }
    __device__ static void stage(const double *in, double *o) { stage_sym(in, o); }
    template <class R>
    __device__ static void eval(R *u, R dx, R dy, R dz, const R *g) {
        const DlFrame<R> a(dx, dy, dz, g);
        u[0] += (a.s * a.s * a.s) * (R(3.0) * a.qh - g[6]);
    }
};

// laplace_quadpgradgrad, LaplaceLayerKernel.hpp:241-311: [p, grad(3), Hessian xx,xy,xz,yy,yz,zz]
//   grad_i = -3 s^7 (5 r_i r.Q.r - r^2 (S r + r trQ)_i) = -3 s^4 ((5 Qh - trQ) n_i - c_i)    post -3
//   H_ij = s^9 (105 r_i r_j r.Q.r - 15 r^2 (d_ij r.Q.r + r_j (S r)_i + r_i (S r)_j + r_i r_j trQ)
//               + 3 r^4 (d_ij trQ + S_ij))
//        = 3 s^5 (n_i n_j (35 Qh - 5 trQ) - 5 (n_j c_i + n_i c_j) + S_ij + d_ij (trQ - 5 Qh))  post 3
struct UkLaplaceQuadPGradGrad {
    static constexpr int KS = 9, KS1 = 9, KL = 7, KT = 10;
    static constexpr unsigned SRC_SCALE_MASK = 0u;
    __host__ __device__ static constexpr int deg(int k) { return k == 0 ? 3 : (k < 4 ? 4 : 5); }
    __host__ __device__ static constexpr double post(int k) { return k == 0 ? 1.0 : (k < 4 ? -3.0 : 3.0); }
    __device__ static void stage(const double *in, double *o) { stage_sym(in, o); }
    template <class R>
    __device__ static void eval(R *u, R dx, R dy, R dz, const R *g) {
        const DlFrame<R> a(dx, dy, dz, g);
        const R trq = g[6];
        const R s2 = a.s * a.s;
        const R s3 = s2 * a.s;
        const R s4 = s2 * s2;
        const R s5 = s4 * a.s;
        u[0] += s3 * (R(3.0) * a.qh - trq);
        const R w = R(5.0) * a.qh - trq;
#pragma unroll
        for (int i = 0; i < 3; ++i)
            u[1 + i] += s4 * (w * a.n[i] - a.c[i]);
        // S_ij of the staged layout: diagonal g[0..2], off-diagonal xy g[3], xz g[4], yz g[5]
        const R A = R(35.0) * a.qh - R(5.0) * trq;
        const R dg = trq - R(5.0) * a.qh;
        R nA[3], c5[3];
#pragma unroll
        for (int i = 0; i < 3; ++i) {
            nA[i] = a.n[i] * A;
            c5[i] = R(5.0) * a.c[i];
        }
        constexpr int I[6] = {0, 0, 0, 1, 1, 2}, J[6] = {0, 1, 2, 1, 2, 2}, SIDX[6] = {0, 3, 4, 1, 5, 2};
#pragma unroll
        for (int k = 0; k < 6; ++k) {
            const int i = I[k], j = J[k];
            R h = nA[i] * a.n[j] - (a.n[j] * c5[i] + a.n[i] * c5[j]) + g[SIDX[k]];
            if (i == j)
                h += dg;
            u[4 + k] += s5 * h;
        }
    }
};

} // namespace stkp2p
