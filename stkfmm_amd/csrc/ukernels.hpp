// ukernels.hpp -- per-pair micro-kernels of the STKFMM P2P path, for gfx950.
//
// One struct per reference uKerEval.  Each provides
//   KS    source values per point as the caller stores them (ker_dim[0] of the reference kernel),
//   KS1   how many of them come from the first value array (the rest from the second; only the
//         fused single+double layer kernel uses two arrays),
//   KL    values per source after stage() (what sits in LDS),
//   KT    target values (ker_dim[1]),
//   SRC_SCALE_MASK  bit k set: source value k has the dimension of a length and is multiplied by
//         lambda when the pair arithmetic runs in coordinates scaled by lambda (FMMData::scaleSrc,
//         Lib/src/FMMData.cpp:285-317),
//   deg(k) homogeneity degree of output k: out(r) = lambda^deg * out(lambda r, scaled sources)
//         (FMMData::scaleTrg, Lib/src/FMMData.cpp:319-470),
//   stage(in, out)  per-source precomputation in FP64.  Only sub-expressions the reference
//         evaluates from source values alone are hoisted here (e.g. sxy + syx, a*a/3), each with
//         the same operands and operation, so hoisting does not change a single rounding,
//   eval<R>(u, dx, dy, dz, f)  u += K(r, f) without the scale factor, r = x_trg - x_src, in the
//         algebraic order of the reference uKerEval (FMA contraction is left to the compiler).
#pragma once

#include <hip/hip_runtime.h>

#ifndef STKP2P_FP64_NEWTON
#define STKP2P_FP64_NEWTON 2
#endif

namespace stkp2p {

// ---- masked reciprocal square root ------------------------------------------------------------
// approx_rsqrt(r2, r2 > 0) of the reference: 0 for coincident points.  v_rsq_{f64,f32} seeds,
// refined by Newton steps y <- y + y (1 - r2 y^2)/2 written as e = 0.5 - (r2/2) y^2, y = y + y e.
__device__ __forceinline__ double rsqrt_masked(double r2) {
    double y = __builtin_amdgcn_rsq(r2);
    const double hr = 0.5 * r2;
#pragma unroll
    for (int i = 0; i < STKP2P_FP64_NEWTON; ++i) {
        const double e = __builtin_fma(-hr, y * y, 0.5);
        y = __builtin_fma(y, e, y);
    }
    return (r2 > 0.0) ? y : 0.0;
}

__device__ __forceinline__ float rsqrt_masked(float r2) {
    float y = __builtin_amdgcn_rsqf(r2);
    const float hr = 0.5f * r2;
    const float e = __builtin_fmaf(-hr, y * y, 0.5f);
    y = __builtin_fmaf(y, e, y);
    return (r2 > 0.0f) ? y : 0.0f;
}

// reference constant one_over_three = 0.3333333333333 (RPYKernel.hpp:30,66)
#define STKP2P_RPY_ONE_OVER_THREE 0.3333333333333

// ---- Laplace ----------------------------------------------------------------------------------

// laplace_p, LaplaceLayerKernel.hpp:58-64
struct UkLaplaceP {
    static constexpr int KS = 1, KS1 = 1, KL = 1, KT = 1;
    static constexpr unsigned SRC_SCALE_MASK = 0u;
    __host__ __device__ static constexpr int deg(int) { return 1; }
    __device__ static void stage(const double *in, double *o) { o[0] = in[0]; }
    template <class R>
    __device__ static void eval(R *u, R dx, R dy, R dz, const R *f) {
        R r2 = dx * dx + dy * dy + dz * dz;
        R rinv = rsqrt_masked(r2);
        u[0] += f[0] * rinv;
    }
};

// laplace_dipolep, LaplaceLayerKernel.hpp:104-112
struct UkLaplaceDipoleP {
    static constexpr int KS = 3, KS1 = 3, KL = 3, KT = 1;
    static constexpr unsigned SRC_SCALE_MASK = 0x7u;
    __host__ __device__ static constexpr int deg(int) { return 1; }
    __device__ static void stage(const double *in, double *o) { o[0] = in[0]; o[1] = in[1]; o[2] = in[2]; }
    template <class R>
    __device__ static void eval(R *u, R dx, R dy, R dz, const R *f) {
        R r2 = dx * dx + dy * dy + dz * dz;
        R rinv = rsqrt_masked(r2);
        R rinv3 = rinv * rinv * rinv;
        R rdotn = f[0] * dx + f[1] * dy + f[2] * dz;
        u[0] += rinv3 * rdotn;
    }
};

// laplace_pgrad, LaplaceLayerKernel.hpp:84-95
struct UkLaplacePGrad {
    static constexpr int KS = 1, KS1 = 1, KL = 1, KT = 4;
    static constexpr unsigned SRC_SCALE_MASK = 0u;
    __host__ __device__ static constexpr int deg(int k) { return k == 0 ? 1 : 2; }
    __device__ static void stage(const double *in, double *o) { o[0] = in[0]; }
    template <class R>
    __device__ static void eval(R *u, R dx, R dy, R dz, const R *f) {
        R r2 = dx * dx + dy * dy + dz * dz;
        R rinv = rsqrt_masked(r2);
        R sv = f[0] * rinv * rinv * rinv;
        u[0] += sv * r2;
        u[1] -= sv * dx;
        u[2] -= sv * dy;
        u[3] -= sv * dz;
    }
};

// laplace_dipolepgrad, LaplaceLayerKernel.hpp:121-135
struct UkLaplaceDipolePGrad {
    static constexpr int KS = 3, KS1 = 3, KL = 3, KT = 4;
    static constexpr unsigned SRC_SCALE_MASK = 0x7u;
    __host__ __device__ static constexpr int deg(int k) { return k == 0 ? 1 : 2; }
    __device__ static void stage(const double *in, double *o) { o[0] = in[0]; o[1] = in[1]; o[2] = in[2]; }
    template <class R>
    __device__ static void eval(R *u, R dx, R dy, R dz, const R *f) {
        R r2 = dx * dx + dy * dy + dz * dz;
        R rinv = rsqrt_masked(r2);
        R rinv2 = rinv * rinv;
        R rinv3 = rinv2 * rinv;
        R rinv5 = rinv3 * rinv2;
        R rdotn = f[0] * dx + f[1] * dy + f[2] * dz;
        const R three = R(3.0);
        u[0] += rdotn * rinv3;
        u[1] += (f[0] * r2 - three * rdotn * dx) * rinv5;
        u[2] += (f[1] * r2 - three * rdotn * dy) * rinv5;
        u[3] += (f[2] * r2 - three * rdotn * dz) * rinv5;
    }
};

// laplace_pgradgrad, LaplaceLayerKernel.hpp:144-165
struct UkLaplacePGradGrad {
    static constexpr int KS = 1, KS1 = 1, KL = 1, KT = 10;
    static constexpr unsigned SRC_SCALE_MASK = 0u;
    __host__ __device__ static constexpr int deg(int k) { return k == 0 ? 1 : (k < 4 ? 2 : 3); }
    __device__ static void stage(const double *in, double *o) { o[0] = in[0]; }
    template <class R>
    __device__ static void eval(R *u, R dx, R dy, R dz, const R *f) {
        R r2 = dx * dx + dy * dy + dz * dz;
        R rinv = rsqrt_masked(r2);
        R rinv2 = rinv * rinv;
        R rinv3 = rinv2 * rinv;
        const R three = R(3.0);
        R sv = f[0] * rinv3;
        u[0] += sv * r2;
        u[1] -= sv * dx;
        u[2] -= sv * dy;
        u[3] -= sv * dz;
        sv *= rinv2;
        u[4] += sv * (three * dx * dx - r2);
        u[5] += sv * three * dx * dy;
        u[6] += sv * three * dx * dz;
        u[7] += sv * (three * dy * dy - r2);
        u[8] += sv * three * dy * dz;
        u[9] += sv * (three * dz * dz - r2);
    }
};

// laplace_dipolepgradgrad, LaplaceLayerKernel.hpp:174-199
struct UkLaplaceDipolePGradGrad {
    static constexpr int KS = 3, KS1 = 3, KL = 3, KT = 10;
    static constexpr unsigned SRC_SCALE_MASK = 0x7u;
    __host__ __device__ static constexpr int deg(int k) { return k == 0 ? 1 : (k < 4 ? 2 : 3); }
    __device__ static void stage(const double *in, double *o) { o[0] = in[0]; o[1] = in[1]; o[2] = in[2]; }
    template <class R>
    __device__ static void eval(R *u, R dx, R dy, R dz, const R *f) {
        R r2 = dx * dx + dy * dy + dz * dz;
        const R three = R(3.0);
        const R threer2 = three * r2;
        const R six = R(6.0);
        const R fifteen = R(15.0);
        R rinv = rsqrt_masked(r2);
        R rinv2 = rinv * rinv;
        R rinv3 = rinv2 * rinv;
        R rinv5 = rinv2 * rinv3;
        R rinv7 = rinv2 * rinv5;
        R rdotn = dx * f[0] + dy * f[1] + dz * f[2];
        u[0] += rdotn * rinv3;
        u[1] += (f[0] * r2 - rdotn * three * dx) * rinv5;
        u[2] += (f[1] * r2 - rdotn * three * dy) * rinv5;
        u[3] += (f[2] * r2 - rdotn * three * dz) * rinv5;
        u[4] += (fifteen * dx * dx * rdotn - r2 * (three * rdotn + six * dx * f[0])) * rinv7;
        u[5] += (fifteen * dx * dy * rdotn - threer2 * (dx * f[1] + dy * f[0])) * rinv7;
        u[6] += (fifteen * dx * dz * rdotn - threer2 * (dx * f[2] + dz * f[0])) * rinv7;
        u[7] += (fifteen * dy * dy * rdotn - r2 * (three * rdotn + six * dy * f[1])) * rinv7;
        u[8] += (fifteen * dy * dz * rdotn - threer2 * (dy * f[2] + dz * f[1])) * rinv7;
        u[9] += (fifteen * dz * dz * rdotn - r2 * (three * rdotn + six * dz * f[2])) * rinv7;
    }
};

// ---- Stokes -----------------------------------------------------------------------------------

// Stokeslet 3 -> 3 (PVFMM stokes_vel, StokesLayerKernel.hpp:38-42; closed form of
// Test/SimpleKernel.cpp:986-1023 with a = 0): u = f/r + r (f.r)/r^3.
struct UkStokesVel {
    static constexpr int KS = 3, KS1 = 3, KL = 3, KT = 3;
    static constexpr unsigned SRC_SCALE_MASK = 0u;
    __host__ __device__ static constexpr int deg(int) { return 1; }
    __device__ static void stage(const double *in, double *o) { o[0] = in[0]; o[1] = in[1]; o[2] = in[2]; }
    template <class R>
    __device__ static void eval(R *u, R dx, R dy, R dz, const R *f) {
        R r2 = dx * dx + dy * dy + dz * dz;
        R rinv = rsqrt_masked(r2);
        R rinv3 = rinv * rinv * rinv;
        R fdotr = f[0] * dx + f[1] * dy + f[2] * dz;
        R c = fdotr * rinv3;
        u[0] += f[0] * rinv + dx * c;
        u[1] += f[1] * rinv + dy * c;
        u[2] += f[2] * rinv + dz * c;
    }
};

// stokes_pvel, StokesSingleLayerKernel.hpp:33-46; source [fx, fy, fz, TrD]
struct UkStokesPVel {
    static constexpr int KS = 4, KS1 = 4, KL = 4, KT = 4;
    static constexpr unsigned SRC_SCALE_MASK = 0x8u;
    __host__ __device__ static constexpr int deg(int k) { return k == 0 ? 2 : 1; }
    __device__ static void stage(const double *in, double *o) {
        o[0] = in[0]; o[1] = in[1]; o[2] = in[2]; o[3] = in[3];
    }
    template <class R>
    __device__ static void eval(R *u, R dx, R dy, R dz, const R *f) {
        R r2 = dx * dx + dy * dy + dz * dz;
        R rinv = rsqrt_masked(r2);
        R rinv3 = rinv * rinv * rinv;
        R commonCoeff = f[0] * dx + f[1] * dy + f[2] * dz;
        const R two = R(2.0);
        u[0] += two * commonCoeff * rinv3;
        commonCoeff -= f[3];
        u[1] += rinv3 * (r2 * f[0] + dx * commonCoeff);
        u[2] += rinv3 * (r2 * f[1] + dy * commonCoeff);
        u[3] += rinv3 * (r2 * f[2] + dz * commonCoeff);
    }
};

// stokes_doublepvel, StokesDoubleLayerKernel.hpp:31-56; source D row-major [Dxx..Dzz].
// Staged: Dxx, Dyy, Dzz, Dxy+Dyx, Dxz+Dzx, Dyz+Dzy, Dxx+Dyy+Dzz (the reference's own source-only
// sub-expressions, :46-48,51).
struct UkStokesDoublePVel {
    static constexpr int KS = 9, KS1 = 9, KL = 7, KT = 4;
    static constexpr unsigned SRC_SCALE_MASK = 0x1FFu;
    __host__ __device__ static constexpr int deg(int k) { return k == 0 ? 2 : 1; }
    __device__ static void stage(const double *in, double *o) {
        o[0] = in[0];
        o[1] = in[4];
        o[2] = in[8];
        o[3] = in[1] + in[3];
        o[4] = in[2] + in[6];
        o[5] = in[5] + in[7];
        o[6] = in[0] + in[4] + in[8];
    }
    template <class R>
    __device__ static void eval_staged(R *u, R dx, R dy, R dz, R r2, R rinv, const R *g) {
        R rinv3 = rinv * rinv * rinv;
        R rinv5 = rinv3 * rinv * rinv;
        const R two = R(2.0);
        R commonCoeff = g[0] * dx * dx + g[1] * dy * dy + g[2] * dz * dz;
        commonCoeff += g[3] * dx * dy;
        commonCoeff += g[4] * dx * dz;
        commonCoeff += g[5] * dy * dz;
        commonCoeff *= R(-3.0) * rinv5;
        u[0] += two * (commonCoeff + r2 * rinv5 * g[6]);
        u[1] += dx * commonCoeff;
        u[2] += dy * commonCoeff;
        u[3] += dz * commonCoeff;
    }
    template <class R>
    __device__ static void eval(R *u, R dx, R dy, R dz, const R *g) {
        R r2 = dx * dx + dy * dy + dz * dz;
        R rinv = rsqrt_masked(r2);
        eval_staged(u, dx, dy, dz, r2, rinv, g);
    }
};

// stokes_pvel + stokes_doublepvel at one source point (StokesLayerKernel.hpp:44-50 with SL points =
// DL points); the displacement, r2 and rsqrt are shared by the two layers.  Source values:
// first array [fx, fy, fz, TrD], second array D row-major.
struct UkStokesPVelFused {
    static constexpr int KS = 13, KS1 = 4, KL = 11, KT = 4;
    static constexpr unsigned SRC_SCALE_MASK = 0x1FF8u;
    __host__ __device__ static constexpr int deg(int k) { return k == 0 ? 2 : 1; }
    __device__ static void stage(const double *in, double *o) {
        o[0] = in[0]; o[1] = in[1]; o[2] = in[2]; o[3] = in[3];
        UkStokesDoublePVel::stage(in + 4, o + 4);
    }
    template <class R>
    __device__ static void eval(R *u, R dx, R dy, R dz, const R *f) {
        R r2 = dx * dx + dy * dy + dz * dz;
        R rinv = rsqrt_masked(r2);
        R rinv3 = rinv * rinv * rinv;
        R commonCoeff = f[0] * dx + f[1] * dy + f[2] * dz;
        const R two = R(2.0);
        u[0] += two * commonCoeff * rinv3;
        commonCoeff -= f[3];
        u[1] += rinv3 * (r2 * f[0] + dx * commonCoeff);
        u[2] += rinv3 * (r2 * f[1] + dy * commonCoeff);
        u[3] += rinv3 * (r2 * f[2] + dz * commonCoeff);
        UkStokesDoublePVel::eval_staged(u, dx, dy, dz, r2, rinv, f + 4);
    }
};

// stokes_regvel, StokesRegSingleLayerKernel.hpp:23-40; source [fx, fy, fz, eps].
// Staged: fx, fy, fz, eps*eps.  The mask is on the regularised r2 = |d|^2 + eps^2.
struct UkStokesRegVel {
    static constexpr int KS = 4, KS1 = 4, KL = 4, KT = 3;
    static constexpr unsigned SRC_SCALE_MASK = 0x8u;
    __host__ __device__ static constexpr int deg(int) { return 1; }
    __device__ static void stage(const double *in, double *o) {
        o[0] = in[0]; o[1] = in[1]; o[2] = in[2]; o[3] = in[3] * in[3];
    }
    template <class R>
    __device__ static void eval(R *u, R dx, R dy, R dz, const R *f) {
        const R reg2 = f[3];
        R r2 = dx * dx + dy * dy + dz * dz + reg2;
        R rinv = rsqrt_masked(r2);
        R rinv3 = rinv * rinv * rinv;
        R r2reg2 = r2 + reg2;
        R commonCoeff = f[0] * dx + f[1] * dy + f[2] * dz;
        u[0] += rinv3 * (r2reg2 * f[0] + dx * commonCoeff);
        u[1] += rinv3 * (r2reg2 * f[1] + dy * commonCoeff);
        u[2] += rinv3 * (r2reg2 * f[2] + dz * commonCoeff);
    }
};

// rpy_u, RPYKernel.hpp:23-44; source [fx, fy, fz, a].  Staged: fx, fy, fz, one_over_three*a*a.
struct UkRpyU {
    static constexpr int KS = 4, KS1 = 4, KL = 4, KT = 3;
    static constexpr unsigned SRC_SCALE_MASK = 0x8u;
    __host__ __device__ static constexpr int deg(int) { return 1; }
    __device__ static void stage(const double *in, double *o) {
        o[0] = in[0]; o[1] = in[1]; o[2] = in[2]; o[3] = STKP2P_RPY_ONE_OVER_THREE * in[3] * in[3];
    }
    template <class R>
    __device__ static void eval(R *u, R dx, R dy, R dz, const R *f) {
        R r2 = dx * dx + dy * dy + dz * dz;
        R rinv = rsqrt_masked(r2);
        R rinv3 = rinv * rinv * rinv;
        R rinv5 = rinv3 * rinv * rinv;
        const R three = R(3.0);
        R fdotr = f[0] * dx + f[1] * dy + f[2] * dz;
        R three_fdotr_rinv5 = three * fdotr * rinv5;
        const R a2_over_three = f[3];
        u[0] += (r2 * f[0] + dx * fdotr) * rinv3 + a2_over_three * (f[0] * rinv3 - three_fdotr_rinv5 * dx);
        u[1] += (r2 * f[1] + dy * fdotr) * rinv3 + a2_over_three * (f[1] * rinv3 - three_fdotr_rinv5 * dy);
        u[2] += (r2 * f[2] + dz * fdotr) * rinv3 + a2_over_three * (f[2] * rinv3 - three_fdotr_rinv5 * dz);
    }
};

// rpy_ulapu, RPYKernel.hpp:58-89; source [fx, fy, fz, a].  Staged as UkRpyU.
struct UkRpyULapU {
    static constexpr int KS = 4, KS1 = 4, KL = 4, KT = 6;
    static constexpr unsigned SRC_SCALE_MASK = 0x8u;
    __host__ __device__ static constexpr int deg(int k) { return k < 3 ? 1 : 3; }
    __device__ static void stage(const double *in, double *o) { UkRpyU::stage(in, o); }
    template <class R>
    __device__ static void eval(R *u, R dx, R dy, R dz, const R *f) {
        R r2 = dx * dx + dy * dy + dz * dz;
        R rinv = rsqrt_masked(r2);
        R rinv3 = rinv * rinv * rinv;
        R rinv5 = rinv3 * rinv * rinv;
        const R two = R(2.0);
        const R three = R(3.0);
        R fdotr = f[0] * dx + f[1] * dy + f[2] * dz;
        R fdotr_rinv3 = fdotr * rinv3;
        R three_fdotr_rinv5 = three * fdotr * rinv5;
        const R a2_over_three = f[3];
        R cx = f[0] * rinv3 - three_fdotr_rinv5 * dx;
        R cy = f[1] * rinv3 - three_fdotr_rinv5 * dy;
        R cz = f[2] * rinv3 - three_fdotr_rinv5 * dz;
        u[0] += f[0] * rinv + dx * fdotr_rinv3 + a2_over_three * cx;
        u[1] += f[1] * rinv + dy * fdotr_rinv3 + a2_over_three * cy;
        u[2] += f[2] * rinv + dz * fdotr_rinv3 + a2_over_three * cz;
        u[3] += two * cx;
        u[4] += two * cy;
        u[5] += two * cz;
    }
};

// stk_ulapu, RPYKernel.hpp:103-128; source [fx, fy, fz]
struct UkStkULapU {
    static constexpr int KS = 3, KS1 = 3, KL = 3, KT = 6;
    static constexpr unsigned SRC_SCALE_MASK = 0u;
    __host__ __device__ static constexpr int deg(int k) { return k < 3 ? 1 : 3; }
    __device__ static void stage(const double *in, double *o) { o[0] = in[0]; o[1] = in[1]; o[2] = in[2]; }
    template <class R>
    __device__ static void eval(R *u, R dx, R dy, R dz, const R *f) {
        R r2 = dx * dx + dy * dy + dz * dz;
        R rinv = rsqrt_masked(r2);
        R rinv3 = rinv * rinv * rinv;
        R rinv5 = rinv3 * rinv * rinv;
        const R two = R(2.0);
        const R three = R(3.0);
        R fdotr = f[0] * dx + f[1] * dy + f[2] * dz;
        R fdotr_rinv3 = fdotr * rinv3;
        R three_fdotr_rinv5 = three * fdotr * rinv5;
        u[0] += f[0] * rinv + dx * fdotr_rinv3;
        u[1] += f[1] * rinv + dy * fdotr_rinv3;
        u[2] += f[2] * rinv + dz * fdotr_rinv3;
        u[3] += two * (f[0] * rinv3 - three_fdotr_rinv5 * dx);
        u[4] += two * (f[1] * rinv3 - three_fdotr_rinv5 * dy);
        u[5] += two * (f[2] * rinv3 - three_fdotr_rinv5 * dz);
    }
};

// compile-time dispatch: micro-kernel id (include/stkfmm_p2p.h STKFMM_UKER_*) -> struct
template <int UK> struct UkerOf;
template <> struct UkerOf<0> { using type = UkLaplaceP; };
template <> struct UkerOf<1> { using type = UkLaplaceDipoleP; };
template <> struct UkerOf<2> { using type = UkLaplacePGrad; };
template <> struct UkerOf<3> { using type = UkLaplaceDipolePGrad; };
template <> struct UkerOf<4> { using type = UkLaplacePGradGrad; };
template <> struct UkerOf<5> { using type = UkLaplaceDipolePGradGrad; };
template <> struct UkerOf<6> { using type = UkStokesVel; };
template <> struct UkerOf<7> { using type = UkStokesPVel; };
template <> struct UkerOf<8> { using type = UkStokesDoublePVel; };
template <> struct UkerOf<9> { using type = UkStokesRegVel; };
template <> struct UkerOf<10> { using type = UkRpyU; };
template <> struct UkerOf<11> { using type = UkRpyULapU; };
template <> struct UkerOf<12> { using type = UkStkULapU; };
template <> struct UkerOf<13> { using type = UkStokesPVelFused; };

constexpr int kNumUker = 14;

} // namespace stkp2p
