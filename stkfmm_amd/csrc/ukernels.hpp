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
//   stage(in, out)  per-source precomputation in FP64.  Sub-expressions the reference evaluates
//         from source values alone are hoisted here (e.g. sxy + syx, a*a/3), each with the same
//         operands and operation, so hoisting does not change a single rounding; the rewritten hot
//         kernels also stage source-only terms of their rewrite (stated next to their code); a
//         kernel whose FP64 and FP32 forms differ declares STAGE_BY_PRECISION and stage_p<PREC>,
//   eval<R, M>(u, dx, dy, dz, f)  u += K(r, f) / post(k) without the scale factor, r = x_trg -
//         x_src; M = false for pairs known to have r2 > 0 (no coincidence mask),
//   post(k) constant factor applied to accumulated output k once per work item (so that per-pair
//         constants such as the 2 of the pressure leave the inner loop),
//   UNMASKED_EVAL (true here: eval<R, false> exists).
// The hot kernels (Stokeslet, PVel single/double/fused, regularised Stokeslet, RPY, Laplace PGrad)
// evaluate the reference expression in an equivalent form with fewer operations (common factors
// pulled out, constants folded into staged source values or post(), each term accumulated by its
// own fused multiply-add); the Laplace dipole-gradient and gradient-of-gradient kernels are written
// with no power of 1/r above 3 (or s^4 in unit-direction form), so that FP32 stays finite at close
// range.  Each rewrite is stated next to its code; the remaining kernels follow the reference's
// algebraic order directly.
#pragma once

#include <hip/hip_runtime.h>

#ifndef STKP2P_FP64_NEWTON
#define STKP2P_FP64_NEWTON 1
#endif

namespace stkp2p {

// ---- masked reciprocal square root ------------------------------------------------------------
// approx_rsqrt(r2, r2 > 0) of the reference: 0 for coincident points.
// FP64: v_rsq_f64 seed (measured on gfx950: 24.2 correct bits over r2 in [1e-30, 1e3]) refined by
// STKP2P_FP64_NEWTON Newton steps y <- y + y (1 - r2 y^2)/2 (one step: 47.8 bits, max relative
// error 4e-15; two: 52.4 bits), written as e = 0.5 - (r2/2) y^2, y = y + y e.
// MASK: 0 unless r2 > 0 -- a select after the refinement (a NaN r2 must give 0 like the reference's
// mask, and a Newton step turns any seed into NaN when r2 is NaN).  MASK = false drops the select,
// for pairs known to have r2 > 0 (the result is then the same).
template <bool MASK = true>
__device__ __forceinline__ double rsqrt_f64(double r2) {
    double y = __builtin_amdgcn_rsq(r2);
    const double hr = 0.5 * r2;
#pragma unroll
    for (int i = 0; i < STKP2P_FP64_NEWTON; ++i) {
        const double e = __builtin_fma(-hr, y * y, 0.5);
        y = __builtin_fma(y, e, y);
    }
    if constexpr (MASK)
        return (r2 > 0.0) ? y : 0.0;
    else
        return y;
}

// 2/r in FP64: rsqrt_f64's last Newton step without its factor 1/2, y (3 - r2 y^2) -- three
// operations (y^2, the FMA, the product) for its four, for micro-kernels whose staged values absorb
// the powers of 2.  MASK as rsqrt_f64.
template <bool MASK = true>
__device__ __forceinline__ double twice_rsqrt_f64(double r2) {
    static_assert(STKP2P_FP64_NEWTON >= 1, "the 2/r form is a Newton step");
    double y = __builtin_amdgcn_rsq(r2);
    const double hr = 0.5 * r2;
#pragma unroll
    for (int i = 1; i < STKP2P_FP64_NEWTON; ++i) {
        const double e = __builtin_fma(-hr, y * y, 0.5);
        y = __builtin_fma(y, e, y);
    }
    const double w = y * __builtin_fma(-r2, y * y, 3.0);
    if constexpr (MASK)
        return (r2 > 0.0) ? w : 0.0;
    else
        return w;
}

// FP32: v_rsq_f32 seed (measured on gfx950: 23.2 correct bits, i.e. within 1.1e-7 relative) and
// one Newton step.
template <bool MASK = true>
__device__ __forceinline__ float rsqrt_f32(float r2) {
    float y = __builtin_amdgcn_rsqf(r2);
    const float hr = 0.5f * r2;
    const float e = __builtin_fmaf(-hr, y * y, 0.5f);
    y = __builtin_fmaf(y, e, y);
    if constexpr (MASK)
        return (r2 > 0.0f) ? y : 0.0f;
    else
        return y;
}

// 2/r in FP32: rsqrt_f32's Newton step without its factor 1/2, as twice_rsqrt_f64.
template <bool MASK = true>
__device__ __forceinline__ float twice_rsqrt_f32(float r2) {
    const float y = __builtin_amdgcn_rsqf(r2);
    const float w = y * __builtin_fmaf(-r2, y * y, 3.0f);
    if constexpr (MASK)
        return (r2 > 0.0f) ? w : 0.0f;
    else
        return w;
}
template <bool M> __device__ __forceinline__ double twice_rsqrt(double r2) { return twice_rsqrt_f64<M>(r2); }
template <bool M> __device__ __forceinline__ float twice_rsqrt(float r2) { return twice_rsqrt_f32<M>(r2); }

// masked and refined (the adjoint kernels and the rest of the family)
__device__ __forceinline__ double rsqrt_masked(double r2) { return rsqrt_f64<true>(r2); }
__device__ __forceinline__ float rsqrt_masked(float r2) { return rsqrt_f32<true>(r2); }

// 1/r and 1/r^2 of a pair, for the micro-kernels with the eval<R, M> interface: masked unless the
// caller has shown r2 > 0 (M = false); s Newton-refined (FP32 and FP64), s2 = s s.
// (Measured on MI355X: the FP32 seed alone, with s2 corrected to FP32 rounding, is 2 VALU operations
// cheaper per pair and leaves small sums at their accuracy, but its error is not zero-mean: summed
// over the ~14,500 sources per target of the 4M-point RPY U-list it raised errorL2 from 4.2e-7 to
// 1.3e-6.  The Newton step stays.)
template <class R> struct RsqPair { R s, s2; };
template <bool M>
__device__ __forceinline__ RsqPair<double> pair_rsqrt(double r2) {
    const double s = rsqrt_f64<M>(r2);
    return {s, s * s};
}
template <bool M>
__device__ __forceinline__ RsqPair<float> pair_rsqrt(float r2) {
    const float s = rsqrt_f32<M>(r2);
    return {s, s * s};
}

// Micro-kernels that declare UNMASKED_EVAL provide eval<R, M> with M = false for tiles in which
// every pair has r2 > 0 (the item kernel checks this per source tile); the others provide eval<R>.
template <class K, class = void> struct HasUnmaskedEval { static constexpr bool value = false; };
template <class K> struct HasUnmaskedEval<K, decltype((void)K::UNMASKED_EVAL)> {
    static constexpr bool value = K::UNMASKED_EVAL;
};

// stage() of a micro-kernel for the precision its pairs run in: micro-kernels that declare
// STAGE_BY_PRECISION stage their values differently for FP64 (PREC 0) and FP32 (PREC 1) pairs
// (stage_p<PREC>); the others have one stage().
template <class K, class = void> struct StagesByPrecision { static constexpr bool value = false; };
template <class K> struct StagesByPrecision<K, decltype((void)K::STAGE_BY_PRECISION)> {
    static constexpr bool value = K::STAGE_BY_PRECISION;
};
// staged values per source in precision PREC: KL_FP64 for PREC 0 when the micro-kernel declares it
// (its FP64 form stages more values than its FP32 one), else KL
template <class K, int PREC, class = void> struct StagedCount { static constexpr int value = K::KL; };
template <class K> struct StagedCount<K, 0, decltype((void)K::KL_FP64)> {
    static constexpr int value = K::KL_FP64;
};

// constant factor of output k in precision PREC: post_p<PREC>(k) for micro-kernels that declare
// POST_BY_PRECISION (their FP64 form accumulates some outputs at another power of 2), else post(k)
template <class K, class = void> struct PostsByPrecision { static constexpr bool value = false; };
template <class K> struct PostsByPrecision<K, decltype((void)K::POST_BY_PRECISION)> {
    static constexpr bool value = K::POST_BY_PRECISION;
};
template <class K, int PREC>
__host__ __device__ constexpr double post_for(int k) {
    if constexpr (PostsByPrecision<K>::value)
        return K::template post_p<PREC>(k);
    else
        return K::post(k);
}

template <class K, int PREC>
__host__ __device__ inline void stage_for(const double *in, double *o) {
    if constexpr (StagesByPrecision<K>::value)
        K::template stage_p<PREC>(in, o);
    else
        K::stage(in, o);
}

// reference constant one_over_three = 0.3333333333333 (RPYKernel.hpp:30,66)
#define STKP2P_RPY_ONE_OVER_THREE 0.3333333333333

// ---- Laplace ----------------------------------------------------------------------------------

// laplace_p, LaplaceLayerKernel.hpp:58-64
struct UkLaplaceP {
    static constexpr int KS = 1, KS1 = 1, KL = 1, KT = 1;
    static constexpr bool UNMASKED_EVAL = true;
    static constexpr unsigned SRC_SCALE_MASK = 0u;
    __host__ __device__ static constexpr int deg(int) { return 1; }
    __host__ __device__ static constexpr double post(int k) { return ((void)k, 1.0); }
    // staged q/2 against w = 2/r (twice_rsqrt, one operation fewer than 1/r)
    __device__ static void stage(const double *in, double *o) { o[0] = 0.5 * in[0]; }
    template <class R, bool M = true>
    __device__ static void eval(R *u, R dx, R dy, R dz, const R *f) {
        R r2 = dx * dx + dy * dy + dz * dz;
        u[0] += f[0] * twice_rsqrt<M>(r2);
    }
};

// laplace_dipolep, LaplaceLayerKernel.hpp:104-112
struct UkLaplaceDipoleP {
    static constexpr int KS = 3, KS1 = 3, KL = 3, KT = 1;
    static constexpr bool UNMASKED_EVAL = true;
    static constexpr unsigned SRC_SCALE_MASK = 0x7u;
    __host__ __device__ static constexpr int deg(int) { return 1; }
    __host__ __device__ static constexpr double post(int k) { return ((void)k, 1.0); }
    // staged n/8 against w = 2/r (twice_rsqrt, one operation fewer than 1/r): (n/8) w^3 = n/r^3
    __device__ static void stage(const double *in, double *o) {
        o[0] = 0.125 * in[0]; o[1] = 0.125 * in[1]; o[2] = 0.125 * in[2];
    }
    template <class R, bool M = true>
    __device__ static void eval(R *u, R dx, R dy, R dz, const R *f) {
        R r2 = dx * dx + dy * dy + dz * dz;
        const R w = twice_rsqrt<M>(r2);
        R w3 = (w * w) * w;
        R rdotn = f[0] * dx + f[1] * dy + f[2] * dz;
        u[0] += w3 * rdotn;
    }
};

// laplace_pgrad, LaplaceLayerKernel.hpp:84-95: p += q/r, grad p -= q r/r^3.
// Rewrite: the reference's sv*r2 (= q rinv^3 r^2) is q*rinv; the gradient is accumulated with a
// positive sign and negated by post().
struct UkLaplacePGrad {
    static constexpr int KS = 1, KS1 = 1, KL = 1, KT = 4;
    static constexpr bool UNMASKED_EVAL = true;
    static constexpr unsigned SRC_SCALE_MASK = 0u;
    __host__ __device__ static constexpr int deg(int k) { return k == 0 ? 1 : 2; }
    // the gradient sums hold 4 q r / r^3 (below): -1/4 once per target
    __host__ __device__ static constexpr double post(int k) { return k == 0 ? 1.0 : -0.25; }
    __device__ static void stage(const double *in, double *o) { o[0] = 0.5 * in[0]; }
    // w = 2/r (the Newton step without its 1/2, one operation fewer), staged q/2: q/r = (q/2) w and
    // (q/r) w^2 = 4 q/r^3 -- every scaling a power of 2
    template <class R, bool M = true>
    __device__ static void eval(R *u, R dx, R dy, R dz, const R *f) {
        R r2 = dx * dx + dy * dy + dz * dz;
        const R w = twice_rsqrt<M>(r2);
        R qr = f[0] * w;
        R sv = qr * (w * w);
        u[0] += qr;
        u[1] += sv * dx;
        u[2] += sv * dy;
        u[3] += sv * dz;
    }
};

// laplace_dipolepgrad, LaplaceLayerKernel.hpp:121-135; dipole n:
//   p = (r.n)/r^3,  grad p_i = (n_i r^2 - 3 (r.n) r_i)/r^5.
// Rewrite: w = (r.n)/r^2; p = w/r, grad p_i = (n_i - 3 w r_i)/r^3 -- no power above 1/r^3, so the
// FP32 form stays finite down to r ~ 1e-12 (1/r^5 overflows FP32 below r ~ 2e-8).
struct UkLaplaceDipolePGrad {
    static constexpr int KS = 3, KS1 = 3, KL = 3, KT = 4;
    static constexpr bool UNMASKED_EVAL = true;
    static constexpr unsigned SRC_SCALE_MASK = 0x7u;
    __host__ __device__ static constexpr int deg(int k) { return k == 0 ? 1 : 2; }
    __host__ __device__ static constexpr double post(int k) { return ((void)k, 1.0); }
    __device__ static void stage(const double *in, double *o) { o[0] = in[0]; o[1] = in[1]; o[2] = in[2]; }
    template <class R, bool M = true>
    __device__ static void eval(R *u, R dx, R dy, R dz, const R *f) {
        R r2 = dx * dx + dy * dy + dz * dz;
        const auto rs = pair_rsqrt<M>(r2);
        R rinv = rs.s, rinv2 = rs.s2;
        R rinv3 = rinv2 * rinv;
        R w = (f[0] * dx + f[1] * dy + f[2] * dz) * rinv2;
        R m = R(-3.0) * w;
        u[0] += w * rinv;
        u[1] += (m * dx + f[0]) * rinv3;
        u[2] += (m * dy + f[1]) * rinv3;
        u[3] += (m * dz + f[2]) * rinv3;
    }
};

// laplace_pgradgrad, LaplaceLayerKernel.hpp:144-165; charge q:
//   p = q/r,  grad p = -q r/r^3,  grad grad p_ij = q (3 r_i r_j - r^2 delta_ij)/r^5.
// Unit-direction form (s = 1/r, n = r s): grad p = -(q s^3) r,  grad grad p_ij = q s^3 (3 n_i n_j -
// delta_ij) -- no power above s^3 (the literal 1/r^5 overflows FP32 below r ~ 2e-8).
struct UkLaplacePGradGrad {
    static constexpr int KS = 1, KS1 = 1, KL = 1, KT = 10;
    static constexpr bool UNMASKED_EVAL = true;
    static constexpr unsigned SRC_SCALE_MASK = 0u;
    __host__ __device__ static constexpr int deg(int k) { return k == 0 ? 1 : (k < 4 ? 2 : 3); }
    __host__ __device__ static constexpr double post(int k) { return ((void)k, 1.0); }
    __device__ static void stage(const double *in, double *o) { o[0] = in[0]; }
    template <class R, bool M = true>
    __device__ static void eval(R *u, R dx, R dy, R dz, const R *f) {
        R r2 = dx * dx + dy * dy + dz * dz;
        const auto rs = pair_rsqrt<M>(r2);
        R s = rs.s;
        R qs = f[0] * s;
        R qs3 = qs * rs.s2;
        u[0] += qs;
        u[1] -= qs3 * dx;
        u[2] -= qs3 * dy;
        u[3] -= qs3 * dz;
        const R nx = dx * s, ny = dy * s, nz = dz * s;
        const R t3 = R(3.0) * qs3;
        const R ax = t3 * nx, ay = t3 * ny, az = t3 * nz;
        u[4] += ax * nx - qs3;
        u[5] += ax * ny;
        u[6] += ax * nz;
        u[7] += ay * ny - qs3;
        u[8] += ay * nz;
        u[9] += az * nz - qs3;
    }
};

// laplace_dipolepgradgrad, LaplaceLayerKernel.hpp:174-199; dipole f, c = n.f (s = 1/r, n = r s):
//   p = c s^2,  grad p_i = (f_i - 3 c n_i) s^3,
//   grad grad p_ij = (15 c n_i n_j - 3 c delta_ij - 3 (n_i f_j + n_j f_i)) s^4
// (the reference's r-polynomials over r^3, r^5, r^7 divided through by r^k; the literal 1/r^7
// overflows FP32 below r ~ 3e-6, the unit-direction form only below r ~ 1e-9).
struct UkLaplaceDipolePGradGrad {
    static constexpr int KS = 3, KS1 = 3, KL = 3, KT = 10;
    static constexpr bool UNMASKED_EVAL = true;
    static constexpr unsigned SRC_SCALE_MASK = 0x7u;
    __host__ __device__ static constexpr int deg(int k) { return k == 0 ? 1 : (k < 4 ? 2 : 3); }
    __host__ __device__ static constexpr double post(int k) { return ((void)k, 1.0); }
    __device__ static void stage(const double *in, double *o) { o[0] = in[0]; o[1] = in[1]; o[2] = in[2]; }
    template <class R, bool M = true>
    __device__ static void eval(R *u, R dx, R dy, R dz, const R *f) {
        R r2 = dx * dx + dy * dy + dz * dz;
        const auto rs = pair_rsqrt<M>(r2);
        const R s = rs.s, s2 = rs.s2;
        const R n[3] = {dx * s, dy * s, dz * s};
        const R c = f[0] * n[0] + f[1] * n[1] + f[2] * n[2];
        const R s3 = s2 * s;
        const R s4 = s2 * s2;
        const R m3 = R(-3.0) * c;
        const R c15 = R(15.0) * c;
        u[0] += c * s2;
#pragma unroll
        for (int i = 0; i < 3; ++i)
            u[1 + i] += (m3 * n[i] + f[i]) * s3;
        R e[3];
#pragma unroll
        for (int i = 0; i < 3; ++i)
            e[i] = c15 * n[i];
        // output order xx, xy, xz, yy, yz, zz
        u[4] += ((e[0] * n[0] + m3) - R(6.0) * n[0] * f[0]) * s4;
        u[5] += (e[0] * n[1] - R(3.0) * (n[0] * f[1] + n[1] * f[0])) * s4;
        u[6] += (e[0] * n[2] - R(3.0) * (n[0] * f[2] + n[2] * f[0])) * s4;
        u[7] += ((e[1] * n[1] + m3) - R(6.0) * n[1] * f[1]) * s4;
        u[8] += (e[1] * n[2] - R(3.0) * (n[1] * f[2] + n[2] * f[1])) * s4;
        u[9] += ((e[2] * n[2] + m3) - R(6.0) * n[2] * f[2]) * s4;
    }
};

// ---- Stokes -----------------------------------------------------------------------------------

// Stokeslet 3 -> 3 (PVFMM stokes_vel, StokesLayerKernel.hpp:38-42; closed form of
// Test/SimpleKernel.cpp:986-1023 with a = 0): u = f/r + r (f.r)/r^3, evaluated as
// u_i = (1/r) (f_i + r_i h) with h = (f.r)/r^2: one multiply per pair fewer than f_i/r + r_i (f.r)/r^3
// (no 1/r^3), the same two FMAs per component.  FP64 takes w = 2/r (twice_rsqrt, one operation
// fewer) with f staged twice: u_i = w (f_i/2 + r_i h'), h' = ((f/8).r) w^2 = (f.r)/(2 r^2) -- every
// scaling a power of 2; FP32 keeps one staged f (a second would cost an LDS row read per source).
struct UkStokesVel {
    static constexpr int KS = 3, KS1 = 3, KL = 3, KL_FP64 = 6, KT = 3;
    static constexpr bool UNMASKED_EVAL = true, STAGE_BY_PRECISION = true;
    static constexpr unsigned SRC_SCALE_MASK = 0u;
    __host__ __device__ static constexpr int deg(int) { return 1; }
    __host__ __device__ static constexpr double post(int k) { return ((void)k, 1.0); }
    template <int PREC>
    __host__ __device__ static void stage_p(const double *in, double *o) {
        if constexpr (PREC == 0) {
            for (int i = 0; i < 3; ++i) {
                o[i] = 0.5 * in[i];
                o[3 + i] = 0.125 * in[i];
            }
        } else {
            o[0] = in[0]; o[1] = in[1]; o[2] = in[2];
        }
    }
    template <class R, bool M = true>
    __device__ static void eval(R *u, R dx, R dy, R dz, const R *f) {
        R r2 = dx * dx + dy * dy + dz * dz;
        if constexpr (sizeof(R) == 8) {
            const R w = twice_rsqrt_f64<M>(r2);
            R h = (f[3] * dx + f[4] * dy + f[5] * dz) * (w * w);
            u[0] += w * (f[0] + dx * h);
            u[1] += w * (f[1] + dy * h);
            u[2] += w * (f[2] + dz * h);
            return;
        }
        const auto rs = pair_rsqrt<M>(r2);
        R rinv = rs.s;
        R h = (f[0] * dx + f[1] * dy + f[2] * dz) * rs.s2;
        u[0] += rinv * (f[0] + dx * h);
        u[1] += rinv * (f[1] + dy * h);
        u[2] += rinv * (f[2] + dz * h);
    }
};

// stokes_pvel, StokesSingleLayerKernel.hpp:33-46; source [fx, fy, fz, TrD].
//   p = 2 (f.r)/r^3,  u_i = (r^2 f_i + r_i (f.r - TrD))/r^3 = f_i/r + r_i (f.r - TrD)/r^3.
// Rewrite: r^2/r^3 = 1/r; the factor 2 of p is applied by post().
struct UkStokesPVel {
    static constexpr int KS = 4, KS1 = 4, KL = 4, KT = 4;
    static constexpr bool UNMASKED_EVAL = true;
    static constexpr bool MIXED_PEEL = true; // item kernel: first source per block on its own (measured)
    static constexpr unsigned SRC_SCALE_MASK = 0x8u;
    __host__ __device__ static constexpr int deg(int k) { return k == 0 ? 2 : 1; }
    __host__ __device__ static constexpr double post(int k) { return k == 0 ? 2.0 : 1.0; }
    __device__ static void stage(const double *in, double *o) {
        o[0] = in[0]; o[1] = in[1]; o[2] = in[2]; o[3] = in[3];
    }
    template <class R, bool M = true>
    __device__ static void eval(R *u, R dx, R dy, R dz, const R *f) {
        R r2 = dx * dx + dy * dy + dz * dz;
        const auto rs = pair_rsqrt<M>(r2);
        R rinv = rs.s;
        R rinv3 = rs.s2 * rinv;
        R cc = f[0] * dx + f[1] * dy + f[2] * dz;
        u[0] += cc * rinv3;
        R b = (cc - f[3]) * rinv3;
        u[1] += f[0] * rinv;
        u[1] += dx * b;
        u[2] += f[1] * rinv;
        u[2] += dy * b;
        u[3] += f[2] * rinv;
        u[3] += dz * b;
    }
};

// stokes_doublepvel, StokesDoubleLayerKernel.hpp:31-56; source D row-major [Dxx..Dzz].
//   Q = r.D.r,  p = 2 (tr D/r^3 - 3 Q/r^5),  u_i = -3 r_i Q/r^5.
// The reference's source-only sub-expressions (:46-48,51) with its factor -3 (:49) folded in: g0..g2 =
// -3 Dxx, -3 Dyy, -3 Dzz, g3..g5 = -3 (Dxy+Dyx), -3 (Dxz+Dzx), -3 (Dyz+Dzy), and tr D.
// Rewrite: t = -3 Q/r^2 (no power above 1/r^3), p/2 = (t + tr D)/r^3, u_i = r_i t/r^3; post() doubles
// p.  -3 Q by precision:
//   FP32: Horner in dx, dy, dz reusing dz^2 of r^2; staged g0..g5, tr D.
//   FP64: with z^2 = r^2 - x^2 - y^2, -3 Q = g2 r^2 + rest, rest = x ((g0 - g2) x + g3 y + g4 z) +
//         y ((g1 - g2) y + g5 z), so t = g2 + rest/r^2 -- one operation fewer; staged g0 - g2,
//         g1 - g2, g2, g3, g4, g5, tr D.  (In FP32 the split adds rounding of size |g2| to t, measured
//         up to 3x the Horner form's error on clustered U-lists, so FP32 keeps the Horner form.)
struct UkStokesDoublePVel {
    static constexpr int KS = 9, KS1 = 9, KL = 7, KT = 4;
    static constexpr bool UNMASKED_EVAL = true, STAGE_BY_PRECISION = true;
    static constexpr unsigned SRC_SCALE_MASK = 0x1FFu;
    __host__ __device__ static constexpr int deg(int k) { return k == 0 ? 2 : 1; }
    __host__ __device__ static constexpr double post(int k) { return k == 0 ? 2.0 : 1.0; }
    template <int PREC>
    __host__ __device__ static void stage_p(const double *in, double *o) {
        const double g0 = -3.0 * in[0], g1 = -3.0 * in[4], g2 = -3.0 * in[8];
        o[0] = (PREC == 0) ? g0 - g2 : g0;
        o[1] = (PREC == 0) ? g1 - g2 : g1;
        o[2] = g2;
        o[3] = -3.0 * (in[1] + in[3]);
        o[4] = -3.0 * (in[2] + in[6]);
        o[5] = -3.0 * (in[5] + in[7]);
        o[6] = in[0] + in[4] + in[8];
    }
    // -3 r.D.r / r^2 from the first six staged values (the layout of stage_p for R's precision), z2 = dz^2
    template <class R>
    __device__ static R t_of(R dx, R dy, R dz, R z2, R rinv2, const R *g) {
        if constexpr (sizeof(R) == 8) {
            const R a1 = g[0] * dx + g[3] * dy + g[4] * dz;
            const R a2 = g[1] * dy + g[5] * dz;
            return (dx * a1 + dy * a2) * rinv2 + g[2];
        } else {
            const R a1 = g[0] * dx + g[3] * dy + g[4] * dz;
            const R a2 = g[1] * dy + g[5] * dz;
            return (dx * a1 + (dy * a2 + g[2] * z2)) * rinv2;
        }
    }
    template <class R, bool M = true>
    __device__ static void eval(R *u, R dx, R dy, R dz, const R *g) {
        const R z2 = dz * dz;
        R r2 = dx * dx + (dy * dy + z2);
        const auto rs = pair_rsqrt<M>(r2);
        R rinv = rs.s, rinv2 = rs.s2;
        R rinv3 = rinv2 * rinv;
        R t = t_of(dx, dy, dz, z2, rinv2, g);
        R q = t * rinv3;
        u[0] += (t + g[6]) * rinv3;
        u[1] += dx * q;
        u[2] += dy * q;
        u[3] += dz * q;
    }
};

// stokes_pvel + stokes_doublepvel at one source point (StokesLayerKernel.hpp:44-50 with SL points =
// DL points): displacement, r2 and rsqrt are shared by the two layers.  Source values: first array
// [fx, fy, fz, TrD], second array D row-major.  With cc = f.r and Q' = -3 Q (Q = r.D.r):
//   p/2 = (cc + tr D)/r^3 + Q'/r^5,   u_i = f_i/r + r_i ((cc - TrD)/r^3 + Q'/r^5);  post() doubles p.
// Rewrite (one factor 1/r^3 pulled out of both layers):
//   t = (cc - TrD) + Q'/r^2,   u_i = f_i/r + r_i (t/r^3),   p/2 = (t + (TrD + tr D))/r^3,
// with cc - TrD formed as one fused multiply-add chain started from the staged constant.  Q'/r^2 as
// in UkStokesDoublePVel, by precision:
//   FP32: Q' by Horner reusing dz^2; staged fx, fy, fz, -TrD, g0..g5, TrD + tr D.
//   FP64: Q'/r^2 = g2 + rest/r^2 with g2 folded into the chain's constant, and w = 2/r from the
//         Newton step without its 1/2 (twice_rsqrt_f64, one operation fewer): with 1/r = w/2,
//         1/r^2 = w^2/4 and 1/r^3 = w^3/8, t/8 = (f/8).r + (-TrD + g2)/8 + rest/32 w^2, and then
//         t/r^3 = (t/8) w^3, p/2 = (t/8 + (TrD + tr D)/8) w^3, f_i/r = (f_i/2) w -- f staged twice
//         (f/8 for the dot product, f/2 for f/r), every scaling a power of 2 (exact); staged f/8,
//         (-TrD + g2)/8, (g0 - g2, g1 - g2, g3, g4, g5)/32, (TrD + tr D)/8, f/2 (KL_FP64 = 13).
// FP32: 26 operations per pair besides the displacement and the reciprocal square root; FP64: 25
// and a cheaper reciprocal square root; the two layers' expressions side by side: 30.
struct UkStokesPVelFused {
    static constexpr int KS = 13, KS1 = 4, KL = 11, KL_FP64 = 13, KT = 4;
    static constexpr bool UNMASKED_EVAL = true, STAGE_BY_PRECISION = true;
    static constexpr int MIXED_UNROLL = 1; // item kernel's mixed-mode source loop (measured fastest)
    static constexpr bool MIXED_PEEL = true; // and its first source per block on its own (measured)
    static constexpr unsigned SRC_SCALE_MASK = 0x1FF8u;
    __host__ __device__ static constexpr int deg(int k) { return k == 0 ? 2 : 1; }
    __host__ __device__ static constexpr double post(int k) { return k == 0 ? 2.0 : 1.0; }
    template <int PREC>
    __host__ __device__ static void stage_p(const double *in, double *o) {
        double d[7];
        UkStokesDoublePVel::stage_p<PREC>(in + 4, d);
        o[0] = in[0]; o[1] = in[1]; o[2] = in[2];
        if constexpr (PREC == 0) {
            o[0] = 0.125 * in[0]; o[1] = 0.125 * in[1]; o[2] = 0.125 * in[2];
            o[3] = 0.125 * (-in[3] + d[2]);
            o[4] = 0.03125 * d[0]; o[5] = 0.03125 * d[1]; o[6] = 0.03125 * d[3]; o[7] = 0.03125 * d[4];
            o[8] = 0.03125 * d[5];
            o[9] = 0.125 * (in[3] + d[6]);
            o[10] = 0.5 * in[0]; o[11] = 0.5 * in[1]; o[12] = 0.5 * in[2];
        } else {
            o[3] = -in[3];
            for (int k = 0; k < 6; ++k)
                o[4 + k] = d[k];
            o[10] = in[3] + d[6];
        }
    }
    template <class R, bool M = true>
    __device__ static void eval(R *u, R dx, R dy, R dz, const R *f) {
        const R z2 = dz * dz;
        R r2 = dx * dx + (dy * dy + z2);
        R cm = f[0] * dx + (f[1] * dy + (f[2] * dz + f[3]));
        if constexpr (sizeof(R) == 8) {
            const R w = twice_rsqrt_f64<M>(r2);
            const R w2 = w * w;
            const R w3 = w2 * w;
            const R a1 = f[4] * dx + f[6] * dy + f[7] * dz;
            const R a2 = f[5] * dy + f[8] * dz;
            const R t8 = (dx * a1 + dy * a2) * w2 + cm; // t/8
            const R b = t8 * w3;                         // t/r^3
            u[0] += (t8 + f[9]) * w3;
            u[1] += f[10] * w;
            u[1] += dx * b;
            u[2] += f[11] * w;
            u[2] += dy * b;
            u[3] += f[12] * w;
            u[3] += dz * b;
            return;
        }
        const auto rs = pair_rsqrt<M>(r2);
        R rinv = rs.s, rinv2 = rs.s2;
        R rinv3 = rinv2 * rinv;
        R t, G;
        {
            const R a1 = f[4] * dx + f[7] * dy + f[8] * dz;
            const R a2 = f[5] * dy + f[9] * dz;
            t = (dx * a1 + (dy * a2 + f[6] * z2)) * rinv2 + cm;
            G = f[10];
        }
        R b = t * rinv3;
        u[0] += (t + G) * rinv3;
        u[1] += f[0] * rinv;
        u[1] += dx * b;
        u[2] += f[1] * rinv;
        u[2] += dy * b;
        u[3] += f[2] * rinv;
        u[3] += dz * b;
    }
};

// stokes_regvel, StokesRegSingleLayerKernel.hpp:23-40; source [fx, fy, fz, eps].
//   rho2 = |d|^2 + eps^2,  u_i = ((rho2 + eps^2) f_i + d_i (f.d))/rho^3.
// Staged: fx, fy, fz, eps*eps.  The mask is on the regularised rho2.
struct UkStokesRegVel {
    static constexpr int KS = 4, KS1 = 4, KL = 4, KT = 3;
    static constexpr bool UNMASKED_EVAL = true;
    static constexpr unsigned SRC_SCALE_MASK = 0x8u;
    __host__ __device__ static constexpr int deg(int) { return 1; }
    __host__ __device__ static constexpr double post(int k) { return ((void)k, 1.0); }
    __device__ static void stage(const double *in, double *o) {
        o[0] = in[0]; o[1] = in[1]; o[2] = in[2]; o[3] = in[3] * in[3];
    }
    template <class R, bool M = true>
    __device__ static void eval(R *u, R dx, R dy, R dz, const R *f) {
        const R reg2 = f[3];
        R r2 = dx * dx + dy * dy + dz * dz + reg2;
        const auto rs = pair_rsqrt<M>(r2);
        R rinv3 = rs.s2 * rs.s;
        R a = (r2 + reg2) * rinv3;
        R b = (f[0] * dx + f[1] * dy + f[2] * dz) * rinv3;
        u[0] += f[0] * a;
        u[0] += dx * b;
        u[1] += f[1] * a;
        u[1] += dy * b;
        u[2] += f[2] * a;
        u[2] += dz * b;
    }
};

// rpy_u, RPYKernel.hpp:23-44; source [fx, fy, fz, a].
//   u_i = f_i/r + r_i (f.r)/r^3 + (a^2/3) (f_i/r^3 - 3 (f.r) r_i/r^5)
//       = f_i (1/r + beta/r^3) + r_i ((f.r)/r^3 - 3 beta (f.r)/r^5),  beta = one_over_three a^2.
// Staged: fx, fy, fz, beta, -3 beta.
struct UkRpyU {
    static constexpr int KS = 4, KS1 = 4, KL = 5, KL_FP64 = 8, KT = 3;
    static constexpr bool UNMASKED_EVAL = true, STAGE_BY_PRECISION = true;
    static constexpr unsigned SRC_SCALE_MASK = 0x8u;
    __host__ __device__ static constexpr int deg(int) { return 1; }
    __host__ __device__ static constexpr double post(int k) { return ((void)k, 1.0); }
    // FP32: f, c = a^2/3, -3c.  FP64 (w = 2/r from twice_rsqrt, one operation fewer than 1/r): f/2,
    // f/8, c/4, -3c/4 -- (f/8).r w^3 = (f.r)/r^3, (c/4) w^3 + w = 2 (c/r^3 + 1/r), so f/2 times it is
    // f (c/r^3 + 1/r), and (-3c/4) g3 w^2 = -3c g3/r^2; every scaling a power of 2
    template <int PREC>
    __host__ __device__ static void stage_p(const double *in, double *o) {
        const double c = STKP2P_RPY_ONE_OVER_THREE * in[3] * in[3];
        if constexpr (PREC == 0) {
            for (int i = 0; i < 3; ++i) {
                o[i] = 0.5 * in[i];
                o[3 + i] = 0.125 * in[i];
            }
            o[6] = 0.25 * c;
            o[7] = 0.25 * (-3.0 * c);
        } else {
            o[0] = in[0]; o[1] = in[1]; o[2] = in[2];
            o[3] = c;
            o[4] = -3.0 * c;
        }
    }
    template <class R, bool M = true>
    __device__ static void eval(R *u, R dx, R dy, R dz, const R *f) {
        R r2 = dx * dx + dy * dy + dz * dz;
        if constexpr (sizeof(R) == 8) {
            const R w = twice_rsqrt_f64<M>(r2);
            const R w2 = w * w;
            const R w3 = w2 * w;
            const R g3 = (f[3] * dx + f[4] * dy + f[5] * dz) * w3;
            const R A2 = f[6] * w3 + w; // 2 A
            const R B = f[7] * (g3 * w2) + g3;
            u[0] += f[0] * A2;
            u[0] += dx * B;
            u[1] += f[1] * A2;
            u[1] += dy * B;
            u[2] += f[2] * A2;
            u[2] += dz * B;
            return;
        }
        const auto rs = pair_rsqrt<M>(r2);
        R rinv = rs.s, rinv2 = rs.s2;
        R rinv3 = rinv2 * rinv;
        R fr = f[0] * dx + f[1] * dy + f[2] * dz;
        R g3 = fr * rinv3;
        R A = f[3] * rinv3 + rinv;
        R B = f[4] * (g3 * rinv2) + g3;
        u[0] += f[0] * A;
        u[0] += dx * B;
        u[1] += f[1] * A;
        u[1] += dy * B;
        u[2] += f[2] * A;
        u[2] += dz * B;
    }
};

// rpy_ulapu, RPYKernel.hpp:58-89; source [fx, fy, fz, a].  With c_i = f_i/r^3 - 3 (f.r) r_i/r^5:
//   u_i = f_i/r + r_i (f.r)/r^3 + beta c_i,  lap u_i = 2 c_i  (beta = one_over_three a^2)
// evaluated as u_i = f_i (1/r + beta/r^3) + r_i (g3 - 3 beta g5) and lap u_i = -6 (f'_i/r^3 + g5 r_i)
// with f' = -f/3 staged (g3 = (f.r)/r^3, g5 = (f.r)/r^5; post() applies the -6).  Staged as UkRpyU,
// then -f/3 (the three extra values fill the 16-byte LDS row that holds the first five).
struct UkRpyULapU {
    static constexpr int KS = 4, KS1 = 4, KL = 8, KL_FP64 = 11, KT = 6;
    static constexpr bool UNMASKED_EVAL = true, STAGE_BY_PRECISION = true, POST_BY_PRECISION = true;
    static constexpr int MIXED_MIN_WAVES = 4; // item kernel's mixed-mode occupancy request (measured)
    // FP32 partial sums over 32 sources (the form is still chosen per 16): the FP64 flush of the six
    // outputs costs less; 4M RPY U-list, same box, twice each: 40.68 -> 40.29 ms, errorL2 4.24e-7 ->
    // 3.16e-7, maxrel 3.38e-7 -> 3.10e-7
    static constexpr int MIXED_ACC_SOURCES = 32;
    static constexpr bool MIXED_PEEL = true; // item kernel: first source per block on its own (measured)
    static constexpr bool MIXED_BLOCK_UNROLL = true; // and its blocks fully unrolled (measured)
    static constexpr unsigned SRC_SCALE_MASK = 0x8u;
    __host__ __device__ static constexpr int deg(int k) { return k < 3 ? 1 : 3; }
    __host__ __device__ static constexpr double post(int k) { return k < 3 ? 1.0 : -6.0; }
    // the FP64 form accumulates the Laplacian at 4 times its value (below): -6/4 there
    template <int PREC>
    __host__ __device__ static constexpr double post_p(int k) {
        return k < 3 ? 1.0 : (PREC == 0 ? -1.5 : -6.0);
    }
    // staged as UkRpyU in each precision, then f' = -f/3 (FP64: f'/2)
    template <int PREC>
    __host__ __device__ static void stage_p(const double *in, double *o) {
        UkRpyU::stage_p<PREC>(in, o);
        constexpr int O = PREC == 0 ? 8 : 5;
        constexpr double S = PREC == 0 ? 0.5 : 1.0;
        for (int i = 0; i < 3; ++i)
            o[O + i] = S * (in[i] * (-1.0 / 3.0));
    }
    template <class R, bool M = true>
    __device__ static void eval(R *u, R dx, R dy, R dz, const R *f) {
        R r2 = dx * dx + dy * dy + dz * dz;
        if constexpr (sizeof(R) == 8) {
            // UkRpyU's FP64 form; g5' = g3 w^2 = 4 g5, and (f'/2) w^3 = 4 f'/r^3: u3..5 hold 4 times
            // their value (post_p<0>)
            const R w = twice_rsqrt_f64<M>(r2);
            const R w2 = w * w;
            const R w3 = w2 * w;
            const R g3 = (f[3] * dx + f[4] * dy + f[5] * dz) * w3;
            const R g5q = g3 * w2;
            const R A2 = f[6] * w3 + w;
            const R B = f[7] * g5q + g3;
            u[0] += f[0] * A2;
            u[0] += dx * B;
            u[1] += f[1] * A2;
            u[1] += dy * B;
            u[2] += f[2] * A2;
            u[2] += dz * B;
            u[3] += f[8] * w3;
            u[3] += dx * g5q;
            u[4] += f[9] * w3;
            u[4] += dy * g5q;
            u[5] += f[10] * w3;
            u[5] += dz * g5q;
            return;
        }
        const auto rs = pair_rsqrt<M>(r2);
        R rinv = rs.s, rinv2 = rs.s2;
        R rinv3 = rinv2 * rinv;
        R fr = f[0] * dx + f[1] * dy + f[2] * dz;
        R g3 = fr * rinv3;
        R g5 = g3 * rinv2;
        R A = f[3] * rinv3 + rinv;
        R B = f[4] * g5 + g3;
        u[0] += f[0] * A;
        u[0] += dx * B;
        u[1] += f[1] * A;
        u[1] += dy * B;
        u[2] += f[2] * A;
        u[2] += dz * B;
        u[3] += f[5] * rinv3;
        u[3] += dx * g5;
        u[4] += f[6] * rinv3;
        u[4] += dy * g5;
        u[5] += f[7] * rinv3;
        u[5] += dz * g5;
    }
};

// stk_ulapu, RPYKernel.hpp:103-128; source [fx, fy, fz]: rpy_ulapu with beta = 0.
struct UkStkULapU {
    static constexpr int KS = 3, KS1 = 3, KL = 3, KT = 6;
    static constexpr bool UNMASKED_EVAL = true;
    static constexpr unsigned SRC_SCALE_MASK = 0u;
    __host__ __device__ static constexpr int deg(int k) { return k < 3 ? 1 : 3; }
    __host__ __device__ static constexpr double post(int k) { return k < 3 ? 1.0 : 2.0; }
    __device__ static void stage(const double *in, double *o) { o[0] = in[0]; o[1] = in[1]; o[2] = in[2]; }
    template <class R, bool M = true>
    __device__ static void eval(R *u, R dx, R dy, R dz, const R *f) {
        R r2 = dx * dx + dy * dy + dz * dz;
        const auto rs = pair_rsqrt<M>(r2);
        R rinv = rs.s, rinv2 = rs.s2;
        R rinv3 = rinv2 * rinv;
        R fr = f[0] * dx + f[1] * dy + f[2] * dz;
        R g3 = fr * rinv3;
        R m = R(-3.0) * (g3 * rinv2);
        u[0] += f[0] * rinv;
        u[0] += dx * g3;
        u[1] += f[1] * rinv;
        u[1] += dy * g3;
        u[2] += f[2] * rinv;
        u[2] += dz * g3;
        u[3] += f[0] * rinv3;
        u[3] += dx * m;
        u[4] += f[1] * rinv3;
        u[4] += dy * m;
        u[5] += f[2] * rinv3;
        u[5] += dz * m;
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

constexpr int kNumUker = 14; // micro-kernels defined in this header (0..13)

} // namespace stkp2p
