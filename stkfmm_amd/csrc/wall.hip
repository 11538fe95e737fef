// wall.hip -- the no-slip wall image system of StkWallFMM (Lib/src/StkWallFMM.cpp:154-292) on top of
// the direct-sum P2P: per-source preparation of the image sources and of the source values of each
// constituent sum, and the per-target assembly of the constituent outputs.
//
// User coordinates: the wall is the plane z = 0; the image of a source at (x, y, z) is (x, y, -z)
// (the reference's setPoints, :40-80, shifts z to [0.5, 1) and places the image at 1 - z); y3 and x3
// of the reference are the source and target heights z.  Sources are stored [originals; images] as
// in srcSLCoordInternal.  Stokes (source f(3) -> u(3)), :154-203:
//   S   = Stokeslet of (f1, f2, 0) at originals and (-f1, -f2, 0) at images
//   L1D = LapPGrad of charges -f3/2 (originals), +f3/2 (images) and dipoles (-z f1, -z f2, z f3) at
//         the images
//   L2  = LapPGradGrad of charges f3 z (originals), -f3 z (images)
//   u_j = S_j + L2_grad_j / 2 + x3 L1D_grad_j,   u_3 -= L1D_p
// RPY (source [f(3), b] -> [u(3), lap u(3)]), :205-292: see the assembly below.
#include <hip/hip_runtime.h>

#include "p2p_internal.hpp"

namespace stkp2p {

// every entry of the source-side arrays is written (zeros included), so the workspace needs no
// clearing; layouts of WallLayout (p2p_internal.hpp)
__global__ void __launch_bounds__(256) wall_prep_stokes_kernel(int n, const double *sc, const double *sv,
                                                              WallStokes w) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n)
        return;
    const double x = sc[3LL * i], y = sc[3LL * i + 1], z = sc[3LL * i + 2];
    const double f0 = sv[3LL * i], f1 = sv[3LL * i + 1], f2 = sv[3LL * i + 2];
    const long long o = i, m = (long long)i + n;
    w.c2[3 * o] = x;
    w.c2[3 * o + 1] = y;
    w.c2[3 * o + 2] = z;
    w.c2[3 * m] = x;
    w.c2[3 * m + 1] = y;
    w.c2[3 * m + 2] = -z;
    w.vS[3 * o] = f0;
    w.vS[3 * o + 1] = f1;
    w.vS[3 * o + 2] = 0.0;
    w.vS[3 * m] = -f0;
    w.vS[3 * m + 1] = -f1;
    w.vS[3 * m + 2] = 0.0;
    w.vL1[o] = -0.5 * f2;
    w.vL1[m] = 0.5 * f2;
    w.vD[3 * o] = -z * f0;
    w.vD[3 * o + 1] = -z * f1;
    w.vD[3 * o + 2] = z * f2;
    w.vL2[o] = f2 * z;
    w.vL2[m] = -(f2 * z);
}

__global__ void __launch_bounds__(256) wall_prep_rpy_kernel(int n, const double *sc, const double *sv, WallRpy w) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n)
        return;
    const double x = sc[3LL * i], y = sc[3LL * i + 1], z = sc[3LL * i + 2];
    const double f0 = sv[4LL * i], f1 = sv[4LL * i + 1], f2 = sv[4LL * i + 2], b = sv[4LL * i + 3];
    const double b2 = b * b, twob2 = 2 * b * b;
    const long long o = i, m = (long long)i + n;
    w.c2[3 * o] = x;
    w.c2[3 * o + 1] = y;
    w.c2[3 * o + 2] = z;
    w.c2[3 * m] = x;
    w.c2[3 * m + 1] = y;
    w.c2[3 * m + 2] = -z;
    w.vR[4 * o] = f0;
    w.vR[4 * o + 1] = f1;
    w.vR[4 * o + 2] = 0.0;
    w.vR[4 * o + 3] = b;
    w.vR[4 * m] = -f0;
    w.vR[4 * m + 1] = -f1;
    w.vR[4 * m + 2] = 0.0;
    w.vR[4 * m + 3] = b;
    w.vLS[o] = f2 * (-0.5);
    w.vLS[m] = -f2 * (-0.5);
    w.vLD[3 * o] = -z * f0;
    w.vLD[3 * o + 1] = -z * f1;
    w.vLD[3 * o + 2] = z * f2;
    w.vLSZ[o] = z * f2 * 0.5;
    w.vLSZ[m] = -z * f2 * 0.5;
    const double dz = b2 * f2 * (1. / 6.);
    w.vLDZ[3 * o] = 0.0;
    w.vLDZ[3 * o + 1] = 0.0;
    w.vLDZ[3 * o + 2] = dz;
    w.vLDZ[3 * m] = 0.0;
    w.vLDZ[3 * m + 1] = 0.0;
    w.vLDZ[3 * m + 2] = dz;
    double *q = w.vQ + 9 * o;
    const double q0 = twob2 * f2 * (1. / 6.);
    q[0] = q0;
    q[1] = 0.0;
    q[2] = 0.0;
    q[3] = 0.0;
    q[4] = q0;
    q[5] = 0.0;
    q[6] = twob2 * f0 * (1. / 6.);
    q[7] = twob2 * f1 * (1. / 6.);
    q[8] = 0.0;
}

__global__ void __launch_bounds__(256) wall_zero_kernel(double *p, long long n) {
    const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n)
        p[i] = 0.0;
}

// trg_value += assembled result (StkWallFMM::evaluateFMM, :103-128)
__global__ void __launch_bounds__(256) wall_assemble_stokes_kernel(int m, const double *tc, WallStokes w, double *out) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= m)
        return;
    const double x3 = tc[3LL * i + 2];
    const double *S = w.oS + 3LL * i, *L1D = w.oL1D + 4LL * i, *L2 = w.oL2 + 10LL * i;
    double u[3];
#pragma unroll
    for (int j = 0; j < 3; ++j)
        u[j] = S[j] + 0.5 * L2[j + 1] + x3 * L1D[j + 1];
    u[2] -= L1D[0];
#pragma unroll
    for (int j = 0; j < 3; ++j)
        out[3LL * i + j] += u[j];
}

// StkWallFMM::evalRPY assembly, :270-291
__global__ void __launch_bounds__(256) wall_assemble_rpy_kernel(int m, const double *tc, WallRpy w, double *out) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= m)
        return;
    const double x3 = tc[3LL * i + 2];
    const double *R = w.oR + 6LL * i, *SD = w.oSD + 10LL * i, *SDZ = w.oSDZ + 4LL * i, *Q = w.oQ + 10LL * i;
    double u[6];
    u[0] = R[0] + SDZ[1] + x3 * SD[1] + x3 * Q[1];
    u[1] = R[1] + SDZ[2] + x3 * SD[2] + x3 * Q[2];
    u[2] = R[2] + SDZ[3] + x3 * SD[3] + x3 * Q[3] - SD[0] - Q[0];
    u[3] = R[3] + 2 * SD[6] + 2 * Q[6];
    u[4] = R[4] + 2 * SD[8] + 2 * Q[8];
    u[5] = R[5] + 2 * SD[9] + 2 * Q[9];
#pragma unroll
    for (int j = 0; j < 6; ++j)
        out[6LL * i + j] += u[j];
}

static unsigned blocks(long long n) { return (unsigned)((n + 255) / 256); }

hipError_t launch_wall_prep_stokes(int n, const double *sc, const double *sv, const WallStokes &w, hipStream_t s) {
    if (n > 0)
        hipLaunchKernelGGL(wall_prep_stokes_kernel, dim3(blocks(n)), dim3(256), 0, s, n, sc, sv, w);
    return hipGetLastError();
}

hipError_t launch_wall_prep_rpy(int n, const double *sc, const double *sv, const WallRpy &w, hipStream_t s) {
    if (n > 0)
        hipLaunchKernelGGL(wall_prep_rpy_kernel, dim3(blocks(n)), dim3(256), 0, s, n, sc, sv, w);
    return hipGetLastError();
}

hipError_t launch_wall_zero(double *p, long long n, hipStream_t s) {
    if (n > 0)
        hipLaunchKernelGGL(wall_zero_kernel, dim3(blocks(n)), dim3(256), 0, s, p, n);
    return hipGetLastError();
}

hipError_t launch_wall_assemble_stokes(int m, const double *tc, const WallStokes &w, double *out, hipStream_t s) {
    if (m > 0)
        hipLaunchKernelGGL(wall_assemble_stokes_kernel, dim3(blocks(m)), dim3(256), 0, s, m, tc, w, out);
    return hipGetLastError();
}

hipError_t launch_wall_assemble_rpy(int m, const double *tc, const WallRpy &w, double *out, hipStream_t s) {
    if (m > 0)
        hipLaunchKernelGGL(wall_assemble_rpy_kernel, dim3(blocks(m)), dim3(256), 0, s, m, tc, w, out);
    return hipGetLastError();
}

} // namespace stkp2p
