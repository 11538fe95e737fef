// box.hip -- the box mapping and the per-kernel value scaling that STKFMM wraps around an FMM
// evaluation, as host and device entries of the C ABI:
//   stkfmm_scale_coord  STKFMM::scaleCoord + wrapCoord (Lib/src/STKFMM.cpp:113-150):
//                       x' = (x - origin) * scaleFactor, then x' - floor(x') on the periodic axes
//   stkfmm_scale_src    FMMData::scaleSrc (Lib/src/FMMData.cpp:285-317): double-layer values, and the
//                       single-layer components with the dimension of a length, times scaleFactor
//   stkfmm_scale_trg    FMMData::scaleTrg (Lib/src/FMMData.cpp:319-470): each output component times
//                       the power of scaleFactor of its dimension
// A caller that hands a scaled tree's near field to the plans (setBox's scaleFactor = 1/len) runs
// scale_src before and scale_trg after, as FMMData::evaluateFMM does (:196-200).  Each component is
// multiplied by the same double the reference multiplies by (its products, e.g. sF * sF, formed in
// the same order), so host and device results equal the reference's bit for bit.
#include <hip/hip_runtime.h>

#include <cmath>

#include "../../include/stkfmm_p2p.h"
#include "p2p_internal.hpp"

namespace stkp2p {
namespace {

constexpr int kMaxComp = 16;
struct Mult {
    double m[kMaxComp];
};

// per-component multipliers of scaleTrg for a kernel; kt = number of components (0: unknown kernel)
int trg_mult(unsigned kernel, double sF, Mult *M) {
    const double s1 = sF, s2 = sF * sF, s3 = sF * sF * sF;
    auto fill = [&](int b, int e, double v) {
        for (int k = b; k < e; ++k)
            M->m[k] = v;
    };
    switch (kernel) {
    case STKFMM_KERNEL_PVel: fill(0, 1, s2); fill(1, 4, s1); return 4;
    case STKFMM_KERNEL_PVelGrad: fill(0, 1, s2); fill(1, 4, s1); fill(4, 7, s3); fill(7, 16, s2); return 16;
    case STKFMM_KERNEL_Traction: fill(0, 9, s2); return 9;
    case STKFMM_KERNEL_PVelLaplacian: fill(0, 1, s2); fill(1, 4, s1); fill(4, 7, s3); return 7;
    case STKFMM_KERNEL_LapPGrad: fill(0, 1, s1); fill(1, 4, s2); return 4;
    case STKFMM_KERNEL_LapPGradGrad: fill(0, 1, s1); fill(1, 4, s2); fill(4, 10, s3); return 10;
    case STKFMM_KERNEL_LapQPGradGrad: {
        const double sf3 = sF * sF * sF, sf4 = sF * sf3, sf5 = sF * sf4;
        fill(0, 1, sf3);
        fill(1, 4, sf4);
        fill(4, 10, sf5);
        return 10;
    }
    case STKFMM_KERNEL_Stokes: fill(0, 3, s1); return 3;
    case STKFMM_KERNEL_StokesRegVel: fill(0, 3, s1); return 3;
    case STKFMM_KERNEL_StokesRegVelOmega: fill(0, 3, s1); fill(3, 6, s2); return 6;
    case STKFMM_KERNEL_RPY: fill(0, 3, s1); fill(3, 6, s3); return 6;
    default: return 0;
    }
}

// scaleSrc: single-layer multipliers (ks components), double layer all times sF
int src_mult(unsigned kernel, double sF, Mult *M) {
    int dims[3];
    if (stkfmm_get_kernel_dimension(kernel, dims) != STKFMM_OK)
        return 0;
    const int ks = dims[0];
    for (int k = 0; k < ks; ++k)
        M->m[k] = 1.0;
    switch (kernel) {
    case STKFMM_KERNEL_PVel:
    case STKFMM_KERNEL_PVelGrad:
    case STKFMM_KERNEL_PVelLaplacian:
    case STKFMM_KERNEL_Traction:
    case STKFMM_KERNEL_RPY:
    case STKFMM_KERNEL_StokesRegVel: M->m[3] = sF; break; // TrD, eps, b (:299-307)
    case STKFMM_KERNEL_StokesRegVelOmega:
        for (int k = 3; k < 7; ++k) // torque and eps (:309-316)
            M->m[k] = sF;
        break;
    default: break;
    }
    return ks;
}

__global__ void __launch_bounds__(256) scale_rows_kernel(double *v, long long n, int k, Mult M) {
    const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n * k)
        return;
    v[i] *= M.m[i % k];
}

__global__ void __launch_bounds__(256) scale_all_kernel(double *v, long long n, double sF) {
    const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n)
        v[i] *= sF;
}

#pragma clang fp contract(off)
__host__ __device__ inline double scale_coord_one(double x, double o, double sF, bool wrap) {
    double y = (x - o) * sF;
    if (wrap)
        y = y - floor(y);
    return y;
}

__global__ void __launch_bounds__(256) scale_coord_kernel(double *c, long long n, double o0, double o1, double o2,
                                                          double sF, int wx, int wy, int wz) {
    const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n)
        return;
    c[3 * i] = scale_coord_one(c[3 * i], o0, sF, wx);
    c[3 * i + 1] = scale_coord_one(c[3 * i + 1], o1, sF, wy);
    c[3 * i + 2] = scale_coord_one(c[3 * i + 2], o2, sF, wz);
}

unsigned nblocks(long long n) { return (unsigned)((n + 255) / 256); }

int check_pbc(unsigned pbc, int *w) {
    if (pbc > 3)
        return report_error(STKFMM_ERR_ARG, "pbc must be PAXIS NONE, PX, PXY or PXYZ (0..3)");
    w[0] = pbc >= 1;
    w[1] = pbc >= 2;
    w[2] = pbc >= 3;
    return STKFMM_OK;
}

int hip_fail(hipError_t e) {
    return e == hipSuccess ? STKFMM_OK : report_error(STKFMM_ERR_HIP, hipGetErrorString(e));
}

} // namespace
} // namespace stkp2p

using namespace stkp2p;

extern "C" int stkfmm_scale_coord(int n, double *coord, const double *origin, double scaleFactor, unsigned pbc) {
    int w[3];
    int rc = check_pbc(pbc, w);
    if (rc)
        return rc;
    if (n < 0 || (n > 0 && (!coord || !origin)))
        return report_error(STKFMM_ERR_ARG, "scale_coord: bad arguments");
    for (long long i = 0; i < n; ++i)
        for (int j = 0; j < 3; ++j)
            coord[3 * i + j] = scale_coord_one(coord[3 * i + j], origin[j], scaleFactor, w[j]);
    return STKFMM_OK;
}

extern "C" int stkfmm_scale_src(unsigned kernel, int nSL, double *srcSLValue, int nDL, double *srcDLValue,
                                double scaleFactor) {
    Mult M;
    const int ks = src_mult(kernel, scaleFactor, &M);
    if (!ks)
        return report_error(STKFMM_ERR_KERNEL, "Error: Kernel not found.");
    int dims[3];
    stkfmm_get_kernel_dimension(kernel, dims);
    if (nSL < 0 || nDL < 0 || (nSL > 0 && !srcSLValue) || (nDL > 0 && !srcDLValue))
        return report_error(STKFMM_ERR_ARG, "scale_src: bad arguments");
    for (long long i = 0; i < (long long)nSL * ks; ++i)
        srcSLValue[i] *= M.m[i % ks];
    for (long long i = 0; i < (long long)nDL * dims[1]; ++i)
        srcDLValue[i] *= scaleFactor;
    return STKFMM_OK;
}

extern "C" int stkfmm_scale_trg(unsigned kernel, int nTrg, double *trgValue, double scaleFactor) {
    Mult M;
    const int kt = trg_mult(kernel, scaleFactor, &M);
    if (!kt)
        return report_error(STKFMM_ERR_KERNEL, "Error: Kernel not found.");
    if (nTrg < 0 || (nTrg > 0 && !trgValue))
        return report_error(STKFMM_ERR_ARG, "scale_trg: bad arguments");
    for (long long i = 0; i < (long long)nTrg * kt; ++i)
        trgValue[i] *= M.m[i % kt];
    return STKFMM_OK;
}

extern "C" int stkfmm_scale_coord_device(int n, double *d_coord, const double *origin, double scaleFactor,
                                         unsigned pbc, void *stream) {
    int w[3];
    int rc = check_pbc(pbc, w);
    if (rc)
        return rc;
    if (n < 0 || (n > 0 && (!d_coord || !origin)))
        return report_error(STKFMM_ERR_ARG, "scale_coord: bad arguments");
    if (n > 0)
        hipLaunchKernelGGL(scale_coord_kernel, dim3(nblocks(n)), dim3(256), 0, (hipStream_t)stream, d_coord,
                           (long long)n, origin[0], origin[1], origin[2], scaleFactor, w[0], w[1], w[2]);
    return hip_fail(hipGetLastError());
}

extern "C" int stkfmm_scale_src_device(unsigned kernel, int nSL, double *d_srcSLValue, int nDL,
                                       double *d_srcDLValue, double scaleFactor, void *stream) {
    Mult M;
    const int ks = src_mult(kernel, scaleFactor, &M);
    if (!ks)
        return report_error(STKFMM_ERR_KERNEL, "Error: Kernel not found.");
    int dims[3];
    stkfmm_get_kernel_dimension(kernel, dims);
    if (nSL < 0 || nDL < 0 || (nSL > 0 && !d_srcSLValue) || (nDL > 0 && !d_srcDLValue))
        return report_error(STKFMM_ERR_ARG, "scale_src: bad arguments");
    hipStream_t s = (hipStream_t)stream;
    if (nSL > 0)
        hipLaunchKernelGGL(scale_rows_kernel, dim3(nblocks((long long)nSL * ks)), dim3(256), 0, s, d_srcSLValue,
                           (long long)nSL, ks, M);
    if (nDL > 0)
        hipLaunchKernelGGL(scale_all_kernel, dim3(nblocks((long long)nDL * dims[1])), dim3(256), 0, s,
                           d_srcDLValue, (long long)nDL * dims[1], scaleFactor);
    return hip_fail(hipGetLastError());
}

extern "C" int stkfmm_scale_trg_device(unsigned kernel, int nTrg, double *d_trgValue, double scaleFactor,
                                       void *stream) {
    Mult M;
    const int kt = trg_mult(kernel, scaleFactor, &M);
    if (!kt)
        return report_error(STKFMM_ERR_KERNEL, "Error: Kernel not found.");
    if (nTrg < 0 || (nTrg > 0 && !d_trgValue))
        return report_error(STKFMM_ERR_ARG, "scale_trg: bad arguments");
    if (nTrg > 0)
        hipLaunchKernelGGL(scale_rows_kernel, dim3(nblocks((long long)nTrg * kt)), dim3(256), 0,
                           (hipStream_t)stream, d_trgValue, (long long)nTrg, kt, M);
    return hip_fail(hipGetLastError());
}
