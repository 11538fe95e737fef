// surface.hip -- equivalent / check surfaces of octree boxes (the point sets of the S2M, L2T, W- and
// X-list evaluations), on the host and for many boxes at once on the device.
//
// The layout is the reference's stkfmm::surface (Lib/include/STKFMM/stkfmm_helpers.hpp:46-85), the
// point order PVFMM's densities are indexed by: point 0 is the corner (-1,-1,-1); then the interior
// and edge points of the three faces x = -1, y = -1, z = -1, (p-1)^2 each, in the index order of
// :53-71; the second half is the point reflection of the first.  Unit coordinates u are mapped to
// (u + 1) * alpha * r + c with r = 0.5^(depth+1), so the points span [c, c + 2 alpha r] per axis.
// A box with lower corner x0 and level L gets c = x0 - r (alpha - 1): its surface is centred on the
// box with half-width alpha r (PVFMM's convention; the reference's M2L precompute places its
// surfaces the same way, M2L/Laplace/Laplace3D3D.cpp:233-243).
//
// Both paths evaluate the same IEEE operations in the same order with contraction off, so the
// device points equal the host points bit for bit.
#include <hip/hip_runtime.h>

#include <cmath>

#include "p2p_internal.hpp"

namespace stkp2p {

#pragma clang fp contract(off)

// unit-cube coordinate k (0, 1, 2) of surface point i, p >= 2
__host__ __device__ inline double surface_unit(int p, int i, int k) {
    const int m = (p - 1) * (p - 1);
    const int half = 3 * m + 1;
    double sign = 1.0;
    if (i >= half) {
        i -= half;
        sign = -1.0;
    }
    double u[3] = {-1.0, -1.0, -1.0};
    if (i > 0) {
        const int f = (i - 1) / m, q = (i - 1) % m;
        const int a = q / (p - 1), b = q % (p - 1);
        const double pm1 = (double)(p - 1);
        if (f == 0) {
            u[1] = (2.0 * (a + 1) - p + 1) / pm1;
            u[2] = (2.0 * b - p + 1) / pm1;
        } else if (f == 1) {
            u[0] = (2.0 * a - p + 1) / pm1;
            u[2] = (2.0 * (b + 1) - p + 1) / pm1;
        } else {
            u[0] = (2.0 * (a + 1) - p + 1) / pm1;
            u[1] = (2.0 * b - p + 1) / pm1;
        }
    }
    return sign < 0.0 ? -u[k] : u[k];
}

__host__ __device__ inline double surface_coord(int p, int i, int k, double c, double alpha, int depth) {
    const double r = 0.5 * ldexp(1.0, -depth);
    const double b = alpha * r;
    return (surface_unit(p, i, k) + 1.0) * b + c;
}

void surface_host(int p, const double *c, double alpha, int depth, double *coord) {
    const int n = surface_count(p);
    for (int i = 0; i < n; ++i)
        for (int k = 0; k < 3; ++k)
            coord[3 * i + k] = surface_coord(p, i, k, c[k], alpha, depth);
}

__global__ void __launch_bounds__(256) box_surfaces_kernel(int p, double alpha, int n_boxes, int ns,
                                                          const double *corner, const int *level, double *out) {
    const long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= (long long)n_boxes * ns)
        return;
    const int box = (int)(t / ns), i = (int)(t % ns);
    const int L = level[box];
    const double r = 0.5 * ldexp(1.0, -L);
#pragma unroll
    for (int k = 0; k < 3; ++k) {
        const double c = corner[3LL * box + k] - r * (alpha - 1.0);
        out[3 * t + k] = surface_coord(p, i, k, c, alpha, L);
    }
}

hipError_t launch_box_surfaces(int p, double alpha, int n_boxes, const double *corner, const int *level, double *out,
                               hipStream_t s) {
    const int ns = surface_count(p);
    const long long n = (long long)n_boxes * ns;
    if (n <= 0)
        return hipSuccess;
    hipLaunchKernelGGL(box_surfaces_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, p, alpha, n_boxes,
                       ns, corner, level, out);
    return hipGetLastError();
}

} // namespace stkp2p
