// ref_shim.cpp -- extern "C" wrappers around the reference's scalar oracle Test/SimpleKernel.cpp.
//
// TEST INFRASTRUCTURE ONLY.  Built by oracle/Makefile together with the reference source file
// where it lies (never copied into this repository) into oracle/_ref/libstkfmm_simplekernel.so.
// The prototypes come from the reference header Test/SimpleKernel.hpp via -I.
//
// simplekernel_direct() follows the reference's own O(N^2) oracle loop runSimpleKernel
// (Test/Test.cpp:436-513): per target, each source's per-pair result is written to a zeroed
// scratch of 20 doubles and its first kdimTrg entries are added to the target.  The kernel is
// chosen by the KERNEL enum value and layer, as in the SL_kernels map (Test/Test.cpp:12-23).
#include "SimpleKernel.hpp"

#include <omp.h>

typedef void (*kernel_func)(double *, double *, double *, double *);

static kernel_func pick(unsigned kernel, int layer) {
    // layer 0 = single layer, 1 = double layer (Test/Test.cpp:12-23)
    switch (kernel) {
    case 1: return layer ? LaplaceDLPGrad : LaplaceSLPGrad;
    case 2: return layer ? LaplaceDLPGradGrad : LaplaceSLPGradGrad;
    case 4: return layer ? nullptr : LaplaceQPGradGrad;
    case 8: return layer ? nullptr : StokesSL;
    case 16: return layer ? nullptr : StokesSLRPY;
    case 32: return layer ? nullptr : StokesRegSLVel;
    case 64: return layer ? nullptr : StokesRegSLVelOmega;
    case 128: return layer ? StokesDLPVel : StokesSLPVel;
    case 256: return layer ? StokesDLPVelGrad : StokesSLPVelGrad;
    case 512: return layer ? StokesDLPVelLaplacian : StokesSLPVelLaplacian;
    case 1024: return layer ? StokesDLTraction : StokesSLTraction;
    default: return nullptr;
    }
}

extern "C" {

int simplekernel_available(unsigned kernel, int layer) { return pick(kernel, layer) != nullptr; }

// sets the OpenMP thread count when n > 0; returns the count simplekernel_direct uses
int simplekernel_threads(int n) {
    if (n > 0) omp_set_num_threads(n);
    return omp_get_max_threads();
}

// one pair: out[0..19] = zeroed then filled by the reference function
int simplekernel_pair(unsigned kernel, int layer, const double *s, const double *t, const double *f, double *out) {
    kernel_func fn = pick(kernel, layer);
    if (!fn) return -1;
    double ss[3] = {s[0], s[1], s[2]}, tt[3] = {t[0], t[1], t[2]};
    double ff[16];
    for (int i = 0; i < 16; i++) ff[i] = f[i];
    for (int i = 0; i < 20; i++) out[i] = 0.0;
    fn(ss, tt, ff, out);
    return 0;
}

// direct sum with the runSimpleKernel accumulation (Test/Test.cpp:478-510); trg_value +=.
// Targets are split over OpenMP threads as in that loop (Test/Test.cpp:478, omp parallel for).
int simplekernel_direct(unsigned kernel, int layer, int ks, int kt, int nsrc, const double *src_coord,
                        const double *src_value, int ntrg, const double *trg_coord, double *trg_value) {
    kernel_func fn = pick(kernel, layer);
    if (!fn || ks > 16 || kt > 20) return -1;
#pragma omp parallel for
    for (int i = 0; i < ntrg; i++) {
        double t[3] = {trg_coord[3 * i], trg_coord[3 * i + 1], trg_coord[3 * i + 2]};
        for (int j = 0; j < nsrc; j++) {
            double result[20] = {0.0};
            double s[3] = {src_coord[3 * j], src_coord[3 * j + 1], src_coord[3 * j + 2]};
            double sval[16];
            for (int k = 0; k < ks; k++) sval[k] = src_value[(long)ks * j + k];
            fn(s, t, sval, result);
            for (int k = 0; k < kt; k++) trg_value[(long)kt * i + k] += result[k];
        }
    }
    return 0;
}
}
