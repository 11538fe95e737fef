// valu_rates.hip -- measures the VALU issue rate of single instruction kinds on the current GPU (used to
// model the P2P item kernel: FP32, FP64, conversion and reciprocal-square-root costs per wave64
// instruction).  Each kernel runs 8 independent chains per lane of one instruction kind for N
// iterations, 64 waves of work per CU; the result is wave-instructions per CU and cycle at the reported
// engine clock.  Not part of the library; built by stkfmm_amd/csrc/Makefile.
#include <hip/hip_runtime.h>
#include <cstdio>
#define CHK(x) do { hipError_t e = (x); if (e != hipSuccess) { printf("err %s line %d\n", hipGetErrorString(e), __LINE__); return 1; } } while (0)

template <int OP>
__global__ void __launch_bounds__(256) k(float *outf, double *outd, int n, float sf, double sd) {
    float f[8]; double d[8];
#pragma unroll
    for (int i = 0; i < 8; ++i) { f[i] = sf + i + threadIdx.x; d[i] = sd + i + threadIdx.x; }
    for (int it = 0; it < n; ++it) {
#pragma unroll
        for (int i = 0; i < 8; ++i) {
            if constexpr (OP == 0) asm volatile("v_fma_f32 %0, %0, %1, %2" : "+v"(f[i]) : "v"(sf), "v"(f[(i + 1) & 7]));
            if constexpr (OP == 1) asm volatile("v_fma_f64 %0, %0, %1, %2" : "+v"(d[i]) : "v"(sd), "v"(d[(i + 1) & 7]));
            if constexpr (OP == 2) asm volatile("v_add_f64 %0, %0, %1" : "+v"(d[i]) : "v"(sd));
            if constexpr (OP == 3) asm volatile("v_cvt_f32_f64 %0, %1" : "=v"(f[i]) : "v"(d[i]));
            if constexpr (OP == 4) asm volatile("v_rsq_f32 %0, %0" : "+v"(f[i]));
            if constexpr (OP == 5) asm volatile("v_mul_f32 %0, %0, %1" : "+v"(f[i]) : "v"(sf));
            if constexpr (OP == 6) asm volatile("v_add_f32 %0, %0, %1" : "+v"(f[i]) : "v"(sf));
            if constexpr (OP == 7) asm volatile("v_mul_f64 %0, %0, %1" : "+v"(d[i]) : "v"(sd));
            if constexpr (OP == 8) asm volatile("v_rsq_f64 %0, %0" : "+v"(d[i]));
            if constexpr (OP == 9) asm volatile("v_fmac_f32 %0, %1, %2" : "+v"(f[i]) : "v"(sf), "v"(f[(i + 1) & 7]));
        }
    }
    float sfv = 0; double sdv = 0;
#pragma unroll
    for (int i = 0; i < 8; ++i) { sfv += f[i]; sdv += d[i]; }
    const int g = blockIdx.x * blockDim.x + threadIdx.x;
    outf[g] = sfv;
    outd[g] = sdv;
}

template <int OP>
static int run(const char *name, int blocks, float *of, double *od, int n) {
    hipEvent_t a, b;
    CHK(hipEventCreate(&a)); CHK(hipEventCreate(&b));
    hipLaunchKernelGGL(k<OP>, dim3(blocks), dim3(256), 0, 0, of, od, n, 1.0f, 1.0);
    CHK(hipGetLastError());
    CHK(hipDeviceSynchronize());
    CHK(hipEventRecord(a, 0));
    for (int r = 0; r < 3; ++r) hipLaunchKernelGGL(k<OP>, dim3(blocks), dim3(256), 0, 0, of, od, n, 1.0f, 1.0);
    CHK(hipEventRecord(b, 0));
    CHK(hipEventSynchronize(b));
    float ms; CHK(hipEventElapsedTime(&ms, a, b));
    ms /= 3;
    const double waves = blocks * 4.0, instr = waves * n * 8.0;
    hipDeviceProp_t p; CHK(hipGetDeviceProperties(&p, 0));
    const double cu_cycles = p.multiProcessorCount * (p.clockRate * 1e3) * ms * 1e-3;
    printf("%-10s %8.3f ms  wave-instr/CU-cycle %.3f  (CUs %d clock %.0f MHz)\n", name, ms, instr / cu_cycles,
           p.multiProcessorCount, p.clockRate / 1e3);
    return 0;
}

int main() {
    hipDeviceProp_t p;
    CHK(hipGetDeviceProperties(&p, 0));
    const int blocks = p.multiProcessorCount * 16, n = 4096; // 16 workgroups of 4 waves per CU
    float *of;
    double *od;
    CHK(hipMalloc(&of, sizeof(float) * blocks * 256));
    CHK(hipMalloc(&od, sizeof(double) * blocks * 256));
    if (run<0>("fma_f32", blocks, of, od, n) || run<9>("fmac_f32", blocks, of, od, n) || run<5>("mul_f32", blocks, of, od, n) ||
        run<6>("add_f32", blocks, of, od, n) || run<4>("rsq_f32", blocks, of, od, n) || run<1>("fma_f64", blocks, of, od, n) ||
        run<7>("mul_f64", blocks, of, od, n) || run<2>("add_f64", blocks, of, od, n) || run<3>("cvt_f32_f64", blocks, of, od, n) ||
        run<8>("rsq_f64", blocks, of, od, n))
        return 1;
    CHK(hipFree(of));
    CHK(hipFree(od));
    return 0;
}
