// direct_driver.cpp -- C++ caller of the STKFMM-compatible P2P interface (include/stkfmm_p2p.hpp).
//
// Follows the reference test driver's direct mode (Test/Test.cpp:542-572 with the inputs of
// genPoint / genSrcValue, Test/Test.cpp:156-203 and :298-404, free space): every active kernel of
// kernelComb is evaluated with STKFMM::evaluateKernel, single layer (SLS2T) then double layer
// (DLS2T), sources coinciding with the targets, source values U(-1, 1) with the radius /
// regularisation slot set to |v| * eps.  Inputs and results are written as raw little-endian
// float64 files K<kernel>_{trg,valSL,valDL,result}.bin in <out_dir> for the parity test.
//
//   direct_driver <out_dir> <n_points> <kernelComb> [eps]
//   direct_driver --inactive-kernel | --unknown-kernel      (error paths, no device work)
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <random>
#include <string>
#include <vector>

#include "../../include/stkfmm_p2p.hpp"

using namespace stkfmm;

static void dump(const std::string &path, const std::vector<double> &v) {
    FILE *f = std::fopen(path.c_str(), "wb");
    if (!f || std::fwrite(v.data(), sizeof(double), v.size(), f) != v.size()) {
        std::fprintf(stderr, "cannot write %s\n", path.c_str());
        std::exit(2);
    }
    std::fclose(f);
}

int main(int argc, char **argv) {
    // error paths of the reference interface; both end the process before any device work
    if (argc == 2 && std::string(argv[1]) == "--inactive-kernel") {
        Stk3DFMM fmm(10, 2000, PAXIS::NONE, asInteger(KERNEL::PVel));
        double x[3] = {0.1, 0.2, 0.3}, v[4] = {1, 1, 1, 0.01}, out[6] = {0};
        fmm.evaluateKernel(KERNEL::RPY, 1, PPKERNEL::SLS2T, 1, x, v, 1, x, out); // exits (STKFMM.cpp:92-95)
        return 0;
    }
    if (argc == 2 && std::string(argv[1]) == "--unknown-kernel") {
        getKernelDimension(static_cast<KERNEL>(3)); // exits (STKFMM.cpp:53-55)
        return 0;
    }
    if (argc < 4) {
        std::fprintf(stderr, "usage: %s <out_dir> <n_points> <kernelComb> [eps]\n", argv[0]);
        return 2;
    }
    const std::string out = argv[1];
    const int n = std::atoi(argv[2]);
    const unsigned comb = (unsigned)std::strtoul(argv[3], nullptr, 10);
    const double eps = argc > 4 ? std::atof(argv[4]) : 0.01;

    std::mt19937_64 gen(20191223);
    std::uniform_real_distribution<double> u01(0.0, 1.0), um11(-1.0, 1.0);
    std::vector<double> trg(3 * (size_t)n);
    for (auto &v : trg)
        v = u01(gen);

    Stk3DFMM fmm(0, 2000, PAXIS::NONE, comb);
    fmm.showActiveKernels();
    for (unsigned k = 1; k <= 1024; k <<= 1) {
        const KERNEL kernel = static_cast<KERNEL>(k);
        if (!(comb & k))
            continue;
        int kdimSL, kdimDL, kdimTrg;
        std::tie(kdimSL, kdimDL, kdimTrg) = getKernelDimension(kernel);
        std::vector<double> valSL((size_t)kdimSL * n), valDL((size_t)kdimDL * n), result((size_t)kdimTrg * n, 0.0);
        for (auto &v : valSL)
            v = um11(gen);
        for (auto &v : valDL)
            v = um11(gen);
        if (kernel == KERNEL::StokesRegVel || kernel == KERNEL::RPY)
            for (int i = 0; i < n; i++)
                valSL[(size_t)kdimSL * i + kdimSL - 1] = std::abs(valSL[(size_t)kdimSL * i + kdimSL - 1]) * eps;
        std::vector<double> srcSL = trg, srcDL = trg; // SL = DL = target points (Test/Test.cpp:183,191)
        fmm.evaluateKernel(kernel, 0, PPKERNEL::SLS2T, n, srcSL.data(), valSL.data(), n, trg.data(),
                           result.data());
        fmm.evaluateKernel(kernel, 0, PPKERNEL::DLS2T, n, srcDL.data(), valDL.data(), n, trg.data(),
                           result.data());
        const std::string tag = out + "/K" + std::to_string(k) + "_";
        dump(tag + "trg.bin", trg);
        dump(tag + "valSL.bin", valSL);
        dump(tag + "valDL.bin", valDL);
        dump(tag + "result.bin", result);
        std::printf("kernel %s done: kdim %d %d %d\n", getKernelName(kernel).c_str(), kdimSL, kdimDL, kdimTrg);
    }
    return 0;
}
