// stkfmm_p2p.hpp -- C++ mirror of STKFMM's P2P interface over the C ABI of libstkfmm_p2p.so.
//
// Same names, argument meaning and error behaviour as the reference for the direct-evaluation
// path, so code written against
//   Lib/include/STKFMM/STKFMM_common.hpp (PAXIS, PPKERNEL, KERNEL, asInteger, getKernelDimension,
//                                         getKernelName)
//   Lib/include/STKFMM/STKFMM.hpp        (STKFMM / Stk3DFMM constructor, evaluateKernel,
//                                         isKernelActive, getMultOrder, showActiveKernels)
// compiles against this header for its evaluateKernel calls.  The FMM part of those classes
// (setBox, setPoints, setupTree, evaluateFMM, clearFMM) is the CPU PVFMM path and is not provided.
// Header-only; link with -lstkfmm_p2p.
#pragma once

#include <cstdio>
#include <cstdlib>
#include <iostream>
#include <string>
#include <tuple>
#include <type_traits>

#include "stkfmm_p2p.h"

namespace stkfmm {

// STKFMM_common.hpp:26-31
enum class PAXIS : unsigned { NONE = 0, PX = 1, PXY = 2, PXYZ = 3 };

// STKFMM_common.hpp:37-41
enum class PPKERNEL : unsigned { SLS2T = 1, DLS2T = 2, L2T = 4 };

// STKFMM_common.hpp:46-62
enum class KERNEL : unsigned {
    LapPGrad = 1,
    LapPGradGrad = 2,
    LapQPGradGrad = 4,
    Stokes = 8,
    RPY = 16,
    StokesRegVel = 32,
    StokesRegVelOmega = 64,
    PVel = 128,
    PVelGrad = 256,
    PVelLaplacian = 512,
    Traction = 1024,
    LapGrad = 2048
};

// STKFMM_common.hpp:102-105
template <typename Enumeration>
auto asInteger(Enumeration const value) -> typename std::underlying_type<Enumeration>::type {
    return static_cast<typename std::underlying_type<Enumeration>::type>(value);
}

// STKFMM.cpp:33-40: [kdimSL, kdimDL, kdimTrg]; an unknown kernel prints and exits (STKFMM.cpp:53-55)
inline std::tuple<int, int, int> getKernelDimension(KERNEL kernel_) {
    int dims[3] = {0, 0, 0};
    if (stkfmm_get_kernel_dimension(asInteger(kernel_), dims) != STKFMM_OK) {
        printf("Error: Kernel not found.\n");
        std::exit(1);
    }
    return std::tuple<int, int, int>(dims[0], dims[1], dims[2]);
}

// STKFMM.cpp:42-46
inline std::string getKernelName(KERNEL kernel_) {
    char buf[64];
    if (stkfmm_get_kernel_name(asInteger(kernel_), buf, sizeof(buf)) != STKFMM_OK) {
        printf("Error: Kernel not found.\n");
        std::exit(1);
    }
    return std::string(buf);
}

// STKFMM.hpp:30-110, P2P subset.  No MPI: the reference's constructor reads the MPI rank
// (STKFMM.cpp:65); here rank is 0.
class STKFMM {
  public:
    STKFMM(int multOrder_, int maxPts_, PAXIS pbc_, unsigned int kernelComb_, bool enableFF_ = true)
        : multOrder(multOrder_), maxPts(maxPts_), pbc(pbc_), kernelComb(kernelComb_) {
        (void)enableFF_;
    }
    virtual ~STKFMM() = default;

    // STKFMM.hpp:108-110 / STKFMM.cpp:88-99 / FMMData.cpp:252-283.  Results are added to
    // trgValuePtr.  nThreads is ignored (the evaluation runs on the current HIP device).
    void evaluateKernel(const KERNEL kernel, const int nThreads, const PPKERNEL p2p, const int nSrc,
                        double *srcCoordPtr, double *srcValuePtr, const int nTrg, double *trgCoordPtr,
                        double *trgValuePtr) {
        if (!isKernelActive(kernel)) {
            std::cout << "Error: no such FMMData exists for kernel " << getKernelName(kernel) << std::endl;
            exit(1);
        }
        const int rc = stkfmm_evaluate_kernel(asInteger(kernel), nThreads, asInteger(p2p), nSrc, srcCoordPtr,
                                              srcValuePtr, nTrg, trgCoordPtr, trgValuePtr);
        if (rc == STKFMM_ERR_PPKERNEL) { // FMMData.cpp:270-273: message, then return
            std::cout << stkfmm_last_error() << std::endl;
            return;
        }
        if (rc != STKFMM_OK) {
            std::cout << "Error: " << stkfmm_last_error() << std::endl;
            exit(1);
        }
    }

    bool isKernelActive(KERNEL kernel_) const { return asInteger(kernel_) & kernelComb; }
    int getMultOrder() const { return multOrder; }

    // STKFMM.cpp:101-111
    void showActiveKernels() const {
        std::cout << "active kernels: ";
        for (unsigned k = 1; k <= 1024; k <<= 1) {
            char buf[64];
            if ((kernelComb & k) && stkfmm_get_kernel_name(k, buf, sizeof(buf)) == STKFMM_OK)
                std::cout << "\t" << buf;
        }
        std::cout << std::endl;
    }

  protected:
    const int multOrder;
    const int maxPts;
    PAXIS pbc;
    const unsigned kernelComb;
};

// STKFMM.hpp:193-204 defaults
class Stk3DFMM : public STKFMM {
  public:
    Stk3DFMM(int multOrder = 10, int maxPts = 2000, PAXIS pbc_ = PAXIS::NONE,
             unsigned int kernelComb_ = asInteger(KERNEL::Stokes) | asInteger(KERNEL::RPY), bool enableFF_ = true)
        : STKFMM(multOrder, maxPts, pbc_, kernelComb_, enableFF_) {}
};

} // namespace stkfmm
