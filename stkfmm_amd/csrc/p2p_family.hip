// p2p_family.hip -- item kernels of the rest of the STKFMM kernel family (micro-kernels 14..24:
// Stokes PVelGrad / PVelLaplacian / Traction single and double layer, regularised force+torque
// velocity/omega, Stokeslet velocity+omega, Laplace quadrupole), in a translation unit of their own
// so that they compile in parallel with the core set.
#include <hip/hip_runtime.h>

#include "p2p_internal.hpp"
#include "ukernels_family.hpp"

namespace stkp2p {

template <> struct UkerOf<14> { using type = UkStokesPVelGrad; };
template <> struct UkerOf<15> { using type = UkStokesDoublePVelGrad; };
template <> struct UkerOf<16> { using type = UkStokesPVelLaplacian; };
template <> struct UkerOf<17> { using type = UkStokesDoubleLaplacian; };
template <> struct UkerOf<18> { using type = UkStokesTraction; };
template <> struct UkerOf<19> { using type = UkStokesDoubleTraction; };
template <> struct UkerOf<20> { using type = UkStokesRegFTVelOmega; };
template <> struct UkerOf<21> { using type = UkStokesVelOmega; };
template <> struct UkerOf<22> { using type = UkLaplaceQuadPGradGrad; };
template <> struct UkerOf<23> { using type = UkLaplaceQuadP; };
template <> struct UkerOf<24> { using type = UkStokesRegFTVel; };

} // namespace stkp2p

#include "p2p_items.hpp"

namespace stkp2p {

namespace {
struct FamilyTable {
    UkerOps row[kNumUkerTotal - kNumCoreUker];
    FamilyTable() {
        row[0] = make_uker_ops<14>();
        row[1] = make_uker_ops<15>();
        row[2] = make_uker_ops<16>();
        row[3] = make_uker_ops<17>();
        row[4] = make_uker_ops<18>();
        row[5] = make_uker_ops<19>();
        row[6] = make_uker_ops<20>();
        row[7] = make_uker_ops<21>();
        row[8] = make_uker_ops<22>();
        row[9] = make_uker_ops<23>();
        row[10] = make_uker_ops<24>();
    }
};
} // namespace

const UkerOps *uker_ops_family(int uk) {
    static const FamilyTable t;
    if (uk >= kNumCoreUker && uk < kNumUkerTotal)
        return &t.row[uk - kNumCoreUker];
    return nullptr;
}

} // namespace stkp2p
