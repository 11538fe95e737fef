// ukernel_host.cpp -- test harness: the library's per-pair micro-kernels (stage, eval, post) and
// per-pair adjoints compiled for the CPU, one pair at a time, so tests/test_ukernels_host.py can check
// their algebra against the oracle in FP64 and FP32 without a GPU.  Built by
// stkfmm_amd/csrc/Makefile with the host shim in tests/cxx/shim.  Not part of the library.
#include "../../stkfmm_amd/csrc/ukernels_family.hpp"
#include "../../stkfmm_amd/csrc/vjp_pairs.hpp"

using namespace stkp2p;

namespace {
template <class K, class R, bool MASK = true>
void pair(const double *r, const double *fin, double scale, double *u) {
    constexpr int P = sizeof(R) == 8 ? 0 : 1, KL = StagedCount<K, P>::value;
    double vst[KL];
    double vin[K::KS];
    for (int k = 0; k < K::KS; ++k)
        vin[k] = fin[k];
    stage_for<K, P>(vin, vst);
    R f[KL];
    for (int k = 0; k < KL; ++k)
        f[k] = (R)vst[k];
    R acc[K::KT];
    for (int k = 0; k < K::KT; ++k)
        acc[k] = R(0);
    if constexpr (HasUnmaskedEval<K>::value)
        K::template eval<R, MASK>(acc, (R)r[0], (R)r[1], (R)r[2], f);
    else
        K::template eval<R>(acc, (R)r[0], (R)r[1], (R)r[2], f);
    for (int k = 0; k < K::KT; ++k)
        u[k] = (double)acc[k] * post_for<K, P>(k) * scale;
}

template <class K>
int run(int prec, const double *r, const double *f, double *u) {
    // prec bit 0: FP32; bit 1: the unmasked form (micro-kernels that offer it; -2 otherwise)
    if ((prec & 2) && !HasUnmaskedEval<K>::value)
        return -2;
    switch (prec) {
    case 0: pair<K, double>(r, f, 1.0, u); break;
    case 1: pair<K, float>(r, f, 1.0, u); break;
    case 2: pair<K, double, false>(r, f, 1.0, u); break;
    default: pair<K, float, false>(r, f, 1.0, u); break;
    }
    return K::KT;
}

// one pair through the adjoint V::vjp in R arithmetic; gr[3] and gf[KS] are overwritten
template <class V, class R>
int vjp_run(const double *r, const double *f, const double *w, double *gr, double *gf) {
    R fr[V::KS], wr[V::KT], g[3] = {R(0), R(0), R(0)}, h[V::KS];
    for (int k = 0; k < V::KS; ++k) {
        fr[k] = (R)f[k];
        h[k] = R(0);
    }
    for (int k = 0; k < V::KT; ++k)
        wr[k] = (R)w[k];
    V::template vjp<R>((R)r[0], (R)r[1], (R)r[2], fr, wr, g, h);
    for (int d = 0; d < 3; ++d)
        gr[d] = (double)g[d];
    for (int k = 0; k < V::KS; ++k)
        gf[k] = (double)h[k];
    return V::KS;
}

template <class V>
int vjp_prec(int prec, const double *r, const double *f, const double *w, double *gr, double *gf) {
    return prec == 0 ? vjp_run<V, double>(r, f, w, gr, gf) : vjp_run<V, float>(r, f, w, gr, gf);
}
} // namespace

// g_r (3) and g_f (KS values) of one pair (no ScaleFactor) through the library's adjoint of micro-
// kernel uk, prec 0 = FP64, 1 = FP32 arithmetic; returns KS, or -1 when uk has no adjoint
extern "C" int vjp_host_pair(int uk, int prec, const double *r, const double *f, const double *w, double *gr,
                             double *gf) {
    switch (uk) {
    case 0: return vjp_prec<VjpLaplaceP>(prec, r, f, w, gr, gf);
    case 2: return vjp_prec<VjpLaplacePGrad>(prec, r, f, w, gr, gf);
    case 6: return vjp_prec<VjpStokesVel>(prec, r, f, w, gr, gf);
    case 7: return vjp_prec<VjpStokesPVel>(prec, r, f, w, gr, gf);
    case 8: return vjp_prec<VjpStokesDoublePVel>(prec, r, f, w, gr, gf);
    case 9: return vjp_prec<VjpStokesRegVel>(prec, r, f, w, gr, gf);
    case 10: return vjp_prec<VjpRpyU>(prec, r, f, w, gr, gf);
    case 11: return vjp_prec<VjpRpyULapU>(prec, r, f, w, gr, gf);
    case 12: return vjp_prec<VjpStkULapU>(prec, r, f, w, gr, gf);
    case 13: return vjp_prec<VjpStokesPVelFused>(prec, r, f, w, gr, gf);
    default: return -1;
    }
}

// u = the micro-kernel's sum for one pair (no ScaleFactor), prec 0 = FP64, 1 = FP32 arithmetic,
// 2 / 3 = the same without the coincidence mask (eval<R, false>, for pairs with r2 > 0); returns the
// number of outputs, -1 for an unknown id, -2 when the micro-kernel has no unmasked form
extern "C" int ukernel_host_pair(int uk, int prec, const double *r, const double *f, double *u) {
    switch (uk) {
    case 0: return run<UkLaplaceP>(prec, r, f, u);
    case 1: return run<UkLaplaceDipoleP>(prec, r, f, u);
    case 2: return run<UkLaplacePGrad>(prec, r, f, u);
    case 3: return run<UkLaplaceDipolePGrad>(prec, r, f, u);
    case 4: return run<UkLaplacePGradGrad>(prec, r, f, u);
    case 5: return run<UkLaplaceDipolePGradGrad>(prec, r, f, u);
    case 6: return run<UkStokesVel>(prec, r, f, u);
    case 7: return run<UkStokesPVel>(prec, r, f, u);
    case 8: return run<UkStokesDoublePVel>(prec, r, f, u);
    case 9: return run<UkStokesRegVel>(prec, r, f, u);
    case 10: return run<UkRpyU>(prec, r, f, u);
    case 11: return run<UkRpyULapU>(prec, r, f, u);
    case 12: return run<UkStkULapU>(prec, r, f, u);
    case 13: return run<UkStokesPVelFused>(prec, r, f, u);
    case 14: return run<UkStokesPVelGrad>(prec, r, f, u);
    case 15: return run<UkStokesDoublePVelGrad>(prec, r, f, u);
    case 16: return run<UkStokesPVelLaplacian>(prec, r, f, u);
    case 17: return run<UkStokesDoubleLaplacian>(prec, r, f, u);
    case 18: return run<UkStokesTraction>(prec, r, f, u);
    case 19: return run<UkStokesDoubleTraction>(prec, r, f, u);
    case 20: return run<UkStokesRegFTVelOmega>(prec, r, f, u);
    case 21: return run<UkStokesVelOmega>(prec, r, f, u);
    case 22: return run<UkLaplaceQuadPGradGrad>(prec, r, f, u);
    case 23: return run<UkLaplaceQuadP>(prec, r, f, u);
    case 24: return run<UkStokesRegFTVel>(prec, r, f, u);
    default: return -1;
    }
}
