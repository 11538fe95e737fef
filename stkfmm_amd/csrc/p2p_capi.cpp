// p2p_capi.cpp -- host side of libstkfmm_p2p.so: the C ABI declared in include/stkfmm_p2p.h.
//
// Kernel table, work-list planning (tiling, chunking, cost-sorted queue), the direct-sum and
// Ker_t-shaped entries, graph capture and the adjoint entry.  The arithmetic lives in
// p2p_kernels.hip; nothing here computes a kernel value on the host.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdlib>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <list>
#include <mutex>
#include <numeric>
#include <string>
#include <vector>

#include "../../include/stkfmm_p2p.h"
#include "p2p_internal.hpp"

using namespace stkp2p;

// ---- errors -------------------------------------------------------------------------------------

static thread_local std::string g_last_error;

static int set_error(int code, const char *fmt, ...) {
    char buf[512];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof(buf), fmt, ap);
    va_end(ap);
    g_last_error = buf;
    return code;
}

// for the entries defined in other translation units (box.hip)
int stkp2p::report_error(int code, const char *msg) { return set_error(code, "%s", msg); }

static int hip_error(hipError_t e, const char *what) {
    return set_error(STKFMM_ERR_HIP, "%s: %s", what, hipGetErrorString(e));
}

#define HIP_TRY(call, what)                                                                        \
    do {                                                                                           \
        hipError_t e_ = (call);                                                                    \
        if (e_ != hipSuccess)                                                                      \
            return hip_error(e_, what);                                                            \
    } while (0)

extern "C" const char *stkfmm_last_error(void) { return g_last_error.c_str(); }

// ---- kernel table (STKFMM.cpp:18-40, STKFMM_common.hpp:46-62, FMMData.cpp:261-268) ---------------

namespace {
struct KernelEntry {
    unsigned kernel;
    const char *name; // ker_name given to BuildKernel
    int kdim_sl, kdim_dl, kdim_trg;
    int uk_sl, uk_dl, uk_l2t; // k_s2t single / double layer, k_l2t; -1: absent in the reference
    int uk_s2m_sl, uk_s2m_dl; // k_s2m (sources -> upward check surface)
    int uk_s2l_sl, uk_s2l_dl; // k_s2l (X-list: sources -> downward check surface)
    int uk_m2t;               // k_m2t (W-list: upward equivalent densities -> targets)
};

// kdimDL of kernels without a double layer is PVFMM's default surf_dim (PVFMM_COORD_DIM +
// ker_dim[0]); PVFMM is not vendored, so those entries are inferred, not pinned.
// Sub-kernels from the BuildKernel lists (argument order s2m, s2l, s2t, m2m, m2l, m2t, l2l, l2t; a
// NULL entry means the kernel itself):
//   LapPGrad / LapPGradGrad: s2m = s2l = laplace_p + laplace_dipolep, m2t and l2t itself
//     (LaplaceLayerKernel.hpp:315-341);
//   LapQPGradGrad: s2m = s2l = laplace_quadp, m2t = l2t = laplace_pgradgrad (LaplaceLayerKernel.hpp:344-355);
//   Stokes: everything stokes_vel (StokesLayerKernel.hpp:38-42);
//   RPY: s2m = s2l = rpy_u, m2t = l2t = stk_ulapu (RPYKernel.hpp:136-152);
//   StokesRegVel: s2m = s2l itself, m2t = l2t PVFMM's Stokes velocity (StokesRegSingleLayerKernel.hpp:211-218);
//   StokesRegVelOmega: s2m = s2l = stokes_regftvel, m2t = l2t = stokes_velomega
//     (StokesRegSingleLayerKernel.hpp:221-232);
//   PVel / PVelGrad / PVelLaplacian / Traction: s2m = s2l = stokes_pvel + stokes_doublepvel, m2t and
//     l2t itself (StokesLayerKernel.hpp:45-88).
#define U(x) STKFMM_UKER_##x
const KernelEntry kTable[] = {
    {1, "laplace_PGrad", 1, 3, 4, U(LAPLACE_PGRAD), U(LAPLACE_DIPOLEPGRAD), U(LAPLACE_PGRAD), U(LAPLACE_P),
     U(LAPLACE_DIPOLEP), U(LAPLACE_P), U(LAPLACE_DIPOLEP), U(LAPLACE_PGRAD)},
    {2, "laplace_PGradGrad", 1, 3, 10, U(LAPLACE_PGRADGRAD), U(LAPLACE_DIPOLEPGRADGRAD), U(LAPLACE_PGRADGRAD),
     U(LAPLACE_P), U(LAPLACE_DIPOLEP), U(LAPLACE_P), U(LAPLACE_DIPOLEP), U(LAPLACE_PGRADGRAD)},
    {4, "laplace_QPGradGrad", 9, 12, 10, U(LAPLACE_QUADPGRADGRAD), -1, U(LAPLACE_PGRADGRAD), U(LAPLACE_QUADP), -1,
     U(LAPLACE_QUADP), -1, U(LAPLACE_PGRADGRAD)},
    {8, "stokes_vel", 3, 6, 3, U(STOKES_VEL), -1, U(STOKES_VEL), U(STOKES_VEL), -1, U(STOKES_VEL), -1, U(STOKES_VEL)},
    {16, "rpy_ulapu", 4, 7, 6, U(RPY_ULAPU), -1, U(STK_ULAPU), U(RPY_U), -1, U(RPY_U), -1, U(STK_ULAPU)},
    {32, "stokes_regvel", 4, 7, 3, U(STOKES_REGVEL), -1, U(STOKES_VEL), U(STOKES_REGVEL), -1, U(STOKES_REGVEL), -1,
     U(STOKES_VEL)},
    {64, "stokes_regftvelomega", 7, 10, 6, U(STOKES_REGFTVELOMEGA), -1, U(STOKES_VELOMEGA), U(STOKES_REGFTVEL), -1,
     U(STOKES_REGFTVEL), -1, U(STOKES_VELOMEGA)},
    {128, "stokes_PVel", 4, 9, 4, U(STOKES_PVEL), U(STOKES_DOUBLEPVEL), U(STOKES_PVEL), U(STOKES_PVEL),
     U(STOKES_DOUBLEPVEL), U(STOKES_PVEL), U(STOKES_DOUBLEPVEL), U(STOKES_PVEL)},
    {256, "stokes_PVelGrad", 4, 9, 16, U(STOKES_PVELGRAD), U(STOKES_DOUBLEPVELGRAD), U(STOKES_PVELGRAD),
     U(STOKES_PVEL), U(STOKES_DOUBLEPVEL), U(STOKES_PVEL), U(STOKES_DOUBLEPVEL), U(STOKES_PVELGRAD)},
    {512, "stokes_PVelLaplacian", 4, 9, 7, U(STOKES_PVELLAPLACIAN), U(STOKES_DOUBLELAPLACIAN),
     U(STOKES_PVELLAPLACIAN), U(STOKES_PVEL), U(STOKES_DOUBLEPVEL), U(STOKES_PVEL), U(STOKES_DOUBLEPVEL),
     U(STOKES_PVELLAPLACIAN)},
    {1024, "stokes_Traction", 4, 9, 9, U(STOKES_TRACTION), U(STOKES_DOUBLETRACTION), U(STOKES_TRACTION),
     U(STOKES_PVEL), U(STOKES_DOUBLEPVEL), U(STOKES_PVEL), U(STOKES_DOUBLEPVEL), U(STOKES_TRACTION)},
};
#undef U

const KernelEntry *find_kernel(unsigned kernel) {
    for (const auto &e : kTable)
        if (e.kernel == kernel)
            return &e;
    return nullptr;
}
} // namespace

extern "C" int stkfmm_get_kernel_dimension(unsigned kernel, int *dims) {
    const KernelEntry *e = find_kernel(kernel);
    if (!e)
        return set_error(STKFMM_ERR_KERNEL, "Error: Kernel not found.");
    if (!dims)
        return set_error(STKFMM_ERR_ARG, "dims is NULL");
    dims[0] = e->kdim_sl;
    dims[1] = e->kdim_dl;
    dims[2] = e->kdim_trg;
    return STKFMM_OK;
}

extern "C" int stkfmm_get_kernel_name(unsigned kernel, char *buf, int buflen) {
    const KernelEntry *e = find_kernel(kernel);
    if (!e)
        return set_error(STKFMM_ERR_KERNEL, "Error: Kernel not found.");
    if (!buf || buflen <= 0)
        return set_error(STKFMM_ERR_ARG, "buffer is NULL or empty");
    snprintf(buf, (size_t)buflen, "%s", e->name);
    return STKFMM_OK;
}

extern "C" int stkfmm_get_uker(unsigned kernel, unsigned p2p, int *uker) {
    const KernelEntry *e = find_kernel(kernel);
    if (!e)
        return set_error(STKFMM_ERR_KERNEL, "Error: Kernel not found.");
    int uk;
    if (p2p == STKFMM_PPKERNEL_SLS2T)
        uk = e->uk_sl;
    else if (p2p == STKFMM_PPKERNEL_DLS2T)
        uk = e->uk_dl;
    else if (p2p == STKFMM_PPKERNEL_L2T)
        uk = e->uk_l2t;
    else
        uk = -1;
    if (uk == -1)
        return set_error(STKFMM_ERR_PPKERNEL, "PPKernel %u not found for direct evaluation", p2p);
    if (uker)
        *uker = uk;
    return STKFMM_OK;
}

extern "C" int stkfmm_get_s2m_uker(unsigned kernel, int double_layer, int *uker) {
    const KernelEntry *e = find_kernel(kernel);
    if (!e)
        return set_error(STKFMM_ERR_KERNEL, "Error: Kernel not found.");
    const int uk = double_layer ? e->uk_s2m_dl : e->uk_s2m_sl;
    if (uk < 0)
        return set_error(STKFMM_ERR_PPKERNEL, "kernel %s has no double-layer S2M sub-kernel", e->name);
    if (uker)
        *uker = uk;
    return STKFMM_OK;
}

extern "C" int stkfmm_get_s2l_uker(unsigned kernel, int double_layer, int *uker) {
    const KernelEntry *e = find_kernel(kernel);
    if (!e)
        return set_error(STKFMM_ERR_KERNEL, "Error: Kernel not found.");
    const int uk = double_layer ? e->uk_s2l_dl : e->uk_s2l_sl;
    if (uk < 0)
        return set_error(STKFMM_ERR_PPKERNEL, "kernel %s has no double-layer S2L sub-kernel", e->name);
    if (uker)
        *uker = uk;
    return STKFMM_OK;
}

extern "C" int stkfmm_get_m2t_uker(unsigned kernel, int *uker) {
    const KernelEntry *e = find_kernel(kernel);
    if (!e)
        return set_error(STKFMM_ERR_KERNEL, "Error: Kernel not found.");
    if (uker)
        *uker = e->uk_m2t;
    return STKFMM_OK;
}

// ---- equivalent / check surfaces (stkfmm_helpers.hpp:46-85) ---------------------------------------

extern "C" int stkfmm_surface_count(int p) { return p >= 2 ? stkp2p::surface_count(p) : STKFMM_ERR_ARG; }

extern "C" int stkfmm_surface(int p, const double *c, double alpha, int depth, double *coord) {
    if (p < 2 || !c || !coord || depth < 0 || depth > 60)
        return set_error(STKFMM_ERR_ARG, "surface: need p >= 2, 0 <= depth <= 60, non-null c and coord");
    stkp2p::surface_host(p, c, alpha, depth, coord);
    return STKFMM_OK;
}

extern "C" int stkfmm_box_surfaces_device(int p, double alpha, int n_boxes, const double *d_corner,
                                          const int *d_level, double *d_coord, void *stream) {
    if (p < 2 || n_boxes < 0 || (n_boxes > 0 && (!d_corner || !d_level || !d_coord)))
        return set_error(STKFMM_ERR_ARG, "box_surfaces: need p >= 2, n_boxes >= 0 and device arrays");
    HIP_TRY(stkp2p::launch_box_surfaces(p, alpha, n_boxes, d_corner, d_level, d_coord, (hipStream_t)stream),
            "box_surfaces launch");
    return STKFMM_OK;
}

extern "C" int stkfmm_uker_dims(int uker, int *kdim_src, int *kdim_trg) {
    if (uker < 0 || uker >= STKFMM_UKER_COUNT)
        return set_error(STKFMM_ERR_ARG, "unknown micro-kernel %d", uker);
    if (kdim_src)
        *kdim_src = uker_ks(uker);
    if (kdim_trg)
        *kdim_trg = uker_kt(uker);
    return STKFMM_OK;
}

// ---- device info ----------------------------------------------------------------------------------

static int device_cu_count(int *n_cu) {
    int dev = 0;
    HIP_TRY(hipGetDevice(&dev), "hipGetDevice");
    HIP_TRY(hipDeviceGetAttribute(n_cu, hipDeviceAttributeMultiprocessorCount, dev), "hipDeviceGetAttribute");
    return STKFMM_OK;
}

extern "C" int stkfmm_device_info(int *n_cu, int *warp_size, char *arch, int arch_len) {
    int dev = 0;
    HIP_TRY(hipGetDevice(&dev), "hipGetDevice");
    hipDeviceProp_t prop;
    HIP_TRY(hipGetDeviceProperties(&prop, dev), "hipGetDeviceProperties");
    if (n_cu)
        *n_cu = prop.multiProcessorCount;
    if (warp_size)
        *warp_size = prop.warpSize;
    if (arch && arch_len > 0)
        snprintf(arch, (size_t)arch_len, "%s", prop.gcnArchName);
    return STKFMM_OK;
}

// ---- plans ------------------------------------------------------------------------------------------

struct stkfmm_plan {
    int uk = 0, prec = 0, T = 2, TT = 128, KS = 0, KS1 = 0, KT = 0;
    int n_src = 0, n_trg = 0;
    int n_items = 0, n_tiles = 0, n_ranges = 0;
    int grid = 0, persistent = 1;
    long long pairs = 0;
    int device = 0;
    Item *d_items = nullptr;
    int2 *d_ranges = nullptr;
    Tile *d_tiles = nullptr;
    double *d_partial = nullptr;
    int *d_counter = nullptr; // [n_queues]
    int *d_qptr = nullptr;    // [n_queues + 1]
    int n_queues = 1;
    hipGraph_t graph = nullptr;
    hipGraphExec_t gexec = nullptr;
    // timing ring: execute k records events ev[3*(k % kRing) + {0,1,2}] around the item and reduce
    // kernels on the execution stream
    static constexpr int kRing = 256;
    bool timing = false;
    std::vector<hipEvent_t> ev;
    long long n_timed = 0;
    std::vector<int> group_items; // mode 3: item index ranges per launched group (host)
    // scalar-broadcast variant: staged source records, one copy per level in [level_min, level_max]
    int broadcast = 0;
    // launch variant of the LDS item kernel: precision, or kPrecMixedFar (mixed with the far-tile
    // displacement) for tree-derived mixed work lists at T = 2
    int kernel_prec = 0;
    int level_min = 0, n_copies = 0, rec_words = 0;
    double *d_recs = nullptr;
    // sort_points: Morton ordering of each execute's points (morton.hip) into the SoA copies below
    bool sorted = false;
    int n_sort_segs = 0, n_sort_blks = 0;
    // every group's target range is its source range: with one coordinate array for both, only the
    // source segments (the first n_src_segs segments, n_src_blks blocks) need ordering
    bool self_ranges = false;
    int n_src_segs = 0, n_src_blks = 0;
    SortSeg *d_sort_segs = nullptr;
    SortBlk *d_sort_blks = nullptr;
    double *d_sort_box = nullptr, *d_src_coord = nullptr, *d_src_value = nullptr, *d_trg_coord = nullptr;
    int *d_sort_hist = nullptr, *d_src_idx = nullptr, *d_trg_idx = nullptr, *d_trg_row = nullptr;
    unsigned short *d_sort_keys = nullptr;
};

static void plan_free(stkfmm_plan *p) {
    if (!p)
        return;
    if (p->gexec)
        (void)hipGraphExecDestroy(p->gexec);
    if (p->graph)
        (void)hipGraphDestroy(p->graph);
    for (auto &e : p->ev)
        (void)hipEventDestroy(e);
    (void)hipFree(p->d_items);
    (void)hipFree(p->d_ranges);
    (void)hipFree(p->d_tiles);
    (void)hipFree(p->d_partial);
    (void)hipFree(p->d_counter);
    (void)hipFree(p->d_qptr);
    (void)hipFree(p->d_recs);
    (void)hipFree(p->d_sort_segs);
    (void)hipFree(p->d_sort_blks);
    (void)hipFree(p->d_sort_box);
    (void)hipFree(p->d_src_coord);
    (void)hipFree(p->d_src_value);
    (void)hipFree(p->d_trg_coord);
    (void)hipFree(p->d_sort_hist);
    (void)hipFree(p->d_src_idx);
    (void)hipFree(p->d_trg_idx);
    (void)hipFree(p->d_trg_row);
    (void)hipFree(p->d_sort_keys);
    delete p;
}

template <class T>
static int upload(T **dst, const std::vector<T> &src, const char *what) {
    *dst = nullptr;
    if (src.empty())
        return STKFMM_OK;
    HIP_TRY(hipMalloc((void **)dst, sizeof(T) * src.size()), what);
    HIP_TRY(hipMemcpy(*dst, src.data(), sizeof(T) * src.size(), hipMemcpyHostToDevice), what);
    return STKFMM_OK;
}

// Target tiles of a group of n targets: whole tiles of TT and a remainder.  A tile of c targets
// occupies TT >> tile_lsplit(c, TT) lane slots (at least TT/16), so a remainder just above a power of
// two leaves up to half its slots idle for every source of the group.  With `pieces` the remainder is
// cut into power-of-two tiles, largest first, until the rest fills its slots to within 1/8 or a cut
// saves no slots (on the 1M PVel U-list: 95.7 % -> 98.5 % of the lane slots busy, 14 % more tiles,
// each staging its sources again).  A target's sources and their order are the same in every layout; the lanes that hold it
// and the butterfly over split lanes are not.
static void group_tiles(int n, int TT, bool pieces, std::vector<int> &counts) {
    counts.clear();
    for (; n >= TT; n -= TT)
        counts.push_back(TT);
    while (n > 0) {
        const int s = TT >> tile_lsplit(n, TT);
        int q = 1;
        while (2 * q <= n)
            q *= 2;
        // stop when the rest fills its slots to within 1/8, or a piece would save no slots
        if (!pieces || 8 * (s - n) <= s || q == n || q + (TT >> tile_lsplit(n - q, TT)) >= s) {
            counts.push_back(n);
            break;
        }
        counts.push_back(q);
        n -= q;
    }
}

extern "C" int stkfmm_plan_create(stkfmm_plan **plan_out, int uk, int precision, int n_src, int n_trg,
                                  int n_groups, const int *grp_trg_begin, const int *grp_trg_end,
                                  const int *grp_level, const int *grp_rng_ptr, const int *rng_src_begin,
                                  const int *rng_src_end, const stkfmm_tune *tune) {
    if (!plan_out)
        return set_error(STKFMM_ERR_ARG, "plan pointer is NULL");
    *plan_out = nullptr;
    if (uk < 0 || uk >= STKFMM_UKER_COUNT)
        return set_error(STKFMM_ERR_ARG, "unknown micro-kernel %d", uk);
    if (precision != STKFMM_PREC_FP64 && precision != STKFMM_PREC_MIXED)
        return set_error(STKFMM_ERR_ARG, "unknown precision %d", precision);
    if (n_src < 0 || n_trg < 0 || n_groups < 0)
        return set_error(STKFMM_ERR_ARG, "negative size");
    if (n_groups > 0 && (!grp_trg_begin || !grp_trg_end || !grp_rng_ptr))
        return set_error(STKFMM_ERR_ARG, "group arrays are NULL");

    stkfmm_tune tn{};
    if (tune)
        tn = *tune;
    // default targets per lane: 2, and 1 in mixed mode for the family members with many outputs or
    // much per-pair state, whose two-target loops run at lower occupancy: 262,144-point U-list, item
    // kernel + reduce (ms), T = 2 -> 1: PVelGrad 2.25 -> 2.03, LapQPGradGrad 2.34 -> 2.13,
    // DoubleLaplacian 1.48 -> 1.40 (DoublePVelGrad, DoubleTraction and RegFTVelOmega 9-12 % faster at
    // T = 1 with the autotuner's item sizes; the others the same or slower at 1)
    const bool t1 = precision == STKFMM_PREC_MIXED &&
                    (uk == STKFMM_UKER_STOKES_PVELGRAD || uk == STKFMM_UKER_STOKES_DOUBLEPVELGRAD ||
                     uk == STKFMM_UKER_STOKES_DOUBLELAPLACIAN || uk == STKFMM_UKER_STOKES_DOUBLETRACTION ||
                     uk == STKFMM_UKER_STOKES_REGFTVELOMEGA || uk == STKFMM_UKER_LAPLACE_QUADPGRADGRAD);
    const int T = tn.targets_per_lane ? tn.targets_per_lane : (t1 ? 1 : 2);
    if (T != 1 && T != 2 && T != 4)
        return set_error(STKFMM_ERR_ARG, "targets_per_lane must be 1, 2 or 4 (got %d)", T);
    const int persistent = tn.persistent ? tn.persistent : 1;
    if (persistent < 1 || persistent > 3)
        return set_error(STKFMM_ERR_ARG, "persistent must be 1, 2 or 3 (got %d)", persistent);
    const int broadcast = tn.broadcast;
    if (broadcast != 0 && broadcast != 1)
        return set_error(STKFMM_ERR_ARG, "broadcast must be 0 (LDS) or 1 (scalar) (got %d)", broadcast);
    if (tn.sort_points != 0 && tn.sort_points != 1)
        return set_error(STKFMM_ERR_ARG, "sort_points must be 0 or 1 (got %d)", tn.sort_points);

    // validate the work list
    long long total_cost = 0; // sum over tiles of the tile's source count
    const int TT = kWave * T;
    for (int g = 0; g < n_groups; ++g) {
        const int b = grp_trg_begin[g], e = grp_trg_end[g];
        if (b < 0 || e < b || e > n_trg)
            return set_error(STKFMM_ERR_ARG, "group %d: target range [%d,%d) outside [0,%d)", g, b, e, n_trg);
        const int r0 = grp_rng_ptr[g], r1 = grp_rng_ptr[g + 1];
        if (r0 < 0 || r1 < r0)
            return set_error(STKFMM_ERR_ARG, "group %d: bad range pointers", g);
        if (r1 > r0 && (!rng_src_begin || !rng_src_end))
            return set_error(STKFMM_ERR_ARG, "range arrays are NULL");
        long long ns = 0;
        for (int r = r0; r < r1; ++r) {
            const int sb = rng_src_begin[r], se = rng_src_end[r];
            if (sb < 0 || se < sb || se > n_src)
                return set_error(STKFMM_ERR_ARG, "range %d: sources [%d,%d) outside [0,%d)", r, sb, se, n_src);
            ns += se - sb;
        }
        if (grp_level && (grp_level[g] < -60 || grp_level[g] > 60))
            return set_error(STKFMM_ERR_ARG, "group %d: level %d out of range", g, grp_level[g]);
        total_cost += ns * (long long)((e - b + TT - 1) / TT);
    }

    // Morton ordering: segments (distinct source ranges, group target ranges) and their blocks
    std::vector<SortSeg> segs;
    std::vector<SortBlk> blks;
    long long hist_len = 0;
    bool self_ranges = false;
    int n_src_segs = 0, n_src_blks = 0;
    if (tn.sort_points) {
        std::vector<std::pair<int, int>> tr, sr;
        self_ranges = true;
        for (int g = 0; g < n_groups; ++g) {
            const int r0 = grp_rng_ptr[g], r1 = grp_rng_ptr[g + 1];
            if (grp_trg_end[g] == grp_trg_begin[g] || r1 == r0)
                continue;
            if (r1 - r0 != 1)
                return set_error(STKFMM_ERR_UNSUPPORTED, "sort_points: group %d has %d source ranges (needs one)", g,
                                 r1 - r0);
            tr.emplace_back(grp_trg_begin[g], grp_trg_end[g]);
            if (rng_src_end[r0] > rng_src_begin[r0])
                sr.emplace_back(rng_src_begin[r0], rng_src_end[r0]);
            self_ranges = self_ranges && rng_src_begin[r0] == grp_trg_begin[g] && rng_src_end[r0] == grp_trg_end[g];
        }
        std::sort(tr.begin(), tr.end());
        std::sort(sr.begin(), sr.end());
        sr.erase(std::unique(sr.begin(), sr.end()), sr.end());
        for (size_t i = 1; i < tr.size(); ++i)
            if (tr[i].first < tr[i - 1].second)
                return set_error(STKFMM_ERR_UNSUPPORTED, "sort_points: target ranges of two groups overlap");
        for (size_t i = 1; i < sr.size(); ++i)
            if (sr[i].first < sr[i - 1].second)
                return set_error(STKFMM_ERR_UNSUPPORTED,
                                 "sort_points: two source ranges overlap without being identical");
        auto add = [&](const std::vector<std::pair<int, int>> &v, int kind) {
            for (const auto &r : v) {
                SortSeg sg{r.first, r.second, (int)blks.size(), 0, sort_level(r.second - r.first), kind};
                for (int b = r.first; b < r.second; b += kSortBlock) {
                    blks.push_back(SortBlk{(int)segs.size(), b, std::min(r.second, b + kSortBlock), hist_len});
                    hist_len += 1LL << (3 * sg.level);
                }
                sg.blk_end = (int)blks.size();
                segs.push_back(sg);
            }
        };
        add(sr, 0);
        n_src_segs = (int)segs.size();
        n_src_blks = (int)blks.size();
        add(tr, 1);
    }

    int n_cu = 0, bpc = 0;
    int rc = device_cu_count(&n_cu);
    if (rc)
        return rc;
    // the far-tile form pays where targets and sources are spatially sorted (groups with octree
    // levels, or points the plan orders itself); for direct sums of unsorted points its extra
    // registers only cost occupancy
    const bool coherent = grp_level || !segs.empty();
    const int kernel_prec = (precision == STKFMM_PREC_MIXED && coherent && T == 2 && !broadcast) ? kPrecMixedFar
                                                                                                  : precision;
    if (broadcast)
        HIP_TRY(items_sgpr_occupancy(uk, precision, T, &bpc), "occupancy query");
    else
        HIP_TRY(items_occupancy(uk, kernel_prec, T, &bpc), "occupancy query");
    if (bpc < 1)
        bpc = 1;
    const int concurrent = n_cu * bpc;

    long long max_src = tn.max_src_per_item;
    if (max_src <= 0 && persistent == 3) {
        max_src = 512; // per-group launches: split like a stand-alone direct sum of that group
    } else if (max_src <= 0) {
        max_src = default_max_src(total_cost, concurrent);
    }
    if (max_src < 1)
        max_src = 1;

    // tiles and items
    std::vector<Tile> tiles;
    std::vector<Item> items;
    std::vector<int> item_group; // work-list group of each item
    std::vector<int2> ranges;
    std::vector<int> group_items(1, 0);
    long long pairs = 0;
    // remainder tiles cut into power-of-two pieces in plans of several groups (a direct sum's one
    // remainder tile is not worth it, and keeps the layout of the no-alloc entry)
    const bool pieces = tn.split_tiles >= 0 && n_groups > 1;
    std::vector<int> tile_counts;
    for (int g = 0; g < n_groups; ++g) {
        const int b = grp_trg_begin[g], e = grp_trg_end[g];
        const int r0 = grp_rng_ptr[g], r1 = grp_rng_ptr[g + 1];
        long long ns = 0;
        for (int r = r0; r < r1; ++r)
            ns += rng_src_end[r] - rng_src_begin[r];
        pairs += ns * (long long)(e - b);
        if (ns == 0 || e == b)
            continue;
        const int level = grp_level ? grp_level[g] : 0;
        group_tiles(e - b, TT, pieces, tile_counts);
        int tb = b;
        for (const int tc : tile_counts) {
            Tile tl;
            tl.trg_begin = tb;
            tl.trg_count = tc;
            tb += tc;
            tl.slot_begin = (int)items.size();
            // lanes per target: the largest power of two (at most 16) that still fits the tile
            const int lsplit = (tn.split_tiles >= 0) ? tile_lsplit(tl.trg_count, TT) : 0;
            Item cur{tl.trg_begin, tl.trg_count, (int)ranges.size(), (int)ranges.size(), 0, level, 0, lsplit};
            long long cur_n = 0;
            for (int r = r0; r < r1; ++r) {
                int sb = rng_src_begin[r];
                const int se = rng_src_end[r];
                while (sb < se) {
                    const int m = (int)std::min<long long>(se - sb, max_src - cur_n);
                    ranges.push_back(make_int2(sb, sb + m));
                    cur.rng_end = (int)ranges.size();
                    cur_n += m;
                    sb += m;
                    if (cur_n == max_src) {
                        cur.cost = (int)cur_n;
                        cur.slot = (int)items.size();
                        items.push_back(cur);
                        item_group.push_back(g);
                        cur.rng_begin = cur.rng_end = (int)ranges.size();
                        cur_n = 0;
                    }
                }
            }
            if (cur_n > 0) {
                cur.cost = (int)cur_n;
                cur.slot = (int)items.size();
                items.push_back(cur);
                item_group.push_back(g);
            }
            tl.slot_end = (int)items.size();
            tiles.push_back(tl);
        }
        if ((int)items.size() > group_items.back())
            group_items.push_back((int)items.size());
    }
    // Item order.  Mode 3 keeps the groups apart in work-list order.  Mode 2, and mode 1 with a single
    // queue: one cost-sorted queue (largest first; ties keep work-list order).  Mode 1 with several
    // queues: groups are dealt to the queues largest-first onto the least-loaded queue (balancing the
    // summed item costs), and each queue lists its groups largest-first with a group's items
    // consecutive (largest first), so the waves of one XCD work through one group's sources at a
    // time.  Item slots (and so results) do not depend on the order.
    int n_queues = 1;
    if (persistent == 1) {
        // default: one queue (measured faster on the VALU-bound kernels); tn.queues = 8 gives one
        // queue per XCD of MI355X (8 XCDs, each with its own L2, workgroups dealt round-robin)
        n_queues = tn.queues > 0 ? tn.queues : 1;
        n_queues = std::max(1, std::min({n_queues, 256, (int)items.size()}));
    }
    std::vector<int> q_ptr(2, 0);
    q_ptr[1] = (int)items.size();
    if (persistent != 3 && n_queues == 1) {
        std::stable_sort(items.begin(), items.end(), [](const Item &x, const Item &y) { return x.cost > y.cost; });
    } else if (persistent == 1) {
        std::vector<long long> gwork(n_groups, 0);
        for (size_t i = 0; i < items.size(); ++i)
            gwork[item_group[i]] += items[i].cost;
        std::vector<int> gorder(n_groups);
        for (int g = 0; g < n_groups; ++g)
            gorder[g] = g;
        std::stable_sort(gorder.begin(), gorder.end(), [&](int x, int y) { return gwork[x] > gwork[y]; });
        std::vector<int> grank(n_groups), gq(n_groups, 0);
        std::vector<long long> qload(n_queues, 0);
        for (int r = 0; r < n_groups; ++r) {
            const int g = gorder[r];
            grank[g] = r;
            int q = 0;
            for (int k = 1; k < n_queues; ++k)
                if (qload[k] < qload[q])
                    q = k;
            gq[g] = q;
            qload[q] += gwork[g];
        }
        std::vector<int> idx(items.size());
        for (size_t i = 0; i < idx.size(); ++i)
            idx[i] = (int)i;
        std::stable_sort(idx.begin(), idx.end(), [&](int x, int y) {
            const int gx = item_group[x], gy = item_group[y];
            if (gq[gx] != gq[gy])
                return gq[gx] < gq[gy];
            if (grank[gx] != grank[gy])
                return grank[gx] < grank[gy];
            return items[x].cost > items[y].cost;
        });
        std::vector<Item> sorted(items.size());
        q_ptr.assign(n_queues + 1, 0);
        for (size_t i = 0; i < idx.size(); ++i) {
            sorted[i] = items[idx[i]];
            ++q_ptr[gq[item_group[idx[i]]] + 1];
        }
        for (int q = 0; q < n_queues; ++q)
            q_ptr[q + 1] += q_ptr[q];
        items.swap(sorted);
    }

    stkfmm_plan *p = new stkfmm_plan();
    p->uk = uk;
    p->prec = precision;
    p->T = T;
    p->TT = TT;
    p->KS = uker_ks(uk);
    p->KS1 = uker_ks1(uk);
    p->KT = uker_kt(uk);
    p->n_src = n_src;
    p->n_trg = n_trg;
    p->n_items = (int)items.size();
    p->n_tiles = (int)tiles.size();
    p->n_ranges = (int)ranges.size();
    p->persistent = persistent;
    p->n_queues = n_queues;
    int wpc = tn.waves_per_cu > 0 ? std::min(tn.waves_per_cu, bpc) : bpc;
    p->grid = (persistent == 1) ? (int)std::min<long long>(p->n_items, (long long)n_cu * wpc) : p->n_items;
    if (persistent == 3)
        p->group_items = group_items;
    p->broadcast = broadcast;
    p->kernel_prec = kernel_prec;
    if (broadcast) {
        int lmin = 0, lmax = 0;
        if (precision != STKFMM_PREC_FP64 && grp_level && n_groups > 0) {
            lmin = lmax = grp_level[0];
            for (int g = 1; g < n_groups; ++g) {
                lmin = std::min(lmin, grp_level[g]);
                lmax = std::max(lmax, grp_level[g]);
            }
        }
        p->level_min = lmin;
        p->n_copies = lmax - lmin + 1;
        p->rec_words = record_words(uk, precision);
    }
    p->pairs = pairs;
    (void)hipGetDevice(&p->device);

    int r;
    if ((r = upload(&p->d_items, items, "plan items")) || (r = upload(&p->d_ranges, ranges, "plan ranges")) ||
        (r = upload(&p->d_tiles, tiles, "plan tiles")) || (r = upload(&p->d_qptr, q_ptr, "plan queues"))) {
        plan_free(p);
        return r;
    }
    if (!segs.empty()) {
        p->sorted = true;
        p->n_sort_segs = (int)segs.size();
        p->n_sort_blks = (int)blks.size();
        p->self_ranges = self_ranges;
        p->n_src_segs = n_src_segs;
        p->n_src_blks = n_src_blks;
        if ((r = upload(&p->d_sort_segs, segs, "sort segments")) || (r = upload(&p->d_sort_blks, blks, "sort blocks"))) {
            plan_free(p);
            return r;
        }
    }
    const size_t partial_elems = (size_t)p->n_items * p->KT * p->TT;
    hipError_t e = hipSuccess;
    if (p->sorted) {
        const size_t ns = (size_t)n_src, nt = (size_t)n_trg;
        e = hipMalloc((void **)&p->d_sort_box, sizeof(double) * 6 * blks.size());
        if (e == hipSuccess)
            e = hipMalloc((void **)&p->d_sort_hist, sizeof(int) * (size_t)hist_len);
        if (e == hipSuccess)
            e = hipMalloc((void **)&p->d_sort_keys, sizeof(unsigned short) * std::max<size_t>(ns + nt, 1));
        if (e == hipSuccess)
            e = hipMalloc((void **)&p->d_src_idx, sizeof(int) * std::max<size_t>(ns, 1));
        if (e == hipSuccess)
            e = hipMalloc((void **)&p->d_trg_idx, sizeof(int) * std::max<size_t>(nt, 1));
        if (e == hipSuccess)
            e = hipMalloc((void **)&p->d_trg_row, sizeof(int) * std::max<size_t>(nt, 1));
        if (e == hipSuccess)
            e = hipMalloc((void **)&p->d_src_coord, sizeof(double) * 3 * std::max<size_t>(ns, 1));
        if (e == hipSuccess)
            e = hipMalloc((void **)&p->d_src_value, sizeof(double) * p->KS * std::max<size_t>(ns, 1));
        if (e == hipSuccess)
            e = hipMalloc((void **)&p->d_trg_coord, sizeof(double) * 3 * std::max<size_t>(nt, 1));
    }
    if (e == hipSuccess && partial_elems)
        e = hipMalloc((void **)&p->d_partial, sizeof(double) * partial_elems);
    if (e == hipSuccess)
        e = hipMalloc((void **)&p->d_counter, sizeof(int) * (size_t)p->n_queues);
    if (e == hipSuccess && p->broadcast && n_src > 0)
        e = hipMalloc((void **)&p->d_recs, sizeof(double) * (size_t)p->n_copies * n_src * p->rec_words);
    if (e != hipSuccess) {
        plan_free(p);
        return hip_error(e, "plan workspace");
    }
    *plan_out = p;
    return STKFMM_OK;
}

static bool arr_ok(const stkfmm_array *a) { return a && a->ptr && a->stride >= 0 && a->fstride >= 0; }

static int plan_enqueue(stkfmm_plan *p, const stkfmm_array *sc, const stkfmm_array *sv, const stkfmm_array *sv2,
                        const stkfmm_array *tc, const stkfmm_out_array *tv, const int *perm, hipStream_t s,
                        bool allow_timing) {
    if (p->n_items == 0)
        return STKFMM_OK;
    if (!arr_ok(sc) || !arr_ok(sv) || !arr_ok(tc) || !tv || !tv->ptr)
        return set_error(STKFMM_ERR_ARG, "NULL device array");
    if (p->KS1 < p->KS && !arr_ok(sv2))
        return set_error(STKFMM_ERR_ARG, "this micro-kernel needs the second source-value array");
    ItemArgs a;
    a.sc = DArr{sc->ptr, sc->stride, sc->fstride};
    a.v1 = DArr{sv->ptr, sv->stride, sv->fstride};
    a.v2 = (p->KS1 < p->KS) ? DArr{sv2->ptr, sv2->stride, sv2->fstride} : a.v1;
    a.tc = DArr{tc->ptr, tc->stride, tc->fstride};
    if (p->sorted) {
        // order the points of this execute, then run on the sorted SoA copies
        SortArgs so;
        so.segs = p->d_sort_segs;
        so.blks = p->d_sort_blks;
        so.sc = a.sc;
        so.v1 = a.v1;
        so.v2 = a.v2;
        so.tc = a.tc;
        so.ks = p->KS;
        so.ks1 = p->KS1;
        so.n_src = p->n_src;
        so.n_trg = p->n_trg;
        so.blk_box = p->d_sort_box;
        so.hist = p->d_sort_hist;
        so.keys = p->d_sort_keys;
        so.src_idx = p->d_src_idx;
        so.trg_idx = p->d_trg_idx;
        so.src_coord = p->d_src_coord;
        so.src_value = p->d_src_value;
        so.trg_coord = p->d_trg_coord;
        so.trg_row = p->d_trg_row;
        so.perm = perm;
        // targets are the sources: same ranges, and the same coordinate array at this execute
        so.shared = p->self_ranges && sc->ptr == tc->ptr && sc->stride == tc->stride && sc->fstride == tc->fstride;
        HIP_TRY(launch_sort_points(so, so.shared ? p->n_src_segs : p->n_sort_segs,
                                   so.shared ? p->n_src_blks : p->n_sort_blks, s),
                "point ordering launch");
        const long long ns = p->n_src, nt = p->n_trg;
        a.sc = DArr{p->d_src_coord, 1, ns};
        a.v1 = DArr{p->d_src_value, 1, ns};
        a.v2 = DArr{p->d_src_value + (long long)p->KS1 * ns, 1, ns};
        a.tc = so.shared ? a.sc : DArr{p->d_trg_coord, 1, nt};
        perm = p->d_trg_row;
    }
    a.items = p->d_items;
    a.ranges = p->d_ranges;
    a.n_items = p->n_items;
    a.counter = (p->persistent == 1) ? p->d_counter : nullptr;
    a.q_ptr = p->d_qptr;
    a.n_queues = p->n_queues;
    a.partial = p->d_partial;
    a.n_src_rec = p->n_src;
    ReduceArgs ra;
    ra.tiles = p->d_tiles;
    ra.n_tiles = p->n_tiles;
    ra.kt = p->KT;
    ra.tt = p->TT;
    ra.partial = p->d_partial;
    ra.out = OArr{tv->ptr, tv->stride, tv->fstride};
    ra.perm = perm;
    ra.scale = uker_scale(p->uk);
    const bool timed = allow_timing && p->timing;
    hipEvent_t *ev = timed ? &p->ev[3 * (p->n_timed % stkfmm_plan::kRing)] : nullptr;
    if (timed) {
        ++p->n_timed;
        HIP_TRY(hipEventRecord(ev[0], s), "event record");
    }
    if (p->persistent == 1)
        HIP_TRY(launch_fill_int(p->d_counter, p->n_queues, 0, s), "counter reset launch");
    if (p->broadcast) {
        StageArgs sa{a.sc, a.v1, a.v2, p->n_src, p->level_min, p->d_recs};
        HIP_TRY(launch_stage_records(p->uk, p->prec, sa, p->n_copies, s), "stage records launch");
    }
    auto launch = [&](int grid, const ItemArgs &ia) {
        return p->broadcast ? launch_items_sgpr(p->uk, p->prec, p->T, grid, ia, p->d_recs, p->level_min, s)
                            : launch_items(p->uk, p->kernel_prec, p->T, grid, ia, s);
    };
    if (p->persistent == 3) {
        // one launch per target group, in work-list order (the per-leaf-launch scheme)
        for (size_t g = 0; g + 1 < p->group_items.size(); ++g) {
            ItemArgs ag = a;
            ag.items = p->d_items + p->group_items[g];
            ag.n_items = p->group_items[g + 1] - p->group_items[g];
            HIP_TRY(launch(ag.n_items, ag), "p2p items launch");
        }
    } else {
        HIP_TRY(launch(p->grid, a), "p2p items launch");
    }
    if (timed)
        HIP_TRY(hipEventRecord(ev[1], s), "event record");
    // permuted rows with many tiles: one contiguous write per target (same sums, same order)
    if (perm && p->n_tiles >= 1024)
        HIP_TRY(launch_reduce_rows(ra, s), "p2p reduce launch");
    else
        HIP_TRY(launch_reduce(ra, s), "p2p reduce launch");
    if (timed)
        HIP_TRY(hipEventRecord(ev[2], s), "event record");
    return STKFMM_OK;
}

extern "C" int stkfmm_plan_execute(stkfmm_plan *p, const stkfmm_array *sc, const stkfmm_array *sv,
                                   const stkfmm_array *sv2, const stkfmm_array *tc, const stkfmm_out_array *tv,
                                   const int *perm, void *stream) {
    if (!p)
        return set_error(STKFMM_ERR_ARG, "plan is NULL");
    return plan_enqueue(p, sc, sv, sv2, tc, tv, perm, (hipStream_t)stream, true);
}

extern "C" int stkfmm_plan_capture_graph(stkfmm_plan *p, const stkfmm_array *sc, const stkfmm_array *sv,
                                         const stkfmm_array *sv2, const stkfmm_array *tc,
                                         const stkfmm_out_array *tv, const int *perm, void *stream) {
    if (!p)
        return set_error(STKFMM_ERR_ARG, "plan is NULL");
    hipStream_t s = (hipStream_t)stream;
    hipStream_t cap = nullptr;
    HIP_TRY(hipStreamCreateWithFlags(&cap, hipStreamNonBlocking), "capture stream");
    if (p->gexec) {
        (void)hipGraphExecDestroy(p->gexec);
        p->gexec = nullptr;
    }
    if (p->graph) {
        (void)hipGraphDestroy(p->graph);
        p->graph = nullptr;
    }
    hipError_t e = hipStreamBeginCapture(cap, hipStreamCaptureModeThreadLocal);
    int rc = STKFMM_OK;
    if (e == hipSuccess) {
        rc = plan_enqueue(p, sc, sv, sv2, tc, tv, perm, cap, false);
        hipGraph_t g = nullptr;
        hipError_t e2 = hipStreamEndCapture(cap, &g);
        if (rc == STKFMM_OK && e2 != hipSuccess)
            rc = hip_error(e2, "end capture");
        if (rc == STKFMM_OK) {
            p->graph = g;
            e2 = hipGraphInstantiate(&p->gexec, g, nullptr, nullptr, 0);
            if (e2 != hipSuccess)
                rc = hip_error(e2, "graph instantiate");
        } else if (g) {
            (void)hipGraphDestroy(g);
        }
    } else {
        rc = hip_error(e, "begin capture");
    }
    (void)hipStreamDestroy(cap);
    (void)s;
    return rc;
}

extern "C" int stkfmm_plan_launch_graph(stkfmm_plan *p, void *stream) {
    if (!p)
        return set_error(STKFMM_ERR_ARG, "plan is NULL");
    if (!p->gexec)
        return set_error(STKFMM_ERR_ARG, "no graph captured for this plan");
    HIP_TRY(hipGraphLaunch(p->gexec, (hipStream_t)stream), "graph launch");
    return STKFMM_OK;
}

extern "C" int stkfmm_plan_stats(const stkfmm_plan *p, long long *pairs, int *n_items, int *n_tiles, int *grid) {
    if (!p)
        return set_error(STKFMM_ERR_ARG, "plan is NULL");
    if (pairs)
        *pairs = p->pairs;
    if (n_items)
        *n_items = p->n_items;
    if (n_tiles)
        *n_tiles = p->n_tiles;
    if (grid)
        *grid = p->grid;
    return STKFMM_OK;
}

extern "C" int stkfmm_plan_set_timing(stkfmm_plan *p, int enable) {
    if (!p)
        return set_error(STKFMM_ERR_ARG, "plan is NULL");
    if (enable && p->ev.empty()) {
        p->ev.resize(3 * stkfmm_plan::kRing, nullptr);
        for (auto &e : p->ev)
            HIP_TRY(hipEventCreate(&e), "event create");
    }
    p->timing = enable != 0;
    p->n_timed = 0;
    return STKFMM_OK;
}

extern "C" int stkfmm_plan_kernel_ms(const stkfmm_plan *p, float *item_ms, float *reduce_ms, int *count) {
    if (!p)
        return set_error(STKFMM_ERR_ARG, "plan is NULL");
    const int n = (int)std::min<long long>(p->n_timed, stkfmm_plan::kRing);
    double sa = 0.0, sb = 0.0;
    int ok = 0;
    for (int k = 0; k < n; ++k) {
        float a = 0.f, b = 0.f;
        if (hipEventElapsedTime(&a, p->ev[3 * k], p->ev[3 * k + 1]) != hipSuccess ||
            hipEventElapsedTime(&b, p->ev[3 * k + 1], p->ev[3 * k + 2]) != hipSuccess)
            return set_error(STKFMM_ERR_HIP, "timing events not complete (synchronise the stream first)");
        sa += a;
        sb += b;
        ++ok;
    }
    if (item_ms)
        *item_ms = ok ? (float)(sa / ok) : -1.f;
    if (reduce_ms)
        *reduce_ms = ok ? (float)(sb / ok) : -1.f;
    if (count)
        *count = ok;
    return STKFMM_OK;
}

extern "C" int stkfmm_plan_destroy(stkfmm_plan *p) {
    plan_free(p);
    return STKFMM_OK;
}

// ---- direct sums (boundary 1) and Ker_t entries (boundary 2) --------------------------------------

namespace {
std::mutex g_mu;

// cache of single-group plans for direct sums, keyed by shape
struct DirectKey {
    int device, uk, prec, n_src, n_trg;
    bool operator==(const DirectKey &o) const {
        return device == o.device && uk == o.uk && prec == o.prec && n_src == o.n_src && n_trg == o.n_trg;
    }
};
std::list<std::pair<DirectKey, stkfmm_plan *>> g_direct_plans;
constexpr size_t kMaxDirectPlans = 16;

int direct_plan(int uk, int prec, int n_src, int n_trg, stkfmm_plan **out) {
    int dev = 0;
    HIP_TRY(hipGetDevice(&dev), "hipGetDevice");
    DirectKey key{dev, uk, prec, n_src, n_trg};
    for (auto it = g_direct_plans.begin(); it != g_direct_plans.end(); ++it) {
        if (it->first == key) {
            g_direct_plans.splice(g_direct_plans.begin(), g_direct_plans, it);
            *out = it->second;
            return STKFMM_OK;
        }
    }
    const int gb = 0, ge = n_trg, gp[2] = {0, 1}, rb = 0, re = n_src;
    // one target group.  From 16,384 sources and targets on, the points are ordered along a Morton
    // curve at every call (sort_points, T = 2): measured on MI355X, whole call including the ordering,
    // 65,536^2 PVel 3.94 -> 3.77 ms FP64 (LDS staging; 3.76 with scalar broadcast on a box where the
    // unordered call took 4.03) and 2.88 -> 2.70 ms mixed, 131,072^2 Stokeslet mixed 9.97 -> 8.43 ms,
    // 16,384^2 about even, 8192^2 slower (0.11 -> 0.14 ms: the ordering's fixed cost).
    // Below that: one target per lane under 32,768 targets (twice the waves for a small sum), and in
    // FP64 the scalar-broadcast staging at or above it (8192^2 LapPGrad 0.127 -> 0.100 ms FP64 and
    // 0.115 -> 0.096 ms mixed at T = 1).  Only the target-to-lane mapping and the staging change, not
    // the order in which a target sums its sources.
    stkfmm_tune tn{};
    const bool small = direct_targets_per_lane(n_trg) == 1;
    tn.targets_per_lane = direct_targets_per_lane(n_trg);
    tn.broadcast = (!small && prec == STKFMM_PREC_FP64) ? 1 : 0;
    if (n_trg >= kDirectSortMin && n_src >= kDirectSortMin) {
        // ordered points: T = 2, and in FP64 the scalar-broadcast records (65,536^2 PVel FP64 4.05 ->
        // 3.76 ms, 131,072^2 14.72 -> 14.24 ms against LDS staging of the ordered points)
        tn.sort_points = 1;
        tn.targets_per_lane = 2;
        tn.broadcast = (prec == STKFMM_PREC_FP64) ? 1 : 0;
    }
    stkfmm_plan *p = nullptr;
    int rc = stkfmm_plan_create(&p, uk, prec, n_src, n_trg, 1, &gb, &ge, nullptr, gp, &rb, &re, &tn);
    if (rc)
        return rc;
    g_direct_plans.emplace_front(key, p);
    while (g_direct_plans.size() > kMaxDirectPlans) {
        plan_free(g_direct_plans.back().second);
        g_direct_plans.pop_back();
    }
    *out = p;
    return STKFMM_OK;
}

// grow-only device buffers for the host-pointer entries
struct Workspace {
    int device = -1;
    // slot 0: host-pointer entries (inputs and output, packed); slots 1-3: unused; slot 4: adjoint
    // partial sums; slot 5: wall images
    double *buf[6] = {nullptr, nullptr, nullptr, nullptr, nullptr, nullptr};
    size_t cap[6] = {0, 0, 0, 0, 0, 0};
};
std::vector<Workspace> g_ws;

int ws_get(int slot, size_t n, double **out) {
    int dev = 0;
    HIP_TRY(hipGetDevice(&dev), "hipGetDevice");
    Workspace *w = nullptr;
    for (auto &x : g_ws)
        if (x.device == dev)
            w = &x;
    if (!w) {
        g_ws.emplace_back();
        w = &g_ws.back();
        w->device = dev;
    }
    if (w->cap[slot] < n) {
        (void)hipFree(w->buf[slot]);
        w->buf[slot] = nullptr;
        w->cap[slot] = 0;
        HIP_TRY(hipMalloc((void **)&w->buf[slot], sizeof(double) * std::max<size_t>(n, 1)), "workspace");
        w->cap[slot] = n;
    }
    *out = w->buf[slot];
    return STKFMM_OK;
}

int eval_device_uk(int uk, int prec, int n_src, const double *sc, const double *sv, int n_trg, const double *tc,
                   double *tv, hipStream_t s) {
    if (n_src < 0 || n_trg < 0)
        return set_error(STKFMM_ERR_ARG, "negative point count");
    if (n_src == 0 || n_trg == 0)
        return STKFMM_OK;
    if (!sc || !sv || !tc || !tv)
        return set_error(STKFMM_ERR_ARG, "NULL buffer");
    stkfmm_plan *p = nullptr;
    int rc = direct_plan(uk, prec, n_src, n_trg, &p);
    if (rc)
        return rc;
    const int ks = uker_ks(uk), ks1 = uker_ks1(uk), kt = uker_kt(uk);
    stkfmm_array a_sc{sc, 3, 1}, a_sv{sv, ks, 1}, a_sv2{sv + ks1, ks, 1}, a_tc{tc, 3, 1};
    stkfmm_out_array a_tv{tv, kt, 1};
    return plan_enqueue(p, &a_sc, &a_sv, &a_sv2, &a_tc, &a_tv, nullptr, s, false);
}

// grow-only pinned host buffer through which the host-pointer entries move their data (one per
// process, used under g_mu; every use ends with a synchronisation, so no copy is in flight when it
// is reallocated)
double *g_pinned = nullptr;
size_t g_pinned_cap = 0;

int pinned_get(size_t n, double **out) {
    if (g_pinned_cap < n) {
        if (g_pinned)
            (void)hipHostFree(g_pinned);
        g_pinned = nullptr;
        g_pinned_cap = 0;
        HIP_TRY(hipHostMalloc((void **)&g_pinned, sizeof(double) * std::max<size_t>(n, 1), hipHostMallocDefault),
                "pinned staging buffer");
        g_pinned_cap = n;
    }
    *out = g_pinned;
    return STKFMM_OK;
}

// A host-pointer evaluation: the three inputs are packed into the pinned buffer and moved in one
// copy; the device output starts at zero and comes back in one copy, after one synchronisation;
// the caller's trgValue += result is then done here (the same additions as on the device, since
// 0 + x is exact).  eval(d_sc, d_sv, d_tc, d_tv) enqueues the device work on the null stream.
template <class Eval>
int host_roundtrip(size_t ks, size_t kt, int n_src, const double *sc, const double *sv, int n_trg,
                   const double *tc, double *tv, Eval eval) {
    const size_t n_sc = 3 * (size_t)n_src, n_sv = ks * n_src, n_tc = 3 * (size_t)n_trg, n_tv = kt * n_trg;
    const size_t n_in = n_sc + n_sv + n_tc;
    double *h = nullptr, *d = nullptr;
    int rc;
    if ((rc = pinned_get(n_in + n_tv, &h)) || (rc = ws_get(0, n_in + n_tv, &d)))
        return rc;
    std::memcpy(h, sc, sizeof(double) * n_sc);
    std::memcpy(h + n_sc, sv, sizeof(double) * n_sv);
    std::memcpy(h + n_sc + n_sv, tc, sizeof(double) * n_tc);
    double *d_tv = d + n_in, *h_tv = h + n_in;
    HIP_TRY(hipMemcpyAsync(d, h, sizeof(double) * n_in, hipMemcpyHostToDevice, nullptr), "copy inputs");
    HIP_TRY(hipMemsetAsync(d_tv, 0, sizeof(double) * n_tv, nullptr), "clear output");
    rc = eval(d, d + n_sc, d + n_sc + n_sv, d_tv);
    if (rc) {
        (void)hipStreamSynchronize(nullptr);
        return rc;
    }
    HIP_TRY(hipMemcpyAsync(h_tv, d_tv, sizeof(double) * n_tv, hipMemcpyDeviceToHost, nullptr), "copy result");
    HIP_TRY(hipStreamSynchronize(nullptr), "synchronise");
    for (size_t i = 0; i < n_tv; ++i)
        tv[i] += h_tv[i];
    return STKFMM_OK;
}

int eval_host_uk(int uk, int prec, int n_src, const double *sc, const double *sv, int n_trg, const double *tc,
                 double *tv) {
    if (n_src < 0 || n_trg < 0)
        return set_error(STKFMM_ERR_ARG, "negative point count");
    if (n_src == 0 || n_trg == 0)
        return STKFMM_OK;
    if (!sc || !sv || !tc || !tv)
        return set_error(STKFMM_ERR_ARG, "NULL buffer");
    return host_roundtrip(uker_ks(uk), uker_kt(uk), n_src, sc, sv, n_trg, tc, tv,
                          [&](const double *d_sc, const double *d_sv, const double *d_tc, double *d_tv) {
                              return eval_device_uk(uk, prec, n_src, d_sc, d_sv, n_trg, d_tc, d_tv, nullptr);
                          });
}

int resolve(unsigned kernel, unsigned p2p, int precision, int *uk) {
    if (precision != STKFMM_PREC_FP64 && precision != STKFMM_PREC_MIXED)
        return set_error(STKFMM_ERR_ARG, "unknown precision %d", precision);
    return stkfmm_get_uker(kernel, p2p, uk);
}
} // namespace

extern "C" int stkfmm_p2p_eval(unsigned kernel, unsigned p2p, int precision, int nSrc, const double *srcCoord,
                               const double *srcValue, int nTrg, const double *trgCoord, double *trgValue) {
    int uk;
    int rc = resolve(kernel, p2p, precision, &uk);
    if (rc)
        return rc;
    std::lock_guard<std::mutex> lk(g_mu);
    return eval_host_uk(uk, precision, nSrc, srcCoord, srcValue, nTrg, trgCoord, trgValue);
}

extern "C" int stkfmm_evaluate_kernel(unsigned kernel, int nThreads, unsigned p2p, int nSrc, const double *srcCoord,
                                      const double *srcValue, int nTrg, const double *trgCoord, double *trgValue) {
    (void)nThreads;
    return stkfmm_p2p_eval(kernel, p2p, STKFMM_PREC_FP64, nSrc, srcCoord, srcValue, nTrg, trgCoord, trgValue);
}

extern "C" int stkfmm_p2p_eval_device(unsigned kernel, unsigned p2p, int precision, int nSrc,
                                      const double *d_srcCoord, const double *d_srcValue, int nTrg,
                                      const double *d_trgCoord, double *d_trgValue, void *stream) {
    int uk;
    int rc = resolve(kernel, p2p, precision, &uk);
    if (rc)
        return rc;
    std::lock_guard<std::mutex> lk(g_mu);
    return eval_device_uk(uk, precision, nSrc, d_srcCoord, d_srcValue, nTrg, d_trgCoord, d_trgValue,
                          (hipStream_t)stream);
}

// ---- wall images (StkWallFMM, Lib/src/StkWallFMM.cpp:154-292) ----------------------------------------

namespace {
// the constituent direct sums of StkWallFMM::evalStokes / evalRPY, with the layouts of wall.hip
int wall_device(unsigned kernel, int prec, int n, const double *sc, const double *sv, int m, const double *tc,
                double *tv, hipStream_t s) {
    const size_t N = (size_t)n, M = (size_t)m;
    double *w = nullptr;
    int rc;
    if (kernel == STKFMM_KERNEL_Stokes) {
        if ((rc = ws_get(5, 6 * N + 6 * N + 2 * N + 3 * N + 2 * N + 17 * M, &w)))
            return rc;
        WallStokes L;
        L.c2 = w;
        L.vS = L.c2 + 6 * N;
        L.vL1 = L.vS + 6 * N;
        L.vD = L.vL1 + 2 * N;
        L.vL2 = L.vD + 3 * N;
        L.oS = L.vL2 + 2 * N;
        L.oL1D = L.oS + 3 * M;
        L.oL2 = L.oL1D + 4 * M;
        HIP_TRY(launch_wall_prep_stokes(n, sc, sv, L, s), "wall prep launch");
        HIP_TRY(launch_wall_zero(L.oS, (long long)(17 * M), s), "wall zero launch");
        const double *cimg = L.c2 + 3 * N;
        if ((rc = eval_device_uk(STKFMM_UKER_STOKES_VEL, prec, 2 * n, L.c2, L.vS, m, tc, L.oS, s)) ||
            (rc = eval_device_uk(STKFMM_UKER_LAPLACE_PGRAD, prec, 2 * n, L.c2, L.vL1, m, tc, L.oL1D, s)) ||
            (rc = eval_device_uk(STKFMM_UKER_LAPLACE_DIPOLEPGRAD, prec, n, cimg, L.vD, m, tc, L.oL1D, s)) ||
            (rc = eval_device_uk(STKFMM_UKER_LAPLACE_PGRADGRAD, prec, 2 * n, L.c2, L.vL2, m, tc, L.oL2, s)))
            return rc;
        HIP_TRY(launch_wall_assemble_stokes(m, tc, L, tv, s), "wall assemble launch");
        return STKFMM_OK;
    }
    if ((rc = ws_get(5, 6 * N + 8 * N + 2 * N + 3 * N + 2 * N + 6 * N + 9 * N + 30 * M, &w)))
        return rc;
    WallRpy L;
    L.c2 = w;
    L.vR = L.c2 + 6 * N;
    L.vLS = L.vR + 8 * N;
    L.vLD = L.vLS + 2 * N;
    L.vLSZ = L.vLD + 3 * N;
    L.vLDZ = L.vLSZ + 2 * N;
    L.vQ = L.vLDZ + 6 * N;
    L.oR = L.vQ + 9 * N;
    L.oSD = L.oR + 6 * M;
    L.oSDZ = L.oSD + 10 * M;
    L.oQ = L.oSDZ + 4 * M;
    HIP_TRY(launch_wall_prep_rpy(n, sc, sv, L, s), "wall prep launch");
    HIP_TRY(launch_wall_zero(L.oR, (long long)(30 * M), s), "wall zero launch");
    const double *cimg = L.c2 + 3 * N;
    if ((rc = eval_device_uk(STKFMM_UKER_RPY_ULAPU, prec, 2 * n, L.c2, L.vR, m, tc, L.oR, s)) ||
        (rc = eval_device_uk(STKFMM_UKER_LAPLACE_PGRADGRAD, prec, 2 * n, L.c2, L.vLS, m, tc, L.oSD, s)) ||
        (rc = eval_device_uk(STKFMM_UKER_LAPLACE_DIPOLEPGRADGRAD, prec, n, cimg, L.vLD, m, tc, L.oSD, s)) ||
        (rc = eval_device_uk(STKFMM_UKER_LAPLACE_PGRAD, prec, 2 * n, L.c2, L.vLSZ, m, tc, L.oSDZ, s)) ||
        (rc = eval_device_uk(STKFMM_UKER_LAPLACE_DIPOLEPGRAD, prec, 2 * n, L.c2, L.vLDZ, m, tc, L.oSDZ, s)) ||
        (rc = eval_device_uk(STKFMM_UKER_LAPLACE_QUADPGRADGRAD, prec, n, cimg, L.vQ, m, tc, L.oQ, s)))
        return rc;
    HIP_TRY(launch_wall_assemble_rpy(m, tc, L, tv, s), "wall assemble launch");
    return STKFMM_OK;
}

int wall_check(unsigned kernel, int precision, int n, int m) {
    if (kernel != STKFMM_KERNEL_Stokes && kernel != STKFMM_KERNEL_RPY)
        return set_error(STKFMM_ERR_KERNEL, "Kernel not supported"); // StkWallFMM.cpp:125-127
    if (precision == STKFMM_PREC_MIXED)
        // the assembled velocity cancels the image sums (u = 0 on the wall), which amplifies the FP32
        // rounding of each constituent sum (~1e-7 of its magnitude) to ~1e-5 of the result: above the
        // 1e-6 bound of the mixed mode, so the wall images are FP64 only
        return set_error(STKFMM_ERR_UNSUPPORTED,
                         "wall images: mixed precision is not supported (the image sums cancel near the wall "
                         "and the FP32 evaluation would exceed 1e-6 relative error); use FP64");
    if (precision != STKFMM_PREC_FP64)
        return set_error(STKFMM_ERR_ARG, "unknown precision %d", precision);
    if (n < 0 || m < 0 || n > (1 << 29))
        return set_error(STKFMM_ERR_ARG, "bad point count");
    return STKFMM_OK;
}
} // namespace

extern "C" int stkfmm_wall_eval_device(unsigned kernel, int precision, int nSL, const double *d_srcCoord,
                                       const double *d_srcValue, int nTrg, const double *d_trgCoord,
                                       double *d_trgValue, void *stream) {
    int rc = wall_check(kernel, precision, nSL, nTrg);
    if (rc)
        return rc;
    if (nSL == 0 || nTrg == 0)
        return STKFMM_OK;
    if (!d_srcCoord || !d_srcValue || !d_trgCoord || !d_trgValue)
        return set_error(STKFMM_ERR_ARG, "NULL buffer");
    std::lock_guard<std::mutex> lk(g_mu);
    return wall_device(kernel, precision, nSL, d_srcCoord, d_srcValue, nTrg, d_trgCoord, d_trgValue,
                       (hipStream_t)stream);
}

extern "C" int stkfmm_wall_evaluate(unsigned kernel, int precision, int nSL, const double *srcCoord,
                                    const double *srcValue, int nTrg, const double *trgCoord, double *trgValue) {
    int rc = wall_check(kernel, precision, nSL, nTrg);
    if (rc)
        return rc;
    if (nSL == 0 || nTrg == 0)
        return STKFMM_OK;
    if (!srcCoord || !srcValue || !trgCoord || !trgValue)
        return set_error(STKFMM_ERR_ARG, "NULL buffer");
    const size_t ks = kernel == STKFMM_KERNEL_Stokes ? 3 : 4, kt = kernel == STKFMM_KERNEL_Stokes ? 3 : 6;
    std::lock_guard<std::mutex> lk(g_mu);
    return host_roundtrip(ks, kt, nSL, srcCoord, srcValue, nTrg, trgCoord, trgValue,
                          [&](const double *d_sc, const double *d_sv, const double *d_tc, double *d_tv) {
                              return wall_device(kernel, precision, nSL, d_sc, d_sv, nTrg, d_tc, d_tv, nullptr);
                          });
}

extern "C" int stkfmm_ker_eval(int uker, const double *r_src, int src_cnt, const double *v_src, int dof,
                               const double *r_trg, int trg_cnt, double *v_trg) {
    if (uker < 0 || uker >= STKFMM_UKER_COUNT)
        return set_error(STKFMM_ERR_ARG, "unknown micro-kernel %d", uker);
    if (dof != 1)
        return set_error(STKFMM_ERR_UNSUPPORTED, "dof=%d: only dof=1 (as passed by FMMData.cpp:280) is supported",
                         dof);
    std::lock_guard<std::mutex> lk(g_mu);
    return eval_host_uk(uker, STKFMM_PREC_FP64, src_cnt, r_src, v_src, trg_cnt, r_trg, v_trg);
}

#define STKFMM_KER_T(NAME, UK)                                                                     \
    extern "C" void stkfmm_ker_##NAME(double *r_src, int src_cnt, double *v_src, int dof, double *r_trg,  \
                                      int trg_cnt, double *v_trg, void *mem_mgr) {                 \
        (void)mem_mgr;                                                                             \
        if (stkfmm_ker_eval(UK, r_src, src_cnt, v_src, dof, r_trg, trg_cnt, v_trg) != STKFMM_OK)  \
            fprintf(stderr, "stkfmm_ker_" #NAME ": %s\n", g_last_error.c_str());                   \
    }

STKFMM_KER_T(laplace_p, STKFMM_UKER_LAPLACE_P)
STKFMM_KER_T(laplace_dipolep, STKFMM_UKER_LAPLACE_DIPOLEP)
STKFMM_KER_T(laplace_pgrad, STKFMM_UKER_LAPLACE_PGRAD)
STKFMM_KER_T(laplace_dipolepgrad, STKFMM_UKER_LAPLACE_DIPOLEPGRAD)
STKFMM_KER_T(laplace_pgradgrad, STKFMM_UKER_LAPLACE_PGRADGRAD)
STKFMM_KER_T(laplace_dipolepgradgrad, STKFMM_UKER_LAPLACE_DIPOLEPGRADGRAD)
STKFMM_KER_T(stokes_vel, STKFMM_UKER_STOKES_VEL)
STKFMM_KER_T(stokes_pvel, STKFMM_UKER_STOKES_PVEL)
STKFMM_KER_T(stokes_doublepvel, STKFMM_UKER_STOKES_DOUBLEPVEL)
STKFMM_KER_T(stokes_regvel, STKFMM_UKER_STOKES_REGVEL)
STKFMM_KER_T(rpy_u, STKFMM_UKER_RPY_U)
STKFMM_KER_T(rpy_ulapu, STKFMM_UKER_RPY_ULAPU)
STKFMM_KER_T(stk_ulapu, STKFMM_UKER_STK_ULAPU)
STKFMM_KER_T(stokes_pvelgrad, STKFMM_UKER_STOKES_PVELGRAD)
STKFMM_KER_T(stokes_doublepvelgrad, STKFMM_UKER_STOKES_DOUBLEPVELGRAD)
STKFMM_KER_T(stokes_pvellaplacian, STKFMM_UKER_STOKES_PVELLAPLACIAN)
STKFMM_KER_T(stokes_doublelaplacian, STKFMM_UKER_STOKES_DOUBLELAPLACIAN)
STKFMM_KER_T(stokes_traction, STKFMM_UKER_STOKES_TRACTION)
STKFMM_KER_T(stokes_doubletraction, STKFMM_UKER_STOKES_DOUBLETRACTION)
STKFMM_KER_T(stokes_regftvelomega, STKFMM_UKER_STOKES_REGFTVELOMEGA)
STKFMM_KER_T(stokes_velomega, STKFMM_UKER_STOKES_VELOMEGA)
STKFMM_KER_T(laplace_quadpgradgrad, STKFMM_UKER_LAPLACE_QUADPGRADGRAD)
STKFMM_KER_T(laplace_quadp, STKFMM_UKER_LAPLACE_QUADP)
STKFMM_KER_T(stokes_regftvel, STKFMM_UKER_STOKES_REGFTVEL)

// ---- adjoint ------------------------------------------------------------------------------------------

extern "C" int stkfmm_p2p_vjp_device(int uker, int precision, int nSrc, const double *d_srcCoord,
                                     const double *d_srcValue, int nTrg, const double *d_trgCoord,
                                     const double *d_w, double *d_g_src_value, double *d_g_src_coord,
                                     double *d_g_trg_coord, void *stream) {
    if (uker < 0 || uker >= STKFMM_UKER_COUNT || !uker_has_vjp(uker))
        return set_error(STKFMM_ERR_UNSUPPORTED, "no adjoint for micro-kernel %d", uker);
    if (precision != STKFMM_PREC_FP64 && precision != STKFMM_PREC_MIXED)
        return set_error(STKFMM_ERR_ARG, "unknown precision %d", precision);
    if (nSrc < 0 || nTrg < 0)
        return set_error(STKFMM_ERR_ARG, "negative point count");
    if ((nSrc > 0 && (!d_srcCoord || !d_srcValue)) || (nTrg > 0 && (!d_trgCoord || !d_w)))
        return set_error(STKFMM_ERR_ARG, "NULL buffer");
    int n_cu = 0;
    int rc = device_cu_count(&n_cu);
    if (rc)
        return rc;
    const int ks = uker_ks(uker);
    // about 64 waves per CU over a (target blocks) x (source chunks) grid, close to square; partials
    // kept within ~2 GiB by fewer, larger blocks for large point counts.  Measured on MI355X,
    // 65,536^2 (ms, FP64 / mixed), waves per CU 16 / 20 / 24 / 32 / 40 / 48 / 64 / 96: PVel 10.7 /
    // 12.0 / 11.3 / 11.6 / 10.7 / 11.1 / 10.4 / 10.7 and 7.5 / 8.5 / 7.9 / 8.2 / 7.6 / 7.7 / 7.1 / 7.3
    // -- the blocks fill whole rounds of the resident waves at 16 and 64 (FP64: 4,096 and 16,384
    // blocks on 4,096 wave slots) and leave a part-empty last round at 32 (86 x 86 blocks); the same
    // order holds for fused PVel and RPY (and again at four / eight targets per lane: 24 / 36 / 48 /
    // 64 / 96 waves per CU, PVel 5.90 / 5.86 / 6.23 / 5.47 / 5.60 ms mixed, 8.64 / 8.74 / 8.89 / 8.26
    // / 8.66 FP64).  T = 1 is slower.  Mixed mode takes more targets per lane
    // once there are 16 tiles of them: fewer LDS reads and source-sum rotations per pair.  65,536^2
    // mixed, same box, T = 2 -> 4 (ms): Stokeslet 6.03 -> 5.52, fused PVel 12.56 -> 11.99, PVel 6.45 ->
    // 6.32, RPY + Laplacian 8.25 -> 8.16; T = 4 -> 8 (3 waves per SIMD, value-gradient terms summed
    // per step): PVel 6.03 -> 5.59, RPY + Laplacian 7.58 -> 7.46, but Stokeslet 5.15 -> 5.26 and fused
    // PVel 11.39 -> 12.08 (spills at 168 VGPRs), so T = 8 only for the first two.  FP64 takes four
    // targets per lane as well (3 waves per SIMD at 152-168 VGPRs), except for the double-layer and
    // fused PVel, which spill there: 65,536^2, same box, T = 2 -> 4 (ms): PVel 8.74 -> 8.35, Stokeslet
    // 8.45 -> 7.95, LapPGrad 6.29 -> 5.90, RPY + Laplacian 11.88 -> 11.62 (fused PVel 18.1 -> 21.7).
    VjpLayout l;
    l.T = (nTrg >= 16 * kWave) ? 2 : 1;
    if (nTrg >= 16 * 4 * kWave) {
        if (precision == STKFMM_PREC_MIXED) {
            const bool t8 = uker == STKFMM_UKER_STOKES_PVEL || uker == STKFMM_UKER_RPY_ULAPU;
            l.T = (t8 && nTrg >= 16 * 8 * kWave) ? 8 : 4;
        } else if (vjp_fp64_four_targets(uker)) {
            l.T = 4;
        }
    }
    if (const char *e = std::getenv("STKFMM_VJP_T")) { // tuning and tests: 1, 2, 4, or 8 (mixed)
        const long t = std::strtol(e, nullptr, 10);
        const bool t4 = t == 4 && (precision == STKFMM_PREC_MIXED || vjp_fp64_four_targets(uker));
        if (t != 1 && t != 2 && !t4 && !(t == 8 && precision == STKFMM_PREC_MIXED))
            return set_error(STKFMM_ERR_ARG,
                             "STKFMM_VJP_T must be 1, 2, 4 (FP64: not for the double-layer or fused PVel) or 8 "
                             "(mixed only) (got %s)", e);
        l.T = (int)t;
    }
    const long long ttiles = std::max(1, (nTrg + kWave * l.T - 1) / (kWave * l.T));
    const long long stiles = std::max(1, (nSrc + kSrcTile - 1) / kSrcTile);
    long long per_cu = 64;
    if (const char *e = std::getenv("STKFMM_VJP_WAVES_PER_CU")) // tuning and tests: other layouts
        per_cu = std::max(1L, std::strtol(e, nullptr, 10));
    const long long waves = per_cu * n_cu;
    const long long root = (long long)std::ceil(std::sqrt((double)waves));
    long long nc = std::min(stiles, root);
    long long nb = std::min(ttiles, (waves + nc - 1) / nc);
    nc = std::min(stiles, (waves + nb - 1) / nb);
    const long long budget = (2LL << 30) / 8;
    while ((nc * 3 * nTrg + nb * (3 + ks) * (long long)nSrc) > budget && (nb > 1 || nc > 1)) {
        if (nb >= nc)
            nb = (nb + 1) / 2;
        else
            nc = (nc + 1) / 2;
    }
    l.trg_tiles_per_block = (int)((ttiles + nb - 1) / nb);
    l.n_blocks = (int)((ttiles + l.trg_tiles_per_block - 1) / l.trg_tiles_per_block);
    l.src_tiles_per_chunk = (int)((stiles + nc - 1) / nc);
    l.n_chunks = (int)((stiles + l.src_tiles_per_chunk - 1) / l.src_tiles_per_chunk);
    std::lock_guard<std::mutex> lk(g_mu);
    double *partial = nullptr;
    const size_t need = (size_t)l.n_chunks * 3 * nTrg + (size_t)l.n_blocks * (3 + ks) * nSrc;
    if ((rc = ws_get(4, std::max<size_t>(need, 1), &partial)))
        return rc;
    VjpArgs a{d_srcCoord, d_srcValue, d_trgCoord, d_w, nSrc, nTrg, uker_scale(uker),
              d_g_src_value, d_g_src_coord, d_g_trg_coord};
    HIP_TRY(launch_vjp(uker, precision, a, l, partial, (hipStream_t)stream), "vjp launch");
    return STKFMM_OK;
}
