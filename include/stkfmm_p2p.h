/*
 * stkfmm_p2p.h -- C ABI of the MI355X (gfx950) P2P near-field library.
 *
 * Drop-in for STKFMM's point-to-point kernel evaluation path.  All entry points are extern "C",
 * take plain pointers and sizes, and return a status code (0 = success, negative = error; the
 * message is available from stkfmm_last_error()).  Reference interfaces each entry replaces are
 * cited as file:line relative to the STKFMM source tree.
 *
 * Conventions (as in the reference):
 *   - r = x_trg - x_src; coincident pairs (r2 == 0) contribute exactly 0;
 *   - results are ADDED to the target buffer (STKFMM.hpp:95);
 *   - the kernel scale factor (1/(4 pi) Laplace, 1/(8 pi) Stokes family) is applied;
 *   - host-pointer entries use AoS layouts: coords [x0,y0,z0,x1,...], values kdim-interleaved.
 *
 * Thread safety: calls that share a plan, or host-pointer calls from several host threads at once,
 * must be serialised by the caller (the reference object is not safe concurrently either,
 * Stk3DFMM.cpp:90-108).
 */
#ifndef STKFMM_P2P_H_
#define STKFMM_P2P_H_

#ifdef __cplusplus
extern "C" {
#endif

/* ---- enumerations ------------------------------------------------------------------------- */

/* stkfmm::KERNEL bit values, STKFMM_common.hpp:46-62 */
enum stkfmm_kernel {
    STKFMM_KERNEL_LapPGrad = 1,
    STKFMM_KERNEL_LapPGradGrad = 2,
    STKFMM_KERNEL_LapQPGradGrad = 4,
    STKFMM_KERNEL_Stokes = 8,
    STKFMM_KERNEL_RPY = 16,
    STKFMM_KERNEL_StokesRegVel = 32,
    STKFMM_KERNEL_StokesRegVelOmega = 64,
    STKFMM_KERNEL_PVel = 128,
    STKFMM_KERNEL_PVelGrad = 256,
    STKFMM_KERNEL_PVelLaplacian = 512,
    STKFMM_KERNEL_Traction = 1024
};

/* stkfmm::PPKERNEL, STKFMM_common.hpp:37-41 */
enum stkfmm_ppkernel { STKFMM_PPKERNEL_SLS2T = 1, STKFMM_PPKERNEL_DLS2T = 2, STKFMM_PPKERNEL_L2T = 4 };

/* arithmetic mode of the HIP kernels (no reference counterpart: the reference is FP64 only) */
enum stkfmm_precision {
    STKFMM_PREC_FP64 = 0,  /* IEEE FP64 throughout (matches the reference formulas to ~1e-15) */
    STKFMM_PREC_MIXED = 1  /* FP64 displacement + accumulation, FP32 rsqrt (+1 Newton) and polynomial;
                              source tiles far from a work item's targets form the displacement in FP32
                              from box-centred coordinates (error <= 5 * 2^-24 r per component) */
};

/* micro-kernels: one per uKerEval of the reference (pvfmm::Kernel<double>::Ker_t entries) */
enum stkfmm_uker {
    STKFMM_UKER_LAPLACE_P = 0,               /* laplace_p               1 -> 1   LaplaceLayerKernel.hpp:41-65   */
    STKFMM_UKER_LAPLACE_DIPOLEP = 1,         /* laplace_dipolep         3 -> 1   LaplaceLayerKernel.hpp:98-113  */
    STKFMM_UKER_LAPLACE_PGRAD = 2,           /* laplace_pgrad           1 -> 4   LaplaceLayerKernel.hpp:67-96   */
    STKFMM_UKER_LAPLACE_DIPOLEPGRAD = 3,     /* laplace_dipolepgrad     3 -> 4   LaplaceLayerKernel.hpp:115-136 */
    STKFMM_UKER_LAPLACE_PGRADGRAD = 4,       /* laplace_pgradgrad       1 -> 10  LaplaceLayerKernel.hpp:138-166 */
    STKFMM_UKER_LAPLACE_DIPOLEPGRADGRAD = 5, /* laplace_dipolepgradgrad 3 -> 10  LaplaceLayerKernel.hpp:168-200 */
    STKFMM_UKER_STOKES_VEL = 6,              /* PVFMM stokes_vel        3 -> 3   StokesLayerKernel.hpp:38-42    */
    STKFMM_UKER_STOKES_PVEL = 7,             /* stokes_pvel             4 -> 4   StokesSingleLayerKernel.hpp:27-47 */
    STKFMM_UKER_STOKES_DOUBLEPVEL = 8,       /* stokes_doublepvel       9 -> 4   StokesDoubleLayerKernel.hpp:25-57 */
    STKFMM_UKER_STOKES_REGVEL = 9,           /* stokes_regvel           4 -> 3   StokesRegSingleLayerKernel.hpp:17-41 */
    STKFMM_UKER_RPY_U = 10,                  /* rpy_u                   4 -> 3   RPYKernel.hpp:17-45 */
    STKFMM_UKER_RPY_ULAPU = 11,              /* rpy_ulapu               4 -> 6   RPYKernel.hpp:52-90 */
    STKFMM_UKER_STK_ULAPU = 12,              /* stk_ulapu               3 -> 6   RPYKernel.hpp:97-129 */
    STKFMM_UKER_STOKES_PVEL_FUSED = 13,      /* stokes_pvel + stokes_doublepvel at one point, 4+9 -> 4 */
    STKFMM_UKER_STOKES_PVELGRAD = 14,        /* stokes_pvelgrad         4 -> 16  StokesSingleLayerKernel.hpp:54-111 */
    STKFMM_UKER_STOKES_DOUBLEPVELGRAD = 15,  /* stokes_doublepvelgrad   9 -> 16  StokesDoubleLayerKernel.hpp:64-143 */
    STKFMM_UKER_STOKES_PVELLAPLACIAN = 16,   /* stokes_pvellaplacian    4 -> 7   StokesSingleLayerKernel.hpp:118-155 */
    STKFMM_UKER_STOKES_DOUBLELAPLACIAN = 17, /* stokes_doublelaplacian  9 -> 7   StokesDoubleLayerKernel.hpp:150-206 */
    STKFMM_UKER_STOKES_TRACTION = 18,        /* stokes_traction         4 -> 9   StokesSingleLayerKernel.hpp:162-195 */
    STKFMM_UKER_STOKES_DOUBLETRACTION = 19,  /* stokes_doubletraction   9 -> 9   StokesDoubleLayerKernel.hpp:213-273 */
    STKFMM_UKER_STOKES_REGFTVELOMEGA = 20,   /* stokes_regftvelomega    7 -> 6   StokesRegSingleLayerKernel.hpp:96-151 */
    STKFMM_UKER_STOKES_VELOMEGA = 21,        /* stokes_velomega         3 -> 6   StokesRegSingleLayerKernel.hpp:158-202 */
    STKFMM_UKER_LAPLACE_QUADPGRADGRAD = 22,  /* laplace_quadpgradgrad   9 -> 10  LaplaceLayerKernel.hpp:234-312 */
    STKFMM_UKER_LAPLACE_QUADP = 23,          /* laplace_quadp           9 -> 1   LaplaceLayerKernel.hpp:202-232 */
    STKFMM_UKER_STOKES_REGFTVEL = 24,        /* stokes_regftvel         7 -> 3   StokesRegSingleLayerKernel.hpp:48-89 */
    STKFMM_UKER_COUNT = 25
};

/* status codes */
enum stkfmm_status {
    STKFMM_OK = 0,
    STKFMM_ERR_KERNEL = -1,      /* unknown KERNEL (reference: exit(1), STKFMM.cpp:53-55) */
    STKFMM_ERR_PPKERNEL = -2,    /* sub-kernel absent (reference: message + return, FMMData.cpp:270-273) */
    STKFMM_ERR_ARG = -3,         /* bad size, pointer or layout */
    STKFMM_ERR_HIP = -4,         /* HIP runtime error */
    STKFMM_ERR_UNSUPPORTED = -5, /* valid request the library does not implement (e.g. dof != 1) */
    STKFMM_ERR_INACTIVE = -6     /* kernel not active in the object (reference: exit(1), STKFMM.cpp:92-95) */
};

/* last error message of the calling thread ("" when none) */
const char *stkfmm_last_error(void);

/* ---- kernel table ------------------------------------------------------------------------- */

/* dims = {kdimSL, kdimDL, kdimTrg}.  Replaces Stk3DFMM_get_kernel_dimension (Stk3DFMM-c.cpp:18-20)
 * and stkfmm::getKernelDimension (STKFMM.cpp:33-40). */
int stkfmm_get_kernel_dimension(unsigned kernel, int *dims);

/* kernel name as ker_name (STKFMM.cpp:42-46); buf receives a NUL-terminated string */
int stkfmm_get_kernel_name(unsigned kernel, char *buf, int buflen);

/* micro-kernel a (KERNEL, PPKERNEL) pair evaluates: the k_s2t->ker_poten / k_s2t->dbl_layer_poten /
 * k_l2t->ker_poten selection of FMMData::evaluateKernel (FMMData.cpp:261-268). */
int stkfmm_get_uker(unsigned kernel, unsigned p2p, int *uker);

/* source and target dimension of a micro-kernel */
int stkfmm_uker_dims(int uker, int *kdim_src, int *kdim_trg);

/* ---- boundary 1: direct O(N^2) evaluation ------------------------------------------------- */

/* Replaces stkfmm::STKFMM::evaluateKernel (STKFMM.hpp:108-110, STKFMM.cpp:88-99, body
 * FMMData.cpp:252-283).  Host AoS buffers; trgValue += result.  nThreads is accepted for
 * signature compatibility and ignored (the work runs on the current HIP device).  FP64. */
int stkfmm_evaluate_kernel(unsigned kernel, int nThreads, unsigned p2p, int nSrc, const double *srcCoord,
                           const double *srcValue, int nTrg, const double *trgCoord, double *trgValue);

/* Same as stkfmm_evaluate_kernel with a precision choice. */
int stkfmm_p2p_eval(unsigned kernel, unsigned p2p, int precision, int nSrc, const double *srcCoord,
                    const double *srcValue, int nTrg, const double *trgCoord, double *trgValue);

/* Device-resident variant: AoS device buffers, enqueued on `stream` (hipStream_t, 0 = null stream),
 * asynchronous.  No host memory is touched.  The device entries (this one, stkfmm_wall_eval_device,
 * stkfmm_p2p_vjp_device) keep per-device scratch (cached plans of each shape, workspaces) that
 * calls share: calls on one device must be ordered (one stream, or synchronised), or they race. */
int stkfmm_p2p_eval_device(unsigned kernel, unsigned p2p, int precision, int nSrc, const double *d_srcCoord,
                           const double *d_srcValue, int nTrg, const double *d_trgCoord, double *d_trgValue,
                           void *stream);

/* No counterpart in the reference: stkfmm_p2p_eval_device for callers that cannot allocate device
 * memory (code under HIP graph capture, tight serving loops, harnesses that forbid hipMalloc).  Same
 * arguments, semantics and kernels; it calls no hipMalloc/hipFree.  Its work list and partial sums
 * live in a fixed 64 MiB device array per GPU, held from the library's load until the process ends,
 * so a sum whose targets alone need more (about 2M targets at 4 output values, 512K at 16) returns
 * STKFMM_ERR_ARG.  Below 16,384 sources or targets, and while the plan-based entry's partial sums fit
 * the array, the results are bitwise identical to it; otherwise the same pairs are summed in another
 * grouping (the plan-based entry orders larger sums' points along a Morton curve; this one keeps the
 * caller's order, and cuts the sources into fewer chunks when its array is the limit).
 * Every call on one GPU uses that GPU's array: calls that overlap on one device (different streams,
 * or several host threads without ordering) race and corrupt each other's results. */
int stkfmm_p2p_eval_device_noalloc(unsigned kernel, unsigned p2p, int precision, int nSrc,
                                   const double *d_srcCoord, const double *d_srcValue, int nTrg,
                                   const double *d_trgCoord, double *d_trgValue, void *stream);

/* ---- boundary 2: pvfmm::Kernel<double>::Ker_t micro-kernel entries ------------------------- */

/* Each has the Ker_t signature (r_src, src_cnt, v_src, dof, r_trg, trg_cnt, v_trg, mem_mgr) whose
 * arity and argument order are read off the calls at FMMData.cpp:261,280-281; the typedef itself is
 * PVFMM's (not vendored).  Host AoS buffers, v_trg += ScaleFactor * sum, FP64, dof must be 1 (the
 * only value STKFMM passes).  mem_mgr is ignored.  Errors are reported through stkfmm_last_error()
 * because Ker_t returns void. */
void stkfmm_ker_laplace_p(double *r_src, int src_cnt, double *v_src, int dof, double *r_trg, int trg_cnt,
                          double *v_trg, void *mem_mgr);
void stkfmm_ker_laplace_dipolep(double *r_src, int src_cnt, double *v_src, int dof, double *r_trg, int trg_cnt,
                                double *v_trg, void *mem_mgr);
void stkfmm_ker_laplace_pgrad(double *r_src, int src_cnt, double *v_src, int dof, double *r_trg, int trg_cnt,
                              double *v_trg, void *mem_mgr);
void stkfmm_ker_laplace_dipolepgrad(double *r_src, int src_cnt, double *v_src, int dof, double *r_trg,
                                    int trg_cnt, double *v_trg, void *mem_mgr);
void stkfmm_ker_laplace_pgradgrad(double *r_src, int src_cnt, double *v_src, int dof, double *r_trg,
                                  int trg_cnt, double *v_trg, void *mem_mgr);
void stkfmm_ker_laplace_dipolepgradgrad(double *r_src, int src_cnt, double *v_src, int dof, double *r_trg,
                                        int trg_cnt, double *v_trg, void *mem_mgr);
void stkfmm_ker_stokes_vel(double *r_src, int src_cnt, double *v_src, int dof, double *r_trg, int trg_cnt,
                           double *v_trg, void *mem_mgr);
void stkfmm_ker_stokes_pvel(double *r_src, int src_cnt, double *v_src, int dof, double *r_trg, int trg_cnt,
                            double *v_trg, void *mem_mgr);
void stkfmm_ker_stokes_doublepvel(double *r_src, int src_cnt, double *v_src, int dof, double *r_trg,
                                  int trg_cnt, double *v_trg, void *mem_mgr);
void stkfmm_ker_stokes_regvel(double *r_src, int src_cnt, double *v_src, int dof, double *r_trg, int trg_cnt,
                              double *v_trg, void *mem_mgr);
void stkfmm_ker_rpy_u(double *r_src, int src_cnt, double *v_src, int dof, double *r_trg, int trg_cnt,
                      double *v_trg, void *mem_mgr);
void stkfmm_ker_rpy_ulapu(double *r_src, int src_cnt, double *v_src, int dof, double *r_trg, int trg_cnt,
                          double *v_trg, void *mem_mgr);
void stkfmm_ker_stk_ulapu(double *r_src, int src_cnt, double *v_src, int dof, double *r_trg, int trg_cnt,
                          double *v_trg, void *mem_mgr);
void stkfmm_ker_stokes_pvelgrad(double *r_src, int src_cnt, double *v_src, int dof, double *r_trg, int trg_cnt,
                                double *v_trg, void *mem_mgr);
void stkfmm_ker_stokes_doublepvelgrad(double *r_src, int src_cnt, double *v_src, int dof, double *r_trg,
                                      int trg_cnt, double *v_trg, void *mem_mgr);
void stkfmm_ker_stokes_pvellaplacian(double *r_src, int src_cnt, double *v_src, int dof, double *r_trg,
                                     int trg_cnt, double *v_trg, void *mem_mgr);
void stkfmm_ker_stokes_doublelaplacian(double *r_src, int src_cnt, double *v_src, int dof, double *r_trg,
                                       int trg_cnt, double *v_trg, void *mem_mgr);
void stkfmm_ker_stokes_traction(double *r_src, int src_cnt, double *v_src, int dof, double *r_trg, int trg_cnt,
                                double *v_trg, void *mem_mgr);
void stkfmm_ker_stokes_doubletraction(double *r_src, int src_cnt, double *v_src, int dof, double *r_trg,
                                      int trg_cnt, double *v_trg, void *mem_mgr);
void stkfmm_ker_stokes_regftvelomega(double *r_src, int src_cnt, double *v_src, int dof, double *r_trg,
                                     int trg_cnt, double *v_trg, void *mem_mgr);
void stkfmm_ker_stokes_velomega(double *r_src, int src_cnt, double *v_src, int dof, double *r_trg, int trg_cnt,
                                double *v_trg, void *mem_mgr);
void stkfmm_ker_laplace_quadpgradgrad(double *r_src, int src_cnt, double *v_src, int dof, double *r_trg,
                                      int trg_cnt, double *v_trg, void *mem_mgr);
void stkfmm_ker_laplace_quadp(double *r_src, int src_cnt, double *v_src, int dof, double *r_trg, int trg_cnt,
                              double *v_trg, void *mem_mgr);
void stkfmm_ker_stokes_regftvel(double *r_src, int src_cnt, double *v_src, int dof, double *r_trg, int trg_cnt,
                                double *v_trg, void *mem_mgr);

/* generic form of the entries above, status-returning */
int stkfmm_ker_eval(int uker, const double *r_src, int src_cnt, const double *v_src, int dof,
                    const double *r_trg, int trg_cnt, double *v_trg);

/* ---- batched evaluation over a work list ---------------------------------------------------- */

/* Strided view of a device array of points: element (i, k) is ptr[i*stride + k*fstride].
 * AoS: stride = kdim, fstride = 1.  SoA: stride = 1, fstride = n. */
typedef struct stkfmm_array {
    const double *ptr;
    long long stride;
    long long fstride;
} stkfmm_array;

typedef struct stkfmm_out_array {
    double *ptr;
    long long stride;
    long long fstride;
} stkfmm_out_array;

/* Tuning knobs (all optional: 0 means "library default / autotuned value"). */
typedef struct stkfmm_tune {
    int targets_per_lane; /* 1, 2 or 4 targets held by each lane */
    int max_src_per_item; /* sources per work item before a target tile's list is split */
    int persistent;       /* 1: persistent waves pull items from a cost-sorted queue (default);
                             2: one launch, one workgroup per item, in queue order;
                             3: one launch per target group (per-leaf launches, no queue) */
    int waves_per_cu;     /* persistent grid = CUs * waves_per_cu (capped by occupancy) */
    int split_tiles;      /* 0: a tile with fewer targets than lanes*T spreads each target over
                             several lanes (default); -1: one lane per target always */
    int broadcast;        /* 0: sources staged through LDS per tile (default); 1: per-source records
                             staged in HBM once per execute, read with scalar loads */
    int queues;           /* persistent mode: number of work queues (default 1: one cost-sorted queue).
                             8 on MI355X = one per XCD (workgroups are dispatched to the 8 XCDs
                             round-robin): all items of a target group go to one queue, so the group's
                             sources stay in that XCD's L2; a wave drains its own queue, then steals
                             from the others.  Measured on the 1M PVel U-list: 22% fewer L2 fetches,
                             1.5% longer (the kernel is VALU-bound) */
    int sort_points;      /* 1: at every execute, order each group's targets and each group's source
                             range along a Morton curve of their own bounding box (on the device, into
                             plan-owned copies; results still go to the original target rows), so that
                             target tiles are compact and most source tiles skip the coincidence mask
                             (mixed mode: and form the displacement in FP32, the far-tile form).  For
                             work lists of unsorted points (direct sums, independent instances); needs
                             exactly one source range per group, pairwise disjoint target ranges and
                             source ranges that are pairwise disjoint or identical.  0: as given */
} stkfmm_tune;

typedef struct stkfmm_plan stkfmm_plan;

/* Build a plan from a work list of target groups (U-list leaves or independent instances).
 * Group g owns targets [grp_trg_begin[g], grp_trg_end[g]) and the source ranges
 * r in [grp_rng_ptr[g], grp_rng_ptr[g+1]): sources [rng_src_begin[r], rng_src_end[r]).
 * grp_level[g] (may be NULL) is the octree level of the group; in MIXED mode the pair arithmetic
 * runs in coordinates scaled by 2^level (exact) and the outputs are rescaled by each component's
 * homogeneity degree -- the scaleSrc/scaleTrg rule of FMMData.cpp:285-470.
 * Every target index must lie in [0, n_trg) and every source index in [0, n_src); the arrays given to
 * stkfmm_plan_execute must hold at least that many points (checked here, trusted there).
 * Host arrays; the plan copies what it needs to the current device.  tune may be NULL. */
int stkfmm_plan_create(stkfmm_plan **plan, int uker, int precision, int n_src, int n_trg, int n_groups,
                       const int *grp_trg_begin, const int *grp_trg_end, const int *grp_level,
                       const int *grp_rng_ptr, const int *rng_src_begin, const int *rng_src_end,
                       const stkfmm_tune *tune);

/* Evaluate the plan: trg_value(perm[i], :) += ScaleFactor * sum over the work list (perm may be
 * NULL: identity).  src_value holds the first kdim values of each source; for
 * STKFMM_UKER_STOKES_PVEL_FUSED src_value holds the 4 single-layer values and src_value2 the 9
 * double-layer values of the same point (src_value2 is ignored for other micro-kernels).
 * Device arrays, enqueued on `stream`, asynchronous. */
int stkfmm_plan_execute(stkfmm_plan *plan, const stkfmm_array *src_coord, const stkfmm_array *src_value,
                        const stkfmm_array *src_value2, const stkfmm_array *trg_coord,
                        const stkfmm_out_array *trg_value, const int *d_trg_perm, void *stream);

/* Capture plan_execute with these exact arguments into a HIP graph held by the plan; subsequent
 * stkfmm_plan_launch_graph calls replay it.  Re-capturing replaces the previous graph. */
int stkfmm_plan_capture_graph(stkfmm_plan *plan, const stkfmm_array *src_coord, const stkfmm_array *src_value,
                              const stkfmm_array *src_value2, const stkfmm_array *trg_coord,
                              const stkfmm_out_array *trg_value, const int *d_trg_perm, void *stream);
int stkfmm_plan_launch_graph(stkfmm_plan *plan, void *stream);

/* statistics: pairs = sum over groups of n_trg * n_src; n_items = work items; n_tiles = target tiles */
int stkfmm_plan_stats(const stkfmm_plan *plan, long long *pairs, int *n_items, int *n_tiles, int *grid);

/* Kernel timing with HIP events on the execution stream.  set_timing(plan, 1) starts recording
 * (and clears earlier records); each later stkfmm_plan_execute records the duration of its item
 * kernel(s) and of its reduction.  kernel_ms returns their averages over the recorded executes (the
 * last 256) and how many there were; the stream must have been synchronised. */
int stkfmm_plan_set_timing(stkfmm_plan *plan, int enable);
int stkfmm_plan_kernel_ms(const stkfmm_plan *plan, float *item_ms, float *reduce_ms, int *count);

int stkfmm_plan_destroy(stkfmm_plan *plan);

/* ---- adaptive octree and U-list (host) ------------------------------------------------------ */

/* Adaptive octree over points in [0,1)^3: a box is split while it holds more than leaf_cap points
 * (the maxPts rule of FMMData::setupTree, FMMData.cpp:111-113).  U-list of leaf B = every leaf whose
 * closed box touches B's (any level, B included).  Points are sorted along the Morton curve; each
 * leaf is a contiguous range of the sorted order. */
typedef struct stkfmm_tree stkfmm_tree;
int stkfmm_tree_build(stkfmm_tree **tree, int n, const double *coord, int leaf_cap);
/* sizes: leaves, U-list source ranges (after merging Morton-adjacent leaves), deepest level, pairs */
int stkfmm_tree_info(const stkfmm_tree *tree, int *n_leaves, int *n_ranges, int *max_level, long long *pairs);
/* perm[i] = original index of sorted point i; leaf arrays [n_leaves]; ulist_ptr [n_leaves+1];
 * rng_begin/rng_end [n_ranges] (sorted-order indices) */
int stkfmm_tree_get(const stkfmm_tree *tree, int *perm, int *leaf_begin, int *leaf_end, int *leaf_level,
                    int *ulist_ptr, int *rng_begin, int *rng_end);
/* lower corner of every leaf box [3*n_leaves] (the box of a level-L leaf has edge 0.5^L) */
int stkfmm_tree_leaf_corners(const stkfmm_tree *tree, double *corner);
/* All nodes (non-empty boxes, root = 0): counts, and per node its lower corner [3*n], level, sorted
 * point range [begin, end), leaf id (-1: internal) and parent (-1: root).  Any output may be NULL. */
int stkfmm_tree_node_info(const stkfmm_tree *tree, int *n_nodes, int *n_w);
int stkfmm_tree_get_nodes(const stkfmm_tree *tree, double *corner, int *level, int *begin, int *end, int *leaf_id,
                          int *parent);
/* W- and X-lists of the adaptive tree (PVFMM's M2T and S2L near-field lists, not vendored; the
 * kernel-independent FMM's definitions): W(leaf B) = nodes deeper than B that do not touch B but whose
 * parent does; X(node C) = the leaves whose W-list holds C.  CSR: w_ptr [n_leaves+1] into w_node
 * [n_w] (node indices); x_ptr [n_nodes+1] into x_leaf [n_w] (leaf ids). */
int stkfmm_tree_get_wx(const stkfmm_tree *tree, int *w_ptr, int *w_node, int *x_ptr, int *x_leaf);
int stkfmm_tree_destroy(stkfmm_tree *tree);

/* ---- equivalent / check surfaces (S2M, L2T, W- and X-list point sets) ------------------------- */

/* stkfmm::surface (Lib/include/STKFMM/stkfmm_helpers.hpp:46-85): the 6(p-1)^2+2 points of the
 * surface of the cube [c, c + alpha 0.5^depth]^3 in the reference's (PVFMM's) order, written to
 * coord[3*count].  Returns STKFMM_OK; count = stkfmm_surface_count(p) (p >= 2). */
int stkfmm_surface_count(int p);
int stkfmm_surface(int p, const double *c, double alpha, int depth, double *coord);
/* Device: the surfaces of n_boxes octree boxes at once.  Box b has lower corner d_corner[3b..3b+2]
 * and level d_level[b]; its surface is centred on the box with half-width alpha * 0.5^(level+1)
 * (stkfmm_surface with c = corner - 0.5^(level+1) (alpha - 1), PVFMM's placement; 1.05 / 2.95 are the
 * equivalent / check radii, M2L/SVD_pvfmm.hpp:37-38).  d_coord [n_boxes][count][3]; bit-identical
 * to stkfmm_surface. */
int stkfmm_box_surfaces_device(int p, double alpha, int n_boxes, const double *d_corner, const int *d_level,
                               double *d_coord, void *stream);
/* k_s2m sub-kernel of a KERNEL (the kernel evaluated from sources to an upward check surface):
 * double_layer 0 / 1.  STKFMM_ERR_PPKERNEL when the kernel has no double layer. */
int stkfmm_get_s2m_uker(unsigned kernel, int double_layer, int *uker);
/* k_s2l (sources -> downward check surface, the X-list kernel) and k_m2t (upward equivalent
 * densities -> targets, the W-list kernel; single layer) of a KERNEL, from the BuildKernel lists. */
int stkfmm_get_s2l_uker(unsigned kernel, int double_layer, int *uker);
int stkfmm_get_m2t_uker(unsigned kernel, int *uker);

/* ---- box mapping and value scaling (the frame an FMM evaluation works in) --------------------- */

/* STKFMM::scaleCoord + wrapCoord (Lib/src/STKFMM.cpp:113-150): coord <- (coord - origin) * scaleFactor,
 * then x - floor(x) on the periodic axes (pbc = PAXIS: 0 none, 1 x, 2 xy, 3 xyz).  In place. */
int stkfmm_scale_coord(int n, double *coord, const double *origin, double scaleFactor, unsigned pbc);
/* FMMData::scaleSrc (Lib/src/FMMData.cpp:285-317): srcDLValue *= scaleFactor; the single-layer
 * components with the dimension of a length (TrD of the PVel family, eps of StokesRegVel, b of RPY,
 * torque and eps of StokesRegVelOmega) *= scaleFactor.  In place; kernel = KERNEL bit. */
int stkfmm_scale_src(unsigned kernel, int nSL, double *srcSLValue, int nDL, double *srcDLValue, double scaleFactor);
/* FMMData::scaleTrg (Lib/src/FMMData.cpp:319-470): every output component times the power of
 * scaleFactor of its dimension (e.g. PVel: p * sF^2, u * sF).  In place. */
int stkfmm_scale_trg(unsigned kernel, int nTrg, double *trgValue, double scaleFactor);
/* The same on device arrays (origin stays a host pointer), enqueued on stream; bit-identical. */
int stkfmm_scale_coord_device(int n, double *d_coord, const double *origin, double scaleFactor, unsigned pbc,
                              void *stream);
int stkfmm_scale_src_device(unsigned kernel, int nSL, double *d_srcSLValue, int nDL, double *d_srcDLValue,
                            double scaleFactor, void *stream);
int stkfmm_scale_trg_device(unsigned kernel, int nTrg, double *d_trgValue, double scaleFactor, void *stream);

/* ---- wall images (StkWallFMM) ---------------------------------------------------------------- */

/* The no-slip wall image system of StkWallFMM (Lib/src/StkWallFMM.cpp:154-292, kernels Stokes and
 * RPY) evaluated with direct sums: trg_value += u over all sources and their images.  Replaces
 * StkWallFMM::setPoints + evaluateFMM (:40-128) for a direct evaluation: the wall is the plane z = 0
 * of the caller's coordinates (the image of (x, y, z) is (x, y, -z)); sources and targets are expected
 * at z >= 0; no periodicity.  Stokes: source f(3) -> u(3); RPY: source [f(3), b] -> [u(3), lap u(3)].
 * AoS arrays; other kernels: STKFMM_ERR_KERNEL ("Kernel not supported", :125-127).  FP64 only:
 * precision STKFMM_PREC_MIXED returns STKFMM_ERR_UNSUPPORTED, because the image sums cancel the
 * direct sums near the wall (u = 0 on it) and amplify FP32 rounding beyond the mixed mode's 1e-6. */
int stkfmm_wall_evaluate(unsigned kernel, int precision, int nSL, const double *srcCoord, const double *srcValue,
                         int nTrg, const double *trgCoord, double *trgValue);
int stkfmm_wall_eval_device(unsigned kernel, int precision, int nSL, const double *d_srcCoord,
                            const double *d_srcValue, int nTrg, const double *d_trgCoord, double *d_trgValue,
                            void *stream);

/* ---- adjoints ----------------------------------------------------------------------------- */

/* Gradients of L = sum_t w_t . u_t for u = direct-sum P2P (same conventions as
 * stkfmm_p2p_eval_device), for STKFMM_UKER_LAPLACE_P, _LAPLACE_PGRAD, _STOKES_VEL, _STOKES_REGVEL
 * (including d/d eps), _RPY_U, _RPY_ULAPU, _STK_ULAPU, _STOKES_PVEL, _STOKES_DOUBLEPVEL and
 * _STOKES_PVEL_FUSED.  No reference counterpart (the reference
 * has no derivatives); callers: the torch.autograd.Function in the Python binding.
 * Device AoS arrays.  Outputs are OVERWRITTEN: g_src_value [nSrc*ks], g_src_coord [nSrc*3],
 * g_trg_coord [nTrg*3]; any of them may be NULL to skip it.  For the fused kernel src_value holds
 * the 13 values [f(3), TrD, D(9)] of each point.  One pass over the pairs; deterministic.  The
 * environment variable STKFMM_VJP_WAVES_PER_CU (default 64) sets how many waves per compute unit the
 * pass is split into (more: better balance, larger partial-sum workspace). */
int stkfmm_p2p_vjp_device(int uker, int precision, int nSrc, const double *d_srcCoord, const double *d_srcValue,
                          int nTrg, const double *d_trgCoord, const double *d_w, double *d_g_src_value,
                          double *d_g_src_coord, double *d_g_trg_coord, void *stream);

/* ---- device information ------------------------------------------------------------------- */
int stkfmm_device_info(int *n_cu, int *warp_size, char *arch, int arch_len);

#ifdef __cplusplus
}
#endif

#endif /* STKFMM_P2P_H_ */
