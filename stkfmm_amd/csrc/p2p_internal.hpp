// p2p_internal.hpp -- types shared by the HIP kernels (p2p_kernels.hip) and the host code
// (p2p_capi.cpp) of libstkfmm_p2p.so.  Not part of the public C ABI.
#pragma once

#include <hip/hip_runtime.h>

namespace stkp2p {

// Target-per-lane tiling: one workgroup = one wavefront of 64 lanes; each lane owns T targets,
// so a work item covers a tile of 64*T consecutive targets.  Sources are staged kSrcTile at a time
// through LDS.
constexpr int kWave = 64;
constexpr int kSrcTile = 64;
// mixed mode: FP32 partial sums over this many sources before they are added to the FP64
// accumulator of the target (build option STKP2P_MIXED_ACC_BLOCK)
#ifndef STKP2P_MIXED_ACC_BLOCK
#define STKP2P_MIXED_ACC_BLOCK 16
#endif
constexpr int kMixedAccBlock = STKP2P_MIXED_ACC_BLOCK;
static_assert(kSrcTile % kMixedAccBlock == 0 && kWave % kMixedAccBlock == 0,
              "source tiles and chunks hold whole partial-sum blocks");

// one work item: a target tile against a list of source ranges
struct Item {
    int trg_begin;  // first target (index into the target arrays)
    int trg_count;  // targets in the tile, 1..64*T
    int rng_begin;  // source ranges [rng_begin, rng_end) of the plan's range array
    int rng_end;
    int slot;       // partial-sum slot this item writes
    int level;      // octree level of the target group (mixed-mode scaling by 2^level)
    int cost;       // sources in the item (scheduling key)
    int lsplit;     // log2 of the lanes per target: a tile of n targets uses 2^lsplit lanes per
                    // target when n * 2^lsplit <= 64*T; each of those lanes takes every
                    // 2^lsplit-th staged source and the partial sums meet by wavefront shuffles
};

// a target tile and the partial-sum slots that cover it
struct Tile {
    int trg_begin;
    int trg_count;
    int slot_begin;
    int slot_end;
};

// element (i, k) of a strided device array is p[i*s + k*fs]
struct DArr {
    const double *p;
    long long s;
    long long fs;
};
struct OArr {
    double *p;
    long long s;
    long long fs;
};

struct ItemArgs {
    DArr sc;   // source coordinates
    DArr v1;   // source values [0, KS1)
    DArr v2;   // source values [KS1, KS) (fused single+double layer only)
    DArr tc;   // target coordinates
    const Item *items;
    const int2 *ranges;   // (src_begin, src_end)
    int n_items;
    int *counter;         // persistent mode: next item of each queue [n_queues]; null: one
                          // workgroup per item
    const int *q_ptr;     // queue q holds items [q_ptr[q], q_ptr[q+1]) (persistent mode)
    int n_queues;
    double *partial;      // [slot][KT][64*T]
    long long n_src_rec;  // scalar-broadcast variant: sources per staged record copy
};

// scalar-broadcast variant: per-source records staged once per execute (one copy per level)
struct StageArgs {
    DArr sc, v1, v2;
    long long n_src;
    int level_min;        // copy c holds sources scaled by 2^(level_min + c) (mixed mode)
    double *rec;          // [copy][n_src][words]
};

struct ReduceArgs {
    const Tile *tiles;
    int n_tiles;
    int kt;
    int tt;               // 64*T
    const double *partial;
    OArr out;
    const int *perm;      // output row of sorted target i (null: identity)
    double scale;         // ScaleFactor of the kernel
};

// XCD-contiguous work mapping.  Workgroups of a launch are dealt to MI355X's 8 XCDs round-robin
// (workgroup b runs on XCD b % 8), each XCD with its own L2.  A launch of 8*ceil(n/8) workgroups that
// takes unit xcd_unit(b, n) gives XCD x the contiguous units [x*q, (x+1)*q), q = ceil(n/8), so work on
// neighbouring data (one instance's tiles or points) shares one L2; the result is -1 for the padding
// workgroups past n.
constexpr int kXcds = 8;
__host__ __device__ inline int xcd_grid(int n) { return kXcds * ((n + kXcds - 1) / kXcds); }
__device__ inline int xcd_unit(int b, int n) {
    const int q = (n + kXcds - 1) / kXcds;
    const int u = (b % kXcds) * q + b / kXcds;
    return u < n ? u : -1;
}

// Morton ordering of a plan's points (morton.hip).  A segment is a group's target range (kind 1) or a
// distinct source range (kind 0), ordered on a 2^level-per-axis grid over its own bounding box; a
// block is a run of at most kSortBlock consecutive points of one segment.
constexpr int kSortBlock = 512;
constexpr int kSortMaxLevel = 4;
constexpr int kSortMaxCells = 1 << (3 * kSortMaxLevel);
struct SortSeg {
    int begin, end;         // point range
    int blk_begin, blk_end; // its blocks
    int level;              // grid level, 1..kSortMaxLevel
    int kind;               // 0 sources, 1 targets
};
struct SortBlk {
    int seg;
    int begin, end;
    long long hist_off; // first of its 8^level histogram entries
};
struct SortArgs {
    const SortSeg *segs;
    const SortBlk *blks;
    DArr sc, v1, v2, tc;    // the caller's arrays
    int ks, ks1;
    long long n_src, n_trg;
    double *blk_box;        // [n_blks][6]
    int *hist;              // per block, per cell: count, then first output position
    unsigned short *keys;   // [n_src + n_trg]: cell of each point
    int *src_idx, *trg_idx; // sorted position -> original point
    double *src_coord;      // SoA copies in sorted order: [3][n_src]
    double *src_value;      // [ks][n_src]
    double *trg_coord;      // [3][n_trg]
    int *trg_row;           // output row of sorted target p (the caller's perm applied)
    const int *perm;        // the caller's output permutation (null: identity)
    int shared;             // targets are the sources (same points, same ranges): the source
                            // ordering serves both, and the source gather writes trg_row
};
// a segment's grid level: about 64 points per cell, at most 3 levels up to 2^18 points (finer
// grids do not make the tiles measurably more compact there, and cost histogram work), 4 above
inline int sort_level(long long n) {
    int l = 1;
    while (l < kSortMaxLevel && (64LL << (3 * l)) < n)
        ++l;
    return (n <= (1LL << 18) && l > 3) ? 3 : l;
}
hipError_t launch_sort_points(const SortArgs &a, int n_segs, int n_blks, hipStream_t s);

// adjoint of a direct sum (one pass over the pairs, vjp_kernels.hpp)
struct VjpArgs {
    const double *sc;  // AoS source coords [n_src*3]
    const double *sv;  // AoS source values [n_src*KS]
    const double *tc;  // AoS target coords [n_trg*3]
    const double *w;   // AoS cotangent [n_trg*KT]
    int n_src;
    int n_trg;
    double scale;
    double *g_sv;      // [n_src*KS] (may be null)
    double *g_sc;      // [n_src*3]  (may be null)
    double *g_tc;      // [n_trg*3]  (may be null)
};

// Every launch entry of one micro-kernel (all PREC x T instantiations), filled by
// make_uker_ops<UK>() in the translation unit that instantiates that micro-kernel's kernels.
struct UkerOps {
    hipError_t (*launch)(int prec, int T, int grid, const ItemArgs &a, hipStream_t s);
    hipError_t (*occ)(int prec, int T, int *blocks);
    hipError_t (*launch_sgpr)(int prec, int T, int grid, const ItemArgs &a, const double *recs, int level_min,
                              hipStream_t s);
    hipError_t (*occ_sgpr)(int prec, int T, int *blocks);
    hipError_t (*stage)(int prec, const StageArgs &a, int n_copies, hipStream_t s);
    int words[2]; // scalar-broadcast record size in 8-byte words, FP64 / mixed
    int ks, ks1, kt;
};

constexpr int kNumCoreUker = 14;   // micro-kernels 0..13 (p2p_kernels.hip)
constexpr int kNumUkerTotal = 25;  // 14..24 in p2p_family.hip

// Sources per work item when the caller does not choose: about 16 items per concurrently resident
// wave keeps the tail of the cost-sorted queue short; total_cost = sum over target tiles of the
// tile's source count, concurrent = resident waves of the item kernel on the device.
inline long long default_max_src(long long total_cost, long long concurrent) {
    const long long target_items = 16LL * concurrent;
    long long m = (total_cost + target_items - 1) / (target_items > 1 ? target_items : 1);
    m = ((m + 511) / 512) * 512;
    return m > 512 ? m : 512;
}

// Targets per lane of a single-group direct sum: one below 32,768 targets (twice the waves for a
// small sum), else two.
inline int direct_targets_per_lane(int n_trg) { return n_trg < 32768 ? 1 : 2; }
// Direct sums with at least this many sources and targets order their points (sort_points) in the
// plan-based entries.
constexpr int kDirectSortMin = 16384;

// lanes per target of a tile of n targets (log2): the largest power of two, at most 16, that still
// fits the tile's 64*T lanes
__host__ __device__ inline int tile_lsplit(int n, int TT) {
    int l = 0;
    while (l < 4 && (n << (l + 1)) <= TT)
        ++l;
    return l;
}

// table row of micro-kernel uk, or nullptr
const UkerOps *uker_ops(int uk);
const UkerOps *uker_ops_family(int uk);

// sets the message stkfmm_last_error() returns; returns code
int report_error(int code, const char *msg);

// kernel-table queries (host)
int uker_ks(int uk);
int uker_ks1(int uk);
int uker_kt(int uk);
double uker_scale(int uk);
bool uker_has_vjp(int uk);
// the FP64 adjoint of micro-kernel uk runs four targets per lane (3 waves per SIMD, no spill):
// measured on MI355X, 65,536^2, four targets against two: 3.7-7.6 % faster for the Laplace, Stokeslet,
// RegVel, single-layer PVel and RPY kernels; the double-layer PVel spills and measured the same, the
// fused PVel spills and measured 20 % slower, so those two keep two
constexpr bool vjp_fp64_four_targets(int uk) { return uk != 8 && uk != 13; }

// launchers (host); return hipSuccess or the launch error.  launch_items / items_occupancy also
// take prec = kPrecMixedFar: the mixed-mode LDS item kernel with the far-tile displacement (T = 2)
constexpr int kPrecMixedFar = 2;
hipError_t launch_items(int uk, int prec, int T, int grid, const ItemArgs &a, hipStream_t s);
hipError_t launch_reduce(const ReduceArgs &a, hipStream_t s);
// the same sums with one thread per target writing all its components (for permuted outputs, whose
// rows are scattered: one contiguous write per target instead of one per component)
hipError_t launch_reduce_rows(const ReduceArgs &a, hipStream_t s);
hipError_t launch_fill_int(int *p, int n, int value, hipStream_t s); // n <= 256
hipError_t items_occupancy(int uk, int prec, int T, int *blocks_per_cu);
// scalar-broadcast variant
int record_words(int uk, int prec);
hipError_t launch_stage_records(int uk, int prec, const StageArgs &a, int n_copies, hipStream_t s);
hipError_t launch_items_sgpr(int uk, int prec, int T, int grid, const ItemArgs &a, const double *recs,
                             int level_min, hipStream_t s);
hipError_t items_sgpr_occupancy(int uk, int prec, int T, int *blocks_per_cu);
// equivalent / check surfaces (surface.hip)
__host__ __device__ inline int surface_count(int p) { return 6 * (p - 1) * (p - 1) + 2; }
void surface_host(int p, const double *c, double alpha, int depth, double *coord);
hipError_t launch_box_surfaces(int p, double alpha, int n_boxes, const double *corner, const int *level, double *out,
                               hipStream_t s);
// wall image system (wall.hip): device views of the workspace of one evaluation.  n sources
// (2n with images: [originals; images]), m targets; v* are source values, o* accumulated outputs.
struct WallStokes {
    double *c2, *vS, *vL1, *vD, *vL2; // [2n*3], [2n*3], [2n], [n*3] (at the images), [2n]
    double *oS, *oL1D, *oL2;          // [m*3], [m*4], [m*10]
};
struct WallRpy {
    double *c2, *vR, *vLS, *vLD, *vLSZ, *vLDZ, *vQ; // [2n*3],[2n*4],[2n],[n*3],[2n],[2n*3],[n*9]
    double *oR, *oSD, *oSDZ, *oQ;                   // [m*6], [m*10], [m*4], [m*10]
};
hipError_t launch_wall_prep_stokes(int n, const double *sc, const double *sv, const WallStokes &w, hipStream_t s);
hipError_t launch_wall_prep_rpy(int n, const double *sc, const double *sv, const WallRpy &w, hipStream_t s);
hipError_t launch_wall_zero(double *p, long long n, hipStream_t s);
hipError_t launch_wall_assemble_stokes(int m, const double *tc, const WallStokes &w, double *out, hipStream_t s);
hipError_t launch_wall_assemble_rpy(int m, const double *tc, const WallRpy &w, double *out, hipStream_t s);
// adjoint work decomposition: T targets per lane; the grid is n_blocks target blocks (of
// trg_tiles_per_block tiles of 64 T targets) x n_chunks source chunks (of src_tiles_per_chunk tiles of
// 64 sources); partial holds n_chunks * 3 * n_trg + n_blocks * (3 + KS) * n_src doubles
struct VjpLayout {
    int T;
    int trg_tiles_per_block, n_blocks;
    int src_tiles_per_chunk, n_chunks;
};
hipError_t launch_vjp(int uk, int prec, const VjpArgs &a, const VjpLayout &l, double *partial, hipStream_t s);
} // namespace stkp2p
