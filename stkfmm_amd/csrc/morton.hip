// morton.hip -- spatial ordering of a plan's points before its item kernel runs.
//
// A direct sum hands the library its points in any order.  A target tile of 64*T consecutive targets
// then spans the whole point cloud, no source tile is ever clear of it, and every pair pays the
// coincidence mask (and, in mixed mode, the FP64 displacement).  A plan created with
// stkfmm_tune.sort_points orders each group's targets, and each group's source range, along a Morton
// curve of the group's own bounding box at every execute, on the device, into plan-owned SoA copies;
// the item kernel runs on the copies (its tiles are then compact boxes, so most source tiles take
// the unmasked -- and in mixed mode the far-tile -- form) and the reduction writes each target's sum
// to its original row.  The order is a stable counting sort on the Morton index of a 2^L-per-axis
// grid over the segment's bounding box (L from the segment's size, sort_level: about 64 points per
// cell, L <= 3 up to 2^18 points, 4 above), so it is deterministic: repeated executes, and graph
// replays, give bitwise identical results.
//
//   sort_box_kernel      per block of at most kSortBlock points: bounding box
//   sort_keys_kernel     per block: the segment's box from its blocks' boxes, each point's cell index,
//                        the block's cell histogram
//   sort_offsets_kernel  per segment: cell-major, block-minor exclusive scan of the histograms
//                        (each block's first output position per cell), in place
//   sort_scatter_kernel  per block, one wave: stable ranks within the wave by ballots on the cell
//                        bits, output position = block offset of the cell + rank
//   sort_gather_kernel   per block: coordinates and values of the sorted points into the SoA copies,
//                        targets' output rows (composed with the caller's permutation)
// When a plan's targets are its sources (each group's target range is its source range, and the
// execute passes one coordinate array for both), only the source segments are ordered: the sorted
// source coordinates serve as the targets' and the source gather writes the output rows.
#include <hip/hip_runtime.h>

#include "p2p_internal.hpp"

namespace stkp2p {

namespace {

// one bit of x every third bit (x < 16)
__device__ __forceinline__ unsigned spread3(unsigned x) {
    x &= 0xFu;
    x = (x | (x << 4)) & 0x0C3u;  // bits 0,1 -> 0,1; bits 2,3 -> 6,7
    x = (x | (x << 2)) & 0x249u;  // 0,1,6,7 -> 0,3,6,9
    return x;
}

// component d of point i of a segment of the given kind (0 sources, 1 targets)
__device__ __forceinline__ double coord(const SortArgs &a, int kind, long long i, int d) {
    const DArr &c = kind ? a.tc : a.sc;
    return c.p[i * c.s + d * c.fs];
}

__device__ __forceinline__ double wave_min(double v) {
    for (int off = 32; off > 0; off >>= 1)
        v = fmin(v, __shfl_xor(v, off));
    return v;
}
__device__ __forceinline__ double wave_max(double v) {
    for (int off = 32; off > 0; off >>= 1)
        v = fmax(v, __shfl_xor(v, off));
    return v;
}

// box of the block's points: [lo x, lo y, lo z, hi x, hi y, hi z]; NaN coordinates are ignored
// (fmin/fmax), an empty or all-NaN block gives +inf / -inf
__global__ void __launch_bounds__(256) sort_box_kernel(SortArgs a, int n_blks) {
    const int bi = xcd_unit(blockIdx.x, n_blks);
    if (bi < 0)
        return;
    const SortBlk b = a.blks[bi];
    const int kind = a.segs[b.seg].kind;
    double lo[3] = {INFINITY, INFINITY, INFINITY}, hi[3] = {-INFINITY, -INFINITY, -INFINITY};
    for (long long i = b.begin + threadIdx.x; i < b.end; i += blockDim.x)
#pragma unroll
        for (int d = 0; d < 3; ++d) {
            const double x = coord(a, kind, i, d);
            lo[d] = fmin(lo[d], x);
            hi[d] = fmax(hi[d], x);
        }
    __shared__ double s[4][6];
    const int w = threadIdx.x >> 6, lane = threadIdx.x & 63;
#pragma unroll
    for (int d = 0; d < 3; ++d) {
        lo[d] = wave_min(lo[d]);
        hi[d] = wave_max(hi[d]);
    }
    if (lane == 0)
#pragma unroll
        for (int d = 0; d < 3; ++d) {
            s[w][d] = lo[d];
            s[w][3 + d] = hi[d];
        }
    __syncthreads();
    if (threadIdx.x < 6) {
        const int k = threadIdx.x;
        double v = s[0][k];
        for (int j = 1; j < 4; ++j)
            v = (k < 3) ? fmin(v, s[j][k]) : fmax(v, s[j][k]);
        a.blk_box[6LL * bi + k] = v;
    }
}

__global__ void __launch_bounds__(256) sort_keys_kernel(SortArgs a, int n_blks) {
    const int bi = xcd_unit(blockIdx.x, n_blks);
    if (bi < 0)
        return;
    const SortBlk b = a.blks[bi];
    const SortSeg g = a.segs[b.seg];
    __shared__ double s_box[4][6];
    __shared__ int s_hist[kSortMaxCells];
    // the segment's box from its blocks' boxes (every block of the segment does the same reduction,
    // so every block sees the same box)
    double v[6];
#pragma unroll
    for (int k = 0; k < 6; ++k)
        v[k] = (k < 3) ? INFINITY : -INFINITY;
    for (int j = g.blk_begin + threadIdx.x; j < g.blk_end; j += blockDim.x)
#pragma unroll
        for (int k = 0; k < 6; ++k) {
            const double x = a.blk_box[6LL * j + k];
            v[k] = (k < 3) ? fmin(v[k], x) : fmax(v[k], x);
        }
    const int w = threadIdx.x >> 6, lane = threadIdx.x & 63;
#pragma unroll
    for (int k = 0; k < 6; ++k)
        v[k] = (k < 3) ? wave_min(v[k]) : wave_max(v[k]);
    if (lane == 0)
#pragma unroll
        for (int k = 0; k < 6; ++k)
            s_box[w][k] = v[k];
    const int cells = 1 << (3 * g.level);
    for (int c = threadIdx.x; c < cells; c += blockDim.x)
        s_hist[c] = 0;
    __syncthreads();
    double lo[3], scale[3];
    const int n_axis = 1 << g.level;
#pragma unroll
    for (int d = 0; d < 3; ++d) {
        double l = s_box[0][d], h = s_box[0][3 + d];
        for (int j = 1; j < 4; ++j) {
            l = fmin(l, s_box[j][d]);
            h = fmax(h, s_box[j][3 + d]);
        }
        lo[d] = l;
        const double ext = h - l;
        scale[d] = (ext > 0.0 && ext < INFINITY) ? n_axis / ext : 0.0;
    }
    for (long long i = b.begin + threadIdx.x; i < b.end; i += blockDim.x) {
        unsigned key = 0;
#pragma unroll
        for (int d = 0; d < 3; ++d) {
            const double q = (coord(a, g.kind, i, d) - lo[d]) * scale[d];
            // NaN and out-of-range values land in the first or last cell
            const int qi = (q >= 0.0) ? ((q < n_axis) ? (int)q : n_axis - 1) : 0;
            key |= spread3((unsigned)qi) << d;
        }
        a.keys[(g.kind ? a.n_src : 0) + i] = (unsigned short)key;
        atomicAdd(&s_hist[key], 1);
    }
    __syncthreads();
    int *h = a.hist + b.hist_off;
    for (int c = threadIdx.x; c < cells; c += blockDim.x)
        h[c] = s_hist[c];
}

// segment g: hist[b][c] <- first output position of cell c's points of block b; cells in Morton
// order, within a cell the blocks in order (so the sort is stable)
__global__ void __launch_bounds__(256) sort_offsets_kernel(SortArgs a) {
    const SortSeg g = a.segs[blockIdx.x];
    const int cells = 1 << (3 * g.level);
    // a segment's blocks hold consecutive histograms of `cells` entries each
    int *const h0 = a.hist + a.blks[g.blk_begin].hist_off;
    const int nb = g.blk_end - g.blk_begin;
    __shared__ int s_scan[256];
    int carry = g.begin;
    for (int c0 = 0; c0 < cells; c0 += blockDim.x) {
        const int c = c0 + threadIdx.x;
        int total = 0;
        if (c < cells)
#pragma unroll 8
            for (int j = 0; j < nb; ++j)
                total += h0[(long long)j * cells + c];
        // inclusive scan of the totals over the block (Hillis-Steele in LDS)
        s_scan[threadIdx.x] = total;
        __syncthreads();
        for (int off = 1; off < (int)blockDim.x; off <<= 1) {
            const int add = (threadIdx.x >= (unsigned)off) ? s_scan[threadIdx.x - off] : 0;
            __syncthreads();
            s_scan[threadIdx.x] += add;
            __syncthreads();
        }
        int run = carry + s_scan[threadIdx.x] - total;
        if (c < cells)
            for (int j = 0; j < nb; ++j) {
                int *h = h0 + (long long)j * cells + c;
                const int n = *h;
                *h = run;
                run += n;
            }
        carry += s_scan[blockDim.x - 1];
        __syncthreads();
    }
}

// one wave per block: the block's points in order, 64 at a time; a point's position is its cell's
// next free position in this block plus its rank among the lanes of the same cell
__global__ void __launch_bounds__(64) sort_scatter_kernel(SortArgs a, int n_blks) {
    const int bi = xcd_unit(blockIdx.x, n_blks);
    if (bi < 0)
        return;
    const SortBlk b = a.blks[bi];
    const SortSeg g = a.segs[b.seg];
    const int cells = 1 << (3 * g.level);
    const int bits = 3 * g.level;
    __shared__ int s_off[kSortMaxCells];
    __shared__ unsigned short s_key[kSortBlock];
    const int lane = threadIdx.x;
    const unsigned short *keys = a.keys + (g.kind ? a.n_src : 0);
    for (int c = lane; c < cells; c += 64)
        s_off[c] = a.hist[b.hist_off + c];
    // the block's keys in one burst of loads, so the ranking loop waits on LDS only
    for (int i = b.begin + lane; i < b.end; i += 64)
        s_key[i - b.begin] = keys[i];
    __syncthreads();
    int *out = g.kind ? a.trg_idx : a.src_idx;
    const unsigned long long below = (lane == 0) ? 0ull : (~0ull >> (64 - lane));
    for (int base = b.begin; base < b.end; base += 64) {
        const int i = base + lane;
        const bool valid = i < b.end;
        const unsigned key = valid ? s_key[i - b.begin] : 0u;
        // lanes holding the same cell: the intersection of the ballots that agree with this lane's bits
        unsigned long long peers = __ballot(valid);
        for (int k = 0; k < bits; ++k) {
            const unsigned long long m = __ballot(valid && ((key >> k) & 1u));
            peers &= ((key >> k) & 1u) ? m : ~m;
        }
        const int rank = __popcll(peers & below);
        const int cnt = __popcll(peers);
        int pos = 0;
        if (valid)
            pos = s_off[key] + rank;
        __syncthreads(); // every lane has read its cell's offset
        if (valid && rank == cnt - 1)
            s_off[key] = pos + 1;
        if (valid)
            out[pos] = i;
        __syncthreads();
    }
}

// sorted position p of the block's segment: copy point idx[p]
__global__ void __launch_bounds__(256) sort_gather_kernel(SortArgs a, int n_blks) {
    const int bi = xcd_unit(blockIdx.x, n_blks);
    if (bi < 0)
        return;
    const SortBlk b = a.blks[bi];
    const int kind = a.segs[b.seg].kind;
    for (long long p = b.begin + threadIdx.x; p < b.end; p += blockDim.x) {
        if (kind == 0) {
            const long long i = a.src_idx[p];
#pragma unroll
            for (int d = 0; d < 3; ++d)
                a.src_coord[d * a.n_src + p] = a.sc.p[i * a.sc.s + d * a.sc.fs];
            for (int k = 0; k < a.ks; ++k)
                a.src_value[k * a.n_src + p] =
                    (k < a.ks1) ? a.v1.p[i * a.v1.s + k * a.v1.fs] : a.v2.p[i * a.v2.s + (k - a.ks1) * a.v2.fs];
            if (a.shared)
                a.trg_row[p] = a.perm ? a.perm[i] : (int)i;
        } else {
            const long long i = a.trg_idx[p];
#pragma unroll
            for (int d = 0; d < 3; ++d)
                a.trg_coord[d * a.n_trg + p] = a.tc.p[i * a.tc.s + d * a.tc.fs];
            a.trg_row[p] = a.perm ? a.perm[i] : (int)i;
        }
    }
}

} // namespace

hipError_t launch_sort_points(const SortArgs &a, int n_segs, int n_blks, hipStream_t s) {
    if (n_segs <= 0 || n_blks <= 0)
        return hipSuccess;
    // per-block kernels: XCD-contiguous blocks, so one segment's points share one L2
    const dim3 gb((unsigned)xcd_grid(n_blks));
    hipLaunchKernelGGL(sort_box_kernel, gb, dim3(256), 0, s, a, n_blks);
    hipLaunchKernelGGL(sort_keys_kernel, gb, dim3(256), 0, s, a, n_blks);
    hipLaunchKernelGGL(sort_offsets_kernel, dim3(n_segs), dim3(256), 0, s, a);
    hipLaunchKernelGGL(sort_scatter_kernel, gb, dim3(64), 0, s, a, n_blks);
    hipLaunchKernelGGL(sort_gather_kernel, gb, dim3(256), 0, s, a, n_blks);
    return hipGetLastError();
}

} // namespace stkp2p
