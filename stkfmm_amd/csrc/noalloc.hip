// noalloc.hip -- stkfmm_p2p_eval_device_noalloc: the device direct sum with no device allocation.
//
// The work list of a single-group direct sum is regular (target tiles of 64*T consecutive targets,
// each against the sources in chunks of max_src), so a small kernel writes it on the device instead
// of the host uploading it, and everything the item and reduce kernels need beyond the caller's
// inputs and output -- items, source ranges, tiles, the queue bounds and counter, the per-item
// partial sums -- lives in one file-scope device array, `scratch`, of kScratchBytes per device.
// The kernels are the ones the cached-plan entry (stkfmm_p2p_eval_device) runs on points in the
// caller's order: below 16,384 sources or targets (where that entry does not order the points along a
// Morton curve) and while the partial sums fit, the tiling and chunking are the plan's as well and the
// two entries give bitwise identical results; otherwise the same pairs are summed in another grouping
// (no point ordering here, which would need copies of the points; above the plan's partial-sum size,
// fewer and longer source chunks).
#include <hip/hip_runtime.h>

#include <cstdio>
#include <mutex>

#include "../../include/stkfmm_p2p.h"
#include "p2p_internal.hpp"

namespace stkp2p {
namespace {

constexpr size_t kScratchBytes = size_t(64) << 20; // 64 MiB per device, held until the process ends
constexpr int kMaxDevices = 64;

static __device__ __align__(256) unsigned char scratch[kScratchBytes];

// device address of `scratch` on each device (the host-side name is not a device address)
void *g_scratch_ptr[kMaxDevices] = {};
std::mutex g_scratch_mu;

int scratch_ptr(int dev, unsigned char **out) {
    if (dev < 0 || dev >= kMaxDevices)
        return report_error(STKFMM_ERR_UNSUPPORTED, "no-alloc direct sum: device index out of range");
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    if (!g_scratch_ptr[dev]) {
        void *p = nullptr;
        const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !p) {
            char buf[256];
            snprintf(buf, sizeof(buf), "no-alloc direct sum: hipGetSymbolAddress(scratch): %s", hipGetErrorString(e));
            return report_error(STKFMM_ERR_HIP, buf);
        }
        g_scratch_ptr[dev] = p;
    }
    *out = (unsigned char *)g_scratch_ptr[dev];
    return STKFMM_OK;
}

constexpr size_t align256(size_t n) { return (n + 255) & ~size_t(255); }

// byte layout of the work list and partial sums inside the scratch array
struct Layout {
    size_t items, ranges, tiles, queue, partial, end;
};
Layout layout(long long n_items, long long n_tiles, int kt, int tt) {
    Layout l;
    l.items = 0;
    l.ranges = l.items + align256(sizeof(Item) * (size_t)n_items);
    l.tiles = l.ranges + align256(sizeof(int2) * (size_t)n_items);
    l.queue = l.tiles + align256(sizeof(Tile) * (size_t)n_tiles);
    l.partial = l.queue + 256; // q_ptr[2], counter[1]
    l.end = l.partial + sizeof(double) * (size_t)n_items * kt * tt;
    return l;
}

// The direct sum's work list: n_tiles target tiles x nc source chunks of max_src (the last chunk
// holds the rest).  Item slot = tile * nc + chunk; the queue lists the full chunks first (in slot
// order), then the last chunks -- the order of the plan's cost-sorted queue.  Also resets the
// queue counter.  One thread per item (and per tile).
struct DirectList {
    Item *items;
    int2 *ranges;
    Tile *tiles;
    int *q_ptr;   // [2]
    int *counter; // [1]
    int n_src, n_trg, tt, n_tiles, nc, max_src;
};

__global__ void __launch_bounds__(256) direct_list_kernel(DirectList d) {
    const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    const long long n_items = (long long)d.n_tiles * d.nc;
    const int nfull = d.n_src / d.max_src; // chunks of exactly max_src sources
    if (i == 0) {
        d.q_ptr[0] = 0;
        d.q_ptr[1] = (int)n_items;
        d.counter[0] = 0;
    }
    if (i < d.n_tiles) {
        const int t = (int)i;
        Tile tl;
        tl.trg_begin = t * d.tt;
        tl.trg_count = min(d.tt, d.n_trg - t * d.tt);
        tl.slot_begin = t * d.nc;
        tl.slot_end = (t + 1) * d.nc;
        d.tiles[t] = tl;
    }
    if (i >= n_items)
        return;
    int t, c;
    if (i < (long long)d.n_tiles * nfull) {
        t = (int)(i / nfull);
        c = (int)(i % nfull);
    } else {
        t = (int)(i - (long long)d.n_tiles * nfull);
        c = nfull;
    }
    const int slot = t * d.nc + c;
    const int sb = c * d.max_src;
    const int se = min(d.n_src, sb + d.max_src);
    d.ranges[slot] = make_int2(sb, se);
    Item it;
    it.trg_begin = t * d.tt;
    it.trg_count = min(d.tt, d.n_trg - t * d.tt);
    it.rng_begin = slot;
    it.rng_end = slot + 1;
    it.slot = slot;
    it.level = 0;
    it.cost = se - sb;
    it.lsplit = tile_lsplit(it.trg_count, d.tt);
    d.items[i] = it;
}

int fail(int code, const char *fmt, long long a = 0, long long b = 0) {
    char buf[320];
    snprintf(buf, sizeof(buf), fmt, a, b);
    return report_error(code, buf);
}

} // namespace
} // namespace stkp2p

using namespace stkp2p;

extern "C" int stkfmm_p2p_eval_device_noalloc(unsigned kernel, unsigned p2p, int precision, int nSrc,
                                              const double *d_srcCoord, const double *d_srcValue, int nTrg,
                                              const double *d_trgCoord, double *d_trgValue, void *stream) {
    if (precision != STKFMM_PREC_FP64 && precision != STKFMM_PREC_MIXED)
        return fail(STKFMM_ERR_ARG, "unknown precision %lld", precision);
    int uk = -1;
    int rc = stkfmm_get_uker(kernel, p2p, &uk);
    if (rc)
        return rc;
    if (nSrc < 0 || nTrg < 0)
        return fail(STKFMM_ERR_ARG, "negative point count");
    if (nSrc == 0 || nTrg == 0)
        return STKFMM_OK;
    if (!d_srcCoord || !d_srcValue || !d_trgCoord || !d_trgValue)
        return fail(STKFMM_ERR_ARG, "NULL buffer");
    const hipStream_t s = (hipStream_t)stream;

    int dev = 0, n_cu = 0, bpc = 0;
    hipError_t e = hipGetDevice(&dev);
    if (e == hipSuccess)
        e = hipDeviceGetAttribute(&n_cu, hipDeviceAttributeMultiprocessorCount, dev);
    const int T = direct_targets_per_lane(nTrg), TT = kWave * T;
    if (e == hipSuccess)
        e = items_occupancy(uk, precision, T, &bpc);
    if (e != hipSuccess)
        return fail(STKFMM_ERR_HIP, "no-alloc direct sum: device query failed");
    unsigned char *base = nullptr;
    if ((rc = scratch_ptr(dev, &base)))
        return rc;
    bpc = bpc < 1 ? 1 : bpc;

    const int ks = uker_ks(uk), ks1 = uker_ks1(uk), kt = uker_kt(uk);
    const long long n_tiles = (nTrg + TT - 1) / TT;
    // the plan's chunking, or fewer and longer chunks when its partial sums do not fit
    long long max_src = default_max_src((long long)nSrc * n_tiles, (long long)n_cu * bpc);
    long long nc = (nSrc + max_src - 1) / max_src;
    if (layout(n_tiles * nc, n_tiles, kt, TT).end > kScratchBytes) {
        const Layout one = layout(n_tiles, n_tiles, kt, TT);
        if (one.end > kScratchBytes)
            return fail(STKFMM_ERR_ARG,
                        "no-alloc direct sum: %lld targets need %lld bytes of scratch at one source chunk per "
                        "target tile, above its fixed size (use stkfmm_p2p_eval_device)",
                        nTrg, (long long)one.end);
        // estimate from the bytes per item, then step down past the alignment padding
        const size_t per_item = sizeof(Item) + sizeof(int2) + sizeof(double) * kt * TT;
        long long nc_max = (long long)((kScratchBytes - one.end) / (per_item * (size_t)n_tiles)) + 1;
        nc_max = nc_max < nc ? nc_max : nc;
        while (nc_max > 1 && layout(n_tiles * nc_max, n_tiles, kt, TT).end > kScratchBytes)
            --nc_max;
        max_src = (nSrc + nc_max - 1) / nc_max;
        nc = (nSrc + max_src - 1) / max_src;
    }
    const long long n_items = n_tiles * nc;
    if (n_items > (1LL << 30))
        return fail(STKFMM_ERR_ARG, "no-alloc direct sum: %lld work items exceed the index range", n_items);
    const Layout L = layout(n_items, n_tiles, kt, TT);

    DirectList d;
    d.items = (Item *)(base + L.items);
    d.ranges = (int2 *)(base + L.ranges);
    d.tiles = (Tile *)(base + L.tiles);
    d.q_ptr = (int *)(base + L.queue);
    d.counter = d.q_ptr + 2;
    d.n_src = nSrc;
    d.n_trg = nTrg;
    d.tt = TT;
    d.n_tiles = (int)n_tiles;
    d.nc = (int)nc;
    d.max_src = (int)max_src;
    const long long n_threads = n_items > n_tiles ? n_items : n_tiles;
    hipLaunchKernelGGL(direct_list_kernel, dim3((unsigned)((n_threads + 255) / 256)), dim3(256), 0, s, d);
    if ((e = hipGetLastError()) != hipSuccess)
        return fail(STKFMM_ERR_HIP, "no-alloc direct sum: work-list launch failed");

    ItemArgs a;
    a.sc = DArr{d_srcCoord, 3, 1};
    a.v1 = DArr{d_srcValue, ks, 1};
    a.v2 = (ks1 < ks) ? DArr{d_srcValue + ks1, ks, 1} : a.v1;
    a.tc = DArr{d_trgCoord, 3, 1};
    a.items = d.items;
    a.ranges = d.ranges;
    a.n_items = (int)n_items;
    a.counter = d.counter;
    a.q_ptr = d.q_ptr;
    a.n_queues = 1;
    a.partial = (double *)(base + L.partial);
    a.n_src_rec = nSrc;
    const int grid = (int)(n_items < (long long)n_cu * bpc ? n_items : (long long)n_cu * bpc);
    if ((e = launch_items(uk, precision, T, grid, a, s)) != hipSuccess)
        return fail(STKFMM_ERR_HIP, "no-alloc direct sum: item kernel launch failed");
    ReduceArgs ra;
    ra.tiles = d.tiles;
    ra.n_tiles = (int)n_tiles;
    ra.kt = kt;
    ra.tt = TT;
    ra.partial = a.partial;
    ra.out = OArr{d_trgValue, kt, 1};
    ra.perm = nullptr;
    ra.scale = uker_scale(uk);
    if ((e = launch_reduce(ra, s)) != hipSuccess)
        return fail(STKFMM_ERR_HIP, "no-alloc direct sum: reduce launch failed");
    return STKFMM_OK;
}
