// octree.cpp -- adaptive octree and U-list (near-field work list) on the host.
//
// PVFMM builds the tree the reference evaluates with (FMMData::setupTree, Lib/src/FMMData.cpp:107-167,
// maxPts rule at :111-113); PVFMM is not vendored, so this is the library's own construction with
// the same two rules: a box is split while it holds more than leaf_cap points, and the near field
// (U-list) of a leaf B is every leaf whose closed box touches B's, at any level, B included.
//
// Points are keyed along the Morton (Z-order) curve at 21 bits per axis and sorted; every node is a
// contiguous range of the sorted order with its children in Morton order, so leaves come out in
// Morton order, and the U-list of a leaf becomes a short list of contiguous source ranges once
// leaves adjacent in that order are merged.
#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <string>
#include <utility>
#include <vector>

#include "../../include/stkfmm_p2p.h"

namespace {

constexpr int kBits = 21; // bits per axis; level-21 boxes are the finest cells

inline uint64_t spread3(uint64_t v) {
    v &= 0x1FFFFFULL;
    v = (v | (v << 32)) & 0x1F00000000FFFFULL;
    v = (v | (v << 16)) & 0x1F0000FF0000FFULL;
    v = (v | (v << 8)) & 0x100F00F00F00F00FULL;
    v = (v | (v << 4)) & 0x10C30C30C30C30C3ULL;
    v = (v | (v << 2)) & 0x1249249249249249ULL;
    return v;
}

// [0,1) -> cell index in [0, 2^21); values outside (and NaN) are clamped to the boundary cells
inline uint32_t quantize(double x) {
    double q = x * (double)(1u << kBits);
    if (!(q >= 0.0))
        q = 0.0;
    if (q > (double)((1u << kBits) - 1))
        q = (double)((1u << kBits) - 1);
    return (uint32_t)q;
}

struct Node {
    uint32_t lo[3];    // lower corner in finest-cell units
    uint32_t size;     // edge in finest-cell units (2^(21-level))
    int level;
    int begin, end;    // sorted point range
    int child_begin;   // into stkfmm_tree::children
    int n_children;    // 0 for a leaf
    int leaf_id;       // -1 for internal nodes
};

} // namespace

struct stkfmm_tree {
    int n = 0;
    int leaf_cap = 0;
    std::vector<int> perm;      // sorted index -> original index
    std::vector<uint64_t> keys; // sorted Morton keys
    std::vector<Node> nodes;
    std::vector<int> children;
    std::vector<int> leaf_node; // leaf id -> node index
    std::vector<int> ulist_ptr; // per leaf, into rng_begin/rng_end
    std::vector<int> rng_begin, rng_end;
    int max_level = 0;
    long long pairs = 0;
    // W-list of every leaf (node indices) and X-list of every node (leaf ids), CSR
    std::vector<int> w_ptr, w_node, x_ptr, x_leaf;
};

namespace {

int set_tree_error(int code) { return code; }

// Build the subtree of points [begin, end) whose keys share the top 3*level bits `prefix`.
// Returns the node index.  Recursion depth is at most kBits.
int build(stkfmm_tree &t, int level, uint64_t prefix, int begin, int end) {
    const int idx = (int)t.nodes.size();
    Node nd;
    std::memset(&nd, 0, sizeof(nd));
    // lower corner: de-interleave the prefix bits
    uint32_t lo[3] = {0, 0, 0};
    for (int l = 0; l < level; ++l) {
        const uint64_t oct = (prefix >> (3 * (kBits - 1 - l))) & 7u;
        lo[0] |= (uint32_t)((oct >> 2) & 1u) << (kBits - 1 - l);
        lo[1] |= (uint32_t)((oct >> 1) & 1u) << (kBits - 1 - l);
        lo[2] |= (uint32_t)(oct & 1u) << (kBits - 1 - l);
    }
    nd.lo[0] = lo[0];
    nd.lo[1] = lo[1];
    nd.lo[2] = lo[2];
    nd.size = 1u << (kBits - level);
    nd.level = level;
    nd.begin = begin;
    nd.end = end;
    nd.child_begin = 0;
    nd.n_children = 0;
    nd.leaf_id = -1;
    t.nodes.push_back(nd);
    if (end - begin <= t.leaf_cap || level == kBits) {
        t.nodes[idx].leaf_id = (int)t.leaf_node.size();
        t.leaf_node.push_back(idx);
        t.max_level = std::max(t.max_level, level);
        return idx;
    }
    const int shift = 3 * (kBits - 1 - level);
    int bnd[9];
    bnd[0] = begin;
    bnd[8] = end;
    for (int c = 1; c < 8; ++c) {
        const uint64_t key = prefix | ((uint64_t)c << shift);
        bnd[c] = (int)(std::lower_bound(t.keys.begin() + bnd[c - 1], t.keys.begin() + end, key) - t.keys.begin());
    }
    int kids[8];
    int nk = 0;
    for (int c = 0; c < 8; ++c)
        if (bnd[c + 1] > bnd[c])
            kids[nk++] = build(t, level + 1, prefix | ((uint64_t)c << shift), bnd[c], bnd[c + 1]);
    t.nodes[idx].child_begin = (int)t.children.size();
    t.nodes[idx].n_children = nk;
    for (int i = 0; i < nk; ++i)
        t.children.push_back(kids[i]);
    return idx;
}

// closed boxes [lo, lo+size] intersect in every axis
inline bool touches(const Node &a, const Node &b) {
    for (int d = 0; d < 3; ++d) {
        const uint64_t alo = a.lo[d], ahi = (uint64_t)a.lo[d] + a.size;
        const uint64_t blo = b.lo[d], bhi = (uint64_t)b.lo[d] + b.size;
        if (alo > bhi || blo > ahi)
            return false;
    }
    return true;
}

void collect(const stkfmm_tree &t, int node, const Node &b, std::vector<int> &out) {
    const Node &nd = t.nodes[node];
    if (!touches(nd, b))
        return;
    if (nd.n_children == 0) {
        out.push_back(nd.leaf_id);
        return;
    }
    for (int i = 0; i < nd.n_children; ++i)
        collect(t, t.children[nd.child_begin + i], b, out);
}

// W-list of leaf b (the M2T list of the kernel-independent FMM): nodes c deeper than b, not touching
// b, whose parent touches b -- descendants of b's colleagues that are separated from b while their
// parents are not.  Walks the nodes that touch b.
void wlist(const stkfmm_tree &t, int node, const Node &b, std::vector<int> &out) {
    const Node &nd = t.nodes[node];
    for (int i = 0; i < nd.n_children; ++i) {
        const int c = t.children[nd.child_begin + i];
        const Node &cn = t.nodes[c];
        if (touches(cn, b))
            wlist(t, c, b, out);
        else if (cn.level > b.level)
            out.push_back(c);
    }
}

} // namespace

extern "C" int stkfmm_tree_build(stkfmm_tree **tree_out, int n, const double *coord, int leaf_cap) {
    if (!tree_out)
        return STKFMM_ERR_ARG;
    *tree_out = nullptr;
    if (n < 0 || leaf_cap < 1 || (n > 0 && !coord))
        return set_tree_error(STKFMM_ERR_ARG);
    stkfmm_tree *t = new stkfmm_tree();
    t->n = n;
    t->leaf_cap = leaf_cap;
    std::vector<std::pair<uint64_t, int>> kv((size_t)n);
    for (int i = 0; i < n; ++i) {
        const uint64_t x = quantize(coord[3 * (size_t)i]), y = quantize(coord[3 * (size_t)i + 1]),
                       z = quantize(coord[3 * (size_t)i + 2]);
        kv[i] = {(spread3(x) << 2) | (spread3(y) << 1) | spread3(z), i};
    }
    std::sort(kv.begin(), kv.end());
    t->perm.resize(n);
    t->keys.resize(n);
    for (int i = 0; i < n; ++i) {
        t->keys[i] = kv[i].first;
        t->perm[i] = kv[i].second;
    }
    if (n > 0)
        build(*t, 0, 0, 0, n);
    // U-lists
    const int nl = (int)t->leaf_node.size();
    t->ulist_ptr.assign(1, 0);
    std::vector<int> nb;
    long long pairs = 0;
    for (int l = 0; l < nl; ++l) {
        const Node &b = t->nodes[t->leaf_node[l]];
        nb.clear();
        collect(*t, 0, b, nb);
        std::sort(nb.begin(), nb.end());
        long long ns = 0;
        int cur_b = -1, cur_e = -1;
        for (int a : nb) {
            const Node &na = t->nodes[t->leaf_node[a]];
            ns += na.end - na.begin;
            if (cur_e == na.begin) {
                cur_e = na.end;
            } else {
                if (cur_b >= 0) {
                    t->rng_begin.push_back(cur_b);
                    t->rng_end.push_back(cur_e);
                }
                cur_b = na.begin;
                cur_e = na.end;
            }
        }
        if (cur_b >= 0) {
            t->rng_begin.push_back(cur_b);
            t->rng_end.push_back(cur_e);
        }
        t->ulist_ptr.push_back((int)t->rng_begin.size());
        pairs += ns * (long long)(b.end - b.begin);
    }
    t->pairs = pairs;
    // W-lists per leaf; X-lists per node are their transpose (X(c) = leaves whose W-list holds c)
    const int nn = (int)t->nodes.size();
    t->w_ptr.assign(1, 0);
    std::vector<int> xcount(nn + 1, 0);
    for (int l = 0; l < nl; ++l) {
        if (!t->nodes.empty())
            wlist(*t, 0, t->nodes[t->leaf_node[l]], t->w_node);
        t->w_ptr.push_back((int)t->w_node.size());
    }
    for (int c : t->w_node)
        ++xcount[c + 1];
    for (int c = 0; c < nn; ++c)
        xcount[c + 1] += xcount[c];
    t->x_ptr = xcount;
    t->x_leaf.assign(t->w_node.size(), 0);
    std::vector<int> fill(xcount.begin(), xcount.end() - 1);
    for (int l = 0; l < nl; ++l)
        for (int k = t->w_ptr[l]; k < t->w_ptr[l + 1]; ++k)
            t->x_leaf[fill[t->w_node[k]]++] = l;
    *tree_out = t;
    return STKFMM_OK;
}

extern "C" int stkfmm_tree_info(const stkfmm_tree *t, int *n_leaves, int *n_ranges, int *max_level,
                                long long *pairs) {
    if (!t)
        return STKFMM_ERR_ARG;
    if (n_leaves)
        *n_leaves = (int)t->leaf_node.size();
    if (n_ranges)
        *n_ranges = (int)t->rng_begin.size();
    if (max_level)
        *max_level = t->max_level;
    if (pairs)
        *pairs = t->pairs;
    return STKFMM_OK;
}

extern "C" int stkfmm_tree_get(const stkfmm_tree *t, int *perm, int *leaf_begin, int *leaf_end, int *leaf_level,
                               int *ulist_ptr, int *rng_begin, int *rng_end) {
    if (!t)
        return STKFMM_ERR_ARG;
    if (perm)
        std::copy(t->perm.begin(), t->perm.end(), perm);
    const int nl = (int)t->leaf_node.size();
    for (int l = 0; l < nl; ++l) {
        const Node &nd = t->nodes[t->leaf_node[l]];
        if (leaf_begin)
            leaf_begin[l] = nd.begin;
        if (leaf_end)
            leaf_end[l] = nd.end;
        if (leaf_level)
            leaf_level[l] = nd.level;
    }
    if (ulist_ptr)
        std::copy(t->ulist_ptr.begin(), t->ulist_ptr.end(), ulist_ptr);
    if (rng_begin)
        std::copy(t->rng_begin.begin(), t->rng_begin.end(), rng_begin);
    if (rng_end)
        std::copy(t->rng_end.begin(), t->rng_end.end(), rng_end);
    return STKFMM_OK;
}

extern "C" int stkfmm_tree_node_info(const stkfmm_tree *t, int *n_nodes, int *n_w) {
    if (!t)
        return STKFMM_ERR_ARG;
    if (n_nodes)
        *n_nodes = (int)t->nodes.size();
    if (n_w)
        *n_w = (int)t->w_node.size();
    return STKFMM_OK;
}

extern "C" int stkfmm_tree_get_nodes(const stkfmm_tree *t, double *corner, int *level, int *begin, int *end,
                                     int *leaf_id, int *parent) {
    if (!t)
        return STKFMM_ERR_ARG;
    const double cell = std::ldexp(1.0, -kBits);
    const int nn = (int)t->nodes.size();
    for (int i = 0; i < nn; ++i) {
        const Node &nd = t->nodes[i];
        if (corner)
            for (int k = 0; k < 3; ++k)
                corner[3 * i + k] = nd.lo[k] * cell;
        if (level)
            level[i] = nd.level;
        if (begin)
            begin[i] = nd.begin;
        if (end)
            end[i] = nd.end;
        if (leaf_id)
            leaf_id[i] = nd.leaf_id;
    }
    if (parent) {
        for (int i = 0; i < nn; ++i)
            parent[i] = -1;
        for (int i = 0; i < nn; ++i) {
            const Node &nd = t->nodes[i];
            for (int c = 0; c < nd.n_children; ++c)
                parent[t->children[nd.child_begin + c]] = i;
        }
    }
    return STKFMM_OK;
}

extern "C" int stkfmm_tree_get_wx(const stkfmm_tree *t, int *w_ptr, int *w_node, int *x_ptr, int *x_leaf) {
    if (!t)
        return STKFMM_ERR_ARG;
    if (w_ptr)
        std::copy(t->w_ptr.begin(), t->w_ptr.end(), w_ptr);
    if (w_node)
        std::copy(t->w_node.begin(), t->w_node.end(), w_node);
    if (x_ptr)
        std::copy(t->x_ptr.begin(), t->x_ptr.end(), x_ptr);
    if (x_leaf)
        std::copy(t->x_leaf.begin(), t->x_leaf.end(), x_leaf);
    return STKFMM_OK;
}

extern "C" int stkfmm_tree_leaf_corners(const stkfmm_tree *t, double *corner) {
    if (!t || !corner)
        return STKFMM_ERR_ARG;
    const double cell = std::ldexp(1.0, -kBits); // finest-cell edge, exact
    const int nl = (int)t->leaf_node.size();
    for (int l = 0; l < nl; ++l) {
        const Node &nd = t->nodes[t->leaf_node[l]];
        for (int k = 0; k < 3; ++k)
            corner[3 * l + k] = nd.lo[k] * cell;
    }
    return STKFMM_OK;
}

extern "C" int stkfmm_tree_destroy(stkfmm_tree *t) {
    delete t;
    return STKFMM_OK;
}
