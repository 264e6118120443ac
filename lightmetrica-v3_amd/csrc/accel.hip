// GPU BVH build and ray-cast entry points of the lmgpu C ABI (gfx950).
//
// Replaces Accel_SAHBVH::build (src/accel/accel_sahbvh.cpp:127-267), a
// top-down full-sort SAH build on host threads, with a fully device-side
// pipeline:
//   1. per-triangle bounds + centroid bounds (ordered-int atomics),
//   2. 63-bit Morton codes of the centroids,
//   3. stable radix sort of (code, index) pairs (sort.hip), checked ordered,
//   4. binary LBVH topology (Karras, HPG 2012), ties broken by index, or
//      PLOC (ploc.hip),
//   5. bottom-up bounds refit, level-synchronous over a frontier, which also
//      fills the SAH cost tables of the collapse (sah_tables),
//   6. collapse into a W-wide tree (W = 2, 4, 8) of least SAH cost: the
//      dynamic program decides which binary subtrees become wide nodes and
//      which become leaves of at most `leaf_size` triangles (a contiguous
//      range of the sorted triangle array); the greedy alternative opens the
//      largest-area child until a node has W children,
//   7. triangles copied into traversal order (48 B each).
// and Accel_SAHBVH::intersect (:269-299) with the ordered wide traversal of
// bvh.h.

#include <algorithm>
#include <type_traits>
#include <chrono>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include "build.h"
#include "bvh.h"
#include "lmgpu_internal.h"

// Rays a wave of the ray-cast kernel takes per atomic on the ray counter
// (128: coherent config-1 rays +10% over 64, incoherent level).
#ifndef LMGPU_TRACE_RAY_CHUNK
#define LMGPU_TRACE_RAY_CHUNK 128
#endif

namespace lmgpu {

// Build defaults: leaves of at most 2 triangles, and a triangle test costing
// half a wide-node visit in the SAH collapse (the watertight test of a
// prefetched triangle is cheaper than a node's 4-8 slab tests plus the sort;
// 0.5 measured 2-3% more incoherent closest-hit rays than 1.0 on a 1.32M
// triangle mesh, with the same hits).
constexpr int kDefaultLeafSize = 2;
constexpr float kDefaultSahTriCost = 0.5f;
// Refit levels launched between two reads of the frontier size.
constexpr int kRefitLevelsPerSync = 6;

// ---------------------------------------------------------------------------
// build kernels
// ---------------------------------------------------------------------------

// Per-triangle AABB (float4 lo, hi); per-block centroid bounds into
// `partial` (6 floats per block), reduced by k_reduce_bounds.
__global__ __launch_bounds__(256) void k_prep(const float* __restrict__ verts, int64_t n, float4* __restrict__ tlo,
                                              float4* __restrict__ thi, float* __restrict__ partial) {
    __shared__ float red[6][4];
    const int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    float cmin[3] = {3e38f, 3e38f, 3e38f}, cmax[3] = {-3e38f, -3e38f, -3e38f};
    if (i < n) {
        const float* v = verts + 9 * i;
        float lo[3], hi[3];
        for (int a = 0; a < 3; a++) {
            lo[a] = fminf(v[a], fminf(v[3 + a], v[6 + a]));
            hi[a] = fmaxf(v[a], fmaxf(v[3 + a], v[6 + a]));
            const float c = (lo[a] + hi[a]) * 0.5f;
            cmin[a] = c;
            cmax[a] = c;
        }
        tlo[i] = make_float4(lo[0], lo[1], lo[2], 0.f);
        thi[i] = make_float4(hi[0], hi[1], hi[2], 0.f);
    }
    const int w = threadIdx.x >> 6;
    for (int a = 0; a < 3; a++) {
        float mn = cmin[a], mx = cmax[a];
        for (int off = 32; off > 0; off >>= 1) {
            mn = fminf(mn, __shfl_xor(mn, off));
            mx = fmaxf(mx, __shfl_xor(mx, off));
        }
        if ((threadIdx.x & 63) == 0) {
            red[a][w] = mn;
            red[3 + a][w] = mx;
        }
    }
    __syncthreads();
    if (threadIdx.x < 6) {
        const int a = threadIdx.x;
        float x = red[a][0];
        for (int k = 1; k < 4; k++) x = a < 3 ? fminf(x, red[a][k]) : fmaxf(x, red[a][k]);
        partial[6 * int64_t(blockIdx.x) + a] = x;
    }
}

// One block reduces the per-block partial bounds into cb[6] (floats).
__global__ __launch_bounds__(256) void k_reduce_bounds(const float* __restrict__ partial, int nparts,
                                                       float* __restrict__ cb) {
    __shared__ float red[6][4];
    float v[6] = {3e38f, 3e38f, 3e38f, -3e38f, -3e38f, -3e38f};
    for (int b = threadIdx.x; b < nparts; b += blockDim.x)
        for (int a = 0; a < 6; a++)
            v[a] = a < 3 ? fminf(v[a], partial[6 * b + a]) : fmaxf(v[a], partial[6 * b + a]);
    const int w = threadIdx.x >> 6;
    for (int a = 0; a < 6; a++) {
        float x = v[a];
        for (int off = 32; off > 0; off >>= 1) x = a < 3 ? fminf(x, __shfl_xor(x, off)) : fmaxf(x, __shfl_xor(x, off));
        if ((threadIdx.x & 63) == 0) red[a][w] = x;
    }
    __syncthreads();
    if (threadIdx.x < 6) {
        const int a = threadIdx.x;
        float x = red[a][0];
        for (int k = 1; k < 4; k++) x = a < 3 ? fminf(x, red[a][k]) : fmaxf(x, red[a][k]);
        cb[a] = x;
    }
}

__device__ __forceinline__ uint64_t spread21(uint64_t x) {
    x &= 0x1fffffull;
    x = (x | x << 32) & 0x1f00000000ffffull;
    x = (x | x << 16) & 0x1f0000ff0000ffull;
    x = (x | x << 8) & 0x100f00f00f00f00full;
    x = (x | x << 4) & 0x10c30c30c30c30c3ull;
    x = (x | x << 2) & 0x1249249249249249ull;
    return x;
}

__global__ __launch_bounds__(256) void k_morton(const float4* __restrict__ tlo, const float4* __restrict__ thi,
                                                int64_t n, const float* __restrict__ cb,
                                                uint64_t* __restrict__ keys, int32_t* __restrict__ vals) {
    const int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const float4 lo = tlo[i], hi = thi[i];
    const float c[3] = {(lo.x + hi.x) * 0.5f, (lo.y + hi.y) * 0.5f, (lo.z + hi.z) * 0.5f};
    uint64_t code = 0;
    for (int a = 0; a < 3; a++) {
        const float mn = cb[a], mx = cb[3 + a];
        const float ext = mx - mn;
        float f = ext > 0.f ? (c[a] - mn) / ext : 0.5f;
        f = fminf(fmaxf(f, 0.f), 1.f);
        const uint64_t q = uint64_t(fminf(f * 2097152.0f, 2097151.0f));
        code |= spread21(q) << (2 - a);
    }
    keys[i] = code;
    vals[i] = int32_t(i);
}


// flag = 1 if keys[i] > keys[i + 1] for some i
__global__ __launch_bounds__(256) void k_check_sorted(const uint64_t* __restrict__ keys, int64_t n,
                                                      int32_t* __restrict__ flag) {
    const int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i + 1 < n && keys[i] > keys[i + 1]) *flag = 1;
}

__device__ __forceinline__ int delta(const uint64_t* __restrict__ k, int64_t n, int64_t i, int64_t j) {
    if (j < 0 || j >= n) return -1;
    const uint64_t a = k[i], b = k[j];
    if (a == b) return 64 + __clzll(uint64_t(i ^ j));
    return __clzll(a ^ b);
}

__global__ __launch_bounds__(256) void k_lbvh(const uint64_t* __restrict__ k, int64_t n, BNode* __restrict__ bn,
                                              int32_t* __restrict__ parent_int, int32_t* __restrict__ parent_leaf) {
    const int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i >= n - 1) return;
    const int d = (delta(k, n, i, i + 1) - delta(k, n, i, i - 1)) >= 0 ? 1 : -1;
    const int dmin = delta(k, n, i, i - d);
    int64_t lmax = 2;
    while (delta(k, n, i, i + lmax * d) > dmin) lmax <<= 1;
    int64_t l = 0;
    for (int64_t t = lmax >> 1; t >= 1; t >>= 1)
        if (delta(k, n, i, i + (l + t) * d) > dmin) l += t;
    const int64_t j = i + l * d;
    const int dnode = delta(k, n, i, j);
    int64_t s = 0;
    int64_t len = l;
    int64_t t = (len + 1) >> 1;
    while (true) {
        if (delta(k, n, i, i + (s + t) * d) > dnode) s += t;
        if (t == 1) break;
        t = (t + 1) >> 1;
    }
    const int64_t g = i + s * d + (d < 0 ? -1 : 0);
    const int64_t lo = i < j ? i : j, hi = i < j ? j : i;
    BNode node;
    node.first = int32_t(lo);
    node.last = int32_t(hi);
    if (lo == g) {
        node.left = ~int32_t(g);
        parent_leaf[g] = int32_t(i);
    } else {
        node.left = int32_t(g);
        parent_int[g] = int32_t(i);
    }
    if (hi == g + 1) {
        node.right = ~int32_t(g + 1);
        parent_leaf[g + 1] = int32_t(i);
    } else {
        node.right = int32_t(g + 1);
        parent_int[g + 1] = int32_t(i);
    }
    bn[i] = node;
}

// One atomic per wave: a lane with `want` gets a slot of *counter (-1
// otherwise). Every lane of the wave must call it.
__device__ __forceinline__ int32_t wave_alloc(int32_t* counter, bool want, int32_t count = 1) {
    const unsigned long long mask = __ballot(want);
    if (mask == 0) return -1;
    const int lane = threadIdx.x & 63;
    const int leader = __ffsll(mask) - 1;
    // inclusive prefix of the counts of the wanting lanes, in lane order
    int32_t c = want ? count : 0, inc = c;
    for (int off = 1; off < 64; off <<= 1) {
        const int32_t y = __shfl_up(inc, off);
        if (lane >= off) inc += y;
    }
    const int32_t total = __shfl(inc, 63);
    int32_t base = 0;
    if (lane == leader) base = atomicAdd(counter, total);
    base = __shfl(base, leader);
    return want ? base + inc - c : -1;
}

// Bottom-up refit without device-scope fences (which write back the
// per-XCD L2s): level-synchronous over a frontier. A node enters the
// frontier of launch k+1 when its second child finished in launch k (atomic
// arrival counter), so both children's bounds were written by earlier
// launches and are visible at kernel boundaries.
__global__ __launch_bounds__(256) void k_refit_leaves(int64_t n, const int32_t* __restrict__ order,
                                                      const float4* __restrict__ tlo, const float4* __restrict__ thi,
                                                      const int32_t* __restrict__ parent_leaf,
                                                      float4* __restrict__ leaf_lo, float4* __restrict__ leaf_hi,
                                                      int32_t* __restrict__ flags, int32_t* __restrict__ frontier,
                                                      int32_t* __restrict__ fcount) {
    const int64_t j = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    bool push = false;
    int32_t p = -1;
    if (j < n) {
        const int32_t ti = order[j];
        leaf_lo[j] = tlo[ti];
        leaf_hi[j] = thi[ti];
        if (n > 1) {
            p = parent_leaf[j];
            push = atomicAdd(&flags[p], 1) == 1;
        }
    }
    const int32_t k = wave_alloc(fcount, push);
    if (push) frontier[k] = p;
}

__device__ __forceinline__ float half_area(float4 lo, float4 hi) {
    const float dx = hi.x - lo.x, dy = hi.y - lo.y, dz = hi.z - lo.z;
    return dx * dy + dy * dz + dz * dx;
}

// Parent links of a binary tree given by its nodes (parent_int[root] stays
// as initialised).
__global__ __launch_bounds__(256) void k_parents(const BNode* __restrict__ bn, int64_t m,
                                                 int32_t* __restrict__ parent_int, int32_t* __restrict__ parent_leaf) {
    const int64_t p = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (p >= m) return;
    const BNode nd = bn[p];
    if (nd.left < 0) parent_leaf[~nd.left] = int32_t(p); else parent_int[nd.left] = int32_t(p);
    if (nd.right < 0) parent_leaf[~nd.right] = int32_t(p); else parent_int[nd.right] = int32_t(p);
}

// Cost tables of the SAH collapse (Ylitie, Karras and Laine, "Efficient
// Incoherent Ray Traversal on GPUs Through Compressed Wide BVHs", HPG 2017,
// sec. 4): for internal binary node p and i = 1..W, C(p, i) is the least SAH
// cost (half areas; a wide-node visit costs 1, a triangle test `ctri`) of
// covering p's subtree with at most i roots, each a wide node or a leaf.
//   C(p, 1) = min(A(p) * N(p) * ctri  [leaf, N(p) <= pmax],
//                 A(p) + D(p, W)       [wide node]),
//   C(p, i) = min(C(p, i - 1), D(p, i)),   D(p, j) = min_k C(L, k) + C(R, j - k),
// and a single triangle j costs A(j) * ctri for every i. choice[p*W + 0] is
// 0xFF for a leaf, else the k of D(p, W); choice[p*W + i - 1] (i >= 2) is 0
// when C(p, i) = C(p, i - 1), else the k of D(p, i).
struct SahDP {
    float* cost;
    uint8_t* choice;
    int W;
    int pmax;
    float ctri;
};

__device__ __forceinline__ void sah_tables(const SahDP& dp, int32_t p, const BNode& nd, float area,
                                           const float4& l0, const float4& h0, const float4& l1, const float4& h1) {
    const int W = dp.W;
    float cl[8], cr[8];
    const float al = half_area(l0, h0) * dp.ctri, ar = half_area(l1, h1) * dp.ctri;
    for (int i = 0; i < W; i++) {
        cl[i] = nd.left < 0 ? al : dp.cost[int64_t(nd.left) * W + i];
        cr[i] = nd.right < 0 ? ar : dp.cost[int64_t(nd.right) * W + i];
    }
    float dist[8];
    int arg[8];
    for (int j = 2; j <= W; j++) {
        float best = __builtin_inff();
        int bk = 1;
        for (int k = 1; k < j; k++) {
            const float c = cl[k - 1] + cr[j - k - 1];
            if (c < best) { best = c; bk = k; }
        }
        dist[j - 1] = best;
        arg[j - 1] = bk;
    }
    const int cnt = nd.last - nd.first + 1;
    const float cint = area + dist[W - 1];
    const float cprim = cnt <= dp.pmax ? area * float(cnt) * dp.ctri : __builtin_inff();
    float c = fminf(cprim, cint);
    float* out = dp.cost + int64_t(p) * W;
    uint8_t* ch = dp.choice + int64_t(p) * W;
    out[0] = c;
    ch[0] = cprim <= cint ? uint8_t(0xFF) : uint8_t(arg[W - 1]);
    for (int i = 2; i <= W; i++) {
        if (c <= dist[i - 1]) {
            ch[i - 1] = 0;
        } else {
            c = dist[i - 1];
            ch[i - 1] = uint8_t(arg[i - 1]);
        }
        out[i - 1] = c;
    }
}

__global__ __launch_bounds__(256) void k_refit_nodes(const int32_t* __restrict__ fin, const int32_t* __restrict__ fin_count,
                                                     const BNode* __restrict__ bn, int32_t root,
                                                     const int32_t* __restrict__ parent_int,
                                                     const float4* __restrict__ leaf_lo,
                                                     const float4* __restrict__ leaf_hi, float4* __restrict__ node_lo,
                                                     float4* __restrict__ node_hi, int32_t* __restrict__ flags,
                                                     int32_t* __restrict__ fout, int32_t* __restrict__ fout_count,
                                                     SahDP dp) {
    const int cnt = *fin_count;
    const int stride = gridDim.x * blockDim.x;
    // whole waves run the same number of iterations (wave-level appends)
    const int rounds = (cnt + stride - 1) / stride;
    for (int r = 0; r < rounds; r++) {
        const int idx = r * stride + blockIdx.x * blockDim.x + threadIdx.x;
        bool push = false;
        int32_t q = -1;
        if (idx < cnt) {
            const int32_t p = fin[idx];
            const BNode nd = bn[p];
            const float4 l0 = nd.left < 0 ? leaf_lo[~nd.left] : node_lo[nd.left];
            const float4 h0 = nd.left < 0 ? leaf_hi[~nd.left] : node_hi[nd.left];
            const float4 l1 = nd.right < 0 ? leaf_lo[~nd.right] : node_lo[nd.right];
            const float4 h1 = nd.right < 0 ? leaf_hi[~nd.right] : node_hi[nd.right];
            const float4 lo = make_float4(fminf(l0.x, l1.x), fminf(l0.y, l1.y), fminf(l0.z, l1.z), 0.f);
            const float4 hi = make_float4(fmaxf(h0.x, h1.x), fmaxf(h0.y, h1.y), fmaxf(h0.z, h1.z), 0.f);
            node_lo[p] = lo;
            node_hi[p] = hi;
            if (dp.cost) sah_tables(dp, p, nd, half_area(lo, hi), l0, h0, l1, h1);
            if (p != root) {
                q = parent_int[p];
                push = atomicAdd(&flags[q], 1) == 1;
            }
        }
        const int32_t k = wave_alloc(fout_count, push);
        if (push) fout[k] = q;
    }
}

__global__ __launch_bounds__(256) void k_reorder(int64_t n, const int32_t* __restrict__ order,
                                                 const float* __restrict__ verts, TriV* __restrict__ tris) {
    const int64_t j = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (j >= n) return;
    const int32_t ti = order[j];
    const float* v = verts + 9 * int64_t(ti);
    TriV t;
    t.v0 = make_float4(v[0], v[1], v[2], __int_as_float(ti));
    t.v1 = make_float4(v[3], v[4], v[5], 0.f);
    t.v2 = make_float4(v[6], v[7], v[8], 0.f);
    tris[j] = t;
}

struct CollapseItem {
    int32_t bref;   // binary ref (internal index or ~leaf)
    int32_t wnode;  // wide node slot
};

// Writes wide node `it.wnode` from its nc <= W binary refs (each a leaf when
// is_leaf(ref), else a wide child). The wide children of one node take
// consecutive slots, so siblings share cache lines; they become next-level
// items. Every lane of the wave calls it (`active` false for lanes without an
// item): slots and queue positions are taken with one atomic per wave and
// the SAH and leaf statistics are summed over the wave first.
template <int W, class IsLeaf>
__device__ __forceinline__ void emit_wide_node(bool active, const CollapseItem& it, const int32_t (&ref)[W], int nc,
                                               IsLeaf is_leaf, const BNode* __restrict__ bn,
                                               const float4* __restrict__ node_lo, const float4* __restrict__ node_hi,
                                               const float4* __restrict__ leaf_lo, const float4* __restrict__ leaf_hi,
                                               CollapseItem* __restrict__ out, int32_t* __restrict__ n_out,
                                               int32_t* __restrict__ node_counter, WNode<W>* __restrict__ wnodes,
                                               double* sah_acc, int32_t* leaf_counter) {
    int nint = 0;
    if (active)
        for (int c = 0; c < nc; c++) nint += is_leaf(ref[c]) ? 0 : 1;
    int32_t slot = wave_alloc(node_counter, nint > 0, nint);
    int32_t q = wave_alloc(n_out, nint > 0, nint);
    double sah = 0.0;
    int nleaf = 0;
    if (active) {
        WNode<W> node;
        for (int c = 0; c < W; c++) {
            if (c < nc) {
                const int32_t r = ref[c];
                const float4 lo = r < 0 ? leaf_lo[~r] : node_lo[r];
                const float4 hi = r < 0 ? leaf_hi[~r] : node_hi[r];
                node.lox[c] = lo.x; node.loy[c] = lo.y; node.loz[c] = lo.z;
                node.hix[c] = hi.x; node.hiy[c] = hi.y; node.hiz[c] = hi.z;
                const int cnt = r < 0 ? 1 : bn[r].last - bn[r].first + 1;
                const double a = double(half_area(lo, hi));
                if (is_leaf(r)) {
                    node.child[c] = r < 0 ? ~r : bn[r].first;
                    node.count[c] = cnt;
                    sah += a * cnt;
                    nleaf++;
                } else {
                    node.child[c] = slot;
                    node.count[c] = 0;
                    sah += a;
                    out[q++] = CollapseItem{r, slot++};
                }
            } else {
                node.lox[c] = node.loy[c] = node.loz[c] = 1e30f;
                node.hix[c] = node.hiy[c] = node.hiz[c] = -1e30f;
                node.child[c] = 0;
                node.count[c] = -1;
            }
        }
        wnodes[it.wnode] = node;
    }
    for (int off = 32; off > 0; off >>= 1) {
        sah += __shfl_xor(sah, off);
        nleaf += __shfl_xor(nleaf, off);
    }
    if ((threadIdx.x & 63) == 0) {
        atomicAdd(sah_acc, sah);
        atomicAdd(leaf_counter, nleaf);
    }
}

// SAH collapse: one work item per wide node; its children are the roots the
// cost tables chose for D(p, W) (sah_tables), found by unfolding the choices
// with a stack of (ref, slots) pairs whose slots sum to at most W.
template <int W>
__global__ __launch_bounds__(256) void k_collapse_sah(const CollapseItem* __restrict__ in,
                                                      const int32_t* __restrict__ n_in, CollapseItem* __restrict__ out,
                                                      int32_t* __restrict__ n_out, int32_t* __restrict__ node_counter,
                                                      const BNode* __restrict__ bn, const float4* __restrict__ node_lo,
                                                      const float4* __restrict__ node_hi,
                                                      const float4* __restrict__ leaf_lo,
                                                      const float4* __restrict__ leaf_hi,
                                                      const uint8_t* __restrict__ choice, WNode<W>* __restrict__ wnodes,
                                                      double* sah_acc, int32_t* leaf_counter) {
    const int idx = blockIdx.x * blockDim.x + threadIdx.x;
    const bool active = idx < *n_in;
    const CollapseItem it = active ? in[idx] : CollapseItem{~0, 0};
    auto is_leaf = [&](int32_t r) { return r < 0 || choice[int64_t(r) * W] == 0xFF; };
    int32_t ref[W];
    int nc = 0;
    if (!active) {
        // no item: only takes part in the wave's allocations
    } else if (is_leaf(it.bref)) {
        ref[nc++] = it.bref;   // the whole tree is one leaf (tiny scenes)
    } else {
        int32_t sr[W];
        int sj[W];
        int sp = 0;
        const BNode b = bn[it.bref];
        const int k = choice[int64_t(it.bref) * W];
        sr[sp] = b.right; sj[sp++] = W - k;
        sr[sp] = b.left; sj[sp++] = k;
        while (sp > 0) {
            --sp;
            const int32_t m = sr[sp];
            int j = sj[sp];
            if (m >= 0)
                while (j > 1 && choice[int64_t(m) * W + j - 1] == 0) j--;
            if (m < 0 || j == 1) {
                ref[nc++] = m;
                continue;
            }
            const int c = choice[int64_t(m) * W + j - 1];
            const BNode bm = bn[m];
            sr[sp] = bm.right; sj[sp++] = j - c;
            sr[sp] = bm.left; sj[sp++] = c;
        }
    }
    emit_wide_node<W>(active, it, ref, nc, is_leaf, bn, node_lo, node_hi, leaf_lo, leaf_hi, out, n_out, node_counter,
                      wnodes, sah_acc, leaf_counter);
}

// Greedy collapse: one work item per wide node; open the largest-area child
// until W, then write the node and emit the internal children as next-level
// items.
template <int W>
__global__ __launch_bounds__(256) void k_collapse(const CollapseItem* __restrict__ in, const int32_t* __restrict__ n_in,
                                                  CollapseItem* __restrict__ out, int32_t* __restrict__ n_out,
                                                  int32_t* __restrict__ node_counter, const BNode* __restrict__ bn,
                                                  const float4* __restrict__ node_lo, const float4* __restrict__ node_hi,
                                                  const float4* __restrict__ leaf_lo, const float4* __restrict__ leaf_hi,
                                                  WNode<W>* __restrict__ wnodes, int leaf_size, double* sah_acc,
                                                  int32_t* leaf_counter) {
    const int idx = blockIdx.x * blockDim.x + threadIdx.x;
    const bool active = idx < *n_in;
    const CollapseItem it = active ? in[idx] : CollapseItem{~0, 0};
    int32_t ref[W];
    int nc = 0;
    auto cnt_of = [&](int32_t r) -> int { return r < 0 ? 1 : bn[r].last - bn[r].first + 1; };
    auto lo_of = [&](int32_t r) { return r < 0 ? leaf_lo[~r] : node_lo[r]; };
    auto hi_of = [&](int32_t r) { return r < 0 ? leaf_hi[~r] : node_hi[r]; };
    if (!active) {
        // no item: only takes part in the wave's allocations
    } else if (it.bref < 0 || cnt_of(it.bref) <= leaf_size) {
        ref[nc++] = it.bref;   // whole range is one leaf (tiny scenes)
    } else {
        ref[nc++] = bn[it.bref].left;
        ref[nc++] = bn[it.bref].right;
        while (nc < W) {
            int best = -1;
            float best_a = -1.f;
            for (int c = 0; c < nc; c++) {
                if (ref[c] >= 0 && cnt_of(ref[c]) > leaf_size) {
                    const float a = half_area(lo_of(ref[c]), hi_of(ref[c]));
                    if (a > best_a) { best_a = a; best = c; }
                }
            }
            if (best < 0) break;
            const BNode b = bn[ref[best]];
            ref[best] = b.left;
            ref[nc++] = b.right;
        }
    }
    auto is_leaf = [&](int32_t r) { return r < 0 || cnt_of(r) <= leaf_size; };
    emit_wide_node<W>(active, it, ref, nc, is_leaf, bn, node_lo, node_hi, leaf_lo, leaf_hi, out, n_out, node_counter,
                      wnodes, sah_acc, leaf_counter);
}

// ---------------------------------------------------------------------------
// ray-cast kernels
// ---------------------------------------------------------------------------

constexpr int kTraceBlock = 256;
constexpr int kLdsDepth = 24;
constexpr int kPrivDepth = 72;

// Persistent closest-hit / any-hit kernel over a ray array (bvh.h,
// trace_persistent). `next` is this launch's ray counter, zeroed by
// k_zero_counter on the same stream.
template <int W, bool ANY, bool COUNT, class Node>
__global__ __launch_bounds__(kTraceBlock) void k_trace(const Node* __restrict__ nodes,
                                                        const TriV* __restrict__ tris,
                                                        const lmgpu_ray* __restrict__ rays,
                                                        lmgpu_hit* __restrict__ hits, uint8_t* __restrict__ occl,
                                                        int64_t n, unsigned long long* __restrict__ next,
                                                        int32_t* __restrict__ status,
                                                        unsigned long long* __restrict__ counters, int refill) {
    __shared__ int lds[kLdsDepth][kTraceBlock];
    auto fetch = [&](int64_t id, f3& o, f3& d, float& tmin, float& tmax) -> int {
        const float4* rp = reinterpret_cast<const float4*>(rays + id);
        const float4 a = rp[0], c = rp[1];
        o = mk(a.x, a.y, a.z);
        d = mk(c.x, c.y, c.z);
        tmin = a.w;
        tmax = c.w;
        return 0;
    };
    auto emit = [&](int64_t id, const TraceResult& r) {
        if (ANY) {
            occl[id] = r.prim >= 0 ? 1 : 0;
        } else {
            reinterpret_cast<float4*>(hits)[id] =
                make_float4(__int_as_float(r.prim), r.prim >= 0 ? r.t : 0.f, r.u, r.v);
        }
    };
    trace_persistent<W, ANY, COUNT, kLdsDepth, kPrivDepth, kTraceBlock>(nodes, tris, n, next, lds, fetch, emit,
                                                                         status, counters, refill, LMGPU_TRACE_RAY_CHUNK);
}

// Quantised copy of a built W-wide node array (bvh.h, quantise_node).
template <int W>
__global__ __launch_bounds__(256) void k_quantise(const WNode<W>* __restrict__ in, QNode<W>* __restrict__ out,
                                                  int64_t n) {
    const int64_t i = int64_t(blockIdx.x) * 256 + threadIdx.x;
    if (i < n) out[i] = quantise_node<W>(in[i]);
}

__global__ void k_zero_i32_1(int32_t* c) {
    if (threadIdx.x == 0) *c = 0;
}

__global__ void k_zero_counter(unsigned long long* c) {
    if (threadIdx.x == 0) *c = 0ull;
}

// ---------------------------------------------------------------------------
// host side
// ---------------------------------------------------------------------------

static inline unsigned nblocks(int64_t n, int b) { return unsigned((n + b - 1) / b); }

template <int W>
static void launch_collapse(bool sah_collapse, const CollapseItem* in, const int32_t* n_in, CollapseItem* out,
                            int32_t* n_out, int32_t* node_counter, const BNode* bn, const float4* node_lo,
                            const float4* node_hi, const float4* leaf_lo, const float4* leaf_hi,
                            const uint8_t* choice, void* wnodes, int leaf_size, double* sah, int32_t* leaves,
                            int64_t count, hipStream_t s) {
    if (sah_collapse)
        hipLaunchKernelGGL(k_collapse_sah<W>, dim3(nblocks(count, 256)), dim3(256), 0, s, in, n_in, out, n_out,
                           node_counter, bn, node_lo, node_hi, leaf_lo, leaf_hi, choice,
                           static_cast<WNode<W>*>(wnodes), sah, leaves);
    else
        hipLaunchKernelGGL(k_collapse<W>, dim3(nblocks(count, 256)), dim3(256), 0, s, in, n_in, out, n_out,
                           node_counter, bn, node_lo, node_hi, leaf_lo, leaf_hi, static_cast<WNode<W>*>(wnodes),
                           leaf_size, sah, leaves);
}

int Accel::build(const float* verts_in, int64_t n, bool on_device, const lmgpu_build_opts* opts, hipStream_t s) {
    int W = opts && opts->width ? opts->width : 4;
    int leaf = opts && opts->leaf_size ? opts->leaf_size : kDefaultLeafSize;
    const int rf = opts && opts->refill ? opts->refill : kDefaultRefill;
    if (rf < 1 || rf > 64) return fail(LMGPU_ERR_INVALID_ARGUMENT, "refill must be in [1, 64]");
    const int builder = opts ? opts->builder : LMGPU_BUILDER_LBVH;
    if (builder != LMGPU_BUILDER_PLOC && builder != LMGPU_BUILDER_LBVH)
        return fail(LMGPU_ERR_INVALID_ARGUMENT, "unknown builder");
    const int collapse = opts ? opts->collapse : LMGPU_COLLAPSE_SAH;
    if (collapse != LMGPU_COLLAPSE_SAH && collapse != LMGPU_COLLAPSE_GREEDY)
        return fail(LMGPU_ERR_INVALID_ARGUMENT, "unknown collapse");
    const float ctri = opts && opts->sah_tri_cost > 0.f ? opts->sah_tri_cost : kDefaultSahTriCost;
    if (!(ctri > 0.f && ctri < 1e6f)) return fail(LMGPU_ERR_INVALID_ARGUMENT, "sah_tri_cost out of range");
    const int radius = opts && opts->ploc_radius ? opts->ploc_radius : 16;
    if (W != 2 && W != 4 && W != 8) return fail(LMGPU_ERR_INVALID_ARGUMENT, "BVH width must be 2, 4 or 8");
    if (leaf < 1 || leaf > 15) return fail(LMGPU_ERR_INVALID_ARGUMENT, "leaf_size must be in [1, 15]");
    if (n <= 0) return fail(LMGPU_ERR_INVALID_ARGUMENT, "accel build needs at least one triangle");
    if (n >= (int64_t(1) << 27)) return fail(LMGPU_ERR_INVALID_ARGUMENT, "too many triangles (max 2^27)");
    if (!verts_in) return fail(LMGPU_ERR_INVALID_ARGUMENT, "verts is NULL");
    release();
    width = W;
    leaf_size = leaf;
    refill = rf;
    num_tris = n;
    DevBuf<float> dverts;
    const float* verts = verts_in;
    if (!on_device) {
        LMGPU_TRY(dverts.alloc(size_t(n) * 9));
        LMGPU_HIP(hipMemcpyAsync(dverts.p, verts_in, sizeof(float) * 9 * n, hipMemcpyHostToDevice, s));
        verts = dverts.p;
    }
    hipEvent_t e0, e1;
    LMGPU_HIP(hipEventCreate(&e0));
    LMGPU_HIP(hipEventCreate(&e1));
    LMGPU_HIP(hipEventRecord(e0, s));

    const int64_t m = std::max<int64_t>(n - 1, 1);   // binary internal nodes
    DevBuf<float4> tlo, thi, leaf_lo, leaf_hi, node_lo, node_hi;
    DevBuf<float> cb, partial, dp_cost;
    DevBuf<uint8_t> dp_choice;
    DevBuf<uint64_t> keys, keys2;
    DevBuf<int32_t> vals, vals2, parent_int, parent_leaf, flags, counters;
    DevBuf<BNode> bn;
    DevBuf<CollapseItem> q0, q1;
    DevBuf<double> sah;
    LMGPU_TRY(tlo.alloc(n)); LMGPU_TRY(thi.alloc(n));
    LMGPU_TRY(leaf_lo.alloc(n)); LMGPU_TRY(leaf_hi.alloc(n));
    LMGPU_TRY(node_lo.alloc(m)); LMGPU_TRY(node_hi.alloc(m));
    LMGPU_TRY(cb.alloc(6));
    const unsigned nb_prep = nblocks(n, 256);
    LMGPU_TRY(partial.alloc(size_t(nb_prep) * 6));
    LMGPU_TRY(keys.alloc(n)); LMGPU_TRY(keys2.alloc(n));
    LMGPU_TRY(vals.alloc(n)); LMGPU_TRY(vals2.alloc(n));
    LMGPU_TRY(parent_int.alloc(m));
    LMGPU_TRY(parent_leaf.alloc(n));
    LMGPU_TRY(flags.alloc(m));
    LMGPU_TRY(counters.alloc(8));
    LMGPU_TRY(bn.alloc(m));
    LMGPU_TRY(q0.alloc(n)); LMGPU_TRY(q1.alloc(n));
    LMGPU_TRY(sah.alloc(1));
    SahDP dp{nullptr, nullptr, W, leaf, ctri};
    if (collapse == LMGPU_COLLAPSE_SAH) {
        LMGPU_TRY(dp_cost.alloc(size_t(m) * W));
        LMGPU_TRY(dp_choice.alloc(size_t(m) * W));
        dp.cost = dp_cost.p;
        dp.choice = dp_choice.p;
    }

    hipLaunchKernelGGL(k_prep, dim3(nb_prep), dim3(256), 0, s, verts, n, tlo.p, thi.p, partial.p);
    hipLaunchKernelGGL(k_reduce_bounds, dim3(1), dim3(256), 0, s, partial.p, int(nb_prep), cb.p);
    hipLaunchKernelGGL(k_morton, dim3(nblocks(n, 256)), dim3(256), 0, s, tlo.p, thi.p, n, cb.p, keys.p, vals.p);
    bool in_alt = false;
    LMGPU_TRY(radix_sort_pairs_u64(keys.p, vals.p, keys2.p, vals2.p, n, 63, &in_alt, s));
    const uint64_t* skeys = in_alt ? keys2.p : keys.p;
    const int32_t* svals = in_alt ? vals2.p : vals.p;
    if (n > 1) {
        // the sort's result must be ordered before the tree is built on it
        LMGPU_HIP(hipMemsetAsync(counters.p, 0, sizeof(int32_t), s));
        hipLaunchKernelGGL(k_check_sorted, dim3(nblocks(n - 1, 256)), dim3(256), 0, s, skeys, n, counters.p);
        int32_t unsorted = 0;
        LMGPU_HIP(hipMemcpyAsync(&unsorted, counters.p, sizeof(int32_t), hipMemcpyDeviceToHost, s));
        LMGPU_HIP(hipStreamSynchronize(s));
        if (unsorted) return fail(LMGPU_ERR_DEVICE, "Morton sort produced an unordered key array");
    }
    // binary tree over the sorted triangles: bn, parent links, and `order`
    // (traversal position -> triangle)
    int32_t root_ref = n > 1 ? 0 : ~0;
    const int32_t* order = svals;
    DevBuf<int32_t> perm;
    if (n > 1) LMGPU_HIP(hipMemsetAsync(parent_int.p, 0xff, sizeof(int32_t) * (n - 1), s));
    if (builder == LMGPU_BUILDER_PLOC) {
        LMGPU_TRY(perm.alloc(n));
        LMGPU_TRY(build_ploc(n, svals, tlo.p, thi.p, radius, bn.p, node_lo.p, node_hi.p, leaf_lo.p, leaf_hi.p,
                             perm.p, &root_ref, s));
        order = perm.p;
        if (n > 1)
            hipLaunchKernelGGL(k_parents, dim3(nblocks(n - 1, 256)), dim3(256), 0, s, bn.p, n - 1, parent_int.p,
                               parent_leaf.p);
    } else if (n > 1) {
        hipLaunchKernelGGL(k_lbvh, dim3(nblocks(n - 1, 256)), dim3(256), 0, s, skeys, n, bn.p, parent_int.p,
                           parent_leaf.p);
    }
    {
        // frontier refit (bounds, and the SAH cost tables when collapsing by
        // SAH): q0/q1 (reinterpreted as int32 lists) hold the frontiers
        if (n > 1) LMGPU_HIP(hipMemsetAsync(flags.p, 0, sizeof(int32_t) * (n - 1), s));
        int32_t* fr[2] = {reinterpret_cast<int32_t*>(q0.p), reinterpret_cast<int32_t*>(q1.p)};
        LMGPU_HIP(hipMemsetAsync(counters.p, 0, 2 * sizeof(int32_t), s));
        hipLaunchKernelGGL(k_refit_leaves, dim3(nblocks(n, 256)), dim3(256), 0, s, n, order, tlo.p, thi.p,
                           parent_leaf.p, leaf_lo.p, leaf_hi.p, flags.p, fr[0], counters.p);
        // A frontier is never larger than the one before it (each parent
        // enters from exactly one child of the previous frontier), so the
        // grid sized for the last count read back serves the next
        // kRefitLevelsPerSync levels, each launch reading its own count.
        int cur = 0;
        int levels_refit = 0;
        while (n > 1) {
            int32_t cnt = 0;
            LMGPU_HIP(hipMemcpyAsync(&cnt, counters.p + cur, sizeof(int32_t), hipMemcpyDeviceToHost, s));
            LMGPU_HIP(hipStreamSynchronize(s));
            if (cnt == 0) break;
            for (int l = 0; l < kRefitLevelsPerSync; l++) {
                hipLaunchKernelGGL(k_zero_i32_1, dim3(1), dim3(64), 0, s, counters.p + (cur ^ 1));
                hipLaunchKernelGGL(k_refit_nodes, dim3(std::min(nblocks(cnt, 256), 2048u)), dim3(256), 0, s, fr[cur],
                                   counters.p + cur, bn.p, root_ref, parent_int.p, leaf_lo.p, leaf_hi.p, node_lo.p,
                                   node_hi.p, flags.p, fr[cur ^ 1], counters.p + (cur ^ 1), dp);
                cur ^= 1;
            }
            levels_refit += kRefitLevelsPerSync;
            if (levels_refit > 4096) return fail(LMGPU_ERR_DEVICE, "BVH refit did not terminate");
        }
    }
    LMGPU_TRY(tris.alloc(n));
    hipLaunchKernelGGL(k_reorder, dim3(nblocks(n, 256)), dim3(256), 0, s, n, order, verts, tris.p);

    // collapse, level by level
    const size_t node_bytes = size_t(32) * W;
    const int64_t max_nodes = std::max<int64_t>(n, 1);
    LMGPU_TRY(nodes.alloc(size_t(max_nodes) * node_bytes / sizeof(float4)));
    const CollapseItem root{root_ref, 0};
    LMGPU_HIP(hipMemcpyAsync(q0.p, &root, sizeof(root), hipMemcpyHostToDevice, s));
    int32_t init[4] = {1, 0, 1, 0};   // n_in, n_out, node_counter, leaves
    LMGPU_HIP(hipMemcpyAsync(counters.p, init, sizeof(init), hipMemcpyHostToDevice, s));
    LMGPU_HIP(hipMemsetAsync(sah.p, 0, sizeof(double), s));
    int64_t count = 1;
    int levels = 0;
    CollapseItem* qin = q0.p;
    CollapseItem* qout = q1.p;
    while (count > 0) {
        levels++;
        LMGPU_HIP(hipMemsetAsync(counters.p + 1, 0, sizeof(int32_t), s));
        const bool by_sah = collapse == LMGPU_COLLAPSE_SAH;
        if (W == 2)
            launch_collapse<2>(by_sah, qin, counters.p, qout, counters.p + 1, counters.p + 2, bn.p, node_lo.p,
                               node_hi.p, leaf_lo.p, leaf_hi.p, dp.choice, nodes.p, leaf, sah.p, counters.p + 3, count,
                               s);
        else if (W == 4)
            launch_collapse<4>(by_sah, qin, counters.p, qout, counters.p + 1, counters.p + 2, bn.p, node_lo.p,
                               node_hi.p, leaf_lo.p, leaf_hi.p, dp.choice, nodes.p, leaf, sah.p, counters.p + 3, count,
                               s);
        else
            launch_collapse<8>(by_sah, qin, counters.p, qout, counters.p + 1, counters.p + 2, bn.p, node_lo.p,
                               node_hi.p, leaf_lo.p, leaf_hi.p, dp.choice, nodes.p, leaf, sah.p, counters.p + 3, count,
                               s);
        int32_t nout = 0;
        LMGPU_HIP(hipMemcpyAsync(&nout, counters.p + 1, sizeof(int32_t), hipMemcpyDeviceToHost, s));
        LMGPU_HIP(hipStreamSynchronize(s));
        LMGPU_HIP(hipMemcpyAsync(counters.p, counters.p + 1, sizeof(int32_t), hipMemcpyDeviceToDevice, s));
        count = nout;
        std::swap(qin, qout);
        if (levels > 4096) return fail(LMGPU_ERR_DEVICE, "BVH collapse did not terminate");
    }
    LMGPU_HIP(hipEventRecord(e1, s));
    int32_t fin[4];
    double sah_sum = 0;
    LMGPU_HIP(hipMemcpyAsync(fin, counters.p, sizeof(fin), hipMemcpyDeviceToHost, s));
    LMGPU_HIP(hipMemcpyAsync(&sah_sum, sah.p, sizeof(double), hipMemcpyDeviceToHost, s));
    LMGPU_HIP(hipStreamSynchronize(s));
    float ms = 0;
    LMGPU_HIP(hipEventElapsedTime(&ms, e0, e1));
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    // root half-area for the SAH normalisation
    float4 rlo, rhi;
    if (root_ref >= 0) {
        LMGPU_HIP(hipMemcpy(&rlo, node_lo.p + root_ref, sizeof(float4), hipMemcpyDeviceToHost));
        LMGPU_HIP(hipMemcpy(&rhi, node_hi.p + root_ref, sizeof(float4), hipMemcpyDeviceToHost));
    } else {
        LMGPU_HIP(hipMemcpy(&rlo, leaf_lo.p, sizeof(float4), hipMemcpyDeviceToHost));
        LMGPU_HIP(hipMemcpy(&rhi, leaf_hi.p, sizeof(float4), hipMemcpyDeviceToHost));
    }
    const double ra = double(rhi.x - rlo.x) * (rhi.y - rlo.y) + double(rhi.y - rlo.y) * (rhi.z - rlo.z) +
                      double(rhi.z - rlo.z) * (rhi.x - rlo.x);
    num_nodes = fin[2];
    num_leaves = fin[3];
    depth = levels;
    sah_cost = ra > 0 ? 1.0 + sah_sum / ra : 1.0;
    build_ms = ms;
    root_lo[0] = rlo.x; root_lo[1] = rlo.y; root_lo[2] = rlo.z;
    root_hi[0] = rhi.x; root_hi[1] = rhi.y; root_hi[2] = rhi.z;
    LMGPU_TRY(make_trav_nodes(s));
    LMGPU_HIP(hipGetLastError());
    built = true;
    return LMGPU_OK;
}

bool quantised_width(int width) {
    return (width == 4 && std::is_same<TravNode<4>::type, QNode<4>>::value) ||
           (width == 8 && std::is_same<TravNode<8>::type, QNode<8>>::value);
}

int quantise_nodes(int width, const float4* wnodes, float4* qnodes, int64_t n, hipStream_t s) {
    if (n <= 0) return LMGPU_OK;
    if (width == 4)
        hipLaunchKernelGGL(k_quantise<4>, dim3(nblocks(n, 256)), dim3(256), 0, s,
                           reinterpret_cast<const WNode<4>*>(wnodes), reinterpret_cast<QNode<4>*>(qnodes), n);
    else if (width == 8)
        hipLaunchKernelGGL(k_quantise<8>, dim3(nblocks(n, 256)), dim3(256), 0, s,
                           reinterpret_cast<const WNode<8>*>(wnodes), reinterpret_cast<QNode<8>*>(qnodes), n);
    else
        return fail(LMGPU_ERR_INVALID_ARGUMENT, "quantised nodes are 4- or 8-wide");
    LMGPU_HIP(hipGetLastError());
    return LMGPU_OK;
}

int Accel::make_trav_nodes(hipStream_t s) {
    qnodes.reset();
    if (!quantised_width(width)) return LMGPU_OK;
    LMGPU_TRY(qnodes.alloc(size_t(num_nodes) * (size_t(width) * 16 / sizeof(float4))));
    LMGPU_TRY(quantise_nodes(width, nodes.p, qnodes.p, num_nodes, s));
    LMGPU_HIP(hipStreamSynchronize(s));
    return LMGPU_OK;
}

// Resident grid of the persistent kernel: CUs x active blocks per CU.
template <int W, bool ANY, bool COUNT, class Node>
static int persistent_grid() {
    static int grid = 0;
    if (grid == 0) {
        int dev = 0, cus = 0, per = 0;
        (void)hipGetDevice(&dev);
        (void)hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev);
        (void)hipOccupancyMaxActiveBlocksPerMultiprocessor(&per, k_trace<W, ANY, COUNT, Node>, kTraceBlock, 0);
        grid = std::max(1, cus) * std::max(1, per);
    }
    return grid;
}

template <int W, bool ANY, bool COUNT>
static void launch_trace(const Accel& a, const lmgpu_ray* rays, lmgpu_hit* hits, uint8_t* occl, int64_t n,
                         unsigned long long* next, int32_t* status, unsigned long long* counters, hipStream_t s) {
    using Node = typename TravNode<W>::type;
    hipLaunchKernelGGL(k_zero_counter, dim3(1), dim3(64), 0, s, next);
    const int grid = int(std::min<int64_t>(persistent_grid<W, ANY, COUNT, Node>(), (n + kTraceBlock - 1) / kTraceBlock));
    hipLaunchKernelGGL((k_trace<W, ANY, COUNT, Node>), dim3(grid), dim3(kTraceBlock), 0, s,
                       static_cast<const Node*>(a.trav_nodes()), a.tris.p, rays, hits, occl, n, next, status, counters,
                       a.refill);
}

int Accel::trace(const lmgpu_ray* rays, lmgpu_hit* hits, uint8_t* occl, int64_t n, bool any,
                 unsigned long long* counters, hipStream_t s) const {
    LMGPU_TRY(check_trace_args(rays, hits, occl, n, any));
    if (n == 0) return LMGPU_OK;
    int32_t* st = status_dev();
    if (!st) return fail(LMGPU_ERR_OUT_OF_MEMORY, "status buffer");
    unsigned long long* next = next_counter();
    if (!next) return fail(LMGPU_ERR_OUT_OF_MEMORY, "ray counter");
    return trace_with(rays, hits, occl, n, any, counters, next, st, s);
}

int Accel::check_trace_args(const lmgpu_ray* rays, const lmgpu_hit* hits, const uint8_t* occl, int64_t n,
                            bool any) const {
    if (!built) return fail(LMGPU_ERR_INVALID_ARGUMENT, "accel is not built");
    if (n < 0) return fail(LMGPU_ERR_INVALID_ARGUMENT, "negative ray count");
    if (n == 0) return LMGPU_OK;
    if (!rays || (!any && !hits) || (any && !occl))
        return fail(LMGPU_ERR_INVALID_ARGUMENT, "NULL ray or result buffer");
    if ((reinterpret_cast<uintptr_t>(rays) & 15) || (hits && (reinterpret_cast<uintptr_t>(hits) & 15)))
        return fail(LMGPU_ERR_INVALID_ARGUMENT, "rays and hits must be 16-byte aligned");
    return LMGPU_OK;
}

int Accel::trace_with(const lmgpu_ray* rays, lmgpu_hit* hits, uint8_t* occl, int64_t n, bool any,
                      unsigned long long* counters, unsigned long long* next, int32_t* st, hipStream_t s) const {
    LMGPU_TRY(check_trace_args(rays, hits, occl, n, any));
    if (n == 0) return LMGPU_OK;
#define LMGPU_DISPATCH(WW)                                                                          \
    if (width == WW) {                                                                              \
        if (any) launch_trace<WW, true, false>(*this, rays, hits, occl, n, next, st, counters, s);  \
        else if (counters) launch_trace<WW, false, true>(*this, rays, hits, occl, n, next, st, counters, s); \
        else launch_trace<WW, false, false>(*this, rays, hits, occl, n, next, st, counters, s);     \
    }
    LMGPU_DISPATCH(2)
    LMGPU_DISPATCH(4)
    LMGPU_DISPATCH(8)
#undef LMGPU_DISPATCH
    LMGPU_HIP(hipGetLastError());
    return LMGPU_OK;
}

}  // namespace lmgpu
