// PLOC binary BVH builder (Meister & Bittner, "Parallel Locally-Ordered
// Clustering for Bounding Volume Hierarchy Construction", IEEE TVCG 2018).
//
// Clusters start as the Morton-sorted triangles. Each iteration:
//   nn      every cluster finds the neighbour within +-radius positions whose
//           merged box has the smallest surface area (LDS-tiled window);
//   flags   mutual nearest neighbours merge (the lower index keeps the new
//           node), the higher index disappears;
//   scan    one exclusive scan of packed (merges, survivors) counts gives both
//           the new node ids and the compacted positions, so node numbering
//           and the tree are deterministic;
//   commit  writes the new nodes and the compacted cluster array.
// Pair keys are symmetric and totally ordered (area, then parity of the lower
// index, then index distance), so equal-area chains pair up (0,1),(2,3),...;
// an iteration that merges nothing falls back to pairing neighbours.
// Afterwards every leaf's depth-first position is found by walking to the
// root, giving contiguous leaf ranges for the wide-BVH collapse.

#include <algorithm>
#include <cstring>
#include <string>
#include <vector>

#include "build.h"
#include "lmgpu_internal.h"

namespace lmgpu {

namespace {

constexpr int kB = 256;
constexpr int kMaxRadius = 64;

struct PNode {
    int32_t left, right, count, pad;
};

__device__ __forceinline__ float half_area3(float4 lo, float4 hi) {
    const float dx = hi.x - lo.x, dy = hi.y - lo.y, dz = hi.z - lo.z;
    return dx * dy + dy * dz + dz * dx;
}

__global__ __launch_bounds__(kB) void k_ploc_init(int64_t n, const int32_t* __restrict__ order,
                                                  const float4* __restrict__ tlo, const float4* __restrict__ thi,
                                                  int32_t* __restrict__ cid, float4* __restrict__ clo,
                                                  float4* __restrict__ chi, float4* __restrict__ slo,
                                                  float4* __restrict__ shi) {
    const int64_t i = int64_t(blockIdx.x) * kB + threadIdx.x;
    if (i >= n) return;
    const int32_t t = order[i];
    const float4 lo = tlo[t], hi = thi[t];
    cid[i] = ~int32_t(i);
    clo[i] = lo;
    chi[i] = hi;
    slo[i] = lo;
    shi[i] = hi;
}

__global__ __launch_bounds__(kB) void k_ploc_nn(int count, const float4* __restrict__ clo,
                                                const float4* __restrict__ chi, int32_t* __restrict__ nn, int r) {
    __shared__ float4 slo[kB + 2 * kMaxRadius], shi[kB + 2 * kMaxRadius];
    const int b0 = blockIdx.x * kB;
    for (int k = threadIdx.x; k < kB + 2 * r; k += kB) {
        const int g = b0 - r + k;
        if (g >= 0 && g < count) {
            slo[k] = clo[g];
            shi[k] = chi[g];
        }
    }
    __syncthreads();
    const int i = b0 + threadIdx.x;
    if (i >= count) return;
    const float4 lo = slo[threadIdx.x + r], hi = shi[threadIdx.x + r];
    uint64_t best = ~0ull;
    int bj = i;
    const int jlo = max(0, i - r), jhi = min(count - 1, i + r);
    for (int j = jlo; j <= jhi; j++) {
        if (j == i) continue;
        const float4 l2 = slo[j - b0 + r], h2 = shi[j - b0 + r];
        const float4 ul = make_float4(fminf(lo.x, l2.x), fminf(lo.y, l2.y), fminf(lo.z, l2.z), 0.f);
        const float4 uh = make_float4(fmaxf(hi.x, h2.x), fmaxf(hi.y, h2.y), fmaxf(hi.z, h2.z), 0.f);
        const float a = fmaxf(half_area3(ul, uh), 0.f);
        const int mn = min(i, j);
        const int dist = i > j ? i - j : j - i;
        const uint64_t key = (uint64_t(__float_as_uint(a)) << 32) | (uint64_t(mn & 1) << 16) | uint64_t(dist);
        if (key < best) {
            best = key;
            bj = j;
        }
    }
    nn[i] = bj;
}

// Neighbour pairing fallback: (0,1), (2,3), ...
__global__ __launch_bounds__(kB) void k_ploc_pairs(int count, int32_t* __restrict__ nn) {
    const int i = blockIdx.x * kB + threadIdx.x;
    if (i >= count) return;
    const int j = i ^ 1;
    nn[i] = j < count ? j : i;
}

__global__ __launch_bounds__(kB) void k_ploc_flags(int count, const int32_t* __restrict__ nn,
                                                   unsigned long long* __restrict__ flags) {
    const int i = blockIdx.x * kB + threadIdx.x;
    if (i >= count) return;
    const int j = nn[i];
    const bool mutual = j != i && nn[j] == i;
    const bool merge = mutual && i < j;
    const bool valid = !(mutual && i > j);
    flags[i] = (uint64_t(merge ? 1 : 0) << 32) | uint64_t(valid ? 1 : 0);
}

__device__ __forceinline__ int ref_count(int32_t ref, const PNode* __restrict__ nodes) {
    return ref < 0 ? 1 : nodes[ref].count;
}

__global__ __launch_bounds__(kB) void k_ploc_commit(int count, const int32_t* __restrict__ nn,
                                                    const int32_t* __restrict__ cid, const float4* __restrict__ clo,
                                                    const float4* __restrict__ chi,
                                                    const unsigned long long* __restrict__ flags,
                                                    const unsigned long long* __restrict__ pre, int node_base,
                                                    PNode* __restrict__ nodes, float4* __restrict__ node_lo,
                                                    float4* __restrict__ node_hi, int32_t* __restrict__ parent_leaf,
                                                    int32_t* __restrict__ parent_int, int32_t* __restrict__ cid2,
                                                    float4* __restrict__ clo2, float4* __restrict__ chi2,
                                                    int32_t* __restrict__ totals) {
    const int i = blockIdx.x * kB + threadIdx.x;
    if (i >= count) return;
    const unsigned long long f = flags[i], p = pre[i];
    const bool merge = (f >> 32) != 0;
    const bool valid = (f & 1ull) != 0;
    int32_t c = cid[i];
    float4 lo = clo[i], hi = chi[i];
    if (merge) {
        const int j = nn[i];
        const int32_t a = c, b = cid[j];
        const int k = node_base + int(p >> 32);
        const float4 l2 = clo[j], h2 = chi[j];
        lo = make_float4(fminf(lo.x, l2.x), fminf(lo.y, l2.y), fminf(lo.z, l2.z), 0.f);
        hi = make_float4(fmaxf(hi.x, h2.x), fmaxf(hi.y, h2.y), fmaxf(hi.z, h2.z), 0.f);
        nodes[k] = PNode{a, b, ref_count(a, nodes) + ref_count(b, nodes), 0};
        node_lo[k] = lo;
        node_hi[k] = hi;
        if (a < 0) parent_leaf[~a] = k; else parent_int[a] = k;
        if (b < 0) parent_leaf[~b] = k; else parent_int[b] = k;
        c = k;
    }
    if (valid) {
        const int dst = int(uint32_t(p));
        cid2[dst] = c;
        clo2[dst] = lo;
        chi2[dst] = hi;
    }
    if (i == count - 1) {
        totals[0] = int(uint32_t(p)) + (valid ? 1 : 0);
        totals[1] = int(p >> 32) + (merge ? 1 : 0);
    }
}

// Depth-first position of a subtree's first leaf: walk to the root adding
// the size of every left sibling passed on the way.
__device__ __forceinline__ int dfs_first(int32_t ref, int32_t root, const PNode* __restrict__ nodes,
                                         const int32_t* __restrict__ parent_leaf,
                                         const int32_t* __restrict__ parent_int) {
    int pos = 0;
    while (ref != root) {
        const int32_t par = ref < 0 ? parent_leaf[~ref] : parent_int[ref];
        const PNode pn = nodes[par];
        if (pn.right == ref) pos += ref_count(pn.left, nodes);
        ref = par;
    }
    return pos;
}

__global__ __launch_bounds__(kB) void k_ploc_leafpos(int64_t n, int32_t root, const PNode* __restrict__ nodes,
                                                     const int32_t* __restrict__ parent_leaf,
                                                     const int32_t* __restrict__ parent_int,
                                                     const int32_t* __restrict__ order, const float4* __restrict__ slo,
                                                     const float4* __restrict__ shi, int32_t* __restrict__ pos,
                                                     float4* __restrict__ leaf_lo, float4* __restrict__ leaf_hi,
                                                     int32_t* __restrict__ perm) {
    const int64_t j = int64_t(blockIdx.x) * kB + threadIdx.x;
    if (j >= n) return;
    const int p = dfs_first(~int32_t(j), root, nodes, parent_leaf, parent_int);
    pos[j] = p;
    leaf_lo[p] = slo[j];
    leaf_hi[p] = shi[j];
    perm[p] = order[j];
}

__global__ __launch_bounds__(kB) void k_ploc_nodes(int64_t ni, int32_t root, const PNode* __restrict__ nodes,
                                                   const int32_t* __restrict__ parent_leaf,
                                                   const int32_t* __restrict__ parent_int,
                                                   const int32_t* __restrict__ pos, BNode* __restrict__ bn) {
    const int64_t k = int64_t(blockIdx.x) * kB + threadIdx.x;
    if (k >= ni) return;
    const int first = dfs_first(int32_t(k), root, nodes, parent_leaf, parent_int);
    const PNode pn = nodes[k];
    BNode b;
    b.left = pn.left < 0 ? ~pos[~pn.left] : pn.left;
    b.right = pn.right < 0 ? ~pos[~pn.right] : pn.right;
    b.first = first;
    b.last = first + pn.count - 1;
    bn[k] = b;
}

inline unsigned nb(int64_t n) { return unsigned((n + kB - 1) / kB); }

}  // namespace

int build_ploc(int64_t n, const int32_t* order, const float4* tlo, const float4* thi, int radius, BNode* bn,
               float4* node_lo, float4* node_hi, float4* leaf_lo, float4* leaf_hi, int32_t* perm, int32_t* root_ref,
               hipStream_t s) {
    if (radius < 1 || radius > kMaxRadius) return fail(LMGPU_ERR_INVALID_ARGUMENT, "PLOC radius out of range");
    DevBuf<int32_t> cid[2], nn, parent_leaf, parent_int, pos, totals;
    DevBuf<float4> clo[2], chi[2], slo, shi;
    DevBuf<unsigned long long> flags, pre;
    DevBuf<PNode> nodes;
    for (int b = 0; b < 2; b++) {
        LMGPU_TRY(cid[b].alloc(n));
        LMGPU_TRY(clo[b].alloc(n));
        LMGPU_TRY(chi[b].alloc(n));
    }
    LMGPU_TRY(nn.alloc(n));
    LMGPU_TRY(parent_leaf.alloc(n));
    LMGPU_TRY(parent_int.alloc(std::max<int64_t>(n - 1, 1)));
    LMGPU_TRY(pos.alloc(n));
    LMGPU_TRY(totals.alloc(2));
    LMGPU_TRY(slo.alloc(n));
    LMGPU_TRY(shi.alloc(n));
    LMGPU_TRY(flags.alloc(n));
    LMGPU_TRY(pre.alloc(n));
    LMGPU_TRY(nodes.alloc(std::max<int64_t>(n - 1, 1)));
    hipLaunchKernelGGL(k_ploc_init, dim3(nb(n)), dim3(kB), 0, s, n, order, tlo, thi, cid[0].p, clo[0].p, chi[0].p,
                       slo.p, shi.p);
    const size_t temp_bytes = exclusive_scan_u64_temp_bytes(n);
    DevBuf<char> temp;
    LMGPU_TRY(temp.alloc(temp_bytes));
    int count = int(n);
    int node_base = 0;
    int cur = 0;
    int iters = 0;
    bool force_pairs = false;
    while (count > 1) {
        if (force_pairs) hipLaunchKernelGGL(k_ploc_pairs, dim3(nb(count)), dim3(kB), 0, s, count, nn.p);
        else hipLaunchKernelGGL(k_ploc_nn, dim3(nb(count)), dim3(kB), 0, s, count, clo[cur].p, chi[cur].p, nn.p, radius);
        hipLaunchKernelGGL(k_ploc_flags, dim3(nb(count)), dim3(kB), 0, s, count, nn.p, flags.p);
        LMGPU_TRY(exclusive_scan_u64(reinterpret_cast<const uint64_t*>(flags.p), reinterpret_cast<uint64_t*>(pre.p),
                                     count, temp.p, temp_bytes, s));
        hipLaunchKernelGGL(k_ploc_commit, dim3(nb(count)), dim3(kB), 0, s, count, nn.p, cid[cur].p, clo[cur].p,
                           chi[cur].p, flags.p, pre.p, node_base, nodes.p, node_lo, node_hi, parent_leaf.p,
                           parent_int.p, cid[cur ^ 1].p, clo[cur ^ 1].p, chi[cur ^ 1].p, totals.p);
        int32_t tot[2];
        LMGPU_HIP(hipMemcpyAsync(tot, totals.p, sizeof(tot), hipMemcpyDeviceToHost, s));
        LMGPU_HIP(hipStreamSynchronize(s));
        if (tot[1] == 0 && !force_pairs) {
            force_pairs = true;   // no mutual pair in any window: pair neighbours once
            continue;
        }
        force_pairs = false;
        node_base += tot[1];
        count = tot[0];
        cur ^= 1;
        if (++iters > 100000) return fail(LMGPU_ERR_DEVICE, "PLOC did not converge");
    }
    int32_t root = ~0;
    LMGPU_HIP(hipMemcpyAsync(&root, cid[cur].p, sizeof(int32_t), hipMemcpyDeviceToHost, s));
    LMGPU_HIP(hipStreamSynchronize(s));
    if (n > 1 && (root < 0 || node_base != n - 1)) return fail(LMGPU_ERR_DEVICE, "PLOC produced an inconsistent tree");
    hipLaunchKernelGGL(k_ploc_leafpos, dim3(nb(n)), dim3(kB), 0, s, n, root, nodes.p, parent_leaf.p, parent_int.p,
                       order, slo.p, shi.p, pos.p, leaf_lo, leaf_hi, perm);
    if (n > 1)
        hipLaunchKernelGGL(k_ploc_nodes, dim3(nb(n - 1)), dim3(kB), 0, s, n - 1, root, nodes.p, parent_leaf.p,
                           parent_int.p, pos.p, bn);
    LMGPU_HIP(hipGetLastError());
    LMGPU_HIP(hipStreamSynchronize(s));
    *root_ref = root;
    return LMGPU_OK;
}

}  // namespace lmgpu
