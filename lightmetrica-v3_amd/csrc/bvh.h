// Wide BVH layout and closest-hit / any-hit traversal (device functions).
//
// Replaces the traversal of accel::sahbvh (src/accel/accel_sahbvh.cpp:269-299:
// FP64 slab test, unordered int[99] stack, Moller-Trumbore) with:
//   * a W-wide BVH (W = 2, 4, 8) whose node is 32*W bytes laid out SoA
//     (lo.x[W], lo.y[W], lo.z[W], hi.x[W], hi.y[W], hi.z[W], child[W], count[W]);
//     W = 4 is exactly one 128-B line,
//   * nearest-first ordered traversal with a per-lane stack in LDS
//     ([depth][lane], bank-conflict free) spilling to private memory,
//   * the watertight FP32 ray/triangle test of Woop, Benthin and Wald
//     (JCGT 2(1), 2013), with the reference's closed edges, closed
//     [tmin, tmax] range and |det| < 1e-8 rejection (accel_sahbvh.cpp:67-71).
#pragma once

#include <algorithm>
#include <cmath>

#include "common.h"
#include "../../include/lmgpu.h"

// Wave-level step vote of the persistent traversal, in quarters (0 = off):
// a wave takes a leaf step when at least LMGPU_LEAF_VOTE_X4 / 4 times as
// many of its busy lanes sit at a leaf as at an inner node (see
// trace_persistent). 2 measured best on the 1.32M-triangle mesh (incoherent
// rays +11%, coherent +5%, path tracing +5% over no vote; 1, 3, 4 and 8
// slower).
#ifndef LMGPU_LEAF_VOTE_X4
#define LMGPU_LEAF_VOTE_X4 2
#endif

// Traversal of a 4- or 8-wide BVH reads a quantised copy of its nodes
// (QNode<W>: 64 B instead of 128 B for W = 4, 128 B instead of 256 B for
// W = 8) when 1: +15% incoherent closest-hit rays on the 1.32M-triangle mesh
// at W = 4 with identical hits.
#ifndef LMGPU_QNODE
#define LMGPU_QNODE 1
#endif

// Rays a wave takes from the global ray counter with one atomic (at least
// 64, so one chunk always covers the idle lanes of a refill); the wave hands
// them to its idle lanes over several refills.
#ifndef LMGPU_RAY_CHUNK
#define LMGPU_RAY_CHUNK 64
#endif
static_assert(LMGPU_RAY_CHUNK >= 64, "a ray chunk must cover a whole wave");

namespace lmgpu {

template <int W>
struct alignas(16) WNode {
    float lox[W], loy[W], loz[W];
    float hix[W], hiy[W], hiz[W];
    int32_t child[W];   // internal: wide node index; leaf: first triangle (sorted order)
    int32_t count[W];   // 0 internal, > 0 leaf triangle count, < 0 empty slot
};
static_assert(sizeof(WNode<4>) == 128, "BVH4 node must be one 128-B line");

// Quantised W-wide node (W = 4: 64 B, half a 128-B line; W = 8: 128 B):
// the child boxes are 8-bit planes on a per-node, per-axis grid,
// plane = fma(q, 2^e, origin), rounded outward when built (quantise_node),
// so the decoded boxes contain the exact ones and traversal visits a
// superset of the nodes of the FP32 tree: hits are unchanged. Child c's
// lower / upper plane on axis a is byte c % 4 of qlo[a][c / 4] / qhi[a][c / 4];
// ent[c] is its stack entry (inner node index >= 0, or
// encode_leaf(first, count) < 0), kQEmpty for an empty slot.
template <int W>
struct alignas(W * 16) QNode {
    float ox, oy, oz;
    uint32_t e;        // biased FP32 exponents of the x, y, z grid steps in bytes 0, 1, 2
    uint32_t qlo[3][W / 4];
    uint32_t qhi[3][W / 4];
    int32_t ent[W];
    int32_t pad[W == 4 ? 2 : 8];
};
using QNode4 = QNode<4>;
static_assert(sizeof(QNode<4>) == 64, "quantised BVH4 node must be 64 B");
static_assert(sizeof(QNode<8>) == 128, "quantised BVH8 node must be 128 B");
constexpr int32_t kQEmpty = INT32_MIN;

// Node type the traversal of a W-wide BVH reads.
template <int W>
struct TravNode {
    using type = WNode<W>;
};
#if LMGPU_QNODE
template <>
struct TravNode<4> {
    using type = QNode<4>;
};
template <>
struct TravNode<8> {
    using type = QNode<8>;
};
#endif

// Triangle in traversal order: 3 x float4 = 48 B; v0.w holds the original
// (flattened) triangle index as raw bits.
struct alignas(16) TriV {
    float4 v0, v1, v2;
};

// Stack entries: internal node index (>= 0) or an encoded leaf (< 0).
LMGPU_HD int encode_leaf(int first, int count) { return -((first << 4) | count) - 1; }
LMGPU_HD int leaf_first(int e) { return (-(e + 1)) >> 4; }
LMGPU_HD int leaf_count(int e) { return (-(e + 1)) & 15; }

struct RayPre {
    f3 o, d, inv;
    float Sx, Sy, Sz;
    int kx, ky, kz;
    float dkz_abs;   // |d[kz]|, relates the watertight det to the MT det
};

LMGPU_HD RayPre ray_prepare(f3 o, f3 d) {
    RayPre r;
    r.o = o;
    r.d = d;
    // Reciprocal direction; a zero component becomes a tiny signed value so
    // the slab distances are +-huge rather than NaN.
    const float tiny = 1e-30f;
    const float dx = fabsf(d.x) < tiny ? copysignf(tiny, d.x) : d.x;
    const float dy = fabsf(d.y) < tiny ? copysignf(tiny, d.y) : d.y;
    const float dz = fabsf(d.z) < tiny ? copysignf(tiny, d.z) : d.z;
    r.inv = mk(1.0f / dx, 1.0f / dy, 1.0f / dz);
    const float ax = fabsf(d.x), ay = fabsf(d.y), az = fabsf(d.z);
    int kz = (ax > ay) ? (ax > az ? 0 : 2) : (ay > az ? 1 : 2);
    int kx = kz + 1 == 3 ? 0 : kz + 1;
    int ky = kx + 1 == 3 ? 0 : kx + 1;
    const float dkz = comp(d, kz);
    if (dkz < 0.f) {   // keep the winding of the projected triangle
        const int t = kx; kx = ky; ky = t;
    }
    r.kx = kx; r.ky = ky; r.kz = kz;
    r.Sx = comp(d, kx) / dkz;
    r.Sy = comp(d, ky) / dkz;
    r.Sz = 1.0f / dkz;
    r.dkz_abs = fabsf(dkz);
    return r;
}

// Watertight ray/triangle test. On acceptance (t in [tmin, tmax], closed
// edges, |det_MT| >= 1e-8 as the reference), writes t and the barycentric
// weights (u, v) of v1 and v2.
// Watertightness needs each edge function to be the exact negative of the
// same edge's function in the neighbouring triangle (a*b - c*d against
// c*d - a*b). Both products must therefore be rounded before the
// subtraction: fusing one of them into an FMA would break the antisymmetry
// and let a ray through a shared edge miss both triangles, so contraction is
// switched off for the three edge functions.
LMGPU_HD bool tri_intersect(const RayPre& r, float4 p0, float4 p1, float4 p2, float tmin, float tmax, float& t,
                            float& u, float& v) {
    const f3 A = mk(p0.x - r.o.x, p0.y - r.o.y, p0.z - r.o.z);
    const f3 B = mk(p1.x - r.o.x, p1.y - r.o.y, p1.z - r.o.z);
    const f3 C = mk(p2.x - r.o.x, p2.y - r.o.y, p2.z - r.o.z);
    const float Akz = comp(A, r.kz), Bkz = comp(B, r.kz), Ckz = comp(C, r.kz);
    const float Ax = fmaf(-r.Sx, Akz, comp(A, r.kx));
    const float Ay = fmaf(-r.Sy, Akz, comp(A, r.ky));
    const float Bx = fmaf(-r.Sx, Bkz, comp(B, r.kx));
    const float By = fmaf(-r.Sy, Bkz, comp(B, r.ky));
    const float Cx = fmaf(-r.Sx, Ckz, comp(C, r.kx));
    const float Cy = fmaf(-r.Sy, Ckz, comp(C, r.ky));
    float U, V, Wt;
    {
#pragma clang fp contract(off)
        U = Cx * By - Cy * Bx;
        V = Ax * Cy - Ay * Cx;
        Wt = Bx * Ay - By * Ax;
    }
    if (U == 0.0f || V == 0.0f || Wt == 0.0f) {
        // Exact edge: recompute the edge functions in FP64 (Woop et al., sec. 3).
        U = float(double(Cx) * double(By) - double(Cy) * double(Bx));
        V = float(double(Ax) * double(Cy) - double(Ay) * double(Cx));
        Wt = float(double(Bx) * double(Ay) - double(By) * double(Ax));
    }
    if ((U < 0.f || V < 0.f || Wt < 0.f) && (U > 0.f || V > 0.f || Wt > 0.f)) return false;
    const float det = U + V + Wt;
    if (det == 0.f) return false;
    // |det_MT| = |d . (e1 x e2)| = |det| |d[kz]| for a unit-length direction
    if (fabsf(det) * r.dkz_abs < 1e-8f) return false;
    const float Az = r.Sz * Akz, Bz = r.Sz * Bkz, Cz = r.Sz * Ckz;
    const float T = U * Az + V * Bz + Wt * Cz;
    const float sdet = fabsf(det);
    const float sT = det < 0.f ? -T : T;
    if (sT < tmin * sdet || sT > tmax * sdet) return false;
    const float rcp = 1.0f / det;
    t = T * rcp;
    u = V * rcp;
    v = Wt * rcp;
    return true;
}

// Slab test of one child box; conservative against FP32 rounding by scaling
// the far distance with 1 + 2*gamma(3) (Ize, "Robust BVH Ray Traversal",
// JCGT 2(2), 2013).
__device__ __forceinline__ float box_near(const RayPre& r, float lx, float ly, float lz, float hx, float hy,
                                          float hz, float tmin, float tmax, bool& hit) {
    const float t1x = (lx - r.o.x) * r.inv.x, t2x = (hx - r.o.x) * r.inv.x;
    const float t1y = (ly - r.o.y) * r.inv.y, t2y = (hy - r.o.y) * r.inv.y;
    const float t1z = (lz - r.o.z) * r.inv.z, t2z = (hz - r.o.z) * r.inv.z;
    const float tn = fmaxf(fmaxf(fminf(t1x, t2x), fminf(t1y, t2y)), fmaxf(fminf(t1z, t2z), tmin));
    const float tf = fminf(fminf(fmaxf(t1x, t2x), fmaxf(t1y, t2y)), fminf(fmaxf(t1z, t2z), tmax));
    hit = tn <= tf * 1.00000036f;
    return tn;
}

// Slab keys of the W children of one node: key[c] is the entry distance of
// a hit child (+inf otherwise), ent[c] its stack entry.
template <int W>
__device__ __forceinline__ void child_keys(const WNode<W>& nd, const RayPre& r, float tmin, float tmax,
                                           float (&key)[W], int (&ent)[W]) {
#pragma unroll
    for (int c = 0; c < W; c++) {
        bool hit;
        const float tn = box_near(r, nd.lox[c], nd.loy[c], nd.loz[c], nd.hix[c], nd.hiy[c], nd.hiz[c], tmin, tmax, hit);
        const int cnt = nd.count[c];
        const int ch = nd.child[c];
        hit = hit && cnt >= 0;
        key[c] = hit ? tn : __builtin_inff();
        ent[c] = cnt == 0 ? ch : encode_leaf(ch, cnt);
    }
}

LMGPU_HD float qplane(uint32_t q, int c, float step, float origin) {
    return std::fma(float((q >> (8 * c)) & 0xffu), step, origin);
}

// Slab keys of the W children of a quantised node. The slab distance of
// grid plane q on axis x is computed as one fma, t = q * ax + bx, with
//   ax = step_x * inv_x   (exact: the step is a power of two),
//   bx = (ox - o_x) * inv_x,
// and no decoded plane: the planes are the exact reals ox + q * step. Each
// axis picks its near and far plane by the sign of inv_x, so one fma gives
// the entry and one the exit distance (no per-axis min / max). Bounds, with
// u = 2^-24: bx is within 2u |bx| of (ox - o_x) * inv_x; the exact plane
// ox + q step lies within u (|ox| + 255 step) of the FP32 plane the builder
// rounded outward (quantise_node); the fma and inv_x round relatively, which
// the 1 + 2 gamma(3) factor on the exit distance covers (box_near). The entry
// offset is lowered and the exit offset raised by
//   Ex = 16u (|bx| + |inv_x| |ox| + 255 |ax|),
// which exceeds the absolute terms with their own rounding, so every child
// box the FP32 test of the decoded planes would hit is hit here too.
template <int W>
LMGPU_HD void child_keys(const QNode<W>& nd, const RayPre& r, float tmin, float tmax, float (&key)[W],
                         int (&ent)[W]) {
    constexpr float kSlack = 0x1p-20f;   // 16u
    const float org[3] = {nd.ox, nd.oy, nd.oz};
    const float o[3] = {r.o.x, r.o.y, r.o.z};
    const float inv[3] = {r.inv.x, r.inv.y, r.inv.z};
    float a[3], bn[3], bf[3];
    uint32_t wn[3][W / 4], wf[3][W / 4];
    // both plane words of every axis as loaded (whole 16-B loads of the node);
    // the empty asm keeps the compiler from turning the sign choice below into
    // per-lane dword loads from chosen addresses
    uint32_t ql[3][W / 4], qh[3][W / 4];
#pragma unroll
    for (int k = 0; k < 3; k++)
#pragma unroll
        for (int w = 0; w < W / 4; w++) {
            ql[k][w] = nd.qlo[k][w];
            qh[k][w] = nd.qhi[k][w];
#if defined(__HIP_DEVICE_COMPILE__)
            asm("" : "+v"(ql[k][w]), "+v"(qh[k][w]));
#endif
        }
#pragma unroll
    for (int k = 0; k < 3; k++) {
        const float step = __builtin_bit_cast(float, ((nd.e >> (8 * k)) & 0xffu) << 23);
        a[k] = step * inv[k];
        const float b = (org[k] - o[k]) * inv[k];
        const float E = (fabsf(b) + fmaf(fabsf(org[k]), fabsf(inv[k]), 255.f * fabsf(a[k]))) * kSlack;
        bn[k] = b - E;
        bf[k] = b + E;
        const bool pos = inv[k] >= 0.f;
#pragma unroll
        for (int w = 0; w < W / 4; w++) {
            wn[k][w] = pos ? ql[k][w] : qh[k][w];
            wf[k][w] = pos ? qh[k][w] : ql[k][w];
        }
    }
#pragma unroll
    for (int c = 0; c < W; c++) {
        const int w = c >> 2, sh = 8 * (c & 3);
        float tn[3], tf[3];
#pragma unroll
        for (int k = 0; k < 3; k++) {
            tn[k] = fmaf(float((wn[k][w] >> sh) & 0xffu), a[k], bn[k]);
            tf[k] = fmaf(float((wf[k][w] >> sh) & 0xffu), a[k], bf[k]);
        }
        const float n = fmaxf(fmaxf(tn[0], tn[1]), fmaxf(tn[2], tmin));
        const float f = fminf(fminf(tf[0], tf[1]), fminf(tf[2], tmax));
        const int e = nd.ent[c];
        const bool hit = n <= f * 1.00000036f && e != kQEmpty;
        key[c] = hit ? n : __builtin_inff();
        ent[c] = e;
    }
}

// Quantises a built W-wide node: per axis the grid starts at the smallest
// child plane, with the smallest power-of-two step that spans the node in
// 255 steps; each plane is rounded outward in the exact FP32 expression the
// traversal decodes (qplane), so every decoded box contains the original.
template <int W>
LMGPU_HD QNode<W> quantise_node(const WNode<W>& n) {
    QNode<W> q;
    const float* lo[3] = {n.lox, n.loy, n.loz};
    const float* hi[3] = {n.hix, n.hiy, n.hiz};
    float org[3];
    uint32_t ebits = 0;
    for (int a = 0; a < 3; a++) {
        float mn = __builtin_inff(), mx = -__builtin_inff();
        for (int c = 0; c < W; c++)
            if (n.count[c] >= 0) {
                mn = std::min(mn, lo[a][c]);
                mx = std::max(mx, hi[a][c]);
            }
        if (!(mn <= mx)) mn = mx = 0.f;   // no child (not produced by the builder)
        org[a] = mn;
        // smallest e with fma(255, 2^e, mn) >= mx; steps of at least 2^-126.
        // Start from ceil(log2(span / 255)), read off the binary exponent:
        // x = m 2^k with m in [0.5, 1) gives ceil(log2 x) = k, or k - 1 when
        // m = 0.5 exactly.
        int e = -126;
        const double span = double(mx) - double(mn);
        if (span > 0.0) {
            int k = 0;
            const double m = std::frexp(span / 255.0, &k);
            e = std::max(-126, m == 0.5 ? k - 1 : k);
        }
        while (e < 127 && std::fma(255.f, std::ldexp(1.f, e), mn) < mx) e++;
        ebits |= uint32_t(e + 127) << (8 * a);
        uint32_t ql[W / 4] = {}, qh[W / 4] = {};
        const float step = std::ldexp(1.f, e);
        for (int c = 0; c < W; c++) {
            int l = 0, h = 0;
            if (n.count[c] >= 0) {
                // division by the power-of-two step is an exact scaling
                l = int(std::floor(std::ldexp(double(lo[a][c]) - double(mn), -e)));
                l = std::min(std::max(l, 0), 255);
                while (l > 0 && std::fma(float(l), step, mn) > lo[a][c]) l--;
                h = int(std::ceil(std::ldexp(double(hi[a][c]) - double(mn), -e)));
                h = std::min(std::max(h, 0), 255);
                while (h < 255 && std::fma(float(h), step, mn) < hi[a][c]) h++;
            }
            ql[c >> 2] |= uint32_t(l) << (8 * (c & 3));
            qh[c >> 2] |= uint32_t(h) << (8 * (c & 3));
        }
        for (int k = 0; k < W / 4; k++) {
            q.qlo[a][k] = ql[k];
            q.qhi[a][k] = qh[k];
        }
    }
    q.ox = org[0];
    q.oy = org[1];
    q.oz = org[2];
    q.e = ebits;
    for (int c = 0; c < W; c++)
        q.ent[c] = n.count[c] < 0 ? kQEmpty : (n.count[c] == 0 ? n.child[c] : encode_leaf(n.child[c], n.count[c]));
    for (int k = 0; k < int(sizeof(q.pad) / sizeof(q.pad[0])); k++) q.pad[k] = 0;
    return q;
}

// Per-lane traversal stack: the first LDS_DEPTH entries live in LDS,
// [entry][lane] so a wave's accesses hit distinct banks, the rest in
// private memory (only reached by deep, rare traversals). The private
// entries are a separate array owned by the caller, so that the stack
// pointer itself stays in a register (an aggregate holding the array would
// be placed in scratch memory whole, stack pointer included).
template <int LDS_DEPTH, int PRIV_DEPTH, int BLOCK>
struct Stack {
    int (*lds)[BLOCK];
    int* priv;
    int sp;
    int lane;
    static constexpr int capacity = LDS_DEPTH + PRIV_DEPTH;
    __device__ __forceinline__ bool push(int e) {
        if (sp < LDS_DEPTH) {
            lds[sp][lane] = e;
        } else {
            if (sp >= capacity) return false;
            priv[sp - LDS_DEPTH] = e;
        }
        sp++;
        return true;
    }
    __device__ __forceinline__ int pop() {
        --sp;
        // an LDS read on every pop (a clamped slot when the entry is
        // private), so the common case stays a ds_read and is not merged
        // with the private read into one generic-address load
        int v = lds[sp < LDS_DEPTH ? sp : 0][lane];
        asm volatile("" : "+v"(v));
        if (sp >= LDS_DEPTH) v = priv[sp - LDS_DEPTH];
        return v;
    }
};

struct TraceResult {
    int prim;
    float t, u, v;
    int best;   // index of the hit triangle in traversal order (-1 if none)
    bool overflow;
    int node_visits, tri_tests;
};

// Hit record of the accepted triangle re-evaluated once in FP64 with the
// reference's Moller-Trumbore formula (accel_sahbvh.cpp:58-75), whose Hit
// carries FP64 t and uv. The hit decision itself stays the FP32 watertight
// test; this only removes the FP32 conditioning error of t at grazing
// incidence (error ~ eps * edge / (t * cos)).
__device__ __forceinline__ void refine_hit(const TriV& tv, f3 o, f3 d, float& t, float& u, float& v) {
    const double p1x = tv.v0.x, p1y = tv.v0.y, p1z = tv.v0.z;
    const double e1x = double(tv.v1.x) - p1x, e1y = double(tv.v1.y) - p1y, e1z = double(tv.v1.z) - p1z;
    const double e2x = double(tv.v2.x) - p1x, e2y = double(tv.v2.y) - p1y, e2z = double(tv.v2.z) - p1z;
    const double dx = d.x, dy = d.y, dz = d.z;
    const double px = dy * e2z - dz * e2y, py = dz * e2x - dx * e2z, pz = dx * e2y - dy * e2x;
    const double tx = double(o.x) - p1x, ty = double(o.y) - p1y, tz = double(o.z) - p1z;
    const double qx = ty * e1z - tz * e1y, qy = tz * e1x - tx * e1z, qz = tx * e1y - ty * e1x;
    const double det = e1x * px + e1y * py + e1z * pz;
    if (det == 0.0) return;
    const double inv = 1.0 / det;
    t = float((e2x * qx + e2y * qy + e2z * qz) * inv);
    u = float((tx * px + ty * py + tz * pz) * inv);
    v = float((dx * qx + dy * qy + dz * qz) * inv);
}

// Sorting network helper: orders (k, e) pairs by k ascending.
__device__ __forceinline__ void cswap(float& ka, int& ea, float& kb, int& eb) {
    const bool s = kb < ka;
    const float tk = s ? kb : ka;
    const int te = s ? eb : ea;
    kb = s ? ka : kb;
    eb = s ? ea : eb;
    ka = tk;
    ea = te;
}

template <int W>
__device__ __forceinline__ void sort_children(float (&k)[W], int (&e)[W]) {
    if constexpr (W == 2) {
        cswap(k[0], e[0], k[1], e[1]);
    } else if constexpr (W == 4) {
        cswap(k[0], e[0], k[1], e[1]);
        cswap(k[2], e[2], k[3], e[3]);
        cswap(k[0], e[0], k[2], e[2]);
        cswap(k[1], e[1], k[3], e[3]);
        cswap(k[1], e[1], k[2], e[2]);
    } else {
        // Batcher odd-even merge sort network for 8 inputs (19 comparators).
        cswap(k[0], e[0], k[1], e[1]); cswap(k[2], e[2], k[3], e[3]);
        cswap(k[4], e[4], k[5], e[5]); cswap(k[6], e[6], k[7], e[7]);
        cswap(k[0], e[0], k[2], e[2]); cswap(k[1], e[1], k[3], e[3]);
        cswap(k[4], e[4], k[6], e[6]); cswap(k[5], e[5], k[7], e[7]);
        cswap(k[1], e[1], k[2], e[2]); cswap(k[5], e[5], k[6], e[6]);
        cswap(k[0], e[0], k[4], e[4]); cswap(k[1], e[1], k[5], e[5]);
        cswap(k[2], e[2], k[6], e[6]); cswap(k[3], e[3], k[7], e[7]);
        cswap(k[2], e[2], k[4], e[4]); cswap(k[3], e[3], k[5], e[5]);
        cswap(k[1], e[1], k[2], e[2]); cswap(k[3], e[3], k[4], e[4]);
        cswap(k[5], e[5], k[6], e[6]);
    }
}

// Persistent traversal: every lane owns one ray at a time; when at least
// `refill` lanes of a wave are idle (or the whole wave is), the idle lanes
// take new rays from the wave's pool, which the wave refills by chunks of
// chunk_rays (default LMGPU_RAY_CHUNK) from one global counter with one atomic (Aila & Laine,
// HPG 2009, "persistent threads"), so lanes do not sit idle behind a wave's
// slowest ray and the counter sees few atomics. refill = 64 keeps whole
// waves of consecutive rays together (best for coherent rays), smaller
// values keep lanes busy.
// Each loop iteration advances the lanes of one kind of step: a wide node or
// a leaf, chosen per wave by vote. Closest hit (ANY = false) keeps shrinking
// tmax; an accepted triangle is replaced by a later one at t == tmax, as the
// reference's closed [tmin, tmax] test does. `fetch(id, o, d, tmin, tmax)` loads ray
// `id` and returns the node its traversal starts from (0, or a root inside a
// multi-BVH arena);
// `emit(id, res)` consumes its result.
template <int W, bool ANY, bool COUNT, int LDS_DEPTH, int PRIV_DEPTH, int BLOCK, class Node, class Fetch, class Emit>
__device__ __forceinline__ void trace_persistent(const Node* __restrict__ nodes, const TriV* __restrict__ tris,
                                                 int64_t n, unsigned long long* next, int (*lds)[BLOCK],
                                                 Fetch fetch, Emit emit, int32_t* status,
                                                 unsigned long long* counters, int refill,
                                                 int chunk_rays = LMGPU_RAY_CHUNK) {
    const int lane = threadIdx.x & 63;
    int priv[PRIV_DEPTH];
    Stack<LDS_DEPTH, PRIV_DEPTH, BLOCK> st;
    st.lds = lds;
    st.priv = priv;
    st.sp = 0;
    st.lane = threadIdx.x;
    RayPre r{};
    TraceResult res{};
    int64_t rid = -1;       // ray of this lane, -1 idle
    bool exhausted = false;
    // wave-uniform: ray ids taken from `next` and not yet handed to a lane
    int64_t pool_next = 0, pool_end = 0;
    float tmin = 0.f;
    int cur = 0;
    unsigned long long nv = 0, nt = 0;
    while (true) {
        const bool idle = rid < 0 && !exhausted;
        const unsigned long long idle_mask = __ballot(idle);
        const unsigned long long busy_mask = __ballot(rid >= 0);
        const int nidle = __popcll(idle_mask);
        if (nidle >= refill || (busy_mask == 0 && nidle > 0)) {
            // idle lanes take, in lane order, the rest of the wave's pool and
            // then the start of a new chunk of LMGPU_RAY_CHUNK ids
            const int64_t avail = pool_end - pool_next;
            int64_t chunk = 0;
            if (avail < nidle) {
                const int leader = __ffsll(idle_mask) - 1;
                unsigned long long b = 0;
                if (lane == leader) b = atomicAdd(next, (unsigned long long)chunk_rays);
                b = (unsigned long long)__shfl((long long)b, leader);
                // wave-uniform from here on: scalar registers for the pool
                chunk = int64_t((uint64_t(uint32_t(__builtin_amdgcn_readfirstlane(uint32_t(b >> 32)))) << 32) |
                                uint32_t(__builtin_amdgcn_readfirstlane(uint32_t(b))));
            }
            if (idle) {
                const int rank = __popcll(idle_mask & ((1ull << lane) - 1ull));
                const int64_t id = rank < avail ? pool_next + rank : chunk + (rank - avail);
                if (id < n) {
                    rid = id;
                    f3 o, d;
                    float tmax;
                    cur = fetch(id, o, d, tmin, tmax);
                    r = ray_prepare(o, d);
                    res.prim = -1;
                    res.t = tmax;
                    res.u = res.v = 0.f;
                    res.best = -1;
                    res.overflow = false;
                    st.sp = 0;
                } else {
                    exhausted = true;
                }
            }
            if (avail < nidle) {
                pool_next = chunk + (nidle - avail);
                pool_end = chunk + chunk_rays;
            } else {
                pool_next += nidle;
            }
        }
        if (__ballot(rid >= 0) == 0) break;
        if (rid < 0) continue;
#if LMGPU_LEAF_VOTE_X4 > 0
        {
            // one kind of step per iteration for the whole wave, so inner-node
            // and leaf code do not both run, each for a part of the lanes:
            // leaves when enough lanes sit at a leaf, else inner nodes; the
            // other lanes wait one iteration
            const int nleaf = __popcll(__ballot(cur < 0));
            const int nnode = __popcll(__ballot(cur >= 0));
            const bool leaf_step = 4 * nleaf >= LMGPU_LEAF_VOTE_X4 * nnode;
            if (leaf_step != (cur < 0)) continue;
        }
#endif
        bool finished = false;
        if (cur >= 0) {
            if (COUNT) nv++;
            float key[W];
            int ent[W];
            child_keys<W>(nodes[cur], r, tmin, res.t, key, ent);
            sort_children<W>(key, ent);
            bool ok = true;
#pragma unroll
            for (int c = W - 1; c >= 1; c--)
                if (key[c] != __builtin_inff()) ok = ok && st.push(ent[c]);
            if (!ok) {
                res.overflow = true;
                finished = true;
            } else if (key[0] != __builtin_inff()) {
                cur = ent[0];
            } else if (st.sp == 0) {
                finished = true;
            } else {
                cur = st.pop();
            }
        } else {
            const int first = leaf_first(cur), cnt = leaf_count(cur);
            for (int k = 0; k < cnt; k++) {
                const TriV tv = tris[first + k];
                if (COUNT) nt++;
                float t, u, v;
                if (tri_intersect(r, tv.v0, tv.v1, tv.v2, tmin, res.t, t, u, v)) {
                    res.t = t;
                    res.u = u;
                    res.v = v;
                    res.prim = __float_as_int(tv.v0.w);
                    res.best = first + k;
                    if (ANY) break;
                }
            }
            if ((ANY && res.prim >= 0) || st.sp == 0) finished = true;
            else cur = st.pop();
        }
        if (finished) {
            if (res.overflow) atomicOr(status, 1);
            if (!ANY && res.best >= 0) refine_hit(tris[res.best], r.o, r.d, res.t, res.u, res.v);
            emit(rid, res);
            rid = -1;
        }
    }
    if (COUNT) {
        for (int off = 32; off > 0; off >>= 1) {
            nv += __shfl_xor(nv, off);
            nt += __shfl_xor(nt, off);
        }
        if (lane == 0) {
            atomicAdd(&counters[0], nv);
            atomicAdd(&counters[1], nt);
        }
    }
}

}  // namespace lmgpu
