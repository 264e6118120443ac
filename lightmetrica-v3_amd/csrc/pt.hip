// Wavefront renderer::pt for gfx950 (C ABI: lmgpu_scene_*, lmgpu_pt_render).
//
// Restates Renderer_PT::render (src/renderer/renderer_pt.cpp:80-315) in
// pixel-sampling mode with scheduler::spp::{sample,time}
// (src/scheduler.cpp:37-106) as a wavefront path tracer. One pass = one
// sample for every pixel; path p is pixel p. A pass is
//   generate  camera rays (camera_pinhole.cpp:128-132), path state reset
//   repeat max_verts-1 times:
//     extend   closest hit of the active extension rays (bvh.h)
//     shade    emitter hit + MIS (renderer_pt.cpp:236-271), env termination,
//              Russian roulette (:281-287), component sampling (:292), NEE
//              light sampling (:142-192) -> shadow queue, BSDF direction
//              sampling (:197-229) -> next extension queue
//     connect  any-hit visibility of the shadow queue; unoccluded
//              contributions added to the path radiance
//   accumulate per-pixel FP64 sum and sum of squares (Film_Bitmap is FP64,
//   src/film/film_bitmap.cpp:19-55)
// Queues are SoA arrays appended with one atomic per wave (ballot + popcount);
// counts live in device memory so the whole pass is captured once as a
// hipGraph and replayed per sample.
//
// Random numbers: Philox4x32-10 keyed by the seed with counter
// (pixel, sample, vertex, block); see common.h. All transport math is FP32.

#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstring>
#include <new>
#include <string>
#include <vector>

#include "bvh.h"
#include "guide.h"
#include "guide_internal.h"
#include "lmgpu_internal.h"
#include "shade.h"

namespace lmgpu {

// Triangle shading record, original (flattened) order: positions, vertex
// normals; p0.w = primitive index bits.
struct alignas(16) ShadeTri {
    float4 p0, p1, p2, n0, n1, n2;
};

struct alignas(16) DevMat {
    float4 a;   // rgb, type (as float)
    float4 b;   // ax, ay, ni, light index (as float, -1 = none)
};

struct alignas(16) DevLight {
    float4 Ke_invA;   // Ke rgb, 1/area
    int32_t cdf_off, count, pad0, pad1;
};

struct SceneDev {
    const WNode<4>* nodes4;
    const WNode<2>* nodes2;
    const WNode<8>* nodes8;
    const TriV* tris;
    const ShadeTri* stri;
    const DevMat* mats;
    const DevLight* lights;
    const float* light_cdf;    // per light: count+1 entries, c[0] = 0, normalised
    const int32_t* light_tri;  // per light: original triangle indices (face order)
    int num_area_lights;
    int num_lights;            // area + env
    int has_env;
    f3 env_Le;
    Camera cam;
    int w, h;
};

// Path state, SoA over pixels.
struct PathState {
    float4* thr;    // throughput xyz
    float4* L;      // radiance xyz
    float4* prev;   // shading normal of the previous vertex xyz, pdf_bsdf (projected SA) w
    int4* info;     // nv, comp, flags (bit0 prev samplable by NEE, bit1 prev degenerate), -
};

struct Queues {
    lmgpu_ray* ext[2];
    lmgpu_hit* hit;
    int32_t* ext_path[2];
    lmgpu_ray* sh;
    float4* sh_contrib;   // xyz contribution, w = path index bits
    int32_t* counts;      // [0],[1] extension queues, [2] shadow queue, [3] sample index, [4..] stats
};

struct PTConsts {
    int max_verts;
    int mode;
    int rr_depth;
    uint32_t seed_lo, seed_hi;
    int npix;
};

// Training record of one guided vertex of one path (slot v * npix + p for
// the v-th surface vertex of path p).
struct alignas(16) GRec {
    float4 a;   // position, mixture pdf of the sampled direction (solid angle; 0 = empty slot)
    float4 b;   // shading normal
    float4 c;   // direction toward the previous vertex
    float4 d;   // sampled direction
    float4 e;   // path radiance when the direction was sampled, plus this vertex's NEE
    float4 f;   // per channel: mixture pdf / throughput before the direction was sampled
};

struct GuideDev {
    const float* mix;   // decoded mixture per extension-queue entry of the bounce [npix][40]
    GRec* rec;          // [rec_depth][npix]
    int32_t* sh_rec;    // record slot of each shadow-queue entry (-1: none)
    float alpha;        // probability of sampling the guide at a non-specular vertex
    int rec_depth;      // surface vertices per path that record (0: no recording)
};

struct GBox {
    float lo[3], inv[3];
};

constexpr int kBlock = 256;
constexpr int kLds = 24;
constexpr int kPriv = 72;

__device__ __forceinline__ Mat load_mat(const DevMat* mats, int prim, int& light) {
    const DevMat d = mats[prim];
    Mat m;
    m.type = int(d.a.w);
    m.c = mk(d.a.x, d.a.y, d.a.z);
    m.ax = d.b.x;
    m.ay = d.b.y;
    m.ni = d.b.z;
    light = int(d.b.w);
    return m;
}

// Per-wave append: returns this lane's slot (lanes with want = false get -1).
__device__ __forceinline__ int wave_append(int32_t* counter, bool want) {
    const unsigned long long mask = __ballot(want);
    if (mask == 0) return -1;
    const int lane = __lane_id();
    const int leader = __ffsll(mask) - 1;
    int base = 0;
    if (lane == leader) base = atomicAdd(counter, __popcll(mask));
    base = __shfl(base, leader);
    const unsigned long long lower = mask & ((1ull << lane) - 1ull);
    return want ? base + __popcll(lower) : -1;
}

__global__ __launch_bounds__(kBlock) void k_generate(SceneDev sc, PathState ps, Queues q, PTConsts c) {
    const int p = blockIdx.x * kBlock + threadIdx.x;
    if (p >= c.npix) return;
    const uint32_t sample = uint32_t(q.counts[3]);
    const int x = p % sc.w, y = p / sc.w;
    const rand4 r = rng(uint32_t(p), sample, 0u, 0u, (uint64_t(c.seed_hi) << 32) | c.seed_lo);
    const float dx = 1.f / sc.w, dy = 1.f / sc.h;
    const f3 d = primary_dir(sc.cam, dx * x + dx * r.u[2], dy * y + dy * r.u[3]);
    ps.thr[p] = make_float4(1.f, 1.f, 1.f, 0.f);
    ps.L[p] = make_float4(0.f, 0.f, 0.f, 0.f);
    ps.prev[p] = make_float4(0.f, 0.f, 0.f, 0.f);
    ps.info[p] = make_int4(1, 0, 2, 0);   // nv = 1, comp 0, prev degenerate (camera), not NEE-samplable
    lmgpu_ray ray;
    ray.org[0] = sc.cam.pos.x; ray.org[1] = sc.cam.pos.y; ray.org[2] = sc.cam.pos.z;
    ray.tmin = kEps;
    ray.dir[0] = d.x; ray.dir[1] = d.y; ray.dir[2] = d.z;
    ray.tmax = kInf;
    q.ext[0][p] = ray;
    q.ext_path[0][p] = p;
    if (p == 0) {
        q.counts[0] = c.npix;
        atomicAdd(reinterpret_cast<unsigned long long*>(q.counts + 8), (unsigned long long)c.npix);
    }
}

__global__ void k_reset_counts(int32_t* counts, int out_q) {
    if (threadIdx.x == 0) {
        counts[out_q] = 0;
        counts[2] = 0;
    }
}

template <int W>
__device__ __forceinline__ const WNode<W>* nodes_of(const SceneDev& sc);
template <>
__device__ __forceinline__ const WNode<2>* nodes_of<2>(const SceneDev& sc) { return sc.nodes2; }
template <>
__device__ __forceinline__ const WNode<4>* nodes_of<4>(const SceneDev& sc) { return sc.nodes4; }
template <>
__device__ __forceinline__ const WNode<8>* nodes_of<8>(const SceneDev& sc) { return sc.nodes8; }

template <int W>
__global__ __launch_bounds__(kBlock) void k_extend(SceneDev sc, Queues q, int in_q, unsigned long long* next, int refill) {
    __shared__ int lds[kLds][kBlock];
    const int n = q.counts[in_q];
    const lmgpu_ray* rays = q.ext[in_q];
    auto fetch = [&](int64_t id, f3& o, f3& d, float& tmin, float& tmax) {
        const float4* rp = reinterpret_cast<const float4*>(rays + id);
        const float4 a = rp[0], b = rp[1];
        o = mk(a.x, a.y, a.z);
        d = mk(b.x, b.y, b.z);
        tmin = a.w;
        tmax = b.w;
    };
    auto emit = [&](int64_t id, const TraceResult& r) {
        reinterpret_cast<float4*>(q.hit)[id] = make_float4(__int_as_float(r.overflow ? -2 : r.prim), r.t, r.u, r.v);
    };
    trace_persistent<W, false, false, kLds, kPriv, kBlock>(nodes_of<W>(sc), sc.tris, n, next, lds, fetch, emit,
                                                           q.counts + 5, nullptr, refill);
    if (blockIdx.x == 0 && threadIdx.x == 0)
        atomicAdd(reinterpret_cast<unsigned long long*>(q.counts + 10), (unsigned long long)n);
}

// Unoccluded shadow rays add their contribution to the path radiance and,
// when the vertex records guide training data, to the record's radiance
// baseline (the NEE of a vertex is not part of its sampled direction's
// incident radiance).
template <int W>
__global__ __launch_bounds__(kBlock) void k_connect(SceneDev sc, PathState ps, Queues q, unsigned long long* next,
                                                   int refill, GRec* __restrict__ rec, const int32_t* __restrict__ sh_rec) {
    __shared__ int lds[kLds][kBlock];
    const int n = q.counts[2];
    auto fetch = [&](int64_t id, f3& o, f3& d, float& tmin, float& tmax) {
        const float4* rp = reinterpret_cast<const float4*>(q.sh + id);
        const float4 a = rp[0], b = rp[1];
        o = mk(a.x, a.y, a.z);
        d = mk(b.x, b.y, b.z);
        tmin = a.w;
        tmax = b.w;
    };
    auto emit = [&](int64_t id, const TraceResult& r) {
        if (r.prim < 0 && !r.overflow) {
            const float4 c = q.sh_contrib[id];
            const int p = __float_as_int(c.w);
            float4 L = ps.L[p];
            L.x += c.x;
            L.y += c.y;
            L.z += c.z;
            ps.L[p] = L;
            if (sh_rec) {
                const int slot = sh_rec[id];
                if (slot >= 0) {
                    float4 e = rec[slot].e;
                    e.x += c.x;
                    e.y += c.y;
                    e.z += c.z;
                    rec[slot].e = e;
                }
            }
        }
    };
    trace_persistent<W, true, false, kLds, kPriv, kBlock>(nodes_of<W>(sc), sc.tris, n, next, lds, fetch, emit,
                                                          q.counts + 5, nullptr, refill);
    if (blockIdx.x == 0 && threadIdx.x == 0)
        atomicAdd(reinterpret_cast<unsigned long long*>(q.counts + 12), (unsigned long long)n);
}

__global__ void k_bounce_begin(unsigned long long* next) {
    if (threadIdx.x < 2) next[threadIdx.x] = 0ull;
}

// pdf of light endpoint `light` seen from the previous vertex (prev normal
// np, degenerate flag) along d (unit, from the previous vertex to the light),
// including the uniform selection probability (path::pdf_direct, path.h:563-583).
__device__ __forceinline__ float pdf_light_area(const SceneDev& sc, int light, f3 np, bool prev_degen, f3 d,
                                                float dist2, f3 nL) {
    const float psel = 1.f / sc.num_lights;
    const float cL = dot(-d, nL);
    if (cL <= 0.f) return 0.f;
    const float c1 = prev_degen ? 1.f : fabsf(dot(np, d));
    const float G = c1 * fabsf(cL) / dist2;
    if (G == 0.f) return 0.f;
    return psel * sc.lights[light].Ke_invA.w / G;
}

__device__ __forceinline__ float pdf_light_env(const SceneDev& sc, f3 np, bool prev_degen, f3 d) {
    const float psel = 1.f / sc.num_lights;
    const float pSA = 1.f / (4.f * kPi);
    if (prev_degen) return psel * pSA;
    const float J = fabsf(dot(np, d));
    return J == 0.f ? 0.f : psel * pSA / J;
}

// Guided vertices (GUIDED, diffuse or glossy) sample the next direction from
// the one-sample MIS mixture alpha * guide + (1 - alpha) * BSDF, whose
// density (solid angle) p_mix enters both the throughput f |cos| / p_mix and,
// as p_mix / |cos|, every MIS weight against light sampling.
template <bool GUIDED>
__global__ __launch_bounds__(kBlock) void k_shade(SceneDev sc, PathState ps, Queues q, PTConsts c, int in_q, GuideDev gd) {
    const int n = q.counts[in_q];
    const int out_q = in_q ^ 1;
    const uint32_t sample = uint32_t(q.counts[3]);
    const uint64_t seed = (uint64_t(c.seed_hi) << 32) | c.seed_lo;
    const int stride = gridDim.x * kBlock;
    // every lane of a wave runs the same number of iterations so the
    // wave-level appends see the whole wave
    const int n_iter = (n + stride - 1) / stride;
    for (int it = 0; it < n_iter; it++) {
        const int i = it * stride + blockIdx.x * kBlock + threadIdx.x;
        bool active = i < n;
        bool emit_ext = false, emit_sh = false;
        lmgpu_ray ext_ray, sh_ray;
        float4 sh_c = make_float4(0, 0, 0, 0);
        int p = 0, sh_slot = -1;
        if (active) {
            p = q.ext_path[in_q][i];
            const float4* rp = reinterpret_cast<const float4*>(q.ext[in_q] + i);
            const float4 ra = rp[0], rb = rp[1];
            const f3 o = mk(ra.x, ra.y, ra.z), wo = mk(rb.x, rb.y, rb.z);
            const float4 h = reinterpret_cast<const float4*>(q.hit)[i];
            const int tri = __float_as_int(h.x);
            float4 thr4 = ps.thr[p];
            float4 L4 = ps.L[p];
            const float4 pv = ps.prev[p];
            int4 info = ps.info[p];
            f3 thr = mk(thr4.x, thr4.y, thr4.z);
            f3 L = mk(L4.x, L4.y, L4.z);
            const f3 np = mk(pv.x, pv.y, pv.z);
            const bool prev_nee = info.z & 1;
            const bool prev_degen = info.z & 2;
            const bool direct_ok = c.mode == LMGPU_SAMPLING_NEE ? !prev_nee : true;
            const bool use_mis = c.mode != LMGPU_SAMPLING_NAIVE && prev_nee;
            if (tri == -2) atomicOr(q.counts + 5, 1);
            if (tri < 0) {
                // miss: environment light endpoint (Scene_::intersect, scene.cpp:321-332)
                if (sc.has_env && direct_ok) {
                    float mis = 1.f;
                    if (use_mis) mis = balance_heuristic(pv.w, pdf_light_env(sc, np, prev_degen, wo));
                    L = L + thr * sc.env_Le * mis;
                }
            } else {
                const ShadeTri st = sc.stri[tri];
                const float b1 = h.z, b2 = h.w, b0 = 1.f - b1 - b2;
                const f3 P0 = mk(st.p0.x, st.p0.y, st.p0.z), P1 = mk(st.p1.x, st.p1.y, st.p1.z),
                         P2 = mk(st.p2.x, st.p2.y, st.p2.z);
                const f3 pos = P0 * b0 + P1 * b1 + P2 * b2;
                const f3 nsh = normalize(mk(st.n0.x, st.n0.y, st.n0.z) * b0 + mk(st.n1.x, st.n1.y, st.n1.z) * b1 +
                                         mk(st.n2.x, st.n2.y, st.n2.z) * b2);
                const f3 gn = normalize(cross(P1 - P0, P2 - P0));
                const Geom g = make_on_surface(pos, nsh, gn);
                const int prim = __float_as_int(st.p0.w);
                int light;
                const Mat m = load_mat(sc.mats, prim, light);
                // emitter hit (renderer_pt.cpp:246-271)
                if (light >= 0 && direct_ok) {
                    const f3 woL = -wo;
                    if (dot(woL, g.n) > 0.f) {
                        const DevLight dl = sc.lights[light];
                        float mis = 1.f;
                        if (use_mis) {
                            const f3 dd = pos - o;
                            const float dist2 = dot(dd, dd);
                            mis = balance_heuristic(pv.w, pdf_light_area(sc, light, np, prev_degen, wo, dist2, g.n));
                        }
                        L = L + thr * mk(dl.Ke_invA.x, dl.Ke_invA.y, dl.Ke_invA.z) * mis;
                    }
                }
                // Russian roulette and component sampling at this vertex
                const int nv = info.x;
                const rand4 r0 = rng(uint32_t(p), sample, uint32_t(nv), 0u, seed);
                bool alive = true;
                if (nv > c.rr_depth) {
                    const float qq = fmaxf(.2f, 1.f - maxc(thr));
                    if (r0.u[0] < qq) alive = false;
                    else thr = thr / (1.f - qq);
                }
                const f3 wi = -wo;
                int comp = 0;
                if (alive) {
                    float cw;
                    comp = mat_sample_component(m, g, wi, r0.u[1], cw);
                    thr = thr * cw;
                }
                const int nv2 = nv + 1;
                if (alive && nv2 < c.max_verts) {
                    const bool spec = is_specular(m);
                    const bool nee = c.mode != LMGPU_SAMPLING_NAIVE && !spec;
                    // guide mixture of this queue entry (guided vertices only)
                    const float* gmix = (GUIDED && !spec) ? gd.mix + int64_t(i) * kOut : nullptr;
                    const int slot = (gmix && nv - 1 < gd.rec_depth) ? (nv - 1) * c.npix + p : -1;
                    if (nee) {
                        // path::sample_direct (path.h:509-528)
                        const rand4 r1 = rng(uint32_t(p), sample, uint32_t(nv), 1u, seed);
                        const int li = min(max(int(r1.u[0] * sc.num_lights), 0), sc.num_lights - 1);
                        const float psel = 1.f / sc.num_lights;
                        f3 dirL;   // from the shading point toward the light
                        float tmaxL = 0.f;
                        f3 weight = mk(0, 0, 0);
                        float p_light = 0.f;
                        bool ok = false;
                        if (li < sc.num_area_lights) {
                            const DevLight dl = sc.lights[li];
                            // Dist::sample: upper_bound on the CDF
                            const float* cdf = sc.light_cdf + dl.cdf_off;
                            int lo = 0, hi = dl.count + 1;
                            const float u = r1.u[1];
                            while (lo < hi) {
                                const int mid = (lo + hi) >> 1;
                                if (cdf[mid] <= u) lo = mid + 1;
                                else hi = mid;
                            }
                            const int k = min(max(lo - 1, 0), dl.count - 1);
                            const int lt = sc.light_tri[dl.cdf_off - li + k];
                            const ShadeTri ls = sc.stri[lt];
                            const f3 A = mk(ls.p0.x, ls.p0.y, ls.p0.z), B = mk(ls.p1.x, ls.p1.y, ls.p1.z),
                                     C = mk(ls.p2.x, ls.p2.y, ls.p2.z);
                            const float s = sqrtf(fmaxf(0.f, r1.u[2]));
                            const float bu = 1.f - s, bv = r1.u[3] * s;
                            const f3 pL = A * (1.f - bu - bv) + B * bu + C * bv;
                            const f3 nL = normalize(cross(B - A, C - A));
                            const f3 wL = normalize(pos - pL);   // from the light to the point
                            if (dot(wL, nL) > 0.f) {
                                const f3 dd = pL - pos;
                                const float d2 = dot(dd, dd);
                                const float G = fabsf(dot(g.n, wL)) * fabsf(dot(nL, -wL)) / d2;
                                if (G != 0.f) {
                                    const float pA = dl.Ke_invA.w / G;
                                    weight = mk(dl.Ke_invA.x, dl.Ke_invA.y, dl.Ke_invA.z) / pA / psel;
                                    p_light = psel * pA;
                                    dirL = normalize(pL - pos);
                                    tmaxL = sqrtf(d2) * (1.f - kEps);
                                    ok = true;
                                }
                            }
                        } else {
                            const f3 wL = sample_uniform_sphere(r1.u[2], r1.u[3]);
                            const float J = fabsf(dot(g.n, -wL));
                            if (J != 0.f) {
                                const float pL = (1.f / (4.f * kPi)) / J;
                                weight = sc.env_Le / pL / psel;
                                p_light = psel * pL;
                                dirL = -wL;
                                tmaxL = kInf - 1.f;
                                ok = true;
                            }
                        }
                        if (ok) {
                            const f3 fs = mat_eval(m, g, wi, dirL) * sn_correction(g, wi, dirL);
                            if (!is_zero(fs)) {
                                float mis = 1.f;
                                if (c.mode == LMGPU_SAMPLING_MIS) {
                                    float pdir = mat_pdf(m, g, wi, dirL);
                                    if (GUIDED && gmix) {
                                        float pg = mixture_pdf(gmix, dirL);
                                        if (!(pg >= 0.f && pg < 3e38f)) pg = 0.f;
                                        pdir = gd.alpha * pg / fabsf(dot(dirL, g.n)) + (1.f - gd.alpha) * pdir;
                                    }
                                    mis = balance_heuristic(p_light, pdir);
                                }
                                const f3 C = thr * fs * weight * mis;
                                if (!is_zero(C)) {
                                    emit_sh = true;
                                    sh_slot = slot;
                                    sh_ray.org[0] = pos.x; sh_ray.org[1] = pos.y; sh_ray.org[2] = pos.z;
                                    sh_ray.tmin = kEps;
                                    sh_ray.dir[0] = dirL.x; sh_ray.dir[1] = dirL.y; sh_ray.dir[2] = dirL.z;
                                    sh_ray.tmax = tmaxL;
                                    sh_c = make_float4(C.x, C.y, C.z, __int_as_float(p));
                                }
                            }
                        }
                    }
                    // path::sample_direction (path.h:410-419)
                    f3 wo2, sw;
                    bool sampled;
                    float pdir = 0.f, pmix = 0.f;
                    if (GUIDED && gmix) {
                        const rand4 r2 = rng(uint32_t(p), sample, uint32_t(nv), 2u, seed);
                        if (r2.u[0] < gd.alpha) {
                            wo2 = mixture_sample(gmix, r2.u[1], r2.u[2], r2.u[3]);
                            sampled = true;
                        } else {
                            sampled = mat_sample_direction(m, g, wi, comp, r0.u[2], r0.u[3], wo2, sw);
                        }
                        sw = mk(0, 0, 0);
                        if (sampled) {
                            const float cs = fabsf(dot(wo2, g.n));
                            float pg = mixture_pdf(gmix, wo2);
                            if (!(pg >= 0.f && pg < 3e38f)) pg = 0.f;
                            pmix = gd.alpha * pg + (1.f - gd.alpha) * mat_pdf(m, g, wi, wo2) * cs;
                            if (pmix > 0.f && cs > 0.f) {
                                sw = mat_eval(m, g, wi, wo2) * (cs / pmix);
                                pdir = pmix / cs;
                            }
                        }
                    } else {
                        sampled = mat_sample_direction(m, g, wi, comp, r0.u[2], r0.u[3], wo2, sw);
                        if (sampled) pdir = mat_pdf(m, g, wi, wo2);
                    }
                    if (sampled) {
                        sw = sw * sn_correction(g, wi, wo2);
                        const f3 thr_before = thr;
                        thr = thr * sw;
                        if (!is_zero(thr)) {
                            emit_ext = true;
                            ext_ray.org[0] = pos.x; ext_ray.org[1] = pos.y; ext_ray.org[2] = pos.z;
                            ext_ray.tmin = kEps;
                            ext_ray.dir[0] = wo2.x; ext_ray.dir[1] = wo2.y; ext_ray.dir[2] = wo2.z;
                            ext_ray.tmax = kInf;
                            ps.prev[p] = make_float4(g.n.x, g.n.y, g.n.z, pdir);
                            info = make_int4(nv2, comp, nee ? 1 : 0, 0);
                            ps.info[p] = info;
                            ps.thr[p] = make_float4(thr.x, thr.y, thr.z, 0.f);
                            if (GUIDED && slot >= 0) {
                                // incident radiance along wo2 is (L_final - e) / throughput after this
                                // vertex; the record keeps p_mix / throughput before it, so that
                                // (L_final - e) * f is the integrand f |cos| L_i estimate
                                GRec r;
                                r.a = make_float4(pos.x, pos.y, pos.z, pmix);
                                r.b = make_float4(g.n.x, g.n.y, g.n.z, 0.f);
                                r.c = make_float4(wi.x, wi.y, wi.z, 0.f);
                                r.d = make_float4(wo2.x, wo2.y, wo2.z, 0.f);
                                r.e = make_float4(L.x, L.y, L.z, 0.f);
                                r.f = make_float4(thr_before.x > 0.f ? pmix / thr_before.x : 0.f,
                                                  thr_before.y > 0.f ? pmix / thr_before.y : 0.f,
                                                  thr_before.z > 0.f ? pmix / thr_before.z : 0.f, 0.f);
                                gd.rec[slot] = r;
                            }
                        }
                    }
                }
            }
            ps.L[p] = make_float4(L.x, L.y, L.z, 0.f);
        }
        const int se = wave_append(q.counts + out_q, emit_ext);
        if (emit_ext) {
            q.ext[out_q][se] = ext_ray;
            q.ext_path[out_q][se] = p;
        }
        const int ss = wave_append(q.counts + 2, emit_sh);
        if (emit_sh) {
            q.sh[ss] = sh_ray;
            q.sh_contrib[ss] = sh_c;
            if (GUIDED) gd.sh_rec[ss] = sh_slot;
        }
    }
}

// Guide inference over the extension queue of a bounce: the query of entry
// i is its hit point, shading normal and the direction back along the ray;
// the decoded mixture goes to mix[i]. Entries without a hit get a zero query.
__global__ __launch_bounds__(256) void k_pt_guide_infer(SceneDev sc, Queues q, int in_q, const uint16_t* __restrict__ Wbf,
                                                        const float* __restrict__ P, GBox box, float* __restrict__ mix) {
    __shared__ __attribute__((aligned(16))) unsigned char lds[kWeightLds + 4 * kWaveLds];
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    uint16_t* wl = reinterpret_cast<uint16_t*>(lds);
    stage_weights(Wbf, wl, threadIdx.x, 256);
    uint16_t* lx = reinterpret_cast<uint16_t*>(lds + kWeightLds + wave * kWaveLds);
    uint16_t* lh = lx + 32 * kPitchX;
    float* ly = reinterpret_cast<float*>(lx);
    const int n = q.counts[in_q];
    for (int base = blockIdx.x * 128; base < n; base += gridDim.x * 128) {
        const int i0 = base + wave * 32;
        if (lane < 32) {
            const int i = i0 + lane;
            uint16_t* row = lx + lane * kPitchX;
            const float4 h = i < n ? reinterpret_cast<const float4*>(q.hit)[i] : make_float4(__int_as_float(-1), 0, 0, 0);
            const int tri = __float_as_int(h.x);
            if (tri >= 0) {
                const ShadeTri st = sc.stri[tri];
                const float b1 = h.z, b2 = h.w, b0 = 1.f - b1 - b2;
                const f3 pos = mk(st.p0.x, st.p0.y, st.p0.z) * b0 + mk(st.p1.x, st.p1.y, st.p1.z) * b1 +
                               mk(st.p2.x, st.p2.y, st.p2.z) * b2;
                const f3 nsh = normalize(mk(st.n0.x, st.n0.y, st.n0.z) * b0 + mk(st.n1.x, st.n1.y, st.n1.z) * b1 +
                                         mk(st.n2.x, st.n2.y, st.n2.z) * b2);
                const lmgpu_ray& r = q.ext[in_q][i];
                encode_input(pos, nsh, mk(-r.dir[0], -r.dir[1], -r.dir[2]), box.lo, box.inv, row);
            } else {
                for (int k = 0; k < kIn; k++) row[k] = 0;
            }
        }
        __syncthreads();
        mlp_tile_forward(wl, P, lx, lh, ly, lane, [](int, int, int, uint16_t) {});
        __syncthreads();
        if (lane < 32) {
            const int i = i0 + lane;
            if (i < n) {
                Mixture m;
                decode_mixture(ly + lane * kPitchY, m);
                float* o = mix + int64_t(i) * kOut;
                for (int k = 0; k < kLobes; k++) {
                    o[k] = m.pi[k];
                    o[kLobes + k] = m.kappa[k];
                    o[2 * kLobes + 3 * k] = m.mu[k].x;
                    o[2 * kLobes + 3 * k + 1] = m.mu[k].y;
                    o[2 * kLobes + 3 * k + 2] = m.mu[k].z;
                }
            }
        }
        __syncthreads();
    }
}

// After a pass: turns the filled record slots into training records
// (pos, nrm, wo, wi, p_mix, f |cos| L_i estimate) appended to train[cap][16]
// when their target weight is positive, and empties the slots. Slots are
// visited in the order j * mul mod total (mul coprime to total) so that a
// capped batch spreads over the film and the path depths.
__global__ __launch_bounds__(kBlock) void k_guide_collect(PathState ps, GuideDev gd, int npix, int64_t total,
                                                          int64_t mul, float* __restrict__ train, int cap,
                                                          int32_t* __restrict__ count,
                                                          unsigned long long* __restrict__ recorded) {
    const int64_t stride = int64_t(gridDim.x) * kBlock;
    const int64_t n_iter = (total + stride - 1) / stride;
    for (int64_t it = 0; it < n_iter; it++) {
        const int64_t j = it * stride + int64_t(blockIdx.x) * kBlock + threadIdx.x;
        bool keep = false;
        float out[16];
        if (j < total) {
            const int64_t slot = (j * mul) % total;
            GRec& r = gd.rec[slot];
            const float4 a = r.a;
            if (a.w > 0.f) {
                const int p = int(slot % npix);
                const float4 Lf = ps.L[p];
                const float4 e = r.e, f = r.f;
                const float lx = fmaxf(0.f, Lf.x - e.x) * f.x, ly = fmaxf(0.f, Lf.y - e.y) * f.y,
                            lz = fmaxf(0.f, Lf.z - e.z) * f.z;
                const float wgt = (0.2126f * lx + 0.7152f * ly + 0.0722f * lz) / a.w;
                r.a.w = 0.f;
                if (wgt > 0.f && wgt < 3e38f) {
                    const float4 b = r.b, c = r.c, d = r.d;
                    keep = true;
                    out[0] = a.x; out[1] = a.y; out[2] = a.z;
                    out[3] = b.x; out[4] = b.y; out[5] = b.z;
                    out[6] = c.x; out[7] = c.y; out[8] = c.z;
                    out[9] = d.x; out[10] = d.y; out[11] = d.z;
                    out[12] = a.w; out[13] = lx; out[14] = ly; out[15] = lz;
                }
            }
        }
        const unsigned long long mask = __ballot(keep);
        if (mask && __lane_id() == __ffsll(mask) - 1) atomicAdd(recorded, (unsigned long long)__popcll(mask));
        const int k = wave_append(count, keep);
        if (keep && k < cap) {
            float4* o = reinterpret_cast<float4*>(train + int64_t(k) * 16);
            o[0] = make_float4(out[0], out[1], out[2], out[3]);
            o[1] = make_float4(out[4], out[5], out[6], out[7]);
            o[2] = make_float4(out[8], out[9], out[10], out[11]);
            o[3] = make_float4(out[12], out[13], out[14], out[15]);
        }
    }
}

// renderer::raycast (src/renderer/renderer_raycast.cpp:62-97): one primary
// ray through each pixel centre, generated in the traversal's fetch; the
// emit shades it: background on a miss (the env radiance when the scene has
// an envconst light), |shading normal|, or reflectance times
// 0.2 + 0.8 |n . -d| (Kd / Ks; 0 for mirror and glass), written with
// set_pixel semantics into film[h][w][3].
struct RaycastConsts {
    f3 bg, color;
    int use_color, constant, normal;
};

template <int W>
__global__ __launch_bounds__(kBlock) void k_raycast(SceneDev sc, unsigned long long* next, int refill, RaycastConsts rc,
                                                    double* __restrict__ film, int32_t* status) {
    __shared__ int lds[kLds][kBlock];
    const int n = sc.w * sc.h;
    auto dir_of = [&](int64_t id) {
        const int x = int(id % sc.w), y = int(id / sc.w);
        return primary_dir(sc.cam, (x + .5f) / sc.w, (y + .5f) / sc.h);
    };
    auto fetch = [&](int64_t id, f3& o, f3& d, float& tmin, float& tmax) {
        o = sc.cam.pos;
        d = dir_of(id);
        tmin = kEps;
        tmax = kInf;
    };
    auto emit = [&](int64_t id, const TraceResult& r) {
        f3 C = rc.bg;
        if (r.prim >= 0) {
            const ShadeTri st = sc.stri[r.prim];
            const float b1 = r.u, b2 = r.v, b0 = 1.f - b1 - b2;
            const f3 nsh = normalize(mk(st.n0.x, st.n0.y, st.n0.z) * b0 + mk(st.n1.x, st.n1.y, st.n1.z) * b1 +
                                     mk(st.n2.x, st.n2.y, st.n2.z) * b2);
            if (rc.normal) {
                C = mk(fabsf(nsh.x), fabsf(nsh.y), fabsf(nsh.z));
            } else {
                int light;
                const Mat m = load_mat(sc.mats, __float_as_int(st.p0.w), light);
                const bool refl = m.type == LMGPU_MAT_DIFFUSE || m.type == LMGPU_MAT_GLOSSY;
                C = rc.use_color ? rc.color : (refl ? m.c : mk(0, 0, 0));
                if (!rc.constant) C = C * (.2f + .8f * fabsf(dot(nsh, -dir_of(id))));
            }
        } else if (sc.has_env) {
            // the env endpoint of Scene_::intersect; it has no normal to visualise
            C = rc.normal ? mk(0, 0, 0) : sc.env_Le;
        }
        film[3 * id + 0] = C.x;
        film[3 * id + 1] = C.y;
        film[3 * id + 2] = C.z;
    };
    trace_persistent<W, false, false, kLds, kPriv, kBlock>(nodes_of<W>(sc), sc.tris, n, next, lds, fetch, emit, status,
                                                           nullptr, refill);
}

__global__ void k_zero_u64(unsigned long long* p) {
    if (threadIdx.x == 0) *p = 0ull;
}

__global__ void k_zero_i32(int32_t* p, int n) {
    if (int(threadIdx.x) < n) p[threadIdx.x] = 0;
}

__global__ __launch_bounds__(kBlock) void k_accumulate(PathState ps, double* __restrict__ sum,
                                                       double* __restrict__ sumsq, int npix) {
    const int p = blockIdx.x * kBlock + threadIdx.x;
    if (p >= npix) return;
    const float4 L = ps.L[p];
    sum[3 * p + 0] += double(L.x);
    sum[3 * p + 1] += double(L.y);
    sum[3 * p + 2] += double(L.z);
    if (sumsq) {
        sumsq[3 * p + 0] += double(L.x) * double(L.x);
        sumsq[3 * p + 1] += double(L.y) * double(L.y);
        sumsq[3 * p + 2] += double(L.z) * double(L.z);
    }
}

__global__ void k_next_sample(int32_t* counts) {
    if (threadIdx.x == 0) counts[3] += 1;
}

__global__ void k_set_sample(int32_t* counts, int s) {
    if (threadIdx.x == 0) {
        counts[3] = s;
        counts[5] = 0;
    }
}

__global__ void k_zero_f64(double* p, int64_t n) {
    const int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i < n) p[i] = 0.0;
}

// ---------------------------------------------------------------------------
// host
// ---------------------------------------------------------------------------

// Guided rendering of one render call (opts resolved, workspace sized).
struct GuidedPass {
    Guide* g = nullptr;
    uint64_t uid = 0;   // guide identity and buffer generation (graph key)
    int gen = 0;
    float alpha = 0.5f, lr = 1e-3f;
    int batch = 0, steps = 0, rec_depth = 0;
    int64_t mul = 1;   // slot visiting order of k_guide_collect
};

struct Scene {
    Accel accel;
    DevBuf<ShadeTri> stri;
    DevBuf<DevMat> mats;
    DevBuf<DevLight> lights;
    DevBuf<float> light_cdf;
    DevBuf<int32_t> light_tri;
    SceneDev dev{};
    int w = 0, h = 0;
    // render workspace
    int ws_npix = 0;
    DevBuf<float4> thr, L, prev, shc;
    DevBuf<int4> info;
    DevBuf<lmgpu_ray> ext0, ext1, sh;
    DevBuf<lmgpu_hit> hit;
    DevBuf<int32_t> path0, path1, counts;
    DevBuf<unsigned long long> next;   // per-bounce ray counters of the persistent trace kernels
    DevBuf<double> fsum, fsq;   // device film when the caller passes host pointers
    // guided rendering workspace
    DevBuf<float> gmix, gtrain;
    DevBuf<GRec> grec;
    DevBuf<int32_t> gshrec, gcount;
    int grec_slots = 0;
    hipGraphExec_t graph = nullptr;
    int refill_primary = 64, refill_secondary = 16;
    int graph_key[4] = {0, 0, 0, 0};
    GuidedPass graph_guide;
    uint64_t graph_seed = 0;
    int graph_sumsq = -1;
    double* graph_sum_ptr = nullptr;
    double* graph_sq_ptr = nullptr;

    ~Scene() {
        if (graph) (void)hipGraphExecDestroy(graph);
    }

    int create(const lmgpu_scene_desc* d, const lmgpu_build_opts* opts);
    int ensure_workspace(int npix);
    int ensure_guided(int npix, const GuidedPass& gp);
    int enqueue_pass(const PTConsts& c, double* sum, double* sumsq, const GuidedPass* gp, hipStream_t s);
    int render(const lmgpu_pt_params* p, double* sum, double* sumsq, bool on_device, const GuidedPass* gp,
               lmgpu_pt_stats* st, hipStream_t s);
    int raycast(const lmgpu_raycast_params* p, double* film, bool on_device, hipStream_t s);
};

static inline float f_or(const float* p, int i) { return p ? p[i] : 0.f; }

int Scene::create(const lmgpu_scene_desc* d, const lmgpu_build_opts* opts) {
    if (!d) return fail(LMGPU_ERR_INVALID_ARGUMENT, "scene desc is NULL");
    if (d->num_tris <= 0 || !d->verts || !d->normals || !d->tri_prim)
        return fail(LMGPU_ERR_INVALID_ARGUMENT, "scene needs triangles, normals and primitive ids");
    if (d->num_prims <= 0 || !d->mat_type || !d->mat_param)
        return fail(LMGPU_ERR_INVALID_ARGUMENT, "scene needs materials");
    if (d->w <= 0 || d->h <= 0) return fail(LMGPU_ERR_INVALID_ARGUMENT, "film size must be positive");
    if (int64_t(d->w) * d->h > (int64_t(1) << 28)) return fail(LMGPU_ERR_INVALID_ARGUMENT, "film too large");
    const int64_t N = d->num_tris;
    const int P = d->num_prims;
    for (int64_t t = 0; t < N; t++)
        if (d->tri_prim[t] < 0 || d->tri_prim[t] >= P)
            return fail(LMGPU_ERR_INVALID_ARGUMENT, "tri_prim out of range");
    for (int i = 0; i < P; i++)
        if (d->mat_type[i] < 0 || d->mat_type[i] > 3) return fail(LMGPU_ERR_INVALID_ARGUMENT, "bad material type");
    LMGPU_TRY(accel.build(d->verts, N, false, opts, nullptr));
    refill_secondary = accel.refill;
    w = d->w;
    h = d->h;
    // shading triangles
    std::vector<ShadeTri> st(N);
    for (int64_t t = 0; t < N; t++) {
        const float* v = d->verts + 9 * t;
        const float* n = d->normals + 9 * t;
        int32_t prim = d->tri_prim[t];
        float pb;
        std::memcpy(&pb, &prim, 4);
        st[t].p0 = make_float4(v[0], v[1], v[2], pb);
        st[t].p1 = make_float4(v[3], v[4], v[5], 0.f);
        st[t].p2 = make_float4(v[6], v[7], v[8], 0.f);
        st[t].n0 = make_float4(n[0], n[1], n[2], 0.f);
        st[t].n1 = make_float4(n[3], n[4], n[5], 0.f);
        st[t].n2 = make_float4(n[6], n[7], n[8], 0.f);
    }
    LMGPU_TRY(stri.alloc(N));
    LMGPU_HIP(hipMemcpy(stri.p, st.data(), sizeof(ShadeTri) * N, hipMemcpyHostToDevice));
    // lights (light_area.cpp:62-74: area CDF over the light mesh's faces)
    const int L = d->num_area_lights;
    if (L < 0 || (L > 0 && !d->light_prims)) return fail(LMGPU_ERR_INVALID_ARGUMENT, "bad light list");
    std::vector<int> prim_light(P, -1);
    std::vector<DevLight> hl(std::max(L, 1));
    std::vector<float> cdf;
    std::vector<int32_t> ltri;
    std::vector<std::vector<int32_t>> per(L);
    for (int l = 0; l < L; l++) {
        const int pr = d->light_prims[l];
        if (pr < 0 || pr >= P) return fail(LMGPU_ERR_INVALID_ARGUMENT, "light primitive out of range");
        if (prim_light[pr] >= 0) return fail(LMGPU_ERR_INVALID_ARGUMENT, "primitive listed twice as a light");
        prim_light[pr] = l;
    }
    for (int64_t t = 0; t < N; t++) {
        const int l = prim_light[d->tri_prim[t]];
        if (l >= 0) per[l].push_back(int32_t(t));
    }
    // cdf_off indexes light_cdf (count+1 entries per light); light_tri uses
    // cdf_off - l (count entries per light).
    for (int l = 0; l < L; l++) {
        if (per[l].empty()) return fail(LMGPU_ERR_INVALID_ARGUMENT, "area light without triangles");
        const int off = int(cdf.size());
        std::vector<double> c{0.0};
        for (int32_t t : per[l]) {
            const float* v = d->verts + 9 * int64_t(t);
            const double e1[3] = {double(v[3]) - v[0], double(v[4]) - v[1], double(v[5]) - v[2]};
            const double e2[3] = {double(v[6]) - v[0], double(v[7]) - v[1], double(v[8]) - v[2]};
            const double cx = e1[1] * e2[2] - e1[2] * e2[1], cy = e1[2] * e2[0] - e1[0] * e2[2],
                         cz = e1[0] * e2[1] - e1[1] * e2[0];
            c.push_back(c.back() + std::sqrt(std::max(0.0, cx * cx + cy * cy + cz * cz)) * .5);
        }
        const double sum = c.back();
        if (!(sum > 0)) return fail(LMGPU_ERR_INVALID_ARGUMENT, "area light with zero area");
        for (double x : c) cdf.push_back(float(x / sum));
        for (int32_t t : per[l]) ltri.push_back(t);
        const float* ke = d->prim_Ke ? d->prim_Ke + 3 * d->light_prims[l] : nullptr;
        hl[l].Ke_invA = make_float4(f_or(ke, 0), f_or(ke, 1), f_or(ke, 2), float(1.0 / sum));
        hl[l].cdf_off = off;
        hl[l].count = int(per[l].size());
    }
    if (L + (d->has_env ? 1 : 0) == 0) return fail(LMGPU_ERR_INVALID_ARGUMENT, "scene has no light");
    std::vector<DevMat> hm(P);
    for (int i = 0; i < P; i++) {
        const float* mp = d->mat_param + 8 * i;
        hm[i].a = make_float4(mp[0], mp[1], mp[2], float(d->mat_type[i]));
        hm[i].b = make_float4(mp[3], mp[4], mp[5], float(prim_light[i]));
        if (d->mat_type[i] == LMGPU_MAT_GLOSSY && (mp[3] <= 0.f || mp[4] <= 0.f))
            return fail(LMGPU_ERR_INVALID_ARGUMENT, "glossy roughness must be positive");
        if (d->mat_type[i] == LMGPU_MAT_GLASS && !(mp[5] > 0.f))
            return fail(LMGPU_ERR_INVALID_ARGUMENT, "glass Ni must be positive");
    }
    LMGPU_TRY(mats.alloc(P));
    LMGPU_HIP(hipMemcpy(mats.p, hm.data(), sizeof(DevMat) * P, hipMemcpyHostToDevice));
    LMGPU_TRY(lights.alloc(hl.size()));
    LMGPU_HIP(hipMemcpy(lights.p, hl.data(), sizeof(DevLight) * hl.size(), hipMemcpyHostToDevice));
    if (cdf.empty()) cdf.push_back(0.f);
    if (ltri.empty()) ltri.push_back(0);
    LMGPU_TRY(light_cdf.alloc(cdf.size()));
    LMGPU_HIP(hipMemcpy(light_cdf.p, cdf.data(), sizeof(float) * cdf.size(), hipMemcpyHostToDevice));
    LMGPU_TRY(light_tri.alloc(ltri.size()));
    LMGPU_HIP(hipMemcpy(light_tri.p, ltri.data(), sizeof(int32_t) * ltri.size(), hipMemcpyHostToDevice));
    // camera::pinhole construct (camera_pinhole.cpp:86-96)
    Camera cam;
    {
        const double px = d->cam_pos[0], py = d->cam_pos[1], pz = d->cam_pos[2];
        double wx = px - d->cam_center[0], wy = py - d->cam_center[1], wz = pz - d->cam_center[2];
        double l = std::sqrt(wx * wx + wy * wy + wz * wz);
        wx /= l; wy /= l; wz /= l;
        const double ux0 = d->cam_up[1] * wz - d->cam_up[2] * wy, uy0 = d->cam_up[2] * wx - d->cam_up[0] * wz,
                     uz0 = d->cam_up[0] * wy - d->cam_up[1] * wx;
        l = std::sqrt(ux0 * ux0 + uy0 * uy0 + uz0 * uz0);
        if (!(l > 0)) return fail(LMGPU_ERR_INVALID_ARGUMENT, "camera up is parallel to the view direction");
        const double ux = ux0 / l, uy = uy0 / l, uz = uz0 / l;
        const double vx = wy * uz - wz * uy, vy = wz * ux - wx * uz, vz = wx * uy - wy * ux;
        cam.pos = mk(float(px), float(py), float(pz));
        cam.u = mk(float(ux), float(uy), float(uz));
        cam.v = mk(float(vx), float(vy), float(vz));
        cam.w = mk(float(wx), float(wy), float(wz));
        cam.tf = float(std::tan(double(d->vfov) * 3.14159265358979323846 / 180.0 * .5));
        cam.aspect = d->aspect;
    }
    dev = SceneDev{};
    if (accel.width == 2) dev.nodes2 = reinterpret_cast<const WNode<2>*>(accel.nodes.p);
    if (accel.width == 4) dev.nodes4 = reinterpret_cast<const WNode<4>*>(accel.nodes.p);
    if (accel.width == 8) dev.nodes8 = reinterpret_cast<const WNode<8>*>(accel.nodes.p);
    dev.tris = accel.tris.p;
    dev.stri = stri.p;
    dev.mats = mats.p;
    dev.lights = lights.p;
    dev.light_cdf = light_cdf.p;
    dev.light_tri = light_tri.p;
    dev.num_area_lights = L;
    dev.num_lights = L + (d->has_env ? 1 : 0);
    dev.has_env = d->has_env ? 1 : 0;
    dev.env_Le = mk(d->env_Le[0], d->env_Le[1], d->env_Le[2]);
    dev.cam = cam;
    dev.w = w;
    dev.h = h;
    return LMGPU_OK;
}

int Scene::ensure_workspace(int npix) {
    if (ws_npix == npix) return LMGPU_OK;
    LMGPU_TRY(thr.alloc(npix)); LMGPU_TRY(L.alloc(npix)); LMGPU_TRY(prev.alloc(npix)); LMGPU_TRY(info.alloc(npix));
    LMGPU_TRY(ext0.alloc(npix)); LMGPU_TRY(ext1.alloc(npix)); LMGPU_TRY(sh.alloc(npix)); LMGPU_TRY(shc.alloc(npix));
    LMGPU_TRY(hit.alloc(npix));
    LMGPU_TRY(path0.alloc(npix)); LMGPU_TRY(path1.alloc(npix));
    LMGPU_TRY(counts.alloc(16));
    LMGPU_TRY(next.alloc(2));
    LMGPU_HIP(hipMemset(counts.p, 0, 16 * sizeof(int32_t)));
    ws_npix = npix;
    if (graph) { (void)hipGraphExecDestroy(graph); graph = nullptr; }
    return LMGPU_OK;
}

int Scene::ensure_guided(int npix, const GuidedPass& gp) {
    // any reallocation here invalidates the captured pass
    auto drop_graph = [&]() {
        if (graph) { (void)hipGraphExecDestroy(graph); graph = nullptr; }
    };
    if (gmix.n != size_t(npix) * kOut) {
        drop_graph();
        LMGPU_TRY(gmix.alloc(size_t(npix) * kOut));
        LMGPU_TRY(gshrec.alloc(npix));
    }
    const int64_t slots = int64_t(gp.rec_depth) * npix;
    if (slots > grec_slots) {
        drop_graph();
        LMGPU_TRY(grec.alloc(slots));
        // every slot starts empty (mixture pdf 0)
        LMGPU_HIP(hipMemset(grec.p, 0, sizeof(GRec) * slots));
        grec_slots = int(slots);
    }
    const size_t tsz = size_t(gp.batch) * std::max(gp.steps, 1) * kRecordFloats;
    if (gp.steps > 0 && gtrain.n < tsz) {
        drop_graph();
        LMGPU_TRY(gtrain.alloc(tsz));
    }
    if (!gcount.p) {
        drop_graph();
        LMGPU_TRY(gcount.alloc(1));
    }
    if (gp.steps > 0) LMGPU_TRY(gp.g->ensure_train(gp.batch));
    return LMGPU_OK;
}

template <int W>
static int pt_trace_grid() {
    static int grid = 0;
    if (grid == 0) {
        int dev = 0, cus = 0, a = 0, b = 0;
        (void)hipGetDevice(&dev);
        (void)hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev);
        (void)hipOccupancyMaxActiveBlocksPerMultiprocessor(&a, k_extend<W>, kBlock, 0);
        (void)hipOccupancyMaxActiveBlocksPerMultiprocessor(&b, k_connect<W>, kBlock, 0);
        grid = std::max(1, cus) * std::max(1, std::min(a, b));
    }
    return grid;
}

static int shade_grid(int n) {
    // enough waves to fill 256 CUs several times over, bounded by the work
    return std::max(1, std::min((n + kBlock - 1) / kBlock, 256 * 8));
}

int Scene::enqueue_pass(const PTConsts& c, double* sum, double* sumsq, const GuidedPass* gp, hipStream_t s) {
    PathState ps{thr.p, L.p, prev.p, info.p};
    Queues q{{ext0.p, ext1.p}, hit.p, {path0.p, path1.p}, sh.p, shc.p, counts.p};
    GuideDev gd{};
    GBox box{};
    if (gp) {
        gd.mix = gmix.p;
        gd.rec = grec.p;
        gd.sh_rec = gshrec.p;
        gd.alpha = gp->alpha;
        gd.rec_depth = gp->rec_depth;
        for (int a = 0; a < 3; a++) {
            box.lo[a] = gp->g->box_lo[a];
            box.inv[a] = gp->g->box_inv[a];
        }
    }
    GRec* crec = (gp && gp->rec_depth > 0) ? grec.p : nullptr;
    const int32_t* cshrec = crec ? gshrec.p : nullptr;
    const int npix = c.npix;
    const unsigned gb = unsigned((npix + kBlock - 1) / kBlock);
    hipLaunchKernelGGL(k_generate, dim3(gb), dim3(kBlock), 0, s, dev, ps, q, c);
    const int sg = shade_grid(npix);
    const int tg = accel.width == 4 ? pt_trace_grid<4>() : (accel.width == 8 ? pt_trace_grid<8>() : pt_trace_grid<2>());
    for (int b = 0; b < c.max_verts - 1; b++) {
        const int in_q = b & 1;
        // camera rays are coherent: keep 64 consecutive rays per wave
        const int pm = b == 0 ? refill_primary : refill_secondary;
        hipLaunchKernelGGL(k_bounce_begin, dim3(1), dim3(64), 0, s, next.p);
        if (accel.width == 4) hipLaunchKernelGGL(k_extend<4>, dim3(tg), dim3(kBlock), 0, s, dev, q, in_q, next.p, pm);
        else if (accel.width == 8) hipLaunchKernelGGL(k_extend<8>, dim3(tg), dim3(kBlock), 0, s, dev, q, in_q, next.p, pm);
        else hipLaunchKernelGGL(k_extend<2>, dim3(tg), dim3(kBlock), 0, s, dev, q, in_q, next.p, pm);
        if (gp)
            hipLaunchKernelGGL(k_pt_guide_infer, dim3(std::max(1, std::min((npix + 127) / 128, 512))), dim3(256), 0, s,
                               dev, q, in_q, gp->g->Wbf.p, gp->g->P.p, box, gmix.p);
        hipLaunchKernelGGL(k_reset_counts, dim3(1), dim3(64), 0, s, counts.p, in_q ^ 1);
        if (gp) hipLaunchKernelGGL(k_shade<true>, dim3(sg), dim3(kBlock), 0, s, dev, ps, q, c, in_q, gd);
        else hipLaunchKernelGGL(k_shade<false>, dim3(sg), dim3(kBlock), 0, s, dev, ps, q, c, in_q, gd);
        if (c.mode != LMGPU_SAMPLING_NAIVE) {
            unsigned long long* nx = next.p + 1;
            if (accel.width == 4)
                hipLaunchKernelGGL(k_connect<4>, dim3(tg), dim3(kBlock), 0, s, dev, ps, q, nx, pm, crec, cshrec);
            else if (accel.width == 8)
                hipLaunchKernelGGL(k_connect<8>, dim3(tg), dim3(kBlock), 0, s, dev, ps, q, nx, pm, crec, cshrec);
            else hipLaunchKernelGGL(k_connect<2>, dim3(tg), dim3(kBlock), 0, s, dev, ps, q, nx, pm, crec, cshrec);
        }
    }
    hipLaunchKernelGGL(k_accumulate, dim3(gb), dim3(kBlock), 0, s, ps, sum, sumsq, npix);
    if (gp && gp->steps > 0) {
        // online training on this pass's records
        const int64_t total = int64_t(gp->rec_depth) * npix;
        const int cap = gp->batch * gp->steps;
        hipLaunchKernelGGL(k_zero_i32, dim3(1), dim3(64), 0, s, gcount.p, 1);
        hipLaunchKernelGGL(k_guide_collect, dim3(unsigned(std::min<int64_t>((total + kBlock - 1) / kBlock, 4096))),
                           dim3(kBlock), 0, s, ps, gd, npix, total, gp->mul, gtrain.p, cap, gcount.p,
                           reinterpret_cast<unsigned long long*>(counts.p + 14));
        for (int t = 0; t < gp->steps; t++)
            LMGPU_TRY(gp->g->enqueue_train(gtrain.p + int64_t(t) * gp->batch * kRecordFloats, gp->batch, gp->batch,
                                           gcount.p, int64_t(t) * gp->batch, gp->lr, nullptr, s));
    }
    hipLaunchKernelGGL(k_next_sample, dim3(1), dim3(64), 0, s, counts.p);
    LMGPU_HIP(hipGetLastError());
    return LMGPU_OK;
}

int Scene::render(const lmgpu_pt_params* p, double* sum_in, double* sumsq_in, bool on_device, const GuidedPass* gp,
                  lmgpu_pt_stats* st, hipStream_t s) {
    if (!p) return fail(LMGPU_ERR_INVALID_ARGUMENT, "params is NULL");
    if (!sum_in) return fail(LMGPU_ERR_INVALID_ARGUMENT, "film_sum is NULL");
    if (p->max_verts < 2 || p->max_verts > 64) return fail(LMGPU_ERR_INVALID_ARGUMENT, "max_verts must be in [2, 64]");
    if (p->sampling_mode < 0 || p->sampling_mode > 2) return fail(LMGPU_ERR_INVALID_ARGUMENT, "bad sampling_mode");
    if (p->render_time <= 0 && p->spp <= 0) return fail(LMGPU_ERR_INVALID_ARGUMENT, "spp must be positive");
    if (p->guiding) return fail(LMGPU_ERR_UNSUPPORTED, "guided rendering goes through lmgpu_pt_render_guided");
    const auto t0 = std::chrono::steady_clock::now();
    const int npix = w * h;
    LMGPU_TRY(ensure_workspace(npix));
    GuidedPass gpv;
    if (gp) {
        LMGPU_TRY(ensure_guided(npix, *gp));
        // buffers as they are after ensure_guided (it may grow the guide's training workspace)
        gpv = *gp;
        gpv.uid = gp->g->uid;
        gpv.gen = gp->g->gen;
        gp = &gpv;
    }
    double* sum = sum_in;
    double* sumsq = sumsq_in;
    if (!on_device) {
        if (fsum.n != size_t(npix) * 3) LMGPU_TRY(fsum.alloc(size_t(npix) * 3));
        sum = fsum.p;
        if (sumsq_in) {
            if (fsq.n != size_t(npix) * 3) LMGPU_TRY(fsq.alloc(size_t(npix) * 3));
            sumsq = fsq.p;
        }
        const unsigned zb = unsigned((int64_t(npix) * 3 + 255) / 256);
        hipLaunchKernelGGL(k_zero_f64, dim3(zb), dim3(256), 0, s, sum, int64_t(npix) * 3);
        if (sumsq) hipLaunchKernelGGL(k_zero_f64, dim3(zb), dim3(256), 0, s, sumsq, int64_t(npix) * 3);
    }
    PTConsts c;
    c.max_verts = p->max_verts;
    c.mode = p->sampling_mode;
    c.rr_depth = p->rr_depth;
    c.seed_lo = uint32_t(p->seed);
    c.seed_hi = uint32_t(p->seed >> 32);
    c.npix = npix;
    // stats counters [8..15] as 64-bit: [8] samples, [10] extension rays, [12] shadow rays,
    // [14] guide training records
    LMGPU_HIP(hipMemsetAsync(counts.p + 8, 0, 8 * sizeof(int32_t), s));
    hipLaunchKernelGGL(k_set_sample, dim3(1), dim3(64), 0, s, counts.p, p->sample_offset);
    // one pass captured as a graph, replayed per sample
    const GuidedPass none;
    const GuidedPass& gk = gp ? *gp : none;
    const bool guide_same = graph_guide.g == gk.g && graph_guide.uid == gk.uid && graph_guide.gen == gk.gen &&
                            graph_guide.alpha == gk.alpha && graph_guide.lr == gk.lr &&
                            graph_guide.batch == gk.batch && graph_guide.steps == gk.steps &&
                            graph_guide.rec_depth == gk.rec_depth && graph_guide.mul == gk.mul;
    const bool key_same = graph && graph_key[0] == c.max_verts && graph_key[1] == c.mode && graph_key[2] == c.rr_depth &&
                          graph_seed == p->seed && graph_sum_ptr == sum && graph_sq_ptr == sumsq && guide_same;
    if (!key_same) {
        if (graph) { (void)hipGraphExecDestroy(graph); graph = nullptr; }
        hipStream_t cs;
        LMGPU_HIP(hipStreamCreateWithFlags(&cs, hipStreamNonBlocking));
        hipGraph_t g;
        LMGPU_HIP(hipStreamBeginCapture(cs, hipStreamCaptureModeThreadLocal));
        const int rc = enqueue_pass(c, sum, sumsq, gp, cs);
        hipError_t e = hipStreamEndCapture(cs, &g);
        if (rc != LMGPU_OK) { (void)hipStreamDestroy(cs); return rc; }
        if (e != hipSuccess) { (void)hipStreamDestroy(cs); return fail(LMGPU_ERR_DEVICE, std::string("graph capture: ") + hipGetErrorString(e)); }
        e = hipGraphInstantiate(&graph, g, nullptr, nullptr, 0);
        (void)hipGraphDestroy(g);
        (void)hipStreamDestroy(cs);
        if (e != hipSuccess) { graph = nullptr; return fail(LMGPU_ERR_DEVICE, std::string("graph instantiate: ") + hipGetErrorString(e)); }
        graph_key[0] = c.max_verts; graph_key[1] = c.mode; graph_key[2] = c.rr_depth;
        graph_seed = p->seed;
        graph_guide = gk;
        graph_sum_ptr = sum;
        graph_sq_ptr = sumsq;
    }
    int64_t passes = 0;
    if (p->render_time > 0) {
        // scheduler::spp::time: 1-spp passes until the budget is spent
        while (true) {
            LMGPU_HIP(hipGraphLaunch(graph, s));
            passes++;
            LMGPU_HIP(hipStreamSynchronize(s));
            const double el = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
            if (el > p->render_time) break;
        }
    } else {
        for (int i = 0; i < p->spp; i++) {
            LMGPU_HIP(hipGraphLaunch(graph, s));
            passes++;
        }
    }
    int32_t cnt[16];
    LMGPU_HIP(hipMemcpyAsync(cnt, counts.p, sizeof(cnt), hipMemcpyDeviceToHost, s));
    if (!on_device) {
        LMGPU_HIP(hipMemcpyAsync(sum_in, sum, sizeof(double) * 3 * npix, hipMemcpyDeviceToHost, s));
        if (sumsq_in) LMGPU_HIP(hipMemcpyAsync(sumsq_in, sumsq, sizeof(double) * 3 * npix, hipMemcpyDeviceToHost, s));
    }
    LMGPU_HIP(hipStreamSynchronize(s));
    if (cnt[5]) return fail(LMGPU_ERR_STACK_OVERFLOW, "BVH traversal stack capacity exceeded during rendering");
    if (st) {
        std::memset(st, 0, sizeof(*st));
        int64_t v[4];
        std::memcpy(v, cnt + 8, sizeof(v));
        st->samples = v[0];
        st->spp = passes;
        st->extension_rays = v[1];
        st->shadow_rays = v[2];
        st->guide_records = gp ? v[3] : 0;
        st->elapsed_s = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    }
    return LMGPU_OK;
}

int Scene::raycast(const lmgpu_raycast_params* p, double* film_in, bool on_device, hipStream_t s) {
    if (!p) return fail(LMGPU_ERR_INVALID_ARGUMENT, "params is NULL");
    if (!film_in) return fail(LMGPU_ERR_INVALID_ARGUMENT, "film is NULL");
    const int npix = w * h;
    double* film = film_in;
    if (!on_device) {
        if (fsum.n != size_t(npix) * 3) LMGPU_TRY(fsum.alloc(size_t(npix) * 3));
        film = fsum.p;
    }
    RaycastConsts rc;
    rc.bg = mk(p->bg_color[0], p->bg_color[1], p->bg_color[2]);
    rc.color = mk(p->color[0], p->color[1], p->color[2]);
    rc.use_color = p->use_color != 0;
    rc.constant = p->use_constant_color != 0;
    rc.normal = p->visualize_normal != 0;
    unsigned long long* next = accel.next_counter();
    int32_t* status = accel.status_dev();
    if (!next || !status) return fail(LMGPU_ERR_OUT_OF_MEMORY, "traversal counters");
    hipLaunchKernelGGL(k_zero_u64, dim3(1), dim3(64), 0, s, next);
    // camera rays are coherent: keep 64 consecutive pixels per wave
    const int refill = 64;
    if (accel.width == 4)
        hipLaunchKernelGGL(k_raycast<4>, dim3(pt_trace_grid<4>()), dim3(kBlock), 0, s, dev, next, refill, rc, film, status);
    else if (accel.width == 8)
        hipLaunchKernelGGL(k_raycast<8>, dim3(pt_trace_grid<8>()), dim3(kBlock), 0, s, dev, next, refill, rc, film, status);
    else
        hipLaunchKernelGGL(k_raycast<2>, dim3(pt_trace_grid<2>()), dim3(kBlock), 0, s, dev, next, refill, rc, film, status);
    LMGPU_HIP(hipGetLastError());
    if (!on_device) LMGPU_HIP(hipMemcpyAsync(film_in, film, sizeof(double) * 3 * npix, hipMemcpyDeviceToHost, s));
    return accel.check_status(s);
}

}  // namespace lmgpu

using lmgpu::fail;

extern "C" {

struct lmgpu_scene {
    lmgpu::Scene s;
};

int lmgpu_scene_create(lmgpu_scene** out, const lmgpu_scene_desc* desc, const lmgpu_build_opts* opts) {
    if (!out) return fail(LMGPU_ERR_INVALID_ARGUMENT, "out is NULL");
    *out = nullptr;
    auto* h = new (std::nothrow) lmgpu_scene();
    if (!h) return fail(LMGPU_ERR_OUT_OF_MEMORY, "scene handle");
    const int rc = h->s.create(desc, opts);
    if (rc != LMGPU_OK) {
        delete h;
        return rc;
    }
    *out = h;
    return LMGPU_OK;
}

int lmgpu_scene_destroy(lmgpu_scene* scene) {
    delete scene;
    return LMGPU_OK;
}

struct lmgpu_accel;
int lmgpu_scene_accel(lmgpu_scene* scene, lmgpu_accel** accel) {
    if (!scene || !accel) return fail(LMGPU_ERR_INVALID_ARGUMENT, "NULL argument");
    // lmgpu_accel wraps exactly one lmgpu::Accel (capi.hip)
    *accel = reinterpret_cast<lmgpu_accel*>(&scene->s.accel);
    return LMGPU_OK;
}

int lmgpu_pt_render(lmgpu_scene* scene, const lmgpu_pt_params* params, double* film_sum, double* film_sumsq,
                    int film_on_device, lmgpu_pt_stats* stats, void* stream) {
    if (!scene) return fail(LMGPU_ERR_INVALID_ARGUMENT, "scene is NULL");
    lmgpu::clear_error();
    return scene->s.render(params, film_sum, film_sumsq, film_on_device != 0, nullptr, stats,
                           static_cast<hipStream_t>(stream));
}

int lmgpu_raycast_render(lmgpu_scene* scene, const lmgpu_raycast_params* params, double* film, int film_on_device,
                         void* stream) {
    if (!scene) return fail(LMGPU_ERR_INVALID_ARGUMENT, "scene is NULL");
    lmgpu::clear_error();
    return scene->s.raycast(params, film, film_on_device != 0, static_cast<hipStream_t>(stream));
}

static int64_t gcd64(int64_t a, int64_t b) {
    while (b) {
        const int64_t t = a % b;
        a = b;
        b = t;
    }
    return a;
}

int lmgpu_pt_render_guided(lmgpu_scene* scene, const lmgpu_pt_params* params, lmgpu_guide* guide,
                           const lmgpu_guide_opts* opts, double* film_sum, double* film_sumsq, int film_on_device,
                           lmgpu_pt_stats* stats, void* stream) {
    if (!scene || !guide) return fail(LMGPU_ERR_INVALID_ARGUMENT, "scene and guide are required");
    lmgpu::clear_error();
    lmgpu_guide_opts o{};
    o.alpha = 0.5f;
    o.learning_rate = 1e-3f;
    o.batch = 1 << 18;
    o.steps_per_pass = 1;
    o.record_depth = 8;
    if (opts) o = *opts;
    if (!(o.alpha >= 0.f && o.alpha <= 1.f)) return fail(LMGPU_ERR_INVALID_ARGUMENT, "alpha must be in [0, 1]");
    if (!(o.learning_rate >= 0.f)) return fail(LMGPU_ERR_INVALID_ARGUMENT, "learning rate must be >= 0");
    if (o.steps_per_pass < 0 || o.steps_per_pass > 64) return fail(LMGPU_ERR_INVALID_ARGUMENT, "steps_per_pass must be in [0, 64]");
    if (o.steps_per_pass > 0 && (o.batch < 1 || o.batch > (1 << 22)))
        return fail(LMGPU_ERR_INVALID_ARGUMENT, "batch must be in [1, 2^22]");
    if (o.record_depth < 0 || o.record_depth > 16) return fail(LMGPU_ERR_INVALID_ARGUMENT, "record_depth must be in [0, 16]");
    lmgpu::GuidedPass gp;
    gp.g = &guide->g;
    gp.alpha = o.alpha;
    gp.lr = o.learning_rate;
    gp.steps = o.steps_per_pass;
    gp.batch = o.steps_per_pass > 0 ? (o.batch + 127) / 128 * 128 : 0;
    gp.rec_depth = o.steps_per_pass > 0 ? o.record_depth : 0;
    if (gp.rec_depth == 0) gp.steps = 0;
    const int64_t npix = int64_t(scene->s.w) * scene->s.h;
    const int64_t total = int64_t(gp.rec_depth) * npix;
    if (total >= (int64_t(1) << 31)) return fail(LMGPU_ERR_INVALID_ARGUMENT, "record_depth * pixels must be < 2^31");
    if (int64_t(gp.batch) * std::max(gp.steps, 1) >= (int64_t(1) << 31))
        return fail(LMGPU_ERR_INVALID_ARGUMENT, "batch * steps_per_pass must be < 2^31");
    if (total > 1) {
        // a multiplier coprime to the slot count: a bijection of the slots
        int64_t m = (total * 5) / 8 + 1;
        while (gcd64(m, total) != 1) m++;
        gp.mul = m % total;
    }
    return scene->s.render(params, film_sum, film_sumsq, film_on_device != 0, &gp, stats,
                           static_cast<hipStream_t>(stream));
}

}  // extern "C"
