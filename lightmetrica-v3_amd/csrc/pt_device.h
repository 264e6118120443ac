// Device data layout and per-vertex transport of the wavefront path tracer,
// shared by the single-scene renderer (pt.hip) and the batched multi-job
// renderer (batch.hip).
#pragma once

#include <hip/hip_runtime.h>

#include "bvh.h"
#include "guide.h"
#include "shade.h"

namespace lmgpu {

// Triangle shading record, original (flattened) order: positions, vertex
// normals; p0.w = primitive index bits.
struct alignas(16) ShadeTri {
    float4 p0, p1, p2, n0, n1, n2;
};

struct alignas(16) DevMat {
    float4 a;   // rgb (Kd; Ks of glossy), type (as float)
    float4 b;   // ax, ay, ni, light index (as float, -1 = none)
    float4 c;   // mixture Ks rgb; w = int bits: texture of Kd (-1 = none), bit 30 = alpha mask
};

constexpr int kTexAlphaBit = 1 << 30;

struct alignas(16) DevLight {
    float4 Ke_invA;   // Ke rgb, 1/area
    int32_t cdf_off, count, pad0, pad1;
};

struct SceneDev {
    const TravNode<4>::type* nodes4;
    const WNode<2>* nodes2;
    const TravNode<8>::type* nodes8;
    const TriV* tris;
    const ShadeTri* stri;
    const DevMat* mats;
    const DevLight* lights;
    const float* light_cdf;    // per light: count+1 entries, c[0] = 0, normalised
    const int32_t* light_tri;  // per light: original triangle indices (face order)
    const float2* uvs;         // [3N] texture coordinates per triangle corner (textured scenes only)
    const float4* texels;      // all textures' texels, RGBA (A = 1 for RGB textures)
    const int4* tex_info;      // per texture: width, height, channels, first texel
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
    int image = 0;   // primary_ray_sampling_mode = image
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
    int32_t* rec_list;  // slots written in this pass
    int32_t* rec_count; // number of entries of rec_list
    float alpha;        // probability of sampling the guide at a non-specular vertex
    int rec_depth;      // surface vertices per path that record (0: no recording)
};

struct GBox {
    float lo[3], inv[3];
};

constexpr int kBlock = 256;
constexpr int kLds = 24;
constexpr int kPriv = 72;

// Material of primitive `prim` at barycentrics (b1, b2) of triangle `tri`.
// A textured Kd is texture::bitmap::eval at the interpolated texture
// coordinate (Mesh_Raw::surface_point, mesh_raw.cpp:87-97): wrapped to
// [0,1), nearest texel (texture_bitmap.cpp:68-75); a mixture with an alpha
// mask also takes the texel's alpha (texture_bitmap.cpp:82-89).
__device__ __forceinline__ Mat load_mat(const SceneDev& sc, int prim, int tri, float b1, float b2, int& light) {
    const DevMat d = sc.mats[prim];
    Mat m;
    m.type = int(d.a.w);
    m.c = mk(d.a.x, d.a.y, d.a.z);
    m.ax = d.b.x;
    m.ay = d.b.y;
    m.ni = d.b.z;
    m.ks = mk(d.c.x, d.c.y, d.c.z);
    m.alpha = 1.f;
    light = int(d.b.w);
    const int tx = __float_as_int(d.c.w);
    if (tx >= 0) {
        const float2 t0 = sc.uvs[3 * int64_t(tri)], t1 = sc.uvs[3 * int64_t(tri) + 1], t2 = sc.uvs[3 * int64_t(tri) + 2];
        const float b0 = 1.f - b1 - b2;
        float u = t0.x * b0 + t1.x * b1 + t2.x * b2;
        float v = t0.y * b0 + t1.y * b1 + t2.y * b2;
        u -= floorf(u);
        v -= floorf(v);
        const int4 ti = sc.tex_info[tx & ~kTexAlphaBit];
        const int x = min(max(int(u * float(ti.x)), 0), ti.x - 1);
        const int y = min(max(int(v * float(ti.y)), 0), ti.y - 1);
        const float4 tv = sc.texels[ti.w + int64_t(y) * ti.x + x];
        m.c = mk(tv.x, tv.y, tv.z);
        if (tx & kTexAlphaBit) m.alpha = tv.w;
    }
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

// Appends to two queues for the whole block: one atomic per block and queue
// on the queue counts (which every block of the launch shares) instead of
// one per wave. i0 / i1 are the entries of the lanes whose flag is set.
// Every thread of the block calls it, the same number of times.
template <int BLOCK>
__device__ __forceinline__ void block_append2(int32_t* c0, int32_t* c1, bool f0, bool f1, int& i0, int& i1) {
    constexpr int kW = BLOCK / 64;
    __shared__ int cnt[2][kW];
    __shared__ int base[2];
    const int lane = __lane_id(), wave = threadIdx.x >> 6;
    const unsigned long long m0 = __ballot(f0), m1 = __ballot(f1);
    if (lane == 0) {
        cnt[0][wave] = __popcll(m0);
        cnt[1][wave] = __popcll(m1);
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        int a = 0, b = 0;
        for (int w = 0; w < kW; w++) {
            const int x = cnt[0][w], y = cnt[1][w];
            cnt[0][w] = a;
            cnt[1][w] = b;
            a += x;
            b += y;
        }
        base[0] = a ? atomicAdd(c0, a) : 0;
        base[1] = b ? atomicAdd(c1, b) : 0;
    }
    __syncthreads();
    const unsigned long long lt = (1ull << lane) - 1ull;
    i0 = f0 ? base[0] + cnt[0][wave] + __popcll(m0 & lt) : -1;
    i1 = f1 ? base[1] + cnt[1][wave] + __popcll(m1 & lt) : -1;
    __syncthreads();   // the next call rewrites cnt and base
}

// As block_append2, for one queue.
template <int BLOCK>
__device__ __forceinline__ int block_append1(int32_t* c0, bool f0) {
    constexpr int kW = BLOCK / 64;
    __shared__ int cnt[kW];
    __shared__ int base;
    const int lane = __lane_id(), wave = threadIdx.x >> 6;
    const unsigned long long m0 = __ballot(f0);
    if (lane == 0) cnt[wave] = __popcll(m0);
    __syncthreads();
    if (threadIdx.x == 0) {
        int a = 0;
        for (int w = 0; w < kW; w++) {
            const int x = cnt[w];
            cnt[w] = a;
            a += x;
        }
        base = a ? atomicAdd(c0, a) : 0;
    }
    __syncthreads();
    const int i0 = f0 ? base + cnt[wave] + __popcll(m0 & ((1ull << lane) - 1ull)) : -1;
    __syncthreads();   // the next call rewrites cnt and base
    return i0;
}

template <int W>
__device__ __forceinline__ const typename TravNode<W>::type* nodes_of(const SceneDev& sc);
template <>
__device__ __forceinline__ const TravNode<2>::type* nodes_of<2>(const SceneDev& sc) { return sc.nodes2; }
template <>
__device__ __forceinline__ const TravNode<4>::type* nodes_of<4>(const SceneDev& sc) { return sc.nodes4; }
template <>
__device__ __forceinline__ const TravNode<8>::type* nodes_of<8>(const SceneDev& sc) { return sc.nodes8; }

// A point on area light `li` and its geometric normal (light::area:
// triangle by the area CDF with Dist::sample's upper_bound, math.h:367-370;
// barycentrics (1 - sqrt(u0), u1 sqrt(u0)), light_area.cpp:47-59).
__device__ __forceinline__ void sample_light_point(const SceneDev& sc, int li, float ut, float u0, float u1, f3& p,
                                                   f3& n) {
    const DevLight dl = sc.lights[li];
    const float* cdf = sc.light_cdf + dl.cdf_off;
    int lo = 0, hi = dl.count + 1;
    while (lo < hi) {
        const int mid = (lo + hi) >> 1;
        if (cdf[mid] <= ut) lo = mid + 1;
        else hi = mid;
    }
    const int k = min(max(lo - 1, 0), dl.count - 1);
    const ShadeTri ls = sc.stri[sc.light_tri[dl.cdf_off - li + k]];
    const f3 A = mk(ls.p0.x, ls.p0.y, ls.p0.z), B = mk(ls.p1.x, ls.p1.y, ls.p1.z), C = mk(ls.p2.x, ls.p2.y, ls.p2.z);
    const float s = sqrtf(fmaxf(0.f, u0));
    const float bu = 1.f - s, bv = u1 * s;
    p = A * (1.f - bu - bv) + B * bu + C * bv;
    n = normalize(cross(B - A, C - A));
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

// Camera ray of `pixel` for `sample` (camera::pinhole primary ray through a
// jittered raster position, src/camera/camera_pinhole.cpp:128-132; the
// jitter is (pixel, sample, 0, 0)[2..3] of the Philox stream). With `image`
// the same two numbers are the raster position over the whole image.
__device__ __forceinline__ lmgpu_ray camera_ray(const SceneDev& sc, int pixel, uint32_t sample, uint64_t seed,
                                                bool image = false) {
    const int x = pixel % sc.w, y = pixel / sc.w;
    const rand4 r = rng(uint32_t(pixel), sample, 0u, 0u, seed);
    const float dx = 1.f / sc.w, dy = 1.f / sc.h;
    const f3 d = image ? primary_dir(sc.cam, r.u[2], r.u[3])
                       : primary_dir(sc.cam, dx * x + dx * r.u[2], dy * y + dy * r.u[3]);
    lmgpu_ray ray;
    ray.org[0] = sc.cam.pos.x; ray.org[1] = sc.cam.pos.y; ray.org[2] = sc.cam.pos.z;
    ray.tmin = kEps;
    ray.dir[0] = d.x; ray.dir[1] = d.y; ray.dir[2] = d.z;
    ray.tmax = kInf;
    return ray;
}

// Registers of one path between two vertices.
struct PathRegs {
    f3 thr, L;     // throughput, radiance
    float4 prev;   // previous vertex: shading normal, pdf of the sampled direction (projected SA)
    int4 info;     // vertex count, component, flags (bit0 prev samplable by NEE, bit1 prev degenerate)
};

// What one vertex emits: the next extension ray and / or an NEE shadow ray
// carrying its unoccluded contribution.
struct VertexOut {
    bool ext = false, sh = false;
    lmgpu_ray ext_ray, sh_ray;
    f3 sh_C = {0, 0, 0};
    int sh_slot = -1;   // guide record slot the NEE contribution belongs to (-1: none)
    int rec_slot = -1;  // guide record slot written by this vertex (-1: none)
};

// One path vertex of Renderer_PT::render (src/renderer/renderer_pt.cpp:
// 133-302): the ray (o, wo) of the path ended at hit h. Adds the emitter
// hit or environment contribution with its MIS weight (:236-271), applies
// Russian roulette (:281-287) and component sampling (:292), emits the NEE
// shadow ray of path::sample_direct (:142-192) and the next direction of
// path::sample_direction (:197-229). Random numbers are keyed by (pixel,
// sample, vertex, block) and the seed. On return P holds the path after the
// vertex (thr, prev, info only meaningful when out.ext).
//
// Guided vertices (GUIDED, diffuse or glossy; `gmix_entry` is the decoded
// guide mixture of this vertex) sample the next direction from the
// one-sample MIS mixture alpha * guide + (1 - alpha) * BSDF, whose density
// (solid angle) p_mix enters both the throughput f |cos| / p_mix and, as
// p_mix / |cos|, every MIS weight against light sampling; they also write
// their training record into slot (vertex - 1) * npix + path of gd.rec.
template <bool GUIDED>
__device__ __forceinline__ void shade_vertex(const SceneDev& sc, const PTConsts& c, uint32_t pixel, uint32_t sample,
                                             uint64_t seed, int path, f3 o, f3 wo, float4 h, PathRegs& P,
                                             const GuideDev& gd, const float* gmix_entry, VertexOut& out) {
    const int tri = __float_as_int(h.x);
    const float4 pv = P.prev;
    int4 info = P.info;
    f3 thr = P.thr;
    f3 L = P.L;
    const f3 np = mk(pv.x, pv.y, pv.z);
    const bool prev_nee = info.z & 1;
    const bool prev_degen = info.z & 2;
    const bool direct_ok = c.mode == LMGPU_SAMPLING_NEE ? !prev_nee : true;
    const bool use_mis = c.mode != LMGPU_SAMPLING_NAIVE && prev_nee;
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
        const Mat m = load_mat(sc, prim, tri, b1, b2, light);
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
        const rand4 r0 = rng(pixel, sample, uint32_t(nv), 0u, seed);
        bool alive = true;
        if (nv > c.rr_depth) {
            const float qq = fmaxf(.2f, 1.f - maxc(thr));
            if (r0.u[0] < qq) alive = false;
            else thr = thr / (1.f - qq);
        }
        const f3 wi = -wo;
        int comp = 0;
        float ul = 0.f;   // lobe selection of a mixture
        if (alive) {
            float cw;
            comp = mat_sample_component(m, g, wi, r0.u[1], cw, ul);
            thr = thr * cw;
        }
        const int nv2 = nv + 1;
        if (alive && nv2 < c.max_verts) {
            const bool spec = is_specular(m, comp);
            const bool nee = c.mode != LMGPU_SAMPLING_NAIVE && !spec;
            // guide mixture of this queue entry (guided vertices only)
            const float* gmix = (GUIDED && !spec) ? gmix_entry : nullptr;
            const int slot = (gmix && nv - 1 < gd.rec_depth) ? (nv - 1) * c.npix + path : -1;
            if (nee) {
                // path::sample_direct (path.h:509-528)
                const rand4 r1 = rng(pixel, sample, uint32_t(nv), 1u, seed);
                const int li = min(max(int(r1.u[0] * sc.num_lights), 0), sc.num_lights - 1);
                const float psel = 1.f / sc.num_lights;
                f3 dirL;   // from the shading point toward the light
                float tmaxL = 0.f;
                f3 weight = mk(0, 0, 0);
                float p_light = 0.f;
                bool ok = false;
                if (li < sc.num_area_lights) {
                    const DevLight dl = sc.lights[li];
                    f3 pL, nL;
                    sample_light_point(sc, li, r1.u[1], r1.u[2], r1.u[3], pL, nL);
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
                    const f3 fs = mat_eval(m, g, wi, dirL, comp) * sn_correction(g, wi, dirL);
                    if (!is_zero(fs)) {
                        float mis = 1.f;
                        if (c.mode == LMGPU_SAMPLING_MIS) {
                            float pdir = mat_pdf(m, g, wi, dirL, comp);
                            if (GUIDED && gmix) {
                                float pg = mixture_pdf(gmix, dirL);
                                if (!(pg >= 0.f && pg < 3e38f)) pg = 0.f;
                                pdir = gd.alpha * pg / fabsf(dot(dirL, g.n)) + (1.f - gd.alpha) * pdir;
                            }
                            mis = balance_heuristic(p_light, pdir);
                        }
                        const f3 C = thr * fs * weight * mis;
                        if (!is_zero(C)) {
                            out.sh = true;
                            out.sh_slot = slot;
                            out.sh_ray.org[0] = pos.x; out.sh_ray.org[1] = pos.y; out.sh_ray.org[2] = pos.z;
                            out.sh_ray.tmin = kEps;
                            out.sh_ray.dir[0] = dirL.x; out.sh_ray.dir[1] = dirL.y; out.sh_ray.dir[2] = dirL.z;
                            out.sh_ray.tmax = tmaxL;
                            out.sh_C = C;
                        }
                    }
                }
            }
            // path::sample_direction (path.h:410-419)
            f3 wo2, sw;
            bool sampled;
            float pdir = 0.f, pmix = 0.f;
            if (GUIDED && gmix) {
                const rand4 r2 = rng(pixel, sample, uint32_t(nv), 2u, seed);
                if (r2.u[0] < gd.alpha) {
                    wo2 = mixture_sample(gmix, r2.u[1], r2.u[2], r2.u[3]);
                    sampled = true;
                } else {
                    sampled = mat_sample_direction(m, g, wi, comp, r0.u[2], r0.u[3], ul, wo2, sw);
                }
                sw = mk(0, 0, 0);
                if (sampled) {
                    const float cs = fabsf(dot(wo2, g.n));
                    float pg = mixture_pdf(gmix, wo2);
                    if (!(pg >= 0.f && pg < 3e38f)) pg = 0.f;
                    pmix = gd.alpha * pg + (1.f - gd.alpha) * mat_pdf(m, g, wi, wo2, comp) * cs;
                    if (pmix > 0.f && cs > 0.f) {
                        sw = mat_eval(m, g, wi, wo2, comp) * (cs / pmix);
                        pdir = pmix / cs;
                    }
                }
            } else {
                sampled = mat_sample_direction(m, g, wi, comp, r0.u[2], r0.u[3], ul, wo2, sw);
                if (sampled) pdir = mat_pdf(m, g, wi, wo2, comp);
            }
            if (sampled) {
                sw = sw * sn_correction(g, wi, wo2);
                const f3 thr_before = thr;
                thr = thr * sw;
                if (!is_zero(thr)) {
                    out.ext = true;
                    out.ext_ray.org[0] = pos.x; out.ext_ray.org[1] = pos.y; out.ext_ray.org[2] = pos.z;
                    out.ext_ray.tmin = kEps;
                    out.ext_ray.dir[0] = wo2.x; out.ext_ray.dir[1] = wo2.y; out.ext_ray.dir[2] = wo2.z;
                    out.ext_ray.tmax = kInf;
                    P.prev = make_float4(g.n.x, g.n.y, g.n.z, pdir);
                    P.info = make_int4(nv2, comp, nee ? 1 : 0, 0);
                    P.thr = thr;
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
                out.rec_slot = slot;
                    }
                }
            }
        }
    }
    P.L = L;
}

// Closest hit of the extension queue `in_q` (persistent traversal), hits
// written as (prim or -1 / -2 on stack overflow, t, u, v); shared by the
// path and light tracers.
template <int W>
__global__ __launch_bounds__(kBlock) void k_extend(SceneDev sc, Queues q, int in_q, unsigned long long* next, int refill,
                                                  int chunk_rays) {
    __shared__ int lds[kLds][kBlock];
    const int n = q.counts[in_q];
    if (n == 0) return;   // an empty queue (whole grid)
    const lmgpu_ray* rays = q.ext[in_q];
    auto fetch = [&](int64_t id, f3& o, f3& d, float& tmin, float& tmax) -> int {
        const float4* rp = reinterpret_cast<const float4*>(rays + id);
        const float4 a = rp[0], b = rp[1];
        o = mk(a.x, a.y, a.z);
        d = mk(b.x, b.y, b.z);
        tmin = a.w;
        tmax = b.w;
        return 0;
    };
    auto emit = [&](int64_t id, const TraceResult& r) {
        reinterpret_cast<float4*>(q.hit)[id] = make_float4(__int_as_float(r.overflow ? -2 : r.prim), r.t, r.u, r.v);
    };
    trace_persistent<W, false, false, kLds, kPriv, kBlock>(nodes_of<W>(sc), sc.tris, n, next, lds, fetch, emit,
                                                           q.counts + 5, nullptr, refill, chunk_rays);
    if (blockIdx.x == 0 && threadIdx.x == 0)
        atomicAdd(reinterpret_cast<unsigned long long*>(q.counts + 10), (unsigned long long)n);
}

// Resident-occupancy grid of k_extend<W> on the current device.
template <int W>
inline int extend_grid() {
    static int grid = 0;
    if (grid == 0) {
        int dev = 0, cus = 0, a = 0;
        (void)hipGetDevice(&dev);
        (void)hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev);
        (void)hipOccupancyMaxActiveBlocksPerMultiprocessor(&a, k_extend<W>, kBlock, 0);
        grid = (cus > 0 ? cus : 1) * (a > 0 ? a : 1);
    }
    return grid;
}

// Guide inference over the extension queue of a bounce: the query of entry
// i is its hit point, shading normal and the direction back along the ray;
// the decoded mixture goes to mix[i]. Entries without a hit get a zero query.
template <class Q>
__global__ __launch_bounds__(256) void k_pt_guide_infer(SceneDev sc, Q q, int in_q, const uint16_t* __restrict__ Wbf,
                                                        const float* __restrict__ P, GBox box, float* __restrict__ mix) {
    __shared__ __attribute__((aligned(16))) unsigned char lds[kWeightLds + 4 * kWaveLds];
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    uint16_t* wl = reinterpret_cast<uint16_t*>(lds);
    stage_weights(Wbf, wl, threadIdx.x, 256);
    uint16_t* lx = reinterpret_cast<uint16_t*>(lds + kWeightLds + wave * kWaveLds);
    uint16_t* lh = lx + 32 * kPitchX;
    float* ly = reinterpret_cast<float*>(lx);
    const int n = q.counts[in_q];
    // two lanes per row: lanes r and r + 32 share row r of the wave's 32-row
    // tile and each encode and decode half of it
    const int r = lane & 31, half = lane >> 5;
    for (int base = blockIdx.x * 128; base < n; base += gridDim.x * 128) {
        const int i0 = base + wave * 32;
        {
            const int i = i0 + r;
            uint16_t* row = lx + r * kPitchX;
            const float4 h = i < n ? reinterpret_cast<const float4*>(q.hit)[i] : make_float4(__int_as_float(-1), 0, 0, 0);
            const int tri = __float_as_int(h.x);
            if (tri >= 0) {
                const ShadeTri st = sc.stri[tri];
                const float b1 = h.z, b2 = h.w, b0 = 1.f - b1 - b2;
                const f3 pos = mk(st.p0.x, st.p0.y, st.p0.z) * b0 + mk(st.p1.x, st.p1.y, st.p1.z) * b1 +
                               mk(st.p2.x, st.p2.y, st.p2.z) * b2;
                const f3 nsh = normalize(mk(st.n0.x, st.n0.y, st.n0.z) * b0 + mk(st.n1.x, st.n1.y, st.n1.z) * b1 +
                                         mk(st.n2.x, st.n2.y, st.n2.z) * b2);
                const lmgpu_ray& ray = q.ext[in_q][i];
                encode_input_half(half, pos, nsh, mk(-ray.dir[0], -ray.dir[1], -ray.dir[2]), box.lo, box.inv, row);
            } else {
                for (int k = half * (kIn / 2); k < (half + 1) * (kIn / 2); k++) row[k] = 0;
            }
        }
        __syncthreads();
        mlp_tile_forward(wl, P, lx, lh, ly, lane, [](int, int, int, uint16_t) {});
        __syncthreads();
        {
            const int i = i0 + r;
            if (i < n) decode_lobes4(ly + r * kPitchY, 4 * half, mix + int64_t(i) * kOut);
        }
        __syncthreads();
    }
}

}  // namespace lmgpu
