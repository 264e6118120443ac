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

#include "guide_internal.h"
#include "lmgpu_internal.h"
#include "pt_device.h"
#include "pt_internal.h"

namespace lmgpu {

// IMAGE is a template parameter so that the pixel-mode camera ray is the
// same code (and the same FP32 rounding) as in the batched renderer.
template <bool IMAGE>
__global__ __launch_bounds__(kBlock) void k_generate(SceneDev sc, PathState ps, Queues q, PTConsts c) {
    const int p = blockIdx.x * kBlock + threadIdx.x;
    if (p >= c.npix) return;
    const uint32_t sample = uint32_t(q.counts[3]);
    const uint64_t seed = (uint64_t(c.seed_hi) << 32) | c.seed_lo;
    const lmgpu_ray ray = camera_ray(sc, p, sample, seed, IMAGE);
    ps.thr[p] = make_float4(1.f, 1.f, 1.f, 0.f);
    ps.L[p] = make_float4(0.f, 0.f, 0.f, 0.f);
    if (IMAGE && c.mode != LMGPU_SAMPLING_NAIVE) {
        // image mode: the camera vertex is NEE-samplable; the pdf of the
        // sampled direction is Camera_Pinhole::pdf_direction = J(wo)
        // (renderer_pt.cpp:133-136, 258-261)
        ps.prev[p] = make_float4(0.f, 0.f, 0.f, cam_J(sc.cam, mk(ray.dir[0], ray.dir[1], ray.dir[2])));
        ps.info[p] = make_int4(1, 0, 3, 0);   // nv = 1, comp 0, prev NEE-samplable and degenerate
    } else {
        ps.prev[p] = make_float4(0.f, 0.f, 0.f, 0.f);
        ps.info[p] = make_int4(1, 0, 2, 0);   // nv = 1, comp 0, prev degenerate (camera), not NEE-samplable
    }
    q.ext[0][p] = ray;
    q.ext_path[0][p] = p;
    if (p == 0) {
        q.counts[0] = c.npix;
        atomicAdd(reinterpret_cast<unsigned long long*>(q.counts + 8), (unsigned long long)c.npix);
    }
}

// Image mode: NEE at the camera vertex (renderer_pt.cpp:138-188 with sp the
// camera): a light point sampled as at every other vertex (Philox block 1 of
// vertex 0), Camera_Pinhole::eval = J(wo) inside the frustum, MIS against
// reaching the light from the camera (balance heuristic with J), queued as a
// shadow ray whose contribution goes to the pixel of the light's raster
// position (contribution w = -(pixel + 1)).
__global__ __launch_bounds__(kBlock) void k_camera_nee(SceneDev sc, Queues q, PTConsts c) {
    const int p = blockIdx.x * kBlock + threadIdx.x;
    bool emit = false;
    lmgpu_ray ray;
    float4 C;
    if (p < c.npix) {
        const uint32_t sample = uint32_t(q.counts[3]);
        const uint64_t seed = (uint64_t(c.seed_hi) << 32) | c.seed_lo;
        const rand4 r1 = rng(uint32_t(p), sample, 0u, 1u, seed);
        const int li = min(max(int(r1.u[0] * sc.num_lights), 0), sc.num_lights - 1);
        const float psel = 1.f / sc.num_lights;
        const f3 pos = sc.cam.pos;
        f3 dirL = mk(0, 0, 0);   // from the camera toward the light
        float tmaxL = 0.f, p_light = 0.f;
        f3 weight = mk(0, 0, 0);
        bool ok = false;
        if (li < sc.num_area_lights) {
            // Light_Area::sample_direct / pdf_direct from a degenerate point: G = |cos_L| / d^2
            const DevLight dl = sc.lights[li];
            f3 pL, nL;
            sample_light_point(sc, li, r1.u[1], r1.u[2], r1.u[3], pL, nL);
            const f3 wL = normalize(pos - pL);
            if (dot(wL, nL) > 0.f) {
                const f3 dd = pL - pos;
                const float d2 = dot(dd, dd);
                const float G = fabsf(dot(nL, -wL)) / d2;
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
            // Light_EnvConst::sample_direct: uniform sphere, solid-angle pdf 1 / (4 pi)
            const f3 wL = sample_uniform_sphere(r1.u[2], r1.u[3]);
            const float pL = 1.f / (4.f * kPi);
            weight = sc.env_Le / pL / psel;
            p_light = psel * pL;
            dirL = -wL;
            tmaxL = kInf - 1.f;
            ok = true;
        }
        float rx, ry;
        if (ok && cam_raster(sc.cam, dirL, rx, ry)) {
            const float J = cam_J(sc.cam, dirL);
            const float mis = c.mode == LMGPU_SAMPLING_MIS ? balance_heuristic(p_light, J) : 1.f;
            const f3 Cv = weight * (J * mis);
            if (!is_zero(Cv)) {
                emit = true;
                ray.org[0] = pos.x; ray.org[1] = pos.y; ray.org[2] = pos.z;
                ray.tmin = kEps;
                ray.dir[0] = dirL.x; ray.dir[1] = dirL.y; ray.dir[2] = dirL.z;
                ray.tmax = tmaxL;
                C = make_float4(Cv.x, Cv.y, Cv.z, __int_as_float(-(raster_pixel(sc.w, sc.h, rx, ry) + 1)));
            }
        }
    }
    const int ss = wave_append(q.counts + 2, emit);
    if (emit) {
        q.sh[ss] = ray;
        q.sh_contrib[ss] = C;
    }
}

__global__ void k_reset_counts(int32_t* counts, int out_q) {
    if (threadIdx.x == 0) {
        counts[out_q] = 0;
        counts[2] = 0;
    }
}


// Unoccluded shadow rays add their contribution to the path radiance and,
// when the vertex records guide training data, to the record's radiance
// baseline (the NEE of a vertex is not part of its sampled direction's
// incident radiance).
template <int W>
__global__ __launch_bounds__(kBlock) void k_connect(SceneDev sc, PathState ps, Queues q, unsigned long long* next,
                                                   int refill, GRec* __restrict__ rec, const int32_t* __restrict__ sh_rec,
                                                   double* __restrict__ splat) {
    __shared__ int lds[kLds][kBlock];
    const int n = q.counts[2];
    if (n == 0) return;   // no shadow rays (whole grid)
    auto fetch = [&](int64_t id, f3& o, f3& d, float& tmin, float& tmax) -> int {
        const float4* rp = reinterpret_cast<const float4*>(q.sh + id);
        const float4 a = rp[0], b = rp[1];
        o = mk(a.x, a.y, a.z);
        d = mk(b.x, b.y, b.z);
        tmin = a.w;
        tmax = b.w;
        return 0;
    };
    auto emit = [&](int64_t id, const TraceResult& r) {
        if (r.prim < 0 && !r.overflow) {
            const float4 c = q.sh_contrib[id];
            const int p = __float_as_int(c.w);
            if (p < 0) {
                // a camera-vertex light sample (image mode): straight into the film at its pixel
                double* f = splat + 3 * int64_t(-p - 1);
                atomicAdd(f + 0, double(c.x));
                atomicAdd(f + 1, double(c.y));
                atomicAdd(f + 2, double(c.z));
                return;
            }
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


// One bounce of shading over the current extension queue (shade_vertex per
// entry), appending the next extension rays and the shadow rays.
template <bool GUIDED>
__global__ __launch_bounds__(kBlock) void k_shade(SceneDev sc, PathState ps, Queues q, PTConsts c, int in_q, GuideDev gd) {
    const int n = q.counts[in_q];
    const int out_q = in_q ^ 1;
    const uint32_t sample = uint32_t(q.counts[3]);
    const uint64_t seed = (uint64_t(c.seed_hi) << 32) | c.seed_lo;
    const int stride = gridDim.x * kBlock;
    // every thread of the grid runs the same number of iterations: the
    // appends are block-wide (wave-wide for the guide's records)
    const int n_iter = (n + stride - 1) / stride;
    for (int it = 0; it < n_iter; it++) {
        const int i = it * stride + blockIdx.x * kBlock + threadIdx.x;
        VertexOut out;
        int p = 0;
        if (i < n) {
            p = q.ext_path[in_q][i];
            const float4* rp = reinterpret_cast<const float4*>(q.ext[in_q] + i);
            const float4 ra = rp[0], rb = rp[1];
            const float4 h = reinterpret_cast<const float4*>(q.hit)[i];
            if (__float_as_int(h.x) == -2) atomicOr(q.counts + 5, 1);
            const float4 thr4 = ps.thr[p], L4 = ps.L[p];
            PathRegs P{mk(thr4.x, thr4.y, thr4.z), mk(L4.x, L4.y, L4.z), ps.prev[p], ps.info[p]};
            shade_vertex<GUIDED>(sc, c, uint32_t(p), sample, seed, p, mk(ra.x, ra.y, ra.z), mk(rb.x, rb.y, rb.z), h,
                                 P, gd, GUIDED ? gd.mix + int64_t(i) * kOut : nullptr, out);
            if (out.ext) {
                ps.prev[p] = P.prev;
                ps.info[p] = P.info;
                ps.thr[p] = make_float4(P.thr.x, P.thr.y, P.thr.z, 0.f);
            }
            ps.L[p] = make_float4(P.L.x, P.L.y, P.L.z, 0.f);
        }
        int se, ss;
        block_append2<kBlock>(q.counts + out_q, q.counts + 2, out.ext, out.sh, se, ss);
        if (out.ext) {
            q.ext[out_q][se] = out.ext_ray;
            q.ext_path[out_q][se] = p;
        }
        if (out.sh) {
            q.sh[ss] = out.sh_ray;
            q.sh_contrib[ss] = make_float4(out.sh_C.x, out.sh_C.y, out.sh_C.z, __int_as_float(p));
            if (GUIDED) gd.sh_rec[ss] = out.sh_slot;
        }
        if (GUIDED) {
            const int rs = block_append1<kBlock>(gd.rec_count, out.rec_slot >= 0);
            if (out.rec_slot >= 0) gd.rec_list[rs] = out.rec_slot;
        }
    }
}

// After a pass: turns the records written in this pass (the slots listed in
// rec_list) into training records (pos, nrm, wo, wi, p_mix, f |cos| L_i
// estimate) appended to train[cap][16] when their target weight is
// positive. The list is visited in the order j * P mod n (P a prime above
// any list length, so a bijection) so that a capped batch spreads over the
// film and the path depths.
__global__ __launch_bounds__(kBlock) void k_guide_collect(PathState ps, GuideDev gd, int npix,
                                                          float* __restrict__ train, int cap,
                                                          int32_t* __restrict__ count,
                                                          unsigned long long* __restrict__ recorded) {
    constexpr int64_t kScramble = 2654435761;   // prime > 2^31
    const int64_t n = *gd.rec_count;
    const int64_t stride = int64_t(gridDim.x) * kBlock;
    const int64_t n_iter = (n + stride - 1) / stride;
    for (int64_t it = 0; it < n_iter; it++) {
        const int64_t j = it * stride + int64_t(blockIdx.x) * kBlock + threadIdx.x;
        bool keep = false;
        float out[16];
        if (j < n) {
            const int64_t slot = gd.rec_list[(j * kScramble) % n];
            const GRec r = gd.rec[slot];
            const int p = int(slot % npix);
            const float4 Lf = ps.L[p];
            const float lx = fmaxf(0.f, Lf.x - r.e.x) * r.f.x, ly = fmaxf(0.f, Lf.y - r.e.y) * r.f.y,
                        lz = fmaxf(0.f, Lf.z - r.e.z) * r.f.z;
            const float wgt = (0.2126f * lx + 0.7152f * ly + 0.0722f * lz) / r.a.w;
            if (wgt > 0.f && wgt < 3e38f) {
                keep = true;
                out[0] = r.a.x; out[1] = r.a.y; out[2] = r.a.z;
                out[3] = r.b.x; out[4] = r.b.y; out[5] = r.b.z;
                out[6] = r.c.x; out[7] = r.c.y; out[8] = r.c.z;
                out[9] = r.d.x; out[10] = r.d.y; out[11] = r.d.z;
                out[12] = r.a.w; out[13] = lx; out[14] = ly; out[15] = lz;
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
    auto fetch = [&](int64_t id, f3& o, f3& d, float& tmin, float& tmax) -> int {
        o = sc.cam.pos;
        d = dir_of(id);
        tmin = kEps;
        tmax = kInf;
        return 0;
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
                const Mat m = load_mat(sc, __float_as_int(st.p0.w), int(r.prim), b1, b2, light);
                // Material::reflectance: Kd (texture) of diffuse and mixture, Ks of glossy, 0 otherwise
                const bool refl = m.type == LMGPU_MAT_DIFFUSE || m.type == LMGPU_MAT_GLOSSY ||
                                  m.type == LMGPU_MAT_MIXTURE;
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

// Adds each path's estimate to the film. Pixel mode: to its own pixel, with
// its square. Image mode: to the pixel of its raster position (the numbers
// of camera_ray), with FP64 atomics since several paths may land there.
__global__ __launch_bounds__(kBlock) void k_accumulate(SceneDev sc, PathState ps, Queues q, PTConsts c,
                                                       double* __restrict__ sum, double* __restrict__ sumsq) {
    const int p = blockIdx.x * kBlock + threadIdx.x;
    if (p >= c.npix) return;
    const float4 L = ps.L[p];
    if (c.image) {
        if (L.x == 0.f && L.y == 0.f && L.z == 0.f) return;
        const rand4 r = rng(uint32_t(p), uint32_t(q.counts[3]), 0u, 0u, (uint64_t(c.seed_hi) << 32) | c.seed_lo);
        double* f = sum + 3 * int64_t(raster_pixel(sc.w, sc.h, r.u[2], r.u[3]));
        atomicAdd(f + 0, double(L.x));
        atomicAdd(f + 1, double(L.y));
        atomicAdd(f + 2, double(L.z));
        return;
    }
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
        if (d->mat_type[i] < 0 || d->mat_type[i] > LMGPU_MAT_MIXTURE)
            return fail(LMGPU_ERR_INVALID_ARGUMENT, "bad material type");
    // textures (texture::bitmap): validated before anything is uploaded
    const int T = d->num_textures;
    if (T < 0 || (T > 0 && (!d->tex_info || !d->tex_data)))
        return fail(LMGPU_ERR_INVALID_ARGUMENT, "bad texture list");
    int64_t texel_count = 0;
    for (int t = 0; t < T; t++) {
        const int32_t* ti = d->tex_info + 4 * t;
        if (ti[0] <= 0 || ti[1] <= 0 || (ti[2] != 3 && ti[2] != 4) || ti[3] < 0 ||
            int64_t(ti[3]) + int64_t(ti[0]) * ti[1] * ti[2] > d->tex_data_len)
            return fail(LMGPU_ERR_INVALID_ARGUMENT, "texture " + std::to_string(t) + ": bad size or offset");
        texel_count += int64_t(ti[0]) * ti[1];
    }
    bool textured = false;
    if (d->mat_tex)
        for (int i = 0; i < P; i++) {
            const int tx = d->mat_tex[i];
            if (tx < -1 || tx >= T) return fail(LMGPU_ERR_INVALID_ARGUMENT, "material texture out of range");
            if (tx >= 0 && d->mat_type[i] != LMGPU_MAT_DIFFUSE && d->mat_type[i] != LMGPU_MAT_MIXTURE)
                return fail(LMGPU_ERR_INVALID_ARGUMENT, "only diffuse and mixture materials take a texture");
            textured = textured || tx >= 0;
        }
    if (textured && !d->uvs) return fail(LMGPU_ERR_INVALID_ARGUMENT, "textured materials need texture coordinates");
    for (int i = 0; i < P; i++)
        if (d->mat_type[i] == LMGPU_MAT_MIXTURE && !d->mat_ks)
            return fail(LMGPU_ERR_INVALID_ARGUMENT, "mixture materials need mat_ks");
    LMGPU_TRY(accel.build(d->verts, N, false, opts, nullptr));
    if (!(opts && opts->refill)) accel.refill = kDefaultSceneRefill;
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
        const int type = d->mat_type[i];
        hm[i].a = make_float4(mp[0], mp[1], mp[2], float(type));
        hm[i].b = make_float4(mp[3], mp[4], mp[5], float(prim_light[i]));
        const float* ks = type == LMGPU_MAT_MIXTURE ? d->mat_ks + 3 * i : nullptr;
        int32_t tx = d->mat_tex ? d->mat_tex[i] : -1;
        // the mixture's alpha mask: a texture with an alpha channel, unless disabled
        if (tx >= 0 && type == LMGPU_MAT_MIXTURE && d->tex_info[4 * tx + 2] == 4 && mp[6] != 0.f) tx |= kTexAlphaBit;
        float txb;
        std::memcpy(&txb, &tx, 4);
        hm[i].c = make_float4(f_or(ks, 0), f_or(ks, 1), f_or(ks, 2), txb);
        if ((type == LMGPU_MAT_GLOSSY || type == LMGPU_MAT_MIXTURE) && (mp[3] <= 0.f || mp[4] <= 0.f))
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
    // texture coordinates and texels (RGBA, A = 1 for RGB textures)
    num_textures = T;
    if (textured) {
        std::vector<float2> huv(3 * size_t(N));
        for (int64_t k = 0; k < 3 * N; k++) huv[k] = make_float2(d->uvs[2 * k], d->uvs[2 * k + 1]);
        LMGPU_TRY(uvs.alloc(huv.size()));
        LMGPU_HIP(hipMemcpy(uvs.p, huv.data(), sizeof(float2) * huv.size(), hipMemcpyHostToDevice));
        std::vector<float4> htx(size_t(std::max<int64_t>(texel_count, 1)));
        std::vector<int4> hti(T);
        int64_t at = 0;
        for (int t = 0; t < T; t++) {
            const int32_t* ti = d->tex_info + 4 * t;
            const float* src = d->tex_data + ti[3];
            const int c = ti[2];
            const int64_t n = int64_t(ti[0]) * ti[1];
            for (int64_t k = 0; k < n; k++)
                htx[at + k] = make_float4(src[c * k], src[c * k + 1], src[c * k + 2], c == 4 ? src[c * k + 3] : 1.f);
            if (at > 0x7fffffff) return fail(LMGPU_ERR_INVALID_ARGUMENT, "textures too large");
            hti[t] = make_int4(ti[0], ti[1], c, int(at));
            at += n;
        }
        LMGPU_TRY(texels.alloc(htx.size()));
        LMGPU_HIP(hipMemcpy(texels.p, htx.data(), sizeof(float4) * htx.size(), hipMemcpyHostToDevice));
        LMGPU_TRY(tex_info.alloc(hti.size()));
        LMGPU_HIP(hipMemcpy(tex_info.p, hti.data(), sizeof(int4) * hti.size(), hipMemcpyHostToDevice));
    }
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
    if (accel.width == 4) dev.nodes4 = static_cast<const TravNode<4>::type*>(accel.trav_nodes());
    if (accel.width == 8) dev.nodes8 = static_cast<const TravNode<8>::type*>(accel.trav_nodes());
    dev.tris = accel.tris.p;
    dev.stri = stri.p;
    dev.mats = mats.p;
    dev.lights = lights.p;
    dev.light_cdf = light_cdf.p;
    dev.light_tri = light_tri.p;
    dev.uvs = uvs.p;
    dev.texels = texels.p;
    dev.tex_info = tex_info.p;
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
        LMGPU_TRY(grec_list.alloc(slots));
        grec_slots = int(slots);
    }
    const size_t tsz = size_t(gp.batch) * std::max(gp.steps, 1) * kRecordFloats;
    if (gp.steps > 0 && gtrain.n < tsz) {
        drop_graph();
        LMGPU_TRY(gtrain.alloc(tsz));
    }
    if (!gcount.p) {
        drop_graph();
        LMGPU_TRY(gcount.alloc(2));
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
        gd.rec_list = grec_list.p;
        gd.rec_count = gcount.p + 1;
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
    if (c.image) hipLaunchKernelGGL(k_generate<true>, dim3(gb), dim3(kBlock), 0, s, dev, ps, q, c);
    else hipLaunchKernelGGL(k_generate<false>, dim3(gb), dim3(kBlock), 0, s, dev, ps, q, c);
    if (gp && gp->rec_depth > 0) hipLaunchKernelGGL(k_zero_i32, dim3(1), dim3(64), 0, s, gcount.p + 1, 1);
    const int sg = shade_grid(npix);
    const int tg = accel.width == 4 ? pt_trace_grid<4>() : (accel.width == 8 ? pt_trace_grid<8>() : pt_trace_grid<2>());
    if (c.image && c.mode != LMGPU_SAMPLING_NAIVE) {
        // the camera vertex's light samples, traced before the first bounce
        hipLaunchKernelGGL(k_zero_i32, dim3(1), dim3(64), 0, s, counts.p + 2, 1);
        hipLaunchKernelGGL(k_camera_nee, dim3(gb), dim3(kBlock), 0, s, dev, q, c);
        hipLaunchKernelGGL(k_bounce_begin, dim3(1), dim3(64), 0, s, next.p);
        unsigned long long* nx = next.p + 1;
        if (accel.width == 4)
            hipLaunchKernelGGL(k_connect<4>, dim3(tg), dim3(kBlock), 0, s, dev, ps, q, nx, refill_primary, nullptr,
                               nullptr, sum);
        else if (accel.width == 8)
            hipLaunchKernelGGL(k_connect<8>, dim3(tg), dim3(kBlock), 0, s, dev, ps, q, nx, refill_primary, nullptr,
                               nullptr, sum);
        else
            hipLaunchKernelGGL(k_connect<2>, dim3(tg), dim3(kBlock), 0, s, dev, ps, q, nx, refill_primary, nullptr,
                               nullptr, sum);
    }
    for (int b = 0; b < c.max_verts - 1; b++) {
        const int in_q = b & 1;
        // camera rays are coherent: keep 64 consecutive rays per wave
        const int pm = b == 0 ? refill_primary : refill_secondary;
        hipLaunchKernelGGL(k_bounce_begin, dim3(1), dim3(64), 0, s, next.p);
        if (accel.width == 4) hipLaunchKernelGGL(k_extend<4>, dim3(tg), dim3(kBlock), 0, s, dev, q, in_q, next.p, pm, LMGPU_RAY_CHUNK);
        else if (accel.width == 8) hipLaunchKernelGGL(k_extend<8>, dim3(tg), dim3(kBlock), 0, s, dev, q, in_q, next.p, pm, LMGPU_RAY_CHUNK);
        else hipLaunchKernelGGL(k_extend<2>, dim3(tg), dim3(kBlock), 0, s, dev, q, in_q, next.p, pm, LMGPU_RAY_CHUNK);
        if (gp)
            hipLaunchKernelGGL(k_pt_guide_infer<Queues>, dim3(std::max(1, std::min((npix + 127) / 128, 512))), dim3(256), 0, s,
                               dev, q, in_q, gp->g->Wbf.p, gp->g->P.p, box, gmix.p);
        hipLaunchKernelGGL(k_reset_counts, dim3(1), dim3(64), 0, s, counts.p, in_q ^ 1);
        if (gp) hipLaunchKernelGGL(k_shade<true>, dim3(sg), dim3(kBlock), 0, s, dev, ps, q, c, in_q, gd);
        else hipLaunchKernelGGL(k_shade<false>, dim3(sg), dim3(kBlock), 0, s, dev, ps, q, c, in_q, gd);
        if (c.mode != LMGPU_SAMPLING_NAIVE) {
            unsigned long long* nx = next.p + 1;
            if (accel.width == 4)
                hipLaunchKernelGGL(k_connect<4>, dim3(tg), dim3(kBlock), 0, s, dev, ps, q, nx, pm, crec, cshrec, sum);
            else if (accel.width == 8)
                hipLaunchKernelGGL(k_connect<8>, dim3(tg), dim3(kBlock), 0, s, dev, ps, q, nx, pm, crec, cshrec, sum);
            else hipLaunchKernelGGL(k_connect<2>, dim3(tg), dim3(kBlock), 0, s, dev, ps, q, nx, pm, crec, cshrec, sum);
        }
    }
    hipLaunchKernelGGL(k_accumulate, dim3(gb), dim3(kBlock), 0, s, dev, ps, q, c, sum, sumsq);
    if (gp && gp->steps > 0) {
        // online training on this pass's records
        const int64_t total = int64_t(gp->rec_depth) * npix;
        const int cap = gp->batch * gp->steps;
        hipLaunchKernelGGL(k_zero_i32, dim3(1), dim3(64), 0, s, gcount.p, 1);
        hipLaunchKernelGGL(k_guide_collect, dim3(unsigned(std::min<int64_t>((total + kBlock - 1) / kBlock, 2048))),
                           dim3(kBlock), 0, s, ps, gd, npix, gtrain.p, cap, gcount.p,
                           reinterpret_cast<unsigned long long*>(counts.p + 14));
        for (int t = 0; t < gp->steps; t++)
            LMGPU_TRY(gp->g->enqueue_train(gtrain.p + int64_t(t) * gp->batch * kRecordFloats, gp->batch, gp->batch,
                                           gcount.p, int64_t(t) * gp->batch, gp->lr, nullptr, s, gp->wclamp));
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
    if (p->guiding) return fail(LMGPU_ERR_UNSUPPORTED, "guided rendering goes through lmgpu_pt_render_guided");
    if (p->primary_mode != LMGPU_PRIMARY_PIXEL && p->primary_mode != LMGPU_PRIMARY_IMAGE)
        return fail(LMGPU_ERR_INVALID_ARGUMENT, "primary_mode must be LMGPU_PRIMARY_PIXEL or LMGPU_PRIMARY_IMAGE");
    const bool image = p->primary_mode == LMGPU_PRIMARY_IMAGE;
    if (!image && p->render_time <= 0 && p->spp <= 0) return fail(LMGPU_ERR_INVALID_ARGUMENT, "spp must be positive");
    if (image && p->render_time <= 0 && p->num_samples <= 0)
        return fail(LMGPU_ERR_INVALID_ARGUMENT, "image sampling mode needs num_samples or render_time");
    if (p->sample_offset < 0) return fail(LMGPU_ERR_INVALID_ARGUMENT, "sample_offset must be >= 0");
    if (!image && p->render_time <= 0 && int64_t(p->sample_offset) + p->spp > int64_t(INT32_MAX))
        return fail(LMGPU_ERR_INVALID_ARGUMENT, "sample index beyond 2^31");
    if (image && sumsq_in)
        return fail(LMGPU_ERR_INVALID_ARGUMENT, "image sampling mode has no per-pixel sum of squares (pass NULL)");
    if (image && gp) return fail(LMGPU_ERR_UNSUPPORTED, "guided rendering uses the pixel sampling mode");
    const auto t0 = std::chrono::steady_clock::now();
    const int npix = w * h;
    // pixel-mode renders go through the regenerating wavefront: fixed-spp
    // renders of 2 or more samples, and the time scheduler in chunks of whole
    // samples per pixel, guided or not; 1-spp calls and the image mode replay
    // the one-sample pass graph
    const bool regen = !image && (p->render_time > 0 || p->spp >= 2);
    if (!regen) LMGPU_TRY(ensure_workspace(npix));
    GuidedPass gpv;
    if (gp && !regen) {
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
    c.image = image ? 1 : 0;
    if (regen) {
        int64_t ext = 0, shd = 0, grecs = 0;
        int64_t done = 0;
        if (p->render_time > 0) {
            // scheduler::spp::time: chunks of k samples per pixel until the
            // budget is spent; k is what the time left allows at the measured
            // time per sample (2 to start, at most 256), so the render ends
            // within about two samples' time of the budget, as the one-sample
            // passes end within one
            double t_spp = 0.0;
            while (true) {
                const double el = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
                if (done > 0 && el > p->render_time) break;
                int k = 2;
                if (t_spp > 0.0) k = int(std::min(256.0, std::max(2.0, std::floor((p->render_time - el) / t_spp))));
                if (int64_t(p->sample_offset) + done + k > int64_t(INT32_MAX))
                    return fail(LMGPU_ERR_INVALID_ARGUMENT, "sample index beyond 2^31");
                int64_t e1 = 0, s1 = 0, g1 = 0;
                LMGPU_TRY(render_regen(c, k, int(p->sample_offset + done), sum, sumsq, &e1, &s1, s, gp, &g1));
                ext += e1;
                shd += s1;
                grecs += g1;
                done += k;
                const double el2 = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
                t_spp = (el2 - el) / k;
            }
        } else {
            LMGPU_TRY(render_regen(c, p->spp, p->sample_offset, sum, sumsq, &ext, &shd, s, gp, &grecs));
            done = p->spp;
        }
        if (!on_device) {
            LMGPU_HIP(hipMemcpyAsync(sum_in, sum, sizeof(double) * 3 * npix, hipMemcpyDeviceToHost, s));
            if (sumsq_in) LMGPU_HIP(hipMemcpyAsync(sumsq_in, sumsq, sizeof(double) * 3 * npix, hipMemcpyDeviceToHost, s));
        }
        LMGPU_HIP(hipStreamSynchronize(s));
        if (st) {
            std::memset(st, 0, sizeof(*st));
            st->samples = int64_t(npix) * done;
            st->spp = done;
            st->extension_rays = ext;
            st->shadow_rays = shd;
            st->guide_records = grecs;
            st->elapsed_s = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
        }
        return LMGPU_OK;
    }
    // stats counters [8..15] as 64-bit: [8] samples, [10] extension rays, [12] shadow rays,
    // [14] guide training records
    LMGPU_HIP(hipMemsetAsync(counts.p + 8, 0, 8 * sizeof(int32_t), s));
    hipLaunchKernelGGL(k_set_sample, dim3(1), dim3(64), 0, s, counts.p, p->sample_offset);
    // one pass captured as a graph, replayed per sample
    const GuidedPass none;
    const GuidedPass& gk = gp ? *gp : none;
    const bool guide_same = graph_guide.g == gk.g && graph_guide.uid == gk.uid && graph_guide.gen == gk.gen &&
                            graph_guide.alpha == gk.alpha && graph_guide.lr == gk.lr && graph_guide.wclamp == gk.wclamp &&
                            graph_guide.batch == gk.batch && graph_guide.steps == gk.steps &&
                            graph_guide.rec_depth == gk.rec_depth;
    const bool key_same = graph && graph_key[0] == c.max_verts && graph_key[1] == c.mode && graph_key[2] == c.rr_depth &&
                          graph_key[3] == c.image &&
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
        graph_key[0] = c.max_verts; graph_key[1] = c.mode; graph_key[2] = c.rr_depth; graph_key[3] = c.image;
        graph_seed = p->seed;
        graph_guide = gk;
        graph_sum_ptr = sum;
        graph_sq_ptr = sumsq;
    }
    int64_t passes = 0;
    if (p->render_time > 0) {
        // scheduler::spi::time (image mode; guided rendering): passes of w*h samples until the budget is spent
        while (true) {
            LMGPU_HIP(hipGraphLaunch(graph, s));
            passes++;
            LMGPU_HIP(hipStreamSynchronize(s));
            const double el = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
            if (el > p->render_time) break;
        }
    } else if (image) {
        // scheduler::spi::sample: full passes from the graph, then the remainder as one shorter pass
        const int64_t full = p->num_samples / npix;
        const int rem = int(p->num_samples % npix);
        for (int64_t i = 0; i < full; i++) {
            LMGPU_HIP(hipGraphLaunch(graph, s));
            passes++;
        }
        if (rem > 0) {
            PTConsts cr = c;
            cr.npix = rem;
            LMGPU_TRY(enqueue_pass(cr, sum, nullptr, nullptr, s));
            passes++;
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

int lmgpu_pt_render_guided(lmgpu_scene* scene, const lmgpu_pt_params* params, lmgpu_guide* guide,
                           const lmgpu_guide_opts* opts, double* film_sum, double* film_sumsq, int film_on_device,
                           lmgpu_pt_stats* stats, void* stream) {
    if (!scene || !guide) return fail(LMGPU_ERR_INVALID_ARGUMENT, "scene and guide are required");
    lmgpu::clear_error();
    lmgpu_guide_opts o{};
    o.alpha = 0.65f;
    o.learning_rate = 1e-2f;
    o.batch = 1 << 14;
    o.steps_per_pass = 4;
    o.record_depth = 8;
    if (opts) o = *opts;
    if (!(o.alpha >= 0.f && o.alpha <= 1.f)) return fail(LMGPU_ERR_INVALID_ARGUMENT, "alpha must be in [0, 1]");
    if (!(o.learning_rate >= 0.f)) return fail(LMGPU_ERR_INVALID_ARGUMENT, "learning rate must be >= 0");
    if (o.steps_per_pass < 0 || o.steps_per_pass > 64) return fail(LMGPU_ERR_INVALID_ARGUMENT, "steps_per_pass must be in [0, 64]");
    if (o.steps_per_pass > 0 && (o.batch < 1 || o.batch > (1 << 22)))
        return fail(LMGPU_ERR_INVALID_ARGUMENT, "batch must be in [1, 2^22]");
    if (o.record_depth < 0 || o.record_depth > 16) return fail(LMGPU_ERR_INVALID_ARGUMENT, "record_depth must be in [0, 16]");
    if (!(o.weight_clamp >= 0.f)) return fail(LMGPU_ERR_INVALID_ARGUMENT, "weight_clamp must be >= 0");
    lmgpu::GuidedPass gp;
    gp.g = &guide->g;
    gp.alpha = o.alpha;
    gp.lr = o.learning_rate;
    gp.wclamp = o.weight_clamp > 0.f ? o.weight_clamp : 50.f;
    gp.steps = o.steps_per_pass;
    gp.batch = o.steps_per_pass > 0 ? (o.batch + 127) / 128 * 128 : 0;
    gp.rec_depth = o.steps_per_pass > 0 ? o.record_depth : 0;
    if (gp.rec_depth == 0) gp.steps = 0;
    const int64_t npix = int64_t(scene->s.w) * scene->s.h;
    const int64_t total = int64_t(gp.rec_depth) * npix;
    if (total >= (int64_t(1) << 31)) return fail(LMGPU_ERR_INVALID_ARGUMENT, "record_depth * pixels must be < 2^31");
    if (int64_t(gp.batch) * std::max(gp.steps, 1) >= (int64_t(1) << 31))
        return fail(LMGPU_ERR_INVALID_ARGUMENT, "batch * steps_per_pass must be < 2^31");
    return scene->s.render(params, film_sum, film_sumsq, film_on_device != 0, &gp, stats,
                           static_cast<hipStream_t>(stream));
}

}  // extern "C"
