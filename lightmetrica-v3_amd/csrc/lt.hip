// Wavefront renderer::lt: light tracing with connections to the pinhole
// camera, a restatement of Renderer_LT_NEE::render
// (src/renderer/renderer_lt.cpp:56-170) on the GPU accel.
//
//   generate  one light subpath per slot: uniform light selection
//             (scene.cpp:374-387), then light::area::sample_ray (a point on
//             the mesh by area, a cosine-weighted direction about the
//             geometric normal; light_area.cpp:78-100) or
//             light::envconst::sample_ray (a uniform direction, a point on
//             the disk of the scene's bounding sphere; light_envconst.cpp:47-74),
//             weight Le / (p_pos p_dir p_sel)
//   extend    closest hit of the light rays (persistent traversal)
//   shade     Russian roulette of the step that reached the vertex
//             (num_verts > 5, renderer_lt.cpp:147-153), component sampling,
//             the camera connection of path::sample_direct with TransDir::EL
//             (camera_pinhole.cpp:178-227: weight J(wo) G, raster position)
//             pushed to the shadow queue with its pixel, and the next
//             direction of path::sample_direction with TransDir::LE: adjoint
//             shading-normal correction (surface.h:424-438), no eta^2 on
//             refraction (material_glass.cpp:85-87)
//   connect   unoccluded camera connections are splatted into the FP64 film
//             at their pixel (Film::raster_to_pixel, film.h:166-171)
// A pass traces min(w*h, remaining) light paths; the reference's film is the
// splat sum times w*h / processed (renderer_lt.cpp:166).

#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstring>

#include "lmgpu_internal.h"
#include "pt_device.h"
#include "pt_internal.h"

namespace lmgpu {

namespace {

struct LTConsts {
    int max_verts;
    int npaths;       // light paths of this pass
    uint32_t pass;    // pass index (second Philox counter word)
    uint32_t seed_lo, seed_hi;
    f3 sph_c;         // bounding sphere of the scene (envconst rays)
    float sph_r;
};

__device__ __forceinline__ uint64_t lt_seed(const LTConsts& c) { return (uint64_t(c.seed_hi) << 32) | c.seed_lo; }

// surface::shading_normal_correction for TransDir::LE (surface.h:424-438):
// (wi.ns)(wo.ng) / ((wo.ns)(wi.ng)), 0 where a direction sees the shading and
// geometric normals with different signs.
__device__ __forceinline__ float sn_correction_le(const Geom& g, f3 wi, f3 wo) {
    const float wi_ng = dot(wi, g.gn), wo_ng = dot(wo, g.gn);
    const float wi_ns = dot(wi, g.n), wo_ns = dot(wo, g.n);
    if (wi_ng * wi_ns <= 0.f || wo_ng * wo_ns <= 0.f) return 0.f;
    return (wi_ns * wo_ng) / (wo_ns * wi_ng);
}

// Light subpath starts. Path state: thr = throughput, info.x = vertex count
// of the vertex the ray will hit. Random numbers: (path, pass, 0, 0) for the
// light selection and the position, (path, pass, 0, 1) for the direction.
__global__ __launch_bounds__(kBlock) void k_lt_generate(SceneDev sc, PathState ps, Queues q, LTConsts c) {
    const int i = blockIdx.x * kBlock + threadIdx.x;
    if (i >= c.npaths) return;
    const uint64_t seed = lt_seed(c);
    const rand4 r0 = rng(uint32_t(i), c.pass, 0u, 0u, seed);
    const rand4 r1 = rng(uint32_t(i), c.pass, 0u, 1u, seed);
    const int li = min(max(int(r0.u[0] * sc.num_lights), 0), sc.num_lights - 1);
    f3 org, d, w;
    if (li < sc.num_area_lights) {
        const DevLight dl = sc.lights[li];
        // Dist::sample: upper_bound on the light's area CDF
        const float* cdf = sc.light_cdf + dl.cdf_off;
        int lo = 0, hi = dl.count + 1;
        while (lo < hi) {
            const int mid = (lo + hi) >> 1;
            if (cdf[mid] <= r0.u[1]) lo = mid + 1;
            else hi = mid;
        }
        const int k = min(max(lo - 1, 0), dl.count - 1);
        const ShadeTri ls = sc.stri[sc.light_tri[dl.cdf_off - li + k]];
        const f3 A = mk(ls.p0.x, ls.p0.y, ls.p0.z), B = mk(ls.p1.x, ls.p1.y, ls.p1.z),
                 C = mk(ls.p2.x, ls.p2.y, ls.p2.z);
        const float s = sqrtf(fmaxf(0.f, r0.u[2]));
        const float bu = 1.f - s, bv = r0.u[3] * s;
        org = A * (1.f - bu - bv) + B * bu + C * bv;
        const f3 nL = normalize(cross(B - A, C - A));
        f3 u, v;
        onb(nL, u, v);
        const f3 dl_ = sample_cosine_weighted(r1.u[0], r1.u[1]);
        d = u * dl_.x + v * dl_.y + nL * dl_.z;
        // Le / (pA pD_projSA) / p_sel = Le A pi n_lights
        w = mk(dl.Ke_invA.x, dl.Ke_invA.y, dl.Ke_invA.z) * (kPi / dl.Ke_invA.w * float(sc.num_lights));
    } else {
        d = sample_uniform_sphere(r1.u[0], r1.u[1]);
        const float rr = sqrtf(fmaxf(0.f, r0.u[2])) * c.sph_r;
        float sn, cs;
        sincosf(2.f * kPi * r0.u[3], &sn, &cs);
        f3 u, v;
        onb(d, u, v);
        org = c.sph_c - d * c.sph_r + u * (rr * cs) + v * (rr * sn);
        // Le / (p_dir p_pos) / p_sel = Le 4 pi (pi r^2) n_lights
        w = sc.env_Le * (4.f * kPi * kPi * c.sph_r * c.sph_r * float(sc.num_lights));
    }
    ps.thr[i] = make_float4(w.x, w.y, w.z, 0.f);
    ps.info[i] = make_int4(2, 0, 0, 0);
    lmgpu_ray ray;
    ray.org[0] = org.x; ray.org[1] = org.y; ray.org[2] = org.z;
    ray.tmin = kEps;
    ray.dir[0] = d.x; ray.dir[1] = d.y; ray.dir[2] = d.z;
    ray.tmax = kInf;
    q.ext[0][i] = ray;
    q.ext_path[0][i] = i;
    if (i == 0) {
        q.counts[0] = c.npaths;
        atomicAdd(reinterpret_cast<unsigned long long*>(q.counts + 8), (unsigned long long)c.npaths);
    }
}

// One light-path vertex per extension-queue entry (the body of the loop of
// renderer_lt.cpp:92-162 for the vertex the entry's ray hit). Random numbers
// (path, pass, vertex, 0) = {RR, component, direction u0, u1}.
__global__ __launch_bounds__(kBlock) void k_lt_shade(SceneDev sc, PathState ps, Queues q, LTConsts c, int in_q) {
    const int n = q.counts[in_q];
    const int out_q = in_q ^ 1;
    const uint64_t seed = lt_seed(c);
    const int stride = gridDim.x * kBlock;
    // every lane of a wave runs the same number of iterations so the
    // wave-level appends see the whole wave
    const int n_iter = (n + stride - 1) / stride;
    for (int it = 0; it < n_iter; it++) {
        const int i = it * stride + blockIdx.x * kBlock + threadIdx.x;
        bool emit_ext = false, emit_sh = false;
        lmgpu_ray ext_ray, sh_ray;
        float4 sh_c = make_float4(0.f, 0.f, 0.f, 0.f);
        int p = 0;
        if (i < n) {
            p = q.ext_path[in_q][i];
            const float4 rb = reinterpret_cast<const float4*>(q.ext[in_q] + i)[1];
            const float4 h = reinterpret_cast<const float4*>(q.hit)[i];
            const int tri = __float_as_int(h.x);
            if (tri == -2) atomicOr(q.counts + 5, 1);
            if (tri >= 0) {
                const float4 thr4 = ps.thr[p];
                f3 thr = mk(thr4.x, thr4.y, thr4.z);
                const int nv = ps.info[p].x;
                const rand4 r0 = rng(uint32_t(p), c.pass, uint32_t(nv), 0u, seed);
                bool alive = true;
                if (nv - 1 > 5) {
                    const float qq = fmaxf(.2f, 1.f - maxc(thr));
                    if (r0.u[0] < qq) alive = false;
                    else thr = thr / (1.f - qq);
                }
                if (alive) {
                    const ShadeTri st = sc.stri[tri];
                    const float b1 = h.z, b2 = h.w, b0 = 1.f - b1 - b2;
                    const f3 P0 = mk(st.p0.x, st.p0.y, st.p0.z), P1 = mk(st.p1.x, st.p1.y, st.p1.z),
                             P2 = mk(st.p2.x, st.p2.y, st.p2.z);
                    const f3 pos = P0 * b0 + P1 * b1 + P2 * b2;
                    const f3 nsh = normalize(mk(st.n0.x, st.n0.y, st.n0.z) * b0 + mk(st.n1.x, st.n1.y, st.n1.z) * b1 +
                                             mk(st.n2.x, st.n2.y, st.n2.z) * b2);
                    const Geom g = make_on_surface(pos, nsh, normalize(cross(P1 - P0, P2 - P0)));
                    int light;
                    const Mat m = load_mat(sc, __float_as_int(st.p0.w), tri, b1, b2, light);
                    const f3 wi = mk(-rb.x, -rb.y, -rb.z);   // toward the light-side vertex
                    float cw, ul;
                    const int comp = mat_sample_component(m, g, wi, r0.u[1], cw, ul);
                    thr = thr * cw;
                    if (nv < c.max_verts) {
                        // camera connection; woc = direction from the camera to the point
                        const Camera& cam = sc.cam;
                        const f3 dc = pos - cam.pos;
                        const float d2 = dot(dc, dc);
                        const float dist = sqrtf(d2);
                        const f3 woc = dc / dist;
                        const float ez = dot(woc, cam.w);
                        if (ez < 0.f) {
                            const float rx = -dot(woc, cam.u) / ez / cam.tf / cam.aspect * .5f + .5f;
                            const float ry = -dot(woc, cam.v) / ez / cam.tf * .5f + .5f;
                            const f3 wo = -woc;
                            const float G = fabsf(dot(g.n, wo)) / d2;
                            if (rx >= 0.f && rx <= 1.f && ry >= 0.f && ry <= 1.f && G > 0.f) {
                                const float ct = -ez;
                                const float J = 1.f / (ct * ct * ct * (4.f * cam.tf * cam.tf * cam.aspect));
                                const f3 fs = mat_eval(m, g, wi, wo, comp) * sn_correction_le(g, wi, wo);
                                const f3 Cc = thr * fs * (J * G);
                                if (!is_zero(Cc)) {
                                    const int px = min(max(int(rx * float(sc.w)), 0), sc.w - 1);
                                    const int py = min(max(int(ry * float(sc.h)), 0), sc.h - 1);
                                    emit_sh = true;
                                    sh_ray.org[0] = pos.x; sh_ray.org[1] = pos.y; sh_ray.org[2] = pos.z;
                                    sh_ray.tmin = kEps;
                                    sh_ray.dir[0] = wo.x; sh_ray.dir[1] = wo.y; sh_ray.dir[2] = wo.z;
                                    sh_ray.tmax = dist * (1.f - kEps);
                                    sh_c = make_float4(Cc.x, Cc.y, Cc.z, __int_as_float(py * sc.w + px));
                                }
                            }
                        }
                    }
                    if (nv + 1 < c.max_verts) {
                        f3 wo2, sw;
                        if (mat_sample_direction(m, g, wi, comp, r0.u[2], r0.u[3], ul, wo2, sw)) {
                            if (m.type == LMGPU_MAT_GLASS && comp == 1) {
                                // refraction carries (1 - Fr) without eta^2 under TransDir::LE
                                f3 wt;
                                float eta;
                                (void)glass_Fr(m, g, wi, wt, eta);
                                sw = sw / (eta * eta);
                            }
                            thr = thr * (sw * sn_correction_le(g, wi, wo2));
                            if (!is_zero(thr)) {
                                emit_ext = true;
                                ext_ray.org[0] = pos.x; ext_ray.org[1] = pos.y; ext_ray.org[2] = pos.z;
                                ext_ray.tmin = kEps;
                                ext_ray.dir[0] = wo2.x; ext_ray.dir[1] = wo2.y; ext_ray.dir[2] = wo2.z;
                                ext_ray.tmax = kInf;
                                ps.thr[p] = make_float4(thr.x, thr.y, thr.z, 0.f);
                                ps.info[p] = make_int4(nv + 1, comp, 0, 0);
                            }
                        }
                    }
                }
            }
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
        }
    }
}

template <int W>
__global__ __launch_bounds__(kBlock) void k_lt_connect(SceneDev sc, Queues q, unsigned long long* next, int refill,
                                                       double* __restrict__ film) {
    __shared__ int lds[kLds][kBlock];
    const int n = q.counts[2];
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
            double* f = film + 3 * int64_t(__float_as_int(c.w));
            atomicAdd(f + 0, double(c.x));
            atomicAdd(f + 1, double(c.y));
            atomicAdd(f + 2, double(c.z));
        }
    };
    trace_persistent<W, true, false, kLds, kPriv, kBlock>(nodes_of<W>(sc), sc.tris, n, next, lds, fetch, emit,
                                                          q.counts + 5, nullptr, refill);
    if (blockIdx.x == 0 && threadIdx.x == 0)
        atomicAdd(reinterpret_cast<unsigned long long*>(q.counts + 12), (unsigned long long)n);
}

__global__ void k_lt_begin(unsigned long long* next, int32_t* counts, int out_q) {
    if (threadIdx.x < 2) next[threadIdx.x] = 0ull;
    if (threadIdx.x == 0) {
        counts[out_q] = 0;
        counts[2] = 0;
    }
}

template <int W>
int lt_connect_grid() {
    static int grid = 0;
    if (grid == 0) {
        int dev = 0, cus = 0, a = 0;
        (void)hipGetDevice(&dev);
        (void)hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev);
        (void)hipOccupancyMaxActiveBlocksPerMultiprocessor(&a, k_lt_connect<W>, kBlock, 0);
        grid = std::max(1, cus) * std::max(1, a);
    }
    return grid;
}

template <int W>
void lt_pass(Scene& S, const LTConsts& c, double* film, int refill, hipStream_t s) {
    PathState ps{S.thr.p, S.L.p, S.prev.p, S.info.p};
    Queues q{{S.ext0.p, S.ext1.p}, S.hit.p, {S.path0.p, S.path1.p}, S.sh.p, S.shc.p, S.counts.p};
    hipLaunchKernelGGL(k_lt_generate, dim3(unsigned((c.npaths + kBlock - 1) / kBlock)), dim3(kBlock), 0, s, S.dev, ps,
                       q, c);
    const int sg = std::max(1, std::min((c.npaths + kBlock - 1) / kBlock, 2048));
    const int tg = extend_grid<W>();
    const int cg = lt_connect_grid<W>();
    for (int b = 0; b < c.max_verts - 1; b++) {
        const int in_q = b & 1;
        hipLaunchKernelGGL(k_lt_begin, dim3(1), dim3(64), 0, s, S.next.p, S.counts.p, in_q ^ 1);
        hipLaunchKernelGGL(k_extend<W>, dim3(tg), dim3(kBlock), 0, s, S.dev, q, in_q, S.next.p, refill);
        hipLaunchKernelGGL(k_lt_shade, dim3(sg), dim3(kBlock), 0, s, S.dev, ps, q, c, in_q);
        hipLaunchKernelGGL(k_lt_connect<W>, dim3(cg), dim3(kBlock), 0, s, S.dev, q, S.next.p + 1, refill, film);
    }
}

}  // namespace

int Scene::lt_render(const lmgpu_lt_params* p, double* film_in, bool on_device, lmgpu_pt_stats* st, hipStream_t s) {
    if (!p) return fail(LMGPU_ERR_INVALID_ARGUMENT, "params is NULL");
    if (!film_in) return fail(LMGPU_ERR_INVALID_ARGUMENT, "film is NULL");
    if (p->max_verts < 2 || p->max_verts > 64) return fail(LMGPU_ERR_INVALID_ARGUMENT, "max_verts must be in [2, 64]");
    if (!(p->render_time > 0) && p->num_samples <= 0)
        return fail(LMGPU_ERR_INVALID_ARGUMENT, "num_samples must be positive");
    if (p->pass_offset < 0) return fail(LMGPU_ERR_INVALID_ARGUMENT, "pass_offset must be >= 0");
    if (dev.num_lights == 0) return fail(LMGPU_ERR_INVALID_ARGUMENT, "light tracing needs a light");
    const auto t0 = std::chrono::steady_clock::now();
    const int npix = w * h;
    LMGPU_TRY(ensure_workspace(npix));
    double* film = film_in;
    if (!on_device) {
        if (fsum.n != size_t(npix) * 3) LMGPU_TRY(fsum.alloc(size_t(npix) * 3));
        film = fsum.p;
        LMGPU_HIP(hipMemsetAsync(film, 0, sizeof(double) * 3 * size_t(npix), s));
    }
    LMGPU_HIP(hipMemsetAsync(counts.p, 0, 16 * sizeof(int32_t), s));
    LTConsts c{};
    c.max_verts = p->max_verts;
    c.seed_lo = uint32_t(p->seed);
    c.seed_hi = uint32_t(p->seed >> 32);
    // bounding sphere of the scene's AABB, radius x1.01 (light_envconst.cpp:38-42)
    const f3 lo = mk(accel.root_lo[0], accel.root_lo[1], accel.root_lo[2]);
    const f3 hi = mk(accel.root_hi[0], accel.root_hi[1], accel.root_hi[2]);
    c.sph_c = (lo + hi) * .5f;
    c.sph_r = length(hi - c.sph_c) * 1.01f;
    int64_t processed = 0, passes = 0;
    uint32_t pass = uint32_t(p->pass_offset);
    const int refill = accel.refill;
    while (true) {
        const int64_t left = p->render_time > 0 ? int64_t(npix) : p->num_samples - processed;
        if (left <= 0) break;
        c.npaths = int(std::min<int64_t>(left, npix));
        c.pass = pass++;
        if (accel.width == 4) lt_pass<4>(*this, c, film, refill, s);
        else if (accel.width == 8) lt_pass<8>(*this, c, film, refill, s);
        else lt_pass<2>(*this, c, film, refill, s);
        LMGPU_HIP(hipGetLastError());
        processed += c.npaths;
        passes++;
        if (p->render_time > 0) {
            LMGPU_HIP(hipStreamSynchronize(s));
            if (std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count() >= p->render_time) break;
        }
    }
    int32_t cnt[16];
    LMGPU_HIP(hipMemcpyAsync(cnt, counts.p, sizeof(cnt), hipMemcpyDeviceToHost, s));
    if (!on_device) LMGPU_HIP(hipMemcpyAsync(film_in, film, sizeof(double) * 3 * npix, hipMemcpyDeviceToHost, s));
    LMGPU_HIP(hipStreamSynchronize(s));
    if (cnt[5]) return fail(LMGPU_ERR_STACK_OVERFLOW, "BVH traversal stack capacity exceeded during light tracing");
    if (st) {
        std::memset(st, 0, sizeof(*st));
        int64_t v[3];
        std::memcpy(v, cnt + 8, sizeof(v));
        st->samples = processed;
        st->spp = passes;
        st->extension_rays = v[1];
        st->shadow_rays = v[2];
        st->elapsed_s = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    }
    return LMGPU_OK;
}

}  // namespace lmgpu

extern "C" int lmgpu_lt_render(lmgpu_scene* scene, const lmgpu_lt_params* params, double* film, int film_on_device,
                               lmgpu_pt_stats* stats, void* stream) {
    if (!scene) return lmgpu::fail(LMGPU_ERR_INVALID_ARGUMENT, "scene is NULL");
    lmgpu::clear_error();
    return scene->s.lt_render(params, film, film_on_device != 0, stats, static_cast<hipStream_t>(stream));
}
