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
#include "wavefront.h"

namespace lmgpu {

namespace {

struct LTConsts {
    int max_verts;
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

// Start of light path `i` of pass `pass` (renderer_lt.cpp:78-90): its first
// ray and weight. Random numbers: (i, pass, 0, 0) for the light selection and
// the position, (i, pass, 0, 1) for the direction.
__device__ __forceinline__ void lt_start(const SceneDev& sc, const LTConsts& c, uint32_t i, uint32_t pass,
                                         lmgpu_ray& ray, f3& w) {
    const uint64_t seed = lt_seed(c);
    const rand4 r0 = rng(i, pass, 0u, 0u, seed);
    const rand4 r1 = rng(i, pass, 0u, 1u, seed);
    const int li = min(max(int(r0.u[0] * sc.num_lights), 0), sc.num_lights - 1);
    f3 org, d;
    if (li < sc.num_area_lights) {
        const DevLight dl = sc.lights[li];
        f3 nL;
        sample_light_point(sc, li, r0.u[1], r0.u[2], r0.u[3], org, nL);
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
    ray.org[0] = org.x; ray.org[1] = org.y; ray.org[2] = org.z;
    ray.tmin = kEps;
    ray.dir[0] = d.x; ray.dir[1] = d.y; ray.dir[2] = d.z;
    ray.tmax = kInf;
}

// Regenerating light tracing (fixed num_samples): frees the slots whose path
// ended (state kFinishing from k_lt_shade) and gives the free slots the next
// light paths, item j being path j mod npix of pass pass0 + j / npix (the
// passes' numbering), their first rays appended to queue in_q. Counters are
// updated once per block (two sweeps over the block's slots, as k_b_regen).
template <int BLOCK>
__global__ __launch_bounds__(BLOCK) void k_lt_regen(SceneDev sc, LTConsts c, int64_t total, int npix, uint32_t pass0,
                                                    int32_t* __restrict__ ppid, int32_t* __restrict__ ppass,
                                                    int32_t* __restrict__ pstate, float4* __restrict__ pthr,
                                                    int4* __restrict__ pinfo, Queues q, int in_q, int cap,
                                                    BatchCtl* ctl) {
    constexpr int kWaves = BLOCK / 64;
    constexpr int kMaxIter = 32;
    __shared__ int s_want[kMaxIter][kWaves];
    __shared__ unsigned long long s_base;
    __shared__ int s_qbase, s_emit, s_idle;
    __shared__ unsigned int s_fin;
    if (threadIdx.x == 0)
        s_idle = __hip_atomic_load(&ctl->remaining, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) == 0;
    __syncthreads();
    if (s_idle) return;
    if (blockIdx.x == 0 && threadIdx.x == 0) atomicAdd(&ctl->iterations, 1ull);
    const int stride = gridDim.x * BLOCK;
    const int n_iter = (cap + stride - 1) / stride;   // <= kMaxIter (host)
    const int lane = __lane_id();
    const int wave = threadIdx.x >> 6;
    if (threadIdx.x == 0) s_fin = 0u;
    for (int it = 0; it < n_iter; it++) {
        const int s = it * stride + blockIdx.x * BLOCK + threadIdx.x;
        bool want = false;
        if (s < cap) {
            const int32_t st = pstate[s];
            want = st == kFinishing || st == kFree;
        }
        const unsigned long long m = __ballot(want);
        if (lane == 0) s_want[it][wave] = __popcll(m);
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        int acc = 0;
        for (int it = 0; it < n_iter; it++)
            for (int w = 0; w < kWaves; w++) {
                const int k = s_want[it][w];
                s_want[it][w] = acc;
                acc += k;
            }
        unsigned long long base = 0;
        int emit = 0, qb = 0;
        if (acc > 0) {
            base = atomicAdd(&ctl->work_next, (unsigned long long)acc);
            if (base < (unsigned long long)total) emit = int(min((unsigned long long)acc, (unsigned long long)total - base));
            if (emit > 0) qb = atomicAdd(q.counts + in_q, emit);
        }
        s_base = base;
        s_emit = emit;
        s_qbase = qb;
    }
    __syncthreads();
    const unsigned long long base = s_base;
    const int emit_n = s_emit, qbase = s_qbase;
    for (int it = 0; it < n_iter; it++) {
        const int s = it * stride + blockIdx.x * BLOCK + threadIdx.x;
        bool want = false;
        if (s < cap) {
            int32_t st = pstate[s];
            if (st == kFinishing) {
                atomicAdd(&s_fin, 1u);
                st = kFree;
                pstate[s] = kFree;
            }
            want = st == kFree;
        }
        const unsigned long long mask = __ballot(want);
        if (want) {
            const int r = s_want[it][wave] + __popcll(mask & ((1ull << lane) - 1ull));
            if (r < emit_n) {
                const unsigned long long j = base + (unsigned long long)r;
                const uint32_t i = uint32_t(j % (unsigned long long)npix);
                const uint32_t pass = pass0 + uint32_t(j / (unsigned long long)npix);
                lmgpu_ray ray;
                f3 w;
                lt_start(sc, c, i, pass, ray, w);
                pthr[s] = make_float4(w.x, w.y, w.z, 0.f);
                pinfo[s] = make_int4(2, 0, 0, 0);
                ppid[s] = int32_t(i);
                ppass[s] = int32_t(pass);
                pstate[s] = kActive;
                q.ext[in_q][qbase + r] = ray;
                q.ext_path[in_q][qbase + r] = s;
            }
        }
    }
    __syncthreads();
    if (threadIdx.x == 0 && s_fin > 0u) atomicAdd(&ctl->remaining, (unsigned long long)(-(long long)s_fin));
}

// One light-path vertex per extension-queue entry (the body of the loop of
// renderer_lt.cpp:92-162 for the vertex the entry's ray hit). Random numbers
// (path, pass, vertex, 0) = {RR, component, direction u0, u1}.
// The queue entries are slots of the regenerating loop, whose light path and
// pass are ppid / ppass; a path that ends marks its slot kFinishing.
__global__ __launch_bounds__(kBlock) void k_lt_shade(SceneDev sc, PathState ps, Queues q, LTConsts c, int in_q,
                                                    const int32_t* __restrict__ ppid,
                                                    const int32_t* __restrict__ ppass, int32_t* __restrict__ pstate) {
    const int n = q.counts[in_q];
    const int out_q = in_q ^ 1;
    const uint64_t seed = lt_seed(c);
    const int stride = gridDim.x * kBlock;
    // every thread of the grid runs the same number of iterations: the
    // appends are block-wide
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
                const uint32_t pid = uint32_t(ppid[p]);
                const uint32_t pass = uint32_t(ppass[p]);
                const rand4 r0 = rng(pid, pass, uint32_t(nv), 0u, seed);
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
        if (i < n && !emit_ext) pstate[p] = kFinishing;
        int se, ss;
        block_append2<kBlock>(q.counts + out_q, q.counts + 2, emit_ext, emit_sh, se, ss);
        if (emit_ext) {
            q.ext[out_q][se] = ext_ray;
            q.ext_path[out_q][se] = p;
        }
        if (emit_sh) {
            q.sh[ss] = sh_ray;
            q.sh_contrib[ss] = sh_c;
        }
    }
}

template <int W>
__global__ __launch_bounds__(kBlock) void k_lt_connect(SceneDev sc, Queues q, unsigned long long* next, int refill,
                                                       double* __restrict__ film, int chunk_rays) {
    __shared__ int lds[kLds][kBlock];
    const int n = q.counts[2];
    if (n == 0) return;   // no connections (whole grid)
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
                                                          q.counts + 5, nullptr, refill, chunk_rays);
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

// One iteration of the regenerating light tracer on queue parity in_q.
template <int W>
void lt_regen_iteration(Scene& S, Wavefront& wf, const LTConsts& c, int64_t total, int npix, uint32_t pass0,
                        double* film, int in_q, hipStream_t s) {
    const int cap = wf.cap;
    Queues q{{wf.ext0.p, wf.ext1.p}, wf.hit.p, {wf.path0.p, wf.path1.p}, wf.sh.p, wf.shc.p, wf.counts.p};
    PathState ps{wf.pthr.p, wf.pL.p, wf.pprev.p, wf.pinfo.p};
    constexpr int kRegenBlock = 128;
    // at most 32 strides of the pool per regenerate block
    const unsigned rg = unsigned(std::max<int64_t>(std::min<int64_t>((cap + kRegenBlock - 1) / kRegenBlock, 4096),
                                                   (cap + kRegenBlock * 32 - 1) / (kRegenBlock * 32)));
    const int sg = std::max(1, std::min((cap + kBlock - 1) / kBlock, 2048));
    const int refill = 32;
    hipLaunchKernelGGL(k_lt_regen<kRegenBlock>, dim3(rg), dim3(kRegenBlock), 0, s, S.dev, c, total, npix, pass0,
                       wf.ppix.p, wf.psample.p, wf.pstate.p, wf.pthr.p, wf.pinfo.p, q, in_q, cap, wf.ctl.p);
    hipLaunchKernelGGL(k_lt_begin, dim3(1), dim3(64), 0, s, wf.next.p, wf.counts.p, in_q ^ 1);
    hipLaunchKernelGGL(k_extend<W>, dim3(extend_grid<W>()), dim3(kBlock), 0, s, S.dev, q, in_q, wf.next.p, refill,
                       LMGPU_REGEN_RAY_CHUNK);
    hipLaunchKernelGGL(k_lt_shade, dim3(sg), dim3(kBlock), 0, s, S.dev, ps, q, c, in_q, wf.ppix.p, wf.psample.p,
                       wf.pstate.p);
    hipLaunchKernelGGL(k_lt_connect<W>, dim3(lt_connect_grid<W>()), dim3(kBlock), 0, s, S.dev, q, wf.next.p + 1, refill,
                       film, LMGPU_REGEN_RAY_CHUNK);
}

// ---------------------------------------------------------------------------
// renderer::bdpt (src/renderer/renderer_bdpt.cpp:339-411, include/lm/bidir.h)
//
// Per sample a light subpath y_0..y_{nL-1} (y_0 on an area light) and an eye
// subpath z_0..z_{nE-1} (z_0 the pinhole, its raster uniform over the image)
// of at most max_verts vertices each, no Russian roulette; every strategy
// (s, t), 2 <= s + t <= max_verts, t >= 1, weighted by the power heuristic
// over all strategies of its full path (Path::eval_mis_weight). The weights
// are accumulated along the subpaths instead of re-walking the path per
// strategy: with q_j = pE(y_j) / pL(y_j) (area pdfs of sampling y_j from the
// eye side over the light side) and u_k = pL(z_k) / pE(z_k),
//   1/w = 1 + q_{s-1}^2 (samp(s-1) + q_{s-2}^2 T(s-2)) + u_{t-1}^2 (samp'(t-1) + u_{t-2}^2 V(t-2)),
//   T(j) = samp(j) + q_{j-1}^2 T(j-1), T(0) = 1,  V(k) = samp'(k) + u_{k-1}^2 V(k-1), V(0) = 0,
// where samp(j) / samp'(k) say whether the strategy connecting the two
// vertices around that edge is samplable (both non-specular;
// Path::is_samplable_bidir) and the connection-dependent pdfs are evaluated
// at the connection. Component pdfs are symmetric and cancel in every ratio.
// Light vertices are stored (LVert) for the eye walk; the eye walk carries
// its accumulator in the path state.

struct alignas(16) LVert {
    float4 a;   // position, primitive bits (light endpoint: -1 - light index)
    float4 b;   // shading normal, triangle bits
    float4 c;   // geometric normal, barycentric b1
    float4 d;   // direction toward the previous vertex, barycentric b2
    float4 e;   // subpath weight (1 / p_comp of this vertex included), w = flags bits
    float4 f;   // pL, pL of the previous vertex, G(previous, this), T(j - 1)
};

constexpr int kFlagSpec = 1 << 8;        // this vertex's component is specular (or it cannot be connected)
constexpr int kFlagSpecPrev = 1 << 9;    // the previous vertex's component is specular (or it cannot be connected)
constexpr int kFlagInfPrev = 1 << 10;    // path state: the previous vertex is the environment, at infinity

// Samples per BDPT pass for max_verts - 1 = lstride stored light vertices
// per path: 2^k, k in [18, 21], the largest with 2^k * lstride * 96 B <= 2 GiB.
inline int64_t bdpt_pass_size(int lstride) {
    const int64_t cap = (int64_t(1) << 31) / (int64_t(lstride) * 96);
    int k = 21;
    while (k > 18 && (int64_t(1) << k) > cap) k--;
    return int64_t(1) << k;
}

struct BDConsts {
    int max_verts;
    int min_verts;    // strategies with fewer vertices are skipped (Renderer_Path_Base, renderer_bdpt.cpp:49)
    int npaths;
    uint32_t pass;
    uint32_t seed_lo, seed_hi;
    int lstride;      // LVert slots per path (max_verts - 1)
    f3 sph_c;         // bounding sphere of the scene (light::envconst rays)
    float sph_r;
};

__device__ __forceinline__ uint64_t bd_seed(const BDConsts& c) { return (uint64_t(c.seed_hi) << 32) | c.seed_lo; }

__device__ __forceinline__ Geom geom_of(const LVert& v) {
    return make_on_surface(mk(v.a.x, v.a.y, v.a.z), mk(v.b.x, v.b.y, v.b.z), mk(v.c.x, v.c.y, v.c.z));
}

// light::envconst's path::pdf_primary_ray: p_sel / (4 pi) / (pi r^2)
// (light_envconst.cpp:70-74, path.h:154-160).
__device__ __forceinline__ float env_pdf_ray(const SceneDev& sc, const BDConsts& c) {
    return 1.f / (float(sc.num_lights) * 4.f * kPi * kPi * c.sph_r * c.sph_r);
}

// Light subpath start with the uniform light selection: light::area::sample_ray
// (light_area.cpp:78-100), or light::envconst::sample_ray (light_envconst.cpp:47-68:
// a uniform direction d and a point of the disk of the scene's bounding
// sphere facing it). y_0 is stored, the ray goes to the queue. The
// environment's y_0 lies at infinity and is not connectable
// (Light_EnvConst::is_connectable): its (d, disk point) sample carries the
// joint pdf p_ray of y_0 and the direction to y_1 (bidir.h:248-258, 399-411),
// so pL(y_0) = 1 and pL(y_1) = p_ray |cos| at y_1.
__global__ __launch_bounds__(kBlock) void k_bd_lgen(SceneDev sc, PathState ps, Queues q, BDConsts c,
                                                    LVert* __restrict__ lv, int32_t* __restrict__ lcount) {
    const int i = blockIdx.x * kBlock + threadIdx.x;
    if (i >= c.npaths) return;
    const uint64_t seed = bd_seed(c);
    const rand4 r0 = rng(uint32_t(i), c.pass, 0u, 0u, seed);
    const rand4 r1 = rng(uint32_t(i), c.pass, 0u, 1u, seed);
    const int li = min(max(int(r0.u[0] * sc.num_lights), 0), sc.num_lights - 1);
    if (i == 0) {
        q.counts[0] = c.npaths;
        atomicAdd(reinterpret_cast<unsigned long long*>(q.counts + 8), (unsigned long long)c.npaths);
    }
    if (li >= sc.num_area_lights) {
        const f3 d = sample_uniform_sphere(r1.u[0], r1.u[1]);
        const float rr = sqrtf(fmaxf(0.f, r0.u[2])) * c.sph_r;
        float sn, cs;
        sincosf(2.f * kPi * r0.u[3], &sn, &cs);
        f3 u, v;
        onb(d, u, v);
        const f3 org = c.sph_c - d * c.sph_r + u * (rr * cs) + v * (rr * sn);
        const float pray = env_pdf_ray(sc, c);
        LVert y;
        y.a = make_float4(org.x, org.y, org.z, __int_as_float(-1 - li));
        y.b = make_float4(d.x, d.y, d.z, __int_as_float(-1));
        y.c = make_float4(d.x, d.y, d.z, 0.f);
        y.d = make_float4(0.f, 0.f, 0.f, 0.f);
        y.e = make_float4(1.f, 1.f, 1.f, __int_as_float(kFlagSpec));   // never connected
        y.f = make_float4(1.f, 0.f, 0.f, 0.f);
        lv[int64_t(i) * c.lstride] = y;
        lcount[i] = 1;
        // weight toward y_1: Le / p_ray (bidir.h:254-256)
        const f3 w = sc.env_Le * (1.f / pray);
        ps.thr[i] = make_float4(w.x, w.y, w.z, 1.f);          // pL(y_0) = 1
        ps.L[i] = make_float4(org.x, org.y, org.z, pray);     // pL(y_1) = p_ray G(y_0, y_1)
        ps.prev[i] = make_float4(1.f, 1.f, 0.f, 0.f);         // cos 1 at infinity, T(0) = 1
        // strategy s = 1 is not samplable: y_1 sees its predecessor as specular
        ps.info[i] = make_int4(1, kFlagSpec | kFlagInfPrev, 0, 0);
        lmgpu_ray ray;
        ray.org[0] = org.x; ray.org[1] = org.y; ray.org[2] = org.z;
        ray.tmin = kEps;
        ray.dir[0] = d.x; ray.dir[1] = d.y; ray.dir[2] = d.z;
        ray.tmax = kInf;
        q.ext[0][i] = ray;
        q.ext_path[0][i] = i;
        return;
    }
    const DevLight dl = sc.lights[li];
    f3 org, nL;
    sample_light_point(sc, li, r0.u[1], r0.u[2], r0.u[3], org, nL);
    f3 u, v;
    onb(nL, u, v);
    const f3 dl_ = sample_cosine_weighted(r1.u[0], r1.u[1]);
    const f3 d = u * dl_.x + v * dl_.y + nL * dl_.z;
    // pL(y_0) = p_sel / area (path::pdf_position of the light endpoint)
    const float pL0 = dl.Ke_invA.w / float(sc.num_lights);
    const float a0 = 1.f / pL0;
    LVert y;
    y.a = make_float4(org.x, org.y, org.z, __int_as_float(-1 - li));
    y.b = make_float4(nL.x, nL.y, nL.z, __int_as_float(-1));
    y.c = make_float4(nL.x, nL.y, nL.z, 0.f);
    y.d = make_float4(0.f, 0.f, 0.f, 0.f);
    y.e = make_float4(a0, a0, a0, __int_as_float(0));
    y.f = make_float4(pL0, 0.f, 0.f, 0.f);
    lv[int64_t(i) * c.lstride] = y;
    lcount[i] = 1;
    // weight toward y_1: alpha_0 Le / pdf_dir, pdf_dir = 1 / pi (projected solid angle)
    const f3 w = mk(dl.Ke_invA.x, dl.Ke_invA.y, dl.Ke_invA.z) * (a0 * kPi);
    ps.thr[i] = make_float4(w.x, w.y, w.z, pL0);
    ps.L[i] = make_float4(org.x, org.y, org.z, 1.f / kPi);
    ps.prev[i] = make_float4(fabsf(dot(nL, d)), 1.f, 0.f, 0.f);   // |cos| at y_0, T(0) = 1
    ps.info[i] = make_int4(1, 0, 0, 0);                          // index of the next vertex, flags
    lmgpu_ray ray;
    ray.org[0] = org.x; ray.org[1] = org.y; ray.org[2] = org.z;
    ray.tmin = kEps;
    ray.dir[0] = d.x; ray.dir[1] = d.y; ray.dir[2] = d.z;
    ray.tmax = kInf;
    q.ext[0][i] = ray;
    q.ext_path[0][i] = i;
    // strategy (s, t) = (1, 1): y_0 seen by the pinhole; the only other
    // strategy of this two-vertex path is the eye ray hitting the light (s = 0)
    const Camera& cam = sc.cam;
    const f3 dc = org - cam.pos;
    const float dc2 = dot(dc, dc);
    const float dist = sqrtf(dc2);
    const f3 woc = dc / dist;
    float rx, ry;
    if (c.min_verts <= 2 && cam_raster(cam, woc, rx, ry) && dot(-woc, nL) > 0.f) {
        const float G = fabsf(dot(nL, woc)) / dc2;
        const float J = cam_J(cam, woc);
        const float q0 = J * G / pL0;
        const f3 Cw = mk(dl.Ke_invA.x, dl.Ke_invA.y, dl.Ke_invA.z) * (a0 * G * J / (1.f + q0 * q0));
        const int px = min(max(int(rx * float(sc.w)), 0), sc.w - 1);
        const int py = min(max(int(ry * float(sc.h)), 0), sc.h - 1);
        const int ss = atomicAdd(q.counts + 2, 1);
        lmgpu_ray sr;
        sr.org[0] = org.x; sr.org[1] = org.y; sr.org[2] = org.z;
        sr.tmin = kEps;
        sr.dir[0] = -woc.x; sr.dir[1] = -woc.y; sr.dir[2] = -woc.z;
        sr.tmax = dist * (1.f - kEps);
        q.sh[ss] = sr;
        q.sh_contrib[ss] = make_float4(Cw.x, Cw.y, Cw.z, __int_as_float(py * sc.w + px));
    }
}

// A light subpath vertex y_j (j >= 1): store it, connect it to the camera
// (strategy t = 1) and sample the next direction (path::sample_direction, LE).
__global__ __launch_bounds__(kBlock) void k_bd_lshade(SceneDev sc, PathState ps, Queues q, BDConsts c, int in_q,
                                                      LVert* __restrict__ lv, int32_t* __restrict__ lcount) {
    const int n = q.counts[in_q];
    const int out_q = in_q ^ 1;
    const uint64_t seed = bd_seed(c);
    const int stride = gridDim.x * kBlock;
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
                const float4 thr4 = ps.thr[p], prv = ps.L[p], acc = ps.prev[p];
                const int4 info = ps.info[p];
                const int j = info.x;
                const bool spec_prev = info.y & kFlagSpec;
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
                const f3 wdir = mk(rb.x, rb.y, rb.z);
                const f3 wi = -wdir;
                const f3 pprev = mk(prv.x, prv.y, prv.z);
                const f3 dv = pos - pprev;
                const float d2 = (info.y & kFlagInfPrev) ? 1.f : dot(dv, dv);   // no distance to infinity
                const float Gin = acc.x * fabsf(dot(nsh, wdir)) / d2;   // G(y_{j-1}, y_j)
                const float pL = prv.w * Gin;                           // area pdf from the light side
                const rand4 r0 = rng(uint32_t(p), c.pass, uint32_t(j), 0u, seed);
                float cw, ul;
                const int comp = mat_sample_component(m, g, wi, r0.u[1], cw, ul);
                const f3 alpha = mk(thr4.x, thr4.y, thr4.z) * cw;
                const bool spec = is_specular(m, comp);
                const float Tprev = acc.y;   // T(j - 1)
                LVert y;
                y.a = make_float4(pos.x, pos.y, pos.z, __int_as_float(prim));
                y.b = make_float4(nsh.x, nsh.y, nsh.z, __int_as_float(tri));
                y.c = make_float4(gn.x, gn.y, gn.z, b1);
                y.d = make_float4(wi.x, wi.y, wi.z, b2);
                y.e = make_float4(alpha.x, alpha.y, alpha.z,
                                  __int_as_float(comp | (spec ? kFlagSpec : 0) | (spec_prev ? kFlagSpecPrev : 0)));
                y.f = make_float4(pL, thr4.w, Gin, Tprev);
                lv[int64_t(p) * c.lstride + j] = y;
                lcount[p] = j + 1;
                // strategy (s, t) = (j + 1, 1): connect y_j to the pinhole (path::sample_direct, EL)
                if (!spec && !is_zero(alpha) && j + 2 >= c.min_verts) {
                    const Camera& cam = sc.cam;
                    const f3 dc = pos - cam.pos;
                    const float dc2 = dot(dc, dc);
                    const float dist = sqrtf(dc2);
                    const f3 woc = dc / dist;
                    float rx, ry;
                    if (cam_raster(cam, woc, rx, ry)) {
                        const f3 wo = -woc;
                        const float G = fabsf(dot(nsh, wo)) / dc2;
                        const float J = cam_J(cam, woc);
                        const f3 fs = mat_eval(m, g, wi, wo, comp) * sn_correction_le(g, wi, wo);
                        const f3 Cc = alpha * fs * (J * G);
                        if (!is_zero(Cc) && G > 0.f) {
                            // MIS: q_j with pE(y_j) = J G; q_{j-1} through y_j's BSDF toward the camera
                            const float qj = J * G / pL;
                            const float pEprev = mat_pdf_full(m, g, wo, wi, comp) * Gin;
                            const float qp = pEprev / thr4.w;
                            const float sampj = spec_prev ? 0.f : 1.f;
                            const float sumL = qj * qj * (sampj + qp * qp * Tprev);
                            const float wgt = 1.f / (1.f + sumL);
                            // a grazing edge can leave a zero pdf: no weight, no NaN
                            emit_sh = wgt >= 0.f && wgt <= 1.f;
                            const int px = min(max(int(rx * float(sc.w)), 0), sc.w - 1);
                            const int py = min(max(int(ry * float(sc.h)), 0), sc.h - 1);
                            sh_ray.org[0] = pos.x; sh_ray.org[1] = pos.y; sh_ray.org[2] = pos.z;
                            sh_ray.tmin = kEps;
                            sh_ray.dir[0] = wo.x; sh_ray.dir[1] = wo.y; sh_ray.dir[2] = wo.z;
                            sh_ray.tmax = dist * (1.f - kEps);
                            const f3 Cw = Cc * wgt;
                            sh_c = make_float4(Cw.x, Cw.y, Cw.z, __int_as_float(py * sc.w + px));
                        }
                    }
                }
                // next light vertex y_{j+1} (usable while j + 2 <= max_verts - 1)
                if (j + 2 <= c.max_verts - 1 && !is_zero(alpha)) {
                    f3 wo2, sw;
                    if (mat_sample_direction(m, g, wi, comp, r0.u[2], r0.u[3], ul, wo2, sw)) {
                        const f3 f = mat_eval_full(m, g, wi, wo2, comp, true) * sn_correction_le(g, wi, wo2);
                        const float pd = mat_pdf_full(m, g, wi, wo2, comp);
                        const f3 a2 = pd > 0.f ? alpha * f / pd : mk(0, 0, 0);
                        if (!is_zero(a2)) {
                            // T(j) = samp(j) + q_{j-1}^2 T(j-1), q_{j-1} through y_j toward y_{j+1}
                            const float pEprev = mat_pdf_full(m, g, wo2, wi, comp) * Gin;
                            const float qp = pEprev / thr4.w;
                            const float Tj = ((spec_prev || spec) ? 0.f : 1.f) + qp * qp * Tprev;
                            ps.thr[p] = make_float4(a2.x, a2.y, a2.z, pL);
                            ps.L[p] = make_float4(pos.x, pos.y, pos.z, pd);
                            ps.prev[p] = make_float4(fabsf(dot(nsh, wo2)), Tj, 0.f, 0.f);
                            ps.info[p] = make_int4(j + 1, spec ? kFlagSpec : 0, 0, 0);
                            emit_ext = true;
                            ext_ray.org[0] = pos.x; ext_ray.org[1] = pos.y; ext_ray.org[2] = pos.z;
                            ext_ray.tmin = kEps;
                            ext_ray.dir[0] = wo2.x; ext_ray.dir[1] = wo2.y; ext_ray.dir[2] = wo2.z;
                            ext_ray.tmax = kInf;
                        }
                    }
                }
            }
        }
        int se, ss;
        block_append2<kBlock>(q.counts + out_q, q.counts + 2, emit_ext, emit_sh, se, ss);
        if (emit_ext) {
            q.ext[out_q][se] = ext_ray;
            q.ext_path[out_q][se] = p;
        }
        if (emit_sh) {
            q.sh[ss] = sh_ray;
            q.sh_contrib[ss] = sh_c;
        }
    }
}

// Eye subpath start: camera::pinhole::sample_ray (camera_pinhole.cpp:134-140),
// a raster position uniform over the image from (sample, pass, 0, 2)[0..1].
__global__ __launch_bounds__(kBlock) void k_bd_egen(SceneDev sc, PathState ps, Queues q, BDConsts c) {
    const int i = blockIdx.x * kBlock + threadIdx.x;
    if (i >= c.npaths) return;
    const rand4 r = rng(uint32_t(i), c.pass, 0u, 2u, bd_seed(c));
    const f3 d = primary_dir(sc.cam, r.u[0], r.u[1]);
    const int px = min(max(int(r.u[0] * float(sc.w)), 0), sc.w - 1);
    const int py = min(max(int(r.u[1] * float(sc.h)), 0), sc.h - 1);
    // z_0: pE = 1 (pdf_position), the sampled direction's pdf is J; V(0) = 0
    ps.thr[i] = make_float4(1.f, 1.f, 1.f, 1.f);
    ps.L[i] = make_float4(sc.cam.pos.x, sc.cam.pos.y, sc.cam.pos.z, cam_J(sc.cam, d));
    ps.prev[i] = make_float4(1.f, 0.f, 0.f, 0.f);
    ps.info[i] = make_int4(1, 0, py * sc.w + px, 0);
    lmgpu_ray ray;
    ray.org[0] = sc.cam.pos.x; ray.org[1] = sc.cam.pos.y; ray.org[2] = sc.cam.pos.z;
    ray.tmin = kEps;
    ray.dir[0] = d.x; ray.dir[1] = d.y; ray.dir[2] = d.z;
    ray.tmax = kInf;
    q.ext[0][i] = ray;
    q.ext_path[0][i] = i;
    if (i == 0) q.counts[0] = c.npaths;
}

// An eye subpath vertex z_{t-1} (t >= 2): the emitter strategy (s = 0),
// the connections to the sample's light vertices (s >= 1) queued as shadow
// rays with their weighted contributions, and the next direction.
__global__ __launch_bounds__(kBlock) void k_bd_eshade(SceneDev sc, PathState ps, Queues q, BDConsts c, int in_q,
                                                      const LVert* __restrict__ lv, const int32_t* __restrict__ lcount,
                                                      double* __restrict__ film) {
    const int n = q.counts[in_q];
    const int out_q = in_q ^ 1;
    const uint64_t seed = bd_seed(c);
    const int stride = gridDim.x * kBlock;
    const int n_iter = (n + stride - 1) / stride;
    for (int it = 0; it < n_iter; it++) {
        const int i = it * stride + blockIdx.x * kBlock + threadIdx.x;
        bool emit_ext = false;
        lmgpu_ray ext_ray;
        int p = 0;
        bool active = false;
        // vertex state shared by the connection loop below
        Geom g;
        Mat m;
        f3 wi = mk(0, 0, 0), alphaE = mk(0, 0, 0);
        float Gin = 0.f, pEcur = 0.f, pEprev = 1.f, Vprev = 0.f;
        int comp = 0, t = 0, nl = 0, pixel = 0;
        bool spec_prev = false, spec = true;
        if (i < n) {
            p = q.ext_path[in_q][i];
            const float4 rb = reinterpret_cast<const float4*>(q.ext[in_q] + i)[1];
            const float4 h = reinterpret_cast<const float4*>(q.hit)[i];
            const int tri = __float_as_int(h.x);
            if (tri == -2) atomicOr(q.counts + 5, 1);
            if (tri >= 0) {
                const float4 thr4 = ps.thr[p], prv = ps.L[p], acc = ps.prev[p];
                const int4 info = ps.info[p];
                t = info.x + 1;   // z_{t-1} is this vertex
                spec_prev = info.y & kFlagSpec;
                pixel = info.z;
                const ShadeTri st = sc.stri[tri];
                const float b1 = h.z, b2 = h.w, b0 = 1.f - b1 - b2;
                const f3 P0 = mk(st.p0.x, st.p0.y, st.p0.z), P1 = mk(st.p1.x, st.p1.y, st.p1.z),
                         P2 = mk(st.p2.x, st.p2.y, st.p2.z);
                const f3 pos = P0 * b0 + P1 * b1 + P2 * b2;
                const f3 nsh = normalize(mk(st.n0.x, st.n0.y, st.n0.z) * b0 + mk(st.n1.x, st.n1.y, st.n1.z) * b1 +
                                         mk(st.n2.x, st.n2.y, st.n2.z) * b2);
                g = make_on_surface(pos, nsh, normalize(cross(P1 - P0, P2 - P0)));
                int light;
                m = load_mat(sc, __float_as_int(st.p0.w), tri, b1, b2, light);
                const f3 wdir = mk(rb.x, rb.y, rb.z);
                wi = -wdir;
                const f3 dv = pos - mk(prv.x, prv.y, prv.z);
                Gin = acc.x * fabsf(dot(nsh, wdir)) / dot(dv, dv);   // G(z_{t-2}, z_{t-1})
                pEcur = prv.w * Gin;
                pEprev = thr4.w;
                Vprev = acc.y;
                const f3 thr = mk(thr4.x, thr4.y, thr4.z);
                // s = 0: the eye subpath hits a light (z_{t-1} becomes the light endpoint)
                if (light >= 0 && dot(wi, nsh) > 0.f && t >= c.min_verts) {
                    const DevLight dl = sc.lights[light];
                    const f3 Le = mk(dl.Ke_invA.x, dl.Ke_invA.y, dl.Ke_invA.z);
                    const float u1 = (dl.Ke_invA.w / float(sc.num_lights)) / pEcur;   // pL(z_{t-1}) / pE
                    const float u2 = (1.f / kPi) * Gin / pEprev;                      // pL(z_{t-2}) / pE
                    const float sumE = u1 * u1 * ((spec_prev ? 0.f : 1.f) + u2 * u2 * Vprev);
                    const float wgt = 1.f / (1.f + sumE);
                    const f3 C = thr * Le * wgt;
                    if (!is_zero(C) && wgt >= 0.f && wgt <= 1.f) {
                        double* f = film + 3 * int64_t(pixel);
                        atomicAdd(f + 0, double(C.x));
                        atomicAdd(f + 1, double(C.y));
                        atomicAdd(f + 2, double(C.z));
                    }
                }
                const rand4 r0 = rng(uint32_t(p), c.pass, uint32_t(t - 1), 2u, seed);
                float cw, ul;
                comp = mat_sample_component(m, g, wi, r0.u[1], cw, ul);
                alphaE = thr * cw;
                spec = is_specular(m, comp);
                nl = lcount[p];
                active = !is_zero(alphaE);
                // next eye vertex z_t (while t + 1 <= max_verts)
                if (active && t + 1 <= c.max_verts) {
                    f3 wo2, sw;
                    if (mat_sample_direction(m, g, wi, comp, r0.u[2], r0.u[3], ul, wo2, sw)) {
                        const f3 f = mat_eval_full(m, g, wi, wo2, comp, false) * sn_correction(g, wi, wo2);
                        const float pd = mat_pdf_full(m, g, wi, wo2, comp);
                        const f3 a2 = pd > 0.f ? alphaE * f / pd : mk(0, 0, 0);
                        if (!is_zero(a2)) {
                            // V(t-1) = samp'(t-1) + u_{t-2}^2 V(t-2), u_{t-2} through z_{t-1} toward z_t
                            const float u2 = mat_pdf_full(m, g, wo2, wi, comp) * Gin / pEprev;
                            const float Vt = ((spec_prev || spec) ? 0.f : 1.f) + u2 * u2 * Vprev;
                            ps.thr[p] = make_float4(a2.x, a2.y, a2.z, pEcur);
                            ps.L[p] = make_float4(pos.x, pos.y, pos.z, pd);
                            ps.prev[p] = make_float4(fabsf(dot(nsh, wo2)), Vt, 0.f, 0.f);
                            ps.info[p] = make_int4(t, spec ? kFlagSpec : 0, pixel, 0);
                            emit_ext = true;
                            ext_ray.org[0] = pos.x; ext_ray.org[1] = pos.y; ext_ray.org[2] = pos.z;
                            ext_ray.tmin = kEps;
                            ext_ray.dir[0] = wo2.x; ext_ray.dir[1] = wo2.y; ext_ray.dir[2] = wo2.z;
                            ext_ray.tmax = kInf;
                        }
                    }
                }
            } else if (tri == -1 && sc.has_env) {
                // s = 0 on light::envconst: the eye subpath leaves the scene and
                // ends on the environment (bidir.h:525-533), z_{t-1} at infinity.
                // Strategy s = 1 is not samplable (the environment cannot be
                // connected), so 1/w = 1 + u1^2 u2^2 V(t-2) with
                // u1 = pL(z_{t-1}) / pE(z_{t-1}) = 1 / (pdf(z_{t-2}) |cos|) and
                // u2 = pL(z_{t-2}) / pE(z_{t-2}) = p_ray |cos| / pE(z_{t-2}).
                const float4 thr4 = ps.thr[p], prv = ps.L[p], acc = ps.prev[p];
                const int4 info = ps.info[p];
                const int te = info.x + 1;
                if (te >= c.min_verts) {
                    const float u12 = env_pdf_ray(sc, c) / (prv.w * thr4.w);
                    const float sumE = acc.y > 0.f ? u12 * u12 * acc.y : 0.f;
                    const float wgt = 1.f / (1.f + sumE);
                    const f3 C = mk(thr4.x, thr4.y, thr4.z) * sc.env_Le * wgt;
                    if (!is_zero(C) && wgt >= 0.f && wgt <= 1.f) {
                        double* f = film + 3 * int64_t(info.z);
                        atomicAdd(f + 0, double(C.x));
                        atomicAdd(f + 1, double(C.y));
                        atomicAdd(f + 2, double(C.z));
                    }
                }
            }
        }
        const int se = block_append1<kBlock>(q.counts + out_q, emit_ext);
        if (emit_ext) {
            q.ext[out_q][se] = ext_ray;
            q.ext_path[out_q][se] = p;
        }
        // connections to light vertices y_j, s = j + 1 <= max_verts - t (every thread
        // runs the same loop: the appends are block-wide)
        const bool can = active && !spec;
        for (int j = 0; j < c.lstride; j++) {
            bool emit_sh = false;
            lmgpu_ray sh_ray;
            float4 sh_c;
            if (can && j < nl && j + 1 + t <= c.max_verts && j + 1 + t >= c.min_verts) {
                const LVert y = lv[int64_t(p) * c.lstride + j];
                const int yflags = __float_as_int(y.e.w);
                if (!(yflags & kFlagSpec)) {
                    const f3 ypos = mk(y.a.x, y.a.y, y.a.z);
                    const f3 yn = mk(y.b.x, y.b.y, y.b.z);
                    const f3 dv = ypos - g.p;
                    const float d2 = dot(dv, dv);
                    const float dist = sqrtf(d2);
                    const f3 dir = dv / dist;   // z -> y
                    const float G = fabsf(dot(g.n, dir)) * fabsf(dot(yn, dir)) / d2;
                    const f3 fsE = mat_eval(m, g, wi, dir, comp) * sn_correction(g, wi, dir);
                    f3 fsL;
                    float pLz;   // pL(z_{t-1}): y's pdf toward z, to area
                    float pEy_prev = 0.f;   // pE(y_{j-1}) through y toward z
                    if (j == 0) {
                        const int li = -1 - __float_as_int(y.a.w);
                        const float4 ke = sc.lights[li].Ke_invA;
                        const bool front = dot(-dir, yn) > 0.f;
                        fsL = front ? mk(ke.x, ke.y, ke.z) : mk(0, 0, 0);
                        pLz = front ? (1.f / kPi) * G : 0.f;
                    } else {
                        int ylight;
                        const Mat ym = load_mat(sc, __float_as_int(y.a.w), __float_as_int(y.b.w), y.c.w, y.d.w, ylight);
                        const Geom yg = geom_of(y);
                        const f3 ywi = mk(y.d.x, y.d.y, y.d.z);
                        const int ycomp = yflags & 0xff;
                        fsL = mat_eval(ym, yg, ywi, -dir, ycomp) * sn_correction_le(yg, ywi, -dir);
                        pLz = mat_pdf_full(ym, yg, ywi, -dir, ycomp) * G;
                        pEy_prev = mat_pdf_full(ym, yg, -dir, ywi, ycomp) * y.f.z;
                    }
                    const f3 C = mk(y.e.x, y.e.y, y.e.z) * fsL * G * fsE * alphaE;
                    if (!is_zero(C) && G > 0.f) {
                        // light side: q_j, q_{j-1}, T(j-1)
                        const float qj = mat_pdf_full(m, g, wi, dir, comp) * G / y.f.x;
                        float inner = 1.f;   // samp(0) for j = 0
                        if (j > 0) {
                            const float qp = pEy_prev / y.f.y;
                            inner = ((yflags & kFlagSpecPrev) ? 0.f : 1.f) + qp * qp * y.f.w;
                        }
                        const float sumL = qj * qj * inner;
                        // eye side: u_{t-1}, u_{t-2}, V(t-2)
                        const float u1 = pLz / pEcur;
                        const float u2 = mat_pdf_full(m, g, dir, wi, comp) * Gin / pEprev;
                        const float sumE = u1 * u1 * ((spec_prev ? 0.f : 1.f) + u2 * u2 * Vprev);
                        const float wgt = 1.f / (1.f + sumL + sumE);
                        const f3 Cw = C * wgt;
                        emit_sh = wgt >= 0.f && wgt <= 1.f;
                        // shadow ray from the light vertex, as Scene::visible(vL, vE)
                        sh_ray.org[0] = ypos.x; sh_ray.org[1] = ypos.y; sh_ray.org[2] = ypos.z;
                        sh_ray.tmin = kEps;
                        sh_ray.dir[0] = -dir.x; sh_ray.dir[1] = -dir.y; sh_ray.dir[2] = -dir.z;
                        sh_ray.tmax = dist * (1.f - kEps);
                        sh_c = make_float4(Cw.x, Cw.y, Cw.z, __int_as_float(pixel));
                    }
                }
            }
            const int ss = block_append1<kBlock>(q.counts + 2, emit_sh);
            if (emit_sh) {
                q.sh[ss] = sh_ray;
                q.sh_contrib[ss] = sh_c;
            }
        }
    }
}

template <int W>
void bd_pass(Scene& S, const BDConsts& c, double* film, int refill, LVert* lv, hipStream_t s) {
    PathState ps{S.thr.p, S.L.p, S.prev.p, S.info.p};
    Queues q{{S.ext0.p, S.ext1.p}, S.hit.p, {S.path0.p, S.path1.p}, S.bsh.p, S.bshc.p, S.counts.p};
    const unsigned gb = unsigned((c.npaths + kBlock - 1) / kBlock);
    const int sg = std::max(1, std::min((c.npaths + kBlock - 1) / kBlock, 2048));
    const int tg = extend_grid<W>();
    const int cg = lt_connect_grid<W>();
    hipLaunchKernelGGL(k_lt_begin, dim3(1), dim3(64), 0, s, S.next.p, S.counts.p, 1);
    hipLaunchKernelGGL(k_bd_lgen, dim3(gb), dim3(kBlock), 0, s, S.dev, ps, q, c, lv, S.blcount.p);
    hipLaunchKernelGGL(k_lt_connect<W>, dim3(cg), dim3(kBlock), 0, s, S.dev, q, S.next.p + 1, refill, film,
                           LMGPU_RAY_CHUNK);
    for (int b = 0; b < c.max_verts - 2; b++) {   // y_1 .. y_{max_verts - 2}
        const int in_q = b & 1;
        hipLaunchKernelGGL(k_lt_begin, dim3(1), dim3(64), 0, s, S.next.p, S.counts.p, in_q ^ 1);
        hipLaunchKernelGGL(k_extend<W>, dim3(tg), dim3(kBlock), 0, s, S.dev, q, in_q, S.next.p, refill, LMGPU_RAY_CHUNK);
        hipLaunchKernelGGL(k_bd_lshade, dim3(sg), dim3(kBlock), 0, s, S.dev, ps, q, c, in_q, lv, S.blcount.p);
        hipLaunchKernelGGL(k_lt_connect<W>, dim3(cg), dim3(kBlock), 0, s, S.dev, q, S.next.p + 1, refill, film,
                           LMGPU_RAY_CHUNK);
    }
    hipLaunchKernelGGL(k_bd_egen, dim3(gb), dim3(kBlock), 0, s, S.dev, ps, q, c);
    for (int b = 0; b < c.max_verts - 1; b++) {   // z_1 .. z_{max_verts - 1}
        const int in_q = b & 1;
        hipLaunchKernelGGL(k_lt_begin, dim3(1), dim3(64), 0, s, S.next.p, S.counts.p, in_q ^ 1);
        hipLaunchKernelGGL(k_extend<W>, dim3(tg), dim3(kBlock), 0, s, S.dev, q, in_q, S.next.p, refill, LMGPU_RAY_CHUNK);
        hipLaunchKernelGGL(k_bd_eshade, dim3(sg), dim3(kBlock), 0, s, S.dev, ps, q, c, in_q, lv, S.blcount.p, film);
        hipLaunchKernelGGL(k_lt_connect<W>, dim3(cg), dim3(kBlock), 0, s, S.dev, q, S.next.p + 1, refill, film,
                           LMGPU_RAY_CHUNK);
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
    // the loop's queues, counters and path state live in the Wavefront pool;
    // the per-pixel buffers of the one-sample passes are not needed
    double* film = film_in;
    if (!on_device) {
        if (fsum.n != size_t(npix) * 3) LMGPU_TRY(fsum.alloc(size_t(npix) * 3));
        film = fsum.p;
        LMGPU_HIP(hipMemsetAsync(film, 0, sizeof(double) * 3 * size_t(npix), s));
    }
    LTConsts c{};
    c.max_verts = p->max_verts;
    c.seed_lo = uint32_t(p->seed);
    c.seed_hi = uint32_t(p->seed >> 32);
    // bounding sphere of the scene's AABB, radius x1.01 (light_envconst.cpp:38-42)
    const f3 lo = mk(accel.root_lo[0], accel.root_lo[1], accel.root_lo[2]);
    const f3 hi = mk(accel.root_hi[0], accel.root_hi[1], accel.root_hi[2]);
    c.sph_c = (lo + hi) * .5f;
    c.sph_r = length(hi - c.sph_c) * 1.01f;
    // scheduler::spi::sample and spi::time: the regenerating loop; a slot
    // whose light path ends takes the next one (path i of pass k as in the
    // passes), so the long paths of a pass no longer hold up the next pass.
    // The time scheduler runs it in chunks of whole passes (k passes of w*h
    // paths, k what the time left allows at the measured time per pass, at
    // most 256), on consecutive pass indices.
    if (!wf) {
        wf = new (std::nothrow) Wavefront();
        if (!wf) return fail(LMGPU_ERR_OUT_OF_MEMORY, "wavefront");
    }
    const int W = accel.width;
    int64_t ext = 0, shd = 0;
    auto run_loop = [&](int64_t total, uint32_t pass0) -> int {
        LMGPU_TRY(wf->ensure_pool(Wavefront::pool_size(total, npix)));
        auto iteration = [&](hipStream_t cs, int in_q) {
            if (W == 4) lt_regen_iteration<4>(*this, *wf, c, total, npix, pass0, film, in_q, cs);
            else if (W == 8) lt_regen_iteration<8>(*this, *wf, c, total, npix, pass0, film, in_q, cs);
            else lt_regen_iteration<2>(*this, *wf, c, total, npix, pass0, film, in_q, cs);
        };
        double sec = 0;
        int64_t iters = 0;
        LMGPU_TRY(wf->loop(iteration, total, p->max_verts, &sec, &iters, s));
        int32_t cnt[16];
        LMGPU_HIP(hipMemcpy(cnt, wf->counts.p, sizeof(cnt), hipMemcpyDeviceToHost));
        int64_t v[3];
        std::memcpy(v, cnt + 8, sizeof(v));
        ext += v[1];
        shd += v[2];
        return LMGPU_OK;
    };
    int64_t total = 0, passes = 0;
    if (!(p->render_time > 0)) {
        total = p->num_samples;
        if (int64_t(p->pass_offset) + (total - 1) / npix > int64_t(UINT32_MAX))
            return fail(LMGPU_ERR_INVALID_ARGUMENT, "pass index beyond 2^32");
        LMGPU_TRY(run_loop(total, uint32_t(p->pass_offset)));
        passes = (total + npix - 1) / npix;
    } else {
        double t_pass = 0.0;
        while (true) {
            const double el = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
            if (passes > 0 && el >= p->render_time) break;
            int k = 1;
            if (t_pass > 0.0) k = int(std::min(256.0, std::max(1.0, std::floor((p->render_time - el) / t_pass))));
            if (int64_t(p->pass_offset) + passes + k > int64_t(UINT32_MAX))
                return fail(LMGPU_ERR_INVALID_ARGUMENT, "pass index beyond 2^32");
            LMGPU_TRY(run_loop(int64_t(k) * npix, uint32_t(p->pass_offset + passes)));
            passes += k;
            total += int64_t(k) * npix;
            const double el2 = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
            t_pass = (el2 - el) / k;
        }
    }
    if (!on_device) LMGPU_HIP(hipMemcpyAsync(film_in, film, sizeof(double) * 3 * npix, hipMemcpyDeviceToHost, s));
    LMGPU_HIP(hipStreamSynchronize(s));
    if (st) {
        std::memset(st, 0, sizeof(*st));
        st->samples = total;
        st->spp = passes;
        st->extension_rays = ext;
        st->shadow_rays = shd;
        st->elapsed_s = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    }
    return LMGPU_OK;
}

int Scene::bdpt_render(const lmgpu_lt_params* p, double* film_in, bool on_device, lmgpu_pt_stats* st, hipStream_t s) {
    if (!p) return fail(LMGPU_ERR_INVALID_ARGUMENT, "params is NULL");
    if (!film_in) return fail(LMGPU_ERR_INVALID_ARGUMENT, "film is NULL");
    if (p->max_verts < 2 || p->max_verts > 32) return fail(LMGPU_ERR_INVALID_ARGUMENT, "max_verts must be in [2, 32]");
    if (!(p->render_time > 0) && p->num_samples <= 0)
        return fail(LMGPU_ERR_INVALID_ARGUMENT, "num_samples must be positive");
    if (p->pass_offset < 0) return fail(LMGPU_ERR_INVALID_ARGUMENT, "pass_offset must be >= 0");
    if (dev.num_lights == 0) return fail(LMGPU_ERR_INVALID_ARGUMENT, "bidirectional path tracing needs a light");
    const auto t0 = std::chrono::steady_clock::now();
    const int npix = w * h;
    LMGPU_TRY(ensure_workspace(npix));
    // samples per pass: the largest power of two in [2^18, 2^21] whose stored
    // light vertices (P (max_verts - 1) LVerts of 96 B) fit in 2 GiB, at most
    // w*h; the connection queue has as many entries
    const int lstride = p->max_verts - 1;
    const int P = int(std::min<int64_t>(npix, bdpt_pass_size(lstride)));
    const size_t nv = size_t(P) * lstride;
    if (blverts.n < nv * sizeof(LVert) / sizeof(float4)) LMGPU_TRY(blverts.alloc(nv * sizeof(LVert) / sizeof(float4)));
    if (bsh.n < nv) {
        LMGPU_TRY(bsh.alloc(nv));
        LMGPU_TRY(bshc.alloc(nv));
    }
    if (blcount.n < size_t(P)) LMGPU_TRY(blcount.alloc(P));
    double* film = film_in;
    if (!on_device) {
        if (fsum.n != size_t(npix) * 3) LMGPU_TRY(fsum.alloc(size_t(npix) * 3));
        film = fsum.p;
        LMGPU_HIP(hipMemsetAsync(film, 0, sizeof(double) * 3 * size_t(npix), s));
    }
    LMGPU_HIP(hipMemsetAsync(counts.p, 0, 16 * sizeof(int32_t), s));
    BDConsts c{};
    c.max_verts = p->max_verts;
    c.min_verts = p->min_verts > 0 ? p->min_verts : 2;
    c.seed_lo = uint32_t(p->seed);
    c.seed_hi = uint32_t(p->seed >> 32);
    c.lstride = lstride;
    // bounding sphere of the scene's AABB, radius x1.01 (light_envconst.cpp:38-42)
    const f3 blo = mk(accel.root_lo[0], accel.root_lo[1], accel.root_lo[2]);
    const f3 bhi = mk(accel.root_hi[0], accel.root_hi[1], accel.root_hi[2]);
    c.sph_c = (blo + bhi) * .5f;
    c.sph_r = length(bhi - c.sph_c) * 1.01f;
    LVert* lv = reinterpret_cast<LVert*>(blverts.p);
    int64_t processed = 0, passes = 0;
    uint32_t pass = uint32_t(p->pass_offset);
    const int refill = accel.refill;
    while (true) {
        const int64_t left = p->render_time > 0 ? int64_t(P) : p->num_samples - processed;
        if (left <= 0) break;
        c.npaths = int(std::min<int64_t>(left, P));
        c.pass = pass++;
        if (accel.width == 4) bd_pass<4>(*this, c, film, refill, lv, s);
        else if (accel.width == 8) bd_pass<8>(*this, c, film, refill, lv, s);
        else bd_pass<2>(*this, c, film, refill, lv, s);
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
    if (cnt[5]) return fail(LMGPU_ERR_STACK_OVERFLOW, "BVH traversal stack capacity exceeded during BDPT");
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

extern "C" int lmgpu_bdpt_render(lmgpu_scene* scene, const lmgpu_lt_params* params, double* film, int film_on_device,
                                 lmgpu_pt_stats* stats, void* stream) {
    if (!scene) return lmgpu::fail(LMGPU_ERR_INVALID_ARGUMENT, "scene is NULL");
    lmgpu::clear_error();
    return scene->s.bdpt_render(params, film, film_on_device != 0, stats, static_cast<hipStream_t>(stream));
}

extern "C" int lmgpu_lt_render(lmgpu_scene* scene, const lmgpu_lt_params* params, double* film, int film_on_device,
                               lmgpu_pt_stats* stats, void* stream) {
    if (!scene) return lmgpu::fail(LMGPU_ERR_INVALID_ARGUMENT, "scene is NULL");
    lmgpu::clear_error();
    return scene->s.lt_render(params, film, film_on_device != 0, stats, static_cast<hipStream_t>(stream));
}
