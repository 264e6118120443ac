/*
 * lmgpu — C ABI of the MI355X (gfx950) path-tracing hot path.
 *
 * A drop-in replacement for Lightmetrica's `accel::*` (behind lm::Accel) and
 * `renderer::pt` (behind lm::Renderer), with lm::Film output unchanged.
 * Plain pointers, sizes and POD structs only; no exceptions cross this
 * boundary. Every entry point returns an int status (LMGPU_OK = 0, < 0 on
 * error) and leaves a message in lmgpu_last_error().
 *
 * Reference interfaces replaced (paths are relative to the reference repo):
 *   lm::Accel::build(const Scene&)                  include/lm/accel.h:38
 *   lm::Accel::intersect(Ray, Float, Float)         include/lm/accel.h:70
 *   lm::Accel::Hit {t, uv, global_transform, primitive, face}
 *                                                   include/lm/accel.h:49-55
 *   lm::Scene::visible(sp1, sp2)                    include/lm/scene.h:343-364
 *   lm::Renderer::render()  (renderer::pt)          include/lm/renderer.h:25,
 *                                                   src/renderer/renderer_pt.cpp:80-315
 *   lm::Film::splat / rescale / buffer              include/lm/film.h:103-205
 *   error model LM_THROW_EXCEPTION(Error::...)      include/lm/exception.h:142-150
 */
#ifndef LMGPU_H
#define LMGPU_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- status codes (mirror lm::Error, include/lm/exception.h) ------------- */
#define LMGPU_OK 0
#define LMGPU_ERR_INVALID_ARGUMENT (-1)   /* Error::InvalidArgument */
#define LMGPU_ERR_UNSUPPORTED (-2)        /* Error::Unsupported */
#define LMGPU_ERR_FAILED_TO_RENDER (-3)   /* Error::FailedToRender */
#define LMGPU_ERR_DEVICE (-4)             /* HIP runtime / kernel error */
#define LMGPU_ERR_OUT_OF_MEMORY (-5)
#define LMGPU_ERR_STACK_OVERFLOW (-6)     /* traversal stack capacity exceeded */

/* Message of the last failing call on this thread ("" if none). */
const char* lmgpu_last_error(void);

/* Library version string. */
const char* lmgpu_version(void);

/* Number of visible HIP devices (0 when no GPU). Does not create a context. */
int lmgpu_device_count(int* count);

/* ---- rays and hits --------------------------------------------------------
 * One ray = 32 B, one hit = 16 B, both 16-B aligned so a wave of 64 lanes
 * reads 2 KiB / writes 1 KiB with full-width vector accesses.
 * Ray: reference lm::Ray {Vec3 o, d} plus the [tmin, tmax] range of
 * Accel::intersect (include/lm/math.h:67-77, include/lm/accel.h:70), in FP32
 * (FP64 -> FP32 at the boundary, as plugin/accel_embree/accel_embree.cpp:135-143).
 * Hit: the global triangle index (flattened order of Accel_SAHBVH::build,
 * src/accel/accel_sahbvh.cpp:132-151) or -1, t, and the barycentric (u, v)
 * weights of the 2nd and 3rd vertex (Accel::Hit::uv). The host maps the
 * triangle index to (primitive node, face) with the tables it flattened. */
typedef struct lmgpu_ray {
    float org[3];
    float tmin;
    float dir[3];
    float tmax;
} lmgpu_ray;

typedef struct lmgpu_hit {
    int32_t prim;
    float t;
    float u;
    float v;
} lmgpu_hit;

/* ---- accel ------------------------------------------------------------------ */
typedef struct lmgpu_accel lmgpu_accel;

typedef struct lmgpu_build_opts {
    int32_t width;       /* BVH branching factor: 2, 4 or 8 (0 = 4) */
    int32_t leaf_size;   /* max triangles per leaf, 1..15 (0 = 2) */
    int32_t refill;      /* traversal: a wave takes new rays once this many of its
                            64 lanes are idle, 1..64 (0 = 8 for lmgpu_accel_*, 32 for a
                            scene's renderers; 64 suits coherent rays) */
    int32_t builder;     /* LMGPU_BUILDER_LBVH (1) or LMGPU_BUILDER_PLOC (0); with opts == NULL: LBVH */
    int32_t ploc_radius; /* PLOC nearest-neighbour window, 1..64 (0 = 16) */
    int32_t collapse;    /* LMGPU_COLLAPSE_SAH (0) or LMGPU_COLLAPSE_GREEDY (1) */
    float sah_tri_cost;  /* SAH collapse: cost of a triangle test relative to a wide-node visit (0 = 0.5) */
    int32_t reserved[1];
} lmgpu_build_opts;

/* Binary tree -> W-wide tree. SAH: the collapse of least SAH cost over all
 * ways of cutting the binary tree into W-wide nodes and leaves of at most
 * leaf_size triangles (a bottom-up dynamic program, Ylitie, Karras and Laine,
 * HPG 2017), which also decides where leaves go. GREEDY: each wide node opens
 * its largest-area child until it has W children; every subtree of at most
 * leaf_size triangles is a leaf. */
#define LMGPU_COLLAPSE_SAH 0
#define LMGPU_COLLAPSE_GREEDY 1

#define LMGPU_BUILDER_PLOC 0  /* locally-ordered clustering over Morton order */
#define LMGPU_BUILDER_LBVH 1  /* Karras radix tree over Morton order */

typedef struct lmgpu_accel_info {
    int64_t num_tris;
    int64_t num_nodes;       /* wide nodes */
    int64_t num_leaves;
    int32_t width;
    int32_t leaf_size;
    int32_t depth;           /* wide-tree depth (root = 1) */
    int32_t stack_capacity;  /* traversal stack entries available per ray */
    int64_t node_bytes;      /* exact nodes (32 * width B each) plus, for width 4 and 8, their
                                quantised traversal copy (16 * width B each) */
    int64_t tri_bytes;
    double sah_cost;         /* SAH cost of the wide tree (node=1, tri=1 units) */
    double build_ms;         /* device time of the last build */
} lmgpu_accel_info;

int lmgpu_accel_create(lmgpu_accel** out);
int lmgpu_accel_destroy(lmgpu_accel* accel);

/* Builds the BVH over `num_tris` triangles given as [num_tris][9] floats
 * (p1, p2, p3 world-space positions, FP32), replacing
 * Accel_SAHBVH::build (src/accel/accel_sahbvh.cpp:127-267).
 * `verts_on_device` != 0: `verts` is a device pointer, else host memory.
 * `stream` is a hipStream_t (NULL = default stream). Synchronous on return. */
int lmgpu_accel_build(lmgpu_accel* accel, const float* verts, int64_t num_tris, int verts_on_device,
                      const lmgpu_build_opts* opts, void* stream);

int lmgpu_accel_info_get(const lmgpu_accel* accel, lmgpu_accel_info* info);

/* Closest hit for `n` rays, device pointers; enqueued on `stream`, asynchronous.
 * Batched form of Accel_SAHBVH::intersect (src/accel/accel_sahbvh.cpp:269-299). */
int lmgpu_accel_intersect(const lmgpu_accel* accel, const lmgpu_ray* rays, lmgpu_hit* hits, int64_t n,
                          void* stream);

/* Same, host pointers; copies in and out and synchronizes. Serves the
 * single-ray lm::Accel::intersect of a host component (n = 1). */
int lmgpu_accel_intersect_host(const lmgpu_accel* accel, const lmgpu_ray* rays, lmgpu_hit* hits,
                               int64_t n);

/* Any-hit occlusion for `n` rays (device pointers): occluded[i] = 1 when a
 * triangle is hit in [tmin, tmax], else 0. Equals !Scene::intersect(ray,
 * tmin, tmax) as used by Scene::visible (include/lm/scene.h:343-364). */
int lmgpu_accel_occluded(const lmgpu_accel* accel, const lmgpu_ray* rays, uint8_t* occluded, int64_t n,
                         void* stream);

/* Allocation-free form of lmgpu_accel_intersect_host. No counterpart in the
 * reference: it is for callers that cannot allocate device memory (a
 * process with HIP graph capture in progress, tight serving loops issuing
 * small ray batches, harnesses that forbid hipMalloc). Host pointers in and
 * out; the rays, hits, ray counter and overflow word are staged through one
 * static device array per GPU (LMGPU_NOALLOC_SCRATCH_BYTES, held until the
 * process ends), so nothing is allocated or freed. At most
 * LMGPU_NOALLOC_MAX_RAYS rays per call: a larger n returns
 * LMGPU_ERR_INVALID_ARGUMENT before any device work. Every call on one GPU
 * shares that GPU's array; calls from threads of one process are serialized
 * by a per-device lock (overlapping uses of the array would race).
 * Synchronous; the results equal lmgpu_accel_intersect_host's. */
#define LMGPU_NOALLOC_MAX_RAYS 65536
#define LMGPU_NOALLOC_SCRATCH_BYTES (LMGPU_NOALLOC_MAX_RAYS * 48 + 256)
int lmgpu_accel_intersect_host_noalloc(const lmgpu_accel* accel, const lmgpu_ray* rays, lmgpu_hit* hits,
                                       int64_t n);

/* Synchronizes `stream`; returns LMGPU_ERR_STACK_OVERFLOW if a traversal on
 * this accel exceeded its stack since the last call. Cannot happen when
 * depth * (width - 1) <= stack_capacity (see lmgpu_accel_info). */
int lmgpu_accel_status(const lmgpu_accel* accel, void* stream);

/* Closest hit plus traversal counters summed over the rays (node visits,
 * triangle tests). Synchronous. */
int lmgpu_accel_intersect_counted(const lmgpu_accel* accel, const lmgpu_ray* rays, lmgpu_hit* hits,
                                  int64_t n, int64_t* node_visits, int64_t* tri_tests, void* stream);

/* Serialized form of a built accel, the counterpart of the reference's
 * serialization of accel::sahbvh (src/accel/accel_sahbvh.cpp:122-124,
 * cereal; own format): a 128-byte header ("LMGPUBVH", version, width, leaf
 * size, depth, counts, SAH cost, bounds) followed by the nodes and the
 * triangles as laid out on the device. With buf == NULL, *size receives the
 * byte count; else *size is the buffer size on input and the bytes written
 * on output. */
int lmgpu_accel_serialize(const lmgpu_accel* accel, void* buf, size_t* size);
/* Replaces the accel's BVH with a serialized one, validated on the host
 * (header, sizes, tree structure, leaf ranges, triangle ids) before upload;
 * opts may give the traversal refill threshold. */
int lmgpu_accel_deserialize(lmgpu_accel* accel, const void* buf, size_t size, const lmgpu_build_opts* opts);

/* ---- scene + renderer::pt ------------------------------------------------------ */

/* Material type codes: material::diffuse, ::glossy, ::mirror, ::glass, and
 * material::mixture_wavefrontobj (src/model/model_wavefrontobj.cpp:340-513: a
 * diffuse Kd (+ map_Kd) and glossy Ks/ax/ay group picked per direction by
 * max(Kd) / (max(Kd) + max(Ks)), plus an alpha-mask pass-through component
 * when the diffuse texture has an alpha channel). */
#define LMGPU_MAT_DIFFUSE 0
#define LMGPU_MAT_GLOSSY 1
#define LMGPU_MAT_MIRROR 2
#define LMGPU_MAT_GLASS 3
#define LMGPU_MAT_MIXTURE 4

/* Flattened scene with identity transforms (world-space vertices). */
typedef struct lmgpu_scene_desc {
    int64_t num_tris;
    const float* verts;        /* [N][9] positions p1,p2,p3 (host) */
    const float* normals;      /* [N][9] vertex normals n1,n2,n3 (host) */
    const int32_t* tri_prim;   /* [N] primitive (scene node) index */
    int32_t num_prims;
    const int32_t* mat_type;   /* [P] LMGPU_MAT_* */
    const float* mat_param;    /* [P][8]: color rgb (Kd of diffuse and mixture, Ks of glossy), ax, ay, Ni,
                                  mixture: alpha mask enabled (0 = no_alpha_mask), - */
    const float* prim_Ke;      /* [P][3] light::area radiance (unused if not a light) */
    int32_t num_area_lights;
    const int32_t* light_prims;/* [L] primitives carrying light::area, in light order */
    int32_t has_env;           /* light::envconst present (light index L) */
    float env_Le[3];
    float cam_pos[3], cam_center[3], cam_up[3];  /* camera::pinhole */
    float vfov;                /* degrees */
    float aspect;              /* width / height */
    int32_t w, h;              /* film size */
    /* material::mixture_wavefrontobj and texture::bitmap (all optional) */
    const float* mat_ks;       /* [P][3] Ks of LMGPU_MAT_MIXTURE materials (NULL when there is none) */
    const int32_t* mat_tex;    /* [P] texture of the material's Kd (map_Kd; diffuse and mixture), -1 = none */
    const float* uvs;          /* [N][6] texture coordinates t1,t2,t3 (Mesh::surface_point's t) */
    int32_t num_textures;
    const int32_t* tex_info;   /* [T][4] width, height, channels (3 or 4), offset into tex_data (floats) */
    const float* tex_data;     /* texels as texture::bitmap holds them: row-major, row 0 = the first row
                                  after its vertical flip, `channels` floats per texel */
    int64_t tex_data_len;      /* floats in tex_data */
} lmgpu_scene_desc;

typedef struct lmgpu_scene lmgpu_scene;

/* Uploads the scene and builds its accel (opts may be NULL). */
int lmgpu_scene_create(lmgpu_scene** out, const lmgpu_scene_desc* desc, const lmgpu_build_opts* opts);
int lmgpu_scene_destroy(lmgpu_scene* scene);
/* Borrowed pointer to the scene's accel (owned by the scene). */
int lmgpu_scene_accel(lmgpu_scene* scene, lmgpu_accel** accel);

/* renderer::pt parameters (src/renderer/renderer_pt.cpp:51-77, scheduler.cpp:32-35,72-75). */
#define LMGPU_SAMPLING_NAIVE 0
#define LMGPU_SAMPLING_NEE 1
#define LMGPU_SAMPLING_MIS 2

typedef struct lmgpu_pt_params {
    int32_t spp;              /* scheduler::spp::sample: samples per pixel */
    int32_t max_verts;        /* max path vertices */
    int32_t sampling_mode;    /* LMGPU_SAMPLING_* */
    int32_t rr_depth;         /* Russian roulette when num_verts > rr_depth (reference 5) */
    uint64_t seed;            /* Philox key */
    int32_t sample_offset;    /* index of the first sample (progressive rendering) */
    int32_t guiding;          /* must be 0: guided rendering is lmgpu_pt_render_guided */
    double render_time;       /* > 0: scheduler::spp::time, render whole samples per pixel for this many
                                 seconds (pixel mode: the regenerating loop in chunks of samples; image
                                 mode: passes of w*h samples) */
    int32_t primary_mode;     /* LMGPU_PRIMARY_PIXEL or LMGPU_PRIMARY_IMAGE (primary_ray_sampling_mode) */
    int32_t reserved0;
    int64_t num_samples;      /* image mode, scheduler::spi::sample: samples over the whole image */
    int32_t reserved[4];
} lmgpu_pt_params;

/* primary_ray_sampling_mode of renderer::pt (renderer_pt.cpp:60-76, 96-107,
 * 133-136, 307-312).
 * PIXEL (scheduler::spp::*): sample s of pixel p jitters its raster position
 * inside pixel p and adds its estimate to pixel p; film value = film_sum / spp.
 * IMAGE (scheduler::spi::*): `num_samples` samples (or passes of w*h samples for
 * `render_time` seconds) in passes of min(w*h, remaining); sample i of pass k
 * (k counted from sample_offset) draws the Philox stream (i, k) and takes a
 * raster position uniform over the whole image; its estimate lands in the
 * pixel of that position. The camera vertex also samples a light (NEE from
 * the camera), whose contribution lands in the pixel of the light's raster
 * position. Film value = film_sum * w*h / stats.samples. IMAGE takes no
 * film_sumsq (a sample adds to two pixels) and no guide; spp is ignored. */
#define LMGPU_PRIMARY_PIXEL 0
#define LMGPU_PRIMARY_IMAGE 1

typedef struct lmgpu_pt_stats {
    int64_t samples;          /* path samples traced */
    int64_t spp;              /* samples per pixel accumulated */
    int64_t extension_rays;
    int64_t shadow_rays;
    double elapsed_s;         /* wall time of the render call */
    int64_t guide_records;    /* guided rendering: training records collected */
    int64_t reserved[3];
} lmgpu_pt_stats;

/* Accumulates, per pixel, the sum and the sum of squares of the per-sample
 * estimate (FP64, [h][w][3], y = 0 the bottom row, as Film_Bitmap) into
 * `film_sum` / `film_sumsq` (either may be NULL... sum is required).
 * `film_on_device` != 0: device pointers, else host pointers (copied back on
 * return). The film value of renderer::pt is film_sum / spp
 * (renderer_pt.cpp:307-308). Synchronous. */
int lmgpu_pt_render(lmgpu_scene* scene, const lmgpu_pt_params* params, double* film_sum, double* film_sumsq,
                    int film_on_device, lmgpu_pt_stats* stats, void* stream);

/* renderer::raycast (src/renderer/renderer_raycast.cpp:48-97): one primary
 * ray through every pixel centre; film[h][w][3] (FP64, y = 0 the bottom row)
 * is set to bg_color on a miss (the envconst radiance when the scene has
 * one), |shading normal| with visualize_normal, else the reflectance (Kd of
 * diffuse, Ks of glossy, 0 for mirror and glass; `color` when use_color) times
 * 0.2 + 0.8 |n . -d| unless use_constant_color. Synchronous for host films. */
typedef struct lmgpu_raycast_params {
    float bg_color[3];
    float color[3];
    int32_t use_color;
    int32_t use_constant_color;
    int32_t visualize_normal;
    int32_t reserved[5];
} lmgpu_raycast_params;

int lmgpu_raycast_render(lmgpu_scene* scene, const lmgpu_raycast_params* params, double* film, int film_on_device,
                         void* stream);

/* renderer::lt (src/renderer/renderer_lt.cpp:22-170, Renderer_LT_NEE::render):
 * light subpaths from the scene's lights (area: light_area.cpp:78-100,
 * envconst: light_envconst.cpp:47-74) connected to the pinhole camera at every
 * surface vertex, splatted into film[h][w][3] (FP64, y = 0 the bottom row;
 * device films accumulate, host films are overwritten). `num_samples` light
 * paths (scheduler::spi::sample); with render_time > 0, whole passes of w*h
 * paths until that many seconds have passed. Both run in a regenerating
 * wavefront loop: a slot whose light path ends takes the next one. Item j is
 * light path j mod (w*h) of pass j / (w*h), so a render equals the passes it
 * spans. The reference's film value is film * (w*h) / stats->samples
 * (renderer_lt.cpp:166); stats->spp counts the passes (begun ones included).
 * Rejects renders whose pass index would pass 2^32. Light paths of pass k
 * (pass_offset + k) draw Philox counters (path, pass, vertex, block), so
 * progressive calls continue with pass_offset += the previous spp. Paths that
 * hit the emitter directly seen by the camera (two vertices) carry no
 * contribution, as in the reference. Synchronous. */
typedef struct lmgpu_lt_params {
    int64_t num_samples;
    int32_t max_verts;        /* max path vertices (camera and light included) */
    int32_t pass_offset;
    uint64_t seed;
    double render_time;
    int32_t min_verts;        /* renderer::bdpt: skip strategies of fewer vertices (0 = 2); unused by LT */
    int32_t reserved[7];
} lmgpu_lt_params;

int lmgpu_lt_render(lmgpu_scene* scene, const lmgpu_lt_params* params, double* film, int film_on_device,
                    lmgpu_pt_stats* stats, void* stream);

/* renderer::bdpt (src/renderer/renderer_bdpt.cpp:339-411, include/lm/bidir.h):
 * per sample an eye subpath (pinhole, raster uniform over the image) and a
 * light subpath (area lights, or light::envconst's disk-and-direction sample,
 * light_envconst.cpp:47-74) of at most max_verts vertices each, without
 * Russian roulette; every connection strategy (s, t), 2 <= s + t <= max_verts,
 * weighted by the power heuristic over all strategies of its full path
 * (Path::eval_mis_weight) and splatted at its raster position. Same film and
 * scheduler contract as lmgpu_lt_render (film value = film * (w*h) /
 * stats->samples; stats->spp counts the passes, of min(w*h, 2^k) samples with
 * 2^k the largest power of two in [2^18, 2^21] for which 2^k (max_verts - 1)
 * stored light vertices of 96 B fit in 2 GiB).
 * An environment light is a non-connectable light endpoint (bidir.h:245-258,
 * 399-411, 585-588): light subpaths may start on it and eye subpaths may end
 * on it (strategy s = 0), but no connection reaches it.
 * max_verts in [2, 32]. Synchronous. */
int lmgpu_bdpt_render(lmgpu_scene* scene, const lmgpu_lt_params* params, double* film, int film_on_device,
                      lmgpu_pt_stats* stats, void* stream);

/* ---- batched rendering ------------------------------------------------------
 * Many renderer::pt jobs (each its own scene, film size, spp and seed) in
 * one wavefront loop over a pool of path slots fed from a device work queue;
 * the jobs' BVHs share one device arena. A job renders what lmgpu_pt_render
 * renders for the same scene, parameters and BVH width. No counterpart in
 * the reference (its renderers run one scene at a time). */
typedef struct lmgpu_batch lmgpu_batch;

typedef struct lmgpu_job_desc {
    const lmgpu_scene_desc* scene;   /* copied by lmgpu_batch_add_job */
    int32_t spp;
    int32_t max_verts;
    int32_t sampling_mode;
    int32_t rr_depth;
    uint64_t seed;
    int32_t reserved[4];
} lmgpu_job_desc;

typedef struct lmgpu_batch_opts {
    int32_t width;          /* BVH width 2, 4, 8; 0 = autotune */
    int32_t path_capacity;  /* paths in flight (queue capacity), 1024..2^24; 0 = autotune */
    int32_t block;          /* block size of the generate / shade kernels, 128 or 256; 0 = autotune */
    int32_t leaf_size;      /* 0 = 2 */
    int32_t film_sumsq;     /* 1: also accumulate per-pixel sums of squares (lmgpu_batch_job_film) */
    int32_t queue_order;    /* LMGPU_QUEUE_JOB_MAJOR (0) or LMGPU_QUEUE_INTERLEAVED (1) */
    int32_t reserved[2];
} lmgpu_batch_opts;

/* Order of the device work queue's (job, sample, pixel) items. JOB_MAJOR:
 * job by job, each sample by sample (a job finishes as early as it can).
 * INTERLEAVED: sample 0 of every job, then sample 1 of every job that has
 * it, and so on (jobs share the device until their last sample). */
#define LMGPU_QUEUE_JOB_MAJOR 0
#define LMGPU_QUEUE_INTERLEAVED 1

typedef struct lmgpu_batch_stats {
    int64_t jobs, samples, extension_rays, shadow_rays, iterations;
    double elapsed_s;       /* device time of the render loop */
    double build_s;         /* BVH builds and arena assembly in this call */
    double tune_s;          /* autotuning trials in this call (builds excluded) */
    int32_t width, path_capacity, block, reserved0;
    int64_t reserved[4];
} lmgpu_batch_stats;

int lmgpu_batch_create(lmgpu_batch** out);
int lmgpu_batch_destroy(lmgpu_batch* batch);
/* Copies the job's scene and parameters; returns its index in job_id. */
int lmgpu_batch_add_job(lmgpu_batch* batch, const lmgpu_job_desc* job, int32_t* job_id);
/* Renders every job (fresh films); autotunes the fields of opts left 0 on
 * the first run and reuses that choice afterwards. Synchronous. */
int lmgpu_batch_run(lmgpu_batch* batch, const lmgpu_batch_opts* opts, lmgpu_batch_stats* stats, void* stream);
/* FP64 sum / sum of squares ([h][w][3], host; the latter needs
 * film_sumsq in the run) of a job's samples in the last run, and its
 * latency: device time from the start of the run to the accumulation of
 * its last sample. Either pointer may be NULL. */
int lmgpu_batch_job_film(const lmgpu_batch* batch, int32_t job, double* film_sum, double* film_sumsq,
                         double* latency_ms);

/* ---- learned path guiding --------------------------------------------------
 * No counterpart in the reference: an MLP maps (position, shading normal,
 * outgoing direction) to a mixture of 8 von Mises-Fisher lobes over incident
 * directions; it is trained online on radiance recorded by the path tracer
 * and mixed with BSDF sampling by one-sample MIS (lmgpu_pt_render_guided: the
 * guide is chosen with probability alpha in [0, 1], default 0.65; 0.5 is an
 * even split), so the rendered estimate stays unbiased for any network state. */
#define LMGPU_GUIDE_NPARAMS 17664   /* packed FP32 parameter slots (15976 trained) */
#define LMGPU_GUIDE_MIX 40          /* decoded mixture: pi[8], kappa[8], mu[8][3] */
#define LMGPU_GUIDE_RECORD 16       /* training record floats: pos3 nrm3 wo3 wi3 pdf L3 */

typedef struct lmgpu_guide lmgpu_guide;

/* box_lo/box_hi: scene bounds used to normalise positions. */
int lmgpu_guide_create(lmgpu_guide** out, const float* box_lo, const float* box_hi, uint64_t seed);
int lmgpu_guide_destroy(lmgpu_guide* guide);
/* Copies the FP32 parameters out (get, host, LMGPU_GUIDE_NPARAMS floats) and/or
 * replaces them (set; resets the Adam state). Either may be NULL. */
int lmgpu_guide_params(lmgpu_guide* guide, float* get, const float* set);
/* Decoded mixtures for n queries: query[n][9] = pos, nrm, wo; mix[n][40]
 * (device pointers, asynchronous on stream). */
int lmgpu_guide_infer(const lmgpu_guide* guide, const float* query, float* mix, int64_t n, void* stream);
/* Samples one direction per decoded mixture (u[n][3] uniforms: lobe, two for
 * the vMF) into dirs[n][3] and, if pdf != NULL, its mixture pdf (solid
 * angle). Device pointers, asynchronous. */
int lmgpu_guide_sample(const float* mix, const float* u, float* dirs, float* pdf, int64_t n, void* stream);
/* Mixture pdf (solid angle) of dirs[n][3] under mix[n][40]. */
int lmgpu_guide_pdf(const float* mix, const float* dirs, float* pdf, int64_t n, void* stream);
/* One Adam step (learning rate lr) on n training records (device pointer,
 * [n][16]); the loss (host) and the gradient before the step (host,
 * LMGPU_GUIDE_NPARAMS floats) are returned when non-NULL. */
int lmgpu_guide_train(lmgpu_guide* guide, const float* records, int64_t n, float lr, float* loss, float* grads,
                      void* stream);
/* Adam steps taken so far and the loss of the last one (synchronous). */
int lmgpu_guide_stats(const lmgpu_guide* guide, int64_t* steps, float* last_loss);

/* Guided renderer::pt. At diffuse and glossy vertices the next direction is
 * drawn from the guide with probability alpha and from the BSDF otherwise;
 * the mixture density is used in the throughput and in all MIS weights, so
 * the estimate is unbiased for any guide. The radiance that arrived along
 * the sampled directions of the first record_depth surface vertices of a
 * path is turned into training records (pos, nrm, wo, wi, p_mix,
 * f|cos|L_i) when the path ends, up to batch * steps_per_pass of them per
 * sample per pixel in a film-spread order, and after every sample per
 * pixel's worth of finished paths the guide takes steps_per_pass Adam steps
 * (the last batches partial or empty); paths started later sample the
 * updated guide. Renders of 2 or more samples per pixel and the time
 * scheduler run in the regenerating loop (the guide's inference, records
 * and training steps are kernels of its captured iteration); a 1-spp call
 * is one captured one-sample pass followed by its training steps. */
typedef struct lmgpu_guide_opts {
    float alpha;              /* guide sampling probability (0.65) */
    float learning_rate;      /* Adam learning rate (1e-2) */
    int32_t batch;            /* records per Adam step, rounded up to 128 (2^14) */
    int32_t steps_per_pass;   /* Adam steps per sample per pixel (4); 0 = frozen guide, nothing recorded */
    int32_t record_depth;     /* recording surface vertices per path (8, at most 16) */
    float weight_clamp;       /* clamp of training weights relative to the batch mean (0 = 50) */
    int32_t reserved[6];
} lmgpu_guide_opts;

/* As lmgpu_pt_render; opts may be NULL (the defaults above). */
int lmgpu_pt_render_guided(lmgpu_scene* scene, const lmgpu_pt_params* params, lmgpu_guide* guide,
                           const lmgpu_guide_opts* opts, double* film_sum, double* film_sumsq, int film_on_device,
                           lmgpu_pt_stats* stats, void* stream);

#ifdef __cplusplus
}  /* extern "C" */
#endif

#endif /* LMGPU_H */
