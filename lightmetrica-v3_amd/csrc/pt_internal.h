// Host side of the wavefront path tracer: the device scene of one render
// (geometry, accel, materials, lights, camera) and its render workspace.
// Defined in pt.hip; the batched renderer (batch.hip) builds each job's
// device scene with Scene::create.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

#include "guide_internal.h"
#include "lmgpu_internal.h"
#include "pt_device.h"

namespace lmgpu {

// Guided rendering of one render call (opts resolved, workspace sized).
struct GuidedPass {
    Guide* g = nullptr;
    uint64_t uid = 0;   // guide identity and buffer generation (graph key)
    int gen = 0;
    float alpha = 0.5f, lr = 1e-3f, wclamp = 50.f;
    int batch = 0, steps = 0, rec_depth = 0;
};

// Path pool and work-queue loop of the regenerating wavefront (batch.hip).
struct Wavefront;
void wavefront_destroy(Wavefront* w);

struct Scene {
    Accel accel;
    DevBuf<ShadeTri> stri;
    DevBuf<DevMat> mats;
    DevBuf<DevLight> lights;
    DevBuf<float> light_cdf;
    DevBuf<int32_t> light_tri;
    DevBuf<float2> uvs;      // textured scenes: texture coordinates per triangle corner
    DevBuf<float4> texels;
    DevBuf<int4> tex_info;
    int num_textures = 0;
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
    DevBuf<int32_t> gshrec, gcount, grec_list;   // gcount[0] training records, gcount[1] written slots
    int grec_slots = 0;
    hipGraphExec_t graph = nullptr;
    int refill_primary = 64, refill_secondary = 16;
    int graph_key[4] = {0, 0, 0, 0};
    GuidedPass graph_guide;
    uint64_t graph_seed = 0;
    int graph_sumsq = -1;
    double* graph_sum_ptr = nullptr;
    double* graph_sq_ptr = nullptr;

    Wavefront* wf = nullptr;   // regenerating path of fixed-spp renders, created on first use

    ~Scene() {
        if (graph) (void)hipGraphExecDestroy(graph);
        wavefront_destroy(wf);
    }

    int create(const lmgpu_scene_desc* d, const lmgpu_build_opts* opts);
    int ensure_workspace(int npix);
    int ensure_guided(int npix, const GuidedPass& gp);
    int enqueue_pass(const PTConsts& c, double* sum, double* sumsq, const GuidedPass* gp, hipStream_t s);
    int render(const lmgpu_pt_params* p, double* sum, double* sumsq, bool on_device, const GuidedPass* gp,
               lmgpu_pt_stats* st, hipStream_t s);
    int raycast(const lmgpu_raycast_params* p, double* film, bool on_device, hipStream_t s);
    // A fixed-spp pixel-mode render through the regenerating wavefront: a pool of
    // path slots that take the render's (sample, pixel) items in order and add a
    // finished sample to the film at once (FP64 atomics), so paths of later samples
    // fill the slots that short paths free instead of each bounce waiting for a
    // pass's longest paths. Same Philox streams and per-vertex code as the passes.
    // With a guide (gp), directions at diffuse and glossy vertices come from
    // the one-sample mixture of guide and BSDF and the guide trains online:
    // the records of paths that finished are collected every iteration and
    // `steps` Adam steps run whenever batch * steps of them are buffered.
    int render_regen(const PTConsts& c, int spp, int sample_offset, double* sum, double* sumsq, int64_t* ext_rays,
                     int64_t* shadow_rays, hipStream_t s, const GuidedPass* gp = nullptr,
                     int64_t* guide_records = nullptr);
    // renderer::lt and renderer::bdpt (lt.hip)
    DevBuf<float4> blverts;            // BDPT light vertices, [samples per pass][max_verts - 1]
    DevBuf<lmgpu_ray> bsh;             // BDPT connection queue
    DevBuf<float4> bshc;
    DevBuf<int32_t> blcount;           // BDPT light subpath lengths
    int lt_render(const lmgpu_lt_params* p, double* film, bool on_device, lmgpu_pt_stats* st, hipStream_t s);
    int bdpt_render(const lmgpu_lt_params* p, double* film, bool on_device, lmgpu_pt_stats* st, hipStream_t s);
};

}  // namespace lmgpu

// The opaque C ABI handle of a scene.
struct lmgpu_scene {
    lmgpu::Scene s;
};
