// Path pool and device work-queue loop of the regenerating wavefront
// renderers (batch.hip: the batch and fixed-spp path tracing; lt.hip:
// light tracing): a fixed pool of path slots whose free slots take the next
// work items, an iteration of regenerate -> extend -> shade -> connect
// kernels, two iterations captured as one hipGraph and replayed until every
// item is done.
#pragma once

#include <hip/hip_runtime.h>

#include <functional>
#include <vector>

#include "lmgpu_internal.h"
#include "pt_device.h"

// Path pool of a fixed-spp render: slots per pixel and at most (1080p
// config 2: 2 per pixel 431, 4: 503, 8: 547, 16: 558 M samples/s; 2^24 slots
// take 3.6 GB).
#ifndef LMGPU_REGEN_SLOTS_PER_PIXEL
#define LMGPU_REGEN_SLOTS_PER_PIXEL 8
#endif
#ifndef LMGPU_REGEN_MAX_SLOTS
#define LMGPU_REGEN_MAX_SLOTS (1 << 24)
#endif
// Rays a wave takes per atomic on the traversal's ray counter in the
// regenerating loops, whose queues stay long (the one-sample passes use
// LMGPU_RAY_CHUNK: their late bounces queue few rays).
// Path pool of a guided render, whose slots also hold training records
// (record_depth x 96 B each): 4 slots per pixel, at most 2^23. With fewer
// paths in flight an iteration finishes about half a sample per pixel's
// worth of paths, so the guide can train after every such sample.
#ifndef LMGPU_REGEN_GUIDED_SLOTS_PER_PIXEL
#define LMGPU_REGEN_GUIDED_SLOTS_PER_PIXEL 4
#endif
#ifndef LMGPU_REGEN_GUIDED_MAX_SLOTS
#define LMGPU_REGEN_GUIDED_MAX_SLOTS (1 << 23)
#endif
#ifndef LMGPU_REGEN_RAY_CHUNK
#define LMGPU_REGEN_RAY_CHUNK 128
#endif

namespace lmgpu {

struct JobDev {
    SceneDev sc;          // job-local pointers into the shading arenas
    int root;             // root of the job's BVH in the node arena
    int npix, spp, max_verts, mode, rr_depth;
    uint32_t seed_lo, seed_hi;
    int64_t film_off;     // first double of the job's film (3 per pixel)
};

// A run of consecutive work items: the pixels 0.. of one sample of one job,
// starting at queue position `start`.
struct WorkSeg {
    int64_t start;
    int32_t job, sample;
};

constexpr int32_t kFree = 0, kActive = 1, kFinishing = 2;

struct PoolDev {
    int32_t *job, *pixel, *sample, *state;
    float4 *thr, *L, *prev;
    int4* info;
};

struct BatchQ {
    lmgpu_ray* ext[2];
    int32_t* ext_path[2];
    lmgpu_hit* hit;
    lmgpu_ray* sh;
    float4* sh_contrib;   // xyz contribution, w = path slot bits
    int32_t* counts;      // [0],[1] extension queues, [2] shadow queue, [5] stack overflow flag
};

struct BatchCtl {
    unsigned long long work_next;   // next work item to hand out
    unsigned long long remaining;   // work items not yet accumulated
    unsigned long long ext_rays, shadow_rays;
    unsigned long long t_start;     // device clock at the start of the run
    unsigned long long iterations;  // iterations that found work left
};


struct Guide;

// Guided path tracing in the regenerating loop: the guide, its options and
// the scene the inference reads hit points from.
struct RegenGuide {
    Guide* g = nullptr;
    SceneDev sc{};
    GBox box{};
    float alpha = 0.5f, lr = 1e-3f, wclamp = 50.f;
    int batch = 0, steps = 0, rec_depth = 0;
    int npix = 0;   // the guide trains after every npix finished paths
};

// Path pool, work queue and the loop that drains it (regenerate -> extend ->
// shade -> connect, two iterations per graph replay), shared by the batch
// and by lmgpu_pt_render's regenerating path.
struct Wavefront {
    int cap = 0;
    DevBuf<int32_t> pjob, ppix, psample, pstate;
    DevBuf<float4> pthr, pL, pprev;
    DevBuf<int4> pinfo;
    DevBuf<lmgpu_ray> ext0, ext1, sh;
    DevBuf<int32_t> path0, path1, counts;
    DevBuf<lmgpu_hit> hit;
    DevBuf<float4> shc;
    DevBuf<BatchCtl> ctl;
    BatchCtl* ctl_mirror = nullptr;   // pinned host copies of ctl, one per chunk parity
    DevBuf<unsigned long long> next;
    // per-run tables: jobs, work segments, per-job completion counters and clocks
    DevBuf<JobDev> djobs;
    DevBuf<WorkSeg> dsegs;
    DevBuf<unsigned long long> job_done, job_time;
    // guided runs: training records per (vertex, slot), decoded mixtures per
    // extension-queue entry, the record slot of each shadow-queue entry, the
    // training buffer with its counters ([0] appended, [1] released to the
    // training steps, [2] paths finished since the last release) and the
    // count of training records
    DevBuf<GRec> grec;
    DevBuf<float> gmix, gtrain;
    DevBuf<int32_t> gshrec, gcnt;
    DevBuf<unsigned long long> grecorded;
    ~Wavefront() {
        if (ctl_mirror) (void)hipHostFree(ctl_mirror);
    }
    int ensure_pool(int cap);
    // Guided buffers for the current pool (records zeroed when allocated).
    int ensure_guided(int rec_depth, int64_t train_cap);
    // Uploads the jobs and the work segments (queue order) of one run.
    int set_work(const std::vector<JobDev>& jobs, const std::vector<WorkSeg>& segs, hipStream_t s);
    // Renders every work item of the tables set last into film_sum / film_sq
    // (FP64 atomics; film_sq may be null) and waits for the end.
    // With a guide, the directions at diffuse and glossy vertices come from
    // the one-sample mixture of guide and BSDF, and the guide trains online
    // on the records of finished paths.
    int run(const void* trav_nodes, const TriV* tris, int width, int block, int64_t total, int max_verts,
            double* film_sum, double* film_sq, double* elapsed, int64_t* iters, hipStream_t s,
            const RegenGuide* guide = nullptr);
    // The loop itself: frees the pool, sets the run's counters for `total`
    // items, captures iteration(stream, 0) and iteration(stream, 1) (the
    // kernels of one iteration on queue parity 0 / 1) as one graph and
    // replays it until every item is done; max_verts bounds the iterations.
    int loop(const std::function<void(hipStream_t, int)>& iteration, int64_t total, int max_verts, double* elapsed,
             int64_t* iters, hipStream_t s);
    // Slot count for `total` items over `npix` pixels: per_pixel slots per
    // pixel, a power of two in [4096, max_slots].
    static int pool_size(int64_t total, int npix, int per_pixel = LMGPU_REGEN_SLOTS_PER_PIXEL,
                         int max_slots = LMGPU_REGEN_MAX_SLOTS) {
        const int64_t want = std::min<int64_t>(total, int64_t(per_pixel) * npix);
        int c = 4096;
        while (c < want && c < max_slots) c <<= 1;
        return c;
    }
};

}  // namespace lmgpu


