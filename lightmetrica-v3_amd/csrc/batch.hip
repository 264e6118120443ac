// Batched multi-job renderer::pt (C ABI: lmgpu_batch_*).
//
// Many render jobs of uneven scene size, resolution and spp are rendered by
// one wavefront loop over a fixed pool of path slots:
//   * the jobs' BVHs live in one node / triangle arena (child indices made
//     global, so one traversal launch serves every job; a ray starts at its
//     job's root), their shading data (triangles, materials, lights) in
//     arenas addressed through per-job base pointers;
//   * one device work queue enumerates (job, sample, pixel) items job by job;
//     free path slots take the next items, finished paths add their sample to
//     their job's FP64 film (atomics) and count it, and the job that counts
//     its last sample records the device clock (job latency);
//   * an iteration is regenerate -> extend (persistent closest hit) -> shade
//     (shade_vertex, the transport of renderer::pt) -> connect (persistent
//     any hit); two iterations (both queue parities) are captured once as a
//     hipGraph and replayed until every work item has been accumulated.
// A job renders exactly what lmgpu_pt_render renders for the same scene,
// seed and BVH: same Philox streams, same per-vertex code.
//
// BVH width, path-pool capacity (queue capacity) and the block size of the
// regenerate / shade kernels are chosen by an autotuner when the caller
// leaves them 0: a coordinate search over {2,4,8} x {2^21,2^22,2^23} x
// {128,256} timing one sample per pixel of every job.

#include <hip/hip_runtime.h>

#include <algorithm>
#include <type_traits>
#include <chrono>
#include <cstring>
#include <new>
#include <string>
#include <vector>

#include "guide_internal.h"
#include "lmgpu_internal.h"
#include "pt_device.h"
#include "pt_internal.h"

#include "wavefront.h"

namespace lmgpu {

namespace {

__device__ __forceinline__ uint64_t job_seed(const JobDev& J) { return (uint64_t(J.seed_hi) << 32) | J.seed_lo; }

// Most grid-stride iterations of k_b_regen per block (the host sizes the grid
// so that a block covers at most this many BLOCK-slot strides of the pool).
constexpr int kRegenMaxIter = 32;
// Entries of k_b_regen's per-block table of finished-sample counts per job.
constexpr int kRegenJobSlots = 64;

// Frees finishing paths (their sample goes to the job film) and hands the
// free slots new work items, appending their camera rays to queue in_q.
// The shared counters are updated once per block, not once per wave: a
// first sweep over the block's slots counts the slots that will take work
// (per stride and wave, in LDS), one atomic on work_next then reserves the
// block's items and one on the queue count its queue entries; finished
// samples are counted per job in an LDS table and added to the job's and
// the run's counters at the end of the block.
template <int BLOCK>
__global__ __launch_bounds__(BLOCK) void k_b_regen(const JobDev* __restrict__ jobs, const WorkSeg* __restrict__ segs,
                                                   int nsegs, int64_t total, PoolDev P, BatchQ q, int in_q, int cap,
                                                   BatchCtl* ctl, unsigned long long* job_done,
                                                   unsigned long long* job_time, double* film_sum, double* film_sq) {
    constexpr int kWaves = BLOCK / 64;
    __shared__ int s_want[kRegenMaxIter][kWaves];   // slots taking work, then their exclusive prefix
    __shared__ unsigned long long s_base;           // first work item of the block
    __shared__ int s_qbase;                         // first queue entry of the block
    __shared__ int s_emit;                          // queue entries of the block
    __shared__ int s_key[kRegenJobSlots];
    __shared__ unsigned int s_cnt[kRegenJobSlots];
    __shared__ unsigned int s_fin;
    __shared__ int s_idle;
    // Every item accumulated: no slot is finishing and nothing is left to
    // hand out, so this and the rest of the iteration's kernels are no-ops
    // (the host enqueues iterations up to a bound without watching). One
    // thread reads the counter, so the whole block takes the same branch
    // (other blocks of this launch may lower it meanwhile).
    if (threadIdx.x == 0)
        s_idle = __hip_atomic_load(&ctl->remaining, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) == 0;
    __syncthreads();
    if (s_idle) return;
    if (blockIdx.x == 0 && threadIdx.x == 0) atomicAdd(&ctl->iterations, 1ull);
    const int stride = gridDim.x * BLOCK;
    const int n_iter = (cap + stride - 1) / stride;   // <= kRegenMaxIter (host)
    const int lane = __lane_id();
    const int wave = threadIdx.x >> 6;
    for (int k = threadIdx.x; k < kRegenJobSlots; k += BLOCK) {
        s_key[k] = -1;
        s_cnt[k] = 0u;
    }
    if (threadIdx.x == 0) s_fin = 0u;
    // sweep 1: slots that take work (free now or freed below)
    for (int it = 0; it < n_iter; it++) {
        const int s = it * stride + blockIdx.x * BLOCK + threadIdx.x;
        bool want = false;
        if (s < cap) {
            const int32_t st = P.state[s];
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
                const int c = s_want[it][w];
                s_want[it][w] = acc;
                acc += c;
            }
        unsigned long long base = 0;
        int emit = 0, qb = 0;
        if (acc > 0) {
            base = atomicAdd(&ctl->work_next, (unsigned long long)acc);
            if (base < (unsigned long long)total)
                emit = int(min((unsigned long long)acc, (unsigned long long)total - base));
            if (emit > 0) qb = atomicAdd(q.counts + in_q, emit);
        }
        s_base = base;
        s_emit = emit;
        s_qbase = qb;
    }
    __syncthreads();
    const unsigned long long base = s_base;
    const int emit_n = s_emit, qbase = s_qbase;
    // sweep 2: finish, then hand out
    for (int it = 0; it < n_iter; it++) {
        const int s = it * stride + blockIdx.x * BLOCK + threadIdx.x;
        bool want = false;
        if (s < cap) {
            int32_t st = P.state[s];
            if (st == kFinishing) {
                const int fj = P.job[s];
                const JobDev& J = jobs[fj];
                const float4 L = P.L[s];
                const int64_t f = J.film_off + 3 * int64_t(P.pixel[s]);
                atomicAdd(film_sum + f + 0, double(L.x));
                atomicAdd(film_sum + f + 1, double(L.y));
                atomicAdd(film_sum + f + 2, double(L.z));
                if (film_sq) {
                    atomicAdd(film_sq + f + 0, double(L.x) * double(L.x));
                    atomicAdd(film_sq + f + 1, double(L.y) * double(L.y));
                    atomicAdd(film_sq + f + 2, double(L.z) * double(L.z));
                }
                // count the finished sample in the block's job table, or
                // straight into the job's counter when its entry is taken
                atomicAdd(&s_fin, 1u);
                const int h = fj & (kRegenJobSlots - 1);
                const int prev = atomicCAS(&s_key[h], -1, fj);
                if (prev == -1 || prev == fj) {
                    atomicAdd(&s_cnt[h], 1u);
                } else {
                    const unsigned long long need = (unsigned long long)J.npix * (unsigned long long)J.spp;
                    if (atomicAdd(job_done + fj, 1ull) + 1ull == need) job_time[fj] = wall_clock64();
                }
                st = kFree;
                P.state[s] = kFree;
            }
            want = st == kFree;
        }
        const unsigned long long mask = __ballot(want);
        if (want) {
            const int r = s_want[it][wave] + __popcll(mask & ((1ull << lane) - 1ull));   // rank in the block
            if (r < emit_n) {
                const unsigned long long w = base + (unsigned long long)r;
                // segment of item w: the last one starting at or before w
                int lo = 0, hi = nsegs - 1;
                while (lo < hi) {
                    const int mid = (lo + hi + 1) >> 1;
                    if (segs[mid].start <= int64_t(w)) lo = mid;
                    else hi = mid - 1;
                }
                const WorkSeg seg = segs[lo];
                const JobDev& J = jobs[seg.job];
                const int sample = seg.sample, pix = int(int64_t(w) - seg.start);
                P.job[s] = seg.job;
                P.pixel[s] = pix;
                P.sample[s] = sample;
                P.state[s] = kActive;
                P.thr[s] = make_float4(1.f, 1.f, 1.f, 0.f);
                P.L[s] = make_float4(0.f, 0.f, 0.f, 0.f);
                P.prev[s] = make_float4(0.f, 0.f, 0.f, 0.f);
                P.info[s] = make_int4(1, 0, 2, 0);   // camera vertex: degenerate, not NEE-samplable
                q.ext[in_q][qbase + r] = camera_ray(J.sc, pix, uint32_t(sample), job_seed(J));
                q.ext_path[in_q][qbase + r] = s;
            }
        }
    }
    // completion counters: once per block for the run, once per job in the table
    __syncthreads();
    if (threadIdx.x == 0 && s_fin > 0u)
        atomicAdd(&ctl->remaining, (unsigned long long)(-(long long)s_fin));
    for (int k = threadIdx.x; k < kRegenJobSlots; k += BLOCK) {
        const int j = s_key[k];
        if (j >= 0 && s_cnt[k] > 0u) {
            const JobDev& J = jobs[j];
            const unsigned long long need = (unsigned long long)J.npix * (unsigned long long)J.spp;
            const unsigned long long c = s_cnt[k];
            if (atomicAdd(job_done + j, c) + c == need) job_time[j] = wall_clock64();
        }
    }
}

template <int W>
__global__ __launch_bounds__(kBlock) void k_b_extend(const typename TravNode<W>::type* __restrict__ nodes,
                                                     const TriV* __restrict__ tris,
                                                     const JobDev* __restrict__ jobs, PoolDev P, BatchQ q, int in_q,
                                                     unsigned long long* next, int refill, BatchCtl* ctl) {
    __shared__ int lds[kLds][kBlock];
    const int n = q.counts[in_q];
    if (n == 0) return;   // an idle iteration after the last work item (whole grid)
    const lmgpu_ray* rays = q.ext[in_q];
    auto fetch = [&](int64_t id, f3& o, f3& d, float& tmin, float& tmax) -> int {
        const float4* rp = reinterpret_cast<const float4*>(rays + id);
        const float4 a = rp[0], b = rp[1];
        o = mk(a.x, a.y, a.z);
        d = mk(b.x, b.y, b.z);
        tmin = a.w;
        tmax = b.w;
        return jobs[P.job[q.ext_path[in_q][id]]].root;
    };
    auto emit = [&](int64_t id, const TraceResult& r) {
        reinterpret_cast<float4*>(q.hit)[id] = make_float4(__int_as_float(r.overflow ? -2 : r.prim), r.t, r.u, r.v);
    };
    trace_persistent<W, false, false, kLds, kPriv, kBlock>(nodes, tris, n, next, lds, fetch, emit, q.counts + 5,
                                                           nullptr, refill, LMGPU_REGEN_RAY_CHUNK);
    if (blockIdx.x == 0 && threadIdx.x == 0) atomicAdd(&ctl->ext_rays, (unsigned long long)n);
}

template <int BLOCK, bool GUIDED>
__global__ __launch_bounds__(BLOCK) void k_b_shade(const JobDev* __restrict__ jobs, PoolDev P, BatchQ q, int in_q,
                                                   GuideDev gd, int cap) {
    const int n = q.counts[in_q];
    const int out_q = in_q ^ 1;
    const int stride = gridDim.x * BLOCK;
    const int n_iter = (n + stride - 1) / stride;
    for (int it = 0; it < n_iter; it++) {
        const int i = it * stride + blockIdx.x * BLOCK + threadIdx.x;
        VertexOut out;
        int s = 0;
        if (i < n) {
            s = q.ext_path[in_q][i];
            const JobDev& J = jobs[P.job[s]];
            const float4* rp = reinterpret_cast<const float4*>(q.ext[in_q] + i);
            const float4 ra = rp[0], rb = rp[1];
            const float4 h = reinterpret_cast<const float4*>(q.hit)[i];
            if (__float_as_int(h.x) == -2) atomicOr(q.counts + 5, 1);
            const float4 thr4 = P.thr[s], L4 = P.L[s];
            PathRegs R{mk(thr4.x, thr4.y, thr4.z), mk(L4.x, L4.y, L4.z), P.prev[s], P.info[s]};
            PTConsts c;
            c.max_verts = J.max_verts;
            c.mode = J.mode;
            c.rr_depth = J.rr_depth;
            c.seed_lo = J.seed_lo;
            c.seed_hi = J.seed_hi;
            // guided: the records of slot s sit at (vertex - 1) * cap + s
            c.npix = GUIDED ? cap : J.npix;
            shade_vertex<GUIDED>(J.sc, c, uint32_t(P.pixel[s]), uint32_t(P.sample[s]), job_seed(J), s,
                                 mk(ra.x, ra.y, ra.z), mk(rb.x, rb.y, rb.z), h, R, gd,
                                 GUIDED ? gd.mix + int64_t(i) * kOut : nullptr, out);
            if (out.ext) {
                P.prev[s] = R.prev;
                P.info[s] = R.info;
                P.thr[s] = make_float4(R.thr.x, R.thr.y, R.thr.z, 0.f);
            } else {
                // the path ends here; its NEE of this vertex lands in connect
                P.state[s] = kFinishing;
            }
            P.L[s] = make_float4(R.L.x, R.L.y, R.L.z, 0.f);
        }
        int se, ss;
        block_append2<BLOCK>(q.counts + out_q, q.counts + 2, out.ext, out.sh, se, ss);
        if (out.ext) {
            q.ext[out_q][se] = out.ext_ray;
            q.ext_path[out_q][se] = s;
        }
        if (out.sh) {
            q.sh[ss] = out.sh_ray;
            q.sh_contrib[ss] = make_float4(out.sh_C.x, out.sh_C.y, out.sh_C.z, __int_as_float(s));
            if (GUIDED) gd.sh_rec[ss] = out.sh_slot;
        }
    }
}

// Training records of the paths that finished in the last iteration (slots
// in kFinishing, before k_b_regen frees them): record d of slot s becomes
// (pos, nrm, wo, wi, p_mix, (L_final - e) f) when its target weight is
// positive, appended to train[tcap][16] while there is room, and every
// record of the slot is cleared for the slot's next path. Each wave takes a
// run of 64 consecutive slots (coalesced loads), the runs in the order
// w * kScramble mod (cap / 64) (an odd multiplier: a bijection on the
// power-of-two count of runs), so that a buffer that fills up holds paths
// from all over the film. A first sweep over a slot's records counts what it keeps,
// one atomic per block and stride reserves the block's entries, a second
// sweep writes them; the finished paths (*nfin) and kept records
// (*recorded) are added once per block.
__device__ __forceinline__ bool grec_target(const GRec* rp, float4 Lf, float4 a, float& lx, float& ly, float& lz) {
    const float4 e = rp->e, f = rp->f;
    lx = fmaxf(0.f, Lf.x - e.x) * f.x;
    ly = fmaxf(0.f, Lf.y - e.y) * f.y;
    lz = fmaxf(0.f, Lf.z - e.z) * f.z;
    const float wgt = (0.2126f * lx + 0.7152f * ly + 0.0722f * lz) / a.w;
    return wgt > 0.f && wgt < 3e38f;
}

__global__ __launch_bounds__(256) void k_b_gcollect(PoolDev P, int cap, GRec* __restrict__ rec, int rec_depth,
                                                    float* __restrict__ train, int tcap, int32_t* __restrict__ tcount,
                                                    int32_t* __restrict__ nfin,
                                                    unsigned long long* __restrict__ recorded) {
    constexpr uint32_t kScramble = 2654435761u;
    __shared__ int s_wave[4];
    __shared__ int s_base;
    __shared__ unsigned long long s_rec[4];
    __shared__ int s_fin[4];
    const int lane = __lane_id(), wave = threadIdx.x >> 6;
    const int stride = gridDim.x * 256;
    const int n_iter = (cap + stride - 1) / stride;
    unsigned long long my_rec = 0;
    int my_fin = 0;
    for (int it = 0; it < n_iter; it++) {
        const int j = it * stride + blockIdx.x * 256 + threadIdx.x;
        const uint32_t run = (uint32_t(j >> 6) * kScramble) & uint32_t((cap >> 6) - 1);
        const int s = int(run << 6) + lane;
        const bool fin = j < cap && P.state[s] == kFinishing;
        const float4 Lf = fin ? P.L[s] : make_float4(0.f, 0.f, 0.f, 0.f);
        // a vertex that sampled a direction left the path's vertex count at
        // its own count + 1, so only records d < count - 1 can have been written
        const int dmax = fin ? min(rec_depth, P.info[s].x - 1) : 0;
        int cnt = 0;
        if (fin) {
            for (int d = 0; d < dmax; d++) {
                const GRec* rp = rec + int64_t(d) * cap + s;
                const float4 a = rp->a;
                float lx, ly, lz;
                if (a.w > 0.f && grec_target(rp, Lf, a, lx, ly, lz)) cnt++;
            }
        }
        my_fin += fin ? 1 : 0;
        my_rec += cnt;
        int incl = cnt;
        for (int o = 1; o < 64; o <<= 1) {
            const int v = __shfl_up(incl, o);
            if (lane >= o) incl += v;
        }
        if (lane == 63) s_wave[wave] = incl;
        __syncthreads();
        if (threadIdx.x == 0) {
            int acc = 0;
            for (int w = 0; w < 4; w++) {
                const int x = s_wave[w];
                s_wave[w] = acc;
                acc += x;
            }
            s_base = acc ? atomicAdd(tcount, acc) : 0;
        }
        __syncthreads();
        int k = s_base + s_wave[wave] + incl - cnt;
        __syncthreads();   // the next stride rewrites s_wave and s_base
        if (fin) {
            for (int d = 0; d < dmax; d++) {
                GRec* rp = rec + int64_t(d) * cap + s;
                const float4 a = rp->a;
                if (a.w > 0.f) {
                    float lx, ly, lz;
                    if (grec_target(rp, Lf, a, lx, ly, lz)) {
                        if (k < tcap) {
                            const float4 b = rp->b, c = rp->c, dd = rp->d;
                            float4* o = reinterpret_cast<float4*>(train + int64_t(k) * kRecordFloats);
                            o[0] = make_float4(a.x, a.y, a.z, b.x);
                            o[1] = make_float4(b.y, b.z, c.x, c.y);
                            o[2] = make_float4(c.z, dd.x, dd.y, dd.z);
                            o[3] = make_float4(a.w, lx, ly, lz);
                        }
                        k++;
                    }
                    rp->a = make_float4(a.x, a.y, a.z, 0.f);
                }
            }
        }
    }
    for (int o = 32; o > 0; o >>= 1) {
        my_rec += __shfl_xor(my_rec, o);
        my_fin += __shfl_xor(my_fin, o);
    }
    if (lane == 0) {
        s_rec[wave] = my_rec;
        s_fin[wave] = my_fin;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        const unsigned long long r = s_rec[0] + s_rec[1] + s_rec[2] + s_rec[3];
        const int f = s_fin[0] + s_fin[1] + s_fin[2] + s_fin[3];
        if (r) atomicAdd(recorded, r);
        if (f) atomicAdd(nfin, f);
    }
}

// Once a sample per pixel's worth of paths (npix) has finished since the
// last release, releases the buffered records to the training steps that
// follow (*tready = min(count, tcap): `steps` batches, the last ones partial
// or empty, as after a one-sample pass) and restarts the buffer; the
// finished-path count keeps its excess (at most another npix), so releases
// keep pace with the samples. Otherwise the steps get no rows (*tready = 0)
// and return at entry.
__global__ void k_b_ggate(int32_t* tcount, int32_t* tready, int32_t* nfin, int npix, int tcap) {
    if (threadIdx.x == 0) {
        if (*nfin >= npix) {
            *tready = min(*tcount, tcap);
            *tcount = 0;
            *nfin = min(*nfin - npix, npix);
        } else {
            *tready = 0;
        }
    }
}

template <int W, bool GUIDED>
__global__ __launch_bounds__(kBlock) void k_b_connect(const typename TravNode<W>::type* __restrict__ nodes,
                                                      const TriV* __restrict__ tris,
                                                      const JobDev* __restrict__ jobs, PoolDev P, BatchQ q,
                                                      unsigned long long* next, int refill, BatchCtl* ctl,
                                                      GRec* __restrict__ rec, const int32_t* __restrict__ sh_rec) {
    __shared__ int lds[kLds][kBlock];
    const int n = q.counts[2];
    if (n == 0) return;   // no shadow rays this iteration (whole grid)
    auto fetch = [&](int64_t id, f3& o, f3& d, float& tmin, float& tmax) -> int {
        const float4* rp = reinterpret_cast<const float4*>(q.sh + id);
        const float4 a = rp[0], b = rp[1];
        o = mk(a.x, a.y, a.z);
        d = mk(b.x, b.y, b.z);
        tmin = a.w;
        tmax = b.w;
        return jobs[P.job[__float_as_int(q.sh_contrib[id].w)]].root;
    };
    auto emit = [&](int64_t id, const TraceResult& r) {
        if (r.prim < 0 && !r.overflow) {
            const float4 c = q.sh_contrib[id];
            const int s = __float_as_int(c.w);
            float4 L = P.L[s];
            L.x += c.x;
            L.y += c.y;
            L.z += c.z;
            P.L[s] = L;
            if (GUIDED) {
                // the vertex's NEE belongs to its record's radiance at sampling time
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
    trace_persistent<W, true, false, kLds, kPriv, kBlock>(nodes, tris, n, next, lds, fetch, emit, q.counts + 5, nullptr,
                                                          refill, LMGPU_REGEN_RAY_CHUNK);
    if (blockIdx.x == 0 && threadIdx.x == 0) atomicAdd(&ctl->shadow_rays, (unsigned long long)n);
}

// Makes a job's BVH node indices global: internal children += node_off,
// leaf triangle ranges += tri_off (empty slots unchanged).
template <int W>
__global__ void k_b_fix_nodes(WNode<W>* nodes, int64_t n, int node_off, int tri_off) {
    const int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i >= n) return;
    WNode<W>& nd = nodes[i];
#pragma unroll
    for (int c = 0; c < W; c++) {
        if (nd.count[c] == 0) nd.child[c] += node_off;
        else if (nd.count[c] > 0) nd.child[c] += tri_off;
    }
}

__global__ void k_b_init(BatchCtl* ctl, int64_t total, int32_t* counts) {
    if (threadIdx.x == 0) {
        ctl->work_next = 0;
        ctl->remaining = (unsigned long long)total;
        ctl->ext_rays = 0;
        ctl->shadow_rays = 0;
        ctl->iterations = 0;
    }
    if (threadIdx.x < 16) counts[threadIdx.x] = 0;
}

__global__ void k_b_clock(BatchCtl* ctl) {
    if (threadIdx.x == 0) ctl->t_start = wall_clock64();
}

__global__ void k_b_fill_i32(int32_t* p, int64_t n, int32_t v) {
    const int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i < n) p[i] = v;
}

__global__ void k_b_fill_u64(unsigned long long* p, int64_t n) {
    const int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i < n) p[i] = 0ull;
}

__global__ void k_b_fill_f64(double* p, int64_t n) {
    const int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i < n) p[i] = 0.0;
}

__global__ void k_b_zero2(unsigned long long* next) {
    if (threadIdx.x < 2) next[threadIdx.x] = 0ull;
}

__global__ void k_b_reset(int32_t* counts, int out_q) {
    if (threadIdx.x == 0) {
        counts[out_q] = 0;
        counts[2] = 0;
    }
}

inline unsigned nblk(int64_t n, int b) { return unsigned((n + b - 1) / b); }

// Graph replays (two iterations each) the host enqueues between looks at the
// device's progress.
constexpr int64_t kChunkReplays = 16;

template <class K>
int occupancy_grid(K kernel, int block) {
    int dev = 0, cus = 0, occ = 0;
    (void)hipGetDevice(&dev);
    (void)hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev);
    (void)hipOccupancyMaxActiveBlocksPerMultiprocessor(&occ, kernel, block, 0);
    return std::max(1, cus) * std::max(1, occ);
}

}  // namespace

// Host copy of one job (the caller's buffers are only valid during add_job).
struct BatchJob {
    std::vector<float> verts, normals, mat_param, prim_Ke, mat_ks, uvs, tex_data;
    std::vector<int32_t> tri_prim, mat_type, light_prims, mat_tex, tex_info;
    lmgpu_scene_desc desc{};
    int spp = 1, max_verts = 2, mode = 2, rr_depth = 5;
    uint64_t seed = 0;
};

struct Batch {
    std::vector<BatchJob> jobs;
    // arena (valid for arena_width / arena_leaf)
    int arena_width = 0, arena_leaf = 0;
    DevBuf<float4> nodes;
    DevBuf<float4> qnodes;   // quantised copy of the arena, when quantised_width(width)
    DevBuf<TriV> tris;
    DevBuf<ShadeTri> stri;
    DevBuf<DevMat> mats;
    DevBuf<DevLight> lights;
    DevBuf<float> cdf;
    DevBuf<int32_t> ltri;
    DevBuf<float2> uvs;       // textured jobs: texture coordinates per triangle corner
    DevBuf<float4> texels;    // textured jobs: texels (RGBA)
    DevBuf<int4> tex_info;    // textured jobs: width, height, channels, first texel within the job's texels
    std::vector<JobDev> jobdev;   // per job, film_off filled per run
    const void* trav_nodes() const { return qnodes.p ? static_cast<const void*>(qnodes.p) : nodes.p; }
    // run state
    Wavefront wf;
    DevBuf<double> film, film_sq;
    int64_t film_doubles = 0;
    std::vector<double> latency_ms;
    bool have_results = false;
    // autotuned choice
    int tuned_width = 0, tuned_cap = 0, tuned_block = 0;

    int build_arena(int width, int leaf, hipStream_t s);
    bool sumsq = false;   // the last run accumulated sums of squares
    int run_once(int width, int block, int spp_cap, bool keep_results, bool with_sumsq, int queue_order,
                 double* elapsed, int64_t* iters, hipStream_t s);
};

int Batch::build_arena(int width, int leaf, hipStream_t s) {
    if (arena_width == width && arena_leaf == leaf) return LMGPU_OK;
    arena_width = 0;
    // element counts per job: nodes, tris, shading tris, materials, lights, cdf entries, light tris,
    // texture coordinates, texels, textures
    constexpr int kParts = 10;
    struct Part {
        int64_t v[kParts];
    };
    std::vector<Part> parts;
    std::vector<Scene*> scenes;
    auto cleanup = [&]() {
        for (Scene* sc : scenes) delete sc;
        scenes.clear();
    };
    lmgpu_build_opts o{};
    o.width = width;
    o.leaf_size = leaf;
    o.builder = LMGPU_BUILDER_LBVH;
    int64_t tot[kParts] = {};
    for (size_t j = 0; j < jobs.size(); j++) {
        Scene* sc = new (std::nothrow) Scene();
        if (!sc) {
            cleanup();
            return fail(LMGPU_ERR_OUT_OF_MEMORY, "job scene");
        }
        scenes.push_back(sc);
        const int rc = sc->create(&jobs[j].desc, &o);
        if (rc != LMGPU_OK) {
            cleanup();
            return fail(rc, "job " + std::to_string(j) + ": " + lmgpu_last_error());
        }
        Part p{{sc->accel.num_nodes, int64_t(sc->accel.tris.n), int64_t(sc->stri.n), int64_t(sc->mats.n),
                int64_t(sc->lights.n), int64_t(sc->light_cdf.n), int64_t(sc->light_tri.n), int64_t(sc->uvs.n),
                int64_t(sc->texels.n), int64_t(sc->tex_info.n)}};
        parts.push_back(p);
        for (int k = 0; k < kParts; k++) tot[k] += p.v[k];
        if (tot[0] >= (int64_t(1) << 31) || tot[1] >= (int64_t(1) << 27)) {
            cleanup();
            return fail(LMGPU_ERR_INVALID_ARGUMENT, "batch too large for one arena (nodes or triangles)");
        }
    }
    const int f4_per_node = 2 * width;   // 32 * W bytes
    int rc = LMGPU_OK;
    if ((rc = nodes.alloc(size_t(tot[0]) * f4_per_node)) || (rc = tris.alloc(tot[1])) || (rc = stri.alloc(tot[2])) ||
        (rc = mats.alloc(tot[3])) || (rc = lights.alloc(tot[4])) || (rc = cdf.alloc(tot[5])) ||
        (rc = ltri.alloc(tot[6]))) {
        cleanup();
        return rc;
    }
    uvs.reset();
    texels.reset();
    tex_info.reset();
    if (tot[7] > 0 && ((rc = uvs.alloc(tot[7])) || (rc = texels.alloc(tot[8])) || (rc = tex_info.alloc(tot[9])))) {
        cleanup();
        return rc;
    }
    jobdev.assign(jobs.size(), JobDev{});
    int64_t off[kParts] = {};
    for (size_t j = 0; j < jobs.size(); j++) {
        Scene* sc = scenes[j];
        const Part& p = parts[j];
        auto cp = [&](void* dst, const void* src, size_t bytes) {
            return bytes ? hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToDevice, s) : hipSuccess;
        };
        hipError_t e = hipSuccess;
        if (e == hipSuccess) e = cp(nodes.p + off[0] * f4_per_node, sc->accel.nodes.p, size_t(p.v[0]) * 32 * width);
        if (e == hipSuccess) e = cp(tris.p + off[1], sc->accel.tris.p, sizeof(TriV) * p.v[1]);
        if (e == hipSuccess) e = cp(stri.p + off[2], sc->stri.p, sizeof(ShadeTri) * p.v[2]);
        if (e == hipSuccess) e = cp(mats.p + off[3], sc->mats.p, sizeof(DevMat) * p.v[3]);
        if (e == hipSuccess) e = cp(lights.p + off[4], sc->lights.p, sizeof(DevLight) * p.v[4]);
        if (e == hipSuccess) e = cp(cdf.p + off[5], sc->light_cdf.p, sizeof(float) * p.v[5]);
        if (e == hipSuccess) e = cp(ltri.p + off[6], sc->light_tri.p, sizeof(int32_t) * p.v[6]);
        if (e == hipSuccess && p.v[7] > 0) e = cp(uvs.p + off[7], sc->uvs.p, sizeof(float2) * p.v[7]);
        if (e == hipSuccess && p.v[7] > 0) e = cp(texels.p + off[8], sc->texels.p, sizeof(float4) * p.v[8]);
        if (e == hipSuccess && p.v[7] > 0) e = cp(tex_info.p + off[9], sc->tex_info.p, sizeof(int4) * p.v[9]);
        if (e != hipSuccess) {
            cleanup();
            return fail(LMGPU_ERR_DEVICE, std::string("arena copy: ") + hipGetErrorString(e));
        }
        void* nbase = nodes.p + off[0] * f4_per_node;
        if (width == 2)
            hipLaunchKernelGGL(k_b_fix_nodes<2>, dim3(nblk(p.v[0], 256)), dim3(256), 0, s,
                               static_cast<WNode<2>*>(nbase), p.v[0], int(off[0]), int(off[1]));
        else if (width == 4)
            hipLaunchKernelGGL(k_b_fix_nodes<4>, dim3(nblk(p.v[0], 256)), dim3(256), 0, s,
                               static_cast<WNode<4>*>(nbase), p.v[0], int(off[0]), int(off[1]));
        else
            hipLaunchKernelGGL(k_b_fix_nodes<8>, dim3(nblk(p.v[0], 256)), dim3(256), 0, s,
                               static_cast<WNode<8>*>(nbase), p.v[0], int(off[0]), int(off[1]));
        JobDev& J = jobdev[j];
        J.sc = sc->dev;
        J.sc.nodes2 = nullptr;
        J.sc.nodes4 = nullptr;
        J.sc.nodes8 = nullptr;
        J.sc.tris = tris.p;
        J.sc.stri = stri.p + off[2];
        J.sc.mats = mats.p + off[3];
        J.sc.lights = lights.p + off[4];
        J.sc.light_cdf = cdf.p + off[5];
        J.sc.light_tri = ltri.p + off[6];
        // a job's texture indices and first texels stay job-local: its pointers start at its own tables
        J.sc.uvs = p.v[7] > 0 ? uvs.p + off[7] : nullptr;
        J.sc.texels = p.v[7] > 0 ? texels.p + off[8] : nullptr;
        J.sc.tex_info = p.v[7] > 0 ? tex_info.p + off[9] : nullptr;
        J.root = int(off[0]);
        for (int k = 0; k < kParts; k++) off[k] += p.v[k];
        // the per-job scene is no longer needed once its data is in the arena
        LMGPU_HIP(hipStreamSynchronize(s));
        delete sc;
        scenes[j] = nullptr;
    }
    LMGPU_HIP(hipGetLastError());
    qnodes.reset();
    if (quantised_width(width)) {
        // after every job's indices are global: quantise the whole arena once
        if ((rc = qnodes.alloc(size_t(tot[0]) * (size_t(width) * 16 / sizeof(float4)))) ||
            (rc = quantise_nodes(width, nodes.p, qnodes.p, tot[0], s))) {
            cleanup();
            return rc;
        }
    }
    LMGPU_HIP(hipStreamSynchronize(s));
    cleanup();
    arena_width = width;
    arena_leaf = leaf;
    return LMGPU_OK;
}

int Wavefront::ensure_pool(int c) {
    if (cap == c) return LMGPU_OK;
    cap = 0;
    LMGPU_TRY(pjob.alloc(c)); LMGPU_TRY(ppix.alloc(c)); LMGPU_TRY(psample.alloc(c)); LMGPU_TRY(pstate.alloc(c));
    LMGPU_TRY(pthr.alloc(c)); LMGPU_TRY(pL.alloc(c)); LMGPU_TRY(pprev.alloc(c)); LMGPU_TRY(pinfo.alloc(c));
    LMGPU_TRY(ext0.alloc(c)); LMGPU_TRY(ext1.alloc(c)); LMGPU_TRY(sh.alloc(c));
    LMGPU_TRY(path0.alloc(c)); LMGPU_TRY(path1.alloc(c)); LMGPU_TRY(hit.alloc(c)); LMGPU_TRY(shc.alloc(c));
    if (!counts.p) LMGPU_TRY(counts.alloc(16));
    if (!ctl.p) LMGPU_TRY(ctl.alloc(1));
    if (!next.p) LMGPU_TRY(next.alloc(2));
    cap = c;
    return LMGPU_OK;
}

int Wavefront::set_work(const std::vector<JobDev>& jobs, const std::vector<WorkSeg>& segs, hipStream_t s) {
    const int nj = int(jobs.size());
    if (djobs.n != size_t(nj)) LMGPU_TRY(djobs.alloc(nj));
    if (dsegs.n != segs.size()) LMGPU_TRY(dsegs.alloc(segs.size()));
    if (job_done.n != size_t(nj)) {
        LMGPU_TRY(job_done.alloc(nj));
        LMGPU_TRY(job_time.alloc(nj));
    }
    LMGPU_HIP(hipStreamSynchronize(s));
    LMGPU_HIP(hipMemcpy(djobs.p, jobs.data(), sizeof(JobDev) * nj, hipMemcpyHostToDevice));
    LMGPU_HIP(hipMemcpy(dsegs.p, segs.data(), sizeof(WorkSeg) * segs.size(), hipMemcpyHostToDevice));
    hipLaunchKernelGGL(k_b_fill_u64, dim3(nblk(nj, 256)), dim3(256), 0, s, job_done.p, int64_t(nj));
    hipLaunchKernelGGL(k_b_fill_u64, dim3(nblk(nj, 256)), dim3(256), 0, s, job_time.p, int64_t(nj));
    LMGPU_HIP(hipGetLastError());
    return LMGPU_OK;
}

int Wavefront::ensure_guided(int rec_depth, int64_t train_cap) {
    const size_t nrec = size_t(std::max(rec_depth, 1)) * size_t(cap);
    if (grec.n < nrec) {
        LMGPU_TRY(grec.alloc(nrec));
        LMGPU_HIP(hipMemset(grec.p, 0, sizeof(GRec) * nrec));
    }
    if (gmix.n < size_t(cap) * kOut) LMGPU_TRY(gmix.alloc(size_t(cap) * kOut));
    if (gshrec.n < size_t(cap)) LMGPU_TRY(gshrec.alloc(cap));
    if (gtrain.n < size_t(std::max<int64_t>(train_cap, 1)) * kRecordFloats)
        LMGPU_TRY(gtrain.alloc(size_t(std::max<int64_t>(train_cap, 1)) * kRecordFloats));
    if (!gcnt.p) LMGPU_TRY(gcnt.alloc(3));
    if (!grecorded.p) LMGPU_TRY(grecorded.alloc(1));
    return LMGPU_OK;
}

int Wavefront::run(const void* trav_nodes, const TriV* tris, int width, int block, int64_t total, int max_verts,
                   double* film_sum, double* fsq, double* elapsed, int64_t* iters, hipStream_t s,
                   const RegenGuide* G) {
    const int nsegs = int(dsegs.n);
    PoolDev P{pjob.p, ppix.p, psample.p, pstate.p, pthr.p, pL.p, pprev.p, pinfo.p};
    BatchQ q{{ext0.p, ext1.p}, {path0.p, path1.p}, hit.p, sh.p, shc.p, counts.p};
    const int refill = 32;
    GuideDev gd{};
    const bool train = G && G->steps > 0 && G->rec_depth > 0;
    const int tcap = G ? G->batch * G->steps : 0;
    if (G) {
        gd.mix = gmix.p;
        gd.rec = G->rec_depth > 0 ? grec.p : nullptr;
        gd.sh_rec = gshrec.p;
        gd.alpha = G->alpha;
        gd.rec_depth = G->rec_depth;
        hipLaunchKernelGGL(k_b_fill_i32, dim3(1), dim3(64), 0, s, gcnt.p, int64_t(3), 0);
        hipLaunchKernelGGL(k_b_fill_u64, dim3(1), dim3(64), 0, s, grecorded.p, int64_t(1));
    }
    auto iteration = [&](hipStream_t cs, int in_q) {
        if (train) {
            // online training: the records of the paths that finished in the
            // last iteration; after every sample per pixel's worth of finished
            // paths, `steps` Adam steps on the buffered records
            hipLaunchKernelGGL(k_b_gcollect, dim3(unsigned(std::min<int64_t>(nblk(cap, 256), 4096))), dim3(256), 0, cs, P,
                               cap, grec.p, G->rec_depth, gtrain.p, tcap, gcnt.p, gcnt.p + 2, grecorded.p);
            hipLaunchKernelGGL(k_b_ggate, dim3(1), dim3(64), 0, cs, gcnt.p, gcnt.p + 1, gcnt.p + 2, G->npix, tcap);
            for (int t = 0; t < G->steps; t++)
                (void)G->g->enqueue_train(gtrain.p + int64_t(t) * G->batch * kRecordFloats, G->batch, G->batch,
                                          gcnt.p + 1, int64_t(t) * G->batch, G->lr, nullptr, cs, G->wclamp);
        }
        // at most kRegenMaxIter strides of the pool per block
        const unsigned rg = unsigned(std::max<int64_t>(std::max<int64_t>(1, std::min<int64_t>(nblk(cap, block), 4096)),
                                                       nblk(cap, block * kRegenMaxIter)));
        if (block == 128)
            hipLaunchKernelGGL(k_b_regen<128>, dim3(rg), dim3(128), 0, cs, djobs.p, dsegs.p, nsegs, total, P, q, in_q, cap,
                               ctl.p, job_done.p, job_time.p, film_sum, fsq);
        else
            hipLaunchKernelGGL(k_b_regen<256>, dim3(rg), dim3(256), 0, cs, djobs.p, dsegs.p, nsegs, total, P, q, in_q, cap,
                               ctl.p, job_done.p, job_time.p, film_sum, fsq);
        hipLaunchKernelGGL(k_b_zero2, dim3(1), dim3(64), 0, cs, next.p);
        if (width == 2)
            hipLaunchKernelGGL(k_b_extend<2>, dim3(occupancy_grid(k_b_extend<2>, kBlock)), dim3(kBlock), 0, cs,
                               static_cast<const TravNode<2>::type*>(trav_nodes), tris, djobs.p, P, q, in_q, next.p, refill, ctl.p);
        else if (width == 4)
            hipLaunchKernelGGL(k_b_extend<4>, dim3(occupancy_grid(k_b_extend<4>, kBlock)), dim3(kBlock), 0, cs,
                               static_cast<const TravNode<4>::type*>(trav_nodes), tris, djobs.p, P, q, in_q, next.p, refill, ctl.p);
        else
            hipLaunchKernelGGL(k_b_extend<8>, dim3(occupancy_grid(k_b_extend<8>, kBlock)), dim3(kBlock), 0, cs,
                               static_cast<const TravNode<8>::type*>(trav_nodes), tris, djobs.p, P, q, in_q, next.p, refill, ctl.p);
        hipLaunchKernelGGL(k_b_reset, dim3(1), dim3(64), 0, cs, counts.p, in_q ^ 1);
        if (G)
            hipLaunchKernelGGL(k_pt_guide_infer<BatchQ>, dim3(unsigned(std::max(1, std::min(cap / 128, 512)))), dim3(256),
                               0, cs, G->sc, q, in_q, G->g->Wbf.p, G->g->P.p, G->box, gmix.p);
        const unsigned sg = unsigned(std::max<int64_t>(1, std::min<int64_t>(nblk(cap, block), 2048 * 256 / block)));
        if (G) {
            if (block == 128)
                hipLaunchKernelGGL((k_b_shade<128, true>), dim3(sg), dim3(128), 0, cs, djobs.p, P, q, in_q, gd, cap);
            else
                hipLaunchKernelGGL((k_b_shade<256, true>), dim3(sg), dim3(256), 0, cs, djobs.p, P, q, in_q, gd, cap);
        } else {
            if (block == 128)
                hipLaunchKernelGGL((k_b_shade<128, false>), dim3(sg), dim3(128), 0, cs, djobs.p, P, q, in_q, gd, cap);
            else
                hipLaunchKernelGGL((k_b_shade<256, false>), dim3(sg), dim3(256), 0, cs, djobs.p, P, q, in_q, gd, cap);
        }
        if (width == 2) {
            const auto* nd = static_cast<const TravNode<2>::type*>(trav_nodes);
            if (G)
                hipLaunchKernelGGL((k_b_connect<2, true>), dim3(occupancy_grid(k_b_connect<2, true>, kBlock)), dim3(kBlock), 0,
                                   cs, nd, tris, djobs.p, P, q, next.p + 1, refill, ctl.p, gd.rec, gd.sh_rec);
            else
                hipLaunchKernelGGL((k_b_connect<2, false>), dim3(occupancy_grid(k_b_connect<2, false>, kBlock)), dim3(kBlock),
                                   0, cs, nd, tris, djobs.p, P, q, next.p + 1, refill, ctl.p, gd.rec, gd.sh_rec);
        } else if (width == 4) {
            const auto* nd = static_cast<const TravNode<4>::type*>(trav_nodes);
            if (G)
                hipLaunchKernelGGL((k_b_connect<4, true>), dim3(occupancy_grid(k_b_connect<4, true>, kBlock)), dim3(kBlock), 0,
                                   cs, nd, tris, djobs.p, P, q, next.p + 1, refill, ctl.p, gd.rec, gd.sh_rec);
            else
                hipLaunchKernelGGL((k_b_connect<4, false>), dim3(occupancy_grid(k_b_connect<4, false>, kBlock)), dim3(kBlock),
                                   0, cs, nd, tris, djobs.p, P, q, next.p + 1, refill, ctl.p, gd.rec, gd.sh_rec);
        } else {
            const auto* nd = static_cast<const TravNode<8>::type*>(trav_nodes);
            if (G)
                hipLaunchKernelGGL((k_b_connect<8, true>), dim3(occupancy_grid(k_b_connect<8, true>, kBlock)), dim3(kBlock), 0,
                                   cs, nd, tris, djobs.p, P, q, next.p + 1, refill, ctl.p, gd.rec, gd.sh_rec);
            else
                hipLaunchKernelGGL((k_b_connect<8, false>), dim3(occupancy_grid(k_b_connect<8, false>, kBlock)), dim3(kBlock),
                                   0, cs, nd, tris, djobs.p, P, q, next.p + 1, refill, ctl.p, gd.rec, gd.sh_rec);
        }
    };
    return loop(iteration, total, max_verts, elapsed, iters, s);
}

int Wavefront::loop(const std::function<void(hipStream_t, int)>& iteration, int64_t total, int max_verts,
                    double* elapsed, int64_t* iters, hipStream_t s) {
    hipLaunchKernelGGL(k_b_fill_i32, dim3(nblk(cap, 256)), dim3(256), 0, s, pstate.p, int64_t(cap), kFree);
    hipLaunchKernelGGL(k_b_init, dim3(1), dim3(64), 0, s, ctl.p, total, counts.p);
    // two iterations (queue parities 0 and 1) captured as one graph
    hipStream_t cs;
    LMGPU_HIP(hipStreamCreateWithFlags(&cs, hipStreamNonBlocking));
    hipGraph_t g;
    LMGPU_HIP(hipStreamBeginCapture(cs, hipStreamCaptureModeThreadLocal));
    iteration(cs, 0);
    iteration(cs, 1);
    hipError_t e = hipStreamEndCapture(cs, &g);
    if (e != hipSuccess) {
        (void)hipStreamDestroy(cs);
        return fail(LMGPU_ERR_DEVICE, std::string("batch graph capture: ") + hipGetErrorString(e));
    }
    hipGraphExec_t ge;
    e = hipGraphInstantiate(&ge, g, nullptr, nullptr, 0);
    (void)hipGraphDestroy(g);
    (void)hipStreamDestroy(cs);
    if (e != hipSuccess) return fail(LMGPU_ERR_DEVICE, std::string("batch graph instantiate: ") + hipGetErrorString(e));
    // every work item needs at most max_verts + 1 iterations in its slot
    const int max_v = std::max(2, max_verts);
    const int64_t iter_cap = (total / cap + 2) * (max_v + 2) + 64;
    hipEvent_t e0, e1;
    LMGPU_HIP(hipEventCreate(&e0));
    LMGPU_HIP(hipEventCreate(&e1));
    LMGPU_HIP(hipEventRecord(e0, s));
    hipLaunchKernelGGL(k_b_clock, dim3(1), dim3(64), 0, s, ctl.p);
    // The host never stalls the device: replays go out in chunks and the host
    // stays one chunk ahead. After each chunk the stream copies the control
    // block into pinned memory and records an event; before deciding on the
    // next chunk the host blocks on the event of the chunk before (no
    // stream synchronization, no spinning) and stops once that copy shows no
    // work left. Replays already queued when the work runs out find the
    // remaining-work counter at 0 and return at entry (k_b_regen). The
    // iteration bound turns a stall into an error.
    int rc = LMGPU_OK;
    if (!ctl_mirror) {
        LMGPU_HIP(hipHostMalloc(reinterpret_cast<void**>(&ctl_mirror), 2 * sizeof(BatchCtl), hipHostMallocDefault));
    }
    hipEvent_t chunk_ev[2];
    LMGPU_HIP(hipEventCreateWithFlags(&chunk_ev[0], hipEventBlockingSync | hipEventDisableTiming));
    LMGPU_HIP(hipEventCreateWithFlags(&chunk_ev[1], hipEventBlockingSync | hipEventDisableTiming));
    const int64_t max_replays = (iter_cap + 1) / 2;
    int64_t enqueued = 0;
    for (int64_t k = 0; enqueued < max_replays; k++) {
        const int64_t nrep = std::min<int64_t>(kChunkReplays, max_replays - enqueued);
        for (int64_t r = 0; r < nrep && rc == LMGPU_OK; r++)
            if ((e = hipGraphLaunch(ge, s)) != hipSuccess)
                rc = fail(LMGPU_ERR_DEVICE, std::string("batch graph launch: ") + hipGetErrorString(e));
        if (rc != LMGPU_OK) break;
        enqueued += nrep;
        if ((e = hipMemcpyAsync(ctl_mirror + (k & 1), ctl.p, sizeof(BatchCtl), hipMemcpyDeviceToHost, s)) !=
                hipSuccess ||
            (e = hipEventRecord(chunk_ev[k & 1], s)) != hipSuccess) {
            rc = fail(LMGPU_ERR_DEVICE, std::string("batch progress copy: ") + hipGetErrorString(e));
            break;
        }
        if (k >= 1) {
            if ((e = hipEventSynchronize(chunk_ev[(k - 1) & 1])) != hipSuccess) {
                rc = fail(LMGPU_ERR_DEVICE, std::string("batch progress: ") + hipGetErrorString(e));
                break;
            }
            if (ctl_mirror[(k - 1) & 1].remaining == 0) break;
        }
    }
    BatchCtl hc{};
    if (rc == LMGPU_OK) {
        if ((e = hipMemcpyAsync(&hc, ctl.p, sizeof(hc), hipMemcpyDeviceToHost, s)) != hipSuccess ||
            (e = hipStreamSynchronize(s)) != hipSuccess)
            rc = fail(LMGPU_ERR_DEVICE, std::string("batch run: ") + hipGetErrorString(e));
        else if (hc.remaining != 0)
            rc = fail(LMGPU_ERR_FAILED_TO_RENDER, "batch did not finish within its iteration bound");
    } else {
        (void)hipStreamSynchronize(s);
    }
    (void)hipEventDestroy(chunk_ev[0]);
    (void)hipEventDestroy(chunk_ev[1]);
    const int64_t it = int64_t(hc.iterations);
    (void)hipEventRecord(e1, s);
    (void)hipEventSynchronize(e1);
    float ms = 0.f;
    (void)hipEventElapsedTime(&ms, e0, e1);
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    (void)hipGraphExecDestroy(ge);
    if (rc != LMGPU_OK) return rc;
    int32_t cnt[16];
    LMGPU_HIP(hipMemcpy(cnt, counts.p, sizeof(cnt), hipMemcpyDeviceToHost));
    if (cnt[5]) return fail(LMGPU_ERR_STACK_OVERFLOW, "BVH traversal stack capacity exceeded during rendering");
    *elapsed = ms * 1e-3;
    *iters = it;
    return LMGPU_OK;
}

// Renders every job once with the current arena and pool: at most spp_cap
// samples per pixel per job when spp_cap > 0 (autotuning trials).
int Batch::run_once(int width, int block, int spp_cap, bool keep_results, bool with_sumsq, int queue_order,
                    double* elapsed, int64_t* iters, hipStream_t s) {
    const int nj = int(jobs.size());
    int64_t total = 0, film_n = 0;
    int max_spp = 0, max_v = 2;
    for (int j = 0; j < nj; j++) {
        JobDev& J = jobdev[j];
        const BatchJob& B = jobs[j];
        J.npix = B.desc.w * B.desc.h;
        J.spp = spp_cap > 0 ? std::min(B.spp, spp_cap) : B.spp;
        J.max_verts = B.max_verts;
        J.mode = B.mode;
        J.rr_depth = B.rr_depth;
        J.seed_lo = uint32_t(B.seed);
        J.seed_hi = uint32_t(B.seed >> 32);
        J.film_off = film_n;
        total += int64_t(J.npix) * J.spp;
        film_n += int64_t(J.npix) * 3;
        max_spp = std::max(max_spp, J.spp);
        max_v = std::max(max_v, B.max_verts);
    }
    // the work queue as segments of one (job, sample) each, in queue order
    std::vector<WorkSeg> segs;
    int64_t cursor = 0;
    auto add_seg = [&](int j, int smp) {
        segs.push_back(WorkSeg{cursor, j, smp});
        cursor += jobdev[j].npix;
    };
    if (queue_order == LMGPU_QUEUE_INTERLEAVED) {
        for (int smp = 0; smp < max_spp; smp++)
            for (int j = 0; j < nj; j++)
                if (smp < jobdev[j].spp) add_seg(j, smp);
    } else {
        for (int j = 0; j < nj; j++)
            for (int smp = 0; smp < jobdev[j].spp; smp++) add_seg(j, smp);
    }
    if (cursor != total || segs.size() >= size_t(1) << 31)
        return fail(LMGPU_ERR_INVALID_ARGUMENT, "batch work queue too large");
    if (film_doubles != film_n) {
        LMGPU_TRY(film.alloc(film_n));
        LMGPU_TRY(film_sq.alloc(film_n));
        film_doubles = film_n;
    }
    LMGPU_TRY(wf.set_work(jobdev, segs, s));
    hipLaunchKernelGGL(k_b_fill_f64, dim3(nblk(film_n, 256)), dim3(256), 0, s, film.p, film_n);
    if (with_sumsq) hipLaunchKernelGGL(k_b_fill_f64, dim3(nblk(film_n, 256)), dim3(256), 0, s, film_sq.p, film_n);
    LMGPU_TRY(wf.run(trav_nodes(), tris.p, width, block, total, max_v, film.p, with_sumsq ? film_sq.p : nullptr,
                     elapsed, iters, s));
    if (keep_results) {
        // job latency = device clock at the job's last sample - clock at the start
        std::vector<unsigned long long> t(nj);
        BatchCtl hc;
        LMGPU_HIP(hipMemcpy(t.data(), wf.job_time.p, sizeof(unsigned long long) * nj, hipMemcpyDeviceToHost));
        LMGPU_HIP(hipMemcpy(&hc, wf.ctl.p, sizeof(hc), hipMemcpyDeviceToHost));
        int dev = 0, khz = 0;
        LMGPU_HIP(hipGetDevice(&dev));
        LMGPU_HIP(hipDeviceGetAttribute(&khz, hipDeviceAttributeWallClockRate, dev));
        latency_ms.assign(nj, 0.0);
        for (int j = 0; j < nj; j++)
            latency_ms[j] = khz > 0 ? double(t[j] - hc.t_start) / double(khz) : 0.0;
        have_results = true;
        sumsq = with_sumsq;
    }
    return LMGPU_OK;
}

void wavefront_destroy(Wavefront* w) { delete w; }

int Scene::render_regen(const PTConsts& c, int spp, int sample_offset, double* sum, double* sumsq, int64_t* ext_rays,
                        int64_t* shadow_rays, hipStream_t s, const GuidedPass* gp, int64_t* guide_records) {
    if (!wf) {
        wf = new (std::nothrow) Wavefront();
        if (!wf) return fail(LMGPU_ERR_OUT_OF_MEMORY, "wavefront");
    }
    JobDev J{};
    J.sc = dev;
    J.root = 0;
    J.npix = c.npix;
    J.spp = spp;
    J.max_verts = c.max_verts;
    J.mode = c.mode;
    J.rr_depth = c.rr_depth;
    J.seed_lo = c.seed_lo;
    J.seed_hi = c.seed_hi;
    J.film_off = 0;
    const int64_t total = int64_t(c.npix) * spp;
    std::vector<WorkSeg> segs(static_cast<size_t>(spp));
    for (int k = 0; k < spp; k++) segs[size_t(k)] = WorkSeg{int64_t(k) * c.npix, 0, sample_offset + k};
    // pool: LMGPU_REGEN_SLOTS_PER_PIXEL paths per pixel (the slots of a
    // sample's short paths take the next sample's pixels)
    const int cap = gp ? Wavefront::pool_size(total, c.npix, LMGPU_REGEN_GUIDED_SLOTS_PER_PIXEL,
                                              LMGPU_REGEN_GUIDED_MAX_SLOTS)
                       : Wavefront::pool_size(total, c.npix);
    LMGPU_TRY(wf->ensure_pool(cap));
    RegenGuide G;
    if (gp) {
        G.g = gp->g;
        G.sc = dev;
        for (int a = 0; a < 3; a++) {
            G.box.lo[a] = gp->g->box_lo[a];
            G.box.inv[a] = gp->g->box_inv[a];
        }
        G.alpha = gp->alpha;
        G.lr = gp->lr;
        G.wclamp = gp->wclamp;
        G.batch = gp->batch;
        G.steps = gp->steps;
        G.rec_depth = gp->rec_depth;
        G.npix = c.npix;
        if (G.steps > 0) LMGPU_TRY(gp->g->ensure_train(G.batch));
        LMGPU_TRY(wf->ensure_guided(G.rec_depth, int64_t(G.batch) * G.steps));
    }
    LMGPU_TRY(wf->set_work(std::vector<JobDev>{J}, segs, s));
    const void* tn = accel.width == 4 ? static_cast<const void*>(dev.nodes4)
                     : accel.width == 8 ? static_cast<const void*>(dev.nodes8) : static_cast<const void*>(dev.nodes2);
    double sec = 0;
    int64_t iters = 0;
    LMGPU_TRY(wf->run(tn, accel.tris.p, accel.width, 128, total, c.max_verts, sum, sumsq, &sec, &iters, s,
                      gp ? &G : nullptr));
    BatchCtl hc;
    LMGPU_HIP(hipMemcpy(&hc, wf->ctl.p, sizeof(hc), hipMemcpyDeviceToHost));
    *ext_rays = int64_t(hc.ext_rays);
    *shadow_rays = int64_t(hc.shadow_rays);
    if (guide_records) {
        unsigned long long r = 0;
        if (gp) LMGPU_HIP(hipMemcpy(&r, wf->grecorded.p, sizeof(r), hipMemcpyDeviceToHost));
        *guide_records = int64_t(r);
    }
    return LMGPU_OK;
}

}  // namespace lmgpu

using lmgpu::fail;

extern "C" {

struct lmgpu_batch {
    lmgpu::Batch b;
};

int lmgpu_batch_create(lmgpu_batch** out) {
    if (!out) return fail(LMGPU_ERR_INVALID_ARGUMENT, "out is NULL");
    *out = new (std::nothrow) lmgpu_batch();
    if (!*out) return fail(LMGPU_ERR_OUT_OF_MEMORY, "batch handle");
    return LMGPU_OK;
}

int lmgpu_batch_destroy(lmgpu_batch* b) {
    delete b;
    return LMGPU_OK;
}

int lmgpu_batch_add_job(lmgpu_batch* b, const lmgpu_job_desc* job, int32_t* job_id) {
    if (!b || !job || !job->scene) return fail(LMGPU_ERR_INVALID_ARGUMENT, "NULL argument");
    const lmgpu_scene_desc& d = *job->scene;
    if (d.num_tris <= 0 || !d.verts || !d.normals || !d.tri_prim || d.num_prims <= 0 || !d.mat_type || !d.mat_param)
        return fail(LMGPU_ERR_INVALID_ARGUMENT, "job scene needs triangles, normals, primitive ids and materials");
    if (d.w <= 0 || d.h <= 0 || int64_t(d.w) * d.h > (int64_t(1) << 26))
        return fail(LMGPU_ERR_INVALID_ARGUMENT, "job film size out of range");
    if (job->spp < 1) return fail(LMGPU_ERR_INVALID_ARGUMENT, "job spp must be positive");
    if (job->max_verts < 2 || job->max_verts > 64) return fail(LMGPU_ERR_INVALID_ARGUMENT, "max_verts must be in [2, 64]");
    if (job->sampling_mode < 0 || job->sampling_mode > 2) return fail(LMGPU_ERR_INVALID_ARGUMENT, "bad sampling_mode");
    if (d.num_area_lights < 0 || (d.num_area_lights > 0 && !d.light_prims))
        return fail(LMGPU_ERR_INVALID_ARGUMENT, "bad light list");
    if (d.num_textures < 0 || (d.num_textures > 0 && (!d.tex_info || !d.tex_data || d.tex_data_len <= 0)))
        return fail(LMGPU_ERR_INVALID_ARGUMENT, "bad texture table");
    // (texture indices, sizes and coordinates are checked in full when the job's scene is built)
    lmgpu::BatchJob J;
    const int64_t N = d.num_tris;
    const int P = d.num_prims;
    J.verts.assign(d.verts, d.verts + 9 * N);
    J.normals.assign(d.normals, d.normals + 9 * N);
    J.tri_prim.assign(d.tri_prim, d.tri_prim + N);
    J.mat_type.assign(d.mat_type, d.mat_type + P);
    J.mat_param.assign(d.mat_param, d.mat_param + 8 * P);
    if (d.prim_Ke) J.prim_Ke.assign(d.prim_Ke, d.prim_Ke + 3 * P);
    else J.prim_Ke.assign(3 * P, 0.f);
    if (d.num_area_lights > 0) J.light_prims.assign(d.light_prims, d.light_prims + d.num_area_lights);
    if (d.mat_ks) J.mat_ks.assign(d.mat_ks, d.mat_ks + 3 * P);
    if (d.mat_tex) J.mat_tex.assign(d.mat_tex, d.mat_tex + P);
    if (d.uvs) J.uvs.assign(d.uvs, d.uvs + 6 * N);
    if (d.num_textures > 0) {
        J.tex_info.assign(d.tex_info, d.tex_info + 4 * int64_t(d.num_textures));
        J.tex_data.assign(d.tex_data, d.tex_data + d.tex_data_len);
    }
    J.desc = d;
    J.spp = job->spp;
    J.max_verts = job->max_verts;
    J.mode = job->sampling_mode;
    J.rr_depth = job->rr_depth;
    J.seed = job->seed;
    b->b.jobs.push_back(std::move(J));
    // point the stored descriptor at the host copies (after the move)
    lmgpu::BatchJob& S = b->b.jobs.back();
    S.desc.verts = S.verts.data();
    S.desc.normals = S.normals.data();
    S.desc.tri_prim = S.tri_prim.data();
    S.desc.mat_type = S.mat_type.data();
    S.desc.mat_param = S.mat_param.data();
    S.desc.prim_Ke = S.prim_Ke.data();
    S.desc.light_prims = S.light_prims.empty() ? nullptr : S.light_prims.data();
    S.desc.mat_ks = S.mat_ks.empty() ? nullptr : S.mat_ks.data();
    S.desc.mat_tex = S.mat_tex.empty() ? nullptr : S.mat_tex.data();
    S.desc.uvs = S.uvs.empty() ? nullptr : S.uvs.data();
    S.desc.tex_info = S.tex_info.empty() ? nullptr : S.tex_info.data();
    S.desc.tex_data = S.tex_data.empty() ? nullptr : S.tex_data.data();
    b->b.arena_width = 0;   // the arena must be rebuilt
    b->b.have_results = false;
    if (job_id) *job_id = int32_t(b->b.jobs.size() - 1);
    return LMGPU_OK;
}

int lmgpu_batch_run(lmgpu_batch* bh, const lmgpu_batch_opts* opts, lmgpu_batch_stats* stats, void* stream) {
    if (!bh) return fail(LMGPU_ERR_INVALID_ARGUMENT, "batch is NULL");
    lmgpu::clear_error();
    lmgpu::Batch& b = bh->b;
    if (b.jobs.empty()) return fail(LMGPU_ERR_INVALID_ARGUMENT, "batch has no jobs");
    hipStream_t s = static_cast<hipStream_t>(stream);
    lmgpu_batch_opts o{};
    if (opts) o = *opts;
    if (o.width != 0 && o.width != 2 && o.width != 4 && o.width != 8)
        return fail(LMGPU_ERR_INVALID_ARGUMENT, "width must be 0 (autotune), 2, 4 or 8");
    if (o.path_capacity != 0 && (o.path_capacity < 1024 || o.path_capacity > (1 << 24)))
        return fail(LMGPU_ERR_INVALID_ARGUMENT, "path_capacity must be 0 (autotune) or in [1024, 2^24]");
    if (o.queue_order != LMGPU_QUEUE_JOB_MAJOR && o.queue_order != LMGPU_QUEUE_INTERLEAVED)
        return fail(LMGPU_ERR_INVALID_ARGUMENT, "unknown queue_order");
    if (o.block != 0 && o.block != 128 && o.block != 256)
        return fail(LMGPU_ERR_INVALID_ARGUMENT, "block must be 0 (autotune), 128 or 256");
    const int leaf = o.leaf_size > 0 ? o.leaf_size : 2;
    if (leaf > 15) return fail(LMGPU_ERR_INVALID_ARGUMENT, "leaf_size must be in [1, 15]");
    double build_s = 0, tune_s = 0;
    auto now = [] { return std::chrono::steady_clock::now(); };
    int width = o.width, capv = o.path_capacity, block = o.block;
    if (!width || !capv || !block) {
        if (b.tuned_width && (!o.width || o.width == b.tuned_width)) {
            width = width ? width : b.tuned_width;
            capv = capv ? capv : b.tuned_cap;
            block = block ? block : b.tuned_block;
        } else {
            // coordinate search on one sample per pixel of every job
            const auto t0 = now();
            double tb = 0;
            auto trial = [&](int w, int c, int bl, double* sec) -> int {
                const auto tb0 = now();
                LMGPU_TRY(b.build_arena(w, leaf, s));
                tb += std::chrono::duration<double>(now() - tb0).count();
                LMGPU_TRY(b.wf.ensure_pool(c));
                int64_t it;
                return b.run_once(w, bl, 1, false, o.film_sumsq != 0, o.queue_order, sec, &it, s);
            };
            const int widths[3] = {4, 8, 2}, caps[3] = {1 << 22, 1 << 21, 1 << 23}, blocks[2] = {256, 128};
            int bw = width ? width : 4, bc = capv ? capv : (1 << 22), bb = block ? block : 256;
            double best = 1e30, sec;
            if (!width)
                for (int w : widths) {
                    LMGPU_TRY(trial(w, bc, bb, &sec));
                    if (sec < best) { best = sec; bw = w; }
                }
            if (!capv)
                for (int c : caps) {
                    LMGPU_TRY(trial(bw, c, bb, &sec));
                    if (sec < best) { best = sec; bc = c; }
                }
            if (!block)
                for (int bl : blocks) {
                    LMGPU_TRY(trial(bw, bc, bl, &sec));
                    if (sec < best) { best = sec; bb = bl; }
                }
            width = bw;
            capv = bc;
            block = bb;
            b.tuned_width = bw;
            b.tuned_cap = bc;
            b.tuned_block = bb;
            build_s += tb;
            tune_s = std::chrono::duration<double>(now() - t0).count() - tb;
        }
    }
    const auto tb0 = now();
    LMGPU_TRY(b.build_arena(width, leaf, s));
    build_s += std::chrono::duration<double>(now() - tb0).count();
    LMGPU_TRY(b.wf.ensure_pool(capv));
    double sec = 0;
    int64_t iters = 0;
    LMGPU_TRY(b.run_once(width, block, 0, true, o.film_sumsq != 0, o.queue_order, &sec, &iters, s));
    if (stats) {
        std::memset(stats, 0, sizeof(*stats));
        lmgpu::BatchCtl hc;
        LMGPU_HIP(hipMemcpy(&hc, b.wf.ctl.p, sizeof(hc), hipMemcpyDeviceToHost));
        int64_t samples = 0;
        for (const auto& J : b.jobs) samples += int64_t(J.desc.w) * J.desc.h * J.spp;
        stats->jobs = int64_t(b.jobs.size());
        stats->samples = samples;
        stats->extension_rays = int64_t(hc.ext_rays);
        stats->shadow_rays = int64_t(hc.shadow_rays);
        stats->iterations = iters;
        stats->elapsed_s = sec;
        stats->build_s = build_s;
        stats->tune_s = tune_s;
        stats->width = width;
        stats->path_capacity = capv;
        stats->block = block;
    }
    return LMGPU_OK;
}

int lmgpu_batch_job_film(const lmgpu_batch* bh, int32_t job, double* film_sum, double* film_sumsq,
                         double* latency_ms) {
    if (!bh) return fail(LMGPU_ERR_INVALID_ARGUMENT, "batch is NULL");
    const lmgpu::Batch& b = bh->b;
    if (job < 0 || size_t(job) >= b.jobs.size()) return fail(LMGPU_ERR_INVALID_ARGUMENT, "job index out of range");
    if (!b.have_results) return fail(LMGPU_ERR_INVALID_ARGUMENT, "the batch has not been run since its last change");
    if (film_sumsq && !b.sumsq) return fail(LMGPU_ERR_INVALID_ARGUMENT, "the last run did not accumulate sums of squares");
    const lmgpu::JobDev& J = b.jobdev[size_t(job)];
    const size_t n = size_t(J.npix) * 3;
    if (film_sum) LMGPU_HIP(hipMemcpy(film_sum, b.film.p + J.film_off, sizeof(double) * n, hipMemcpyDeviceToHost));
    if (film_sumsq)
        LMGPU_HIP(hipMemcpy(film_sumsq, b.film_sq.p + J.film_off, sizeof(double) * n, hipMemcpyDeviceToHost));
    if (latency_ms) *latency_ms = b.latency_ms[size_t(job)];
    return LMGPU_OK;
}

}  // extern "C"
