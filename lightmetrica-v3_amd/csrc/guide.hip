// Learned path guiding: MLP forward (inference and training), vMF-mixture
// loss gradient, backward pass, weight gradients and Adam, all on gfx950
// MFMA (BF16 operands, FP32 accumulation). C ABI: lmgpu_guide_*.
//
// Training minimises the Monte Carlo estimate of the KL divergence between
// the normalised target (incident radiance luminance along w_i, recorded by
// the path tracer, over the density p_mix that sampled w_i) and the guide:
//   loss = -(1/B) sum_i wt_i log p_guide(w_i | x_i),  wt_i = w_i / mean(w),
//   w_i = lum(L_i) / p_mix(w_i),
// the estimator of Neural Importance Sampling (Mueller et al., TOG 2019) with
// a vMF mixture output as in neural parametric mixtures (Dong et al., 2023).
// The estimator of the rendered image stays unbiased for any guide because
// the guide is only ever mixed with BSDF sampling by one-sample MIS.

#include <hip/hip_runtime.h>

#include <algorithm>
#include <atomic>
#include <cmath>
#include <cstring>
#include <new>
#include <random>
#include <string>
#include <vector>

#include "guide.h"
#include "guide_internal.h"
#include "lmgpu_internal.h"

namespace lmgpu {

namespace {

// Samples per block of the weight-gradient kernel (a multiple of 128).
#ifndef LMGPU_DW_SLICE
#define LMGPU_DW_SLICE 1024
#endif

constexpr int kGB = 256;                  // 4 waves x 32 rows per block tile
constexpr int kRowsPerBlock = 128;

struct Box {
    float lo[3], inv[3];
};

__device__ __forceinline__ float lum(float r, float g, float b) { return 0.2126f * r + 0.7152f * g + 0.0722f * b; }

// Rows of a training batch actually present: n, or, when the count lives in
// device memory (batches compacted by the path tracer inside a captured
// graph), clamp(*ndev - off, 0, n).
__device__ __forceinline__ int64_t batch_rows(const int32_t* ndev, int64_t off, int64_t n) {
    if (!ndev) return n;
    return std::min<int64_t>(n, std::max<int64_t>(0, int64_t(*ndev) - off));
}
__device__ __forceinline__ int64_t round128(int64_t n) { return (n + 127) / 128 * 128; }

// Forward over query arrays. query[i] = pos(3), nrm(3), wo(3) (stride qs
// floats) -> decoded mixtures mix[i][40] = pi[8], kappa[8], mu[8][3].
__global__ __launch_bounds__(kGB) void k_guide_forward(int64_t n, const float* __restrict__ query, int qs,
                                                       const uint16_t* __restrict__ Wbf, const float* __restrict__ P,
                                                       Box box, float* __restrict__ mix) {
    __shared__ __attribute__((aligned(16))) unsigned char lds[kWeightLds + 4 * kWaveLds];
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    uint16_t* wl = reinterpret_cast<uint16_t*>(lds);
    stage_weights(Wbf, wl, threadIdx.x, kGB);
    uint16_t* lx = reinterpret_cast<uint16_t*>(lds + kWeightLds + wave * kWaveLds);
    uint16_t* lh = lx + 32 * kPitchX;
    float* ly = reinterpret_cast<float*>(lx);
    // two lanes per row: lanes r and r + 32 share row r of the wave's 32-row
    // tile and each encode and decode half of it
    const int r = lane & 31, half = lane >> 5;
    const int k_lo = half * (kIn / 2), k_hi = k_lo + kIn / 2;   // feature half for zeroing
    for (int64_t base = int64_t(blockIdx.x) * kRowsPerBlock; base < n; base += int64_t(gridDim.x) * kRowsPerBlock) {
        const int64_t i = base + wave * 32 + r;
        {
            uint16_t* row = lx + r * kPitchX;
            if (i < n) {
                const float* q = query + i * qs;
                encode_input_half(half, mk(q[0], q[1], q[2]), mk(q[3], q[4], q[5]), mk(q[6], q[7], q[8]), box.lo,
                                  box.inv, row);
            } else {
                for (int k = k_lo; k < k_hi; k++) row[k] = 0;
            }
        }
        __syncthreads();
        mlp_tile_forward(wl, P, lx, lh, ly, lane, [](int, int, int, uint16_t) {});
        __syncthreads();
        if (i < n) decode_lobes4(ly + r * kPitchY, 4 * half, mix + i * kOut);
        __syncthreads();
    }
}

// Target weights w_i = lum(L_i) / p_mix_i of training records
// rec[i] = pos(3), nrm(3), wo(3), wi(3), p_mix, L(3); sum into *wsum (zero
// on entry); clears *loss, which the row kernel accumulates.
__global__ __launch_bounds__(256) void k_guide_weights(int64_t n_max, const int32_t* ndev, int64_t off,
                                                       const float* __restrict__ rec, float* __restrict__ w,
                                                       float* __restrict__ wsum, float* __restrict__ loss) {
    __shared__ float red[4];
    const int64_t n = batch_rows(ndev, off, n_max);
    if (blockIdx.x == 0 && threadIdx.x == 0) *loss = 0.f;
    if (int64_t(blockIdx.x) * 256 >= n) return;   // a block past the batch (whole block)
    const int64_t i = int64_t(blockIdx.x) * 256 + threadIdx.x;
    float v = 0.f;
    if (i < n) {
        const float* r = rec + i * kRecordFloats;
        const float p = r[12];
        const float l = lum(r[13], r[14], r[15]);
        v = (p > 0.f && l > 0.f && isfinite(l / p)) ? l / p : 0.f;
        w[i] = v;
    }
    for (int sh = 32; sh > 0; sh >>= 1) v += __shfl_xor(v, sh);
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = v;
    __syncthreads();
    if (threadIdx.x == 0) atomicAdd(wsum, red[0] + red[1] + red[2] + red[3]);
}

// One training pass over the 128-row tile of a block (32 rows per wave): input encoding,
// forward (saving xT and hT, feature major, for the weight gradients), the
// gradient of the loss -wt log p_guide(w_i) w.r.t. the 40 raw outputs of
// every row (scaled by 1/n), and backward through layers 4..2 (saving d4T,
// d3T, d2T, d1T). The raw outputs and the deltas stay in LDS; rows from n to
// the end of the last 128-row tile get zero deltas. Lanes r and r + 32 share
// row r: each encodes half of it and differentiates four of the eight lobes.
__global__ __launch_bounds__(kGB, 2) void k_guide_train_rows(int64_t n_max, const int32_t* ndev, int64_t off, int64_t B,
                                                          const float* __restrict__ rec,
                                                          const uint16_t* __restrict__ Wbf,
                                                          const float* __restrict__ P,
                                                          const uint16_t* __restrict__ WnBf,
                                                          const float* __restrict__ zero64, Box box,
                                                          const float* __restrict__ w, const float* __restrict__ wsum,
                                                          float wclamp, uint16_t* __restrict__ xT,
                                                          uint16_t* __restrict__ hT, uint16_t* __restrict__ dT,
                                                          float* __restrict__ loss) {
    __shared__ __attribute__((aligned(16))) unsigned char lds[kWeightLds + 4 * kWaveLds];
    __shared__ float red[4];
    const int64_t n = batch_rows(ndev, off, n_max);
    // one 128-row tile per block (no loop over tiles: values the tiles would
    // share stay out of registers, so two waves fit per SIMD)
    const int64_t base = int64_t(blockIdx.x) * kRowsPerBlock;
    if (base >= n) return;   // whole block
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    uint16_t* wl = reinterpret_cast<uint16_t*>(lds);
    stage_weights(Wbf, wl, threadIdx.x, kGB);
    uint16_t* lx = reinterpret_cast<uint16_t*>(lds + kWeightLds + wave * kWaveLds);
    uint16_t* lh = lx + 32 * kPitchX;
    float* ly = reinterpret_cast<float*>(lx);
    const int r = lane & 31, half = lane >> 5;
    const int k_lo = half * (kIn / 2), k_hi = k_lo + kIn / 2;
    const float mean = n > 0 ? *wsum / float(n) : 0.f;
    float lsum = 0.f;
    const int64_t i0 = base + wave * 32;
    const int64_t i = i0 + r;
    // MFMA result element e of half t of this lane is row 8(e/4) + e%4 +
    // 4 * half, column 32t + r (dense64); its bit in a ReLU mask is 2e + t
    // (the backward pass produces delta_{l-1} in the layout in which the
    // forward produced h_{l-1}, so the masks stay in registers)
    const int64_t ib = i0 + 4 * half;
    {
        uint16_t* row = lx + r * kPitchX;
        if (i < n) {
            const float* q = rec + i * kRecordFloats;
            encode_input_half(half, mk(q[0], q[1], q[2]), mk(q[3], q[4], q[5]), mk(q[6], q[7], q[8]), box.lo,
                              box.inv, row);
        } else {
            for (int k = k_lo; k < k_hi; k++) row[k] = 0;
        }
    }
    __syncthreads();
    {
        // the feature-major copy for the weight gradients, after both halves of each row are written
        const uint16_t* row = lx + r * kPitchX;
        if (i < B)
            for (int k = k_lo; k < k_hi; k++) xT[int64_t(k) * B + i] = row[k];
    }
    uint32_t m1 = 0, m2 = 0, m3 = 0;   // ReLU masks of h1, h2, h3
    mlp_tile_forward(wl, P, lx, lh, ly, lane, [&](int layer, int e, int t, uint16_t v) {
        const int rc = 8 * (e >> 2) + (e & 3);
        if (ib + rc < B) hT[(int64_t(layer - 1) * kHid + 32 * t + r) * B + ib + rc] = v;
        const uint32_t bit = bf2f(v) > 0.f ? 1u << (2 * e + t) : 0u;
        if (layer == 1) m1 |= bit;
        else if (layer == 2) m2 |= bit;
        else m3 |= bit;
    });
    __syncthreads();
    // loss gradient of row r, lobes 4 * half .. 4 * half + 3
    float g[5 * 4];   // [0..3] logits, [4..7] concentrations, [8..19] mean directions
    for (int k = 0; k < 20; k++) g[k] = 0.f;
    {
        float pi[4], f[4], kap[4], raw_k[4], len[4], s = 0.f, wt = 0.f, pp = 0.f;
        f3 mu[4], rv[4];
        f3 wi = mk(0.f, 0.f, 1.f);
        if (i < n) {
            const float* y = ly + r * kPitchY;
            const float* q = rec + i * kRecordFloats;
            wt = mean > 0.f ? fminf(w[i] / mean, wclamp) : 0.f;
            s = wt / float(n);
            wi = normalize(mk(q[9], q[10], q[11]));
            float mx = y[0];
            for (int k = 1; k < kLobes; k++) mx = fmaxf(mx, y[k]);
            float se = 0.f;
            for (int k = 0; k < kLobes; k++) se += expf(y[k] - mx);
            const float inv = 1.f / se;
            for (int j = 0; j < 4; j++) {
                const int k = 4 * half + j;
                pi[j] = expf(y[k] - mx) * inv;
                raw_k[j] = y[kLobes + k];
                kap[j] = fminf(fmaxf(softplus(raw_k[j]), kKappaMin), kKappaMax);
                rv[j] = mk(y[2 * kLobes + 3 * k], y[2 * kLobes + 3 * k + 1], y[2 * kLobes + 3 * k + 2]);
                len[j] = length(rv[j]);
                mu[j] = len[j] > 1e-8f ? rv[j] / len[j] : mk(0.f, 0.f, 1.f);
                f[j] = vmf_pdf(kap[j], mu[j], wi);
                pp += pi[j] * f[j];
            }
        }
        const float p = pp + __shfl_xor(pp, 32);
        if (i < n && s > 0.f && p > 1e-30f) {
            if (half == 0) lsum += -wt * logf(p);
            for (int j = 0; j < 4; j++) {
                const float rk = pi[j] * f[j] / p;
                g[j] = s * (pi[j] - rk);
                const float sp = softplus(raw_k[j]);
                if (sp > kKappaMin && sp < kKappaMax) {
                    const float dlogf = 1.f / kap[j] - 2.f / expm1f(2.f * kap[j]) + (dot(mu[j], wi) - 1.f);
                    const float sig = 1.f / (1.f + expf(-raw_k[j]));
                    g[4 + j] = -s * rk * dlogf * sig;
                }
                if (len[j] > 1e-8f) {
                    const f3 gm = wi * (-s * rk * kap[j]);
                    const f3 gr = (gm - mu[j] * dot(gm, mu[j])) / len[j];
                    g[8 + 3 * j] = gr.x;
                    g[8 + 3 * j + 1] = gr.y;
                    g[8 + 3 * j + 2] = gr.z;
                }
            }
        }
    }
    __syncthreads();   // every lane has read the raw outputs that the deltas overwrite
    {
        // delta of the 64 (padded) outputs as the A operand of the backward pass, and d4T
        uint16_t* row = lx + r * kPitchX;
        uint16_t* d4 = dT + 3 * int64_t(kHid) * B;
        auto put = [&](int col, float v) {
            const uint16_t b = f2bf(v);
            row[col] = b;
            if (i < B) d4[int64_t(col) * B + i] = b;
        };
        for (int j = 0; j < 4; j++) {
            put(4 * half + j, g[j]);
            put(kLobes + 4 * half + j, g[4 + j]);
            for (int c = 0; c < 3; c++) put(2 * kLobes + 12 * half + 3 * j + c, g[8 + 3 * j + c]);
        }
        for (int c = kOut + 12 * half; c < kOut + 12 * (half + 1); c++) row[c] = 0;
    }
    __syncthreads();
    // backward: delta_{l-1} = (delta_l W_l) * relu'(h_{l-1}), l = 4, 3, 2
    uint16_t* in = lx;
    uint16_t* out = lh;
    int pin = kPitchX, pout = kPitchH;
#pragma unroll 1
    for (int l = 4; l >= 2; l--) {
        const uint32_t m = l == 4 ? m3 : (l == 3 ? m2 : m1);
        uint16_t* dsave = dT + int64_t(l - 2) * kHid * B;
        dense64<kHid, kHid, 1>(in, pin, WnBf + (l - 2) * kHid * kHid, zero64, lane,
                               [&](int row, int col, float v, int e, int t) {
            const uint16_t b = f2bf((m >> (2 * e + t)) & 1u ? v : 0.f);
            out[row * pout + col] = b;
            const int rc = 8 * (e >> 2) + (e & 3);
            if (ib + rc < B) dsave[int64_t(32 * t + r) * B + ib + rc] = b;
        });
        __syncthreads();
        uint16_t* t = in;
        in = out;
        out = t;
        const int tp = pin;
        pin = pout;
        pout = tp;
    }

    for (int sh = 32; sh > 0; sh >>= 1) lsum += __shfl_xor(lsum, sh);
    if (lane == 0) red[wave] = lsum;
    __syncthreads();
    if (threadIdx.x == 0 && n > 0) atomicAdd(loss, (red[0] + red[1] + red[2] + red[3]) / float(n));
}

// Weight gradients dWt_l[n][k] = sum_i d_l[n][i] a_{l-1}[k][i] for one
// (layer, 32x32 tile) job per blockIdx.y over a slice of samples per
// blockIdx.x; MFMA with the sample index as the reduction dimension,
// accumulated into grad with one atomic per element per slice. The jobs of
// input tile 0 also sum their delta operand, the bias gradient
// db_l[n] = sum_i d_l[n][i]. Block (0, 0) advances the Adam step count of a
// non-empty batch (the Adam kernel that follows reads it).
struct DwJob {
    int layer, nt, kt;
};
__constant__ DwJob c_dw_jobs[18];

__global__ __launch_bounds__(64) void k_guide_dw(int64_t n_max, const int32_t* ndev, int64_t off, int64_t B, int slice,
                                                 const uint16_t* __restrict__ xT, const uint16_t* __restrict__ hT,
                                                 const uint16_t* __restrict__ dT, float* __restrict__ grad,
                                                 int32_t* __restrict__ stepc) {
    const DwJob job = c_dw_jobs[blockIdx.y];
    const int lane = threadIdx.x, r = lane & 31, h = lane >> 5;
    const int K = job.layer == 1 ? kIn : kHid;
    const int N = job.layer == 4 ? kOut : kHid;
    const uint16_t* A = dT + int64_t(job.layer - 1) * kHid * B + int64_t(job.nt * 32 + r) * B;
    const uint16_t* Bm = (job.layer == 1 ? xT : hT + int64_t(job.layer - 2) * kHid * B) + int64_t(job.kt * 32 + r) * B;
    const int64_t rows = batch_rows(ndev, off, n_max);
    if (blockIdx.x == 0 && blockIdx.y == 0 && lane == 0 && rows > 0) *stepc += 1;
    const int64_t i_beg = int64_t(blockIdx.x) * slice;
    const int64_t i_end = std::min<int64_t>(round128(rows), i_beg + slice);
    if (i_end <= i_beg) return;   // a block past the batch (whole block)
    const bool bias = job.kt == 0;
    f32x16 acc = {};
    float bsum = 0.f;
    for (int64_t i = i_beg; i < i_end; i += 16) {
        const bf16x8 a = *reinterpret_cast<const bf16x8*>(A + i + 8 * h);
        const bf16x8 b = *reinterpret_cast<const bf16x8*>(Bm + i + 8 * h);
        acc = mfma_bf16(a, b, acc);
        if (bias) {
#pragma unroll
            for (int e = 0; e < 8; e++) bsum += float(a[e]);
        }
    }
    const int poff = job.layer == 1 ? kOffW1 : (job.layer == 2 ? kOffW2 : (job.layer == 3 ? kOffW3 : kOffW4));
#pragma unroll
    for (int e = 0; e < 16; e++) {
        const int n = job.nt * 32 + 8 * (e >> 2) + 4 * h + (e & 3);
        const int k = job.kt * 32 + r;
        if (n < N && k < K) atomicAdd(&grad[poff + n * K + k], acc[e]);
    }
    if (bias) {
        bsum += __shfl_xor(bsum, 32);
        const int boff = job.layer == 1 ? kOffB1 : (job.layer == 2 ? kOffB2 : (job.layer == 3 ? kOffB3 : kOffB4));
        const int n = job.nt * 32 + r;
        if (h == 0 && n < N) atomicAdd(&grad[boff + n], bsum);
    }
}

// Refreshes the BF16 copies: Wbf (packed, transposed weights) and WnBf
// (layers 2..4 as W[k][n], used by the backward pass).
__device__ __forceinline__ void write_bf16(int j, float p, uint16_t* Wbf, uint16_t* WnBf) {
    const uint16_t b = f2bf(p);
    Wbf[j] = b;
    int off = -1, l = 0;
    if (j >= kOffW2 && j < kOffB2) { off = kOffW2; l = 0; }
    else if (j >= kOffW3 && j < kOffB3) { off = kOffW3; l = 1; }
    else if (j >= kOffW4 && j < kOffB4) { off = kOffW4; l = 2; }
    if (off >= 0) {
        const int n = (j - off) / kHid, k = (j - off) % kHid;
        WnBf[l * kHid * kHid + k * kHid + n] = b;
    }
}

__global__ __launch_bounds__(256) void k_guide_sync(const float* __restrict__ P, uint16_t* __restrict__ Wbf,
                                                    uint16_t* __restrict__ WnBf) {
    const int j = blockIdx.x * 256 + threadIdx.x;
    if (j < kNParams) write_bf16(j, P[j], Wbf, WnBf);
}

// Adam (Kingma & Ba 2015) on the FP32 master weights; refreshes the BF16
// copies and clears the gradient and the target-weight sum for the next step
// (both are zero between steps). The step count t lives in device memory (k_guide_dw
// has advanced it); an empty batch leaves everything unchanged.
__global__ __launch_bounds__(256) void k_guide_adam(int64_t n_max, const int32_t* ndev, int64_t off,
                                                    const int32_t* __restrict__ stepc, float* __restrict__ P,
                                                    float* __restrict__ M, float* __restrict__ V,
                                                    float* __restrict__ G, float lr, float b1, float b2, float eps,
                                                    uint16_t* __restrict__ Wbf, uint16_t* __restrict__ WnBf,
                                                    float* __restrict__ wsum) {
    const int j = blockIdx.x * 256 + threadIdx.x;
    if (j == 0) *wsum = 0.f;
    if (j >= kNParams || batch_rows(ndev, off, n_max) == 0) return;
    const float t = float(*stepc);
    const float c1 = 1.f - powf(b1, t), c2 = 1.f - powf(b2, t);
    const float g = G[j];
    G[j] = 0.f;
    const float m = b1 * M[j] + (1.f - b1) * g;
    const float v = b2 * V[j] + (1.f - b2) * g * g;
    M[j] = m;
    V[j] = v;
    const float p = P[j] - lr * (m / c1) / (sqrtf(v / c2) + eps);
    P[j] = p;
    write_bf16(j, p, Wbf, WnBf);
}

__global__ void k_zero_f32(float* p, int64_t n) {
    const int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i < n) p[i] = 0.f;
}

__global__ __launch_bounds__(256) void k_guide_sample(int64_t n, const float* __restrict__ mix,
                                                      const float* __restrict__ u, float* __restrict__ dirs,
                                                      float* __restrict__ pdf) {
    const int64_t i = int64_t(blockIdx.x) * 256 + threadIdx.x;
    if (i >= n) return;
    const float* o = mix + i * kOut;
    const f3 d = mixture_sample(o, u[3 * i], u[3 * i + 1], u[3 * i + 2]);
    dirs[3 * i] = d.x;
    dirs[3 * i + 1] = d.y;
    dirs[3 * i + 2] = d.z;
    if (pdf) pdf[i] = mixture_pdf(o, d);
}

__global__ __launch_bounds__(256) void k_guide_pdf(int64_t n, const float* __restrict__ mix,
                                                   const float* __restrict__ dirs, float* __restrict__ pdf) {
    const int64_t i = int64_t(blockIdx.x) * 256 + threadIdx.x;
    if (i >= n) return;
    pdf[i] = mixture_pdf(mix + i * kOut, mk(dirs[3 * i], dirs[3 * i + 1], dirs[3 * i + 2]));
}

inline unsigned nb(int64_t n, int b) { return unsigned((n + b - 1) / b); }

}  // namespace

int Guide::create(const float* lo, const float* hi, uint64_t seed) {
    static std::atomic<uint64_t> next_uid{1};
    uid = next_uid.fetch_add(1);
    for (int a = 0; a < 3; a++) {
        if (!(hi[a] > lo[a])) return fail(LMGPU_ERR_INVALID_ARGUMENT, "guide box must have positive extent");
        box_lo[a] = lo[a];
        box_inv[a] = 1.f / (hi[a] - lo[a]);
    }
    LMGPU_TRY(P.alloc(kNParams));
    LMGPU_TRY(M.alloc(kNParams));
    LMGPU_TRY(V.alloc(kNParams));
    LMGPU_TRY(G.alloc(kNParams));
    LMGPU_TRY(Wbf.alloc(kNParams));
    LMGPU_TRY(WnBf.alloc(3 * kHid * kHid));
    LMGPU_TRY(zero64.alloc(64));
    LMGPU_TRY(scal.alloc(4));
    LMGPU_TRY(stepc.alloc(1));
    // (layer, output tile, input tile) of every 32x32 weight-gradient tile
    DwJob jobs[18];
    int c = 0;
    for (int l = 1; l <= 4; l++)
        for (int nt = 0; nt < 2; nt++)
            for (int kt = 0; kt < (l == 1 ? 3 : 2); kt++) jobs[c++] = DwJob{l, nt, kt};
    LMGPU_HIP(hipMemcpyToSymbol(HIP_SYMBOL(c_dw_jobs), jobs, sizeof(jobs)));
    LMGPU_HIP(hipMemset(zero64.p, 0, 64 * sizeof(float)));
    LMGPU_HIP(hipMemset(WnBf.p, 0, 3 * kHid * kHid * sizeof(uint16_t)));
    // torch.nn.Linear-style init: U(-1/sqrt(fan_in), 1/sqrt(fan_in)) for W and b
    std::vector<float> h(kNParams, 0.f);
    std::mt19937_64 rng(seed);
    auto fill = [&](int off, int nout, int nin, int nin_real, int nout_real) {
        const float b = 1.f / std::sqrt(float(nin_real));
        std::uniform_real_distribution<float> U(-b, b);
        for (int n = 0; n < nout; n++)
            for (int k = 0; k < nin; k++) h[off + n * nin + k] = (n < nout_real && k < nin_real) ? U(rng) : 0.f;
    };
    auto fillb = [&](int off, int nout_real, int nin_real) {
        const float b = 1.f / std::sqrt(float(nin_real));
        std::uniform_real_distribution<float> U(-b, b);
        for (int n = 0; n < nout_real; n++) h[off + n] = U(rng);
    };
    fill(kOffW1, kHid, kIn, 78, kHid);
    fillb(kOffB1, kHid, 78);
    fill(kOffW2, kHid, kHid, kHid, kHid);
    fillb(kOffB2, kHid, kHid);
    fill(kOffW3, kHid, kHid, kHid, kHid);
    fillb(kOffB3, kHid, kHid);
    fill(kOffW4, kOutPad, kHid, kHid, kOut);
    fillb(kOffB4, kOut, kHid);
    return set_params(h.data());
}

int Guide::set_params(const float* host) {
    std::vector<float> h(host, host + kNParams);
    // padding slots stay zero
    for (int n = 0; n < kHid; n++)
        for (int k = 78; k < kIn; k++) h[kOffW1 + n * kIn + k] = 0.f;
    for (int n = kOut; n < kOutPad; n++) {
        for (int k = 0; k < kHid; k++) h[kOffW4 + n * kHid + k] = 0.f;
        h[kOffB4 + n] = 0.f;
    }
    LMGPU_HIP(hipMemcpy(P.p, h.data(), sizeof(float) * kNParams, hipMemcpyHostToDevice));
    LMGPU_HIP(hipMemset(M.p, 0, sizeof(float) * kNParams));
    LMGPU_HIP(hipMemset(V.p, 0, sizeof(float) * kNParams));
    LMGPU_HIP(hipMemset(stepc.p, 0, sizeof(int32_t)));
    LMGPU_HIP(hipMemset(G.p, 0, sizeof(float) * kNParams));
    LMGPU_HIP(hipMemset(scal.p, 0, 4 * sizeof(float)));
    hipLaunchKernelGGL(k_guide_sync, dim3(nb(kNParams, 256)), dim3(256), 0, nullptr, P.p, Wbf.p, WnBf.p);
    LMGPU_HIP(hipGetLastError());
    LMGPU_HIP(hipDeviceSynchronize());
    return LMGPU_OK;
}

int Guide::get_params(float* host) const {
    LMGPU_HIP(hipMemcpy(host, P.p, sizeof(float) * kNParams, hipMemcpyDeviceToHost));
    return LMGPU_OK;
}

// persistent grid over 128-row tiles: two blocks per CU (LDS bound)
static int grid_for(int64_t n) {
    return int(std::max<int64_t>(1, std::min<int64_t>((n + kRowsPerBlock - 1) / kRowsPerBlock, 256 * 2)));
}

int Guide::infer(const float* query, int qs, float* mix, int64_t n, hipStream_t s) const {
    if (n <= 0) return LMGPU_OK;
    Box b;
    for (int a = 0; a < 3; a++) { b.lo[a] = box_lo[a]; b.inv[a] = box_inv[a]; }
    hipLaunchKernelGGL(k_guide_forward, dim3(grid_for(n)), dim3(kGB), 0, s, n, query, qs, Wbf.p, P.p, b, mix);
    LMGPU_HIP(hipGetLastError());
    return LMGPU_OK;
}

int Guide::ensure_train(int64_t B) {
    if (B % 128) return fail(LMGPU_ERR_INVALID_ARGUMENT, "training capacity must be a multiple of 128");
    if (B > (int64_t(1) << 22)) return fail(LMGPU_ERR_INVALID_ARGUMENT, "training batch above 2^22 records");
    if (B <= train_cap) return LMGPU_OK;
    LMGPU_TRY(xT.alloc(size_t(96) * B));
    LMGPU_TRY(hT.alloc(size_t(3) * kHid * B));
    LMGPU_TRY(dT.alloc(size_t(4) * kHid * B));
    LMGPU_TRY(w.alloc(B));
    // Input rows 80..95 and output-delta rows 40..63 are read by the 32-wide
    // weight-gradient tiles and never written: they stay zero.
    LMGPU_HIP(hipMemset(xT.p, 0, sizeof(uint16_t) * 96 * B));
    LMGPU_HIP(hipMemset(dT.p, 0, sizeof(uint16_t) * 4 * kHid * B));
    LMGPU_HIP(hipMemset(hT.p, 0, sizeof(uint16_t) * 3 * kHid * B));
    train_cap = B;
    gen++;
    return LMGPU_OK;
}

int Guide::enqueue_train(const float* rec, int64_t n, int64_t B, const int32_t* ndev, int64_t off, float lr,
                         float* grad_host, hipStream_t s, float clamp) {
    if (B > train_cap || n > B) return fail(LMGPU_ERR_INVALID_ARGUMENT, "training batch exceeds the workspace");
    // four launches per step: target weights summed (and the loss cleared);
    // the rows (forward, loss gradient, backward); weight and bias gradients
    // with the step count; Adam, which also clears the gradient and the weight
    // sum for the next step
    hipLaunchKernelGGL(k_guide_weights, dim3(nb(n, 256)), dim3(256), 0, s, n, ndev, off, rec, w.p, scal.p,
                       scal.p + 1);
    Box b;
    for (int a = 0; a < 3; a++) { b.lo[a] = box_lo[a]; b.inv[a] = box_inv[a]; }
    hipLaunchKernelGGL(k_guide_train_rows, dim3(unsigned(B / kRowsPerBlock)), dim3(kGB), 0, s, n, ndev, off, B, rec, Wbf.p, P.p,
                       WnBf.p, zero64.p, b, w.p, scal.p, clamp > 0.f ? clamp : wclamp, xT.p, hT.p, dT.p, scal.p + 1);
    const int slice = LMGPU_DW_SLICE;
    hipLaunchKernelGGL(k_guide_dw, dim3(nb(B, slice), 18), dim3(64), 0, s, n, ndev, off, B, slice, xT.p, hT.p, dT.p,
                       G.p, stepc.p);
    if (grad_host) LMGPU_HIP(hipMemcpyAsync(grad_host, G.p, sizeof(float) * kNParams, hipMemcpyDeviceToHost, s));
    hipLaunchKernelGGL(k_guide_adam, dim3(nb(kNParams, 256)), dim3(256), 0, s, n, ndev, off, stepc.p, P.p, M.p, V.p,
                       G.p, lr, 0.9f, 0.999f, 1e-8f, Wbf.p, WnBf.p, scal.p);
    LMGPU_HIP(hipGetLastError());
    return LMGPU_OK;
}

int Guide::train(const float* rec, int64_t n, float lr, float* loss_out, float* grad_out, hipStream_t s) {
    if (n <= 0) return fail(LMGPU_ERR_INVALID_ARGUMENT, "empty training batch");
    // batch padded to whole 128-row tiles
    const int64_t B = (n + 127) / 128 * 128;
    LMGPU_TRY(ensure_train(B));
    LMGPU_TRY(enqueue_train(rec, n, B, nullptr, 0, lr, grad_out, s));
    if (loss_out) {
        float l = 0.f;
        LMGPU_HIP(hipMemcpyAsync(&l, scal.p + 1, sizeof(float), hipMemcpyDeviceToHost, s));
        LMGPU_HIP(hipStreamSynchronize(s));
        *loss_out = l;
    } else if (grad_out) {
        LMGPU_HIP(hipStreamSynchronize(s));
    }
    return LMGPU_OK;
}

}  // namespace lmgpu

using lmgpu::fail;

extern "C" {

int lmgpu_guide_create(lmgpu_guide** out, const float* box_lo, const float* box_hi, uint64_t seed) {
    if (!out || !box_lo || !box_hi) return fail(LMGPU_ERR_INVALID_ARGUMENT, "NULL argument");
    *out = nullptr;
    auto* h = new (std::nothrow) lmgpu_guide();
    if (!h) return fail(LMGPU_ERR_OUT_OF_MEMORY, "guide handle");
    const int rc = h->g.create(box_lo, box_hi, seed);
    if (rc != LMGPU_OK) {
        delete h;
        return rc;
    }
    *out = h;
    return LMGPU_OK;
}

int lmgpu_guide_destroy(lmgpu_guide* guide) {
    delete guide;
    return LMGPU_OK;
}

int lmgpu_guide_params(lmgpu_guide* guide, float* get, const float* set) {
    if (!guide) return fail(LMGPU_ERR_INVALID_ARGUMENT, "guide is NULL");
    if (set) LMGPU_TRY(guide->g.set_params(set));
    if (get) LMGPU_TRY(guide->g.get_params(get));
    return LMGPU_OK;
}

int lmgpu_guide_infer(const lmgpu_guide* guide, const float* query, float* mix, int64_t n, void* stream) {
    if (!guide || (n > 0 && (!query || !mix))) return fail(LMGPU_ERR_INVALID_ARGUMENT, "NULL argument");
    if (n < 0) return fail(LMGPU_ERR_INVALID_ARGUMENT, "negative count");
    return guide->g.infer(query, 9, mix, n, static_cast<hipStream_t>(stream));
}

int lmgpu_guide_sample(const float* mix, const float* u, float* dirs, float* pdf, int64_t n, void* stream) {
    if (n < 0 || (n > 0 && (!mix || !u || !dirs))) return fail(LMGPU_ERR_INVALID_ARGUMENT, "bad arguments");
    if (n == 0) return LMGPU_OK;
    hipLaunchKernelGGL(lmgpu::k_guide_sample, dim3(lmgpu::nb(n, 256)), dim3(256), 0, static_cast<hipStream_t>(stream),
                       n, mix, u, dirs, pdf);
    LMGPU_HIP(hipGetLastError());
    return LMGPU_OK;
}

int lmgpu_guide_pdf(const float* mix, const float* dirs, float* pdf, int64_t n, void* stream) {
    if (n < 0 || (n > 0 && (!mix || !dirs || !pdf))) return fail(LMGPU_ERR_INVALID_ARGUMENT, "bad arguments");
    if (n == 0) return LMGPU_OK;
    hipLaunchKernelGGL(lmgpu::k_guide_pdf, dim3(lmgpu::nb(n, 256)), dim3(256), 0, static_cast<hipStream_t>(stream), n,
                       mix, dirs, pdf);
    LMGPU_HIP(hipGetLastError());
    return LMGPU_OK;
}

int lmgpu_guide_stats(const lmgpu_guide* guide, int64_t* steps, float* last_loss) {
    if (!guide) return fail(LMGPU_ERR_INVALID_ARGUMENT, "guide is NULL");
    if (steps) {
        int32_t t = 0;
        LMGPU_HIP(hipMemcpy(&t, guide->g.stepc.p, sizeof(t), hipMemcpyDeviceToHost));
        *steps = t;
    }
    if (last_loss) LMGPU_HIP(hipMemcpy(last_loss, guide->g.scal.p + 1, sizeof(float), hipMemcpyDeviceToHost));
    return LMGPU_OK;
}

int lmgpu_guide_train(lmgpu_guide* guide, const float* records, int64_t n, float lr, float* loss, float* grads,
                      void* stream) {
    if (!guide || !records) return fail(LMGPU_ERR_INVALID_ARGUMENT, "NULL argument");
    if (!(lr >= 0.f)) return fail(LMGPU_ERR_INVALID_ARGUMENT, "learning rate must be >= 0");
    return guide->g.train(records, n, lr, loss, grads, static_cast<hipStream_t>(stream));
}

}  // extern "C"
