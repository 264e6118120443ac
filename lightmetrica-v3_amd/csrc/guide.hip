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
// floats). INFER: decoded mixtures mix[i][40] = pi[8], kappa[8], mu[8][3].
// TRAIN: saves xT[96][B], hT[3][64][B] (bf16, feature major) and raw y[i][40].
template <bool TRAIN>
__global__ __launch_bounds__(kGB) void k_guide_forward(int64_t n_max, const int32_t* ndev, int64_t off, int64_t B,
                                                       const float* __restrict__ query, int qs,
                                                       const uint16_t* __restrict__ Wbf, const float* __restrict__ P,
                                                       Box box, float* __restrict__ mix, uint16_t* __restrict__ xT,
                                                       uint16_t* __restrict__ hT, float* __restrict__ yout) {
    __shared__ __attribute__((aligned(16))) unsigned char lds[kWeightLds + 4 * kWaveLds];
    const int64_t n = batch_rows(ndev, off, n_max);
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    uint16_t* wl = reinterpret_cast<uint16_t*>(lds);
    stage_weights(Wbf, wl, threadIdx.x, kGB);
    uint16_t* lx = reinterpret_cast<uint16_t*>(lds + kWeightLds + wave * kWaveLds);
    uint16_t* lh = lx + 32 * kPitchX;
    float* ly = reinterpret_cast<float*>(lx);
    // two lanes per row: lanes r and r + 32 share row r of the wave's 32-row
    // tile and each encode, save and decode half of it
    const int r = lane & 31, half = lane >> 5;
    const int k_lo = half * (kIn / 2), k_hi = k_lo + kIn / 2;   // feature half for zeroing and saving
    for (int64_t base = int64_t(blockIdx.x) * kRowsPerBlock; base < n; base += int64_t(gridDim.x) * kRowsPerBlock) {
        const int64_t i0 = base + wave * 32;
        {
            const int64_t i = i0 + r;
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
        if (TRAIN) {
            // the feature-major copy for the weight gradients, after both halves of each row are written
            const int64_t i = i0 + r;
            const uint16_t* row = lx + r * kPitchX;
            if (i < B)
                for (int k = k_lo; k < k_hi; k++) xT[int64_t(k) * B + i] = row[k];
        }
        mlp_tile_forward(wl, P, lx, lh, ly, lane, [&](int layer, int row, int col, uint16_t v) {
            if (TRAIN) {
                const int64_t i = i0 + row;
                if (i < B) hT[(int64_t(layer - 1) * kHid + col) * B + i] = v;
            }
        });
        __syncthreads();
        {
            const int64_t i = i0 + r;
            if (i < n) {
                const float* y = ly + r * kPitchY;
                if (TRAIN) {
                    for (int k = half * (kOut / 2); k < (half + 1) * (kOut / 2); k++) yout[i * kOut + k] = y[k];
                } else {
                    decode_lobes4(y, 4 * half, mix + i * kOut);
                }
            }
        }
        __syncthreads();
    }
}

// Target weights w_i = lum(L_i) / p_mix_i of training records
// rec[i] = pos(3), nrm(3), wo(3), wi(3), p_mix, L(3); sum into *wsum.
__global__ __launch_bounds__(256) void k_guide_weights(int64_t n_max, const int32_t* ndev, int64_t off,
                                                       const float* __restrict__ rec, float* __restrict__ w,
                                                       float* __restrict__ wsum) {
    __shared__ float red[4];
    const int64_t n = batch_rows(ndev, off, n_max);
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

// Per-sample gradient of -wt log p_guide(w_i) w.r.t. the 40 raw outputs,
// scaled by 1/n, written feature-major as bf16 into d4T[64][B]; rows from n
// to the end of the last 128-row tile are zeroed.
__global__ __launch_bounds__(256) void k_guide_loss(int64_t n_max, const int32_t* ndev, int64_t off, int64_t B,
                                                    const float* __restrict__ y, const float* __restrict__ rec,
                                                    const float* __restrict__ w, const float* __restrict__ wsum,
                                                    float wclamp, uint16_t* __restrict__ d4T,
                                                    float* __restrict__ loss) {
    __shared__ float red[4];
    const int64_t n = batch_rows(ndev, off, n_max);
    const int64_t i = int64_t(blockIdx.x) * 256 + threadIdx.x;
    float li = 0.f;
    if (i >= n && i < round128(n))
        for (int k = 0; k < kOut; k++) d4T[int64_t(k) * B + i] = 0;
    if (i < n) {
        const float mean = *wsum / float(n);
        const float wt = mean > 0.f ? fminf(w[i] / mean, wclamp) : 0.f;
        const float s = wt / float(n);
        const float* yi = y + i * kOut;
        const float* r = rec + i * kRecordFloats;
        const f3 wi = normalize(mk(r[9], r[10], r[11]));
        Mixture m;
        decode_mixture(yi, m);
        float f[kLobes];
        float p = 0.f;
        for (int k = 0; k < kLobes; k++) {
            f[k] = vmf_pdf(m.kappa[k], m.mu[k], wi);
            p += m.pi[k] * f[k];
        }
        float g[kOut];
        for (int k = 0; k < kOut; k++) g[k] = 0.f;
        if (s > 0.f && p > 1e-30f) {
            li = -wt * logf(p);
            for (int k = 0; k < kLobes; k++) {
                const float rk = m.pi[k] * f[k] / p;
                g[k] = s * (m.pi[k] - rk);
                const float kap = m.kappa[k];
                const float raw = yi[kLobes + k];
                const float sp = softplus(raw);
                if (sp > kKappaMin && sp < kKappaMax) {
                    const float dlogf = 1.f / kap - 2.f / expm1f(2.f * kap) + (dot(m.mu[k], wi) - 1.f);
                    const float sig = 1.f / (1.f + expf(-raw));
                    g[kLobes + k] = -s * rk * dlogf * sig;
                }
                const f3 rv = mk(yi[2 * kLobes + 3 * k], yi[2 * kLobes + 3 * k + 1], yi[2 * kLobes + 3 * k + 2]);
                const float len = length(rv);
                if (len > 1e-8f) {
                    const f3 gm = wi * (-s * rk * kap);
                    const f3 gr = (gm - m.mu[k] * dot(gm, m.mu[k])) / len;
                    g[2 * kLobes + 3 * k] = gr.x;
                    g[2 * kLobes + 3 * k + 1] = gr.y;
                    g[2 * kLobes + 3 * k + 2] = gr.z;
                }
            }
        }
        for (int k = 0; k < kOut; k++) d4T[int64_t(k) * B + i] = f2bf(g[k]);
    }
    for (int sh = 32; sh > 0; sh >>= 1) li += __shfl_xor(li, sh);
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = li;
    __syncthreads();
    if (threadIdx.x == 0 && n > 0) atomicAdd(loss, (red[0] + red[1] + red[2] + red[3]) / float(n));
}

// Backward through layers 4..2 for 32-row wave tiles: delta_{l-1} =
// (delta_l W_l^T) * relu'(h_{l-1}); saves d3T, d2T, d1T (bf16, feature major).
__global__ __launch_bounds__(kGB) void k_guide_backward(int64_t n_max, const int32_t* ndev, int64_t off, int64_t B,
                                                        const uint16_t* __restrict__ WnBf,
                                                        const float* __restrict__ zero64,
                                                        const uint16_t* __restrict__ hT, uint16_t* __restrict__ dT) {
    constexpr int kWnLds = 3 * kHid * kPWh * 2;   // W2..W4 as [k][n], rows padded to kPWh
    __shared__ __attribute__((aligned(16))) unsigned char lds[kWnLds + 4 * kWaveLds];
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    uint16_t* wn = reinterpret_cast<uint16_t*>(lds);
    for (int c = threadIdx.x; c < 3 * kHid * (kHid / 8); c += kGB) {
        const int r = c / (kHid / 8), k = (c % (kHid / 8)) * 8;   // r over the 3 x 64 rows
        *reinterpret_cast<uint4*>(wn + r * kPWh + k) = *reinterpret_cast<const uint4*>(WnBf + r * kHid + k);
    }
    __syncthreads();
    uint16_t* lx = reinterpret_cast<uint16_t*>(lds + kWnLds + wave * kWaveLds);
    uint16_t* lh = lx + 32 * kPitchX;
    const int64_t rows = round128(batch_rows(ndev, off, n_max));
    for (int64_t base = int64_t(blockIdx.x) * kRowsPerBlock; base < rows; base += int64_t(gridDim.x) * kRowsPerBlock) {
        const int64_t i0 = base + wave * 32;
        if (lane < 32) {
            const int64_t i = i0 + lane;
            for (int k = 0; k < kOutPad; k++)
                lx[lane * kPitchX + k] = (i < B && k < kOut) ? dT[(3 * int64_t(kHid) + k) * B + i] : uint16_t(0);
        }
        __syncthreads();
        // layer l = 4, 3, 2: input delta in `in`, output to `out`, saved as d_{l-1}
        uint16_t* in = lx;
        uint16_t* out = lh;
        int pin = kPitchX, pout = kPitchH;
        for (int l = 4; l >= 2; l--) {
            const uint16_t* Wn = wn + (l - 2) * kHid * kPWh;
            const uint16_t* hprev = hT + int64_t(l - 2) * kHid * B;   // h_{l-1}
            uint16_t* dsave = dT + int64_t(l - 2) * kHid * B;          // d_{l-1}
            dense64<kHid, kPWh, 1>(in, pin, Wn, zero64, lane, [&](int row, int col, float v) {
                const int64_t i = i0 + row;
                float g = 0.f;
                if (i < B && bf2f(hprev[int64_t(col) * B + i]) > 0.f) g = v;
                const uint16_t b = f2bf(g);
                out[row * pout + col] = b;
                if (i < B) dsave[int64_t(col) * B + i] = b;
            });
            __syncthreads();
            uint16_t* t = in;
            in = out;
            out = t;
            const int tp = pin;
            pin = pout;
            pout = tp;
        }
    }
}

// Weight gradients dWt_l[n][k] = sum_i d_l[n][i] a_{l-1}[k][i] for one
// (layer, 32x32 tile) job per blockIdx.y over a slice of samples per
// blockIdx.x; MFMA with the sample index as the reduction dimension,
// accumulated into grad with one atomic per element per slice.
struct DwJob {
    int layer, nt, kt;
};
__constant__ DwJob c_dw_jobs[18];

__global__ __launch_bounds__(64) void k_guide_dw(int64_t n_max, const int32_t* ndev, int64_t off, int64_t B, int slice,
                                                 const uint16_t* __restrict__ xT, const uint16_t* __restrict__ hT,
                                                 const uint16_t* __restrict__ dT, float* __restrict__ grad) {
    const DwJob job = c_dw_jobs[blockIdx.y];
    const int lane = threadIdx.x, r = lane & 31, h = lane >> 5;
    const int K = job.layer == 1 ? kIn : kHid;
    const int N = job.layer == 4 ? kOut : kHid;
    const uint16_t* A = dT + int64_t(job.layer - 1) * kHid * B + int64_t(job.nt * 32 + r) * B;
    const uint16_t* Bm = (job.layer == 1 ? xT : hT + int64_t(job.layer - 2) * kHid * B) + int64_t(job.kt * 32 + r) * B;
    const int64_t i_beg = int64_t(blockIdx.x) * slice;
    const int64_t i_end = std::min<int64_t>(round128(batch_rows(ndev, off, n_max)), i_beg + slice);
    f32x16 acc = {};
    for (int64_t i = i_beg; i < i_end; i += 16) {
        const bf16x8 a = *reinterpret_cast<const bf16x8*>(A + i + 8 * h);
        const bf16x8 b = *reinterpret_cast<const bf16x8*>(Bm + i + 8 * h);
        acc = mfma_bf16(a, b, acc);
    }
    const int poff = job.layer == 1 ? kOffW1 : (job.layer == 2 ? kOffW2 : (job.layer == 3 ? kOffW3 : kOffW4));
#pragma unroll
    for (int e = 0; e < 16; e++) {
        const int n = job.nt * 32 + 8 * (e >> 2) + 4 * h + (e & 3);
        const int k = job.kt * 32 + r;
        if (n < N && k < K) atomicAdd(&grad[poff + n * K + k], acc[e]);
    }
}

// Bias gradients: db_l[n] = sum_i d_l[n][i]; blocks (layer * 64 + n, chunk)
// read 8 bf16 per lane and add their partial sum atomically (G starts at 0).
__global__ __launch_bounds__(256) void k_guide_db(int64_t n_max, const int32_t* ndev, int64_t off, int64_t B,
                                                  const uint16_t* __restrict__ dT, float* __restrict__ grad) {
    __shared__ float red[4];
    const int layer = blockIdx.x / kHid + 1, n = blockIdx.x % kHid;
    if (layer == 4 && n >= kOut) return;   // whole block: padded output rows
    const uint4* d = reinterpret_cast<const uint4*>(dT + int64_t(blockIdx.x) * B);
    const int64_t nvec = round128(batch_rows(ndev, off, n_max)) / 8;
    float s = 0.f;
    for (int64_t v = int64_t(blockIdx.y) * 256 + threadIdx.x; v < nvec; v += int64_t(gridDim.y) * 256) {
        const uint4 q = d[v];
        const uint32_t w[4] = {q.x, q.y, q.z, q.w};
#pragma unroll
        for (int k = 0; k < 4; k++) {
            float lo, hi;
            const uint32_t a = w[k] << 16, b = w[k] & 0xffff0000u;
            memcpy(&lo, &a, 4);
            memcpy(&hi, &b, 4);
            s += lo + hi;
        }
    }
    for (int sh = 32; sh > 0; sh >>= 1) s += __shfl_xor(s, sh);
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = s;
    __syncthreads();
    if (threadIdx.x == 0) {
        const int poff = layer == 1 ? kOffB1 : (layer == 2 ? kOffB2 : (layer == 3 ? kOffB3 : kOffB4));
        atomicAdd(&grad[poff + n], red[0] + red[1] + red[2] + red[3]);
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
// copies. The step count t lives in device memory (k_guide_step advances it
// afterwards); an empty batch leaves everything unchanged.
__global__ __launch_bounds__(256) void k_guide_adam(int64_t n_max, const int32_t* ndev, int64_t off,
                                                    const int32_t* __restrict__ stepc, float* __restrict__ P,
                                                    float* __restrict__ M, float* __restrict__ V,
                                                    const float* __restrict__ G, float lr, float b1, float b2, float eps,
                                                    uint16_t* __restrict__ Wbf, uint16_t* __restrict__ WnBf) {
    const int j = blockIdx.x * 256 + threadIdx.x;
    if (j >= kNParams || batch_rows(ndev, off, n_max) == 0) return;
    const float t = float(*stepc + 1);
    const float c1 = 1.f - powf(b1, t), c2 = 1.f - powf(b2, t);
    const float g = G[j];
    const float m = b1 * M[j] + (1.f - b1) * g;
    const float v = b2 * V[j] + (1.f - b2) * g * g;
    M[j] = m;
    V[j] = v;
    const float p = P[j] - lr * (m / c1) / (sqrtf(v / c2) + eps);
    P[j] = p;
    write_bf16(j, p, Wbf, WnBf);
}

__global__ void k_guide_step(int64_t n_max, const int32_t* ndev, int64_t off, int32_t* stepc) {
    if (threadIdx.x == 0 && batch_rows(ndev, off, n_max) > 0) *stepc += 1;
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
    hipLaunchKernelGGL(k_guide_forward<false>, dim3(grid_for(n)), dim3(kGB), 0, s, n, nullptr, int64_t(0), int64_t(0),
                       query, qs, Wbf.p, P.p, b, mix, nullptr, nullptr, nullptr);
    LMGPU_HIP(hipGetLastError());
    return LMGPU_OK;
}

int Guide::ensure_train(int64_t B) {
    if (B % 128) return fail(LMGPU_ERR_INVALID_ARGUMENT, "training capacity must be a multiple of 128");
    if (B <= train_cap) return LMGPU_OK;
    LMGPU_TRY(xT.alloc(size_t(96) * B));
    LMGPU_TRY(hT.alloc(size_t(3) * kHid * B));
    LMGPU_TRY(dT.alloc(size_t(4) * kHid * B));
    LMGPU_TRY(y.alloc(size_t(kOut) * B));
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
    hipLaunchKernelGGL(k_zero_f32, dim3(1), dim3(64), 0, s, scal.p, int64_t(4));
    hipLaunchKernelGGL(k_zero_f32, dim3(nb(kNParams, 256)), dim3(256), 0, s, G.p, int64_t(kNParams));
    hipLaunchKernelGGL(k_guide_weights, dim3(nb(n, 256)), dim3(256), 0, s, n, ndev, off, rec, w.p, scal.p);
    Box b;
    for (int a = 0; a < 3; a++) { b.lo[a] = box_lo[a]; b.inv[a] = box_inv[a]; }
    hipLaunchKernelGGL(k_guide_forward<true>, dim3(grid_for(B)), dim3(kGB), 0, s, n, ndev, off, B, rec, kRecordFloats,
                       Wbf.p, P.p, b, nullptr, xT.p, hT.p, y.p);
    hipLaunchKernelGGL(k_guide_loss, dim3(nb(B, 256)), dim3(256), 0, s, n, ndev, off, B, y.p, rec, w.p, scal.p,
                       clamp > 0.f ? clamp : wclamp,
                       dT.p + 3 * kHid * B, scal.p + 1);
    hipLaunchKernelGGL(k_guide_backward, dim3(grid_for(B)), dim3(kGB), 0, s, n, ndev, off, B, WnBf.p, zero64.p, hT.p,
                       dT.p);
    const int slice = 4096;
    hipLaunchKernelGGL(k_guide_dw, dim3(nb(B, slice), 18), dim3(64), 0, s, n, ndev, off, B, slice, xT.p, hT.p, dT.p,
                       G.p);
    hipLaunchKernelGGL(k_guide_db, dim3(4 * kHid, unsigned(std::min<int64_t>(16, (B / 8 + 255) / 256))), dim3(256), 0,
                       s, n, ndev, off, B, dT.p, G.p);
    if (grad_host) LMGPU_HIP(hipMemcpyAsync(grad_host, G.p, sizeof(float) * kNParams, hipMemcpyDeviceToHost, s));
    hipLaunchKernelGGL(k_guide_adam, dim3(nb(kNParams, 256)), dim3(256), 0, s, n, ndev, off, stepc.p, P.p, M.p, V.p,
                       G.p, lr, 0.9f, 0.999f, 1e-8f, Wbf.p, WnBf.p);
    hipLaunchKernelGGL(k_guide_step, dim3(1), dim3(64), 0, s, n, ndev, off, stepc.p);
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
