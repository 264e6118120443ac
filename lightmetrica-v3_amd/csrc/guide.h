// Learned path-guiding distribution: device-side definitions.
//
// The guide maps a surface point (position, shading normal, outgoing
// direction) to a mixture of kLobes von Mises-Fisher lobes over incident
// directions. Network: 12-frequency sin/cos encoding of the position in the
// scene box (72) + normal (3) + outgoing direction (3) = 78 inputs (padded to
// 80), three 64-wide ReLU layers, a 40-wide linear output: per lobe a weight
// logit, a concentration (softplus) and a mean direction (normalised).
// Matrix products run on MFMA (v_mfma_f32_32x32x16_bf16: BF16 operands, FP32
// accumulation); master weights and Adam state are FP32.
#pragma once

#include "common.h"

namespace lmgpu {

constexpr int kLobes = 8;
constexpr int kFreq = 12;
constexpr int kIn = 80;      // 78 inputs + 2 zero pads
constexpr int kHid = 64;
constexpr int kOut = 40;     // logits[8], kappa[8], mu[8][3]
constexpr int kOutPad = 64;
// Packed parameter layout (FP32 master copy), weights stored transposed
// (output neuron major): Wt1[64][80], b1[64], Wt2[64][64], b2[64],
// Wt3[64][64], b3[64], Wt4[64][64] (rows >= 40 zero), b4[64] (>= 40 zero).
constexpr int kOffW1 = 0;
constexpr int kOffB1 = kOffW1 + kHid * kIn;
constexpr int kOffW2 = kOffB1 + kHid;
constexpr int kOffB2 = kOffW2 + kHid * kHid;
constexpr int kOffW3 = kOffB2 + kHid;
constexpr int kOffB3 = kOffW3 + kHid * kHid;
constexpr int kOffW4 = kOffB3 + kHid;
constexpr int kOffB4 = kOffW4 + kOutPad * kHid;
constexpr int kNParams = kOffB4 + kOutPad;   // 17664 slots, 15976 trained
constexpr float kKappaMin = 0.05f, kKappaMax = 4000.f;

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float f32x16 __attribute__((ext_vector_type(16)));

// Round-to-nearest-even FP32 -> BF16 bits (finite inputs): one
// v_cvt_pk_bf16_f32 on the device, the same rounding by integer arithmetic
// on the host.
LMGPU_HD uint16_t f2bf(float f) {
#if defined(__HIP_DEVICE_COMPILE__)
    const __bf16 b = static_cast<__bf16>(f);
    uint16_t h;
    __builtin_memcpy(&h, &b, 2);
    return h;
#else
    uint32_t u;
    memcpy(&u, &f, 4);
    u += 0x7fffu + ((u >> 16) & 1u);
    return uint16_t(u >> 16);
#endif
}
LMGPU_HD float bf2f(uint16_t h) {
    const uint32_t u = uint32_t(h) << 16;
    float f;
    memcpy(&f, &u, 4);
    return f;
}

// Decoded mixture of one query point.
struct Mixture {
    float pi[kLobes];
    float kappa[kLobes];
    f3 mu[kLobes];
};

__device__ __forceinline__ float softplus(float x) { return x > 20.f ? x : log1pf(expf(x)); }

// raw network outputs (40 floats) -> mixture parameters
__device__ __forceinline__ void decode_mixture(const float* y, Mixture& m) {
    float mx = y[0];
    for (int k = 1; k < kLobes; k++) mx = fmaxf(mx, y[k]);
    float s = 0.f;
    for (int k = 0; k < kLobes; k++) {
        m.pi[k] = expf(y[k] - mx);
        s += m.pi[k];
    }
    const float inv = 1.f / s;
    for (int k = 0; k < kLobes; k++) {
        m.pi[k] *= inv;
        m.kappa[k] = fminf(fmaxf(softplus(y[kLobes + k]), kKappaMin), kKappaMax);
        const f3 r = mk(y[2 * kLobes + 3 * k], y[2 * kLobes + 3 * k + 1], y[2 * kLobes + 3 * k + 2]);
        const float l = length(r);
        m.mu[k] = l > 1e-8f ? r / l : mk(0.f, 0.f, 1.f);
    }
}

// vMF density on the unit sphere (solid angle), stable for large kappa:
// kappa / (2 pi (1 - exp(-2 kappa))) * exp(kappa (mu.w - 1)).
__device__ __forceinline__ float vmf_pdf(float kappa, f3 mu, f3 w) {
    return kappa / (2.f * 3.14159265358979f * (-expm1f(-2.f * kappa))) * expf(kappa * (dot(mu, w) - 1.f));
}

// Density (solid angle) and sampling of a decoded mixture (pi[8], kappa[8],
// mu[8][3]) read lobe by lobe from memory. Sampling picks the lobe by u0 and
// the direction by (u1, u2) with the stable vMF inversion of W. Jakob,
// "Numerically stable sampling of the von Mises-Fisher distribution on S^2"
// (2012).
__device__ __forceinline__ float mixture_pdf(const float* __restrict__ o, f3 w) {
    float p = 0.f;
    for (int k = 0; k < kLobes; k++) {
        const f3 mu = mk(o[2 * kLobes + 3 * k], o[2 * kLobes + 3 * k + 1], o[2 * kLobes + 3 * k + 2]);
        p += o[k] * vmf_pdf(o[kLobes + k], mu, w);
    }
    return p;
}

__device__ __forceinline__ f3 mixture_sample(const float* __restrict__ o, float u0, float u1, float u2) {
    int k = 0;
    float c = o[0];
    while (k < kLobes - 1 && u0 >= c) {
        k++;
        c += o[k];
    }
    const float kap = o[kLobes + k];
    const float W = fmaxf(-1.f, 1.f + logf(u1 + (1.f - u1) * expf(-2.f * kap)) / kap);
    const float sw = sqrtf(fmaxf(0.f, 1.f - W * W));
    float s, co;
    sincosf(2.f * 3.14159265358979f * u2, &s, &co);
    const f3 mu = mk(o[2 * kLobes + 3 * k], o[2 * kLobes + 3 * k + 1], o[2 * kLobes + 3 * k + 2]);
    const float sg = copysignf(1.0f, mu.z);
    const float a = -1.0f / (sg + mu.z);
    const float b = mu.x * mu.y * a;
    const f3 t1 = mk(1.0f + sg * mu.x * mu.x * a, sg * b, -sg * mu.x);
    const f3 t2 = mk(b, sg + mu.y * mu.y * a, -mu.y);
    return normalize(mu * W + (t1 * co + t2 * s) * sw);
}

// Input encoding of one query point into 80 features (bf16 bits).
// Position normalised to [0,1]^3 in the scene box; sin and cos of
// 2^j * f, f = pi * q rounded to FP32, for j < kFreq. sincos is evaluated
// at every third octave and the two octaves after it follow from the
// double-angle identities sin 2x = 2 sin x cos x, cos 2x = 1 - 2 sin^2 x,
// which at most quadruple the FP32 error of the anchor (below 1e-6).
// Layout: features 24a .. 24a + 23 for axis a, then normal (72..74),
// outgoing direction (75..77) and two zero pads.
constexpr int kTailFeature = 6 * kFreq;   // 72

__device__ __forceinline__ void scene_coords(f3 p, const float* box_lo, const float* box_inv, float (&q)[3]) {
    q[0] = (p.x - box_lo[0]) * box_inv[0];
    q[1] = (p.y - box_lo[1]) * box_inv[1];
    q[2] = (p.z - box_lo[2]) * box_inv[2];
}

// features of axes [A0, A1) into out[2 * kFreq * A0 ...]
template <int A0, int A1>
__device__ __forceinline__ void encode_axes(const float (&q)[3], uint16_t* out) {
    static_assert(kFreq % 3 == 0, "octaves come in groups of three");
    int o = 2 * kFreq * A0;
#pragma unroll
    for (int a = A0; a < A1; a++) {
        const float f = 3.14159265358979f * q[a];
#pragma unroll
        for (int j0 = 0; j0 < kFreq; j0 += 3) {
            float s, c;
            sincosf(ldexpf(f, j0), &s, &c);
#pragma unroll
            for (int j = 0; j < 3; j++) {
                out[o++] = f2bf(s);
                out[o++] = f2bf(c);
                const float s2 = 2.f * s * c;
                c = fmaf(-2.f * s, s, 1.f);
                s = s2;
            }
        }
    }
}

// normal, outgoing direction and pads into out[kTailFeature ...]
__device__ __forceinline__ void encode_tail(f3 n, f3 wo, uint16_t* out) {
    uint16_t* t = out + kTailFeature;
    t[0] = f2bf(n.x);
    t[1] = f2bf(n.y);
    t[2] = f2bf(n.z);
    t[3] = f2bf(wo.x);
    t[4] = f2bf(wo.y);
    t[5] = f2bf(wo.z);
    t[6] = 0;
    t[7] = 0;
}

// Half of the encoding for a pair of lanes sharing one row: half 0 writes
// axes x and y (features 0..47), half 1 axis z, normal, direction and pads
// (48..79).
__device__ __forceinline__ void encode_input_half(int half, f3 p, f3 n, f3 wo, const float* box_lo,
                                                  const float* box_inv, uint16_t* out) {
    float q[3];
    scene_coords(p, box_lo, box_inv, q);
    if (half == 0) {
        encode_axes<0, 2>(q, out);
    } else {
        encode_axes<2, 3>(q, out);
        encode_tail(n, wo, out);
    }
}

// Decoded lobes [k0, k0 + 4) of raw outputs y (40 floats) written as
// o[k] = pi, o[kLobes + k] = kappa, o[2 kLobes + 3k ..] = mu: the same
// arithmetic as decode_mixture, so that two lanes can decode one row.
__device__ __forceinline__ void decode_lobes4(const float* y, int k0, float* __restrict__ o) {
    float mx = y[0];
    for (int k = 1; k < kLobes; k++) mx = fmaxf(mx, y[k]);
    float s = 0.f;
    for (int k = 0; k < kLobes; k++) s += expf(y[k] - mx);
    const float inv = 1.f / s;
#pragma unroll
    for (int kk = 0; kk < 4; kk++) {
        const int k = k0 + kk;
        o[k] = expf(y[k] - mx) * inv;
        o[kLobes + k] = fminf(fmaxf(softplus(y[kLobes + k]), kKappaMin), kKappaMax);
        const f3 r = mk(y[2 * kLobes + 3 * k], y[2 * kLobes + 3 * k + 1], y[2 * kLobes + 3 * k + 2]);
        const float l = length(r);
        const f3 mu = l > 1e-8f ? r / l : mk(0.f, 0.f, 1.f);
        o[2 * kLobes + 3 * k] = mu.x;
        o[2 * kLobes + 3 * k + 1] = mu.y;
        o[2 * kLobes + 3 * k + 2] = mu.z;
    }
}

}  // namespace lmgpu

namespace lmgpu {

// LDS pitches (in bf16 / f32 elements) of the per-wave tile buffers. Row
// pitches of 176 B and 144 B put the 16 rows read by one 16-lane phase of a
// ds_read_b128 on 16 distinct 4-bank groups.
constexpr int kPitchX = 88;    // [32][88] bf16: inputs / even layers
constexpr int kPitchH = 72;    // [32][72] bf16: odd layers
constexpr int kPitchY = 68;    // [32][68] f32: outputs
constexpr int kWaveLds = 32 * kPitchX * 2 + 32 * kPitchH * 2;   // 10240 B; the f32 output tile aliases it
static_assert(32 * kPitchY * 4 <= kWaveLds, "output tile must fit the aliased buffers");

// BF16 weights staged in LDS once per block, rows padded as the tiles:
// W1 [64][88], W2..W4 [64][72].
constexpr int kPW1 = kPitchX, kPWh = kPitchH;
constexpr int kWOff2 = kHid * kPW1, kWOffH = kPWh * kHid;
constexpr int kWeightLds = (kHid * kPW1 + 3 * kOutPad * kPWh) * 2;   // 38912 B

__device__ __forceinline__ void stage_weights(const uint16_t* __restrict__ Wbf, uint16_t* wl, int tid, int nthr) {
    for (int c = tid; c < kHid * (kIn / 8); c += nthr) {
        const int r = c / (kIn / 8), k = (c % (kIn / 8)) * 8;
        *reinterpret_cast<uint4*>(wl + r * kPW1 + k) = *reinterpret_cast<const uint4*>(Wbf + kOffW1 + r * kIn + k);
    }
    const int offs[3] = {kOffW2, kOffW3, kOffW4};
    for (int l = 0; l < 3; l++)
        for (int c = tid; c < kOutPad * (kHid / 8); c += nthr) {
            const int r = c / (kHid / 8), k = (c % (kHid / 8)) * 8;
            *reinterpret_cast<uint4*>(wl + kWOff2 + l * kWOffH + r * kPWh + k) =
                *reinterpret_cast<const uint4*>(Wbf + offs[l] + r * kHid + k);
        }
}

__device__ __forceinline__ f32x16 mfma_bf16(bf16x8 a, bf16x8 b, f32x16 c) {
    return __builtin_amdgcn_mfma_f32_32x32x16_bf16(a, b, c, 0, 0, 0);
}

// One dense layer for the 32 rows of a wave tile: out = act(in[32][K] W + b)
// with in from LDS (bf16, pitch pin), weights Wt[N][K] bf16 (output-neuron
// major, row pitch PW, in LDS or global memory) and N = 64. The MFMA result
// layout (lane l, element r: row 8(r/4) + 4(l/32) + r%4, column 32t + l%32)
// is written back by `store(row, col, value, e, t)`, e and t being the
// accumulator element and the 32-column half (row = 8(e/4) + 4(l/32) + e%4,
// col = 32t + l%32).
template <int K, int PW, int UNROLL, class Store>
__device__ __forceinline__ void dense64(const uint16_t* __restrict__ in, int pin, const uint16_t* __restrict__ Wt,
                                        const float* __restrict__ bias, int lane, Store store) {
    f32x16 acc0 = {}, acc1 = {};
    const int r = lane & 31, h = lane >> 5;
#pragma unroll UNROLL
    for (int k0 = 0; k0 < K; k0 += 16) {
        const bf16x8 a = *reinterpret_cast<const bf16x8*>(in + r * pin + k0 + 8 * h);
        const bf16x8 b0 = *reinterpret_cast<const bf16x8*>(Wt + r * PW + k0 + 8 * h);
        const bf16x8 b1 = *reinterpret_cast<const bf16x8*>(Wt + (32 + r) * PW + k0 + 8 * h);
        acc0 = mfma_bf16(a, b0, acc0);
        acc1 = mfma_bf16(a, b1, acc1);
    }
    const float bb0 = bias[r], bb1 = bias[32 + r];
#pragma unroll
    for (int e = 0; e < 16; e++) {
        const int row = 8 * (e >> 2) + 4 * h + (e & 3);
        store(row, r, acc0[e] + bb0, e, 0);
        store(row, 32 + r, acc1[e] + bb1, e, 1);
    }
}

// Forward pass of the guide MLP for the 32 rows of a wave tile whose encoded
// inputs are in lx ([32][kPitchX] bf16), with the weights staged in LDS by
// stage_weights (wl). Raw outputs land in ly ([32][kPitchY] f32, columns
// 0..39). `save(layer, e, t, bf16)` sees every hidden activation
// (layer 1..3) for training. Every wave of the block calls this the same
// number of times (block barriers between the layers); callers put a block
// barrier after staging the weights and writing lx, and after this call
// before reading ly.
template <class Save>
__device__ __forceinline__ void mlp_tile_forward(const uint16_t* wl, const float* __restrict__ P, uint16_t* lx,
                                                 uint16_t* lh, float* ly, int lane, Save save) {
    dense64<kIn, kPW1, kIn / 16>(lx, kPitchX, wl, P + kOffB1, lane, [&](int row, int col, float v, int e, int t) {
        const uint16_t b = f2bf(fmaxf(v, 0.f));
        lh[row * kPitchH + col] = b;
        save(1, e, t, b);
    });
    __syncthreads();
    dense64<kHid, kPWh, kHid / 16>(lh, kPitchH, wl + kWOff2, P + kOffB2, lane, [&](int row, int col, float v, int e, int t) {
        const uint16_t b = f2bf(fmaxf(v, 0.f));
        lx[row * kPitchX + col] = b;
        save(2, e, t, b);
    });
    __syncthreads();
    dense64<kHid, kPWh, kHid / 16>(lx, kPitchX, wl + kWOff2 + kWOffH, P + kOffB3, lane, [&](int row, int col, float v, int e, int t) {
        const uint16_t b = f2bf(fmaxf(v, 0.f));
        lh[row * kPitchH + col] = b;
        save(3, e, t, b);
    });
    __syncthreads();
    dense64<kHid, kPWh, kHid / 16>(lh, kPitchH, wl + kWOff2 + 2 * kWOffH, P + kOffB4, lane,
                           [&](int row, int col, float v, int, int) { ly[row * kPitchY + col] = v; });
}

}  // namespace lmgpu
