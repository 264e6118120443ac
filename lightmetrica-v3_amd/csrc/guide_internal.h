// Host-side guide handle behind lmgpu_guide (guide.hip).
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

#include "lmgpu_internal.h"

namespace lmgpu {

// Training record written by the path tracer and read by Guide::train:
// pos(3), shading normal(3), outgoing direction(3), sampled incident
// direction(3), mixture pdf of that direction (solid angle), incident
// radiance estimate L(3).
constexpr int kRecordFloats = 16;

struct Guide {
    DevBuf<float> P, M, V, G;       // parameters, Adam moments, gradients (packed layout; G is zero between steps)
    DevBuf<uint16_t> Wbf, WnBf;     // BF16 copies: packed transposed weights; layers 2..4 as W[k][n]
    DevBuf<float> zero64, scal;     // zero bias; scalars: [0] weight sum, [1] loss of the last step
    DevBuf<int32_t> stepc;          // Adam step count
    DevBuf<uint16_t> xT, hT, dT;    // training activations / deltas, feature major
    DevBuf<float> w;                // target weights
    int64_t train_cap = 0;
    // identity of the device buffers, for callers that capture them in graphs:
    // uid is unique per guide in the process, gen grows with every reallocation
    uint64_t uid = 0;
    int gen = 0;
    float box_lo[3] = {0, 0, 0}, box_inv[3] = {1, 1, 1};
    float wclamp = 50.f;            // clamp of normalised target weights

    int create(const float* lo, const float* hi, uint64_t seed);
    int set_params(const float* host);
    int get_params(float* host) const;
    int ensure_train(int64_t B);
    // query: n rows of `qs` floats starting with pos(3), nrm(3), wo(3)
    int infer(const float* query, int qs, float* mix, int64_t n, hipStream_t s) const;
    int train(const float* rec, int64_t n, float lr, float* loss, float* grad, hipStream_t s);
    // One Adam step on the first min(n, max(0, *ndev - off)) records of rec
    // (all n when ndev is NULL); rec points at row `off` of a compacted
    // buffer whose total count *ndev is in device memory. B >= n is the
    // padded row capacity (multiple of 128, prepared by ensure_train).
    // Enqueues kernels only (graph-capturable) unless grad_host is given.
    // wclamp <= 0: the guide's own clamp of normalised training weights
    int enqueue_train(const float* rec, int64_t n, int64_t B, const int32_t* ndev, int64_t off, float lr,
                      float* grad_host, hipStream_t s, float wclamp = 0.f);
};

}  // namespace lmgpu

// Opaque handle of the C ABI.
struct lmgpu_guide {
    lmgpu::Guide g;
};
