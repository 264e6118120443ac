// Shared device/host definitions of the lmgpu library (gfx950).
#pragma once

#include <hip/hip_runtime.h>
#include <cstdint>

#define LMGPU_HD __host__ __device__ __forceinline__

namespace lmgpu {

struct f3 {
    float x, y, z;
};
LMGPU_HD f3 mk(float x, float y, float z) { return {x, y, z}; }
LMGPU_HD f3 operator+(f3 a, f3 b) { return {a.x + b.x, a.y + b.y, a.z + b.z}; }
LMGPU_HD f3 operator-(f3 a, f3 b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
LMGPU_HD f3 operator-(f3 a) { return {-a.x, -a.y, -a.z}; }
LMGPU_HD f3 operator*(f3 a, float s) { return {a.x * s, a.y * s, a.z * s}; }
LMGPU_HD f3 operator*(float s, f3 a) { return {a.x * s, a.y * s, a.z * s}; }
LMGPU_HD f3 operator*(f3 a, f3 b) { return {a.x * b.x, a.y * b.y, a.z * b.z}; }
LMGPU_HD f3 operator/(f3 a, float s) { return {a.x / s, a.y / s, a.z / s}; }
LMGPU_HD float dot(f3 a, f3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
LMGPU_HD f3 cross(f3 a, f3 b) { return {a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x}; }
LMGPU_HD float length(f3 a) { return sqrtf(dot(a, a)); }
LMGPU_HD f3 normalize(f3 a) { return a * (1.0f / sqrtf(dot(a, a))); }
LMGPU_HD float maxc(f3 a) { return fmaxf(a.x, fmaxf(a.y, a.z)); }
LMGPU_HD f3 fmin3(f3 a, f3 b) { return {fminf(a.x, b.x), fminf(a.y, b.y), fminf(a.z, b.z)}; }
LMGPU_HD f3 fmax3(f3 a, f3 b) { return {fmaxf(a.x, b.x), fmaxf(a.y, b.y), fmaxf(a.z, b.z)}; }
LMGPU_HD float comp(f3 a, int i) { return i == 0 ? a.x : (i == 1 ? a.y : a.z); }
LMGPU_HD bool is_zero(f3 a) { return a.x == 0.f && a.y == 0.f && a.z == 0.f; }

// Philox4x32-10 (Salmon et al., SC'11): multipliers 0xD2511F53, 0xCD9E8D57,
// Weyl key increments 0x9E3779B9, 0xBB67AE85.
struct u4 {
    uint32_t x, y, z, w;
};
LMGPU_HD u4 philox4x32_10(u4 c, uint32_t k0, uint32_t k1) {
#pragma unroll
    for (int r = 0; r < 10; r++) {
        const uint64_t p0 = uint64_t(0xD2511F53u) * c.x;
        const uint64_t p1 = uint64_t(0xCD9E8D57u) * c.z;
        const uint32_t hi0 = uint32_t(p0 >> 32), lo0 = uint32_t(p0);
        const uint32_t hi1 = uint32_t(p1 >> 32), lo1 = uint32_t(p1);
        c = u4{hi1 ^ c.y ^ k0, lo1, hi0 ^ c.w ^ k1, lo0};
        k0 += 0x9E3779B9u;
        k1 += 0xBB67AE85u;
    }
    return c;
}
// 24-bit uniform in [0,1), exactly representable in FP32.
LMGPU_HD float u24(uint32_t x) { return float(x >> 8) * (1.0f / 16777216.0f); }

struct rand4 {
    float u[4];
};
// Stream layout (shared with the CPU oracle): counter (pixel, sample,
// vertex, block), key = seed. Vertex k >= 1 of a path draws block 0
// {RR, component, direction u0, direction u1} and block 1 {light select,
// light triangle, position u0, position u1}; the camera vertex (k = 1, index
// 0) uses block 0 [2..3] for the pixel jitter; block 2 feeds path guiding.
LMGPU_HD rand4 rng(uint32_t pixel, uint32_t sample, uint32_t vertex, uint32_t block, uint64_t seed) {
    const u4 o = philox4x32_10(u4{pixel, sample, vertex, block}, uint32_t(seed), uint32_t(seed >> 32));
    return rand4{{u24(o.x), u24(o.y), u24(o.z), u24(o.w)}};
}

}  // namespace lmgpu
