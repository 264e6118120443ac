// Internal host-side definitions of the lmgpu library: error state, device
// buffers, and the accel / scene handles behind the opaque C ABI types.
#pragma once

#include <hip/hip_runtime.h>

#include <atomic>
#include <cstdint>
#include <string>
#include <utility>

#include "../../include/lmgpu.h"
#include "bvh.h"

namespace lmgpu {

// Default refill thresholds of the persistent traversal: lmgpu_accel_* ray
// casts (incoherent config-1 rays measured fastest at 8 of 64 idle lanes,
// +3% over 32) and the renderers of a scene, whose bounce kernels measured
// fastest at 32.
constexpr int kDefaultRefill = 8;
constexpr int kDefaultSceneRefill = 32;

// Sets the thread-local last error and returns `code`.
int fail(int code, const std::string& msg);
void clear_error();

#define LMGPU_HIP(expr)                                                                        \
    do {                                                                                       \
        hipError_t e_ = (expr);                                                                \
        if (e_ != hipSuccess) {                                                                \
            return ::lmgpu::fail(e_ == hipErrorOutOfMemory ? LMGPU_ERR_OUT_OF_MEMORY          \
                                                           : LMGPU_ERR_DEVICE,                 \
                                 std::string(#expr) + ": " + hipGetErrorString(e_));          \
        }                                                                                      \
    } while (0)

#define LMGPU_TRY(expr)              \
    do {                             \
        int r_ = (expr);             \
        if (r_ != LMGPU_OK) return r_; \
    } while (0)

// Owning device allocation (hipMalloc / hipFree).
template <typename T>
struct DevBuf {
    T* p = nullptr;
    size_t n = 0;
    DevBuf() = default;
    DevBuf(const DevBuf&) = delete;
    DevBuf& operator=(const DevBuf&) = delete;
    DevBuf(DevBuf&& o) noexcept : p(o.p), n(o.n) { o.p = nullptr; o.n = 0; }
    DevBuf& operator=(DevBuf&& o) noexcept {
        if (this != &o) { reset(); p = o.p; n = o.n; o.p = nullptr; o.n = 0; }
        return *this;
    }
    ~DevBuf() { reset(); }
    void reset() {
        if (p) (void)hipFree(p);
        p = nullptr;
        n = 0;
    }
    int alloc(size_t count) {
        reset();
        if (count == 0) count = 1;
        hipError_t e = hipMalloc(reinterpret_cast<void**>(&p), count * sizeof(T));
        if (e != hipSuccess) {
            p = nullptr;
            return fail(e == hipErrorOutOfMemory ? LMGPU_ERR_OUT_OF_MEMORY : LMGPU_ERR_DEVICE,
                        std::string("hipMalloc(") + std::to_string(count * sizeof(T)) + "): " + hipGetErrorString(e));
        }
        n = count;
        return LMGPU_OK;
    }
};

// True when the traversal of a `width`-wide BVH reads quantised nodes
// (TravNode<width> = QNode<width>, 16 * width bytes each).
bool quantised_width(int width);
// qnodes[i] = quantise_node<width>(wnodes[i]) for i < n (bvh.h), enqueued on s.
int quantise_nodes(int width, const float4* wnodes, float4* qnodes, int64_t n, hipStream_t s);

struct Accel {
    DevBuf<float4> nodes;   // W-wide nodes, 32*W bytes each
    DevBuf<float4> qnodes;  // quantised copy of the nodes (QNode<width>), when quantised_width(width)
    DevBuf<TriV> tris;      // triangles in traversal order
    DevBuf<int32_t> status; // bit 0: traversal stack overflow
    DevBuf<unsigned long long> ring;  // per-launch ray counters of the persistent kernel
    mutable std::atomic<uint32_t> ring_pos{0};
    int width = 4;
    int leaf_size = 4;
    int refill = 16;        // persistent-kernel refill threshold (1..64)
    int64_t num_tris = 0;
    int64_t num_nodes = 0;
    int64_t num_leaves = 0;
    int depth = 0;
    double sah_cost = 0;
    double build_ms = 0;
    float root_lo[3] = {0, 0, 0}, root_hi[3] = {0, 0, 0};
    bool built = false;

    // node array the traversal reads (TravNode<width>::type)
    const void* trav_nodes() const { return qnodes.p ? static_cast<const void*>(qnodes.p) : nodes.p; }
    // builds the traversal copy of the nodes where TravNode differs from WNode
    int make_trav_nodes(hipStream_t s);

    void release() {
        nodes.reset();
        qnodes.reset();
        tris.reset();
        built = false;
    }
    int32_t* status_dev() const {
        if (!status.p) {
            auto* self = const_cast<Accel*>(this);
            if (self->status.alloc(1) != LMGPU_OK) return nullptr;
            if (hipMemset(self->status.p, 0, sizeof(int32_t)) != hipSuccess) return nullptr;
        }
        return status.p;
    }
    // Ray counter for one persistent launch: a ring of slots so launches that
    // overlap on different streams use different counters (up to kRing in
    // flight at once).
    static constexpr int kRing = 256;
    unsigned long long* next_counter() const {
        if (!ring.p) {
            auto* self = const_cast<Accel*>(this);
            if (self->ring.alloc(kRing) != LMGPU_OK) return nullptr;
        }
        return ring.p + (ring_pos.fetch_add(1) % kRing);
    }
    // Stack entries needed in the worst case: (W-1) pushes per level.
    int stack_bound() const { return (width - 1) * depth; }

    int build(const float* verts, int64_t n, bool on_device, const lmgpu_build_opts* opts, hipStream_t s);
    int trace(const lmgpu_ray* rays, lmgpu_hit* hits, uint8_t* occl, int64_t n, bool any,
              unsigned long long* counters, hipStream_t s) const;
    // As trace, with the caller's ray counter (zeroed by the launch itself)
    // and overflow status word (bit 0 set on overflow, never cleared here)
    // instead of the accel's own, so nothing is allocated.
    int trace_with(const lmgpu_ray* rays, lmgpu_hit* hits, uint8_t* occl, int64_t n, bool any,
                   unsigned long long* counters, unsigned long long* next, int32_t* status, hipStream_t s) const;
    int check_trace_args(const lmgpu_ray* rays, const lmgpu_hit* hits, const uint8_t* occl, int64_t n,
                         bool any) const;
    // Synchronizes and reports a traversal stack overflow since the last check.
    int check_status(hipStream_t s) const;
};

}  // namespace lmgpu
