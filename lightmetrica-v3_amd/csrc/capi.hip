// C ABI entry points for the accel (include/lmgpu.h). No C++ exception or
// HIP error crosses this boundary: every call returns a status code and sets
// the thread-local message read by lmgpu_last_error(), mirroring the
// reference's LM_THROW_EXCEPTION(Error::...) contract at the host component.

#include <hip/hip_runtime.h>

#include <cstdlib>
#include <cstring>
#include <mutex>
#include <new>
#include <string>
#include <vector>

#include "lmgpu_internal.h"

namespace lmgpu {


static thread_local std::string g_last_error;

int fail(int code, const std::string& msg) {
    g_last_error = msg;
    return code;
}
void clear_error() { g_last_error.clear(); }

int Accel::check_status(hipStream_t s) const {
    if (!status.p) return LMGPU_OK;
    int32_t v = 0;
    LMGPU_HIP(hipMemcpyAsync(&v, status.p, sizeof(v), hipMemcpyDeviceToHost, s));
    LMGPU_HIP(hipStreamSynchronize(s));
    if (v) {
        LMGPU_HIP(hipMemsetAsync(status.p, 0, sizeof(int32_t), s));
        return fail(LMGPU_ERR_STACK_OVERFLOW, "BVH traversal stack capacity exceeded");
    }
    return LMGPU_OK;
}

}  // namespace lmgpu

using lmgpu::fail;

extern "C" {

const char* lmgpu_last_error(void) { return lmgpu::g_last_error.c_str(); }

const char* lmgpu_version(void) { return "lmgpu 0.1 (gfx950)"; }

int lmgpu_device_count(int* count) {
    if (!count) return fail(LMGPU_ERR_INVALID_ARGUMENT, "count is NULL");
    int c = 0;
    if (hipGetDeviceCount(&c) != hipSuccess) c = 0;
    *count = c;
    return LMGPU_OK;
}

struct lmgpu_accel {
    lmgpu::Accel a;
};

int lmgpu_accel_create(lmgpu_accel** out) {
    if (!out) return fail(LMGPU_ERR_INVALID_ARGUMENT, "out is NULL");
    *out = new (std::nothrow) lmgpu_accel();
    if (!*out) return fail(LMGPU_ERR_OUT_OF_MEMORY, "accel handle");
    return LMGPU_OK;
}

int lmgpu_accel_destroy(lmgpu_accel* accel) {
    delete accel;
    return LMGPU_OK;
}

int lmgpu_accel_build(lmgpu_accel* accel, const float* verts, int64_t num_tris, int verts_on_device,
                      const lmgpu_build_opts* opts, void* stream) {
    if (!accel) return fail(LMGPU_ERR_INVALID_ARGUMENT, "accel is NULL");
    lmgpu::clear_error();
    return accel->a.build(verts, num_tris, verts_on_device != 0, opts, static_cast<hipStream_t>(stream));
}

int lmgpu_accel_info_get(const lmgpu_accel* accel, lmgpu_accel_info* info) {
    if (!accel || !info) return fail(LMGPU_ERR_INVALID_ARGUMENT, "NULL argument");
    const lmgpu::Accel& a = accel->a;
    if (!a.built) return fail(LMGPU_ERR_INVALID_ARGUMENT, "accel is not built");
    std::memset(info, 0, sizeof(*info));
    info->num_tris = a.num_tris;
    info->num_nodes = a.num_nodes;
    info->num_leaves = a.num_leaves;
    info->width = a.width;
    info->leaf_size = a.leaf_size;
    info->depth = a.depth;
    info->stack_capacity = 24 + 72;
    info->node_bytes = a.num_nodes * 32 * a.width + int64_t(a.qnodes.n) * int64_t(sizeof(float4));
    info->tri_bytes = a.num_tris * int64_t(sizeof(lmgpu::TriV));
    info->sah_cost = a.sah_cost;
    info->build_ms = a.build_ms;
    return LMGPU_OK;
}

int lmgpu_accel_intersect(const lmgpu_accel* accel, const lmgpu_ray* rays, lmgpu_hit* hits, int64_t n,
                          void* stream) {
    if (!accel) return fail(LMGPU_ERR_INVALID_ARGUMENT, "accel is NULL");
    return accel->a.trace(rays, hits, nullptr, n, false, nullptr, static_cast<hipStream_t>(stream));
}

int lmgpu_accel_intersect_counted(const lmgpu_accel* accel, const lmgpu_ray* rays, lmgpu_hit* hits, int64_t n,
                                  int64_t* node_visits, int64_t* tri_tests, void* stream) {
    if (!accel) return fail(LMGPU_ERR_INVALID_ARGUMENT, "accel is NULL");
    hipStream_t s = static_cast<hipStream_t>(stream);
    lmgpu::DevBuf<unsigned long long> c;
    LMGPU_TRY(c.alloc(2));
    LMGPU_HIP(hipMemsetAsync(c.p, 0, 2 * sizeof(unsigned long long), s));
    LMGPU_TRY(accel->a.trace(rays, hits, nullptr, n, false, c.p, s));
    unsigned long long h[2] = {0, 0};
    LMGPU_HIP(hipMemcpyAsync(h, c.p, sizeof(h), hipMemcpyDeviceToHost, s));
    LMGPU_HIP(hipStreamSynchronize(s));
    if (node_visits) *node_visits = int64_t(h[0]);
    if (tri_tests) *tri_tests = int64_t(h[1]);
    return accel->a.check_status(s);
}

int lmgpu_accel_intersect_host(const lmgpu_accel* accel, const lmgpu_ray* rays, lmgpu_hit* hits, int64_t n) {
    if (!accel) return fail(LMGPU_ERR_INVALID_ARGUMENT, "accel is NULL");
    if (n < 0 || !rays || !hits) return fail(LMGPU_ERR_INVALID_ARGUMENT, "bad ray buffers");
    if (n == 0) return LMGPU_OK;
    lmgpu::DevBuf<lmgpu_ray> dr;
    lmgpu::DevBuf<lmgpu_hit> dh;
    LMGPU_TRY(dr.alloc(n));
    LMGPU_TRY(dh.alloc(n));
    LMGPU_HIP(hipMemcpy(dr.p, rays, sizeof(lmgpu_ray) * n, hipMemcpyHostToDevice));
    LMGPU_TRY(accel->a.trace(dr.p, dh.p, nullptr, n, false, nullptr, nullptr));
    LMGPU_HIP(hipMemcpy(hits, dh.p, sizeof(lmgpu_hit) * n, hipMemcpyDeviceToHost));
    return accel->a.check_status(nullptr);
}

int lmgpu_accel_occluded(const lmgpu_accel* accel, const lmgpu_ray* rays, uint8_t* occluded, int64_t n,
                         void* stream) {
    if (!accel) return fail(LMGPU_ERR_INVALID_ARGUMENT, "accel is NULL");
    return accel->a.trace(rays, nullptr, occluded, n, true, nullptr, static_cast<hipStream_t>(stream));
}

int lmgpu_accel_status(const lmgpu_accel* accel, void* stream) {
    if (!accel) return fail(LMGPU_ERR_INVALID_ARGUMENT, "accel is NULL");
    return accel->a.check_status(static_cast<hipStream_t>(stream));
}

}  // extern "C"

// ---- allocation-free host ray cast -------------------------------------------
// lmgpu_accel_intersect_host_noalloc stages rays and hits through `scratch`,
// a file-scope device array (one copy per GPU, held until the process ends):
//   [0, 32*M)          rays  (lmgpu_ray, M = LMGPU_NOALLOC_MAX_RAYS)
//   [32*M, 48*M)       hits  (lmgpu_hit)
//   [48*M, 48*M + 8)   ray counter of the persistent traversal (zeroed by its launch)
//   [48*M + 8, +4)     traversal overflow word (cleared before each call)
// In host code `scratch` names a host-side stand-in, not device memory: its
// device address is looked up once per device (hipGetSymbolAddress, with that
// device current) and only that address is used from then on.
namespace lmgpu {

constexpr size_t kNoallocRayBytes = size_t(LMGPU_NOALLOC_MAX_RAYS) * sizeof(lmgpu_ray);
constexpr size_t kNoallocHitBytes = size_t(LMGPU_NOALLOC_MAX_RAYS) * sizeof(lmgpu_hit);
static_assert(kNoallocRayBytes + kNoallocHitBytes + 256 == size_t(LMGPU_NOALLOC_SCRATCH_BYTES),
              "scratch layout and LMGPU_NOALLOC_SCRATCH_BYTES disagree");

static __device__ __align__(256) unsigned char scratch[LMGPU_NOALLOC_SCRATCH_BYTES];

constexpr int kNoallocMaxDevices = 64;
static std::mutex g_noalloc_mu[kNoallocMaxDevices];
static unsigned char* g_noalloc_ptr[kNoallocMaxDevices];  // device address of `scratch` per device

// Device address of `scratch` on `dev` (the current device); called with
// g_noalloc_mu[dev] held.
static int noalloc_scratch(int dev, unsigned char** out) {
    if (!g_noalloc_ptr[dev]) {
        void* p = nullptr;
        LMGPU_HIP(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
        if (!p) return fail(LMGPU_ERR_DEVICE, "hipGetSymbolAddress(scratch) returned NULL");
        g_noalloc_ptr[dev] = static_cast<unsigned char*>(p);
    }
    *out = g_noalloc_ptr[dev];
    return LMGPU_OK;
}

}  // namespace lmgpu

extern "C" int lmgpu_accel_intersect_host_noalloc(const lmgpu_accel* accel, const lmgpu_ray* rays,
                                                  lmgpu_hit* hits, int64_t n) {
    using namespace lmgpu;
    if (!accel) return fail(LMGPU_ERR_INVALID_ARGUMENT, "accel is NULL");
    if (n < 0 || (n > 0 && (!rays || !hits))) return fail(LMGPU_ERR_INVALID_ARGUMENT, "bad ray buffers");
    if (n > LMGPU_NOALLOC_MAX_RAYS)
        return fail(LMGPU_ERR_INVALID_ARGUMENT,
                    "at most " + std::to_string(LMGPU_NOALLOC_MAX_RAYS) +
                        " rays per call (static staging area); got " + std::to_string(n));
    if (!accel->a.built) return fail(LMGPU_ERR_INVALID_ARGUMENT, "accel is not built");
    if (n == 0) return LMGPU_OK;
    int dev = 0;
    LMGPU_HIP(hipGetDevice(&dev));
    if (dev < 0 || dev >= kNoallocMaxDevices) return fail(LMGPU_ERR_UNSUPPORTED, "device index beyond 63");
    std::lock_guard<std::mutex> lock(g_noalloc_mu[dev]);
    unsigned char* base = nullptr;
    LMGPU_TRY(noalloc_scratch(dev, &base));
    auto* drays = reinterpret_cast<lmgpu_ray*>(base);
    auto* dhits = reinterpret_cast<lmgpu_hit*>(base + kNoallocRayBytes);
    auto* next = reinterpret_cast<unsigned long long*>(base + kNoallocRayBytes + kNoallocHitBytes);
    auto* status = reinterpret_cast<int32_t*>(next + 1);
    hipStream_t s = nullptr;
    LMGPU_HIP(hipMemsetAsync(status, 0, sizeof(int32_t), s));
    LMGPU_HIP(hipMemcpyAsync(drays, rays, sizeof(lmgpu_ray) * size_t(n), hipMemcpyHostToDevice, s));
    LMGPU_TRY(accel->a.trace_with(drays, dhits, nullptr, n, false, nullptr, next, status, s));
    LMGPU_HIP(hipMemcpyAsync(hits, dhits, sizeof(lmgpu_hit) * size_t(n), hipMemcpyDeviceToHost, s));
    int32_t overflow = 0;
    LMGPU_HIP(hipMemcpyAsync(&overflow, status, sizeof(int32_t), hipMemcpyDeviceToHost, s));
    LMGPU_HIP(hipStreamSynchronize(s));
    if (overflow) return fail(LMGPU_ERR_STACK_OVERFLOW, "BVH traversal stack capacity exceeded");
    return LMGPU_OK;
}

// ---- serialization ------------------------------------------------------------
// A built accel as bytes: a 128-byte little-endian header followed by the
// wide nodes and the triangles exactly as laid out on the device. Loading
// validates the whole structure on the host first (a malformed BVH would
// send the traversal out of bounds or round a cycle).

namespace {

struct SerialHeader {
    char magic[8];          // "LMGPUBVH"
    uint32_t version;       // 1
    uint32_t width, leaf_size, depth;
    int64_t num_tris, num_nodes, num_leaves;
    double sah_cost;
    float root_lo[3], root_hi[3];
    uint8_t pad[128 - 8 - 16 - 24 - 8 - 24];
};
static_assert(sizeof(SerialHeader) == 128, "header layout");

constexpr uint32_t kSerialVersion = 1;

template <int W>
int validate_nodes(const std::vector<unsigned char>& nb, int64_t num_nodes, int64_t num_tris, int* depth_out) {
    const auto* nodes = reinterpret_cast<const lmgpu::WNode<W>*>(nb.data());
    std::vector<unsigned char> seen(size_t(num_nodes), 0), covered(size_t(num_tris), 0);
    std::vector<std::pair<int64_t, int>> st{{0, 1}};
    int64_t visited = 0;
    int depth = 0;
    while (!st.empty()) {
        const auto [i, d] = st.back();
        st.pop_back();
        if (seen[size_t(i)]) return fail(LMGPU_ERR_INVALID_ARGUMENT, "serialized BVH is not a tree");
        seen[size_t(i)] = 1;
        visited++;
        depth = std::max(depth, d);
        const lmgpu::WNode<W>& n = nodes[i];
        for (int c = 0; c < W; c++) {
            const int cnt = n.count[c], ch = n.child[c];
            if (cnt == 0) {
                if (ch <= 0 || ch >= num_nodes) return fail(LMGPU_ERR_INVALID_ARGUMENT, "serialized BVH: child out of range");
                st.push_back({ch, d + 1});
            } else if (cnt > 0) {
                if (cnt > 15 || ch < 0 || int64_t(ch) + cnt > num_tris)
                    return fail(LMGPU_ERR_INVALID_ARGUMENT, "serialized BVH: leaf range out of range");
                for (int k = 0; k < cnt; k++) {
                    if (covered[size_t(ch) + k])
                        return fail(LMGPU_ERR_INVALID_ARGUMENT, "serialized BVH: a triangle is in two leaves");
                    covered[size_t(ch) + k] = 1;
                }
            }
        }
    }
    if (visited != num_nodes) return fail(LMGPU_ERR_INVALID_ARGUMENT, "serialized BVH has unreachable nodes");
    for (unsigned char c : covered)
        if (!c) return fail(LMGPU_ERR_INVALID_ARGUMENT, "serialized BVH: a triangle is in no leaf");
    *depth_out = depth;
    return LMGPU_OK;
}

}  // namespace

extern "C" {

int lmgpu_accel_serialize(const lmgpu_accel* accel, void* buf, size_t* size) {
    if (!accel || !size) return fail(LMGPU_ERR_INVALID_ARGUMENT, "NULL argument");
    const lmgpu::Accel& a = accel->a;
    if (!a.built) return fail(LMGPU_ERR_INVALID_ARGUMENT, "accel is not built");
    const size_t nbytes = size_t(a.num_nodes) * 32 * a.width, tbytes = size_t(a.num_tris) * sizeof(lmgpu::TriV);
    const size_t total = sizeof(SerialHeader) + nbytes + tbytes;
    if (!buf) {
        *size = total;
        return LMGPU_OK;
    }
    if (*size < total) return fail(LMGPU_ERR_INVALID_ARGUMENT, "serialization buffer too small");
    SerialHeader h;
    std::memset(&h, 0, sizeof(h));
    std::memcpy(h.magic, "LMGPUBVH", 8);
    h.version = kSerialVersion;
    h.width = uint32_t(a.width);
    h.leaf_size = uint32_t(a.leaf_size);
    h.depth = uint32_t(a.depth);
    h.num_tris = a.num_tris;
    h.num_nodes = a.num_nodes;
    h.num_leaves = a.num_leaves;
    h.sah_cost = a.sah_cost;
    for (int k = 0; k < 3; k++) {
        h.root_lo[k] = a.root_lo[k];
        h.root_hi[k] = a.root_hi[k];
    }
    auto* out = static_cast<unsigned char*>(buf);
    std::memcpy(out, &h, sizeof(h));
    LMGPU_HIP(hipMemcpy(out + sizeof(h), a.nodes.p, nbytes, hipMemcpyDeviceToHost));
    LMGPU_HIP(hipMemcpy(out + sizeof(h) + nbytes, a.tris.p, tbytes, hipMemcpyDeviceToHost));
    *size = total;
    return LMGPU_OK;
}

int lmgpu_accel_deserialize(lmgpu_accel* accel, const void* buf, size_t size, const lmgpu_build_opts* opts) {
    if (!accel || !buf) return fail(LMGPU_ERR_INVALID_ARGUMENT, "NULL argument");
    lmgpu::clear_error();
    if (size < sizeof(SerialHeader)) return fail(LMGPU_ERR_INVALID_ARGUMENT, "serialized accel too short");
    SerialHeader h;
    std::memcpy(&h, buf, sizeof(h));
    if (std::memcmp(h.magic, "LMGPUBVH", 8) != 0) return fail(LMGPU_ERR_INVALID_ARGUMENT, "not a serialized lmgpu accel");
    if (h.version != kSerialVersion) return fail(LMGPU_ERR_UNSUPPORTED, "unsupported serialized accel version");
    if (h.width != 2 && h.width != 4 && h.width != 8) return fail(LMGPU_ERR_INVALID_ARGUMENT, "bad BVH width");
    if (h.num_tris <= 0 || h.num_nodes <= 0 || h.num_tris >= (int64_t(1) << 27) || h.num_nodes >= (int64_t(1) << 31))
        return fail(LMGPU_ERR_INVALID_ARGUMENT, "bad BVH sizes");
    const size_t nbytes = size_t(h.num_nodes) * 32 * h.width, tbytes = size_t(h.num_tris) * sizeof(lmgpu::TriV);
    if (size != sizeof(SerialHeader) + nbytes + tbytes) return fail(LMGPU_ERR_INVALID_ARGUMENT, "serialized accel size mismatch");
    const auto* in = static_cast<const unsigned char*>(buf) + sizeof(SerialHeader);
    std::vector<unsigned char> nb(in, in + nbytes);
    int depth = 0;
    int rc = h.width == 2 ? validate_nodes<2>(nb, h.num_nodes, h.num_tris, &depth)
             : h.width == 4 ? validate_nodes<4>(nb, h.num_nodes, h.num_tris, &depth)
                            : validate_nodes<8>(nb, h.num_nodes, h.num_tris, &depth);
    if (rc != LMGPU_OK) return rc;
    // every triangle's original index (v0.w bits) must name a triangle
    const auto* tv = reinterpret_cast<const lmgpu::TriV*>(in + nbytes);
    for (int64_t t = 0; t < h.num_tris; t++) {
        int32_t id;
        std::memcpy(&id, &tv[t].v0.w, 4);
        if (id < 0 || id >= h.num_tris) return fail(LMGPU_ERR_INVALID_ARGUMENT, "serialized BVH: triangle id out of range");
    }
    lmgpu::Accel& a = accel->a;
    a.release();
    LMGPU_TRY(a.nodes.alloc(size_t(h.num_nodes) * 2 * h.width));
    LMGPU_TRY(a.tris.alloc(size_t(h.num_tris)));
    LMGPU_HIP(hipMemcpy(a.nodes.p, in, nbytes, hipMemcpyHostToDevice));
    LMGPU_HIP(hipMemcpy(a.tris.p, in + nbytes, tbytes, hipMemcpyHostToDevice));
    a.width = int(h.width);
    a.num_nodes = h.num_nodes;
    LMGPU_TRY(a.make_trav_nodes(nullptr));
    a.leaf_size = int(h.leaf_size);
    a.refill = (opts && opts->refill > 0) ? std::min(opts->refill, 64) : lmgpu::kDefaultRefill;
    a.num_tris = h.num_tris;
    a.num_nodes = h.num_nodes;
    a.num_leaves = h.num_leaves;
    a.depth = depth;
    a.sah_cost = h.sah_cost;
    a.build_ms = 0;
    for (int k = 0; k < 3; k++) {
        a.root_lo[k] = h.root_lo[k];
        a.root_hi[k] = h.root_hi[k];
    }
    a.built = true;
    return LMGPU_OK;
}

}  // extern "C"
