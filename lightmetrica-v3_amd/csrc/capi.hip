// C ABI entry points for the accel (include/lmgpu.h). No C++ exception or
// HIP error crosses this boundary: every call returns a status code and sets
// the thread-local message read by lmgpu_last_error(), mirroring the
// reference's LM_THROW_EXCEPTION(Error::...) contract at the host component.

#include <hip/hip_runtime.h>

#include <cstdlib>
#include <cstring>
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
    info->node_bytes = a.num_nodes * 32 * a.width;
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
