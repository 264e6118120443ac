// Shared definitions of the device BVH builders (accel.hip, ploc.hip).
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

namespace lmgpu {

// Binary BVH node over a contiguous range [first, last] of the triangle
// array in traversal order. Child refs: >= 0 internal node, < 0 leaf ~j.
struct BNode {
    int32_t left, right;
    int32_t first, last;
};

// Stable LSD radix sort of n (key, value) pairs by the low `key_bits` bits
// of the keys (8-bit digits), ping-ponging between (keys, vals) and
// (keys_alt, vals_alt); *in_alt tells which pair holds the result.
// Synchronizes s before returning (its temporary buffers are freed).
int radix_sort_pairs_u64(uint64_t* keys, int32_t* vals, uint64_t* keys_alt, int32_t* vals_alt, int64_t n,
                         int key_bits, bool* in_alt, hipStream_t s);

// Exclusive prefix sum of n uint64 values (in == out allowed), enqueued on s,
// with a caller-owned temporary buffer of exclusive_scan_u64_temp_bytes(n).
size_t exclusive_scan_u64_temp_bytes(int64_t n);
int exclusive_scan_u64(const uint64_t* in, uint64_t* out, int64_t n, void* temp, size_t temp_bytes, hipStream_t s);

// Builds a binary BVH by parallel locally-ordered clustering (PLOC,
// Meister & Bittner, IEEE TVCG 24(3), 2018) over the Morton-sorted
// triangles. In: n, `order` (Morton position -> triangle), per-triangle
// bounds tlo/thi. Out (device arrays, n-1 internal nodes, n leaves):
// bn + node_lo/node_hi with contiguous leaf ranges, leaf_lo/leaf_hi and
// `perm` (traversal position -> triangle) in depth-first leaf order, and the
// root ref (`*root_ref`, host). `radius` is the nearest-neighbour window.
int build_ploc(int64_t n, const int32_t* order, const float4* tlo, const float4* thi, int radius, BNode* bn,
               float4* node_lo, float4* node_hi, float4* leaf_lo, float4* leaf_hi, int32_t* perm, int32_t* root_ref,
               hipStream_t s);

}  // namespace lmgpu
