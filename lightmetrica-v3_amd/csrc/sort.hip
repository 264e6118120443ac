// Device-wide exclusive scan and LSD radix sort of (uint64 key, int32 value)
// pairs, used by the BVH builders (Morton-order sort, PLOC compaction).
//
// Scan: reduce-then-scan over tiles of 4096 elements (256 threads x 16),
// each tile scanned in 16 rounds of a 256-wide block scan (wave shuffles,
// one LDS exchange of the 4 wave totals); each block adds the sum of the
// tile sums before its own (up to 4096 tiles; beyond, the tile sums are
// scanned recursively).
//
// Sort: one pass per 8-bit digit, three launches per pass and no spin-waits:
//   upsweep    per-tile digit histogram (LDS atomics) -> hist[digit][tile],
//   scan       exclusive scan of hist: the first output slot of every
//              (digit, tile) pair,
//   downsweep  each wave ranks its 16 x 64 keys stably by digit with ballots
//              (8 ballots give the lanes sharing a digit), per-wave digit
//              counts in LDS give the tile-local order, the tile is reordered
//              in LDS by digit and written out in digit runs, so global
//              writes are contiguous within each run.
// Stable in every pass, so the result is ordered by key and, among equal
// keys, by input position.

#include <algorithm>

#include "build.h"
#include "lmgpu_internal.h"

namespace lmgpu {

namespace {

constexpr int kT = 256;           // threads per block (4 waves)
constexpr int kItems = 16;        // elements per thread
constexpr int kTile = kT * kItems;
constexpr int kRadix = 256;

inline unsigned tiles(int64_t n) { return unsigned((n + kTile - 1) / kTile); }

// Exclusive scan of one value per thread over a 256-thread block; `total`
// receives the block sum. Uses lds[0..3]; all threads must call it.
template <class T>
__device__ __forceinline__ T block_scan_256(T x, T& total, T* lds) {
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    T inc = x;
#pragma unroll
    for (int off = 1; off < 64; off <<= 1) {
        const T y = __shfl_up(inc, off);
        if (lane >= off) inc += y;
    }
    if (lane == 63) lds[w] = inc;
    __syncthreads();
    T pre = 0;
    for (int k = 0; k < w; k++) pre += lds[k];
    total = lds[0] + lds[1] + lds[2] + lds[3];
    __syncthreads();
    return pre + inc - x;
}

template <class T>
__global__ __launch_bounds__(kT) void k_scan_reduce(const T* __restrict__ in, int64_t m, T* __restrict__ sums) {
    __shared__ T lds[4];
    const int64_t base = int64_t(blockIdx.x) * kTile;
    T acc = 0;
    for (int i = 0; i < kItems; i++) {
        const int64_t idx = base + int64_t(i) * kT + threadIdx.x;
        if (idx < m) acc += in[idx];
    }
    T total;
    (void)block_scan_256<T>(acc, total, lds);
    if (threadIdx.x == 0) sums[blockIdx.x] = total;
}

// out[i] = offset(tile) + exclusive prefix of in within the tile; in == out
// is allowed. offsets == nullptr: offset 0 (a single tile).
template <class T>
__global__ __launch_bounds__(kT) void k_scan_down(const T* in, int64_t m, const T* __restrict__ offsets, T* out) {
    __shared__ T lds[4];
    const int64_t base = int64_t(blockIdx.x) * kTile;
    T carry = offsets ? offsets[blockIdx.x] : T(0);
    for (int i = 0; i < kItems; i++) {
        const int64_t idx = base + int64_t(i) * kT + threadIdx.x;
        const T x = idx < m ? in[idx] : T(0);
        T total;
        const T ex = block_scan_256<T>(x, total, lds);
        if (idx < m) out[idx] = carry + ex;
        carry += total;
    }
}

// As k_scan_down, with each tile's offset summed by its own block from the
// tile sums before it (for up to kTile tiles; saves the scan of the sums).
template <class T>
__global__ __launch_bounds__(kT) void k_scan_down_sums(const T* in, int64_t m, const T* __restrict__ sums, T* out) {
    __shared__ T lds[4];
    T part = 0;
    for (unsigned b = threadIdx.x; b < blockIdx.x; b += kT) part += sums[b];
    T carry;
    (void)block_scan_256<T>(part, carry, lds);
    const int64_t base = int64_t(blockIdx.x) * kTile;
    for (int i = 0; i < kItems; i++) {
        const int64_t idx = base + int64_t(i) * kT + threadIdx.x;
        const T x = idx < m ? in[idx] : T(0);
        T total;
        const T ex = block_scan_256<T>(x, total, lds);
        if (idx < m) out[idx] = carry + ex;
        carry += total;
    }
}

// Elements of temporary storage scan_rec needs for m elements.
int64_t scan_temp(int64_t m) {
    const int64_t nb = (m + kTile - 1) / kTile;
    return nb <= 1 ? 0 : nb + scan_temp(nb);
}

// Exclusive scan of in[0..m) into out (in == out allowed), enqueued on s;
// temp holds scan_temp(m) elements.
template <class T>
void scan_rec(const T* in, T* out, int64_t m, T* temp, hipStream_t s) {
    if (m <= 0) return;
    const unsigned nb = tiles(m);
    if (nb == 1) {
        hipLaunchKernelGGL(k_scan_down<T>, dim3(1), dim3(kT), 0, s, in, m, static_cast<const T*>(nullptr), out);
        return;
    }
    T* sums = temp;
    hipLaunchKernelGGL(k_scan_reduce<T>, dim3(nb), dim3(kT), 0, s, in, m, sums);
    if (nb <= unsigned(kTile)) {
        hipLaunchKernelGGL(k_scan_down_sums<T>, dim3(nb), dim3(kT), 0, s, in, m, static_cast<const T*>(sums), out);
        return;
    }
    scan_rec<T>(sums, sums, int64_t(nb), temp + nb, s);
    hipLaunchKernelGGL(k_scan_down<T>, dim3(nb), dim3(kT), 0, s, in, m, static_cast<const T*>(sums), out);
}

// ---- radix sort ----------------------------------------------------------------

__global__ __launch_bounds__(kT) void k_sort_upsweep(const uint64_t* __restrict__ keys, int64_t n, int shift,
                                                     uint32_t* __restrict__ hist, unsigned ntiles) {
    __shared__ uint32_t h[kRadix];
    h[threadIdx.x] = 0;
    __syncthreads();
    const int64_t base = int64_t(blockIdx.x) * kTile;
    for (int i = 0; i < kItems; i++) {
        const int64_t idx = base + int64_t(i) * kT + threadIdx.x;
        if (idx < n) atomicAdd(&h[uint32_t(keys[idx] >> shift) & 0xffu], 1u);
    }
    __syncthreads();
    hist[int64_t(threadIdx.x) * ntiles + blockIdx.x] = h[threadIdx.x];
}

__global__ __launch_bounds__(kT) void k_sort_downsweep(const uint64_t* __restrict__ kin,
                                                       const int32_t* __restrict__ vin, uint64_t* __restrict__ kout,
                                                       int32_t* __restrict__ vout, int64_t n, int shift,
                                                       const uint32_t* __restrict__ offs, unsigned ntiles) {
    __shared__ uint64_t sk[kTile];
    __shared__ int32_t sv[kTile];
    __shared__ uint32_t wcnt[4][kRadix];   // per-wave digit counts, then their exclusive prefix over waves
    __shared__ uint32_t goff[kRadix];      // first output slot of each digit for this tile
    __shared__ uint32_t dstart[kRadix];    // first LDS slot of each digit within the tile
    __shared__ uint32_t red[4];
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    for (int k = 0; k < 4; k++) wcnt[k][threadIdx.x] = 0;
    goff[threadIdx.x] = offs[int64_t(threadIdx.x) * ntiles + blockIdx.x];
    __syncthreads();
    const int64_t base = int64_t(blockIdx.x) * kTile + int64_t(w) * (kItems * 64);
    const uint64_t lt_mask = (1ull << lane) - 1ull;
    uint64_t key[kItems];
    int32_t val[kItems];
    uint32_t rank[kItems];
    for (int i = 0; i < kItems; i++) {
        const int64_t idx = base + int64_t(i) * 64 + lane;
        const bool valid = idx < n;
        key[i] = valid ? kin[idx] : 0ull;
        val[i] = valid ? vin[idx] : 0;
        const uint32_t d = uint32_t(key[i] >> shift) & 0xffu;
        uint64_t peers = __ballot(valid);
#pragma unroll
        for (int b = 0; b < 8; b++) {
            const bool bit = (d >> b) & 1u;
            const uint64_t m = __ballot(bit);
            peers &= bit ? m : ~m;
        }
        const uint32_t below = uint32_t(__popcll(peers & lt_mask));
        const uint32_t prev = valid ? wcnt[w][d] : 0u;
        // all lanes of the wave read before the group's lowest lane writes
        __builtin_amdgcn_wave_barrier();
        if (valid && below == 0) wcnt[w][d] = prev + uint32_t(__popcll(peers));
        __builtin_amdgcn_wave_barrier();
        rank[i] = prev + below;
    }
    __syncthreads();
    // thread = digit: prefix over the 4 waves, then over the digits
    uint32_t run = 0;
    for (int k = 0; k < 4; k++) {
        const uint32_t c = wcnt[k][threadIdx.x];
        wcnt[k][threadIdx.x] = run;
        run += c;
    }
    uint32_t tot;
    dstart[threadIdx.x] = block_scan_256<uint32_t>(run, tot, red);
    __syncthreads();
    for (int i = 0; i < kItems; i++) {
        const int64_t idx = base + int64_t(i) * 64 + lane;
        if (idx < n) {
            const uint32_t d = uint32_t(key[i] >> shift) & 0xffu;
            const uint32_t p = dstart[d] + wcnt[w][d] + rank[i];
            sk[p] = key[i];
            sv[p] = val[i];
        }
    }
    __syncthreads();
    const int64_t tile_n = std::min<int64_t>(kTile, n - int64_t(blockIdx.x) * kTile);
    for (int j = threadIdx.x; j < tile_n; j += kT) {
        const uint64_t k = sk[j];
        const uint32_t d = uint32_t(k >> shift) & 0xffu;
        const int64_t o = int64_t(goff[d]) + (j - int64_t(dstart[d]));
        kout[o] = k;
        vout[o] = sv[j];
    }
}

}  // namespace

size_t exclusive_scan_u64_temp_bytes(int64_t n) { return size_t(std::max<int64_t>(scan_temp(n), 1)) * 8; }

int exclusive_scan_u64(const uint64_t* in, uint64_t* out, int64_t n, void* temp, size_t temp_bytes, hipStream_t s) {
    if (temp_bytes < exclusive_scan_u64_temp_bytes(n)) return fail(LMGPU_ERR_INVALID_ARGUMENT, "scan: temp too small");
    scan_rec<unsigned long long>(reinterpret_cast<const unsigned long long*>(in),
                                 reinterpret_cast<unsigned long long*>(out), n,
                                 static_cast<unsigned long long*>(temp), s);
    LMGPU_HIP(hipGetLastError());
    return LMGPU_OK;
}

int radix_sort_pairs_u64(uint64_t* keys, int32_t* vals, uint64_t* keys_alt, int32_t* vals_alt, int64_t n,
                         int key_bits, bool* in_alt, hipStream_t s) {
    *in_alt = false;
    if (n <= 1) return LMGPU_OK;
    if (int64_t(tiles(n)) * kRadix >= (int64_t(1) << 31) || n >= (int64_t(1) << 32))
        return fail(LMGPU_ERR_INVALID_ARGUMENT, "radix sort: too many keys");
    const unsigned nt = tiles(n);
    DevBuf<uint32_t> hist, temp;
    LMGPU_TRY(hist.alloc(size_t(nt) * kRadix));
    LMGPU_TRY(temp.alloc(size_t(std::max<int64_t>(scan_temp(int64_t(nt) * kRadix), 1))));
    uint64_t* kbuf[2] = {keys, keys_alt};
    int32_t* vbuf[2] = {vals, vals_alt};
    int cur = 0;
    for (int shift = 0; shift < key_bits; shift += 8) {
        hipLaunchKernelGGL(k_sort_upsweep, dim3(nt), dim3(kT), 0, s, kbuf[cur], n, shift, hist.p, nt);
        scan_rec<uint32_t>(hist.p, hist.p, int64_t(nt) * kRadix, temp.p, s);
        hipLaunchKernelGGL(k_sort_downsweep, dim3(nt), dim3(kT), 0, s, kbuf[cur], vbuf[cur], kbuf[cur ^ 1],
                           vbuf[cur ^ 1], n, shift, hist.p, nt);
        cur ^= 1;
    }
    LMGPU_HIP(hipGetLastError());
    *in_alt = cur == 1;
    // hist and temp are freed on return: wait for the passes that read them
    LMGPU_HIP(hipStreamSynchronize(s));
    return LMGPU_OK;
}

}  // namespace lmgpu
