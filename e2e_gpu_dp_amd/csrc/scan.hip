// scan.hip — device-wide exclusive scan of uint32 row counts into CSR row offsets.
//
// Replaces the thrust::reduce / thrust::exclusive_scan pair of src/build_model.cu:829-831
// (which also summed the counts in float, exact only below 2^24) with an integer
// reduce-then-scan in three launches:
//   1. tile_sums     : one 256-thread block per 2048-element tile -> 64-bit tile sum
//   2. tile_offsets  : one block scans the tile sums (carry loop), writes the total
//   3. tile_scan     : each tile rescans its 2048 elements from its offset
// Reads are 32 contiguous bytes per thread (two 16-byte loads), so each wave moves
// 2 KiB per instruction pair.
#include "internal.h"

namespace e2e {

namespace {

constexpr int kThreads = 256;
constexpr int kItems = 8;
constexpr int kTile = kThreads * kItems;

__device__ __forceinline__ unsigned long long wave_incl_scan(unsigned long long x, int lane) {
#pragma unroll
    for (int o = 1; o < 64; o <<= 1) {
        unsigned long long y = __shfl_up(x, o, 64);
        if (lane >= o) x += y;
    }
    return x;
}

// Exclusive scan across the 256 threads of a block; *total gets the block sum.
__device__ __forceinline__ unsigned long long block_excl_scan(unsigned long long v,
                                                              unsigned long long* lds,
                                                              unsigned long long* total) {
    const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
    unsigned long long inc = wave_incl_scan(v, lane);
    if (lane == 63) lds[wid] = inc;
    __syncthreads();
    unsigned long long pre = 0, tot = 0;
#pragma unroll
    for (int w = 0; w < kThreads / 64; w++) {
        unsigned long long s = lds[w];
        if (w < wid) pre += s;
        tot += s;
    }
    __syncthreads();
    *total = tot;
    return pre + inc - v;
}

__device__ __forceinline__ void load_items(const uint32_t* in, int64_t n, int64_t base,
                                           uint32_t (&x)[kItems]) {
    int64_t i0 = base + (int64_t)threadIdx.x * kItems;
    if (i0 + kItems <= n && ((reinterpret_cast<uintptr_t>(in + i0) & 15) == 0)) {
        const uint4* p = reinterpret_cast<const uint4*>(in + i0);
        uint4 a = p[0], b = p[1];
        x[0] = a.x; x[1] = a.y; x[2] = a.z; x[3] = a.w;
        x[4] = b.x; x[5] = b.y; x[6] = b.z; x[7] = b.w;
    } else {
#pragma unroll
        for (int k = 0; k < kItems; k++) x[k] = (i0 + k < n) ? in[i0 + k] : 0u;
    }
}

__global__ __launch_bounds__(kThreads) void tile_sums(const uint32_t* __restrict__ in, int64_t n,
                                                      unsigned long long* __restrict__ sums) {
    __shared__ unsigned long long lds[kThreads / 64];
    uint32_t x[kItems];
    load_items(in, n, (int64_t)blockIdx.x * kTile, x);
    unsigned long long s = 0;
#pragma unroll
    for (int k = 0; k < kItems; k++) s += x[k];
    unsigned long long tot;
    block_excl_scan(s, lds, &tot);
    if (threadIdx.x == 0) sums[blockIdx.x] = tot;
}

__global__ __launch_bounds__(kThreads) void tile_offsets(unsigned long long* __restrict__ sums,
                                                         int64_t nt, unsigned long long* total,
                                                         uint32_t* __restrict__ out, int64_t n) {
    __shared__ unsigned long long lds[kThreads / 64];
    unsigned long long carry = 0;
    for (int64_t base = 0; base < nt; base += kThreads) {
        int64_t i = base + threadIdx.x;
        unsigned long long v = (i < nt) ? sums[i] : 0ull;
        unsigned long long tot;
        unsigned long long ex = block_excl_scan(v, lds, &tot);
        if (i < nt) sums[i] = carry + ex;
        carry += tot;
    }
    if (threadIdx.x == 0) {
        *total = carry;
        out[n] = (uint32_t)carry;
    }
}

__global__ __launch_bounds__(kThreads) void tile_scan(const uint32_t* __restrict__ in, int64_t n,
                                                      const unsigned long long* __restrict__ offs,
                                                      uint32_t* __restrict__ out) {
    __shared__ unsigned long long lds[kThreads / 64];
    uint32_t x[kItems];
    const int64_t base = (int64_t)blockIdx.x * kTile;
    load_items(in, n, base, x);
    unsigned long long s = 0;
#pragma unroll
    for (int k = 0; k < kItems; k++) s += x[k];
    unsigned long long tot;
    unsigned long long run = offs[blockIdx.x] + block_excl_scan(s, lds, &tot);
    int64_t i0 = base + (int64_t)threadIdx.x * kItems;
#pragma unroll
    for (int k = 0; k < kItems; k++) {
        if (i0 + k < n) out[i0 + k] = (uint32_t)run;
        run += x[k];
    }
}

}  // namespace

size_t scan_tmp_words(int64_t n) { return (size_t)ceil_div(n > 0 ? n : 1, kTile); }

e2e_status scan_u32(const uint32_t* in, uint32_t* out, int64_t n, unsigned long long* tmp,
                    unsigned long long* total_dev, hipStream_t s) {
    int64_t nt = ceil_div(n > 0 ? n : 1, kTile);
    E2E_CHECK(nt < (1ll << 31), E2E_ERR_CAPACITY, "scan: %lld elements exceed the grid limit",
              (long long)n);
    hipLaunchKernelGGL(tile_sums, dim3((unsigned)nt), dim3(kThreads), 0, s, in, n, tmp);
    hipLaunchKernelGGL(tile_offsets, dim3(1), dim3(kThreads), 0, s, tmp, nt, total_dev, out, n);
    hipLaunchKernelGGL(tile_scan, dim3((unsigned)nt), dim3(kThreads), 0, s, in, n, tmp, out);
    E2E_HIP(hipGetLastError());
    return E2E_OK;
}

}  // namespace e2e
