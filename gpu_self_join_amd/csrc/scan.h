// scan.h -- device-wide exclusive prefix sums (2048-item tiles).
// Used for CSR offsets, cell heads, neighbor runs, tiles and radix digits.
#pragma once
#include "common.h"

namespace gdsj {

constexpr int SCAN_BLOCK = 256;
constexpr int SCAN_ITEMS = 8;
constexpr int SCAN_TILE = SCAN_BLOCK * SCAN_ITEMS;
// tiles up to which each down-sweep block sums the tile sums before it
constexpr uint32_t SCAN_DIRECT_TILES = 1024;

template <typename Acc>
__device__ __forceinline__ Acc wave_inclusive_scan(Acc v) {
    const uint32_t lane = lane_id();
#pragma unroll
    for (int off = 1; off < GDSJ_WAVE; off <<= 1) {
        Acc o = __shfl_up(v, off, GDSJ_WAVE);
        if (lane >= (uint32_t)off) v += o;
    }
    return v;
}

// exclusive scan of one value per thread over a 256-thread block
template <typename Acc>
__device__ __forceinline__ Acc block_exclusive_scan(Acc v, Acc *block_total) {
    __shared__ Acc wsum[SCAN_BLOCK / GDSJ_WAVE];
    const int w = threadIdx.x / GDSJ_WAVE;
    const uint32_t lane = lane_id();
    Acc inc = wave_inclusive_scan(v);
    if (lane == GDSJ_WAVE - 1) wsum[w] = inc;
    __syncthreads();
    Acc base = 0, tot = 0;
#pragma unroll
    for (int i = 0; i < SCAN_BLOCK / GDSJ_WAVE; i++) {
        Acc s = wsum[i];
        if (i < w) base += s;
        tot += s;
    }
    __syncthreads();
    if (block_total) *block_total = tot;
    return base + inc - v;
}

// host entry points (scan.hip).  tmp must hold scan_tmp_bytes(n) bytes.
size_t scan_tmp_bytes(uint64_t n);
// out[i] = sum_{j<i} in[j] for i < n; out[n] = total when write_total.
int exclusive_scan_u32_u64(const uint32_t *in, uint64_t *out, uint64_t n, bool write_total,
                           void *tmp, hipStream_t s);
int exclusive_scan_u32_u32(const uint32_t *in, uint32_t *out, uint64_t n, bool write_total,
                           void *tmp, hipStream_t s);

// *out = sum of in[0..n) (uint64), on stream s.
int sum_u64(const uint64_t *in, uint64_t n, uint64_t *out, hipStream_t s);

}  // namespace gdsj
