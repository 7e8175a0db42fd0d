// radix.hip -- stable LSD radix sort of (uint64 key, uint32 value) pairs,
// 8-bit digits.  Used to order points by linearized grid-cell id (the
// reference builds this order with a std::set + lower_bound on the host,
// main.cu:707-763); stability keeps the points of one cell in ascending
// point id, as the reference's per-cell push_back does (main.cu:762).
//
// Per pass: (1) per-tile digit histogram in LDS, (2) exclusive scan of the
// digit-major histogram, (3) scatter with a stable in-tile rank built from
// wave64 ballots (peer mask per digit) and per-wave LDS counts.
#include "radix.h"
#include "scan.h"

namespace gdsj {

constexpr int RADIX_BLOCK = 256;
constexpr int RADIX_ROUNDS = 16;
constexpr int RADIX_TILE = RADIX_BLOCK * RADIX_ROUNDS;
constexpr int RADIX_BINS = 256;

__global__ __launch_bounds__(RADIX_BLOCK) void k_radix_hist(const uint64_t *__restrict__ keys, uint32_t n,
                                                             int shift, uint32_t *__restrict__ hist,
                                                             uint32_t nb) {
    __shared__ uint32_t h[RADIX_BINS];
    h[threadIdx.x] = 0;
    __syncthreads();
    const uint64_t base = (uint64_t)blockIdx.x * RADIX_TILE;
#pragma unroll 4
    for (int r = 0; r < RADIX_ROUNDS; r++) {
        uint64_t idx = base + (uint64_t)r * RADIX_BLOCK + threadIdx.x;
        if (idx < n) atomicAdd(&h[(uint32_t)(keys[idx] >> shift) & 255u], 1u);
    }
    __syncthreads();
    hist[(uint64_t)threadIdx.x * nb + blockIdx.x] = h[threadIdx.x];
}

__global__ __launch_bounds__(RADIX_BLOCK) void k_radix_scatter(const uint64_t *__restrict__ kin,
                                                                const uint32_t *__restrict__ vin,
                                                                uint64_t *__restrict__ kout,
                                                                uint32_t *__restrict__ vout, uint32_t n,
                                                                int shift, const uint32_t *__restrict__ offs,
                                                                uint32_t nb) {
    constexpr int NW = RADIX_BLOCK / GDSJ_WAVE;
    __shared__ uint32_t run[RADIX_BINS];
    __shared__ uint32_t base_off[RADIX_BINS];
    __shared__ uint32_t wcnt[NW][RADIX_BINS];
    const int tid = threadIdx.x;
    const int w = tid / GDSJ_WAVE;
    const uint32_t lane = lane_id();
    const uint64_t lt_mask = (lane == 0) ? 0ull : (~0ull >> (64 - lane));
    run[tid] = 0;
    base_off[tid] = offs[(uint64_t)tid * nb + blockIdx.x];
#pragma unroll
    for (int i = 0; i < NW; i++) wcnt[i][tid] = 0;
    __syncthreads();
    const uint64_t base = (uint64_t)blockIdx.x * RADIX_TILE;
    for (int r = 0; r < RADIX_ROUNDS; r++) {
        if (base + (uint64_t)r * RADIX_BLOCK >= n) break;  // uniform across the block
        const uint64_t idx = base + (uint64_t)r * RADIX_BLOCK + tid;
        const bool valid = idx < n;
        const uint64_t key = valid ? kin[idx] : 0ull;
        const uint32_t dig = (uint32_t)(key >> shift) & 255u;
        uint64_t peers = ballot(valid);
#pragma unroll
        for (int b = 0; b < 8; b++) {
            const bool bit = (dig >> b) & 1u;
            const uint64_t m = ballot(bit);
            peers &= bit ? m : ~m;
        }
        const uint64_t below = peers & lt_mask;
        const uint32_t rank = popc64(below);
        if (valid && below == 0) wcnt[w][dig] = popc64(peers);
        __syncthreads();
        if (valid) {
            uint32_t pos = base_off[dig] + run[dig] + rank;
#pragma unroll
            for (int i = 0; i < NW; i++)
                if (i < w) pos += wcnt[i][dig];
            kout[pos] = key;
            vout[pos] = vin[idx];
        }
        __syncthreads();
        uint32_t add = 0;
#pragma unroll
        for (int i = 0; i < NW; i++) {
            add += wcnt[i][tid];
            wcnt[i][tid] = 0;
        }
        run[tid] += add;
        __syncthreads();
    }
}

size_t radix_tmp_bytes(uint32_t n) {
    uint32_t nb = (n + RADIX_TILE - 1) / RADIX_TILE;
    uint64_t hn = (uint64_t)nb * RADIX_BINS;
    size_t hist = hn * sizeof(uint32_t) * 2 + 256;
    return hist + scan_tmp_bytes(hn) + 256;
}

int radix_sort_pairs(uint64_t *keys, uint32_t *vals, uint64_t *keys_alt, uint32_t *vals_alt, uint32_t n,
                     int bits, void *tmp, hipStream_t s, int *result_in_alt) {
    *result_in_alt = 0;
    if (n <= 1 || bits <= 0) return GDSJ_OK;
    const uint32_t nb = (n + RADIX_TILE - 1) / RADIX_TILE;
    const uint64_t hn = (uint64_t)nb * RADIX_BINS;
    uint32_t *hist = (uint32_t *)tmp;
    uint32_t *offs = hist + hn;
    void *stmp = (void *)(((uintptr_t)(offs + hn) + 255) & ~(uintptr_t)255);
    uint64_t *ki = keys, *ko = keys_alt;
    uint32_t *vi = vals, *vo = vals_alt;
    int alt = 0;
    for (int shift = 0; shift < bits; shift += 8) {
        k_radix_hist<<<nb, RADIX_BLOCK, 0, s>>>(ki, n, shift, hist, nb);
        GDSJ_TRY(exclusive_scan_u32_u32(hist, offs, hn, false, stmp, s));
        k_radix_scatter<<<nb, RADIX_BLOCK, 0, s>>>(ki, vi, ko, vo, n, shift, offs, nb);
        GDSJ_HIP(hipGetLastError());
        uint64_t *tk = ki; ki = ko; ko = tk;
        uint32_t *tv = vi; vi = vo; vo = tv;
        alt ^= 1;
    }
    *result_in_alt = alt;
    return GDSJ_OK;
}

}  // namespace gdsj
