// sortrows.hip -- ascending order inside each CSR row (optional output
// form; the reference's row order follows its atomics and is unspecified).
// Rows of <= 64 ids: one wave, bitonic network over shuffles.  Longer rows:
// one workgroup each, bitonic network in LDS (<= 8192 ids) or in place in
// global memory, with out-of-row slots treated as +infinity.
#include "sortrows.h"

namespace gdsj {

constexpr uint32_t SR_LDS = 8192;

__global__ __launch_bounds__(256) void k_sort_rows_small(const uint64_t *__restrict__ offs, uint32_t n,
                                                          uint32_t *__restrict__ ids, uint32_t *__restrict__ big,
                                                          uint32_t *__restrict__ nbig) {
    const uint32_t row = blockIdx.x * 4 + threadIdx.x / 64;
    if (row >= n) return;
    const uint32_t lane = lane_id();
    const uint64_t b = offs[row], e = offs[row + 1];
    const uint64_t len = e - b;
    if (len <= 1) return;
    if (len > 64) {
        if (lane == 0) big[atomicAdd(nbig, 1u)] = row;
        return;
    }
    uint32_t v = lane < len ? ids[b + lane] : 0xFFFFFFFFu;
    for (uint32_t k = 2; k <= 64; k <<= 1) {
        for (uint32_t j = k >> 1; j > 0; j >>= 1) {
            uint32_t o = __shfl_xor(v, (int)j, 64);
            bool up = (lane & k) == 0;
            bool lower = (lane & j) == 0;
            uint32_t mn = v < o ? v : o, mx = v < o ? o : v;
            v = (lower == up) ? mn : mx;
        }
    }
    if (lane < len) ids[b + lane] = v;
}

__global__ __launch_bounds__(256) void k_sort_rows_big(const uint64_t *__restrict__ offs,
                                                        uint32_t *__restrict__ ids,
                                                        const uint32_t *__restrict__ big,
                                                        const uint32_t *__restrict__ nbig) {
    __shared__ uint32_t sh[SR_LDS];
    const uint32_t cnt = *nbig;
    for (uint32_t bi = blockIdx.x; bi < cnt; bi += gridDim.x) {
        const uint32_t row = big[bi];
        const uint64_t b = offs[row];
        const uint64_t len = offs[row + 1] - b;
        uint64_t p2 = 1;
        while (p2 < len) p2 <<= 1;
        if (p2 <= SR_LDS) {
            for (uint64_t i = threadIdx.x; i < p2; i += blockDim.x) sh[i] = i < len ? ids[b + i] : 0xFFFFFFFFu;
            __syncthreads();
            for (uint64_t k = 2; k <= p2; k <<= 1) {
                for (uint64_t j = k >> 1; j > 0; j >>= 1) {
                    for (uint64_t i = threadIdx.x; i < p2; i += blockDim.x) {
                        uint64_t l = i ^ j;
                        if (l > i) {
                            bool up = (i & k) == 0;
                            uint32_t x = sh[i], y = sh[l];
                            if ((x > y) == up) {
                                sh[i] = y;
                                sh[l] = x;
                            }
                        }
                    }
                    __syncthreads();
                }
            }
            for (uint64_t i = threadIdx.x; i < len; i += blockDim.x) ids[b + i] = sh[i];
            __syncthreads();
        } else {
            uint32_t *r = ids + b;
            for (uint64_t k = 2; k <= p2; k <<= 1) {
                for (uint64_t j = k >> 1; j > 0; j >>= 1) {
                    for (uint64_t i = threadIdx.x; i < p2; i += blockDim.x) {
                        uint64_t l = i ^ j;
                        if (l > i && i < len) {
                            bool up = (i & k) == 0;
                            uint32_t x = r[i];
                            uint32_t y = l < len ? r[l] : 0xFFFFFFFFu;
                            if ((x > y) == up) {
                                // y is +inf only when l >= len; then x > y is false
                                r[i] = y;
                                r[l] = x;
                            }
                        }
                    }
                    __syncthreads();
                }
            }
        }
    }
}

int sort_rows(const uint64_t *offs, uint32_t n, uint32_t *ids, hipStream_t s, DevBuf &work) {
    if (n == 0) return GDSJ_OK;
    GDSJ_TRY(work.ensure((size_t)n * 4 + 256));
    uint32_t *nbig = work.as<uint32_t>();
    uint32_t *big = nbig + 64;
    GDSJ_HIP(hipMemsetAsync(nbig, 0, 4, s));
    k_sort_rows_small<<<(n + 3) / 4, 256, 0, s>>>(offs, n, ids, big, nbig);
    k_sort_rows_big<<<1024, 256, 0, s>>>(offs, ids, big, nbig);
    GDSJ_HIP(hipGetLastError());
    return GDSJ_OK;
}

}  // namespace gdsj
