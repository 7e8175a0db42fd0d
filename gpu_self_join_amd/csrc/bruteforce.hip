// bruteforce.hip -- the all-pairs self-join: the reference's search mode 1
// (makeDistanceTableGPUBruteForce GPU.cu:1285-1480, kernelBruteForce
// kernel.cu:486-515), which tests every ordered pair (i, j), i itself
// included, and only counts them with one global atomic per pair.
//
// Here: the points are copied once to SoA [d][ld]; one wave per query, 64
// consecutive candidates per step (lane = candidate, coalesced loads, the
// query's coordinates wave-uniform).  The distance is the join's (dist_mem:
// the AccMode order of kernel.cu:307-357) and the predicate acc <= thr the
// join's (kernel.cu:360-365).  A counting pass, a scan and a fill pass give
// per-point counts and CSR rows; a hit's slot is the running row cursor plus
// the ballot rank, so every row comes out ascending without a sort.
#include "bruteforce.h"
#include "join_kernels.h"

namespace gdsj {

template <typename T>
__global__ __launch_bounds__(256) void k_bf_soa(const T *__restrict__ in, uint32_t n, uint32_t d, uint64_t ld,
                                                T *__restrict__ xs) {
    const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= (uint64_t)n * d) return;
    const uint64_t p = i / d, j = i - p * d;
    xs[j * ld + p] = in[i];
}

template <typename T, int PASS>
__global__ __launch_bounds__(256) void k_bruteforce(const T *__restrict__ xs, uint64_t ld, uint32_t n, uint32_t d,
                                                    int mode, uint32_t ilp, double thrd,
                                                    uint32_t *__restrict__ counts,
                                                    const uint64_t *__restrict__ offsets,
                                                    uint32_t *__restrict__ out) {
    const uint32_t q = uniform((uint32_t)(blockIdx.x * (blockDim.x / GDSJ_WAVE) + threadIdx.x / GDSJ_WAVE));
    if (q >= n) return;
    const uint32_t lane = lane_id();
    const T thr = (T)thrd;
    uint32_t cnt = 0;
    uint64_t pos = PASS == PASS_FILL ? uniform64(offsets[q]) : 0ull;
    for (uint32_t c0 = 0; c0 < n; c0 += GDSJ_WAVE) {
        const uint32_t c = c0 + lane;
        const bool valid = c < n;
        // lanes past n test the query against itself and are masked out
        const T acc = dist_mem<T>(xs, ld, q, valid ? c : q, d, mode, ilp);
        const bool hit = valid && acc <= thr;
        const uint64_t m = ballot(hit);
        if (PASS == PASS_FILL) {
            if (hit) out[pos + mbcnt(m)] = c;
            pos += popc64(m);
        } else {
            cnt += popc64(m);
        }
    }
    if (PASS == PASS_COUNT && lane == 0) counts[q] = cnt;
}

int bruteforce_soa(const void *in, uint32_t n, uint32_t d, int dtype, uint64_t ld, void *xs, hipStream_t s) {
    const uint64_t tot = (uint64_t)n * d;
    if (tot == 0) return GDSJ_OK;
    const uint32_t blocks = (uint32_t)((tot + 255) / 256);
    if (dtype == GDSJ_DTYPE_F32)
        k_bf_soa<float><<<blocks, 256, 0, s>>>((const float *)in, n, d, ld, (float *)xs);
    else
        k_bf_soa<double><<<blocks, 256, 0, s>>>((const double *)in, n, d, ld, (double *)xs);
    GDSJ_HIP(hipGetLastError());
    return GDSJ_OK;
}

int bruteforce_pass(const void *xs, uint64_t ld, uint32_t n, uint32_t d, int dtype, int mode, uint32_t ilp,
                    double thr, int pass, uint32_t *counts, const uint64_t *offsets, uint32_t *out, hipStream_t s) {
    if (n == 0) return GDSJ_OK;
    const uint32_t blocks = (n + 3) / 4;
    if (dtype == GDSJ_DTYPE_F32) {
        if (pass == PASS_FILL)
            k_bruteforce<float, PASS_FILL><<<blocks, 256, 0, s>>>((const float *)xs, ld, n, d, mode, ilp, thr,
                                                                  counts, offsets, out);
        else
            k_bruteforce<float, PASS_COUNT><<<blocks, 256, 0, s>>>((const float *)xs, ld, n, d, mode, ilp, thr,
                                                                   counts, offsets, out);
    } else {
        if (pass == PASS_FILL)
            k_bruteforce<double, PASS_FILL><<<blocks, 256, 0, s>>>((const double *)xs, ld, n, d, mode, ilp, thr,
                                                                   counts, offsets, out);
        else
            k_bruteforce<double, PASS_COUNT><<<blocks, 256, 0, s>>>((const double *)xs, ld, n, d, mode, ilp, thr,
                                                                    counts, offsets, out);
    }
    GDSJ_HIP(hipGetLastError());
    return GDSJ_OK;
}

}  // namespace gdsj
