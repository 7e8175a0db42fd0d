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
#include <cmath>
#include <mutex>

#include "bruteforce.h"
#include "join_kernels.h"
#include "scan.h"

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

// ------------------------------------------------- allocation-free form --
// The scratch of gdsj_bruteforce_device_noalloc: the SoA copy of the points
// and the scan's partial sums, in a fixed array that every device the
// library runs on holds from module load to process end.  Host code never
// uses the array's name as an address: its device address comes from
// hipGetSymbolAddress, once per device (noalloc_scratch below).
static __device__ __align__(256) unsigned char scratch[GDSJ_NOALLOC_SCRATCH_BYTES];

static std::mutex g_noalloc_mu;
static void *g_noalloc_ptr[64] = {};

// This device's address of `scratch` (looked up on the device's first call).
static int noalloc_scratch(int dev, void **out) {
    if (dev < 0 || dev >= 64) {
        set_error("device %d out of range for the allocation-free entry point", dev);
        return GDSJ_EINVAL;
    }
    std::lock_guard<std::mutex> lk(g_noalloc_mu);
    if (!g_noalloc_ptr[dev]) {
        void *p = nullptr;
        GDSJ_HIP(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
        if (!p) {
            set_error("hipGetSymbolAddress returned no address for the scratch array");
            return GDSJ_EDEVICE;
        }
        g_noalloc_ptr[dev] = p;
    }
    *out = g_noalloc_ptr[dev];
    return GDSJ_OK;
}

}  // namespace gdsj

using namespace gdsj;

extern "C" int gdsj_bruteforce_device_noalloc(const void *d_data, uint32_t n, uint32_t dims, int dtype,
                                              double epsilon, const gdsj_params *pin, uint32_t *d_counts,
                                              uint64_t *d_offsets, uint32_t *d_ids, uint64_t ids_cap,
                                              uint64_t *total, void *stream) {
    if ((!d_data && n) || (!d_counts && n) || !d_offsets || !total) return GDSJ_EINVAL;
    if (dtype != GDSJ_DTYPE_F32 && dtype != GDSJ_DTYPE_F64) {
        set_error("dtype must be GDSJ_DTYPE_F32 or GDSJ_DTYPE_F64");
        return GDSJ_EINVAL;
    }
    if (dims == 0 || dims > (uint32_t)MAXD) {
        set_error("dims=%u must be in 1..%d", dims, MAXD);
        return GDSJ_EINVAL;
    }
    if (!(epsilon > 0) || !std::isfinite(epsilon)) {
        set_error("epsilon must be positive and finite");
        return GDSJ_EINVAL;
    }
    gdsj_params p;
    if (pin) p = *pin; else gdsj_params_default(&p);
    if (p.ilp > 64 || p.ilp_fold > 1) {
        set_error("ILP=%u exceeds 64 or ilp_fold=%u is not 0 or 1", p.ilp, p.ilp_fold);
        return GDSJ_EINVAL;
    }
    double thr;
    if (dtype == GDSJ_DTYPE_F32) {
        const float e = (float)epsilon;
        if ((double)e != epsilon) {
            set_error("epsilon %.17g is not a float value", epsilon);
            return GDSJ_EINVAL;
        }
        thr = (double)gdsj_eps_threshold_f32(e);
    } else {
        thr = gdsj_eps_threshold_f64(epsilon);
    }
    // scratch layout: SoA copy [dims][ld], then the scan's partials
    const size_t es = dtype == GDSJ_DTYPE_F64 ? 8 : 4;
    const uint64_t ld = ((uint64_t)n + 63) / 64 * 64 + 64;
    const uint64_t soaBytes = (ld * dims * es + 255) / 256 * 256;
    const uint64_t need = soaBytes + scan_tmp_bytes((uint64_t)n + 1) + 256;
    if (need > (uint64_t)GDSJ_NOALLOC_SCRATCH_BYTES) {
        set_error("n=%u x dims=%u needs %llu bytes of scratch; the allocation-free entry point holds %llu", n,
                  dims, (unsigned long long)need, (unsigned long long)GDSJ_NOALLOC_SCRATCH_BYTES);
        return GDSJ_EINVAL;
    }
    int dev = 0;
    GDSJ_HIP(hipGetDevice(&dev));
    void *base = nullptr;
    GDSJ_TRY(noalloc_scratch(dev, &base));
    unsigned char *xs = (unsigned char *)base;
    void *scanTmp = xs + soaBytes;
    hipStream_t s = (hipStream_t)stream;
    const int mode = p.cpu_sum ? ACC_UNFUSED : acc_mode(p.ilp, p.shortcircuit, p.ilp_fold);
    GDSJ_TRY(bruteforce_soa(d_data, n, dims, dtype, ld, xs, s));
    GDSJ_TRY(bruteforce_pass(xs, ld, n, dims, dtype, mode, p.ilp, thr, PASS_COUNT, d_counts, nullptr, nullptr, s));
    GDSJ_TRY(exclusive_scan_u32_u64(d_counts, d_offsets, n, true, scanTmp, s));
    uint64_t tot = 0;
    GDSJ_HIP(hipMemcpyAsync(&tot, d_offsets + n, 8, hipMemcpyDeviceToHost, s));
    GDSJ_HIP(hipStreamSynchronize(s));
    *total = tot;
    if (!d_ids) return GDSJ_OK;
    if (tot > ids_cap) {
        set_error("result has %llu pairs, output buffer holds %llu", (unsigned long long)tot,
                  (unsigned long long)ids_cap);
        return GDSJ_EOVERFLOW;
    }
    GDSJ_TRY(bruteforce_pass(xs, ld, n, dims, dtype, mode, p.ilp, thr, PASS_FILL, nullptr, d_offsets, d_ids, s));
    GDSJ_HIP(hipStreamSynchronize(s));
    return GDSJ_OK;
}
