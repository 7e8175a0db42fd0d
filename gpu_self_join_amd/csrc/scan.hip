// scan.hip -- exclusive scan: tile reduce, then tile scan plus carry-in,
// the carry summed by each block itself (up to SCAN_DIRECT_TILES tiles) or
// taken from a scan of the tile sums.
#include "scan.h"

namespace gdsj {

template <typename In, typename Acc>
__global__ __launch_bounds__(SCAN_BLOCK) void k_scan_reduce(const In *__restrict__ in, uint64_t n,
                                                             Acc *__restrict__ partial) {
    const uint64_t base = (uint64_t)blockIdx.x * SCAN_TILE;
    Acc v = 0;
#pragma unroll
    for (int i = 0; i < SCAN_ITEMS; i++) {
        uint64_t idx = base + (uint64_t)i * SCAN_BLOCK + threadIdx.x;
        if (idx < n) v += (Acc)in[idx];
    }
    Acc tot;
    (void)block_exclusive_scan<Acc>(v, &tot);
    if (threadIdx.x == 0) partial[blockIdx.x] = tot;
}

// single block: exclusive scan of nb partial sums in place; partial[nb] = total
template <typename Acc>
__global__ __launch_bounds__(SCAN_BLOCK) void k_scan_partials(Acc *__restrict__ partial, uint32_t nb) {
    Acc carry = 0;
    for (uint32_t base = 0; base < nb; base += SCAN_BLOCK * SCAN_ITEMS) {
        Acc v[SCAN_ITEMS];
        Acc loc = 0;
#pragma unroll
        for (int i = 0; i < SCAN_ITEMS; i++) {
            uint32_t idx = base + threadIdx.x * SCAN_ITEMS + i;
            v[i] = idx < nb ? partial[idx] : 0;
            loc += v[i];
        }
        Acc tot;
        Acc ex = block_exclusive_scan<Acc>(loc, &tot) + carry;
#pragma unroll
        for (int i = 0; i < SCAN_ITEMS; i++) {
            uint32_t idx = base + threadIdx.x * SCAN_ITEMS + i;
            if (idx < nb) partial[idx] = ex;
            ex += v[i];
        }
        carry += tot;
        __syncthreads();
    }
    if (threadIdx.x == 0) partial[nb] = carry;
}

// down-sweep whose block sums the tile sums before its tile itself (no
// separate scan of the tile sums: one launch less, for up to
// SCAN_DIRECT_TILES tiles)
template <typename In, typename Out, typename Acc>
__global__ __launch_bounds__(SCAN_BLOCK) void k_scan_down_direct(const In *__restrict__ in, Out *__restrict__ out,
                                                                  uint64_t n, const Acc *__restrict__ partial,
                                                                  int write_total) {
    Acc c = 0;
    for (uint32_t i = threadIdx.x; i < blockIdx.x; i += SCAN_BLOCK) c += partial[i];
    Acc carry;
    (void)block_exclusive_scan<Acc>(c, &carry);
    const uint64_t base = (uint64_t)blockIdx.x * SCAN_TILE + (uint64_t)threadIdx.x * SCAN_ITEMS;
    Acc v[SCAN_ITEMS];
    Acc loc = 0;
#pragma unroll
    for (int i = 0; i < SCAN_ITEMS; i++) {
        uint64_t idx = base + i;
        v[i] = idx < n ? (Acc)in[idx] : 0;
        loc += v[i];
    }
    Acc tot;
    Acc ex = block_exclusive_scan<Acc>(loc, &tot) + carry;
#pragma unroll
    for (int i = 0; i < SCAN_ITEMS; i++) {
        uint64_t idx = base + i;
        if (idx < n) out[idx] = (Out)ex;
        ex += v[i];
    }
    if (write_total && blockIdx.x == gridDim.x - 1 && threadIdx.x == 0) out[n] = (Out)(carry + tot);
}

// writes the total of an empty scan (a kernel, not a memset, so that the
// scan can be captured in a graph)
template <typename Out>
__global__ void k_scan_empty(Out *__restrict__ out) {
    if (threadIdx.x == 0) out[0] = 0;
}

template <typename In, typename Out, typename Acc>
__global__ __launch_bounds__(SCAN_BLOCK) void k_scan_down(const In *__restrict__ in, Out *__restrict__ out,
                                                           uint64_t n, const Acc *__restrict__ partial,
                                                           uint32_t nb, int write_total) {
    // each thread owns SCAN_ITEMS consecutive items of the tile
    const uint64_t base = (uint64_t)blockIdx.x * SCAN_TILE + (uint64_t)threadIdx.x * SCAN_ITEMS;
    Acc v[SCAN_ITEMS];
    Acc loc = 0;
#pragma unroll
    for (int i = 0; i < SCAN_ITEMS; i++) {
        uint64_t idx = base + i;
        v[i] = idx < n ? (Acc)in[idx] : 0;
        loc += v[i];
    }
    Acc ex = block_exclusive_scan<Acc>(loc, nullptr) + partial[blockIdx.x];
#pragma unroll
    for (int i = 0; i < SCAN_ITEMS; i++) {
        uint64_t idx = base + i;
        if (idx < n) out[idx] = (Out)ex;
        ex += v[i];
    }
    if (write_total && blockIdx.x == 0 && threadIdx.x == 0) out[n] = (Out)partial[nb];
}

__global__ __launch_bounds__(SCAN_BLOCK) void k_sum_u64(const uint64_t *__restrict__ in, uint64_t n,
                                                         unsigned long long *__restrict__ out) {
    uint64_t v = 0;
    for (uint64_t i = (uint64_t)blockIdx.x * SCAN_BLOCK + threadIdx.x; i < n; i += (uint64_t)gridDim.x * SCAN_BLOCK)
        v += in[i];
    uint64_t tot;
    (void)block_exclusive_scan<uint64_t>(v, &tot);
    if (threadIdx.x == 0 && tot) atomicAdd(out, (unsigned long long)tot);
}

int sum_u64(const uint64_t *in, uint64_t n, uint64_t *out, hipStream_t s) {
    GDSJ_HIP(hipMemsetAsync(out, 0, 8, s));
    if (n == 0) return GDSJ_OK;
    uint32_t blocks = (uint32_t)((n + SCAN_BLOCK * 16 - 1) / (SCAN_BLOCK * 16));
    if (blocks > 1024) blocks = 1024;
    k_sum_u64<<<blocks, SCAN_BLOCK, 0, s>>>(in, n, (unsigned long long *)out);
    GDSJ_HIP(hipGetLastError());
    return GDSJ_OK;
}

size_t scan_tmp_bytes(uint64_t n) {
    uint64_t nb = (n + SCAN_TILE - 1) / SCAN_TILE;
    return (nb + 2) * sizeof(uint64_t);
}

template <typename Out, typename Acc>
static int scan_impl(const uint32_t *in, Out *out, uint64_t n, bool write_total, void *tmp,
                     hipStream_t s) {
    uint32_t nb = (uint32_t)((n + SCAN_TILE - 1) / SCAN_TILE);
    Acc *partial = (Acc *)tmp;
    if (n == 0) {
        if (write_total) k_scan_empty<Out><<<1, 64, 0, s>>>(out);
        GDSJ_HIP(hipGetLastError());
        return GDSJ_OK;
    }
    k_scan_reduce<uint32_t, Acc><<<nb, SCAN_BLOCK, 0, s>>>(in, n, partial);
    if (nb <= SCAN_DIRECT_TILES) {
        k_scan_down_direct<uint32_t, Out, Acc><<<nb, SCAN_BLOCK, 0, s>>>(in, out, n, partial, write_total ? 1 : 0);
    } else {
        k_scan_partials<Acc><<<1, SCAN_BLOCK, 0, s>>>(partial, nb);
        k_scan_down<uint32_t, Out, Acc><<<nb, SCAN_BLOCK, 0, s>>>(in, out, n, partial, nb, write_total ? 1 : 0);
    }
    GDSJ_HIP(hipGetLastError());
    return GDSJ_OK;
}

int exclusive_scan_u32_u64(const uint32_t *in, uint64_t *out, uint64_t n, bool write_total, void *tmp,
                           hipStream_t s) {
    return scan_impl<uint64_t, uint64_t>(in, out, n, write_total, tmp, s);
}

int exclusive_scan_u32_u32(const uint32_t *in, uint32_t *out, uint64_t n, bool write_total, void *tmp,
                           hipStream_t s) {
    return scan_impl<uint32_t, uint64_t>(in, out, n, write_total, tmp, s);
}

}  // namespace gdsj
