// grid.hip -- device epsilon-grid build (see grid.h for the HBM layout).
//
// Reference steps restated (all arithmetic in DTYPE, IEEE round-to-nearest,
// compiled with -ffp-contract=off):
//   REORDER      main.cu:997-1067  variance over rows 0,100,200,... and a
//                                  descending sort of the dimensions
//   grid extent  main.cu:899-946   min/max per indexed dim, -/+ eps,
//                                  nCells = ceil((max-min)/eps)
//   cell ids     main.cu:713,744   (unsigned)((x-min)/eps), linear id with
//                                  dim 0 fastest (main.cu:867-883)
//   cell lists   main.cu:741-794   points of a cell in ascending id order
#include <math.h>

#include <cmath>
#include <string.h>

#include <algorithm>
#include <vector>

#include "grid.h"
#include "radix.h"
#include "scan.h"

namespace gdsj {

struct KeyArgs {
    uint32_t k;
    uint32_t pk[MAXK];  // source column of indexed dim j (after REORDER)
    uint32_t nCells[MAXK];
    uint64_t mult[MAXK];
    double minArr[MAXK];
};

// ------------------------------------------------------------- REORDER --
template <typename T>
__global__ void k_sample_gather(const T *__restrict__ data, uint32_t n, uint32_t d, uint32_t m,
                                T *__restrict__ out) {
    uint64_t idx = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (idx >= (uint64_t)m * d) return;
    uint32_t j = (uint32_t)(idx / d), l = (uint32_t)(idx % d);
    out[(uint64_t)l * m + j] = data[(uint64_t)j * 100u * d + l];
}

// one thread per dimension: the reference's sequential sums, unfused
template <typename T>
__global__ void k_variance(const T *__restrict__ samples, uint32_t m, uint32_t d, T denom,
                           T *__restrict__ var) {
    uint32_t l = blockIdx.x * blockDim.x + threadIdx.x;
    if (l >= d) return;
    const T *col = samples + (uint64_t)l * m;
    T sum = 0;
    for (uint32_t j = 0; j < m; j++) sum = sum + col[j];
    T avg = sum / denom;
    sum = 0;
    for (uint32_t j = 0; j < m; j++) {
        T t = col[j] - avg;
        T sq = t * t;
        sum = sum + sq;
    }
    var[l] = sum / denom;
}

// ------------------------------------------------------------ extent --
template <typename T>
__global__ __launch_bounds__(256) void k_minmax(const T *__restrict__ data, uint32_t n, uint32_t d,
                                                 KeyArgs a, T *__restrict__ partial) {
    __shared__ T smin[MAXK][256 / GDSJ_WAVE], smax[MAXK][256 / GDSJ_WAVE];
    T mn[MAXK], mx[MAXK];
    const uint32_t k = a.k;
    for (uint32_t j = 0; j < k; j++) {
        mn[j] = data[a.pk[j]];
        mx[j] = mn[j];
    }
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
         i += (uint64_t)gridDim.x * blockDim.x) {
        for (uint32_t j = 0; j < k; j++) {
            T v = data[i * d + a.pk[j]];
            mn[j] = v < mn[j] ? v : mn[j];
            mx[j] = v > mx[j] ? v : mx[j];
        }
    }
    const int w = threadIdx.x / GDSJ_WAVE;
    for (uint32_t j = 0; j < k; j++) {
        T a0 = mn[j], a1 = mx[j];
        for (int off = 32; off > 0; off >>= 1) {
            T b0 = __shfl_xor(a0, off, GDSJ_WAVE), b1 = __shfl_xor(a1, off, GDSJ_WAVE);
            a0 = b0 < a0 ? b0 : a0;
            a1 = b1 > a1 ? b1 : a1;
        }
        if (lane_id() == 0) {
            smin[j][w] = a0;
            smax[j][w] = a1;
        }
    }
    __syncthreads();
    if (threadIdx.x < k) {
        uint32_t j = threadIdx.x;
        T a0 = smin[j][0], a1 = smax[j][0];
        for (int i = 1; i < 256 / GDSJ_WAVE; i++) {
            a0 = smin[j][i] < a0 ? smin[j][i] : a0;
            a1 = smax[j][i] > a1 ? smax[j][i] : a1;
        }
        partial[((uint64_t)blockIdx.x * k + j) * 2 + 0] = a0;
        partial[((uint64_t)blockIdx.x * k + j) * 2 + 1] = a1;
    }
}

// ----------------------------------------------------------- cell keys --
template <typename T>
__global__ void k_cell_keys(const T *__restrict__ data, uint32_t n, uint32_t d, KeyArgs a, T eps,
                            uint64_t *__restrict__ keys, uint32_t *__restrict__ vals) {
    uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    uint64_t lin = 0;
    for (uint32_t j = 0; j < a.k; j++) {
        T x = data[(uint64_t)i * d + a.pk[j]];
        T q = (x - (T)a.minArr[j]) / eps;  // IEEE quotient, truncated below
        uint32_t c = (uint32_t)q;
        lin += (uint64_t)c * a.mult[j];
    }
    keys[i] = lin;
    vals[i] = i;
}

__global__ void k_heads(const uint64_t *__restrict__ keys, uint32_t n, uint32_t *__restrict__ flag) {
    uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    flag[i] = (i == 0 || keys[i] != keys[i - 1]) ? 1u : 0u;
}

__global__ void k_cells(const uint64_t *__restrict__ keys, const uint32_t *__restrict__ flag,
                        const uint32_t *__restrict__ cidx, uint32_t n, uint64_t *__restrict__ cellKey,
                        uint32_t *__restrict__ cellStart) {
    uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    if (flag[i]) {
        uint32_t c = cidx[i];
        cellKey[c] = keys[i];
        cellStart[c] = i;
    }
    if (i == n - 1) cellStart[cidx[n]] = n;
}

// SoA gather in sorted order with the dimension permutation applied
template <typename T>
__global__ void k_soa(const T *__restrict__ data, uint32_t n, uint32_t d, const uint32_t *__restrict__ perm,
                      const uint32_t *__restrict__ order, T *__restrict__ xs, uint64_t ld,
                      uint32_t *__restrict__ ids, uint32_t *__restrict__ nonfinite) {
    uint32_t s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= n) return;
    uint32_t id = order[s];
    ids[s] = id;
    const T *row = data + (uint64_t)id * d;
    bool bad = false;
    for (uint32_t l = 0; l < d; l++) {
        T v = row[perm[l]];
        bad |= !isfinite(v);
        xs[(uint64_t)l * ld + s] = v;
    }
    if (bad) atomicOr(nonfinite, 1u);
}

// ------------------------------------------------------ neighbor runs --
__device__ __forceinline__ uint32_t lower_bound_u64(const uint64_t *__restrict__ a, uint32_t lo, uint32_t hi,
                                                    uint64_t v) {
    while (lo < hi) {
        uint32_t mid = (lo + hi) >> 1;
        if (a[mid] < v) lo = mid + 1; else hi = mid;
    }
    return lo;
}

__device__ __forceinline__ uint32_t upper_bound_u64(const uint64_t *__restrict__ a, uint32_t lo, uint32_t hi,
                                                    uint64_t v) {
    while (lo < hi) {
        uint32_t mid = (lo + hi) >> 1;
        if (a[mid] <= v) lo = mid + 1; else hi = mid;
    }
    return lo;
}

// Per non-empty cell: every non-empty cell whose coordinates differ by at
// most one in each indexed dim (the reference's loop nest kernel.cu:292-300
// with its per-dimension clamps kernel.cu:183-187; its occupancy masks
// kernel.cu:208-253 only skip coordinates no point occupies, which cannot
// change the set).  Walks the sorted linear ids from the most significant
// dim down, pruning empty prefixes; each leaf is a run of up to three cells
// consecutive in dim 0, emitted as a sorted-position range.
template <bool FILL>
__global__ void k_runs(const uint64_t *__restrict__ cellKey, const uint32_t *__restrict__ cellStart, uint32_t nc,
                       KeyArgs a, const uint32_t *__restrict__ runOff, uint32_t *__restrict__ runCount,
                       uint2 *__restrict__ runs, uint32_t *__restrict__ cellCand,
                       unsigned long long *__restrict__ candTotal) {
    uint32_t cell = blockIdx.x * blockDim.x + threadIdx.x;
    uint64_t myCand = 0;
    if (cell < nc) {
        const uint32_t k = a.k;
        uint32_t lo[MAXK], hi[MAXK], v[MAXK], A[MAXK], B[MAXK];
        uint64_t P[MAXK];
        uint64_t key = cellKey[cell];
        for (uint32_t j = 0; j < k; j++) {
            uint32_t c = (uint32_t)(key % a.nCells[j]);
            key /= a.nCells[j];
            lo[j] = c == 0 ? 0 : c - 1;
            uint32_t h = c + 1;
            hi[j] = (a.nCells[j] - 1 < h) ? a.nCells[j] - 1 : h;
        }
        uint32_t nr = 0;
        uint32_t out = FILL ? runOff[cell] : 0;
        int j = (int)k - 1;
        A[j] = 0;
        B[j] = nc;
        P[j] = 0;
        v[j] = lo[j];
        for (;;) {
            if (j == 0) {
                uint32_t x = lower_bound_u64(cellKey, A[0], B[0], P[0] + lo[0]);
                uint32_t y = upper_bound_u64(cellKey, x, B[0], P[0] + hi[0]);
                if (x < y) {
                    uint32_t s0 = cellStart[x], s1 = cellStart[y];
                    if (FILL) runs[out + nr] = make_uint2(s0, s1);
                    myCand += s1 - s0;
                    nr++;
                }
                j = 1;
                if (j > (int)k - 1) break;
                v[j]++;
                continue;
            }
            if (v[j] > hi[j]) {
                if (j == (int)k - 1) break;
                j++;
                v[j]++;
                continue;
            }
            uint64_t lowId = P[j] + (uint64_t)v[j] * a.mult[j];
            uint64_t highId = lowId + a.mult[j] - 1;
            uint32_t x = lower_bound_u64(cellKey, A[j], B[j], lowId);
            uint32_t y = upper_bound_u64(cellKey, x, B[j], highId);
            if (x < y) {
                j--;
                A[j] = x;
                B[j] = y;
                P[j] = lowId;
                v[j] = lo[j];
            } else {
                v[j]++;
            }
        }
        if (!FILL) runCount[cell] = nr;
        if (FILL) {
            cellCand[cell] = (uint32_t)myCand;
            myCand *= (uint64_t)(cellStart[cell + 1] - cellStart[cell]);
        }
    }
    if (FILL) {
        for (int off = 32; off > 0; off >>= 1) myCand += __shfl_xor(myCand, off, GDSJ_WAVE);
        if (lane_id() == 0 && myCand) atomicAdd(candTotal, (unsigned long long)myCand);
    }
}

// ------------------------------------------------------------- tiles --
__global__ void k_tiles_count(const uint32_t *__restrict__ cellStart, uint32_t nc, uint32_t tileq,
                              uint32_t *__restrict__ cnt) {
    uint32_t c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= nc) return;
    uint32_t m = cellStart[c + 1] - cellStart[c];
    cnt[c] = (m + tileq - 1) / tileq;
}

__global__ void k_tiles_fill(const uint32_t *__restrict__ cellStart, const uint32_t *__restrict__ cellCand,
                             uint32_t nc, uint32_t tileq, const uint32_t *__restrict__ tileOff,
                             uint4 *__restrict__ tiles, uint64_t *__restrict__ tileKey,
                             uint32_t *__restrict__ tileVal, int order) {
    uint32_t c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= nc) return;
    uint32_t s0 = cellStart[c], s1 = cellStart[c + 1];
    uint32_t t = tileOff[c];
    for (uint32_t q = s0; q < s1; q += tileq, t++) {
        uint32_t q1 = q + tileq < s1 ? q + tileq : s1;
        tiles[t] = make_uint4(c, q, q1, cellCand[c]);
        if (order) {
            // heaviest tiles first (QUERYREORDER, GPU.cu:1491-1525, per tile)
            uint64_t cost = (uint64_t)(q1 - q) * cellCand[c];
            tileKey[t] = ~cost & 0xFFFFFFFFFFull;
            tileVal[t] = t;
        }
    }
}

__global__ void k_tiles_permute(const uint4 *__restrict__ in, const uint32_t *__restrict__ order, uint32_t nt,
                                uint4 *__restrict__ out) {
    uint32_t t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= nt) return;
    out[t] = in[order[t]];
}

// ---------------------------------------------------------------- host --
void Grid::release() {
    DevBuf *b[] = {&xs, &ids, &cellKey, &cellStart, &runOff, &runs, &tiles, &keyA, &keyB, &valB,
                   &tmp, &aux, &flag, &tileKey, &tileVal, &cellAux, &scalars};
    for (DevBuf *p : b) p->release();
}

static int bit_length(uint64_t v) {
    int b = 0;
    while (v) { b++; v >>= 1; }
    return b;
}

// host restatement of generateNDGridDimensions' DTYPE arithmetic
template <typename T>
static int host_extent(Grid &G, const T *mn, const T *mx) {
    uint64_t total = 1;
    T eps = (T)G.eps;
    for (uint32_t j = 0; j < G.k; j++) {
        T lo = mn[j] - eps;
        T hi = mx[j] + eps;
        T q = (hi - lo) / eps;
        T c = sizeof(T) == 4 ? (T)ceilf((float)q) : (T)ceil((double)q);
        if (!(c >= 1) || c > (T)4294967295.0) {
            set_error("grid extent out of range in dim %u", j);
            return GDSJ_ERANGE;
        }
        G.g.nCells[j] = (uint32_t)c;
        G.g.minArr[j] = (double)lo;
        G.g.mult[j] = total;
        uint64_t nt = total * (uint64_t)G.g.nCells[j];
        if (nt / G.g.nCells[j] != total) {
            set_error("total grid cells overflow uint64 (dims 0..%u)", j);
            return GDSJ_ERANGE;
        }
        total = nt;
    }
    G.totalCells = total;
    return GDSJ_OK;
}

template <typename T>
static int grid_build_t(Grid &G, const T *data, hipStream_t s, int reorder, int order_tiles) {
    const uint32_t n = G.n, d = G.d, k = G.k;
    // ---- REORDER
    for (uint32_t l = 0; l < d; l++) G.perm[l] = l;
    if (reorder && n > 0) {
        uint32_t m = (n + 99) / 100;
        GDSJ_TRY(G.aux.ensure((size_t)m * d * sizeof(T) + d * sizeof(T) + 256));
        T *samp = G.aux.as<T>();
        T *var = samp + (size_t)m * d;
        uint64_t tot = (uint64_t)m * d;
        k_sample_gather<T><<<(uint32_t)((tot + 255) / 256), 256, 0, s>>>(data, n, d, m, samp);
        const T inv_sample = (T)(1.0 / (100 * 1.0));
        const T denom = (T)n * inv_sample;
        k_variance<T><<<(d + 63) / 64, 64, 0, s>>>(samp, m, d, denom, var);
        std::vector<T> hv(d);
        GDSJ_HIP(hipMemcpyAsync(hv.data(), var, d * sizeof(T), hipMemcpyDeviceToHost, s));
        GDSJ_HIP(hipStreamSynchronize(s));
        // descending variance, ties in ascending dim order
        for (uint32_t i = 1; i < d; i++) {
            uint32_t pi = G.perm[i];
            T vv = hv[pi];
            int64_t j = (int64_t)i - 1;
            while (j >= 0 && hv[G.perm[j]] < vv) {
                G.perm[j + 1] = G.perm[j];
                j--;
            }
            G.perm[j + 1] = pi;
        }
    }
    KeyArgs a{};
    a.k = k;
    for (uint32_t j = 0; j < k; j++) a.pk[j] = G.perm[j];
    if (n == 0) {
        G.ncell = G.nruns = G.ntiles = 0;
        return GDSJ_OK;
    }
    // ---- extent
    const uint32_t mmBlocks = std::min<uint32_t>(1024, ceil_div(n, 256 * 4));
    GDSJ_TRY(G.aux.ensure((size_t)mmBlocks * k * 2 * sizeof(T) + MAXD * 4 + 512));
    T *part = G.aux.as<T>();
    k_minmax<T><<<mmBlocks, 256, 0, s>>>(data, n, d, a, part);
    std::vector<T> hp((size_t)mmBlocks * k * 2);
    GDSJ_HIP(hipMemcpyAsync(hp.data(), part, hp.size() * sizeof(T), hipMemcpyDeviceToHost, s));
    GDSJ_HIP(hipStreamSynchronize(s));
    T mn[MAXK], mx[MAXK];
    for (uint32_t j = 0; j < k; j++) {
        mn[j] = hp[j * 2];
        mx[j] = hp[j * 2 + 1];
        for (uint32_t b = 1; b < mmBlocks; b++) {
            T v0 = hp[((size_t)b * k + j) * 2], v1 = hp[((size_t)b * k + j) * 2 + 1];
            mn[j] = v0 < mn[j] ? v0 : mn[j];
            mx[j] = v1 > mx[j] ? v1 : mx[j];
        }
        if (!std::isfinite((double)mn[j]) || !std::isfinite((double)mx[j])) {
            set_error("non-finite coordinate in indexed dim %u", j);
            return GDSJ_EINVAL;
        }
    }
    GDSJ_TRY(host_extent<T>(G, mn, mx));
    for (uint32_t j = 0; j < k; j++) {
        a.nCells[j] = G.g.nCells[j];
        a.mult[j] = G.g.mult[j];
        a.minArr[j] = G.g.minArr[j];
    }
    // ---- keys + sort
    GDSJ_TRY(G.keyA.ensure((size_t)n * 8));
    GDSJ_TRY(G.keyB.ensure((size_t)n * 8));
    GDSJ_TRY(G.ids.ensure((size_t)n * 4 + 256));
    GDSJ_TRY(G.valB.ensure((size_t)n * 4 + 256));
    GDSJ_TRY(G.tmp.ensure(std::max(radix_tmp_bytes(n), scan_tmp_bytes((uint64_t)n + 1)) + 1024));
    uint64_t *kA = G.keyA.as<uint64_t>(), *kB = G.keyB.as<uint64_t>();
    uint32_t *vA = G.ids.as<uint32_t>(), *vB = G.valB.as<uint32_t>();
    k_cell_keys<T><<<ceil_div(n, 256), 256, 0, s>>>(data, n, d, a, (T)G.eps, kA, vA);
    int alt = 0;
    int bits = bit_length(G.totalCells - 1);
    GDSJ_TRY(radix_sort_pairs(kA, vA, kB, vB, n, bits, G.tmp.p, s, &alt));
    uint64_t *skeys = alt ? kB : kA;
    uint32_t *svals = alt ? vB : vA;
    // ---- cells
    GDSJ_TRY(G.flag.ensure(((size_t)n + 1) * 4 * 2 + 64));
    uint32_t *flag = G.flag.as<uint32_t>();
    uint32_t *cidx = flag + n + 1;
    k_heads<<<ceil_div(n, 256), 256, 0, s>>>(skeys, n, flag);
    GDSJ_TRY(exclusive_scan_u32_u32(flag, cidx, n, true, G.tmp.p, s));
    uint32_t nc = 0;
    GDSJ_HIP(hipMemcpyAsync(&nc, cidx + n, 4, hipMemcpyDeviceToHost, s));
    GDSJ_HIP(hipStreamSynchronize(s));
    G.ncell = nc;
    GDSJ_TRY(G.cellKey.ensure((size_t)nc * 8 + 64));
    GDSJ_TRY(G.cellStart.ensure(((size_t)nc + 1) * 4 + 64));
    k_cells<<<ceil_div(n, 256), 256, 0, s>>>(skeys, flag, cidx, n, G.cellKey.as<uint64_t>(),
                                              G.cellStart.as<uint32_t>());
    // ---- SoA (sorted order, dims permuted)
    G.ld = ((uint64_t)n + 63) / 64 * 64;
    GDSJ_TRY(G.xs.ensure((size_t)G.ld * d * sizeof(T) + 256));
    // perm to device (in aux, after the min/max partials)
    uint32_t *dperm = (uint32_t *)((char *)G.aux.p + (((size_t)mmBlocks * k * 2 * sizeof(T) + 255) & ~(size_t)255));
    GDSJ_HIP(hipMemcpyAsync(dperm, G.perm, d * 4, hipMemcpyHostToDevice, s));
    GDSJ_TRY(G.scalars.ensure(256));
    uint32_t *nonfinite = G.scalars.as<uint32_t>();
    // ids must not alias the sort's value buffer when it is read: write the
    // final ids into whichever value buffer is not svals
    uint32_t *idsOut = alt ? vA : vB;
    GDSJ_HIP(hipMemsetAsync(nonfinite, 0, 4, s));
    k_soa<T><<<ceil_div(n, 256), 256, 0, s>>>(data, n, d, dperm, svals, G.xs.as<T>(), G.ld, idsOut, nonfinite);
    if (idsOut != G.ids.as<uint32_t>()) std::swap(G.ids, G.valB);
    // ---- runs
    GDSJ_TRY(G.runOff.ensure(((size_t)nc + 1) * 4 + 64));
    GDSJ_TRY(G.cellAux.ensure(((size_t)nc + 1) * 4 * 4 + 256));
    uint32_t *runCount = G.cellAux.as<uint32_t>();
    uint32_t *cellCand = runCount + (nc + 1);
    uint32_t *tileCnt = cellCand + (nc + 1);
    uint32_t *tileOff = tileCnt + (nc + 1);
    unsigned long long *candTotal = (unsigned long long *)(G.scalars.as<uint64_t>() + 1);
    k_runs<false><<<ceil_div(nc, 64), 64, 0, s>>>(G.cellKey.as<uint64_t>(), G.cellStart.as<uint32_t>(), nc, a,
                                                  nullptr, runCount, nullptr, nullptr, nullptr);
    GDSJ_TRY(exclusive_scan_u32_u32(runCount, G.runOff.as<uint32_t>(), nc, true, G.tmp.p, s));
    uint32_t nr = 0, hbad = 0;
    GDSJ_HIP(hipMemcpyAsync(&nr, G.runOff.as<uint32_t>() + nc, 4, hipMemcpyDeviceToHost, s));
    GDSJ_HIP(hipMemcpyAsync(&hbad, nonfinite, 4, hipMemcpyDeviceToHost, s));
    GDSJ_HIP(hipStreamSynchronize(s));
    if (hbad) {
        set_error("non-finite coordinate in the data");
        return GDSJ_EINVAL;
    }
    G.nruns = nr;
    GDSJ_TRY(G.runs.ensure((size_t)nr * 8 + 64));
    GDSJ_HIP(hipMemsetAsync(candTotal, 0, 8, s));
    k_runs<true><<<ceil_div(nc, 64), 64, 0, s>>>(G.cellKey.as<uint64_t>(), G.cellStart.as<uint32_t>(), nc, a,
                                                 G.runOff.as<uint32_t>(), nullptr, G.runs.as<uint2>(), cellCand,
                                                 candTotal);
    // ---- tiles
    k_tiles_count<<<ceil_div(nc, 256), 256, 0, s>>>(G.cellStart.as<uint32_t>(), nc, G.tileq, tileCnt);
    GDSJ_TRY(exclusive_scan_u32_u32(tileCnt, tileOff, nc, true, G.tmp.p, s));
    uint32_t nt = 0;
    GDSJ_HIP(hipMemcpyAsync(&nt, tileOff + nc, 4, hipMemcpyDeviceToHost, s));
    GDSJ_HIP(hipMemcpyAsync(&G.candidates, candTotal, 8, hipMemcpyDeviceToHost, s));
    GDSJ_HIP(hipStreamSynchronize(s));
    G.ntiles = nt;
    GDSJ_TRY(G.tiles.ensure((size_t)nt * 16 * 2 + 64));
    uint4 *tilesA = G.tiles.as<uint4>();
    if (order_tiles && nt > 1) {
        GDSJ_TRY(G.tileKey.ensure((size_t)nt * 8 * 2 + 64));
        GDSJ_TRY(G.tileVal.ensure((size_t)nt * 4 * 2 + 64));
        uint64_t *tk = G.tileKey.as<uint64_t>();
        uint32_t *tv = G.tileVal.as<uint32_t>();
        uint4 *tilesB = tilesA + nt;
        k_tiles_fill<<<ceil_div(nc, 256), 256, 0, s>>>(G.cellStart.as<uint32_t>(), cellCand, nc, G.tileq,
                                                        tileOff, tilesB, tk, tv, 1);
        GDSJ_TRY(G.tmp.ensure(radix_tmp_bytes(nt) + 1024));
        int alt2 = 0;
        GDSJ_TRY(radix_sort_pairs(tk, tv, tk + nt, tv + nt, nt, 40, G.tmp.p, s, &alt2));
        k_tiles_permute<<<ceil_div(nt, 256), 256, 0, s>>>(tilesB, alt2 ? tv + nt : tv, nt, tilesA);
    } else {
        k_tiles_fill<<<ceil_div(nc, 256), 256, 0, s>>>(G.cellStart.as<uint32_t>(), cellCand, nc, G.tileq,
                                                        tileOff, tilesA, nullptr, nullptr, 0);
    }
    GDSJ_HIP(hipGetLastError());
    return GDSJ_OK;
}

int grid_build(Grid &G, const void *d_data, uint32_t n, uint32_t d, int dtype, double eps, uint32_t k,
               int reorder, uint32_t tileq, int order_tiles, hipStream_t s) {
    if (d == 0 || d > MAXD || k == 0 || k > MAXK || k > d) {
        set_error("invalid dims=%u k=%u (need 1 <= k <= min(dims,%d), dims <= %d)", d, k, MAXK, MAXD);
        return GDSJ_EINVAL;
    }
    if (!(eps > 0) || !std::isfinite(eps)) {
        set_error("epsilon must be positive and finite");
        return GDSJ_EINVAL;
    }
    if (tileq == 0 || tileq > 64) tileq = 64;
    G.n = n;
    G.d = d;
    G.k = k;
    G.dtype = dtype;
    G.eps = eps;
    G.tileq = tileq;
    G.candidates = 0;
    if (dtype == GDSJ_DTYPE_F32) return grid_build_t<float>(G, (const float *)d_data, s, reorder, order_tiles);
    if (dtype == GDSJ_DTYPE_F64) return grid_build_t<double>(G, (const double *)d_data, s, reorder, order_tiles);
    set_error("unknown dtype %d", dtype);
    return GDSJ_EINVAL;
}

}  // namespace gdsj
