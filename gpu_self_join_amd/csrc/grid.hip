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
// rows 0, 100, 200, ... of the points, row-major: the reference's variance
// sample (main.cu:997-1009)
template <typename T>
__global__ void k_sample_gather(const T *__restrict__ data, uint32_t n, uint32_t d, uint32_t m,
                                T *__restrict__ out) {
    uint64_t idx = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (idx >= (uint64_t)m * d) return;
    uint32_t j = (uint32_t)(idx / d), l = (uint32_t)(idx % d);
    out[idx] = data[(uint64_t)j * 100u * d + l];
}

// The reference's per-dim mean and variance of the sample (main.cu:997-1022):
// sequential sums in DTYPE, unfused, in sample order, each divided by
// denom.  One chain per dim; the dims' chains are independent, so the
// inner loop over dims may run in vector lanes without changing any sum.
// Host code of a translation unit built with -ffp-contract=off.
template <typename T>
static void sample_variance(const T *x, uint32_t m, uint32_t d, T denom, T *var) {
    std::vector<T> sum(d, (T)0), avg(d);
    for (uint32_t j = 0; j < m; j++) {
        const T *row = x + (size_t)j * d;
        for (uint32_t l = 0; l < d; l++) sum[l] = sum[l] + row[l];
    }
    for (uint32_t l = 0; l < d; l++) {
        avg[l] = sum[l] / denom;
        sum[l] = (T)0;
    }
    for (uint32_t j = 0; j < m; j++) {
        const T *row = x + (size_t)j * d;
        for (uint32_t l = 0; l < d; l++) {
            const T t = row[l] - avg[l];
            const T q = t * t;
            sum[l] = sum[l] + q;
        }
    }
    for (uint32_t l = 0; l < d; l++) var[l] = sum[l] / denom;
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
// Order-preserving integer key of a DTYPE value (-0 and +0 equal)
__device__ __forceinline__ uint64_t value_order_key(float v) {
    const uint32_t u = __float_as_uint(v == 0.0f ? 0.0f : v);
    return (u & 0x80000000u) ? (uint64_t)(~u) : (uint64_t)(u | 0x80000000u);
}
__device__ __forceinline__ uint64_t value_order_key(double v) {
    const uint64_t u = (uint64_t)__double_as_longlong(v == 0.0 ? 0.0 : v);
    return (u >> 63) ? ~u : (u | (1ull << 63));
}

// SORT: keys = the points' values at column col, vals = point ids
template <typename T>
__global__ void k_value_keys(const T *__restrict__ data, uint32_t n, uint32_t d, uint32_t col,
                             uint64_t *__restrict__ keys, uint32_t *__restrict__ vals) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    keys[i] = value_order_key(data[(uint64_t)i * d + col]);
    vals[i] = i;
}

// cell linear ids of the points order[0 .. n) (order nullable: the
// identity); vals[i] = the point of slot i (vals may alias order)
template <typename T>
__global__ void k_cell_keys(const T *__restrict__ data, uint32_t n, uint32_t d, KeyArgs a, T eps,
                            uint64_t *__restrict__ keys, uint32_t *vals,
                            uint32_t *__restrict__ outside, const uint32_t *order) {
    uint32_t slot = blockIdx.x * blockDim.x + threadIdx.x;
    if (slot >= n) return;
    const uint32_t i = order ? order[slot] : slot;
    uint64_t lin = 0;
    bool bad = false;
    for (uint32_t j = 0; j < a.k; j++) {
        T x = data[(uint64_t)i * d + a.pk[j]];
        T q = (x - (T)a.minArr[j]) / eps;  // IEEE quotient, truncated below
        uint32_t c = (uint32_t)q;
        bad |= !(q < (T)a.nCells[j]);      // only when eps is below the data's precision
        lin += (uint64_t)c * a.mult[j];
    }
    keys[slot] = lin;
    vals[slot] = i;
    if (bad) atomicOr(outside, 1u);
}

// ------------------------------------------------ principal-axes index --
// An index on the principal axes of the data (Grid::indexWant) instead of
// the reference's k dims of largest variance: the candidate set changes, the
// pair test does not.  The grid is built on y = R x (R: k orthonormal rows,
// the top principal directions of the REORDER sample) with cells wide
// enough that every pair the reference's arithmetic accepts lies in adjacent
// cells, and it is used only when no point sits within rounding of a face of
// the reference's own grid, so that the reference's candidate set holds
// every accepted pair as well: both tables are then the set of pairs the
// reference's arithmetic accepts.

// scatter matrix of the m sample rows about mean (m times their covariance;
// rows row-major, d columns); one block per (i, j >= i), written to both halves
template <typename T>
__global__ __launch_bounds__(256) void k_cov(const T *__restrict__ samp, uint32_t m, uint32_t d,
                                             const double *__restrict__ mean, double *__restrict__ cov) {
    const uint32_t i = blockIdx.y, j = blockIdx.x;
    if (j < i) return;  // uniform per block
    double acc = 0.0;
    for (uint32_t p = threadIdx.x; p < m; p += blockDim.x)
        acc += ((double)samp[(uint64_t)p * d + i] - mean[i]) * ((double)samp[(uint64_t)p * d + j] - mean[j]);
    __shared__ double red[256];
    red[threadIdx.x] = acc;
    __syncthreads();
    for (uint32_t st = 128; st > 0; st >>= 1) {
        if (threadIdx.x < st) red[threadIdx.x] += red[threadIdx.x + st];
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        cov[(uint64_t)i * d + j] = red[0];
        cov[(uint64_t)j * d + i] = red[0];
    }
}

// Points whose quotient q = (x - min) / eps, computed as k_cell_keys computes
// it, lies within tau of an integer in one of the reference's indexed dims:
// list entry (dim << 56 | below << 55 | floor(q)), below = q less than tau
// under the next integer, else q at most tau above floor(q).  Two points
// whose quotients differ by at most 1 + tau in a dim fall two cells apart
// there only as such a pair: q_a < m + 1 and q_b >= m + 2 give q_a >= m + 1
// - tau (a below, floor m) and q_b < m + 2 + tau (b above, floor m + 2).
template <typename T>
__global__ void k_face_risk(const T *__restrict__ data, uint32_t n, uint32_t d, KeyArgs a, T eps, double tau,
                            uint32_t cap, uint32_t *__restrict__ count, uint64_t *__restrict__ list) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    for (uint32_t j = 0; j < a.k; j++) {
        const T x = data[(uint64_t)i * d + a.pk[j]];
        const T q = (x - (T)a.minArr[j]) / eps;
        const double qd = (double)q, fl = floor(qd), f = qd - fl;
        const bool below = !(f < 1.0 - tau), above = f < tau;
        if (!below && !above) continue;
        const uint32_t slot = atomicAdd(count, 1u);
        if (slot < cap)
            list[slot] = ((uint64_t)j << 56) | ((uint64_t)(below ? 1u : 0u) << 55) | (uint64_t)fmax(fl, 0.0);
    }
}

// y = R x for every point (R: k x d doubles, summed in double by FMAs,
// stored in DTYPE), and the largest |x|^2 (double bits, non-negative, so
// their integer order is their value order)
template <typename T>
__global__ __launch_bounds__(256) void k_project(const T *__restrict__ data, uint32_t n, uint32_t d, uint32_t k,
                                                 const double *__restrict__ R, T *__restrict__ y,
                                                 unsigned long long *__restrict__ norm2max) {
    // R (k x d), then the block's 256 rows, PJ dims at a time: the rows are
    // read as runs of PJ contiguous elements instead of one element per
    // lane 8 d bytes apart
    extern __shared__ double sR[];
    constexpr uint32_t PJ = 8;
    double *tile = sR + (((size_t)k * d + 1) & ~(size_t)1);  // [256][PJ + 1]
    for (uint32_t t = threadIdx.x; t < k * d; t += blockDim.x) sR[t] = R[t];
    const uint32_t r0 = blockIdx.x * 256u;
    const uint32_t nr = n - r0 < 256u ? n - r0 : 256u;
    const uint32_t i = r0 + threadIdx.x;
    double nn = 0.0;
    double acc[MAXK];
#pragma unroll
    for (int c = 0; c < MAXK; c++) acc[c] = 0.0;
    for (uint32_t j0 = 0; j0 < d; j0 += PJ) {
        const uint32_t jc = d - j0 < PJ ? d - j0 : PJ;
        __syncthreads();  // R is in place; the previous chunk is consumed
        for (uint32_t e = threadIdx.x; e < 256u * PJ; e += blockDim.x) {
            const uint32_t r = e / PJ, jj = e % PJ;
            tile[r * (PJ + 1) + jj] = (r < nr && jj < jc) ? (double)data[(uint64_t)(r0 + r) * d + j0 + jj] : 0.0;
        }
        __syncthreads();
        if (threadIdx.x < nr) {
            for (uint32_t jj = 0; jj < jc; jj++) {
                const double x = tile[threadIdx.x * (PJ + 1) + jj];
                nn = fma(x, x, nn);
#pragma unroll
                for (int c = 0; c < MAXK; c++)
                    if ((uint32_t)c < k) acc[c] = fma(sR[(uint32_t)c * d + j0 + jj], x, acc[c]);
            }
        }
    }
    if (i < n) {
#pragma unroll
        for (int c = 0; c < MAXK; c++)
            if ((uint32_t)c < k) y[(uint64_t)i * k + c] = (T)acc[c];
    }
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) nn = fmax(nn, __shfl_xor(nn, off, GDSJ_WAVE));
    if (lane_id() == 0) atomicMax(norm2max, (unsigned long long)__double_as_longlong(nn));
}

// SoA copy of the index coordinates in sorted order (NaN past n)
template <typename T>
__global__ void k_gather_index(const T *__restrict__ y, uint32_t n, uint32_t k, const uint32_t *__restrict__ order,
                               uint64_t ld, T *__restrict__ ixs) {
    const uint64_t p = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= ld) return;
    for (uint32_t c = 0; c < k; c++)
        ixs[(uint64_t)c * ld + p] = p < n ? y[(uint64_t)order[p] * k + c] : (T)__builtin_nan("");
}

// Eigen-decomposition of the symmetric k x k matrix A (row-major) by cyclic
// Jacobi rotations: eigenvalues on return in A's diagonal, eigenvectors in
// the columns of V.
static void jacobi_sym(std::vector<double> &A, std::vector<double> &V, uint32_t k) {
    V.assign((size_t)k * k, 0.0);
    for (uint32_t i = 0; i < k; i++) V[(size_t)i * k + i] = 1.0;
    for (int sweep = 0; sweep < 30; sweep++) {
        // converged when the off-diagonal part is at rounding level next to
        // the diagonal (relative 1e-14 in norm)
        double off = 0.0, dg = 0.0;
        for (uint32_t i = 0; i < k; i++) {
            dg += A[(size_t)i * k + i] * A[(size_t)i * k + i];
            for (uint32_t j = i + 1; j < k; j++) off += A[(size_t)i * k + j] * A[(size_t)i * k + j];
        }
        if (off <= 1e-28 * dg || off < 1e-300) break;
        for (uint32_t p = 0; p < k; p++)
            for (uint32_t q = p + 1; q < k; q++) {
                const double apq = A[(size_t)p * k + q];
                if (apq == 0.0) continue;
                const double th = (A[(size_t)q * k + q] - A[(size_t)p * k + p]) / (2.0 * apq);
                const double t = (th >= 0 ? 1.0 : -1.0) / (fabs(th) + sqrt(th * th + 1.0));
                const double c = 1.0 / sqrt(t * t + 1.0), sn = t * c;
                for (uint32_t r = 0; r < k; r++) {  // A <- A J
                    const double arp = A[(size_t)r * k + p], arq = A[(size_t)r * k + q];
                    A[(size_t)r * k + p] = c * arp - sn * arq;
                    A[(size_t)r * k + q] = sn * arp + c * arq;
                }
                for (uint32_t r = 0; r < k; r++) {  // A <- J^T A
                    const double apr = A[(size_t)p * k + r], aqr = A[(size_t)q * k + r];
                    A[(size_t)p * k + r] = c * apr - sn * aqr;
                    A[(size_t)q * k + r] = sn * apr + c * aqr;
                }
                for (uint32_t r = 0; r < k; r++) {  // V <- V J
                    const double vrp = V[(size_t)r * k + p], vrq = V[(size_t)r * k + q];
                    V[(size_t)r * k + p] = c * vrp - sn * vrq;
                    V[(size_t)r * k + q] = sn * vrp + c * vrq;
                }
            }
    }
}

// The top-k principal directions of the covariance C (d x d): subspace
// iteration from the unit vectors of the k dims of largest variance
// (perm[0..k)), modified Gram-Schmidt after every product, then the
// Rayleigh-Ritz rotation that orders them by variance.  R: all k of them
// (k x d, rows orthonormal up to rounding, by decreasing variance); ritz:
// their variances (Ritz values, decreasing); *total: trace of C.
static void pca_basis(const std::vector<double> &C, uint32_t d, uint32_t k, const uint32_t *perm,
                      std::vector<double> &R, std::vector<double> &ritz, double *total) {
    std::vector<double> Q((size_t)d * k, 0.0), Z((size_t)d * k);
    for (uint32_t c = 0; c < k; c++) Q[(size_t)perm[c] * k + c] = 1.0;
    auto orth = [&](std::vector<double> &M) {
        for (uint32_t c = 0; c < k; c++) {
            for (uint32_t b = 0; b < c; b++) {
                double dot = 0.0;
                for (uint32_t j = 0; j < d; j++) dot += M[(size_t)j * k + c] * M[(size_t)j * k + b];
                for (uint32_t j = 0; j < d; j++) M[(size_t)j * k + c] -= dot * M[(size_t)j * k + b];
            }
            double nn = 0.0;
            for (uint32_t j = 0; j < d; j++) nn += M[(size_t)j * k + c] * M[(size_t)j * k + c];
            const double inv = nn > 0.0 ? 1.0 / sqrt(nn) : 0.0;
            for (uint32_t j = 0; j < d; j++) M[(size_t)j * k + c] *= inv;
            if (nn == 0.0) M[(size_t)perm[c] * k + c] = 1.0;  // degenerate: keep the unit vector
        }
    };
    // Out = C In, each entry summed over j in ascending order; the k
    // columns advance together (contiguous rows of In)
    auto mulC = [&](const std::vector<double> &In, std::vector<double> &Out) {
        for (uint32_t i = 0; i < d; i++) {
            double *o = &Out[(size_t)i * k];
            for (uint32_t c = 0; c < k; c++) o[c] = 0.0;
            for (uint32_t j = 0; j < d; j++) {
                const double cij = C[(size_t)i * d + j];
                const double *in = &In[(size_t)j * k];
                for (uint32_t c = 0; c < k; c++) o[c] += cij * in[c];
            }
        }
    };
    for (int it = 0; it < 12; it++) {
        mulC(Q, Z);
        orth(Z);
        Q.swap(Z);
    }
    // Rayleigh-Ritz: A = Q^T C Q, A = V diag V^T, Q <- Q V (by variance)
    mulC(Q, Z);
    std::vector<double> A((size_t)k * k), V;
    for (uint32_t a0 = 0; a0 < k; a0++)
        for (uint32_t b0 = 0; b0 < k; b0++) {
            double acc = 0.0;
            for (uint32_t j = 0; j < d; j++) acc += Q[(size_t)j * k + a0] * Z[(size_t)j * k + b0];
            A[(size_t)a0 * k + b0] = 0.5 * acc;
        }
    for (uint32_t a0 = 0; a0 < k; a0++)  // symmetrise
        for (uint32_t b0 = a0 + 1; b0 < k; b0++) {
            const double m = A[(size_t)a0 * k + b0] + A[(size_t)b0 * k + a0];
            A[(size_t)a0 * k + b0] = A[(size_t)b0 * k + a0] = m;
        }
    for (uint32_t a0 = 0; a0 < k; a0++) A[(size_t)a0 * k + a0] *= 2.0;
    jacobi_sym(A, V, k);
    std::vector<uint32_t> order(k);
    for (uint32_t c = 0; c < k; c++) order[c] = c;
    std::sort(order.begin(), order.end(),
              [&](uint32_t x, uint32_t y) { return A[(size_t)x * k + x] > A[(size_t)y * k + y]; });
    ritz.resize(k);
    for (uint32_t c = 0; c < k; c++) ritz[c] = A[(size_t)order[c] * k + order[c]];
    R.assign((size_t)k * d, 0.0);
    for (uint32_t c = 0; c < k; c++)
        for (uint32_t j = 0; j < d; j++) {
            double acc = 0.0;
            for (uint32_t b = 0; b < k; b++) acc += Q[(size_t)j * k + b] * V[(size_t)b * k + order[c]];
            R[(size_t)c * d + j] = acc;
        }
    *total = 0.0;
    for (uint32_t j = 0; j < d; j++) *total += C[(size_t)j * d + j];
}

// SORT: the cell of every sorted position (inclusive count of cell heads - 1)
__global__ void k_pos_cell(const uint32_t *__restrict__ flag, const uint32_t *__restrict__ cidx, uint32_t n,
                           uint32_t *__restrict__ posCell) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    posCell[i] = cidx[i] + flag[i] - 1u;
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
    // positions n .. ld-1 (at least one) hold NaN: the join's lanes past the
    // end of a candidate window read them, and NaN fails every compare
    if ((uint64_t)s < ld - n)
        for (uint32_t l = 0; l < d; l++) xs[(uint64_t)l * ld + n + s] = (T)__builtin_nan("");
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

// The same as k_soa for a tile of TP sorted positions per block, transposed
// through LDS: the rows are read with the threads along each row (coalesced;
// one thread per point reads 64 rows per load instruction) and the SoA
// columns are written as runs of TP.  tile = [d][TP + 1] (dynamic LDS).
template <typename T, uint32_t TP>
__global__ __launch_bounds__(256) void k_soa_tile(const T *__restrict__ data, uint32_t n, uint32_t d,
                                                  const uint32_t *__restrict__ perm,
                                                  const uint32_t *__restrict__ order, T *__restrict__ xs,
                                                  uint64_t ld, uint32_t *__restrict__ ids,
                                                  uint32_t *__restrict__ nonfinite, T *__restrict__ xa,
                                                  uint32_t as) {
    extern __shared__ __align__(16) unsigned char soa_lds[];
    T *tile = reinterpret_cast<T *>(soa_lds);
    __shared__ uint32_t s_id[TP];
    __shared__ uint32_t s_perm[MAXD];
    __shared__ uint32_t s_bad;
    if (blockIdx.x == 0) {  // NaN padding at positions n .. ld-1
        for (uint64_t t = threadIdx.x; t < (ld - n) * d; t += blockDim.x)
            xs[(t / (ld - n)) * ld + n + t % (ld - n)] = (T)__builtin_nan("");
        if (xa)
            for (uint64_t t = threadIdx.x; t < (ld - n) * as; t += blockDim.x)
                xa[(uint64_t)n * as + t] = t % as < d ? (T)__builtin_nan("") : (T)0;
    }
    const uint32_t s0 = blockIdx.x * TP;
    const uint32_t np = n - s0 < TP ? n - s0 : TP;
    if (threadIdx.x == 0) s_bad = 0;
    for (uint32_t p = threadIdx.x; p < np; p += blockDim.x) {
        const uint32_t id = order[s0 + p];
        s_id[p] = id;
        ids[s0 + p] = id;
    }
    for (uint32_t l = threadIdx.x; l < d; l += blockDim.x) s_perm[l] = perm[l];
    __syncthreads();
    bool bad = false;
    for (uint32_t t = threadIdx.x; t < np * d; t += blockDim.x) {
        const uint32_t p = t / d, l = t - p * d;
        const T v = data[(uint64_t)s_id[p] * d + s_perm[l]];
        bad |= !isfinite(v);
        tile[l * (TP + 1) + p] = v;
    }
    if (bad) atomicOr(&s_bad, 1u);
    __syncthreads();
    for (uint32_t t = threadIdx.x; t < d * TP; t += blockDim.x) {
        const uint32_t l = t / TP, p = t % TP;
        if (p < np) xs[(uint64_t)l * ld + s0 + p] = tile[l * (TP + 1) + p];
    }
    if (xa)
        for (uint32_t t = threadIdx.x; t < np * as; t += blockDim.x) {
            const uint32_t p = t / as, l = t - p * as;
            xa[(uint64_t)(s0 + p) * as + l] = l < d ? tile[l * (TP + 1) + p] : (T)0;
        }
    if (threadIdx.x == 0 && s_bad) atomicOr(nonfinite, 1u);
}

// ------------------------------------------------- cell hash + coords --
// Open-addressing table row -> first cell of the row, the replacement for
// the per-candidate-cell binary search of the reference (kernel.cu:388-399).
// Rows: cells equal in dims 1..k-1 (linear id minus c_0); each row's first
// cell c is inserted under row + 1 with value c | (cells in the row) << 32
// (a row's cells are contiguous in cell order).
__global__ void k_hash_build(const uint64_t *__restrict__ cellKey, uint32_t nc, uint32_t ncells0, uint32_t bits,
                             unsigned long long *__restrict__ hkeys) {
    uint32_t c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= nc) return;
    const uint64_t row = cellKey[c] - cellKey[c] % ncells0;
    if (c > 0 && cellKey[c - 1] - cellKey[c - 1] % ncells0 == row) return;  // not the row's first cell
    uint32_t len = 1;
    while (c + len < nc && cellKey[c + len] < row + ncells0) len++;
    const uint32_t mask = (1u << bits) - 1u;
    const unsigned long long key = (unsigned long long)row + 1ull;
    uint32_t h = cell_hash(row, bits);
    for (uint32_t probe = 0; probe <= mask; probe++) {
        unsigned long long prev = atomicCAS(&hkeys[2 * (uint64_t)h], 0ull, key);
        if (prev == 0ull || prev == key) {
            hkeys[2 * (uint64_t)h + 1] = (unsigned long long)c | ((unsigned long long)len << 32);
            return;
        }
        h = (h + 1) & mask;
    }
}

__global__ void k_dense_build(const uint64_t *__restrict__ cellKey, uint32_t nc, uint32_t *__restrict__ table) {
    uint32_t c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= nc) return;
    table[cellKey[c]] = c;
}

// coordinates of each non-empty cell (decomposition of its linear id)
__global__ void k_cell_coords(const uint64_t *__restrict__ cellKey, uint32_t nc, GridArgs a,
                              uint32_t *__restrict__ coord) {
    uint32_t c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= nc) return;
    uint64_t key = cellKey[c];
    for (uint32_t j = 0; j < a.k; j++) {
        coord[(uint64_t)j * nc + c] = (uint32_t)(key % a.nCells[j]);
        key /= a.nCells[j];
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

__global__ void k_tiles_fill(const uint32_t *__restrict__ cellStart, uint32_t nc, uint32_t tileq,
                             const uint32_t *__restrict__ tileOff, uint4 *__restrict__ tiles,
                             uint64_t *__restrict__ tileKey, uint32_t *__restrict__ tileVal, int order) {
    uint32_t c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= nc) return;
    uint32_t s0 = cellStart[c], s1 = cellStart[c + 1];
    uint32_t t = tileOff[c];
    for (uint32_t q = s0; q < s1; q += tileq, t++) {
        uint32_t q1 = q + tileq < s1 ? q + tileq : s1;
        tiles[t] = make_uint4(c, q, q1, 0u);
        if (order) {
            // QUERYREORDER (GPU.cu:1491-1525): queries of the most occupied
            // cells first; stable, so ties keep cell order.  order 2: by
            // occupancy class (power of two) only, so each class is one
            // spatially ordered run of tiles (neighboring tiles share
            // candidates in L2)
            const uint32_t occ = s1 - s0;
            const uint32_t cls = 31u - (uint32_t)__builtin_clz(occ | 1u);  // floor(log2(occ))
            tileKey[t] = order == 2 ? (uint64_t)(31u - cls) : (uint64_t)(~occ);
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
    DevBuf *b[] = {&xs, &ids, &cellKey, &cellStart, &cellCoord, &hkeys, &hvals, &tiles, &keyA, &keyB,
                   &valB, &tmp, &aux, &flag, &tileKey, &tileVal, &cellAux, &scalars, &runs, &nruns, &cellCand, &cellHist,
                   &comboTab, &posCell, &xa, &ixs, &ybuf, &pcaBuf};
    for (DevBuf *p : b) p->release();
    hostSample.release();
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
    T eps = (T)G.cellW;
    for (uint32_t j = 0; j < G.k; j++) {
        T lo = mn[j] - eps;
        T hi = mx[j] + eps;
        T q = (hi - lo) / eps;
        T c = sizeof(T) == 4 ? (T)ceilf((float)q) : (T)ceil((double)q);
        if (!(c >= 1) || c > (T)4294967295.0) {
            set_error("grid extent out of range in dim %u", j);
            return GDSJ_ERANGE;
        }
        G.ga.nCells[j] = (uint32_t)c;
        G.minArr[j] = (double)lo;
        G.ga.mult[j] = total;
        uint64_t nt = total * (uint64_t)G.ga.nCells[j];
        if (nt / G.ga.nCells[j] != total) {
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
    const uint32_t n = G.n, d = G.d;
    uint32_t k = G.k;  // the principal-axes grid may index fewer dims (kp)
    G.ga.k = k;
    // ---- REORDER
    for (uint32_t l = 0; l < d; l++) G.perm[l] = l;
    if (reorder && n > 0) {
        uint32_t m = (n + 99) / 100;
        const size_t sb = (size_t)m * d * sizeof(T);
        GDSJ_TRY(G.aux.ensure(sb + 256));
        GDSJ_TRY(G.hostSample.ensure(sb + 256));
        T *samp = G.aux.as<T>();
        uint64_t tot = (uint64_t)m * d;
        k_sample_gather<T><<<(uint32_t)((tot + 255) / 256), 256, 0, s>>>(data, n, d, m, samp);
        GDSJ_HIP(hipMemcpyAsync(G.hostSample.p, samp, sb, hipMemcpyDeviceToHost, s));
        GDSJ_HIP(hipStreamSynchronize(s));
        const T inv_sample = (T)(1.0 / (100 * 1.0));
        const T denom = (T)n * inv_sample;
        std::vector<T> hv(d);
        sample_variance<T>((const T *)G.hostSample.p, m, d, denom, hv.data());
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
    // ---- principal-axes index: the basis from the REORDER sample rows
    bool pca = false;
    std::vector<double> R;
    double delta = 0.0;
    G.cellW = G.eps;
    G.indexUsed = 1;
    G.riskyPoints = 0;
    const uint32_t kp = G.indexDims && G.indexDims < k ? G.indexDims : k;  // principal axes indexed
    G.gramAxes = 0;
    G.gramNorm = 0.0;
    G.gramP.clear();
    if ((G.indexWant != 1 || G.gramAxesWant != 1) && n > 0 && d > k && d <= 256 && !G.sortCells) {
        const uint32_t m = (n + 99) / 100;
        const size_t sb = (size_t)m * d * sizeof(T);
        if (!reorder) {
            GDSJ_TRY(G.aux.ensure(sb + 256));
            GDSJ_TRY(G.hostSample.ensure(sb + 256));
            const uint64_t tot = (uint64_t)m * d;
            k_sample_gather<T><<<(uint32_t)((tot + 255) / 256), 256, 0, s>>>(data, n, d, m, G.aux.as<T>());
            GDSJ_HIP(hipMemcpyAsync(G.hostSample.p, G.aux.p, sb, hipMemcpyDeviceToHost, s));
            GDSJ_HIP(hipStreamSynchronize(s));
        }
        std::vector<double> mean(d, 0.0);
        const T *hs = (const T *)G.hostSample.p;
        for (uint32_t p = 0; p < m; p++)
            for (uint32_t j = 0; j < d; j++) mean[j] += (double)hs[(size_t)p * d + j];
        for (uint32_t j = 0; j < d; j++) mean[j] /= (double)m;
        GDSJ_TRY(G.pcaBuf.ensure(((size_t)d * d + d + (size_t)MAXK * d) * 8 + 256));
        double *dmean = G.pcaBuf.as<double>(), *dcov = dmean + d;
        GDSJ_HIP(hipMemcpyAsync(dmean, mean.data(), (size_t)d * 8, hipMemcpyHostToDevice, s));
        k_cov<T><<<dim3(d, d), 256, 0, s>>>(G.aux.as<T>(), m, d, dmean, dcov);
        std::vector<double> C((size_t)d * d);
        GDSJ_HIP(hipMemcpyAsync(C.data(), dcov, (size_t)d * d * 8, hipMemcpyDeviceToHost, s));
        GDSJ_HIP(hipStreamSynchronize(s));
        double capP = 0, total = 0, capR = 0;
        // the index's basis does not depend on the filter, so that every
        // filter's rows come in the same order
        std::vector<double> Rk, ritz;
        pca_basis(C, d, k, G.perm, Rk, ritz, &total);
        for (uint32_t c = 0; c < k; c++) capP += ritz[c];
        R.assign(Rk.begin(), Rk.begin() + (size_t)kp * d);
        for (uint32_t c = 0; c < k; c++) capR += C[(size_t)G.perm[c] * d + G.perm[c]];  // the reference's dims
        // auto: when k principal axes leave at most a twentieth of the
        // variance that the reference's k dims leave out (data near a
        // rotated subspace of at most k dims)
        pca = G.indexWant == 2 || (G.indexWant == 0 && (total - capP) < 0.05 * (total - capR));
        // the Gram filter's operands: gm axes where the distance they leave
        // out (twice the variance off them, for a pair) is under a hundredth
        // of eps^2, so that they reject nearly as many pairs as all d dims
        const uint32_t gm = G.gramAxesWant == 0 ? 16u : (uint32_t)G.gramAxesWant;
        if (G.gramAxesWant != 1 && gm < d && d >= 32 && d <= 128) {
            std::vector<double> P, ritzM;
            double capM = 0;
            pca_basis(C, d, gm, G.perm, P, ritzM, &total);
            for (uint32_t c = 0; c < gm; c++) capM += ritzM[c];
            // (C sums the products over the m sample rows)
            if (G.gramAxesWant != 0 || 2.0 * (total - capM) / (double)m < 0.01 * G.eps * G.eps) {
                // |P|_2 <= sqrt(1 + |P P^T - I|_F)
                double f = 0.0;
                for (uint32_t a0 = 0; a0 < gm; a0++)
                    for (uint32_t b0 = 0; b0 < gm; b0++) {
                        double dot = 0.0;
                        for (uint32_t j = 0; j < d; j++) dot += P[(size_t)a0 * d + j] * P[(size_t)b0 * d + j];
                        const double e = dot - (a0 == b0 ? 1.0 : 0.0);
                        f += e * e;
                    }
                G.gramNorm = sqrt(1.0 + sqrt(f)) * (1.0 + 1e-12);
                G.gramP.assign((size_t)gm * d, 0.0);
                for (uint32_t c = 0; c < gm; c++)
                    for (uint32_t l = 0; l < d; l++) G.gramP[(size_t)c * d + l] = P[(size_t)c * d + G.perm[l]];
                G.gramAxes = gm;
            }
        }
        if (pca) {
            // |R|_2 <= sqrt(1 + |R R^T - I|_F)
            double f = 0.0;
            for (uint32_t a0 = 0; a0 < kp; a0++)
                for (uint32_t b0 = 0; b0 < kp; b0++) {
                    double dot = 0.0;
                    for (uint32_t j = 0; j < d; j++) dot += R[(size_t)a0 * d + j] * R[(size_t)b0 * d + j];
                    const double e = dot - (a0 == b0 ? 1.0 : 0.0);
                    f += e * e;
                }
            delta = sqrt(1.0 + sqrt(f)) - 1.0 + 1e-12;
        }
    }
    KeyArgs a{};
    a.k = k;
    for (uint32_t j = 0; j < k; j++) a.pk[j] = G.perm[j];
    if (n == 0) {
        G.ncell = G.ntiles = 0;
        G.ga.ncell = 0;
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
        a.nCells[j] = G.ga.nCells[j];
        a.mult[j] = G.ga.mult[j];
        a.minArr[j] = G.minArr[j];
    }
    const double uT = sizeof(T) == 4 ? 0x1p-24 : 0x1p-53;
    if (pca) {
        // a point near a face of the reference's grid: the reference may
        // miss a pair of it that the principal-axes grid finds
        double qmax = 0.0;
        for (uint32_t j = 0; j < k; j++) qmax = std::max(qmax, (double)G.ga.nCells[j]);
        const double tau = ((double)d + (double)G.ilpTerms + 8.0) * uT + 4.1 * uT * (qmax + 1.0);
        const uint32_t cap = 1u << 16;
        GDSJ_TRY(G.scalars.ensure(256));
        GDSJ_TRY(G.ybuf.ensure((size_t)cap * 8 + 256));
        uint64_t *list = G.ybuf.as<uint64_t>();
        uint32_t *cnt = G.scalars.as<uint32_t>() + 2;
        GDSJ_HIP(hipMemsetAsync(cnt, 0, 4, s));
        k_face_risk<T><<<ceil_div(n, 256), 256, 0, s>>>(data, n, d, a, (T)G.eps, tau, cap, cnt, list);
        uint32_t nl = 0;
        GDSJ_HIP(hipMemcpyAsync(&nl, cnt, 4, hipMemcpyDeviceToHost, s));
        GDSJ_HIP(hipStreamSynchronize(s));
        uint64_t risky = 0;
        if (nl > cap) {
            risky = nl;  // too many points on faces to check: the reference's dims
        } else if (nl > 0) {
            std::vector<uint64_t> hl(nl);
            GDSJ_HIP(hipMemcpy(hl.data(), list, (size_t)nl * 8, hipMemcpyDeviceToHost));
            std::sort(hl.begin(), hl.end());
            const uint64_t belowBit = 1ull << 55;
            for (uint64_t e : hl) {
                if (!(e & belowBit)) continue;
                // an entry above the face two cells up in the same dim
                const uint64_t want = (e & ~belowBit) + 2u;
                if (std::binary_search(hl.begin(), hl.end(), want)) risky++;
            }
        }
        G.riskyPoints = risky;
        if (risky) pca = false;
    }
    const T *keyData = data;  // the coordinates the cells are keyed on
    uint32_t keyD = d;
    double pcaPrune = 0.0;
    if (pca) {
        k = kp;
        G.k = kp;
        G.ga.k = kp;
        double *dR = G.pcaBuf.as<double>() + (size_t)d * d + d;
        GDSJ_HIP(hipMemcpyAsync(dR, R.data(), (size_t)k * d * 8, hipMemcpyHostToDevice, s));
        GDSJ_TRY(G.ybuf.ensure((size_t)n * k * sizeof(T) + 256));
        T *y = G.ybuf.as<T>();
        unsigned long long *nmax = G.scalars.as<unsigned long long>() + 8;
        GDSJ_HIP(hipMemsetAsync(nmax, 0, 8, s));
        const size_t projLds = ((((size_t)k * d + 1) & ~(size_t)1) + 256 * 9) * 8;
        k_project<T><<<ceil_div(n, 256), 256, projLds, s>>>(data, n, d, k, dR, y, nmax);
        KeyArgs ay{};
        ay.k = k;
        for (uint32_t j = 0; j < k; j++) ay.pk[j] = j;
        k_minmax<T><<<mmBlocks, 256, 0, s>>>(y, n, k, ay, part);
        GDSJ_HIP(hipMemcpyAsync(hp.data(), part, hp.size() * sizeof(T), hipMemcpyDeviceToHost, s));
        unsigned long long nbits = 0;
        GDSJ_HIP(hipMemcpyAsync(&nbits, nmax, 8, hipMemcpyDeviceToHost, s));
        GDSJ_HIP(hipStreamSynchronize(s));
        T ymn[MAXK], ymx[MAXK];
        double yabs = 0.0;
        for (uint32_t j = 0; j < k; j++) {
            ymn[j] = hp[j * 2];
            ymx[j] = hp[j * 2 + 1];
            for (uint32_t b = 1; b < mmBlocks; b++) {
                T v0 = hp[((size_t)b * k + j) * 2], v1 = hp[((size_t)b * k + j) * 2 + 1];
                ymn[j] = v0 < ymn[j] ? v0 : ymn[j];
                ymx[j] = v1 > ymx[j] ? v1 : ymx[j];
            }
            yabs = std::max(yabs, std::max(fabs((double)ymn[j]), fabs((double)ymx[j])));
        }
        double n2;
        memcpy(&n2, &nbits, 8);
        const double xnorm = sqrt(n2) * (1.0 + 1e-12);
        // projection error per coordinate: the double FMA sum (gamma_d of
        // sum |R_cj x_j| <= |R_c| |x|) and the rounding to DTYPE
        auto gam = [](double m, double u) { return m * u / (1.0 - m * u); };
        const double err = gam((double)d + 1.0, 0x1p-53) * (1.0 + delta) * xnorm + uT * yabs * (1.0 + 1e-9);
        // |x_a - x_b| of a pair the reference accepts (acc <= thr, acc within
        // gamma of the sum of its rounded squared differences)
        const double thr = sizeof(T) == 4 ? (double)gdsj_eps_threshold_f32((float)G.eps)
                                          : gdsj_eps_threshold_f64(G.eps);
        const double nt = (double)d + (double)G.ilpTerms + 2.0;
        const double epsP = sqrt(thr / ((1.0 - gam(nt, uT)) * (1.0 - uT) * (1.0 - uT))) * (1.0 + 1e-12);
        // per coordinate |y_a - y_b| <= w0; cells of width w >= w0 plus the
        // rounding of the quotient (y - min) / w keep such pairs adjacent
        const double w0 = (1.0 + delta) * epsP + 2.0 * err;
        double span = 0.0;
        for (uint32_t j = 0; j < k; j++) span = std::max(span, (double)ymx[j] - (double)ymn[j]);
        const double qmax = span / w0 + 4.0;
        if (!(4.1 * uT * qmax < 0.01)) {
            set_error("principal-axes index: cells too fine for DTYPE");
            return GDSJ_ERANGE;
        }
        double w = w0 / (1.0 - 4.1 * uT * qmax) * (1.0 + 4.0 * uT);
        T wt = (T)w;
        if ((double)wt < w) wt = sizeof(T) == 4 ? (T)nextafterf((float)wt, INFINITY) : (T)nextafter((double)wt, INFINITY);
        G.cellW = (double)wt;
        GDSJ_TRY(host_extent<T>(G, ymn, ymx));
        a = ay;
        for (uint32_t j = 0; j < k; j++) {
            a.nCells[j] = G.ga.nCells[j];
            a.mult[j] = G.ga.mult[j];
            a.minArr[j] = G.minArr[j];
        }
        keyData = y;
        keyD = k;
        // the Euclidean |y_a - y_b| over the k axes of such a pair
        const double w2 = (1.0 + delta) * epsP + 2.0 * sqrt((double)k) * err;
        pcaPrune = w2 * w2 * (1.0 + 1e-9);
        G.indexUsed = 2;
    }
    // ---- keys + sort
    GDSJ_TRY(G.keyA.ensure((size_t)n * 8));
    GDSJ_TRY(G.keyB.ensure((size_t)n * 8));
    GDSJ_TRY(G.ids.ensure((size_t)n * 4 + 256));
    GDSJ_TRY(G.valB.ensure((size_t)n * 4 + 256));
    GDSJ_TRY(G.tmp.ensure(std::max(radix_tmp_bytes(n), scan_tmp_bytes((uint64_t)n + 1)) + 1024));
    GDSJ_TRY(G.scalars.ensure(256));
    uint64_t *kA = G.keyA.as<uint64_t>(), *kB = G.keyB.as<uint64_t>();
    uint32_t *vA = G.ids.as<uint32_t>(), *vB = G.valB.as<uint32_t>();
    uint32_t *outside = G.scalars.as<uint32_t>() + 1;
    GDSJ_HIP(hipMemsetAsync(outside, 0, 4, s));
    uint32_t *v0 = vA, *v1 = vB;  // the value buffer the cell sort starts from, and its alternate
    if (G.sortCells) {
        // SORT (GPU.cu:351-391): a stable sort by the value at the sort dim
        // first, so that the stable cell sort leaves each cell in value order
        const uint32_t col = G.perm[k == d ? 0 : k];
        k_value_keys<T><<<ceil_div(n, 256), 256, 0, s>>>(data, n, d, col, kA, vA);
        int altv = 0;
        GDSJ_TRY(radix_sort_pairs(kA, vA, kB, vB, n, (int)(8 * sizeof(T)), G.tmp.p, s, &altv));
        if (altv) {
            v0 = vB;
            v1 = vA;
        }
        k_cell_keys<T><<<ceil_div(n, 256), 256, 0, s>>>(data, n, d, a, (T)G.eps, kA, v0, outside, v0);
    } else {
        k_cell_keys<T><<<ceil_div(n, 256), 256, 0, s>>>(keyData, n, keyD, a, (T)G.cellW, kA, v0, outside,
                                                         nullptr);
    }
    int alt = 0;
    int bits = bit_length(G.totalCells - 1);
    GDSJ_TRY(radix_sort_pairs(kA, v0, kB, v1, n, bits, G.tmp.p, s, &alt));
    uint64_t *skeys = alt ? kB : kA;
    uint32_t *svals = alt ? v1 : v0;
    // ---- cells
    GDSJ_TRY(G.flag.ensure(((size_t)n + 1) * 4 * 2 + 64));
    uint32_t *flag = G.flag.as<uint32_t>();
    uint32_t *cidx = flag + n + 1;
    k_heads<<<ceil_div(n, 256), 256, 0, s>>>(skeys, n, flag);
    GDSJ_TRY(exclusive_scan_u32_u32(flag, cidx, n, true, G.tmp.p, s));
    if (G.sortCells) {
        GDSJ_TRY(G.posCell.ensure((size_t)n * 4 + 256));
        k_pos_cell<<<ceil_div(n, 256), 256, 0, s>>>(flag, cidx, n, G.posCell.as<uint32_t>());
    }
    // ---- SoA (sorted order, dims permuted), launched before the cell count sync
    G.ld = ((uint64_t)n + 1 + 63) / 64 * 64;  // n points + NaN padding
    GDSJ_TRY(G.xs.ensure((size_t)G.ld * d * sizeof(T) + 256));
    uint32_t *dperm = (uint32_t *)((char *)G.aux.p + (((size_t)mmBlocks * k * 2 * sizeof(T) + 255) & ~(size_t)255));
    GDSJ_HIP(hipMemcpyAsync(dperm, G.perm, d * 4, hipMemcpyHostToDevice, s));
    uint32_t *nonfinite = G.scalars.as<uint32_t>();
    uint32_t *idsOut = svals == vA ? vB : vA;  // the value buffer the sort did not end in
    GDSJ_HIP(hipMemsetAsync(nonfinite, 0, 4, s));
    G.xaStride = G.aosWant && d <= 16 ? xa_stride(d) : 0u;
    T *xa = nullptr;
    if (G.xaStride) {
        GDSJ_TRY(G.xa.ensure((size_t)G.ld * G.xaStride * sizeof(T) + 256));
        xa = G.xa.as<T>();
    }
    if ((size_t)d * 65 * sizeof(T) <= 48 * 1024)
        k_soa_tile<T, 64><<<ceil_div(n, 64), 256, (size_t)d * 65 * sizeof(T), s>>>(data, n, d, dperm, svals,
                                                                                     G.xs.as<T>(), G.ld, idsOut,
                                                                                     nonfinite, xa, G.xaStride);
    else if ((size_t)d * 33 * sizeof(T) <= 48 * 1024)
        k_soa_tile<T, 32><<<ceil_div(n, 32), 256, (size_t)d * 33 * sizeof(T), s>>>(data, n, d, dperm, svals,
                                                                                     G.xs.as<T>(), G.ld, idsOut,
                                                                                     nonfinite, xa, G.xaStride);
    else
        k_soa<T><<<ceil_div(n, 256), 256, 0, s>>>(data, n, d, dperm, svals, G.xs.as<T>(), G.ld, idsOut, nonfinite);
    GDSJ_HIP(hipGetLastError());
    if (pca) {
        // the index coordinates in sorted order, for the run pruning
        GDSJ_TRY(G.ixs.ensure((size_t)G.ld * k * sizeof(T) + 256));
        k_gather_index<T><<<(uint32_t)((G.ld + 255) / 256), 256, 0, s>>>(G.ybuf.as<T>(), n, k, svals, G.ld,
                                                                         G.ixs.as<T>());
    }
    if (idsOut != G.ids.as<uint32_t>()) std::swap(G.ids, G.valB);
    uint32_t hdr[3] = {0, 0, 0};
    GDSJ_HIP(hipMemcpyAsync(&hdr[0], cidx + n, 4, hipMemcpyDeviceToHost, s));
    GDSJ_HIP(hipMemcpyAsync(&hdr[1], nonfinite, 8, hipMemcpyDeviceToHost, s));
    GDSJ_HIP(hipStreamSynchronize(s));
    if (hdr[1]) {
        set_error("non-finite coordinate in the data");
        return GDSJ_EINVAL;
    }
    if (hdr[2]) {
        set_error("a cell coordinate falls outside the grid extent: epsilon is below the precision of the data");
        return GDSJ_ERANGE;
    }
    const uint32_t nc = hdr[0];
    G.ncell = nc;
    G.ga.ncell = nc;
    GDSJ_TRY(G.cellKey.ensure((size_t)nc * 8 + 64));
    GDSJ_TRY(G.cellStart.ensure(((size_t)nc + 1) * 4 + 64));
    k_cells<<<ceil_div(n, 256), 256, 0, s>>>(skeys, flag, cidx, n, G.cellKey.as<uint64_t>(),
                                              G.cellStart.as<uint32_t>());
    // ---- cell lookup table + coordinates
    GDSJ_TRY(G.cellCoord.ensure((size_t)nc * k * 4 + 64));
    if (G.totalCells <= std::max<uint64_t>(4ull * n, 1ull << 20) && G.totalCells <= (1ull << 28)) {
        // direct table: linear id -> cell index
        G.ga.hbits = 0;
        GDSJ_TRY(G.hvals.ensure((size_t)G.totalCells * 4 + 64));
        GDSJ_HIP(hipMemsetAsync(G.hvals.p, 0xFF, (size_t)G.totalCells * 4, s));
        k_dense_build<<<ceil_div(nc, 256), 256, 0, s>>>(G.cellKey.as<uint64_t>(), nc, G.hvals.as<uint32_t>());
        GDSJ_TRY(G.hkeys.ensure(64));
    } else {
        uint32_t hb = 6;
        while ((1ull << hb) < 2ull * nc) hb++;
        G.ga.hbits = hb;
        GDSJ_TRY(G.hkeys.ensure(((size_t)1 << hb) * 16));
        GDSJ_TRY(G.hvals.ensure(64));
        GDSJ_HIP(hipMemsetAsync(G.hkeys.p, 0, ((size_t)1 << hb) * 16, s));
        k_hash_build<<<ceil_div(nc, 256), 256, 0, s>>>(G.cellKey.as<uint64_t>(), nc, G.ga.nCells[0], hb,
                                                        G.hkeys.as<unsigned long long>());
    }
    k_cell_coords<<<ceil_div(nc, 256), 256, 0, s>>>(G.cellKey.as<uint64_t>(), nc, G.ga, G.cellCoord.as<uint32_t>());
    // pruning constants: a neighbor cell whose squared gap to every query of a
    // tile exceeds eps^2 (1 + 8 (D + 4) u) cannot hold a pair the DTYPE
    // predicate accepts (the accumulated sum is within (D + 3) u of the exact one)
    // (principal axes: the squared Euclidean bound over the k axes of
    // |y_a - y_b| for such a pair)
    for (uint32_t j = 0; j < k; j++) G.ga.minArr[j] = G.minArr[j];
    G.ga.eps = G.cellW;
    G.ga.ulp = sizeof(T) == 4 ? 0x1p-24 : 0x1p-53;
    G.ga.prune_lim = pca ? pcaPrune : G.eps * G.eps * (1.0 + 8.0 * (d + 4) * G.ga.ulp);
    // ---- tiles
    GDSJ_TRY(G.cellAux.ensure(((size_t)nc + 1) * 4 * 2 + 256));
    uint32_t *tileCnt = G.cellAux.as<uint32_t>();
    uint32_t *tileOff = tileCnt + (nc + 1);
    k_tiles_count<<<ceil_div(nc, 256), 256, 0, s>>>(G.cellStart.as<uint32_t>(), nc, G.tileq, tileCnt);
    GDSJ_TRY(exclusive_scan_u32_u32(tileCnt, tileOff, nc, true, G.tmp.p, s));
    uint32_t nt = 0;
    GDSJ_HIP(hipMemcpyAsync(&nt, tileOff + nc, 4, hipMemcpyDeviceToHost, s));
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
        k_tiles_fill<<<ceil_div(nc, 256), 256, 0, s>>>(G.cellStart.as<uint32_t>(), nc, G.tileq, tileOff, tilesB,
                                                        tk, tv, order_tiles == 2 ? 2 : 1);
        GDSJ_TRY(G.tmp.ensure(radix_tmp_bytes(nt) + 1024));
        int alt2 = 0;
        GDSJ_TRY(radix_sort_pairs(tk, tv, tk + nt, tv + nt, nt, order_tiles == 2 ? 8 : 32, G.tmp.p, s, &alt2));
        k_tiles_permute<<<ceil_div(nt, 256), 256, 0, s>>>(tilesB, alt2 ? tv + nt : tv, nt, tilesA);
    } else {
        k_tiles_fill<<<ceil_div(nc, 256), 256, 0, s>>>(G.cellStart.as<uint32_t>(), nc, G.tileq, tileOff, tilesA,
                                                        nullptr, nullptr, 0);
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
    if (tileq == 0 || tileq > 128) tileq = 64;
    G.n = n;
    G.d = d;
    G.k = k;
    G.dtype = dtype;
    G.eps = eps;
    G.tileq = tileq;
    if (dtype == GDSJ_DTYPE_F32) return grid_build_t<float>(G, (const float *)d_data, s, reorder, order_tiles);
    if (dtype == GDSJ_DTYPE_F64) return grid_build_t<double>(G, (const double *)d_data, s, reorder, order_tiles);
    set_error("unknown dtype %d", dtype);
    return GDSJ_EINVAL;
}

}  // namespace gdsj
