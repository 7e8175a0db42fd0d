// gram.hip -- MFMA Gram-form filter join (see gram.h for the bound).
//
// One wave per tile of up to RB blocks of 32 queries of one cell.  Each
// block's h vectors are the A operand of v_mfma_f32_32x32x16_{bf16,f16}
// (lane l: row l%32, k = 8*(l/32)..+7 of each 16-wide k-step); 32
// candidates at a time are the B operand (lane l: column l%32), loaded once
// and used by all RB row blocks.  Output register i of lane l holds
// S[row (i/4)*8 + (l/32)*4 + i%4][column l%32].  In the split form the
// operands are h + l (two bf16 terms) and S = h.h' + h.l' + l.h'; the
// neglected l.l' is inside the bound.  Survivors of S >= W_q + W_p (one wave
// mask per accumulator entry) queue in LDS across blocks and are rechecked
// exactly, 64 at a time, from a row-major copy of the points; hits are
// counted / written per row in queue order, which is candidate order, so the
// rows come out identical to the plain join's.
#include <math.h>

#include "gram_kernels.h"

namespace gdsj {

void GramData::release() {
    h.release();
    l.release();
    xr.release();
    pm.release();
    axes = 0;
    xrs = 0;
    w.release();
    n2.release();
    scratch.release();
    dpad = 0;
}

static uint32_t pad_dims(uint32_t d) {
    const uint32_t ks = (d + 15) / 16;
    const uint32_t kp = ks <= 4 ? ks : (ks <= 6 ? 6 : 8);
    return kp * 16;
}

bool gram_supported(uint32_t d) { return d >= 16 && d <= 128; }

uint32_t gram_tile_queries(uint32_t d, int lowp) {
    const int ks = (int)(pad_dims(d) / 16);
    return 32u * (uint32_t)rb_max(ks, lowp == 0 ? 1 : 0);
}

// ------------------------------------------------------------- preparation --
// Per-dim mean and max |x - mean|: GS blocks per dim write partials, one
// block per dim reduces them in a fixed order (deterministic).
constexpr uint32_t GS = 64;

template <typename T>
__global__ __launch_bounds__(256) void k_dim_partial(const T *__restrict__ xs, uint64_t ld, uint32_t n,
                                                      const double *__restrict__ mean, double *__restrict__ part) {
    const uint32_t l = blockIdx.y, g = blockIdx.x;
    const uint32_t per = (n + GS - 1) / GS;
    const uint32_t i0 = g * per, i1 = i0 + per < n ? i0 + per : n;
    const bool dev = mean != nullptr;  // second phase: max deviation from the mean
    const double o = dev ? mean[l] : 0.0;
    double v = 0.0;
    for (uint32_t i = i0 + threadIdx.x; i < i1; i += blockDim.x) {
        const double x = (double)xs[(uint64_t)l * ld + i];
        v = dev ? fmax(v, fabs(x - o)) : v + x;
    }
    __shared__ double red[256];
    red[threadIdx.x] = v;
    __syncthreads();
    for (uint32_t q = 128; q > 0; q >>= 1) {
        if (threadIdx.x < q) red[threadIdx.x] = dev ? fmax(red[threadIdx.x], red[threadIdx.x + q])
                                                    : red[threadIdx.x] + red[threadIdx.x + q];
        __syncthreads();
    }
    if (threadIdx.x == 0) part[(uint64_t)l * GS + g] = red[0];
}

__global__ void k_dim_finish(const double *__restrict__ part, uint32_t d, uint32_t n, int dev,
                             double *__restrict__ out) {
    const uint32_t l = blockIdx.x * blockDim.x + threadIdx.x;
    if (l >= d) return;
    double v = 0.0;
    for (uint32_t g = 0; g < GS; g++) v = dev ? fmax(v, part[(uint64_t)l * GS + g]) : v + part[(uint64_t)l * GS + g];
    out[l] = dev ? v : (n ? v / n : 0.0);
}

// h = lp(sigma (x - o)), l = lp(sigma (x - o) - h) in the split form
// (subnormals flushed to zero here, so the MFMA sees exactly the values the
// bound was computed from); e = |sigma (x - o) - h - l| measured in double
// and widened for the double rounding of sigma (x - o).  With N = |h + l|^2,
// N1 = | |h| + |l| |^2 (bounds the magnitude of the MFMA's products),
// L = |l|^2 (bounds the neglected l.l') and k = sigma eps'/2 + e:
//     W = (N - c N1 - L) / 2 - (1 + 2^-20) k^2, rounded down.
template <int LP>
__device__ __forceinline__ double lp_round(double v, typename Lp<LP>::scalar &out) {
    typename Lp<LP>::scalar h = (typename Lp<LP>::scalar)(float)v;
    double hd = (double)(float)h;
    if (fabs(hd) < Lp<LP>::min_normal) {
        hd = 0.0;
        h = (typename Lp<LP>::scalar)0.0f;
    }
    out = h;
    return hd;
}

// W from the point's N, N1, L, e and sigma eps' / 2.  cfac covers the MFMA's
// FP32 accumulation of the products (and the initial -W_q it accumulates
// onto, |W_q| <= N1_q / 2 + k_q^2: cfac holds twice the term count's
// roundings, and the k^2 factor the same).
__device__ __forceinline__ float gram_weight(double N, double N1, double L, double e, double half_eps, double cfac) {
    const double k = half_eps + e;
    const double w = 0.5 * (N * (1.0 - 0x1p-40) - cfac * N1 - L) - (1.0 + 0x1p-20 + cfac) * k * k;
    return __double2float_rd(w);
}

template <typename T, int LP>
__global__ void k_gram_convert(const T *__restrict__ xs, uint64_t ld, uint32_t n, uint32_t d, uint32_t dpad,
                               const double *__restrict__ mean, const double *__restrict__ maxdev, double eps_p,
                               double cfac, int split, typename Lp<LP>::scalar *__restrict__ H,
                               typename Lp<LP>::scalar *__restrict__ Lo, float *__restrict__ W,
                               double *__restrict__ N2, double *__restrict__ sigOut) {
    const uint32_t s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= n) return;
    double sigma = 1.0;
    if (LP == 1) {
        double mx = 0.0;
        for (uint32_t l = 0; l < d; l++) mx = fmax(mx, maxdev[l]);
        if (mx > 0.0) sigma = exp2(floor(log2(16384.0 / mx)));
        sigma = fmin(fmax(sigma, 0x1p-60), 0x1p60);
    }
    double N = 0.0, N1 = 0.0, L = 0.0, E2 = 0.0, X2 = 0.0;
    typename Lp<LP>::scalar *hp = H + (uint64_t)s * dpad;
    typename Lp<LP>::scalar *lp = split ? Lo + (uint64_t)s * dpad : nullptr;
    // 8 dims per step: the operands leave as one 16-byte store per step
    // (dpad is a multiple of 16)
    for (uint32_t l0 = 0; l0 < dpad; l0 += 8) {
        typename Lp<LP>::scalar hb[8], lb[8];
#pragma unroll
        for (uint32_t u = 0; u < 8; u++) {
            const uint32_t l = l0 + u;
            double xp = 0.0;
            if (l < d) xp = ((double)xs[(uint64_t)l * ld + s] - mean[l]) * sigma;
            typename Lp<LP>::scalar hv, lv = {};
            const double hd = lp_round<LP>(xp, hv);
            double ld2 = 0.0;
            hb[u] = hv;
            if (split) ld2 = lp_round<LP>(xp - hd, lv);
            lb[u] = lv;
            const double a = hd + ld2;  // exact: ld2 is the leading bits of xp - hd
            const double r = xp - a;
            N += a * a;
            N1 += (fabs(hd) + fabs(ld2)) * (fabs(hd) + fabs(ld2));
            L += ld2 * ld2;
            E2 += r * r;
            X2 += xp * xp;
        }
        uint4 hv4, lv4;
        __builtin_memcpy(&hv4, hb, 16);
        *reinterpret_cast<uint4 *>(hp + l0) = hv4;
        if (split) {
            __builtin_memcpy(&lv4, lb, 16);
            *reinterpret_cast<uint4 *>(lp + l0) = lv4;
        }
    }
    const double e = sqrt(E2) * (1.0 + 0x1p-30) + sqrt(X2) * 0x1p-50;
    W[s] = gram_weight(N, N1, L, e, sigma * eps_p * 0.5, cfac);
    N2[s] = N;
    if (s == 0) sigOut[0] = sigma;
}

// The scale of the projected operands: |y_c| <= sum_l |P_cl| maxdev_l.
__global__ void k_proj_sigma(const double *__restrict__ P, uint32_t m, uint32_t d, const double *__restrict__ maxdev,
                             int lowp, double *__restrict__ sigOut) {
    __shared__ double red[64];
    double v = 0.0;
    for (uint32_t c = threadIdx.x; c < m; c += 64) {
        double b = 0.0;
        for (uint32_t l = 0; l < d; l++) b += fabs(P[(size_t)c * d + l]) * maxdev[l];
        v = fmax(v, b);
    }
    red[threadIdx.x] = v;
    __syncthreads();
    for (uint32_t q = 32; q > 0; q >>= 1) {
        if (threadIdx.x < q) red[threadIdx.x] = fmax(red[threadIdx.x], red[threadIdx.x + q]);
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        double sigma = 1.0;
        const double mx = red[0] * (1.0 + 1e-6);
        if (lowp == 1) {
            if (mx > 0.0) sigma = exp2(floor(log2(16384.0 / mx)));
            sigma = fmin(fmax(sigma, 0x1p-60), 0x1p60);
        }
        sigOut[0] = sigma;
    }
}

// h (and l) of y = P (x - o) on M principal axes, in double: v_l = fl(x_l -
// o_l), y_c = FMA sum of P_cl v_l, so |y_c - P_c (x - o)| <= gamma(d + 1)
// A_c with A_c = sum |P_cl| |v_l| (accumulated beside y_c); the 2-norm of
// that error over the M axes, scaled by sigma, joins the measured rounding
// e of the operands.  eps_p arrives widened by |P|_2.
template <typename T, int LP, uint32_t M>
__global__ __launch_bounds__(256) void k_gram_convert_proj(
    const T *__restrict__ xs, uint64_t ld, uint32_t n, uint32_t d, const double *__restrict__ P,
    const double *__restrict__ mean, const double *__restrict__ sigIn, double eps_p, double cfac, int split,
    typename Lp<LP>::scalar *__restrict__ H, typename Lp<LP>::scalar *__restrict__ Lo, float *__restrict__ W,
    double *__restrict__ N2) {
    const uint32_t s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= n) return;
    const double sigma = sigIn[0];
    double y[M], A[M];
#pragma unroll
    for (uint32_t c = 0; c < M; c++) {
        y[c] = 0.0;
        A[c] = 0.0;
    }
    for (uint32_t l = 0; l < d; l++) {
        const double v = (double)xs[(uint64_t)l * ld + s] - mean[l];
        const double av = fabs(v);
#pragma unroll
        for (uint32_t c = 0; c < M; c++) {
            const double pc = P[(size_t)c * d + l];
            y[c] = fma(pc, v, y[c]);
            A[c] = fma(fabs(pc), av, A[c]);
        }
    }
    double N = 0.0, N1 = 0.0, L = 0.0, E2 = 0.0, X2 = 0.0, A2 = 0.0;
    typename Lp<LP>::scalar *hp = H + (uint64_t)s * M;
    typename Lp<LP>::scalar *lp = split ? Lo + (uint64_t)s * M : nullptr;
#pragma unroll
    for (uint32_t l0 = 0; l0 < M; l0 += 8) {
        typename Lp<LP>::scalar hb[8], lb[8];
#pragma unroll
        for (uint32_t u = 0; u < 8; u++) {
            const double xp = y[l0 + u] * sigma;
            typename Lp<LP>::scalar hv, lv = {};
            const double hd = lp_round<LP>(xp, hv);
            double ld2 = 0.0;
            hb[u] = hv;
            if (split) ld2 = lp_round<LP>(xp - hd, lv);
            lb[u] = lv;
            const double a = hd + ld2;
            const double r = xp - a;
            N += a * a;
            N1 += (fabs(hd) + fabs(ld2)) * (fabs(hd) + fabs(ld2));
            L += ld2 * ld2;
            E2 += r * r;
            X2 += xp * xp;
            A2 += A[l0 + u] * A[l0 + u];
        }
        uint4 hv4, lv4;
        __builtin_memcpy(&hv4, hb, 16);
        *reinterpret_cast<uint4 *>(hp + l0) = hv4;
        if (split) {
            __builtin_memcpy(&lv4, lb, 16);
            *reinterpret_cast<uint4 *>(lp + l0) = lv4;
        }
    }
    const double gd1 = ((double)d + 2.0) * 0x1p-53 / (1.0 - ((double)d + 2.0) * 0x1p-53);
    const double eproj = sigma * gd1 * sqrt(A2) * (1.0 + 0x1p-20);
    const double e = sqrt(E2) * (1.0 + 0x1p-30) + sqrt(X2) * 0x1p-50 + eproj;
    W[s] = gram_weight(N, N1, L, e, sigma * eps_p * 0.5, cfac);
    N2[s] = N;
}

// row-major copy of the sorted points for the exact rechecks (padding zero),
// TP points per block transposed through LDS (tile = [xrs][TP + 1]): the SoA
// columns are read and the rows written as contiguous runs
template <typename T, uint32_t TP>
__global__ __launch_bounds__(256) void k_rows(const T *__restrict__ xs, uint64_t ld, uint32_t n, uint32_t d,
                                              uint32_t xrs, T *__restrict__ xr) {
    extern __shared__ __align__(16) unsigned char rows_lds[];
    T *tile = reinterpret_cast<T *>(rows_lds);
    const uint32_t s0 = blockIdx.x * TP;
    const uint32_t np = n - s0 < TP ? n - s0 : TP;
    for (uint32_t t = threadIdx.x; t < xrs * TP; t += blockDim.x) {
        const uint32_t l = t / TP, p = t % TP;
        tile[l * (TP + 1) + p] = (l < d && p < np) ? xs[(uint64_t)l * ld + s0 + p] : (T)0;
    }
    __syncthreads();
    for (uint32_t t = threadIdx.x; t < np * xrs; t += blockDim.x) {
        const uint32_t p = t / xrs, l = t - p * xrs;
        xr[(uint64_t)(s0 + p) * xrs + l] = tile[l * (TP + 1) + p];
    }
}

template <typename T, int LP>
static int prepare_t(GramData &gd, const Grid &G, hipStream_t s) {
    const uint32_t n = G.n, d = G.d, dpad = gd.dpad;
    GDSJ_TRY(gd.h.ensure((size_t)n * dpad * 2 + 256));
    if (gd.split) GDSJ_TRY(gd.l.ensure((size_t)n * dpad * 2 + 256));
    GDSJ_TRY(gd.w.ensure((size_t)n * 4 + 256));
    GDSJ_TRY(gd.n2.ensure((size_t)n * 8 + 256));
    GDSJ_TRY(gd.scratch.ensure((size_t)d * 16 + 256 + (size_t)d * GS * 8));
    gd.xrs = (d + 15) / 16 * 16;
    GDSJ_TRY(gd.xr.ensure((size_t)n * gd.xrs * sizeof(T) + 256));
    if (n) {
        if ((size_t)gd.xrs * 65 * sizeof(T) <= 48 * 1024)
            k_rows<T, 64><<<ceil_div(n, 64), 256, (size_t)gd.xrs * 65 * sizeof(T), s>>>(G.xs.as<T>(), G.ld, n, d,
                                                                                       gd.xrs, gd.xr.as<T>());
        else
            k_rows<T, 32><<<ceil_div(n, 32), 256, (size_t)gd.xrs * 33 * sizeof(T), s>>>(G.xs.as<T>(), G.ld, n, d,
                                                                                       gd.xrs, gd.xr.as<T>());
        GDSJ_HIP(hipGetLastError());
    }
    double *mean = gd.scratch.as<double>();
    double *maxdev = mean + d;
    double *part = mean + 2 * (size_t)d + 8;  // after mean, maxdev and sigma
    const T *xs = G.xs.as<T>();
    k_dim_partial<T><<<dim3(GS, d), 256, 0, s>>>(xs, G.ld, n, nullptr, part);
    k_dim_finish<<<ceil_div(d, 64), 64, 0, s>>>(part, d, n, 0, mean);
    k_dim_partial<T><<<dim3(GS, d), 256, 0, s>>>(xs, G.ld, n, mean, part);
    k_dim_finish<<<ceil_div(d, 64), 64, 0, s>>>(part, d, n, 1, maxdev);
    // eps' = eps (1 + 2 (D+5) u_T): beyond it the DTYPE sum cannot pass the predicate
    const double uT = sizeof(T) == 4 ? 0x1p-24 : 0x1p-53;
    const double eps_p = G.eps * (1.0 + 2.0 * (d + 5) * uT);
    // MFMA FP32 accumulation of (1 or 3) x dpad exact products, with margin
    // for the FP32 rounding of W_q + W_p and of the stored N
    // (and, as the kernel accumulates onto -W_q, once more for that term)
    const double cfac = (4.0 * (gd.split ? 3.0 : 1.0) * dpad + 24.0) * 0x1p-24;
    auto *Ho = (typename Lp<LP>::scalar *)gd.h.p;
    auto *Lo = gd.split ? (typename Lp<LP>::scalar *)gd.l.p : nullptr;
    if (gd.axes) {
        // operands on the principal axes: eps' widened by |P|_2
        double *sig = mean + 2 * (size_t)d;
        k_proj_sigma<<<1, 64, 0, s>>>(gd.pm.as<double>(), gd.axes, d, maxdev, LP, sig);
        const double eps_pp = eps_p * G.gramNorm;
        if (gd.axes == 16)
            k_gram_convert_proj<T, LP, 16><<<ceil_div(n, 256), 256, 0, s>>>(
                xs, G.ld, n, d, gd.pm.as<double>(), mean, sig, eps_pp, cfac, gd.split, Ho, Lo, gd.w.as<float>(),
                gd.n2.as<double>());
        else
            k_gram_convert_proj<T, LP, 32><<<ceil_div(n, 256), 256, 0, s>>>(
                xs, G.ld, n, d, gd.pm.as<double>(), mean, sig, eps_pp, cfac, gd.split, Ho, Lo, gd.w.as<float>(),
                gd.n2.as<double>());
    } else {
        k_gram_convert<T, LP><<<ceil_div(n, 256), 256, 0, s>>>(xs, G.ld, n, d, dpad, mean, maxdev, eps_p, cfac,
                                                               gd.split, Ho, Lo, gd.w.as<float>(), gd.n2.as<double>(),
                                                               mean + 2 * (size_t)d);
    }
    GDSJ_HIP(hipGetLastError());
    return GDSJ_OK;
}

int gram_prepare(GramData &gd, const Grid &G, int lowp, hipStream_t s) {
    if (!gram_supported(G.d)) {
        set_error("the Gram filter handles 16 to 128 dims");
        return GDSJ_EUNSUPPORTED;
    }
    // principal axes (16 or 32) where the grid build chose them, else the dims
    gd.axes = (G.gramAxes == 16 || G.gramAxes == 32) && G.gramAxes < pad_dims(G.d) &&
                      G.gramP.size() == (size_t)G.gramAxes * G.d
                  ? G.gramAxes
                  : 0u;
    gd.dpad = gd.axes ? gd.axes : pad_dims(G.d);
    // fp16 operands by default (one MFMA per k-step; measured faster on
    // FP32 data than bf16 hi + lo: half the operand bytes, the band only
    // slightly wider); bf16 only for FP32 data on request
    if (G.dtype == GDSJ_DTYPE_F64) lowp = 1;
    gd.lowp = lowp;
    gd.split = lowp == 0 ? 1 : 0;  // bf16 carries 8 bits: two terms; fp16's 11 bits suffice alone
    if (G.n == 0) return GDSJ_OK;
    if (gd.axes) {
        GDSJ_TRY(gd.pm.ensure(G.gramP.size() * 8 + 256));
        GDSJ_HIP(hipMemcpyAsync(gd.pm.p, G.gramP.data(), G.gramP.size() * 8, hipMemcpyHostToDevice, s));
    }
    if (G.dtype == GDSJ_DTYPE_F32) return lowp ? prepare_t<float, 1>(gd, G, s) : prepare_t<float, 0>(gd, G, s);
    return lowp ? prepare_t<double, 1>(gd, G, s) : prepare_t<double, 0>(gd, G, s);
}

int gram_dispatch_f32(const GramData &gd, const JoinArgs &a, int mode, int pass, unsigned long long *stats,
                      bool err, hipStream_t s);
int gram_dispatch_f64(const GramData &gd, const JoinArgs &a, int mode, int pass, unsigned long long *stats,
                      bool err, hipStream_t s);

int gram_join_pass(const Grid &G, const GramData &gd, const JoinArgs &a, int mode, int pass,
                   unsigned long long *stats, hipStream_t s, bool measure_err) {
    const uint32_t opd = gd.axes ? gd.axes : G.d;  // the operands' dims
    if (G.tileq > gram_tile_queries(opd, gd.lowp)) {
        set_error("Gram tiles hold at most %u queries here", gram_tile_queries(opd, gd.lowp));
        return GDSJ_EINVAL;
    }
    // FP32 data: fp16 or bf16 (split in two terms); FP64 data: fp16
    if (G.dtype == GDSJ_DTYPE_F32) return gram_dispatch_f32(gd, a, mode, pass, stats, measure_err, s);
    return gram_dispatch_f64(gd, a, mode, pass, stats, measure_err, s);
}

}  // namespace gdsj
