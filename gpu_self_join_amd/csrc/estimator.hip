// estimator.hip -- learned and sampling result-size estimators (see
// estimator.h).  The MLP kernels are plain FP32 VALU code: the network has
// 1,633 parameters and runs once per tile, far from any roofline.
#include <math.h>
#include <string.h>

#include <algorithm>
#include <vector>

#include "estimator.h"
#include "join.h"
#include "scan.h"

namespace gdsj {

void Estimator::release() {
    DevBuf *b[] = {&params, &adam_m, &adam_v, &partial, &feat, &target, &pred, &sample, &work};
    for (DevBuf *p : b) p->release();
    init = trained = false;
    step = 0;
}

// ---------------------------------------------------------------- features --
// [0] log2(1+queries of the tile)       [1] log2(1+points of its cell)
// [2] log2(1+candidates per query)      [3] log2(1+non-empty neighbor runs)
// [4] log2(1+queries x candidates)      [5..10] neighbor-run length histogram
// [11] dims/64  [12] k/12  [13] log2(1+n)/32  [14] log2(1+cell size^2/candidates)
// [15] 1 (bias feature).  Log features are scaled by 1/16.
__global__ void k_tile_features(const uint4 *__restrict__ tiles, uint32_t nt, const uint32_t *__restrict__ cellStart,
                                const uint32_t *__restrict__ cellCand, const uint32_t *__restrict__ nruns,
                                const float *__restrict__ cellHist, uint32_t d, uint32_t k, uint32_t n,
                                float *__restrict__ feat) {
    const uint32_t t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= nt) return;
    const uint4 tl = tiles[t];
    const uint32_t c = tl.x;
    const float nq = (float)(tl.z - tl.y);
    const float cs = (float)(cellStart[c + 1] - cellStart[c]);
    const float cand = (float)cellCand[c];
    const float nr = (float)nruns[c];
    float *f = feat + (uint64_t)t * MLP_F;
    const float s = 1.0f / 16.0f;
    f[0] = log2f(1.0f + nq) * s;
    f[1] = log2f(1.0f + cs) * s;
    f[2] = log2f(1.0f + cand) * s;
    f[3] = log2f(1.0f + nr) * s;
    f[4] = log2f(1.0f + nq * cand) * s;
    for (int b = 0; b < RUN_HIST_BINS; b++) f[5 + b] = cellHist[(uint64_t)c * RUN_HIST_BINS + b];
    f[11] = (float)d / 64.0f;
    f[12] = (float)k / 12.0f;
    f[13] = log2f(1.0f + (float)n) / 32.0f;
    f[14] = log2f(1.0f + cs * cs / fmaxf(cand, 1.0f)) * s;
    f[15] = 1.0f;
}

int tile_features(const Grid &G, float *feat, hipStream_t s) {
    if (G.ntiles == 0) return GDSJ_OK;
    k_tile_features<<<ceil_div(G.ntiles, 256), 256, 0, s>>>(G.tiles.as<uint4>(), G.ntiles,
                                                            G.cellStart.as<uint32_t>(), G.cellCand.as<uint32_t>(),
                                                            G.nruns.as<uint32_t>(), G.cellHist.as<float>(), G.d,
                                                            G.k, G.n, feat);
    GDSJ_HIP(hipGetLastError());
    return GDSJ_OK;
}

__global__ void k_tile_targets(const uint4 *__restrict__ tiles, uint32_t nt, const uint32_t *__restrict__ ids,
                               const uint32_t *__restrict__ counts, float *__restrict__ target) {
    const uint32_t t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= nt) return;
    const uint4 tl = tiles[t];
    uint64_t sum = 0;
    for (uint32_t q = tl.y; q < tl.z; q++) sum += counts[ids[q]];
    target[t] = log2f(1.0f + (float)sum);
}

int tile_targets(const Grid &G, const uint32_t *counts, float *target, hipStream_t s) {
    if (G.ntiles == 0) return GDSJ_OK;
    k_tile_targets<<<ceil_div(G.ntiles, 256), 256, 0, s>>>(G.tiles.as<uint4>(), G.ntiles, G.ids.as<uint32_t>(),
                                                           counts, target);
    GDSJ_HIP(hipGetLastError());
    return GDSJ_OK;
}

// --------------------------------------------------------------------- MLP --
__device__ __forceinline__ float mlp_eval(const float *__restrict__ P, const float (&x)[MLP_F], float (&h1)[MLP_H],
                                          float (&h2)[MLP_H]) {
#pragma unroll 4
    for (int j = 0; j < MLP_H; j++) {
        float a = P[MLP_OFF_B1 + j];
#pragma unroll
        for (int i = 0; i < MLP_F; i++) a = fmaf(P[MLP_OFF_W1 + j * MLP_F + i], x[i], a);
        h1[j] = fmaxf(a, 0.0f);
    }
#pragma unroll 4
    for (int j = 0; j < MLP_H; j++) {
        float a = P[MLP_OFF_B2 + j];
#pragma unroll
        for (int i = 0; i < MLP_H; i++) a = fmaf(P[MLP_OFF_W2 + j * MLP_H + i], h1[i], a);
        h2[j] = fmaxf(a, 0.0f);
    }
    float y = P[MLP_OFF_B3];
#pragma unroll
    for (int i = 0; i < MLP_H; i++) y = fmaf(P[MLP_OFF_W3 + i], h2[i], y);
    return y;
}

__global__ __launch_bounds__(256) void k_mlp_forward(const float *__restrict__ params, const float *__restrict__ x,
                                                     uint32_t n, float *__restrict__ y) {
    __shared__ float P[MLP_P];
    for (uint32_t i = threadIdx.x; i < (uint32_t)MLP_P; i += blockDim.x) P[i] = params[i];
    __syncthreads();
    const uint32_t t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= n) return;
    float xv[MLP_F], h1[MLP_H], h2[MLP_H];
#pragma unroll
    for (int i = 0; i < MLP_F; i++) xv[i] = x[(uint64_t)t * MLP_F + i];
    y[t] = mlp_eval(P, xv, h1, h2);
}

int mlp_forward(const float *params, const float *x, uint32_t n, float *y, hipStream_t s) {
    if (n == 0) return GDSJ_OK;
    k_mlp_forward<<<ceil_div(n, 256), 256, 0, s>>>(params, x, n, y);
    GDSJ_HIP(hipGetLastError());
    return GDSJ_OK;
}

constexpr int GB_S = 64;  // samples per block in the gradient kernel

// Per block: forward + backward of GB_S samples (one thread each), then the
// parameter gradients summed over those samples as small products over LDS
// (one thread per parameter, fixed order), written to partial[block].
// Sample j of the minibatch is row (a * (base + j) + c) mod m of (x, t).
__global__ __launch_bounds__(256) void k_mlp_grad(const float *__restrict__ params, const float *__restrict__ x,
                                                  const float *__restrict__ t, uint32_t m, uint32_t base,
                                                  uint32_t count, uint32_t pa, uint32_t pc, float inv_n,
                                                  float *__restrict__ partial) {
    __shared__ float P[MLP_P];
    __shared__ float sx[GB_S][MLP_F];
    __shared__ float sh1[GB_S][MLP_H], sh2[GB_S][MLP_H], sd1[GB_S][MLP_H], sd2[GB_S][MLP_H];
    __shared__ float sdy[GB_S], sl[GB_S];
    for (uint32_t i = threadIdx.x; i < (uint32_t)MLP_P; i += blockDim.x) P[i] = params[i];
    __syncthreads();
    const uint32_t j = threadIdx.x;
    if (j < (uint32_t)GB_S) {
        const uint32_t b = blockIdx.x * GB_S + j;
        float xv[MLP_F], h1[MLP_H], h2[MLP_H];
        if (b < count) {
            const uint32_t row = (uint32_t)(((uint64_t)pa * (base + b) + pc) % m);
#pragma unroll
            for (int i = 0; i < MLP_F; i++) xv[i] = x[(uint64_t)row * MLP_F + i];
            const float y = mlp_eval(P, xv, h1, h2);
            const float e = y - t[row];
            const float dy = 2.0f * e * inv_n;
            sdy[j] = dy;
            sl[j] = e * e;
            float d2[MLP_H];
#pragma unroll
            for (int i = 0; i < MLP_H; i++) d2[i] = h2[i] > 0.0f ? dy * P[MLP_OFF_W3 + i] : 0.0f;
#pragma unroll 4
            for (int i = 0; i < MLP_H; i++) {
                float a = 0.0f;
#pragma unroll
                for (int r = 0; r < MLP_H; r++) a = fmaf(d2[r], P[MLP_OFF_W2 + r * MLP_H + i], a);
                sd1[j][i] = h1[i] > 0.0f ? a : 0.0f;
            }
#pragma unroll
            for (int i = 0; i < MLP_H; i++) {
                sh1[j][i] = h1[i];
                sh2[j][i] = h2[i];
                sd2[j][i] = d2[i];
            }
#pragma unroll
            for (int i = 0; i < MLP_F; i++) sx[j][i] = xv[i];
        } else {
            sdy[j] = 0.0f;
            sl[j] = 0.0f;
#pragma unroll
            for (int i = 0; i < MLP_H; i++) sh1[j][i] = sh2[j][i] = sd1[j][i] = sd2[j][i] = 0.0f;
#pragma unroll
            for (int i = 0; i < MLP_F; i++) sx[j][i] = 0.0f;
        }
    }
    __syncthreads();
    float *out = partial + (uint64_t)blockIdx.x * (MLP_P + 1);
    for (uint32_t p = threadIdx.x; p < (uint32_t)MLP_P + 1; p += blockDim.x) {
        float g = 0.0f;
        if (p < (uint32_t)MLP_OFF_B1) {
            const uint32_t r = p / MLP_F, c = p % MLP_F;
            for (int s = 0; s < GB_S; s++) g = fmaf(sd1[s][r], sx[s][c], g);
        } else if (p < (uint32_t)MLP_OFF_W2) {
            const uint32_t r = p - MLP_OFF_B1;
            for (int s = 0; s < GB_S; s++) g += sd1[s][r];
        } else if (p < (uint32_t)MLP_OFF_B2) {
            const uint32_t q = p - MLP_OFF_W2, r = q / MLP_H, c = q % MLP_H;
            for (int s = 0; s < GB_S; s++) g = fmaf(sd2[s][r], sh1[s][c], g);
        } else if (p < (uint32_t)MLP_OFF_W3) {
            const uint32_t r = p - MLP_OFF_B2;
            for (int s = 0; s < GB_S; s++) g += sd2[s][r];
        } else if (p < (uint32_t)MLP_OFF_B3) {
            const uint32_t c = p - MLP_OFF_W3;
            for (int s = 0; s < GB_S; s++) g = fmaf(sdy[s], sh2[s][c], g);
        } else if (p == (uint32_t)MLP_OFF_B3) {
            for (int s = 0; s < GB_S; s++) g += sdy[s];
        } else {
            for (int s = 0; s < GB_S; s++) g += sl[s];  // squared-error sum (loss slot)
        }
        out[p] = g;
    }
}

// grad[p] = sum over blocks (fixed order); grad[MLP_P] = mean loss
__global__ void k_mlp_reduce(const float *__restrict__ partial, uint32_t nblocks, float inv_n,
                             float *__restrict__ grad) {
    const uint32_t p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p > (uint32_t)MLP_P) return;
    float g = 0.0f;
    for (uint32_t b = 0; b < nblocks; b++) g += partial[(uint64_t)b * (MLP_P + 1) + p];
    grad[p] = p == (uint32_t)MLP_P ? g * inv_n : g;
}

size_t mlp_grad_scratch_floats(uint32_t n) { return (size_t)ceil_div(n, GB_S) * (MLP_P + 1) + 64; }

int mlp_grad_rows(const float *params, const float *x, const float *t, uint32_t m, uint32_t base, uint32_t count,
                  uint32_t pa, uint32_t pc, float *grad, float *partial, hipStream_t s) {
    if (count == 0) return GDSJ_OK;
    const uint32_t nb = ceil_div(count, GB_S);
    k_mlp_grad<<<nb, 256, 0, s>>>(params, x, t, m, base, count, pa, pc, 1.0f / (float)count, partial);
    k_mlp_reduce<<<ceil_div(MLP_P + 1, 256), 256, 0, s>>>(partial, nb, 1.0f / (float)count, grad);
    GDSJ_HIP(hipGetLastError());
    return GDSJ_OK;
}

int mlp_grad(const float *params, const float *x, const float *t, uint32_t n, float *grad, float *loss,
             float *partial, hipStream_t s) {
    GDSJ_TRY(mlp_grad_rows(params, x, t, n, 0, n, 1, 0, grad, partial, s));
    if (loss) GDSJ_HIP(hipMemcpyAsync(loss, grad + MLP_P, 4, hipMemcpyDeviceToDevice, s));
    return GDSJ_OK;
}

__global__ void k_adam(float *__restrict__ params, const float *__restrict__ grad, float *__restrict__ m,
                       float *__restrict__ v, float lr, float b1, float b2, float c1, float c2) {
    const uint32_t p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= (uint32_t)MLP_P) return;
    const float g = grad[p];
    const float mm = b1 * m[p] + (1.0f - b1) * g;
    const float vv = b2 * v[p] + (1.0f - b2) * g * g;
    m[p] = mm;
    v[p] = vv;
    params[p] -= lr * (mm / c1) / (sqrtf(vv / c2) + 1e-8f);
}

int mlp_adam(float *params, const float *grad, float *m, float *v, uint32_t step, float lr, hipStream_t s) {
    const float b1 = 0.9f, b2 = 0.999f;
    const float c1 = 1.0f - powf(b1, (float)step), c2 = 1.0f - powf(b2, (float)step);
    k_adam<<<ceil_div(MLP_P, 256), 256, 0, s>>>(params, grad, m, v, lr, b1, b2, c1, c2);
    GDSJ_HIP(hipGetLastError());
    return GDSJ_OK;
}

static uint64_t splitmix(uint64_t &st) {
    uint64_t z = (st += 0x9E3779B97F4A7C15ull);
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}

int mlp_init(float *params, uint64_t seed, hipStream_t s) {
    std::vector<float> h(MLP_P, 0.0f);
    uint64_t st = seed;
    auto uni = [&](float a) { return a * (2.0f * (float)((splitmix(st) >> 40) * 0x1p-24) - 1.0f); };
    for (int i = 0; i < MLP_H * MLP_F; i++) h[MLP_OFF_W1 + i] = uni(sqrtf(6.0f / MLP_F));
    for (int i = 0; i < MLP_H * MLP_H; i++) h[MLP_OFF_W2 + i] = uni(sqrtf(6.0f / MLP_H));
    for (int i = 0; i < MLP_H; i++) h[MLP_OFF_W3 + i] = uni(sqrtf(3.0f / MLP_H));
    GDSJ_HIP(hipMemcpyAsync(params, h.data(), MLP_P * 4, hipMemcpyHostToDevice, s));
    GDSJ_HIP(hipStreamSynchronize(s));
    return GDSJ_OK;
}

// ------------------------------------------------------------- training --
__global__ void k_gather_rows(const float *__restrict__ feat, const float *__restrict__ target, uint32_t nt,
                              uint32_t m, float *__restrict__ xs, float *__restrict__ ts) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= m) return;
    const uint32_t t = (uint32_t)(((uint64_t)i * nt) / m);
    for (int f = 0; f < MLP_F; f++) xs[(uint64_t)i * MLP_F + f] = feat[(uint64_t)t * MLP_F + f];
    ts[i] = target[t];
}

static uint32_t coprime_step(uint32_t m) {
    const uint32_t cands[] = {40503u, 48271u, 69621u, 16807u, 2654435761u % 1000003u};
    for (uint32_t c : cands) {
        uint32_t a = c % m, b = m;
        if (a == 0) continue;
        uint32_t x = a, y = b;
        while (y) { uint32_t r = x % y; x = y; y = r; }
        if (x == 1) return a;
    }
    return 1;
}

int estimator_train(Estimator &E, const Grid &G, const uint32_t *counts, hipStream_t s, double *ms) {
    const uint32_t nt = G.ntiles;
    if (nt == 0 || !G.haveRuns) return GDSJ_OK;
    GDSJ_TRY(E.feat.ensure((size_t)nt * MLP_F * 4));
    GDSJ_TRY(E.target.ensure((size_t)nt * 4));
    GDSJ_TRY(tile_features(G, E.feat.as<float>(), s));
    GDSJ_TRY(tile_targets(G, counts, E.target.as<float>(), s));
    return estimator_train_rows(E, E.feat.as<float>(), E.target.as<float>(), nt, s, ms);
}

int estimator_train_rows(Estimator &E, const float *feat, const float *target, uint32_t nt, hipStream_t s,
                         double *ms) {
    if (nt == 0) return GDSJ_OK;
    hipEvent_t e0, e1;
    (void)hipEventCreate(&e0);
    (void)hipEventCreate(&e1);
    GDSJ_HIP(hipEventRecord(e0, s));
    GDSJ_TRY(E.params.ensure(MLP_P * 4));
    GDSJ_TRY(E.adam_m.ensure(MLP_P * 4));
    GDSJ_TRY(E.adam_v.ensure(MLP_P * 4));
    const uint32_t m = std::min<uint32_t>(nt, 65536);
    GDSJ_TRY(E.sample.ensure((size_t)m * (MLP_F + 1) * 4 + 256));
    float *sx = E.sample.as<float>();
    float *st = sx + (size_t)m * MLP_F;
    k_gather_rows<<<ceil_div(m, 256), 256, 0, s>>>(feat, target, nt, m, sx, st);
    if (!E.init) {
        GDSJ_TRY(mlp_init(E.params.as<float>(), 0x5eed1234ull, s));
        // output bias starts at the mean target
        std::vector<float> ht(m);
        GDSJ_HIP(hipMemcpyAsync(ht.data(), st, (size_t)m * 4, hipMemcpyDeviceToHost, s));
        GDSJ_HIP(hipStreamSynchronize(s));
        double mean = 0;
        for (float v : ht) mean += v;
        float b3 = (float)(mean / m);
        GDSJ_HIP(hipMemcpyAsync(E.params.as<float>() + MLP_OFF_B3, &b3, 4, hipMemcpyHostToDevice, s));
        GDSJ_HIP(hipMemsetAsync(E.adam_m.p, 0, MLP_P * 4, s));
        GDSJ_HIP(hipMemsetAsync(E.adam_v.p, 0, MLP_P * 4, s));
        GDSJ_HIP(hipStreamSynchronize(s));
        E.init = true;
        E.step = 0;
    }
    // minibatch Adam: a fixed number of steps (first training: 600, later
    // joins: 150), cycling over the sample in a shuffled order per epoch
    const uint32_t batch = std::min<uint32_t>(2048, m);
    const uint32_t steps = E.trained ? 150 : 600;
    GDSJ_TRY(E.partial.ensure(mlp_grad_scratch_floats(batch) * 4 + (MLP_P + 8) * 4));
    float *partial = E.partial.as<float>();
    float *grad = partial + mlp_grad_scratch_floats(batch);
    const uint32_t pa = coprime_step(m);
    uint32_t b = 0, ep = 0;
    for (uint32_t it = 0; it < steps; it++) {
        if (b >= m) {
            b = 0;
            ep++;
        }
        const uint32_t pc = (uint32_t)((ep * 2654435761ull) % m);
        const uint32_t cnt = std::min(batch, m - b);
        GDSJ_TRY(mlp_grad_rows(E.params.as<float>(), sx, st, m, b, cnt, pa, pc, grad, partial, s));
        E.step++;
        const float lr = it < steps / 2 ? 3e-3f : 1e-3f;
        GDSJ_TRY(mlp_adam(E.params.as<float>(), grad, E.adam_m.as<float>(), E.adam_v.as<float>(), E.step, lr, s));
        b += cnt;
    }
    // slot headroom: 99th percentile of actual / predicted hits on the sample
    GDSJ_TRY(E.pred.ensure((size_t)std::max(nt, m) * 4));
    GDSJ_TRY(mlp_forward(E.params.as<float>(), sx, m, E.pred.as<float>(), s));
    std::vector<float> hp(m), ht(m);
    GDSJ_HIP(hipMemcpyAsync(hp.data(), E.pred.p, (size_t)m * 4, hipMemcpyDeviceToHost, s));
    GDSJ_HIP(hipMemcpyAsync(ht.data(), st, (size_t)m * 4, hipMemcpyDeviceToHost, s));
    GDSJ_HIP(hipEventRecord(e1, s));
    GDSJ_HIP(hipStreamSynchronize(s));
    std::vector<double> ratio(m);
    for (uint32_t i = 0; i < m; i++) {
        const double y = exp2((double)ht[i]) - 1.0, p = std::max(exp2((double)hp[i]) - 1.0, 1.0);
        ratio[i] = y / p;
    }
    const double rmax = *std::max_element(ratio.begin(), ratio.end());
    const size_t q = (size_t)(0.99 * (m - 1));
    std::nth_element(ratio.begin(), ratio.begin() + q, ratio.end());
    E.alpha = std::min(4.0, std::max(0.05, ratio[q] - 1.0));
    E.alpha_heavy = std::min(16.0, std::max(E.alpha, rmax - 1.0));
    E.trained = true;
    float t = 0;
    (void)hipEventElapsedTime(&t, e0, e1);
    if (ms) *ms = t;
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    return GDSJ_OK;
}

// Sum (or max) over a block of up to 1024 threads; the result in thread 0.
// One atomic per block instead of one per wave on the shared totals.
template <typename V, bool MAX>
__device__ __forceinline__ V block_reduce(V v) {
    for (int off = 32; off > 0; off >>= 1) {
        const V o = __shfl_xor(v, off, GDSJ_WAVE);
        v = MAX ? (o > v ? o : v) : v + o;
    }
    __shared__ V part[1024 / GDSJ_WAVE];
    const uint32_t w = threadIdx.x / GDSJ_WAVE, nw = (blockDim.x + GDSJ_WAVE - 1) / GDSJ_WAVE;
    if (lane_id() == 0) part[w] = v;
    __syncthreads();
    if (threadIdx.x == 0)
        for (uint32_t i = 1; i < nw; i++) v = MAX ? (part[i] > v ? part[i] : v) : v + part[i];
    __syncthreads();
    return v;
}

constexpr uint32_t RED_BLOCK = 1024;

// ------------------------------------------------------------- capacities --
// Slot capacity per tile: pred (1 + alpha) + beta, between nq and every
// candidate a hit.  Heavy tiles (at least HEAVY_PAIRS candidate pairs) take
// alpha_heavy: redoing one of them after an overflow costs a full pass over
// its candidates, while unused slots cost only memory.
constexpr double HEAVY_PAIRS = 1048576.0;

__global__ void k_capacity(const float *__restrict__ predLog2, const uint4 *__restrict__ tiles, uint32_t nt,
                           const uint32_t *__restrict__ cellCand, float alpha, float alpha_heavy, float beta,
                           double perQuery, uint32_t *__restrict__ cap, unsigned long long *__restrict__ predSum,
                           float *__restrict__ predOut) {
    const uint32_t t = blockIdx.x * blockDim.x + threadIdx.x;
    uint64_t mine = 0;
    if (t < nt) {
        const uint4 tl = tiles[t];
        const double nq = (double)(tl.z - tl.y);
        const double upper = nq * (double)cellCand[tl.x];  // every candidate a hit
        const double pred = predLog2 ? fmax(exp2((double)predLog2[t]) - 1.0, 0.0) : nq * perQuery;
        const double al = upper >= HEAVY_PAIRS ? (double)alpha_heavy : (double)alpha;
        double c = ceil(pred * (1.0 + al)) + beta;
        c = fmin(fmax(c, nq), upper);
        cap[t] = (uint32_t)c;
        mine = (uint64_t)pred;
        if (predOut) predOut[t] = (float)log2(1.0 + pred);
    }
    mine = block_reduce<uint64_t, false>(mine);
    if (threadIdx.x == 0 && mine) atomicAdd(predSum, (unsigned long long)mine);
}

int estimator_capacity_learned(Estimator &E, const Grid &G, uint32_t *capOut, uint64_t *predicted,
                               hipStream_t s) {
    const uint32_t nt = G.ntiles;
    GDSJ_TRY(E.feat.ensure((size_t)nt * MLP_F * 4 + 64));
    GDSJ_TRY(E.pred.ensure((size_t)nt * 4 + 64));
    GDSJ_TRY(E.work.ensure(64));
    GDSJ_TRY(tile_features(G, E.feat.as<float>(), s));
    GDSJ_TRY(mlp_forward(E.params.as<float>(), E.feat.as<float>(), nt, E.pred.as<float>(), s));
    GDSJ_HIP(hipMemsetAsync(E.work.p, 0, 8, s));
    if (nt)
        k_capacity<<<ceil_div(nt, RED_BLOCK), RED_BLOCK, 0, s>>>(E.pred.as<float>(), G.tiles.as<uint4>(), nt,
                                                     G.cellCand.as<uint32_t>(), (float)E.alpha,
                                                     (float)E.alpha_heavy, 8.0f, 0.0, capOut,
                                                     E.work.as<unsigned long long>(), nullptr);
    GDSJ_HIP(hipGetLastError());
    GDSJ_HIP(hipMemcpyAsync(predicted, E.work.p, 8, hipMemcpyDeviceToHost, s));
    return GDSJ_OK;
}

// sample tiles: one query each, the reference's sample ids 0, 66, 132, ...
// (kernel.cu:545-547 with QUERYREORDER=0; N*0.015 samples, GPU.cu:100)
__global__ void k_inverse_perm(const uint32_t *__restrict__ ids, uint32_t n, uint32_t *__restrict__ inv) {
    const uint32_t s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s < n) inv[ids[s]] = s;
}

__global__ void k_sample_tiles(const uint32_t *__restrict__ inv, const uint32_t *__restrict__ flag,
                               const uint32_t *__restrict__ cidx, uint32_t ns, uint32_t stride,
                               uint4 *__restrict__ tiles) {
    const uint32_t m = blockIdx.x * blockDim.x + threadIdx.x;
    if (m >= ns) return;
    const uint32_t sp = inv[(uint64_t)m * stride];
    const uint32_t c = cidx[sp] + flag[sp] - 1;
    tiles[m] = make_uint4(c, sp, sp + 1, 0u);
}

__global__ void k_sample_sum(const uint32_t *__restrict__ counts, uint32_t ns, uint32_t stride,
                             unsigned long long *__restrict__ out) {
    const uint32_t m = blockIdx.x * blockDim.x + threadIdx.x;
    uint64_t v = m < ns ? counts[(uint64_t)m * stride] : 0;
    for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, GDSJ_WAVE);
    if (lane_id() == 0 && v) atomicAdd(out, (unsigned long long)v);
}

int estimator_capacity_sampling(const Grid &G, const void *joinArgs, int mode, uint32_t *countsScratch,
                                uint32_t *capOut, uint64_t *predicted, float *predLog2Out, hipStream_t s,
                                DevBuf &scratch) {
    const uint32_t n = G.n, nt = G.ntiles;
    const double sampleRate = 0.015;             // SAMPLERATE, params.h:93
    const uint32_t stride = (uint32_t)(1.0 / sampleRate);  // offsetRate = 66, GPU.cu:86
    const uint32_t ns = (uint32_t)(n * sampleRate);        // N_batchEst, GPU.cu:100
    GDSJ_TRY(scratch.ensure((size_t)n * 4 + (size_t)ns * 16 + 256));
    uint32_t *inv = scratch.as<uint32_t>();
    uint4 *stiles = (uint4 *)(((uintptr_t)(inv + n) + 255) & ~(uintptr_t)255);
    unsigned long long *sum = (unsigned long long *)G.scalars.as<uint64_t>() + 2;
    GDSJ_HIP(hipMemsetAsync(sum, 0, 8, s));
    uint64_t est = 0;
    if (ns > 0) {
        k_inverse_perm<<<ceil_div(n, 256), 256, 0, s>>>(G.ids.as<uint32_t>(), n, inv);
        const uint32_t *flag = G.flag.as<uint32_t>();
        k_sample_tiles<<<ceil_div(ns, 256), 256, 0, s>>>(inv, flag, flag + n + 1, ns, stride, stiles);
        JoinArgs a = *(const JoinArgs *)joinArgs;
        a.tiles = stiles;
        a.tileList = nullptr;
        a.mjList = nullptr;
        a.ntiles = ns;
        a.counts = countsScratch;
        a.tileWork = nullptr;
        a.hmask = nullptr;
        GDSJ_TRY(join_pass(G, a, mode, PASS_COUNT, s));
        k_sample_sum<<<ceil_div(ns, 256), 256, 0, s>>>(countsScratch, ns, stride, sum);
        uint64_t cnt = 0;
        GDSJ_HIP(hipMemcpyAsync(&cnt, sum, 8, hipMemcpyDeviceToHost, s));
        GDSJ_HIP(hipStreamSynchronize(s));
        est = cnt * (uint64_t)stride;  // estimatedNeighbors, GPU.cu:214
    }
    *predicted = est;
    const double alpha = 0.10;  // PYTHON overestimation factor, GPU.cu:224-226
    GDSJ_HIP(hipMemsetAsync(sum, 0, 8, s));
    if (nt)
        k_capacity<<<ceil_div(nt, RED_BLOCK), RED_BLOCK, 0, s>>>(nullptr, G.tiles.as<uint4>(), nt, G.cellCand.as<uint32_t>(),
                                                     (float)alpha, (float)alpha, 8.0f, n ? (double)est / n : 0.0, capOut, sum,
                                                     predLog2Out);
    GDSJ_HIP(hipGetLastError());
    return GDSJ_OK;
}

// ----------------------------------------------------------------- error --
// one thread per tile: the tile's hits (a loop over its queries, 3 on
// average at 6 dims) against the model's prediction, in float; each block
// writes its sum and maximum to part[block] (no contended atomics), and one
// block reduces them
constexpr uint32_t EST_BLOCK = 256;

__global__ __launch_bounds__(EST_BLOCK) void k_est_error(const float *__restrict__ predLog2,
                                                         const uint4 *__restrict__ tiles, uint32_t nt,
                                                         const uint32_t *__restrict__ ids,
                                                         const uint32_t *__restrict__ counts,
                                                         float2 *__restrict__ part) {
    const uint32_t t = blockIdx.x * blockDim.x + threadIdx.x;
    float rel = 0.0f;
    if (t < nt) {
        const uint4 tl = tiles[t];
        uint32_t y = 0;
        for (uint32_t q = tl.y; q < tl.z; q++) y += counts[ids[q]];
        const float p = fmaxf(exp2f(predLog2[t]) - 1.0f, 0.0f);
        rel = fabsf(p - (float)y) / fmaxf((float)y, 1.0f);
    }
    float sm = rel, mx = rel;
    for (int off = 32; off > 0; off >>= 1) {
        sm += __shfl_xor(sm, off, GDSJ_WAVE);
        mx = fmaxf(mx, __shfl_xor(mx, off, GDSJ_WAVE));
    }
    __shared__ float ps[EST_BLOCK / GDSJ_WAVE], pm[EST_BLOCK / GDSJ_WAVE];
    const uint32_t w = threadIdx.x / GDSJ_WAVE;
    if (lane_id() == 0) {
        ps[w] = sm;
        pm[w] = mx;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        float a = 0.0f, m = 0.0f;
        for (uint32_t i = 0; i < EST_BLOCK / GDSJ_WAVE; i++) {
            a += ps[i];
            m = fmaxf(m, pm[i]);
        }
        part[blockIdx.x] = make_float2(a, m);
    }
}

__global__ __launch_bounds__(EST_BLOCK) void k_est_error_reduce(const float2 *__restrict__ part, uint32_t np,
                                                                double *__restrict__ out) {
    double sm = 0.0;
    float mx = 0.0f;
    for (uint32_t i = threadIdx.x; i < np; i += blockDim.x) {
        sm += (double)part[i].x;
        mx = fmaxf(mx, part[i].y);
    }
    for (int off = 32; off > 0; off >>= 1) {
        sm += __shfl_xor(sm, off, GDSJ_WAVE);
        mx = fmaxf(mx, __shfl_xor(mx, off, GDSJ_WAVE));
    }
    __shared__ double ps[EST_BLOCK / GDSJ_WAVE];
    __shared__ float pm[EST_BLOCK / GDSJ_WAVE];
    const uint32_t w = threadIdx.x / GDSJ_WAVE;
    if (lane_id() == 0) {
        ps[w] = sm;
        pm[w] = mx;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        double a = 0.0;
        float m = 0.0f;
        for (uint32_t i = 0; i < EST_BLOCK / GDSJ_WAVE; i++) {
            a += ps[i];
            m = fmaxf(m, pm[i]);
        }
        out[0] = a;
        out[1] = (double)m;
    }
}

int estimator_error(const Grid &G, const float *predLog2, const uint32_t *counts, double *meanRel, double *maxRel,
                    hipStream_t s, DevBuf &scratch) {
    const uint32_t nt = G.ntiles;
    *meanRel = *maxRel = 0.0;
    if (nt == 0) return GDSJ_OK;
    const uint32_t nb = ceil_div(nt, EST_BLOCK);
    GDSJ_TRY(scratch.ensure(64 + (size_t)nb * sizeof(float2)));
    double *out = scratch.as<double>();
    float2 *part = (float2 *)(scratch.as<char>() + 64);
    k_est_error<<<nb, EST_BLOCK, 0, s>>>(predLog2, G.tiles.as<uint4>(), nt, G.ids.as<uint32_t>(), counts, part);
    k_est_error_reduce<<<1, EST_BLOCK, 0, s>>>(part, nb, out);
    GDSJ_HIP(hipGetLastError());
    double h[2] = {0, 0};
    GDSJ_HIP(hipMemcpyAsync(h, out, 16, hipMemcpyDeviceToHost, s));
    GDSJ_HIP(hipStreamSynchronize(s));
    *meanRel = h[0] / nt;
    *maxRel = h[1];
    return GDSJ_OK;
}

}  // namespace gdsj
