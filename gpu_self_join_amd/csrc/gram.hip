// gram.hip -- MFMA Gram-form filter join (see gram.h for the bound).
//
// One wave per tile of <= 32 queries of one cell.  The queries' h vectors
// are the A operand of v_mfma_f32_32x32x16_{bf16,f16} (lane l: row l%32,
// k = 8*(l/32)..+7 of each 16-wide k-step); 32 candidates at a time are the
// B operand (lane l: column l%32).  Output register i of lane l holds
// S[row (i/4)*8 + (l/32)*4 + i%4][column l%32].  Survivors of
// S >= W_q + W_p are rechecked exactly; hits are counted / written per row
// with the ballot + mbcnt scheme of the plain join, in candidate order, so
// the rows come out identical to the plain join's.
#include <math.h>

#include "gram.h"
#include "join_kernels.h"
#include "neighbors.h"

namespace gdsj {

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef _Float16 f16x8 __attribute__((ext_vector_type(8)));
typedef float f32x16 __attribute__((ext_vector_type(16)));

template <int LP> struct Lp;
template <> struct Lp<0> {
    using scalar = __bf16;
    using vec = bf16x8;
    static constexpr double min_normal = 0x1p-126;
};
template <> struct Lp<1> {
    using scalar = _Float16;
    using vec = f16x8;
    static constexpr double min_normal = 0x1p-14;
};

template <int LP>
__device__ __forceinline__ f32x16 mfma32(typename Lp<LP>::vec a, typename Lp<LP>::vec b, f32x16 c) {
    if constexpr (LP == 0)
        return __builtin_amdgcn_mfma_f32_32x32x16_bf16(a, b, c, 0, 0, 0);
    else
        return __builtin_amdgcn_mfma_f32_32x32x16_f16(a, b, c, 0, 0, 0);
}

template <int LP>
__device__ __forceinline__ typename Lp<LP>::vec load8(const typename Lp<LP>::scalar *p) {
    const uint4 raw = *reinterpret_cast<const uint4 *>(p);
    return __builtin_bit_cast(typename Lp<LP>::vec, raw);
}

void GramData::release() {
    h.release();
    w.release();
    scratch.release();
    dpad = 0;
}

static uint32_t pad_dims(uint32_t d) {
    const uint32_t ks = (d + 15) / 16;
    const uint32_t kp = ks <= 4 ? ks : (ks <= 6 ? 6 : 8);
    return kp * 16;
}

bool gram_supported(uint32_t d) { return d >= 16 && d <= 128; }

// ------------------------------------------------------------- preparation --
template <typename T>
__global__ __launch_bounds__(256) void k_dim_mean(const T *__restrict__ xs, uint64_t ld, uint32_t n,
                                                   double *__restrict__ mean) {
    const uint32_t l = blockIdx.x;
    double s = 0.0;
    for (uint32_t i = threadIdx.x; i < n; i += blockDim.x) s += (double)xs[(uint64_t)l * ld + i];
    __shared__ double red[256];
    red[threadIdx.x] = s;
    __syncthreads();
    for (uint32_t o = 128; o > 0; o >>= 1) {
        if (threadIdx.x < o) red[threadIdx.x] += red[threadIdx.x + o];
        __syncthreads();
    }
    if (threadIdx.x == 0) mean[l] = n ? red[0] / n : 0.0;
}

template <typename T>
__global__ __launch_bounds__(256) void k_dim_maxdev(const T *__restrict__ xs, uint64_t ld, uint32_t n,
                                                     const double *__restrict__ mean, double *__restrict__ maxdev) {
    const uint32_t l = blockIdx.x;
    const double o = mean[l];
    double m = 0.0;
    for (uint32_t i = threadIdx.x; i < n; i += blockDim.x) m = fmax(m, fabs((double)xs[(uint64_t)l * ld + i] - o));
    __shared__ double red[256];
    red[threadIdx.x] = m;
    __syncthreads();
    for (uint32_t q = 128; q > 0; q >>= 1) {
        if (threadIdx.x < q) red[threadIdx.x] = fmax(red[threadIdx.x], red[threadIdx.x + q]);
        __syncthreads();
    }
    if (threadIdx.x == 0) maxdev[l] = red[0];
}

// h = lp(sigma (x - o)) (subnormals flushed to zero here, so the MFMA sees
// exactly the values the bound was computed from); e = |sigma (x - o) - h|
// measured in double and widened for the double rounding of sigma (x - o);
// W = (1-c)/2 |h|^2 - (1 + 2^-20)(sigma eps'/2 + e)^2, rounded down.
template <typename T, int LP>
__global__ void k_gram_convert(const T *__restrict__ xs, uint64_t ld, uint32_t n, uint32_t d, uint32_t dpad,
                               const double *__restrict__ mean, const double *__restrict__ maxdev, double eps_p,
                               double cfac, typename Lp<LP>::scalar *__restrict__ H, float *__restrict__ W) {
    const uint32_t s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= n) return;
    double sigma = 1.0;
    if (LP == 1) {
        double mx = 0.0;
        for (uint32_t l = 0; l < d; l++) mx = fmax(mx, maxdev[l]);
        if (mx > 0.0) sigma = exp2(floor(log2(16384.0 / mx)));
        sigma = fmin(fmax(sigma, 0x1p-60), 0x1p60);
    }
    double N = 0.0, E2 = 0.0, X2 = 0.0;
    typename Lp<LP>::scalar *hp = H + (uint64_t)s * dpad;
    for (uint32_t l = 0; l < dpad; l++) {
        double xp = 0.0;
        if (l < d) xp = ((double)xs[(uint64_t)l * ld + s] - mean[l]) * sigma;
        typename Lp<LP>::scalar h = (typename Lp<LP>::scalar)(float)xp;
        double hd = (double)(float)h;
        if (fabs(hd) < Lp<LP>::min_normal) {
            hd = 0.0;
            h = (typename Lp<LP>::scalar)0.0f;
        }
        hp[l] = h;
        const double r = xp - hd;
        N += hd * hd;
        E2 += r * r;
        X2 += xp * xp;
    }
    const double e = sqrt(E2) * (1.0 + 0x1p-30) + sqrt(X2) * 0x1p-50;
    const double k = sigma * eps_p * 0.5 + e;
    const double w = 0.5 * (1.0 - cfac) * N - (1.0 + 0x1p-20) * k * k;
    W[s] = __double2float_rd(w);
}

template <typename T, int LP>
static int prepare_t(GramData &gd, const Grid &G, hipStream_t s) {
    const uint32_t n = G.n, d = G.d, dpad = gd.dpad;
    GDSJ_TRY(gd.h.ensure((size_t)n * dpad * 2 + 256));
    GDSJ_TRY(gd.w.ensure((size_t)n * 4 + 256));
    GDSJ_TRY(gd.scratch.ensure((size_t)d * 16 + 256));
    double *mean = gd.scratch.as<double>();
    double *maxdev = mean + d;
    const T *xs = G.xs.as<T>();
    k_dim_mean<T><<<d, 256, 0, s>>>(xs, G.ld, n, mean);
    k_dim_maxdev<T><<<d, 256, 0, s>>>(xs, G.ld, n, mean, maxdev);
    // eps' = eps (1 + 2 (D+5) u_T): beyond it the DTYPE sum cannot pass the predicate
    const double uT = sizeof(T) == 4 ? 0x1p-24 : 0x1p-53;
    const double eps_p = G.eps * (1.0 + 2.0 * (d + 5) * uT);
    // MFMA FP32 accumulation of dpad exact products plus the stored |h|^2,
    // with margin for the FP32 rounding of W_q + W_p
    const double cfac = (2.0 * dpad + 16.0) * 0x1p-24;
    k_gram_convert<T, LP><<<ceil_div(n, 256), 256, 0, s>>>(xs, G.ld, n, d, dpad, mean, maxdev, eps_p, cfac,
                                                           (typename Lp<LP>::scalar *)gd.h.p, gd.w.as<float>());
    GDSJ_HIP(hipGetLastError());
    return GDSJ_OK;
}

int gram_prepare(GramData &gd, const Grid &G, int lowp, hipStream_t s) {
    if (!gram_supported(G.d)) {
        set_error("the Gram filter handles 16 to 128 dims");
        return GDSJ_EUNSUPPORTED;
    }
    gd.dpad = pad_dims(G.d);
    gd.lowp = lowp;
    if (G.n == 0) return GDSJ_OK;
    if (G.dtype == GDSJ_DTYPE_F32) return lowp ? prepare_t<float, 1>(gd, G, s) : prepare_t<float, 0>(gd, G, s);
    return lowp ? prepare_t<double, 1>(gd, G, s) : prepare_t<double, 0>(gd, G, s);
}

// ---------------------------------------------------------------- the join --
template <typename T, int LP, int KS, int PASS>
__global__ __launch_bounds__(256) void k_gram(const T *__restrict__ xs, const uint32_t *__restrict__ ids,
                                              const typename Lp<LP>::scalar *__restrict__ H,
                                              const float *__restrict__ W, const JoinArgs a, int mode,
                                              unsigned long long *__restrict__ stats) {
    using vec = typename Lp<LP>::vec;
    constexpr uint32_t DP = KS * 16;
    const uint32_t lane = lane_id();
    const uint32_t w = threadIdx.x / GDSJ_WAVE;
    const uint32_t lblock = a.xcd_map ? xcd_block(blockIdx.x, gridDim.x) : blockIdx.x;
    const uint32_t tslot = uniform((uint32_t)(lblock * (blockDim.x / GDSJ_WAVE) + w));
    if (tslot >= a.ntiles) return;
    const uint32_t tile = a.tileList ? uniform(a.tileList[tslot]) : tslot;
    const uint64_t ld = a.ld;
    const T thr = (T)a.thr;
    const uint4 tl = a.tiles[tile];
    const uint32_t cell = uniform(tl.x), q0 = uniform(tl.y), nq = uniform(tl.z - tl.y);
    const GridArgs &ga = a.ga;
    const bool pre = a.runs != nullptr;
    const uint32_t row_l = lane & 31u, half = lane >> 5;
    const uint64_t ltm = lane == 0 ? 0ull : (~0ull >> (64 - lane));

    __shared__ double s_gap[256 / GDSJ_WAVE][2][MAXK];
    __shared__ uint32_t s_cnt[256 / GDSJ_WAVE][32];
    __shared__ unsigned long long s_cur[256 / GDSJ_WAVE][32];
    double *glo = s_gap[w][0];
    double *ghi = s_gap[w][1];
    uint32_t cc[MAXK];
    uint32_t nwin;
    if (pre) {
        nwin = uniform(a.nruns[cell]);
    } else {
        nwin = a.ncombo;
        for (uint32_t j = 0; j < ga.k; j++) cc[j] = uniform(a.cellCoord[(uint64_t)j * ga.ncell + cell]);
        if (a.prune) {
            tile_gaps<T>(xs, ld, q0, nq, cc, ga, glo, ghi);
        } else {
            if (lane < MAXK) {
                glo[lane] = 0.0;
                ghi[lane] = 0.0;
            }
            __builtin_amdgcn_wave_barrier();
        }
    }
    if (lane < 32) {
        s_cnt[w][lane] = 0;
        s_cur[w][lane] = (PASS == PASS_FILL && lane < nq) ? a.offsets[ids[q0 + lane]] : 0ull;
    }
    __builtin_amdgcn_wave_barrier();

    vec A[KS];
#pragma unroll
    for (int ks = 0; ks < KS; ks++) {
        if (row_l < nq)
            A[ks] = load8<LP>(H + (uint64_t)(q0 + row_l) * DP + ks * 16 + half * 8);
        else
            A[ks] = (vec){};
    }
    float wq[16];
#pragma unroll
    for (int i = 0; i < 16; i++) {
        const uint32_t r = (i / 4) * 8 + half * 4 + (i % 4);
        wq[i] = r < nq ? W[q0 + r] : INFINITY;
    }
    uint32_t scur = 0, scap = 0;
    uint64_t sbase = 0;
    if (PASS == PASS_SLOT) {
        scap = uniform(a.slotCap[tile]);
        sbase = uniform64(a.slotOff[tile]);
    }
    uint64_t work = 0;
    uint32_t nrech = 0, nhit = 0;
    bool overflow = false;

    for (uint32_t cw = 0; cw < nwin; cw += GDSJ_WAVE) {
        const uint32_t nrw = uniform(nwin - cw < (uint32_t)GDSJ_WAVE ? nwin - cw : (uint32_t)GDSJ_WAVE);
        uint2 run = make_uint2(0u, 0u);
        if (lane < nrw)
            run = pre ? a.runs[(uint64_t)cell * a.runStride + cw + lane]
                      : combo_run(cw + lane, cc, glo, ghi, ga, a.cellStart, a.hkeys, a.hvals);
        const uint32_t len = run.y - run.x;
        uint32_t incl = len;
#pragma unroll
        for (int off = 1; off < GDSJ_WAVE; off <<= 1) {
            const uint32_t o = __shfl_up(incl, off, GDSJ_WAVE);
            if (lane >= (uint32_t)off) incl += o;
        }
        const uint32_t total = uniform(shfl_u32(incl, nrw - 1));
        if (total == 0) continue;
        work += total;
        const uint32_t rexcl = incl - len;
        const uint32_t rstart = run.x;
        const uint32_t fpos = uniform(cand_pos(0u, nrw, rstart, rexcl));
        for (uint32_t cb = 0; cb < total; cb += 32) {
            const uint32_t g = cb + row_l;
            const bool valid = g < total;
            uint32_t sp = cand_pos(valid ? g : 0u, nrw, rstart, rexcl);
            sp = valid ? sp : fpos;
            f32x16 acc = {};
#pragma unroll
            for (int ks = 0; ks < KS; ks++) {
                const vec B = load8<LP>(H + (uint64_t)sp * DP + ks * 16 + half * 8);
                acc = mfma32<LP>(A[ks], B, acc);
            }
            const float wc = valid ? W[sp] : INFINITY;
            uint32_t smask = 0;
#pragma unroll
            for (int i = 0; i < 16; i++) smask |= (acc[i] >= wq[i] + wc) ? (1u << i) : 0u;
            if (ballot(smask != 0) == 0) continue;
            const uint32_t cid = ids[sp];
#pragma unroll
            for (int i = 0; i < 16; i++) {
                const bool surv = (smask >> i) & 1u;
                if (ballot(surv) == 0) continue;
                const uint32_t r = (i / 4) * 8 + half * 4 + (i % 4);
                bool hit = false;
                if (surv) {
                    nrech++;
                    const T e = dist_mem<T>(xs, ld, q0 + r, sp, a.d, mode, a.ilp);
                    hit = e <= thr;
                    nhit += hit ? 1u : 0u;
                }
                const uint64_t m = ballot(hit);
                if (m == 0) continue;
                const uint32_t hm = half ? (uint32_t)(m >> 32) : (uint32_t)m;
                const uint32_t hlt = half ? (uint32_t)((m & ltm) >> 32) : (uint32_t)(m & ltm);
                if (PASS == PASS_FILL) {
                    const uint64_t pos = s_cur[w][r] + (uint64_t)__popc(hlt);
                    if (hit) {
                        if (pos < a.cap) a.out[pos] = cid; else overflow = true;
                    }
                    __builtin_amdgcn_wave_barrier();
                    if (row_l == 0) s_cur[w][r] += (unsigned long long)__popc(hm);
                } else {
                    if (row_l == 0) s_cnt[w][r] += (uint32_t)__popc(hm);
                    if (PASS == PASS_SLOT) {
                        const uint32_t p = scur + mbcnt(m);
                        if (hit && p < scap) {
                            a.slotIds[sbase + p] = cid;
                            a.slotQ[sbase + p] = (uint8_t)r;
                        }
                        scur += popc64(m);
                    }
                }
                __builtin_amdgcn_wave_barrier();
            }
        }
    }
    __builtin_amdgcn_wave_barrier();
    if (PASS != PASS_FILL) {
        if (lane < nq) a.counts[ids[q0 + lane]] = s_cnt[w][lane];
        if (a.tileWork && lane == 0) a.tileWork[tile] = work * nq;
    }
    if (PASS == PASS_SLOT && lane == 0) {
        a.slotUsed[tile] = scur;
        if (scur > scap) a.ovfList[atomicAdd(a.ovfCount, 1u)] = tile;
    }
    if (PASS == PASS_FILL && overflow) atomicOr(a.status, 1u);
    if (stats) {
        unsigned long long r1 = nrech, r2 = nhit;
#pragma unroll
        for (int off = 32; off > 0; off >>= 1) {
            r1 += __shfl_xor(r1, off, GDSJ_WAVE);
            r2 += __shfl_xor(r2, off, GDSJ_WAVE);
        }
        if (lane == 0 && r1) {
            atomicAdd(&stats[0], r1);
            atomicAdd(&stats[1], r2);
        }
    }
}

template <typename T, int LP, int KS>
static int launch_gram(const GramData &gd, const JoinArgs &a, int mode, int pass, unsigned long long *stats,
                       hipStream_t s) {
    const uint32_t blocks = (a.ntiles + 3) / 4;
    if (blocks == 0) return GDSJ_OK;
    const T *xs = (const T *)a.xs;
    const auto *H = (const typename Lp<LP>::scalar *)gd.h.p;
    const float *W = gd.w.as<float>();
    if (pass == PASS_FILL)
        k_gram<T, LP, KS, PASS_FILL><<<blocks, 256, 0, s>>>(xs, a.ids, H, W, a, mode, stats);
    else if (pass == PASS_SLOT)
        k_gram<T, LP, KS, PASS_SLOT><<<blocks, 256, 0, s>>>(xs, a.ids, H, W, a, mode, stats);
    else
        k_gram<T, LP, KS, PASS_COUNT><<<blocks, 256, 0, s>>>(xs, a.ids, H, W, a, mode, stats);
    GDSJ_HIP(hipGetLastError());
    return GDSJ_OK;
}

template <typename T, int LP>
static int gram_dispatch(const GramData &gd, const JoinArgs &a, int mode, int pass, unsigned long long *stats,
                         hipStream_t s) {
    switch (gd.dpad / 16) {
        case 1: return launch_gram<T, LP, 1>(gd, a, mode, pass, stats, s);
        case 2: return launch_gram<T, LP, 2>(gd, a, mode, pass, stats, s);
        case 3: return launch_gram<T, LP, 3>(gd, a, mode, pass, stats, s);
        case 4: return launch_gram<T, LP, 4>(gd, a, mode, pass, stats, s);
        case 6: return launch_gram<T, LP, 6>(gd, a, mode, pass, stats, s);
        case 8: return launch_gram<T, LP, 8>(gd, a, mode, pass, stats, s);
        default: break;
    }
    set_error("unsupported padded dims %u for the Gram filter", gd.dpad);
    return GDSJ_EUNSUPPORTED;
}

int gram_join_pass(const Grid &G, const GramData &gd, const JoinArgs &a, int mode, int pass,
                   unsigned long long *stats, hipStream_t s) {
    if (G.tileq > 32) {
        set_error("the Gram filter needs tiles of at most 32 queries");
        return GDSJ_EINVAL;
    }
    // FP32 data uses bf16, FP64 data fp16 (the filter's operand format)
    if (G.dtype == GDSJ_DTYPE_F32)
        return gd.lowp ? gram_dispatch<float, 1>(gd, a, mode, pass, stats, s)
                       : gram_dispatch<float, 0>(gd, a, mode, pass, stats, s);
    return gd.lowp ? gram_dispatch<double, 1>(gd, a, mode, pass, stats, s)
                   : gram_dispatch<double, 0>(gd, a, mode, pass, stats, s);
}

}  // namespace gdsj
