// join.hip -- dispatch of the join pass by DTYPE.
#include <math.h>

#include "join.h"
#include "scan.h"

namespace gdsj {

int join_pass_f32(const Grid &G, const JoinArgs &a, int mode, int pass, hipStream_t s);
int join_pass_f64(const Grid &G, const JoinArgs &a, int mode, int pass, hipStream_t s);
bool join_specialised_f32(uint32_t d, int mode, uint32_t ilp);
bool join_specialised_f64(uint32_t d, int mode, uint32_t ilp);

int join_pass(const Grid &G, const JoinArgs &a, int mode, int pass, hipStream_t s) {
    if (G.dtype == GDSJ_DTYPE_F32) return join_pass_f32(G, a, mode, pass, s);
    return join_pass_f64(G, a, mode, pass, s);
}

// The Gram-form filter's band.  A query q and a candidate c of one tile
// are taken relative to the tile's first query o (a = fl(c - o),
// b = fl(q - o)); every dim is indexed, so |c - o| <= Rc = 2 eps and
// |q - o| <= Rq = eps up to the rounding of the cell coordinates (slack).
// The kernel forms g = fl(|a|^2 - 2 a.b) by FMAs and compares it with
// fl(lo - |b|^2) and fl(hi - |b|^2).  With u the unit roundoff and
// gamma(m) = m u / (1 - m u), the differences between
//   the reference's acc (rounded differences and sums, kernel.cu:307-357),
//   the exact sum of squared differences X,
//   the exact Y = |a - b|^2 of the rounded local coordinates,
//   and the computed g + |b|^2
// are bounded by
//   |acc - X|        <= gamma(D+3) X
//   |Y - X|          <= 2.01 u D (Rc + Rq)^2
//   |g - (Y - |b|^2)| <= gamma(D) (2.01 D Rc^2 + 2 D Rc Rq)
//   |qn - |b|^2|     <= gamma(D) D Rq^2,  threshold rounding <= u (thr + D Rq^2)
// so g <= fl(lo - |b|^2) implies acc <= thr and g > fl(hi - |b|^2) implies
// acc > thr for lo = thr - E, hi = thr + E with E twice the sum of these
// bounds at X ~ thr.  Pairs between the two are decided by the reference's
// arithmetic in the kernel.
bool gram_valu_band(const Grid &G, double thr, float *lo, float *mid, float *half) {
    if (G.dtype != GDSJ_DTYPE_F32 || G.k != G.d || G.d < 3 || G.d > 8) return false;
    const double u = 0x1p-24, eps = G.eps, D = (double)G.d;
    double xmax = 0.0;
    for (uint32_t j = 0; j < G.k; j++) {
        xmax = fmax(xmax, fabs(G.minArr[j]));
        xmax = fmax(xmax, fabs(G.minArr[j] + (double)G.ga.nCells[j] * eps));
    }
    const double slack = 16.0 * u * (xmax + 2.0 * eps);
    if (!(slack < 0.01 * eps)) return false;  // cells too fine for the float grid
    const double rc = 2.0 * eps + 2.0 * slack, rq = eps + 2.0 * slack;
    auto gamma = [&](double m) { return m * u / (1.0 - m * u); };
    const double e = gamma(D + 3) * thr * 1.01 + 2.01 * u * D * (rc + rq) * (rc + rq) +
                     gamma(D) * (2.01 * D * rc * rc + 2.0 * D * rc * rq) + gamma(D) * D * rq * rq +
                     u * (thr + D * rq * rq);
    const double E = 2.0 * e;
    // round outward to float; the kernel also subtracts |q'|^2 from lo and
    // mid in float, so the half-width gets a relative margin that covers
    // those roundings and that of g - mid
    float l = (float)(thr - E);
    if ((double)l > thr - E) l = nextafterf(l, -INFINITY);
    const float m = (float)thr;
    const double hw = ((double)m - (double)l) * (1.0 + 1e-3) + 4.0 * u * (thr + 9.0 * D * eps * eps);
    float hf = (float)hw;
    if ((double)hf < hw) hf = nextafterf(hf, INFINITY);
    *lo = l;
    *mid = m;
    *half = hf;
    return true;
}

// The matrix-core join's band.  A query q and a candidate c of one tile are
// taken relative to the tile's first query o (every dim indexed, so
// |q - o| <= rq = eps and |c - o| <= rc = 2 eps up to the rounding of the
// cell coordinates, slack): a = fl32(fl(q - o)), b = fl32(fl(c - o)), each
// within 2u of the exact difference (u = 2^-24 >= the DTYPE's roundoff)
// plus 2^-126 where the matrix core flushes a subnormal.  The kernel forms
// na = fl |a|^2, nb = fl |b|^2 and F, the MFMA's sum of the K = D + 2
// products (-2 a_j) b_j, 1 * nb and na * 1.  With X the exact squared
// distance and acc the reference's sum (|acc - X| <= g X, g = gamma(D+3),
// kernel.cu:307-357):
//   |(a - b)^2 - X|                    <= E1 = 2 em D (rq + rc) + D em^2,  em = 2u (rq + rc) + 2^-125
//   |F - (na + nb - 2 a.b)|            <= gamma(K+1) 2.01 D (Aq^2 + Ac^2) + K 2^-126
//   |na - |a|^2| + |nb - |b|^2|        <= gamma(D) D (Aq^2 + Ac^2)
// so |F - X| <= E (their sum).  acc <= thr implies X <= thr / (1 - g) and
// F <= thr / (1 - g) + E <= hi; F <= lo implies X <= lo + E <=
// thr / (1 + g) and acc <= thr.  Both margins are taken twice.
bool mjoin_band(const Grid &G, double thr, float *hi, float *lo) {
    if (G.k != G.d || G.d < 3 || G.d > 7) return false;
    const double u = 0x1p-24, eps = G.eps, D = (double)G.d;
    const double K = 2.0 * (double)((G.d + 3) / 2);
    if (!(eps > 0x1p-40) || !(eps < 0x1p40)) return false;  // squares stay normal floats
    double xmax = 0.0;
    for (uint32_t j = 0; j < G.k; j++) {
        xmax = fmax(xmax, fabs(G.minArr[j]));
        xmax = fmax(xmax, fabs(G.minArr[j] + (double)G.ga.nCells[j] * eps));
    }
    const double slack = 16.0 * u * (xmax + 2.0 * eps);
    if (!(slack < 0.01 * eps)) return false;  // cells too fine for float coordinates
    const double rq = eps + 2.0 * slack, rc = 2.0 * eps + 2.0 * slack;
    const double aq = rq * (1.0 + 2.0 * u) + 0x1p-126, ac = rc * (1.0 + 2.0 * u) + 0x1p-126;
    auto gamma = [&](double m) { return m * u / (1.0 - m * u); };
    const double em = 2.0 * u * (rq + rc) + 0x1p-125;
    const double e1 = 2.0 * em * D * (rq + rc) + D * em * em;
    const double e2 = gamma(K + 1.0) * 2.01 * D * (aq * aq + ac * ac) + K * 0x1p-126;
    const double e3 = gamma(D) * D * (aq * aq + ac * ac);
    const double E = e1 + e2 + e3;
    const double g = gamma(D + 3.0);
    const double th = thr / (1.0 - g) + 2.0 * E;
    const double tl = thr / (1.0 + g) - 2.0 * E;
    if (!(th - tl < 0.05 * thr) || !(tl > 0.0)) return false;
    float hf = (float)th;
    if ((double)hf < th) hf = nextafterf(hf, INFINITY);
    float lf = (float)tl;
    if ((double)lf > tl) lf = nextafterf(lf, -INFINITY);
    *hi = hf;
    *lo = lf;
    return true;
}

// The FMA-chain filter's band.  Both the filter's f (a product, then one
// FMA per dim: every term t_j^2 of S = sum t_j^2, t_j = fl(c_j - q_j), meets
// at most n = d + ilp roundings) and the reference's acc (kernel.cu:307-357:
// products and adds, fused first dims, ILP partials and their fold; again at
// most n roundings per term) satisfy |x - S| <= g S + n eta, g = gamma(n),
// eta the spacing of subnormals, so
//   f <= lo = (thr - n eta)(1 - g)/(1 + g) - n eta  implies acc <= thr,
//   f >  hi = (thr + n eta)(1 + g)/(1 - g) + n eta  implies acc >  thr.
// g is doubled (plus 8 u) and lo / hi are rounded outward to DTYPE for the
// host's own roundings.  Partial sums only grow in the reference's arithmetic, so a
// partial f above hi also decides the early exit of SHORTCIRCUIT.
void fma_band(int dtype, uint32_t d, uint32_t ilp, double thr, uint32_t mode, double *lo, double *hi) {
    if (mode == 2 || !(thr >= 0.0) || !(thr < 1e300)) {
        // every pair goes to the reference's arithmetic
        *lo = -1.0;
        *hi = __builtin_inf();
        return;
    }
    const bool f32 = dtype == GDSJ_DTYPE_F32;
    const double u = f32 ? 0x1p-24 : 0x1p-53, eta = f32 ? 0x1p-149 : 0x1p-1074;
    const double n = (double)d + (double)ilp + 1.0;
    const double g = 2.0 * (n * u / (1.0 - n * u)) + 8.0 * u;
    double l = (thr - n * eta) * (1.0 - g) / (1.0 + g) - n * eta;
    double h = (thr + n * eta) * (1.0 + g) / (1.0 - g) + n * eta;
    if (f32) {
        float lf = (float)l, hf = (float)h;
        if ((double)lf > l) lf = nextafterf(lf, -INFINITY);
        if ((double)hf < h) hf = nextafterf(hf, INFINITY);
        l = lf;
        h = hf;
    } else {
        l = nextafter(l, -INFINITY);
        h = nextafter(h, INFINITY);
    }
    *lo = l < 0.0 ? -1.0 : l;
    *hi = h;
}

__global__ void k_mj_flags(const uint4 *__restrict__ tiles, uint32_t nt, const uint32_t *__restrict__ cellStart,
                           uint32_t minOcc, uint32_t *__restrict__ flags) {
    const uint32_t t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= nt) return;
    const uint32_t c = tiles[t].x;
    flags[t] = cellStart[c + 1] - cellStart[c] >= minOcc ? 1u : 0u;
}

__global__ void k_mj_scatter(const uint32_t *__restrict__ flags, const uint32_t *__restrict__ pre, uint32_t nt,
                             uint32_t *__restrict__ list) {
    const uint32_t t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= nt) return;
    const uint32_t nbig = pre[nt];
    list[flags[t] ? pre[t] : nbig + (t - pre[t])] = t;
}

int mjoin_partition(const Grid &G, uint32_t minOcc, uint32_t *list, uint32_t *flags, void *tmp, uint32_t *nbig,
                    hipStream_t s) {
    const uint32_t nt = G.ntiles;
    *nbig = 0;
    if (nt == 0) return GDSJ_OK;
    // flags[0 .. nt), their exclusive prefix (and the total) behind them
    uint32_t *pre = flags + nt + 1;
    k_mj_flags<<<ceil_div(nt, 256), 256, 0, s>>>(G.tiles.as<uint4>(), nt, G.cellStart.as<uint32_t>(), minOcc, flags);
    GDSJ_TRY(exclusive_scan_u32_u32(flags, pre, nt, true, tmp, s));
    k_mj_scatter<<<ceil_div(nt, 256), 256, 0, s>>>(flags, pre, nt, list);
    GDSJ_HIP(hipGetLastError());
    GDSJ_HIP(hipMemcpyAsync(nbig, pre + nt, 4, hipMemcpyDeviceToHost, s));
    GDSJ_HIP(hipStreamSynchronize(s));
    return GDSJ_OK;
}

const char *join_variant(int dtype, uint32_t d, int mode, uint32_t ilp) {
    bool sp = dtype == GDSJ_DTYPE_F32 ? join_specialised_f32(d, mode, ilp) : join_specialised_f64(d, mode, ilp);
    return sp ? "k_join<register-resident>" : "k_join<generic>";
}

}  // namespace gdsj
