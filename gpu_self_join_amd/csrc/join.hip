// join.hip -- dispatch of the join pass by DTYPE.
#include <math.h>

#include "join.h"

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

const char *join_variant(int dtype, uint32_t d, int mode, uint32_t ilp) {
    bool sp = dtype == GDSJ_DTYPE_F32 ? join_specialised_f32(d, mode, ilp) : join_specialised_f64(d, mode, ilp);
    return sp ? "k_join<register-resident>" : "k_join<generic>";
}

}  // namespace gdsj
