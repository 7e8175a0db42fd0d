// gram.h -- MFMA Gram-form distance filter for dims >= 16.
//
// Every point x (dims permuted, sorted order) is stored once per join as a
// low-precision vector h = lp(sigma (x - o)) (fp16, or bf16 in two terms for
// FP32 data on request; o = per-dim mean, sigma a power of two keeping fp16
// in range)
// plus one float W = (1-c)/2 |h|^2 - (eps'/2 + e)^2, where e bounds |x' - h|
// (measured exactly per point) and c bounds the FP32 accumulation error of
// the MFMA dot products.  For a query q and candidate p the 32x32x16 MFMA
// gives S ~ h_q . h_p, and
//     S < W_q + W_p   =>   |x_q - x_p| > eps'   (no pair possible)
// so only pairs with S >= W_q + W_p are rechecked by direct differencing in
// DTYPE with the reference's accumulation order (kernel.cu:307-365).  The
// filter never decides a pair: the result equals the plain join's.
// For data near a low-dimensional subspace h may instead hold the point's
// coordinates on m principal axes, y = P (x - o): |P v| <= |P|_2 |v| for
// every v, so with eps' widened by |P|_2 and e by the rounding of the
// projection the same test keeps every pair within eps' (a filter on m
// instead of d dims: m / d of the operand bytes and MFMAs).
#pragma once
#include "common.h"
#include "grid.h"
#include "join.h"

namespace gdsj {

struct GramData {
    DevBuf h;        // [n][dpad] bf16 or fp16: leading part of sigma (x - o)
    DevBuf l;        // [n][dpad] the remainder in the same format (split form only)
    DevBuf w;        // [n] float
    DevBuf n2;       // [n] double: N = |h + l|^2 (filter-error measurement)
    DevBuf scratch;  // per-dim means [d], max deviations [d], sigma
    DevBuf xr;       // [n][xrs] DTYPE: the sorted points row-major (the exact rechecks' loads)
    DevBuf pm;       // [axes][d] double: the principal axes h is taken on (axes > 0)
    uint32_t axes = 0;  // 0: h holds the points' own dims; else y = P (x - o) on this many axes
    uint32_t xrs = 0;  // row stride: dims rounded up to a multiple of 16
    uint32_t dpad = 0;
    int lowp = 0;    // 0 bf16, 1 fp16
    int split = 0;   // 1: S = h.h' + h.l' + l.h' (three MFMAs per k-step)
    void release();
};

// dims the filter handles (padded to a multiple of 16)
bool gram_supported(uint32_t d);
// queries per Gram tile (a multiple of 32) for dims d and operand format lowp
uint32_t gram_tile_queries(uint32_t d, int lowp);
// Builds h and w for the grid's points; lowp: 0 bf16, 1 fp16.
int gram_prepare(GramData &gd, const Grid &G, int lowp, hipStream_t s);
// One pass (JoinPass) of the filtered join.  stats[0] += rechecked pairs,
// stats[1] += rechecked pairs that are hits; with measure_err, stats[2] =
// max over rechecked pairs of |Gram estimate - FP64 squared distance| / eps^2
// (double bits).
int gram_join_pass(const Grid &G, const GramData &gd, const JoinArgs &a, int mode, int pass,
                   unsigned long long *stats, hipStream_t s, bool measure_err = false);

}  // namespace gdsj
