// bruteforce.h -- the all-pairs self-join (the reference's search mode 1).
#pragma once
#include "common.h"

namespace gdsj {

// SoA copy of row-major points: xs[j * ld + i] = in[i * d + j] (ld >= n).
int bruteforce_soa(const void *in, uint32_t n, uint32_t d, int dtype, uint64_t ld, void *xs, hipStream_t s);

// One pass over every ordered pair (pass: PASS_COUNT writes counts[n],
// PASS_FILL writes each row at offsets[q], ascending).  mode / ilp select the
// accumulation order (AccMode), thr the predicate acc <= thr.
int bruteforce_pass(const void *xs, uint64_t ld, uint32_t n, uint32_t d, int dtype, int mode, uint32_t ilp,
                    double thr, int pass, uint32_t *counts, const uint64_t *offsets, uint32_t *out, hipStream_t s);

}  // namespace gdsj
