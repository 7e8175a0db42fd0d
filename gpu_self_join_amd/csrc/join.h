// join.h -- the distance join over work tiles (count pass and fill pass).
#pragma once
#include "common.h"
#include "grid.h"

namespace gdsj {

// accumulation modes of the reference's distance loop (kernel.cu:307-357)
enum AccMode : int {
    ACC_SEQ = 0,       // ILP == 0: one FMA accumulator, dims in order
    ACC_ILP_SC = 1,    // ILP > 0, SHORTCIRCUIT == 1: fold after every ILP dims
    ACC_ILP_NOSC = 2,  // ILP > 0, SHORTCIRCUIT == 0: ILP partials, fold at end
};

struct JoinArgs {
    const void *xs;
    uint64_t ld;
    const uint32_t *ids;
    const uint4 *tiles;
    uint32_t ntiles;
    const uint32_t *cellStart;
    const uint32_t *cellCoord;
    const uint64_t *hkeys;
    const uint32_t *hvals;
    GridArgs ga;
    uint32_t ncombo;   // 3^(k-1) neighbor combinations of dims 1..k-1
    double thr;        // largest DTYPE value x with sqrt(x) <= eps (exact in double)
    uint32_t d;        // dims
    uint32_t ilp;      // ILP
    uint32_t sc;       // SHORTCIRCUIT early exit
    uint32_t *counts;  // count pass: counts[original id]
    uint64_t *tileWork;  // count pass: queries x candidates per tile (nullable)
    const uint64_t *offsets;  // fill pass: CSR offsets by original id
    uint32_t *out;     // fill pass: neighbor ids
    uint64_t cap;      // capacity of out
    uint32_t *status;  // [0] set when a write would exceed cap
    uint32_t xcd_map;  // 1: XCD-contiguous block->tile mapping
    uint32_t prune;    // 1: query-box pruning of neighbor cells
};

// Launches the count (fill=false) or fill (fill=true) pass on stream s.
int join_pass(const Grid &G, const JoinArgs &a, int mode, bool fill, hipStream_t s);

// Name of the kernel variant join_pass picks (for profiles and tests).
const char *join_variant(int dtype, uint32_t d, int mode, uint32_t ilp);

}  // namespace gdsj
