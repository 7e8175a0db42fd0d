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
    ACC_UNFUSED = 3,   // the reference's CPU paths: acc = acc + round(t*t), dims in order
                       // (main.cu:963-967; all-pairs join only)
};

// The accumulation a join runs for ILP, SHORTCIRCUIT and the fold reading
// (gdsj_params.ilp_fold): with every dim fused into the running sum, ILP > 0
// with SHORTCIRCUIT is the sequential sum of ILP = 0.
inline int acc_mode(uint32_t ilp, uint32_t sc, uint32_t fold) {
    if (ilp == 0 || (sc && fold)) return ACC_SEQ;
    return sc ? ACC_ILP_SC : ACC_ILP_NOSC;
}

enum JoinPass : int { PASS_COUNT = 0, PASS_FILL = 1, PASS_SLOT = 2 };

// Everything the join kernels read besides the coordinates and ids
// (passed by value as one kernel argument).
// slot positions of a tile stay below this (their byte offsets fit 32 bits);
// a tile with more hits than min(slot capacity, limit) has overflowed
constexpr uint32_t SLOT_POS_LIMIT = 1u << 30;

struct JoinArgs {
    const void *xs;
    uint64_t ld;
    // the index coordinates in sorted order ([k][ld]): xs itself, or the
    // principal-axes coordinates (Grid::ixs); the run pruning reads them
    const void *ixs;
    // nullable: the row-major copy of xs (Grid::xa; up to 8 dims), from which
    // the kernels load candidate coordinates; xaStride elements per point
    const void *xa;
    uint32_t xaStride;
    const uint32_t *ids;
    const uint4 *tiles;
    const uint32_t *tileList;  // nullable: process tiles[tileList[t]] for t < ntiles
    uint32_t ntiles;
    const uint32_t *cellStart;
    const uint32_t *cellCoord;
    const uint64_t *cellKey;
    const uint64_t *hkeys;
    const uint32_t *hvals;
    const uint2 *runs;         // nullable: precomputed neighbor runs per cell
    const uint32_t *nruns;     //   number of runs of each cell
    uint64_t runStride;        //   runs of cell c start at c * runStride
    GridArgs ga;
    uint32_t ncombo;   // 3^(k-1) neighbor combinations of dims 1..k-1
    double thr;        // largest DTYPE value x with sqrt(x) <= eps (exact in double)
    // FMA-chain filter of the plain kernels' register-resident pair tests
    // (ILP > 0, counting passes): f = the squared differences summed by one
    // FMA each decides
    // a pair when f <= ffLo (a neighbor) or f > ffHi (not one); a chunk with
    // a pair in between is evaluated in the reference's order (fma_band).
    // Zero-initialised args are unsafe here: every JoinArgs gets its band.
    double ffLo, ffHi;
    // Gram-form filter of the plain kernels (every dim indexed): a pair is a
    // neighbor when |c'|^2 - 2 q'.c' <= gvLo - |q'|^2 and not one when it
    // exceeds gvHi - |q'|^2 (c', q': coordinates relative to the tile's
    // first query); in between the reference's arithmetic decides
    // (gram_valu_band).  gv = 0: the reference's arithmetic for every pair.
    uint32_t gv;
    float gvLo, gvMid, gvHalf;  // band [gvLo, gvLo + 2 half'], gvHalf >= half' with room
    // SORT (kernel.cu:400-467; generic kernels): 1 every dim indexed (sort
    // dim 0), 2 the first unindexed dim sortDim; cells in value order,
    // posCell = cell of each sorted position
    uint32_t sortMode, sortDim;
    const uint32_t *posCell;
    uint32_t d;        // dims
    uint32_t ilp;      // ILP
    uint32_t sc;       // SHORTCIRCUIT early exit
    uint32_t xcd_map;  // 1: XCD-contiguous block->tile mapping
    uint32_t prune;    // 1: query-box pruning of neighbor cells (in-kernel enumeration)
    // count / slot pass
    uint32_t *counts;    // counts[original id]
    uint64_t *tileWork;  // queries x candidates per tile (nullable)
    // fill pass
    const uint64_t *offsets;  // CSR offsets by original id
    uint32_t *out;            // neighbor ids
    uint64_t cap;             // capacity of out
    // slot pass: hits go to a per-tile slot sized by the estimator
    const uint64_t *slotOff;  // slot start per tile
    const uint32_t *slotCap;  // slot capacity per tile
    uint32_t *slotUsed;       // hits per tile (may exceed slotCap: overflow)
    uint32_t *slotIds;
    uint8_t *slotQ;           // query index inside the tile of each slot entry
    uint32_t *ovfList;        // tiles whose slot overflowed (tileBase + tile)
    uint32_t *ovfCount;
    uint32_t tileBase;        // first global tile id of this join in a batch (0 alone)
    // unicomp (STAMP): each adjacent cell pair is evaluated from its owner's
    // tiles only and both directions are emitted; counts[] is accumulated
    // with atomics (zeroed before the counting pass), fcounts[q] = the
    // query's own-side count, rcur[j] = reverse entries placed so far
    uint32_t stamp;  // 0 off, 1 STAMP, 2 LINEARSTAMP ownership
    uint32_t *fcounts;
    uint32_t *rcur;
    uint32_t *status;  // [0] set when a fill write would exceed cap
    // hit masks (exact passes over precomputed runs, no unicomp): the
    // counting pass stores the hit ballot of every (query, 64-candidate
    // chunk), the fill pass reads it back instead of evaluating distances.
    // Masks of the query at sorted position p of cell c start at word
    // hmCell[c] + (p - cellStart[c]) * stride(c), one word per chunk, with
    // stride(c) = ceil(cellCand[c] / 64) + ceil(nruns[c] / 64) (a window of
    // runs leaves at most one partial chunk).
    uint64_t *hmask;            // nullable: no masks
    const uint64_t *hmCell;
    const uint32_t *cellCand;   // candidates per query of each cell
    // candidate splits (Gram kernel, over precomputed runs: the exact fill
    // of overflowed tiles, whose neighborhoods are the largest):
    // tiles[t].w = s + 1 restricts tile t to the s-th of winSplit equal
    // ranges of its cell's candidate order (cellCand: the total).  The
    // counting pass then writes row r's count to splitCnt[t * SPLIT_ROWS + r]
    // and the fill pass starts row r at splitOff[t * SPLIT_ROWS + r].
    uint32_t winSplit;          // 0: tiles are not split
    uint32_t *splitCnt;
    const uint64_t *splitOff;
    // matrix-core join of low-dimensional data (every dim indexed, 3 to 7
    // dims; k_mjoin): the tiles mjList[0 .. mjN) run on it, the tiles
    // vjList[0 .. vjN) on the plain kernels (mjN + vjN = ntiles, both lists in
    // tile order).  A pair whose Gram-form estimate F exceeds mjHi is not a
    // neighbor, one with F <= mjLo is (mjoin_band); mjList = nullptr: the
    // plain kernels for every tile.
    const uint32_t *mjList;
    uint32_t mjN;
    const uint32_t *vjList;
    uint32_t vjN;
    float mjHi, mjLo;
    // (host side) the matrix-core tiles run on the stream mjStream (a
    // hipStream_t), forked from and joined back into the pass's stream by
    // the events mjFork and mjJoin, so that the plain kernels' many small
    // tiles fill the SIMDs the few heavy matrix-core tiles leave idle;
    // nullptr: both on the pass's stream, one after the other
    void *mjStream, *mjFork, *mjJoin;
};

constexpr uint32_t SPLIT_ROWS = 128;  // row stride of splitCnt / splitOff

// mask words per query of a cell with cand candidates in nruns runs
__host__ __device__ __forceinline__ uint32_t hmask_stride(uint32_t cand, uint32_t nruns) {
    return (cand + 63u) / 64u + (nruns + 63u) / 64u;
}

// Band around thr for the Gram-form filter of a join whose dims are all
// indexed (FP32): pairs with g <= lo are neighbors, pairs with g > hi are
// not, and |g - mid| <= half holds for every g in (lo, hi] even after the
// rounding of g - mid.  False when the filter does not apply.
bool gram_valu_band(const Grid &G, double thr, float *lo, float *mid, float *half);

// Thresholds of the matrix-core join (k_mjoin) of G (every dim indexed, 3 to
// 7 dims): pairs with F > hi are not neighbors, pairs with F <= lo are.
// False when it does not apply (other dims, cells too fine for float
// coordinates, a band wider than 5 % of thr).
bool mjoin_band(const Grid &G, double thr, float *hi, float *lo);
// The FMA-chain filter's band for DTYPE dtype, d dims and ILP ilp; mode 2
// gives a band that holds every pair (each decided by the reference's
// arithmetic).
void fma_band(int dtype, uint32_t d, uint32_t ilp, double thr, uint32_t mode, double *lo, double *hi);

// list[0 .. *nbig) = the tiles whose cell holds at least minOcc points,
// list[*nbig .. ntiles) = the others, both in tile order (tmp: scan scratch
// of scan_tmp_bytes(ntiles + 1) bytes, flags: 2 (ntiles + 1) words).
// Waits for the stream (the count comes back to the host).
int mjoin_partition(const Grid &G, uint32_t minOcc, uint32_t *list, uint32_t *flags, void *tmp, uint32_t *nbig,
                    hipStream_t s);

// Launches one pass (JoinPass) of the join over a.ntiles tiles on stream s.
int join_pass(const Grid &G, const JoinArgs &a, int mode, int pass, hipStream_t s);

// Builds the compacted neighbor-run list of every cell (runs, nruns,
// cellCand = candidates per query of the cell) when it fits the memory cap;
// returns false (and leaves the lists unset) otherwise.
// stamp: only the neighbor cells each cell owns (unicomp: 1 STAMP, 2
// LINEARSTAMP ownership, combo_decode in neighbors.h).
// maskW (nullable): per cell, its queries' hit-mask words (saturated at
// 2^32 - 1).
int build_runs(Grid &G, hipStream_t s, bool prune, int stamp, bool *built, uint32_t *maskW = nullptr);

// out[0 .. *nout) = the tiles of list cut into sub-tiles of <= subq queries
// (tileq: the tiles' maximum query count; out holds nlist * ceil(tileq/subq)).
int split_tiles(const uint32_t *list, uint32_t nlist, const uint4 *tiles, uint32_t tileq, uint32_t subq,
                uint4 *out, uint32_t *nout, hipStream_t s);

// out[i * S + s] = sub-tile i restricted to the s-th of S candidate ranges
// (w = s + 1; JoinArgs::winSplit).
int split_windows(const uint4 *sub, uint32_t nsub, uint32_t S, uint4 *out, hipStream_t s);

// off[(i * S + s) * SPLIT_ROWS + r] = where row r of window split s of
// sub-tile i starts: the row's CSR offset plus the counts (cnt, same
// layout) of the earlier splits.
int split_offsets(const uint4 *tiles, uint32_t nsub, uint32_t S, const uint32_t *ids, const uint64_t *offsets,
                  const uint32_t *cnt, uint64_t *off, hipStream_t s);

// Moves every non-overflowed tile's slot entries into CSR rows.
int compact_slots(const JoinArgs &a, hipStream_t s);

// Name of the kernel variant join_pass picks (for profiles and tests).
const char *join_variant(int dtype, uint32_t d, int mode, uint32_t ilp);

}  // namespace gdsj
