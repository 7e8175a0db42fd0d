// grid.h -- device epsilon-grid: the MI355X counterpart of the reference's
// host index build (ReorderByDimension main.cu:981-1075,
// generateNDGridDimensions main.cu:891-948,
// populateNDGridIndexAndLookupArray main.cu:690-857) plus the per-cell
// neighbor-run lists and work tiles the join kernels consume.
//
// HBM layout after build (sorted = points ordered by (cell linear id, id)):
//   xs      [D][ld] DTYPE   SoA coordinates in sorted order, dims permuted
//   ids     [n]     uint32  sorted position -> original point id
//   cellKey [nc]    uint64  sorted unique linear ids (gridCellLookup)
//   cellStart[nc+1] uint32  sorted-position range of each cell (grid)
//   runOff  [nc+1]  uint32  per cell: range into runs
//   runs    [nr]    uint2   {start,end} sorted positions of up to three
//                           cells consecutive in dimension 0
//   tiles   [nt]    uint4   {cell, q0, q1, candidates}
#pragma once
#include "common.h"

namespace gdsj {

constexpr int MAXK = 12;
constexpr int MAXD = 1024;

struct GridDims {
    uint32_t k;
    uint32_t nCells[MAXK];
    uint64_t mult[MAXK];  // product of nCells below each dim
    double minArr[MAXK];  // DTYPE value held exactly in a double
};

struct Grid {
    // inputs
    uint32_t n = 0, d = 0, k = 0;
    int dtype = 0;
    double eps = 0;
    uint32_t perm[MAXD];
    GridDims g{};
    uint64_t totalCells = 0;
    uint32_t ncell = 0, nruns = 0, ntiles = 0, tileq = 64;
    uint64_t candidates = 0;  // point comparisons of a full join
    uint64_t ld = 0;          // leading dimension of xs
    // device buffers
    DevBuf xs, ids, cellKey, cellStart, runOff, runs, tiles;
    DevBuf keyA, keyB, valB, tmp, aux, flag, tileKey, tileVal, cellAux, scalars;
    void release();
};

// Builds the grid for d_data (n x d row-major DTYPE) on stream s.
// reorder: REORDER; tileq: queries per tile (<= 64); order_tiles: QUERYREORDER.
int grid_build(Grid &G, const void *d_data, uint32_t n, uint32_t d, int dtype, double eps, uint32_t k,
               int reorder, uint32_t tileq, int order_tiles, hipStream_t s);

}  // namespace gdsj
