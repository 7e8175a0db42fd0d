// grid.h -- device epsilon-grid: the MI355X counterpart of the reference's
// host index build (ReorderByDimension main.cu:981-1075,
// generateNDGridDimensions main.cu:891-948,
// populateNDGridIndexAndLookupArray main.cu:690-857) plus the cell hash
// table and work tiles the join kernels consume.
//
// HBM layout after build (sorted = points ordered by (cell linear id, id)):
//   xs        [D][ld] DTYPE  SoA coordinates in sorted order, dims permuted;
//                            positions n .. ld-1 (ld > n) hold NaN
//   ids       [n]     uint32 sorted position -> original point id
//   cellKey   [nc]    uint64 sorted unique linear ids (gridCellLookup)
//   cellStart [nc+1]  uint32 sorted-position range of each cell (grid)
//   cellCoord [k][nc] uint32 cell coordinates per indexed dim
//   hkeys     [cap][2] uint64 open-addressing table, one 16-byte slot per
//                     row: {row (linear id without dim 0) + 1, first cell of
//                     the row | number of cells of the row << 32}; or
//   hvals     [cells] uint32 a direct table linear id -> cell (small grids)
//                     (both replace the binary searches of kernel.cu:395-398)
//   tiles     [nt]    uint4  {cell, q0, q1, 0}: up to 64 queries of one cell
#pragma once
#include <vector>

#include "common.h"

namespace gdsj {

constexpr int MAXK = 12;
constexpr int MAXD = 1024;
constexpr int RUN_HIST_BINS = 6;  // neighbor-run length histogram per cell

// what the join kernels need to enumerate adjacent cells
struct GridArgs {
    uint32_t k;
    uint32_t ncell;
    uint32_t hbits;  // log2 of the hash capacity; 0 = direct table indexed by linear id
    uint32_t nCells[MAXK];
    uint64_t mult[MAXK];   // product of nCells below each dim
    double minArr[MAXK];   // grid origin per indexed dim (exact DTYPE values)
    double eps;            // epsilon (exact DTYPE value)
    double ulp;            // unit roundoff of DTYPE (2^-24 or 2^-53)
    double prune_lim;      // eps^2 * (1 + margin): squared gap beyond which no pair can match
};

// Row stride of the row-major copy Grid::xa for d dims: d (no padding, so a
// run of candidates spans as few cache lines as its coordinates need).
__host__ __device__ constexpr uint32_t xa_stride(uint32_t d) { return d; }

struct Grid {
    // inputs
    uint32_t n = 0, d = 0, k = 0;
    int dtype = 0;
    double eps = 0;
    uint32_t perm[MAXD];
    double minArr[MAXK];  // DTYPE value held exactly in a double
    GridArgs ga{};
    uint64_t totalCells = 0;
    uint32_t ncell = 0, ntiles = 0, tileq = 64;
    uint64_t ld = 0;  // leading dimension of xs
    // device buffers
    DevBuf xs, ids, cellKey, cellStart, cellCoord, hkeys, hvals, tiles;
    DevBuf runs, nruns, cellCand, cellHist;  // per-cell neighbor runs (build_runs)
    DevBuf comboTab;                         // neighbor combinations (build_runs)
    std::vector<uint4> hComboTab;            //   its host copy (source of the upload)
    uint64_t runStride = 0;
    bool haveRuns = false;
    // SORT (params.h:46; set before grid_build): each cell's points ordered by
    // their value at dim 0 when every dim is indexed, else at dim k (after
    // REORDER); posCell maps a sorted position to its cell
    int sortCells = 0;
    DevBuf posCell;
    // row-major copy of xs (set aosWant before grid_build; up to 16 dims):
    // point p's coordinates at xa[p * xaStride ..], xaStride = dims
    // (xa_stride), NaN in the padding rows n .. ld-1.  A candidate's coordinates then share one or two cache lines
    // instead of one per dim (runs of adjacent cells are short from 6 dims).
    int aosWant = 0;
    uint32_t xaStride = 0;
    DevBuf xa;
    // index (set indexWant and ilpTerms before grid_build): 1 the reference's
    // k dims of largest variance; 2 the top k principal axes of the REORDER
    // sample (y = R x, cells of width cellW > eps so that every pair the
    // reference's arithmetic accepts lies in adjacent cells), used only when
    // no point lies within rounding of a face of the reference's grid
    // (riskyPoints = 0); 0 auto (principal axes when they leave at most a
    // twentieth of the variance the reference's dims leave).  ixs: the index
    // coordinates in sorted order ([k][ld], principal axes only), which the
    // run pruning reads instead of xs; ybuf: the same in input order.
    int indexWant = 1;
    uint32_t indexDims = 0;  // principal axes indexed (0: k); the grid's k becomes it
    uint32_t ilpTerms = 0;
    int indexUsed = 1;
    uint64_t riskyPoints = 0;
    double cellW = 0;
    DevBuf ixs, ybuf, pcaBuf;
    // principal axes for the Gram filter's operands (set gramAxesWant before
    // grid_build; computed only where the index's basis is): 1 the points'
    // own dims; 0 auto (16 axes where the variance they leave out is small
    // next to eps^2); 16 or 32: that many axes.  gramP: gramAxes x d, columns
    // in xs's dim order; gramNorm >= the spectral norm of gramP.
    int gramAxesWant = 1;
    uint32_t gramAxes = 0;
    double gramNorm = 0;
    std::vector<double> gramP;
    DevBuf keyA, keyB, valB, tmp, aux, flag, tileKey, tileVal, cellAux, scalars;
    HostBuf hostSample;  // REORDER's sample rows, pinned (the variance is summed on the host)
    void release();
};

// Builds the grid for d_data (n x d row-major DTYPE) on stream s.
// reorder: REORDER; tileq: queries per tile (<= 64); order_tiles: QUERYREORDER.
int grid_build(Grid &G, const void *d_data, uint32_t n, uint32_t d, int dtype, double eps, uint32_t k,
               int reorder, uint32_t tileq, int order_tiles, hipStream_t s);

__device__ __forceinline__ uint32_t cell_hash(uint64_t key, uint32_t bits) {
    return (uint32_t)((key * 0x9E3779B97F4A7C15ull) >> (64 - bits));
}

// Direct table (hbits == 0): cell index of linear id lin, hvals indexed by
// linear id, 0xFFFFFFFF = empty.  Hashed (hbits > 0): the table holds rows
// (cells equal in dims 1..k-1; their linear ids are row + c_0, so a row's
// cells are contiguous in cell order): key row + 1 -> first cell of the row,
// linear probing in a table at most half full (a miss ends at an empty slot),
// key and value side by side in 16-byte slots.
// Returns the cell (direct) or the row's first cell (hashed), or -1.
__device__ __forceinline__ int32_t cell_find(const uint64_t *__restrict__ hkeys, const uint32_t *__restrict__ hvals,
                                             uint32_t bits, uint64_t lin) {
    if (bits == 0) return (int32_t)hvals[lin];
    const uint32_t mask = (1u << bits) - 1u;
    uint32_t h = cell_hash(lin, bits);
    const uint64_t want = lin + 1;
    if (bits <= 27) {
        // key and cell in one 16-byte buffer load (the compiler would
        // otherwise fetch the cell by a second, dependent load after the
        // key compare); slot offsets fit the 32-bit buffer offset
        const __amdgpu_buffer_rsrc_t rs = __builtin_amdgcn_make_buffer_rsrc(
            (void *)hkeys, (short)0, (int)((mask + 1u) * 16u), 0x00020000);
        for (uint32_t probe = 0; probe <= mask; probe++) {
            const auto e = __builtin_amdgcn_raw_buffer_load_b128(rs, h * 16u, 0, 0);
            const uint64_t key = ((uint64_t)e[1] << 32) | e[0];
            if (key == want) return (int32_t)e[2];
            if (key == 0) return -1;
            h = (h + 1) & mask;
        }
        return -1;
    }
    for (uint32_t probe = 0; probe <= mask; probe++) {
        const ulonglong2 e = reinterpret_cast<const ulonglong2 *>(hkeys)[h];
        if (e.x == want) return (int32_t)e.y;
        if (e.x == 0) return -1;
        h = (h + 1) & mask;
    }
    return -1;
}

// first index i in [lo, hi) with keys[i] >= want (hi if none); galloping
// from lo, so a short distance costs few loads
// First index in [lo, hi) of a sorted uint32 range whose value is >= want
// (hi if none).
__device__ __forceinline__ uint32_t u32_lower_bound(const uint32_t *__restrict__ v, uint32_t lo, uint32_t hi,
                                                    uint32_t want) {
    while (lo < hi) {
        const uint32_t m = lo + (hi - lo) / 2;
        if (v[m] < want) lo = m + 1; else hi = m;
    }
    return lo;
}

__host__ __device__ __forceinline__ uint32_t key_lower_bound(const uint64_t *__restrict__ keys, uint32_t lo, uint32_t hi,
                                                    uint64_t want) {
    uint32_t step = 1, a = lo, b = lo;
    while (b < hi && keys[b] < want) {
        a = b + 1;
        b = (hi - b > step) ? b + step : hi;
        step <<= 1;
    }
    while (a < b) {
        const uint32_t m = a + (b - a) / 2;
        if (keys[m] < want) a = m + 1; else b = m;
    }
    return a;
}

}  // namespace gdsj
