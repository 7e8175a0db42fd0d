// runs.hip -- per-cell neighbor-run lists and slot compaction.
//
// build_runs: one wave per non-empty cell enumerates the 3^(k-1) rows of
// adjacent cells once (lookup table, clamps of kernel.cu:183-187, pruning
// against the cell's point box) and stores the non-empty runs compacted at
// a fixed stride, so the join passes read them instead of re-enumerating.
//
// compact_slots: one wave per tile moves the tile's slot entries (neighbor
// id + query index, written in candidate order) into the CSR rows with a
// stable partition by query index.
#include "compact.h"
#include "join.h"
#include "neighbors.h"

namespace gdsj {

template <typename T>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(8, 8))) void k_cell_runs(const T *__restrict__ xs, uint64_t ld, GridArgs ga,
                                                    const uint32_t *__restrict__ cellStart,
                                                    const uint32_t *__restrict__ cellCoord,
                                                    const uint64_t *__restrict__ cellKey,
                                                    const uint64_t *__restrict__ hkeys,
                                                    const uint32_t *__restrict__ hvals, uint32_t ncombo,
                                                    uint64_t stride, int prune, int stamp,
                                                    uint2 *__restrict__ runs,
                                                    uint32_t *__restrict__ nruns, uint32_t *__restrict__ cellCand,
                                                    float *__restrict__ cellHist) {
    const uint32_t lane = lane_id();
    // each XCD takes a contiguous range of cells: cells near in key order
    // share neighbor rows, whose hash slots and keys then meet in one L2
    const uint32_t lblock = xcd_block(blockIdx.x, gridDim.x);
    const uint32_t cell = uniform((uint32_t)(lblock * (blockDim.x / GDSJ_WAVE) + threadIdx.x / GDSJ_WAVE));
    if (cell >= ga.ncell) return;
    __shared__ double s_gap[256 / GDSJ_WAVE][2][MAXK];
    double *glo = s_gap[threadIdx.x / GDSJ_WAVE][0];
    double *ghi = s_gap[threadIdx.x / GDSJ_WAVE][1];
    uint32_t cc[MAXK];
    load_cell_coords(cellCoord, ga.ncell, cell, ga.k, cc);
    const uint32_t s0 = uniform(cellStart[cell]), s1 = uniform(cellStart[cell + 1]);
    if (prune) {
        tile_gaps<T>(xs, ld, s0, s1 - s0, cc, ga, glo, ghi);
    } else {
        if (lane < MAXK) {
            glo[lane] = 0.0;
            ghi[lane] = 0.0;
        }
        __builtin_amdgcn_wave_barrier();
    }
    uint32_t cursor = 0;
    uint32_t cand = 0;
    uint32_t hist[RUN_HIST_BINS] = {};
    uint2 *mine = runs + (uint64_t)cell * stride;
    for (uint32_t cw = 0; cw < ncombo; cw += GDSJ_WAVE) {
        uint2 run = make_uint2(0u, 0u);
        if (cw + lane < ncombo)
            run = combo_run(cw + lane, cc, glo, ghi, ga, cellStart, cellKey, hkeys, hvals, stamp != 0);
        const bool ne = run.y > run.x;
        const uint64_t m = ballot(ne);
        if (ne) mine[cursor + mbcnt(m)] = run;
        cursor += popc64(m);
        uint32_t len = run.y - run.x;
        // run-length histogram (estimator features): 1, 2-3, 4-7, 8-15, 16-63, 64+
        const uint32_t bin = len < 2 ? 0u : len < 4 ? 1u : len < 8 ? 2u : len < 16 ? 3u : len < 64 ? 4u : 5u;
#pragma unroll
        for (int b = 0; b < RUN_HIST_BINS; b++) hist[b] += popc64(ballot(ne && bin == (uint32_t)b));
#pragma unroll
        for (int off = 32; off > 0; off >>= 1) len += __shfl_xor(len, off, GDSJ_WAVE);
        cand += uniform(len);
    }
    if (lane == 0) {
        nruns[cell] = cursor;
        cellCand[cell] = cand;
        const float inv = cursor ? 1.0f / (float)cursor : 0.0f;
#pragma unroll
        for (int b = 0; b < RUN_HIST_BINS; b++) cellHist[(uint64_t)cell * RUN_HIST_BINS + b] = (float)hist[b] * inv;
    }
}

int build_runs(Grid &G, hipStream_t s, bool prune, bool stamp, bool *built) {
    *built = false;
    uint64_t ncombo = 1;
    for (uint32_t j = 1; j < G.k; j++) ncombo *= 3;
    const uint64_t entries = (uint64_t)G.ncell * ncombo;
    if (G.ncell == 0 || entries * 8 > (6ull << 30)) return GDSJ_OK;
    G.runStride = ncombo;
    GDSJ_TRY(G.runs.ensure(entries * 8 + 64));
    GDSJ_TRY(G.nruns.ensure((size_t)G.ncell * 4 + 64));
    GDSJ_TRY(G.cellCand.ensure((size_t)G.ncell * 4 + 64));
    GDSJ_TRY(G.cellHist.ensure((size_t)G.ncell * RUN_HIST_BINS * 4 + 64));
    const uint32_t blocks = (G.ncell + 3) / 4;
    if (G.dtype == GDSJ_DTYPE_F32)
        k_cell_runs<float><<<blocks, 256, 0, s>>>(G.xs.as<float>(), G.ld, G.ga, G.cellStart.as<uint32_t>(),
                                                  G.cellCoord.as<uint32_t>(), G.cellKey.as<uint64_t>(), G.hkeys.as<uint64_t>(),
                                                  G.hvals.as<uint32_t>(), (uint32_t)ncombo, ncombo, prune ? 1 : 0,
                                                  stamp ? 1 : 0, G.runs.as<uint2>(), G.nruns.as<uint32_t>(),
                                                  G.cellCand.as<uint32_t>(), G.cellHist.as<float>());
    else
        k_cell_runs<double><<<blocks, 256, 0, s>>>(G.xs.as<double>(), G.ld, G.ga, G.cellStart.as<uint32_t>(),
                                                   G.cellCoord.as<uint32_t>(), G.cellKey.as<uint64_t>(), G.hkeys.as<uint64_t>(),
                                                   G.hvals.as<uint32_t>(), (uint32_t)ncombo, ncombo,
                                                   prune ? 1 : 0, stamp ? 1 : 0, G.runs.as<uint2>(),
                                                   G.nruns.as<uint32_t>(),
                                                   G.cellCand.as<uint32_t>(), G.cellHist.as<float>());
    GDSJ_HIP(hipGetLastError());
    *built = true;
    return GDSJ_OK;
}

__global__ __launch_bounds__(256) void k_compact(const JoinArgs a) {
    const uint32_t tile = uniform((uint32_t)(blockIdx.x * (blockDim.x / GDSJ_WAVE) + threadIdx.x / GDSJ_WAVE));
    if (tile >= a.ntiles) return;
    compact_tile(a, tile);
}

// Sub-tiles of at most subq queries of the tiles in list (per = tileq / subq
// sub-tiles each; the unused ones are empty).  Rows do not depend on the
// tiling, so the fill pass over the sub-tiles writes the same rows with more
// waves in flight.
__global__ void k_split_tiles(const uint32_t *__restrict__ list, uint32_t nlist, const uint4 *__restrict__ tiles,
                              uint32_t subq, uint32_t per, uint4 *__restrict__ out) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= nlist * per) return;
    const uint4 tl = tiles[list[i / per]];
    const uint32_t j = i % per;
    uint32_t a = tl.y + j * subq;
    if (a > tl.z) a = tl.z;
    const uint32_t b = a + subq < tl.z ? a + subq : tl.z;
    out[i] = make_uint4(tl.x, a, b, 0u);
}

int split_tiles(const uint32_t *list, uint32_t nlist, const uint4 *tiles, uint32_t tileq, uint32_t subq,
                uint4 *out, uint32_t *nout, hipStream_t s) {
    const uint32_t per = (tileq + subq - 1) / subq;
    *nout = nlist * per;
    if (*nout == 0) return GDSJ_OK;
    k_split_tiles<<<ceil_div(*nout, 256), 256, 0, s>>>(list, nlist, tiles, subq, per, out);
    GDSJ_HIP(hipGetLastError());
    return GDSJ_OK;
}

int compact_slots(const JoinArgs &a, hipStream_t s) {
    const uint32_t blocks = (a.ntiles + 3) / 4;
    if (blocks == 0) return GDSJ_OK;
    k_compact<<<blocks, 256, 0, s>>>(a);
    GDSJ_HIP(hipGetLastError());
    return GDSJ_OK;
}

}  // namespace gdsj
