// runs.hip -- per-cell neighbor-run lists and slot compaction.
//
// build_runs: one wave per non-empty cell enumerates the 3^(k-1) rows of
// adjacent cells once (lookup table, clamps of kernel.cu:183-187, pruning
// against the cell's point box) and stores the non-empty runs compacted at
// a fixed stride, so the join passes read them instead of re-enumerating.
// The runs are combo_run's (neighbors.h), in combination order.
//
// compact_slots: one wave per tile moves the tile's slot entries (neighbor
// id + query index, written in candidate order) into the CSR rows with a
// stable partition by query index.
#include "compact.h"
#include "join.h"
#include "neighbors.h"

namespace gdsj {

// Neighbor combination t of a cell (digits d_j of t in base 3 = offsets
// -1/0/+1 of dims 1..k-1, dim 1 least significant) is described once per
// join, independently of the cell, by a table entry {delta lo, delta hi,
// info, 0}: delta = sum_j (d_j - 1) * mult[j] (two's complement), the
// offset of the neighbor row's linear id from the cell's; info = ia |
// ib << 12 | top << 24 | d_top << 28 with ia / ib the base-3 index of the
// digits of dims 1..a / a+1..k-1 (a = (k-1)/2), top the highest dim with a
// nonzero offset (0: none) and d_top its digit.  Per cell, the squared gaps of the 3^a and 3^(k-1-a)
// digit groups (+inf where a coordinate leaves the grid) sit in LDS, so a
// combination's gap is two LDS reads and an add instead of a loop over the
// dims.
static void combo_table(Grid &G, uint32_t *na, uint32_t *nb, uint32_t *split) {
    const uint32_t m = G.k - 1, a = m / 2;
    uint32_t pa = 1, pb = 1;
    for (uint32_t j = 0; j < a; j++) pa *= 3u;
    for (uint32_t j = a; j < m; j++) pb *= 3u;
    const uint32_t ncombo = pa * pb;
    G.hComboTab.resize(ncombo);
    for (uint32_t t = 0; t < ncombo; t++) {
        uint32_t tt = t, ia = 0, ib = 0, top = 0, topd = 1, wa = 1, wb = 1;
        uint64_t delta = 0;
        for (uint32_t j = 1; j <= m; j++) {
            const uint32_t d = tt % 3u;
            tt /= 3u;
            delta += (uint64_t)((int64_t)d - 1) * G.ga.mult[j];
            if (j <= a) {
                ia += d * wa;
                wa *= 3u;
            } else {
                ib += d * wb;
                wb *= 3u;
            }
            if (d != 1) {
                top = j;
                topd = d;
            }
        }
        G.hComboTab[t] = make_uint4((uint32_t)delta, (uint32_t)(delta >> 32),
                                    ia | (ib << 12) | (top << 24) | (topd << 28), 0u);
    }
    *na = pa;
    *nb = pb;
    *split = a;
}

// R: combinations per lane and loop step (their lookups overlap)
template <typename T, int R>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(R > 1 ? 6 : 8, 8))) void k_cell_runs(
    const T *__restrict__ xs, uint64_t ld, GridArgs ga, const uint32_t *__restrict__ cellStart,
    const uint32_t *__restrict__ cellCoord, const uint64_t *__restrict__ cellKey,
    const uint64_t *__restrict__ hkeys, const uint32_t *__restrict__ hvals, const uint4 *__restrict__ tab,
    uint32_t ncombo, uint32_t na, uint32_t nb, uint32_t split, uint64_t stride, int prune, int stamp,
    uint2 *__restrict__ runs, uint32_t *__restrict__ nruns, uint32_t *__restrict__ cellCand,
    float *__restrict__ cellHist, uint32_t *__restrict__ maskW) {
    const uint32_t lane = lane_id();
    // each XCD takes a contiguous range of cells: cells near in key order
    // share neighbor rows, whose hash slots and keys then meet in one L2
    const uint32_t lblock = xcd_block(blockIdx.x, gridDim.x);
    const uint32_t cell = uniform((uint32_t)(lblock * (blockDim.x / GDSJ_WAVE) + threadIdx.x / GDSJ_WAVE));
    if (cell >= ga.ncell) return;
    __shared__ double s_gap[256 / GDSJ_WAVE][2][MAXK];
    extern __shared__ double s_grp[];  // per wave: na + nb group gaps
    double *glo = s_gap[threadIdx.x / GDSJ_WAVE][0];
    double *ghi = s_gap[threadIdx.x / GDSJ_WAVE][1];
    double *gA = s_grp + (threadIdx.x / GDSJ_WAVE) * (na + nb);
    double *gB = gA + na;
    uint32_t cc[MAXK];
    load_cell_coords(cellCoord, ga.ncell, cell, ga.k, cc);
    const uint32_t s0 = uniform(cellStart[cell]), s1 = uniform(cellStart[cell + 1]);
    if (prune) {
        tile_gaps<T, true>(xs, ld, s0, s1 - s0, cc, ga, glo, ghi);
    } else {
        if (lane < MAXK) {
            glo[lane] = 0.0;
            ghi[lane] = 0.0;
        }
        __builtin_amdgcn_wave_barrier();
    }
    // the cell's row (linear id without dim 0) and coordinate parities
    uint64_t rowbase = 0;
    uint32_t odd = 0;
    for (uint32_t j = 1; j < ga.k; j++) {
        rowbase += (uint64_t)cc[j] * ga.mult[j];
        odd |= (cc[j] & 1u) << j;
    }
    // squared gaps of the digit groups, in dim order within a group
    for (uint32_t i = lane; i < na + nb; i += GDSJ_WAVE) {
        const bool ga_ = i < na;
        uint32_t idx = ga_ ? i : i - na;
        const uint32_t j0 = ga_ ? 1u : split + 1u, j1 = ga_ ? split + 1u : ga.k;
        double g = 0.0;
        bool valid = true;
        for (uint32_t j = j0; j < j1; j++) {
            const uint32_t d = idx % 3u;
            idx /= 3u;
            const int64_t cj = (int64_t)cc[j] + (int64_t)d - 1;
            valid = valid && cj >= 0 && cj < (int64_t)ga.nCells[j];
            g += d == 0 ? glo[j] : (d == 2 ? ghi[j] : 0.0);
        }
        gA[i] = valid ? g : __builtin_inf();
    }
    __builtin_amdgcn_wave_barrier();
    const uint32_t c0 = cc[0];
    const uint32_t x0d = c0 == 0 ? 0u : c0 - 1u;
    const uint32_t x1d = c0 + 1u < ga.nCells[0] ? c0 + 1u : ga.nCells[0] - 1u;
    const double g0lo = glo[0], g0hi = ghi[0];
    uint32_t cursor = 0;
    uint32_t cand = 0;
    uint32_t hist[RUN_HIST_BINS] = {};
    uint2 *mine = runs + (uint64_t)cell * stride;
    for (uint32_t cw = 0; cw < ncombo; cw += GDSJ_WAVE * R) {
        ComboRow w[R];
        uint4 te[R];  // table entries, loaded together
#pragma unroll
        for (int r = 0; r < R; r++) {
            const uint32_t t = cw + (uint32_t)r * GDSJ_WAVE + lane;
            te[r] = tab[t < ncombo ? t : 0u];
        }
#pragma unroll
        for (int r = 0; r < R; r++) {
            const uint32_t t = cw + (uint32_t)r * GDSJ_WAVE + lane;
            const uint4 e = te[r];
            const double g = gA[e.z & 0xFFFu] + gB[(e.z >> 12) & 0xFFFu];
            const uint32_t top = (e.z >> 24) & 0xFu, topd = e.z >> 28;
            // unicomp ownership (combo_decode's rules)
            const bool owned = stamp == 1 ? !(top > 0 && !((odd >> top) & 1u))
                                          : (stamp == 2 ? !(top > 0 && topd == 0) : true);
            w[r].base = rowbase + (((uint64_t)e.y << 32) | e.x);
            w[r].live = t < ncombo && !(g > ga.prune_lim) && owned;
            w[r].x0 = x0d;
            w[r].x1 = x1d;
            if (x0d < c0 && g + g0lo > ga.prune_lim) w[r].x0 = c0;
            if (x1d > c0 && g + g0hi > ga.prune_lim) w[r].x1 = c0;
            if (stamp == 1 && top == 0 && !(c0 & 1u)) w[r].x0 = w[r].x1 = c0;
            if (stamp == 2 && top == 0) w[r].x0 = c0;
        }
        uint2 rr[R];
        combo_lookup<R>(w, ga, cellStart, cellCoord, hkeys, hvals, rr);
#pragma unroll
        for (int r = 0; r < R; r++) {
            const uint2 run = rr[r];
            const bool ne = run.y > run.x;
            const uint64_t m = ballot(ne);
            if (ne) mine[cursor + mbcnt(m)] = run;
            cursor += popc64(m);
            uint32_t len = run.y - run.x;
            // run-length histogram (estimator features): 1, 2-3, 4-7, 8-15, 16-63, 64+
            const uint32_t bin = len < 2 ? 0u : len < 4 ? 1u : len < 8 ? 2u : len < 16 ? 3u : len < 64 ? 4u : 5u;
#pragma unroll
            for (int b = 0; b < RUN_HIST_BINS; b++) hist[b] += popc64(ballot(ne && bin == (uint32_t)b));
            cand += len;  // this lane's candidates; summed over the wave below
        }
    }
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) cand += __shfl_xor(cand, off, GDSJ_WAVE);
    cand = uniform(cand);
    if (lane == 0) {
        nruns[cell] = cursor;
        cellCand[cell] = cand;
        const float inv = cursor ? 1.0f / (float)cursor : 0.0f;
#pragma unroll
        for (int b = 0; b < RUN_HIST_BINS; b++) cellHist[(uint64_t)cell * RUN_HIST_BINS + b] = (float)hist[b] * inv;
        if (maskW) {
            const uint64_t w = (uint64_t)(s1 - s0) * hmask_stride(cand, cursor);
            maskW[cell] = w > 0xFFFFFFFFull ? 0xFFFFFFFFu : (uint32_t)w;
        }
    }
}

template <typename T, int R>
static void launch_cell_runs(Grid &G, uint32_t blocks, uint32_t ncombo, uint32_t na, uint32_t nb, uint32_t split,
                             bool prune, int stamp, uint32_t *maskW, hipStream_t s) {
    const size_t lds = (size_t)(256 / GDSJ_WAVE) * (na + nb) * sizeof(double);
    const T *ixs = G.indexUsed == 2 ? G.ixs.as<T>() : G.xs.as<T>();  // the index coordinates
    k_cell_runs<T, R><<<blocks, 256, lds, s>>>(ixs, G.ld, G.ga, G.cellStart.as<uint32_t>(),
                                               G.cellCoord.as<uint32_t>(), G.cellKey.as<uint64_t>(),
                                               G.hkeys.as<uint64_t>(), G.hvals.as<uint32_t>(), G.comboTab.as<uint4>(),
                                               ncombo, na, nb, split, G.runStride, prune ? 1 : 0, stamp,
                                               G.runs.as<uint2>(), G.nruns.as<uint32_t>(), G.cellCand.as<uint32_t>(),
                                               G.cellHist.as<float>(), maskW);
}

int build_runs(Grid &G, hipStream_t s, bool prune, int stamp, bool *built, uint32_t *maskW) {
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
    uint32_t na = 1, nb = 1, split = 0;
    combo_table(G, &na, &nb, &split);
    GDSJ_TRY(G.comboTab.ensure((size_t)ncombo * 16 + 64));
    GDSJ_HIP(hipMemcpyAsync(G.comboTab.p, G.hComboTab.data(), (size_t)ncombo * 16, hipMemcpyHostToDevice, s));
    const uint32_t blocks = (G.ncell + 3) / 4;
    // 3^(k-1) rows per cell: up to four 64-row steps overlap their lookups
    const int R = ncombo <= 64 ? 1 : (ncombo <= 128 ? 2 : 4);
    const uint32_t nc = (uint32_t)ncombo;
    if (G.dtype == GDSJ_DTYPE_F32) {
        if (R == 1) launch_cell_runs<float, 1>(G, blocks, nc, na, nb, split, prune, stamp, maskW, s);
        else if (R == 2) launch_cell_runs<float, 2>(G, blocks, nc, na, nb, split, prune, stamp, maskW, s);
        else launch_cell_runs<float, 4>(G, blocks, nc, na, nb, split, prune, stamp, maskW, s);
    } else {
        if (R == 1) launch_cell_runs<double, 1>(G, blocks, nc, na, nb, split, prune, stamp, maskW, s);
        else if (R == 2) launch_cell_runs<double, 2>(G, blocks, nc, na, nb, split, prune, stamp, maskW, s);
        else launch_cell_runs<double, 4>(G, blocks, nc, na, nb, split, prune, stamp, maskW, s);
    }
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

__global__ void k_split_windows(const uint4 *__restrict__ sub, uint32_t nsub, uint32_t S, uint4 *__restrict__ out) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= nsub * S) return;
    const uint4 t = sub[i / S];
    out[i] = make_uint4(t.x, t.y, t.z, i % S + 1u);
}

int split_windows(const uint4 *sub, uint32_t nsub, uint32_t S, uint4 *out, hipStream_t s) {
    if (nsub * S == 0) return GDSJ_OK;
    k_split_windows<<<ceil_div(nsub * S, 256), 256, 0, s>>>(sub, nsub, S, out);
    GDSJ_HIP(hipGetLastError());
    return GDSJ_OK;
}

__global__ void k_split_offsets(const uint4 *__restrict__ tiles, uint32_t nsub, uint32_t S,
                                const uint32_t *__restrict__ ids, const uint64_t *__restrict__ offsets,
                                const uint32_t *__restrict__ cnt, uint64_t *__restrict__ off) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;  // (sub-tile, row)
    if (i >= nsub * SPLIT_ROWS) return;
    const uint32_t sub = i / SPLIT_ROWS, r = i % SPLIT_ROWS;
    const uint4 t = tiles[(uint64_t)sub * S];
    if (r >= t.z - t.y) return;
    uint64_t base = offsets[ids[t.y + r]];
    for (uint32_t k = 0; k < S; k++) {
        const uint64_t at = ((uint64_t)sub * S + k) * SPLIT_ROWS + r;
        off[at] = base;
        base += cnt[at];
    }
}

int split_offsets(const uint4 *tiles, uint32_t nsub, uint32_t S, const uint32_t *ids, const uint64_t *offsets,
                  const uint32_t *cnt, uint64_t *off, hipStream_t s) {
    if (nsub == 0) return GDSJ_OK;
    k_split_offsets<<<ceil_div(nsub * SPLIT_ROWS, 256), 256, 0, s>>>(tiles, nsub, S, ids, offsets, cnt, off);
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
