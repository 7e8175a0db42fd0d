// neighbors.h -- adjacent-cell enumeration shared by the run-list kernel
// and the join kernels: the reference's candidate cells (kernel.cu:180-300
// with kernelloops.h) found through the cell lookup table, plus the
// query-box pruning that skips cells no query of a tile can reach.
#pragma once
#include "common.h"
#include "grid.h"

namespace gdsj {

__device__ __forceinline__ uint32_t shfl_u32(uint32_t v, uint32_t src) { return __shfl(v, (int)src, GDSJ_WAVE); }

// lane -> candidate sorted position for chunk-relative index g within a
// window of up to 64 runs (run i held by lane i: start, exclusive prefix)
__device__ __forceinline__ uint32_t cand_pos(uint32_t g, uint32_t nrw, uint32_t rstart, uint32_t rexcl) {
    uint32_t idx = 0;
#pragma unroll
    for (uint32_t step = 32; step > 0; step >>= 1) {
        uint32_t cnd = idx + step;
        uint32_t e = shfl_u32(rexcl, cnd < 64 ? cnd : 63);
        if (cnd < nrw && e <= g) idx = cnd;
    }
    return shfl_u32(rstart, idx) + (g - shfl_u32(rexcl, idx));
}

// Neighbor run of combination t (digits of t in base 3 = offsets -1/0/+1 of
// dims 1..k-1, dim k-1 most significant, so runs come in ascending linear
// id): the up to three cells (c0-1, c0, c0+1) of that row, which are
// consecutive in the sorted cell order.  Coordinates outside
// [0, nCells-1] are skipped, as the reference's clamps do (kernel.cu:183-187).
// Cells whose squared gap to the tile's query box exceeds prune_lim are
// skipped too: no pair with them can pass the predicate.
//
// stamp (unicomp): keep only the cells this cell owns, so each unordered
// pair of adjacent cells is evaluated from one side only.
//   1 (STAMP, stamploops.h:2-201): a pair is owned by the cell whose
//     coordinate is odd in the highest dimension where they differ: a row of
//     dims 1..k-1 whose highest nonzero offset is in dim j is kept iff c_j is
//     odd; the cell's own row keeps the dim-0 neighbors iff c_0 is odd.
//   2 (LINEARSTAMP, kernel.cu:272-289): a pair is owned by the cell with the
//     smaller linear id: a row is kept iff its highest nonzero offset is +1
//     (the row's linear ids exceed the cell's); the own row keeps x + 1.
// The own cell is always kept.
// The row of combination t: its linear id without dim 0 (base) and the
// dim-0 range [x0, x1] left after the clamps and the pruning; live is false
// when the whole row is skipped.
struct ComboRow {
    uint64_t base;
    uint32_t x0, x1;
    bool live;
};

__device__ __forceinline__ ComboRow combo_decode(uint32_t t, const uint32_t (&cc)[MAXK], const double *glo,
                                                 const double *ghi, const GridArgs &ga, int stamp) {
    ComboRow w;
    w.base = 0;
    bool ok = true;
    double gap = 0.0;
    uint32_t tt = t;
    int top = -1;       // highest dim (>= 1) with a nonzero offset
    uint32_t topd = 1;  // its offset digit (0: -1, 2: +1)
    for (uint32_t j = 1; j < ga.k; j++) {
        const uint32_t dlt = tt % 3u;
        tt /= 3u;
        const int64_t cj = (int64_t)cc[j] + (int64_t)dlt - 1;
        ok &= (cj >= 0) && (cj < (int64_t)ga.nCells[j]);
        gap += dlt == 0 ? glo[j] : (dlt == 2 ? ghi[j] : 0.0);
        w.base += (uint64_t)cj * ga.mult[j];
        if (dlt != 1) {
            top = (int)j;
            topd = dlt;
        }
    }
    const bool owned = stamp == 1 ? !(top >= 0 && !(cc[top] & 1u)) : (stamp == 2 ? !(top >= 0 && topd == 0) : true);
    w.live = ok && !(gap > ga.prune_lim) && owned;
    w.x0 = cc[0] == 0 ? 0u : cc[0] - 1u;
    w.x1 = cc[0] + 1u < ga.nCells[0] ? cc[0] + 1u : ga.nCells[0] - 1u;
    if (w.x0 < cc[0] && gap + glo[0] > ga.prune_lim) w.x0 = cc[0];
    if (w.x1 > cc[0] && gap + ghi[0] > ga.prune_lim) w.x1 = cc[0];
    if (stamp == 1 && top < 0 && !(cc[0] & 1u)) w.x0 = w.x1 = cc[0];
    if (stamp == 2 && top < 0) w.x0 = cc[0];
    return w;
}

__device__ __forceinline__ uint2 combo_run(uint32_t t, const uint32_t (&cc)[MAXK], const double *glo,
                                           const double *ghi, const GridArgs &ga,
                                           const uint32_t *__restrict__ cellStart,
                                           const uint64_t *__restrict__ cellKey,
                                           const uint64_t *__restrict__ hkeys,
                                           const uint32_t *__restrict__ hvals, int stamp = 0) {
    const ComboRow w = combo_decode(t, cc, glo, ghi, ga, stamp);
    uint2 run = make_uint2(0u, 0u);
    if (!w.live) return run;
    const uint64_t base = w.base;
    const uint32_t x0 = w.x0, x1 = w.x1;
    if (ga.hbits) {
        // one probe for the row, then the cells x0..x1 of it by key: the
        // row's first four keys are read together (sorted, so the number
        // below a bound is its lower bound); longer rows continue galloping
        const int32_t r0 = cell_find(hkeys, hvals, ga.hbits, base);
        if (r0 < 0) return run;
        const uint32_t r = (uint32_t)r0;
        const uint64_t w0 = base + x0, w1 = base + x1 + 1;
        uint32_t nf = 0, ne = 0;
#pragma unroll
        for (uint32_t i = 0; i < 4; i++) {
            const uint64_t key = r + i < ga.ncell ? cellKey[r + i] : ~0ull;
            nf += key < w0 ? 1u : 0u;
            ne += key < w1 ? 1u : 0u;
        }
        uint32_t f = r + nf, e = r + ne;
        if (nf == 4) f = key_lower_bound(cellKey, r + 4, ga.ncell, w0);
        if (ne == 4) e = key_lower_bound(cellKey, f, ga.ncell, w1);
        if (f < e) run = make_uint2(cellStart[f], cellStart[e]);
        return run;
    }
    int32_t first = -1, last = -1;
    for (uint32_t x = x0; x <= x1; x++) {
        const int32_t ci = cell_find(hkeys, hvals, ga.hbits, base + x);
        if (ci >= 0) {
            if (first < 0) first = ci;
            last = ci;
        }
    }
    if (first >= 0) run = make_uint2(cellStart[first], cellStart[last + 1]);
    return run;
}

// The runs of R decoded rows w[0..R) (combo_run's lookup half): each
// lookup step (row probe, row cells, cell starts) issues its loads for all R
// before the first of them is used, so a wave waits for R dependent chains
// together instead of one after another.  Hashed grids: the row's slot gives
// its first cell and its length; the dim-0 coordinates of the row's first 8
// cells (cellCoord's dim-0 slice, sorted within the row) come in two 16-byte
// loads and the run's bounds are counts over them; rows past 8 cells continue
// by a binary search.
template <int R>
__device__ __forceinline__ void combo_lookup(const ComboRow (&w)[R], const GridArgs &ga,
                                             const uint32_t *__restrict__ cellStart,
                                             const uint32_t *__restrict__ cellC0,
                                             const uint64_t *__restrict__ hkeys,
                                             const uint32_t *__restrict__ hvals, uint2 (&run)[R]) {
#pragma unroll
    for (int r = 0; r < R; r++) run[r] = make_uint2(0u, 0u);
    int32_t first[R], last[R];  // cell range of the run, -1: none
    if (ga.hbits) {
        const uint32_t mask = (1u << ga.hbits) - 1u;
        const ulonglong2 *slots = reinterpret_cast<const ulonglong2 *>(hkeys);
        uint32_t h[R], len[R];
        ulonglong2 e[R];
        bool pend[R];
        // first probe of every row (a table at most half full mostly
        // answers there)
#pragma unroll
        for (int r = 0; r < R; r++) {
            h[r] = cell_hash(w[r].base, ga.hbits);
            e[r] = make_ulonglong2(0ull, 0ull);
            if (w[r].live) e[r] = slots[h[r]];
        }
        // the longer probe sequences, all rows of the wave in lock step (a
        // wave waits for its longest sequence once, not once per r); every
        // sequence ends at the row's slot or at an empty slot, and at least
        // half of the slots are empty
        for (uint32_t probe = 0;; probe++) {
            bool any = false;
#pragma unroll
            for (int r = 0; r < R; r++) {
                first[r] = -1;
                len[r] = 0;
                pend[r] = false;
                if (!w[r].live || e[r].x == 0) continue;
                if (e[r].x == w[r].base + 1) {
                    first[r] = (int32_t)(uint32_t)e[r].y;
                    len[r] = (uint32_t)(e[r].y >> 32);
                    continue;
                }
                pend[r] = true;
                any = true;
            }
            if (!ballot(any) || probe > mask) break;
#pragma unroll
            for (int r = 0; r < R; r++)
                if (pend[r]) {
                    h[r] = (h[r] + 1u) & mask;
                    e[r] = slots[h[r]];
                }
        }
        // dim-0 coordinates of the row's first 8 cells (sorted); the slice
        // is followed by the other dims' and 64 bytes of padding, so the
        // loads stay inside the buffer
        uint32_t cv[R][8];
#pragma unroll
        for (int r = 0; r < R; r++) {
#pragma unroll
            for (int i = 0; i < 8; i++) cv[r][i] = 0xFFFFFFFFu;
            if (first[r] >= 0) {
                uint4 v;
                __builtin_memcpy(&v, cellC0 + first[r], 16);
                cv[r][0] = v.x;
                cv[r][1] = v.y;
                cv[r][2] = v.z;
                cv[r][3] = v.w;
                if (len[r] > 4) {
                    __builtin_memcpy(&v, cellC0 + first[r] + 4, 16);
                    cv[r][4] = v.x;
                    cv[r][5] = v.y;
                    cv[r][6] = v.z;
                    cv[r][7] = v.w;
                }
            }
        }
#pragma unroll
        for (int r = 0; r < R; r++) {
            if (first[r] < 0) continue;
            const uint32_t rr = (uint32_t)first[r], n8 = len[r] < 8u ? len[r] : 8u;
            uint32_t nf = 0, ne = 0;
#pragma unroll
            for (uint32_t i = 0; i < 8; i++) {
                nf += i < n8 && cv[r][i] < w[r].x0 ? 1u : 0u;
                ne += i < n8 && cv[r][i] <= w[r].x1 ? 1u : 0u;
            }
            uint32_t f = rr + nf, en = rr + ne;
            if (nf == 8u) f = u32_lower_bound(cellC0, rr + 8u, rr + len[r], w[r].x0);
            if (ne == 8u) en = u32_lower_bound(cellC0, f > rr + 8u ? f : rr + 8u, rr + len[r], w[r].x1 + 1u);
            // as a cell range [first, last]
            first[r] = f < en ? (int32_t)f : -1;
            last[r] = (int32_t)en - 1;
        }
    } else {
        // direct table: the up to three cells of every row
        int32_t ci[R][3];
#pragma unroll
        for (int r = 0; r < R; r++)
#pragma unroll
            for (uint32_t x = 0; x < 3; x++) {
                ci[r][x] = -1;
                if (w[r].live && w[r].x0 + x <= w[r].x1) ci[r][x] = (int32_t)hvals[w[r].base + w[r].x0 + x];
            }
#pragma unroll
        for (int r = 0; r < R; r++) {
            first[r] = -1;
            last[r] = -1;
#pragma unroll
            for (uint32_t x = 0; x < 3; x++) {
                if (ci[r][x] >= 0) {
                    if (first[r] < 0) first[r] = ci[r][x];
                    last[r] = ci[r][x];
                }
            }
        }
    }
    // the cell starts of every run
#pragma unroll
    for (int r = 0; r < R; r++)
        if (first[r] >= 0) run[r] = make_uint2(cellStart[first[r]], cellStart[last[r] + 1]);
}

// Cell coordinates of a cell in the k indexed dims, all loads issued
// together (entries past k repeat dim 0 and are not used).
__device__ __forceinline__ void load_cell_coords(const uint32_t *__restrict__ cellCoord, uint32_t ncell,
                                                 uint32_t cell, uint32_t k, uint32_t (&cc)[MAXK]) {
#pragma unroll
    for (int j = 0; j < MAXK; j++) {
        const uint32_t jj = (uint32_t)j < k ? (uint32_t)j : 0u;
        cc[j] = uniform(cellCoord[(uint64_t)jj * ncell + cell]);
    }
}

// Squared gaps from the bounding box of the n query points at sorted
// positions [q0, q0+n) to the faces of their cell in each indexed dim:
// glo[j] towards coordinate c_j - 1, ghi[j] towards c_j + 1.  Face
// positions are widened by the classification rounding of (x - min) / eps
// (a few units of DTYPE roundoff), so the gaps never overstate the distance
// to any point of the neighbor cell.  Written by lane 0 (wave-uniform).
// FEW (the run-list kernel: a wave per cell, typically a few points): the
// min / max in DTYPE over ceil(log2 n) butterfly steps instead of six steps
// in double; it needs more registers, which the join kernels cannot spare.
template <typename T>
__device__ __forceinline__ void store_gaps(double mn, double mx, uint32_t j, const uint32_t (&cc)[MAXK],
                                           const GridArgs &ga, double *glo, double *ghi) {
    const double lo_face = ga.minArr[j] + (double)cc[j] * ga.eps;
    const double hi_face = lo_face + ga.eps;
    const double tol = 8.0 * ga.ulp * (fabs(ga.minArr[j]) + fabs(hi_face) + 2.0 * ga.eps);
    const double gl = fmax(0.0, mn - lo_face - tol);
    const double gh = fmax(0.0, hi_face - tol - mx);
    if (lane_id() == 0) {
        glo[j] = gl * gl;
        ghi[j] = gh * gh;
    }
}

template <typename T, bool FEW = false>
__device__ __forceinline__ void tile_gaps(const T *__restrict__ xs, uint64_t ld, uint32_t q0, uint32_t n,
                                          const uint32_t (&cc)[MAXK], const GridArgs &ga, double *glo,
                                          double *ghi) {
    const uint32_t lane = lane_id();
    if (n <= (uint32_t)GDSJ_WAVE) {
        // one point per lane: every dim's loads are issued before the first
        // reduction (dims past k read dim 0 again and are not used)
        // (lanes past n read point 0 of the range, n >= 1, and are masked below)
        const uint32_t pl = lane < n ? lane : 0u;
        T raw[MAXK];
#pragma unroll
        for (int j = 0; j < MAXK; j++) {
            const uint32_t jj = (uint32_t)j < ga.k ? (uint32_t)j : 0u;
            raw[j] = xs[(uint64_t)jj * ld + q0 + pl];
        }
        if constexpr (FEW) {
            // exact in DTYPE, so the double conversion after it gives the
            // same bounds; lanes [0, 2^steps) cover the n points, lanes past
            // n hold the neutral values
            const uint32_t steps = n <= 1u ? 0u : 32u - (uint32_t)__builtin_clz(n - 1u);
#pragma unroll
            for (int j = 0; j < MAXK; j++) {
                if ((uint32_t)j >= ga.k) break;
                T mn = lane < n ? raw[j] : (T)__builtin_inf(), mx = lane < n ? raw[j] : -(T)__builtin_inf();
                for (uint32_t st = 0; st < steps; st++) {
                    const int off = 1 << st;
                    mn = fmin(mn, __shfl_xor(mn, off, GDSJ_WAVE));
                    mx = fmax(mx, __shfl_xor(mx, off, GDSJ_WAVE));
                }
                store_gaps<T>((double)mn, (double)mx, (uint32_t)j, cc, ga, glo, ghi);
            }
        } else {
            double v[MAXK];
#pragma unroll
            for (int j = 0; j < MAXK; j++) v[j] = (double)raw[j];
#pragma unroll
            for (int j = 0; j < MAXK; j++) {
                if ((uint32_t)j >= ga.k) break;
                double mn = lane < n ? v[j] : 1e300, mx = lane < n ? v[j] : -1e300;
#pragma unroll
                for (int off = 32; off > 0; off >>= 1) {
                    mn = fmin(mn, __shfl_xor(mn, off, GDSJ_WAVE));
                    mx = fmax(mx, __shfl_xor(mx, off, GDSJ_WAVE));
                }
                store_gaps<T>(mn, mx, (uint32_t)j, cc, ga, glo, ghi);
            }
        }
        __builtin_amdgcn_wave_barrier();
        return;
    }
    // large ranges (whole cells of tens of thousands of points): 8 loads per
    // lane issued before their min / max, so a dim costs n / 512 memory
    // round trips instead of n / 64 (min and max are exact in any order)
    constexpr uint32_t U = 8;
    for (uint32_t j = 0; j < ga.k; j++) {
        double mn = 1e300, mx = -1e300;
        const T *col = xs + (uint64_t)j * ld + q0;
        uint32_t b = 0;
        for (; b + U * GDSJ_WAVE <= n; b += U * GDSJ_WAVE) {
            T v[U];
#pragma unroll
            for (uint32_t u = 0; u < U; u++) v[u] = col[b + u * GDSJ_WAVE + lane];
#pragma unroll
            for (uint32_t u = 0; u < U; u++) {
                mn = fmin(mn, (double)v[u]);
                mx = fmax(mx, (double)v[u]);
            }
        }
        for (; b < n; b += GDSJ_WAVE) {
            if (b + lane < n) {
                const double v = (double)col[b + lane];
                mn = fmin(mn, v);
                mx = fmax(mx, v);
            }
        }
#pragma unroll
        for (int off = 32; off > 0; off >>= 1) {
            mn = fmin(mn, __shfl_xor(mn, off, GDSJ_WAVE));
            mx = fmax(mx, __shfl_xor(mx, off, GDSJ_WAVE));
        }
        store_gaps<T>(mn, mx, j, cc, ga, glo, ghi);
    }
    __builtin_amdgcn_wave_barrier();
}

}  // namespace gdsj
