// join_kernels.h -- distance join kernel templates (instantiated per DTYPE
// in join_f32.hip / join_f64.hip).
//
// Work decomposition (replaces one thread per query, kernel.cu:132-305):
// one wave64 per tile = up to 64 query points of one grid cell.  The tile's
// candidates are the concatenated neighbor runs of its cell; they are packed
// 64 to a chunk (lane = candidate), up to CMAX chunks held in VGPRs as SoA
// coordinates, and every query of the tile (its coordinates wave-uniform)
// is tested against them.  A hit ballot gives the count (popcount) or, in
// the fill pass, each hit's slot (mbcnt) inside the query's CSR row, so the
// table is written without global atomics and without a sort.
//
// Distance arithmetic follows kernel.cu:307-365 exactly (AccMode); the
// predicate sqrt(acc) <= eps is evaluated as acc <= thr with thr the
// largest DTYPE value whose correctly rounded square root is <= eps.
#pragma once
#include "join.h"

namespace gdsj {

template <typename T, int D>
struct CMaxFor {
    static constexpr int words = (int)sizeof(T) / 4;
    static constexpr int raw = D > 0 ? 48 / (D * words) : 1;
    static constexpr int value = raw < 1 ? 1 : (raw > 8 ? 8 : raw);
};

// Squared distance in the reference's order.  q: query coords (uniform),
// c: candidate coords (this lane).  Returns false early (acc > thr on every
// lane of the wave that is still a candidate) when short-circuiting; the
// early exit never changes a result because partial sums only grow.
template <typename T, int D, int MODE, int ILPV>
__device__ __forceinline__ T dist_reg(const T (&q)[D], const T (&c)[D], T thr, bool live, bool sc,
                                      bool &alive) {
    alive = true;
    if (MODE == ACC_SEQ) {
        T acc = 0;
#pragma unroll
        for (int l = 0; l < D; l++) {
            T t = c[l] - q[l];
            acc = fma_rn(t, t, acc);
            if (sc && ((l & 7) == 7) && l + 1 < D) {
                if (ballot(live && acc <= thr) == 0) {
                    alive = false;
                    return acc;
                }
            }
        }
        return acc;
    } else if (MODE == ACC_ILP_SC) {
        // r0 = fma(t_l,t_l,r0); r_j = round(t*t); r0 = r0 + r_j (j=1..ILP-1)
        T r0 = 0;
#pragma unroll
        for (int l = 0; l < D; l += ILPV) {
            T t0 = c[l] - q[l];
            r0 = fma_rn(t0, t0, r0);
#pragma unroll
            for (int j = 1; j < ILPV; j++) {
                if (l + j < D) {
                    T t = c[l + j] - q[l + j];
                    T p = t * t;
                    r0 = r0 + p;
                }
            }
            if (sc && l + ILPV < D) {
                if (ballot(live && r0 <= thr) == 0) {
                    alive = false;
                    return r0;
                }
            }
        }
        return r0;
    } else {
        T r[ILPV];
#pragma unroll
        for (int j = 0; j < ILPV; j++) r[j] = 0;
#pragma unroll
        for (int l = 0; l < D; l += ILPV) {
#pragma unroll
            for (int j = 0; j < ILPV; j++)
                if (l + j < D) {
                    T t = c[l + j] - q[l + j];
                    r[j] = fma_rn(t, t, r[j]);
                }
        }
        T acc = r[0];
#pragma unroll
        for (int j = 1; j < ILPV; j++) acc = acc + r[j];
        return acc;
    }
}

// Same contract with runtime dims and ILP; coordinates read from memory.
template <typename T>
__device__ __forceinline__ T dist_mem(const T *__restrict__ xs, uint64_t ld, uint32_t qpos, uint32_t cpos,
                                      uint32_t d, int mode, uint32_t ilp) {
    if (mode == ACC_SEQ) {
        T acc = 0;
        for (uint32_t l = 0; l < d; l++) {
            T t = xs[l * ld + cpos] - xs[l * ld + qpos];
            acc = fma_rn(t, t, acc);
        }
        return acc;
    }
    if (mode == ACC_ILP_SC) {
        T r0 = 0;
        for (uint32_t l = 0; l < d; l += ilp) {
            T t0 = xs[l * ld + cpos] - xs[l * ld + qpos];
            r0 = fma_rn(t0, t0, r0);
            for (uint32_t j = 1; j < ilp && l + j < d; j++) {
                T t = xs[(l + j) * ld + cpos] - xs[(l + j) * ld + qpos];
                T p = t * t;
                r0 = r0 + p;
            }
        }
        return r0;
    }
    // ACC_ILP_NOSC: partial j accumulates dims j, j+ilp, j+2*ilp, ...; fold in j order
    T acc = 0;
    for (uint32_t j = 0; j < ilp; j++) {
        T rj = 0;
        for (uint32_t l = j; l < d; l += ilp) {
            T t = xs[l * ld + cpos] - xs[l * ld + qpos];
            rj = fma_rn(t, t, rj);
        }
        acc = (j == 0) ? rj : acc + rj;
    }
    return acc;
}

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
__device__ __forceinline__ uint2 combo_run(uint32_t t, const uint32_t (&cc)[MAXK], const double *glo,
                                           const double *ghi, const GridArgs &ga,
                                           const uint32_t *__restrict__ cellStart,
                                           const uint64_t *__restrict__ hkeys,
                                           const uint32_t *__restrict__ hvals) {
    uint64_t base = 0;
    bool ok = true;
    double gap = 0.0;
    uint32_t tt = t;
    for (uint32_t j = 1; j < ga.k; j++) {
        const uint32_t dlt = tt % 3u;
        tt /= 3u;
        const int64_t cj = (int64_t)cc[j] + (int64_t)dlt - 1;
        ok &= (cj >= 0) && (cj < (int64_t)ga.nCells[j]);
        gap += dlt == 0 ? glo[j] : (dlt == 2 ? ghi[j] : 0.0);
        base += (uint64_t)cj * ga.mult[j];
    }
    uint2 run = make_uint2(0u, 0u);
    if (!ok || gap > ga.prune_lim) return run;
    uint32_t x0 = cc[0] == 0 ? 0u : cc[0] - 1u;
    uint32_t x1 = cc[0] + 1u < ga.nCells[0] ? cc[0] + 1u : ga.nCells[0] - 1u;
    if (x0 < cc[0] && gap + glo[0] > ga.prune_lim) x0 = cc[0];
    if (x1 > cc[0] && gap + ghi[0] > ga.prune_lim) x1 = cc[0];
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

// Squared gaps from the tile's query bounding box to the cell faces in each
// indexed dim: glo[j] towards coordinate c_j - 1, ghi[j] towards c_j + 1.
// Face positions are widened by the classification rounding of
// (x - min) / eps (a few units of DTYPE roundoff), so the gaps never
// overstate the distance to any point of the neighbor cell.
template <typename T>
__device__ __forceinline__ void tile_gaps(const T *__restrict__ xs, uint64_t ld, uint32_t q0, uint32_t nq,
                                          const uint32_t (&cc)[MAXK], const GridArgs &ga, double *glo,
                                          double *ghi) {
    const uint32_t lane = lane_id();
    for (uint32_t j = 0; j < ga.k; j++) {
        double v = lane < nq ? (double)xs[(uint64_t)j * ld + q0 + lane] : 0.0;
        double mn = lane < nq ? v : 1e300, mx = lane < nq ? v : -1e300;
#pragma unroll
        for (int off = 32; off > 0; off >>= 1) {
            mn = fmin(mn, __shfl_xor(mn, off, GDSJ_WAVE));
            mx = fmax(mx, __shfl_xor(mx, off, GDSJ_WAVE));
        }
        const double lo_face = ga.minArr[j] + (double)cc[j] * ga.eps;
        const double hi_face = lo_face + ga.eps;
        const double tol = 8.0 * ga.ulp * (fabs(ga.minArr[j]) + fabs(hi_face) + 2.0 * ga.eps);
        const double gl = fmax(0.0, mn - lo_face - tol);
        const double gh = fmax(0.0, hi_face - tol - mx);
        if (lane == 0) {
            glo[j] = gl * gl;
            ghi[j] = gh * gh;
        }
    }
    __builtin_amdgcn_wave_barrier();
}

template <typename T, int D, int MODE, int ILPV, bool FILL>
__global__ __launch_bounds__(256) void k_join(const T *__restrict__ xs, const uint32_t *__restrict__ ids,
                                              const uint4 *__restrict__ tiles,
                                              const uint32_t *__restrict__ cellStart,
                                              const uint32_t *__restrict__ cellCoord,
                                              const uint64_t *__restrict__ hkeys,
                                              const uint32_t *__restrict__ hvals, GridArgs ga, uint32_t ncombo,
                                              uint32_t *__restrict__ counts, uint64_t *__restrict__ tileWork,
                                              const uint64_t *__restrict__ offsets, uint32_t *__restrict__ out,
                                              uint32_t *__restrict__ status, uint64_t ld, uint64_t cap,
                                              uint32_t ntiles, double thr_d, uint32_t dd, uint32_t ilp_rt,
                                              uint32_t sc, uint32_t xcd_map, uint32_t prune) {
    constexpr int CMAX = CMaxFor<T, D>::value;
    const uint32_t lane = lane_id();
    const uint32_t lblock = xcd_map ? xcd_block(blockIdx.x, gridDim.x) : blockIdx.x;
    const uint32_t tile = uniform((uint32_t)(lblock * (blockDim.x / GDSJ_WAVE) + threadIdx.x / GDSJ_WAVE));
    if (tile >= ntiles) return;
    const T thr = (T)thr_d;
    const uint4 tl = tiles[tile];
    const uint32_t cell = uniform(tl.x), q0 = uniform(tl.y), nq = uniform(tl.z - tl.y);
    uint32_t cc[MAXK];
    for (uint32_t j = 0; j < ga.k; j++) cc[j] = uniform(cellCoord[(uint64_t)j * ga.ncell + cell]);
    // per-wave squared face gaps of this tile (wave-uniform, kept in LDS)
    __shared__ double s_gap[256 / GDSJ_WAVE][2][MAXK];
    double *glo = s_gap[threadIdx.x / GDSJ_WAVE][0];
    double *ghi = s_gap[threadIdx.x / GDSJ_WAVE][1];
    if (prune) {
        tile_gaps<T>(xs, ld, q0, nq, cc, ga, glo, ghi);
    } else {
        if (lane < MAXK) {
            glo[lane] = 0.0;
            ghi[lane] = 0.0;
        }
        __builtin_amdgcn_wave_barrier();
    }
    const uint32_t qid_l = lane < nq ? ids[q0 + lane] : 0u;
    uint32_t cnt_l = 0;
    uint64_t cur_l = 0;
    uint64_t work = 0;
    if (FILL && lane < nq) cur_l = offsets[qid_l];
    bool overflow = false;

    for (uint32_t cw = 0; cw < ncombo; cw += GDSJ_WAVE) {
        const uint32_t nrw = uniform(ncombo - cw < (uint32_t)GDSJ_WAVE ? ncombo - cw : (uint32_t)GDSJ_WAVE);
        uint2 run = lane < nrw ? combo_run(cw + lane, cc, glo, ghi, ga, cellStart, hkeys, hvals)
                               : make_uint2(0u, 0u);
        uint32_t len = run.y - run.x;
        uint32_t incl = len;
#pragma unroll
        for (int off = 1; off < GDSJ_WAVE; off <<= 1) {
            uint32_t o = __shfl_up(incl, off, GDSJ_WAVE);
            if (lane >= (uint32_t)off) incl += o;
        }
        const uint32_t total = uniform(shfl_u32(incl, nrw - 1));
        if (total == 0) continue;
        work += total;
        const uint32_t rexcl = incl - len;
        const uint32_t rstart = run.x;
        // a valid position for lanes past the end: first candidate of the window
        const uint32_t fpos = uniform(cand_pos(0u, nrw, rstart, rexcl));

        for (uint32_t cb = 0; cb < total; cb += GDSJ_WAVE * CMAX) {
            const uint32_t nch = uniform((total - cb + GDSJ_WAVE - 1) / GDSJ_WAVE < (uint32_t)CMAX
                                             ? (total - cb + GDSJ_WAVE - 1) / GDSJ_WAVE
                                             : (uint32_t)CMAX);
            T c[CMAX][D > 0 ? D : 1];
            uint32_t cpos[CMAX];
            uint32_t cid[CMAX];
            bool valid[CMAX];
#pragma unroll
            for (int ch = 0; ch < CMAX; ch++) {
                const uint32_t g = cb + ch * GDSJ_WAVE + lane;
                valid[ch] = g < total;
                uint32_t sp = cand_pos(valid[ch] ? g : 0u, nrw, rstart, rexcl);
                sp = valid[ch] ? sp : fpos;
                cpos[ch] = sp;
                if (D > 0 && ch < (int)nch) {
#pragma unroll
                    for (int l = 0; l < (D > 0 ? D : 1); l++) c[ch][l] = xs[(uint64_t)l * ld + sp];
                }
                if (FILL) cid[ch] = (ch < (int)nch) ? ids[sp] : 0u;
            }
            for (uint32_t qi = 0; qi < nq; qi++) {
                const uint32_t qpos = q0 + qi;
                T q[D > 0 ? D : 1];
                if (D > 0) {
#pragma unroll
                    for (int l = 0; l < (D > 0 ? D : 1); l++) q[l] = xs[(uint64_t)l * ld + qpos];
                }
                uint32_t tot = 0;
                uint64_t cur = 0;
                if (FILL) {
                    uint32_t lo = __builtin_amdgcn_readlane((uint32_t)cur_l, qi);
                    uint32_t hi = __builtin_amdgcn_readlane((uint32_t)(cur_l >> 32), qi);
                    cur = ((uint64_t)hi << 32) | lo;
                }
#pragma unroll
                for (int ch = 0; ch < CMAX; ch++) {
                    if (ch < (int)nch) {
                        bool hit;
                        if (D > 0) {
                            bool alive;
                            T acc = dist_reg<T, (D > 0 ? D : 1), MODE, (ILPV > 0 ? ILPV : 1)>(
                                q, c[ch], thr, valid[ch], sc != 0, alive);
                            hit = alive && valid[ch] && acc <= thr;
                        } else {
                            T acc = dist_mem<T>(xs, ld, qpos, cpos[ch], dd, MODE, ilp_rt);
                            hit = valid[ch] && acc <= thr;
                        }
                        const uint64_t m = ballot(hit);
                        if (FILL) {
                            if (hit) {
                                const uint64_t pos = cur + mbcnt(m);
                                if (pos < cap) out[pos] = cid[ch]; else overflow = true;
                            }
                            cur += popc64(m);
                        } else {
                            tot += popc64(m);
                        }
                    }
                }
                if (FILL) {
                    if (lane == qi) cur_l = cur;
                } else {
                    if (lane == qi) cnt_l += tot;
                }
            }
        }
    }
    if (!FILL && lane < nq) counts[qid_l] = cnt_l;
    if (!FILL && tileWork && lane == 0) tileWork[tile] = work * nq;
    if (FILL && overflow) atomicOr(status, 1u);
}

}  // namespace gdsj
