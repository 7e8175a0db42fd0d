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
#include "neighbors.h"

namespace gdsj {

// candidate chunks held in VGPRs: about GDSJ_CMAX_WORDS coordinate words
#ifndef GDSJ_CMAX_WORDS
#define GDSJ_CMAX_WORDS 48
#endif
// the slotted pass: fewer (5 waves per SIMD instead of 4 at 6 dims; measured
// 6 % faster on gmm6d, where the hit-mask passes are faster with 48)
#ifndef GDSJ_CMAX_WORDS_SLOT
#define GDSJ_CMAX_WORDS_SLOT 36
#endif
template <typename T, int D, int W = GDSJ_CMAX_WORDS>
struct CMaxFor {
    static constexpr int words = (int)sizeof(T) / 4;
    static constexpr int raw = D > 0 ? W / (D * words) : 1;
    static constexpr int value = raw < 1 ? 1 : (raw > 8 ? 8 : raw);
};

// Row stride of Grid::xa for D dims (xa_stride, grid.h).
template <int D>
struct XaStride {
    static constexpr int value = (int)xa_stride((uint32_t)D);
};

template <int N>
struct LoadDw;  // N consecutive dwords
template <> struct LoadDw<1> { typedef uint32_t v __attribute__((ext_vector_type(1))); };
template <> struct LoadDw<2> { typedef uint32_t v __attribute__((ext_vector_type(2))); };
template <> struct LoadDw<3> { typedef uint32_t v __attribute__((ext_vector_type(3))); };
template <> struct LoadDw<4> { typedef uint32_t v __attribute__((ext_vector_type(4))); };

// Coordinates of the point at sorted position p from the row-major copy xa
// (stride XaStride<D> = D, so rows are only dword-aligned): loads of up to
// 16 bytes that cover the row (the hardware takes multi-dword loads at any
// dword alignment), instead of a load per dim from D columns.
template <typename T, int D>
__device__ __forceinline__ void load_row(const T *__restrict__ xa, uint64_t p, T (&c)[D]) {
    constexpr int S = XaStride<D>::value;
    constexpr int W = (int)sizeof(T) / 4;  // dwords per coordinate
    constexpr int NDW = D * W;              // dwords of the row
    const uint32_t *row = reinterpret_cast<const uint32_t *>(xa + p * (uint64_t)S);
    uint32_t dw[NDW];
#pragma unroll
    for (int i = 0; i < NDW; i += 4) {
        const int len = NDW - i < 4 ? NDW - i : 4;
        if (len == 4) {
            typename LoadDw<4>::v v;
            __builtin_memcpy(&v, row + i, 16);
#pragma unroll
            for (int k = 0; k < 4; k++) dw[(i + k) < NDW ? i + k : 0] = v[k];
        } else if (len == 3) {
            typename LoadDw<3>::v v;
            __builtin_memcpy(&v, row + i, 12);
#pragma unroll
            for (int k = 0; k < 3; k++) dw[(i + k) < NDW ? i + k : 0] = v[k];
        } else if (len == 2) {
            typename LoadDw<2>::v v;
            __builtin_memcpy(&v, row + i, 8);
#pragma unroll
            for (int k = 0; k < 2; k++) dw[(i + k) < NDW ? i + k : 0] = v[k];
        } else {
            dw[i] = row[i];
        }
    }
#pragma unroll
    for (int l = 0; l < D; l++) {
        if constexpr (sizeof(T) == 4) {
            c[l] = __uint_as_float(dw[l]);
        } else {
            const uint64_t b = ((uint64_t)dw[2 * l + 1] << 32) | dw[2 * l];
            c[l] = __builtin_bit_cast(T, b);
        }
    }
}

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

// The FMA-chain filter value of a pair: t_0^2, then one FMA per dim (2 D
// operations against the reference's 3 D - 1 with ILP), compared with the
// band of JoinArgs::ffLo / ffHi by the caller.  In dist_reg's ILP_SC order
// the wave leaves early (alive = false) after a group of ILPV dims when no
// live lane's partial sum is at most hi: the reference's partial sums then
// exceed the threshold, and they only grow.
template <typename T, int D, int MODE, int ILPV>
__device__ __forceinline__ T fma_chain(const T (&q)[D], const T (&c)[D], T hi, bool live, bool sc, bool &alive) {
    alive = true;
    T f = 0;
#pragma unroll
    for (int l = 0; l < D; l++) {
        const T t = c[l] - q[l];
        f = fma_rn(t, t, f);
        if (MODE == ACC_ILP_SC && sc && (l % ILPV) == ILPV - 1 && l + 1 < D) {
            if (ballot(live && f <= hi) == 0) {
                alive = false;
                return f;
            }
        }
    }
    return f;
}

// Same contract with runtime dims and ILP; coordinates read from memory.
template <typename T>
__device__ __forceinline__ T dist_mem(const T *__restrict__ xs, uint64_t ld, uint32_t qpos, uint32_t cpos,
                                      uint32_t d, int mode, uint32_t ilp) {
    if (mode == ACC_UNFUSED) {
        T acc = 0;
        for (uint32_t l = 0; l < d; l++) {
            T t = xs[l * ld + cpos] - xs[l * ld + qpos];
            T p = t * t;
            acc = acc + p;
        }
        return acc;
    }
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

// dist_mem with the coordinates of up to C dims loaded before their
// arithmetic (2 C loads in flight instead of a load round trip per dim); the
// floating-point operations and their order are dist_mem's.  Used by the
// Gram filter's exact rechecks, where each lane checks a different pair.
template <typename T, uint32_t C>
__device__ __forceinline__ T dist_mem_batch(const T *__restrict__ xs, uint64_t ld, uint32_t qpos, uint32_t cpos,
                                            uint32_t d, int mode, uint32_t ilp) {
    if (mode != ACC_ILP_NOSC) {
        // SEQ: an FMA per dim.  ILP_SC: an FMA for the first dim of each
        // group of ilp dims, a rounded product and an add for the others.
        T acc = 0;
        uint32_t g = 0;  // position of the dim within its group
        for (uint32_t l0 = 0; l0 < d; l0 += C) {
            T tv[C];
#pragma unroll
            for (uint32_t u = 0; u < C; u++) {
                const uint64_t l = l0 + u < d ? l0 + u : l0;  // in range; unused past d
                tv[u] = xs[l * ld + cpos] - xs[l * ld + qpos];
            }
#pragma unroll
            for (uint32_t u = 0; u < C; u++) {
                if (l0 + u < d) {
                    if (mode == ACC_SEQ || g == 0) {
                        acc = fma_rn(tv[u], tv[u], acc);
                    } else {
                        const T p = tv[u] * tv[u];
                        acc = acc + p;
                    }
                    g = g + 1 == ilp ? 0u : g + 1;
                }
            }
        }
        return acc;
    }
    // ACC_ILP_NOSC: partial j accumulates dims j, j+ilp, ...; fold in j order
    T acc = 0;
    for (uint32_t j = 0; j < ilp; j++) {
        T rj = 0;
        for (uint32_t l0 = j; l0 < d; l0 += C * ilp) {
            T tv[C];
#pragma unroll
            for (uint32_t u = 0; u < C; u++) {
                const uint64_t l = l0 + u * ilp < d ? l0 + u * ilp : l0;
                tv[u] = xs[l * ld + cpos] - xs[l * ld + qpos];
            }
#pragma unroll
            for (uint32_t u = 0; u < C; u++)
                if (l0 + u * ilp < d) rj = fma_rn(tv[u], tv[u], rj);
        }
        acc = (j == 0) ? rj : acc + rj;
    }
    return acc;
}

// dist_mem's arithmetic (in the same order) on two points stored as rows of
// a row-major copy, padded to a multiple of C dims: a point's coordinates
// share one or a few cache lines, where the SoA layout puts every dim in its
// own line.  Used by the Gram filter's exact rechecks, whose candidates are
// scattered.
template <typename T, uint32_t C>
__device__ __forceinline__ T dist_row(const T *__restrict__ qrow, const T *__restrict__ crow, uint32_t d, int mode,
                                      uint32_t ilp) {
    if (mode != ACC_ILP_NOSC) {
        T acc = 0;
        uint32_t g = 0;  // position of the dim within its group
        for (uint32_t l0 = 0; l0 < d; l0 += C) {
            T tv[C];
#pragma unroll
            for (uint32_t u = 0; u < C; u++) tv[u] = crow[l0 + u] - qrow[l0 + u];  // padded rows: in range
#pragma unroll
            for (uint32_t u = 0; u < C; u++) {
                if (l0 + u < d) {
                    if (mode == ACC_SEQ || g == 0) {
                        acc = fma_rn(tv[u], tv[u], acc);
                    } else {
                        const T p = tv[u] * tv[u];
                        acc = acc + p;
                    }
                    g = g + 1 == ilp ? 0u : g + 1;
                }
            }
        }
        return acc;
    }
    // ACC_ILP_NOSC: partial j accumulates dims j, j+ilp, ...; fold in j order
    T acc = 0;
    for (uint32_t j = 0; j < ilp; j++) {
        T rj = 0;
        for (uint32_t l = j; l < d; l += ilp) {
            const T t = crow[l] - qrow[l];
            rj = fma_rn(t, t, rj);
        }
        acc = (j == 0) ? rj : acc + rj;
    }
    return acc;
}

// dist_row with the next C dims of both rows loaded while the current C
// are summed (two batches in flight), for rechecks whose rows come from
// beyond L1.  Rows are padded to a multiple of 16 >= 2 C dims, so the batch
// after l0 (l0 + C < d) stays inside the row.
template <typename T, uint32_t C>
__device__ __forceinline__ T dist_row_pf(const T *__restrict__ qrow, const T *__restrict__ crow, uint32_t d,
                                         int mode, uint32_t ilp) {
    if (mode == ACC_ILP_NOSC) return dist_row<T, C>(qrow, crow, d, mode, ilp);
    T acc = 0;
    uint32_t g = 0;  // position of the dim within its group
    T cn[C], qn[C];
#pragma unroll
    for (uint32_t u = 0; u < C; u++) {
        cn[u] = crow[u];
        qn[u] = qrow[u];
    }
    for (uint32_t l0 = 0; l0 < d; l0 += C) {
        T tv[C];
#pragma unroll
        for (uint32_t u = 0; u < C; u++) tv[u] = cn[u] - qn[u];
        if (l0 + C < d) {
#pragma unroll
            for (uint32_t u = 0; u < C; u++) {
                cn[u] = crow[l0 + C + u];
                qn[u] = qrow[l0 + C + u];
            }
        }
#pragma unroll
        for (uint32_t u = 0; u < C; u++) {
            if (l0 + u < d) {
                if (mode == ACC_SEQ || g == 0) {
                    acc = fma_rn(tv[u], tv[u], acc);
                } else {
                    const T p = tv[u] * tv[u];
                    acc = acc + p;
                }
                g = g + 1 == ilp ? 0u : g + 1;
            }
        }
    }
    return acc;
}

// SORT=1 (kernel.cu:400-467): whether the reference evaluates candidate c
// (sorted position cpos, in cell ccell) for query q (position qpos) at all.
// Its cells are ordered by the value at the sort dim.
// sortMode 1 (every dim indexed, sort dim 0, kernel.cu:403-431): a cell with
// the query cell's dim-0 coordinate qc0 is walked whole; a cell on either
// side is walked from the query's side and the walk stops at the first value
// with |value - query| > eps.  The values on one side lie all below (above)
// the query's, so |value - query| grows along the walk and the walked points
// are those with |value - query| <= eps.
// sortMode 2 (sort dim k, kernel.cu:433-465): the reference's stepping
// search gives a start index; the points from it on are walked while
// value - query < eps, which, the values ascending, is every point at or
// past the start with value - query < eps.
template <typename T>
__device__ __forceinline__ bool sort_evaluated(const T *__restrict__ xs, uint64_t ld, const JoinArgs &a,
                                               uint32_t qpos, uint32_t cpos, uint32_t ccell, uint32_t qc0) {
    const T eps = (T)a.ga.eps;
    if (a.sortMode == 1) {
        if (a.cellCoord[ccell] == qc0) return true;
        const T diff = xs[cpos] - xs[qpos];
        return (diff < (T)0 ? -diff : diff) <= eps;
    }
    const T *col = xs + (uint64_t)a.sortDim * ld;
    const uint32_t s = a.cellStart[ccell];
    const int length = (int)(a.cellStart[ccell + 1] - s) - 1;
    const T q = col[qpos];
    int idx = length / 2;
    for (int step = (length + 3) / 4; step > 1; step = (step + 1) / 2) {
        const T dq = q - col[s + (uint32_t)idx];
        if (dq >= eps) {
            idx += step;
            if (idx > length) idx = length;
        } else {
            idx -= step;
            if (idx < 0) {
                idx = 0;
                step = 0;
            }
        }
    }
    if (idx > 0 && q - col[s + (uint32_t)idx - 1] <= eps) idx--;
    if (idx > 0 && q - col[s + (uint32_t)idx - 1] <= eps) idx--;
    return (int)(cpos - s) >= idx && col[cpos] - q < eps;
}

// One work tile of the join (one wave): the body shared by the per-join
// kernel below and the persistent batched kernel (batch.hip).
// CM: candidate chunks held in VGPRs (0: CMaxFor's default).
// STAMP (unicomp; count and fill passes): the tile's candidates are its own
// cell and the neighbor cells it owns.  A hit (q, p) with p outside the own
// cell is also the hit (p, q), which no other tile evaluates: the counting
// pass adds each candidate's hits over the tile's queries to counts[p] (one
// atomic per candidate and chunk batch), the fill pass collects the hitting
// queries of each candidate in a 64-bit mask and places them behind p's own
// entries with one atomic on p's cursor (their order within a row depends
// on the order the tiles run).
// QRW: most query coordinate words held in VGPRs (0: queries read from memory).
// UNSW: two copies of the fill pass's query loop (with and without the
// capacity check); a single copy keeps the batch kernel's registers down.
// BMAP: run-start bitmap for the candidate positions of precomputed runs
// (else the binary search over the window's runs only).
// XA: candidates' coordinates from the row-major copy (JoinArgs::xa, which
// must be set), else from the SoA columns.
// HM (precomputed runs, no unicomp): hit masks (JoinArgs::hmask).  The
// counting pass stores each query's chunk ballots; the fill pass takes its
// hits from them and loads no coordinates.  Both passes cut the candidates
// into the same windows and chunk batches, so chunk numbers agree.
template <typename T, int D, int MODE, int ILPV, int PASS, int CM = 0, bool STAMP = false, int QRW = 16,
          bool UNSW = true, bool BMAP = true, bool HM = false, bool GV = false, bool SRT = false, bool XA = false>
__device__ __forceinline__ void join_tile(const T *__restrict__ xs, const uint32_t *__restrict__ ids,
                                          const JoinArgs &a, const uint32_t tile,
                                          const uint64_t *__restrict__ hmr = nullptr) {
    constexpr int CMAX = CM > 0 ? CM : CMaxFor<T, D>::value;
    constexpr bool NEED_IDS = PASS != PASS_COUNT || STAMP;
    static_assert(!STAMP || PASS != PASS_SLOT, "unicomp runs with the counting and fill passes");
    static_assert(!HM || (!STAMP && PASS != PASS_SLOT), "hit masks: exact passes without unicomp");
    constexpr bool MFILL = HM && PASS == PASS_FILL;  // hits read from the masks
    const uint32_t lane = lane_id();
    const uint64_t ld = a.ld;
    // the candidates' coordinates by buffer loads (every column at a 32-bit
    // byte offset from xs) and their ids at 32-bit offsets, where that fits
    const bool off32 = D > 0 && (uint64_t)(D > 0 ? D : 1) * ld * sizeof(T) <= 0xFFFFFFFFull;
    const __amdgpu_buffer_rsrc_t xrs = buffer_rsrc(xs);
    const T thr = (T)a.thr;
    const T ffl = (T)a.ffLo, ffh = (T)a.ffHi;  // FMA-chain filter band (fma_band)
    const uint4 tl = a.tiles[tile];
    const uint32_t cell = uniform(tl.x), q0 = uniform(tl.y), nq = uniform(tl.z - tl.y);
    if (nq == 0) return;  // an empty sub-tile (split overflowed tiles)
    const GridArgs &ga = a.ga;
    const bool pre = a.runs != nullptr;
    uint32_t cc[MAXK];
    __shared__ double s_gap[256 / GDSJ_WAVE][2][MAXK];
    double *glo = s_gap[threadIdx.x / GDSJ_WAVE][0];
    double *ghi = s_gap[threadIdx.x / GDSJ_WAVE][1];
    uint32_t nwin;
    if (pre) {
        nwin = uniform(a.nruns[cell]);
    } else {
        nwin = a.ncombo;
        load_cell_coords(a.cellCoord, ga.ncell, cell, ga.k, cc);
        if (a.prune) {
            tile_gaps<T>((const T *)a.ixs, ld, q0, nq, cc, ga, glo, ghi);
        } else {
            if (lane < MAXK) {
                glo[lane] = 0.0;
                ghi[lane] = 0.0;
            }
            __builtin_amdgcn_wave_barrier();
        }
    }
    const uint32_t qid_l = lane < nq ? ids[q0 + lane] : 0u;
    const uint32_t sqc0 = SRT ? uniform(a.cellCoord[cell]) : 0u;  // SORT: the query cell's dim-0 coordinate
    __shared__ uint32_t s_qid[STAMP ? 256 / GDSJ_WAVE : 1][GDSJ_WAVE];
    constexpr int BMW = 64;  // run-start bitmap words per wave
    __shared__ uint64_t s_bm[BMAP ? 256 / GDSJ_WAVE : 1][BMAP ? BMW : 1];
    uint32_t own0 = 0, own1 = 0;  // sorted-position range of the tile's own cell
    if (STAMP) {
        own0 = uniform(a.cellStart[cell]);
        own1 = uniform(a.cellStart[cell + 1]);
        s_qid[threadIdx.x / GDSJ_WAVE][lane] = qid_l;
        __builtin_amdgcn_wave_barrier();
    }
    // low dims: lane l holds query l's coordinates; the query loop reads them
    // with readlane (no memory access per query)
    constexpr bool QREG = !MFILL && D > 0 && D * (int)(sizeof(T) / 4) <= QRW;
    // Gram-form filter (GV): coordinates relative to the tile's first query
    // (go, wave-uniform); lane l holds -2 (q_l - go) of query l (qv), and
    // gvLo and the band's midpoint gvMid less |q_l - go|^2 (tqa, tqm)
    constexpr bool GVF = GV && QREG && !MFILL;
    T qv[QREG ? D : 1];
    T go[GVF ? D : 1];
    T tqa = 0, tqm = 0;
    if (QREG) {
#pragma unroll
        for (int l = 0; l < (QREG ? D : 1); l++) qv[l] = lane < nq ? xs[(uint64_t)l * ld + q0 + lane] : (T)0;
    }
    if (GVF) {
        T qn = 0;
#pragma unroll
        for (int l = 0; l < (GVF ? D : 1); l++) {
            go[l] = lane_value(qv[l], 0u);
            const T b = qv[l] - go[l];
            qn = l == 0 ? b * b : fma_rn(b, b, qn);
            qv[l] = (T)-2 * b;
        }
        tqa = (T)a.gvLo - qn;
        tqm = (T)a.gvMid - qn;
    }
    uint32_t cnt_l = 0;
    uint64_t cur_l = 0;  // fill pass: this lane's query's next CSR slot
    uint32_t inc_l = 0;  // fill pass: its entries written in the current batch
    uint64_t work = 0;
    if (PASS == PASS_FILL && lane < nq) cur_l = a.offsets[qid_l];
    uint32_t scur = 0, scap = 0;
    uint32_t *sid = nullptr;  // the tile's slot entries (wave-uniform bases)
    uint8_t *sq = nullptr;
    if (PASS == PASS_SLOT) {
        // positions stay below SLOT_POS_LIMIT, so their byte offsets fit
        // 32 bits; a tile with more hits is an overflowed tile
        scap = uniform(a.slotCap[tile]);
        scap = scap < SLOT_POS_LIMIT ? scap : SLOT_POS_LIMIT;
        const uint64_t sbase = uniform64(a.slotOff[tile]);
        sid = a.slotIds + sbase;
        sq = a.slotQ + sbase;
    }
    bool overflow = false;
    // hit masks: word of (query qi, chunk j) = hbase + qi * hstride + j
    uint64_t hbase = 0;
    uint32_t hstride = 0, gch = 0;
    if (HM) {
        hstride = hmask_stride(uniform(a.cellCand[cell]), nwin);
        hbase = uniform64(a.hmCell[cell]) + (uint64_t)(q0 - uniform(a.cellStart[cell])) * hstride;
    }

    for (uint32_t cw = 0; cw < nwin; cw += GDSJ_WAVE) {
        const uint32_t nrw = uniform(nwin - cw < (uint32_t)GDSJ_WAVE ? nwin - cw : (uint32_t)GDSJ_WAVE);
        uint2 run = make_uint2(0u, 0u);
        if (lane < nrw)
            run = pre ? a.runs[(uint64_t)cell * a.runStride + cw + lane]
                      : combo_run(cw + lane, cc, glo, ghi, ga, a.cellStart, a.cellKey, a.hkeys, a.hvals,
                                  STAMP ? (int)a.stamp : 0);
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
        // fill pass: when every query's row can take all of this window's
        // candidates below the output capacity, the writes skip the check
        bool roomy = false;
        if (PASS == PASS_FILL) {
            uint64_t mx = lane < nq ? cur_l : 0ull;
#pragma unroll
            for (int off = 32; off > 0; off >>= 1) {
                const uint64_t o = __shfl_xor(mx, off, GDSJ_WAVE);
                mx = o > mx ? o : mx;
            }
            roomy = uniform64(mx) + total <= a.cap;
        }
        const uint32_t rexcl = incl - len;
        const uint32_t rstart = run.x;
        // Precomputed runs are non-empty: a window of up to BMW * 64
        // candidates marks its run starts in an LDS bitmap (bit rexcl - 1 of
        // each run after the first), and a chunk's lanes find their run from
        // the chunk's bitmap word instead of a binary search over the runs.
        const bool bmap = BMAP && pre && total <= (uint32_t)(BMW * GDSJ_WAVE);
        uint64_t *bm = s_bm[BMAP ? threadIdx.x / GDSJ_WAVE : 0];
        if (bmap) {
            if (lane < (total + GDSJ_WAVE - 1) / GDSJ_WAVE) bm[lane] = 0ull;
            __builtin_amdgcn_wave_barrier();
            if (lane < nrw && rexcl > 0)
                atomicOr(reinterpret_cast<uint32_t *>(bm) + ((rexcl - 1u) >> 5), 1u << ((rexcl - 1u) & 31u));
            __builtin_amdgcn_wave_barrier();
        }
        const uint32_t delta = rstart - rexcl;  // position - window index, per run
        // a valid position for lanes past the end: first candidate of the window
        const uint32_t fpos = bmap ? uniform(shfl_u32(rstart, 0u)) : uniform(cand_pos(0u, nrw, rstart, rexcl));

        for (uint32_t cb = 0; cb < total; cb += GDSJ_WAVE * CMAX) {
            const uint32_t nch = uniform((total - cb + GDSJ_WAVE - 1) / GDSJ_WAVE < (uint32_t)CMAX
                                             ? (total - cb + GDSJ_WAVE - 1) / GDSJ_WAVE
                                             : (uint32_t)CMAX);
            T c[CMAX][D > 0 ? D : 1];
            uint32_t cpos[CMAX];
            uint32_t cid[CMAX];
            uint32_t scell[SRT ? CMAX : 1];  // SORT: the candidate's cell
            T cn[GVF ? CMAX : 1];  // GV: |c'|^2 of this lane's candidate
            bool valid[CMAX];
            bool other[CMAX];                                    // STAMP: candidate outside the own cell
            uint32_t jc[STAMP && PASS == PASS_COUNT ? CMAX : 1];  // STAMP: hits of this candidate
            uint64_t jm[STAMP && PASS == PASS_FILL ? CMAX : 1];   // STAMP: queries hitting it
#pragma unroll
            for (int ch = 0; ch < CMAX; ch++) {
                const uint32_t g = cb + ch * GDSJ_WAVE + lane;
                valid[ch] = g < total;
                if (STAMP && PASS == PASS_COUNT) jc[ch] = 0;
                if (STAMP && PASS == PASS_FILL) jm[ch] = 0;
                uint32_t sp = fpos;
                if (bmap) {
                    if (ch < (int)nch) {
                        // run of g = (runs starting at or before the chunk's
                        // first index) + (starts in (base, g]) - 1
                        const uint32_t base = cb + (uint32_t)ch * GDSJ_WAVE;
                        const uint64_t mw = uniform64(bm[base >> 6]);
                        const uint32_t le = popc64(ballot(rexcl <= base));
                        sp = g + shfl_u32(delta, mbcnt_add(mw, le - 1u));
                    }
                } else {
                    sp = cand_pos(valid[ch] ? g : 0u, nrw, rstart, rexcl);
                }
                sp = valid[ch] ? sp : fpos;
                cpos[ch] = sp;
                other[ch] = STAMP && valid[ch] && (sp < own0 || sp >= own1);
                if (!MFILL && D > 0 && ch < (int)nch) {
                    // lanes past the end read the NaN padding behind the last
                    // point (position ld - 1), which fails every compare
                    if (XA) {
                        T row[D > 0 ? D : 1];
                        load_row<T, (D > 0 ? D : 1)>((const T *)a.xa, valid[ch] ? (uint64_t)sp : ld - 1, row);
#pragma unroll
                        for (int l = 0; l < (D > 0 ? D : 1); l++) c[ch][l] = row[l];
                    } else if (off32) {
                        const uint32_t bo = (valid[ch] ? sp : (uint32_t)(ld - 1)) * (uint32_t)sizeof(T);
#pragma unroll
                        for (int l = 0; l < (D > 0 ? D : 1); l++)
                            c[ch][l] = buffer_load(xrs, bo, (uint32_t)((uint64_t)l * ld * sizeof(T)), T());
                    } else {
                        const uint64_t xp = valid[ch] ? (uint64_t)sp : ld - 1;
#pragma unroll
                        for (int l = 0; l < (D > 0 ? D : 1); l++) c[ch][l] = xs[(uint64_t)l * ld + xp];
                    }
                    if (GVF) {
                        T nn = 0;
#pragma unroll
                        for (int l = 0; l < (GVF ? D : 1); l++) {
                            c[ch][l] = c[ch][l] - go[l];
                            nn = l == 0 ? c[ch][l] * c[ch][l] : fma_rn(c[ch][l], c[ch][l], nn);
                        }
                        cn[ch] = nn;
                    }
                }
                if (NEED_IDS) cid[ch] = (ch < (int)nch) ? (off32 ? load_at(ids, sp * 4u) : ids[sp]) : 0u;
                if (SRT) scell[ch] = (ch < (int)nch) ? a.posCell[sp] : 0u;
            }
            // unswitched on the fill pass's capacity check where it pays
            // (the single-join kernel); one copy with the check otherwise
            if (UNSW && PASS == PASS_FILL && roomy) {
#define GDSJ_RY true
#define GDSJ_FULL false
#include "join_query_loop.inc"
#undef GDSJ_FULL
#undef GDSJ_RY
            } else if (GVF && nch == (uint32_t)CMAX) {
#define GDSJ_RY (UNSW ? false : roomy)
#define GDSJ_FULL true
#include "join_query_loop.inc"
#undef GDSJ_FULL
#undef GDSJ_RY
            } else {
#define GDSJ_RY (UNSW ? false : roomy)
#define GDSJ_FULL false
#include "join_query_loop.inc"
#undef GDSJ_FULL
#undef GDSJ_RY
            }
            if (PASS == PASS_FILL) cur_l += inc_l;
            if (HM) gch += nch;
            if (STAMP) {
                // the reverse direction of this chunk batch's hits
#pragma unroll
                for (int ch = 0; ch < CMAX; ch++) {
                    if (ch >= (int)nch || !other[ch]) continue;
                    if (PASS == PASS_COUNT) {
                        if (jc[ch]) atomicAdd(&a.counts[cid[ch]], jc[ch]);
                    } else if (PASS == PASS_FILL) {
                        uint64_t mm = jm[ch];
                        if (mm) {
                            const uint32_t p = cid[ch];
                            const uint32_t b = atomicAdd(&a.rcur[p], (uint32_t)popc64(mm));
                            uint64_t pos = a.offsets[p] + a.fcounts[p] + b;
                            while (mm) {
                                const int bit = __builtin_ctzll(mm);
                                mm &= mm - 1;
                                if (pos < a.cap) a.out[pos] = s_qid[threadIdx.x / GDSJ_WAVE][bit];
                                else overflow = true;
                                pos++;
                            }
                        }
                    }
                }
            }
        }
    }
    if (PASS != PASS_FILL && lane < nq) {
        if (STAMP) {
            a.fcounts[qid_l] = cnt_l;
            atomicAdd(&a.counts[qid_l], cnt_l);
        } else {
            a.counts[qid_l] = cnt_l;
        }
    }
    if (PASS != PASS_FILL && a.tileWork && lane == 0) a.tileWork[tile] = work * nq;
    if (PASS == PASS_SLOT && lane == 0) {
        a.slotUsed[tile] = scur;
        if (scur > scap) a.ovfList[atomicAdd(a.ovfCount, 1u)] = a.tileBase + tile;
    }
    if (PASS == PASS_FILL && overflow) atomicOr(a.status, 1u);
}

// hm: the hit masks (HM; a.hmask), read-only here in the fill pass
// GV: the Gram-form filter (JoinArgs::gv) decides the pairs outside its band;
// SRT: SORT's candidate windows (generic kernels only)
template <typename T, int D, int MODE, int ILPV, int PASS, bool STAMP = false, bool HM = false, bool GV = false,
          bool SRT = false>
__global__ __launch_bounds__(256) void k_join(const T *__restrict__ xs, const uint32_t *__restrict__ ids,
                                              const JoinArgs a, const uint64_t *__restrict__ hm) {
    const uint32_t lblock = a.xcd_map ? xcd_block(blockIdx.x, gridDim.x) : blockIdx.x;
    const uint32_t tslot = uniform((uint32_t)(lblock * (blockDim.x / GDSJ_WAVE) + threadIdx.x / GDSJ_WAVE));
    if (tslot >= a.ntiles) return;
    const uint32_t tile = a.tileList ? uniform(a.tileList[tslot]) : tslot;
    // the Gram-form filter holds one more word per candidate (|c'|^2); the
    // hit-mask passes keep the default so that the counting pass and the
    // fill from its masks cut the candidates into the same chunks
    constexpr int CM = PASS == PASS_SLOT ? CMaxFor<T, GV ? D + 1 : D, GDSJ_CMAX_WORDS_SLOT>::value
                                         : (GV && !HM ? CMaxFor<T, D + 1>::value : 0);
    // up to 16 dims the grid keeps the row-major copy for these kernels
    constexpr bool XA = D > 0 && D <= 16;
    join_tile<T, D, MODE, ILPV, PASS, CM, STAMP, 16, true, true, HM, GV, SRT, XA>(xs, ids, a, tile, hm);
}

}  // namespace gdsj
