// mjoin_kernels.h -- the join of low-dimensional data on the matrix cores
// (every dim indexed, 3 to 7 dims): instantiated per DTYPE in
// join_f32.hip / join_f64.hip.
//
// A pair's squared distance is the Gram form |a|^2 + |b|^2 - 2 a.b of the
// coordinates relative to the tile's first query o (a = q - o, b = c - o,
// rounded to float).  One wave takes a tile of up to 64 queries of one
// cell, as one or two row blocks of 32; its candidates (the cell's
// precomputed neighbor runs, in run order) come 64 at a time, one per lane,
// as two column blocks of 32.  Each query and each candidate is a vector of
// K = D + 2 entries, (-2 a, 1, |a|^2) and (b, |b|^2, 1), so one chain of
// v_mfma_f32_32x32x2f32 (K = 2 per instruction) gives F = |a|^2 + |b|^2 -
// 2 a.b for a 32 x 32 block.  In that instruction lane l supplies A[l%32]
// [l/32] and B[l/32][l%32], which v_permlane32_swap makes out of the
// natural layouts (lane = query, lane = candidate): swapping the upper half
// of entry 2m with the lower half of entry 2m + 1 gives the operands of
// both 32-row (or 32-column) blocks.  Result register r of lane l is
// F[8 (r/4) + 4 (l/32) + r%4][l%32]: the ballot of register r's compare in
// the two column blocks holds, for each of its two rows, one bit per
// candidate of the chunk.  The 64 ballot words of a row block are written
// into one register (v_writelane, word 4 r + 2 column block + half) and
// read back by each row's lane (ds_bpermute), so that lane R holds the
// 64-bit hit mask of query R over the chunk, candidate j at bit j.  Counts,
// slot entries and rows are then written per lane, in candidate order, so
// the table is the plain join's, row for row.
//
// Exactness: F > Thi proves the reference's acc exceeds its threshold,
// F <= Tlo proves it is within it (mjoin_band bounds every rounding), and
// the pairs in between (a band of a few ulps of eps^2) are decided by the
// reference's arithmetic on the stored coordinates (kernel.cu:307-365).
#pragma once
#include "join_kernels.h"

namespace gdsj {

typedef float mj_f32x16 __attribute__((ext_vector_type(16)));

// chunks of 64 candidates whose loads are issued together
#ifndef GDSJ_MJ_CB
#define GDSJ_MJ_CB 4
#endif
constexpr int MJ_CB = GDSJ_MJ_CB;

// (lower halves of x and y, upper halves of x and y): lane l of the first
// result holds x[l] (l < 32) or y[l - 32], of the second x[l + 32] or y[l]
__device__ __forceinline__ void mj_swap(float x, float y, float &lo, float &hi) {
    const auto p = __builtin_amdgcn_permlane32_swap(__float_as_uint(x), __float_as_uint(y), false, false);
    lo = __uint_as_float(p[0]);
    hi = __uint_as_float(p[1]);
}

template <typename T, int D, int PASS>
__global__ __launch_bounds__(256) void k_mjoin(const T *__restrict__ xs, const uint32_t *__restrict__ ids,
                                               const JoinArgs a, int mode) {
    static_assert(D >= 1 && D <= 14, "K = D + 2 entries");
    constexpr int NE = D + 2;         // entries per vector
    constexpr int KM = (NE + 1) / 2;  // MFMAs per 32 x 32 block
    const uint32_t lane = lane_id();
    const uint32_t w = threadIdx.x / GDSJ_WAVE;
    const uint32_t tslot = uniform((uint32_t)(blockIdx.x * (blockDim.x / GDSJ_WAVE) + w));
    if (tslot >= a.ntiles) return;
    const uint32_t tile = a.tileList ? uniform(a.tileList[tslot]) : tslot;
    const uint4 tl = a.tiles[tile];
    const uint32_t cell = uniform(tl.x), q0 = uniform(tl.y), nq = uniform(tl.z - tl.y);
    if (nq == 0) return;
    const uint64_t ld = a.ld;
    const T thr = (T)a.thr;
    const float thi = a.mjHi, tlo = a.mjLo;
    const uint32_t nwin = uniform(a.nruns[cell]);
    const bool off32 = (uint64_t)D * ld * sizeof(T) <= 0xFFFFFFFFull;
    const __amdgpu_buffer_rsrc_t xrs = buffer_rsrc(xs);

    // origin: the tile's first query (exact DTYPE values)
    T o[D];
#pragma unroll
    for (int j = 0; j < D; j++) o[j] = lane_value(xs[(uint64_t)j * ld + q0], 0u);
    // lane = query: its vector (-2 a, 1, |a|^2)
    float A[2][KM];
    {
        const uint32_t qs = lane < nq ? lane : 0u;
        float e[2 * KM];
        float na = 0.0f;
#pragma unroll
        for (int j = 0; j < D; j++) {
            const float u = (float)(xs[(uint64_t)j * ld + q0 + qs] - o[j]);
            na = j == 0 ? u * u : fma_rn(u, u, na);
            e[j] = -2.0f * u;
        }
        e[D] = 1.0f;
        e[D + 1] = lane < nq ? na : INFINITY;  // rows past nq: F = +inf, never a hit
        if (2 * KM > NE) e[2 * KM - 1] = 0.0f;
#pragma unroll
        for (int m = 0; m < KM; m++) mj_swap(e[2 * m], e[2 * m + 1], A[0][m], A[1][m]);
    }
    const uint32_t nrb = nq > 32u ? 2u : 1u;
    const uint32_t qid_l = lane < nq ? ids[q0 + lane] : 0u;
    uint32_t cnt_l = 0;
    uint64_t cur_l = 0;  // fill pass: this lane's query's next CSR slot
    if (PASS == PASS_FILL && lane < nq) cur_l = a.offsets[qid_l];
    uint32_t scur = 0, scap = 0;
    uint32_t *sid = nullptr;
    uint8_t *sq = nullptr;
    if (PASS == PASS_SLOT) {
        scap = uniform(a.slotCap[tile]);
        scap = scap < SLOT_POS_LIMIT ? scap : SLOT_POS_LIMIT;
        const uint64_t sbase = uniform64(a.slotOff[tile]);
        sid = a.slotIds + sbase;
        sq = a.slotQ + sbase;
    }
    bool overflow = false;
    uint64_t work = 0;
    __shared__ uint64_t s_bm[256 / GDSJ_WAVE][GDSJ_WAVE];
    __shared__ uint32_t s_pos[256 / GDSJ_WAVE][GDSJ_WAVE];  // the chunk's candidates: positions
    __shared__ uint32_t s_cid[256 / GDSJ_WAVE][PASS != PASS_COUNT ? GDSJ_WAVE : 1];  // and ids
    uint64_t *bm = s_bm[w];

    for (uint32_t cw = 0; cw < nwin; cw += GDSJ_WAVE) {
        const uint32_t nrw = uniform(nwin - cw < (uint32_t)GDSJ_WAVE ? nwin - cw : (uint32_t)GDSJ_WAVE);
        uint2 run = make_uint2(0u, 0u);
        if (lane < nrw) run = a.runs[(uint64_t)cell * a.runStride + cw + lane];
        const uint32_t len = run.y - run.x;
        uint32_t incl = len;
#pragma unroll
        for (int off = 1; off < GDSJ_WAVE; off <<= 1) {
            const uint32_t t = __shfl_up(incl, off, GDSJ_WAVE);
            if (lane >= (uint32_t)off) incl += t;
        }
        const uint32_t total = uniform(shfl_u32(incl, nrw - 1));
        if (total == 0) continue;
        work += total;
        const uint32_t rexcl = incl - len;
        const uint32_t rstart = run.x;
        // run starts in an LDS bitmap (windows of up to 4096 candidates),
        // else a binary search over the window's runs
        const bool bmap = total <= (uint32_t)(GDSJ_WAVE * GDSJ_WAVE);
        if (bmap) {
            if (lane < (total + GDSJ_WAVE - 1) / GDSJ_WAVE) bm[lane] = 0ull;
            __builtin_amdgcn_wave_barrier();
            if (lane < nrw && rexcl > 0)
                atomicOr(reinterpret_cast<uint32_t *>(bm) + ((rexcl - 1u) >> 5), 1u << ((rexcl - 1u) & 31u));
            __builtin_amdgcn_wave_barrier();
        }
        const uint32_t delta = rstart - rexcl;
        const uint32_t fpos = uniform(shfl_u32(rstart, 0u));

        for (uint32_t cb0 = 0; cb0 < total; cb0 += GDSJ_WAVE * MJ_CB) {
            // ---- a batch of up to MJ_CB chunks, lane = candidate: every
            // chunk's loads are issued before the first chunk's arithmetic
            const uint32_t nch = uniform((total - cb0 + GDSJ_WAVE - 1) / GDSJ_WAVE < (uint32_t)MJ_CB
                                             ? (total - cb0 + GDSJ_WAVE - 1) / GDSJ_WAVE
                                             : (uint32_t)MJ_CB);
            T cx[MJ_CB][D];
            uint32_t cpos[MJ_CB], cids[MJ_CB];
#pragma unroll
            for (int ch = 0; ch < MJ_CB; ch++) {
                const uint32_t cb = cb0 + (uint32_t)ch * GDSJ_WAVE;
                const uint32_t g = cb + lane;
                const bool valid = g < total;
                uint32_t sp = fpos;
                if (ch < (int)nch) {
                    if (bmap) {
                        const uint64_t mw = uniform64(bm[cb >> 6]);
                        const uint32_t le = popc64(ballot(rexcl <= cb));
                        sp = g + shfl_u32(delta, mbcnt_add(mw, le - 1u));
                    } else {
                        sp = cand_pos(valid ? g : 0u, nrw, rstart, rexcl);
                    }
                    sp = valid ? sp : fpos;
                    // lanes past the end read the NaN padding behind the last
                    // point (position ld - 1): their entries are NaN and fail
                    // every compare
                    const uint64_t xp = valid ? (uint64_t)sp : ld - 1;
                    if (a.xa) {
                        load_row<T, D>((const T *)a.xa, xp, cx[ch]);
                    } else {
#pragma unroll
                        for (int j = 0; j < D; j++) {
                            if (off32)
                                cx[ch][j] = buffer_load(xrs, (uint32_t)xp * (uint32_t)sizeof(T),
                                                        (uint32_t)((uint64_t)j * ld * sizeof(T)), T());
                            else
                                cx[ch][j] = xs[(uint64_t)j * ld + xp];
                        }
                    }
                    cids[ch] = PASS != PASS_COUNT ? (off32 ? load_at(ids, sp * 4u) : ids[sp]) : 0u;
                } else {
#pragma unroll
                    for (int j = 0; j < D; j++) cx[ch][j] = (T)0;
                    cids[ch] = 0u;
                }
                cpos[ch] = sp;
            }
#pragma unroll
            for (int ch = 0; ch < MJ_CB; ch++) {
                if (ch >= (int)nch) break;
                float e[2 * KM];
                float nb = 0.0f;
#pragma unroll
                for (int j = 0; j < D; j++) {
                    const float b = (float)(cx[ch][j] - o[j]);
                    nb = j == 0 ? b * b : fma_rn(b, b, nb);
                    e[j] = b;
                }
                e[D] = nb;
                e[D + 1] = 1.0f;
                if (2 * KM > NE) e[2 * KM - 1] = 0.0f;
                float B0[KM], B1[KM];
#pragma unroll
                for (int m = 0; m < KM; m++) mj_swap(e[2 * m], e[2 * m + 1], B0[m], B1[m]);
                s_pos[w][lane] = cpos[ch];
                if (PASS != PASS_COUNT) s_cid[w][lane] = cids[ch];
                __builtin_amdgcn_wave_barrier();

                // ballot words of the sure hits (F <= Tlo) of row blocks 0
                // and 1, and of the possible ones (F <= Thi) where a block
                // has a pair in the band
                uint32_t X = 0, Y = 0, Xm = 0, Ym = 0;
                bool anyband = false;
#pragma unroll
                for (int rb = 0; rb < 2; rb++) {
                    if (rb == 1 && nrb < 2) break;
                    mj_f32x16 F0 = {}, F1 = {};
#pragma unroll
                    for (int m = 0; m < KM; m++) {
                        const float am = rb ? A[1][m] : A[0][m];
                        F0 = __builtin_amdgcn_mfma_f32_32x32x2f32(am, B0[m], F0, 0, 0, 0);
                        F1 = __builtin_amdgcn_mfma_f32_32x32x2f32(am, B1[m], F1, 0, 0, 0);
                    }
                    uint32_t st = 0;
                    uint64_t band = 0;
#pragma unroll
                    for (int r = 0; r < 16; r++) {
                        const uint64_t S0 = ballot(F0[r] <= tlo), S1 = ballot(F1[r] <= tlo);
                        const uint64_t M0 = ballot(F0[r] <= thi), M1 = ballot(F1[r] <= thi);
                        band |= (M0 & ~S0) | (M1 & ~S1);
                        st = writelane_u32((uint32_t)S0, 4u * r, st);
                        st = writelane_u32((uint32_t)(S0 >> 32), 4u * r + 1u, st);
                        st = writelane_u32((uint32_t)S1, 4u * r + 2u, st);
                        st = writelane_u32((uint32_t)(S1 >> 32), 4u * r + 3u, st);
                    }
                    uint32_t sm = 0;
                    if (__builtin_expect(band != 0, 0)) {
                        anyband = true;
#pragma unroll
                        for (int r = 0; r < 16; r++) {
                            const uint64_t M0 = ballot(F0[r] <= thi), M1 = ballot(F1[r] <= thi);
                            sm = writelane_u32((uint32_t)M0, 4u * r, sm);
                            sm = writelane_u32((uint32_t)(M0 >> 32), 4u * r + 1u, sm);
                            sm = writelane_u32((uint32_t)M1, 4u * r + 2u, sm);
                            sm = writelane_u32((uint32_t)(M1 >> 32), 4u * r + 3u, sm);
                        }
                    }
                    if (rb == 0) {
                        X = st;
                        Xm = sm;
                    } else {
                        Y = st;
                        Ym = sm;
                    }
                }
                // lane R = row R of the tile (R < 32: block 0, else block
                // 1): row i = R % 32 is register 4 (i/8) + i%4, half (i/4)%2
                const uint32_t ri = lane & 31u;
                const uint32_t wlo = 4u * (4u * (ri >> 3) + (ri & 3u)) + ((ri >> 2) & 1u);
                const uint32_t mlo_x = (uint32_t)__shfl((int)X, (int)wlo, GDSJ_WAVE);
                const uint32_t mhi_x = (uint32_t)__shfl((int)X, (int)(wlo + 2u), GDSJ_WAVE);
                const uint32_t mlo_y = (uint32_t)__shfl((int)Y, (int)wlo, GDSJ_WAVE);
                const uint32_t mhi_y = (uint32_t)__shfl((int)Y, (int)(wlo + 2u), GDSJ_WAVE);
                uint64_t m = lane < 32u ? (((uint64_t)mhi_x << 32) | mlo_x) : (((uint64_t)mhi_y << 32) | mlo_y);
                if (__builtin_expect(anyband, 0)) {
                    // the reference's arithmetic decides the pairs in the band
                    const uint32_t plo_x = (uint32_t)__shfl((int)Xm, (int)wlo, GDSJ_WAVE);
                    const uint32_t phi_x = (uint32_t)__shfl((int)Xm, (int)(wlo + 2u), GDSJ_WAVE);
                    const uint32_t plo_y = (uint32_t)__shfl((int)Ym, (int)wlo, GDSJ_WAVE);
                    const uint32_t phi_y = (uint32_t)__shfl((int)Ym, (int)(wlo + 2u), GDSJ_WAVE);
                    const uint64_t pm = lane < 32u ? (((uint64_t)phi_x << 32) | plo_x) : (((uint64_t)phi_y << 32) | plo_y);
                    uint64_t bb = pm & ~m;
                    while (bb) {
                        const uint32_t j = (uint32_t)__builtin_ctzll(bb);
                        bb &= bb - 1ull;
                        if (dist_mem<T>(xs, ld, q0 + lane, s_pos[w][j], (uint32_t)D, mode, a.ilp) <= thr) m |= 1ull << j;
                    }
                }
                if (lane >= nq) m = 0;  // rows past nq
                const uint32_t pc = popc64(m);
                if (PASS == PASS_FILL) {
                    while (m) {
                        const uint32_t j = (uint32_t)__builtin_ctzll(m);
                        m &= m - 1ull;
                        if (cur_l < a.cap) a.out[cur_l] = s_cid[w][j];
                        else overflow = true;
                        cur_l++;
                    }
                } else {
                    cnt_l += pc;
                    if (PASS == PASS_SLOT) {
                        // each row's entries of this chunk at scur + (the
                        // entries of the rows before it)
                        uint32_t incl = pc;
#pragma unroll
                        for (int off = 1; off < GDSJ_WAVE; off <<= 1) {
                            const uint32_t t = __shfl_up(incl, off, GDSJ_WAVE);
                            if (lane >= (uint32_t)off) incl += t;
                        }
                        uint32_t p = scur + incl - pc;
                        scur += uniform(shfl_u32(incl, GDSJ_WAVE - 1));
                        while (m) {
                            const uint32_t j = (uint32_t)__builtin_ctzll(m);
                            m &= m - 1ull;
                            if (p < scap) {
                                store_at(sid, p * 4u, s_cid[w][j]);
                                store_at(sq, p, (uint8_t)lane);
                            }
                            p++;
                        }
                    }
                }
                __builtin_amdgcn_wave_barrier();
            }
        }
    }
    if (PASS != PASS_FILL && lane < nq) a.counts[qid_l] = cnt_l;
    if (PASS != PASS_FILL && a.tileWork && lane == 0) a.tileWork[tile] = work * nq;
    if (PASS == PASS_SLOT && lane == 0) {
        a.slotUsed[tile] = scur;
        if (scur > scap) a.ovfList[atomicAdd(a.ovfCount, 1u)] = a.tileBase + tile;
    }
    if (PASS == PASS_FILL && overflow) atomicOr(a.status, 1u);
}

}  // namespace gdsj
