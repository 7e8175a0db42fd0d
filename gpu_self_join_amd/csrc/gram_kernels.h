// gram_kernels.h -- the Gram-filter join kernel template (instantiated per
// DTYPE in gram_f32.hip / gram_f64.hip); see gram.hip for the algorithm.
#pragma once
#include "gram.h"
#include "join_kernels.h"
#include "neighbors.h"

namespace gdsj {

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef _Float16 f16x8 __attribute__((ext_vector_type(8)));
typedef float f32x16 __attribute__((ext_vector_type(16)));

template <int LP> struct Lp;
template <> struct Lp<0> {
    using scalar = __bf16;
    using vec = bf16x8;
    static constexpr double min_normal = 0x1p-126;
};
template <> struct Lp<1> {
    using scalar = _Float16;
    using vec = f16x8;
    static constexpr double min_normal = 0x1p-14;
};

template <int LP>
__device__ __forceinline__ f32x16 mfma32(typename Lp<LP>::vec a, typename Lp<LP>::vec b, f32x16 c) {
    if constexpr (LP == 0)
        return __builtin_amdgcn_mfma_f32_32x32x16_bf16(a, b, c, 0, 0, 0);
    else
        return __builtin_amdgcn_mfma_f32_32x32x16_f16(a, b, c, 0, 0, 0);
}

template <int LP>
__device__ __forceinline__ typename Lp<LP>::vec load8(const typename Lp<LP>::scalar *p) {
    const uint4 raw = *reinterpret_cast<const uint4 *>(p);
    return __builtin_bit_cast(typename Lp<LP>::vec, raw);
}

// row blocks of 32 queries per wave: A operands of all blocks stay in VGPRs
constexpr int rb_max(int ks, int split) {
    const int r = 16 / (ks * (split ? 2 : 1));
    return r < 1 ? 1 : (r > 4 ? 4 : r);
}

// Optional measurement of the filter's error on the rechecked pairs (a
// separate counting-pass variant, ERR, so the timed kernels keep their
// register allocation): the
// Gram estimate N_q + N_p - 2 S of the scaled squared distance against the
// squared distance in FP64, max of |difference| / (sigma eps)^2 into stats[2]
// (the bits of a non-negative double, ordered as an integer).
struct GramAux {
    const double *nrm;  // N = |h + l|^2 per point (scaled units)
    const double *sig;  // sigma
    uint32_t err;       // 1: measure
};

// ---------------------------------------------------------------- the join --
template <typename T, int LP, int KS, int SPLIT, int PASS, bool ERR>
__global__ __launch_bounds__(256) void k_gram(const T *__restrict__ xs, const uint32_t *__restrict__ ids,
                                              const typename Lp<LP>::scalar *__restrict__ H,
                                              const typename Lp<LP>::scalar *__restrict__ Lo,
                                              const float *__restrict__ W, const JoinArgs a, int mode,
                                              unsigned long long *__restrict__ stats, const GramAux aux,
                                              const T *__restrict__ XR, uint32_t xrs) {
    using vec = typename Lp<LP>::vec;
    constexpr uint32_t DP = KS * 16;
    constexpr int RB = rb_max(KS, SPLIT);
    constexpr uint32_t TQ = 32 * RB;
    const uint32_t lane = lane_id();
    const uint32_t w = threadIdx.x / GDSJ_WAVE;
    const uint32_t lblock = a.xcd_map ? xcd_block(blockIdx.x, gridDim.x) : blockIdx.x;
    const uint32_t tslot = uniform((uint32_t)(lblock * (blockDim.x / GDSJ_WAVE) + w));
    if (tslot >= a.ntiles) return;
    const uint32_t tile = a.tileList ? uniform(a.tileList[tslot]) : tslot;
    const uint64_t ld = a.ld;
    const T thr = (T)a.thr;
    const uint4 tl = a.tiles[tile];
    const uint32_t cell = uniform(tl.x), q0 = uniform(tl.y), nq = uniform(tl.z - tl.y);
    if (nq == 0) return;  // an empty sub-tile (split overflowed tiles)
    const uint32_t nrb = (nq + 31) / 32;
    const GridArgs &ga = a.ga;
    const bool pre = a.runs != nullptr;
    const uint32_t row_l = lane & 31u, half = lane >> 5;
    const uint64_t ltm = lane == 0 ? 0ull : (~0ull >> (64 - lane));

    __shared__ double s_gap[256 / GDSJ_WAVE][2][MAXK];
    __shared__ uint32_t s_cnt[256 / GDSJ_WAVE][TQ];
    __shared__ unsigned long long s_cur[256 / GDSJ_WAVE][TQ];
    __shared__ __align__(16) float s_wq[256 / GDSJ_WAVE][TQ];  // query weights
    // survivors waiting for their exact recheck, in block order (and so in
    // candidate order per row): candidate position and row.  Rechecks run 64
    // at a time, so that one memory round trip serves a wave of pairs rather
    // than the few survivors of one block.  A row block adds at most 32 x 32
    // entries; the queue holds them on top of a partial round.
    constexpr uint32_t QCAP = 32 * 32 + GDSJ_WAVE;
    __shared__ uint32_t s_qp[256 / GDSJ_WAVE][QCAP];
    __shared__ uint8_t s_qr[256 / GDSJ_WAVE][QCAP];
    double *glo = s_gap[w][0];
    double *ghi = s_gap[w][1];
    uint32_t cc[MAXK];
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
    for (uint32_t r = lane; r < TQ; r += GDSJ_WAVE) {
        s_cnt[w][r] = 0;
        s_cur[w][r] = (PASS == PASS_FILL && r < nq)
                          ? (a.splitOff ? a.splitOff[(uint64_t)tile * SPLIT_ROWS + r] : a.offsets[ids[q0 + r]])
                          : 0ull;
    }
    __builtin_amdgcn_wave_barrier();

    // negated query weights (-INFINITY past nq) in LDS: row r at s_wq[w][r];
    // the rows a lane's accumulator entries 4j..4j+3 cover are consecutive,
    // so they load as the MFMAs' initial accumulator: S - W_q
    for (uint32_t r = lane; r < TQ; r += GDSJ_WAVE) s_wq[w][r] = r < nq ? -W[q0 + r] : -INFINITY;
    __builtin_amdgcn_wave_barrier();
    vec A[RB][KS];
    vec Al[RB][KS];
#pragma unroll
    for (int rb = 0; rb < RB; rb++) {
        const uint32_t qr = rb * 32 + row_l;
#pragma unroll
        for (int ks = 0; ks < KS; ks++) {
            if (qr < nq) {
                A[rb][ks] = load8<LP>(H + (uint64_t)(q0 + qr) * DP + ks * 16 + half * 8);
                if (SPLIT) Al[rb][ks] = load8<LP>(Lo + (uint64_t)(q0 + qr) * DP + ks * 16 + half * 8);
            } else {
                A[rb][ks] = (vec){};
                if (SPLIT) Al[rb][ks] = (vec){};
            }
        }
    }
    uint32_t scur = 0, scap = 0;
    uint64_t sbase = 0;
    if (PASS == PASS_SLOT) {
        scap = uniform(a.slotCap[tile]);
        scap = scap < SLOT_POS_LIMIT ? scap : SLOT_POS_LIMIT;  // as the plain kernels' slotted pass
        sbase = uniform64(a.slotOff[tile]);
    }
    uint64_t work = 0;
    uint32_t nrech = 0, nhit = 0;
    double errmax = 0.0;
    bool overflow = false;
    static_assert(TQ <= 128, "rows fit the queue's 8-bit row field and the 7-bit peer match");
    uint32_t qn = 0;  // queued survivors (uniform)
    uint32_t *qp = s_qp[w];
    uint8_t *qr = s_qr[w];
    // rechecks the first cnt queued pairs (cnt a multiple of 64, or all of
    // them at the end of the tile) and emits the hits in queue order; the
    // rest move to the front
    auto flush = [&](uint32_t cnt) {
        __builtin_amdgcn_wave_barrier();
        for (uint32_t e0 = 0; e0 < cnt; e0 += GDSJ_WAVE) {
            const uint32_t e = e0 + lane;
            const bool v = e < cnt;
            const uint32_t pc = v ? qp[e] : 0u;
            const uint32_t r = v ? (uint32_t)qr[e] : 0u;
            bool hit = false;
            if (v) {
                nrech++;
                // dims per batch of loads (8 floats or 4 doubles: 8 VGPRs
                // per row; 8 doubles at one k-step of operands: 10.7 ->
                // 10.1 ms per slotted pass on 64-D FP64 data on 16 axes,
                // at 3 waves per SIMD instead of 4)
                constexpr uint32_t RC = sizeof(T) == 8 && KS > 1 ? 4u : 8u;
                // two batches in flight where the registers keep the
                // occupancy (FP64 rows, or one k-step of fp16 operands):
                // measured on 64-D FP64 data on 16 axes, 13.4 -> 10.8 ms per
                // slotted pass; FP32 at two k-steps loses a wave per SIMD
                // (9.1 -> 10.7 ms)
                constexpr bool PF = sizeof(T) == 8 || (KS == 1 && !SPLIT);
                const T *qrow = XR + (uint64_t)(q0 + r) * xrs, *crow = XR + (uint64_t)pc * xrs;
                const T dv = PF ? dist_row_pf<T, RC>(qrow, crow, a.d, mode, a.ilp)
                                : dist_row<T, RC>(qrow, crow, a.d, mode, a.ilp);
                hit = dv <= thr;
            }
            const uint64_t hm = ballot(hit);
            if (hm == 0) continue;
            nhit += hit ? 1u : 0u;
            if (PASS == PASS_COUNT) {
                if (hit) atomicAdd(&s_cnt[w][r], 1u);
            } else if (PASS == PASS_SLOT) {
                if (hit) {
                    atomicAdd(&s_cnt[w][r], 1u);
                    const uint32_t p = scur + mbcnt(hm);
                    if (p < scap) {
                        a.slotIds[sbase + p] = ids[pc];
                        a.slotQ[sbase + p] = (uint8_t)r;
                    }
                }
                scur += popc64(hm);
            } else {
                // lanes hitting the same row write in lane (= queue) order
                uint64_t peers = hm;
#pragma unroll
                for (int b = 0; b < 7; b++) {
                    const bool bit = (r >> b) & 1u;
                    const uint64_t mb = ballot(bit);
                    peers &= bit ? mb : ~mb;
                }
                const uint64_t below = peers & ltm;
                uint64_t pos = 0;
                if (hit) pos = s_cur[w][r] + popc64(below);
                __builtin_amdgcn_wave_barrier();
                if (hit) {
                    if (pos < a.cap) a.out[pos] = ids[pc]; else overflow = true;
                    if (below == 0) s_cur[w][r] += popc64(peers);
                }
                __builtin_amdgcn_wave_barrier();
            }
        }
        const uint32_t rest = qn - cnt;  // < 64 by the callers
        uint32_t tp = 0, tr = 0;
        if (lane < rest) {
            tp = qp[cnt + lane];
            tr = qr[cnt + lane];
        }
        __builtin_amdgcn_wave_barrier();
        if (lane < rest) {
            qp[lane] = tp;
            qr[lane] = (uint8_t)tr;
        }
        __builtin_amdgcn_wave_barrier();
        qn = rest;
    };

    // candidates of this tile: all, or the (tl.w - 1)-th of a.winSplit
    // equal ranges of the cell's candidate order (a.cellCand: its total)
    uint32_t clo = 0, chi = 0xFFFFFFFFu;
    if (a.winSplit && tl.w) {
        const uint64_t tot = a.cellCand[cell], sidx = tl.w - 1;
        clo = uniform((uint32_t)(sidx * tot / a.winSplit));
        chi = uniform((uint32_t)((sidx + 1) * tot / a.winSplit));
    }
    uint32_t cbase = 0;  // candidate index of the window's first candidate
    for (uint32_t cw = 0; cw < nwin && cbase < chi; cw += GDSJ_WAVE) {
        const uint32_t nrw = uniform(nwin - cw < (uint32_t)GDSJ_WAVE ? nwin - cw : (uint32_t)GDSJ_WAVE);
        uint2 run = make_uint2(0u, 0u);
        if (lane < nrw)
            run = pre ? a.runs[(uint64_t)cell * a.runStride + cw + lane]
                      : combo_run(cw + lane, cc, glo, ghi, ga, a.cellStart, a.cellKey, a.hkeys, a.hvals);
        const uint32_t len = run.y - run.x;
        uint32_t incl = len;
#pragma unroll
        for (int off = 1; off < GDSJ_WAVE; off <<= 1) {
            const uint32_t o = __shfl_up(incl, off, GDSJ_WAVE);
            if (lane >= (uint32_t)off) incl += o;
        }
        const uint32_t total = uniform(shfl_u32(incl, nrw - 1));
        if (total == 0) continue;
        work += total;
        // this window's share of the tile's candidate range: [wlo, whi)
        const uint32_t wlo = clo > cbase ? clo - cbase : 0u;
        const uint32_t whi = chi - cbase < total ? chi - cbase : total;
        cbase += total;
        if (wlo >= whi) continue;
        const uint32_t rexcl = incl - len;
        const uint32_t rstart = run.x;
        const uint32_t fpos = uniform(cand_pos(0u, nrw, rstart, rexcl));
        // block cb + 32's position and operands are loaded while block cb's
        // MFMAs and rechecks run
        uint32_t sp_n;
        vec Bn[KS], Bln[KS];
        float wc_n;
        // blocks advance through the candidate order, so the run of a
        // block's first candidate is tracked by a scalar cursor (ri); runs
        // average hundreds of candidates, and a block usually lies in one
        // or two of them
        uint32_t ri = 0;
        auto load_block = [&](uint32_t cbl) {
            while (ri + 1 < nrw && (uint32_t)__builtin_amdgcn_readlane(rexcl, ri + 1) <= cbl) ri++;
            const uint32_t gl = cbl + row_l;
            const bool vl = gl >= wlo && gl < whi;
            uint32_t p = (uint32_t)__builtin_amdgcn_readlane(rstart, ri) +
                         (gl - (uint32_t)__builtin_amdgcn_readlane(rexcl, ri));
            // lanes past the end of run k take the next run's positions
            for (uint32_t k = ri + 1; k < nrw; k++) {
                const uint32_t nx = (uint32_t)__builtin_amdgcn_readlane(rexcl, k);
                if (nx > cbl + 31u) break;
                if (gl >= nx) p = (uint32_t)__builtin_amdgcn_readlane(rstart, k) + (gl - nx);
            }
            p = vl ? p : fpos;
            sp_n = p;
#pragma unroll
            for (int ks = 0; ks < KS; ks++) {
                Bn[ks] = load8<LP>(H + (uint64_t)p * DP + ks * 16 + half * 8);
                if (SPLIT) Bln[ks] = load8<LP>(Lo + (uint64_t)p * DP + ks * 16 + half * 8);
            }
            wc_n = vl ? W[p] : INFINITY;
        };
        const uint32_t cb0 = wlo & ~31u;
        load_block(cb0);
        for (uint32_t cb = cb0; cb < whi; cb += 32) {
            const uint32_t sp = sp_n;
            const float wc = wc_n;
            vec B[KS], Bl[KS];
#pragma unroll
            for (int ks = 0; ks < KS; ks++) {
                B[ks] = Bn[ks];
                if (SPLIT) Bl[ks] = Bln[ks];
            }
            if (cb + 32 < whi) load_block(cb + 32);
            // per row block: the MFMAs, accumulating onto -W_q, then one wave
            // mask per accumulator entry: bit l of the mask = the pair (row of
            // the entry in lane l, column l & 31) survives the filter,
            // S - W_q >= W_p
            auto gram_rb = [&](int rb, const float4 (&wv)[4]) {
                f32x16 acc;
#pragma unroll
                for (int j = 0; j < 4; j++) {
                    acc[4 * j + 0] = wv[j].x;
                    acc[4 * j + 1] = wv[j].y;
                    acc[4 * j + 2] = wv[j].z;
                    acc[4 * j + 3] = wv[j].w;
                }
#pragma unroll
                for (int ks = 0; ks < KS; ks++) {
                    acc = mfma32<LP>(A[rb][ks], B[ks], acc);
                    if (SPLIT) {
                        acc = mfma32<LP>(A[rb][ks], Bl[ks], acc);
                        acc = mfma32<LP>(Al[rb][ks], B[ks], acc);
                    }
                }
                return acc;
            };
            // two accumulators, alternating by row block
            // one row block's accumulator at a time: issuing the next row
            // block's MFMAs early costs 32 VGPRs, and without them the kernel
            // runs 4 waves per SIMD at 32 dims and 3 at 64 (measured faster)
#pragma unroll
            for (int rb = 0; rb < RB; rb++) {
                if (rb >= (int)nrb) break;
                float4 wv[4];  // -W_q of the rows of this lane's entries
#pragma unroll
                for (int j = 0; j < 4; j++) wv[j] = *reinterpret_cast<const float4 *>(&s_wq[w][rb * 32 + j * 8 + half * 4]);
                const f32x16 acc = gram_rb(rb, wv);
                // the common case, no survivor in the row block, costs the
                // lane's largest entry (three-input max steps), one compare
                // and one branch (entries are finite or -inf)
                float mx = acc[0];
#pragma unroll
                for (int i = 1; i + 1 < 16; i += 2) mx = fmaxf(fmaxf(mx, acc[i]), acc[i + 1]);
                mx = fmaxf(mx, acc[15]);
                if (ballot(mx >= wc) == 0) continue;
                if (qn + 32 * 32 > QCAP) flush(qn & ~(GDSJ_WAVE - 1u));
#pragma unroll
                for (int j = 0; j < 4; j++) {
#pragma unroll
                    for (int e = 0; e < 4; e++) {
                        const int i = 4 * j + e;
                        const bool keep = acc[i] >= wc;
                        const uint64_t m = ballot(keep);
                        if (m == 0) continue;
                        const uint32_t r = rb * 32 + j * 8 + half * 4 + e;
                        if (keep) {
                            // the row is formed here, on the rare survivor
                            // path: hoisted, the 16 x RB rows would each hold
                            // a register through the loop
                            uint32_t hb = half * 4u;
                            asm volatile("" : "+v"(hb));
                            const uint32_t at = qn + mbcnt(m);
                            qp[at] = sp;
                            qr[at] = (uint8_t)(hb + (uint32_t)(rb * 32 + j * 8 + e));
                            if (ERR) {
                                const double s2 = aux.sig[0] * aux.sig[0];
                                const float nwq = e == 0 ? wv[j].x : (e == 1 ? wv[j].y : (e == 2 ? wv[j].z : wv[j].w));
                                const double est = aux.nrm[q0 + r] + aux.nrm[sp] - 2.0 * ((double)acc[i] - (double)nwq);
                                double ex = 0.0;
                                for (uint32_t l = 0; l < a.d; l++) {
                                    const double t = (double)xs[(uint64_t)l * ld + sp] -
                                                     (double)xs[(uint64_t)l * ld + q0 + r];
                                    ex += t * t;
                                }
                                errmax = fmax(errmax, fabs(est - s2 * ex) / (s2 * a.ga.eps * a.ga.eps));
                            }
                        }
                        qn += popc64(m);
                    }
                }
            }
            if (qn >= GDSJ_WAVE) flush(qn & ~(GDSJ_WAVE - 1u));
        }
    }
    if (qn) flush(qn);
    __builtin_amdgcn_wave_barrier();
    if (PASS != PASS_FILL) {
        for (uint32_t r = lane; r < nq; r += GDSJ_WAVE) {
            if (a.splitCnt)
                a.splitCnt[(uint64_t)tile * SPLIT_ROWS + r] = s_cnt[w][r];
            else
                a.counts[ids[q0 + r]] = s_cnt[w][r];
        }
        if (a.tileWork && lane == 0) a.tileWork[tile] = work * nq;
    }
    if (PASS == PASS_SLOT && lane == 0) {
        a.slotUsed[tile] = scur;
        if (scur > scap) a.ovfList[atomicAdd(a.ovfCount, 1u)] = a.tileBase + tile;
    }
    if (PASS == PASS_FILL && overflow) atomicOr(a.status, 1u);
    if (stats) {
        unsigned long long r1 = nrech, r2 = nhit;
#pragma unroll
        for (int off = 32; off > 0; off >>= 1) {
            r1 += __shfl_xor(r1, off, GDSJ_WAVE);
            r2 += __shfl_xor(r2, off, GDSJ_WAVE);
        }
        if (lane == 0 && r1) {
            atomicAdd(&stats[0], r1);
            atomicAdd(&stats[1], r2);
        }
        if (ERR) {
#pragma unroll
            for (int off = 32; off > 0; off >>= 1) errmax = fmax(errmax, __shfl_xor(errmax, off, GDSJ_WAVE));
            if (lane == 0 && errmax > 0.0) atomicMax(&stats[2], (unsigned long long)__double_as_longlong(errmax));
        }
    }
}

template <typename T, int LP, int KS>
inline int launch_gram(const GramData &gd, const JoinArgs &a, int mode, int pass, unsigned long long *stats,
                       bool err, hipStream_t s) {
    const uint32_t blocks = (a.ntiles + 3) / 4;
    if (blocks == 0) return GDSJ_OK;
    const T *xs = (const T *)a.xs;
    const auto *H = (const typename Lp<LP>::scalar *)gd.h.p;
    const auto *Lo = (const typename Lp<LP>::scalar *)(gd.split ? gd.l.p : gd.h.p);
    const float *W = gd.w.as<float>();
    constexpr int SP = LP == 0 ? 1 : 0;
    const GramAux aux{gd.n2.as<double>(), gd.scratch.as<double>() + 2 * (size_t)a.d, err && stats ? 1u : 0u};
    const T *XR = gd.xr.as<T>();
    const uint32_t xrs = gd.xrs;
    if (pass == PASS_FILL)
        k_gram<T, LP, KS, SP, PASS_FILL, false><<<blocks, 256, 0, s>>>(xs, a.ids, H, Lo, W, a, mode, stats, aux, XR,
                                                                      xrs);
    else if (pass == PASS_SLOT)
        k_gram<T, LP, KS, SP, PASS_SLOT, false><<<blocks, 256, 0, s>>>(xs, a.ids, H, Lo, W, a, mode, stats, aux, XR,
                                                                      xrs);
    else if (aux.err)  // measurement variant (counting pass only)
        k_gram<T, LP, KS, SP, PASS_COUNT, true><<<blocks, 256, 0, s>>>(xs, a.ids, H, Lo, W, a, mode, stats, aux, XR,
                                                                      xrs);
    else
        k_gram<T, LP, KS, SP, PASS_COUNT, false><<<blocks, 256, 0, s>>>(xs, a.ids, H, Lo, W, a, mode, stats, aux, XR,
                                                                       xrs);
    GDSJ_HIP(hipGetLastError());
    return GDSJ_OK;
}

template <typename T, int LP>
inline int gram_dispatch(const GramData &gd, const JoinArgs &a, int mode, int pass, unsigned long long *stats,
                         bool err, hipStream_t s) {
    switch (gd.dpad / 16) {
        case 1: return launch_gram<T, LP, 1>(gd, a, mode, pass, stats, err, s);
        case 2: return launch_gram<T, LP, 2>(gd, a, mode, pass, stats, err, s);
        case 3: return launch_gram<T, LP, 3>(gd, a, mode, pass, stats, err, s);
        case 4: return launch_gram<T, LP, 4>(gd, a, mode, pass, stats, err, s);
        case 6: return launch_gram<T, LP, 6>(gd, a, mode, pass, stats, err, s);
        case 8: return launch_gram<T, LP, 8>(gd, a, mode, pass, stats, err, s);
        default: break;
    }
    set_error("unsupported padded dims %u for the Gram filter", gd.dpad);
    return GDSJ_EUNSUPPORTED;
}

}  // namespace gdsj
