// batch.hip -- batched joins (include/gdsjoin.h, gdsj_batch_*): many
// independent self-joins of uneven n, dims and epsilon run as one job.
//
// Set-up (index construction, once): every instance gets its own grid,
// neighbor runs and tiles, exactly as a single join builds them; the tiles
// of all instances are numbered globally and ordered for the work queue
// (heaviest first: queries x candidates).  Each run is one hipGraph:
//   exact:   reset -> persistent counting pass -> scan of all counts
//            -> persistent fill pass
//   learned: reset -> persistent slotted pass (per-tile result slots sized
//            by the MLP) -> scan -> persistent compaction of the slots
//            -> persistent fill pass over the tiles whose slot overflowed
// with the output rows of instance i at rows rowBase[i] .. of the caller's
// CSR.  The graph is captured on the first run for a given output and
// replayed afterwards.  An autotuner times the counting pass over tile
// sizes, block sizes, queue orders (and ILP when the caller leaves it to
// the tuner) and keeps the fastest.
#include <string.h>

#include <algorithm>
#include <memory>
#include <vector>

#include "batch.h"
#include "estimator.h"
#include "grid.h"
#include "join.h"
#include "radix.h"
#include "scan.h"

using namespace gdsj;

namespace gdsj {

// queue heads and flags: [0] count, [1] fill, [2] slot, [3] compact,
// [4] overflow fill, [6] overflow count, [7] status
constexpr int BQ_COUNT = 0, BQ_FILL = 1, BQ_SLOT = 2, BQ_COMPACT = 3, BQ_OVF = 4, BQ_NOVF = 6, BQ_STATUS = 7;
constexpr int BQ_N = 8;

__global__ void k_batch_reset(uint32_t *ctr) {
    if (threadIdx.x < BQ_N) ctr[threadIdx.x] = 0u;
}

// items in a fixed wave-strided order: a tile's compaction is short, so a
// queue atomic per item would set the pace
__global__ __launch_bounds__(256) void k_batch_compact(const JoinArgs *__restrict__ inst, const BatchLaunch L) {
    const uint32_t nitems = batch_items(L);
    const uint32_t nw = gridDim.x * (blockDim.x / GDSJ_WAVE);
    for (uint32_t it = uniform(blockIdx.x * (blockDim.x / GDSJ_WAVE) + threadIdx.x / GDSJ_WAVE); it < nitems;
         it += nw) {
        uint32_t i, tile;
        batch_item(L, it, i, tile);
        const JoinArgs &a = inst_args(inst, i);
        compact_tile(a, tile);
    }
}

// queue-order key of every tile of one instance: descending queries x
// candidates (candidates per query of the tile's cell from its run list,
// else the cell's own occupancy), ties in global tile order
__global__ void k_batch_weights(const uint4 *__restrict__ tiles, uint32_t nt, const uint32_t *__restrict__ cellCand,
                                const uint32_t *__restrict__ cellStart, uint32_t base, uint64_t *__restrict__ keys,
                                uint32_t *__restrict__ vals) {
    const uint32_t t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= nt) return;
    const uint4 tl = tiles[t];
    const uint64_t nq = tl.z - tl.y;
    const uint64_t cand = cellCand ? cellCand[tl.x] : (uint64_t)(cellStart[tl.x + 1] - cellStart[tl.x]);
    const uint64_t w = nq * cand;
    keys[base + t] = 0xFFFFFFFFull - (w > 0xFFFFFFFFull ? 0xFFFFFFFFull : w);
    vals[base + t] = base + t;
}

}  // namespace gdsj

struct gdsj_batch {
    int device = 0;
    hipStream_t own = nullptr;
    int dtype = GDSJ_DTYPE_F32;
    gdsj_params p{};
    bool ilp_auto = false;
    uint32_t ninst = 0;
    std::vector<gdsj_instance> in;
    std::vector<std::unique_ptr<Grid>> G;
    std::vector<double> thr;
    std::vector<uint32_t> k;
    std::vector<uint64_t> rowBase;   // [ninst + 1]
    std::vector<uint32_t> tileBase;  // [ninst + 1]
    uint64_t npts = 0;
    uint32_t ntiles = 0;
    uint32_t built_tq = 0;
    bool all_runs = false;
    DevBuf dArgs, dTileBase, order, ctr, work, scanTmp, keys, vals, keys2, vals2, tcounts, toffsets;
    std::vector<JoinArgs> hArgs;
    // learned slots
    Estimator est;
    DevBuf feat, target, slotCap, slotOff, slotUsed, slotOvf, slotIds, slotQ;
    bool learned_ready = false;
    uint64_t slots = 0, predicted = 0;
    // tuning in force
    gdsj_batch_tuning tune{};
    bool tuned = false;
    // captured graph and what it was captured for
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    const void *gkey[3] = {nullptr, nullptr, nullptr};
    uint64_t gcap = 0;
    int gslotted = -1;
    bool args_valid = false;
    hipEvent_t ev[4] = {};
    gdsj_batch_stats st{};
};

namespace {

void drop_graph(gdsj_batch *b) {
    if (b->exec) (void)hipGraphExecDestroy(b->exec);
    if (b->graph) (void)hipGraphDestroy(b->graph);
    b->exec = nullptr;
    b->graph = nullptr;
    b->gslotted = -1;
}

// (Re)builds every instance's grid with tq queries per tile, the global
// tile numbering and the queue order.
int build_all(gdsj_batch *b, uint32_t tq) {
    hipStream_t s = b->own;
    drop_graph(b);
    b->args_valid = false;
    b->learned_ready = false;
    b->tileBase.assign(b->ninst + 1, 0);
    b->all_runs = true;
    for (uint32_t i = 0; i < b->ninst; i++) {
        const gdsj_instance &x = b->in[i];
        Grid &g = *b->G[i];
        // the specialised bodies (2, 4, 8, 16 dims) read the row-major copy
        g.aosWant = x.dims == 2 || x.dims == 4 || x.dims == 8 || x.dims == 16;
        GDSJ_TRY(grid_build(g, x.d_data, x.n, x.dims, b->dtype, x.epsilon, b->k[i], b->p.reorder, tq, 0, s));
        bool built = false;
        const bool prune = b->p.prune == 1 || (b->p.prune == 0 && g.k >= 4);
        GDSJ_TRY(build_runs(g, s, prune, false, &built));
        g.haveRuns = built;
        if (!built && g.ntiles) b->all_runs = false;
        const uint64_t next = (uint64_t)b->tileBase[i] + g.ntiles;
        if (next >= 0xFFFFFFFFull) {
            set_error("batch has more than 2^32 - 1 tiles");
            return GDSJ_ERANGE;
        }
        b->tileBase[i + 1] = (uint32_t)next;
    }
    b->ntiles = b->tileBase[b->ninst];
    b->built_tq = tq;
    const uint32_t nt = b->ntiles;
    GDSJ_TRY(b->dTileBase.ensure(((size_t)b->ninst + 1) * 4 + 64));
    GDSJ_HIP(hipMemcpyAsync(b->dTileBase.p, b->tileBase.data(), ((size_t)b->ninst + 1) * 4, hipMemcpyHostToDevice,
                            s));
    GDSJ_TRY(b->order.ensure((size_t)nt * 4 + 64));
    GDSJ_TRY(b->keys.ensure((size_t)nt * 8 + 64));
    GDSJ_TRY(b->vals.ensure((size_t)nt * 4 + 64));
    GDSJ_TRY(b->keys2.ensure((size_t)nt * 8 + 64));
    GDSJ_TRY(b->vals2.ensure((size_t)nt * 4 + 64));
    GDSJ_TRY(b->work.ensure((size_t)nt * 8 + 64));
    for (uint32_t i = 0; i < b->ninst; i++) {
        const Grid &g = *b->G[i];
        if (!g.ntiles) continue;
        k_batch_weights<<<ceil_div(g.ntiles, 256), 256, 0, s>>>(
            g.tiles.as<uint4>(), g.ntiles, g.haveRuns ? g.cellCand.as<uint32_t>() : nullptr,
            g.cellStart.as<uint32_t>(), b->tileBase[i], b->keys.as<uint64_t>(), b->vals.as<uint32_t>());
    }
    GDSJ_HIP(hipGetLastError());
    if (nt) {
        DevBuf tmp;
        GDSJ_TRY(tmp.ensure(radix_tmp_bytes(nt) + 256));
        int alt = 0;
        GDSJ_TRY(radix_sort_pairs(b->keys.as<uint64_t>(), b->vals.as<uint32_t>(), b->keys2.as<uint64_t>(),
                                  b->vals2.as<uint32_t>(), nt, 32, tmp.p, s, &alt));
        GDSJ_HIP(hipMemcpyAsync(b->order.p, alt ? b->vals2.p : b->vals.p, (size_t)nt * 4, hipMemcpyDeviceToDevice,
                                s));
        GDSJ_HIP(hipStreamSynchronize(s));
        tmp.release();
    }
    GDSJ_HIP(hipStreamSynchronize(s));
    return GDSJ_OK;
}

// per-instance kernel arguments for the given outputs
void make_args(gdsj_batch *b, uint32_t *counts, uint64_t *offsets, uint32_t *ids, uint64_t cap) {
    b->hArgs.assign(b->ninst, JoinArgs{});
    uint32_t *ctr = b->ctr.as<uint32_t>();
    for (uint32_t i = 0; i < b->ninst; i++) {
        const Grid &g = *b->G[i];
        JoinArgs &a = b->hArgs[i];
        a.xs = g.xs.p;
        a.ld = g.ld;
        a.ixs = g.xs.p;
        a.xa = g.xaStride ? g.xa.p : nullptr;
        a.xaStride = g.xaStride;
        a.ids = g.ids.as<uint32_t>();
        a.tiles = g.tiles.as<uint4>();
        a.ntiles = g.ntiles;
        a.cellStart = g.cellStart.as<uint32_t>();
        a.cellCoord = g.cellCoord.as<uint32_t>();
        a.cellKey = g.cellKey.as<uint64_t>();
        a.hkeys = g.hkeys.as<uint64_t>();
        a.hvals = g.hvals.as<uint32_t>();
        a.ga = g.ga;
        a.ncombo = 1;
        for (uint32_t j = 1; j < g.k; j++) a.ncombo *= 3;
        a.thr = b->thr[i];
        a.d = g.d;
        a.ilp = b->tune.ilp;
        fma_band(g.dtype, g.d, a.ilp, a.thr, b->p.ffilter, &a.ffLo, &a.ffHi);
        a.sc = b->p.shortcircuit;
        a.xcd_map = 0;
        a.prune = b->p.prune == 1 || (b->p.prune == 0 && g.k >= 4) ? 1u : 0u;
        if (g.haveRuns) {
            a.runs = g.runs.as<uint2>();
            a.nruns = g.nruns.as<uint32_t>();
            a.runStride = g.runStride;
        }
        a.counts = counts + b->rowBase[i];
        a.tileWork = b->work.as<uint64_t>() + b->tileBase[i];
        a.offsets = offsets ? offsets + b->rowBase[i] : nullptr;
        a.out = ids;
        a.cap = cap;
        if (b->learned_ready) {
            a.slotOff = b->slotOff.as<uint64_t>() + b->tileBase[i];
            a.slotCap = b->slotCap.as<uint32_t>() + b->tileBase[i];
            a.slotUsed = b->slotUsed.as<uint32_t>() + b->tileBase[i];
            a.slotIds = b->slotIds.as<uint32_t>();
            a.slotQ = b->slotQ.as<uint8_t>();
            a.ovfList = b->slotOvf.as<uint32_t>();
            a.ovfCount = ctr + BQ_NOVF;
        }
        a.tileBase = b->tileBase[i];
        a.status = ctr + BQ_STATUS;
    }
}

int upload_args(gdsj_batch *b, uint32_t *counts, uint64_t *offsets, uint32_t *ids, uint64_t cap) {
    make_args(b, counts, offsets, ids, cap);
    GDSJ_TRY(b->dArgs.ensure((size_t)b->ninst * sizeof(JoinArgs) + 64));
    GDSJ_HIP(hipMemcpyAsync(b->dArgs.p, b->hArgs.data(), (size_t)b->ninst * sizeof(JoinArgs),
                            hipMemcpyHostToDevice, b->own));
    GDSJ_HIP(hipStreamSynchronize(b->own));
    b->args_valid = true;
    return GDSJ_OK;
}

BatchLaunch launch_of(gdsj_batch *b, int q, bool lpt) {
    BatchLaunch L{};
    L.tileBase = b->dTileBase.as<uint32_t>();
    L.ninst = b->ninst;
    L.order = lpt ? b->order.as<uint32_t>() : nullptr;
    L.nitems = b->ntiles;
    L.nitemsDev = nullptr;
    L.queue = b->ctr.as<uint32_t>() + q;
    return L;
}

int pass(gdsj_batch *b, const BatchLaunch &L, int ps, hipStream_t s, uint32_t *blocks, bool launch = true) {
    const JoinArgs *inst = b->dArgs.as<JoinArgs>();
    const int mode = acc_mode(b->tune.ilp, b->p.shortcircuit, b->p.ilp_fold);
    const uint32_t threads = b->tune.waves_per_block * GDSJ_WAVE;
    if (b->dtype == GDSJ_DTYPE_F32)
        return batch_pass_f32(inst, L, mode, b->tune.ilp, b->tune.reg_budget, ps, threads, blocks, s, launch);
    return batch_pass_f64(inst, L, mode, b->tune.ilp, b->tune.reg_budget, ps, threads, blocks, s, launch);
}

// occupancy queries of the passes a run makes, outside any capture
int size_passes(gdsj_batch *b) {
    const BatchLaunch L = launch_of(b, BQ_COUNT, false);
    const int ps[3] = {PASS_COUNT, PASS_FILL, PASS_SLOT};
    for (int q : ps) GDSJ_TRY(pass(b, L, q, b->own, nullptr, false));
    return GDSJ_OK;
}

// the launches of one run (captured into the graph)
int enqueue_run(gdsj_batch *b, uint32_t *counts, uint64_t *offsets, bool slotted, hipStream_t s) {
    k_batch_reset<<<1, 64, 0, s>>>(b->ctr.as<uint32_t>());
    GDSJ_HIP(hipGetLastError());
    const bool lpt = b->tune.lpt != 0;
    uint32_t blocks = 0;
    if (!slotted) {
        GDSJ_TRY(pass(b, launch_of(b, BQ_COUNT, lpt), PASS_COUNT, s, &blocks));
        GDSJ_TRY(exclusive_scan_u32_u64(counts, offsets, b->npts, true, b->scanTmp.p, s));
        GDSJ_TRY(pass(b, launch_of(b, BQ_FILL, lpt), PASS_FILL, s, nullptr));
    } else {
        GDSJ_TRY(pass(b, launch_of(b, BQ_SLOT, lpt), PASS_SLOT, s, &blocks));
        GDSJ_TRY(exclusive_scan_u32_u64(counts, offsets, b->npts, true, b->scanTmp.p, s));
        // the compaction is latency-bound (a few dependent loads per tile):
        // its own grid of 4-wave blocks, 8 per CU, whatever the join passes'
        // tuned shape (one wave per block at 3 per SIMD held it to 3 waves
        // per CU)
        BatchLaunch C = launch_of(b, BQ_COMPACT, lpt);
        k_batch_compact<<<2048, 256, 0, s>>>(b->dArgs.as<JoinArgs>(), C);
        GDSJ_HIP(hipGetLastError());
        BatchLaunch O = launch_of(b, BQ_OVF, true);
        O.order = b->slotOvf.as<uint32_t>();
        O.nitemsDev = b->ctr.as<uint32_t>() + BQ_NOVF;
        GDSJ_TRY(pass(b, O, PASS_FILL, s, nullptr));
    }
    b->tune.blocks = blocks;
    return GDSJ_OK;
}

// count pass alone, timed (autotuner)
int time_count(gdsj_batch *b, float *ms) {
    hipStream_t s = b->own;
    float best = 1e30f;
    for (int r = 0; r < 3; r++) {
        k_batch_reset<<<1, 64, 0, s>>>(b->ctr.as<uint32_t>());
        GDSJ_HIP(hipEventRecord(b->ev[0], s));
        uint32_t blocks = 0;
        GDSJ_TRY(pass(b, launch_of(b, BQ_COUNT, b->tune.lpt != 0), PASS_COUNT, s, &blocks));
        GDSJ_HIP(hipEventRecord(b->ev[1], s));
        GDSJ_HIP(hipEventSynchronize(b->ev[1]));
        float t = 0;
        GDSJ_HIP(hipEventElapsedTime(&t, b->ev[0], b->ev[1]));
        if (r > 0) best = std::min(best, t);  // the first launch warms caches
    }
    *ms = best;
    return GDSJ_OK;
}

// trains the estimator on the counts of an exact run and sizes every tile's slot
int learn(gdsj_batch *b, const uint32_t *counts) {
    hipStream_t s = b->own;
    const uint32_t nt = b->ntiles;
    GDSJ_TRY(b->feat.ensure((size_t)nt * MLP_F * 4 + 64));
    GDSJ_TRY(b->target.ensure((size_t)nt * 4 + 64));
    for (uint32_t i = 0; i < b->ninst; i++) {
        const Grid &g = *b->G[i];
        if (!g.ntiles) continue;
        GDSJ_TRY(tile_features(g, b->feat.as<float>() + (size_t)b->tileBase[i] * MLP_F, s));
        GDSJ_TRY(tile_targets(g, counts + b->rowBase[i], b->target.as<float>() + b->tileBase[i], s));
    }
    double tms = 0;
    GDSJ_TRY(estimator_train_rows(b->est, b->feat.as<float>(), b->target.as<float>(), nt, s, &tms));
    b->st.ms_train = tms;
    GDSJ_TRY(b->slotCap.ensure((size_t)nt * 4 + 64));
    GDSJ_TRY(b->slotOff.ensure(((size_t)nt + 1) * 8 + 64));
    GDSJ_TRY(b->slotUsed.ensure((size_t)nt * 4 + 64));
    GDSJ_TRY(b->slotOvf.ensure((size_t)nt * 4 + 64));
    std::vector<uint64_t> pred(b->ninst, 0);
    for (uint32_t i = 0; i < b->ninst; i++) {
        const Grid &g = *b->G[i];
        if (!g.ntiles) continue;
        GDSJ_TRY(estimator_capacity_learned(b->est, g, b->slotCap.as<uint32_t>() + b->tileBase[i], &pred[i], s));
        GDSJ_HIP(hipStreamSynchronize(s));
    }
    b->predicted = 0;
    for (uint64_t v : pred) b->predicted += v;
    DevBuf tmp;
    GDSJ_TRY(tmp.ensure(scan_tmp_bytes((uint64_t)nt + 1) + 256));
    GDSJ_TRY(exclusive_scan_u32_u64(b->slotCap.as<uint32_t>(), b->slotOff.as<uint64_t>(), nt, true, tmp.p, s));
    uint64_t slots = 0;
    GDSJ_HIP(hipMemcpyAsync(&slots, b->slotOff.as<uint64_t>() + nt, 8, hipMemcpyDeviceToHost, s));
    GDSJ_HIP(hipStreamSynchronize(s));
    tmp.release();
    b->slots = slots;
    GDSJ_TRY(b->slotIds.ensure((size_t)slots * 4 + 256));
    GDSJ_TRY(b->slotQ.ensure((size_t)slots + 256));
    b->learned_ready = true;
    b->args_valid = false;
    drop_graph(b);
    return GDSJ_OK;
}

}  // namespace

extern "C" {

int gdsj_batch_create(int device, const gdsj_instance *inst, uint32_t ninst, int dtype, const gdsj_params *pin,
                      gdsj_batch **out) {
    if (!out || (!inst && ninst)) return GDSJ_EINVAL;
    *out = nullptr;
    if (dtype != GDSJ_DTYPE_F32 && dtype != GDSJ_DTYPE_F64) {
        set_error("dtype must be GDSJ_DTYPE_F32 or GDSJ_DTYPE_F64");
        return GDSJ_EINVAL;
    }
    gdsj_params p;
    if (pin) p = *pin; else gdsj_params_default(&p);
    if (p.sort) {
        set_error("batched joins run without SORT");
        return GDSJ_EUNSUPPORTED;
    }
    if (p.filter == GDSJ_FILTER_GRAM || p.filter == GDSJ_FILTER_GRAM_BF16) {
        set_error("batched joins use the plain distance kernels (filter PLAIN or AUTO)");
        return GDSJ_EUNSUPPORTED;
    }
    if (p.estimator == GDSJ_EST_SAMPLE) {
        set_error("batched joins size their results exactly or with the learned estimator");
        return GDSJ_EUNSUPPORTED;
    }
    const bool ilp_auto = p.ilp == GDSJ_ILP_AUTO;
    if (!ilp_auto && p.ilp > 64) {
        set_error("ILP=%u exceeds 64", p.ilp);
        return GDSJ_EINVAL;
    }
    if (p.ilp_fold > 1) {
        set_error("ilp_fold=%u: 0 or 1", p.ilp_fold);
        return GDSJ_EINVAL;
    }
    int ndev = 0;
    GDSJ_HIP(hipGetDeviceCount(&ndev));
    if (device < 0 || device >= ndev) {
        set_error("device %d out of range (%d devices)", device, ndev);
        return GDSJ_EINVAL;
    }
    GDSJ_HIP(hipSetDevice(device));
    std::unique_ptr<gdsj_batch> b(new gdsj_batch());
    b->device = device;
    b->dtype = dtype;
    b->p = p;
    b->ilp_auto = ilp_auto;
    b->ninst = ninst;
    b->in.assign(inst, inst + ninst);
    b->rowBase.assign((size_t)ninst + 1, 0);
    for (uint32_t i = 0; i < ninst; i++) {
        const gdsj_instance &x = inst[i];
        if ((!x.d_data && x.n) || x.dims == 0) {
            set_error("instance %u: data pointer or dims missing", i);
            return GDSJ_EINVAL;
        }
        uint32_t kk = p.k ? p.k : (x.dims < 6 ? x.dims : 6);
        if (kk > x.dims || kk > (uint32_t)MAXK) {
            set_error("instance %u: NUMINDEXEDDIM k=%u must be <= dims=%u and <= %d", i, kk, x.dims, MAXK);
            return GDSJ_EINVAL;
        }
        b->k.push_back(kk);
        double t;
        if (dtype == GDSJ_DTYPE_F32) {
            const float e = (float)x.epsilon;
            if ((double)e != x.epsilon) {
                set_error("instance %u: epsilon %.17g is not a float value", i, x.epsilon);
                return GDSJ_EINVAL;
            }
            t = (double)gdsj_eps_threshold_f32(e);
        } else {
            t = gdsj_eps_threshold_f64(x.epsilon);
        }
        b->thr.push_back(t);
        b->rowBase[i + 1] = b->rowBase[i] + x.n;
        b->G.emplace_back(new Grid());
    }
    b->npts = b->rowBase[ninst];
    if (hipStreamCreateWithFlags(&b->own, hipStreamNonBlocking) != hipSuccess) {
        set_error("hipStreamCreate failed");
        return GDSJ_EDEVICE;
    }
    for (int i = 0; i < 4; i++) GDSJ_HIP(hipEventCreate(&b->ev[i]));
    GDSJ_TRY(b->ctr.ensure(256));
    GDSJ_TRY(b->scanTmp.ensure(scan_tmp_bytes(b->npts + 1) + 256));
    b->tune.tile_queries = p.tile_queries && p.tile_queries <= 64 ? p.tile_queries : 64;
    b->tune.waves_per_block = 4;
    b->tune.ilp = ilp_auto ? 8 : p.ilp;
    b->tune.lpt = 1;
    b->tune.reg_budget = 24;
    GDSJ_HIP(hipEventRecord(b->ev[2], b->own));
    int rc = build_all(b.get(), b->tune.tile_queries);
    if (rc != GDSJ_OK) {
        gdsj_batch_destroy(b.release());
        return rc;
    }
    GDSJ_HIP(hipEventRecord(b->ev[3], b->own));
    GDSJ_HIP(hipEventSynchronize(b->ev[3]));
    float ms = 0;
    (void)hipEventElapsedTime(&ms, b->ev[2], b->ev[3]);
    b->st.ms_build = ms;
    *out = b.release();
    return GDSJ_OK;
}

int gdsj_batch_destroy(gdsj_batch *b) {
    if (!b) return GDSJ_OK;
    (void)hipSetDevice(b->device);
    if (b->own) (void)hipStreamSynchronize(b->own);
    drop_graph(b);
    for (auto &g : b->G) g->release();
    DevBuf *bufs[] = {&b->dArgs, &b->dTileBase, &b->order,   &b->ctr,      &b->work,     &b->scanTmp,
                      &b->keys,  &b->vals,      &b->keys2,   &b->vals2,    &b->tcounts,  &b->toffsets,
                      &b->feat,  &b->target,    &b->slotCap, &b->slotOff,  &b->slotUsed, &b->slotOvf,
                      &b->slotIds, &b->slotQ};
    for (DevBuf *x : bufs) x->release();
    b->est.release();
    for (int i = 0; i < 4; i++)
        if (b->ev[i]) (void)hipEventDestroy(b->ev[i]);
    if (b->own) (void)hipStreamDestroy(b->own);
    delete b;
    return GDSJ_OK;
}

int gdsj_batch_autotune(gdsj_batch *b, gdsj_batch_tuning *chosen) {
    if (!b) return GDSJ_EINVAL;
    GDSJ_HIP(hipSetDevice(b->device));
    GDSJ_HIP(hipEventRecord(b->ev[2], b->own));
    // counts and offsets of the timed passes go to scratch of the batch
    GDSJ_TRY(b->tcounts.ensure((size_t)b->npts * 4 + 256));
    const uint32_t tqs[3] = {16, 32, 64};
    const uint32_t wpbs[3] = {1, 2, 4};
    const uint32_t budgets[2] = {24, 48};
    std::vector<uint32_t> ilps;
    if (b->ilp_auto) ilps = {0, 8}; else ilps = {b->p.ilp};
    gdsj_batch_tuning best{};
    best.ms = 1e30;
    const bool learned = b->learned_ready;
    for (uint32_t tq : tqs) {
        if (b->p.tile_queries && tq != b->p.tile_queries) continue;
        if (b->built_tq != tq) GDSJ_TRY(build_all(b, tq));
        for (uint32_t ilp : ilps) {
            b->tune.tile_queries = tq;
            b->tune.ilp = ilp;
            GDSJ_TRY(upload_args(b, b->tcounts.as<uint32_t>(), nullptr, nullptr, 0));
            for (uint32_t budget : budgets)
                for (uint32_t w : wpbs)
                    for (uint32_t lpt = 0; lpt < 2; lpt++) {
                        b->tune.waves_per_block = w;
                        b->tune.reg_budget = budget;
                        b->tune.lpt = lpt;
                        float ms = 0;
                        GDSJ_TRY(time_count(b, &ms));
                        if (ms < best.ms) {
                            best = b->tune;
                            best.ms = ms;
                        }
                    }
        }
    }
    if (best.ms >= 1e30) {
        set_error("autotuner found no candidate");
        return GDSJ_EINVAL;
    }
    if (b->built_tq != best.tile_queries) GDSJ_TRY(build_all(b, best.tile_queries));
    const double ms_keep = best.ms;
    b->tune = best;
    b->tune.ms = ms_keep;
    b->tuned = true;
    b->args_valid = false;
    drop_graph(b);
    (void)learned;  // a rebuild drops the slots; they are learned again on the next run
    GDSJ_HIP(hipEventRecord(b->ev[3], b->own));
    GDSJ_HIP(hipEventSynchronize(b->ev[3]));
    float ms = 0;
    (void)hipEventElapsedTime(&ms, b->ev[2], b->ev[3]);
    b->st.ms_tune = ms;
    if (chosen) *chosen = b->tune;
    return GDSJ_OK;
}

int gdsj_batch_run(gdsj_batch *b, uint32_t *d_counts, uint64_t *d_offsets, uint32_t *d_ids, uint64_t ids_cap,
                   uint64_t *total, void *stream) {
    if (!b || !total || (!d_counts && b->npts) || !d_offsets) return GDSJ_EINVAL;
    GDSJ_HIP(hipSetDevice(b->device));
    *total = 0;
    if (b->npts == 0) {
        GDSJ_HIP(hipMemsetAsync(d_offsets, 0, 8, b->own));
        GDSJ_HIP(hipStreamSynchronize(b->own));
        return GDSJ_OK;
    }
    hipStream_t s = stream ? (hipStream_t)stream : b->own;
    const bool slotted = b->p.estimator == GDSJ_EST_LEARNED && b->learned_ready;
    const uint64_t cap = d_ids ? ids_cap : 0;
    const bool same = b->args_valid && b->gkey[0] == d_counts && b->gkey[1] == d_offsets && b->gkey[2] == d_ids &&
                      b->gcap == cap;
    if (!same) {
        drop_graph(b);
        GDSJ_TRY(upload_args(b, d_counts, d_offsets, d_ids, cap));
        b->gkey[0] = d_counts;
        b->gkey[1] = d_offsets;
        b->gkey[2] = d_ids;
        b->gcap = cap;
    }
    if (!b->exec || b->gslotted != (slotted ? 1 : 0)) {
        drop_graph(b);
        GDSJ_TRY(size_passes(b));
        GDSJ_HIP(hipStreamBeginCapture(b->own, hipStreamCaptureModeThreadLocal));
        int rc = enqueue_run(b, d_counts, d_offsets, slotted, b->own);
        hipGraph_t g = nullptr;
        hipError_t e = hipStreamEndCapture(b->own, &g);
        if (rc != GDSJ_OK) {
            if (g) (void)hipGraphDestroy(g);
            return rc;
        }
        GDSJ_HIP(e);
        b->graph = g;
        GDSJ_HIP(hipGraphInstantiate(&b->exec, b->graph, nullptr, nullptr, 0));
        b->gslotted = slotted ? 1 : 0;
        b->st.graph_captures++;
    }
    GDSJ_HIP(hipEventRecord(b->ev[0], s));
    GDSJ_HIP(hipGraphLaunch(b->exec, s));
    GDSJ_HIP(hipEventRecord(b->ev[1], s));
    uint64_t tot = 0;
    uint32_t c[BQ_N] = {};
    GDSJ_HIP(hipMemcpyAsync(&tot, d_offsets + b->npts, 8, hipMemcpyDeviceToHost, s));
    GDSJ_HIP(hipMemcpyAsync(c, b->ctr.p, sizeof(c), hipMemcpyDeviceToHost, s));
    GDSJ_HIP(hipStreamSynchronize(s));
    float ms = 0;
    (void)hipEventElapsedTime(&ms, b->ev[0], b->ev[1]);
    *total = tot;
    b->st.total_pairs = tot;
    b->st.n_points = b->npts;
    b->st.n_tiles = b->ntiles;
    b->st.ms_last_run = ms;
    b->st.graph_replays++;
    b->st.slotted = slotted ? 1 : 0;
    b->st.est_predicted = slotted ? b->predicted : 0;
    b->st.est_capacity = slotted ? b->slots : 0;
    b->st.est_overflow_tiles = slotted ? c[BQ_NOVF] : 0;
    b->st.tuning = b->tune;
    if (b->st.candidates == 0 && b->ntiles) {
        // queries x candidates of every tile (written by the counting or slotted pass)
        uint64_t *sum = (uint64_t *)(b->ctr.as<uint32_t>() + 16);
        GDSJ_TRY(sum_u64(b->work.as<uint64_t>(), b->ntiles, sum, b->own));
        uint64_t cand = 0;
        GDSJ_HIP(hipMemcpyAsync(&cand, sum, 8, hipMemcpyDeviceToHost, b->own));
        GDSJ_HIP(hipStreamSynchronize(b->own));
        b->st.candidates = cand;
    }
    if (tot > cap) {
        set_error("batch result has %llu pairs, output buffer holds %llu", (unsigned long long)tot,
                  (unsigned long long)cap);
        return GDSJ_EOVERFLOW;
    }
    if (c[BQ_STATUS]) {
        set_error("batch fill exceeded the output capacity (internal inconsistency)");
        return GDSJ_EDEVICE;
    }
    // learned sizing: the first exact run's counts train the model and size the slots
    if (b->p.estimator == GDSJ_EST_LEARNED && !b->learned_ready) {
        if (!b->all_runs) return GDSJ_OK;  // features need every instance's run lists: stay exact
        if (s != b->own) GDSJ_HIP(hipStreamSynchronize(s));
        GDSJ_TRY(learn(b, d_counts));
    }
    return GDSJ_OK;
}

int gdsj_batch_get_stats(gdsj_batch *b, gdsj_batch_stats *out) {
    if (!b || !out) return GDSJ_EINVAL;
    *out = b->st;
    return GDSJ_OK;
}

}  // extern "C"
