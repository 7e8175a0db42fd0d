// api.hip -- the C ABI (include/gdsjoin.h): handle API and the legacy
// GDSJoinPy / copyResultIntoPythonArray drop-in symbols.
//
// Host driver flow per join (replaces GDSJoinPy main.cu:360-505 and
// distanceTableNDGridBatches GPU.cu:270-1088):
//   grid_build (REORDER, extent, keys, radix sort, cells, SoA, runs, tiles)
//   -> count pass -> exclusive scan of counts (CSR offsets, total)
//   -> fill pass straight into the caller's CSR buffer.
// No sort of result pairs, no per-batch host tables: the device CSR is the
// neighbor table (the reference's neighborTableLookup, structs.h:55-85).
#include <math.h>
#include <stdarg.h>
#include <stdlib.h>
#include <string.h>

#include <atomic>
#include <cmath>
#include <mutex>
#include <thread>
#include <vector>

#include "bruteforce.h"
#include "estimator.h"
#include "gram.h"
#include "grid.h"
#include "join.h"
#include "scan.h"
#include "sortrows.h"

namespace gdsj {

static thread_local char g_err[512] = "";

void set_error(const char *fmt, ...) {
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(g_err, sizeof(g_err), fmt, ap);
    va_end(ap);
}

}  // namespace gdsj

using namespace gdsj;

// Host-buffer copies (gdsjoin's path) run through a ring of STAGE_RING
// pinned buffers of STAGE_CHUNK bytes: one stream moves the chunks in order
// (the link stays busy) while up to STAGE_THREADS host threads copy chunks
// out of (or into) the buffers, chunk c through buffer c % STAGE_RING by
// thread c % threads, and hand each buffer back for chunk c + STAGE_RING.
constexpr int STAGE_THREADS = 8;
constexpr int STAGE_RING = 16;
constexpr size_t STAGE_CHUNK = 4u << 20;

struct gdsj_handle {
    int device = 0;
    hipStream_t own = nullptr;
    hipStream_t stream = nullptr;
    Grid G;
    DevBuf counts, offsets, ids, data, scanTmp, status, work;
    // slotted single pass (estimator-sized per-tile result slots)
    DevBuf slotCap, slotOff, slotUsed, slotOvf, slotIds, slotQ, scalars, estPred, estScratch, subTiles;
    DevBuf winTiles, winCnt, winOff;  // window splits of the overflowed Gram tiles
    DevBuf sortWork, sampleScratch;   // row sort and sampling-estimator scratch
    Estimator est;
    GramData gram;
    DevBuf gramStats;
    bool use_gram = false;
    bool stamp = false;  // unicomp in force for this join
    DevBuf fcnt, rcur;   // unicomp: own-side counts, reverse-entry cursors
    DevBuf maskW, hmCell, hmask;  // hit masks of the exact passes
    // matrix-core join (gdsj_params.mfma): tile lists of the last grid
    DevBuf mjList, mjFlags;
    hipStream_t mjs = nullptr;  // the matrix-core tiles' stream (default priority)
    hipEvent_t mjFork = nullptr, mjJoin = nullptr;
    bool use_mj = false;
    uint32_t mj_nbig = 0;
    float mj_hi = 0.0f, mj_lo = 0.0f;
    DevBuf bfXs;                  // all-pairs join: SoA copy of the points
    bool use_hm = false;
    bool slotted = false;
    bool have_grid = false;  // a grid, runs and tiles of the last join (gdsj_rejoin_device)
    uint32_t novf = 0;
    // the last all-pairs join's setting, for gdsj_bruteforce_rows_host
    bool bf_ok = false;
    uint32_t bf_n = 0, bf_dims = 0, bf_ilp = 0;
    int bf_dtype = 0, bf_mode = 0;
    double bf_thr = 0;
    uint64_t bf_ld = 0, bf_total = 0;
    gdsj_params p{};
    int mode = 0;
    double thr = 0;
    bool have_counts = false;
    bool have_table = false;
    uint64_t total = 0;
    uint32_t *cur_counts = nullptr;
    uint64_t *cur_offsets = nullptr;
    gdsj_stats st{};
    hipEvent_t ev[9] = {};  // 8: start of the row sort
    // host-buffer copies: per copy thread a stream, two pinned chunk buffers
    // and their events (staged_copy)
    hipStream_t cstream = nullptr;
    hipEvent_t cev[STAGE_RING] = {};
    HostBuf pin[STAGE_RING];
};

static int staged_copy(gdsj_handle *h, void *dst, const void *src, size_t bytes, bool d2h);

extern "C" {

const char *gdsj_last_error(void) { return gdsj::g_err; }

int gdsj_params_default(gdsj_params *p) {
    if (!p) return GDSJ_EINVAL;
    memset(p, 0, sizeof(*p));
    // params.h as shipped (params.h:23,40-60): NUMINDEXEDDIM 6, STAMP 0,
    // REORDER 1, ILP 8, SHORTCIRCUIT 1, QUERYREORDER 1
    p->k = 0;
    p->ilp = 8;
    p->shortcircuit = 1;
    p->reorder = 1;
    p->stamp = 0;
    p->queryreorder = 1;
    p->filter = GDSJ_FILTER_AUTO;
    p->estimator = GDSJ_EST_EXACT;
    p->sort_neighbors = 0;
    p->verbose = 0;
    p->tile_queries = 0;
    p->schedule = 0;
    p->prune = 0;
    p->train = 0;
    return GDSJ_OK;
}

float gdsj_eps_threshold_f32(float eps) {
    if (!(eps >= 0)) return NAN;
    float t = eps * eps;
    while (t > 0 && sqrtf(t) > eps) t = nextafterf(t, 0.0f);
    for (;;) {
        float u = nextafterf(t, INFINITY);
        if (std::isinf(u) || !(sqrtf(u) <= eps)) break;
        t = u;
    }
    return t;
}

double gdsj_eps_threshold_f64(double eps) {
    if (!(eps >= 0)) return NAN;
    double t = eps * eps;
    while (t > 0 && sqrt(t) > eps) t = nextafter(t, 0.0);
    for (;;) {
        double u = nextafter(t, INFINITY);
        if (std::isinf(u) || !(sqrt(u) <= eps)) break;
        t = u;
    }
    return t;
}

int gdsj_create(int device, gdsj_handle **out) {
    if (!out) return GDSJ_EINVAL;
    *out = nullptr;
    int ndev = 0;
    GDSJ_HIP(hipGetDeviceCount(&ndev));
    if (device < 0 || device >= ndev) {
        set_error("device %d out of range (%d devices)", device, ndev);
        return GDSJ_EINVAL;
    }
    GDSJ_HIP(hipSetDevice(device));
    gdsj_handle *h = new gdsj_handle();
    h->device = device;
    hipError_t e = hipStreamCreateWithFlags(&h->own, hipStreamNonBlocking);
    if (e != hipSuccess) {
        set_error("hipStreamCreate: %s", hipGetErrorString(e));
        delete h;
        return GDSJ_EDEVICE;
    }
    for (int i = 0; i < 9; i++) (void)hipEventCreate(&h->ev[i]);
    h->stream = h->own;
    gdsj_params_default(&h->p);
    *out = h;
    return GDSJ_OK;
}

int gdsj_destroy(gdsj_handle *h) {
    if (!h) return GDSJ_OK;
    (void)hipSetDevice(h->device);
    if (h->own) (void)hipStreamSynchronize(h->own);
    h->G.release();
    h->fcnt.release();
    h->rcur.release();
    h->maskW.release();
    h->bfXs.release();
    h->hmCell.release();
    h->hmask.release();
    h->mjList.release();
    h->mjFlags.release();
    DevBuf *b[] = {&h->counts,  &h->offsets,  &h->ids,     &h->data,    &h->scanTmp, &h->status,
                   &h->work,    &h->slotCap,  &h->slotOff, &h->slotUsed, &h->slotOvf, &h->slotIds,
                   &h->slotQ,   &h->scalars,  &h->estPred, &h->estScratch, &h->subTiles,
                   &h->winTiles, &h->winCnt,  &h->winOff, &h->sortWork, &h->sampleScratch};
    h->est.release();
    h->gram.release();
    h->gramStats.release();
    for (DevBuf *x : b) x->release();
    for (int i = 0; i < 9; i++)
        if (h->ev[i]) (void)hipEventDestroy(h->ev[i]);
    if (h->cstream) (void)hipStreamSynchronize(h->cstream);
    for (int k = 0; k < STAGE_RING; k++) {
        if (h->cev[k]) (void)hipEventDestroy(h->cev[k]);
        h->pin[k].release();
    }
    if (h->cstream) (void)hipStreamDestroy(h->cstream);
    if (h->mjs) {
        (void)hipStreamSynchronize(h->mjs);
        (void)hipStreamDestroy(h->mjs);
    }
    if (h->mjFork) (void)hipEventDestroy(h->mjFork);
    if (h->mjJoin) (void)hipEventDestroy(h->mjJoin);
    if (h->own) (void)hipStreamDestroy(h->own);
    delete h;
    return GDSJ_OK;
}

static int resolve_params(const gdsj_params *in, uint32_t dims, gdsj_params *out) {
    if (in) *out = *in; else gdsj_params_default(out);
    if (out->k == 0) out->k = dims < 6 ? dims : 6;
    if (out->k > dims || out->k > (uint32_t)MAXK) {
        set_error("NUMINDEXEDDIM k=%u must be <= dims=%u and <= %d", out->k, dims, MAXK);
        return GDSJ_EINVAL;
    }
    if (out->ilp > 64) {
        set_error("ILP=%u exceeds 64", out->ilp);
        return GDSJ_EINVAL;
    }
    if (out->cpu_sum) {
        set_error("cpu_sum applies to the all-pairs join only");
        return GDSJ_EUNSUPPORTED;
    }
    if (out->sort > 1) {
        set_error("sort=%u: 0 off, 1 on", out->sort);
        return GDSJ_EINVAL;
    }
    if (out->sort && out->stamp) {
        set_error("SORT runs without unicomp (stamp 0)");
        return GDSJ_EUNSUPPORTED;
    }
    if (out->sort && (out->filter == GDSJ_FILTER_GRAM || out->filter == GDSJ_FILTER_GRAM_BF16)) {
        set_error("SORT runs with the plain distance kernels (filter PLAIN or AUTO)");
        return GDSJ_EUNSUPPORTED;
    }
    if (out->vfilter > 2) {
        set_error("vfilter=%u: 0 auto, 1 on, 2 off", out->vfilter);
        return GDSJ_EINVAL;
    }
    if (out->index > 2) {
        set_error("index=%u: 0 auto, 1 the reference's dims, 2 principal axes", out->index);
        return GDSJ_EINVAL;
    }
    if (out->ffilter > 2) {
        set_error("ffilter=%u: 0 auto, 1 on, 2 every pair by the reference's arithmetic", out->ffilter);
        return GDSJ_EINVAL;
    }
    if (out->gram_axes != 0 && out->gram_axes != 1 && out->gram_axes != 16 && out->gram_axes != 32) {
        set_error("gram_axes=%u: 0 auto, 1 the points' dims, 16 or 32 principal axes", out->gram_axes);
        return GDSJ_EINVAL;
    }
    if (out->ilp_fold > 1) {
        set_error("ilp_fold=%u: 0 or 1", out->ilp_fold);
        return GDSJ_EINVAL;
    }
    if (out->stamp > 2) {
        set_error("stamp=%u: 0 off, 1 STAMP, 2 LINEARSTAMP", out->stamp);
        return GDSJ_EINVAL;
    }
    return GDSJ_OK;
}

static float ms_between(hipEvent_t a, hipEvent_t b) {
    float ms = 0;
    if (hipEventElapsedTime(&ms, a, b) != hipSuccess) return 0;
    return ms;
}

static JoinArgs make_args(gdsj_handle *h) {
    JoinArgs a{};
    a.xs = h->G.xs.p;
    a.ld = h->G.ld;
    a.ixs = h->G.indexUsed == 2 ? h->G.ixs.p : h->G.xs.p;
    if (h->G.xaStride) {
        a.xa = h->G.xa.p;
        a.xaStride = h->G.xaStride;
    }
    a.ids = h->G.ids.as<uint32_t>();
    a.tiles = h->G.tiles.as<uint4>();
    a.ntiles = h->G.ntiles;
    a.cellStart = h->G.cellStart.as<uint32_t>();
    a.cellCoord = h->G.cellCoord.as<uint32_t>();
    a.cellKey = h->G.cellKey.as<uint64_t>();
    a.hkeys = h->G.hkeys.as<uint64_t>();
    a.hvals = h->G.hvals.as<uint32_t>();
    a.ga = h->G.ga;
    a.ncombo = 1;
    for (uint32_t j = 1; j < h->G.k; j++) a.ncombo *= 3;
    a.thr = h->thr;
    fma_band(h->G.dtype, h->G.d, h->p.ilp, h->thr, h->p.ffilter, &a.ffLo, &a.ffHi);
    a.d = h->G.d;
    a.ilp = h->p.ilp;
    a.sc = h->p.shortcircuit;
    a.status = h->status.as<uint32_t>();
    // auto: round-robin workgroup order (XCD-contiguous ranges unbalance
    // clustered data); query-box pruning from k >= 4, where it pays
    a.xcd_map = h->p.schedule == 2 ? 1u : 0u;
    a.prune = h->p.prune == 1 || (h->p.prune == 0 && h->G.k >= 4) ? 1u : 0u;
    if (h->stamp) {
        a.stamp = h->p.stamp;
        a.fcounts = h->fcnt.as<uint32_t>();
        a.rcur = h->rcur.as<uint32_t>();
    }
    // the Gram-form prefilter on request (auto leaves it off: on gmm6d's
    // slotted pass it issues 30 % fewer VALU instructions but runs 15 %
    // slower, its per-chunk band test and branches serialising the waves)
    if (h->p.vfilter == 1 && !h->p.sort && gram_valu_band(h->G, h->thr, &a.gvLo, &a.gvMid, &a.gvHalf)) a.gv = 1;
    a.mjHi = h->mj_hi;
    a.mjLo = h->mj_lo;
    if (h->p.sort) {
        a.sortMode = h->G.k == h->G.d ? 1u : 2u;
        a.sortDim = h->G.k == h->G.d ? 0u : h->G.k;
        a.posCell = h->G.posCell.as<uint32_t>();
    }
    if (h->G.haveRuns) {
        a.runs = h->G.runs.as<uint2>();
        a.nruns = h->G.nruns.as<uint32_t>();
        a.runStride = h->G.runStride;
    }
    if (h->use_hm) {
        a.hmask = h->hmask.as<uint64_t>();
        a.hmCell = h->hmCell.as<uint64_t>();
        a.cellCand = h->G.cellCand.as<uint32_t>();
    }
    return a;
}

// The passes over the grid's own tiles split them between the matrix-core
// join and the plain kernels (sub-tiles and sampled tiles keep the plain
// kernels).
static void set_mj_lists(gdsj_handle *h, JoinArgs &a) {
    if (!h->use_mj) return;
    a.mjList = h->mjList.as<uint32_t>();
    a.mjN = h->mj_nbig;
    a.vjList = a.mjList + h->mj_nbig;
    a.vjN = h->G.ntiles - h->mj_nbig;
    a.mjStream = h->mjs;
    a.mjFork = h->mjFork;
    a.mjJoin = h->mjJoin;
}

// Writes the table of the last join into d_ids: the exact fill pass, or,
// after a slotted pass, compaction of the slots plus the fill pass for the
// tiles whose slot overflowed.
// one join pass through the plain kernels or the MFMA Gram filter
static int run_pass(gdsj_handle *h, const JoinArgs &a, int pass) {
    if (h->use_gram)
        return gram_join_pass(h->G, h->gram, a, h->mode, pass,
                              pass == PASS_FILL ? nullptr : h->gramStats.as<unsigned long long>(), h->stream,
                              h->p.filter_stats != 0);
    return join_pass(h->G, a, h->mode, pass, h->stream);
}

static void set_slot_args(gdsj_handle *h, JoinArgs &a) {
    a.slotOff = h->slotOff.as<uint64_t>();
    a.slotCap = h->slotCap.as<uint32_t>();
    a.slotUsed = h->slotUsed.as<uint32_t>();
    a.slotIds = h->slotIds.as<uint32_t>();
    a.slotQ = h->slotQ.as<uint8_t>();
    a.ovfList = h->slotOvf.as<uint32_t>();
    a.ovfCount = h->scalars.as<uint32_t>();
}

// Queues the table's passes (compaction and fills) on the handle's stream;
// fill_finish waits for them.  Writes stop at cap in any case.
static int fill_enqueue(gdsj_handle *h, uint32_t *d_ids, uint64_t cap) {
    hipStream_t s = h->stream;
    JoinArgs a = make_args(h);
    a.offsets = h->cur_offsets;
    a.out = d_ids;
    a.cap = cap;
    GDSJ_HIP(hipMemsetAsync(h->status.p, 0, 4, s));
    GDSJ_HIP(hipEventRecord(h->ev[4], s));
    if (h->stamp && h->G.n) GDSJ_HIP(hipMemsetAsync(h->rcur.p, 0, (size_t)h->G.n * 4, s));
    if (h->slotted) {
        set_slot_args(h, a);
        GDSJ_TRY(compact_slots(a, s));
        GDSJ_HIP(hipEventRecord(h->ev[7], s));
        if (h->novf) {
            // the overflowed tiles are the heaviest and few: their exact fill
            // runs on sub-tiles (32 queries for the Gram kernel, 16 for the
            // plain one) so that it is not bound by one wave's latency
            const uint32_t subq = h->use_gram ? 32u : 16u;
            const uint32_t per = (h->G.tileq + subq - 1) / subq;
            GDSJ_TRY(h->subTiles.ensure((size_t)h->novf * per * 16 + 64));
            uint32_t nsub = 0;
            GDSJ_TRY(split_tiles(h->slotOvf.as<uint32_t>(), h->novf, h->G.tiles.as<uint4>(), h->G.tileq, subq,
                                 h->subTiles.as<uint4>(), &nsub, s));
            a.tiles = h->subTiles.as<uint4>();
            a.tileList = nullptr;
            a.ntiles = nsub;
            constexpr uint32_t WSPLIT = 32;
            if (h->use_gram && !h->stamp && h->G.haveRuns && (uint64_t)nsub * WSPLIT <= 65536) {
                // a Gram sub-tile still streams its cell's whole
                // neighborhood, the largest of the join: its candidates are
                // split WSPLIT ways as well (equal ranges of the cell's
                // candidate order), counted, placed by a prefix over the
                // splits, then filled (up to 65,536 split tiles, 100 MB of
                // split counts and offsets)
                const uint64_t nw = (uint64_t)nsub * WSPLIT;
                GDSJ_TRY(h->winTiles.ensure(nw * 16 + 64));
                GDSJ_TRY(h->winCnt.ensure(nw * SPLIT_ROWS * 4 + 64));
                GDSJ_TRY(h->winOff.ensure(nw * SPLIT_ROWS * 8 + 64));
                GDSJ_TRY(split_windows(h->subTiles.as<uint4>(), nsub, WSPLIT, h->winTiles.as<uint4>(), s));
                JoinArgs b = a;
                b.tiles = h->winTiles.as<uint4>();
                b.ntiles = (uint32_t)nw;
                b.winSplit = WSPLIT;
                b.cellCand = h->G.cellCand.as<uint32_t>();
                b.tileWork = nullptr;
                b.splitCnt = h->winCnt.as<uint32_t>();
                GDSJ_TRY(gram_join_pass(h->G, h->gram, b, h->mode, PASS_COUNT, nullptr, s, false));
                GDSJ_TRY(split_offsets(b.tiles, nsub, WSPLIT, a.ids, a.offsets, b.splitCnt, h->winOff.as<uint64_t>(),
                                       s));
                b.splitCnt = nullptr;
                b.splitOff = h->winOff.as<uint64_t>();
                GDSJ_TRY(gram_join_pass(h->G, h->gram, b, h->mode, PASS_FILL, nullptr, s, false));
            } else {
                GDSJ_TRY(run_pass(h, a, PASS_FILL));
            }
        }
    } else {
        JoinArgs f = a;
        set_mj_lists(h, f);
        GDSJ_TRY(run_pass(h, f, PASS_FILL));
        GDSJ_HIP(hipEventRecord(h->ev[7], s));
    }
    GDSJ_HIP(hipEventRecord(h->ev[5], s));
    return GDSJ_OK;
}

// Sorts the rows when asked, waits for the stream and checks the fill.
static int fill_finish(gdsj_handle *h, uint32_t *d_ids) {
    hipStream_t s = h->stream;
    if (h->p.sort_neighbors) {
        GDSJ_HIP(hipEventRecord(h->ev[8], s));
        GDSJ_TRY(sort_rows(h->cur_offsets, h->G.n, d_ids, s, h->sortWork));
        GDSJ_HIP(hipEventRecord(h->ev[6], s));
    }
    uint32_t flag = 0;
    GDSJ_HIP(hipMemcpyAsync(&flag, h->status.p, 4, hipMemcpyDeviceToHost, s));
    GDSJ_HIP(hipStreamSynchronize(s));
    if (flag) {
        set_error("fill pass exceeded the output capacity (internal inconsistency)");
        return GDSJ_EDEVICE;
    }
    h->have_table = true;
    if (h->slotted) {
        h->st.ms_compact = ms_between(h->ev[4], h->ev[7]);
        h->st.ms_fill = ms_between(h->ev[7], h->ev[5]);
    } else {
        h->st.ms_fill = ms_between(h->ev[4], h->ev[5]);
    }
    h->st.ms_sort = h->p.sort_neighbors ? ms_between(h->ev[8], h->ev[6]) : 0.0;
    return GDSJ_OK;
}

static int do_fill(gdsj_handle *h, uint32_t *d_ids, uint64_t cap) {
    GDSJ_TRY(fill_enqueue(h, d_ids, cap));
    return fill_finish(h, d_ids);
}

// Single pass with per-tile result slots sized by an estimator.  Returns
// GDSJ_OK with *used = true when the slotted pass ran (counts written).
static int slotted_pass(gdsj_handle *h, JoinArgs &a, bool *used) {
    *used = false;
    hipStream_t s = h->stream;
    Grid &G = h->G;
    const uint32_t nt = G.ntiles;
    if (!G.haveRuns || nt == 0) return GDSJ_OK;
    const bool learned = h->p.estimator == GDSJ_EST_LEARNED;
    if (learned && !h->est.trained) return GDSJ_OK;  // untrained: the exact passes (and training) run
    GDSJ_TRY(h->slotCap.ensure((size_t)nt * 4 + 64));
    GDSJ_TRY(h->slotOff.ensure(((size_t)nt + 1) * 8 + 64));
    GDSJ_TRY(h->slotUsed.ensure((size_t)nt * 4 + 64));
    GDSJ_TRY(h->slotOvf.ensure((size_t)nt * 4 + 64));
    GDSJ_TRY(h->estPred.ensure((size_t)nt * 4 + 64));
    uint64_t predicted = 0;
    GDSJ_HIP(hipEventRecord(h->ev[1], s));
    if (learned) {
        GDSJ_TRY(estimator_capacity_learned(h->est, G, h->slotCap.as<uint32_t>(), &predicted, s));
        GDSJ_HIP(hipMemcpyAsync(h->estPred.p, h->est.pred.p, (size_t)nt * 4, hipMemcpyDeviceToDevice, s));
    } else {
        GDSJ_TRY(estimator_capacity_sampling(G, &a, h->mode, a.counts, h->slotCap.as<uint32_t>(), &predicted,
                                             h->estPred.as<float>(), s, h->sampleScratch));
    }
    GDSJ_TRY(h->scanTmp.ensure(scan_tmp_bytes((uint64_t)nt + 1) + scan_tmp_bytes((uint64_t)G.n + 1) + 256));
    GDSJ_TRY(exclusive_scan_u32_u64(h->slotCap.as<uint32_t>(), h->slotOff.as<uint64_t>(), nt, true,
                                    h->scanTmp.p, s));
    uint64_t slots = 0;
    GDSJ_HIP(hipMemcpyAsync(&slots, h->slotOff.as<uint64_t>() + nt, 8, hipMemcpyDeviceToHost, s));
    GDSJ_HIP(hipStreamSynchronize(s));
    h->st.est_predicted = predicted;
    h->st.est_capacity = slots;
    h->st.est_tiles = nt;
    GDSJ_TRY(h->slotIds.ensure((size_t)slots * 4 + 256));
    GDSJ_TRY(h->slotQ.ensure((size_t)slots + 256));
    GDSJ_TRY(h->scalars.ensure(64));
    GDSJ_HIP(hipMemsetAsync(h->scalars.p, 0, 4, s));
    GDSJ_HIP(hipEventRecord(h->ev[2], s));
    set_slot_args(h, a);
    JoinArgs m = a;
    set_mj_lists(h, m);
    GDSJ_TRY(run_pass(h, m, PASS_SLOT));
    *used = true;
    return GDSJ_OK;
}

static int join_passes(gdsj_handle *h, uint32_t n, uint32_t dims, uint32_t *d_counts, uint64_t *d_offsets,
                       uint32_t *d_ids, uint64_t ids_cap, uint64_t *total);

int gdsj_join_device(gdsj_handle *h, const void *d_data, uint32_t n, uint32_t dims, int dtype, double epsilon,
                     const gdsj_params *pin, uint32_t *d_counts, uint64_t *d_offsets, uint32_t *d_ids,
                     uint64_t ids_cap, uint64_t *total, void *stream) {
    if (!h || (!d_data && n) || (!d_counts && n) || !total) return GDSJ_EINVAL;
    if (dtype != GDSJ_DTYPE_F32 && dtype != GDSJ_DTYPE_F64) {
        set_error("dtype must be GDSJ_DTYPE_F32 or GDSJ_DTYPE_F64");
        return GDSJ_EINVAL;
    }
    GDSJ_HIP(hipSetDevice(h->device));
    h->have_counts = h->have_table = false;
    h->have_grid = false;
    h->slotted = false;
    h->novf = 0;
    GDSJ_TRY(resolve_params(pin, dims, &h->p));
    h->stream = stream ? (hipStream_t)stream : h->own;
    hipStream_t s = h->stream;
    if (dtype == GDSJ_DTYPE_F32) {
        float e = (float)epsilon;
        if ((double)e != epsilon) {
            set_error("epsilon %.17g is not a float value", epsilon);
            return GDSJ_EINVAL;
        }
        h->thr = (double)gdsj_eps_threshold_f32(e);
    } else {
        h->thr = gdsj_eps_threshold_f64(epsilon);
    }
    // ILP > 0 splits the sum (kernel.cu:312-342); its fold depends on SHORTCIRCUIT
    h->mode = acc_mode(h->p.ilp, h->p.shortcircuit, h->p.ilp_fold);
    memset(&h->st, 0, sizeof(h->st));
    GDSJ_TRY(h->status.ensure(256));
    GDSJ_HIP(hipEventRecord(h->ev[0], s));
    // the MFMA Gram filter for dims >= 16 (AUTO) or on request
    if (h->p.filter > GDSJ_FILTER_GRAM_BF16) {
        set_error("filter=%u is not a GDSJ_FILTER_* mode", h->p.filter);
        return GDSJ_EINVAL;
    }
    const bool gram_asked = h->p.filter == GDSJ_FILTER_GRAM || h->p.filter == GDSJ_FILTER_GRAM_BF16;
    h->use_gram = (gram_asked || h->p.filter == GDSJ_FILTER_AUTO) && gram_supported(dims) && !h->p.sort;
    // operands: fp16 (one MFMA per k-step) unless bf16 hi + lo is asked for FP32 data
    const int gram_lowp = (h->p.filter == GDSJ_FILTER_GRAM_BF16 && dtype == GDSJ_DTYPE_F32) ? 0 : 1;
    if (gram_asked && !h->use_gram) {
        set_error("the Gram filter handles 16 to 128 dims (dims=%u)", dims);
        return GDSJ_EUNSUPPORTED;
    }
    // unicomp (STAMP) with the plain kernels and the exact passes; with the
    // Gram filter or an estimator the full neighborhood is evaluated instead
    // (the result is the same either way)
    h->stamp = h->p.stamp && !h->use_gram && h->p.estimator == GDSJ_EST_EXACT;
    if (h->stamp) {
        GDSJ_TRY(h->fcnt.ensure((size_t)n * 4 + 256));
        GDSJ_TRY(h->rcur.ensure((size_t)n * 4 + 256));
    }
    uint32_t tq = h->p.tile_queries ? h->p.tile_queries : 64;
    if (h->use_gram) {
        const uint32_t gq = gram_tile_queries(dims, gram_lowp);
        if (!h->p.tile_queries || tq > gq) tq = gq;
    } else if (tq > 64) {
        tq = 64;  // the plain kernels keep one query per lane
    }
    h->G.sortCells = h->p.sort ? 1 : 0;
    // the index: the reference's dims, or the principal axes (not with
    // unicomp, whose ownership rules are stated on the reference's cells)
    h->G.indexWant = h->stamp ? 1 : (h->p.index == 0 ? 0 : (int)h->p.index);
    // principal axes indexed: two fewer than k from k = 6, for cells that
    // hold more queries sharing each candidate (measured with the Gram filter
    // on 64-D data near an 8-D subspace, k = 8: 5 to 7 axes 22.4 ms per
    // join, 8 axes 32.4 ms); the same for every filter, so that their rows
    // come in the same order
    h->G.indexDims = h->p.index_dims ? h->p.index_dims : (h->p.k >= 6 ? h->p.k - 2 : h->p.k);
    h->G.ilpTerms = h->p.ilp;
    // the Gram filter's operands on principal axes (its basis comes with the index's)
    h->G.gramAxesWant = h->use_gram ? (int)h->p.gram_axes : 1;
    // the row-major copy for the plain kernels' candidate loads (up to 16 dims)
    h->G.aosWant = !h->use_gram && dims <= 16 ? 1 : 0;
    GDSJ_TRY(grid_build(h->G, d_data, n, dims, dtype, epsilon, h->p.k, h->p.reorder, tq, h->p.queryreorder, s));
    if (h->use_gram) {
        GDSJ_TRY(gram_prepare(h->gram, h->G, gram_lowp, s));
        GDSJ_TRY(h->gramStats.ensure(64));
    }
    // hit masks: exact passes over precomputed runs, plain kernels, no unicomp
    // (auto: from 4 dims, where a pair's distance costs more than reading
    // its mask bit back; at 2 dims the fill pass is faster recomputing)
    h->use_hm = false;
    const bool hm_ok = !h->use_gram && !h->stamp && h->p.estimator == GDSJ_EST_EXACT && !h->p.filter_stats &&
                       !h->p.sort;
    if (h->p.hitmask == 1 && !hm_ok) {
        set_error("hit masks need the exact passes without the Gram filter and without unicomp");
        return GDSJ_EINVAL;
    }
    const bool want_hm = hm_ok && n > 0 && (h->p.hitmask == 1 || (h->p.hitmask == 0 && dims >= 4));
    {
        bool built = false;
        const bool prune = h->p.prune == 1 || (h->p.prune == 0 && h->G.k >= 4);
        if (want_hm) GDSJ_TRY(h->maskW.ensure((size_t)h->G.ncell * 4 + 256));
        GDSJ_TRY(build_runs(h->G, s, prune, h->stamp ? (int)h->p.stamp : 0, &built,
                            want_hm ? h->maskW.as<uint32_t>() : nullptr));
        h->G.haveRuns = built;
    }
    if (want_hm && h->G.haveRuns) {
        // per-cell first mask word; the total decides whether they fit
        const uint32_t nc = h->G.ncell;
        GDSJ_TRY(h->hmCell.ensure(((size_t)nc + 1) * 8 + 256));
        GDSJ_TRY(h->scanTmp.ensure(scan_tmp_bytes((uint64_t)nc + 1) + scan_tmp_bytes((uint64_t)n + 1) + 256));
        GDSJ_TRY(exclusive_scan_u32_u64(h->maskW.as<uint32_t>(), h->hmCell.as<uint64_t>(), nc, true, h->scanTmp.p,
                                        s));
        uint64_t words = 0;
        GDSJ_HIP(hipMemcpyAsync(&words, h->hmCell.as<uint64_t>() + nc, 8, hipMemcpyDeviceToHost, s));
        GDSJ_HIP(hipStreamSynchronize(s));
        const uint64_t limit = 8ull << 30;
        if (words * 8 <= limit) {
            GDSJ_TRY(h->hmask.ensure((size_t)words * 8 + 256));
            h->use_hm = true;
        } else if (h->p.hitmask == 1) {
            set_error("hit masks need %llu bytes (limit %llu)", (unsigned long long)(words * 8),
                      (unsigned long long)limit);
            return GDSJ_EUNSUPPORTED;
        }
    }
    // the matrix-core join: every dim indexed, 3 to 7 dims, plain kernels
    // over precomputed runs, tiles of at most 64 queries
    h->use_mj = false;
    h->mj_nbig = 0;
    if (h->p.mfma != 0 && h->p.mfma != 2 && !h->use_gram && !h->stamp && !h->p.sort && h->G.haveRuns &&
        h->p.vfilter != 1 && tq <= 64 && h->G.ntiles > 0 && mjoin_band(h->G, h->thr, &h->mj_hi, &h->mj_lo)) {
        const uint32_t nt = h->G.ntiles;
        const uint32_t minOcc = h->p.mfma == 1 ? 1u : h->p.mfma;
        GDSJ_TRY(h->mjList.ensure((size_t)nt * 4 + 256));
        GDSJ_TRY(h->mjFlags.ensure(((size_t)nt + 1) * 8 + 256));
        GDSJ_TRY(h->scanTmp.ensure(scan_tmp_bytes((uint64_t)nt + 1) + scan_tmp_bytes((uint64_t)n + 1) + 256));
        GDSJ_TRY(mjoin_partition(h->G, minOcc, h->mjList.as<uint32_t>(), h->mjFlags.as<uint32_t>(), h->scanTmp.p,
                                 &h->mj_nbig, s));
        h->use_mj = h->mj_nbig > 0;
        if (h->use_mj && !h->mjs) {
            GDSJ_HIP(hipStreamCreateWithFlags(&h->mjs, hipStreamNonBlocking));
            GDSJ_HIP(hipEventCreateWithFlags(&h->mjFork, hipEventDisableTiming));
            GDSJ_HIP(hipEventCreateWithFlags(&h->mjJoin, hipEventDisableTiming));
        }
    }
    h->have_grid = true;
    return join_passes(h, n, dims, d_counts, d_offsets, d_ids, ids_cap, total);
}

// The passes of a join over the handle's grid (built by the last
// gdsj_join_device): counting or slotted pass, scan, fill, training.
static int join_passes(gdsj_handle *h, uint32_t n, uint32_t dims, uint32_t *d_counts, uint64_t *d_offsets,
                       uint32_t *d_ids, uint64_t ids_cap, uint64_t *total) {
    hipStream_t s = h->stream;
    if (h->use_gram) GDSJ_HIP(hipMemsetAsync(h->gramStats.p, 0, 24, s));
    uint64_t *offs = d_offsets;
    if (!offs) {
        GDSJ_TRY(h->offsets.ensure(((size_t)n + 1) * 8));
        offs = h->offsets.as<uint64_t>();
    }
    h->cur_counts = d_counts;
    h->cur_offsets = offs;
    JoinArgs a = make_args(h);
    a.counts = d_counts;
    GDSJ_TRY(h->scanTmp.ensure(scan_tmp_bytes((uint64_t)n + 1) + 256));
    GDSJ_TRY(h->work.ensure((size_t)h->G.ntiles * 8 + 256));
    a.tileWork = h->work.as<uint64_t>();
    // ---- counting: slotted single pass when an estimator sizes the slots,
    //      else the exact counting pass
    bool slotted = false;
    // filter-error measurement runs in the exact counting pass
    if (h->p.estimator != GDSJ_EST_EXACT && !(h->use_gram && h->p.filter_stats))
        GDSJ_TRY(slotted_pass(h, a, &slotted));
    if (!slotted) {
        GDSJ_HIP(hipEventRecord(h->ev[1], s));
        GDSJ_HIP(hipEventRecord(h->ev[2], s));
        if (h->stamp && n) GDSJ_HIP(hipMemsetAsync(d_counts, 0, (size_t)n * 4, s));  // accumulated by atomics
        JoinArgs m = a;
        set_mj_lists(h, m);
        GDSJ_TRY(run_pass(h, m, PASS_COUNT));
    }
    h->slotted = slotted;
    GDSJ_HIP(hipEventRecord(h->ev[3], s));
    GDSJ_TRY(exclusive_scan_u32_u64(d_counts, offs, n, true, h->scanTmp.p, s));
    GDSJ_HIP(hipEventRecord(h->ev[4], s));
    uint64_t *candSum = h->status.as<uint64_t>() + 4;
    GDSJ_TRY(sum_u64(a.tileWork, h->G.ntiles, candSum, s));
    // exact passes into a caller's buffer: the fill is queued behind the
    // scan before the total is known (its writes stop at ids_cap), so the
    // join waits for the device once
    const bool queued = d_ids != nullptr && !slotted;
    if (queued) GDSJ_TRY(fill_enqueue(h, d_ids, ids_cap));
    uint64_t hres[3] = {0, 0, 0}, gst[3] = {0, 0, 0};
    if (h->use_gram) GDSJ_HIP(hipMemcpyAsync(gst, h->gramStats.p, 24, hipMemcpyDeviceToHost, s));
    GDSJ_HIP(hipMemcpyAsync(&hres[0], offs + n, 8, hipMemcpyDeviceToHost, s));
    GDSJ_HIP(hipMemcpyAsync(&hres[1], candSum, 8, hipMemcpyDeviceToHost, s));
    if (slotted) GDSJ_HIP(hipMemcpyAsync(&hres[2], h->scalars.p, 4, hipMemcpyDeviceToHost, s));
    GDSJ_HIP(hipStreamSynchronize(s));
    uint64_t tot = hres[0];
    h->novf = (uint32_t)hres[2];
    h->total = tot;
    *total = tot;
    h->have_counts = true;
    h->st.total_pairs = tot;
    h->st.candidates = hres[1];
    h->st.n_nonempty_cells = h->G.ncell;
    h->st.total_cells = h->G.totalCells;
    h->st.index_used = (uint64_t)h->G.indexUsed;
    h->st.index_risky = h->G.riskyPoints;
    h->st.gram_axes = h->use_gram ? h->gram.axes : 0u;
    h->st.n_tiles = h->G.ntiles;
    h->st.n_runs = h->G.haveRuns ? 0 : (uint64_t)h->G.ntiles * a.ncombo;
    h->st.est_overflow_tiles = h->novf;
    h->st.mfma_tiles = h->use_mj ? h->mj_nbig : 0;
    h->st.filter_rechecks = gst[0];
    h->st.filter_hits = gst[1];
    memcpy(&h->st.filter_max_err, &gst[2], 8);
    for (uint32_t i = 0; i < dims && i < 64; i++) h->st.perm[i] = h->G.perm[i];
    h->st.ms_grid = ms_between(h->ev[0], h->ev[1]);
    h->st.ms_estimate = ms_between(h->ev[1], h->ev[2]);
    h->st.ms_count = ms_between(h->ev[2], h->ev[3]);
    h->st.ms_scan = ms_between(h->ev[3], h->ev[4]);
    if (slotted)
        GDSJ_TRY(estimator_error(h->G, h->estPred.as<float>(), d_counts, &h->st.est_mean_rel_err,
                                 &h->st.est_max_rel_err, s, h->estScratch));
    int rc = GDSJ_OK;
    if (!d_ids) {
        rc = tot ? GDSJ_EOVERFLOW : GDSJ_OK;
    } else if (tot > ids_cap) {
        set_error("result has %llu pairs, output buffer holds %llu", (unsigned long long)tot,
                  (unsigned long long)ids_cap);
        rc = GDSJ_EOVERFLOW;
    } else if (queued) {
        GDSJ_TRY(fill_finish(h, d_ids));
    } else {
        GDSJ_TRY(do_fill(h, d_ids, ids_cap));
    }
    // ---- the estimator learns from the counts this join produced
    const bool train = h->p.estimator == GDSJ_EST_LEARNED &&
                       (h->p.train == 1 || (h->p.train == 0 && !h->est.trained));
    if (train && h->G.haveRuns) {
        double tms = 0;
        GDSJ_TRY(estimator_train(h->est, h->G, d_counts, s, &tms));
        h->st.ms_train = tms;
    }
    h->st.ms_total = h->st.ms_grid + h->st.ms_estimate + h->st.ms_count + h->st.ms_scan + h->st.ms_compact +
                     h->st.ms_fill + h->st.ms_sort;
    return rc;
}

int gdsj_rejoin_device(gdsj_handle *h, uint32_t *d_counts, uint64_t *d_offsets, uint32_t *d_ids, uint64_t ids_cap,
                       uint64_t *total, void *stream) {
    if (!h || !total) return GDSJ_EINVAL;
    if (!h->have_grid) {
        set_error("gdsj_rejoin_device needs a preceding gdsj_join_device on this handle");
        return GDSJ_ESTATE;
    }
    if (!d_counts && h->G.n) return GDSJ_EINVAL;
    GDSJ_HIP(hipSetDevice(h->device));
    h->have_counts = h->have_table = false;
    h->slotted = false;
    h->novf = 0;
    h->stream = stream ? (hipStream_t)stream : h->own;
    memset(&h->st, 0, sizeof(h->st));
    GDSJ_HIP(hipEventRecord(h->ev[0], h->stream));
    return join_passes(h, h->G.n, h->G.d, d_counts, d_offsets, d_ids, ids_cap, total);
}

// All-pairs join (the reference's search mode 1, GPU.cu:1285-1480,
// kernel.cu:486-515): counts, offsets and, when d_ids holds them, the rows.
int gdsj_bruteforce_device(gdsj_handle *h, const void *d_data, uint32_t n, uint32_t dims, int dtype,
                           double epsilon, const gdsj_params *pin, uint32_t *d_counts, uint64_t *d_offsets,
                           uint32_t *d_ids, uint64_t ids_cap, uint64_t *total, void *stream) {
    if (!h || (!d_data && n) || (!d_counts && n) || !total) return GDSJ_EINVAL;
    if (dtype != GDSJ_DTYPE_F32 && dtype != GDSJ_DTYPE_F64) {
        set_error("dtype must be GDSJ_DTYPE_F32 or GDSJ_DTYPE_F64");
        return GDSJ_EINVAL;
    }
    if (dims == 0 || dims > (uint32_t)MAXD) {
        set_error("dims=%u must be in 1..%d", dims, MAXD);
        return GDSJ_EINVAL;
    }
    if (!(epsilon > 0) || !std::isfinite(epsilon)) {
        set_error("epsilon must be positive and finite");
        return GDSJ_EINVAL;
    }
    gdsj_params p;
    if (pin) p = *pin; else gdsj_params_default(&p);
    if (p.ilp > 64 || p.ilp_fold > 1) {
        set_error("ILP=%u exceeds 64 or ilp_fold=%u is not 0 or 1", p.ilp, p.ilp_fold);
        return GDSJ_EINVAL;
    }
    double thr;
    if (dtype == GDSJ_DTYPE_F32) {
        const float e = (float)epsilon;
        if ((double)e != epsilon) {
            set_error("epsilon %.17g is not a float value", epsilon);
            return GDSJ_EINVAL;
        }
        thr = (double)gdsj_eps_threshold_f32(e);
    } else {
        thr = gdsj_eps_threshold_f64(epsilon);
    }
    const int mode = p.cpu_sum ? ACC_UNFUSED : acc_mode(p.ilp, p.shortcircuit, p.ilp_fold);
    GDSJ_HIP(hipSetDevice(h->device));
    h->have_counts = h->have_table = false;
    h->bf_ok = false;
    hipStream_t s = stream ? (hipStream_t)stream : h->own;
    const size_t es = dtype == GDSJ_DTYPE_F64 ? 8 : 4;
    const uint64_t ld = ((uint64_t)n + 63) / 64 * 64 + 64;
    GDSJ_TRY(h->bfXs.ensure((size_t)ld * dims * es + 256));
    uint64_t *offs = d_offsets;
    if (!offs) {
        GDSJ_TRY(h->offsets.ensure(((size_t)n + 1) * 8));
        offs = h->offsets.as<uint64_t>();
    }
    GDSJ_TRY(h->scanTmp.ensure(scan_tmp_bytes((uint64_t)n + 1) + 256));
    GDSJ_TRY(bruteforce_soa(d_data, n, dims, dtype, ld, h->bfXs.p, s));
    GDSJ_TRY(bruteforce_pass(h->bfXs.p, ld, n, dims, dtype, mode, p.ilp, thr, PASS_COUNT, d_counts, nullptr, nullptr,
                             s));
    GDSJ_TRY(exclusive_scan_u32_u64(d_counts, offs, n, true, h->scanTmp.p, s));
    uint64_t tot = 0;
    GDSJ_HIP(hipMemcpyAsync(&tot, offs + n, 8, hipMemcpyDeviceToHost, s));
    GDSJ_HIP(hipStreamSynchronize(s));
    *total = tot;
    // what gdsj_bruteforce_rows_host needs to fill the rows later
    h->bf_ok = !d_offsets;
    h->bf_n = n;
    h->bf_dims = dims;
    h->bf_dtype = dtype;
    h->bf_mode = mode;
    h->bf_ilp = p.ilp;
    h->bf_thr = thr;
    h->bf_ld = ld;
    h->bf_total = tot;
    if (!d_ids) return GDSJ_OK;
    if (tot > ids_cap) {
        set_error("result has %llu pairs, output buffer holds %llu", (unsigned long long)tot,
                  (unsigned long long)ids_cap);
        return GDSJ_EOVERFLOW;
    }
    GDSJ_TRY(bruteforce_pass(h->bfXs.p, ld, n, dims, dtype, mode, p.ilp, thr, PASS_FILL, nullptr, offs, d_ids, s));
    GDSJ_HIP(hipStreamSynchronize(s));
    return GDSJ_OK;
}

// Host-buffer form (the CLI's mode 1): counts[n] to host and the total.
int gdsj_bruteforce_host(gdsj_handle *h, const void *data, uint32_t n, uint32_t dims, int dtype, double epsilon,
                         const gdsj_params *p, uint32_t *counts, uint64_t *total) {
    if (!h || (!data && n) || (!counts && n) || !total) return GDSJ_EINVAL;
    GDSJ_HIP(hipSetDevice(h->device));
    const size_t bytes = (size_t)n * dims * (dtype == GDSJ_DTYPE_F64 ? 8 : 4);
    GDSJ_TRY(h->data.ensure(bytes + 256));
    GDSJ_TRY(h->counts.ensure((size_t)n * 4 + 256));
    hipStream_t s = h->own;
    GDSJ_HIP(hipStreamSynchronize(s));
    GDSJ_TRY(staged_copy(h, h->data.p, data, bytes, false));
    GDSJ_TRY(gdsj_bruteforce_device(h, h->data.p, n, dims, dtype, epsilon, p, h->counts.as<uint32_t>(), nullptr,
                                    nullptr, 0, total, s));
    GDSJ_TRY(staged_copy(h, counts, h->counts.p, (size_t)n * 4, true));
    return GDSJ_OK;
}

int gdsj_bruteforce_rows_host(gdsj_handle *h, uint32_t *ids, uint64_t cap) {
    if (!h) return GDSJ_EINVAL;
    if (!h->bf_ok) {
        set_error("gdsj_bruteforce_rows_host needs a preceding gdsj_bruteforce_host on this handle");
        return GDSJ_ESTATE;
    }
    if (cap < h->bf_total || (!ids && h->bf_total)) {
        set_error("buffer holds %llu ids, the table has %llu", (unsigned long long)cap,
                  (unsigned long long)h->bf_total);
        return GDSJ_EOVERFLOW;
    }
    GDSJ_HIP(hipSetDevice(h->device));
    hipStream_t s = h->own;
    GDSJ_TRY(h->ids.ensure((size_t)h->bf_total * 4 + 256));
    GDSJ_TRY(bruteforce_pass(h->bfXs.p, h->bf_ld, h->bf_n, h->bf_dims, h->bf_dtype, h->bf_mode, h->bf_ilp, h->bf_thr,
                             PASS_FILL, nullptr, h->offsets.as<uint64_t>(), h->ids.as<uint32_t>(), s));
    GDSJ_HIP(hipStreamSynchronize(s));
    return staged_copy(h, ids, h->ids.p, (size_t)h->bf_total * 4, true);
}

int gdsj_fill_device(gdsj_handle *h, uint32_t *d_ids, uint64_t ids_cap) {
    if (!h) return GDSJ_EINVAL;
    if (!h->have_counts) {
        set_error("gdsj_fill_device called without a preceding join");
        return GDSJ_ESTATE;
    }
    if (h->total > ids_cap || (!d_ids && h->total)) {
        set_error("output buffer too small: need %llu", (unsigned long long)h->total);
        return GDSJ_EOVERFLOW;
    }
    GDSJ_HIP(hipSetDevice(h->device));
    GDSJ_TRY(do_fill(h, d_ids, ids_cap));
    h->st.ms_total = h->st.ms_grid + h->st.ms_estimate + h->st.ms_count + h->st.ms_scan + h->st.ms_compact +
                     h->st.ms_fill + h->st.ms_sort;
    return GDSJ_OK;
}

// Copies bytes between a host buffer and device memory of h's device
// (d2h: device src -> host dst) through the pinned staging buffers; returns
// when the copy is complete.  Small copies go directly.
static int staged_copy(gdsj_handle *h, void *dst, const void *src, size_t bytes, bool d2h) {
    if (bytes == 0) return GDSJ_OK;
    if (bytes < 2 * STAGE_CHUNK) {
        GDSJ_HIP(hipMemcpy(dst, src, bytes, d2h ? hipMemcpyDeviceToHost : hipMemcpyHostToDevice));
        return GDSJ_OK;
    }
    const int64_t nch = (int64_t)((bytes + STAGE_CHUNK - 1) / STAGE_CHUNK);
    if (!h->cstream) GDSJ_HIP(hipStreamCreateWithFlags(&h->cstream, hipStreamNonBlocking));
    for (int k = 0; k < STAGE_RING; k++) {
        if (!h->cev[k]) GDSJ_HIP(hipEventCreateWithFlags(&h->cev[k], hipEventDisableTiming));
        GDSJ_TRY(h->pin[k].ensure(STAGE_CHUNK));
    }
    const unsigned hc = std::thread::hardware_concurrency();
    int nt = hc >= 2 && hc - 1 < (unsigned)STAGE_THREADS ? (int)hc - 1 : STAGE_THREADS;
    if ((int64_t)nt > nch) nt = (int)nch;
    // per buffer: the last chunk the stream took (issued) and the last chunk
    // the host side finished with it (done); -1: none
    std::atomic<int64_t> issued[STAGE_RING], done[STAGE_RING];
    for (int k = 0; k < STAGE_RING; k++) {
        issued[k].store(-1);
        done[k].store(-1);
    }
    std::atomic<int> failed{0};
    hipStream_t cs = h->cstream;
    auto chunk_len = [&](int64_t c) {
        const size_t off = (size_t)c * STAGE_CHUNK;
        return bytes - off < STAGE_CHUNK ? bytes - off : STAGE_CHUNK;
    };
    auto wait_for = [&](std::atomic<int64_t> &v, int64_t want) {
        while (v.load(std::memory_order_acquire) < want) {
            if (failed.load(std::memory_order_relaxed)) return false;
            std::this_thread::yield();
        }
        return true;
    };
    auto host_side = [&](int t) {
        if (hipSetDevice(h->device) != hipSuccess) {
            failed = 1;
            return;
        }
        for (int64_t c = t; c < nch; c += nt) {
            const int k = (int)(c % STAGE_RING);
            const size_t off = (size_t)c * STAGE_CHUNK, len = chunk_len(c);
            if (d2h) {
                // the stream has taken chunk c into buffer k: wait for its copy
                if (!wait_for(issued[k], c)) return;
                if (hipEventSynchronize(h->cev[k]) != hipSuccess) {
                    failed = 1;
                    return;
                }
                memcpy((char *)dst + off, h->pin[k].p, len);
            } else {
                // buffer k is free once the stream has sent chunk c - STAGE_RING
                if (c >= STAGE_RING) {
                    if (!wait_for(issued[k], c - STAGE_RING)) return;
                    if (hipEventSynchronize(h->cev[k]) != hipSuccess) {
                        failed = 1;
                        return;
                    }
                }
                memcpy(h->pin[k].p, (const char *)src + off, len);
            }
            done[k].store(c, std::memory_order_release);
        }
    };
    std::vector<std::thread> th;
    for (int t = 0; t < nt; t++) th.emplace_back(host_side, t);
    // this thread drives the stream, chunks in order
    for (int64_t c = 0; c < nch && !failed.load(); c++) {
        const int k = (int)(c % STAGE_RING);
        const size_t off = (size_t)c * STAGE_CHUNK, len = chunk_len(c);
        bool ok;
        if (d2h) {
            // buffer k is free once the host side has copied chunk c - STAGE_RING out
            ok = c < STAGE_RING || wait_for(done[k], c - STAGE_RING);
            ok = ok && hipMemcpyAsync(h->pin[k].p, (const char *)src + off, len, hipMemcpyDeviceToHost, cs) ==
                           hipSuccess;
        } else {
            // chunk c is in buffer k once the host side has filled it
            ok = wait_for(done[k], c);
            ok = ok && hipMemcpyAsync((char *)dst + off, h->pin[k].p, len, hipMemcpyHostToDevice, cs) == hipSuccess;
        }
        ok = ok && hipEventRecord(h->cev[k], cs) == hipSuccess;
        if (!ok) {
            failed = 1;
            break;
        }
        issued[k].store(c, std::memory_order_release);
    }
    for (auto &x : th) x.join();
    if (hipStreamSynchronize(cs) != hipSuccess) failed = 1;
    GDSJ_HIP(hipSetDevice(h->device));
    if (failed) {
        set_error("staged %s copy of %zu bytes failed", d2h ? "device-to-host" : "host-to-device", bytes);
        return GDSJ_EDEVICE;
    }
    return GDSJ_OK;
}

int gdsj_join_host(gdsj_handle *h, const void *data, uint32_t n, uint32_t dims, int dtype, double epsilon,
                   const gdsj_params *p, uint32_t *counts, uint64_t *total) {
    if (!h || (!data && n) || (!counts && n) || !total) return GDSJ_EINVAL;
    GDSJ_HIP(hipSetDevice(h->device));
    size_t es = dtype == GDSJ_DTYPE_F64 ? 8 : 4;
    size_t bytes = (size_t)n * dims * es;
    GDSJ_TRY(h->data.ensure(bytes + 256));
    GDSJ_TRY(h->counts.ensure((size_t)n * 4 + 256));
    hipStream_t s = h->own;
    GDSJ_HIP(hipStreamSynchronize(s));  // the previous join no longer reads data
    GDSJ_TRY(staged_copy(h, h->data.p, data, bytes, false));
    uint64_t tot = 0;
    int rc = gdsj_join_device(h, h->data.p, n, dims, dtype, epsilon, p, h->counts.as<uint32_t>(), nullptr,
                              nullptr, 0, &tot, s);
    if (rc != GDSJ_OK && rc != GDSJ_EOVERFLOW) return rc;
    GDSJ_TRY(h->ids.ensure((size_t)tot * 4 + 256));
    GDSJ_TRY(gdsj_fill_device(h, h->ids.as<uint32_t>(), tot));
    GDSJ_TRY(staged_copy(h, counts, h->counts.p, (size_t)n * 4, true));
    *total = tot;
    return GDSJ_OK;
}

int gdsj_result_copy(gdsj_handle *h, uint32_t *ids, uint64_t cap) {
    if (!h) return GDSJ_EINVAL;
    if (!h->have_table) {
        set_error("no table: run gdsj_join_host first");
        return GDSJ_ESTATE;
    }
    if (cap < h->total || (!ids && h->total)) {
        set_error("buffer holds %llu ids, table has %llu", (unsigned long long)cap, (unsigned long long)h->total);
        return GDSJ_EOVERFLOW;
    }
    GDSJ_HIP(hipSetDevice(h->device));
    GDSJ_TRY(staged_copy(h, ids, h->ids.p, (size_t)h->total * 4, true));
    return GDSJ_OK;
}

int gdsj_get_stats(gdsj_handle *h, gdsj_stats *out) {
    if (!h || !out) return GDSJ_EINVAL;
    *out = h->st;
    return GDSJ_OK;
}

// ------------------------------------------------------- learned estimator --
uint32_t gdsj_mlp_param_count(void) { return (uint32_t)MLP_P; }
uint32_t gdsj_mlp_feature_count(void) { return (uint32_t)MLP_F; }

// host-buffer wrappers of the MLP kernels (used to check them against a CPU
// restatement); run on the current device
static int mlp_host_call(const float *params, const float *x, const float *t, uint32_t n, float *y, float *grad,
                         float *loss) {
    if (!params || (!x && n)) return GDSJ_EINVAL;
    hipStream_t s = nullptr;
    DevBuf dp, dx, dt, dy, dpart;
    GDSJ_TRY(dp.ensure(MLP_P * 4));
    GDSJ_TRY(dx.ensure((size_t)n * MLP_F * 4 + 64));
    GDSJ_HIP(hipMemcpy(dp.p, params, MLP_P * 4, hipMemcpyHostToDevice));
    if (n) GDSJ_HIP(hipMemcpy(dx.p, x, (size_t)n * MLP_F * 4, hipMemcpyHostToDevice));
    int rc = GDSJ_OK;
    if (y) {
        rc = dy.ensure((size_t)n * 4 + 64);
        if (!rc) rc = mlp_forward(dp.as<float>(), dx.as<float>(), n, dy.as<float>(), s);
        if (!rc && n) rc = hipMemcpy(y, dy.p, (size_t)n * 4, hipMemcpyDeviceToHost) == hipSuccess ? 0 : GDSJ_EDEVICE;
    } else {
        rc = dt.ensure((size_t)n * 4 + 64);
        if (!rc) rc = dpart.ensure(mlp_grad_scratch_floats(n) * 4 + (MLP_P + 8) * 4);
        if (!rc && n) rc = hipMemcpy(dt.p, t, (size_t)n * 4, hipMemcpyHostToDevice) == hipSuccess ? 0 : GDSJ_EDEVICE;
        float *g = dpart.as<float>() + mlp_grad_scratch_floats(n);
        if (!rc) rc = mlp_grad(dp.as<float>(), dx.as<float>(), dt.as<float>(), n, g, nullptr, dpart.as<float>(), s);
        std::vector<float> hg(MLP_P + 1);
        if (!rc) rc = hipMemcpy(hg.data(), g, (MLP_P + 1) * 4, hipMemcpyDeviceToHost) == hipSuccess ? 0 : GDSJ_EDEVICE;
        if (!rc) {
            memcpy(grad, hg.data(), MLP_P * 4);
            if (loss) *loss = hg[MLP_P];
        }
    }
    DevBuf *b[] = {&dp, &dx, &dt, &dy, &dpart};
    for (DevBuf *q : b) q->release();
    return rc;
}

int gdsj_mlp_forward(const float *params, const float *x, uint32_t n, float *y) {
    if (!y && n) return GDSJ_EINVAL;
    return mlp_host_call(params, x, nullptr, n, y ? y : (float *)params, nullptr, nullptr);
}

int gdsj_mlp_grad(const float *params, const float *x, const float *t, uint32_t n, float *grad, float *loss) {
    if (!grad || (!t && n) || n == 0) return GDSJ_EINVAL;
    return mlp_host_call(params, x, t, n, nullptr, grad, loss);
}

int gdsj_model_get(gdsj_handle *h, float *params, uint32_t cap, double *alpha) {
    if (!h || !params || cap < (uint32_t)MLP_P) return GDSJ_EINVAL;
    if (!h->est.init) {
        set_error("the estimator has not been initialised (no learned join yet)");
        return GDSJ_ESTATE;
    }
    GDSJ_HIP(hipSetDevice(h->device));
    GDSJ_HIP(hipMemcpy(params, h->est.params.p, MLP_P * 4, hipMemcpyDeviceToHost));
    if (alpha) *alpha = h->est.alpha;
    return GDSJ_OK;
}

int gdsj_model_set(gdsj_handle *h, const float *params, uint32_t count, double alpha) {
    if (!h || !params || count != (uint32_t)MLP_P || !(alpha >= 0)) return GDSJ_EINVAL;
    GDSJ_HIP(hipSetDevice(h->device));
    GDSJ_TRY(h->est.params.ensure(MLP_P * 4));
    GDSJ_TRY(h->est.adam_m.ensure(MLP_P * 4));
    GDSJ_TRY(h->est.adam_v.ensure(MLP_P * 4));
    GDSJ_HIP(hipMemcpy(h->est.params.p, params, MLP_P * 4, hipMemcpyHostToDevice));
    GDSJ_HIP(hipMemset(h->est.adam_m.p, 0, MLP_P * 4));
    GDSJ_HIP(hipMemset(h->est.adam_v.p, 0, MLP_P * 4));
    h->est.alpha = alpha;
    h->est.alpha_heavy = alpha;
    h->est.init = h->est.trained = true;
    h->est.step = 0;
    return GDSJ_OK;
}

// ---------------------------------------------------------------- legacy --
// The reference keeps the pending table in a process-global pointer
// (main.cu:359) until copyResultIntoPythonArray copies and frees it
// (main.cu:507-511).  Here the pending table stays in the device's legacy
// handle (a mutex-protected global) and is copied from HBM straight into the
// caller's array, with no host staging copy.
static std::mutex g_legacy_mu;
static gdsj_params g_legacy_params;
static bool g_legacy_params_set = false;
static gdsj_handle *g_legacy_handle[64] = {};
static gdsj_handle *g_pending = nullptr;  // handle holding the pending table
static uint64_t g_pending_n = 0;
static int g_legacy_status = GDSJ_OK;

int gdsj_set_legacy_params(const gdsj_params *p) {
    std::lock_guard<std::mutex> lk(g_legacy_mu);
    if (!p) {
        g_legacy_params_set = false;
        return GDSJ_OK;
    }
    g_legacy_params = *p;
    g_legacy_params_set = true;
    return GDSJ_OK;
}

int gdsj_legacy_status(void) { return g_legacy_status; }

static void legacy_join(const void *data, unsigned int n, double eps, unsigned int ndim, int dtype, int ctx,
                        unsigned int *counts) {
    std::lock_guard<std::mutex> lk(g_legacy_mu);
    if (counts && n) memset(counts, 0, (size_t)n * 4);
    g_pending = nullptr;
    g_pending_n = 0;
    g_legacy_status = GDSJ_OK;
    if (ctx < 0 || ctx >= 64) {
        set_error("gpuContext %d out of range", ctx);
        g_legacy_status = GDSJ_EINVAL;
        return;
    }
    if (!g_legacy_handle[ctx]) {
        int rc = gdsj_create(ctx, &g_legacy_handle[ctx]);
        if (rc) {
            g_legacy_status = rc;
            return;
        }
    }
    gdsj_handle *h = g_legacy_handle[ctx];
    gdsj_params p;
    if (g_legacy_params_set) p = g_legacy_params; else gdsj_params_default(&p);
    uint64_t total = 0;
    int rc = gdsj_join_host(h, data, n, ndim, dtype, eps, &p, counts, &total);
    if (rc) {
        if (counts && n) memset(counts, 0, (size_t)n * 4);
        g_legacy_status = rc;
        return;
    }
    g_pending = h;
    g_pending_n = total;
}

void GDSJoinPy(float *dataset, unsigned int NUMPOINTS, float epsilon, unsigned int NDIM, int gpuContext,
               unsigned int *outNumNeighborsWithinEps) {
    legacy_join(dataset, NUMPOINTS, (double)epsilon, NDIM, GDSJ_DTYPE_F32, gpuContext, outNumNeighborsWithinEps);
}

void GDSJoinPyDouble(double *dataset, unsigned int NUMPOINTS, double epsilon, unsigned int NDIM, int gpuContext,
                     unsigned int *outNumNeighborsWithinEps) {
    legacy_join(dataset, NUMPOINTS, epsilon, NDIM, GDSJ_DTYPE_F64, gpuContext, outNumNeighborsWithinEps);
}

void copyResultIntoPythonArray(unsigned int *outPythonNeighborTable, unsigned int numResults) {
    std::lock_guard<std::mutex> lk(g_legacy_mu);
    if (!g_pending) return;
    uint64_t m = numResults < g_pending_n ? numResults : g_pending_n;
    if (outPythonNeighborTable && m) {
        gdsj_handle *h = g_pending;
        const hipError_t e = hipSetDevice(h->device);
        const int rc = e == hipSuccess ? staged_copy(h, outPythonNeighborTable, h->ids.p, (size_t)m * 4, true)
                                       : GDSJ_EDEVICE;
        if (rc != GDSJ_OK) {
            if (e != hipSuccess) set_error("copyResultIntoPythonArray: %s", hipGetErrorString(e));
            g_legacy_status = rc;
        }
    }
    g_pending = nullptr;
    g_pending_n = 0;
}

}  // extern "C"
