// join_inst.h -- per-DTYPE dispatch of the join kernel variants.
// Included by join_f32.hip / join_f64.hip with JT (DTYPE) and JNAME set.
#include "mjoin_kernels.h"

namespace gdsj {

template <int D, int MODE, int ILPV, bool GV>
static int launch_join_t(const Grid &G, const JoinArgs &a, int pass, hipStream_t s) {
    const uint32_t waves = 256 / GDSJ_WAVE;
    const uint32_t blocks = (a.ntiles + waves - 1) / waves;
    if (blocks == 0) return GDSJ_OK;
    const JT *xs = (const JT *)a.xs;
    if (a.stamp && pass == PASS_SLOT) {
        set_error("unicomp runs with the counting and fill passes only");
        return GDSJ_EINVAL;
    }
    // hit masks with the exact passes over precomputed runs (no unicomp)
    const bool hm = a.hmask != nullptr && !a.stamp && a.runs != nullptr;
    const uint64_t *hmp = a.hmask;
    if (pass == PASS_FILL) {
        if (a.stamp) k_join<JT, D, MODE, ILPV, PASS_FILL, true, false, GV><<<blocks, 256, 0, s>>>(xs, a.ids, a, hmp);
        else if (hm) k_join<JT, D, MODE, ILPV, PASS_FILL, false, true><<<blocks, 256, 0, s>>>(xs, a.ids, a, hmp);
        else k_join<JT, D, MODE, ILPV, PASS_FILL, false, false, GV><<<blocks, 256, 0, s>>>(xs, a.ids, a, hmp);
    } else if (pass == PASS_SLOT) {
        k_join<JT, D, MODE, ILPV, PASS_SLOT, false, false, GV><<<blocks, 256, 0, s>>>(xs, a.ids, a, hmp);
    } else {
        if (a.stamp) k_join<JT, D, MODE, ILPV, PASS_COUNT, true, false, GV><<<blocks, 256, 0, s>>>(xs, a.ids, a, hmp);
        else if (hm) k_join<JT, D, MODE, ILPV, PASS_COUNT, false, true, GV><<<blocks, 256, 0, s>>>(xs, a.ids, a, hmp);
        else k_join<JT, D, MODE, ILPV, PASS_COUNT, false, false, GV><<<blocks, 256, 0, s>>>(xs, a.ids, a, hmp);
    }
    GDSJ_HIP(hipGetLastError());
    return GDSJ_OK;
}

// SORT (JoinArgs::sortMode): the generic kernels, counting, fill and
// slotted passes (no unicomp, no hit masks)
template <int MODE>
static int launch_join_sort(const JoinArgs &a, int pass, hipStream_t s) {
    const uint32_t blocks = (a.ntiles + 3) / 4;
    if (blocks == 0) return GDSJ_OK;
    if (a.stamp || a.hmask) {
        set_error("SORT runs without unicomp and without hit masks");
        return GDSJ_EINVAL;
    }
    const JT *xs = (const JT *)a.xs;
    if (pass == PASS_FILL)
        k_join<JT, 0, MODE, 0, PASS_FILL, false, false, false, true><<<blocks, 256, 0, s>>>(xs, a.ids, a, nullptr);
    else if (pass == PASS_SLOT)
        k_join<JT, 0, MODE, 0, PASS_SLOT, false, false, false, true><<<blocks, 256, 0, s>>>(xs, a.ids, a, nullptr);
    else
        k_join<JT, 0, MODE, 0, PASS_COUNT, false, false, false, true><<<blocks, 256, 0, s>>>(xs, a.ids, a, nullptr);
    GDSJ_HIP(hipGetLastError());
    return GDSJ_OK;
}

// the Gram-form filter variants exist for FP32 and the dims JOIN_GV_DIM accepts
template <int D, int MODE, int ILPV>
static int launch_join_gv(const Grid &G, const JoinArgs &a, int pass, hipStream_t s) {
    if constexpr (JOIN_GV_DIM(D)) {
        if (a.gv) return launch_join_t<D, MODE, ILPV, true>(G, a, pass, s);
    }
    return launch_join_t<D, MODE, ILPV, false>(G, a, pass, s);
}

// the matrix-core join (every dim indexed, 3 to 7 dims) over a.tileList
template <int D>
static int launch_mjoin_t(const JoinArgs &a, int mode, int pass, hipStream_t s) {
    const uint32_t blocks = (a.ntiles + 3) / 4;
    if (blocks == 0) return GDSJ_OK;
    const JT *xs = (const JT *)a.xs;
    if (pass == PASS_FILL) k_mjoin<JT, D, PASS_FILL><<<blocks, 256, 0, s>>>(xs, a.ids, a, mode);
    else if (pass == PASS_SLOT) k_mjoin<JT, D, PASS_SLOT><<<blocks, 256, 0, s>>>(xs, a.ids, a, mode);
    else k_mjoin<JT, D, PASS_COUNT><<<blocks, 256, 0, s>>>(xs, a.ids, a, mode);
    GDSJ_HIP(hipGetLastError());
    return GDSJ_OK;
}

static int launch_mjoin(const JoinArgs &a, int mode, int pass, hipStream_t s) {
    switch (a.d) {
        case 3: return launch_mjoin_t<3>(a, mode, pass, s);
        case 4: return launch_mjoin_t<4>(a, mode, pass, s);
        case 5: return launch_mjoin_t<5>(a, mode, pass, s);
        case 6: return launch_mjoin_t<6>(a, mode, pass, s);
        case 7: return launch_mjoin_t<7>(a, mode, pass, s);
        default:
            set_error("the matrix-core join takes 3 to 7 dims (dims=%u)", a.d);
            return GDSJ_EINVAL;
    }
}

#define GDSJ_DIMCASE(DV)                                                              \
    case DV:                                                                          \
        if (mode == ACC_SEQ) return launch_join_gv<DV, ACC_SEQ, 1>(G, a, pass, s);    \
        if (mode == ACC_ILP_SC && a.ilp == 8) return launch_join_gv<DV, ACC_ILP_SC, 8>(G, a, pass, s); \
        break;

int JNAME(const Grid &G, const JoinArgs &a, int mode, int pass, hipStream_t s) {
    if (a.mjList) {
        // the listed tiles on the matrix cores, the others on the plain kernels
        if (a.stamp || a.sortMode || !a.runs || a.tileList || a.mjN + a.vjN != a.ntiles) {
            set_error("the matrix-core join runs over precomputed runs, without unicomp and SORT");
            return GDSJ_EINVAL;
        }
        JoinArgs m = a;
        m.tileList = a.mjList;
        m.ntiles = a.mjN;
        m.mjList = nullptr;
        m.hmask = nullptr;
        hipStream_t ms = s;
        if (a.mjStream) {
            ms = (hipStream_t)a.mjStream;
            GDSJ_HIP(hipEventRecord((hipEvent_t)a.mjFork, s));
            GDSJ_HIP(hipStreamWaitEvent(ms, (hipEvent_t)a.mjFork, 0));
        }
        GDSJ_TRY(launch_mjoin(m, mode, pass, ms));
        if (a.mjStream) GDSJ_HIP(hipEventRecord((hipEvent_t)a.mjJoin, ms));
        JoinArgs v = a;
        v.tileList = a.vjList;
        v.ntiles = a.vjN;
        v.mjList = nullptr;
        GDSJ_TRY(JNAME(G, v, mode, pass, s));
        if (a.mjStream) GDSJ_HIP(hipStreamWaitEvent(s, (hipEvent_t)a.mjJoin, 0));
        return GDSJ_OK;
    }
    if (a.sortMode) {
        if (mode == ACC_SEQ) return launch_join_sort<ACC_SEQ>(a, pass, s);
        if (mode == ACC_ILP_SC) return launch_join_sort<ACC_ILP_SC>(a, pass, s);
        return launch_join_sort<ACC_ILP_NOSC>(a, pass, s);
    }
    switch (a.d) {
        JOIN_DIMS
        default:
            break;
    }
    if (mode == ACC_SEQ) return launch_join_t<0, ACC_SEQ, 0, false>(G, a, pass, s);
    if (mode == ACC_ILP_SC) return launch_join_t<0, ACC_ILP_SC, 0, false>(G, a, pass, s);
    return launch_join_t<0, ACC_ILP_NOSC, 0, false>(G, a, pass, s);
}

bool JNAME_SPECIALISED(uint32_t d, int mode, uint32_t ilp) {
    switch (d) {
#define GDSJ_DIMQ(DV) case DV: return mode == ACC_SEQ || (mode == ACC_ILP_SC && ilp == 8);
        JOIN_DIMQ
#undef GDSJ_DIMQ
        default:
            return false;
    }
}

}  // namespace gdsj
