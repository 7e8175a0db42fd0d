// join_inst.h -- per-DTYPE dispatch of the join kernel variants.
// Included by join_f32.hip / join_f64.hip with JT (DTYPE) and JNAME set.
#include "join_kernels.h"

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

#define GDSJ_DIMCASE(DV)                                                              \
    case DV:                                                                          \
        if (mode == ACC_SEQ) return launch_join_gv<DV, ACC_SEQ, 1>(G, a, pass, s);    \
        if (mode == ACC_ILP_SC && a.ilp == 8) return launch_join_gv<DV, ACC_ILP_SC, 8>(G, a, pass, s); \
        break;

int JNAME(const Grid &G, const JoinArgs &a, int mode, int pass, hipStream_t s) {
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
