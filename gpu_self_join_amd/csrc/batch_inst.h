// batch_inst.h -- per-DTYPE launch of the persistent batch kernels.
// Included by batch_f32.hip / batch_f64.hip with BT (DTYPE) and BNAME set.
#include "batch_kernels.h"

namespace gdsj {

// Persistent grid: as many workgroups as the GPU holds at once (occupancy
// of this kernel at this block size times the CU count), cached per size.
// launch == false only fills the cache (done before a graph capture).
template <int MODE, int ILPV, int PASS, bool SPEC, int BUDGET>
static int batch_launch_t(const JoinArgs *inst, const BatchLaunch &L, uint32_t threads, uint32_t *blocksOut,
                          hipStream_t s, bool launch) {
    static uint32_t cache[5] = {0, 0, 0, 0, 0};
    const uint32_t slot = threads / GDSJ_WAVE;
    if (slot == 0 || slot > 4 || threads % GDSJ_WAVE) {
        set_error("batch kernels take 64, 128, 192 or 256 threads per block");
        return GDSJ_EINVAL;
    }
    if (!cache[slot]) {
        int nb = 0, dev = 0, cus = 0;
        GDSJ_HIP(hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, k_batch_join<BT, MODE, ILPV, PASS, SPEC, BUDGET>,
                                                              (int)threads, 0));
        GDSJ_HIP(hipGetDevice(&dev));
        GDSJ_HIP(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev));
        cache[slot] = (uint32_t)(nb > 0 ? nb : 1) * (uint32_t)(cus > 0 ? cus : 1);
    }
    if (blocksOut) *blocksOut = cache[slot];
    if (!launch) return GDSJ_OK;
    k_batch_join<BT, MODE, ILPV, PASS, SPEC, BUDGET><<<cache[slot], threads, 0, s>>>(inst, L);
    GDSJ_HIP(hipGetLastError());
    return GDSJ_OK;
}

template <int PASS, int BUDGET>
static int batch_pick(const JoinArgs *inst, const BatchLaunch &L, int mode, uint32_t ilp, uint32_t threads,
                      uint32_t *blocks, hipStream_t s, bool launch) {
    if (mode == ACC_SEQ) return batch_launch_t<ACC_SEQ, 1, PASS, true, BUDGET>(inst, L, threads, blocks, s, launch);
    if (mode == ACC_ILP_SC && ilp == 8)
        return batch_launch_t<ACC_ILP_SC, 8, PASS, true, BUDGET>(inst, L, threads, blocks, s, launch);
    if (mode == ACC_ILP_SC) return batch_launch_t<ACC_ILP_SC, 0, PASS, false, 0>(inst, L, threads, blocks, s, launch);
    return batch_launch_t<ACC_ILP_NOSC, 0, PASS, false, 0>(inst, L, threads, blocks, s, launch);
}

template <int PASS>
static int batch_budget(const JoinArgs *inst, const BatchLaunch &L, int mode, uint32_t ilp, uint32_t budget,
                        uint32_t threads, uint32_t *blocks, hipStream_t s, bool launch) {
    if (budget <= 24) return batch_pick<PASS, 24>(inst, L, mode, ilp, threads, blocks, s, launch);
    return batch_pick<PASS, 48>(inst, L, mode, ilp, threads, blocks, s, launch);
}

int BNAME(const JoinArgs *inst, const BatchLaunch &L, int mode, uint32_t ilp, uint32_t budget, int pass,
          uint32_t threads, uint32_t *blocks, hipStream_t s, bool launch) {
    if (pass == PASS_FILL) return batch_budget<PASS_FILL>(inst, L, mode, ilp, budget, threads, blocks, s, launch);
    if (pass == PASS_SLOT) return batch_budget<PASS_SLOT>(inst, L, mode, ilp, budget, threads, blocks, s, launch);
    return batch_budget<PASS_COUNT>(inst, L, mode, ilp, budget, threads, blocks, s, launch);
}

}  // namespace gdsj
