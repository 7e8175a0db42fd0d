// batch_kernels.h -- persistent work-queue kernels of batched joins
// (instantiated per DTYPE in batch_f32.hip / batch_f64.hip).
//
// A batch is many independent self-joins (uneven n, dims, epsilon), each
// with its own grid and tiles.  Their tiles are numbered globally
// (instance i owns tiles tileBase[i] .. tileBase[i+1]-1) and listed in the
// order the queue hands them out.  The kernels are launched with about as
// many waves as the GPU holds at once; each wave takes the next item with
// one atomic on the queue head and runs the same per-tile body as the
// single-join kernels, specialised by the instance's dims at run time, until
// the list is exhausted.  Every item is taken exactly once, so the results
// are those of the per-instance joins.
#pragma once
#include "compact.h"
#include "join_kernels.h"

namespace gdsj {

struct BatchLaunch {
    const uint32_t *tileBase;   // [ninst + 1] first global tile of each instance
    uint32_t ninst;
    const uint32_t *order;      // [nitems] global tile ids in queue order (nullable: identity)
    uint32_t nitems;            // number of items, unless
    const uint32_t *nitemsDev;  //   this (device) holds it
    uint32_t *queue;            // queue head, zero before the launch
};

// Takes the next item: false once the queue is exhausted.  The instance is
// the last i with tileBase[i] <= g (empty instances share their successor's
// base and are skipped by the search).
// the (instance, tile) of queue item it
__device__ __forceinline__ void batch_item(const BatchLaunch &L, uint32_t it, uint32_t &inst, uint32_t &tile) {
    const uint32_t g = L.order ? uniform(L.order[it]) : it;
    uint32_t lo = 0, hi = L.ninst;
    while (hi - lo > 1) {
        const uint32_t mid = (lo + hi) >> 1;
        if (uniform(L.tileBase[mid]) <= g) lo = mid; else hi = mid;
    }
    inst = lo;
    tile = g - uniform(L.tileBase[lo]);
}

__device__ __forceinline__ bool batch_next(const BatchLaunch &L, uint32_t nitems, uint32_t &inst,
                                           uint32_t &tile) {
    uint32_t it = 0;
    if (lane_id() == 0) it = atomicAdd(L.queue, 1u);
    it = uniform(it);
    if (it >= nitems) return false;
    const uint32_t g = L.order ? uniform(L.order[it]) : it;
    uint32_t lo = 0, hi = L.ninst;
    while (hi - lo > 1) {
        const uint32_t mid = (lo + hi) >> 1;
        if (uniform(L.tileBase[mid]) <= g) lo = mid; else hi = mid;
    }
    inst = lo;
    tile = g - uniform(L.tileBase[lo]);
    return true;
}

// Instance arguments never change during a launch: read through the
// constant address space they are scalar loads, kept in SGPRs.
typedef const __attribute__((address_space(4))) JoinArgs ConstJoinArgs;
__device__ __forceinline__ const JoinArgs &inst_args(const JoinArgs *inst, uint32_t i) {
    ConstJoinArgs *c = (ConstJoinArgs *)inst;
    return *(const JoinArgs *)(c + i);
}

__device__ __forceinline__ uint32_t batch_items(const BatchLaunch &L) {
    return L.nitemsDev ? uniform(*L.nitemsDev) : L.nitems;
}

// Candidate chunks held in VGPRs: BUDGET coordinate words per lane; query
// coordinates are read from memory (no register copy of the tile's queries).  Inlined
// side by side, the dims' bodies share one register allocation, larger than
// any of them needs alone: budget 24 keeps the kernel near 110 VGPRs (4 waves
// per SIMD), 48 is the single-join default (more reuse, 3 waves per SIMD).
template <typename T, int D, int BUDGET>
struct BatchCMax {
    static constexpr int raw = BUDGET / (D * (int)(sizeof(T) / 4));
    static constexpr int value = raw < 1 ? 1 : (raw > 8 ? 8 : raw);
};

// SPEC: specialised bodies for dims 2, 4, 8, 16 (MODE/ILPV fixed; they read
// candidates from the instance's row-major copy, JoinArgs::xa), generic
// body for the others; !SPEC: generic body only (runtime ILP).
template <typename T, int MODE, int ILPV, int PASS, bool SPEC, int BUDGET>
__global__ __launch_bounds__(256) void k_batch_join(const JoinArgs *__restrict__ inst, const BatchLaunch L) {
    const uint32_t nitems = batch_items(L);
    uint32_t i, tile;
    while (batch_next(L, nitems, i, tile)) {
        const JoinArgs &a = inst_args(inst, i);
        const T *xs = (const T *)a.xs;
        if (SPEC) {
            switch (a.d) {
                case 2:
                    join_tile<T, 2, MODE, ILPV, PASS, BatchCMax<T, 2, BUDGET>::value, false, 0, false, false, false, false,
                              false, true>(xs, a.ids, a, tile);
                    continue;
                case 4:
                    join_tile<T, 4, MODE, ILPV, PASS, BatchCMax<T, 4, BUDGET>::value, false, 0, false, false, false, false,
                              false, true>(xs, a.ids, a, tile);
                    continue;
                case 8:
                    join_tile<T, 8, MODE, ILPV, PASS, BatchCMax<T, 8, BUDGET>::value, false, 0, false, false, false, false,
                              false, true>(xs, a.ids, a, tile);
                    continue;
                case 16:
                    join_tile<T, 16, MODE, ILPV, PASS, BatchCMax<T, 16, BUDGET>::value, false, 0, false, false, false, false,
                              false, true>(xs, a.ids, a, tile);
                    continue;
                default:
                    break;
            }
        }
        join_tile<T, 0, MODE, 0, PASS, 0, false, 0, false, false>(xs, a.ids, a, tile);
    }
}

}  // namespace gdsj
