// compact.h -- moves one work tile's slot entries into CSR rows.
//
// After the slotted pass, a tile whose slot did not overflow holds its hits
// as (neighbor id, query index) entries in candidate order.  One wave per
// tile re-reads them 64 at a time; entries of the same query are matched by
// ballots over the 7 bits of the query index, and each entry's rank among
// them gives its position behind the query's running row cursor (in LDS), so
// every row keeps the candidate order of the exact fill pass.
#pragma once
#include "join.h"

namespace gdsj {

// tiles hold at most COMPACT_TQ queries (64 for the plain join, up to 128
// for the Gram filter)
constexpr uint32_t COMPACT_TQ = 128;

__device__ __forceinline__ void compact_tile(const JoinArgs &a, const uint32_t tile) {
    __shared__ unsigned long long scur[256 / GDSJ_WAVE][COMPACT_TQ];
    __shared__ uint32_t scnt[256 / GDSJ_WAVE][COMPACT_TQ];
    const uint32_t lane = lane_id();
    const uint32_t w = threadIdx.x / GDSJ_WAVE;
    // the tile's loads are issued together where they do not depend on each
    // other: tiles average a few queries on some workloads, so their latency,
    // not their bytes, sets the time
    const uint32_t used = uniform(a.slotUsed[tile]);
    const uint32_t cap = uniform(a.slotCap[tile]);
    const uint4 tl = a.tiles[tile];
    const uint64_t sbase = uniform64(a.slotOff[tile]);
    if (used > cap || used > SLOT_POS_LIMIT) return;  // overflowed: completed by the fill pass
    const uint32_t q0 = uniform(tl.y), nq = uniform(tl.z - tl.y);
    const uint64_t lt = lane == 0 ? 0ull : (~0ull >> (64 - lane));
    const uint32_t qbits = nq > 1 ? 32u - (uint32_t)__builtin_clz(nq - 1u) : 0u;
    // first batch of entries, loaded beside the rows' offsets
    bool vn = lane < used;
    uint32_t qn = vn ? (uint32_t)a.slotQ[sbase + lane] : 0u;
    uint32_t idn = vn ? a.slotIds[sbase + lane] : 0u;
    for (uint32_t q = lane; q < nq; q += GDSJ_WAVE) {
        scur[w][q] = a.offsets[a.ids[q0 + q]];
        scnt[w][q] = 0;
    }
    __builtin_amdgcn_wave_barrier();
    bool overflow = false;
    for (uint32_t e = 0; e < used; e += GDSJ_WAVE) {
        const bool valid = vn;
        const uint32_t qi = qn, id = idn;
        // the next batch is loaded while this one is placed
        vn = e + GDSJ_WAVE + lane < used;
        if (vn) {
            qn = (uint32_t)a.slotQ[sbase + e + GDSJ_WAVE + lane];
            idn = a.slotIds[sbase + e + GDSJ_WAVE + lane];
        }
        uint64_t peers = ballot(valid);
        // the query indices are below nq: only their low qbits bits differ
        for (uint32_t b = 0; b < qbits; b++) {
            const bool bit = (qi >> b) & 1u;
            const uint64_t mb = ballot(bit);
            peers &= bit ? mb : ~mb;
        }
        const uint64_t below = peers & lt;
        if (valid) {
            const uint64_t pos = scur[w][qi] + popc64(below);
            if (pos < a.cap) a.out[pos] = id; else overflow = true;
            if (below == 0) scnt[w][qi] = popc64(peers);
        }
        __builtin_amdgcn_wave_barrier();
        for (uint32_t q = lane; q < nq; q += GDSJ_WAVE) {
            scur[w][q] += scnt[w][q];
            scnt[w][q] = 0;
        }
        __builtin_amdgcn_wave_barrier();
    }
    if (overflow) atomicOr(a.status, 1u);
}

}  // namespace gdsj
