// batch.h -- host entry points of the persistent batch kernels.
#pragma once
#include "batch_kernels.h"

namespace gdsj {

// One pass (PASS_COUNT / PASS_FILL / PASS_SLOT) over the queue's items;
// *blocks receives the persistent grid size.  launch == false only sizes the
// grid (occupancy queries are made before a graph capture, not inside it).
// budget: coordinate words per lane of candidate chunks in VGPRs (24 or 48).
int batch_pass_f32(const JoinArgs *inst, const BatchLaunch &L, int mode, uint32_t ilp, uint32_t budget, int pass,
                   uint32_t threads, uint32_t *blocks, hipStream_t s, bool launch = true);
int batch_pass_f64(const JoinArgs *inst, const BatchLaunch &L, int mode, uint32_t ilp, uint32_t budget, int pass,
                   uint32_t threads, uint32_t *blocks, hipStream_t s, bool launch = true);

}  // namespace gdsj
