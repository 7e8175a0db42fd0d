// estimator.h -- result-size estimators for the single-pass (slotted) join.
//
// Learned: a 16-32-32-1 ReLU MLP predicts log2(1 + hits) of every work
// tile from cell-occupancy and neighbor-cell features; it is trained on the
// per-tile counts the join itself produces (forward, backward and Adam are
// the kernels below).  Sampling: the reference's estimator
// (kernelNDGridIndexBatchEstimator kernel.cu:522-694, callGPUBatchEst
// GPU.cu:57-268) -- exact counts of every 66th query, scaled by 66.
// Both only size per-tile result slots; a tile whose slot overflows is
// completed by the exact fill pass, so the result never depends on them.
#pragma once
#include "common.h"
#include "grid.h"

namespace gdsj {

constexpr int MLP_F = 16;  // features per tile
constexpr int MLP_H = 32;  // hidden width (two hidden layers)
constexpr int MLP_OFF_W1 = 0;
constexpr int MLP_OFF_B1 = MLP_OFF_W1 + MLP_H * MLP_F;
constexpr int MLP_OFF_W2 = MLP_OFF_B1 + MLP_H;
constexpr int MLP_OFF_B2 = MLP_OFF_W2 + MLP_H * MLP_H;
constexpr int MLP_OFF_W3 = MLP_OFF_B2 + MLP_H;
constexpr int MLP_OFF_B3 = MLP_OFF_W3 + MLP_H;
constexpr int MLP_P = MLP_OFF_B3 + 1;  // 1633 parameters

struct Estimator {
    DevBuf params, adam_m, adam_v, partial, feat, target, pred, sample, work;
    bool init = false;     // parameters initialised
    bool trained = false;  // trained on at least one join
    uint32_t step = 0;     // Adam step
    double alpha = 0.25;   // slot headroom: cap = pred * (1 + alpha) + beta
    double alpha_heavy = 0.25;  // headroom of tiles with many candidate pairs
    void release();
};

// features[t][MLP_F] of every tile (needs the per-cell run lists)
int tile_features(const Grid &G, float *feat, hipStream_t s);
// target[t] = log2(1 + sum of counts of tile t's queries)
int tile_targets(const Grid &G, const uint32_t *counts, float *target, hipStream_t s);

// y[i] = MLP(x[i]) for i < n (device pointers)
int mlp_forward(const float *params, const float *x, uint32_t n, float *y, hipStream_t s);
// grad[MLP_P] = d/dparams of mean_i (MLP(x[i]) - t[i])^2 ; *loss = that mean
// (partial: scratch of mlp_grad_scratch_floats(n) floats)
size_t mlp_grad_scratch_floats(uint32_t n);
int mlp_grad(const float *params, const float *x, const float *t, uint32_t n, float *grad, float *loss,
             float *partial, hipStream_t s);
// one Adam step on params from grad
int mlp_adam(float *params, const float *grad, float *m, float *v, uint32_t step, float lr, hipStream_t s);
// deterministic He initialisation (host-generated, copied to params)
int mlp_init(float *params, uint64_t seed, hipStream_t s);

// Trains the estimator on (features, targets) of the current grid's tiles.
int estimator_train(Estimator &E, const Grid &G, const uint32_t *counts, hipStream_t s, double *ms);
// Trains on nt rows of features feat[nt][MLP_F] and targets target[nt] (device).
int estimator_train_rows(Estimator &E, const float *feat, const float *target, uint32_t nt, hipStream_t s,
                         double *ms);

// Per-tile slot capacities from the learned model; *predicted = sum of
// predicted hits.  capOut[t] >= number of queries of t, <= queries x candidates.
int estimator_capacity_learned(Estimator &E, const Grid &G, uint32_t *capOut, uint64_t *predicted,
                               hipStream_t s);
// The reference's sampling estimator restated, spread uniformly per query.
// scratch: n + 4 * (number of sampled queries) words, kept by the caller.
int estimator_capacity_sampling(const Grid &G, const void *joinArgs, int mode, uint32_t *countsScratch,
                                uint32_t *capOut, uint64_t *predicted, float *predLog2Out, hipStream_t s,
                                DevBuf &scratch);

// mean and max of |pred_t - y_t| / max(y_t, 1) over tiles (pred from capacities' model output)
int estimator_error(const Grid &G, const float *predLog2, const uint32_t *counts, double *meanRel,
                    double *maxRel, hipStream_t s, DevBuf &scratch);

}  // namespace gdsj
