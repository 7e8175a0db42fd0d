/*
 * cpu_baselines.c -- the reference's two CPU self-join paths, restated, for
 * timing beside the GPU (bench.py's CPU baseline leg).
 *
 * TEST / BENCHMARK INFRASTRUCTURE ONLY: nothing in the product links this.
 *
 * Both use the reference CPU paths' arithmetic (host code built without FMA
 * contraction): run += (q_j - c_j) * (q_j - c_j) in DTYPE, unfused, dims in
 * order, then sqrt(run) <= eps (tree_index.cpp:209-220, main.cu:953-978).
 * That is NOT the GPU path's contract (the GPU kernels fuse), so pairs at
 * the eps boundary can differ; the counts are reported, not compared.
 *
 * cpu_bruteforce_*: every query against every point (main.cu:953-978),
 *   on nthreads OpenMP threads (1 = the reference's sequential loop).
 * cpu_rtree_*: the sequential R-tree search (tree_index.cpp:35-160):
 *   points inserted one at a time as degenerate boxes into an R-tree with
 *   at most 8 and at least 4 entries per node, Guttman's quadratic split,
 *   the volume measure of a box being the volume of its circumscribed sphere
 *   (the reference builds its tree with that option, RTree.h:31); each query
 *   searches the box [q - eps, q + eps] (inclusive overlap,
 *   tree_index.cpp:169-176) and filters the candidates as above.
 * This R-tree is written from Guttman's published algorithm (A. Guttman,
 * "R-trees: a dynamic index structure for spatial searching", SIGMOD 1984).
 */
#include <math.h>
#include <omp.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>
#include <time.h>

#define RT_MAX 8
#define RT_MIN 4

static double now_s(void) {
    struct timespec ts;
    clock_gettime(CLOCK_MONOTONIC, &ts);
    return ts.tv_sec + 1e-9 * ts.tv_nsec;
}

/* unit-sphere volume in d dims (the sphere-volume measure's constant) */
static double unit_sphere_volume(uint32_t d) {
    return pow(M_PI, d / 2.0) / tgamma(d / 2.0 + 1.0);
}

#define DEFINE_BASELINES(T, SFX, SQRT)                                                        \
                                                                                              \
static inline int SFX(within)(const T *q, const T *c, uint32_t d, T eps) {                    \
    T run = 0;                                                                                \
    for (uint32_t j = 0; j < d; j++) {                                                        \
        T t = q[j] - c[j];                                                                    \
        T sq = t * t;                                                                         \
        run = run + sq;                                                                       \
    }                                                                                         \
    return SQRT(run) <= eps;                                                                  \
}                                                                                             \
                                                                                              \
int SFX(cpu_bruteforce)(const T *data, uint32_t n, uint32_t d, T eps, const uint32_t *qids,   \
                        uint32_t nq, int nthreads, uint32_t *counts, double *seconds) {      \
    double t0 = now_s();                                                                      \
    if (nthreads < 1) nthreads = 1;                                                           \
    _Pragma("omp parallel for schedule(dynamic, 8) num_threads(nthreads)")                   \
    for (int64_t s = 0; s < (int64_t)nq; s++) {                                               \
        const T *q = data + (size_t)qids[s] * d;                                              \
        uint32_t cnt = 0;                                                                     \
        for (uint32_t j = 0; j < n; j++) cnt += (uint32_t)SFX(within)(q, data + (size_t)j * d, d, eps); \
        counts[s] = cnt;                                                                      \
    }                                                                                         \
    *seconds = now_s() - t0;                                                                  \
    return 0;                                                                                 \
}                                                                                             \
                                                                                              \
/* node pool: node i has count entries; entry e's box is bmin/bmax at      */               \
/* ((size_t)i * RT_MAX + e) * d; child = node index (inner) or point id    */               \
typedef struct {                                                                              \
    uint32_t d, nnodes, cap, root;                                                            \
    int32_t *level, *count;                                                                   \
    uint32_t *child;                                                                          \
    T *bmin, *bmax;                                                                           \
    float *sph; /* unit sphere volume constant (float, like the reference's ELEMTYPEREAL) */ \
    float sphk;                                                                               \
} SFX(rtree);                                                                                 \
                                                                                              \
static uint32_t SFX(rt_new)(SFX(rtree) *t, int32_t level) {                                   \
    if (t->nnodes == t->cap) {                                                                \
        t->cap = t->cap ? t->cap * 2 : 1024;                                                  \
        t->level = (int32_t *)realloc(t->level, sizeof(int32_t) * t->cap);                    \
        t->count = (int32_t *)realloc(t->count, sizeof(int32_t) * t->cap);                    \
        t->child = (uint32_t *)realloc(t->child, sizeof(uint32_t) * t->cap * RT_MAX);         \
        t->bmin = (T *)realloc(t->bmin, sizeof(T) * (size_t)t->cap * RT_MAX * t->d);          \
        t->bmax = (T *)realloc(t->bmax, sizeof(T) * (size_t)t->cap * RT_MAX * t->d);          \
    }                                                                                         \
    uint32_t i = t->nnodes++;                                                                 \
    t->level[i] = level;                                                                      \
    t->count[i] = 0;                                                                          \
    return i;                                                                                 \
}                                                                                             \
                                                                                              \
/* volume of the sphere around a box: (half diagonal)^d * unit volume */                      \
static float SFX(rt_volume)(const SFX(rtree) *t, const T *lo, const T *hi) {                  \
    float r2 = 0;                                                                             \
    for (uint32_t j = 0; j < t->d; j++) {                                                     \
        float h = ((float)hi[j] - (float)lo[j]) * 0.5f;                                       \
        r2 += h * h;                                                                          \
    }                                                                                         \
    float r = sqrtf(r2);                                                                      \
    return powf(r, (float)t->d) * t->sphk;                                                    \
}                                                                                             \
                                                                                              \
static void SFX(rt_union)(uint32_t d, const T *alo, const T *ahi, const T *blo, const T *bhi, \
                          T *lo, T *hi) {                                                     \
    for (uint32_t j = 0; j < d; j++) {                                                        \
        lo[j] = alo[j] < blo[j] ? alo[j] : blo[j];                                            \
        hi[j] = ahi[j] > bhi[j] ? ahi[j] : bhi[j];                                            \
    }                                                                                         \
}                                                                                             \
                                                                                              \
static void SFX(rt_node_box)(const SFX(rtree) *t, uint32_t nd, T *lo, T *hi) {                \
    const uint32_t d = t->d;                                                                  \
    memcpy(lo, t->bmin + (size_t)nd * RT_MAX * d, sizeof(T) * d);                             \
    memcpy(hi, t->bmax + (size_t)nd * RT_MAX * d, sizeof(T) * d);                             \
    for (int32_t e = 1; e < t->count[nd]; e++) {                                              \
        const T *elo = t->bmin + ((size_t)nd * RT_MAX + e) * d;                               \
        const T *ehi = t->bmax + ((size_t)nd * RT_MAX + e) * d;                               \
        SFX(rt_union)(d, lo, hi, elo, ehi, lo, hi);                                           \
    }                                                                                         \
}                                                                                             \
                                                                                              \
static void SFX(rt_put)(SFX(rtree) *t, uint32_t nd, const T *lo, const T *hi, uint32_t ch) {  \
    const uint32_t d = t->d;                                                                  \
    int32_t e = t->count[nd]++;                                                               \
    memcpy(t->bmin + ((size_t)nd * RT_MAX + e) * d, lo, sizeof(T) * d);                       \
    memcpy(t->bmax + ((size_t)nd * RT_MAX + e) * d, hi, sizeof(T) * d);                       \
    t->child[(size_t)nd * RT_MAX + e] = ch;                                                   \
}                                                                                             \
                                                                                              \
/* Quadratic split of node nd's RT_MAX entries plus (lo, hi, ch) into nd and */             \
/* a new node (returned).                                                    */             \
static uint32_t SFX(rt_split)(SFX(rtree) *t, uint32_t nd, const T *lo, const T *hi, uint32_t ch) { \
    const uint32_t d = t->d;                                                                  \
    enum { M = RT_MAX + 1 };                                                                  \
    T *elo = (T *)malloc(sizeof(T) * M * d), *ehi = (T *)malloc(sizeof(T) * M * d);           \
    T *glo = (T *)malloc(sizeof(T) * 2 * d), *ghi = (T *)malloc(sizeof(T) * 2 * d);           \
    T *ulo = (T *)malloc(sizeof(T) * d), *uhi = (T *)malloc(sizeof(T) * d);                   \
    uint32_t ech[M];                                                                          \
    float vol[M];                                                                             \
    int grp[M];                                                                               \
    for (int e = 0; e < RT_MAX; e++) {                                                        \
        memcpy(elo + (size_t)e * d, t->bmin + ((size_t)nd * RT_MAX + e) * d, sizeof(T) * d);  \
        memcpy(ehi + (size_t)e * d, t->bmax + ((size_t)nd * RT_MAX + e) * d, sizeof(T) * d);  \
        ech[e] = t->child[(size_t)nd * RT_MAX + e];                                           \
    }                                                                                         \
    memcpy(elo + (size_t)RT_MAX * d, lo, sizeof(T) * d);                                      \
    memcpy(ehi + (size_t)RT_MAX * d, hi, sizeof(T) * d);                                      \
    ech[RT_MAX] = ch;                                                                         \
    for (int e = 0; e < M; e++) {                                                             \
        vol[e] = SFX(rt_volume)(t, elo + (size_t)e * d, ehi + (size_t)e * d);                 \
        grp[e] = -1;                                                                          \
    }                                                                                         \
    /* pick seeds: the pair wasting the most volume when combined */                          \
    int sa = 0, sb = 1;                                                                       \
    float worst = -INFINITY;                                                                  \
    for (int a = 0; a < M - 1; a++)                                                           \
        for (int b = a + 1; b < M; b++) {                                                     \
            SFX(rt_union)(d, elo + (size_t)a * d, ehi + (size_t)a * d, elo + (size_t)b * d,   \
                          ehi + (size_t)b * d, ulo, uhi);                                     \
            float w = SFX(rt_volume)(t, ulo, uhi) - vol[a] - vol[b];                          \
            if (w > worst) { worst = w; sa = a; sb = b; }                                     \
        }                                                                                     \
    int cnt[2] = {1, 1};                                                                      \
    grp[sa] = 0;                                                                              \
    grp[sb] = 1;                                                                              \
    memcpy(glo, elo + (size_t)sa * d, sizeof(T) * d);                                         \
    memcpy(ghi, ehi + (size_t)sa * d, sizeof(T) * d);                                         \
    memcpy(glo + d, elo + (size_t)sb * d, sizeof(T) * d);                                     \
    memcpy(ghi + d, ehi + (size_t)sb * d, sizeof(T) * d);                                     \
    int left = M - 2;                                                                         \
    while (left > 0) {                                                                        \
        /* a group that needs every remaining entry to reach the minimum takes them */       \
        int force = -1;                                                                       \
        if (cnt[0] + left == RT_MIN) force = 0;                                               \
        else if (cnt[1] + left == RT_MIN) force = 1;                                          \
        if (force >= 0) {                                                                     \
            for (int e = 0; e < M; e++)                                                       \
                if (grp[e] < 0) { grp[e] = force; cnt[force]++; }                             \
            break;                                                                            \
        }                                                                                     \
        float gv0 = SFX(rt_volume)(t, glo, ghi), gv1 = SFX(rt_volume)(t, glo + d, ghi + d);   \
        /* pick next: the entry with the largest preference for one group */                 \
        int best = -1, bestg = 0;                                                             \
        float bestdiff = -1.0f;                                                               \
        for (int e = 0; e < M; e++) {                                                         \
            if (grp[e] >= 0) continue;                                                        \
            SFX(rt_union)(d, glo, ghi, elo + (size_t)e * d, ehi + (size_t)e * d, ulo, uhi);   \
            float g0 = SFX(rt_volume)(t, ulo, uhi) - gv0;                                     \
            SFX(rt_union)(d, glo + d, ghi + d, elo + (size_t)e * d, ehi + (size_t)e * d, ulo, uhi); \
            float g1 = SFX(rt_volume)(t, ulo, uhi) - gv1;                                     \
            float diff = fabsf(g1 - g0);                                                      \
            if (diff > bestdiff) {                                                            \
                bestdiff = diff;                                                              \
                best = e;                                                                     \
                if (g0 < g1) bestg = 0;                                                       \
                else if (g1 < g0) bestg = 1;                                                  \
                else if (gv0 < gv1) bestg = 0;                                                \
                else if (gv1 < gv0) bestg = 1;                                                \
                else bestg = cnt[0] <= cnt[1] ? 0 : 1;                                        \
            }                                                                                 \
        }                                                                                     \
        grp[best] = bestg;                                                                    \
        cnt[bestg]++;                                                                         \
        SFX(rt_union)(d, glo + (size_t)bestg * d, ghi + (size_t)bestg * d, elo + (size_t)best * d, \
                      ehi + (size_t)best * d, glo + (size_t)bestg * d, ghi + (size_t)bestg * d); \
        left--;                                                                               \
    }                                                                                         \
    uint32_t other = SFX(rt_new)(t, t->level[nd]);                                            \
    t->count[nd] = 0;                                                                         \
    for (int e = 0; e < M; e++)                                                               \
        SFX(rt_put)(t, grp[e] == 0 ? nd : other, elo + (size_t)e * d, ehi + (size_t)e * d, ech[e]); \
    free(elo); free(ehi); free(glo); free(ghi); free(ulo); free(uhi);                          \
    return other;                                                                             \
}                                                                                             \
                                                                                              \
/* Inserts (lo, hi, ch) below node nd (at its level's leaf); returns a new   */             \
/* sibling of nd when nd split, else UINT32_MAX.                             */             \
static uint32_t SFX(rt_insert_at)(SFX(rtree) *t, uint32_t nd, const T *lo, const T *hi, uint32_t ch) { \
    const uint32_t d = t->d;                                                                  \
    if (t->level[nd] == 0) {                                                                  \
        if (t->count[nd] < RT_MAX) {                                                          \
            SFX(rt_put)(t, nd, lo, hi, ch);                                                   \
            return UINT32_MAX;                                                                \
        }                                                                                     \
        return SFX(rt_split)(t, nd, lo, hi, ch);                                              \
    }                                                                                         \
    /* choose the child needing the least enlargement (ties: smaller volume) */               \
    T *ulo = (T *)malloc(sizeof(T) * d), *uhi = (T *)malloc(sizeof(T) * d);                   \
    int best = 0;                                                                             \
    float bestInc = INFINITY, bestVol = INFINITY;                                             \
    for (int32_t e = 0; e < t->count[nd]; e++) {                                              \
        const T *elo = t->bmin + ((size_t)nd * RT_MAX + e) * d;                               \
        const T *ehi = t->bmax + ((size_t)nd * RT_MAX + e) * d;                               \
        float v = SFX(rt_volume)(t, elo, ehi);                                                \
        SFX(rt_union)(d, elo, ehi, lo, hi, ulo, uhi);                                         \
        float inc = SFX(rt_volume)(t, ulo, uhi) - v;                                          \
        if (inc < bestInc || (inc == bestInc && v < bestVol)) {                               \
            best = e;                                                                         \
            bestInc = inc;                                                                    \
            bestVol = v;                                                                      \
        }                                                                                     \
    }                                                                                         \
    const uint32_t c = t->child[(size_t)nd * RT_MAX + best];                                  \
    const uint32_t sib = SFX(rt_insert_at)(t, c, lo, hi, ch);                                 \
    /* refresh the chosen entry's box (it may have split) */                                  \
    SFX(rt_node_box)(t, c, ulo, uhi);                                                         \
    memcpy(t->bmin + ((size_t)nd * RT_MAX + best) * d, ulo, sizeof(T) * d);                   \
    memcpy(t->bmax + ((size_t)nd * RT_MAX + best) * d, uhi, sizeof(T) * d);                   \
    uint32_t ret = UINT32_MAX;                                                                \
    if (sib != UINT32_MAX) {                                                                  \
        SFX(rt_node_box)(t, sib, ulo, uhi);                                                   \
        if (t->count[nd] < RT_MAX) SFX(rt_put)(t, nd, ulo, uhi, sib);                          \
        else ret = SFX(rt_split)(t, nd, ulo, uhi, sib);                                       \
    }                                                                                         \
    free(ulo);                                                                                \
    free(uhi);                                                                                \
    return ret;                                                                               \
}                                                                                             \
                                                                                              \
static void SFX(rt_insert)(SFX(rtree) *t, const T *p, uint32_t id) {                          \
    const uint32_t sib = SFX(rt_insert_at)(t, t->root, p, p, id);                             \
    if (sib != UINT32_MAX) {                                                                  \
        const uint32_t d = t->d;                                                              \
        T *lo = (T *)malloc(sizeof(T) * d), *hi = (T *)malloc(sizeof(T) * d);                 \
        const uint32_t old = t->root;                                                         \
        const uint32_t nr = SFX(rt_new)(t, t->level[old] + 1);                                \
        SFX(rt_node_box)(t, old, lo, hi);                                                     \
        SFX(rt_put)(t, nr, lo, hi, old);                                                      \
        SFX(rt_node_box)(t, sib, lo, hi);                                                     \
        SFX(rt_put)(t, nr, lo, hi, sib);                                                      \
        t->root = nr;                                                                         \
        free(lo);                                                                             \
        free(hi);                                                                             \
    }                                                                                         \
}                                                                                             \
                                                                                              \
/* inclusive overlap search; appends leaf ids to *out */                                      \
static void SFX(rt_search)(const SFX(rtree) *t, uint32_t nd, const T *lo, const T *hi,       \
                           uint32_t **out, uint32_t *n, uint32_t *cap) {                      \
    const uint32_t d = t->d;                                                                  \
    for (int32_t e = 0; e < t->count[nd]; e++) {                                              \
        const T *elo = t->bmin + ((size_t)nd * RT_MAX + e) * d;                               \
        const T *ehi = t->bmax + ((size_t)nd * RT_MAX + e) * d;                               \
        int hit = 1;                                                                          \
        for (uint32_t j = 0; j < d && hit; j++)                                               \
            if (elo[j] > hi[j] || lo[j] > ehi[j]) hit = 0;                                    \
        if (!hit) continue;                                                                   \
        const uint32_t c = t->child[(size_t)nd * RT_MAX + e];                                 \
        if (t->level[nd] == 0) {                                                              \
            if (*n == *cap) {                                                                 \
                *cap *= 2;                                                                    \
                *out = (uint32_t *)realloc(*out, sizeof(uint32_t) * *cap);                    \
            }                                                                                 \
            (*out)[(*n)++] = c;                                                               \
        } else {                                                                              \
            SFX(rt_search)(t, c, lo, hi, out, n, cap);                                        \
        }                                                                                     \
    }                                                                                         \
}                                                                                             \
                                                                                              \
int SFX(cpu_rtree)(const T *data, uint32_t n, uint32_t d, T eps, const uint32_t *qids,        \
                   uint32_t nq, uint32_t *counts, double *build_seconds, double *query_seconds) { \
    SFX(rtree) t;                                                                             \
    memset(&t, 0, sizeof(t));                                                                 \
    t.d = d;                                                                                  \
    t.sphk = (float)unit_sphere_volume(d);                                                    \
    double t0 = now_s();                                                                      \
    t.root = SFX(rt_new)(&t, 0);                                                              \
    for (uint32_t i = 0; i < n; i++) SFX(rt_insert)(&t, data + (size_t)i * d, i);             \
    double t1 = now_s();                                                                      \
    T *lo = (T *)malloc(sizeof(T) * d), *hi = (T *)malloc(sizeof(T) * d);                     \
    uint32_t cap = 1024, m = 0;                                                               \
    uint32_t *cand = (uint32_t *)malloc(sizeof(uint32_t) * cap);                              \
    for (uint32_t s = 0; s < nq; s++) {                                                       \
        const T *q = data + (size_t)qids[s] * d;                                              \
        for (uint32_t j = 0; j < d; j++) {                                                    \
            lo[j] = q[j] - eps;                                                               \
            hi[j] = q[j] + eps;                                                               \
        }                                                                                     \
        m = 0;                                                                                \
        SFX(rt_search)(&t, t.root, lo, hi, &cand, &m, &cap);                                  \
        uint32_t cnt = 0;                                                                     \
        for (uint32_t c = 0; c < m; c++) cnt += (uint32_t)SFX(within)(q, data + (size_t)cand[c] * d, d, eps); \
        counts[s] = cnt;                                                                      \
    }                                                                                         \
    double t2 = now_s();                                                                      \
    *build_seconds = t1 - t0;                                                                 \
    *query_seconds = t2 - t1;                                                                 \
    free(cand); free(lo); free(hi);                                                           \
    free(t.level); free(t.count); free(t.child); free(t.bmin); free(t.bmax);                  \
    return 0;                                                                                 \
}

#define SFX_F32(x) x##_f32
#define SFX_F64(x) x##_f64
DEFINE_BASELINES(float, SFX_F32, sqrtf)
DEFINE_BASELINES(double, SFX_F64, sqrt)
