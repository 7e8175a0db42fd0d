/*
 * oracle_impl.h -- body of the CPU oracle, instantiated once per DTYPE by
 * gdsj_oracle.c with:
 *   T      float | double
 *   SFX(x) x##_f32 | x##_f64
 *   FMA    fmaf | fma           (correctly rounded fused multiply-add)
 *   SQRT   sqrtf | sqrt         (correctly rounded square root)
 *   CEIL   ceilf | ceil
 *
 * TEST INFRASTRUCTURE ONLY.  This is a plain restatement of the reference
 * algorithm ("algorithm 3", the exact Euclidean epsilon-grid self-join) used
 * as the checker for the HIP product path.  Nothing in the product links it.
 *
 * Every arithmetic step follows the reference line it cites; the file is
 * compiled with -ffp-contract=off so that only the explicit FMA calls fuse.
 */

/* ---- a1: ReorderByDimension (main.cu:981-1075) ------------------------- */
/* Mean and variance over rows 0,100,200,... in DTYPE, unfused (host code,
 * main.cu:1002-1022).  The denominator is size()*inv_sample evaluated in
 * DTYPE (size_t*DTYPE converts the size to DTYPE, main.cu:998,1012,1022).
 * Dimensions are then ordered by descending variance (main.cu:33-36,1040).
 * std::sort leaves ties unspecified; libstdc++ sorts arrays of <= 16
 * elements by insertion sort, which keeps ties in ascending dimension order,
 * and this restatement keeps that order for every size. */
void SFX(orc_reorder_perm)(const T *data, uint32_t n, uint32_t d,
                           uint32_t *perm, T *variance_out)
{
    const int sample = 100;
    const T inv_sample = (T)(1.0 / (sample * 1.0));
    const T denom = (T)n * inv_sample;
    T *var = (T *)malloc(sizeof(T) * (d ? d : 1));
    for (uint32_t i = 0; i < d; i++) {
        T sum = 0;
        for (uint64_t j = 0; j < n; j += sample)
            sum = sum + data[j * d + i];
        T avg = sum / denom;
        sum = 0;
        for (uint64_t j = 0; j < n; j += sample) {
            T t = data[j * d + i] - avg;
            T sq = t * t;
            sum = sum + sq;
        }
        var[i] = sum / denom;
        perm[i] = i;
    }
    /* stable insertion sort, descending variance */
    for (uint32_t i = 1; i < d; i++) {
        uint32_t pi = perm[i];
        T v = var[pi];
        int64_t j = (int64_t)i - 1;
        while (j >= 0 && var[perm[j]] < v) {
            perm[j + 1] = perm[j];
            j--;
        }
        perm[j + 1] = pi;
    }
    if (variance_out)
        for (uint32_t i = 0; i < d; i++) variance_out[i] = var[i];
    free(var);
}

/* ---- distance contract (kernel.cu:307-365) ----------------------------- */
/* t = candidate - query (kernel.cu:323,349).
 * ILP == 0: acc = fma(t,t,acc) per dimension in order (kernel.cu:347-349,
 *   nvcc contracts "acc += t*t" to one FMA; pinned by the README count).
 * ILP == m, SHORTCIRCUIT == 1 (kernel.cu:320-334): per chunk of m dims,
 *   r0 = fma(t,t,r0); r_j = fma(t,t,+0) = round(t*t) for j >= 1; then
 *   r0 = r0 + r_j for j = 1..m-1 (fold_fused = 0, the default reading) or
 *   r0 = fma(t,t,r0) for those dims (fold_fused = 1, the other reading).
 * ILP == m, SHORTCIRCUIT == 0 (kernel.cu:320-340): r_j = fma(t,t,r_j)
 *   over all chunks, then r0 = ((r0 + r1) + r2) ... + r_{m-1}.
 * SHORTCIRCUIT's early exit cannot change the outcome (partial sums of
 * non-negative terms never decrease under round-to-nearest).  orc_dist2
 * forms the full sum; orc_dist2_bounded (the join's checker) stops once the
 * running sum exceeds stop, which returns a value > stop and so cannot
 * change a predicate decided below stop. */
static inline T SFX(orc_dist2_bounded)(const T *p, const T *q, uint32_t d,
                                       uint32_t ilp, uint32_t sc, uint32_t fold_fused, T stop)
{
    if (ilp == 0) {
        T acc = 0;
        for (uint32_t l = 0; l < d; l++) {
            T t = q[l] - p[l];
            acc = FMA(t, t, acc);
            if ((l & 7) == 7 && acc > stop) return acc;
        }
        return acc;
    }
    if (ilp == 8 && sc && !fold_fused) {
        /* the same steps as the general loop below with the group size
         * fixed (params.h's ILP): fma(t,t,+0) of a later dim of a group is
         * the rounded product, and adding a missing dim's +0 leaves the
         * non-negative sum unchanged */
        T r0 = 0;
        uint32_t l = 0;
        for (; l + 8 <= d; l += 8) {
            const T t0 = q[l] - p[l], t1 = q[l + 1] - p[l + 1], t2 = q[l + 2] - p[l + 2],
                    t3 = q[l + 3] - p[l + 3], t4 = q[l + 4] - p[l + 4], t5 = q[l + 5] - p[l + 5],
                    t6 = q[l + 6] - p[l + 6], t7 = q[l + 7] - p[l + 7];
            const T s1 = t1 * t1, s2 = t2 * t2, s3 = t3 * t3, s4 = t4 * t4, s5 = t5 * t5, s6 = t6 * t6,
                    s7 = t7 * t7;
            r0 = FMA(t0, t0, r0);
            r0 = r0 + s1;
            r0 = r0 + s2;
            r0 = r0 + s3;
            r0 = r0 + s4;
            r0 = r0 + s5;
            r0 = r0 + s6;
            r0 = r0 + s7;
            if (r0 > stop) return r0;
        }
        if (l < d) {
            const T t0 = q[l] - p[l];
            r0 = FMA(t0, t0, r0);
            for (uint32_t j = l + 1; j < d; j++) {
                const T t = q[j] - p[j];
                const T sq = t * t;
                r0 = r0 + sq;
            }
        }
        return r0;
    }
    T r[64];
    for (uint32_t j = 0; j < ilp; j++) r[j] = 0;
    if (sc) {
        for (uint32_t l = 0; l < d; l += ilp) {
            for (uint32_t j = 0; j < ilp && l + j < d; j++) {
                T t = q[l + j] - p[l + j];
                if (j == 0 || fold_fused)
                    r[0] = FMA(t, t, r[0]);
                else
                    r[j] = FMA(t, t, r[j]);
            }
            if (!fold_fused)
                for (uint32_t j = 1; j < ilp; j++) {
                    r[0] = r[0] + r[j];
                    r[j] = 0;
                }
            if (r[0] > stop) return r[0];
        }
        return r[0];
    }
    for (uint32_t l = 0; l < d; l += ilp)
        for (uint32_t j = 0; j < ilp && l + j < d; j++) {
            T t = q[l + j] - p[l + j];
            r[j] = FMA(t, t, r[j]);
        }
    for (uint32_t j = 1; j < ilp; j++) r[0] = r[0] + r[j];
    return r[0];
}

static inline T SFX(orc_dist2)(const T *p, const T *q, uint32_t d,
                               uint32_t ilp, uint32_t sc, uint32_t fold_fused)
{
    return SFX(orc_dist2_bounded)(p, q, d, ilp, sc, fold_fused, (T)INFINITY);
}

/* exported form of the accumulation, for arithmetic tests */
T SFX(orc_acc)(const T *p, const T *q, uint32_t d, uint32_t ilp, uint32_t sc, uint32_t fold_fused)
{
    return SFX(orc_dist2)(p, q, d, ilp, sc, fold_fused);
}

/* predicate sqrt(acc) <= eps (kernel.cu:361,364) */
static inline int SFX(orc_pred)(T acc, T eps) { return SQRT(acc) <= eps; }

/* ---- grid (main.cu:690-948) -------------------------------------------- */
typedef struct {
    uint32_t n, d, k;
    T eps;
    T *pts;              /* n x d, dimensions already permuted (AoS)      */
    T *spts;             /* the same rows in cell order (order[])          */
    T minArr[ORC_MAXK];
    uint32_t nCells[ORC_MAXK];
    uint64_t totalCells;
    uint32_t *cellOf;    /* n: cell coordinates, n x k                     */
    uint32_t *order;     /* n: point ids sorted by (linear id, point id)   */
    uint64_t *cellId;    /* ncell: sorted unique linear ids                */
    uint32_t *cellStart; /* ncell+1: ranges into order                     */
    uint32_t ncell;
} SFX(orc_grid_t);

typedef struct { uint64_t key; T val; uint32_t id; } SFX(orc_kv_t);

/* cell linear id, then (SORT only; val = 0 otherwise) the value at the sort
 * dim, then point id (std::sort in GPU.cu:378 leaves ties unspecified) */
static int SFX(orc_kv_cmp)(const void *a, const void *b)
{
    const SFX(orc_kv_t) *x = (const SFX(orc_kv_t) *)a, *y = (const SFX(orc_kv_t) *)b;
    if (x->key != y->key) return x->key < y->key ? -1 : 1;
    if (x->val != y->val) return x->val < y->val ? -1 : 1;
    return x->id < y->id ? -1 : (x->id > y->id);
}

static void SFX(orc_grid_free)(SFX(orc_grid_t) *g)
{
    free(g->pts); free(g->spts); free(g->cellOf); free(g->order); free(g->cellId); free(g->cellStart);
    memset(g, 0, sizeof(*g));
}

/* builds the grid; returns 0 or a negative status */
static int SFX(orc_grid_build)(SFX(orc_grid_t) *g, const T *data, uint32_t n,
                               uint32_t d, T eps, const orc_params *p)
{
    memset(g, 0, sizeof(*g));
    uint32_t k = p->k;
    if (d == 0 || k < 1 || k > d || k > ORC_MAXK || !(eps > 0)) return ORC_EINVAL;
    g->n = n; g->d = d; g->k = k; g->eps = eps;
    uint32_t perm[ORC_MAXD];
    if (d > ORC_MAXD) return ORC_EINVAL;
    if (p->reorder)
        SFX(orc_reorder_perm)(data, n, d, perm, NULL);
    else
        for (uint32_t i = 0; i < d; i++) perm[i] = i;
    g->pts = (T *)malloc(sizeof(T) * (size_t)n * d + 1);
    for (uint64_t i = 0; i < n; i++)
        for (uint32_t j = 0; j < d; j++) g->pts[i * d + j] = data[i * d + perm[j]];
    if (n == 0) return 0;
    /* generateNDGridDimensions main.cu:899-946 */
    T maxArr[ORC_MAXK];
    for (uint32_t j = 0; j < k; j++) g->minArr[j] = maxArr[j] = g->pts[j];
    for (uint64_t i = 1; i < n; i++)
        for (uint32_t j = 0; j < k; j++) {
            T v = g->pts[i * d + j];
            if (v < g->minArr[j]) g->minArr[j] = v;
            if (v > maxArr[j]) maxArr[j] = v;
        }
    for (uint32_t j = 0; j < k; j++) {
        g->minArr[j] = g->minArr[j] - eps;
        maxArr[j] = maxArr[j] + eps;
    }
    uint64_t total = 1;
    for (uint32_t j = 0; j < k; j++) {
        g->nCells[j] = (uint32_t)CEIL((maxArr[j] - g->minArr[j]) / eps);
        uint64_t nt = total * (uint64_t)g->nCells[j];
        if (g->nCells[j] && nt / g->nCells[j] != total) return ORC_ERANGE;
        total = nt;
    }
    g->totalCells = total;
    /* populateNDGridIndexAndLookupArray main.cu:710-794: cell coordinate =
     * truncation of the quotient, linear id with dimension 0 fastest
     * (getLinearID_nDimensions main.cu:867-883); points of a cell kept in
     * ascending id order (main.cu:741-763). */
    g->cellOf = (uint32_t *)malloc(sizeof(uint32_t) * (size_t)n * k);
    SFX(orc_kv_t) *kv = (SFX(orc_kv_t) *)malloc(sizeof(SFX(orc_kv_t)) * n);
#pragma omp parallel for schedule(static)
    for (int64_t i = 0; i < (int64_t)n; i++) {
        uint64_t lin = 0, mult = 1;
        for (uint32_t j = 0; j < k; j++) {
            uint32_t c = (uint32_t)((g->pts[i * d + j] - g->minArr[j]) / eps);
            g->cellOf[i * k + j] = c;
            lin += (uint64_t)c * mult;
            mult *= g->nCells[j];
        }
        kv[i].key = lin;
        kv[i].id = (uint32_t)i;
        /* SORT (GPU.cu:351-391): a cell's points ordered by their value at
         * dim 0 when every dim is indexed, else at the first unindexed dim
         * (after REORDER) */
        kv[i].val = p->sort ? g->pts[i * d + (k == d ? 0 : k)] : (T)0;
    }
    qsort(kv, n, sizeof(SFX(orc_kv_t)), SFX(orc_kv_cmp));
    g->order = (uint32_t *)malloc(sizeof(uint32_t) * n);
    g->cellId = (uint64_t *)malloc(sizeof(uint64_t) * n);
    g->cellStart = (uint32_t *)malloc(sizeof(uint32_t) * (n + 1));
    uint32_t nc = 0;
    for (uint32_t i = 0; i < n; i++) {
        g->order[i] = kv[i].id;
        if (i == 0 || kv[i].key != kv[i - 1].key) {
            g->cellId[nc] = kv[i].key;
            g->cellStart[nc] = i;
            nc++;
        }
    }
    g->cellStart[nc] = n;
    g->ncell = nc;
    free(kv);
    /* the points again in cell order: a cell's candidates are read as one
     * contiguous block */
    g->spts = (T *)malloc(sizeof(T) * (size_t)n * d + 1);
#pragma omp parallel for schedule(static)
    for (int64_t i = 0; i < (int64_t)n; i++)
        memcpy(g->spts + (size_t)i * d, g->pts + (size_t)g->order[i] * d, sizeof(T) * d);
    return 0;
}

static inline int64_t SFX(orc_find_cell)(const SFX(orc_grid_t) *g, uint64_t lin)
{
    int64_t lo = 0, hi = (int64_t)g->ncell - 1;
    while (lo <= hi) {
        int64_t mid = (lo + hi) >> 1;
        uint64_t v = g->cellId[mid];
        if (v == lin) return mid;
        if (v < lin) lo = mid + 1; else hi = mid - 1;
    }
    return -1;
}

/* Non-empty cells adjacent to the cell with coordinates c (kernel.cu:180-253
 * and the loop nest kernel.cu:292-300 + kernelloops.h): per dimension the
 * range is [c-1, c+1] clamped to [0, nCells-1]; the per-dimension mask
 * of the reference only drops coordinates that hold no point at all, so the
 * candidate set is every non-empty cell in that box.  Returns the count;
 * cells are written in ascending linear id order. */
static uint32_t SFX(orc_adjacent)(const SFX(orc_grid_t) *g, const uint32_t *c,
                                  int64_t *out)
{
    uint32_t k = g->k;
    uint32_t lo[ORC_MAXK], hi[ORC_MAXK], cur[ORC_MAXK];
    for (uint32_t j = 0; j < k; j++) {
        lo[j] = c[j] == 0 ? 0 : c[j] - 1;
        uint32_t h = c[j] + 1;
        hi[j] = (g->nCells[j] - 1 < h) ? g->nCells[j] - 1 : h;
        cur[j] = lo[j];
        if (lo[j] > hi[j]) return 0;
    }
    uint32_t cnt = 0;
    for (;;) {
        uint64_t lin = 0, mult = 1;
        for (uint32_t j = 0; j < k; j++) {
            lin += (uint64_t)cur[j] * mult;
            mult *= g->nCells[j];
        }
        int64_t ci = SFX(orc_find_cell)(g, lin);
        if (ci >= 0) out[cnt++] = ci;
        /* odometer, dimension 0 fastest => ascending linear ids */
        uint32_t j = 0;
        while (j < k) {
            if (cur[j] < hi[j]) { cur[j]++; break; }
            cur[j] = lo[j];
            j++;
        }
        if (j == k) break;
    }
    /* ascending linear id already holds when no aliasing occurs; sort to be safe */
    for (uint32_t a = 1; a < cnt; a++) {
        int64_t v = out[a];
        int64_t b = (int64_t)a - 1;
        while (b >= 0 && out[b] > v) { out[b + 1] = out[b]; b--; }
        out[b + 1] = v;
    }
    return cnt;
}

static int SFX(orc_u32_cmp)(const void *a, const void *b)
{
    uint32_t x = *(const uint32_t *)a, y = *(const uint32_t *)b;
    return x < y ? -1 : (x > y);
}

/* SORT=1 (kernel.cu:400-467): the part [*s, *e) of candidate cell c, whose
 * points are ordered by their value at the sort dim, that the reference
 * evaluates for query qi.
 * Every dim indexed (sort dim 0, kernel.cu:403-431): a cell with query qi's
 * coordinate in dim 0 is evaluated whole; a cell to the left is walked from
 * its last point down, one to the right from its first point up, and the
 * walk stops at the first point whose |value - query value| exceeds eps.
 * Otherwise (sort dim k, "sortidu", kernel.cu:433-465): a stepping search
 * for the first point whose value is within eps below the query's, two steps
 * back while still within eps, then points while value - query < eps. */
static void SFX(orc_sort_window)(const SFX(orc_grid_t) *g, int64_t c, uint32_t qi, uint32_t *s, uint32_t *e)
{
    const uint32_t lo = *s, hi = *e;
    const T eps = g->eps;
    if (g->k == g->d) {
        const T q = g->pts[(size_t)qi * g->d];
        const uint32_t qc = g->cellOf[(size_t)qi * g->k];
        const uint32_t cc = (uint32_t)(g->cellId[c] % g->nCells[0]);
        if (cc == qc) return;
        if (qc > cc) { /* left: from the last point down */
            uint32_t t = hi;
            while (t > lo) {
                const T v = g->spts[(size_t)(t - 1) * g->d];
                const T diff = v - q;
                if ((diff < 0 ? -diff : diff) > eps) break;
                t--;
            }
            *s = t;
        } else {
            uint32_t t = lo;
            while (t < hi) {
                const T v = g->spts[(size_t)t * g->d];
                const T diff = v - q;
                if ((diff < 0 ? -diff : diff) > eps) break;
                t++;
            }
            *e = t;
        }
        return;
    }
    const uint32_t sd = g->k;
    const T q = g->pts[(size_t)qi * g->d + sd];
#define ORC_V(i) (g->spts[(size_t)(lo + (uint32_t)(i)) * g->d + sd])
    const int length = (int)(hi - lo) - 1;
    int idx = length / 2;
    for (int step = (length + 3) / 4; step > 1; step = (step + 1) / 2) {
        const T dq = q - ORC_V(idx);
        if (dq >= eps) {
            idx += step;
            if (idx > length) idx = length;
        } else {
            idx -= step;
            if (idx < 0) {
                idx = 0;
                step = 0;
            }
        }
    }
    if (idx > 0 && q - ORC_V(idx - 1) <= eps) idx--;
    if (idx > 0 && q - ORC_V(idx - 1) <= eps) idx--;
    int end = idx;
    while (end <= length && ORC_V(end) - q < eps) end++;
#undef ORC_V
    *s = lo + (uint32_t)idx;
    *e = lo + (uint32_t)end;
}

/* core: evaluates query point qi against every candidate of its adjacent
 * cells.  Returns the count; optionally appends ids and accumulates stats. */
ORC_HOT static uint32_t SFX(orc_query)(const SFX(orc_grid_t) *g, const int64_t *adj,
                               uint32_t nadj, uint32_t qi, const orc_params *p,
                               uint64_t *hash, uint32_t *ids, uint64_t *ncand,
                               uint64_t *nambig)
{
    const T *pq = g->pts + (size_t)qi * g->d;
    uint32_t cnt = 0;
    uint64_t h = 0, cand = 0, amb = 0;
    int check_amb = (p->ilp > 0 && p->shortcircuit && nambig != NULL);
    /* The two fold readings differ by a few ulps of the sum: a sum outside
     * eps^2 * (1 -+ 2^-10) decides both alike, and a running sum past the
     * upper end cannot come back below eps^2 (the sums stop there). */
    const T e2 = g->eps * g->eps;
    const T band = e2 * (T)(1.0 / 1024.0);
    const T stop = e2 + band;
    for (uint32_t a = 0; a < nadj; a++) {
        uint32_t s = g->cellStart[adj[a]], e = g->cellStart[adj[a] + 1];
        if (p->sort) SFX(orc_sort_window)(g, adj[a], qi, &s, &e);
        cand += e - s;
        for (uint32_t t = s; t < e; t++) {
            uint32_t cj = g->order[t];
            const T *pc = g->spts + (size_t)t * g->d;
            T acc = SFX(orc_dist2_bounded)(pq, pc, g->d, p->ilp, p->shortcircuit, p->fold_fused, stop);
            int hit = SFX(orc_pred)(acc, g->eps);
            if (check_amb && acc >= e2 - band && acc <= stop) {
                T acc2 = SFX(orc_dist2)(pq, pc, g->d, p->ilp, p->shortcircuit, !p->fold_fused);
                if (SFX(orc_pred)(acc2, g->eps) != hit) amb++;
            }
            if (hit) {
                if (ids) ids[cnt] = cj;
                h += orc_mix64(cj);
                cnt++;
            }
        }
    }
    if (ids && cnt > 1) qsort(ids, cnt, sizeof(uint32_t), SFX(orc_u32_cmp));
    if (hash) *hash = h;
    if (ncand) *ncand += cand;
    if (nambig) *nambig += amb;
    return cnt;
}

/* Full self-join.  counts[n] always written; hashes[n] (per-point sum of
 * orc_mix64(neighbor id)) if non-NULL; ids (CSR in point order, each list
 * ascending) if non-NULL and ids_cap >= total.  stats (nullable):
 * [0] total pairs [1] candidate comparisons [2] non-empty cells
 * [3] pairs on which the two ILP fold readings disagree
 * [4] total cells (incl. empty) */
int SFX(orc_join)(const T *data, uint32_t n, uint32_t d, T eps,
                  const orc_params *p, uint32_t *counts, uint64_t *hashes,
                  uint32_t *ids, uint64_t ids_cap, uint64_t *stats)
{
    SFX(orc_grid_t) g;
    int rc = SFX(orc_grid_build)(&g, data, n, d, eps, p);
    if (rc) { SFX(orc_grid_free)(&g); return rc; }
    uint64_t total = 0, cand = 0, amb = 0;
    uint64_t *offs = NULL;
    int nth = p->nthreads > 0 ? p->nthreads : omp_get_max_threads();
    int pass_max = ids ? 2 : 1;
    for (int pass = 0; pass < pass_max; pass++) {
        if (pass == 1) {
            if (total > ids_cap) { rc = ORC_ERANGE; break; }
            offs = (uint64_t *)malloc(sizeof(uint64_t) * ((size_t)n + 1));
            uint64_t run = 0;
            for (uint32_t i = 0; i < n; i++) { offs[i] = run; run += counts[i]; }
            offs[n] = run;
        }
        total = 0;
#pragma omp parallel num_threads(nth) reduction(+:total,cand,amb)
        {
            int64_t *adj = (int64_t *)malloc(sizeof(int64_t) * orc_pow3(g.k));
#pragma omp for schedule(dynamic, 16)
            for (int64_t c = 0; c < (int64_t)g.ncell; c++) {
                uint32_t s = g.cellStart[c], e = g.cellStart[c + 1];
                uint32_t q0 = g.order[s];
                uint32_t nadj = SFX(orc_adjacent)(&g, g.cellOf + (size_t)q0 * g.k, adj);
                for (uint32_t t = s; t < e; t++) {
                    uint32_t qi = g.order[t];
                    uint64_t h = 0;
                    uint32_t cnt = SFX(orc_query)(&g, adj, nadj, qi, p, &h,
                                                  pass ? ids + offs[qi] : NULL,
                                                  pass ? NULL : &cand,
                                                  (pass || !stats) ? NULL : &amb);
                    if (!pass) {
                        counts[qi] = cnt;
                        if (hashes) hashes[qi] = h;
                    }
                    total += cnt;
                }
            }
            free(adj);
        }
        if (pass == 0 && stats) {
            stats[0] = total; stats[1] = cand; stats[2] = g.ncell; stats[3] = amb;
            stats[4] = g.totalCells;
        }
    }
    free(offs);
    SFX(orc_grid_free)(&g);
    return rc;
}

/* Join restricted to the query ids qids[0..nq): counts/hashes per listed
 * query (the grid is built over all n points). */
int SFX(orc_join_subset)(const T *data, uint32_t n, uint32_t d, T eps,
                         const orc_params *p, const uint32_t *qids, uint32_t nq,
                         uint32_t *counts, uint64_t *hashes, uint64_t *stats)
{
    SFX(orc_grid_t) g;
    double t0 = omp_get_wtime();
    int rc = SFX(orc_grid_build)(&g, data, n, d, eps, p);
    double t1 = omp_get_wtime();
    if (rc) { SFX(orc_grid_free)(&g); return rc; }
    for (uint32_t i = 0; i < nq; i++)
        if (qids[i] >= n) { SFX(orc_grid_free)(&g); return ORC_EINVAL; }
    uint64_t cand = 0, amb = 0, total = 0;
    int nth = p->nthreads > 0 ? p->nthreads : omp_get_max_threads();
#pragma omp parallel num_threads(nth) reduction(+:cand,amb,total)
    {
        int64_t *adj = (int64_t *)malloc(sizeof(int64_t) * orc_pow3(g.k));
#pragma omp for schedule(dynamic, 4)
        for (int64_t i = 0; i < (int64_t)nq; i++) {
            uint32_t qi = qids[i];
            uint32_t nadj = SFX(orc_adjacent)(&g, g.cellOf + (size_t)qi * g.k, adj);
            uint64_t h = 0;
            counts[i] = SFX(orc_query)(&g, adj, nadj, qi, p, &h, NULL, &cand, &amb);
            if (hashes) hashes[i] = h;
            total += counts[i];
        }
        free(adj);
    }
    double t2 = omp_get_wtime();
    if (stats) {
        stats[0] = total; stats[1] = cand; stats[2] = g.ncell; stats[3] = amb;
        stats[4] = g.totalCells;
        stats[5] = (uint64_t)((t1 - t0) * 1e6); /* grid build, microseconds */
        stats[6] = (uint64_t)((t2 - t1) * 1e6); /* queries, microseconds */
    }
    SFX(orc_grid_free)(&g);
    return 0;
}

/* All-pairs scan with the same arithmetic and no grid (tiny n only):
 * counts and hashes per point; the grid's candidate restriction must agree
 * with it except where rounding puts a true neighbor two cells away. */
int SFX(orc_bruteforce)(const T *data, uint32_t n, uint32_t d, T eps,
                        const orc_params *p, uint32_t *counts, uint64_t *hashes)
{
    uint32_t perm[ORC_MAXD];
    if (d == 0 || d > ORC_MAXD || !(eps > 0)) return ORC_EINVAL;
    if (p->reorder) SFX(orc_reorder_perm)(data, n, d, perm, NULL);
    else for (uint32_t i = 0; i < d; i++) perm[i] = i;
    T *pts = (T *)malloc(sizeof(T) * (size_t)n * d + 1);
    for (uint64_t i = 0; i < n; i++)
        for (uint32_t j = 0; j < d; j++) pts[i * d + j] = data[i * d + perm[j]];
#pragma omp parallel for schedule(dynamic, 16)
    for (int64_t i = 0; i < (int64_t)n; i++) {
        uint32_t cnt = 0;
        uint64_t h = 0;
        for (uint32_t j = 0; j < n; j++) {
            T acc = SFX(orc_dist2)(pts + i * d, pts + (size_t)j * d, d, p->ilp,
                                   p->shortcircuit, p->fold_fused);
            if (SFX(orc_pred)(acc, eps)) { cnt++; h += orc_mix64(j); }
        }
        counts[i] = cnt;
        if (hashes) hashes[i] = h;
    }
    free(pts);
    return 0;
}

/* Grid summary for cross-checks of the device grid: minArr/nCells (k),
 * totalCells, number of non-empty cells, and the dimension permutation. */
int SFX(orc_grid_info)(const T *data, uint32_t n, uint32_t d, T eps,
                       const orc_params *p, T *minArr, uint32_t *nCells,
                       uint64_t *totalCells, uint32_t *ncell, uint32_t *perm,
                       uint64_t *cellIds /* nullable, cap n */,
                       uint32_t *order /* nullable, n */)
{
    SFX(orc_grid_t) g;
    int rc = SFX(orc_grid_build)(&g, data, n, d, eps, p);
    if (rc) { SFX(orc_grid_free)(&g); return rc; }
    for (uint32_t j = 0; j < g.k; j++) { minArr[j] = g.minArr[j]; nCells[j] = g.nCells[j]; }
    *totalCells = g.totalCells;
    *ncell = g.ncell;
    if (perm) {
        if (p->reorder) SFX(orc_reorder_perm)(data, n, d, perm, NULL);
        else for (uint32_t i = 0; i < d; i++) perm[i] = i;
    }
    if (cellIds) memcpy(cellIds, g.cellId, sizeof(uint64_t) * g.ncell);
    if (order) memcpy(order, g.order, sizeof(uint32_t) * n);
    SFX(orc_grid_free)(&g);
    return 0;
}

/* CSV/whitespace text -> DTYPE values, parsing each decimal token directly
 * into DTYPE as "stringstream >> DTYPE" does (import_dataset.cpp:49-66).
 * Returns the number of values written (stops at cap). */
size_t SFX(orc_parse_text)(const char *text, size_t len, T *out, size_t cap)
{
    size_t cnt = 0;
    const char *s = text, *end = text + len;
    char buf[128];
    while (s < end && cnt < cap) {
        while (s < end && (*s == ',' || *s == ' ' || *s == '\n' || *s == '\r' || *s == '\t')) s++;
        if (s >= end) break;
        size_t m = 0;
        while (s < end && *s != ',' && *s != ' ' && *s != '\n' && *s != '\r' && *s != '\t' && m < sizeof(buf) - 1)
            buf[m++] = *s++;
        buf[m] = 0;
        out[cnt++] = ORC_STRTO(buf, NULL);
    }
    return cnt;
}
