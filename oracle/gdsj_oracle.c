/*
 * gdsj_oracle.c -- CPU oracle for the exact Euclidean epsilon-grid self-join.
 *
 * TEST INFRASTRUCTURE ONLY: imported by tests/, by __graft_entry__.smoke()
 * as the checker, and by bench.py's cpu_baseline leg (timed as the CPU
 * baseline).  The product (gpu_self_join_amd) never links or calls it.
 *
 * It restates the reference's result contract on the CPU:
 *   REORDER            main.cu:981-1075
 *   grid dimensions    main.cu:891-948
 *   cell ids / lookup  main.cu:690-883
 *   candidate cells    kernel.cu:180-300, kernelloops.h, stamploops.h
 *   distance/predicate kernel.cu:307-365
 *   CSV parse          import_dataset.cpp:49-66
 * Output is CSR (per-point counts, neighbor ids ascending per point), the
 * table GDSJoinPy hands to Python (main.cu:488-489, 514-542, 630-636).
 *
 * Pinning (each has a test under tests/):
 *   - ILP=0 arithmetic, predicate, self-inclusion and candidate set: the
 *     README known answer 27,982,028 on test_dat.txt at eps=1.0 (README.md:82-84).
 *   - candidate restriction: equality with the all-pairs scan
 *     (orc_bruteforce) on small inputs, and the closed forms "eps below the
 *     minimum distance => every count is 1" and "all points in adjacent cells
 *     => count = n".
 *   - ILP>0 with SHORTCIRCUIT=1: the fold order follows kernel.cu:320-342;
 *     whether nvcc fuses the fold adds is NOT pinned by any reference vector
 *     (parity unpinned); both readings are implemented and their
 *     disagreements counted (stats[3]).
 *   - REORDER ties: std::sort leaves them unspecified; this keeps ties in
 *     ascending dimension order (parity unpinned for exact ties).
 *
 * Built with -ffp-contract=off: only the explicit fmaf/fma calls fuse.
 */
#include <math.h>
#include <omp.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#define ORC_MAXK 12
#define ORC_MAXD 4096
#define ORC_EINVAL (-1)
#define ORC_ERANGE (-2)

typedef struct {
    uint32_t k;            /* NUMINDEXEDDIM                                   */
    uint32_t ilp;          /* ILP: 0 = one sequential FMA accumulator         */
    uint32_t shortcircuit; /* SHORTCIRCUIT                                    */
    uint32_t reorder;      /* REORDER                                         */
    uint32_t fold_fused;   /* ILP>0 && SHORTCIRCUIT: 1 = fused fold reading   */
    int32_t nthreads;      /* OpenMP threads, 0 = default                     */
    uint32_t sort;         /* SORT (params.h:46): cells ordered by one dim    */
} orc_params;

/* order-independent per-point hash term (splitmix64 finalizer) */
uint64_t orc_mix64(uint64_t z)
{
    z += 0x9E3779B97F4A7C15ull;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}

static size_t orc_pow3(uint32_t k)
{
    size_t r = 1;
    for (uint32_t i = 0; i < k; i++) r *= 3;
    return r;
}

int orc_max_threads(void) { return omp_get_max_threads(); }

#define T float
#define SFX(x) x##_f32
/* The per-query loop is built twice, for hosts with and without the FMA
 * instruction (picked at load time); fmaf / fma stay correctly rounded
 * either way, the instruction only avoids a library call per term. */
#if defined(__x86_64__)
#define ORC_HOT __attribute__((target_clones("fma", "default")))
#else
#define ORC_HOT
#endif

#define FMA fmaf
#define SQRT sqrtf
#define CEIL ceilf
#define ORC_STRTO strtof
#include "oracle_impl.h"
#undef T
#undef SFX
#undef FMA
#undef SQRT
#undef CEIL
#undef ORC_STRTO

#define T double
#define SFX(x) x##_f64
#define FMA fma
#define SQRT sqrt
#define CEIL ceil
#define ORC_STRTO strtod
#include "oracle_impl.h"
#undef T
#undef SFX
#undef FMA
#undef SQRT
#undef CEIL
#undef ORC_STRTO
