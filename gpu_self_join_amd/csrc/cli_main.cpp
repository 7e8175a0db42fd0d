// cli_main.cpp -- the command-line join: gdsjoin_main <file> <eps> <dim> 3|1|0
//
// Drop-in for the reference's `main <dataset> <epsilon> <dims> <mode>`
// (main.cu:66-120, mode 3 at main.cu:178-331, mode 1 at main.cu:150-175):
// mode 3 is the grid join, mode 1 the all-pairs GPU join (it prints the
// total as the reference's brute force does), mode 0 the table of the
// reference's CPU brute force (main.cu:121-141: its unfused host sums,
// computed by the all-pairs kernels, listed point by point).  Mode 3 imports the text dataset
// (gdsj_csv_load, import_dataset.cpp:49-95), runs the grid self-join on the
// GPU through the C ABI, prints the time and the total, and appends one line
// to gpu_stats.txt in the reference's format (main.cu:257-262).  The
// reference fixes dims, NUMINDEXEDDIM, DTYPE and the optimisation switches
// at compile time (params.h); here they are options:
//   --dtype float|double (default float)  --k K (default min(dim, 6))
//   --ilp I  --shortcircuit 0|1  --ilp-fold 0|1  --reorder 0|1  --stamp 0|1  --linearstamp 0|1
//   --sort 0|1 (SORT, params.h:46)
//   --queryreorder 0|1
//   --estimator exact|learned|sample  --filter plain|gram|gram-bf16|auto  --device D
//   --print   print every point's sorted neighbor list (PRINTNEIGHBORTABLE
//             with STAMP, main.cu:299-322)
//   --table-file    write DSSJ_out.txt (printNeighborTable main.cu:544-565)
//   --outlier-scores  write DSSJ_outlier_scores.txt (printOutlierScores
//             main.cu:568-626)
// Errors exit with a non-zero status (the reference prints and returns 0).
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include <algorithm>
#include <chrono>
#include <fstream>
#include <string>
#include <vector>

#include "gdsjoin.h"

static int usage() {
    fprintf(stderr,
            "usage: gdsjoin_main <dataset file> <epsilon> <dims> <mode> [options]\n"
            "  mode 3: GPU grid self-join; mode 1: all-pairs GPU join (total only);\n"
            "  mode 0: the CPU brute force's neighbor table (all pairs, unfused sums)\n"
            "  options: --dtype float|double --k K --ilp I --shortcircuit 0|1 --ilp-fold 0|1 --reorder 0|1 --sort 0|1\n"
            "           --stamp 0|1 --linearstamp 0|1 --queryreorder 0|1 --estimator exact|learned|sample\n"
            "           --filter plain|gram|gram-bf16|auto --device D --print --table-file\n"
            "           --outlier-scores\n");
    return 2;
}

// DSSJ_out.txt as printNeighborTable writes it (main.cu:544-565): a header,
// then one line per point id with its neighbor ids ascending, each followed
// by ", ".  ids holds the CSR rows in point order.
static void write_table(const std::vector<uint32_t> &counts, std::vector<uint32_t> &ids, uint64_t npts) {
    const char *fname = "DSSJ_out.txt";
    FILE *f = fopen(fname, "w");
    if (!f) {
        fprintf(stderr, "\nERROR: cannot write %s\n", fname);
        return;
    }
    printf("\n\nOutputting neighbors to: %s\n", fname);
    fprintf(f, "#data point (line is the point id), neighbor point ids\n");
    uint64_t off = 0;
    for (uint64_t i = 0; i < npts; i++) {
        std::sort(ids.begin() + (long)off, ids.begin() + (long)(off + counts[i]));
        for (uint32_t j = 0; j < counts[i]; j++) fprintf(f, "%u, ", ids[off + j]);
        fprintf(f, "\n");
        off += counts[i];
    }
    fclose(f);
}

// DSSJ_outlier_scores.txt as printOutlierScores writes it (main.cu:568-626):
// per point id, its neighbor count and its rank when the points are ordered
// by ascending count (0 = fewest neighbors).  The reference orders with
// std::sort, which leaves the order of equal counts unspecified; here equal
// counts keep ascending point ids (a stable order).
static void write_outlier_scores(const std::vector<uint32_t> &counts, uint64_t npts) {
    const char *fname = "DSSJ_outlier_scores.txt";
    std::vector<uint32_t> order(npts);
    for (uint64_t i = 0; i < npts; i++) order[i] = (uint32_t)i;
    std::stable_sort(order.begin(), order.end(), [&](uint32_t a, uint32_t b) { return counts[a] < counts[b]; });
    std::vector<uint32_t> rank(npts);
    for (uint64_t r = 0; r < npts; r++) rank[order[r]] = (uint32_t)r;
    FILE *f = fopen(fname, "w");
    if (!f) {
        fprintf(stderr, "\nERROR: cannot write %s\n", fname);
        return;
    }
    printf("\nOutputting outlier scores to: %s\n", fname);
    fprintf(f, "#data point (line is the point id), col0: Number of neighbors the point has within epsilon, ");
    fprintf(f, "col1: outlier ranking for col0\n");
    for (uint64_t i = 0; i < npts; i++) fprintf(f, "%u, %u\n", counts[i], rank[i]);
    fclose(f);
}

static int fail(const char *what, int rc) {
    fprintf(stderr, "\nERROR: %s: %s (status %d)\n", what, gdsj_last_error(), rc);
    return 1;
}

int main(int argc, char **argv) {
    if (argc < 5) {
        fprintf(stderr, "\n\nIncorrect number of input parameters.  \n"
                        "Should be dataset file, epsilon, number of dimensions, searchmode\n");
        return usage();
    }
    const char *fname = argv[1];
    const double eps_arg = atof(argv[2]);  // main.cu:95: DTYPE epsilon = atof(...)
    const int ndim = atoi(argv[3]);
    const int mode = atoi(argv[4]);
    int dtype = GDSJ_DTYPE_F32, device = 0, print = 0, table_file = 0, outlier = 0;
    gdsj_params p;
    gdsj_params_default(&p);
    for (int i = 5; i < argc; i++) {
        const char *a = argv[i];
        const char *v = i + 1 < argc ? argv[i + 1] : nullptr;
        if (!strcmp(a, "--print")) {
            print = 1;
            continue;
        }
        if (!strcmp(a, "--table-file")) {
            table_file = 1;
            continue;
        }
        if (!strcmp(a, "--outlier-scores")) {
            outlier = 1;
            continue;
        }
        if (!v) return usage();
        i++;
        if (!strcmp(a, "--dtype")) {
            if (!strcmp(v, "float")) dtype = GDSJ_DTYPE_F32;
            else if (!strcmp(v, "double")) dtype = GDSJ_DTYPE_F64;
            else return usage();
        } else if (!strcmp(a, "--k")) p.k = (uint32_t)atoi(v);
        else if (!strcmp(a, "--ilp")) p.ilp = (uint32_t)atoi(v);
        else if (!strcmp(a, "--shortcircuit")) p.shortcircuit = (uint32_t)atoi(v);
        else if (!strcmp(a, "--ilp-fold")) p.ilp_fold = (uint32_t)atoi(v);
        else if (!strcmp(a, "--sort")) p.sort = (uint32_t)atoi(v);
        else if (!strcmp(a, "--reorder")) p.reorder = (uint32_t)atoi(v);
        else if (!strcmp(a, "--stamp")) p.stamp = atoi(v) ? 1u : (p.stamp == 2 ? 2u : 0u);
        else if (!strcmp(a, "--linearstamp")) p.stamp = atoi(v) ? 2u : (p.stamp == 1 ? 1u : 0u);
        else if (!strcmp(a, "--queryreorder")) p.queryreorder = (uint32_t)atoi(v);
        else if (!strcmp(a, "--device")) device = atoi(v);
        else if (!strcmp(a, "--estimator")) {
            if (!strcmp(v, "exact")) p.estimator = GDSJ_EST_EXACT;
            else if (!strcmp(v, "learned")) p.estimator = GDSJ_EST_LEARNED;
            else if (!strcmp(v, "sample")) p.estimator = GDSJ_EST_SAMPLE;
            else return usage();
        } else if (!strcmp(a, "--filter")) {
            if (!strcmp(v, "plain")) p.filter = GDSJ_FILTER_PLAIN;
            else if (!strcmp(v, "gram")) p.filter = GDSJ_FILTER_GRAM;
            else if (!strcmp(v, "auto")) p.filter = GDSJ_FILTER_AUTO;
            else if (!strcmp(v, "gram-bf16")) p.filter = GDSJ_FILTER_GRAM_BF16;
            else return usage();
        } else {
            return usage();
        }
    }
    if (ndim <= 0) {
        fprintf(stderr, "\nERROR: the number of dimensions must be positive\n");
        return 1;
    }
    if (mode != 3 && mode != 1 && mode != 0) {
        fprintf(stderr, "\nERROR: search mode %d is not part of this build (3: GPU grid join, 1: all-pairs "
                        "GPU join, 0: the CPU brute force's table)\n", mode);
        return 2;
    }
    // DTYPE epsilon = atof(...): the double value rounded once to DTYPE
    const double eps = dtype == GDSJ_DTYPE_F32 ? (double)(float)eps_arg : eps_arg;
    printf("\nDataset file: %s", fname);
    printf("\nEpsilon: %f", eps);
    printf("\nNumber of dimensions (NDIM): %d\n", ndim);

    uint64_t nvals = 0;
    int rc = gdsj_csv_load(fname, dtype, nullptr, 0, &nvals);
    if (rc) return fail("import", rc);
    const size_t es = dtype == GDSJ_DTYPE_F64 ? 8 : 4;
    std::vector<unsigned char> data((size_t)nvals * es + 8);
    rc = gdsj_csv_load(fname, dtype, data.data(), nvals, &nvals);
    if (rc) return fail("import", rc);
    const uint64_t npts = nvals / (uint64_t)ndim;
    if (npts > 0xFFFFFFFFull) {
        fprintf(stderr, "\nERROR: more than 2^32 - 1 points\n");
        return 1;
    }
    printf("\nData import: Total size of all data (1-D) vect (number of points * GPUNUMDIM): %llu",
           (unsigned long long)nvals);
    printf("\nData import: Total data points: %llu", (unsigned long long)npts);

    gdsj_handle *h = nullptr;
    rc = gdsj_create(device, &h);
    if (rc) return fail("gdsj_create", rc);
    // warm-up (main.cu:193): a tiny join initialises the device and the kernels
    {
        float w[8] = {0, 0, 1, 1, 2, 2, 3, 3};
        uint32_t wc[4];
        uint64_t wt = 0;
        gdsj_params wp;
        gdsj_params_default(&wp);
        rc = gdsj_join_host(h, w, 4, 2, GDSJ_DTYPE_F32, 1.5, &wp, wc, &wt);
        if (rc) return fail("warm-up", rc);
    }
    std::vector<uint32_t> counts(npts ? npts : 1);
    uint64_t total = 0;
    if (mode == 1) {
        // main.cu:150-175: the all-pairs join in the reference kernel's
        // sequential accumulation (kernel.cu:500-503)
        printf("\nBrute force GPU (NO BATCHING):");
        gdsj_params bp = p;
        bp.ilp = 0;
        const auto b0 = std::chrono::steady_clock::now();
        rc = gdsj_bruteforce_host(h, data.data(), (uint32_t)npts, (uint32_t)ndim, dtype, eps, &bp, counts.data(),
                                  &total);
        const auto b1 = std::chrono::steady_clock::now();
        if (rc) {
            gdsj_destroy(h);
            return fail("brute force", rc);
        }
        printf("\nBRUTE FORCE Time on GPU: %f", std::chrono::duration<double>(b1 - b0).count());
        printf("\nTotal neighbours in table: %llu", (unsigned long long)total);
        printf("\n*********************************\n");
        if (outlier) write_outlier_scores(counts, npts);
        gdsj_destroy(h);
        return 0;
    }
    if (mode == 0) {
        // main.cu:121-141: the all-pairs table of CPUBruteForceTable
        // (main.cu:953-978), whose sums are the host's sequential unfused
        // runningDist += t*t; computed here by the all-pairs kernels with that
        // arithmetic, then listed point by point in the reference's format
        printf("\nBrute force CPU arithmetic (all pairs on the GPU): ");
        gdsj_params bp = p;
        bp.cpu_sum = 1;
        rc = gdsj_bruteforce_host(h, data.data(), (uint32_t)npts, (uint32_t)ndim, dtype, eps, &bp, counts.data(),
                                  &total);
        std::vector<uint32_t> ids(total ? total : 1);
        if (!rc) rc = gdsj_bruteforce_rows_host(h, ids.data(), total);
        if (rc) {
            gdsj_destroy(h);
            return fail("brute force", rc);
        }
        printf("\nTotal neighbors in table: %u", (unsigned)total);
        uint64_t off = 0;
        for (uint64_t i = 0; i < npts; i++) {
            printf("\npoint id: %d, neighbors: ", (int)i);
            for (uint32_t j = 0; j < counts[i]; j++) printf("%d,", (int)ids[off + j]);
            off += counts[i];
        }
        printf("\n");
        gdsj_destroy(h);
        return 0;
    }
    const auto t0 = std::chrono::steady_clock::now();
    rc = gdsj_join_host(h, data.data(), (uint32_t)npts, (uint32_t)ndim, dtype, eps, &p, counts.data(), &total);
    const auto t1 = std::chrono::steady_clock::now();
    if (rc) {
        gdsj_destroy(h);
        return fail("join", rc);
    }
    const double secs = std::chrono::duration<double>(t1 - t0).count();
    gdsj_stats st;
    gdsj_get_stats(h, &st);
    printf("\nTotal neighbors: %llu", (unsigned long long)total);
    printf("\nTime: %f", secs);
    printf("\nDevice time (ms): grid %.3f, count %.3f, scan %.3f, fill %.3f, total %.3f\n", st.ms_grid,
           st.ms_count, st.ms_scan, st.ms_fill, st.ms_total);

    const uint32_t k = p.k ? p.k : (uint32_t)std::min(ndim, 6);
    std::ofstream stats("gpu_stats.txt", std::ios::app);
    stats << secs << ", " << fname << ", " << eps << ", " << total
          << ", GPUNUMDIM/NUMINDEXEDDIM/ILP/STAMP/SORT/REORDER/SHORTCIRCUIT/QUERYREORDER/DTYPE(float/double): "
          << ndim << ", " << k << ", " << p.ilp << ", " << (p.stamp == 1 ? 1 : 0) << ", " << p.sort << ", " << p.reorder << ", "
          << p.shortcircuit << ", " << p.queryreorder << ", " << (dtype == GDSJ_DTYPE_F64 ? "double" : "float")
          << std::endl;
    stats.close();

    if (outlier) write_outlier_scores(counts, npts);
    if (print || table_file) {
        std::vector<uint32_t> ids(total ? total : 1);
        rc = gdsj_result_copy(h, ids.data(), total);
        if (rc) {
            gdsj_destroy(h);
            return fail("result copy", rc);
        }
        if (table_file) write_table(counts, ids, npts);
    }
    if (print) {
        std::vector<uint32_t> ids(total ? total : 1);
        rc = gdsj_result_copy(h, ids.data(), total);
        if (rc) {
            gdsj_destroy(h);
            return fail("result copy", rc);
        }
        uint64_t off = 0;
        for (uint64_t i = 0; i < npts; i++) {
            std::sort(ids.begin() + (long)off, ids.begin() + (long)(off + counts[i]));
            printf("\npoint id: %llu, neighbors: ", (unsigned long long)i);
            for (uint32_t j = 0; j < counts[i]; j++) printf("%u,", ids[off + j]);
            off += counts[i];
        }
        printf("\n");
    }
    gdsj_destroy(h);
    return 0;
}
