// Command line front end with the reference's argv and result-file format
// (main.cu:43-181):
//
//   gts <data> <query|update file> <op> <k|r> <result file> [order max_size max_h]
//
//   op 0: kNN with k = atoi(argv[4]); the result file lists each query's k-th
//         NN distance ("%f ", main.cu:89-111)
//   op 1: range with r = (float)stod(argv[4]); the result file lists counts
//   op 2: the update stream of argv[2] with r; one count per query op
//
// The optional trailing arguments set the fan-out, leaf capacity and height
// (the reference's commented-out overrides, main.cu:54-61); height 0 sizes
// the tree to the data.
#include <chrono>
#include <cstdio>
#include <cstdlib>
#include <string>
#include <vector>

#include "gts.h"

static int die(const char *what, int rc)
{
    fprintf(stderr, "%s: %s\n", what, gts_last_error());
    return rc == GTS_E_IO ? 255 : 1;
}

int main(int argc, char **argv)
{
    if (argc < 6) {
        fprintf(stderr,
                "usage: %s <data> <query|update file> <0 knn|1 range|2 update> <k|r> <result file>"
                " [order max_size max_h]\n",
                argv[0]);
        return 2;
    }
    gts_build_params bp = {10, 20, 0, 0};
    if (argc >= 9) {
        bp.order = atoi(argv[6]);
        bp.max_size = atoi(argv[7]);
        bp.max_h = atoi(argv[8]);
    }
    gts_dataset *ds = nullptr;
    printf("Loading data file...\n");
    int rc = gts_dataset_load(argv[1], &ds);
    if (rc) return die("load", rc);
    const int op = atoi(argv[3]);
    const int k = atoi(argv[4]);
    const float r = (float)std::stod(argv[4]);
    std::vector<int32_t> qids;
    std::vector<gts_update_op> ops;
    if (op != 2) {
        printf("Loading query file...\n");
        int64_t nq = 0;
        rc = gts_load_queries(argv[2], nullptr, 0, &nq);
        if (rc) return die("load queries", rc);
        qids.resize(nq);
        rc = gts_load_queries(argv[2], qids.data(), nq, &nq);
        if (rc) return die("load queries", rc);
    } else {
        printf("Loading update file...\n");
        int64_t nu = 0;
        rc = gts_load_updates(argv[2], nullptr, 0, &nu);
        if (rc) return die("load updates", rc);
        ops.resize(nu);
        rc = gts_load_updates(argv[2], ops.data(), nu, &nu);
        if (rc) return die("load updates", rc);
    }

    printf("Index construction...\n");
    auto t0 = std::chrono::high_resolution_clock::now();
    gts_index *idx = nullptr;
    rc = gts_build(ds, &bp, &idx);
    if (rc) return die("build", rc);
    auto t1 = std::chrono::high_resolution_clock::now();
    const float time_index = std::chrono::duration<float>(t1 - t0).count();
    int height = 0;
    gts_index_info(idx, &height, nullptr, nullptr);
    printf("Tree height: %d\n", height);

    FILE *f = fopen(argv[5], "w");
    if (!f) {
        fprintf(stderr, "cannot open %s\n", argv[5]);
        return 255;
    }
    const int64_t nq = (int64_t)qids.size();
    if (op == 0) {
        if (k < 1) {
            fprintf(stderr, "k must be >= 1\n");
            return 2;
        }
        fprintf(f, "Knn search num: %d\nResult radius: \n", k);
        std::vector<float> kth(nq);
        printf("Searching...\n");
        auto s = std::chrono::high_resolution_clock::now();
        rc = gts_knn_query(idx, qids.data(), nq, k, nullptr, nullptr, kth.data());
        auto e = std::chrono::high_resolution_clock::now();
        if (rc) return die("knn", rc);
        const float ts = std::chrono::duration<float>(e - s).count();
        for (int64_t i = 0; i < nq; i++) fprintf(f, "%f ", kth[i]);
        printf("Time of index construction: %f\n", time_index);
        printf("Search time: %f\n", ts / (nq > 0 ? nq : 1));
        fprintf(f, "\nTime of index construction: %f\n", time_index);
        fprintf(f, "Search time: %f\n", ts / (nq > 0 ? nq : 1));
    } else if (op == 1) {
        fprintf(f, "Range search radius: %f\nResult num: \n", r);
        std::vector<int64_t> counts(nq);
        printf("Searching...\n");
        auto s = std::chrono::high_resolution_clock::now();
        rc = gts_range_query(idx, qids.data(), nq, r, counts.data(), nullptr, nullptr, nullptr, 0);
        auto e = std::chrono::high_resolution_clock::now();
        if (rc) return die("range", rc);
        const float ts = std::chrono::duration<float>(e - s).count();
        for (int64_t i = 0; i < nq; i++) fprintf(f, "%lld ", (long long)counts[i]);
        printf("\nTime of index construction: %f\n", time_index);
        printf("Search time: %f\n", ts / (nq > 0 ? nq : 1));
        fprintf(f, "\nTime of index construction: %f\n", time_index);
        fprintf(f, "Search time: %f\n", ts / (nq > 0 ? nq : 1));
    } else if (op == 2) {
        fprintf(f, "Range search radius (check for updates): %f\nResult num: \n", r);
        printf("Updating...\n");
        std::vector<int64_t> counts(ops.size() + 1);
        int64_t nres = 0;
        auto s = std::chrono::high_resolution_clock::now();
        rc = gts_update(idx, ops.data(), (int64_t)ops.size(), r, 10, nullptr, 0, counts.data(),
                        (int64_t)counts.size(), &nres);
        auto e = std::chrono::high_resolution_clock::now();
        if (rc) return die("update", rc);
        const float tu = std::chrono::duration<float>(e - s).count();
        for (int64_t i = 0; i < nres; i++) fprintf(f, "%lld ", (long long)counts[i]);
        const double per = tu / (ops.empty() ? 1 : (double)ops.size());
        printf("Time of index construction: %f\n", time_index);
        printf("Total update time: %f\n", per);
        fprintf(f, "\nTime of index construction: %f\n", time_index);
        fprintf(f, "Total update time: %f\n", per);
    } else {
        fprintf(stderr, "op must be 0, 1 or 2\n");
        return 2;
    }
    fclose(f);
    gts_index_free(idx);
    gts_dataset_free(ds);
    return 0;
}
