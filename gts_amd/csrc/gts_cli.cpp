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
// the tree to the data.  Options anywhere after the five positional
// arguments: --ids FILE and --dists FILE write every result's IDs and
// distances (ops 0 and 1) with the reference's writers saveK / saveR
// (file.cuh:154-199, defined there but not called by main.cu).
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
                " [order max_size max_h] [--ids FILE] [--dists FILE]\n",
                argv[0]);
        return 2;
    }
    gts_build_params bp = {10, 20, 0, 0};
    const char *ids_path = nullptr, *dists_path = nullptr;
    std::vector<char *> pos;
    for (int i = 1; i < argc; i++) {
        const std::string a = argv[i];
        if ((a == "--ids" || a == "--dists") && i + 1 < argc) {
            (a == "--ids" ? ids_path : dists_path) = argv[++i];
        } else {
            pos.push_back(argv[i]);
        }
    }
    if (pos.size() < 5) {
        fprintf(stderr, "missing positional arguments\n");
        return 2;
    }
    if (pos.size() >= 8) {
        bp.order = atoi(pos[5]);
        bp.max_size = atoi(pos[6]);
        bp.max_h = atoi(pos[7]);
    }
    argv = pos.data() - 1;   // argv[1..5]: the positional arguments
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
        const bool keep = ids_path || dists_path;
        std::vector<int32_t> rid(keep ? nq * k : 0);
        std::vector<float> rdist(keep ? nq * k : 0);
        printf("Searching...\n");
        auto s = std::chrono::high_resolution_clock::now();
        rc = gts_knn_query(idx, qids.data(), nq, k, keep ? rid.data() : nullptr, keep ? rdist.data() : nullptr,
                           kth.data());
        auto e = std::chrono::high_resolution_clock::now();
        if (rc) return die("knn", rc);
        if (keep && (rc = gts_save_knn(ids_path, dists_path, rid.data(), rdist.data(), k, nq))) return die("saveK", rc);
        const float ts = std::chrono::duration<float>(e - s).count();
        for (int64_t i = 0; i < nq; i++) fprintf(f, "%f ", kth[i]);
        printf("Time of index construction: %f\n", time_index);
        printf("Search time: %f\n", ts / (nq > 0 ? nq : 1));
        fprintf(f, "\nTime of index construction: %f\n", time_index);
        fprintf(f, "Search time: %f\n", ts / (nq > 0 ? nq : 1));
    } else if (op == 1) {
        fprintf(f, "Range search radius: %f\nResult num: \n", r);
        std::vector<int64_t> counts(nq);
        const bool keep = ids_path || dists_path;
        std::vector<int64_t> offs(nq + 1);
        std::vector<int32_t> rid;
        std::vector<float> rdist;
        printf("Searching...\n");
        auto s = std::chrono::high_resolution_clock::now();
        rc = gts_range_query(idx, qids.data(), nq, r, counts.data(), nullptr, nullptr, nullptr, 0);
        auto e = std::chrono::high_resolution_clock::now();
        if (rc) return die("range", rc);
        if (keep) {   // the ID lists, outside the timed search (the reference times counts only)
            int64_t total = 0;
            for (int64_t i = 0; i < nq; i++) total += counts[i];
            rid.resize(std::max<int64_t>(total, 1));
            rdist.resize(std::max<int64_t>(total, 1));
            rc = gts_range_query(idx, qids.data(), nq, r, counts.data(), offs.data(), rid.data(), rdist.data(),
                                 total);
            if (rc) return die("range", rc);
            if ((rc = gts_save_range(ids_path, dists_path, rid.data(), rdist.data(), counts.data(), offs.data(), nq)))
                return die("saveR", rc);
        }
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
        rc = gts_update(idx, ops.data(), (int64_t)ops.size(), r, 10, nullptr, 0, counts.data(),
                        (int64_t)counts.size(), &nres);
        if (rc) return die("update", rc);
        for (int64_t i = 0; i < nres; i++) fprintf(f, "%lld ", (long long)counts[i]);
        // per-op averages as main.cu:152-158 forms them: search time per query
        // op, update time per insert or delete op, and their sum
        gts_stats st{};
        gts_last_stats(idx, &st);
        const double ts = st.update_queries ? st.ms_update_search / 1e3 / (double)st.update_queries : 0.0;
        const double tu = st.update_changes ? st.ms_update_apply / 1e3 / (double)st.update_changes : 0.0;
        printf("Time of index construction: %f\n", time_index);
        printf("Total update time: %f\n", ts + tu);
        printf("Search time in update: %f\n", ts);
        printf("Update time in update: %f\n", tu);
        fprintf(f, "\nTime of index construction: %f\n", time_index);
        fprintf(f, "Total update time: %f\n", ts + tu);
        fprintf(f, "Search time in update : % f\n", ts);
        fprintf(f, "Update time in update: %f\n", tu);
    } else {
        fprintf(stderr, "op must be 0, 1 or 2\n");
        return 2;
    }
    fclose(f);
    gts_index_free(idx);
    gts_dataset_free(ds);
    return 0;
}
