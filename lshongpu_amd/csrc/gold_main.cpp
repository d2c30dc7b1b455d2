// gold_main.cpp -- truth files on the GPU (GoldStandard/goldstandard.cpp:
// 175-200 computes them on the CPU for -cos only):
//
//   goldstandard K dataFile queryFile outputFile [distanceFunc]
//
// distanceFunc 0 (default) ranks by -cos as goldstandard.cpp:111-121 does;
// 1, 2, 3 are the angular, generalized-Jaccard and Euclidean distances of
// implementation 2.  The neighbours are the exact k-NN with (distance, ID)
// order (FP64-verified); the file keeps the query IDs of the query file.
#include <cstdio>
#include <cstdlib>
#include <vector>

#include "../../include/lshgpu.h"

static int die(const char* what) {
    std::fprintf(stderr, "%s: %s\n", what, lsh_last_error());
    return 1;
}

int main(int argc, char** argv) {
    if (argc != 5 && argc != 6) {
        std::fprintf(stderr, "usage: %s K dataFile queryFile outputFile [distanceFunc]\n", argv[0]);
        return 2;
    }
    const int K = std::atoi(argv[1]);
    const int func = argc == 6 ? std::atoi(argv[5]) : 0;
    if (K < 1 || func < 0 || func > 3) {
        std::fprintf(stderr, "K must be >= 1 and distanceFunc 0..3\n");
        return 2;
    }
    float *data = nullptr, *queries = nullptr;
    int32_t* qids = nullptr;
    int64_t n = 0, nq = 0;
    int d = 0, dq = 0;
    if (lsh_read_text_dataset(argv[2], &data, &n, &d, nullptr)) return die("reading data");
    if (lsh_read_text_dataset(argv[3], &queries, &nq, &dq, &qids)) return die("reading queries");
    if (d != dq) {
        std::fprintf(stderr, "data has %d dimensions, queries %d\n", d, dq);
        return 1;
    }
    lsh_index_config_t cfg;
    lsh_index_config_default(&cfg);
    cfg.implementation = func == 0 ? 1 : 2;
    cfg.distanceFunc = func == 0 ? 1 : func;
    cfg.d = d;
    cfg.N_data = n;
    lsh_index_t* ix = nullptr;
    if (lsh_index_create(&cfg, data, 0, &ix)) return die("building the index");
    const int kq = lsh_calculate_k(K);
    std::vector<lsh_point_t> res((size_t)nq * kq);
    if (lsh_index_search(ix, queries, (int)nq, K, 0, res.data(), 0, nullptr)) return die("search");
    if (lsh_write_truth(argv[4], res.data(), qids, (int)nq, kq, K)) return die("writing the truth file");
    lsh_index_destroy(ix);
    lsh_free(data);
    lsh_free(queries);
    lsh_free(qids);
    std::printf("Wrote %lld queries x %d neighbours to %s\n", (long long)nq, K, argv[4]);
    return 0;
}
