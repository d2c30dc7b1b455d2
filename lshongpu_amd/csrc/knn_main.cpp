// knn_main.cpp -- the reference runner (LinearScan/kernel.cu:119-196) on the
// lshgpu C ABI, with the same 15 arguments:
//
//   knn data queries truth runValidation writeResults k implementation sketchDim
//       distanceFunc framework bucketKeyBits tables keysImplementation
//       runWithSketchedData resultFile
//
// It parses both datasets, runs linearScans (framework 0) or LSH (framework
// 1) with k rounded by calculateK, optionally validates against the truth
// file and prints the results, and appends the 13-column performance line.
// Errors print the library's message and exit with status 1 (the reference
// exits with -1 from inside the failing call).
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstdlib>

#include "../../include/lshgpu.h"

static int die(const char* what) {
    std::fprintf(stderr, "%s: %s\n", what, lsh_last_error());
    return 1;
}

int main(int argc, char** argv) {
    if (argc != 16) {
        std::fprintf(stderr,
                     "usage: %s data queries truth runValidation writeResults k implementation sketchDim "
                     "distanceFunc framework bucketKeyBits tables keysImplementation runWithSketchedData "
                     "resultFile\n",
                     argv[0]);
        return 2;
    }
    const char* data_path = argv[1];
    const char* query_path = argv[2];
    const char* truth_path = argv[3];
    const int run_validation = std::atoi(argv[4]);
    const int write_res = std::atoi(argv[5]);
    const int report_k = std::atoi(argv[6]);
    const int implementation = std::atoi(argv[7]);
    const int k = lsh_calculate_k(report_k);
    const int sketch_dim = std::atoi(argv[8]);
    const int distance_func = std::atoi(argv[9]);
    const int framework = std::atoi(argv[10]);
    const int bucket_key_bits = std::atoi(argv[11]);
    const int tables = std::atoi(argv[12]);
    const int keys_implementation = std::atoi(argv[13]);
    const int with_sketch = std::atoi(argv[14]) != 0;
    const char* result_path = argv[15];

    const auto t0 = std::chrono::steady_clock::now();
    float *queries = nullptr, *data = nullptr;
    int64_t n_query = 0, n_data = 0;
    int d = 0, dq = 0;
    if (lsh_read_text_dataset(query_path, &queries, &n_query, &dq, nullptr)) return die("reading queries");
    if (lsh_read_text_dataset(data_path, &data, &n_data, &d, nullptr)) return die("reading data");
    const double read_ms =
        std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
    if (d != dq) {
        std::fprintf(stderr, "data has %d dimensions, queries %d\n", d, dq);
        return 1;
    }
    std::printf("Time to read data files: %.0f \n", read_ms);
    std::printf("N_Query = %lld \nN_Data = %lld \n", (long long)n_query, (long long)n_data);
    std::printf("k is set to: %d\n", k);

    lsh_result_t res;
    int st;
    if (framework == 0)
        st = lsh_linear_scans(implementation, k, d, (int)n_query, (int)n_data, data, queries, sketch_dim,
                              distance_func, &res);
    else
        st = lsh_framework(implementation, keys_implementation, k, d, (int)n_query, (int)n_data, data, queries,
                           sketch_dim, distance_func, bucket_key_bits, tables, with_sketch, &res);
    if (st) return die("search");

    if (run_validation) {
        if (lsh_validate(truth_path, res.results, (int)n_query, k, report_k, &res.recall, &res.avgDistance))
            return die("validation");
        std::printf("Averge recall: %g\nTotal average: %g\n", res.recall, res.avgDistance);
    }
    if (write_res && lsh_write_results(nullptr, res.results, k, (int)n_query, report_k)) return die("writing results");
    if (lsh_write_performance(result_path, &res, implementation, keys_implementation, sketch_dim, k, bucket_key_bits,
                              tables))
        return die("writing the performance line");
    lsh_free(queries);
    lsh_free(data);
    lsh_result_free(&res);
    std::printf("Success. Program exiting. \n");
    return 0;
}
