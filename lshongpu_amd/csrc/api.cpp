// api.cpp -- the C ABI of include/lshgpu.h.  Every entry point converts
// failures into status codes (the reference prints and calls exit(-1),
// cudaHelpers.cuh:8-15) and records the message for lsh_last_error().
#include <chrono>
#include <cmath>
#include <cstdlib>
#include <new>

#include "engine.hpp"

namespace {

using lshgpu::guarded;

double now_ms() {
    return std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now().time_since_epoch())
        .count();
}

void check_counts(int d, int64_t N_query, int64_t N_data, const float* data, const float* queries) {
    LSH_REQUIRE(d >= 1, LSH_ERR_INVALID, "d must be >= 1");
    LSH_REQUIRE(N_query >= 0 && N_data >= 0, LSH_ERR_INVALID, "N_query and N_data must be >= 0");
    LSH_REQUIRE(N_data < 0x7FFFFFFF, LSH_ERR_INVALID, "N_data must fit a 32-bit point ID");
    LSH_REQUIRE(N_data == 0 || data, LSH_ERR_INVALID, "data is null");
    LSH_REQUIRE(N_query == 0 || queries, LSH_ERR_INVALID, "queries is null");
}

void configure(lsh_index& ix, const lsh_index_config_t& c) {
    LSH_REQUIRE(c.framework == 0 || c.framework == 1, LSH_ERR_INVALID, "framework must be 0 or 1");
    if (c.framework == 0) {
        LSH_REQUIRE(c.implementation >= 1 && c.implementation <= 6, LSH_ERR_INVALID,
                    "Invalid implementation selected (1..6)");
    } else {
        LSH_REQUIRE(c.implementation >= 2 && c.implementation <= 6, LSH_ERR_INVALID,
                    "Invalid implementation selected for LSH (2..6)");
        LSH_REQUIRE(c.keysImplementation == 3 || c.keysImplementation == 4 || c.keysImplementation == 5 ||
                        c.keysImplementation == 7 || c.keysImplementation == 8,
                    LSH_ERR_INVALID, "Invalid keys implementation selected for LSH (3, 4, 5, 7, 8)");
        LSH_REQUIRE(c.tables >= 1, LSH_ERR_INVALID, "tables must be >= 1");
        LSH_REQUIRE(c.bucketKeyBits >= 1 && c.bucketKeyBits <= 32, LSH_ERR_INVALID,
                    "bucketKeyBits must be in 1..32");
    }
    if (c.implementation == 2) {
        LSH_REQUIRE(c.distanceFunc >= 1 && c.distanceFunc <= 3, LSH_ERR_INVALID,
                    "Invalid operation selected for distance function (1..3)");
    }
    if (c.implementation >= 3) LSH_REQUIRE(c.sketchDim >= 1, LSH_ERR_INVALID, "sketchDim must be >= 1");
    ix.impl = c.implementation;
    ix.dfunc = c.distanceFunc;
    ix.d = c.d;
    ix.N = c.N_data;
    ix.sketchDim = c.sketchDim;
    ix.framework = c.framework;
    ix.keysImpl = c.keysImplementation;
    ix.bucketKeyBits = c.bucketKeyBits;
    ix.tables = c.tables;
    ix.withSketch = c.runWithSketchedData != 0;
    ix.candidates = c.candidates > 0 ? c.candidates : 4096;
    ix.margin = c.rescan_margin > 0 ? c.rescan_margin : 32;
    ix.device = c.device;
    LSH_REQUIRE(c.projection == 0 || c.projection == 1, LSH_ERR_INVALID, "projection must be 0 or 1");
    ix.projection = c.projection;
    LSH_REQUIRE(ix.candidates <= lshgpu::kMaxTopK, LSH_ERR_UNSUPPORTED, "candidates must be <= 4096");
}

struct StreamGuard {
    hipStream_t s = nullptr;
    ~StreamGuard() {
        if (s) (void)hipStreamDestroy(s);
    }
};

int one_shot(const lsh_index_config_t& cfg, int k, int N_query, const float* data, const float* queries,
             lsh_result_t* out) {
    return guarded([&] {
        LSH_REQUIRE(out, LSH_ERR_INVALID, "out is null");
        out->results = nullptr;
        out->preprocessTime = out->constructionTime = out->scanTime = 0;
        out->recall = NAN;
        out->avgDistance = NAN;
        check_counts(cfg.d, N_query, cfg.N_data, data, queries);
        LSH_REQUIRE(k >= 1, LSH_ERR_INVALID, "k must be >= 1");
        lsh_index ix;
        configure(ix, cfg);
        LSH_HIP(hipSetDevice(ix.device));
        StreamGuard sg;
        LSH_HIP(hipStreamCreateWithFlags(&sg.s, hipStreamNonBlocking));
        ix.stream = sg.s;
        const int kq = lshgpu::calculate_k(k);
        double t0 = now_ms();
        // The reference derives the generalized-Jaccard shift and the weighted
        // MinHash bounds from data and queries together
        // (processingUtils.cuh:19-29, weightedMinHash.cuh:47-57).
        ix.build(data, false, queries, N_query);
        double t1 = now_ms();
        lsh_point_t* res = (lsh_point_t*)std::malloc(sizeof(lsh_point_t) * (size_t)std::max(N_query, 1) * kq);
        LSH_REQUIRE(res, LSH_ERR_NOMEM, "result allocation failed");
        try {
            ix.search(queries, N_query, k, false, res, false, nullptr);
        } catch (...) {
            std::free(res);
            throw;
        }
        double t2 = now_ms();
        out->results = res;
        out->preprocessTime = 0;
        out->constructionTime = (long)std::llround(t1 - t0);
        out->scanTime = (long)std::llround(t2 - t1);
    });
}

}  // namespace

extern "C" {

int lsh_calculate_k(int k) { return lshgpu::calculate_k(k); }

void lsh_options_default(lsh_options_t* opt) {
    if (!opt) return;
    std::memset(opt, 0, sizeof(*opt));
    opt->candidates = 4096;
    opt->device = 0;
    opt->rescan_margin = 32;
}

void lsh_index_config_default(lsh_index_config_t* cfg) {
    if (!cfg) return;
    std::memset(cfg, 0, sizeof(*cfg));
    cfg->implementation = 2;
    cfg->distanceFunc = 1;
    cfg->candidates = 4096;
    cfg->rescan_margin = 32;
}

const char* lsh_last_error(void) { return lshgpu::last_error().c_str(); }

const char* lsh_version(void) { return "lshgpu 0.1.0 gfx950"; }

int lsh_linear_scans_ex(int implementation, int k, int d, int N_query, int N_data, const float* data,
                        const float* queries, int sketchDim, int distanceFunc, const lsh_options_t* opt,
                        lsh_result_t* out) {
    lsh_options_t o;
    if (opt)
        o = *opt;
    else
        lsh_options_default(&o);
    lsh_index_config_t cfg;
    lsh_index_config_default(&cfg);
    cfg.implementation = implementation;
    cfg.distanceFunc = distanceFunc;
    cfg.d = d;
    cfg.N_data = N_data;
    cfg.sketchDim = sketchDim;
    cfg.framework = 0;
    cfg.candidates = o.candidates;
    cfg.device = o.device;
    cfg.rescan_margin = o.rescan_margin;
    return one_shot(cfg, k, N_query, data, queries, out);
}

int lsh_linear_scans(int implementation, int k, int d, int N_query, int N_data, const float* data,
                     const float* queries, int sketchDim, int distanceFunc, lsh_result_t* out) {
    return lsh_linear_scans_ex(implementation, k, d, N_query, N_data, data, queries, sketchDim, distanceFunc,
                               nullptr, out);
}

int lsh_framework_ex(int implementation, int keysImplementation, int k, int d, int N_query, int N_data,
                     const float* data, const float* queries, int sketchDim, int distanceFunc,
                     int bucketKeyBits, int tables, int runWithSketchedData, const lsh_options_t* opt,
                     lsh_result_t* out) {
    lsh_options_t o;
    if (opt)
        o = *opt;
    else
        lsh_options_default(&o);
    lsh_index_config_t cfg;
    lsh_index_config_default(&cfg);
    cfg.implementation = implementation;
    cfg.keysImplementation = keysImplementation;
    cfg.distanceFunc = distanceFunc;
    cfg.d = d;
    cfg.N_data = N_data;
    cfg.sketchDim = sketchDim;
    cfg.framework = 1;
    cfg.bucketKeyBits = bucketKeyBits;
    cfg.tables = tables;
    cfg.runWithSketchedData = runWithSketchedData;
    cfg.candidates = o.candidates;
    cfg.device = o.device;
    cfg.rescan_margin = o.rescan_margin;
    return one_shot(cfg, k, N_query, data, queries, out);
}

int lsh_framework(int implementation, int keysImplementation, int k, int d, int N_query, int N_data,
                  const float* data, const float* queries, int sketchDim, int distanceFunc, int bucketKeyBits,
                  int tables, int runWithSketchedData, lsh_result_t* out) {
    return lsh_framework_ex(implementation, keysImplementation, k, d, N_query, N_data, data, queries, sketchDim,
                            distanceFunc, bucketKeyBits, tables, runWithSketchedData, nullptr, out);
}

void lsh_result_free(lsh_result_t* r) {
    if (!r) return;
    std::free(r->results);
    r->results = nullptr;
}

int lsh_random_normals(float* out, int64_t n, double stddev) {
    return guarded([&] {
        LSH_REQUIRE(out || n == 0, LSH_ERR_INVALID, "out is null");
        std::vector<float> v = lshgpu::normal_stream(n, stddev);
        std::memcpy(out, v.data(), sizeof(float) * (size_t)n);
    });
}

int lsh_random_signs(int32_t* out, int64_t n) {
    return guarded([&] {
        LSH_REQUIRE(out || n == 0, LSH_ERR_INVALID, "out is null");
        std::vector<int32_t> v = lshgpu::sign_stream(n);
        std::memcpy(out, v.data(), sizeof(int32_t) * (size_t)n);
    });
}

int lsh_simhash_sketch(const float* x, int64_t n, int d, int words, int hashBits, const float* planes,
                       uint32_t* out) {
    return guarded([&] {
        LSH_REQUIRE((x && out) || n == 0, LSH_ERR_INVALID, "null pointer");
        if (n == 0) return;
        lshgpu::simhash_host(x, n, d, words, hashBits, planes, out, 0);
    });
}

int lsh_jl_sketch(const float* x, int64_t n, int d, int sketchDim, const float* planes, float* out) {
    return guarded([&] {
        LSH_REQUIRE((x && out) || n == 0, LSH_ERR_INVALID, "null pointer");
        if (n == 0) return;
        lshgpu::jl_host(x, n, d, sketchDim, planes, out, 0);
    });
}

int lsh_gj_transform(float* data, int64_t N_data, float* queries, int64_t N_query, int d) {
    return guarded([&] {
        check_counts(d, N_query, N_data, data, queries);
        if (N_data == 0 && N_query == 0) return;
        lshgpu::gj_transform_host(data, N_data, queries, N_query, d, 0);
    });
}

int lsh_hamming_topc(const uint32_t* data_sketch, int64_t N_data, const uint32_t* query_sketch, int64_t N_query,
                     int words, int C, int32_t* ids, int32_t* distances) {
    return guarded([&] {
        LSH_REQUIRE((data_sketch || N_data == 0) && (query_sketch || N_query == 0) && (ids || N_query == 0),
                    LSH_ERR_INVALID, "null pointer");
        lshgpu::hamming_topc_host(data_sketch, N_data, query_sketch, N_query, words, C, ids, distances, 0);
    });
}

int lsh_hamming_topc_noalloc(const uint32_t* data_sketch, int64_t N_data, const uint32_t* query_sketch,
                             int64_t N_query, int words, int C, int32_t* ids, int32_t* distances, void* stream) {
    return guarded([&] {
        lshgpu::hamming_topc_device_noalloc(data_sketch, N_data, query_sketch, N_query, words, C, ids, distances,
                                            (hipStream_t)stream);
    });
}

int lsh_wmh_sketch(const float* x, int64_t n, int d, int sketchDim, int oneBit, uint32_t* out) {
    return guarded([&] {
        LSH_REQUIRE((x && out) || n == 0, LSH_ERR_INVALID, "null pointer");
        if (n == 0) return;
        lshgpu::wmh_host(x, n, d, sketchDim, oneBit, out, 0);
    });
}

int lsh_crosspoly_keys(const float* x, int64_t n, int d, int tables, const int32_t* signs, uint16_t* keys) {
    return guarded([&] {
        LSH_REQUIRE((x && keys) || n == 0, LSH_ERR_INVALID, "null pointer");
        LSH_REQUIRE(d >= 1 && tables >= 1, LSH_ERR_INVALID, "d and tables must be >= 1");
        if (n == 0) return;
        lshgpu::crosspoly_host(x, n, d, tables, signs, keys, 0);
    });
}

int lsh_crosspoly_dense_keys(const float* x, int64_t n, int d, int tables, int K, uint32_t* keys) {
    return guarded([&] {
        LSH_REQUIRE((x && keys) || n == 0, LSH_ERR_INVALID, "null pointer");
        lshgpu::crosspoly_dense_host(x, n, d, tables, K, keys, 0);
    });
}

int lsh_sort_table_keys(const uint32_t* keys, int64_t N, int tables, int bits, uint32_t* sorted_keys,
                        uint32_t* ids) {
    return guarded([&] {
        LSH_REQUIRE((keys && sorted_keys && ids) || N == 0, LSH_ERR_INVALID, "null pointer");
        LSH_REQUIRE((int64_t)tables * N < 0xFFFFFFFFll, LSH_ERR_UNSUPPORTED, "tables * N must be below 2^32");
        lshgpu::sort_tables_host(keys, N, tables, bits, sorted_keys, ids, 0);
    });
}

int lsh_bucket_distribution(const uint8_t* hashes, int64_t n, int64_t* counts256) {
    return guarded([&] {
        LSH_REQUIRE(counts256 && (hashes || n == 0) && n >= 0, LSH_ERR_INVALID, "null pointer or negative size");
        for (int i = 0; i < 256; ++i) counts256[i] = 0;
        if (n == 0) return;
        lshgpu::byte_distribution_host(hashes, n, counts256, 0);
    });
}

int lsh_bucket_distribution_full_key(const uint8_t* hashes, int64_t n, int sketchDim, uint8_t** keys,
                                     int64_t** counts, int64_t* n_distinct) {
    return guarded([&] {
        LSH_REQUIRE(keys && counts && n_distinct && (hashes || n == 0), LSH_ERR_INVALID, "null pointer");
        *keys = nullptr;
        *counts = nullptr;
        *n_distinct = 0;
        LSH_REQUIRE(sketchDim >= 1, LSH_ERR_INVALID, "sketchDim must be >= 1");
        LSH_REQUIRE(n >= 0 && n < (int64_t(1) << 31), LSH_ERR_INVALID, "n must be in [0, 2^31)");
        if (n == 0) return;
        std::vector<uint8_t> k;
        std::vector<int64_t> c;
        lshgpu::full_key_distribution_host(hashes, n, sketchDim, k, c, 0);
        auto* kp = static_cast<uint8_t*>(std::malloc(std::max<size_t>(k.size(), 1)));
        auto* cp = static_cast<int64_t*>(std::malloc(std::max<size_t>(c.size(), 1) * sizeof(int64_t)));
        if (!kp || !cp) {
            std::free(kp);
            std::free(cp);
            throw std::bad_alloc();
        }
        std::copy(k.begin(), k.end(), kp);
        std::copy(c.begin(), c.end(), cp);
        *keys = kp;
        *counts = cp;
        *n_distinct = (int64_t)c.size();
    });
}

int lsh_index_create(const lsh_index_config_t* cfg, const float* data, int data_on_device, lsh_index_t** out) {
    return guarded([&] {
        LSH_REQUIRE(cfg && out, LSH_ERR_INVALID, "null pointer");
        *out = nullptr;
        check_counts(cfg->d, 0, cfg->N_data, data, nullptr);
        auto* ix = new lsh_index();
        try {
            configure(*ix, *cfg);
            LSH_HIP(hipSetDevice(ix->device));
            LSH_HIP(hipStreamCreateWithFlags(&ix->stream, hipStreamNonBlocking));
            ix->build(data, data_on_device != 0, nullptr, 0);
        } catch (...) {
            if (ix->stream) (void)hipStreamDestroy(ix->stream);
            delete ix;
            throw;
        }
        *out = ix;
    });
}

int lsh_index_search(lsh_index_t* index, const float* queries, int N_query, int k, int queries_on_device,
                     lsh_point_t* results, int results_on_device, void* stream) {
    return guarded([&] {
        LSH_REQUIRE(index, LSH_ERR_INVALID, "index is null");
        LSH_REQUIRE(N_query == 0 || (queries && results), LSH_ERR_INVALID, "null pointer");
        std::lock_guard<std::mutex> lk(index->mu);
        index->search(queries, N_query, k, queries_on_device != 0, results, results_on_device != 0,
                      (hipStream_t)stream);
    });
}

int lsh_index_search_batches(lsh_index_t* index, const float* queries, const int* batch_sizes, int n_batches,
                             int k, int queries_on_device, lsh_point_t* results, int results_on_device,
                             void* stream) {
    return guarded([&] {
        LSH_REQUIRE(index, LSH_ERR_INVALID, "index is null");
        LSH_REQUIRE(n_batches == 0 || (queries && results && batch_sizes), LSH_ERR_INVALID, "null pointer");
        std::lock_guard<std::mutex> lk(index->mu);
        index->search_batches(queries, batch_sizes, n_batches, k, queries_on_device != 0, results,
                              results_on_device != 0, (hipStream_t)stream);
    });
}

int lsh_index_set_profiling(lsh_index_t* index, int on) {
    return guarded([&] {
        LSH_REQUIRE(index, LSH_ERR_INVALID, "index is null");
        LSH_REQUIRE(on >= 0 && on <= 3, LSH_ERR_INVALID, "profiling mode must be 0, 1, 2 or 3");
        std::lock_guard<std::mutex> lk(index->mu);
        // Mode 2 times only the rerank / finalisation phase (index 3), mode 3
        // only the pivot-cell search's appending pass (index 7).
        index->timer.enable(on == 0 ? 0u : on == 1 ? ~0u : on == 2 ? (1u << 3) : (1u << 7));
        index->prof_mfma_launches = 0;
        index->prof_mfma_points = 0;
        index->prof_pass2_tiles = 0;
    });
}

int lsh_index_phase_times_n(lsh_index_t* index, double* out, int n) {
    return guarded([&] {
        LSH_REQUIRE(index && out && n >= 0, LSH_ERR_INVALID, "null pointer or negative count");
        std::lock_guard<std::mutex> lk(index->mu);
        double ms[8];
        index->timer.collect(ms, 8);
        const double all[11] = {ms[0], ms[1], ms[2], ms[3], ms[4], ms[5], (double)index->prof_mfma_launches,
                                (double)index->prof_mfma_points, ms[6], ms[7], (double)index->prof_pass2_tiles};
        for (int i = 0; i < n; ++i) out[i] = i < 11 ? all[i] : 0.0;
        index->prof_mfma_launches = 0;
        index->prof_mfma_points = 0;
        index->prof_pass2_tiles = 0;
    });
}

int lsh_index_batch_stats(lsh_index_t* index, int64_t* graph_batches, int64_t* redone_queries) {
    return guarded([&] {
        LSH_REQUIRE(index && graph_batches && redone_queries, LSH_ERR_INVALID, "null pointer");
        std::lock_guard<std::mutex> lk(index->mu);
        *graph_batches = index->batch_graphs;
        *redone_queries = index->batch_redone;
    });
}

int lsh_index_phase_times(lsh_index_t* index, double* out8) { return lsh_index_phase_times_n(index, out8, 8); }

int lsh_index_bucket_stats(lsh_index_t* index, int table, lsh_bucket_stats_t* out, int64_t* hist, int hist_len) {
    return guarded([&] {
        LSH_REQUIRE(index && out, LSH_ERR_INVALID, "null pointer");
        std::lock_guard<std::mutex> lk(index->mu);
        int64_t s4[4];
        lshgpu::index_bucket_stats(index, table, s4, hist, hist_len);
        out->buckets = s4[0];
        out->nonempty = s4[1];
        out->points = s4[2];
        out->max_size = s4[3];
    });
}

int64_t lsh_index_device_bytes(const lsh_index_t* index) {
    if (!index) return -1;
    std::lock_guard<std::mutex> lk(index->mu);
    return (int64_t)index->device_bytes();
}

int lsh_index_destroy(lsh_index_t* index) {
    return guarded([&] {
        if (!index) return;
        { std::lock_guard<std::mutex> lk(index->mu); }  // no call is inside the index any more
        int prev = 0;
        (void)hipGetDevice(&prev);
        (void)hipSetDevice(index->device);
        hipStream_t s = index->stream;
        if (s) (void)hipStreamSynchronize(s);
        // A search on a caller's stream may still be running.
        if (index->done_ev) {
            (void)hipEventSynchronize(index->done_ev);
            (void)hipEventDestroy(index->done_ev);
        }
        delete index;
        if (s) (void)hipStreamDestroy(s);
        (void)hipSetDevice(prev);
    });
}

}  // extern "C"
