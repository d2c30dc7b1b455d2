/*
 * lshgpu.h -- C ABI of the MI355X-native k-NN search engine.
 *
 * The first block mirrors the reference's launch/result surface
 * (LinearScan/kernel.cu, point.h, resultDTO.h, launchHelper.cuh) so a caller
 * of `linearScans` / `LSH` can switch by renaming the call.  Inputs are host
 * float32 row-major arrays that are never modified; results are allocated by
 * the library with malloc (the reference's main() frees them with free(),
 * kernel.cu:191).  Every entry point returns a status code (0 = success) and
 * never exits the process; lsh_last_error() gives the message of the last
 * failure on the calling thread.
 *
 * The second block exposes the per-stage kernels of the path (sketch
 * builders, key builders, sketch-distance top-C) so that the integer stages
 * can be checked bit for bit.  The third block is the resident index used for
 * batched search and serving; it has no counterpart in the reference, whose
 * every call resets the device (cudaHelpers.cu:12-14).
 */
#ifndef LSHGPU_H
#define LSHGPU_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ------------------------------------------------------------------ types */

/* point.h:3-7 -- {int ID; float distance;}, 8 bytes. */
typedef struct lsh_point {
    int32_t ID;
    float distance;
} lsh_point_t;

/* resultDTO.h:5-32 on LP64 (clock_t is long).  The three times are in
 * milliseconds, as the reference stores them (time_lapsed*1000/CLOCKS_PER_SEC);
 * here they are host wall-clock times around the device work (the index
 * build, then the search).  recall and avgDistance are NaN until a
 * validation step fills them. */
typedef struct lsh_result {
    lsh_point_t* results;   /* [N_query][k], k already rounded by lsh_calculate_k */
    long preprocessTime;
    long constructionTime;
    long scanTime;
    float recall;
    float avgDistance;
} lsh_result_t;

/* status codes */
enum {
    LSH_OK = 0,
    LSH_ERR_INVALID = 1,     /* bad id, size or argument */
    LSH_ERR_UNSUPPORTED = 2, /* valid combination the engine does not provide */
    LSH_ERR_DEVICE = 3,      /* HIP runtime error */
    LSH_ERR_NOMEM = 4,       /* host or device allocation failed / over budget */
    LSH_ERR_INTERNAL = 5,
    LSH_ERR_IO = 6           /* file cannot be opened, read or written */
};

/* Optional knobs of the _ex entry points.  lsh_options_default() fills the
 * values the plain entry points use. */
typedef struct lsh_options {
    int candidates;     /* C: size of the sketch-distance candidate set that is
                           reranked exactly.  Default 4096 = 32 lanes x
                           THREAD_QUEUE_SIZE 128, the width of the reference's
                           rerank queue (constants.cuh:6, candidateSetScanner.cuh:21). */
    int device;         /* HIP device ordinal, default 0 (cudaHelpers.cu:16-18) */
    int rescan_margin;  /* exact scans: extra FP32 candidates kept past k for the
                           FP64 recheck (default 32).  The sketch paths' rerank
                           rechecks in FP64 every candidate within the error
                           band above the k-th approximate distance instead. */
    int reserved[8];
} lsh_options_t;

/* -------------------------------------------------- reference surface */

/* launchHelper.cuh:21-25 calculateK: round k up to a multiple of 32. */
int lsh_calculate_k(int k);

void lsh_options_default(lsh_options_t* opt);

/* kernel.cu:27 linearScans(implementation, k, d, N_query, N_data, data,
 * queries, sketchDim, distanceFunc).
 *   implementation 1: exact scan, distance -cos, padding (-1, 2.0f)
 *                    (optimizedLinearScan.cu:22-91)
 *   implementation 2: exact scan with distanceFunc 1 angular, 2 generalized
 *                    Jaccard (after shift+normalise), 3 Euclidean
 *                    (memOptimizedLinearScan.cu:129-191)
 *   implementation 3: SimHash, 32*sketchDim bits, Hamming top-C, angular rerank
 *                    (simHash.cu:61-144)
 *   implementation 4: weighted MinHash, sketchDim hashes, match-count top-C,
 *                    generalized-Jaccard rerank (weightedMinHash.cuh:344-429)
 *   implementation 5: 1-bit weighted MinHash, 32*sketchDim bits, Hamming top-C,
 *                    generalized-Jaccard rerank
 *   implementation 6: SimHash-JL, sketchDim floats, squared-L2 top-C, angular
 *                    rerank (simHashJL.cuh:238-281)
 * k is rounded with lsh_calculate_k; out->results is [N_query][k] ascending by
 * (distance, ID), padded with (-1, 2147483648.0f). */
int lsh_linear_scans(int implementation, int k, int d, int N_query, int N_data,
                     const float* data, const float* queries, int sketchDim,
                     int distanceFunc, lsh_result_t* out);

/* kernel.cu:91 LSH(...): (K,L) bucket framework.  keysImplementation 3 SimHash
 * (bucketKeyBits bits per table, 16-bit key type), 4 weighted MinHash red
 * count (8-bit key type; needs bucketKeyBits >= 8), 5 1-bit weighted MinHash
 * (bucketKeyBits bits, 16-bit key type), 7 cross-polytope (one hash per
 * table; needs 2^bucketKeyBits >= 2*P, P = d rounded up to a power of two,
 * d <= 256), 8 (no reference counterpart) dense-rotation cross-polytope,
 * K = bucketKeyBits / log2(2P) hashes per table in a 32-bit key (see
 * lsh_crosspoly_dense_keys; bucketKeyBits a multiple of log2(2P), d <= 128);
 * bucketKeyBits 1..32.  Tables of up to 16 key bits are 2^bits bucket
 * offsets as the reference builds them; wider keys and keysImplementation 8
 * keep each table as (key, ID) pairs sorted on the device, looked up by
 * binary search.  Each key family preprocesses the rows
 * first (4, 5: shift + normalise; 7: normalise), then the sketch family of
 * `implementation` does (4, 5: shift + normalise; 6: normalise), in that
 * order (lshFramework.cuh:330, 391).  The candidate set of a query is the
 * union of its L buckets without duplicates.  With runWithSketchedData and
 * implementation 3..6 it is cut to the top-C by sketch distance (ID ties)
 * and reranked exactly by distanceFunc; otherwise all of it is ranked by
 * distanceFunc (lshFramework.cuh:142-280).  Rows with fewer candidates than
 * k are padded with (-1, 2147483648.0f). */
int lsh_framework(int implementation, int keysImplementation, int k, int d,
                  int N_query, int N_data, const float* data,
                  const float* queries, int sketchDim, int distanceFunc,
                  int bucketKeyBits, int tables, int runWithSketchedData,
                  lsh_result_t* out);

int lsh_linear_scans_ex(int implementation, int k, int d, int N_query,
                        int N_data, const float* data, const float* queries,
                        int sketchDim, int distanceFunc,
                        const lsh_options_t* opt, lsh_result_t* out);

int lsh_framework_ex(int implementation, int keysImplementation, int k, int d,
                     int N_query, int N_data, const float* data,
                     const float* queries, int sketchDim, int distanceFunc,
                     int bucketKeyBits, int tables, int runWithSketchedData,
                     const lsh_options_t* opt, lsh_result_t* out);

/* Frees out->results (allocated with malloc) and clears the pointer. */
void lsh_result_free(lsh_result_t* r);

/* Message of the last failed call on this thread ("" if none). */
const char* lsh_last_error(void);

/* Library version string. */
const char* lsh_version(void);

/* ------------------------------------------------------ stage entry points
 * Host pointers in, host pointers out; each runs its kernels on the device. */

/* randomVectorGenerator.cpp:32-53 generateRandomVectors(float*, N): N draws of
 * std::normal_distribution<double>(0, 1) on a default-seeded
 * std::default_random_engine, cast to float.  stddev != 1 gives the
 * (N, sketchDim) variant of randomVectorGenerator.cpp:56-74 with
 * stddev = 1/sketchDim. */
int lsh_random_normals(float* out, int64_t n, double stddev);

/* randomVectorGenerator.cpp:76-87 generateRandomOnePlusMinusVector. */
int lsh_random_signs(int32_t* out, int64_t n);

/* simHash.cuh:20-46 sketchDataGeneric / simHash.cu:22-47 sketch on the GPU:
 * word w, bit b of point i is dot(x_i, planes[(w*hashBits+b)*d ...]) >= 0 with
 * the dot product a sequential FP32 fused multiply-add chain over dimensions
 * in index order; bits are packed LSB first.  planes == NULL uses the
 * reference stream (lsh_random_normals with stddev 1). */
int lsh_simhash_sketch(const float* x, int64_t n, int d, int words,
                       int hashBits, const float* planes, uint32_t* out);

/* simHashJL.cuh:34-51: float JL projections with planes drawn from
 * N(0, 1/sketchDim) (planes == NULL uses the reference stream). */
int lsh_jl_sketch(const float* x, int64_t n, int d, int sketchDim,
                  const float* planes, float* out);

/* processingUtils.cuh:12-67 as used by memOptimizedLinearScan.cu:149-163:
 * shift every dimension by |min(0, floor(min over data and queries))|, then
 * scale every row to unit L2 norm (FP32, fused multiply-add norm).  In place. */
int lsh_gj_transform(float* data, int64_t N_data, float* queries,
                     int64_t N_query, int d);

/* Top-C candidate ids per query by (Hamming distance over `words` packed
 * words, id), ascending.  distances may be NULL. */
int lsh_hamming_topc(const uint32_t* data_sketch, int64_t N_data,
                     const uint32_t* query_sketch, int64_t N_query, int words,
                     int C, int32_t* ids, int32_t* distances);

/* No counterpart in the reference.  lsh_hamming_topc for callers that cannot
 * allocate device memory (code under graph capture, tight serving loops,
 * harnesses that forbid hipMalloc): the same top-C (IDs and distances, equal
 * to lsh_hamming_topc's), but every pointer is a device pointer on the
 * calling thread's current device, the work is queued on `stream` (a
 * hipStream_t, NULL = the null stream) and the call returns without waiting.
 * It allocates nothing: its working memory (the query sketches padded to 8
 * words and 256 rows, and a top-C state of 8*(C + capacity) + 16 bytes per
 * query, capacity = 2C + 256 for C <= 1024, else C + 512) comes from one
 * 32 MiB array per GPU that the library holds from its first use until the
 * process ends.  Calls that need more return LSH_ERR_NOMEM.  Every call on a
 * GPU shares that array: calls that overlap on one GPU (other streams or
 * threads) race, so the caller orders them.  C <= 4096 (C <= 2048 keeps an
 * unordered best list; larger C is rounded up to a power of two inside). */
int lsh_hamming_topc_noalloc(const uint32_t* data_sketch, int64_t N_data, const uint32_t* query_sketch,
                             int64_t N_query, int words, int C, int32_t* ids, int32_t* distances, void* stream);

/* weightedMinHash.cuh:41-166 on rows that are already shifted and
 * normalised: bounds from x, then per hash h (seed h*1234 + 92138) the number
 * of red draws before the first green one.  oneBit == 0: out is [n][sketchDim]
 * counts; oneBit != 0: out is [n][sketchDim] words whose bit b of word w is
 * bitmap[count of hash w*32+b] (sketchDataOneBit).  The draws come from
 * Philox4x32-10 (key (seed, 0x5EED5EED), counter t/4, word t%4, U = (word>>8)
 * * 2^-24), not the reference's cuRAND XORWOW stream; rows with no green draw
 * in 2^16 draws record 2^16. */
int lsh_wmh_sketch(const float* x, int64_t n, int d, int sketchDim, int oneBit,
                   uint32_t* out);

/* crossPolyTope.cuh:25-85 generateHashCrossPoly: pad to P = next power of two
 * >= d, three rounds of (+-1 diagonal, unnormalised Walsh-Hadamard
 * butterfly), key = argmax |v| encoded as i (negative) or i + P.  Rows are
 * L2-normalised first (crossPolyTope.cuh:17-21).  signs == NULL uses the
 * reference stream (lsh_random_signs over tables*3*P values). */
int lsh_crosspoly_keys(const float* x, int64_t n, int d, int tables,
                       const int32_t* signs, uint16_t* keys);

/* No counterpart in the reference (keysImplementation 8): dense-rotation
 * cross-polytope keys.  Rows are L2-normalised, then hash h = t*K + k of row
 * x is the index of the largest |y_j| of y = G_h x (the first on ties),
 * encoded j when y_j < 0 and j + P otherwise, as crossPolyTope.cuh:64-76
 * encodes its HD^3 hash; y_j is the sequential FP32 fused multiply-add chain
 * over dimensions in index order.  G_h is P x d (P = d rounded up to a power
 * of two), its entries the reference normal stream (lsh_random_normals with
 * stddev 1, tables*K*P*d values, row-major [h][j][i]) rounded to bfloat16.
 * Key t packs hashes t*K .. t*K+K-1 into log2(2P)-bit fields, hash k at bit
 * k*log2(2P).  keys is [n][tables].  d <= 128, K*log2(2P) <= 32. */
int lsh_crosspoly_dense_keys(const float* x, int64_t n, int d, int tables, int K, uint32_t* keys);

/* The bucket-table build of keys wider than 16 bits (and of keysImplementation
 * 8), which replaces the 2^bits offsets of lshFramework.cuh:102-140: per table
 * t, keys[t][0..N) (the low `bits` bits; 1..32) sorted ascending with their
 * point IDs, equal keys in ascending ID order (a stable sort).  keys,
 * sorted_keys and ids are [tables][N]; tables*N < 2^32. */
int lsh_sort_table_keys(const uint32_t* keys, int64_t N, int tables, int bits, uint32_t* sorted_keys,
                        uint32_t* ids);

/* statistics.cuh:10-16 bucketDistribution, as weightedMinHash.cuh:215-229
 * runBucketStatistics runs it over byte sketches: counts256[v] = the number
 * of the n bytes of `hashes` equal to v, for every v in 0..255 (the
 * reference's 255-entry array has no slot for 255). */
int lsh_bucket_distribution(const uint8_t* hashes, int64_t n, int64_t* counts256);

/* statisticsCpu.cpp:11-28 bucketDistributionFullKey: the distinct keys among
 * n keys of sketchDim bytes each (`hashes` is [n][sketchDim]) with the number
 * of keys equal to each, in byte-lexicographic key order (the order of the
 * reference's std::map over the keys' bit strings).  *keys
 * ([n_distinct][sketchDim]) and *counts ([n_distinct]) are allocated with
 * malloc; release them with lsh_free.  n < 2^31. */
int lsh_bucket_distribution_full_key(const uint8_t* hashes, int64_t n, int sketchDim,
                                     uint8_t** keys, int64_t** counts,
                                     int64_t* n_distinct);

/* ------------------------------------- datasets, truth files and reports
 * The runner's I/O (kernel.cu:119-196): host code, no device needed.  Arrays
 * returned through pointer arguments are allocated with malloc; release them
 * with lsh_free. */

void lsh_free(void* p);

/* gloveparser.cu:9-86 parseFile, the format DataParsing/Writer/writer.cpp:
 * 10-29 writes: "N\nd\n", then one row per line "ID v1 ... vd".  Values are
 * read as C floats (decimal or exponent form; the reference's digit loop
 * keeps only the last digit of the integer part and drops a row's last value
 * when no space follows it).  *data is [N][d]; *ids (may be NULL) receives
 * the row IDs, which the search does not use. */
int lsh_read_text_dataset(const char* path, float** data, int64_t* n, int* d,
                          int32_t** ids);

/* DataParsing/Parsers/siftParser.cpp:8-28: .fvecs records (int32 d, then d
 * float32), all of one dimension. */
int lsh_read_fvecs(const char* path, float** data, int64_t* n, int* d);

/* DataParsing/Writer/writer.cpp:10-29: N, d, then "i v1 ... vd " per row
 * (IDs 0..N-1), values with 9 significant digits so they read back exactly. */
int lsh_write_text_dataset(const char* path, const float* data, int64_t n,
                           int d);

/* Truth files, GoldStandard/goldstandard.cpp:123-134 (writer) and
 * LinearScan/validation.cpp:62-107 (reader): "n\nk\n", then per query
 * "qid:" and k lines "ID distance".  lsh_write_truth writes the first k of
 * each result row (rows k_stride apart); query_ids may be NULL (0..n-1). */
int lsh_read_truth(const char* path, int* n_query, int* k, int32_t** ids,
                   float** distances);
int lsh_write_truth(const char* path, const lsh_point_t* results,
                    const int32_t* query_ids, int n_query, int k_stride,
                    int k);

/* validation.cpp:129-208 runValidation(truth, container, results, N_query,
 * k, reportK): recall = mean over the truth file's queries of |first reportK
 * truth IDs found among the first reportK result IDs| / reportK; avgDistance
 * = mean of (sum of the distances of the distinct result IDs, in ID order) /
 * (sum of the first reportK truth distances), in float as the reference
 * sums them.  results is [N_query][k]. */
int lsh_validate(const char* truth_path, const lsh_point_t* results,
                 int N_query, int k, int reportK, float* recall,
                 float* avgDistance);

/* resultWriter.cpp:7-20 writeResult: "N_query\nreportK\n", then "i:" and
 * reportK lines "ID distance" (printf %d %f).  path NULL writes to stdout. */
int lsh_write_results(const char* path, const lsh_point_t* results, int k,
                      int N_query, int reportK);

/* resultWriter.cpp:37-60 writePerformanceResults: appends one line of 13
 * comma-separated columns: implementation, keysImplementation, sketchDim, k,
 * 128 (THREAD_QUEUE_SIZE), bucketKeyBits, tables, 0 (WITH_TQ_OR_BUFFER),
 * preprocessTime, constructionTime, scanTime, recall, avgDistance. */
int lsh_write_performance(const char* path, const lsh_result_t* r,
                          int implementation, int keysImplementation,
                          int sketchDim, int k, int bucketKeyBits,
                          int tables);

/* ---------------------------------------------------------- resident index
 * No counterpart in the reference.  The index keeps the data, its sketches
 * and bucket tables resident in HBM; searches take host or device query
 * pointers and write host or device result arrays. */

typedef struct lsh_index lsh_index_t;

typedef struct lsh_index_config {
    int implementation;      /* as lsh_linear_scans / lsh_framework */
    int distanceFunc;        /* 1, 2, 3 (implementation 2) */
    int d;
    int64_t N_data;
    int sketchDim;
    int framework;           /* 0 linear scan, 1 (K,L) buckets */
    int keysImplementation;
    int bucketKeyBits;
    int tables;
    int runWithSketchedData;
    int candidates;          /* C; 0 = default */
    int device;
    int rescan_margin;       /* 0 = default 32 */
    int projection;          /* SimHash sketches (implementation 3): 0 = bf16 matrix
                                cores with an FP32 recheck of every bit within the
                                error bound (d <= 128; bit-identical), 1 = FP32
                                vector units only */
    int reserved[7];
} lsh_index_config_t;

void lsh_index_config_default(lsh_index_config_t* cfg);

/* data_on_device != 0: `data` is a device pointer on cfg->device. */
int lsh_index_create(const lsh_index_config_t* cfg, const float* data,
                     int data_on_device, lsh_index_t** out);

/* A handle may be used from several threads: the calls on one index are
 * serialised.
 *
 * Search N_query queries for their k nearest (k rounded up to a multiple of
 * 32, as lsh_calculate_k).  results is [N_query][lsh_calculate_k(k)].
 * stream: a hipStream_t (NULL = the index's own stream).  With device results
 * on a given stream the results are complete in that stream's order (the
 * call may return while the last kernels run); otherwise the call returns
 * when the results are complete.  Device queries may be read in place and
 * device results are written in place, so neither buffer may be touched by
 * other work until the search is complete. */
int lsh_index_search(lsh_index_t* index, const float* queries, int N_query,
                     int k, int queries_on_device, lsh_point_t* results,
                     int results_on_device, void* stream);

/* n_batches searches, one per batch: batch i is the next batch_sizes[i] rows
 * of `queries`, and its results the next rows of `results` ([total][k
 * rounded]), as consecutive lsh_index_search calls would give them.  Each
 * batch replays a captured hipGraph of its query-count class (classes about
 * 1.25x apart up to 4096 queries; a larger batch is an ordinary search),
 * with no host wait between batches, on the exact (K,L) rescan of
 * keysImplementation 3, 4, 5 or 8 (without runWithSketchedData; rounded k +
 * rescan margin <= 512) and on the SimHash index (implementation 3) with
 * pivot cells (rounded k <= 128, candidates <= 2048); queries a graph leaves
 * incomplete are redone by ordinary searches before the call returns.  Other
 * index kinds run the batches as ordinary searches.  The call returns when
 * every result is complete (stream: as lsh_index_search). */
int lsh_index_search_batches(lsh_index_t* index, const float* queries, const int* batch_sizes, int n_batches,
                             int k, int queries_on_device, lsh_point_t* results, int results_on_device,
                             void* stream);

/* Diagnostics of the last lsh_index_search_batches call on the index (and of
 * the request loops' batches since): the batches that replayed a captured
 * graph, and the queries redone afterwards by ordinary searches.  No counterpart in the reference (it has no batched
 * search); for callers that check which path their batches took. */
int lsh_index_batch_stats(lsh_index_t* index, int64_t* graph_batches, int64_t* redone_queries);

/* Per-phase device times (ms), summed over the searches since profiling was
 * turned on or since the previous lsh_index_phase_times(_n) call (which waits
 * for those searches and starts a new sum): [0] query preparation and
 * sketch, [1] candidate scans (all rounds), [2] candidate selection, [3]
 * rerank / finalisation (the rerank's filter kernel alone when the FP64
 * recheck runs as its own kernel), [4] total, [5] the matrix-core Hamming
 * scan launches among [1]; [6] the number of those launches and [7] the
 * (query, point) pairs they multiplied per query (summed over launches); [8]
 * the rerank's separate FP64 recheck; [9] the pivot-cell search's appending
 * pass (among [1]) and [10] its tile products (32 queries x 32 points each).
 * Recorded only while profiling is on: on = 1 times every phase, on = 2 only
 * [3] and on = 3 only [9] (two events per search each), 0 turns it off.
 * lsh_index_phase_times fills [0, 8), lsh_index_phase_times_n the first n
 * (entries past [10] read 0). */
int lsh_index_set_profiling(lsh_index_t* index, int on);
int lsh_index_phase_times(lsh_index_t* index, double* out8);
int lsh_index_phase_times_n(lsh_index_t* index, double* out, int n);

/* Bucket occupancy of table `table` (0 <= table < tables) of an index built
 * with framework 1: the figures that guide the choice of bucketKeyBits and
 * tables (statistics.cuh:10-16 counts hash values; weightedMinHash.cuh:
 * 215-229 prints them).  hist (may be NULL with hist_len 0) receives, for
 * s < hist_len - 1, the number of buckets holding exactly s points, and in
 * hist[hist_len - 1] the number holding hist_len - 1 or more. */
typedef struct lsh_bucket_stats {
    int64_t buckets;   /* 2^bucketKeyBits */
    int64_t nonempty;  /* buckets holding at least one point */
    int64_t points;    /* entries of the table (N_data) */
    int64_t max_size;  /* points in the largest bucket */
} lsh_bucket_stats_t;

int lsh_index_bucket_stats(lsh_index_t* index, int table, lsh_bucket_stats_t* out,
                           int64_t* hist, int hist_len);

/* No counterpart in the reference (its run scripts sweep (K,L), sketch sizes
 * and queue lengths by hand, LinearScan/run.sh).  Builds each of the
 * n_candidates configurations over `data` (all with the same d, N_data and
 * device), searches the n_queries host queries with k, and measures recall@k
 * against the exact scan of candidates[0].distanceFunc (1..3; angular
 * otherwise) and the median host time of three searches.  *chosen = the
 * fastest candidate whose recall reaches target_recall, or among the
 * candidates that reach it within 3% of that time the one of highest recall
 * (the most accurate one when none reaches it).  reports (may be NULL) receives one entry per candidate;
 * status is the candidate's build/search status (0 = measured).  The choice
 * is cached for the process per (data, queries, k, target, candidate list). */
typedef struct lsh_tune_report {
    double recall;
    double search_ms;
    double build_ms;
    int64_t device_bytes;
    int status;
} lsh_tune_report_t;

int lsh_autotune(const lsh_index_config_t* candidates, int n_candidates, const float* data, int data_on_device,
                 const float* queries, int n_queries, int k, double target_recall, lsh_tune_report_t* reports,
                 int* chosen);

/* ----------------------------------------------------------------- serving
 * No counterpart in the reference.  A request loop over a resident index:
 * requests (host queries and k) are submitted from any thread and returned
 * through lsh_server_wait.  One worker thread batches every queued request
 * that fits (continuous batching): at most max_batch_queries queries and, for
 * two batches of its size, memory_budget_bytes of device memory (0 = half
 * of the free device memory; one query is budgeted at 2*(4d + 8*128) + 8*(3C +
 * 768) + 4096 bytes, C the index's candidates).  A batch runs as one search
 * at the largest rounded k of its requests; each request receives the first
 * lsh_calculate_k(k) results of each of its rows, which equal those of
 * searching its queries alone.  Up to four batches are in flight.  On the
 * SimHash index with pivot cells (rounded k <= 128, at most 4096 queries a
 * batch) each is enqueued as a replay of the index's captured graphs, with
 * no host wait, so several run on the device together; queries a replay
 * leaves incomplete are redone by an ordinary search before their request
 * is answered.  Otherwise one search runs at a time while the next batch is
 * packed and its queries copied on a second stream.  The index must stay
 * alive until lsh_server_destroy, which serves every request already
 * submitted before it returns; other calls on the index (from other threads)
 * are serialised with the server's. */
typedef struct lsh_server lsh_server_t;

typedef struct lsh_server_config {
    int max_batch_queries;        /* default 4096 */
    int max_query_size;           /* largest request accepted; default 2048 */
    int64_t memory_budget_bytes;  /* 0 = half of the free device memory */
    int reserved[8];
} lsh_server_config_t;

void lsh_server_config_default(lsh_server_config_t* cfg);
int lsh_server_create(lsh_index_t* index, const lsh_server_config_t* cfg, lsh_server_t** out);
/* queries [n_query][d] and results [n_query][lsh_calculate_k(k)] are host
 * buffers the caller keeps until lsh_server_wait returns for the ticket.
 * 1 <= k <= 128. */
int lsh_server_submit(lsh_server_t* server, const float* queries, int n_query, int k, lsh_point_t* results,
                      int64_t* ticket);
/* Blocks until the request is served; returns its status.  latency_ms (may
 * be NULL): from submission to delivery. */
int lsh_server_wait(lsh_server_t* server, int64_t ticket, double* latency_ms);
/* Batches and queries served so far, and the largest batch the budget admits. */
int lsh_server_stats(lsh_server_t* server, int64_t* batches, int64_t* queries, int* max_batch);
int lsh_server_destroy(lsh_server_t* server);

/* Device bytes held by the index. */
int64_t lsh_index_device_bytes(const lsh_index_t* index);

int lsh_index_destroy(lsh_index_t* index);

#ifdef __cplusplus
}
#endif

#endif /* LSHGPU_H */
