// io.cpp -- datasets, truth files, validation and report formats of the
// reference's runner (host code).
//
//   text datasets   gloveparser.cu:9-86 (reader), DataParsing/Writer/writer.cpp:10-29 (writer)
//   .fvecs          DataParsing/Parsers/siftParser.cpp:8-28
//   truth files     GoldStandard/goldstandard.cpp:123-134 (writer), LinearScan/validation.cpp:62-107 (reader)
//   validation      LinearScan/validation.cpp:129-208
//   result dump     LinearScan/resultWriter.cpp:7-20
//   performance CSV LinearScan/resultWriter.cpp:37-60
#include <cerrno>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <fstream>
#include <map>
#include <set>
#include <sstream>
#include <string>
#include <vector>

#include "lshgpu_internal.hpp"

using lshgpu::fail;
using lshgpu::guarded;

namespace {

struct File {
    FILE* f = nullptr;
    File(const char* path, const char* mode) {
        if (path) f = std::fopen(path, mode);
    }
    ~File() {
        if (f) std::fclose(f);
    }
};

template <class T>
T* malloc_array(size_t n) {
    T* p = static_cast<T*>(std::malloc(sizeof(T) * std::max<size_t>(n, 1)));
    if (!p) fail(LSH_ERR_NOMEM, "host allocation failed");
    return p;
}

// Reads one text line (without the newline); false at end of file.
bool read_line(FILE* f, std::string& line) {
    line.clear();
    int c;
    bool any = false;
    while ((c = std::fgetc(f)) != EOF) {
        any = true;
        if (c == '\n') return true;
        line.push_back((char)c);
    }
    return any;
}

bool blank(const std::string& s) {
    for (char c : s)
        if (c != ' ' && c != '\t' && c != '\r') return false;
    return true;
}

// One entry of a truth file or a result row, compared by ID as the
// reference's PointValidation is (validation.cpp:12-28).
struct ById {
    int32_t id;
    float dist;
    bool operator<(const ById& o) const { return id < o.id; }
};

struct Truth {
    int n = 0, k = 0;
    std::vector<int32_t> ids;   // [n][k]
    std::vector<float> dists;   // [n][k]
};

Truth read_truth(const char* path) {
    std::ifstream in(path ? path : "");
    if (!in) fail(LSH_ERR_IO, std::string("cannot open truth file ") + (path ? path : "(null)"));
    Truth t;
    in >> t.n >> t.k;
    if (!in || t.n < 0 || t.k < 0) fail(LSH_ERR_IO, "truth file header must be the query count and k");
    t.ids.resize((size_t)t.n * t.k);
    t.dists.resize((size_t)t.n * t.k);
    std::string qid;
    for (int q = 0; q < t.n; ++q) {
        in >> qid;
        for (int j = 0; j < t.k; ++j) in >> t.ids[(size_t)q * t.k + j] >> t.dists[(size_t)q * t.k + j];
        if (!in) fail(LSH_ERR_IO, "truth file ends early at query " + std::to_string(q));
    }
    return t;
}

}  // namespace

extern "C" {

void lsh_free(void* p) { std::free(p); }

int lsh_read_text_dataset(const char* path, float** out, int64_t* n_out, int* d_out, int32_t** ids_out) {
    return guarded([&] {
        LSH_REQUIRE(out && n_out && d_out, LSH_ERR_INVALID, "null output pointer");
        *out = nullptr;
        if (ids_out) *ids_out = nullptr;
        File fp(path, "r");
        if (!fp.f) fail(LSH_ERR_IO, std::string("cannot open dataset ") + (path ? path : "(null)"));
        std::string line;
        // Header: N and d on the first two lines (gloveparser.cu:26-31).
        if (!read_line(fp.f, line)) fail(LSH_ERR_IO, "dataset has no header");
        const long long n = std::atoll(line.c_str());
        if (!read_line(fp.f, line)) fail(LSH_ERR_IO, "dataset header has no dimension line");
        const int d = std::atoi(line.c_str());
        LSH_REQUIRE(n >= 0 && d >= 1, LSH_ERR_IO, "dataset header must give N >= 0 and d >= 1");
        float* x = malloc_array<float>((size_t)n * d);
        int32_t* ids = malloc_array<int32_t>((size_t)n);
        int64_t row = 0;
        while (row < n && read_line(fp.f, line)) {
            if (blank(line)) continue;
            const char* s = line.c_str();
            char* end = nullptr;
            // Row: ID then d values; values are parsed as C floats (the
            // reference's digit loop keeps only the last integer digit,
            // gloveparser.cu:58-66).
            errno = 0;
            const long id = std::strtol(s, &end, 10);
            if (end == s) {
                std::free(x);
                std::free(ids);
                fail(LSH_ERR_IO, "dataset row " + std::to_string(row) + " has no ID");
            }
            ids[row] = (int32_t)id;
            s = end;
            for (int j = 0; j < d; ++j) {
                const float v = std::strtof(s, &end);
                if (end == s) {
                    std::free(x);
                    std::free(ids);
                    fail(LSH_ERR_IO, "dataset row " + std::to_string(row) + " has " + std::to_string(j) +
                                         " values, expected " + std::to_string(d));
                }
                x[(size_t)row * d + j] = v;
                s = end;
            }
            ++row;
        }
        if (row != n) {
            std::free(x);
            std::free(ids);
            fail(LSH_ERR_IO, "dataset has " + std::to_string(row) + " rows, header says " + std::to_string(n));
        }
        *out = x;
        *n_out = n;
        *d_out = d;
        if (ids_out)
            *ids_out = ids;
        else
            std::free(ids);
    });
}

int lsh_read_fvecs(const char* path, float** out, int64_t* n_out, int* d_out) {
    return guarded([&] {
        LSH_REQUIRE(out && n_out && d_out, LSH_ERR_INVALID, "null output pointer");
        *out = nullptr;
        File fp(path, "rb");
        if (!fp.f) fail(LSH_ERR_IO, std::string("cannot open ") + (path ? path : "(null)"));
        std::vector<float> x;
        int d = -1;
        int64_t n = 0;
        int32_t dim;
        while (std::fread(&dim, sizeof(dim), 1, fp.f) == 1) {
            LSH_REQUIRE(dim >= 1 && (d < 0 || dim == d), LSH_ERR_IO, ".fvecs records must share one dimension");
            d = dim;
            const size_t at = x.size();
            x.resize(at + (size_t)d);
            if (std::fread(x.data() + at, sizeof(float), (size_t)d, fp.f) != (size_t)d)
                fail(LSH_ERR_IO, ".fvecs record " + std::to_string(n) + " is truncated");
            ++n;
        }
        LSH_REQUIRE(n > 0, LSH_ERR_IO, ".fvecs file holds no record");
        float* p = malloc_array<float>(x.size());
        std::memcpy(p, x.data(), sizeof(float) * x.size());
        *out = p;
        *n_out = n;
        *d_out = d;
    });
}

int lsh_write_text_dataset(const char* path, const float* data, int64_t n, int d) {
    return guarded([&] {
        LSH_REQUIRE((data || n == 0) && d >= 1 && n >= 0, LSH_ERR_INVALID, "bad dataset arguments");
        File fp(path, "w");
        if (!fp.f) fail(LSH_ERR_IO, std::string("cannot write ") + (path ? path : "(null)"));
        // writer.cpp:15-26 layout: N, d, then "ID v1 ... vd " rows with IDs
        // 0..N-1; 9 significant digits so that values read back exactly.
        std::fprintf(fp.f, "%lld\n%d\n", (long long)n, d);
        for (int64_t i = 0; i < n; ++i) {
            std::fprintf(fp.f, "%lld ", (long long)i);
            for (int j = 0; j < d; ++j) std::fprintf(fp.f, "%.9g ", data[(size_t)i * d + j]);
            std::fputc('\n', fp.f);
        }
        if (std::ferror(fp.f)) fail(LSH_ERR_IO, "write failed");
    });
}

int lsh_read_truth(const char* path, int* n_query, int* k, int32_t** ids, float** distances) {
    return guarded([&] {
        LSH_REQUIRE(n_query && k && ids && distances, LSH_ERR_INVALID, "null output pointer");
        Truth t = read_truth(path);
        int32_t* pi = malloc_array<int32_t>(t.ids.size());
        float* pd = malloc_array<float>(t.dists.size());
        std::memcpy(pi, t.ids.data(), sizeof(int32_t) * t.ids.size());
        std::memcpy(pd, t.dists.data(), sizeof(float) * t.dists.size());
        *n_query = t.n;
        *k = t.k;
        *ids = pi;
        *distances = pd;
    });
}

int lsh_write_truth(const char* path, const lsh_point_t* results, const int32_t* query_ids, int n_query,
                    int k_stride, int k) {
    return guarded([&] {
        LSH_REQUIRE((results || n_query == 0) && n_query >= 0 && k >= 0 && k <= k_stride, LSH_ERR_INVALID,
                    "bad truth arguments");
        std::ofstream out(path ? path : "");
        if (!out) fail(LSH_ERR_IO, std::string("cannot write ") + (path ? path : "(null)"));
        // goldstandard.cpp:123-134: n, k, then "qid:" and k lines "ID distance"
        // (default stream formatting, as the reference writes them).
        out << n_query << "\n" << k << "\n";
        for (int q = 0; q < n_query; ++q) {
            out << (query_ids ? query_ids[q] : q) << ":\n";
            for (int j = 0; j < k; ++j) {
                const lsh_point_t& p = results[(size_t)q * k_stride + j];
                out << p.ID << " " << p.distance << "\n";
            }
        }
        if (!out) fail(LSH_ERR_IO, "write failed");
    });
}

int lsh_validate(const char* truth_path, const lsh_point_t* results, int n_query, int k, int reportK,
                 float* recall, float* avg_distance) {
    return guarded([&] {
        LSH_REQUIRE(results && recall && avg_distance && reportK >= 0 && reportK <= k, LSH_ERR_INVALID,
                    "bad validation arguments");
        const Truth t = read_truth(truth_path);
        LSH_REQUIRE(t.n <= n_query, LSH_ERR_INVALID, "the truth file has more queries than the results");
        // validation.cpp:39-60: the first reportK results of a query as a set
        // keyed by ID (a repeated ID keeps its first entry).
        float total_recall = 0.0f, total_ratio = 0.0f;
        for (int q = 0; q < t.n; ++q) {
            std::set<ById> nn;
            for (int i = 0; i < reportK; ++i) {
                const lsh_point_t& p = results[(size_t)q * k + i];
                nn.insert(ById{p.ID, p.distance});
            }
            // :129-158 recall: truth entries i < reportK found in the set.
            float recalled = 0.0f, truth_sum = 0.0f, result_sum = 0.0f;
            for (int i = 0; i < t.k && i < reportK; ++i) {
                const int32_t id = t.ids[(size_t)q * t.k + i];
                if (nn.count(ById{id, 0.0f})) recalled += 1.0f;
                truth_sum += t.dists[(size_t)q * t.k + i];
            }
            total_recall += recalled / (float)reportK;
            // :160-198 distance ratio: the set's distances in ID order over
            // the first reportK truth distances.
            for (const ById& p : nn) result_sum += p.dist;
            total_ratio += result_sum / truth_sum;
        }
        *recall = t.n ? total_recall / (float)t.n : NAN;
        *avg_distance = t.n ? total_ratio / (float)t.n : NAN;
    });
}

int lsh_write_results(const char* path, const lsh_point_t* results, int k, int n_query, int reportK) {
    return guarded([&] {
        LSH_REQUIRE((results || n_query == 0) && reportK >= 0 && reportK <= k, LSH_ERR_INVALID,
                    "bad result arguments");
        File fp(path, "w");
        FILE* f = path ? fp.f : stdout;
        if (!f) fail(LSH_ERR_IO, std::string("cannot write ") + path);
        // resultWriter.cpp:7-20.
        std::fprintf(f, "%d\n%d\n", n_query, reportK);
        for (int i = 0; i < n_query; ++i) {
            std::fprintf(f, "%d:\n", i);
            for (int j = 0; j < reportK; ++j) {
                const lsh_point_t& p = results[(size_t)i * k + j];
                std::fprintf(f, "%d %f\n", p.ID, p.distance);
            }
        }
        if (std::ferror(f)) fail(LSH_ERR_IO, "write failed");
    });
}

int lsh_write_performance(const char* path, const lsh_result_t* r, int implementation, int keysImplementation,
                          int sketchDim, int k, int bucketKeyBits, int tables) {
    return guarded([&] {
        LSH_REQUIRE(r, LSH_ERR_INVALID, "null result");
        std::ofstream out(path ? path : "", std::ios_base::app);
        if (!out) fail(LSH_ERR_IO, std::string("cannot append to ") + (path ? path : "(null)"));
        // resultWriter.cpp:43-57, 13 columns.  Columns 5 and 8 hold the
        // reference's compile-time THREAD_QUEUE_SIZE (128) and
        // WITH_TQ_OR_BUFFER (0) (constants.cuh:6, 11), kept so that existing
        // readers of these files see the same layout.
        out << implementation << "," << keysImplementation << "," << sketchDim << "," << k << "," << 128 << ","
            << bucketKeyBits << "," << tables << "," << 0 << "," << r->preprocessTime << "," << r->constructionTime
            << "," << r->scanTime << "," << r->recall << "," << r->avgDistance << std::endl;
        if (!out) fail(LSH_ERR_IO, "write failed");
    });
}

}  // extern "C"
