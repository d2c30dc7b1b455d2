// oracle.cpp -- CPU restatement of the reference's k-NN path.
//
// TEST INFRASTRUCTURE ONLY.  Nothing in the product (lshongpu_amd/) links,
// loads or calls this file; only tests/, __graft_entry__.smoke() and the
// cpu_baseline leg of bench.py do, and only as the checker.
//
// Each function cites the reference file:line it restates (paths relative to
// the reference repository root).  Conventions fixed here and shared with the
// GPU path:
//   * the order of results is (distance ascending, ID ascending); distances
//     are compared through an order-preserving integer image of the IEEE bits,
//     so NaN sorts after +inf;
//   * "exact" distances are FP64: products and sums of the float inputs
//     promoted to double, accumulated sequentially in dimension order with
//     fma, square roots and quotients correctly rounded;
//   * FP32 stages that must be bit-exact (normalisation, SimHash dot
//     products, cross-polytope butterflies, weighted-MinHash draws) are
//     written as the explicit sequence of correctly rounded single-precision
//     operations the reference's device code performs (nvcc contracts
//     `s += a*b` into fmaf, and uses IEEE sqrtf and division by default).
//
// Parity pins (what checks each function) are listed in tests/ and DESIGN.md.

#include <algorithm>
#include <atomic>
#include <chrono>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <functional>
#include <numeric>
#include <random>
#include <thread>
#include <vector>

#define EXPORT extern "C" __attribute__((visibility("default")))

namespace {

inline uint32_t f32_key(float f) {
    if (f != f) return 0xFFFFFFFFu;  // every NaN ranks last
    uint32_t u;
    std::memcpy(&u, &f, 4);
    return (u & 0x80000000u) ? ~u : (u | 0x80000000u);
}

inline uint64_t f64_key(double f) {
    if (f != f) return ~0ull;  // every NaN ranks last
    uint64_t u;
    std::memcpy(&u, &f, 8);
    return (u & 0x8000000000000000ull) ? ~u : (u | 0x8000000000000000ull);
}

constexpr float kPadDistance = 2147483648.0f;  // (float)INT_MAX, candidateSetScanner.cuh:30

int hw_threads() {
    unsigned n = std::thread::hardware_concurrency();
    return n == 0 ? 1 : (int)std::min(n, 16u);
}

template <class F>
void parallel_for(int64_t n, F&& f, int threads = 0) {
    if (threads <= 0) threads = hw_threads();
    if (threads == 1 || n < 2) {
        for (int64_t i = 0; i < n; ++i) f(i);
        return;
    }
    std::atomic<int64_t> next{0};
    std::vector<std::thread> pool;
    for (int t = 0; t < threads; ++t)
        pool.emplace_back([&] {
            for (;;) {
                int64_t i = next.fetch_add(1);
                if (i >= n) break;
                f(i);
            }
        });
    for (auto& th : pool) th.join();
}

// distanceFunctions.cuh:10-71 restated in FP64.  func 0 is the -cos of
// optimizedLinearScan.cu:56-57 (implementation 1).
double dist64(int func, const float* q, const float* x, int d) {
    switch (func) {
        case 0:
        case 1: {
            double dot = 0, nq = 0, nx = 0;
            for (int j = 0; j < d; ++j) {
                double a = q[j], b = x[j];
                dot = std::fma(a, b, dot);
                nq = std::fma(a, a, nq);
                nx = std::fma(b, b, nx);
            }
            double c = dot / (std::sqrt(nq) * std::sqrt(nx));
            return func == 0 ? -c : 1.0 - c;
        }
        case 2: {
            double smin = 0, smax = 0;
            for (int j = 0; j < d; ++j) {
                double a = q[j], b = x[j];
                smin += std::min(a, b);
                smax += std::max(a, b);
            }
            return 1.0 - smin / smax;
        }
        case 3: {
            double s = 0;
            for (int j = 0; j < d; ++j) {
                double t = (double)q[j] - (double)x[j];
                s = std::fma(t, t, s);
            }
            return std::sqrt(s);
        }
        default:
            return NAN;
    }
}

struct Cand {
    uint64_t key;
    int32_t id;
    bool operator<(const Cand& o) const { return key != o.key ? key < o.key : id < o.id; }
};

}  // namespace

// ----------------------------------------------------------------- basics

// launchHelper.cuh:21-25
EXPORT int orc_calculate_k(int k) {
    int divisor = (k - 1) / 32;
    return divisor * 32 + 32;
}

EXPORT double orc_dist64(int func, const float* q, const float* x, int d) {
    return dist64(func, q, x, d);
}

// goldstandard.cpp:136-173 (per query: distance to every point, order,
// keep the first k), with the distances of distanceFunctions.cuh:10-71 in FP64
// and the (distance, ID) order.  ids/dists are [Q][k]; rows with fewer than k
// points are padded with (-1, (float)INT_MAX), or (-1, 2.0f) for func 0
// (optimizedLinearScan.cu:28-31).
EXPORT void orc_exact_knn(const float* data, int64_t N, const float* queries,
                          int64_t Q, int d, int func, int k, int32_t* ids,
                          float* dists) {
    parallel_for(Q, [&](int64_t q) {
        std::vector<Cand> c((size_t)N);
        const float* qv = queries + q * d;
        for (int64_t i = 0; i < N; ++i) {
            c[i].key = f64_key(dist64(func, qv, data + i * d, d));
            c[i].id = (int32_t)i;
        }
        int64_t m = std::min<int64_t>(k, N);
        std::partial_sort(c.begin(), c.begin() + m, c.end());
        for (int64_t j = 0; j < k; ++j) {
            if (j < m) {
                ids[q * k + j] = c[j].id;
                dists[q * k + j] = (float)dist64(func, qv, data + (int64_t)c[j].id * d, d);
            } else {
                ids[q * k + j] = -1;
                dists[q * k + j] = func == 0 ? 2.0f : kPadDistance;
            }
        }
    });
}

// processingUtils.cuh:48-67 transformToUnitVectors in FP32: magnitude as a
// fused multiply-add chain, IEEE sqrtf, IEEE division.
EXPORT void orc_normalize_rows(float* x, int64_t n, int d) {
    parallel_for(n, [&](int64_t i) {
        float* r = x + i * d;
        float m = 0.0f;
        for (int j = 0; j < d; ++j) m = std::fmaf(r[j], r[j], m);
        m = std::sqrt(m);
        for (int j = 0; j < d; ++j) r[j] = r[j] / m;
    });
}

// processingUtils.cuh:12-46 transformData: minValues[j] starts at 0 and takes
// atomicMin of (int)floor(x) over data and queries; every value is then
// shifted by abs(minValues[j]).  Followed by the normalisation of
// memOptimizedLinearScan.cu:159 / weightedMinHash.cuh:207.
EXPORT void orc_gj_transform(float* data, int64_t N, float* queries, int64_t Q, int d) {
    std::vector<int> mins((size_t)d, 0);
    auto scan = [&](const float* a, int64_t n) {
        for (int64_t i = 0; i < n; ++i)
            for (int j = 0; j < d; ++j) mins[j] = std::min(mins[j], (int)std::floor(a[i * d + j]));
    };
    scan(data, N);
    scan(queries, Q);
    auto shift = [&](float* a, int64_t n) {
        for (int64_t i = 0; i < n; ++i)
            for (int j = 0; j < d; ++j) a[i * d + j] += (float)std::abs(mins[j]);
    };
    shift(data, N);
    shift(queries, Q);
    orc_normalize_rows(queries, Q, d);
    orc_normalize_rows(data, N, d);
}

// ------------------------------------------------------------ random streams

// randomVectorGenerator.cpp:32-53 (stddev 1) and :56-74 (stddev 1/sketchDim):
// default-seeded std::default_random_engine, std::normal_distribution<double>,
// each draw cast to float.
EXPORT void orc_normal_stream(float* out, int64_t n, double stddev) {
    std::default_random_engine gen;
    std::normal_distribution<double> dist(0.0, stddev);
    for (int64_t i = 0; i < n; ++i) out[i] = (float)dist(gen);
}

// randomVectorGenerator.cpp:76-87: uniform_int_distribution<int>(0,2), 0 -> -1.
EXPORT void orc_sign_stream(int32_t* out, int64_t n) {
    std::default_random_engine gen;
    std::uniform_int_distribution<int> dist(0, 2);
    for (int64_t i = 0; i < n; ++i) out[i] = dist(gen) == 0 ? -1 : 1;
}

// weightedMinHash.cuh:231-250: uniform_int_distribution<int>(0,1).
EXPORT void orc_bit_stream(uint8_t* out, int64_t n) {
    std::default_random_engine gen;
    std::uniform_int_distribution<int> dist(0, 1);
    for (int64_t i = 0; i < n; ++i) out[i] = (uint8_t)dist(gen);
}

// ------------------------------------------------------------------ SimHash

// simHash.cuh:20-46 sketchDataGeneric (simHash.cu:22-47 is the hashBits = 32
// case): bit b of word w is dot(x, planes[(w*hashBits + b)*d ...]) >= 0, the
// dot product a sequential fmaf chain from 0.0f; bits packed LSB first.
EXPORT void orc_simhash(const float* x, int64_t n, int d, const float* planes,
                        int words, int hashBits, uint32_t* out) {
    parallel_for(n, [&](int64_t i) {
        const float* r = x + i * d;
        for (int w = 0; w < words; ++w) {
            uint32_t s = 0;
            for (int b = 0; b < hashBits; ++b) {
                const float* p = planes + ((int64_t)w * hashBits + b) * d;
                float dot = 0.0f;
                for (int j = 0; j < d; ++j) dot = std::fmaf(r[j], p[j], dot);
                s |= (dot >= 0.0f ? 1u : 0u) << b;
            }
            out[i * words + w] = s;
        }
    });
}

// simHashJL.cuh:34-51: sketch value s = sequential fmaf chain of x . planes[s].
EXPORT void orc_jl(const float* x, int64_t n, int d, const float* planes, int sdim, float* out) {
    parallel_for(n, [&](int64_t i) {
        const float* r = x + i * d;
        for (int s = 0; s < sdim; ++s) {
            const float* p = planes + (int64_t)s * d;
            float dot = 0.0f;
            for (int j = 0; j < d; ++j) dot = std::fmaf(r[j], p[j], dot);
            out[i * sdim + s] = dot;
        }
    });
}

// distanceFunctions.cuh:73-85 hammingDistanceFunc (popcount of xor).
EXPORT int orc_hamming(const uint32_t* a, const uint32_t* b, int words) {
    int h = 0;
    for (int w = 0; w < words; ++w) h += __builtin_popcount(a[w] ^ b[w]);
    return h;
}

// The candidate set of a sketch scan: the C smallest (sketch distance, ID)
// pairs over all data points (the reference keeps an implicit 32*TQ-slot queue,
// sketchedDistanceScanners.cuh:14-100; the build makes C explicit).
// kind 0: Hamming over `words` uint32 (implementations 3 and 5,
// distanceFunctions.cuh:73-85 and :118-134 order identically);
// kind 1: the number of unequal uint32 hash values, i.e. H minus the matches
// of jaccardDistanceFunc (implementation 4, distanceFunctions.cuh:102-116).
// ids/dist are [Q][C], padded with -1.
EXPORT void orc_sketch_topc(int kind, const uint32_t* ds, int64_t N,
                            const uint32_t* qs, int64_t Q, int words, int C,
                            int32_t* ids, int32_t* dist) {
    parallel_for(Q, [&](int64_t q) {
        std::vector<Cand> c((size_t)N);
        const uint32_t* qv = qs + q * words;
        for (int64_t i = 0; i < N; ++i) {
            const uint32_t* dv = ds + i * words;
            int32_t v = 0;
            if (kind == 0) {
                for (int w = 0; w < words; ++w) v += __builtin_popcount(dv[w] ^ qv[w]);
            } else {
                for (int w = 0; w < words; ++w) v += (dv[w] != qv[w]) ? 1 : 0;
            }
            c[i].key = (uint64_t)(uint32_t)v;
            c[i].id = (int32_t)i;
        }
        int64_t m = std::min<int64_t>(C, N);
        std::partial_sort(c.begin(), c.begin() + m, c.end());
        for (int64_t j = 0; j < C; ++j) {
            ids[q * C + j] = j < m ? c[j].id : -1;
            if (dist) dist[q * C + j] = j < m ? (int32_t)c[j].key : -1;
        }
    });
}

// Float sketch candidate set (implementation 6): squared L2 of JL sketches as
// distanceFunctions.cuh:87-99 euclideanDistanceFunc sums it in FP32.  Each
// term is pow(q - x, 2): the square of a float is exact in double (48
// significant bits), so whichever precision pow runs in, the term stored in
// the float `dist` is the correctly rounded fl(t*t); it is then added to the
// running sum by a separate FP32 add.  (Assumption: the reference's compiler
// did not fuse that add with the square; the statement boundary and the pow
// call stand between them.)  Candidates are ordered by (sum, ID).
EXPORT void orc_jl_topc(const float* ds, int64_t N, const float* qs, int64_t Q,
                        int sdim, int C, int32_t* ids) {
    parallel_for(Q, [&](int64_t q) {
        std::vector<Cand> c((size_t)N);
        const float* qv = qs + q * sdim;
        for (int64_t i = 0; i < N; ++i) {
            const float* dv = ds + i * sdim;
            float s = 0.0f;
            for (int j = 0; j < sdim; ++j) {
                const float t = qv[j] - dv[j];
                const float sq = t * t;  // -ffp-contract=off: not fused with the add
                s = s + sq;
            }
            c[i].key = f32_key(s);
            c[i].id = (int32_t)i;
        }
        int64_t m = std::min<int64_t>(C, N);
        std::partial_sort(c.begin(), c.begin() + m, c.end());
        for (int64_t j = 0; j < C; ++j) ids[q * C + j] = j < m ? c[j].id : -1;
    });
}

// candidateSetScanner.cuh:10-32 followed by the final sort and write of
// sketchedDistanceScanners.cuh:87-99: exact distance of every candidate, then
// the k smallest by (FP64 distance, ID).  cand is [Q][C] (entries < 0 skipped,
// duplicates counted once).
EXPORT void orc_rerank(const float* data, const float* queries, int64_t Q, int d,
                       int func, const int32_t* cand, int C, int k, int32_t* ids,
                       float* dists) {
    parallel_for(Q, [&](int64_t q) {
        std::vector<Cand> c;
        c.reserve(C);
        const float* qv = queries + q * d;
        std::vector<int32_t> seen(cand + q * C, cand + q * C + C);
        std::sort(seen.begin(), seen.end());
        seen.erase(std::unique(seen.begin(), seen.end()), seen.end());
        for (int32_t id : seen) {
            if (id < 0) continue;
            c.push_back({f64_key(dist64(func, qv, data + (int64_t)id * d, d)), id});
        }
        int64_t m = std::min<int64_t>(k, (int64_t)c.size());
        std::partial_sort(c.begin(), c.begin() + m, c.end());
        for (int64_t j = 0; j < k; ++j) {
            if (j < m) {
                ids[q * k + j] = c[j].id;
                dists[q * k + j] = (float)dist64(func, qv, data + (int64_t)c[j].id * d, d);
            } else {
                ids[q * k + j] = -1;
                dists[q * k + j] = kPadDistance;
            }
        }
    });
}

// ---------------------------------------------------------------- validation

// validation.cpp:129-158 calculateRecall: per query, the share of the first
// reportK truth IDs found in the set of the first reportK result IDs (result
// rows have stride k); averaged over queries.
EXPORT float orc_recall(const int32_t* truth, int truth_k, const int32_t* res,
                        int k, int64_t Q, int reportK) {
    float total = 0;
    for (int64_t q = 0; q < Q; ++q) {
        std::vector<int32_t> s(res + q * k, res + q * k + reportK);
        std::sort(s.begin(), s.end());
        float found = 0;
        for (int i = 0; i < truth_k && i < reportK; ++i)
            if (std::binary_search(s.begin(), s.end(), truth[q * truth_k + i])) found++;
        total += found / reportK;
    }
    return total / (float)Q;
}

// validation.cpp:160-198 calculateDistanceRatio: (sum of the distances of the
// distinct result IDs among the first reportK) / (sum of the first reportK
// truth distances), averaged over queries.  A repeated ID keeps the distance
// of its first occurrence (std::set::insert).
EXPORT float orc_distance_ratio(const float* truth_d, int truth_k, const int32_t* res_ids,
                                const float* res_d, int k, int64_t Q, int reportK) {
    float total = 0;
    for (int64_t q = 0; q < Q; ++q) {
        float td = 0, rd = 0;
        for (int i = 0; i < truth_k && i < reportK; ++i) td += truth_d[q * truth_k + i];
        std::vector<std::pair<int32_t, float>> s;
        for (int i = 0; i < reportK; ++i) {
            int32_t id = res_ids[q * k + i];
            bool dup = false;
            for (auto& p : s)
                if (p.first == id) dup = true;
            if (!dup) s.push_back({id, res_d[q * k + i]});
        }
        std::sort(s.begin(), s.end());
        for (auto& p : s) rd += p.second;
        total += rd / td;
    }
    return total / (float)Q;
}

// ------------------------------------------------------------ cross-polytope

// crossPolyTope.cuh:95-104 logDim: smallest power of two >= d.
EXPORT int orc_pow2_at_least(int d) {
    int p = 1;
    while (p < d) p *= 2;
    return p;
}

// crossPolyTope.cuh:25-85 generateHashCrossPoly for already normalised rows:
// per table, pad to P, three rounds of (multiply by the +-1 diagonal
// signs[t*3P + round*P + i], butterflies for offset P/2 ... 1 with
// v'[i] = v[i] + v[i^o] when i < i^o, else -v[i] + v[i^o]); the key is the
// first index of the largest |v| (strictly greater wins), i for negative v and
// i + P otherwise.
// Dense-rotation cross-polytope keys (keysImplementation 8; the reference has
// no counterpart: it is the hash of crossPolyTope.cuh:64-76 applied to a dense
// Gaussian rotation instead of the HD^3 transform).  Hash h of row x:
// y_j = sequential FP32 fmaf chain over i < d of G[h][j][i] * x[i] from 0,
// j < P; the index is the first j of largest |y_j| by the reference's
// comparison (crossPolyTope.cuh:66-75), encoded j when y_j < 0, else j + P.
// Table t concatenates hashes h = t*K .. t*K+K-1: key = sum_k hash << (k*b),
// b = log2(2P).  G is [tables*K][P][d].
EXPORT void orc_crosspoly_dense(const float* x, int64_t n, int d, int P, int tables, int K,
                                const float* G, uint32_t* keys) {
    int b = 0;
    while ((1 << b) < 2 * P) ++b;
    parallel_for(n, [&](int64_t i) {
        const float* r = x + i * d;
        for (int t = 0; t < tables; ++t) {
            uint32_t key = 0;
            for (int k = 0; k < K; ++k) {
                const float* g = G + (int64_t)(t * K + k) * P * d;
                int index = P;  // no finite value: j = 0, non-negative
                float best = (float)INT32_MIN;
                for (int j = 0; j < P; ++j) {
                    float y = 0.0f;
                    for (int q = 0; q < d; ++q) y = std::fmaf(g[(int64_t)j * d + q], r[q], y);
                    const bool neg = y < 0;
                    const float m = neg ? std::fmax(best, -y) : std::fmax(best, y);
                    if (m > best) {
                        index = neg ? j : j + P;
                        best = m;
                    }
                }
                key |= (uint32_t)index << (k * b);
            }
            keys[i * tables + t] = key;
        }
    });
}

EXPORT void orc_crosspoly(const float* x, int64_t n, int d, int tables,
                          const int32_t* signs, uint16_t* keys) {
    const int P = orc_pow2_at_least(d);
    parallel_for(n, [&](int64_t i) {
        std::vector<float> v(P), t(P);
        for (int tb = 0; tb < tables; ++tb) {
            for (int j = 0; j < P; ++j) v[j] = j < d ? x[i * d + j] : 0.0f;
            for (int r = 0; r < 3; ++r) {
                for (int j = 0; j < P; ++j) v[j] *= (float)signs[(int64_t)tb * 3 * P + r * P + j];
                for (int o = P / 2; o > 0; o /= 2) {
                    for (int j = 0; j < P; ++j) {
                        int other = j ^ o;
                        t[j] = j < other ? v[j] + v[other] : -v[j] + v[other];
                    }
                    std::swap(v, t);
                }
            }
            int index = 0;
            float best = (float)INT32_MIN;
            for (int j = 0; j < P; ++j) {
                bool neg = v[j] < 0;
                float m = neg ? std::fmax(best, -v[j]) : std::fmax(best, v[j]);
                if (m > best) {
                    index = neg ? j : j + P;
                    best = m;
                }
            }
            keys[i * tables + tb] = (uint16_t)index;
        }
    });
}

// ---------------------------------------------------------- weighted MinHash

// Philox4x32-10 (Salmon, Moraes, Dror, Shaw, "Parallel random numbers: as easy
// as 1, 2, 3", SC'11): the counter-based generator that replaces the
// reference's cuRAND XORWOW stream (weightedMinHash.cuh:92-96, 119-120,
// 153-154), whose output cannot be reproduced here.
EXPORT void orc_philox4x32_10(const uint32_t ctr[4], const uint32_t key[2], uint32_t out[4]) {
    uint32_t c0 = ctr[0], c1 = ctr[1], c2 = ctr[2], c3 = ctr[3];
    uint32_t k0 = key[0], k1 = key[1];
    for (int r = 0; r < 10; ++r) {
        uint64_t p0 = (uint64_t)0xD2511F53u * c0;
        uint64_t p1 = (uint64_t)0xCD9E8D57u * c2;
        uint32_t n0 = (uint32_t)(p1 >> 32) ^ c1 ^ k0;
        uint32_t n2 = (uint32_t)(p0 >> 32) ^ c3 ^ k1;
        c1 = (uint32_t)p1;
        c3 = (uint32_t)p0;
        c0 = n0;
        c2 = n2;
        k0 += 0x9E3779B9u;
        k1 += 0xBB67AE85u;
    }
    out[0] = c0;
    out[1] = c1;
    out[2] = c2;
    out[3] = c3;
}

// Draw t of hash seed s: Philox4x32-10 with key (s, 0x5EED5EED) and counter
// (t/4, 0, 0, 0), word t%4; U = (word >> 8) * 2^-24 in [0, 1).
EXPORT float orc_wmh_uniform(uint32_t seed, uint32_t t) {
    uint32_t ctr[4] = {t >> 2, 0, 0, 0};
    uint32_t key[2] = {seed, 0x5EED5EEDu};
    uint32_t o[4];
    orc_philox4x32_10(ctr, key, o);
    return (float)(o[t & 3] >> 8) * (1.0f / 16777216.0f);
}

// weightedMinHash.cuh:41-90 preprocess + setupMapIndex: bounds[j] =
// max(0, max ceil(x_j)) over the given rows; start[j] = exclusive prefix sum;
// returns M = sum of bounds.
EXPORT int orc_wmh_bounds(const float* x, int64_t n, const float* y, int64_t m, int d,
                          int32_t* bounds, int32_t* start) {
    for (int j = 0; j < d; ++j) bounds[j] = 0;
    for (int64_t i = 0; i < n; ++i)
        for (int j = 0; j < d; ++j) bounds[j] = std::max(bounds[j], (int)std::ceil(x[i * d + j]));
    for (int64_t i = 0; i < m; ++i)
        for (int j = 0; j < d; ++j) bounds[j] = std::max(bounds[j], (int)std::ceil(y[i * d + j]));
    int M = 0;
    for (int j = 0; j < d; ++j) {
        start[j] = M;
        M += bounds[j];
    }
    return M;
}

// weightedMinHash.cuh:98-106 isGreen and :144-166 sketchDataOriginal: draw
// r = M*U (FP32) until r <= start[c] + x[c] with c the component owning
// floor(r); the hash is the number of red draws.  Seeds are
// i*1234 + 92138 (weightedMinHash.cuh:252-262).  mode 0: one uint32 count per
// hash (H hashes); mode 1 (weightedMinHash.cuh:109-140): word w bit b is
// bitmap[count % bitmap_len] of seed index w*hbits + b, b < hbits (H words).
EXPORT void orc_wmh_sketch(const float* x, int64_t n, int d, const int32_t* bounds,
                           const int32_t* start, int M, int H, int mode, int hbits,
                           const uint8_t* bitmap, int bitmap_len, uint32_t* out) {
    std::vector<int32_t> owner((size_t)std::max(M, 1));
    for (int j = 0; j < d; ++j)
        for (int r = start[j]; r < start[j] + bounds[j]; ++r) owner[r] = j;
    const float Mf = (float)M;
    // A row that sees no green draw in 2^16 draws (a zero or NaN row) records 2^16.
    auto count = [&](const float* row, uint32_t seed) -> uint32_t {
        if (M <= 0) return 0;
        for (uint32_t t = 0; t < 65536u; ++t) {
            float r = Mf * orc_wmh_uniform(seed, t);
            int ri = std::min((int)r, M - 1);
            int c = owner[ri];
            if (r <= (float)start[c] + row[c]) return t;
        }
        return 65536u;
    };
    parallel_for(n, [&](int64_t i) {
        const float* row = x + i * d;
        if (mode == 0) {
            for (int h = 0; h < H; ++h) out[i * H + h] = count(row, (uint32_t)(h * 1234 + 92138));
        } else {
            for (int w = 0; w < H; ++w) {
                uint32_t s = 0;
                for (int b = 0; b < hbits; ++b) {
                    uint32_t c = count(row, (uint32_t)((w * hbits + b) * 1234 + 92138));
                    s |= (uint32_t)bitmap[c % (uint32_t)bitmap_len] << b;
                }
                out[i * H + w] = s;
            }
        }
    });
}

// --------------------------------------------------------------------- LSH

// lshFramework.cuh:102-140 (bucket histogram, offsets, scatter) and
// :142-280 (scan of the query's L buckets, optional rerank, de-duplication),
// restated as sets: the candidates of query q are the distinct IDs i with
// dkeys[i][t] == qkeys[q][t] for some table t, in ascending ID order.
// Returns the candidate count; writes at most cap IDs.
EXPORT int64_t orc_lsh_candidates(const uint32_t* dkeys, int64_t N, const uint32_t* qkeys,
                                  int tables, int64_t q, int32_t* out, int64_t cap) {
    int64_t c = 0;
    for (int64_t i = 0; i < N; ++i) {
        bool hit = false;
        for (int t = 0; t < tables && !hit; ++t) hit = dkeys[i * tables + t] == qkeys[q * tables + t];
        if (hit) {
            if (c < cap) out[c] = (int32_t)i;
            ++c;
        }
    }
    return c;
}

// ---------------------------------------------------------- CPU baseline

namespace gs {
// goldstandard.cpp:17-22, 95-121, 136-173, restated with the same data
// structures and by-value calls (each point is copied into the distance
// function, and the whole array is re-sorted with std::sort per query).
struct point {
    std::vector<float> elements;
    float cachedDistance;
    int id;
};
float magnitude(point x) {
    return std::sqrt(std::accumulate(x.elements.begin(), x.elements.end(), 0.0,
                                     [](float state, float xi) { return state + (xi * xi); }));
}
float innerProduct(point x, point y) {
    float dotProd = std::inner_product(x.elements.begin(), x.elements.end(), y.elements.begin(), 0.0);
    float xMagnitude = magnitude(x);
    float yMagnitude = magnitude(y);
    float sim = dotProd / (xMagnitude * yMagnitude);
    return -sim;
}
}  // namespace gs

// Times the reference's CPU exact k-NN (goldstandard.cpp:136-173) on
// `threads` threads, each with its own copy of the point array (the reference
// sorts that array in place), over queries [0, Q).  Writes the first K IDs per
// query and returns the elapsed wall time in seconds.
EXPORT double orc_goldstandard_port(const float* data, int64_t N, const float* queries,
                                    int64_t Q, int d, int K, int threads, int32_t* ids) {
    if (threads < 1) threads = 1;
    std::vector<std::vector<gs::point>> copies((size_t)threads);
    parallel_for(threads, [&](int64_t t) {
        auto& v = copies[t];
        v.resize((size_t)N);
        for (int64_t i = 0; i < N; ++i) {
            v[i].elements.assign(data + i * d, data + (i + 1) * d);
            v[i].id = (int)i;
            v[i].cachedDistance = 0;
        }
    }, threads);
    std::vector<gs::point> qs((size_t)Q);
    for (int64_t i = 0; i < Q; ++i) {
        qs[i].elements.assign(queries + i * d, queries + (i + 1) * d);
        qs[i].id = (int)i;
    }
    auto t0 = std::chrono::steady_clock::now();
    std::vector<std::thread> pool;
    for (int t = 0; t < threads; ++t)
        pool.emplace_back([&, t] {
            auto& pts = copies[t];
            for (int64_t qi = t; qi < Q; qi += threads) {
                gs::point q = qs[qi];
                auto compare = [](gs::point x, gs::point y) { return x.cachedDistance < y.cachedDistance; };
                for (int64_t pi = 0; pi < N; ++pi) {
                    gs::point p = pts[pi];
                    p.cachedDistance = gs::innerProduct(q, p);
                    pts[pi] = p;
                }
                std::sort(pts.begin(), pts.end(), compare);
                for (int i = 0; i < K && i < N; ++i) ids[qi * K + i] = pts[i].id;
            }
        });
    for (auto& th : pool) th.join();
    auto t1 = std::chrono::steady_clock::now();
    return std::chrono::duration<double>(t1 - t0).count();
}
