// lshgpu_internal.hpp -- shared declarations of the engine (host and device).
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <cstring>
#include <new>
#include <stdexcept>
#include <string>
#include <vector>

#include "../../include/lshgpu.h"

namespace lshgpu {

// ------------------------------------------------------------------ errors

struct Error : std::runtime_error {
    int code;
    Error(int c, const std::string& m) : std::runtime_error(m), code(c) {}
};

[[noreturn]] void fail(int code, const std::string& msg);

// Message of the last failed entry point on this thread.
std::string& last_error();

// Runs an entry point body: exceptions become status codes and the message
// is kept for lsh_last_error().
template <class F>
int guarded(F&& f) {
    try {
        f();
        last_error().clear();
        return LSH_OK;
    } catch (const Error& e) {
        last_error() = e.what();
        return e.code;
    } catch (const std::bad_alloc&) {
        last_error() = "host allocation failed";
        return LSH_ERR_NOMEM;
    } catch (const std::exception& e) {
        last_error() = e.what();
        return LSH_ERR_INTERNAL;
    }
}

#define LSH_HIP(expr)                                                              \
    do {                                                                           \
        hipError_t _e = (expr);                                                    \
        if (_e != hipSuccess)                                                      \
            ::lshgpu::fail(LSH_ERR_DEVICE, std::string(#expr) + ": " +             \
                                               hipGetErrorString(_e) + " at " +    \
                                               __FILE__ + ":" + std::to_string(__LINE__)); \
    } while (0)

#define LSH_REQUIRE(cond, code, msg) \
    do {                             \
        if (!(cond)) ::lshgpu::fail(code, msg); \
    } while (0)

// Checks the launch just issued on `s` (configuration errors only; faults
// surface at the next synchronisation).
void check_launch(const char* what);

// Raises the dynamic LDS limit of `kernel` to `bytes` when it exceeds the
// default 64 KiB (CDNA4 workgroups may use up to 160 KiB).
void set_dynamic_lds(const void* kernel, size_t bytes);

// ------------------------------------------------------------ device memory

template <class T>
class DevBuf {
   public:
    DevBuf() = default;
    explicit DevBuf(size_t n) { alloc(n); }
    ~DevBuf() { release(); }
    DevBuf(const DevBuf&) = delete;
    DevBuf& operator=(const DevBuf&) = delete;
    DevBuf(DevBuf&& o) noexcept : p_(o.p_), n_(o.n_) { o.p_ = nullptr; o.n_ = 0; }
    DevBuf& operator=(DevBuf&& o) noexcept {
        if (this != &o) {
            release();
            p_ = o.p_;
            n_ = o.n_;
            o.p_ = nullptr;
            o.n_ = 0;
        }
        return *this;
    }
    void alloc(size_t n) {
        if (p_) (void)hipDeviceSynchronize();  // work queued on the old buffer has finished
        release();
        if (n == 0) return;
        void* p = nullptr;
        hipError_t e = hipMalloc(&p, n * sizeof(T));
        if (e != hipSuccess) {
            (void)hipGetLastError();
            fail(LSH_ERR_NOMEM, "hipMalloc of " + std::to_string(n * sizeof(T)) + " bytes failed: " +
                                    hipGetErrorString(e));
        }
        p_ = static_cast<T*>(p);
        n_ = n;
    }
    // Grow-only reuse.
    void ensure(size_t n) {
        if (n > n_) alloc(n);
    }
    void release() {
        if (p_) (void)hipFree(p_);
        p_ = nullptr;
        n_ = 0;
    }
    T* get() const { return p_; }
    size_t size() const { return n_; }
    size_t bytes() const { return n_ * sizeof(T); }

   private:
    T* p_ = nullptr;
    size_t n_ = 0;
};

// ------------------------------------------------------------ ordered keys
// A candidate is a 64-bit key: an order-preserving image of its distance in
// the high 32 bits and its ID in the low 32 bits, so that comparing keys as
// unsigned integers is the (distance, ID) order.  kEmptyKey marks no candidate.

constexpr uint64_t kEmptyKey = ~0ull;
constexpr float kPadDistance = 2147483648.0f;  // (float)INT_MAX, candidateSetScanner.cuh:30

__host__ __device__ inline uint32_t f32_key(float f) {
    if (f != f) return 0xFFFFFFFFu;  // every NaN ranks last
    uint32_t u = __builtin_bit_cast(uint32_t, f);
    return (u & 0x80000000u) ? ~u : (u | 0x80000000u);
}
__host__ __device__ inline float key_f32(uint32_t k) {
    uint32_t u = (k & 0x80000000u) ? (k & 0x7FFFFFFFu) : ~k;
    return __builtin_bit_cast(float, u);
}
__host__ __device__ inline uint64_t f64_key(double f) {
    if (f != f) return ~0ull;  // every NaN ranks last
    uint64_t u = __builtin_bit_cast(uint64_t, f);
    return (u & 0x8000000000000000ull) ? ~u : (u | 0x8000000000000000ull);
}
__host__ __device__ inline double key_f64(uint64_t k) {
    uint64_t u = (k & 0x8000000000000000ull) ? (k & 0x7FFFFFFFFFFFFFFFull) : ~k;
    return __builtin_bit_cast(double, u);
}
__host__ __device__ inline uint64_t make_key(uint32_t dkey, uint32_t id) {
    return ((uint64_t)dkey << 32) | id;
}

// --------------------------------------------------------- distance kinds
// Exact distances of distanceFunctions.cuh:10-71.  kNegCos is the -cos of
// implementation 1 (optimizedLinearScan.cu:56-57); kSqL2 the squared L2 of
// the JL sketch scan (distanceFunctions.cuh:87-99).
enum DistKind : int { kNegCos = 0, kAngular = 1, kGenJaccard = 2, kEuclid = 3, kSqL2 = 4 };

// Bound on |FP32 distance - exact distance| used to decide when the FP32
// ranking may differ from the FP64 one (absolute for kNegCos/kAngular/
// kGenJaccard, relative for kEuclid).  d is the dimension.
double fp32_error_bound(int kind, int d);

// ------------------------------------------------------------- top-K state
// Per query: keys[q*stride .. +K) hold the best K keys so far in ascending
// order (kEmptyKey padded); keys[q*stride + K .. + newcap) receive the
// candidates appended by a scan round; cnt[q] counts the appends (it may
// exceed newcap, which is recorded as an overflow); thr[q] is the K-th best
// key or kEmptyKey while fewer than K are known.
struct TopKView {
    uint64_t* keys;
    uint32_t* cnt;
    uint64_t* thr;
    uint32_t* ovf;
    int K;
    int newcap;
    int stride;  // K + newcap
    int Q;
    int int_keys;  // 1: integer distances, best list unordered (select_hist_kernel)
    int fresh;     // 1: the best lists are empty and not stored (lazy reset; the next select reads only appends)
    // 1: an appended key's low word is the point's data row (its position in
    // the pivot-cell layout) in place of its ID; the select, given the
    // layout's IDs, turns them into (distance, ID) keys.  Set after a lazy
    // reset, by the pivot-cell search's caller.
    int row_keys;
    // 1 (with row_keys): an appended key's high word is (distance << 16 |
    // q), q the pair's 16-bit approximate angular distance from the fused
    // cell scan; the select orders by (distance, ID) as ever and keeps such
    // keys unchanged in the best list, for the rerank's filter.
    int fused_keys;
};

class TopKState {
   public:
    // int_keys: the distances are small integers; the best list is then an
    // unordered set of exactly K keys and K need not be a power of two.
    // lazy (integer keys only): the best lists are not written; the next
    // select treats them as empty and must be followed by selected().
    // launch = false: only the view is set up; the caller clears it (clear(),
    // or a kernel of its own that does what clear() does).
    void reset(int Q, int K, int newcap, hipStream_t s, bool int_keys = false, bool lazy = false,
               bool launch = true);
    // The reset's clearing: thresholds empty, counts and overflow flags zero,
    // and the best lists empty unless the view is fresh.
    void clear(hipStream_t s);
    TopKView view() const { return v_; }
    // After the select that follows a lazy reset: the best lists are stored.
    void selected() { v_.fresh = 0; }
    // After a lazy reset: the appends will carry rows (TopKView::row_keys);
    // the next reset clears it.
    void rows_in_keys() { v_.row_keys = 1; }
    // After rows_in_keys: the appends will carry approximate distances
    // (TopKView::fused_keys); the next reset clears it.
    void fused_in_keys() { v_.fused_keys = 1; }
    size_t bytes() const { return keys_.bytes() + cnt_.bytes() + thr_.bytes() + ovf_.bytes(); }

   private:
    DevBuf<uint64_t> keys_;
    DevBuf<uint32_t> cnt_;
    DevBuf<uint64_t> thr_;
    DevBuf<uint32_t> ovf_;
    TopKView v_{};
};

// Largest K the select kernels support (keys of the best list plus the sort
// buffer must fit the 160 KiB LDS of a CDNA4 workgroup); integer-key lists
// above kMaxIntTopK use the float-key (sorted) kernel.
constexpr int kMaxTopK = 4096;
constexpr int kMaxIntTopK = 2048;

int pow2_ceil(int x);
int newcap_for(int K);

// ----------------------------------------------------------- kernel entry
// (defined in the .hip files; all launch on `s` and do not synchronise)

// select.hip
// flags (may be null, may be pinned host memory): after the select,
// flags[q] = 1 for each query whose list is incomplete, and flags[Q + i] =
// extra[i] for i < n_extra (n_extra <= 256).  row_ids: with v.row_keys, the
// pivot-cell layout's IDs (row -> ID); the select then also writes the data
// row of each best-list key i as int32 i of the query's append space (-1
// for an empty key), for the rerank.
// The clearing TopKState::reset does, over a view of memory the caller owns.
void launch_topk_clear(const TopKView& v, hipStream_t s);
// fb (may be null; needs v.fused_keys and row_ids): the select also does the
// rerank filter's work for fused keys (TopKView::fused_keys): with q_k the
// k-th smallest 16-bit distance among a query's kept keys, every kept key
// whose q is at most q_k + delta goes to band[query][0, m) as row << 32 (the
// recheck looks the IDs up), and bcnt[query] = m; every kept key when qbad[query] (the query's
// FP16 parts are zero) or fewer than k are kept.  delta = floor(2 eps
// 2^15), eps the fused keys' error bound: the band of launch_rerank.
struct FusedBand {
    uint64_t* band;  // [queries][stride]
    uint32_t* bcnt;
    const uint32_t* qbad;
    int stride;
    int k;
    int delta;
};
void launch_select_merge(const TopKView& v, hipStream_t s, uint32_t* flags = nullptr,
                         const uint32_t* extra = nullptr, int n_extra = 0, const int32_t* row_ids = nullptr,
                         const FusedBand* fb = nullptr);
// Whether a view of K best keys and newcap appends can carry rows in its
// keys (the select then holds every key in registers).
bool row_keys_supported(int K, int newcap);
void launch_sort_rows(const TopKView& v, hipStream_t s);  // orders the best list of every query
void launch_count_overflow(const TopKView& v, uint32_t* out_count, hipStream_t s);
// dst.thr[q] = (integer distance of src.thr[q]) + 1 with ID 0: a scan with it
// appends exactly the pairs at or below that distance.
void launch_threshold_above(const TopKView& src, const TopKView& dst, hipStream_t s);
// flags[q] = 1 when query q overflowed or holds fewer than K keys.
void launch_incomplete_flags(const TopKView& v, uint32_t* flags, hipStream_t s, const uint32_t* extra = nullptr,
                             int n_extra = 0);

// scan_exact.hip: distances of queries [0,Q) to points [p0,p1) of data.
struct ExactScanArgs {
    const float* data;       // [N][d]
    const float* data_inv;   // [N] 1/||x|| (FP32) for kNegCos/kAngular, else null
    const float* qT;         // [ceil(Q/QG)][d][QG] transposed query groups
    const float* q_inv;      // [Q] 1/||q|| (FP32) for kNegCos/kAngular
    int d;
    int Q;
    int64_t p0, p1;
    int kind;
};
int exact_scan_query_group();
// scan_exact_mfma.hip: the angular / -cos exact scan on the FP32 matrix
// cores.  Rows in operand order (launch_tile_rows: tiles of 32 rows,
// exact_mfma_steps(d) steps of 2 dimensions, 64 floats a step, zero padded);
// dinv has a multiple of 32 entries (zero past N); the FP64 finalisation of
// its results takes mfma_error_bound(d).
bool exact_mfma_supported(int kind, int d);
int exact_mfma_steps(int d);
void launch_tile_rows(const float* x, int64_t n, int d, float* out, hipStream_t s);
void launch_exact_mfma(const float* dT, const float* dinv, int d, const float* qB, const float* qinv, int Q,
                       int64_t p0, int64_t p1, int kind, const TopKView& v, hipStream_t s);
double mfma_error_bound(int d);
void launch_exact_scan(const ExactScanArgs& a, const TopKView& v, hipStream_t s);
void launch_transpose_queries(const float* q, int Q, int d, float* qT, hipStream_t s);
void launch_row_inv_norms(const float* x, int64_t n, int d, float* inv, hipStream_t s);

// FP64 finalisation of an exact scan: keys hold the best K' by FP32; writes
// k results per query (id, (float)FP64 distance) in (FP64 distance, id)
// order, and flags[q] = 1 when the FP32 margin does not prove the set exact.
void launch_exact_finalize(const float* data, const float* queries, int d, int kind,
                           const TopKView& v, int k, double eps, lsh_point_t* out,
                           uint32_t* flags, hipStream_t s);

// rerank.hip: candidates = low 32 bits of the first C keys of each query
// (or `cand` ids when non-null, [Q][C]).  Approximate distances with error
// at most eps (FP32 over data, or FP32 over the FP16 image hdata of the unit
// rows, angular kinds only); the candidates within the band that eps leaves
// above the k-th approximate distance are recomputed in FP64, and the k
// best by (FP64 distance, ID) are written.
// qorder (may be null): a permutation of the queries, taken in that order
// per XCD so that queries with shared candidates run close together.
// rowmap (may be null): candidate ID i is row rowmap[i] of data and hdata.
// rb (may be null): workspace for the band; with it and k <= 64 the rerank
// runs as two kernels, the filter (which writes each query's band) and a
// one-wave-per-query FP64 recheck, the filter then needing less LDS.
// Returns whether it split; with recheck_too false the caller launches the
// recheck itself (launch_rerank_recheck, same queries and order).
// rows_in (may be null): the candidates' rows, rows_in[q * rows_stride + i]
// for key i of query q (as the select writes them with row_ids), in place
// of rowmap lookups.
// hdata_elems: the FP16 image's length in halves (0: unknown); an image
// under 4 GiB is read through 32-bit offsets.
// fused_ids (angular kinds, k <= 64 with rb): the keys are fused keys
// (TopKView::fused_keys) whose low word is the candidate's row, and
// fused_ids[row] its ID; the filter takes each candidate's approximate
// distance from its key (no row is read) and eps must be
// fused_error_bound(d).
struct RerankBand {
    DevBuf<uint64_t> band;  // [queries][Cp]: (row << 32 | ID) of the band's members
    DevBuf<uint32_t> cnt;   // [queries]: members
};
bool launch_rerank(const float* data, const float* queries, int d, int kind, const uint64_t* keys,
                   int key_stride, const int32_t* cand, int C, int Q, int k, double eps,
                   lsh_point_t* out, hipStream_t s, const int32_t* qorder = nullptr,
                   const int32_t* rowmap = nullptr, const _Float16* hdata = nullptr, RerankBand* rb = nullptr,
                   bool recheck_too = true, const int32_t* rows_in = nullptr, int rows_stride = 0,
                   size_t hdata_elems = 0, const int32_t* fused_ids = nullptr);
// band_ids (may be null): the band's entries hold rows alone and
// band_ids[row] is the member's ID.
void launch_rerank_recheck(const float* data, const float* queries, int d, int kind, int C, int Q, int k,
                           lsh_point_t* out, hipStream_t s, const int32_t* qorder, RerankBand& rb,
                           const int32_t* band_ids = nullptr);

// FP16 image of npos rows for the angular rerank: row i is x_i * 2^14 /
// ||x_i|| (FP64 norm), rounded to FP16 and zero padded to
// half_row_stride(d) halves; rows with ids[i] < 0 are zero.  Returns false
// (and releases out) when a real row's norm is zero or not finite.
bool build_unit_rows_half(const float* rows, const int32_t* ids, int64_t npos, int d, DevBuf<_Float16>& out,
                          hipStream_t s);
int half_row_stride(int d);
// Error bound of the angular distance computed over that image.
double half_error_bound(int d);
// K-steps of 16 halves of the fused cell scan's products at dimension d
// (0 when d is too large for it), the bound of its approximate angular
// distances (the 16-bit keys included), and the queries' FP16 parts it
// reads: per query 2 x 16 ksteps halves, the high parts of the unit query
// scaled by 2^14, then the low parts (zero past d; all zero for a query
// whose norm is zero, not finite or outside [2^-60, 2^60]).
// After the parts, launch_query_half writes one uint32 per query, 1 when its
// parts are zero (query_half_flags); query_half_elems is the buffer's length
// in halves for Q queries.
int fused_k_steps(int d);
double fused_error_bound(int d);
int fused_band_delta(int d);  // floor(2 fused_error_bound(d) 2^15)
void launch_query_half(const float* q, int Q, int d, int ksteps, _Float16* out, hipStream_t s);
inline size_t query_half_elems(int Q, int ksteps) { return (size_t)Q * 32 * ksteps + 2 * (size_t)Q; }
inline uint32_t* query_half_flags(_Float16* q16, int Q, int ksteps) {
    return reinterpret_cast<uint32_t*>(q16 + (size_t)Q * 32 * ksteps);
}

// simhash_mfma.hip: SimHash bits (32 hyperplanes a word) on the bf16 matrix
// cores with an FP32 recheck of every bit within the error bound of zero;
// bit-identical to launch_simhash.  planes is the host [words*32][d]; the
// device planesT ([words][d][32], FP32) is kept by the caller.  d <= 128.
bool simhash_mfma_supported(int d, int hashBits);
struct SimhashMfma {
    int d = 0, words = 0, KS = 0;
    const float* planesT = nullptr;
    DevBuf<uint16_t> ph, pl;  // bfloat16 high and low parts, matrix-core operand order
    DevBuf<float> pnorm;      // ||p_j||, rounded up
    DevBuf<float> prow;       // [words*32][d] FP32 hyperplanes (the recheck's rows)
    mutable DevBuf<uint64_t> fix;   // the recheck list of a sketch call (grow-only)
    mutable DevBuf<uint32_t> fixn;  // its length
    void build(const float* planes, int d, int words, const float* planesT_dev, hipStream_t s);
    // Grows the recheck list when n needs more than it holds (not while a
    // graph that uses it may replay).
    void sketch(const float* x, int64_t n, uint32_t* out, int out_stride, hipStream_t s) const;
    bool built() const { return words > 0; }
    size_t bytes() const { return ph.bytes() + pl.bytes() + pnorm.bytes() + prow.bytes() + fix.bytes() + fixn.bytes(); }
};

// sketch.hip (out rows have `out_stride` words; planesT is [words][d][hashBits])
void launch_simhash(const float* x, int64_t n, int d, const float* planesT, int words,
                    int hashBits, uint32_t* out, int out_stride, hipStream_t s);
void launch_jl(const float* x, int64_t n, int d, const float* planesT, int sdim, float* out,
               hipStream_t s);
void launch_normalize_rows(float* x, int64_t n, int d, hipStream_t s);
void launch_gj_mins(const float* x, int64_t n, int d, int* mins, hipStream_t s);
void launch_gj_shift(float* x, int64_t n, int d, const int* mins, hipStream_t s);
void launch_crosspoly(const float* x, int64_t n, int d, int tables, const int32_t* signs,
                      uint16_t* keys, hipStream_t s);

// scan_sketch.hip: integer sketch distances of queries [0,Q) to points
// [p0,p1).  kind 0: Hamming over `words` words; kind 1: words minus the
// number of equal words (weighted-MinHash match count).
struct SketchScanArgs {
    const uint32_t* data;   // [N][dwords]
    const uint32_t* qs;     // [ceil(Q/64)*64][words], rows past Q zero
    int words;              // a multiple of 8 (query rows zero padded)
    int Q;
    int64_t p0, p1;
    int kind;
    int dwords = 0;         // words per data row, <= words (0: words); rows of fewer
                            // words, or not 16-byte aligned, are read word by word
};
void launch_sketch_scan(const SketchScanArgs& a, const TopKView& v, hipStream_t s);

// scan_mfma.hip: Hamming scan on the matrix cores (FP4 +-1 operands).  qs
// has ceil(Q/hamming_mfma_query_rows())*hamming_mfma_query_rows() rows; the
// data sketches are the FP4 operand image written by launch_expand_fp4
// (hamming_fp4_words(N, words) uint32 words) from the packed [N][words].
int hamming_mfma_query_rows();
bool hamming_mfma_supported(int words);
size_t hamming_fp4_words(int64_t N, int words);
void launch_expand_fp4(const uint32_t* ds, int64_t N, int words, uint32_t* frag, hipStream_t s);
void launch_hamming_mfma(const uint32_t* frag, const uint32_t* qs, int words, int Q, int64_t p0, int64_t p1,
                         const TopKView& v, hipStream_t s);

// minhash.hip
int wmh_max_draws();
void launch_wmh_bounds(const float* x, int64_t n, int d, int* bounds, hipStream_t s);
void launch_wmh_draws(int nhash, int R, int M, const int* owner, float* draws, int* downer, hipStream_t s);
// mode 0: one count per hash; mode 1: words of hbits bits (nhash = words*hbits).
void launch_wmh_sketch(const float* x, int64_t n, int d, const int* owner, const int* start, int M, int nhash,
                       int mode, int hbits, const uint8_t* bitmap, int bitmap_len, const float* draws,
                       const int* draw_owner, int R, uint32_t* out, int out_stride, hipStream_t s);

// util.hip
// Rows src + 1 .. src + n of `rows` (cols floats each) become copies of row
// src (the padding of a captured batch).
void launch_repeat_row_f32(float* rows, int cols, int src, int n, hipStream_t s);
void launch_gather_rows_f32(const float* src, int cols, const int32_t* idx, int m, float* dst,
                            hipStream_t s);
void launch_gather_rows_u32(const uint32_t* src, int cols, const int32_t* idx, int m, uint32_t* dst,
                            hipStream_t s);
void launch_scatter_rows_u64(const uint64_t* src, int src_stride, int cols, const int32_t* idx, int m,
                             uint64_t* dst, int dst_stride, hipStream_t s);
void launch_scatter_rows_pts(const lsh_point_t* src, int cols, const int32_t* idx, int m,
                             lsh_point_t* dst, hipStream_t s);
void launch_pad_rows_u32(const uint32_t* src, int64_t n, int cols, int64_t npad, int pcols,
                         uint32_t* dst, hipStream_t s);
void launch_fill_i32(int32_t* p, size_t n, int32_t v, hipStream_t s);

// crosspoly_dense.hip: dense-rotation cross-polytope keys (keysImplementation
// 8).  Rotation h = t*K + k is P x d (P = d rounded up to a power of two),
// its entries the reference normal stream rounded to bfloat16; a table's key
// packs K hashes into b = log2(2P)-bit fields.  keys() writes [n][L] and/or
// [L][n] (either may be null) for unit rows x [n][d].  d <= 128.
int dense_key_field_bits(int d);
struct DenseRotations {
    int d = 0, P = 0, dpad = 0, K = 0, L = 0, H = 0, b = 0;
    DevBuf<uint16_t> G;   // [H][P][dpad] bfloat16 bits
    DevBuf<float> gmax;   // [H] largest row norm, rounded up
    void build(int d, int tables, int K, hipStream_t s);
    void keys(const float* x, int64_t n, uint32_t* keys_nl, uint32_t* keys_ln, hipStream_t s) const;
    size_t bytes() const { return G.bytes() + gmax.bytes(); }
};

// lsh.hip: (K,L) bucket tables and the candidate scan.
// keys [n][L] -> keys & mask (in place); u16 keys widened to u32.
void launch_mask_keys(uint32_t* keys, size_t n, uint32_t mask, hipStream_t s);
void launch_widen_keys(const uint16_t* in, uint32_t* out, size_t n, hipStream_t s);
// Counting sort of the N points of every table by key: off is [L][T+1] with
// off[t][key] the absolute start of the bucket in perm ([L][N], table t at
// t*N) and off[t][T] = t*N + N; perm holds the IDs of each bucket (in no
// particular order).  cursor is [L][T+1] scratch.
void launch_bucket_build(const uint32_t* keys, int64_t N, int L, uint32_t T, uint32_t* off, uint32_t* cursor,
                         uint32_t* perm, hipStream_t s);
// Per query: lo[q][t] = start of the query's bucket in table t, pre[q][0..L]
// = prefix sums of the bucket sizes; maxS = max over queries of pre[q][L]
// (maxS must be zero on entry).
void launch_query_ranges(const uint32_t* qkeys, int Q, int L, uint32_t T, const uint32_t* off, uint32_t* lo,
                         uint32_t* pre, uint32_t* maxS, hipStream_t s);
// Sorted tables (any key width): a stable LSD radix sort (8-bit digits over
// the low `bits` bits) of every table's keys keys_ln [L][N] with their point
// IDs: sorted keys skeys [L][N] and IDs perm [L][N] (equal keys in ID order).
// tk, ti: [L][N] scratch; counts: sort_table_counts(N, L) words.
size_t sort_table_counts(int64_t N, int L);
void launch_sort_tables(const uint32_t* keys_ln, int64_t N, int L, int bits, uint32_t* skeys, uint32_t* perm,
                        uint32_t* tk, uint32_t* ti, uint32_t* counts, hipStream_t s);
// launch_query_ranges over sorted tables: the bucket of a key is its run in
// skeys (binary search).  lo holds absolute positions t*N + start; len
// ([Q][L]) receives the run lengths.
void launch_query_ranges_sorted(const uint32_t* qkeys, int Q, int L, int64_t N, const uint32_t* skeys,
                                uint32_t* lo, uint32_t* len, uint32_t* pre, uint32_t* maxS, hipStream_t s);
// [n][L] -> [L][n]
void launch_transpose_keys(const uint32_t* nl, int64_t n, int L, uint32_t* ln, hipStream_t s);

// Candidate positions [p0,p1) of every query row's concatenated L buckets;
// a point is scored only under the first table in which it shares the
// query's key, so the union is visited without duplicates.
struct LshScanArgs {
    const uint32_t* perm;   // [L][N]
    const uint32_t* dkeys;  // [N][L]
    const uint32_t* qkeys;  // [Qall][L]
    const uint32_t* qlo;    // [Qall][L]
    const uint32_t* qpre;   // [Qall][L+1]
    const int32_t* qmap;    // [Q] row -> query, or null for identity
    int L;
    int Q;
    int64_t p0, p1;
    int kind;               // kAngular/kGenJaccard/kEuclid exact; 0/1 integer sketch (with sketch); kSqL2 JL
    bool sketch;            // score by sketch distance
    const float* data;      // exact: [N][d]
    const float* queries;   // exact: [Qall][d]
    int d;
    const uint32_t* dsk;    // integer sketch [N][words]
    const uint32_t* qsk;    // [Qall][words]
    int words;
    const float* djl;       // JL [N][sdim]
    const float* qjl;       // [Qall][sdim]
    int sdim;
};
void launch_lsh_scan(const LshScanArgs& a, const TopKView& v, hipStream_t s);
// The exact rescan (a.sketch false, a.qmap null) of every query's whole
// candidate list without rounds or host reads: positions cut into items of
// lsh_item_positions(), item offsets in ibase ([Q+1]), at most cap items
// kept (KP keys each in partials, [cap][KP]), and each query's best KP (a
// power of two in 32..512, <= v.K) written sorted to v.keys (the rest of
// its K slots empty) with v.thr its KP-th key.  mflags[q] = 1 when query q
// had items past cap (its list is then incomplete).
int lsh_item_positions();
void launch_lsh_exact_worklist(const LshScanArgs& a, int KP, uint32_t* ibase, uint32_t cap, uint64_t* partials,
                               const TopKView& v, uint32_t* mflags, hipStream_t s);

// ---------------------------------------------------------- host helpers

std::vector<float> normal_stream(int64_t n, double stddev);
std::vector<int32_t> sign_stream(int64_t n);
std::vector<uint8_t> bit_stream(int64_t n);

// Phase timing with HIP events on one stream.
class PhaseTimer {
   public:
    explicit PhaseTimer(bool on) : mask_(on ? ~0u : 0u) {}
    ~PhaseTimer();
    PhaseTimer(const PhaseTimer&) = delete;
    PhaseTimer& operator=(const PhaseTimer&) = delete;
    bool on() const { return mask_ != 0; }
    // Records the phases whose bit is set in mask (0: none) and drops what
    // was recorded.
    void enable(uint32_t mask);
    void begin(int phase, hipStream_t s);
    void end(int phase, hipStream_t s);
    // Sum per phase (ms) of the intervals recorded since the last collect;
    // synchronises their events, then recycles them.
    void collect(double* ms, int nphase);

   private:
    hipEvent_t take();
    uint32_t mask_;
    struct Rec {
        int phase;
        hipEvent_t a, b;
    };
    std::vector<Rec> recs_;
    std::vector<std::pair<int, hipEvent_t>> open_;
    std::vector<hipEvent_t> pool_;
};

}  // namespace lshgpu
