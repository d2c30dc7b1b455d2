// engine.hpp -- the resident index and its search paths.
#pragma once

#include <map>
#include <memory>
#include <mutex>
#include <utility>

#include "lshgpu_internal.hpp"

namespace lshgpu {

// One preprocessing step applied in place to data and queries.
// kGjStep: shift by |min(0, floor(min))| per dimension, then unit L2 rows
// (processingUtils.cuh:12-67); kNormStep: unit L2 rows only.
enum StepKind : int { kGjStep = 0, kNormStep = 1 };
struct Transform {
    int kind = kNormStep;
    DevBuf<int> mins;  // kGjStep: per-dimension floor minimum (frozen at build)
};

// Every stride-th row of a packed sketch array (or of the previous level)
// and its FP4 image.
struct SampleLevel {
    DevBuf<uint32_t> ds;
    DevBuf<uint32_t> fp4;
    int64_t N = 0;
};
void build_sample_levels(const uint32_t* ds, int64_t N, int wpad, int C, std::vector<SampleLevel>& out,
                         hipStream_t s);

// Weighted MinHash tables (weightedMinHash.cuh:41-90, 231-262) and the
// sketch of rows with them.  mode 0: one red count per hash; mode 1: words of
// hbits bits, bit b of word w from hash w*hbits + b.
struct WmhHasher {
    int mode = 0;
    int nhash = 0;
    int hbits = 32;
    int M = 0;
    int R = 0;
    int bitmap_len = 0;
    DevBuf<int> owner;       // [M] dimension of each unit of mass
    DevBuf<int> start;       // [d] first unit of each dimension
    DevBuf<uint8_t> bitmap;  // bit of each red count (one-bit mode)
    DevBuf<float> draws;     // [nhash][R] first draws r = M*U of each hash
    DevBuf<int> draw_owner;  // [nhash][R] dimension owning floor(r)

    // Bounds from the rows of x (and extra, when given).
    void build(const float* x, int64_t n, const float* extra, int64_t n_extra, int d, int mode_, int nhash_,
               int hbits_, hipStream_t s);
    // out[i][*] with row stride out_stride (words for mode 1, hashes for mode 0).
    void sketch(const float* x, int64_t n, int d, uint32_t* out, int out_stride, hipStream_t s) const;
    size_t bytes() const {
        return owner.bytes() + start.bytes() + bitmap.bytes() + draws.bytes() + draw_owner.bytes();
    }
};

// Device scratch reused across searches (grow-only).  A call at recursion
// depth `depth` uses raw slots 4*depth .. 4*depth+3 and top-K state `depth`.
struct Scratch {
    static constexpr int kDepth = 8;
    DevBuf<uint8_t> raw[4 * kDepth];
    TopKState tk[kDepth];
    // Pinned host words (grow-only) and an event, for readbacks that the
    // host waits on while later work is already queued behind them.
    uint32_t* host = nullptr;
    size_t host_n = 0;
    hipEvent_t ev = nullptr;
    Scratch() = default;
    Scratch(const Scratch&) = delete;
    Scratch& operator=(const Scratch&) = delete;
    ~Scratch() {
        if (host) (void)hipHostFree(host);
        if (ev) (void)hipEventDestroy(ev);
    }
    uint32_t* pinned(size_t n) {
        if (n > host_n) {
            if (host) {
                (void)hipDeviceSynchronize();  // no copy into the old buffer is still queued
                (void)hipHostFree(host);
                host = nullptr;
                host_n = 0;
            }
            void* p = nullptr;
            LSH_HIP(hipHostMalloc(&p, n * sizeof(uint32_t), hipHostMallocDefault));
            host = static_cast<uint32_t*>(p);
            host_n = n;
        }
        return host;
    }
    hipEvent_t event() {
        if (!ev) LSH_HIP(hipEventCreateWithFlags(&ev, hipEventDisableTiming));
        return ev;
    }
    template <class T>
    T* get(int depth, int slot, size_t n) {
        LSH_REQUIRE(depth < kDepth, LSH_ERR_INTERNAL, "scratch depth exceeded");
        DevBuf<uint8_t>& b = raw[4 * depth + slot];
        b.ensure(std::max<size_t>(n, 1) * sizeof(T));
        return reinterpret_cast<T*>(b.get());
    }
    size_t bytes() const {
        size_t n = 0;
        for (const auto& b : raw) n += b.bytes();
        for (const auto& t : tk) n += t.bytes();
        return n;
    }
};

// Pivot cells of a packed Hamming sketch array (cells.hip): the data in
// 32-point tiles grouped by nearest pivot and sorted by distance to it, with
// each tile's distance range, so that a pass can skip the tiles the triangle
// inequality rules out.  Built for kCellMinPoints points or more.
constexpr int64_t kCellMinPoints = 65536;
int cell_pivots_for(int64_t N);

struct CellIndex {
    int P = 0;       // pivots; 0 when no cells were built
    int words = 0;   // sketch words per row (8, 16, 24 or 32)
    int64_t ntiles = 0;
    int64_t nparts = 0;            // sum over cells of their tile ranges (chunks of one query group)
    DevBuf<uint32_t> pfp4;         // FP4 image of the pivot sketches
    DevBuf<uint32_t> sk;           // [ntiles*32][W] packed sketches in cell order, even words then odd
    DevBuf<int32_t> ids;           // [ntiles*32] data row of each position, -1 for padding
    DevBuf<uint32_t> tinfo;        // per tile: min | max << 11 | rows << 22 (distances to the pivot)
    DevBuf<uint32_t> ctile;        // [P+1] first tile of each cell
    DevBuf<uint32_t> csize;        // [P] points of each cell
    DevBuf<uint16_t> cmin, cmax;   // per cell: distance range to the pivot (min > max: empty)
    DevBuf<uint2> crec;            // per cell: (size, min | max << 16)
    void build(const uint32_t* ds, int64_t N, int wpad, hipStream_t s);
    size_t bytes() const;
    bool built() const { return P > 0; }
};

// The float rows [N][d] in the cells' order (rows: [ntiles*32][d], zero
// for padding) and pos[id] = the position of data row id.
void cell_rows(const CellIndex& c, const float* data, int64_t N, int d, DevBuf<float>& rows, DevBuf<int32_t>& pos,
               hipStream_t s);

// Threshold search over the cells (Q query rows of qs, top-K): clears the
// fresh view v (as TopKState::clear; v.Q == Q) in its first launch, sets
// v.thr to a valid threshold from each query's nearest cells, then appends every
// pair below it to v, as launch_hamming_mfma over all points would.  Nothing
// waits on the host.  Returns a device array stat[3]: stat[0] = 1 when the
// thresholds left too many tiles for pruning to pay and nothing was appended
// (v.thr is still valid; the caller then runs the full pass); stat[1] and
// stat[2] = the tile products of the two passes (32-query units), an upper
// bound.  Uses scratch slots 0-3 of `depth`.  qorder (may be null; Q
// entries) receives the queries grouped by their nearest cell.
// fuse (may be null; needs v.row_keys and v.fused_keys): the appending pass
// also writes each pair's approximate angular distance into its key (see
// TopKView::fused_keys) from the FP16 image of the unit rows in cell order
// (chalf, hs halves a row) and the queries' FP16 parts (q16,
// launch_query_half with 16 ksteps halves a part).
struct CellFuse {
    const _Float16* chalf;
    int hs;
    const _Float16* q16;
    int ksteps;
};
// pt (may be null) times the appending pass as phase `phase`.
const uint32_t* cell_search(const CellIndex& c, const uint32_t* qs, int Q, int K, const TopKView& v, Scratch& sc,
                            int depth, hipStream_t s, int32_t* qorder, const CellFuse* fuse = nullptr,
                            PhaseTimer* pt = nullptr, int phase = 0);
// Sizes the scratch slots of depth `depth` for a cell search of Q queries,
// so that searches of up to Q queries allocate nothing (graph capture).
void cell_search_reserve(const CellIndex& c, int Q, Scratch& sc, int depth);

}  // namespace lshgpu

// The per-search buffers of the exact (K,L) rescan: the index's own, or a
// batch slot's.
struct lsh_work {
    lshgpu::DevBuf<float>& q;
    lshgpu::DevBuf<uint32_t>& qkeys;
    lshgpu::DevBuf<uint32_t>& qlo;
    lshgpu::DevBuf<uint32_t>& qpre;
    lshgpu::DevBuf<uint32_t>& qmax;
    lshgpu::DevBuf<uint32_t>& qlen;
    lshgpu::DevBuf<uint32_t>& wl_ibase;
    lshgpu::DevBuf<uint32_t>& wl_flags;
    lshgpu::DevBuf<uint64_t>& wl_part;
    lshgpu::DevBuf<uint32_t>& flags;
    lshgpu::TopKState& st;
    lsh_point_t* outp;
};

// One batch slot of the captured batch searches: a stream of its own
// (default priority), the buffers its graphs read and write, and the graphs
// captured over them, one per (query-count class, rounded k).  Batches dealt
// to different slots run concurrently.  The SimHash index with pivot cells
// uses the sketch buffers, the (K,L) rescan the lsh ones.
struct batch_slot {
    hipStream_t stream = nullptr;
    lshgpu::DevBuf<float> q;
    lshgpu::DevBuf<lsh_point_t> out;
    lshgpu::TopKState st;
    lshgpu::DevBuf<uint32_t> flags;  // sketch: the select's flags, then the cell search's three status words
    // sketch
    lshgpu::DevBuf<uint32_t> qsk;
    bool qsk_zeroed = false;
    lshgpu::Scratch scratch;
    lshgpu::DevBuf<int32_t> qorder;
    lshgpu::RerankBand rband;
    lshgpu::DevBuf<_Float16> q16;  // the queries' FP16 parts of the fused cell scan
    // (K,L) rescan
    lshgpu::DevBuf<uint32_t> qkeys, qlo, qpre, qmax, qlen, wl_ibase, wl_flags;
    lshgpu::DevBuf<uint64_t> wl_part;
    std::map<std::pair<int, int>, hipGraphExec_t> graphs;
    std::vector<uintptr_t> sig;  // the buffers' layout when the graphs were captured
    batch_slot() = default;
    batch_slot(const batch_slot&) = delete;
    batch_slot& operator=(const batch_slot&) = delete;
    ~batch_slot();
    void drop_graphs();
    size_t bytes() const;
    lsh_work work() {
        return {q, qkeys, qlo, qpre, qmax, qlen, wl_ibase, wl_flags, wl_part, flags, st, out.get()};
    }
};

struct lsh_index {
    // Held by every entry point that uses the index (searches share its
    // scratch), so a handle may be used from several threads.
    mutable std::mutex mu;
    // Resolved configuration.
    int impl = 2;
    int dfunc = 1;
    int d = 0;
    int64_t N = 0;
    int sketchDim = 0;
    int framework = 0;
    int keysImpl = 0;
    int bucketKeyBits = 0;
    int tables = 0;
    bool withSketch = false;
    int candidates = 4096;
    int margin = 32;
    int device = 0;

    hipStream_t stream = nullptr;
    // Phase timing while profiling is on: intervals and counters summed over
    // the searches since the last read.
    lshgpu::PhaseTimer timer{false};
    int64_t prof_mfma_launches = 0;
    int64_t prof_mfma_points = 0;
    int64_t prof_pass2_tiles = 0;

    // Working copy of the data after the path's preprocessing.
    lshgpu::DevBuf<float> data;
    lshgpu::DevBuf<float> data_inv;  // 1/||x|| for the angular kinds
    lshgpu::DevBuf<float> dtile;     // rows in matrix-core operand order (exact angular scans)

    // Sketches.
    int words = 0;   // uint32 words per point
    int wpad = 0;    // words padded to a multiple of 8
    lshgpu::DevBuf<uint32_t> dsk;    // [N][wpad]
    lshgpu::DevBuf<uint32_t> dfp4;   // FP4 operand image of dsk for the matrix-core Hamming scan
    std::vector<lshgpu::SampleLevel> samples;  // strided samples of dsk (threshold estimates)
    lshgpu::CellIndex cells;         // pivot cells of dsk (pruned threshold pass)
    lshgpu::DevBuf<float> cdata;     // data rows in cell order (rerank reads)
    lshgpu::DevBuf<int32_t> cpos;    // position of each data row in cdata
    lshgpu::DevBuf<_Float16> chalf;  // FP16 image of the unit rows of cdata (angular rerank filter)
    lshgpu::DevBuf<_Float16> q16;    // per search: the queries' FP16 parts for the fused cell scan
    lshgpu::RerankBand rband;        // the rerank's band workspace (grow-only)
    // Recorded at the end of each search: the next search's stream waits on it,
    // since a search may return while its last kernels still run.
    hipEvent_t done_ev = nullptr;
    bool done_rec = false;
    lshgpu::DevBuf<float> planesT;   // SimHash [words][d][32], JL [groups][d][32]
    int projection = 0;              // 0: SimHash on the matrix cores (FP32 recheck), 1: FP32 only
    lshgpu::SimhashMfma smf;         // the matrix-core SimHash of planesT, when used
    int sdim = 0;                    // JL sketch length
    lshgpu::DevBuf<float> djl;       // [N][sdim]

    // Weighted MinHash hashers: `wmh` for the sketches, `kwmh` for bucket keys.
    lshgpu::WmhHasher wmh, kwmh;

    // (K,L) bucket tables (framework 1).
    int keyBitsMask = 0;                  // key width of the reference's key type (8 or 16 bits)
    uint32_t tableSize = 0;               // 2^bucketKeyBits
    std::vector<lshgpu::Transform> key_steps, sketch_steps;  // preprocessing before keys / sketches
    lshgpu::DevBuf<float> kplanesT;       // SimHash key hyperplanes [tables][d][bucketKeyBits]
    lshgpu::DevBuf<int32_t> ksigns;       // cross-polytope diagonals [tables][3][P]
    lshgpu::DenseRotations dense;         // dense-rotation cross-polytope keys (keysImplementation 8)
    bool sorted = false;                  // tables as sorted (key, ID) runs instead of 2^bits offsets
    int sortBits = 0;                     // key bits the sorted tables are ordered by
    lshgpu::DevBuf<uint32_t> bkeys;       // sorted tables: [tables][N] keys in ascending order (bperm alongside)
    lshgpu::DevBuf<uint32_t> dkeys;       // [N][tables]
    lshgpu::DevBuf<uint32_t> boff;        // [tables][tableSize+1] absolute positions into bperm
    lshgpu::DevBuf<uint32_t> bperm;       // [tables][N] point IDs grouped by key
    lshgpu::DevBuf<uint32_t> qkeys, qlo, qpre, qmax, qlen;
    lshgpu::DevBuf<uint32_t> wl_ibase, wl_flags;  // exact rescan work list: item offsets, capacity flags
    lshgpu::DevBuf<uint64_t> wl_part;             // its per-item best lists
    int64_t last_candidates = 0;          // sum over queries of bucket entries walked (last search)

    // Per-search scratch (grow-only).
    lshgpu::DevBuf<float> q, qT, q_inv, qjl, qjlT;
    lshgpu::DevBuf<uint32_t> qsk;
    const uint32_t* qsk_zeroed = nullptr;  // qsk was cleared at this address for qsk_zeroed_q queries
    int qsk_zeroed_q = -1;
    lshgpu::DevBuf<lsh_point_t> out;
    // The current search's queries (the caller's device buffer, or q) and
    // results (the caller's device buffer, or out).
    const float* qp = nullptr;
    lsh_point_t* outp = nullptr;
    lshgpu::DevBuf<uint32_t> flags;
    lshgpu::DevBuf<int32_t> qorder;  // queries grouped by nearest cell (rerank order)
    lshgpu::TopKState st;
    lshgpu::Scratch scratch;
    int64_t scan_launches = 0;
    int64_t mfma_launches = 0;   // matrix-core Hamming scan launches of the last search
    int64_t mfma_points = 0;     // data points those launches covered
    int64_t pass2_tiles = 0;     // the pivot-cell appending pass's tile products (last search)

    // Batched search through captured graphs (search_batches): batch slots
    // (created by the first batch call) whose graphs cover query-count
    // classes up to kGraphMaxQ; whether a batch of the SimHash index found
    // the cells' second pass too dense (later batch calls then run ordinary
    // searches, whose fallback handles that).
    static constexpr int kGraphMaxQ = 4096;
    static constexpr int kGraphMaxK = 128;  // rounded k of the SimHash index's graphs
    static constexpr int kBatchSlots = 4;
    std::vector<std::unique_ptr<batch_slot>> slots;
    int next_slot = 0;               // the slot of the next batch
    std::vector<int64_t> slots_key;  // the layout the slots are sized for
    bool cells_dense = false;
    uint32_t* bflags = nullptr;  // pinned: per-query flags of a batch call
    size_t bflags_n = 0;
    // The last batch call: batches that replayed a graph, queries redone.
    int64_t batch_graphs = 0, batch_redone = 0;

    ~lsh_index();
    size_t device_bytes() const;
    void build(const float* src, bool on_device, const float* extra_host, int64_t n_extra);
    void search(const float* queries, int Q, int k, bool q_on_device, lsh_point_t* results,
                bool r_on_device, hipStream_t s);
    // n_batches searches of consecutive slices of `queries` (sizes[i] queries
    // each) into consecutive rows of `results`, each as its own search; with
    // one host wait for all of them.
    void search_batches(const float* queries, const int* sizes, int n_batches, int k, bool q_on_device,
                        lsh_point_t* results, bool r_on_device, hipStream_t s);
    void drop_graphs();
    // One batch of n <= kGraphMaxQ queries of the SimHash index with pivot
    // cells, enqueued on the next batch slot with no host wait (for a
    // request loop; host buffers should be pinned): it starts after `ready`
    // (may be null), its results go to res (n x calculateK(k)) and its flags
    // to flags_out (pinned host: n incomplete-query flags, then the cell
    // search's three status words, the first nonzero when the second pass was
    // too dense); `done` (may be null) is recorded once all are complete.  A
    // flagged query, or every query of a too-dense batch, needs an ordinary
    // search.  Returns false, enqueuing nothing, when the index has no
    // captured form for the batch.
    bool enqueue_batch(const float* q, bool q_on_device, int n, int k, lsh_point_t* res, bool r_on_device,
                       uint32_t* flags_out, hipEvent_t ready, hipEvent_t done);

   private:
    void build_fp4(hipStream_t s);
    void build_lsh(const float* extra, int64_t n_extra, hipStream_t s);
    void lsh_keys(const float* x, int64_t n, uint32_t* keys, hipStream_t s);
    // SimHash words of rows x (32 hyperplanes a word) by smf when it is built.
    void sketch_simhash(const float* x, int64_t n, uint32_t* out, int stride, hipStream_t s) const;
    void build_simhash_planes(hipStream_t s);
    bool lsh_graphable(int kq) const;
    // The index's own buffers for search_lsh.
    lsh_work own_work() {
        return {q, qkeys, qlo, qpre, qmax, qlen, wl_ibase, wl_flags, wl_part, flags, st, outp};
    }
    void search_lsh(const lsh_work& w, int Q, int kq, hipStream_t s, lshgpu::PhaseTimer& pt,
                    bool enqueue_only = false);
    // The SimHash index's search of Qc queries held in a slot's q, enqueued
    // with no host wait: sketch, cell search, select (flags to the slot's
    // flags) and rerank (results to its out).
    bool sketch_graphable(int kq) const;
    bool fuse_ok(int kq) const;
    void enqueue_sketch_search(batch_slot& sl, int Qc, int kq);
    // Creates the batch slots on first use and sizes every buffer their
    // graphs touch for the largest class (nothing is allocated while
    // capturing); graphs captured over another layout are dropped.
    void prepare_slots(bool sketch, int kq);
    // One batch of b <= kGraphMaxQ queries (q, copied with qk) through slot
    // sl's graph of its class (captured on first use): its results to res
    // (rk), its flags to flags_out and to extra_out (sketch: the cell
    // search's three status words; (K,L): the work list's capacity flags),
    // both pinned host memory; all on the slot's stream, with no host wait.
    void enqueue_slot_batch(batch_slot& sl, bool sketch, const float* q, hipMemcpyKind qk, int b, int kq,
                            lsh_point_t* res, hipMemcpyKind rk, uint32_t* flags_out, uint32_t* extra_out);
    // The batches through the slots' graphs (sketch: the SimHash index path,
    // else the (K,L) rescan).
    void search_batches_graphs(bool sketch, const float* queries, const int* sizes, int n_batches, int k,
                               int64_t total, bool q_on_device, lsh_point_t* results, bool r_on_device,
                               hipStream_t s);
};

namespace lshgpu {

// One-shot helpers behind the reference-shaped entry points.
int calculate_k(int k);

// Sketch-stage entry points (host in, host out).
void simhash_host(const float* x, int64_t n, int d, int words, int hashBits, const float* planes,
                  uint32_t* out, int device);
void jl_host(const float* x, int64_t n, int d, int sdim, const float* planes, float* out, int device);
void gj_transform_host(float* data, int64_t N, float* queries, int64_t Q, int d, int device);
void hamming_topc_host(const uint32_t* ds, int64_t N, const uint32_t* qs, int64_t Q, int words, int C,
                       int32_t* ids, int32_t* dists, int device);
void wmh_host(const float* x, int64_t n, int d, int sketchDim, int one_bit, uint32_t* out, int device);
// Device sketches in, device IDs and distances out, working memory from the
// per-GPU scratch array of noalloc.hip (no allocation; on the current device).
void hamming_topc_device_noalloc(const uint32_t* ds, int64_t N, const uint32_t* qs, int64_t Q, int words, int C,
                                 int32_t* ids, int32_t* dists, hipStream_t s);
void crosspoly_host(const float* x, int64_t n, int d, int tables, const int32_t* signs, uint16_t* keys,
                    int device);
void crosspoly_dense_host(const float* x, int64_t n, int d, int tables, int K, uint32_t* keys, int device);
void sort_tables_host(const uint32_t* keys_ln, int64_t N, int L, int bits, uint32_t* skeys, uint32_t* ids,
                      int device);

// Statistics (stats.hip).  summary4: buckets, non-empty buckets, points, largest bucket.
void byte_distribution_host(const uint8_t* hashes, int64_t n, int64_t* counts, int device);
void full_key_distribution_host(const uint8_t* hashes, int64_t n, int sd, std::vector<uint8_t>& keys,
                                std::vector<int64_t>& counts, int device);
void index_bucket_stats(lsh_index* idx, int table, int64_t* summary4, int64_t* hist, int hist_len);

}  // namespace lshgpu
