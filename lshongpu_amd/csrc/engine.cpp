// engine.cpp -- resident index: preprocessing, sketch builds and search.
//
// Search is organised in rounds (see select.hip): a scan kernel sweeps a
// range of data points for every query of the batch and appends the pairs
// that beat the query's current threshold; a select kernel folds them into
// the query's best-K list and tightens the threshold.  Range sizes grow
// geometrically so that each round is expected to append about half of the
// append capacity per query; a query whose appends overflow is re-run on its
// own with ranges no larger than the capacity, which cannot overflow.
#include <functional>

#include "engine.hpp"

#include <algorithm>
#include <cmath>
#include <cstdlib>

using namespace lshgpu;

namespace {

enum Phase {
    PH_PREP = 0,
    PH_SCAN = 1,
    PH_SELECT = 2,
    PH_FINAL = 3,
    PH_TOTAL = 4,
    PH_MFMA = 5,
    PH_RECHECK = 6,
    PH_PASS2 = 7  // the pivot-cell search's appending pass
};

struct DeviceGuard {
    int prev = 0;
    explicit DeviceGuard(int dev) {
        LSH_HIP(hipGetDevice(&prev));
        if (prev != dev) LSH_HIP(hipSetDevice(dev));
    }
    ~DeviceGuard() { (void)hipSetDevice(prev); }
};

int round_up(int x, int m) { return (x + m - 1) / m * m; }

bool angular_kind(int kind) { return kind == kNegCos || kind == kAngular; }

// LSHGPU_VALU_HAMMING=1 selects the VALU Hamming scan instead of the matrix
// cores (both give identical results; used to compare them).
bool force_valu_scan() {
    static const bool on = [] {
        const char* e = std::getenv("LSHGPU_VALU_HAMMING");
        return e && e[0] == '1';
    }();
    return on;
}

// LSHGPU_FULL_PASS=1 makes the threshold pass cover every point instead of
// the pivot cells (both give identical results; used to compare them).
bool force_full_pass() {
    static const bool on = [] {
        const char* e = std::getenv("LSHGPU_FULL_PASS");
        return e && e[0] == '1';
    }();
    return on;
}

// Indices of the nonzero entries of a device flag array (and, with extra,
// n_extra more device words copied in the same synchronisation).
std::vector<int32_t> flagged(const uint32_t* dflags, int Q, hipStream_t s, const uint32_t* extra = nullptr,
                             uint32_t* h_extra = nullptr, int n_extra = 0) {
    std::vector<uint32_t> h((size_t)Q);
    LSH_HIP(hipMemcpyAsync(h.data(), dflags, sizeof(uint32_t) * (size_t)Q, hipMemcpyDeviceToHost, s));
    if (extra) LSH_HIP(hipMemcpyAsync(h_extra, extra, sizeof(uint32_t) * n_extra, hipMemcpyDeviceToHost, s));
    LSH_HIP(hipStreamSynchronize(s));
    std::vector<int32_t> idx;
    for (int i = 0; i < Q; ++i)
        if (h[i]) idx.push_back(i);
    return idx;
}

// Indices q with a[q] or b[q] nonzero (both device arrays of Q words).
std::vector<int32_t> flagged_either(const uint32_t* a, const uint32_t* b, int Q, hipStream_t s) {
    std::vector<uint32_t> ha((size_t)Q), hb((size_t)Q);
    LSH_HIP(hipMemcpyAsync(ha.data(), a, sizeof(uint32_t) * (size_t)Q, hipMemcpyDeviceToHost, s));
    LSH_HIP(hipMemcpyAsync(hb.data(), b, sizeof(uint32_t) * (size_t)Q, hipMemcpyDeviceToHost, s));
    LSH_HIP(hipStreamSynchronize(s));
    std::vector<int32_t> idx;
    for (int i = 0; i < Q; ++i)
        if (ha[i] || hb[i]) idx.push_back(i);
    return idx;
}

// Items the exact rescan's work list keeps per search (256 MiB of best
// lists).  LSHGPU_WORKLIST_CAP overrides it (a small value exercises the
// re-run of the queries past it).
uint32_t worklist_cap(int K) {
    const char* e = std::getenv("LSHGPU_WORKLIST_CAP");
    if (e && *e) return (uint32_t)std::max(1L, std::strtol(e, nullptr, 10));
    return (uint32_t)((size_t(256) << 20) / (sizeof(uint64_t) * K));
}

// LSHGPU_EXACT_VALU=1 at index build keeps the exact angular scans on the
// vector units (the FP32 FMA chain) instead of the matrix cores.
bool exact_valu_forced() {
    const char* e = std::getenv("LSHGPU_EXACT_VALU");
    return e && e[0] == '1';
}

// LSHGPU_LSH_ROUNDS=1 runs the exact (K,L) rescan in rounds instead of the
// work list (both give identical results; used to compare them).
bool force_rounds_lsh() {
    static const bool on = [] {
        const char* e = std::getenv("LSHGPU_LSH_ROUNDS");
        return e && e[0] == '1';
    }();
    return on;
}

// Rounds over points [0, N) (see the file comment).  `scan(p0, p1, view)`
// launches the scan kernel of one range.  kNormal sizes the ranges for
// about newcap/2 appends per query, kCautious for newcap/8 (the re-run of
// queries that overflowed), kFixed keeps every range at newcap points, which
// cannot overflow.
enum RoundsMode : int { kNormal = 0, kCautious = 1, kFixed = 2 };

template <class ScanFn>
void run_rounds(TopKState& st, int Q, int K, int64_t N, int mode, ScanFn&& scan, hipStream_t s,
                PhaseTimer& pt, int64_t& launches, bool int_keys = false) {
    const int newcap = newcap_for(K);
    st.reset(Q, K, newcap, s, int_keys);
    const TopKView v = st.view();
    int64_t p = 0;
    int64_t chunk = std::min<int64_t>(N, newcap);
    const double target = mode == kCautious ? 8.0 : 2.0;
    while (p < N) {
        const int64_t p1 = std::min<int64_t>(N, p + chunk);
        pt.begin(PH_SCAN, s);
        scan(p, p1, v);
        pt.end(PH_SCAN, s);
        ++launches;
        pt.begin(PH_SELECT, s);
        launch_select_merge(v, s);
        pt.end(PH_SELECT, s);
        if (mode != kFixed)
            chunk = std::max<int64_t>(std::min<int64_t>(chunk, newcap),
                                      (int64_t)((double)p1 * newcap / (target * K)));
        p = p1;
    }
}

// Prepared queries of an exact-distance scan over `data` (rows of length d).
struct ExactQueries {
    const float* rows;  // [Q][d]
    int Q;
};

// `dtile` (the rows in matrix-core operand order, or null) selects the
// matrix-core scan of the angular kinds; its results take mfma_error_bound.
void exact_topk(const float* data, const float* data_inv, const float* dtile, int64_t N, int d, int kind,
                const ExactQueries& eq, int K, int mode, TopKState& st, Scratch& sc, int depth, hipStream_t s,
                PhaseTimer& pt, int64_t& launches) {
    float* qinv = sc.get<float>(depth, 1, (size_t)eq.Q);
    if (angular_kind(kind)) launch_row_inv_norms(eq.rows, eq.Q, d, qinv, s);
    if (dtile) {
        float* qB = sc.get<float>(depth, 0, (size_t)(eq.Q + 31) / 32 * exact_mfma_steps(d) * 64);
        launch_tile_rows(eq.rows, eq.Q, d, qB, s);
        auto scan = [&](int64_t p0, int64_t p1, const TopKView& v) {
            launch_exact_mfma(dtile, data_inv, d, qB, qinv, eq.Q, p0, p1, kind, v, s);
        };
        run_rounds(st, eq.Q, K, N, mode, scan, s, pt, launches);
    } else {
        const int QG = exact_scan_query_group();
        const int groups = (eq.Q + QG - 1) / QG;
        float* qT = sc.get<float>(depth, 0, (size_t)groups * d * QG);
        launch_transpose_queries(eq.rows, eq.Q, d, qT, s);
        auto scan = [&](int64_t p0, int64_t p1, const TopKView& v) {
            ExactScanArgs a{data, data_inv, qT, qinv, d, eq.Q, p0, p1, kind};
            launch_exact_scan(a, v, s);
        };
        run_rounds(st, eq.Q, K, N, mode, scan, s, pt, launches);
    }
    if (mode == kFixed) return;
    const std::vector<int32_t> ovf = flagged(st.view().ovf, eq.Q, s);
    if (ovf.empty()) return;
    const int m = (int)ovf.size();
    int32_t* idx = sc.get<int32_t>(depth, 2, (size_t)m);
    LSH_HIP(hipMemcpyAsync(idx, ovf.data(), sizeof(int32_t) * m, hipMemcpyHostToDevice, s));
    float* sub = sc.get<float>(depth, 3, (size_t)m * d);
    launch_gather_rows_f32(eq.rows, d, idx, m, sub, s);
    TopKState& sst = sc.tk[depth];
    exact_topk(data, data_inv, dtile, N, d, kind, ExactQueries{sub, m}, K, mode == kNormal ? kCautious : kFixed, sst,
               sc, depth + 1, s, pt, launches);
    const TopKView sv = sst.view(), v = st.view();
    launch_scatter_rows_u64(sv.keys, sv.stride, K, idx, m, v.keys, v.stride, s);
}

// Packed integer sketches [N][wpad] and, for Hamming scans, their FP4 image.
struct SketchSet {
    const uint32_t* ds;
    const uint32_t* fp4;
    int64_t N;
    const CellIndex* cells = nullptr;  // pivot cells of ds, for the threshold pass
    int32_t* qorder = nullptr;         // out: queries grouped by nearest cell, when the cells ran
    bool* ordered = nullptr;           // out: qorder was written
    // Work to enqueue once the threshold pass and its select are enqueued,
    // before the host waits for the incomplete-query flags (the rerank, so
    // that it does not wait on the host).  `ran` tells whether it was
    // enqueued over the final candidates; `redone` lists the queries whose
    // candidates changed afterwards.
    struct Hook {
        std::function<void()> fn;
        bool ran = false;
        std::vector<int32_t> redone;
        // The pivot-cell layout's IDs: with them the cell scan appends rows
        // in place of IDs, and the select leaves each kept key's row beside
        // the best list (rows = true for fn).
        const int32_t* row_ids = nullptr;
        bool rows = false;
        // With fuse (and row_ids), the cell scan also writes each appended
        // pair's approximate angular distance into its key (fused = true
        // for fn when it did).
        const CellFuse* fuse = nullptr;
        bool fused = false;
        // Out: the appending pass's tile products (32 x 32 pairs each).
        int64_t pass2_tiles = 0;
        // With fused keys the select also writes the rerank's band (band
        // may be null); banded = true for fn when it did.
        const FusedBand* band = nullptr;
        bool banded = false;
    }* hook = nullptr;
};

// Rows of a level per row of its sample.
int sample_step(int64_t n, int64_t sample_n) { return (int)((n + sample_n - 1) / sample_n); }

// Sample candidates Ks for K over a level sampled every `stride` rows: with
// Ks*stride points expected at or below the sample's Ks-th distance and a
// standard deviation of about sqrt(Ks)*stride, Ks solves
// Ks*stride = K + 4*sqrt(Ks)*stride, so fewer than K points is a 4-sigma
// event (the query is then redone).
int sample_k(int K, int stride) {
    const double r = (double)K / stride;
    const double x = (4.0 + std::sqrt(16.0 + 4.0 * r)) / 2.0;
    return std::min(kMaxIntTopK, std::max(1, (int)std::ceil(x * x)));
}

// Rows the query sketch buffer of an integer sketch scan must have.
int sketch_query_rows(int Q) { return round_up(std::max(Q, 1), 256); }

// Integer sketch scan: qs has sketch_query_rows(Q) rows of wpad words (zero
// past Q).  Hamming scans of up to 1024 bits run on the matrix cores.
//
// With strided samples of the data (samples[0] every stride-th row of d,
// samples[1] of samples[0], ...), a Hamming scan first finds the exact
// Ks-th nearest sample distance T of every query (Ks from sample_k, the
// same way one level down), then makes one pass over all points appending
// the pairs at distance <= T (about Ks*stride per query instead of the
// ~K*ln(N/K) of the geometric rounds), and selects the K best of them.  A
// query whose pass overflowed, or found fewer than K pairs, is redone with
// the rounds, so the result is the exact top-K either way.
void sketch_topk(const SketchSet& d, const SketchSet* samples, int nsamples, int wpad, int kind,
                 const uint32_t* qs, int Q, int K, int mode, TopKState& st, Scratch& sc, int depth, hipStream_t s,
                 PhaseTimer& pt, int64_t& launches, int64_t* mfma_launches = nullptr,
                 int64_t* mfma_points = nullptr) {
    const bool mfma = kind == 0 && hamming_mfma_supported(wpad) && !force_valu_scan();
    const bool int_keys = K <= kMaxIntTopK;
    // Re-runs of the queries in `bad` (rows of qs) with the rounds in `rmode`,
    // scattered back into st.
    auto redo = [&](const std::vector<int32_t>& bad, int rmode) {
        const int m = (int)bad.size();
        const int mpad = sketch_query_rows(m);
        std::vector<int32_t> hidx(bad);
        hidx.resize((size_t)mpad, bad[0]);
        int32_t* idx = sc.get<int32_t>(depth, 1, (size_t)mpad);
        LSH_HIP(hipMemcpyAsync(idx, hidx.data(), sizeof(int32_t) * mpad, hipMemcpyHostToDevice, s));
        uint32_t* sub = sc.get<uint32_t>(depth, 2, (size_t)mpad * wpad);
        launch_gather_rows_u32(qs, wpad, idx, mpad, sub, s);
        TopKState& sst = sc.tk[depth];
        sketch_topk(SketchSet{d.ds, rmode == kFixed ? nullptr : d.fp4, d.N}, nullptr, 0, wpad, kind, sub, m, K,
                    rmode, sst, sc, depth + 1, s, pt, launches, mfma_launches, mfma_points);
        const TopKView sv = sst.view(), v = st.view();
        launch_scatter_rows_u64(sv.keys, sv.stride, K, idx, m, v.keys, v.stride, s);
    };
    if (mfma && mode == kNormal && nsamples > 0 && samples[0].N > 0 && d.fp4 && int_keys) {
        // Threshold from the sample, then one pass over all points.
        auto sampled_pass = [&]() {
            const int Ks = sample_k(K, sample_step(d.N, samples[0].N));
            TopKState& sa = sc.tk[depth];
            sketch_topk(samples[0], samples + 1, nsamples - 1, wpad, kind, qs, Q, Ks, kNormal, sa, sc, depth + 1, s,
                        pt, launches, mfma_launches, mfma_points);
            launch_threshold_above(sa.view(), st.view(), s);
            pt.begin(PH_SCAN, s);
            pt.begin(PH_MFMA, s);
            launch_hamming_mfma(d.fp4, qs, wpad, Q, 0, d.N, st.view(), s);
            pt.end(PH_MFMA, s);
            pt.end(PH_SCAN, s);
            ++launches;
            if (mfma_launches) ++*mfma_launches;
            if (mfma_points) *mfma_points += d.N;
        };
        auto select_and_flag = [&](const uint32_t* extra, uint32_t* h_extra, int n_extra) {
            // The select writes the flags (and extra words) to pinned memory
            // ahead of the hook's work, so the host reads them while that
            // work runs.
            uint32_t* hb = sc.pinned((size_t)Q + n_extra);
            pt.begin(PH_SELECT, s);
            const bool rk = st.view().row_keys != 0;
            const bool fk = rk && st.view().fused_keys != 0;
            const FusedBand* fb = fk ? d.hook->band : nullptr;
            launch_select_merge(st.view(), s, hb, extra, extra ? n_extra : 0, rk ? d.hook->row_ids : nullptr, fb);
            if (d.hook) {
                d.hook->rows = rk;
                d.hook->fused = fk;
                d.hook->banded = fb != nullptr;
            }
            st.selected();
            pt.end(PH_SELECT, s);
            const hipEvent_t ev = sc.event();
            LSH_HIP(hipEventRecord(ev, s));
            if (d.hook && d.hook->fn) {
                d.hook->fn();
                d.hook->ran = true;
            }
            LSH_HIP(hipEventSynchronize(ev));
            std::vector<int32_t> bad;
            for (int i = 0; i < Q; ++i)
                if (hb[i]) bad.push_back(i);
            for (int i = 0; i < n_extra; ++i) h_extra[i] = hb[Q + i];
            return bad;
        };
        // One select follows each reset here, so the best lists need no
        // clearing (lazy reset).  The cell search clears the view in its
        // first launch.
        const bool cells_path = d.cells && !force_full_pass();
        st.reset(Q, K, 2 * K + 512, s, true, true, !cells_path);
        if (cells_path && d.hook && d.hook->row_ids && row_keys_supported(K, 2 * K + 512)) {
            st.rows_in_keys();
            if (d.hook->fuse) st.fused_in_keys();
        }
        const uint32_t* cstat = nullptr;
        if (cells_path) {
            // Threshold from the nearest pivot cells, then the cells the
            // triangle inequality leaves.
            pt.begin(PH_SCAN, s);
            pt.begin(PH_MFMA, s);
            cstat = cell_search(*d.cells, qs, Q, K, st.view(), sc, depth, s, d.qorder,
                                st.view().fused_keys ? d.hook->fuse : nullptr, &pt, PH_PASS2);
            pt.end(PH_MFMA, s);
            pt.end(PH_SCAN, s);
            if (cstat && d.ordered) *d.ordered = d.qorder != nullptr;
        }
        if (!cstat) {
            if (cells_path) st.clear(s);  // the cell search ran nothing
            sampled_pass();
        }
        uint32_t hstat[3] = {0, 0, 0};
        std::vector<int32_t> bad = select_and_flag(cstat, hstat, cstat ? 3 : 0);
        if (cstat) {
            if (hstat[0]) {
                // Too dense for the cells: nothing was appended.
                st.reset(Q, K, 2 * K + 512, s, true, true);
                sampled_pass();
                bad = select_and_flag(nullptr, nullptr, 0);  // runs the hook again over the new candidates
            } else {
                ++launches;
                if (mfma_launches) ++*mfma_launches;
                // Pairs multiplied per query: the cells' tile products cover 32 x 32 pairs.
                if (mfma_points) *mfma_points += ((int64_t)hstat[1] + hstat[2]) * 1024 / std::max(Q, 1);
                if (d.hook) d.hook->pass2_tiles += hstat[2];
            }
        }
        if (std::getenv("LSHGPU_DEBUG_CELLS"))
            std::fprintf(stderr, "threshold pass (%s): %zu of %d queries redone\n",
                         cstat && !hstat[0] ? "cells" : "full", bad.size(), Q);
        if (!bad.empty()) {
            redo(bad, kCautious);
            if (d.hook) d.hook->redone = bad;
        }
        return;
    }
    auto scan = [&](int64_t p0, int64_t p1, const TopKView& v) {
        // The matrix-core scan queues candidates in LDS, which suits ranges
        // where few pairs pass: once 16*K points are behind every threshold,
        // a 256-query x 32-point tile yields about 256*32*K/p0 <= 512 of them.
        if (mfma && d.fp4 && mode != kFixed && p0 >= 16 * (int64_t)K) {
            pt.begin(PH_MFMA, s);
            launch_hamming_mfma(d.fp4, qs, wpad, Q, p0, p1, v, s);
            pt.end(PH_MFMA, s);
            if (mfma_launches) ++*mfma_launches;
            if (mfma_points) *mfma_points += p1 - p0;
        } else {
            SketchScanArgs a{d.ds, qs, wpad, Q, p0, p1, kind};
            launch_sketch_scan(a, v, s);
        }
    };
    run_rounds(st, Q, K, d.N, mode, scan, s, pt, launches, int_keys);
    if (mode == kFixed) return;
    const std::vector<int32_t> ovf = flagged(st.view().ovf, Q, s);
    if (!ovf.empty()) redo(ovf, mode == kNormal ? kCautious : kFixed);
}

}  // namespace

// Sample ladder of a packed sketch array: level 0 is every stride-th row of
// the data, level i+1 every stride-th row of level i, while a level keeps at
// least 32*8 rows per candidate (stride >= 8).  One level is kept: at
// 1.2M points a second one (4.7k rows, 28 candidates) cost more in launches
// and re-runs than it saved.
void lshgpu::build_sample_levels(const uint32_t* ds, int64_t N, int wpad, int C, std::vector<SampleLevel>& out,
                                 hipStream_t s) {
    out.clear();
    const uint32_t* cur = ds;
    int64_t n = N;
    int K = C;
    constexpr int kMaxLevels = 1;
    while ((int)out.size() < kMaxLevels) {
        const int64_t st = n / (32 * (int64_t)std::max(K, 1));
        if (st < 8) break;
        const int stride = (int)std::min<int64_t>(16, st);
        SampleLevel L;
        L.N = (n + stride - 1) / stride;
        std::vector<int32_t> rows((size_t)L.N);
        for (int64_t i = 0; i < L.N; ++i) rows[i] = (int32_t)(i * stride);
        DevBuf<int32_t> idx((size_t)L.N);
        LSH_HIP(hipMemcpyAsync(idx.get(), rows.data(), sizeof(int32_t) * rows.size(), hipMemcpyHostToDevice, s));
        L.ds.alloc((size_t)L.N * wpad);
        launch_gather_rows_u32(cur, wpad, idx.get(), (int)L.N, L.ds.get(), s);
        L.fp4.alloc(hamming_fp4_words(L.N, wpad));
        launch_expand_fp4(L.ds.get(), L.N, wpad, L.fp4.get(), s);
        LSH_HIP(hipStreamSynchronize(s));
        K = sample_k(K, sample_step(n, L.N));
        cur = L.ds.get();
        n = L.N;
        out.push_back(std::move(L));
    }
}

namespace {

std::vector<SketchSet> sample_views(const std::vector<SampleLevel>& lv) {
    std::vector<SketchSet> v;
    for (const SampleLevel& L : lv) v.push_back(SketchSet{L.ds.get(), L.fp4.get(), L.N});
    return v;
}

// Hyperplanes of the reference stream, transposed to [words][d][hashBits].
std::vector<float> transpose_planes(const std::vector<float>& planes, int words, int hb, int d) {
    std::vector<float> t((size_t)words * d * hb);
    for (int w = 0; w < words; ++w)
        for (int b = 0; b < hb; ++b)
            for (int j = 0; j < d; ++j)
                t[((size_t)w * d + j) * hb + b] = planes[((size_t)w * hb + b) * d + j];
    return t;
}

// JL planes [sdim][d] to [groups][d][32] (zero padded).
std::vector<float> transpose_jl(const std::vector<float>& planes, int sdim, int d) {
    const int groups = (sdim + 31) / 32;
    std::vector<float> t((size_t)groups * d * 32, 0.0f);
    for (int s = 0; s < sdim; ++s)
        for (int j = 0; j < d; ++j) t[((size_t)(s / 32) * d + j) * 32 + (s % 32)] = planes[(size_t)s * d + j];
    return t;
}

}  // namespace

// ---------------------------------------------------------------- index

namespace lshgpu {

void WmhHasher::build(const float* x, int64_t n, const float* extra, int64_t n_extra, int d, int mode_,
                      int nhash_, int hbits_, hipStream_t s) {
    mode = mode_;
    nhash = nhash_;
    hbits = hbits_;
    DevBuf<int> b((size_t)d);
    launch_fill_i32(b.get(), (size_t)d, 0, s);
    launch_wmh_bounds(x, n, d, b.get(), s);
    if (extra && n_extra > 0) launch_wmh_bounds(extra, n_extra, d, b.get(), s);
    std::vector<int> hb((size_t)d);
    LSH_HIP(hipMemcpyAsync(hb.data(), b.get(), sizeof(int) * (size_t)d, hipMemcpyDeviceToHost, s));
    LSH_HIP(hipStreamSynchronize(s));
    // weightedMinHash.cuh:62-90: M = sum of bounds, start = exclusive prefix
    // sum, owner[r] = the dimension whose interval holds r.
    std::vector<int> st((size_t)d);
    int64_t m = 0;
    for (int j = 0; j < d; ++j) {
        st[j] = (int)m;
        m += hb[j];
    }
    LSH_REQUIRE(m < (int64_t)1 << 30, LSH_ERR_UNSUPPORTED, "weighted MinHash bounds sum too large");
    M = (int)m;
    std::vector<int> own((size_t)std::max<int64_t>(m, 1), 0);
    for (int j = 0; j < d; ++j)
        for (int r = st[j]; r < st[j] + hb[j]; ++r) own[r] = j;
    owner.alloc(own.size());
    start.alloc((size_t)d);
    LSH_HIP(hipMemcpyAsync(owner.get(), own.data(), sizeof(int) * own.size(), hipMemcpyHostToDevice, s));
    LSH_HIP(hipMemcpyAsync(start.get(), st.data(), sizeof(int) * (size_t)d, hipMemcpyHostToDevice, s));
    // weightedMinHash.cuh:231-250 generates 255 bits of the default-seeded
    // stream; the table continues that stream so counts past 254 stay defined.
    std::vector<uint8_t> bm = bit_stream(4096);
    bitmap_len = (int)bm.size();
    bitmap.alloc(bm.size());
    LSH_HIP(hipMemcpyAsync(bitmap.get(), bm.data(), bm.size(), hipMemcpyHostToDevice, s));
    R = 256;
    draws.alloc((size_t)nhash * R);
    draw_owner.alloc((size_t)nhash * R);
    launch_wmh_draws(nhash, R, M, owner.get(), draws.get(), draw_owner.get(), s);
    LSH_HIP(hipStreamSynchronize(s));
}

void WmhHasher::sketch(const float* x, int64_t n, int d, uint32_t* out, int out_stride, hipStream_t s) const {
    launch_wmh_sketch(x, n, d, owner.get(), start.get(), M, nhash, mode, hbits, bitmap.get(), bitmap_len,
                      draws.get(), draw_owner.get(), R, out, out_stride, s);
}

}  // namespace lshgpu

namespace {

// Applies `steps` to x (and extra) at build time, freezing the shifts.
void apply_steps_build(std::vector<Transform>& steps, float* x, int64_t n, float* extra, int64_t n_extra, int d,
                       hipStream_t s) {
    for (Transform& t : steps) {
        if (t.kind == kGjStep) {
            t.mins.alloc((size_t)d);
            launch_fill_i32(t.mins.get(), (size_t)d, 0, s);
            launch_gj_mins(x, n, d, t.mins.get(), s);
            if (extra && n_extra > 0) launch_gj_mins(extra, n_extra, d, t.mins.get(), s);
            launch_gj_shift(x, n, d, t.mins.get(), s);
            if (extra && n_extra > 0) launch_gj_shift(extra, n_extra, d, t.mins.get(), s);
        }
        launch_normalize_rows(x, n, d, s);
        if (extra && n_extra > 0) launch_normalize_rows(extra, n_extra, d, s);
    }
}

void apply_steps_query(const std::vector<Transform>& steps, float* q, int64_t n, int d, hipStream_t s) {
    for (const Transform& t : steps) {
        if (t.kind == kGjStep) launch_gj_shift(q, n, d, t.mins.get(), s);
        launch_normalize_rows(q, n, d, s);
    }
}

std::vector<Transform> one_step(int kind) {
    std::vector<Transform> v(1);
    v[0].kind = kind;
    return v;
}

// Top-K over the unique candidate sets of the rows of `a` (rows = a.Q).
// hmap: host copy of a.qmap (empty when a.qmap is null).  Queries whose
// appends overflowed a round are re-run on their own with ranges no larger
// than the append capacity.
void lsh_topk(const LshScanArgs& base, const std::vector<int32_t>& hmap, int64_t maxS, int K, int mode,
              bool int_keys, TopKState& st, Scratch& sc, int depth, hipStream_t s, PhaseTimer& pt,
              int64_t& launches) {
    auto scan = [&](int64_t p0, int64_t p1, const TopKView& v) {
        LshScanArgs a = base;
        a.p0 = p0;
        a.p1 = p1;
        launch_lsh_scan(a, v, s);
    };
    run_rounds(st, base.Q, K, std::max<int64_t>(maxS, 1), mode, scan, s, pt, launches, int_keys);
    if (mode == kFixed) return;
    const std::vector<int32_t> ovf = flagged(st.view().ovf, base.Q, s);
    if (ovf.empty()) return;
    const int m = (int)ovf.size();
    std::vector<int32_t> sub_map((size_t)m);
    for (int i = 0; i < m; ++i) sub_map[i] = hmap.empty() ? ovf[i] : hmap[ovf[i]];
    int32_t* dmap = sc.get<int32_t>(depth, 0, (size_t)m);
    int32_t* drows = sc.get<int32_t>(depth, 1, (size_t)m);
    LSH_HIP(hipMemcpyAsync(dmap, sub_map.data(), sizeof(int32_t) * m, hipMemcpyHostToDevice, s));
    LSH_HIP(hipMemcpyAsync(drows, ovf.data(), sizeof(int32_t) * m, hipMemcpyHostToDevice, s));
    LshScanArgs a = base;
    a.qmap = dmap;
    a.Q = m;
    TopKState& sst = sc.tk[depth];
    lsh_topk(a, sub_map, maxS, K, mode == kNormal ? kCautious : kFixed, int_keys, sst, sc, depth + 1, s, pt,
             launches);
    const TopKView sv = sst.view(), v = st.view();
    launch_scatter_rows_u64(sv.keys, sv.stride, K, drows, m, v.keys, v.stride, s);
}

}  // namespace

size_t lsh_index::device_bytes() const {
    size_t b = data.bytes() + data_inv.bytes() + dtile.bytes() + dsk.bytes() + dfp4.bytes() + planesT.bytes() + smf.bytes() +
               djl.bytes() + wmh.bytes() +
               kwmh.bytes() + scratch.bytes() + kplanesT.bytes() + ksigns.bytes() + dkeys.bytes() + boff.bytes() + bperm.bytes() +
               qkeys.bytes() + qlo.bytes() + qpre.bytes() + qmax.bytes() + qlen.bytes() + bkeys.bytes() +
               wl_ibase.bytes() + wl_flags.bytes() + wl_part.bytes() +
               dense.bytes() + q.bytes() + qT.bytes() +
               q_inv.bytes() + qjl.bytes() + qjlT.bytes() + qsk.bytes() + out.bytes() + flags.bytes() + st.bytes();
    for (const SampleLevel& L : samples) b += L.ds.bytes() + L.fp4.bytes();
    b += cells.bytes() + qorder.bytes() + cdata.bytes() + cpos.bytes() + chalf.bytes() + q16.bytes() + rband.band.bytes() +
         rband.cnt.bytes();
    for (const Transform& t : key_steps) b += t.mins.bytes();
    for (const Transform& t : sketch_steps) b += t.mins.bytes();
    for (const auto& sl : slots) b += sl->bytes();
    return b;
}

void lsh_index::build(const float* src, bool on_device, const float* extra_host, int64_t n_extra) {
    DeviceGuard g(device);
    hipStream_t s = stream;
    data.alloc((size_t)N * d);
    if (N > 0)
        LSH_HIP(hipMemcpyAsync(data.get(), src, sizeof(float) * (size_t)N * d,
                               on_device ? hipMemcpyDeviceToDevice : hipMemcpyHostToDevice, s));
    DevBuf<float> extra;
    if (extra_host && n_extra > 0) {
        extra.alloc((size_t)n_extra * d);
        LSH_HIP(hipMemcpyAsync(extra.get(), extra_host, sizeof(float) * (size_t)n_extra * d,
                               hipMemcpyHostToDevice, s));
    }
    if (framework == 1) {
        build_lsh(extra.get(), n_extra, s);
        LSH_HIP(hipStreamSynchronize(s));
        return;
    }

    // processingUtils.cuh:12-67 as memOptimizedLinearScan.cu:149-163 and
    // weightedMinHash.cuh:201-213 apply it; simHashJL.cuh:28-32 normalises.
    // The shifts are frozen at build time; the one-shot entry points include
    // the queries, as the reference does.
    if ((impl == 2 && dfunc == 2) || impl == 4 || impl == 5)
        sketch_steps = one_step(kGjStep);
    else if (impl == 6)
        sketch_steps = one_step(kNormStep);
    apply_steps_build(sketch_steps, data.get(), N, extra.get(), n_extra, d, s);

    const int kind = impl == 1 ? kNegCos : (impl == 2 ? dfunc : (impl == 4 || impl == 5 ? kGenJaccard : kAngular));
    if (impl <= 2 && angular_kind(kind)) {
        // Zero past N to a whole tile of 32 rows (the matrix-core scan reads
        // the inverse norms of whole tiles).
        data_inv.alloc((size_t)((N + 31) / 32 * 32));
        if (N > 0) LSH_HIP(hipMemsetAsync(data_inv.get(), 0, data_inv.bytes(), s));
        launch_row_inv_norms(data.get(), N, d, data_inv.get(), s);
        if (exact_mfma_supported(kind, d) && !exact_valu_forced()) {
            dtile.alloc((size_t)((N + 31) / 32) * exact_mfma_steps(d) * 64);
            launch_tile_rows(data.get(), N, d, dtile.get(), s);
        }
    }

    if (impl == 3) {
        // simHash.cu:61-99: 32*sketchDim hyperplanes of the reference stream.
        words = sketchDim;
        wpad = round_up(std::max(words, 1), 8);
        build_simhash_planes(s);
        dsk.alloc((size_t)N * wpad);
        if (N > 0) LSH_HIP(hipMemsetAsync(dsk.get(), 0, dsk.bytes(), s));
        sketch_simhash(data.get(), N, dsk.get(), wpad, s);
        build_fp4(s);
    } else if (impl == 6) {
        // simHashJL.cuh:238-262: planes from N(0, 1/sketchDim).
        sdim = sketchDim;
        std::vector<float> planes = normal_stream((int64_t)sdim * d, 1.0 / sdim);
        std::vector<float> pt = transpose_jl(planes, sdim, d);
        planesT.alloc(pt.size());
        LSH_HIP(hipMemcpyAsync(planesT.get(), pt.data(), sizeof(float) * pt.size(), hipMemcpyHostToDevice, s));
        djl.alloc((size_t)N * sdim);
        launch_jl(data.get(), N, d, planesT.get(), sdim, djl.get(), s);
    } else if (impl == 4 || impl == 5) {
        // weightedMinHash.cuh:344-416: bounds, index map, sketches (the bounds
        // include the one-shot queries, as the reference computes them).
        words = sketchDim;
        wpad = round_up(std::max(words, 1), 8);
        wmh.build(data.get(), N, extra.get(), n_extra, d, impl == 5 ? 1 : 0, impl == 5 ? 32 * words : words, 32, s);
        dsk.alloc((size_t)N * wpad);
        if (N > 0) LSH_HIP(hipMemsetAsync(dsk.get(), 0, dsk.bytes(), s));
        wmh.sketch(data.get(), N, d, dsk.get(), wpad, s);
        if (impl == 5) build_fp4(s);
    }
    LSH_HIP(hipStreamSynchronize(s));
}

constexpr int64_t kSimhashMfmaMinRows = 65536;

// simHash.cu:61-99: 32*words hyperplanes of the reference stream, FP32
// transposed for the vector kernel (and the matrix-core kernel's recheck),
// and the matrix-core operands unless projection 1 asks for FP32 only.
void lsh_index::build_simhash_planes(hipStream_t s) {
    std::vector<float> planes = normal_stream((int64_t)words * 32 * d, 1.0);
    std::vector<float> pt = transpose_planes(planes, words, 32, d);
    planesT.alloc(pt.size());
    LSH_HIP(hipMemcpyAsync(planesT.get(), pt.data(), sizeof(float) * pt.size(), hipMemcpyHostToDevice, s));
    if (projection == 0 && simhash_mfma_supported(d, 32)) smf.build(planes.data(), d, words, planesT.get(), s);
}

void lsh_index::sketch_simhash(const float* x, int64_t n, uint32_t* out, int stride, hipStream_t s) const {
    // The matrix cores pay for large row counts (the data: 1.2M rows of 1024
    // bits in 1.9 ms against 4.3 ms on the vector units); a query batch of
    // 10k rows is faster on the vector units (50 us against 70 us).
    if (smf.built() && n >= kSimhashMfmaMinRows)
        smf.sketch(x, n, out, stride, s);
    else
        launch_simhash(x, n, d, planesT.get(), words, 32, out, stride, s);
}

// FP4 operand image of the packed sketches for the matrix-core Hamming scan.
void lsh_index::build_fp4(hipStream_t s) {
    if (!hamming_mfma_supported(wpad) || N < 16 * 1024) return;
    dfp4.alloc(hamming_fp4_words(N, wpad));
    launch_expand_fp4(dsk.get(), N, wpad, dfp4.get(), s);
    build_sample_levels(dsk.get(), N, wpad, candidates, samples, s);
    if (!samples.empty() && framework == 0) {
        cells.build(dsk.get(), N, wpad, s);
        if (cells.built()) {
            cell_rows(cells, data.get(), N, d, cdata, cpos, s);
            // The angular rerank (SimHash) filters on an FP16 image of the
            // unit rows; data with a zero or non-finite row, or more than
            // 2^32 sixteen-byte pieces of image, keeps FP32.
            // LSHGPU_RERANK_F32=1 keeps the FP32 filter (for comparisons).
            const char* f32 = std::getenv("LSHGPU_RERANK_F32");
            if (impl == 3 && !(f32 && f32[0] == '1'))
                build_unit_rows_half(cdata.get(), cells.ids.get(), cells.ntiles * 32, d, chalf, s);
        }
    }
}

// lshFramework.cuh:312-409 runLsh up to the scan: keys (generateHashes with
// isHashKeys), bucket tables, then the sketches of the in-bucket distance.
// Each hash family preprocesses the rows in place before hashing, exactly as
// the reference's does (weightedMinHash.cuh:451, crossPolyTope.cuh:109,
// simHashJL.cuh:59), so a second family sees the first one's output.
void lsh_index::build_lsh(const float* extra_in, int64_t n_extra, hipStream_t s) {
    float* extra = const_cast<float*>(extra_in);
    const int L = tables;
    LSH_REQUIRE(bucketKeyBits >= 1 && bucketKeyBits <= 32, LSH_ERR_UNSUPPORTED, "bucketKeyBits must be in 1..32");
    LSH_REQUIRE((int64_t)L * std::max<int64_t>(N, 1) < 0xFFFFFFFFll, LSH_ERR_UNSUPPORTED,
                "tables * N_data must be below 2^32");
    if (keysImpl == 4 || keysImpl == 5) key_steps = one_step(kGjStep);
    else if (keysImpl == 7 || keysImpl == 8) key_steps = one_step(kNormStep);
    apply_steps_build(key_steps, data.get(), N, extra, n_extra, d, s);

    // The reference's key types (kernel.cu:67-80): unsigned short for 3, 5
    // and 7, unsigned char for 4.  The dense-rotation keys (8) are 32-bit.
    keyBitsMask = keysImpl == 4 ? 0xFF : (keysImpl == 8 ? -1 : 0xFFFF);
    // Tables wider than 16 bits (and the dense-rotation keys) are sorted
    // (key, ID) arrays searched by binary search; narrower ones keep the
    // reference's 2^bits bucket offsets (lshFramework.cuh:102-140).
    sorted = keysImpl == 8 || bucketKeyBits > 16;
    sortBits = keysImpl == 8 ? bucketKeyBits : std::min(bucketKeyBits, keysImpl == 4 ? 8 : 16);
    tableSize = sorted ? 0u : 1u << bucketKeyBits;
    if (keysImpl == 8) {
        // K = bucketKeyBits / log2(2P) hashes of P outputs per table.
        const int b = dense_key_field_bits(d);
        LSH_REQUIRE(d <= 128, LSH_ERR_UNSUPPORTED, "dense cross-polytope keys support d <= 128");
        LSH_REQUIRE(bucketKeyBits % b == 0, LSH_ERR_UNSUPPORTED,
                    "dense cross-polytope keys: bucketKeyBits must be a multiple of log2(2P) = " + std::to_string(b) +
                        " (P = d rounded up to a power of two)");
        dense.build(d, L, bucketKeyBits / b, s);
    } else if (keysImpl == 3) {
        // lshFramework.cuh:22-45: tables*d*bucketKeyBits hyperplanes; table t
        // takes bits b < bucketKeyBits from hyperplane t*bucketKeyBits + b.
        LSH_REQUIRE(bucketKeyBits <= 32, LSH_ERR_UNSUPPORTED, "SimHash keys take at most 32 bits");
        std::vector<float> planes = normal_stream((int64_t)L * bucketKeyBits * d, 1.0);
        std::vector<float> pt = transpose_planes(planes, L, bucketKeyBits, d);
        kplanesT.alloc(pt.size());
        LSH_HIP(hipMemcpyAsync(kplanesT.get(), pt.data(), sizeof(float) * pt.size(), hipMemcpyHostToDevice, s));
    } else if (keysImpl == 4) {
        // weightedMinHashGeneric with sketchDim = tables, original counts.
        LSH_REQUIRE(bucketKeyBits >= 8, LSH_ERR_UNSUPPORTED,
                    "weighted MinHash count keys are 8-bit: bucketKeyBits must be >= 8");
        kwmh.build(data.get(), N, extra, n_extra, d, 0, L, 32, s);
    } else if (keysImpl == 5) {
        // weightedMinHashGeneric with sketchDim = tables, hashBits = bucketKeyBits.
        kwmh.build(data.get(), N, extra, n_extra, d, 1, L * bucketKeyBits, bucketKeyBits, s);
    } else {
        // crossPolyTope.cuh:106-130: one cross-polytope hash per table.
        int P = 1;
        while (P < d) P *= 2;
        LSH_REQUIRE(P <= 256, LSH_ERR_UNSUPPORTED, "cross-polytope keys support d <= 256");
        LSH_REQUIRE(bucketKeyBits >= 31 || (int64_t)(2 * P) <= ((int64_t)1 << bucketKeyBits), LSH_ERR_UNSUPPORTED,
                    "cross-polytope keys need 2^bucketKeyBits >= 2 * (d rounded up to a power of two)");
        std::vector<int32_t> sg = sign_stream((int64_t)L * 3 * P);
        ksigns.alloc(sg.size());
        LSH_HIP(hipMemcpyAsync(ksigns.get(), sg.data(), sizeof(int32_t) * sg.size(), hipMemcpyHostToDevice, s));
    }
    dkeys.alloc((size_t)N * L);
    bperm.alloc((size_t)L * N);
    if (sorted) {
        // Keys per table ([L][N]), then a stable radix sort of (key, ID).
        DevBuf<uint32_t> kln((size_t)L * N), tk((size_t)L * N), ti((size_t)L * N),
            counts(sort_table_counts(N, L));
        if (keysImpl == 8) {
            dense.keys(data.get(), N, dkeys.get(), kln.get(), s);
        } else {
            lsh_keys(data.get(), N, dkeys.get(), s);
            launch_transpose_keys(dkeys.get(), N, L, kln.get(), s);
        }
        bkeys.alloc((size_t)L * N);
        launch_sort_tables(kln.get(), N, L, sortBits, bkeys.get(), bperm.get(), tk.get(), ti.get(), counts.get(), s);
        LSH_HIP(hipStreamSynchronize(s));
    } else {
        lsh_keys(data.get(), N, dkeys.get(), s);
        boff.alloc((size_t)L * (tableSize + 1));
        DevBuf<uint32_t> cursor((size_t)L * (tableSize + 1));
        launch_bucket_build(dkeys.get(), N, L, tableSize, boff.get(), cursor.get(), bperm.get(), s);
        LSH_HIP(hipStreamSynchronize(s));
    }

    // generateHashes(isHashKeys = false), lshFramework.cuh:391.
    if (impl == 4 || impl == 5) sketch_steps = one_step(kGjStep);
    else if (impl == 6) sketch_steps = one_step(kNormStep);
    apply_steps_build(sketch_steps, data.get(), N, extra, n_extra, d, s);
    if (!withSketch || impl == 2) return;
    if (impl == 3) {
        words = sketchDim;
        wpad = words;
        build_simhash_planes(s);
        dsk.alloc((size_t)N * wpad);
        sketch_simhash(data.get(), N, dsk.get(), wpad, s);
    } else if (impl == 4 || impl == 5) {
        words = sketchDim;
        wpad = words;
        wmh.build(data.get(), N, extra, n_extra, d, impl == 5 ? 1 : 0, impl == 5 ? 32 * words : words, 32, s);
        dsk.alloc((size_t)N * wpad);
        wmh.sketch(data.get(), N, d, dsk.get(), wpad, s);
    } else {
        sdim = sketchDim;
        std::vector<float> planes = normal_stream((int64_t)sdim * d, 1.0 / sdim);
        std::vector<float> pt = transpose_jl(planes, sdim, d);
        planesT.alloc(pt.size());
        LSH_HIP(hipMemcpyAsync(planesT.get(), pt.data(), sizeof(float) * pt.size(), hipMemcpyHostToDevice, s));
        djl.alloc((size_t)N * sdim);
        launch_jl(data.get(), N, d, planesT.get(), sdim, djl.get(), s);
    }
}

void lsh_index::lsh_keys(const float* x, int64_t n, uint32_t* keys, hipStream_t s) {
    const int L = tables;
    if (n <= 0) return;
    if (keysImpl == 8) {
        dense.keys(x, n, keys, nullptr, s);
        return;
    }
    if (keysImpl == 3) {
        launch_simhash(x, n, d, kplanesT.get(), L, bucketKeyBits, keys, L, s);
    } else if (keysImpl == 4 || keysImpl == 5) {
        kwmh.sketch(x, n, d, keys, L, s);
    } else {
        DevBuf<uint16_t> k16((size_t)n * L);
        launch_crosspoly(x, n, d, L, ksigns.get(), k16.get(), s);
        launch_widen_keys(k16.get(), keys, (size_t)n * L, s);
        LSH_HIP(hipStreamSynchronize(s));
    }
    launch_mask_keys(keys, (size_t)n * L, (uint32_t)keyBitsMask, s);
}

void lsh_index::search(const float* queries, int Q, int k, bool q_on_device, lsh_point_t* results,
                       bool r_on_device, hipStream_t user_stream) {
    DeviceGuard g(device);
    LSH_REQUIRE(Q >= 0, LSH_ERR_INVALID, "N_query must be >= 0");
    LSH_REQUIRE(k >= 1, LSH_ERR_INVALID, "k must be >= 1");
    const int kq = calculate_k(k);
    hipStream_t s = user_stream ? user_stream : stream;
    if (done_rec) LSH_HIP(hipStreamWaitEvent(s, done_ev, 0));
    PhaseTimer& pt = timer;
    scan_launches = 0;
    mfma_launches = 0;
    mfma_points = 0;
    pass2_tiles = 0;
    last_candidates = 0;
    if (Q == 0) return;
    pt.begin(PH_TOTAL, s);
    pt.begin(PH_PREP, s);
    // Device queries that need no transform are read where they are; device
    // results are written in place.
    if (q_on_device && framework == 0 && sketch_steps.empty()) {
        qp = queries;
    } else {
        q.ensure((size_t)Q * d);
        LSH_HIP(hipMemcpyAsync(q.get(), queries, sizeof(float) * (size_t)Q * d,
                               q_on_device ? hipMemcpyDeviceToDevice : hipMemcpyHostToDevice, s));
        if (framework == 0) apply_steps_query(sketch_steps, q.get(), Q, d, s);
        qp = q.get();
    }
    pt.end(PH_PREP, s);
    if (r_on_device) {
        outp = results;
    } else {
        out.ensure((size_t)Q * kq);
        outp = out.get();
    }
    flags.ensure((size_t)Q);

    if (framework == 1) {
        search_lsh(own_work(), Q, kq, s, pt);
    } else if (impl == 1 || impl == 2) {
        const int kind = impl == 1 ? kNegCos : dfunc;
        LSH_REQUIRE(kq <= kMaxTopK, LSH_ERR_UNSUPPORTED, "exact scan supports k <= 4096");
        int K = std::min(kMaxTopK, pow2_ceil(kq + margin));
        const float* dt = dtile.get();
        const double eps = dt ? mfma_error_bound(d) : fp32_error_bound(kind, d);
        exact_topk(data.get(), data_inv.get(), dt, N, d, kind, ExactQueries{qp, Q}, K, kNormal, st, scratch, 0, s,
                   pt, scan_launches);
        pt.begin(PH_FINAL, s);
        launch_exact_finalize(data.get(), qp, d, kind, st.view(), kq, eps, outp, flags.get(), s);
        pt.end(PH_FINAL, s);
        std::vector<int32_t> bad = flagged(flags.get(), Q, s);
        if (!bad.empty()) {
            // More than K-k points lie within the FP32 error band of the k-th
            // distance: redo those queries with the widest list.
            LSH_REQUIRE(K < kMaxTopK, LSH_ERR_UNSUPPORTED,
                        "more than 4096 points within the FP32 error band of the k-th distance");
            const int m = (int)bad.size();
            DevBuf<int32_t> idx((size_t)m);
            LSH_HIP(hipMemcpyAsync(idx.get(), bad.data(), sizeof(int32_t) * m, hipMemcpyHostToDevice, s));
            DevBuf<float> sub((size_t)m * d);
            launch_gather_rows_f32(qp, d, idx.get(), m, sub.get(), s);
            TopKState sst;
            exact_topk(data.get(), data_inv.get(), dt, N, d, kind, ExactQueries{sub.get(), m}, kMaxTopK, kNormal,
                       sst, scratch, 0, s, pt, scan_launches);
            DevBuf<lsh_point_t> sout((size_t)m * kq);
            DevBuf<uint32_t> sflags((size_t)m);
            launch_exact_finalize(data.get(), sub.get(), d, kind, sst.view(), kq, eps, sout.get(), sflags.get(), s);
            LSH_REQUIRE(flagged(sflags.get(), m, s).empty(), LSH_ERR_UNSUPPORTED,
                        "more than 4096 points within the FP32 error band of the k-th distance");
            launch_scatter_rows_pts(sout.get(), kq, idx.get(), m, outp, s);
            LSH_HIP(hipStreamSynchronize(s));
        }
    } else if (impl >= 3 && impl <= 5) {
        const int Qpad = sketch_query_rows(Q);
        qsk.ensure((size_t)Qpad * wpad);
        pt.begin(PH_PREP, s);
        // Rows past Q and words past the sketch stay zero: the sketch writes
        // neither, so the buffer is cleared only when it or Q changes.
        if (qsk.get() != qsk_zeroed || Q != qsk_zeroed_q) {
            LSH_HIP(hipMemsetAsync(qsk.get(), 0, sizeof(uint32_t) * (size_t)Qpad * wpad, s));
            qsk_zeroed = qsk.get();
            qsk_zeroed_q = Q;
        }
        if (impl == 3) sketch_simhash(qp, Q, qsk.get(), wpad, s);
        else wmh.sketch(qp, Q, d, qsk.get(), wpad, s);
        // The queries' FP16 parts for the fused cell scan.
        const bool fuse = fuse_ok(kq);
        if (fuse) {
            q16.ensure(query_half_elems(Q, fused_k_steps(d)));
            launch_query_half(qp, Q, d, fused_k_steps(d), q16.get(), s);
        }
        pt.end(PH_PREP, s);
        const int Ceff = (int)std::max<int64_t>(1, std::min<int64_t>(candidates, N));
        const int K = Ceff <= kMaxIntTopK ? Ceff : pow2_ceil(Ceff);
        LSH_REQUIRE(K <= kMaxTopK, LSH_ERR_UNSUPPORTED, "candidate sets larger than 4096 are not supported");
        bool ordered = false;
        if (cells.built()) qorder.ensure((size_t)Q);
        const int kind = impl == 3 ? kAngular : kGenJaccard;
        // The rerank of queries `order` (all queries when null; m of them);
        // rows: the select left each best-list key's row beside the list;
        // fused: the keys are fused keys (approximate distance and row);
        // banded: the select wrote the band too (indexed by query), and only
        // the FP64 recheck is left.
        auto rerank = [&](const int32_t* order, int m, bool rows, bool fused, bool banded) {
            if (banded) {
                pt.begin(PH_RECHECK, s);
                launch_rerank_recheck(cdata.get(), qp, d, kind, Ceff, m, kq, outp, s, nullptr, rband, cells.ids.get());
                pt.end(PH_RECHECK, s);
                return;
            }
            pt.begin(PH_FINAL, s);
            // The rerank's filter (timed as the rerank phase), then its FP64
            // recheck when it splits.
            const TopKView v = st.view();
            const int32_t* rows_in =
                rows && !fused ? reinterpret_cast<const int32_t*>(v.keys) + 2 * (size_t)v.K : nullptr;
            const double eps = fused ? fused_error_bound(d) : chalf.get() ? half_error_bound(d) : fp32_error_bound(kind, d);
            const bool split =
                cells.built()
                    ? launch_rerank(cdata.get(), qp, d, kind, v.keys, v.stride, nullptr, Ceff, m, kq, eps, outp, s,
                                    order, cpos.get(), chalf.get(), &rband, false, rows_in, 2 * v.stride,
                                    chalf.size(), fused ? cells.ids.get() : nullptr)
                    : launch_rerank(data.get(), qp, d, kind, st.view().keys, st.view().stride, nullptr, Ceff, m,
                                    kq, fp32_error_bound(kind, d), outp, s, order, nullptr, nullptr, &rband,
                                    false);
            pt.end(PH_FINAL, s);
            if (split) {
                pt.begin(PH_RECHECK, s);
                launch_rerank_recheck(cells.built() ? cdata.get() : data.get(), qp, d, kind, Ceff, m, kq,
                                      outp, s, order, rband);
                pt.end(PH_RECHECK, s);
            }
        };
        SketchSet::Hook hook;
        hook.row_ids = cells.built() ? cells.ids.get() : nullptr;
        // Approximate distances fused into the cell scan's keys: the angular
        // index with its FP16 image, k within the filter's split form.
        CellFuse cfuse{};
        FusedBand fband{};
        if (fuse) {
            cfuse = CellFuse{chalf.get(), half_row_stride(d), q16.get(), fused_k_steps(d)};
            hook.fuse = &cfuse;
            const int Cp = pow2_ceil(std::max(Ceff, 2));
            rband.band.ensure((size_t)Q * Cp);
            rband.cnt.ensure((size_t)Q);
            fband = FusedBand{rband.band.get(), rband.cnt.get(), query_half_flags(q16.get(), Q, fused_k_steps(d)),
                              Cp, kq, fused_band_delta(d)};
            hook.band = &fband;
        }
        hook.fn = [&] { rerank(ordered ? qorder.get() : nullptr, Q, hook.rows, hook.fused, hook.banded); };
        const SketchSet full{dsk.get(), dfp4.get(), N, &cells, cells.built() ? qorder.get() : nullptr, &ordered,
                             &hook};
        const std::vector<SketchSet> sv = sample_views(samples);
        sketch_topk(full, sv.data(), (int)sv.size(), wpad, impl == 4 ? 1 : 0, qsk.get(), Q, K, kNormal, st,
                    scratch, 0, s, pt, scan_launches,
                    &mfma_launches, &mfma_points);
        pass2_tiles = hook.pass2_tiles;
        if (!hook.ran) {
            rerank(ordered ? qorder.get() : nullptr, Q, false, false, false);
        } else if (!hook.redone.empty()) {
            // Queries whose candidates were redone after the rerank was
            // enqueued: rerank them again.
            const int m = (int)hook.redone.size();
            flags.ensure((size_t)m);
            int32_t* idx = reinterpret_cast<int32_t*>(flags.get());
            LSH_HIP(hipMemcpyAsync(idx, hook.redone.data(), sizeof(int32_t) * m, hipMemcpyHostToDevice, s));
            rerank(idx, m, false, false, false);
        }
    } else {
        qjl.ensure((size_t)Q * sdim);
        pt.begin(PH_PREP, s);
        launch_jl(qp, Q, d, planesT.get(), sdim, qjl.get(), s);
        pt.end(PH_PREP, s);
        const int Ceff = (int)std::min<int64_t>(candidates, N);
        const int K = pow2_ceil(std::max(Ceff, 1));
        LSH_REQUIRE(K <= kMaxTopK, LSH_ERR_UNSUPPORTED, "candidate sets larger than 4096 are not supported");
        exact_topk(djl.get(), nullptr, nullptr, N, sdim, kSqL2, ExactQueries{qjl.get(), Q}, K, kNormal, st, scratch, 0,
                   s, pt, scan_launches);
        pt.begin(PH_FINAL, s);
        launch_rerank(data.get(), qp, d, kAngular, st.view().keys, st.view().stride, nullptr, Ceff, Q, kq,
                      fp32_error_bound(kAngular, d), outp, s, nullptr, nullptr, nullptr, &rband);
        pt.end(PH_FINAL, s);
    }

    if (!r_on_device)
        LSH_HIP(hipMemcpyAsync(results, outp, sizeof(lsh_point_t) * (size_t)Q * kq, hipMemcpyDeviceToHost, s));
    pt.end(PH_TOTAL, s);
    // Device results on the caller's stream are ordered on it; otherwise the
    // call returns with the results complete.
    if (!done_ev) LSH_HIP(hipEventCreateWithFlags(&done_ev, hipEventDisableTiming));
    LSH_HIP(hipEventRecord(done_ev, s));
    done_rec = true;
    if (!r_on_device || !user_stream) LSH_HIP(hipStreamSynchronize(s));
    if (pt.on()) {
        prof_mfma_launches += mfma_launches;
        prof_mfma_points += mfma_points;
        prof_pass2_tiles += pass2_tiles;
    }
}

// ------------------------------------------------------- batched search
// Query-count classes of the captured graphs: a batch of b queries runs the
// graph of the smallest class >= b, the rows past b zero (a zero row
// normalises to NaN, whose keys match no real point, so it costs nearly
// nothing).  Classes grow by about 1.25x from 1 to 4096.
namespace {

int graph_class(int b) {
    int c = 1;
    while (c < b) c = std::max(c + 1, (int)std::ceil(c * 1.25));
    return std::min(c, lsh_index::kGraphMaxQ);
}

}  // namespace

lsh_index::~lsh_index() {
    slots.clear();
    if (bflags) (void)hipHostFree(bflags);
}

void lsh_index::drop_graphs() {
    for (auto& sl : slots) sl->drop_graphs();
}

batch_slot::~batch_slot() {
    drop_graphs();
    if (stream) (void)hipStreamDestroy(stream);
}

void batch_slot::drop_graphs() {
    for (auto& g : graphs) (void)hipGraphExecDestroy(g.second);
    graphs.clear();
    sig.clear();
}

size_t batch_slot::bytes() const {
    return q.bytes() + out.bytes() + st.bytes() + flags.bytes() + qsk.bytes() + scratch.bytes() + qorder.bytes() +
           rband.band.bytes() + rband.cnt.bytes() + qkeys.bytes() + qlo.bytes() + qpre.bytes() + qmax.bytes() +
           qlen.bytes() + wl_ibase.bytes() + wl_flags.bytes() + wl_part.bytes() + q16.bytes();
}

void lsh_index::search_batches(const float* queries, const int* sizes, int n_batches, int k, bool q_on_device,
                               lsh_point_t* results, bool r_on_device, hipStream_t user_stream) {
    DeviceGuard g(device);
    LSH_REQUIRE(n_batches >= 0 && (n_batches == 0 || sizes), LSH_ERR_INVALID, "bad batch list");
    LSH_REQUIRE(k >= 1, LSH_ERR_INVALID, "k must be >= 1");
    int64_t total = 0;
    for (int i = 0; i < n_batches; ++i) {
        LSH_REQUIRE(sizes[i] >= 0, LSH_ERR_INVALID, "batch sizes must be >= 0");
        total += sizes[i];
    }
    const int kq = calculate_k(k);
    batch_graphs = 0;
    batch_redone = 0;
    const bool sketch = sketch_graphable(kq);
    if (!sketch && !lsh_graphable(kq)) {
        // No captured form for this path: one search per batch.
        int64_t off = 0;
        for (int i = 0; i < n_batches; ++i) {
            if (sizes[i] > 0)
                search(queries + off * d, sizes[i], k, q_on_device, results + off * kq, r_on_device, user_stream);
            off += sizes[i];
        }
        if (r_on_device && user_stream) LSH_HIP(hipStreamSynchronize(user_stream));
        return;
    }
    search_batches_graphs(sketch, queries, sizes, n_batches, k, total, q_on_device, results, r_on_device,
                          user_stream);
}

// The SimHash index with pivot cells (impl 3) in captured graphs: the cell
// search needs no host decision until its select has run, so sketch, cell
// search, select and rerank are one graph per (query-count class, k).  A
// batch that leaves a query incomplete (flagged by the select) or finds the
// cells' second pass too dense is redone afterwards by ordinary searches,
// whose results replace the graph's.
// Whether a search with k rounded to kq fuses approximate distances into the
// cell scan's keys: the angular index's FP16 image, d within the fused
// products' 128 halves, k within the filter's split form (LSHGPU_FUSED=0
// keeps the filter's row gather).
bool lsh_index::fuse_ok(int kq) const {
    const char* off = std::getenv("LSHGPU_FUSED");
    return framework == 0 && impl == 3 && cells.built() && chalf.get() && kq <= 64 && fused_k_steps(d) > 0 &&
           sketch_steps.empty() && !(off && off[0] == '0');
}

bool lsh_index::sketch_graphable(int kq) const {
    const int64_t ceff = std::max<int64_t>(1, std::min<int64_t>(candidates, N));
    return framework == 0 && impl == 3 && cells.built() && !cells_dense && sketch_steps.empty() && kq <= kGraphMaxK &&
           ceff <= kMaxIntTopK && hamming_mfma_supported(wpad) && !force_valu_scan() && !force_full_pass();
}

void lsh_index::enqueue_sketch_search(batch_slot& sl, int Qc, int kq) {
    const hipStream_t s = sl.stream;
    const int Ceff = (int)std::max<int64_t>(1, std::min<int64_t>(candidates, N));
    const int K = Ceff;
    sketch_simhash(sl.q.get(), Qc, sl.qsk.get(), wpad, s);
    // Lazy reset: the pivot-matrix launch of the cell search clears the view.
    // The cell scan appends rows in place of IDs where the select can turn
    // them back; the select then leaves the rows beside the best lists.
    // With fused keys the scan also writes each pair's approximate distance
    // into its key, and the rerank's filter reads no rows.
    sl.st.reset(Qc, K, 2 * K + 512, s, true, true, false);
    const bool rk = row_keys_supported(K, 2 * K + 512);
    const bool fused = rk && fuse_ok(kq) && sl.q16.get();
    CellFuse cf{};
    if (rk) sl.st.rows_in_keys();
    FusedBand fb{};
    if (fused) {
        launch_query_half(sl.q.get(), Qc, d, fused_k_steps(d), sl.q16.get(), s);
        sl.st.fused_in_keys();
        cf = CellFuse{chalf.get(), half_row_stride(d), sl.q16.get(), fused_k_steps(d)};
        fb = FusedBand{sl.rband.band.get(), sl.rband.cnt.get(), query_half_flags(sl.q16.get(), Qc, fused_k_steps(d)),
                       pow2_ceil(std::max(Ceff, 2)), kq, fused_band_delta(d)};
    }
    const uint32_t* cstat = cell_search(cells, sl.qsk.get(), Qc, K, sl.st.view(), sl.scratch, 0, s, sl.qorder.get(),
                                        fused ? &cf : nullptr);
    LSH_REQUIRE(cstat != nullptr, LSH_ERR_INTERNAL, "batched search: the cell search did not run");
    // Flags of the Qc queries, then the cell search's three status words.
    launch_select_merge(sl.st.view(), s, sl.flags.get(), cstat, 3, rk ? cells.ids.get() : nullptr,
                        fused ? &fb : nullptr);
    sl.st.selected();
    if (fused) {
        // The select wrote the band (indexed by query): the FP64 recheck is left.
        launch_rerank_recheck(cdata.get(), sl.q.get(), d, kAngular, Ceff, Qc, kq, sl.out.get(), s, nullptr,
                              sl.rband, cells.ids.get());
        return;
    }
    const double eps = fused ? fused_error_bound(d) : chalf.get() ? half_error_bound(d) : fp32_error_bound(kAngular, d);
    const TopKView v = sl.st.view();
    const bool split = launch_rerank(cdata.get(), sl.q.get(), d, kAngular, v.keys, v.stride, nullptr, Ceff, Qc, kq,
                                     eps, sl.out.get(), s, sl.qorder.get(), cpos.get(), chalf.get(), &sl.rband, false,
                                     rk && !fused ? reinterpret_cast<const int32_t*>(v.keys) + 2 * (size_t)v.K
                                                  : nullptr,
                                     2 * v.stride, chalf.size(), fused ? cells.ids.get() : nullptr);
    if (split)
        launch_rerank_recheck(cdata.get(), sl.q.get(), d, kAngular, Ceff, Qc, kq, sl.out.get(), s, sl.qorder.get(),
                              sl.rband);
}

void lsh_index::prepare_slots(bool sketch, int kq) {
    const int QM = kGraphMaxQ;
    const int L = tables;
    const int Ceff = (int)std::max<int64_t>(1, std::min<int64_t>(candidates, N));
    const int Cp = pow2_ceil(std::max(Ceff, 2));
    const int K = sketch ? Ceff : std::min(kMaxTopK, pow2_ceil(kq + margin));
    // Sized already for this layout: nothing to do (the common case, once a
    // batch per call of a request loop).  Otherwise batches still in flight
    // on the slots finish before any buffer or graph is replaced.
    const std::vector<int64_t> key = {sketch, K, Cp, d, wpad, L, QM, cells.built() ? (int64_t)cells.P : -1,
                                      cells.built() ? (int64_t)cells.words : -1};
    if (key == slots_key && (int)slots.size() == kBatchSlots) return;
    for (auto& sl : slots) LSH_HIP(hipStreamSynchronize(sl->stream));
    while ((int)slots.size() < kBatchSlots) {
        auto sl = std::make_unique<batch_slot>();
        LSH_HIP(hipStreamCreateWithFlags(&sl->stream, hipStreamNonBlocking));
        slots.push_back(std::move(sl));
    }
    for (auto& slp : slots) {
        batch_slot& sl = *slp;
        sl.q.ensure((size_t)QM * d);
        sl.out.ensure((size_t)QM * kGraphMaxK);
        std::vector<uintptr_t> sig = {(uintptr_t)sketch, (uintptr_t)K, (uintptr_t)sl.q.get(), (uintptr_t)sl.out.get()};
        if (sketch) {
            const int qrows = sketch_query_rows(QM);
            const uint32_t* old_qsk = sl.qsk.get();
            sl.qsk.ensure((size_t)qrows * wpad);
            // Words past the sketch are never written and must stay zero.
            if (!sl.qsk_zeroed || sl.qsk.get() != old_qsk) {
                LSH_HIP(hipMemsetAsync(sl.qsk.get(), 0, sizeof(uint32_t) * (size_t)qrows * wpad, sl.stream));
                sl.qsk_zeroed = true;
            }
            sl.st.reset(QM, K, 2 * K + 512, sl.stream, true, true, false);
            sl.qorder.ensure((size_t)QM);
            cell_search_reserve(cells, QM, sl.scratch, 0);
            sl.rband.band.ensure((size_t)QM * Cp);
            sl.rband.cnt.ensure((size_t)QM);
            sl.flags.ensure((size_t)QM + 3);
            if (fuse_ok(1)) sl.q16.ensure(query_half_elems(QM, fused_k_steps(d)));
            sig.insert(sig.end(), {(uintptr_t)sl.q16.get(), (uintptr_t)sl.qsk.get(), (uintptr_t)sl.qorder.get(),
                                   (uintptr_t)sl.scratch.raw[0].get(), (uintptr_t)sl.scratch.raw[1].get(),
                                   (uintptr_t)sl.scratch.raw[2].get(), (uintptr_t)sl.scratch.raw[3].get(),
                                   (uintptr_t)sl.rband.band.get(), (uintptr_t)sl.rband.cnt.get()});
        } else {
            sl.qkeys.ensure((size_t)QM * L);
            sl.qlo.ensure((size_t)QM * L);
            sl.qlen.ensure((size_t)QM * L);
            sl.qpre.ensure((size_t)QM * (L + 1));
            sl.qmax.ensure(1);
            sl.wl_ibase.ensure((size_t)QM + 1);
            sl.wl_flags.ensure((size_t)QM);
            sl.wl_part.ensure((size_t)worklist_cap(K) * K);
            sl.flags.ensure((size_t)QM);
            sl.st.reset(QM, K, newcap_for(K), sl.stream, false, false, false);
            sig.insert(sig.end(), {(uintptr_t)sl.qkeys.get(), (uintptr_t)sl.qlo.get(), (uintptr_t)sl.qlen.get(),
                                   (uintptr_t)sl.qpre.get(), (uintptr_t)sl.qmax.get(), (uintptr_t)sl.wl_ibase.get(),
                                   (uintptr_t)sl.wl_flags.get(), (uintptr_t)sl.wl_part.get(),
                                   (uintptr_t)worklist_cap(K)});
        }
        const TopKView sv = sl.st.view();
        sig.insert(sig.end(), {(uintptr_t)sl.flags.get(), (uintptr_t)sv.keys, (uintptr_t)sv.cnt, (uintptr_t)sv.thr,
                               (uintptr_t)sv.ovf});
        if (sig != sl.sig) {
            sl.drop_graphs();
            sl.sig = sig;
        }
    }
    slots_key = key;
}

void lsh_index::enqueue_slot_batch(batch_slot& sl, bool sketch, const float* q, hipMemcpyKind qk, int b, int kq,
                                   lsh_point_t* res, hipMemcpyKind rk, uint32_t* flags_out, uint32_t* extra_out) {
    const hipStream_t ss = sl.stream;
    const int Qc = graph_class(b);
    LSH_HIP(hipMemcpyAsync(sl.q.get(), q, sizeof(float) * (size_t)b * d, qk, ss));
    // The class's rows past the batch: for the SimHash index copies of its
    // last query (a search of ordinary cost), for the (K,L) rescan zero rows
    // (they normalise to NaN, whose keys match no real point, so they cost
    // nearly nothing); their results are not copied out.
    if (sketch) launch_repeat_row_f32(sl.q.get(), d, b - 1, Qc - b, ss);
    else if (Qc > b) LSH_HIP(hipMemsetAsync(sl.q.get() + (size_t)b * d, 0, sizeof(float) * (size_t)(Qc - b) * d, ss));
    auto it = sl.graphs.find({Qc, kq});
    if (it == sl.graphs.end()) {
        hipGraph_t graph = nullptr;
        LSH_HIP(hipStreamBeginCapture(ss, hipStreamCaptureModeThreadLocal));
        try {
            if (sketch) {
                enqueue_sketch_search(sl, Qc, kq);
            } else {
                PhaseTimer none(false);
                search_lsh(sl.work(), Qc, kq, ss, none, true);
            }
        } catch (...) {
            (void)hipStreamEndCapture(ss, &graph);
            if (graph) (void)hipGraphDestroy(graph);
            throw;
        }
        LSH_HIP(hipStreamEndCapture(ss, &graph));
        hipGraphExec_t exec = nullptr;
        const hipError_t e = hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0);
        (void)hipGraphDestroy(graph);
        LSH_HIP(e);
        it = sl.graphs.emplace(std::make_pair(Qc, kq), exec).first;
    }
    LSH_HIP(hipGraphLaunch(it->second, ss));
    ++batch_graphs;
    LSH_HIP(hipMemcpyAsync(res, sl.out.get(), sizeof(lsh_point_t) * (size_t)b * kq, rk, ss));
    LSH_HIP(hipMemcpyAsync(flags_out, sl.flags.get(), sizeof(uint32_t) * (size_t)b, hipMemcpyDeviceToHost, ss));
    if (sketch)
        LSH_HIP(hipMemcpyAsync(extra_out, sl.flags.get() + Qc, sizeof(uint32_t) * 3, hipMemcpyDeviceToHost, ss));
    else
        LSH_HIP(hipMemcpyAsync(extra_out, sl.wl_flags.get(), sizeof(uint32_t) * (size_t)b, hipMemcpyDeviceToHost, ss));
}

bool lsh_index::enqueue_batch(const float* q, bool q_on_device, int n, int k, lsh_point_t* res, bool r_on_device,
                              uint32_t* flags_out, hipEvent_t ready, hipEvent_t done) {
    DeviceGuard g(device);
    if (n < 1 || n > kGraphMaxQ || k < 1) return false;
    const int kq = calculate_k(k);
    if (!sketch_graphable(kq)) return false;
    prepare_slots(true, kq);
    batch_slot& sl = *slots[next_slot];
    next_slot = (next_slot + 1) % (int)slots.size();
    if (ready) LSH_HIP(hipStreamWaitEvent(sl.stream, ready, 0));
    if (done_rec) LSH_HIP(hipStreamWaitEvent(sl.stream, done_ev, 0));
    enqueue_slot_batch(sl, true, q, q_on_device ? hipMemcpyDeviceToDevice : hipMemcpyHostToDevice, n, kq, res,
                       r_on_device ? hipMemcpyDeviceToDevice : hipMemcpyDeviceToHost, flags_out, flags_out + n);
    if (done) LSH_HIP(hipEventRecord(done, sl.stream));
    return true;
}

void lsh_index::search_batches_graphs(bool sketch, const float* queries, const int* sizes, int n_batches, int k,
                                      int64_t total, bool q_on_device, lsh_point_t* results, bool r_on_device,
                                      hipStream_t user_stream) {
    const int kq = calculate_k(k);
    const int QM = kGraphMaxQ;
    // The call's work is ordered after what was queued on the caller's
    // stream (or the index's own one; the legacy null stream cannot be
    // captured).
    hipStream_t s = user_stream ? user_stream : stream;
    if (done_rec) LSH_HIP(hipStreamWaitEvent(s, done_ev, 0));
    prepare_slots(sketch, kq);
    if (!done_ev) LSH_HIP(hipEventCreateWithFlags(&done_ev, hipEventDisableTiming));
    LSH_HIP(hipEventRecord(done_ev, s));
    for (auto& sl : slots) LSH_HIP(hipStreamWaitEvent(sl->stream, done_ev, 0));
    auto sync_slots = [&] {
        for (auto& sl : slots) LSH_HIP(hipStreamSynchronize(sl->stream));
    };
    // Pinned per call: the flags of every query, then per batch three words
    // (sketch: the cell search's status; (K,L): none) and, for the (K,L)
    // rescan, the work list's capacity flags of every query.
    const size_t need = 2 * (size_t)total + 3 * (size_t)n_batches;
    if (need > bflags_n) {
        sync_slots();
        LSH_HIP(hipStreamSynchronize(s));
        if (bflags) (void)hipHostFree(bflags);
        bflags = nullptr;
        bflags_n = 0;
        void* p = nullptr;
        LSH_HIP(hipHostMalloc(&p, sizeof(uint32_t) * std::max<size_t>(need, 1), hipHostMallocDefault));
        bflags = static_cast<uint32_t*>(p);
        bflags_n = std::max<size_t>(need, 1);
    }
    uint32_t* const wflags = bflags + total;                  // (K,L): capacity flags
    uint32_t* const status = bflags + 2 * (size_t)total;      // three words per batch
    const hipMemcpyKind qk = q_on_device ? hipMemcpyDeviceToDevice : hipMemcpyHostToDevice;
    const hipMemcpyKind rk = r_on_device ? hipMemcpyDeviceToDevice : hipMemcpyDeviceToHost;
    int64_t off = 0;
    for (int i = 0; i < n_batches; ++i) {
        const int b = sizes[i];
        uint32_t* st3 = status + 3 * (size_t)i;
        st3[0] = st3[1] = st3[2] = 0u;
        if (b == 0) continue;
        if (b > QM) {
            // Larger than any class: an ordinary search of its own (complete
            // when it returns).
            sync_slots();
            search(queries + off * d, b, k, q_on_device, results + off * kq, r_on_device, nullptr);
            for (int j = 0; j < b; ++j) bflags[off + j] = wflags[off + j] = 0u;
            off += b;
            continue;
        }
        batch_slot& sl = *slots[next_slot];
        next_slot = (next_slot + 1) % (int)slots.size();
        enqueue_slot_batch(sl, sketch, queries + off * d, qk, b, kq, results + off * kq, rk, bflags + off,
                           sketch ? st3 : wflags + off);
        off += b;
    }
    sync_slots();
    // SimHash batches whose second pass was too dense ran no scan: each is
    // redone by an ordinary search (its sampled threshold pass), and later
    // batch calls on this index skip the graphs.  Flagged queries of the
    // other batches (incomplete, or the (K,L) work list's capacity or FP32
    // band) are redone together by one ordinary search.
    std::vector<int64_t> bad;
    off = 0;
    for (int i = 0; i < n_batches; ++i) {
        const int b = sizes[i];
        if (sketch && b > 0 && b <= QM && status[3 * (size_t)i]) {
            cells_dense = true;
            batch_redone += b;
            search(queries + off * d, b, k, q_on_device, results + off * kq, r_on_device, nullptr);
        } else {
            for (int j = 0; j < b; ++j)
                if (bflags[off + j] || (!sketch && wflags[off + j])) bad.push_back(off + j);
        }
        off += b;
    }
    batch_redone += (int64_t)bad.size();
    if (!bad.empty()) {
        const int m = (int)bad.size();
        std::vector<float> hq((size_t)m * d);
        std::vector<lsh_point_t> hr((size_t)m * kq);
        for (int i = 0; i < m; ++i)
            LSH_HIP(hipMemcpy(hq.data() + (size_t)i * d, queries + bad[i] * d, sizeof(float) * d,
                              q_on_device ? hipMemcpyDeviceToHost : hipMemcpyHostToHost));
        search(hq.data(), m, k, false, hr.data(), false, nullptr);
        for (int i = 0; i < m; ++i)
            LSH_HIP(hipMemcpy(results + bad[i] * kq, hr.data() + (size_t)i * kq, sizeof(lsh_point_t) * kq,
                              r_on_device ? hipMemcpyHostToDevice : hipMemcpyHostToHost));
    }
    LSH_HIP(hipEventRecord(done_ev, s));
    done_rec = true;
}

// lshFramework.cuh:142-280 scan + removeDuplicates over the unique candidate
// set of each query: either all of it ranked by the exact distance
// (distanceFunc), or its top-C by sketch distance reranked exactly
// (runWithSketchedData, candidateSetScanner.cuh:10-32).
bool lsh_index::lsh_graphable(int kq) const {
    return framework == 1 && !(withSketch && impl >= 3) && keysImpl != 7 &&
           std::min(kMaxTopK, pow2_ceil(kq + margin)) <= 512 && !force_rounds_lsh();
}

void lsh_index::search_lsh(const lsh_work& w, int Q, int kq, hipStream_t s, PhaseTimer& pt, bool enqueue_only) {
    const int L = tables;
    pt.begin(PH_PREP, s);
    apply_steps_query(key_steps, w.q.get(), Q, d, s);
    w.qkeys.ensure((size_t)Q * L);
    lsh_keys(w.q.get(), Q, w.qkeys.get(), s);
    apply_steps_query(sketch_steps, w.q.get(), Q, d, s);
    w.qlo.ensure((size_t)Q * L);
    w.qpre.ensure((size_t)Q * (L + 1));
    w.qmax.ensure(1);
    launch_fill_i32(reinterpret_cast<int32_t*>(w.qmax.get()), 1, 0, s);
    if (sorted) {
        w.qlen.ensure((size_t)Q * L);
        launch_query_ranges_sorted(w.qkeys.get(), Q, L, N, bkeys.get(), w.qlo.get(), w.qlen.get(), w.qpre.get(), w.qmax.get(),
                                   s);
    } else {
        launch_query_ranges(w.qkeys.get(), Q, L, tableSize, boff.get(), w.qlo.get(), w.qpre.get(), w.qmax.get(), s);
    }
    const bool sketched = withSketch && impl >= 3;
    if (sketched) {
        if (impl == 6) {
            qjl.ensure((size_t)Q * sdim);
            launch_jl(w.q.get(), Q, d, planesT.get(), sdim, qjl.get(), s);
        } else {
            qsk.ensure((size_t)Q * wpad);
            if (impl == 3) sketch_simhash(w.q.get(), Q, qsk.get(), wpad, s);
            else wmh.sketch(w.q.get(), Q, d, qsk.get(), wpad, s);
        }
    }
    pt.end(PH_PREP, s);
    const bool worklist = !sketched && std::min(kMaxTopK, pow2_ceil(kq + margin)) <= 512 && !force_rounds_lsh();
    uint32_t maxS = 0;
    if (!worklist) {
        LSH_HIP(hipMemcpyAsync(&maxS, w.qmax.get(), sizeof(uint32_t), hipMemcpyDeviceToHost, s));
        LSH_HIP(hipStreamSynchronize(s));
    }

    LshScanArgs a{};
    a.perm = bperm.get();
    a.dkeys = dkeys.get();
    a.qkeys = w.qkeys.get();
    a.qlo = w.qlo.get();
    a.qpre = w.qpre.get();
    a.qmap = nullptr;
    a.L = L;
    a.Q = Q;
    a.data = data.get();
    a.queries = w.q.get();
    a.d = d;
    a.dsk = dsk.get();
    a.qsk = qsk.get();
    a.words = wpad;
    a.djl = djl.get();
    a.qjl = qjl.get();
    a.sdim = sdim;
    const std::vector<int32_t> none;
    const int kind = dfunc;  // candidateSetScan / runDistanceFunction use distanceFunc
    const double eps = fp32_error_bound(kind, d);
    if (sketched) {
        a.sketch = true;
        a.kind = impl == 6 ? kSqL2 : (impl == 4 ? 1 : 0);
        const bool int_keys = impl != 6;
        const int Ceff = (int)std::max<int64_t>(1, std::min<int64_t>(candidates, N));
        const int K = int_keys && Ceff <= kMaxIntTopK ? Ceff : pow2_ceil(Ceff);
        LSH_REQUIRE(K <= kMaxTopK, LSH_ERR_UNSUPPORTED, "candidate sets larger than 4096 are not supported");
        lsh_topk(a, none, maxS, K, kNormal, int_keys && K <= kMaxIntTopK, w.st, scratch, 0, s, pt, scan_launches);
        pt.begin(PH_FINAL, s);
        launch_rerank(data.get(), w.q.get(), d, kind, w.st.view().keys, w.st.view().stride, nullptr, Ceff, Q, kq, eps,
                      w.outp, s, nullptr, nullptr, nullptr, &rband);
        pt.end(PH_FINAL, s);
        return;
    }
    a.sketch = false;
    a.kind = kind;
    LSH_REQUIRE(kq <= kMaxTopK, LSH_ERR_UNSUPPORTED, "k must be <= 4096");
    const int K = std::min(kMaxTopK, pow2_ceil(kq + margin));
    std::vector<int32_t> bad;
    if (worklist) {
        // Every query's whole list in one pass of items, no host reads; the
        // queries flagged by the merge (item capacity) or the finalisation
        // (FP32 band) are redone below with the rounds.
        const uint32_t cap = worklist_cap(K);
        w.wl_ibase.ensure((size_t)Q + 1);
        w.wl_part.ensure((size_t)cap * K);
        w.wl_flags.ensure((size_t)Q);
        w.st.reset(Q, K, newcap_for(K), s, false, false, false);
        pt.begin(PH_SCAN, s);
        launch_lsh_exact_worklist(a, K, w.wl_ibase.get(), cap, w.wl_part.get(), w.st.view(), w.wl_flags.get(), s);
        pt.end(PH_SCAN, s);
        ++scan_launches;
        pt.begin(PH_FINAL, s);
        launch_exact_finalize(data.get(), w.q.get(), d, kind, w.st.view(), kq, eps, w.outp, w.flags.get(), s);
        pt.end(PH_FINAL, s);
        if (enqueue_only) return;  // the caller reads w.flags and w.wl_flags later
        bad = flagged_either(w.flags.get(), w.wl_flags.get(), Q, s);
        if (bad.empty()) return;
        LSH_HIP(hipMemcpyAsync(&maxS, w.qmax.get(), sizeof(uint32_t), hipMemcpyDeviceToHost, s));
        LSH_HIP(hipStreamSynchronize(s));
    } else {
        lsh_topk(a, none, maxS, K, kNormal, false, w.st, scratch, 0, s, pt, scan_launches);
        pt.begin(PH_FINAL, s);
        launch_exact_finalize(data.get(), w.q.get(), d, kind, w.st.view(), kq, eps, w.outp, w.flags.get(), s);
        pt.end(PH_FINAL, s);
        bad = flagged(w.flags.get(), Q, s);
        if (bad.empty()) return;
    }
    LSH_REQUIRE(K < kMaxTopK, LSH_ERR_UNSUPPORTED,
                "more than 4096 candidates within the FP32 error band of the k-th distance");
    const int m = (int)bad.size();
    DevBuf<int32_t> idx((size_t)m);
    LSH_HIP(hipMemcpyAsync(idx.get(), bad.data(), sizeof(int32_t) * m, hipMemcpyHostToDevice, s));
    DevBuf<float> sub((size_t)m * d);
    launch_gather_rows_f32(w.q.get(), d, idx.get(), m, sub.get(), s);
    LshScanArgs b = a;
    b.qmap = idx.get();
    b.Q = m;
    TopKState sst;
    lsh_topk(b, bad, maxS, kMaxTopK, kNormal, false, sst, scratch, 0, s, pt, scan_launches);
    DevBuf<lsh_point_t> sout((size_t)m * kq);
    DevBuf<uint32_t> sflags((size_t)m);
    launch_exact_finalize(data.get(), sub.get(), d, kind, sst.view(), kq, eps, sout.get(), sflags.get(), s);
    LSH_REQUIRE(flagged(sflags.get(), m, s).empty(), LSH_ERR_UNSUPPORTED,
                "more than 4096 candidates within the FP32 error band of the k-th distance");
    launch_scatter_rows_pts(sout.get(), kq, idx.get(), m, w.outp, s);
    LSH_HIP(hipStreamSynchronize(s));
}

// ---------------------------------------------------------- stage helpers

namespace lshgpu {

int calculate_k(int k) {
    // launchHelper.cuh:21-25
    int divisor = (k - 1) / 32;
    return divisor * 32 + 32;
}

void simhash_host(const float* x, int64_t n, int d, int words, int hashBits, const float* planes,
                  uint32_t* out, int device) {
    DeviceGuard g(device);
    LSH_REQUIRE(n >= 0 && d > 0 && words > 0 && hashBits >= 1 && hashBits <= 32, LSH_ERR_INVALID,
                "bad SimHash arguments");
    std::vector<float> pl;
    if (planes)
        pl.assign(planes, planes + (size_t)words * hashBits * d);
    else
        pl = normal_stream((int64_t)words * hashBits * d, 1.0);
    std::vector<float> pt = transpose_planes(pl, words, hashBits, d);
    DevBuf<float> dx((size_t)n * d), dp(pt.size());
    DevBuf<uint32_t> dout((size_t)n * words);
    LSH_HIP(hipMemcpy(dx.get(), x, sizeof(float) * (size_t)n * d, hipMemcpyHostToDevice));
    LSH_HIP(hipMemcpy(dp.get(), pt.data(), sizeof(float) * pt.size(), hipMemcpyHostToDevice));
    // The matrix-core kernel where it applies (as an index sketches), unless
    // LSHGPU_SIMHASH_VALU=1 asks for the FP32 vector kernel.
    const char* valu = std::getenv("LSHGPU_SIMHASH_VALU");
    if (simhash_mfma_supported(d, hashBits) && !(valu && valu[0] == '1')) {
        SimhashMfma m;
        m.build(pl.data(), d, words, dp.get(), nullptr);
        m.sketch(dx.get(), n, dout.get(), words, nullptr);
    } else {
        launch_simhash(dx.get(), n, d, dp.get(), words, hashBits, dout.get(), words, nullptr);
    }
    LSH_HIP(hipDeviceSynchronize());
    LSH_HIP(hipMemcpy(out, dout.get(), sizeof(uint32_t) * (size_t)n * words, hipMemcpyDeviceToHost));
}

void jl_host(const float* x, int64_t n, int d, int sdim, const float* planes, float* out, int device) {
    DeviceGuard g(device);
    LSH_REQUIRE(n >= 0 && d > 0 && sdim > 0, LSH_ERR_INVALID, "bad JL arguments");
    std::vector<float> pl;
    if (planes)
        pl.assign(planes, planes + (size_t)sdim * d);
    else
        pl = normal_stream((int64_t)sdim * d, 1.0 / sdim);
    std::vector<float> pt = transpose_jl(pl, sdim, d);
    DevBuf<float> dx((size_t)n * d), dp(pt.size()), dout((size_t)n * sdim);
    LSH_HIP(hipMemcpy(dx.get(), x, sizeof(float) * (size_t)n * d, hipMemcpyHostToDevice));
    LSH_HIP(hipMemcpy(dp.get(), pt.data(), sizeof(float) * pt.size(), hipMemcpyHostToDevice));
    launch_jl(dx.get(), n, d, dp.get(), sdim, dout.get(), nullptr);
    LSH_HIP(hipDeviceSynchronize());
    LSH_HIP(hipMemcpy(out, dout.get(), sizeof(float) * (size_t)n * sdim, hipMemcpyDeviceToHost));
}

void gj_transform_host(float* data, int64_t N, float* queries, int64_t Q, int d, int device) {
    DeviceGuard g(device);
    DevBuf<float> dd((size_t)N * d), dq((size_t)Q * d);
    DevBuf<int> m((size_t)d);
    LSH_HIP(hipMemcpy(dd.get(), data, sizeof(float) * (size_t)N * d, hipMemcpyHostToDevice));
    if (Q > 0) LSH_HIP(hipMemcpy(dq.get(), queries, sizeof(float) * (size_t)Q * d, hipMemcpyHostToDevice));
    launch_fill_i32(m.get(), (size_t)d, 0, nullptr);
    launch_gj_mins(dd.get(), N, d, m.get(), nullptr);
    if (Q > 0) launch_gj_mins(dq.get(), Q, d, m.get(), nullptr);
    launch_gj_shift(dd.get(), N, d, m.get(), nullptr);
    launch_normalize_rows(dd.get(), N, d, nullptr);
    if (Q > 0) {
        launch_gj_shift(dq.get(), Q, d, m.get(), nullptr);
        launch_normalize_rows(dq.get(), Q, d, nullptr);
    }
    LSH_HIP(hipDeviceSynchronize());
    LSH_HIP(hipMemcpy(data, dd.get(), sizeof(float) * (size_t)N * d, hipMemcpyDeviceToHost));
    if (Q > 0) LSH_HIP(hipMemcpy(queries, dq.get(), sizeof(float) * (size_t)Q * d, hipMemcpyDeviceToHost));
}

void hamming_topc_host(const uint32_t* ds, int64_t N, const uint32_t* qs, int64_t Q, int words, int C,
                       int32_t* ids, int32_t* dists, int device) {
    DeviceGuard g(device);
    LSH_REQUIRE(N >= 0 && Q >= 0 && words > 0 && C >= 1, LSH_ERR_INVALID, "bad Hamming top-C arguments");
    const int K = C <= kMaxIntTopK ? C : pow2_ceil(C);
    LSH_REQUIRE(K <= kMaxTopK, LSH_ERR_UNSUPPORTED, "C must be <= 4096");
    if (Q == 0) return;
    const int wpad = round_up(words, 8);
    const int Qpad = sketch_query_rows((int)Q);
    DevBuf<uint32_t> draw((size_t)std::max<int64_t>(N, 1) * words), qraw((size_t)Q * words);
    DevBuf<uint32_t> dpad((size_t)std::max<int64_t>(N, 1) * wpad), qpad((size_t)Qpad * wpad);
    if (N > 0) LSH_HIP(hipMemcpy(draw.get(), ds, sizeof(uint32_t) * (size_t)N * words, hipMemcpyHostToDevice));
    LSH_HIP(hipMemcpy(qraw.get(), qs, sizeof(uint32_t) * (size_t)Q * words, hipMemcpyHostToDevice));
    launch_pad_rows_u32(draw.get(), N, words, N, wpad, dpad.get(), nullptr);
    launch_pad_rows_u32(qraw.get(), Q, words, Qpad, wpad, qpad.get(), nullptr);
    TopKState st;
    PhaseTimer pt(false);
    int64_t launches = 0;
    DevBuf<uint32_t> fp4;
    std::vector<SampleLevel> lv;
    CellIndex cells;
    if (hamming_mfma_supported(wpad) && N >= 16 * 1024) {
        fp4.alloc(hamming_fp4_words(N, wpad));
        launch_expand_fp4(dpad.get(), N, wpad, fp4.get(), nullptr);
        build_sample_levels(dpad.get(), N, wpad, C, lv, nullptr);
        if (!lv.empty()) cells.build(dpad.get(), N, wpad, nullptr);
    }
    const SketchSet full{dpad.get(), fp4.get(), N, &cells, nullptr, nullptr};
    const std::vector<SketchSet> sv = sample_views(lv);
    Scratch sc;
    sketch_topk(full, sv.data(), (int)sv.size(), wpad, 0, qpad.get(), (int)Q, K, kNormal, st, sc, 0, nullptr, pt,
                launches);
    launch_sort_rows(st.view(), nullptr);
    LSH_HIP(hipDeviceSynchronize());
    const TopKView v = st.view();
    std::vector<uint64_t> keys((size_t)Q * v.stride);
    LSH_HIP(hipMemcpy(keys.data(), v.keys, sizeof(uint64_t) * keys.size(), hipMemcpyDeviceToHost));
    for (int64_t q = 0; q < Q; ++q)
        for (int c = 0; c < C; ++c) {
            const uint64_t key = keys[(size_t)q * v.stride + c];
            ids[q * C + c] = key == kEmptyKey ? -1 : (int32_t)(uint32_t)key;
            if (dists) dists[q * C + c] = key == kEmptyKey ? -1 : (int32_t)(key >> 32);
        }
}

void wmh_host(const float* x, int64_t n, int d, int sketchDim, int one_bit, uint32_t* out, int device) {
    DeviceGuard g(device);
    LSH_REQUIRE(n >= 0 && d > 0 && sketchDim > 0, LSH_ERR_INVALID, "bad weighted MinHash arguments");
    DevBuf<float> dx((size_t)n * d);
    DevBuf<uint32_t> dout((size_t)n * sketchDim);
    LSH_HIP(hipMemcpy(dx.get(), x, sizeof(float) * (size_t)n * d, hipMemcpyHostToDevice));
    WmhHasher h;
    h.build(dx.get(), n, nullptr, 0, d, one_bit ? 1 : 0, one_bit ? 32 * sketchDim : sketchDim, 32, nullptr);
    h.sketch(dx.get(), n, d, dout.get(), sketchDim, nullptr);
    LSH_HIP(hipDeviceSynchronize());
    LSH_HIP(hipMemcpy(out, dout.get(), sizeof(uint32_t) * (size_t)n * sketchDim, hipMemcpyDeviceToHost));
}

void crosspoly_dense_host(const float* x, int64_t n, int d, int tables, int K, uint32_t* keys, int device) {
    DeviceGuard g(device);
    LSH_REQUIRE(n >= 0 && d >= 1 && tables >= 1 && K >= 1, LSH_ERR_INVALID, "bad dense cross-polytope arguments");
    LSH_REQUIRE(d <= 128, LSH_ERR_UNSUPPORTED, "dense cross-polytope keys support d <= 128");
    LSH_REQUIRE(K * dense_key_field_bits(d) <= 32, LSH_ERR_UNSUPPORTED, "K hashes do not fit a 32-bit key");
    if (n == 0) return;
    DenseRotations dr;
    dr.build(d, tables, K, nullptr);
    DevBuf<float> dx((size_t)n * d);
    DevBuf<uint32_t> dk((size_t)n * tables);
    LSH_HIP(hipMemcpy(dx.get(), x, sizeof(float) * (size_t)n * d, hipMemcpyHostToDevice));
    launch_normalize_rows(dx.get(), n, d, nullptr);
    dr.keys(dx.get(), n, dk.get(), nullptr, nullptr);
    LSH_HIP(hipDeviceSynchronize());
    LSH_HIP(hipMemcpy(keys, dk.get(), sizeof(uint32_t) * (size_t)n * tables, hipMemcpyDeviceToHost));
}

void sort_tables_host(const uint32_t* keys_ln, int64_t N, int L, int bits, uint32_t* skeys, uint32_t* ids,
                      int device) {
    DeviceGuard g(device);
    LSH_REQUIRE(N >= 0 && L >= 1 && bits >= 1 && bits <= 32, LSH_ERR_INVALID, "bad table sort arguments");
    if (N == 0) return;
    const size_t n = (size_t)L * N;
    DevBuf<uint32_t> in(n), ok(n), oi(n), tk(n), ti(n), counts(sort_table_counts(N, L));
    LSH_HIP(hipMemcpy(in.get(), keys_ln, sizeof(uint32_t) * n, hipMemcpyHostToDevice));
    launch_sort_tables(in.get(), N, L, bits, ok.get(), oi.get(), tk.get(), ti.get(), counts.get(), nullptr);
    LSH_HIP(hipDeviceSynchronize());
    LSH_HIP(hipMemcpy(skeys, ok.get(), sizeof(uint32_t) * n, hipMemcpyDeviceToHost));
    LSH_HIP(hipMemcpy(ids, oi.get(), sizeof(uint32_t) * n, hipMemcpyDeviceToHost));
}

void crosspoly_host(const float* x, int64_t n, int d, int tables, const int32_t* signs, uint16_t* keys,
                    int device) {
    DeviceGuard g(device);
    int P = 1;
    while (P < d) P *= 2;
    std::vector<int32_t> sg;
    if (signs)
        sg.assign(signs, signs + (size_t)tables * 3 * P);
    else
        sg = sign_stream((int64_t)tables * 3 * P);
    DevBuf<float> dx((size_t)n * d);
    DevBuf<int32_t> ds(sg.size());
    DevBuf<uint16_t> dk((size_t)n * tables);
    LSH_HIP(hipMemcpy(dx.get(), x, sizeof(float) * (size_t)n * d, hipMemcpyHostToDevice));
    LSH_HIP(hipMemcpy(ds.get(), sg.data(), sizeof(int32_t) * sg.size(), hipMemcpyHostToDevice));
    launch_normalize_rows(dx.get(), n, d, nullptr);
    launch_crosspoly(dx.get(), n, d, tables, ds.get(), dk.get(), nullptr);
    LSH_HIP(hipDeviceSynchronize());
    LSH_HIP(hipMemcpy(keys, dk.get(), sizeof(uint16_t) * (size_t)n * tables, hipMemcpyDeviceToHost));
}

}  // namespace lshgpu
