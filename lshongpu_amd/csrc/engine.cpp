// engine.cpp -- resident index: preprocessing, sketch builds and search.
//
// Search is organised in rounds (see select.hip): a scan kernel sweeps a
// range of data points for every query of the batch and appends the pairs
// that beat the query's current threshold; a select kernel folds them into
// the query's best-K list and tightens the threshold.  Range sizes grow
// geometrically so that each round is expected to append about half of the
// append capacity per query; a query whose appends overflow is re-run on its
// own with ranges no larger than the capacity, which cannot overflow.
#include "engine.hpp"

#include <algorithm>
#include <cmath>
#include <cstdlib>

using namespace lshgpu;

namespace {

enum Phase { PH_PREP = 0, PH_SCAN = 1, PH_SELECT = 2, PH_FINAL = 3, PH_TOTAL = 4, PH_MFMA = 5 };

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

// Indices of the nonzero entries of a device flag array.
std::vector<int32_t> flagged(const uint32_t* dflags, int Q, hipStream_t s) {
    std::vector<uint32_t> h((size_t)Q);
    LSH_HIP(hipMemcpyAsync(h.data(), dflags, sizeof(uint32_t) * (size_t)Q, hipMemcpyDeviceToHost, s));
    LSH_HIP(hipStreamSynchronize(s));
    std::vector<int32_t> idx;
    for (int i = 0; i < Q; ++i)
        if (h[i]) idx.push_back(i);
    return idx;
}

// Rounds over points [0, N) (see the file comment).  `scan(p0, p1, view)`
// launches the scan kernel of one range.
template <class ScanFn>
void run_rounds(TopKState& st, int Q, int K, int64_t N, bool fixed, ScanFn&& scan, hipStream_t s,
                PhaseTimer& pt, int64_t& launches, bool int_keys = false) {
    const int newcap = newcap_for(K);
    st.reset(Q, K, newcap, s, int_keys);
    const TopKView v = st.view();
    int64_t p = 0;
    int64_t chunk = std::min<int64_t>(N, newcap);
    while (p < N) {
        const int64_t p1 = std::min<int64_t>(N, p + chunk);
        pt.begin(PH_SCAN, s);
        scan(p, p1, v);
        pt.end(PH_SCAN, s);
        ++launches;
        pt.begin(PH_SELECT, s);
        launch_select_merge(v, s);
        pt.end(PH_SELECT, s);
        if (!fixed) chunk = std::max<int64_t>(256, (int64_t)((double)p1 * newcap / (2.0 * K)));
        p = p1;
    }
}

// Prepared queries of an exact-distance scan over `data` (rows of length d).
struct ExactQueries {
    const float* rows;  // [Q][d]
    int Q;
};

void exact_topk(const float* data, const float* data_inv, int64_t N, int d, int kind,
                const ExactQueries& eq, int K, bool fixed, TopKState& st, hipStream_t s, PhaseTimer& pt,
                int64_t& launches) {
    const int QG = exact_scan_query_group();
    const int groups = (eq.Q + QG - 1) / QG;
    DevBuf<float> qT((size_t)groups * d * QG);
    DevBuf<float> qinv(angular_kind(kind) ? (size_t)eq.Q : 0);
    launch_transpose_queries(eq.rows, eq.Q, d, qT.get(), s);
    if (angular_kind(kind)) launch_row_inv_norms(eq.rows, eq.Q, d, qinv.get(), s);
    auto scan = [&](int64_t p0, int64_t p1, const TopKView& v) {
        ExactScanArgs a{data, data_inv, qT.get(), qinv.get(), d, eq.Q, p0, p1, kind};
        launch_exact_scan(a, v, s);
    };
    run_rounds(st, eq.Q, K, N, fixed, scan, s, pt, launches);
    if (fixed) return;
    const std::vector<int32_t> ovf = flagged(st.view().ovf, eq.Q, s);
    if (ovf.empty()) return;
    const int m = (int)ovf.size();
    DevBuf<int32_t> idx((size_t)m);
    LSH_HIP(hipMemcpyAsync(idx.get(), ovf.data(), sizeof(int32_t) * m, hipMemcpyHostToDevice, s));
    DevBuf<float> sub((size_t)m * d);
    launch_gather_rows_f32(eq.rows, d, idx.get(), m, sub.get(), s);
    TopKState sst;
    exact_topk(data, data_inv, N, d, kind, ExactQueries{sub.get(), m}, K, true, sst, s, pt, launches);
    const TopKView sv = sst.view(), v = st.view();
    launch_scatter_rows_u64(sv.keys, sv.stride, K, idx.get(), m, v.keys, v.stride, s);
    LSH_HIP(hipStreamSynchronize(s));
}

// Rows the query sketch buffer of an integer sketch scan must have.
int sketch_query_rows(int Q) { return round_up(std::max(Q, 1), 256); }

// Integer sketch scan: qs has sketch_query_rows(Q) rows of wpad words (zero
// past Q).  Hamming scans of up to 1024 bits run on the matrix cores.
void sketch_topk(const uint32_t* ds, int64_t N, int wpad, int kind, const uint32_t* qs, int Q, int K,
                 bool fixed, TopKState& st, hipStream_t s, PhaseTimer& pt, int64_t& launches,
                 int64_t* mfma_launches = nullptr, int64_t* mfma_points = nullptr) {
    const bool mfma = kind == 0 && hamming_mfma_supported(wpad) && !force_valu_scan();
    auto scan = [&](int64_t p0, int64_t p1, const TopKView& v) {
        // The matrix-core scan queues candidates in LDS, which suits ranges
        // where few pairs pass: once 16*K points are behind every threshold,
        // a 256-query x 32-point tile yields about 256*32*K/p0 <= 512 of them.
        // The re-run of overflowed queries (fixed) appends straight to global
        // memory, whose capacity per round it never exceeds.
        if (mfma && !fixed && p0 >= 16 * (int64_t)K) {
            pt.begin(PH_MFMA, s);
            launch_hamming_mfma(ds, qs, wpad, Q, p0, p1, v, s);
            pt.end(PH_MFMA, s);
            if (mfma_launches) ++*mfma_launches;
            if (mfma_points) *mfma_points += p1 - p0;
        } else {
            SketchScanArgs a{ds, qs, wpad, Q, p0, p1, kind};
            launch_sketch_scan(a, v, s);
        }
    };
    run_rounds(st, Q, K, N, fixed, scan, s, pt, launches, K <= kMaxIntTopK);
    if (fixed) return;
    const std::vector<int32_t> ovf = flagged(st.view().ovf, Q, s);
    if (ovf.empty()) return;
    const int m = (int)ovf.size();
    const int mpad = sketch_query_rows(m);
    std::vector<int32_t> hidx(ovf);
    hidx.resize((size_t)mpad, ovf[0]);
    DevBuf<int32_t> idx((size_t)mpad);
    LSH_HIP(hipMemcpyAsync(idx.get(), hidx.data(), sizeof(int32_t) * mpad, hipMemcpyHostToDevice, s));
    DevBuf<uint32_t> sub((size_t)mpad * wpad);
    launch_gather_rows_u32(qs, wpad, idx.get(), mpad, sub.get(), s);
    TopKState sst;
    sketch_topk(ds, N, wpad, kind, sub.get(), m, K, true, sst, s, pt, launches);
    const TopKView sv = sst.view(), v = st.view();
    launch_scatter_rows_u64(sv.keys, sv.stride, K, idx.get(), m, v.keys, v.stride, s);
    LSH_HIP(hipStreamSynchronize(s));
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

size_t lsh_index::device_bytes() const {
    return data.bytes() + data_inv.bytes() + mins.bytes() + dsk.bytes() + planesT.bytes() + djl.bytes() +
           wmh_owner.bytes() + wmh_start.bytes() + wmh_bitmap.bytes() + q.bytes() + qT.bytes() +
           q_inv.bytes() + qjl.bytes() + qjlT.bytes() + qsk.bytes() + out.bytes() + flags.bytes() +
           st.bytes();
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

    gj_transform = (framework == 0 && impl == 2 && dfunc == 2) || impl == 4 || impl == 5 ||
                   (framework == 1 && (keysImpl == 4 || keysImpl == 5));
    normalize = impl == 6 || (framework == 1 && keysImpl == 7);
    if (gj_transform) {
        // processingUtils.cuh:12-46 then :48-67 (memOptimizedLinearScan.cu:149-163,
        // weightedMinHash.cuh:201-213).  The shift is frozen at build time; the
        // one-shot entry points include the queries, as the reference does.
        mins.alloc((size_t)d);
        launch_fill_i32(mins.get(), (size_t)d, 0, s);
        launch_gj_mins(data.get(), N, d, mins.get(), s);
        if (extra.get()) launch_gj_mins(extra.get(), n_extra, d, mins.get(), s);
        launch_gj_shift(data.get(), N, d, mins.get(), s);
        launch_normalize_rows(data.get(), N, d, s);
        if (extra.get()) {
            launch_gj_shift(extra.get(), n_extra, d, mins.get(), s);
            launch_normalize_rows(extra.get(), n_extra, d, s);
        }
    } else if (normalize) {
        // simHashJL.cuh:28-32 / crossPolyTope.cuh:17-21.
        launch_normalize_rows(data.get(), N, d, s);
    }

    const int kind = impl == 1 ? kNegCos : (impl == 2 ? dfunc : (impl == 4 || impl == 5 ? kGenJaccard : kAngular));
    if (angular_kind(kind)) {
        data_inv.alloc((size_t)N);
        launch_row_inv_norms(data.get(), N, d, data_inv.get(), s);
    }

    if (framework == 0 && impl == 3) {
        // simHash.cu:61-99: 32*sketchDim hyperplanes of the reference stream.
        words = sketchDim;
        wpad = round_up(std::max(words, 1), 8);
        std::vector<float> planes = normal_stream((int64_t)words * 32 * d, 1.0);
        std::vector<float> pt = transpose_planes(planes, words, 32, d);
        planesT.alloc(pt.size());
        LSH_HIP(hipMemcpyAsync(planesT.get(), pt.data(), sizeof(float) * pt.size(), hipMemcpyHostToDevice, s));
        dsk.alloc((size_t)N * wpad);
        if (N > 0) LSH_HIP(hipMemsetAsync(dsk.get(), 0, dsk.bytes(), s));
        launch_simhash(data.get(), N, d, planesT.get(), words, 32, dsk.get(), wpad, s);
        LSH_HIP(hipStreamSynchronize(s));
    } else if (framework == 0 && impl == 6) {
        // simHashJL.cuh:238-262: planes from N(0, 1/sketchDim).
        sdim = sketchDim;
        std::vector<float> planes = normal_stream((int64_t)sdim * d, 1.0 / sdim);
        std::vector<float> pt = transpose_jl(planes, sdim, d);
        planesT.alloc(pt.size());
        LSH_HIP(hipMemcpyAsync(planesT.get(), pt.data(), sizeof(float) * pt.size(), hipMemcpyHostToDevice, s));
        djl.alloc((size_t)N * sdim);
        launch_jl(data.get(), N, d, planesT.get(), sdim, djl.get(), s);
        LSH_HIP(hipStreamSynchronize(s));
    } else if (framework == 0 && (impl == 4 || impl == 5)) {
        // weightedMinHash.cuh:344-416: bounds, index map, sketches (the bounds
        // include the one-shot queries, as the reference computes them).
        wmh_mode = impl == 5 ? 1 : 0;
        words = sketchDim;
        wpad = round_up(std::max(words, 1), 8);
        build_wmh_tables(data.get(), N, extra.get(), n_extra, s);
        dsk.alloc((size_t)N * wpad);
        if (N > 0) LSH_HIP(hipMemsetAsync(dsk.get(), 0, dsk.bytes(), s));
        wmh_sketch_rows(data.get(), N, dsk.get(), s);
        LSH_HIP(hipStreamSynchronize(s));
    } else if (framework == 1) {
        fail(LSH_ERR_UNSUPPORTED, "the (K,L) framework is not built yet");
    }
    LSH_HIP(hipStreamSynchronize(s));
}

void lsh_index::build_wmh_tables(const float* x, int64_t n, const float* extra, int64_t n_extra, hipStream_t s) {
    DevBuf<int> b((size_t)d);
    launch_fill_i32(b.get(), (size_t)d, 0, s);
    launch_wmh_bounds(x, n, d, b.get(), s);
    if (extra && n_extra > 0) launch_wmh_bounds(extra, n_extra, d, b.get(), s);
    std::vector<int> hb((size_t)d);
    LSH_HIP(hipMemcpyAsync(hb.data(), b.get(), sizeof(int) * (size_t)d, hipMemcpyDeviceToHost, s));
    LSH_HIP(hipStreamSynchronize(s));
    // weightedMinHash.cuh:62-90: M = sum of bounds, start = exclusive prefix
    // sum, owner[r] = the dimension whose interval holds r.
    std::vector<int> start((size_t)d);
    int64_t M = 0;
    for (int j = 0; j < d; ++j) {
        start[j] = (int)M;
        M += hb[j];
    }
    LSH_REQUIRE(M < (int64_t)1 << 30, LSH_ERR_UNSUPPORTED, "weighted MinHash bounds sum too large");
    wmh_M = (int)M;
    std::vector<int> owner((size_t)std::max<int64_t>(M, 1), 0);
    for (int j = 0; j < d; ++j)
        for (int r = start[j]; r < start[j] + hb[j]; ++r) owner[r] = j;
    wmh_owner.alloc(owner.size());
    wmh_start.alloc((size_t)d);
    LSH_HIP(hipMemcpyAsync(wmh_owner.get(), owner.data(), sizeof(int) * owner.size(), hipMemcpyHostToDevice, s));
    LSH_HIP(hipMemcpyAsync(wmh_start.get(), start.data(), sizeof(int) * (size_t)d, hipMemcpyHostToDevice, s));
    // weightedMinHash.cuh:231-250 generates 255 bits of the default-seeded
    // stream; the table continues that stream so counts past 254 stay defined.
    std::vector<uint8_t> bm = bit_stream(4096);
    wmh_bitmap_len = (int)bm.size();
    wmh_bitmap.alloc(bm.size());
    LSH_HIP(hipMemcpyAsync(wmh_bitmap.get(), bm.data(), bm.size(), hipMemcpyHostToDevice, s));
    const int nhash = wmh_mode == 1 ? 32 * words : words;
    wmh_R = 256;
    wmh_draws.alloc((size_t)nhash * wmh_R);
    wmh_draw_owner.alloc((size_t)nhash * wmh_R);
    launch_wmh_draws(nhash, wmh_R, wmh_M, wmh_owner.get(), wmh_draws.get(), wmh_draw_owner.get(), s);
    LSH_HIP(hipStreamSynchronize(s));
}

void lsh_index::wmh_sketch_rows(const float* x, int64_t n, uint32_t* out, hipStream_t s) {
    const int nhash = wmh_mode == 1 ? 32 * words : words;
    launch_wmh_sketch(x, n, d, wmh_owner.get(), wmh_start.get(), wmh_M, nhash, wmh_mode, wmh_bitmap.get(),
                      wmh_bitmap_len, wmh_draws.get(), wmh_draw_owner.get(), wmh_R, out, wpad, s);
}

void lsh_index::search(const float* queries, int Q, int k, bool q_on_device, lsh_point_t* results,
                       bool r_on_device, hipStream_t user_stream) {
    DeviceGuard g(device);
    LSH_REQUIRE(Q >= 0, LSH_ERR_INVALID, "N_query must be >= 0");
    LSH_REQUIRE(k >= 1, LSH_ERR_INVALID, "k must be >= 1");
    const int kq = calculate_k(k);
    hipStream_t s = user_stream ? user_stream : stream;
    PhaseTimer pt(profiling);
    scan_launches = 0;
    mfma_launches = 0;
    mfma_points = 0;
    if (Q == 0) return;
    pt.begin(PH_TOTAL, s);
    pt.begin(PH_PREP, s);
    q.ensure((size_t)Q * d);
    LSH_HIP(hipMemcpyAsync(q.get(), queries, sizeof(float) * (size_t)Q * d,
                           q_on_device ? hipMemcpyDeviceToDevice : hipMemcpyHostToDevice, s));
    if (gj_transform) {
        launch_gj_shift(q.get(), Q, d, mins.get(), s);
        launch_normalize_rows(q.get(), Q, d, s);
    } else if (normalize) {
        launch_normalize_rows(q.get(), Q, d, s);
    }
    pt.end(PH_PREP, s);
    out.ensure((size_t)Q * kq);
    flags.ensure((size_t)Q);

    if (framework == 0 && (impl == 1 || impl == 2)) {
        const int kind = impl == 1 ? kNegCos : dfunc;
        LSH_REQUIRE(kq <= kMaxTopK, LSH_ERR_UNSUPPORTED, "exact scan supports k <= 4096");
        int K = std::min(kMaxTopK, pow2_ceil(kq + margin));
        const double eps = fp32_error_bound(kind, d);
        exact_topk(data.get(), data_inv.get(), N, d, kind, ExactQueries{q.get(), Q}, K, false, st, s, pt,
                   scan_launches);
        pt.begin(PH_FINAL, s);
        launch_exact_finalize(data.get(), q.get(), d, kind, st.view(), kq, eps, out.get(), flags.get(), s);
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
            launch_gather_rows_f32(q.get(), d, idx.get(), m, sub.get(), s);
            TopKState sst;
            exact_topk(data.get(), data_inv.get(), N, d, kind, ExactQueries{sub.get(), m}, kMaxTopK, false, sst,
                       s, pt, scan_launches);
            DevBuf<lsh_point_t> sout((size_t)m * kq);
            DevBuf<uint32_t> sflags((size_t)m);
            launch_exact_finalize(data.get(), sub.get(), d, kind, sst.view(), kq, eps, sout.get(), sflags.get(), s);
            LSH_REQUIRE(flagged(sflags.get(), m, s).empty(), LSH_ERR_UNSUPPORTED,
                        "more than 4096 points within the FP32 error band of the k-th distance");
            launch_scatter_rows_pts(sout.get(), kq, idx.get(), m, out.get(), s);
            LSH_HIP(hipStreamSynchronize(s));
        }
    } else if (framework == 0 && (impl == 3 || impl == 5 || impl == 4)) {
        const int Qpad = sketch_query_rows(Q);
        qsk.ensure((size_t)Qpad * wpad);
        pt.begin(PH_PREP, s);
        LSH_HIP(hipMemsetAsync(qsk.get(), 0, sizeof(uint32_t) * (size_t)Qpad * wpad, s));
        if (impl == 3) launch_simhash(q.get(), Q, d, planesT.get(), words, 32, qsk.get(), wpad, s);
        else wmh_sketch_rows(q.get(), Q, qsk.get(), s);
        pt.end(PH_PREP, s);
        const int Ceff = (int)std::max<int64_t>(1, std::min<int64_t>(candidates, N));
        const int K = Ceff <= kMaxIntTopK ? Ceff : pow2_ceil(Ceff);
        LSH_REQUIRE(K <= kMaxTopK, LSH_ERR_UNSUPPORTED, "candidate sets larger than 4096 are not supported");
        sketch_topk(dsk.get(), N, wpad, impl == 4 ? 1 : 0, qsk.get(), Q, K, false, st, s, pt, scan_launches,
                    &mfma_launches, &mfma_points);
        const int kind = impl == 3 ? kAngular : kGenJaccard;
        pt.begin(PH_FINAL, s);
        launch_rerank(data.get(), q.get(), d, kind, st.view().keys, st.view().stride, nullptr, Ceff, Q, kq,
                      kq + margin, fp32_error_bound(kind, d), out.get(), s);
        pt.end(PH_FINAL, s);
    } else if (framework == 0 && impl == 6) {
        qjl.ensure((size_t)Q * sdim);
        pt.begin(PH_PREP, s);
        launch_jl(q.get(), Q, d, planesT.get(), sdim, qjl.get(), s);
        pt.end(PH_PREP, s);
        const int Ceff = (int)std::min<int64_t>(candidates, N);
        const int K = pow2_ceil(std::max(Ceff, 1));
        LSH_REQUIRE(K <= kMaxTopK, LSH_ERR_UNSUPPORTED, "candidate sets larger than 4096 are not supported");
        exact_topk(djl.get(), nullptr, N, sdim, kSqL2, ExactQueries{qjl.get(), Q}, K, false, st, s, pt,
                   scan_launches);
        pt.begin(PH_FINAL, s);
        launch_rerank(data.get(), q.get(), d, kAngular, st.view().keys, st.view().stride, nullptr, Ceff, Q, kq,
                      kq + margin, fp32_error_bound(kAngular, d), out.get(), s);
        pt.end(PH_FINAL, s);
    } else {
        fail(LSH_ERR_UNSUPPORTED, "this implementation / framework combination is not built yet");
    }

    LSH_HIP(hipMemcpyAsync(results, out.get(), sizeof(lsh_point_t) * (size_t)Q * kq,
                           r_on_device ? hipMemcpyDeviceToDevice : hipMemcpyDeviceToHost, s));
    pt.end(PH_TOTAL, s);
    LSH_HIP(hipStreamSynchronize(s));
    pt.collect(phase_ms, 6);
    phase_ms[6] = (double)mfma_launches;
    phase_ms[7] = (double)mfma_points;
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
    launch_simhash(dx.get(), n, d, dp.get(), words, hashBits, dout.get(), words, nullptr);
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
    sketch_topk(dpad.get(), N, wpad, 0, qpad.get(), (int)Q, K, false, st, nullptr, pt, launches);
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
    lsh_index ix;
    ix.d = d;
    ix.N = n;
    ix.device = device;
    ix.wmh_mode = one_bit ? 1 : 0;
    ix.words = sketchDim;
    ix.wpad = sketchDim;
    DevBuf<float> dx((size_t)n * d);
    DevBuf<uint32_t> dout((size_t)n * sketchDim);
    LSH_HIP(hipMemcpy(dx.get(), x, sizeof(float) * (size_t)n * d, hipMemcpyHostToDevice));
    ix.build_wmh_tables(dx.get(), n, nullptr, 0, nullptr);
    ix.wmh_sketch_rows(dx.get(), n, dout.get(), nullptr);
    LSH_HIP(hipDeviceSynchronize());
    LSH_HIP(hipMemcpy(out, dout.get(), sizeof(uint32_t) * (size_t)n * sketchDim, hipMemcpyDeviceToHost));
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
