// select.hip -- per-query top-K maintenance between scan rounds.
//
// A scan round appends every candidate whose key beats the query's current
// threshold (the K-th best key so far).  This kernel folds those appends into
// the sorted best-K list: the new keys are bitonic-sorted in LDS, the K
// smallest of (best list, new keys) are taken as the element-wise minimum of
// the best list and the reversed sorted new keys (a bitonic sequence), and a
// half-cleaner network sorts them.  This replaces the reference's per-lane
// insertion sort plus warp bitonic network (sortingFramework.cuh:163-254)
// with exact (distance, ID) keys instead of its EPSILON comparator
// (pointExtensions.cuh:8-33).
#include "device_common.hpp"
#include "select_int.hpp"

namespace lshgpu {

namespace {

constexpr int kSelThreads = 256;

__global__ __launch_bounds__(kSelThreads) void select_merge_kernel(TopKView v) {
    extern __shared__ uint64_t sm[];
    const int q = blockIdx.x;
    const uint32_t n_raw = v.cnt[q];
    if (n_raw == 0) return;
    const int K = v.K;
    const int n = (int)min(n_raw, (uint32_t)v.newcap);
    int P2 = K;
    while (P2 < n) P2 <<= 1;
    uint64_t* A = sm;
    uint64_t* B = sm + K;
    uint64_t* row = v.keys + (size_t)q * v.stride;
    for (int i = threadIdx.x; i < P2; i += blockDim.x) B[i] = i < n ? row[K + i] : kEmptyKey;
    for (int i = threadIdx.x; i < K; i += blockDim.x) A[i] = row[i];
    bitonic_sort_lds(B, P2);
    for (int i = threadIdx.x; i < K; i += blockDim.x) {
        uint64_t a = A[i], b = B[K - 1 - i];
        A[i] = a < b ? a : b;
    }
    for (int stride = K >> 1; stride > 0; stride >>= 1) {
        __syncthreads();
        for (int t = threadIdx.x; t < (K >> 1); t += blockDim.x) {
            int i = 2 * t - (t & (stride - 1));
            cmpx_u64(A, i, i + stride, true);
        }
    }
    __syncthreads();
    for (int i = threadIdx.x; i < K; i += blockDim.x) row[i] = A[i];
    if (threadIdx.x == 0) {
        v.thr[q] = A[K - 1];
        v.cnt[q] = 0;
        if (n_raw > (uint32_t)v.newcap) v.ovf[q] = 1;
    }
}

// Integer-key variant (sketch distances are small integers): the best list
// is kept as an unordered set of exactly K keys (K need not be a power of
// two).  A histogram of the distances of (best list + new keys) in LDS gives
// the bin T that holds the K-th key; a radix select over the keys of that
// bin (8-bit digits of the ID, or of the whole key in the last bin, which
// also holds every larger distance) finds the K-th key itself, and every key
// at or below it is kept.  thr becomes that key.
//
// RK > 0: every thread loads its keys (at most RK, i = thread + 256 r) into
// registers once, and every pass reads them there; the host picks it when
// the view can hold at most 256 RK keys per query.
template <int RK>
struct KeysInRegisters {
    uint64_t key[RK];
    __device__ KeysInRegisters(const uint64_t* in, int n) {
#pragma unroll
        for (int r = 0; r < RK; ++r) {
            const int i = (int)threadIdx.x + kSelThreads * r;
            key[r] = i < n ? in[i] : kEmptyKey;
        }
    }
    template <class F>
    __device__ void operator()(F&& f) const {
#pragma unroll
        for (int r = 0; r < RK; ++r)
            if (key[r] != kEmptyKey) f(key[r]);
    }
};

//
// flags (may be null, may be pinned host memory): flags[q] = 1 when query q's
// list is incomplete after the select (overflowed or fewer than K keys),
// flags[Q + i] = extra[i] for i < n_extra.
// Wave 0 of the workgroup: the bin of hist[0, 256) at which the running
// count reaches rank (1-based; rank <= the total), and rank within it, to
// out[0] and out[1].
__device__ inline void wave_bin_of_rank(const uint32_t* hist, uint32_t rank, uint32_t* out) {
    if (threadIdx.x >= 64) return;
    const int lane = threadIdx.x;
    uint32_t c[4], sum = 0;
#pragma unroll
    for (int e = 0; e < 4; ++e) {
        c[e] = hist[4 * lane + e];
        sum += c[e];
    }
    uint32_t inc = sum;
#pragma unroll
    for (int o = 1; o < 64; o <<= 1) {
        const uint32_t x = __shfl_up(inc, o);
        if (lane >= o) inc += x;
    }
    uint32_t run = inc - sum;
#pragma unroll
    for (int e = 0; e < 4; ++e) {
        if (run < rank && run + c[e] >= rank) {
            out[0] = 4u * lane + e;
            out[1] = rank - run;
        }
        run += c[e];
    }
}

// The rerank filter's work over the nkeep fused keys in kept[] (LDS) of
// query q: the k-th smallest 16-bit distance q_k (two 8-bit histogram
// rounds; qh[0, 256) already counts the kept keys' high bytes, qh[256, 512)
// is zero), then the band, every key with q <= q_k + delta, as (row << 32 |
// 0), the IDs left to the recheck.  sv: 4 words of LDS.  Every thread must
// call it, after a barrier that follows the counts.
__device__ __forceinline__ void fused_band(const uint64_t* kept, uint32_t nkeep, int q, const FusedBand& fb,
                           uint32_t* qh, uint32_t* sv) {
    const int tid = threadIdx.x, bs = blockDim.x;
    auto qof = [](uint64_t key) { return (uint32_t)(key >> 32) & 0xFFFFu; };
    const bool all = fb.qbad[q] != 0u || nkeep < (uint32_t)fb.k;
    uint32_t qub = 0xFFFFFFFFu;
    if (!all) {
        wave_bin_of_rank(qh, (uint32_t)fb.k, sv);
        __syncthreads();
        const uint32_t hb = sv[0], rank = sv[1];
        for (uint32_t i = tid; i < nkeep; i += bs) {
            const uint32_t x = qof(kept[i]);
            if ((x >> 8) == hb) atomicAdd(&qh[256 + (x & 255u)], 1u);
        }
        __syncthreads();
        wave_bin_of_rank(qh + 256, rank, sv + 2);
        __syncthreads();
        qub = ((hb << 8) | sv[2]) + (uint32_t)fb.delta;
    }
    if (tid == 0) sv[3] = 0;
    __syncthreads();
    uint64_t* __restrict__ bq = fb.band + (size_t)q * fb.stride;
    for (uint32_t i = tid; i < nkeep; i += bs) {
        const uint64_t key = kept[i];
        if (qof(key) <= qub) bq[atomicAdd(&sv[3], 1u)] = (uint64_t)(uint32_t)key << 32;
    }
    __syncthreads();
    if (tid == 0) fb.bcnt[q] = sv[3];
}

template <int RK>
__global__ __launch_bounds__(kSelThreads) void select_hist_kernel(TopKView v, uint32_t* flags,
                                                                  const uint32_t* extra, int n_extra,
                                                                  const int32_t* __restrict__ row_ids,
                                                                  FusedBand fb) {
    extern __shared__ uint64_t kept[];  // K
    __shared__ uint32_t hist[kSelBins];
    __shared__ uint32_t qh[512];        // fused band: histograms of the kept keys' 16-bit distances
    __shared__ uint64_t ties[kTieCap];
    __shared__ uint32_t sv[kSelScratch];
    __shared__ uint32_t s_nkeep;
    const int q = blockIdx.x;
    if (flags && blockIdx.x == 0 && (int)threadIdx.x < n_extra) flags[v.Q + threadIdx.x] = extra[threadIdx.x];
    const uint32_t n_raw = v.cnt[q];
    if (n_raw == 0 && !v.fresh) {
        if (flags && threadIdx.x == 0) flags[q] = (v.ovf[q] != 0u || v.thr[q] == kEmptyKey) ? 1u : 0u;
        return;
    }
    const int K = v.K;
    const int n = (int)min(n_raw, (uint32_t)v.newcap);
    uint64_t* row = v.keys + (size_t)q * v.stride;
    // A fresh view has no stored best list: only the appends are read (the
    // K written below are disjoint from them).
    const uint64_t* in = v.fresh ? row + K : row;
    const int total = v.fresh ? n : K + n;
    if (threadIdx.x == 0) s_nkeep = 0;
    // Cleared here, counted in the keep loop: the K-th search's barriers
    // order the two.
    if (fb.band)
        for (int i = threadIdx.x; i < 512; i += blockDim.x) qh[i] = 0;
    auto run = [&](const auto& each) {
        const uint64_t kth = int_kth_key(each, (uint32_t)K, hist, ties, sv);  // the K-th key (none: keep all)
        each([&](uint64_t key) {
            if (key != kEmptyKey && key <= kth) {
                const uint32_t at = atomicAdd(&s_nkeep, 1u);
                if (at < (uint32_t)K) kept[at] = key;
            }
        });
        return kth;
    };
    // Row keys (fresh views only, held in registers): each key's row is
    // looked up in the layout's IDs (a query's rows come from a few cells, so
    // the lookups are local) and the key becomes (distance, ID); each kept
    // key's row is stored at its slot of the best list, as int32 from the
    // list's end (the appends were all read before the k-th search's
    // barriers, so that space is free by then).
    int32_t* rws = reinterpret_cast<int32_t*>(row + K);
    uint64_t kth;
    if constexpr (RK > 0) {
        if (v.row_keys && v.fused_keys) {
            // Fused keys: the K-th by (distance, ID) with the IDs looked up
            // for the boundary bin's keys alone (the others are kept or
            // dropped by distance, whatever their low word); the kept keys
            // stay (distance << 16 | q, row), and their IDs are looked up
            // later only where needed.
            KeysInRegisters<RK> kr(in, total);
            uint32_t fq[(RK + 1) / 2] = {};  // the 16-bit q of each key, two a word
#pragma unroll
            for (int r = 0; r < RK; ++r)
                if (kr.key[r] != kEmptyKey) {
                    const uint32_t hi = (uint32_t)(kr.key[r] >> 32);
                    fq[r >> 1] |= (hi & 0xFFFFu) << (16 * (r & 1));
                    kr.key[r] = make_key(hi >> 16, (uint32_t)kr.key[r]);
                }
            // The boundary bin's keys get their IDs (rows kept aside), all
            // lookups issued together.
            uint32_t brow[RK];
            const auto to_ids = [&](uint32_t T) {
#pragma unroll
                for (int r = 0; r < RK; ++r) {
                    const uint64_t key = kr.key[r];
                    const bool b = key != kEmptyKey && min((uint32_t)(key >> 32), (uint32_t)(kSelBins - 1)) == T;
                    brow[r] = (uint32_t)key;
                    if (b) kr.key[r] = (key & 0xFFFFFFFF00000000ull) | (uint32_t)row_ids[(uint32_t)key];
                }
            };
            kth = int_kth_key(kr, (uint32_t)K, hist, ties, sv, to_ids);
#pragma unroll
            for (int r = 0; r < RK; ++r) {
                const uint64_t key = kr.key[r];
                if (key != kEmptyKey && key <= kth) {
                    const uint32_t at = atomicAdd(&s_nkeep, 1u);
                    const uint32_t fqr = (fq[r >> 1] >> (16 * (r & 1))) & 0xFFFFu;
                    if (at < (uint32_t)K) {
                        kept[at] = make_key(((uint32_t)(key >> 32) << 16) | fqr, brow[r]);
                        if (fb.band) atomicAdd(&qh[fqr >> 8], 1u);
                    }
                }
            }
        } else if (v.row_keys) {
            KeysInRegisters<RK> kr(in, total);
            int32_t rw[RK];
#pragma unroll
            for (int r = 0; r < RK; ++r) rw[r] = kr.key[r] != kEmptyKey ? (int32_t)(uint32_t)kr.key[r] : -1;
#pragma unroll
            for (int r = 0; r < RK; ++r)
                if (rw[r] >= 0) kr.key[r] = (kr.key[r] & 0xFFFFFFFF00000000ull) | (uint32_t)row_ids[rw[r]];
            kth = int_kth_key(kr, (uint32_t)K, hist, ties, sv);
#pragma unroll
            for (int r = 0; r < RK; ++r) {
                const uint64_t key = kr.key[r];
                if (key != kEmptyKey && key <= kth) {
                    const uint32_t at = atomicAdd(&s_nkeep, 1u);
                    if (at < (uint32_t)K) {
                        kept[at] = key;
                        rws[at] = rw[r];
                    }
                }
            }
        } else {
            kth = run(KeysInRegisters<RK>(in, total));
        }
    } else {
        kth = run(KeysInMemory{in, total});
    }
    __syncthreads();
    const uint32_t nkeep = min(s_nkeep, (uint32_t)K);
    for (uint32_t i = threadIdx.x; i < nkeep; i += blockDim.x) row[i] = kept[i];
    for (uint32_t i = nkeep + threadIdx.x; i < (uint32_t)K; i += blockDim.x) row[i] = kEmptyKey;
    if (RK > 0 && v.row_keys)
        for (uint32_t i = nkeep + threadIdx.x; i < (uint32_t)K; i += blockDim.x) rws[i] = -1;
    if (threadIdx.x == 0) {
        v.thr[q] = nkeep == (uint32_t)K ? kth : kEmptyKey;
        v.cnt[q] = 0;
        if (n_raw > (uint32_t)v.newcap) v.ovf[q] = 1;
        if (flags) flags[q] = (v.ovf[q] != 0u || nkeep != (uint32_t)K) ? 1u : 0u;
    }
    // sv is free (int_kth_key ended with a barrier); the barrier after the
    // keep loop ordered the counts in qh.
    if (fb.band) fused_band(kept, nkeep, q, fb, qh, sv);
}

// Sorts the first K keys of every row ascending (K <= 4096).
__global__ __launch_bounds__(kSelThreads) void sort_rows_kernel(uint64_t* keys, int stride, int K, int P) {
    extern __shared__ uint64_t sm[];
    uint64_t* row = keys + (size_t)blockIdx.x * stride;
    for (int i = threadIdx.x; i < P; i += blockDim.x) sm[i] = i < K ? row[i] : kEmptyKey;
    bitonic_sort_lds(sm, P);
    for (int i = threadIdx.x; i < K; i += blockDim.x) row[i] = sm[i];
}

// Reset of a top-K state: the first K keys of every row (the best lists,
// unless the view is fresh) and the threshold empty, the counts and overflow
// flags zero.
__global__ void reset_state_kernel(TopKView v) {
    const size_t total = v.fresh ? (size_t)v.Q : (size_t)v.Q * v.K;
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < total; i += (size_t)gridDim.x * blockDim.x) {
        if (!v.fresh) {
            const size_t q = i / v.K;
            v.keys[q * v.stride + (i - q * v.K)] = kEmptyKey;
        }
        if (i < (size_t)v.Q) {
            v.thr[i] = kEmptyKey;
            v.cnt[i] = 0u;
            v.ovf[i] = 0u;
        }
    }
}

__global__ void fill_u64(uint64_t* p, size_t n, uint64_t val) {
    size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    size_t st = (size_t)gridDim.x * blockDim.x;
    for (; i < n; i += st) p[i] = val;
}

__global__ void fill_u32(uint32_t* p, size_t n, uint32_t val) {
    size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    size_t st = (size_t)gridDim.x * blockDim.x;
    for (; i < n; i += st) p[i] = val;
}

__global__ void count_flags_kernel(const uint32_t* f, int n, uint32_t* out) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    uint32_t c = (i < n && f[i]) ? 1u : 0u;
    unsigned long long m = __ballot(c);
    if ((threadIdx.x & 63) == 0 && m) atomicAdd(out, (uint32_t)__popcll(m));
}

__global__ void threshold_above_kernel(TopKView src, TopKView dst) {
    const int q = blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= dst.Q) return;
    const uint64_t t = src.thr[q];
    dst.thr[q] = t == kEmptyKey ? kEmptyKey : make_key((uint32_t)(t >> 32) + 1u, 0u);
}

// flags[q] = 1 when query q's list is incomplete; flags[Q + i] = extra[i].
// flags may be pinned host memory (the host reads it after an event).
__global__ void incomplete_flags_kernel(TopKView v, uint32_t* flags, const uint32_t* extra, int n_extra) {
    const int q = blockIdx.x * blockDim.x + threadIdx.x;
    if (q < v.Q) flags[q] = (v.ovf[q] != 0u || v.thr[q] == kEmptyKey) ? 1u : 0u;
    if (q < n_extra) flags[v.Q + q] = extra[q];
}

unsigned grid_for(size_t n) {
    size_t b = (n + 255) / 256;
    return (unsigned)std::min<size_t>(std::max<size_t>(b, 1), 65535 * 8);
}

}  // namespace

int pow2_ceil(int x) {
    int p = 1;
    while (p < x) p <<= 1;
    return p;
}

int newcap_for(int K) { return K <= 1024 ? 2 * K + 256 : K + 512; }

void TopKState::reset(int Q, int K, int newcap, hipStream_t s, bool int_keys, bool lazy, bool launch) {
    LSH_REQUIRE(K >= 1 && K <= kMaxTopK && (int_keys || (K & (K - 1)) == 0), LSH_ERR_INVALID,
                "top-K size must be a power of two <= 4096");
    LSH_REQUIRE(!lazy || (int_keys && K <= kMaxIntTopK), LSH_ERR_INTERNAL, "lazy reset needs the integer-key select");
    const int stride = K + newcap;
    keys_.ensure((size_t)Q * stride);
    cnt_.ensure((size_t)Q);
    thr_.ensure((size_t)Q);
    ovf_.ensure((size_t)Q);
    v_ = TopKView{keys_.get(), cnt_.get(), thr_.get(), ovf_.get(), K, newcap, stride, Q, int_keys ? 1 : 0,
                  lazy ? 1 : 0, 0, 0};
    if (launch) clear(s);
}

void TopKState::clear(hipStream_t s) { launch_topk_clear(v_, s); }

void launch_topk_clear(const TopKView& v, hipStream_t s) {
    // A fill kernel rather than memsets keeps the sequence capturable in a
    // graph.  Only the best lists need clearing: the selects read the
    // appended slots up to each query's count.
    if (v.Q <= 0) return;
    const size_t n = v.fresh ? (size_t)v.Q : (size_t)v.Q * v.K;
    hipLaunchKernelGGL(reset_state_kernel, dim3(grid_for(n)), dim3(256), 0, s, v);
    check_launch("top-K state reset");
}

// Keys per thread the register select holds.
constexpr int kSelRegKeys = 12;

bool row_keys_supported(int K, int newcap) {
    return 2 * (size_t)newcap >= (size_t)K && newcap <= kSelThreads * kSelRegKeys;
}

void launch_select_merge(const TopKView& v, hipStream_t s, uint32_t* flags, const uint32_t* extra, int n_extra,
                         const int32_t* row_ids, const FusedBand* fbp) {
    LSH_REQUIRE(n_extra >= 0 && n_extra <= kSelThreads && (n_extra == 0 || extra), LSH_ERR_INTERNAL,
                "select: bad extra words");
    LSH_REQUIRE(!v.row_keys || (v.int_keys && v.fresh && row_ids && row_keys_supported(v.K, v.newcap)),
                LSH_ERR_INTERNAL, "select: row keys need a fresh integer-key view, the layout's IDs and room");
    LSH_REQUIRE(!fbp || (v.fused_keys && v.K <= kMaxIntTopK && row_ids && fbp->band && fbp->bcnt && fbp->qbad &&
                         fbp->k >= 1 &&
                         fbp->stride >= v.K && fbp->delta >= 0),
                LSH_ERR_INTERNAL, "select: the fused band needs fused keys and its buffers");
    const FusedBand fb = fbp ? *fbp : FusedBand{};
    if (v.int_keys && v.Q > 0) {
        const size_t lds = (size_t)v.K * sizeof(uint64_t);
        const int most = v.fresh ? v.newcap : v.K + v.newcap;  // keys a query can hold
        constexpr int RK = kSelRegKeys;
        if (most <= kSelThreads * RK) {
            set_dynamic_lds((const void*)select_hist_kernel<RK>, lds + sizeof(uint32_t) * (kSelBins + 8));
            hipLaunchKernelGGL(select_hist_kernel<RK>, dim3(v.Q), dim3(kSelThreads), lds, s, v, flags, extra,
                               n_extra, row_ids, fb);
        } else {
            set_dynamic_lds((const void*)select_hist_kernel<0>, lds + sizeof(uint32_t) * (kSelBins + 8));
            LSH_REQUIRE(!fbp, LSH_ERR_INTERNAL, "select: the fused band needs the register select");
            hipLaunchKernelGGL(select_hist_kernel<0>, dim3(v.Q), dim3(kSelThreads), lds, s, v, flags, extra,
                               n_extra, row_ids, fb);
        }
        check_launch("select_hist_kernel");
        return;
    }
    if (v.Q > 0) {
        const size_t lds = (size_t)(v.K + pow2_ceil(std::max(v.newcap, v.K))) * sizeof(uint64_t);
        set_dynamic_lds((const void*)select_merge_kernel, lds);
        hipLaunchKernelGGL(select_merge_kernel, dim3(v.Q), dim3(kSelThreads), lds, s, v);
        check_launch("select_merge_kernel");
    }
    if (flags) launch_incomplete_flags(v, flags, s, extra, n_extra);
}

void launch_sort_rows(const TopKView& v, hipStream_t s) {
    const int P = pow2_ceil(std::max(v.K, 2));
    const size_t lds = (size_t)P * sizeof(uint64_t);
    set_dynamic_lds((const void*)sort_rows_kernel, lds);
    hipLaunchKernelGGL(sort_rows_kernel, dim3(v.Q), dim3(kSelThreads), lds, s, v.keys, v.stride, v.K, P);
    check_launch("sort_rows_kernel");
}

void launch_count_overflow(const TopKView& v, uint32_t* out_count, hipStream_t s) {
    hipLaunchKernelGGL(fill_u32, dim3(1), dim3(64), 0, s, out_count, (size_t)1, 0u);
    hipLaunchKernelGGL(count_flags_kernel, dim3((v.Q + 255) / 256), dim3(256), 0, s, v.ovf, v.Q,
                       out_count);
    check_launch("count_flags_kernel");
}

void launch_threshold_above(const TopKView& src, const TopKView& dst, hipStream_t s) {
    hipLaunchKernelGGL(threshold_above_kernel, dim3((dst.Q + 255) / 256), dim3(256), 0, s, src, dst);
    check_launch("threshold_above_kernel");
}

void launch_incomplete_flags(const TopKView& v, uint32_t* flags, hipStream_t s, const uint32_t* extra, int n_extra) {
    const int n = std::max(v.Q, n_extra);
    if (n <= 0) return;
    hipLaunchKernelGGL(incomplete_flags_kernel, dim3((n + 255) / 256), dim3(256), 0, s, v, flags, extra, n_extra);
    check_launch("incomplete_flags_kernel");
}

}  // namespace lshgpu
