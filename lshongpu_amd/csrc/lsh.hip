// lsh.hip -- (K,L) bucket tables and the scan of a query's candidate set
// (lshFramework.cuh:102-280).
//
// Build: for every table a counting sort of the point IDs by bucket key.  The
// reference counts with atomics, scans the counts serially per table and
// scatters with one warp per table (lshFramework.cuh:102-140); here the count
// runs over (point, table) pairs with per-workgroup LDS histograms when the
// table is small, one workgroup per table computes the exclusive scan, and the
// scatter reserves slots with atomics.  off[t][key] is the absolute start of
// the bucket in perm, as the reference's hashKeys array holds it
// (lshFramework.cuh:116-124), with one sentinel entry per table so that the
// end of the last bucket needs no special case (:183-184).
//
// Query: a query's candidates are the points of its L buckets.  They form a
// virtual list of positions 0 .. S-1 (bucket of table 0, then table 1, ...);
// a scan round covers positions [p0, p1) of every query.  A point that shares
// the query's key in several tables appears several times in that list; it is
// scored only at its first table (a table t' < t with equal keys means it was
// already seen), which visits the union once without any shared state.  The
// reference instead scores every occurrence and drops adjacent duplicates
// after the final sort (lshFramework.cuh:255-280).
#include <type_traits>

#include "device_common.hpp"

namespace lshgpu {

namespace {

constexpr int kTPB = 256;
constexpr int kItems = 4;                 // positions per thread per launch tile
constexpr int kTile = kTPB * kItems;      // positions per workgroup
constexpr uint32_t kLdsHistMax = 16384;   // table sizes counted in LDS

unsigned grid_for(size_t n, int tpb) {
    size_t b = (n + tpb - 1) / tpb;
    return (unsigned)std::min<size_t>(std::max<size_t>(b, 1), 65535u * 16u);
}

__global__ void mask_keys_kernel(uint32_t* __restrict__ k, size_t n, uint32_t mask) {
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x)
        k[i] &= mask;
}

__global__ void widen_keys_kernel(const uint16_t* __restrict__ in, uint32_t* __restrict__ out, size_t n) {
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x)
        out[i] = in[i];
}

// Counts of small tables: grid.y = table, each workgroup a slice of points.
__global__ __launch_bounds__(kTPB) void bucket_count_lds_kernel(const uint32_t* __restrict__ keys, int64_t N,
                                                                int L, uint32_t T1, int64_t per_block,
                                                                uint32_t* __restrict__ cnt) {
    extern __shared__ uint32_t hist[];
    const int t = blockIdx.y;
    for (uint32_t j = threadIdx.x; j < T1; j += kTPB) hist[j] = 0;
    __syncthreads();
    const int64_t i0 = (int64_t)blockIdx.x * per_block;
    const int64_t i1 = min(N, i0 + per_block);
    for (int64_t i = i0 + threadIdx.x; i < i1; i += kTPB) {
        const uint32_t key = keys[(size_t)i * L + t];
        if (key < T1 - 1) atomicAdd(&hist[key], 1u);  // keys are below T by construction
    }
    __syncthreads();
    uint32_t* row = cnt + (size_t)t * T1;
    for (uint32_t j = threadIdx.x; j < T1; j += kTPB)
        if (hist[j]) atomicAdd(&row[j], hist[j]);
}

__global__ void bucket_count_kernel(const uint32_t* __restrict__ keys, int64_t N, int L, uint32_t T1,
                                    uint32_t* __restrict__ cnt) {
    const int64_t total = N * L;
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < total;
         i += (int64_t)gridDim.x * blockDim.x) {
        const int t = (int)(i % L);
        const uint32_t key = keys[i];
        if (key < T1 - 1) atomicAdd(&cnt[(size_t)t * T1 + key], 1u);
    }
}

// One workgroup of 1024 per table: in-place exclusive scan of T1 counts
// (the last is zero) plus the table's base t*N.
__global__ __launch_bounds__(1024) void bucket_offsets_kernel(uint32_t* __restrict__ a, int64_t N, uint32_t T1) {
    __shared__ uint32_t wsum[16];
    __shared__ uint32_t total_s;
    uint32_t* row = a + (size_t)blockIdx.x * T1;
    uint32_t carry = (uint32_t)((int64_t)blockIdx.x * N);
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    for (uint32_t c0 = 0; c0 < T1; c0 += 4096) {
        uint32_t v[4];
        uint32_t sum = 0;
#pragma unroll
        for (int e = 0; e < 4; ++e) {
            const uint32_t j = c0 + threadIdx.x * 4 + e;
            v[e] = j < T1 ? row[j] : 0u;
            sum += v[e];
        }
        uint32_t inc = sum;
#pragma unroll
        for (int o = 1; o < 64; o <<= 1) {
            const uint32_t t = __shfl_up(inc, o);
            if (lane >= o) inc += t;
        }
        if (lane == 63) wsum[w] = inc;
        __syncthreads();
        if (w == 0) {
            const uint32_t x = lane < 16 ? wsum[lane] : 0u;
            uint32_t xi = x;
#pragma unroll
            for (int o = 1; o < 16; o <<= 1) {
                const uint32_t t = __shfl_up(xi, o);
                if (lane >= o) xi += t;
            }
            if (lane < 16) wsum[lane] = xi - x;
            if (lane == 15) total_s = xi;
        }
        __syncthreads();
        uint32_t base = carry + wsum[w] + inc - sum;
#pragma unroll
        for (int e = 0; e < 4; ++e) {
            const uint32_t j = c0 + threadIdx.x * 4 + e;
            if (j < T1) row[j] = base;
            base += v[e];
        }
        carry += total_s;
        __syncthreads();
    }
}

__global__ void bucket_scatter_kernel(const uint32_t* __restrict__ keys, int64_t N, int L, uint32_t T1,
                                      uint32_t* __restrict__ cursor, uint32_t* __restrict__ perm) {
    const int64_t total = N * L;
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < total;
         i += (int64_t)gridDim.x * blockDim.x) {
        const int t = (int)(i % L);
        const uint32_t key = keys[i];
        if (key >= T1 - 1) continue;
        const uint32_t pos = atomicAdd(&cursor[(size_t)t * T1 + key], 1u);
        perm[pos] = (uint32_t)(i / L);
    }
}

// ------------------------------------------------ sorted tables (radix sort)
// One pass of a stable LSD radix sort over 8-bit digits, all tables at once
// (grid.y = table).  Tiles of kSortTile elements: the histogram kernel counts
// each tile's digits into counts[t][digit][tile]; bucket_offsets_kernel turns
// each table's row of counts (digit-major, so digit d of tile i follows every
// smaller digit and digit d of every earlier tile) into absolute positions;
// the scatter kernel ranks each element among the equal digits before it in
// its tile, by 64-lane match masks (eight ballots) per wave and per-wave
// counts in LDS, in element order, so the sort is stable.
constexpr int kSortItems = 8;
constexpr int kSortTile = kTPB * kSortItems;

__global__ __launch_bounds__(kTPB) void radix_hist_kernel(const uint32_t* __restrict__ keys, int64_t N, int shift,
                                                          int ntiles, uint32_t* __restrict__ counts) {
    __shared__ uint32_t hist[256];
    const int t = blockIdx.y;
    hist[threadIdx.x] = 0;
    __syncthreads();
    const uint32_t* __restrict__ k = keys + (size_t)t * N;
    const int64_t base = (int64_t)blockIdx.x * kSortTile;
#pragma unroll
    for (int it = 0; it < kSortItems; ++it) {
        const int64_t e = base + it * kTPB + threadIdx.x;
        if (e < N) atomicAdd(&hist[(k[e] >> shift) & 255u], 1u);
    }
    __syncthreads();
    counts[((size_t)t * 256 + threadIdx.x) * ntiles + blockIdx.x] = hist[threadIdx.x];
}

__global__ __launch_bounds__(kTPB) void radix_scatter_kernel(const uint32_t* __restrict__ keys,
                                                             const uint32_t* __restrict__ ids, int64_t N, int shift,
                                                             int ntiles, const uint32_t* __restrict__ counts,
                                                             uint32_t* __restrict__ okeys,
                                                             uint32_t* __restrict__ oids) {
    __shared__ uint32_t gofs[256];
    __shared__ uint32_t run[256];
    __shared__ uint32_t wcnt[kTPB / 64][256];
    const int t = blockIdx.y;
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    gofs[threadIdx.x] = counts[((size_t)t * 256 + threadIdx.x) * ntiles + blockIdx.x];
    run[threadIdx.x] = 0;
    const uint32_t* __restrict__ k = keys + (size_t)t * N;
    const uint32_t* __restrict__ id = ids ? ids + (size_t)t * N : nullptr;
    const int64_t base = (int64_t)blockIdx.x * kSortTile;
    const uint64_t lt = (1ull << lane) - 1ull;
    for (int it = 0; it < kSortItems; ++it) {
#pragma unroll
        for (int ww = 0; ww < kTPB / 64; ++ww) wcnt[ww][threadIdx.x] = 0;
        const int64_t e = base + it * kTPB + threadIdx.x;
        const bool valid = e < N;
        const uint32_t key = valid ? k[e] : 0u;
        const uint32_t pid = valid ? (id ? id[e] : (uint32_t)e) : 0u;
        const uint32_t dig = (key >> shift) & 255u;
        uint64_t m = __ballot(valid);
#pragma unroll
        for (int bit = 0; bit < 8; ++bit) {
            const bool on = (dig >> bit) & 1u;
            const uint64_t bb = __ballot(on);
            m &= on ? bb : ~bb;
        }
        const uint32_t rank = (uint32_t)__popcll(m & lt);
        __syncthreads();  // wcnt cleared; the previous round's reads of run are done
        if (valid && rank == 0) wcnt[w][dig] = (uint32_t)__popcll(m);
        __syncthreads();
        {
            uint32_t r = run[threadIdx.x];
#pragma unroll
            for (int ww = 0; ww < kTPB / 64; ++ww) {
                const uint32_t c = wcnt[ww][threadIdx.x];
                wcnt[ww][threadIdx.x] = r;
                r += c;
            }
            run[threadIdx.x] = r;
        }
        __syncthreads();
        if (valid) {
            const uint32_t dst = gofs[dig] + wcnt[w][dig] + rank;
            okeys[dst] = key;
            oids[dst] = pid;
        }
        __syncthreads();  // wcnt is cleared again next round
    }
}

__global__ void transpose_keys_kernel(const uint32_t* __restrict__ nl, int64_t n, int L, uint32_t* __restrict__ ln) {
    const size_t total = (size_t)n * L;
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < total; i += (size_t)gridDim.x * blockDim.x) {
        const size_t t = i / (size_t)n, r = i - t * (size_t)n;
        ln[i] = nl[r * L + t];
    }
}

// First position of a value >= key in the ascending array a[0, n).
__device__ inline uint32_t lower_bound_u32(const uint32_t* __restrict__ a, uint32_t n, uint32_t key) {
    uint32_t lo = 0, len = n;
    while (len > 0) {
        const uint32_t half = len >> 1;
        if (a[lo + half] < key) {
            lo += half + 1;
            len -= half + 1;
        } else {
            len = half;
        }
    }
    return lo;
}

// One thread per (query, table): the run of the query's key in the table.
__global__ void sorted_ranges_kernel(const uint32_t* __restrict__ qkeys, int Q, int L, int64_t N,
                                     const uint32_t* __restrict__ skeys, uint32_t* __restrict__ lo,
                                     uint32_t* __restrict__ len) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= (int64_t)Q * L) return;
    const int t = (int)(i % L);
    const uint32_t key = qkeys[i];
    const uint32_t* __restrict__ a = skeys + (size_t)t * N;
    const uint32_t b = lower_bound_u32(a, (uint32_t)N, key);
    const uint32_t e = key == 0xFFFFFFFFu ? (uint32_t)N : lower_bound_u32(a, (uint32_t)N, key + 1u);
    lo[i] = (uint32_t)((int64_t)t * N) + b;
    len[i] = e - b;
}

// Prefix sums of the run lengths per query.
__global__ void ranges_prefix_kernel(const uint32_t* __restrict__ len, int Q, int L, uint32_t* __restrict__ pre,
                                     uint32_t* __restrict__ maxS) {
    const int q = blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= Q) return;
    uint32_t acc = 0;
    pre[(size_t)q * (L + 1)] = 0;
    for (int t = 0; t < L; ++t) {
        acc += len[(size_t)q * L + t];
        pre[(size_t)q * (L + 1) + t + 1] = acc;
    }
    atomicMax(maxS, acc);
}

__global__ void query_ranges_kernel(const uint32_t* __restrict__ qkeys, int Q, int L, uint32_t T,
                                    const uint32_t* __restrict__ off, uint32_t* __restrict__ lo,
                                    uint32_t* __restrict__ pre, uint32_t* __restrict__ maxS) {
    const int q = blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= Q) return;
    uint32_t acc = 0;
    pre[(size_t)q * (L + 1)] = 0;
    for (int t = 0; t < L; ++t) {
        const uint32_t key = qkeys[(size_t)q * L + t];
        const uint32_t* o = off + (size_t)t * (T + 1);
        const uint32_t a = key < T ? o[key] : o[T];
        const uint32_t b = key < T ? o[key + 1] : o[T];
        lo[(size_t)q * L + t] = a;
        acc += b - a;
        pre[(size_t)q * (L + 1) + t + 1] = acc;
    }
    atomicMax(maxS, acc);
}

// Position of the n-th (from 0) set bit of m; m has more than n set bits.
__device__ inline int nth_set_bit(uint64_t m, int n) {
    int pos = 0;
#pragma unroll
    for (int w = 32; w > 0; w >>= 1) {
        const uint64_t low = m & ((1ull << w) - 1ull);
        const int c = __popcll(low);
        if (n >= c) {
            n -= c;
            m >>= w;
            pos += w;
        } else {
            m = low;
        }
    }
    return pos;
}

// MODE 0: exact distance (kind), 16-lane groups per candidate row.
// MODE 1: integer sketch distance (kind 0 Hamming, 1 unequal words).
// MODE 2: squared L2 over JL sketches.
// Whether point id's key equals the query's at a table before t (the point
// is then scored at that earlier table, so the union is visited once).  The
// point's keys are one row of L words; with L a multiple of 4 (at most 32)
// its first t keys are read as 16-byte loads issued together, so the test
// waits for one memory round trip instead of one per table.
__device__ inline bool seen_earlier(const uint32_t* __restrict__ dkeys, uint32_t id, int L, int t,
                                    const uint32_t* s_qk) {
    const uint32_t* __restrict__ dk = dkeys + (size_t)id * L;
    if ((L & 3) == 0 && L <= 32) {
        const uint4* __restrict__ d4 = reinterpret_cast<const uint4*>(dk);
        const int n4 = (t + 3) >> 2;  // <= L / 4
        uint4 kv[8];
#pragma unroll
        for (int c = 0; c < 8; ++c)
            if (c < n4) kv[c] = d4[c];
        bool hit = false;
#pragma unroll
        for (int c = 0; c < 8; ++c)
            if (c < n4) {
                const int b = 4 * c;
                hit |= (b + 0 < t && kv[c].x == s_qk[b + 0]);
                hit |= (b + 1 < t && kv[c].y == s_qk[b + 1]);
                hit |= (b + 2 < t && kv[c].z == s_qk[b + 2]);
                hit |= (b + 3 < t && kv[c].w == s_qk[b + 3]);
            }
        return hit;
    }
    for (int t2 = 0; t2 < t; ++t2)
        if (dk[t2] == s_qk[t2]) return true;
    return false;
}

// Exact distances of a wave's owned candidates (mask m over the lanes, each
// lane's candidate in id; n = popcount(m)): 16-lane groups take one
// candidate each and read its row with coalesced float4 loads.  For rows of
// at most 128 floats (a multiple of 4) the loads of four candidates per group
// are issued before any is used, so a round of 16 candidates costs one memory
// round trip.  Every lane calls emit(leader, cid, dist) for every slot
// (leader: the slot holds a candidate and this is its group's first lane), so
// emit may hold wave-wide operations.  The per-candidate arithmetic is the
// same in both paths (components in index order per lane, then the 16-lane
// tree), so the distances are too.
template <class Emit>
__device__ inline void score_owned(const LshScanArgs& a, const float* s_qf, float qn, uint64_t m, uint32_t id,
                                   int n, Emit&& emit) {
    const int lane = threadIdx.x & 63;
    const int g = lane >> 4, gl = lane & 15;
    if ((a.d & 3) == 0 && a.d <= 128) {
        constexpr int U = 4;
        const int j0 = gl * 4, j1 = gl * 4 + 64;
        const bool in0 = j0 < a.d, in1 = j1 < a.d;
        for (int it0 = 0; it0 * 4 < n; it0 += U) {
            float4 xa[U], xb[U];
            uint32_t cids[U];
            bool hs[U];
#pragma unroll
            for (int u = 0; u < U; ++u) {
                const int j = (it0 + u) * 4 + g;
                hs[u] = j < n;
                const int src = hs[u] ? nth_set_bit(m, j) : 0;
                cids[u] = (uint32_t)__shfl((int)id, src);
                const float* __restrict__ row = a.data + (size_t)cids[u] * a.d;
                xa[u] = hs[u] && in0 ? *reinterpret_cast<const float4*>(row + j0) : make_float4(0.f, 0.f, 0.f, 0.f);
                xb[u] = hs[u] && in1 ? *reinterpret_cast<const float4*>(row + j1) : make_float4(0.f, 0.f, 0.f, 0.f);
            }
#pragma unroll
            for (int u = 0; u < U; ++u) {
                float x = 0.0f, y = 0.0f;
                if (hs[u]) {
                    if (in0) {
                        group_accumulate(a.kind, xa[u].x, s_qf[j0], x, y);
                        group_accumulate(a.kind, xa[u].y, s_qf[j0 + 1], x, y);
                        group_accumulate(a.kind, xa[u].z, s_qf[j0 + 2], x, y);
                        group_accumulate(a.kind, xa[u].w, s_qf[j0 + 3], x, y);
                    }
                    if (in1) {
                        group_accumulate(a.kind, xb[u].x, s_qf[j1], x, y);
                        group_accumulate(a.kind, xb[u].y, s_qf[j1 + 1], x, y);
                        group_accumulate(a.kind, xb[u].z, s_qf[j1 + 2], x, y);
                        group_accumulate(a.kind, xb[u].w, s_qf[j1 + 3], x, y);
                    }
                }
#pragma unroll
                for (int o = 8; o > 0; o >>= 1) {
                    x += __shfl_xor(x, o, 16);
                    y += __shfl_xor(y, o, 16);
                }
                emit(hs[u] && gl == 0, cids[u], group_distance(a.kind, x, y, qn));
            }
        }
        return;
    }
    for (int it = 0; it * 4 < n; ++it) {
        const int j = it * 4 + g;
        const bool has = j < n;
        const int src = has ? nth_set_bit(m, j) : 0;
        const uint32_t cid = (uint32_t)__shfl((int)id, src);
        float x = 0.0f, y = 0.0f;
        if (has) {
            const float* __restrict__ row = a.data + (size_t)cid * a.d;
            if ((a.d & 3) == 0) {
                for (int jj = gl * 4; jj < a.d; jj += 64) {
                    const float4 xv = *reinterpret_cast<const float4*>(row + jj);
                    group_accumulate(a.kind, xv.x, s_qf[jj], x, y);
                    group_accumulate(a.kind, xv.y, s_qf[jj + 1], x, y);
                    group_accumulate(a.kind, xv.z, s_qf[jj + 2], x, y);
                    group_accumulate(a.kind, xv.w, s_qf[jj + 3], x, y);
                }
            } else {
                for (int jj = gl; jj < a.d; jj += 16) group_accumulate(a.kind, row[jj], s_qf[jj], x, y);
            }
        }
#pragma unroll
        for (int o = 8; o > 0; o >>= 1) {
            x += __shfl_xor(x, o, 16);
            y += __shfl_xor(y, o, 16);
        }
        emit(has && gl == 0, cid, group_distance(a.kind, x, y, qn));
    }
}

template <int MODE>
__global__ __launch_bounds__(kTPB) void lsh_scan_kernel(LshScanArgs a, TopKView v) {
    extern __shared__ uint32_t lds[];
    const int L = a.L;
    uint32_t* s_pre = lds;              // L+1
    uint32_t* s_lo = lds + (L + 1);     // L
    uint32_t* s_qk = s_lo + L;          // L
    float* s_qf = reinterpret_cast<float*>(s_qk + L);  // query vector (d / sdim floats)
    uint32_t* s_qw = s_qk + L;                          // query sketch (words)

    for (int r = blockIdx.y; r < a.Q; r += gridDim.y) {
        const int q = a.qmap ? a.qmap[r] : r;
        const uint32_t* pre = a.qpre + (size_t)q * (L + 1);
        const int64_t S = pre[L];
        const int64_t start = a.p0 + (int64_t)blockIdx.x * kTile;
        const int64_t end = min(min(a.p1, S), start + (int64_t)kTile);
        if (start >= end) continue;  // uniform over the workgroup
        __syncthreads();             // previous row's LDS reads are done
        for (int t = threadIdx.x; t <= L; t += kTPB) s_pre[t] = pre[t];
        for (int t = threadIdx.x; t < L; t += kTPB) {
            s_lo[t] = a.qlo[(size_t)q * L + t];
            s_qk[t] = a.qkeys[(size_t)q * L + t];
        }
        if (MODE == 0) {
            for (int j = threadIdx.x; j < a.d; j += kTPB) s_qf[j] = a.queries[(size_t)q * a.d + j];
        } else if (MODE == 1) {
            for (int j = threadIdx.x; j < a.words; j += kTPB) s_qw[j] = a.qsk[(size_t)q * a.words + j];
        } else {
            for (int j = threadIdx.x; j < a.sdim; j += kTPB) s_qf[j] = a.qjl[(size_t)q * a.sdim + j];
        }
        __syncthreads();
        const uint64_t thr = v.thr[r];
        float qn = 0.0f;
        if (MODE == 0 && (a.kind == kAngular || a.kind == kNegCos)) {
            for (int j = 0; j < a.d; ++j) qn = __builtin_fmaf(s_qf[j], s_qf[j], qn);
            qn = __builtin_sqrtf(qn);
        }
        for (int64_t base = start; base < end; base += kTPB) {
            const int64_t pos = base + threadIdx.x;
            bool owned = false;
            uint32_t id = 0;
            if (pos < end) {
                int lo = 0, hi = L;  // pre[lo] <= pos < pre[hi]
                while (hi - lo > 1) {
                    const int mid = (lo + hi) >> 1;
                    if ((int64_t)s_pre[mid] <= pos) lo = mid;
                    else hi = mid;
                }
                id = a.perm[s_lo[lo] + (uint32_t)(pos - s_pre[lo])];
                owned = !seen_earlier(a.dkeys, id, L, lo, s_qk);
            }
            if (MODE == 0) {
                const uint64_t m = __ballot(owned);
                score_owned(a, s_qf, qn, m, id, __popcll(m), [&](bool leader, uint32_t cid, float dist) {
                    const uint64_t key = make_key(f32_key(dist), cid);
                    append_key(v, r, leader && key < thr, key);
                });
            } else if (MODE == 1) {
                uint32_t dist = 0;
                if (owned) {
                    const uint32_t* __restrict__ row = a.dsk + (size_t)id * a.words;
                    if (a.kind == 0) {
                        for (int w = 0; w < a.words; ++w) dist += __popc(row[w] ^ s_qw[w]);
                    } else {
                        for (int w = 0; w < a.words; ++w) dist += row[w] != s_qw[w] ? 1u : 0u;
                    }
                }
                const uint64_t key = make_key(dist, id);
                append_key(v, r, owned && key < thr, key);
            } else {
                float dist = 0.0f;
                if (owned) {
                    const float* __restrict__ row = a.djl + (size_t)id * a.sdim;
                    for (int j = 0; j < a.sdim; ++j) {
                        // distanceFunctions.cuh:87-99: fl(t*t), then a separate add.
                        const float t = s_qf[j] - row[j];
                        const float sq = t * t;
                        dist = dist + sq;
                    }
                }
                const uint64_t key = make_key(f32_key(dist), id);
                append_key(v, r, owned && key < thr, key);
            }
        }
    }
}

// ------------------------------------------ exact rescan as a work list
// The exact (K,L) rescan without rounds: every query's list of S positions
// is cut into items of kItemPos positions; one launch turns the per-query
// item counts into item offsets (ibase), a persistent grid takes items from
// that list (a grid-stride loop, so the grid needs no host-side count) and
// keeps each item's KP best (distance, ID) keys, and one workgroup per query
// merges its items' lists into the query's sorted best KP.  Nothing is read
// back by the host; a query whose items do not fit the item capacity is
// flagged (mflags) for a re-run.
constexpr int kItemPos = 2048;

__global__ __launch_bounds__(1024) void lsh_items_kernel(const uint32_t* __restrict__ pre, int Q, int L,
                                                         uint32_t* __restrict__ ibase) {
    __shared__ uint32_t wsum[16];
    __shared__ uint32_t total_s;
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    uint32_t carry = 0;
    for (int c0 = 0; c0 < Q; c0 += 1024) {
        const int q = c0 + threadIdx.x;
        const uint32_t S = q < Q ? pre[(size_t)q * (L + 1) + L] : 0u;
        const uint32_t n = (S + kItemPos - 1) / kItemPos;
        uint32_t inc = n;
#pragma unroll
        for (int o = 1; o < 64; o <<= 1) {
            const uint32_t t = __shfl_up(inc, o);
            if (lane >= o) inc += t;
        }
        if (lane == 63) wsum[w] = inc;
        __syncthreads();
        if (w == 0) {
            const uint32_t x = lane < 16 ? wsum[lane] : 0u;
            uint32_t xi = x;
#pragma unroll
            for (int o = 1; o < 16; o <<= 1) {
                const uint32_t t = __shfl_up(xi, o);
                if (lane >= o) xi += t;
            }
            if (lane < 16) wsum[lane] = xi - x;
            if (lane == 15) total_s = xi;
        }
        __syncthreads();
        if (q < Q) ibase[q] = carry + wsum[w] + inc - n;
        carry += total_s;
        __syncthreads();
    }
    if (threadIdx.x == 0) ibase[Q] = carry;
}

template <int KP>
__global__ __launch_bounds__(kTPB) void lsh_partial_kernel(LshScanArgs a, const uint32_t* __restrict__ ibase,
                                                           uint32_t cap, uint64_t* __restrict__ partials) {
    extern __shared__ uint32_t lds[];
    const int L = a.L;
    uint64_t* s_keys = reinterpret_cast<uint64_t*>(lds);  // kItemPos
    uint32_t* s_pre = lds + 2 * kItemPos;                  // L+1
    uint32_t* s_lo = s_pre + (L + 1);                      // L
    uint32_t* s_qk = s_lo + L;                             // L
    float* s_qf = reinterpret_cast<float*>(s_qk + L);      // d
    __shared__ uint32_t s_n;
    const uint32_t total = min(ibase[a.Q], cap);
    for (uint32_t item = blockIdx.x; item < total; item += gridDim.x) {
        int lo = 0, hi = a.Q;  // ibase[lo] <= item < ibase[hi]
        while (hi - lo > 1) {
            const int mid = (lo + hi) >> 1;
            if (ibase[mid] <= item) lo = mid;
            else hi = mid;
        }
        const int q = lo;
        const uint32_t* pre = a.qpre + (size_t)q * (L + 1);
        const int64_t S = pre[L];
        const int64_t c0 = (int64_t)(item - ibase[q]) * kItemPos;
        const int64_t c1 = min<int64_t>(S, c0 + kItemPos);
        __syncthreads();  // the previous item's LDS reads are done
        for (int t = threadIdx.x; t <= L; t += kTPB) s_pre[t] = pre[t];
        for (int t = threadIdx.x; t < L; t += kTPB) {
            s_lo[t] = a.qlo[(size_t)q * L + t];
            s_qk[t] = a.qkeys[(size_t)q * L + t];
        }
        for (int j = threadIdx.x; j < a.d; j += kTPB) s_qf[j] = a.queries[(size_t)q * a.d + j];
        if (threadIdx.x == 0) s_n = 0;
        __syncthreads();
        float qn = 0.0f;
        if (a.kind == kAngular || a.kind == kNegCos) {
            for (int j = 0; j < a.d; ++j) qn = __builtin_fmaf(s_qf[j], s_qf[j], qn);
            qn = __builtin_sqrtf(qn);
        }
        for (int64_t base = c0; base < c1; base += kTPB) {
            const int64_t pos = base + threadIdx.x;
            bool owned = false;
            uint32_t id = 0;
            if (pos < c1) {
                int tl = 0, th = L;  // pre[tl] <= pos < pre[th]
                while (th - tl > 1) {
                    const int mid = (tl + th) >> 1;
                    if ((int64_t)s_pre[mid] <= pos) tl = mid;
                    else th = mid;
                }
                id = a.perm[s_lo[tl] + (uint32_t)(pos - s_pre[tl])];
                owned = !seen_earlier(a.dkeys, id, L, tl, s_qk);
            }
            const uint64_t m = __ballot(owned);
            score_owned(a, s_qf, qn, m, id, __popcll(m), [&](bool leader, uint32_t cid, float dist) {
                if (leader) s_keys[atomicAdd(&s_n, 1u)] = make_key(f32_key(dist), cid);
            });
        }
        __syncthreads();
        // The item's best KP: all of them when there are at most KP, else
        // the first KP of the sorted list.
        const int n = (int)s_n;
        int np = KP;
        while (np < n) np <<= 1;
        for (int t = n + threadIdx.x; t < np; t += kTPB) s_keys[t] = kEmptyKey;
        if (n > KP) bitonic_sort_lds(s_keys, np);
        else __syncthreads();
        uint64_t* out = partials + (size_t)item * KP;
        for (int t = threadIdx.x; t < KP; t += kTPB) out[t] = s_keys[t];
    }
}

template <int KP>
__global__ __launch_bounds__(kTPB) void lsh_merge_kernel(const uint32_t* __restrict__ ibase, uint32_t cap,
                                                         const uint64_t* __restrict__ partials, TopKView v,
                                                         uint32_t* __restrict__ mflags) {
    constexpr int B = 4096;  // LDS buffer: the best KP so far, then the next keys
    __shared__ uint64_t buf[B];
    __shared__ uint32_t s_n;
    const int q = blockIdx.x;
    const uint32_t i0 = ibase[q], i1 = ibase[q + 1];
    const uint32_t e1 = min(i1, cap);
    for (int t = threadIdx.x; t < KP; t += kTPB) buf[t] = kEmptyKey;
    if (threadIdx.x == 0) s_n = KP;
    __syncthreads();
    uint64_t thr = kEmptyKey;  // the current KP-th best: larger keys are dropped
    const size_t k0 = (size_t)i0 * KP, k1 = e1 > i0 ? (size_t)e1 * KP : k0;
    for (size_t c = k0; c < k1; c += B - KP) {
        const size_t ce = min(k1, c + (size_t)(B - KP));
        for (size_t t = c + threadIdx.x; t < ce; t += kTPB) {
            const uint64_t key = partials[t];
            if (key < thr) buf[atomicAdd(&s_n, 1u)] = key;
        }
        __syncthreads();
        const int n = (int)s_n;
        int np = KP;
        while (np < n) np <<= 1;
        for (int t = n + threadIdx.x; t < np; t += kTPB) buf[t] = kEmptyKey;
        bitonic_sort_lds(buf, np);
        thr = buf[KP - 1];
        if (threadIdx.x == 0) s_n = KP;
        __syncthreads();
    }
    uint64_t* row = v.keys + (size_t)q * v.stride;
    for (int t = threadIdx.x; t < v.K; t += kTPB) row[t] = t < KP ? buf[t] : kEmptyKey;
    if (threadIdx.x == 0) {
        v.thr[q] = thr;
        mflags[q] = i1 > cap ? 1u : 0u;
    }
}

}  // namespace

void launch_mask_keys(uint32_t* keys, size_t n, uint32_t mask, hipStream_t s) {
    if (n == 0) return;
    hipLaunchKernelGGL(mask_keys_kernel, dim3(grid_for(n, 256)), dim3(256), 0, s, keys, n, mask);
    check_launch("mask_keys_kernel");
}

void launch_widen_keys(const uint16_t* in, uint32_t* out, size_t n, hipStream_t s) {
    if (n == 0) return;
    hipLaunchKernelGGL(widen_keys_kernel, dim3(grid_for(n, 256)), dim3(256), 0, s, in, out, n);
    check_launch("widen_keys_kernel");
}

void launch_bucket_build(const uint32_t* keys, int64_t N, int L, uint32_t T, uint32_t* off, uint32_t* cursor,
                         uint32_t* perm, hipStream_t s) {
    const uint32_t T1 = T + 1;
    const size_t nrow = (size_t)L * T1;
    launch_fill_i32(reinterpret_cast<int32_t*>(off), nrow, 0, s);
    if (N > 0) {
        if (T1 <= kLdsHistMax) {
            const int64_t per = std::max<int64_t>(4096, (N + 1023) / 1024);
            const unsigned gx = (unsigned)((N + per - 1) / per);
            const size_t lds = sizeof(uint32_t) * T1;
            set_dynamic_lds((const void*)bucket_count_lds_kernel, lds);
            hipLaunchKernelGGL(bucket_count_lds_kernel, dim3(gx, (unsigned)L), dim3(kTPB), lds, s, keys, N, L, T1,
                               per, off);
            check_launch("bucket_count_lds_kernel");
        } else {
            hipLaunchKernelGGL(bucket_count_kernel, dim3(grid_for((size_t)N * L, 256)), dim3(256), 0, s, keys, N,
                               L, T1, off);
            check_launch("bucket_count_kernel");
        }
    }
    hipLaunchKernelGGL(bucket_offsets_kernel, dim3((unsigned)L), dim3(1024), 0, s, off, N, T1);
    check_launch("bucket_offsets_kernel");
    if (N > 0) {
        LSH_HIP(hipMemcpyAsync(cursor, off, sizeof(uint32_t) * nrow, hipMemcpyDeviceToDevice, s));
        hipLaunchKernelGGL(bucket_scatter_kernel, dim3(grid_for((size_t)N * L, 256)), dim3(256), 0, s, keys, N, L,
                           T1, cursor, perm);
        check_launch("bucket_scatter_kernel");
    }
}

void launch_query_ranges(const uint32_t* qkeys, int Q, int L, uint32_t T, const uint32_t* off, uint32_t* lo,
                         uint32_t* pre, uint32_t* maxS, hipStream_t s) {
    if (Q <= 0) return;
    hipLaunchKernelGGL(query_ranges_kernel, dim3((unsigned)((Q + 255) / 256)), dim3(256), 0, s, qkeys, Q, L, T,
                       off, lo, pre, maxS);
    check_launch("query_ranges_kernel");
}

size_t sort_table_counts(int64_t N, int L) {
    const int64_t ntiles = std::max<int64_t>(1, (N + kSortTile - 1) / kSortTile);
    return (size_t)L * 256 * ntiles;
}

void launch_sort_tables(const uint32_t* keys_ln, int64_t N, int L, int bits, uint32_t* skeys, uint32_t* perm,
                        uint32_t* tk, uint32_t* ti, uint32_t* counts, hipStream_t s) {
    if (N <= 0 || L <= 0) return;
    LSH_REQUIRE(bits >= 1 && bits <= 32, LSH_ERR_INTERNAL, "sort: key bits must be 1..32");
    LSH_REQUIRE((int64_t)L * N < 0xFFFFFFFFll, LSH_ERR_UNSUPPORTED, "tables * N_data must be below 2^32");
    const int ntiles = (int)((N + kSortTile - 1) / kSortTile);
    LSH_REQUIRE(ntiles <= 65535 * 64, LSH_ERR_UNSUPPORTED, "too many points for the table sort");
    const int passes = (bits + 7) / 8;
    const uint32_t T1 = 256u * (uint32_t)ntiles;
    for (int p = 0; p < passes; ++p) {
        // The last pass writes skeys/perm; earlier ones alternate so that it does.
        const bool to_out = ((passes - 1 - p) & 1) == 0;
        const uint32_t* ik = p == 0 ? keys_ln : (to_out ? tk : skeys);
        const uint32_t* ii = p == 0 ? nullptr : (to_out ? ti : perm);
        uint32_t* ok = to_out ? skeys : tk;
        uint32_t* oi = to_out ? perm : ti;
        const dim3 grid((unsigned)ntiles, (unsigned)L);
        hipLaunchKernelGGL(radix_hist_kernel, grid, dim3(kTPB), 0, s, ik, N, 8 * p, ntiles, counts);
        check_launch("radix_hist_kernel");
        // Exclusive scan of each table's counts plus the table base t*N.
        hipLaunchKernelGGL(bucket_offsets_kernel, dim3((unsigned)L), dim3(1024), 0, s, counts, N, T1);
        check_launch("bucket_offsets_kernel");
        hipLaunchKernelGGL(radix_scatter_kernel, grid, dim3(kTPB), 0, s, ik, ii, N, 8 * p, ntiles, counts, ok, oi);
        check_launch("radix_scatter_kernel");
    }
}

void launch_query_ranges_sorted(const uint32_t* qkeys, int Q, int L, int64_t N, const uint32_t* skeys,
                                uint32_t* lo, uint32_t* len, uint32_t* pre, uint32_t* maxS, hipStream_t s) {
    if (Q <= 0) return;
    const size_t n = (size_t)Q * L;
    hipLaunchKernelGGL(sorted_ranges_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, qkeys, Q, L, N,
                       skeys, lo, len);
    check_launch("sorted_ranges_kernel");
    hipLaunchKernelGGL(ranges_prefix_kernel, dim3((unsigned)((Q + 255) / 256)), dim3(256), 0, s, len, Q, L, pre,
                       maxS);
    check_launch("ranges_prefix_kernel");
}

void launch_transpose_keys(const uint32_t* nl, int64_t n, int L, uint32_t* ln, hipStream_t s) {
    if (n <= 0) return;
    hipLaunchKernelGGL(transpose_keys_kernel, dim3(grid_for((size_t)n * L, 256)), dim3(256), 0, s, nl, n, L, ln);
    check_launch("transpose_keys_kernel");
}

int lsh_item_positions() { return kItemPos; }

void launch_lsh_exact_worklist(const LshScanArgs& a, int KP, uint32_t* ibase, uint32_t cap, uint64_t* partials,
                               const TopKView& v, uint32_t* mflags, hipStream_t s) {
    if (a.Q <= 0) return;
    LSH_REQUIRE(!a.sketch && a.qmap == nullptr && v.K >= KP, LSH_ERR_INTERNAL, "work-list rescan: bad arguments");
    hipLaunchKernelGGL(lsh_items_kernel, dim3(1), dim3(1024), 0, s, a.qpre, a.Q, a.L, ibase);
    check_launch("lsh_items_kernel");
    const size_t lds = sizeof(uint64_t) * kItemPos + sizeof(uint32_t) * (3 * (size_t)a.L + 1) + sizeof(float) * a.d;
    // A persistent grid: 8 workgroups per CU of the 256.
    const unsigned grid = 2048;
    auto go = [&](auto kp) {
        constexpr int K_ = decltype(kp)::value;
        set_dynamic_lds((const void*)lsh_partial_kernel<K_>, lds);
        hipLaunchKernelGGL(lsh_partial_kernel<K_>, dim3(grid), dim3(kTPB), lds, s, a, ibase, cap, partials);
        check_launch("lsh_partial_kernel");
        hipLaunchKernelGGL(lsh_merge_kernel<K_>, dim3((unsigned)a.Q), dim3(kTPB), 0, s, ibase, cap, partials, v,
                           mflags);
        check_launch("lsh_merge_kernel");
    };
    switch (KP) {
        case 32: go(std::integral_constant<int, 32>()); break;
        case 64: go(std::integral_constant<int, 64>()); break;
        case 128: go(std::integral_constant<int, 128>()); break;
        case 256: go(std::integral_constant<int, 256>()); break;
        case 512: go(std::integral_constant<int, 512>()); break;
        default: fail(LSH_ERR_INTERNAL, "work-list rescan: KP must be a power of two in 32..512");
    }
}

void launch_lsh_scan(const LshScanArgs& a, const TopKView& v, hipStream_t s) {
    if (a.p1 <= a.p0 || a.Q <= 0) return;
    const int64_t tiles = (a.p1 - a.p0 + kTile - 1) / kTile;
    LSH_REQUIRE(tiles < (int64_t)1 << 31, LSH_ERR_UNSUPPORTED, "candidate range too large");
    const dim3 grid((unsigned)tiles, (unsigned)std::min(a.Q, 65535));
    const int mode = !a.sketch ? 0 : (a.kind == kSqL2 ? 2 : 1);
    const int qlen = mode == 0 ? a.d : (mode == 1 ? a.words : a.sdim);
    const size_t lds = sizeof(uint32_t) * ((size_t)3 * a.L + 1 + (size_t)qlen);
    switch (mode) {
        case 0:
            set_dynamic_lds((const void*)lsh_scan_kernel<0>, lds);
            hipLaunchKernelGGL(lsh_scan_kernel<0>, grid, dim3(kTPB), lds, s, a, v);
            break;
        case 1:
            set_dynamic_lds((const void*)lsh_scan_kernel<1>, lds);
            hipLaunchKernelGGL(lsh_scan_kernel<1>, grid, dim3(kTPB), lds, s, a, v);
            break;
        default:
            set_dynamic_lds((const void*)lsh_scan_kernel<2>, lds);
            hipLaunchKernelGGL(lsh_scan_kernel<2>, grid, dim3(kTPB), lds, s, a, v);
            break;
    }
    check_launch("lsh_scan_kernel");
}

}  // namespace lshgpu
