// stats.hip -- bucket and sketch statistics on the device.
//
// statistics.cuh:10-16 bucketDistribution: a histogram of byte-valued hashes
// (weightedMinHash.cuh:215-229 runs it over the byte sketches and prints the
// non-zero bins).  Here each workgroup counts into one LDS histogram per wave
// from 16-byte loads and flushes it with one 64-bit atomic per non-empty bin;
// all 256 values are counted (the reference's 255-entry result array has no
// slot for the value 255).
//
// statisticsCpu.cpp:11-28 bucketDistributionFullKey: the number of points per
// distinct sketchDim-byte key, keyed by the concatenated 8-bit strings of the
// bytes in a std::map, so the keys come out in byte-lexicographic order.
// Here the keys are packed into big-endian 64-bit words (numeric order of the
// words is the byte order), the point IDs are sorted by (words, ID) with a
// bitonic network over the next power of two (padding sorts last), and run
// boundaries, a prefix sum over them and the run starts give the distinct
// keys and their counts.  The reference function has no return
// statement and prints every key; neither is reproduced.
//
// Bucket occupancy of a (K,L) index table: with off[t][key] the absolute start
// of each bucket, bucket b holds off[b+1] - off[b] points.  One pass over the
// T buckets builds the histogram of bucket sizes (the last bin collects the
// sizes at or above it), the number of non-empty buckets and the largest one:
// the occupancy figures that choose bucketKeyBits and tables.
#include "device_common.hpp"
#include "engine.hpp"

namespace lshgpu {

namespace {

constexpr int kTPB = 256;
constexpr int kWaves = kTPB / 64;
constexpr int kLdsBins = 8192;  // occupancy histograms up to this length are built in LDS

unsigned blocks_for(size_t n, int per_block, unsigned cap) {
    const size_t b = (n + per_block - 1) / per_block;
    return (unsigned)std::min<size_t>(std::max<size_t>(b, 1), cap);
}

__global__ __launch_bounds__(kTPB) void byte_hist_kernel(const uint8_t* __restrict__ h, int64_t n,
                                                          unsigned long long* __restrict__ counts) {
    __shared__ uint32_t hist[kWaves][256];
    for (int i = threadIdx.x; i < kWaves * 256; i += kTPB) (&hist[0][0])[i] = 0;
    __syncthreads();
    uint32_t* mine = hist[threadIdx.x / 64];
    const int64_t n16 = n / 16;
    const uint4* v = reinterpret_cast<const uint4*>(h);
    for (int64_t i = (int64_t)blockIdx.x * kTPB + threadIdx.x; i < n16; i += (int64_t)gridDim.x * kTPB) {
        const uint4 w = v[i];
        const uint32_t parts[4] = {w.x, w.y, w.z, w.w};
#pragma unroll
        for (int p = 0; p < 4; ++p)
#pragma unroll
            for (int b = 0; b < 4; ++b) atomicAdd(&mine[(parts[p] >> (8 * b)) & 0xFF], 1u);
    }
    if (blockIdx.x == 0)
        for (int64_t i = n16 * 16 + threadIdx.x; i < n; i += kTPB) atomicAdd(&mine[h[i]], 1u);
    __syncthreads();
    for (int b = threadIdx.x; b < 256; b += kTPB) {
        uint32_t c = 0;
#pragma unroll
        for (int w = 0; w < kWaves; ++w) c += hist[w][b];
        if (c) atomicAdd(&counts[b], (unsigned long long)c);
    }
}

// Word w of key i: bytes 8w .. 8w+7 of the key, byte 8w in bits 63..56;
// bytes past the key's end are zero.  Layout [W][n].
__global__ void pack_key_words_kernel(const uint8_t* __restrict__ h, int64_t n, int sd, int W,
                                      uint64_t* __restrict__ words) {
    const int64_t total = n * W;
    for (int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; t < total;
         t += (int64_t)gridDim.x * blockDim.x) {
        const int64_t i = t / W;
        const int w = (int)(t - i * W);
        const uint8_t* k = h + i * sd;
        uint64_t x = 0;
        for (int b = 0; b < 8; ++b) {
            const int j = 8 * w + b;
            x = (x << 8) | (j < sd ? k[j] : 0u);
        }
        words[(int64_t)w * n + i] = x;
    }
}

// p[i] = i below n, -1 (sorts after every key) up to m.
__global__ void iota_pad_kernel(int32_t* __restrict__ p, int64_t n, int64_t m) {
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < m; i += (int64_t)gridDim.x * blockDim.x)
        p[i] = i < n ? (int32_t)i : -1;
}

// (key words, ID) order; -1 is larger than every ID.
__device__ inline bool key_greater(const uint64_t* __restrict__ words, int64_t n, int W, int32_t a, int32_t b) {
    if (a < 0 || b < 0) return a < 0 && b >= 0;
    for (int w = 0; w < W; ++w) {
        const uint64_t x = words[(int64_t)w * n + a], y = words[(int64_t)w * n + b];
        if (x != y) return x > y;
    }
    return a > b;
}

// One step (k, j) of the bitonic network over p[0, m): thread t handles the
// pair (i, i + j) with i = 2j*(t/j) + t%j, ascending where i & k == 0.
__global__ void bitonic_step_kernel(int32_t* __restrict__ p, int64_t m, int64_t k, int64_t j,
                                    const uint64_t* __restrict__ words, int64_t n, int W) {
    for (int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; t < m / 2; t += (int64_t)gridDim.x * blockDim.x) {
        const int64_t i = 2 * j * (t / j) + t % j;
        const int32_t a = p[i], b = p[i + j];
        const bool up = (i & k) == 0;
        if (key_greater(words, n, W, a, b) == up) {
            p[i] = b;
            p[i + j] = a;
        }
    }
}

// Inclusive prefix sum of a[0, n) in place, by one workgroup of 1024 threads
// walking the array in chunks of 4096 with a running carry.
__global__ __launch_bounds__(1024) void inclusive_scan_kernel(int32_t* __restrict__ a, int64_t n) {
    __shared__ int32_t wsum[16];
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    int32_t carry = 0;
    for (int64_t c0 = 0; c0 < n; c0 += 4096) {
        int32_t v[4];
        int32_t sum = 0;
#pragma unroll
        for (int e = 0; e < 4; ++e) {
            const int64_t j = c0 + threadIdx.x * 4 + e;
            v[e] = j < n ? a[j] : 0;
            sum += v[e];
        }
        int32_t inc = sum;
#pragma unroll
        for (int o = 1; o < 64; o <<= 1) {
            const int32_t t = __shfl_up(inc, o);
            if (lane >= o) inc += t;
        }
        if (lane == 63) wsum[w] = inc;
        __syncthreads();
        int32_t before = 0, total = 0;
#pragma unroll
        for (int q = 0; q < 16; ++q) {
            before += q < w ? wsum[q] : 0;
            total += wsum[q];
        }
        int32_t run = carry + before + inc - sum;
#pragma unroll
        for (int e = 0; e < 4; ++e) {
            const int64_t j = c0 + threadIdx.x * 4 + e;
            run += v[e];
            if (j < n) a[j] = run;
        }
        carry += total;
        __syncthreads();
    }
}

// flag[i] = 1 where sorted position i starts a new key.
__global__ void run_flags_kernel(const uint64_t* __restrict__ words, const int32_t* __restrict__ perm, int64_t n,
                                 int W, int32_t* __restrict__ flag) {
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
        int f = i == 0;
        if (!f) {
            const int64_t a = perm[i], b = perm[i - 1];
            for (int w = 0; w < W && !f; ++w) f = words[(int64_t)w * n + a] != words[(int64_t)w * n + b];
        }
        flag[i] = f;
    }
}

// seg is the inclusive prefix sum of flag: run seg[i]-1 starts at position i.
__global__ void run_starts_kernel(const int32_t* __restrict__ flag, const int32_t* __restrict__ seg,
                                  const int32_t* __restrict__ perm, int64_t n, int32_t* __restrict__ start,
                                  int32_t* __restrict__ rep) {
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
        if (flag[i]) {
            start[seg[i] - 1] = (int32_t)i;
            rep[seg[i] - 1] = perm[i];
        }
}

__global__ void run_output_kernel(const uint8_t* __restrict__ h, int sd, const int32_t* __restrict__ start,
                                  const int32_t* __restrict__ rep, int64_t runs, int64_t n,
                                  uint8_t* __restrict__ keys, int64_t* __restrict__ counts) {
    for (int64_t r = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; r < runs; r += (int64_t)gridDim.x * blockDim.x) {
        counts[r] = (r + 1 < runs ? (int64_t)start[r + 1] : n) - start[r];
        const uint8_t* src = h + (int64_t)rep[r] * sd;
        for (int b = 0; b < sd; ++b) keys[r * sd + b] = src[b];
    }
}

// summary[0] += non-empty buckets, summary[1] = max(size) (as an unsigned
// 64-bit atomic max).
template <bool kLds>
__global__ __launch_bounds__(kTPB) void bucket_occupancy_kernel(const uint32_t* __restrict__ off, uint32_t T,
                                                                int hist_len, unsigned long long* __restrict__ hist,
                                                                unsigned long long* __restrict__ summary) {
    extern __shared__ uint32_t lhist[];
    if (kLds) {
        for (int i = threadIdx.x; i < hist_len; i += kTPB) lhist[i] = 0;
        __syncthreads();
    }
    uint32_t nonempty = 0, maxsz = 0;
    for (uint32_t b = blockIdx.x * kTPB + threadIdx.x; b < T; b += gridDim.x * kTPB) {
        const uint32_t sz = off[b + 1] - off[b];
        nonempty += sz != 0;
        maxsz = sz > maxsz ? sz : maxsz;
        if (hist_len > 0) {
            const uint32_t bin = sz < (uint32_t)(hist_len - 1) ? sz : (uint32_t)(hist_len - 1);
            if (kLds)
                atomicAdd(&lhist[bin], 1u);
            else
                atomicAdd(&hist[bin], 1ull);
        }
    }
    for (int o = 32; o > 0; o >>= 1) {
        nonempty += __shfl_xor(nonempty, o);
        const uint32_t m = __shfl_xor(maxsz, o);
        maxsz = m > maxsz ? m : maxsz;
    }
    if ((threadIdx.x & 63) == 0) {
        if (nonempty) atomicAdd(&summary[0], (unsigned long long)nonempty);
        atomicMax(&summary[1], (unsigned long long)maxsz);
    }
    if (kLds) {
        __syncthreads();
        for (int i = threadIdx.x; i < hist_len; i += kTPB)
            if (lhist[i]) atomicAdd(&hist[i], (unsigned long long)lhist[i]);
    }
}

// The same figures for a sorted table (keys ascending, N of them): each run
// of equal keys is one non-empty bucket, found at its first position; its
// size comes from a binary search for the run's end.
template <bool kLds>
__global__ __launch_bounds__(kTPB) void sorted_occupancy_kernel(const uint32_t* __restrict__ k, uint32_t N,
                                                                int hist_len, unsigned long long* __restrict__ hist,
                                                                unsigned long long* __restrict__ summary) {
    extern __shared__ uint32_t lhist[];
    if (kLds) {
        for (int i = threadIdx.x; i < hist_len; i += kTPB) lhist[i] = 0;
        __syncthreads();
    }
    uint32_t nonempty = 0, maxsz = 0;
    for (uint32_t i = blockIdx.x * kTPB + threadIdx.x; i < N; i += gridDim.x * kTPB) {
        const uint32_t key = k[i];
        if (i > 0 && k[i - 1] == key) continue;
        uint32_t lo = i + 1, len = N - lo;  // first position past the run
        while (len > 0) {
            const uint32_t half = len >> 1;
            if (k[lo + half] == key) {
                lo += half + 1;
                len -= half + 1;
            } else {
                len = half;
            }
        }
        const uint32_t sz = lo - i;
        ++nonempty;
        maxsz = sz > maxsz ? sz : maxsz;
        if (hist_len > 0) {
            const uint32_t bin = sz < (uint32_t)(hist_len - 1) ? sz : (uint32_t)(hist_len - 1);
            if (kLds)
                atomicAdd(&lhist[bin], 1u);
            else
                atomicAdd(&hist[bin], 1ull);
        }
    }
    for (int o = 32; o > 0; o >>= 1) {
        nonempty += __shfl_xor(nonempty, o);
        const uint32_t m = __shfl_xor(maxsz, o);
        maxsz = m > maxsz ? m : maxsz;
    }
    if ((threadIdx.x & 63) == 0) {
        if (nonempty) atomicAdd(&summary[0], (unsigned long long)nonempty);
        atomicMax(&summary[1], (unsigned long long)maxsz);
    }
    if (kLds) {
        __syncthreads();
        for (int i = threadIdx.x; i < hist_len; i += kTPB)
            if (lhist[i]) atomicAdd(&hist[i], (unsigned long long)lhist[i]);
    }
}

struct DeviceGuard {
    int prev = 0;
    explicit DeviceGuard(int dev) {
        LSH_HIP(hipGetDevice(&prev));
        if (prev != dev) LSH_HIP(hipSetDevice(dev));
    }
    ~DeviceGuard() { (void)hipSetDevice(prev); }
};

}  // namespace

void byte_distribution_host(const uint8_t* hashes, int64_t n, int64_t* counts, int device) {
    DeviceGuard g(device);
    DevBuf<uint8_t> dh((size_t)n);
    DevBuf<unsigned long long> dc(256);
    LSH_HIP(hipMemcpy(dh.get(), hashes, (size_t)n, hipMemcpyHostToDevice));
    LSH_HIP(hipMemset(dc.get(), 0, 256 * sizeof(unsigned long long)));
    hipLaunchKernelGGL(byte_hist_kernel, dim3(blocks_for((size_t)(n / 16), kTPB * 4, 2048)), dim3(kTPB), 0, nullptr,
                       dh.get(), n, dc.get());
    check_launch("byte_hist_kernel");
    LSH_HIP(hipDeviceSynchronize());
    LSH_HIP(hipMemcpy(counts, dc.get(), 256 * sizeof(int64_t), hipMemcpyDeviceToHost));
}

void full_key_distribution_host(const uint8_t* hashes, int64_t n, int sd, std::vector<uint8_t>& keys,
                                std::vector<int64_t>& counts, int device) {
    DeviceGuard g(device);
    keys.clear();
    counts.clear();
    if (n == 0) return;
    const int W = (sd + 7) / 8;
    int64_t m = 1;
    while (m < n) m *= 2;
    const int64_t nbytes = n * sd;
    DevBuf<uint8_t> dh((size_t)nbytes);
    DevBuf<uint64_t> words((size_t)n * W);
    DevBuf<int32_t> perm((size_t)m), flag((size_t)n), seg((size_t)n), start((size_t)n), rep((size_t)n);
    LSH_HIP(hipMemcpy(dh.get(), hashes, (size_t)nbytes, hipMemcpyHostToDevice));
    const unsigned cap = 65535u * 16u;
    hipLaunchKernelGGL(pack_key_words_kernel, dim3(blocks_for((size_t)n * W, kTPB, cap)), dim3(kTPB), 0, nullptr,
                       dh.get(), n, sd, W, words.get());
    check_launch("pack_key_words_kernel");
    hipLaunchKernelGGL(iota_pad_kernel, dim3(blocks_for((size_t)m, kTPB, cap)), dim3(kTPB), 0, nullptr, perm.get(),
                       n, m);
    check_launch("iota_pad_kernel");
    const unsigned gp = blocks_for((size_t)(m / 2), kTPB, cap);
    for (int64_t k = 2; k <= m; k *= 2)
        for (int64_t j = k / 2; j >= 1; j /= 2) {
            hipLaunchKernelGGL(bitonic_step_kernel, dim3(gp), dim3(kTPB), 0, nullptr, perm.get(), m, k, j,
                               words.get(), n, W);
            check_launch("bitonic_step_kernel");
        }
    const unsigned gn = blocks_for((size_t)n, kTPB, cap);
    hipLaunchKernelGGL(run_flags_kernel, dim3(gn), dim3(kTPB), 0, nullptr, words.get(), perm.get(), n, W,
                       flag.get());
    check_launch("run_flags_kernel");
    LSH_HIP(hipMemcpyAsync(seg.get(), flag.get(), sizeof(int32_t) * (size_t)n, hipMemcpyDeviceToDevice, nullptr));
    hipLaunchKernelGGL(inclusive_scan_kernel, dim3(1), dim3(1024), 0, nullptr, seg.get(), n);
    check_launch("inclusive_scan_kernel");
    hipLaunchKernelGGL(run_starts_kernel, dim3(gn), dim3(kTPB), 0, nullptr, flag.get(), seg.get(), perm.get(), n,
                       start.get(), rep.get());
    check_launch("run_starts_kernel");
    int32_t runs = 0;
    LSH_HIP(hipMemcpy(&runs, seg.get() + (n - 1), sizeof(int32_t), hipMemcpyDeviceToHost));
    LSH_REQUIRE(runs >= 1 && runs <= n, LSH_ERR_INTERNAL, "full-key distribution: bad run count");
    DevBuf<uint8_t> dkeys((size_t)runs * sd);
    DevBuf<int64_t> dcounts((size_t)runs);
    hipLaunchKernelGGL(run_output_kernel, dim3(blocks_for((size_t)runs, kTPB, cap)), dim3(kTPB), 0, nullptr,
                       dh.get(), sd, start.get(), rep.get(), (int64_t)runs, n, dkeys.get(), dcounts.get());
    check_launch("run_output_kernel");
    LSH_HIP(hipDeviceSynchronize());
    keys.resize((size_t)runs * sd);
    counts.resize((size_t)runs);
    LSH_HIP(hipMemcpy(keys.data(), dkeys.get(), keys.size(), hipMemcpyDeviceToHost));
    LSH_HIP(hipMemcpy(counts.data(), dcounts.get(), counts.size() * sizeof(int64_t), hipMemcpyDeviceToHost));
}

void index_bucket_stats(lsh_index* idx, int table, int64_t* summary4, int64_t* hist, int hist_len) {
    LSH_REQUIRE(idx->framework == 1 && (idx->boff.get() || idx->bkeys.get()), LSH_ERR_INVALID,
                "bucket statistics need an index built with framework 1");
    LSH_REQUIRE(table >= 0 && table < idx->tables, LSH_ERR_INVALID, "table out of range");
    LSH_REQUIRE(hist_len >= 0 && (hist || hist_len == 0), LSH_ERR_INVALID, "bad histogram buffer");
    DeviceGuard g(idx->device);
    hipStream_t s = idx->stream;
    if (idx->sorted) {
        const uint32_t N = (uint32_t)idx->N;
        const uint32_t* k = idx->bkeys.get() + (size_t)table * N;
        DevBuf<unsigned long long> dh((size_t)hist_len + 2);
        LSH_HIP(hipMemsetAsync(dh.get(), 0, sizeof(unsigned long long) * ((size_t)hist_len + 2), s));
        unsigned long long* summary = dh.get() + hist_len;
        const unsigned grid = blocks_for(N, kTPB * 4, 2048);
        if (N > 0) {
            if (hist_len > 0 && hist_len <= kLdsBins)
                hipLaunchKernelGGL(sorted_occupancy_kernel<true>, dim3(grid), dim3(kTPB), sizeof(uint32_t) * hist_len,
                                   s, k, N, hist_len, dh.get(), summary);
            else
                hipLaunchKernelGGL(sorted_occupancy_kernel<false>, dim3(grid), dim3(kTPB), 0, s, k, N, hist_len,
                                   dh.get(), summary);
            check_launch("sorted_occupancy_kernel");
        }
        std::vector<unsigned long long> h((size_t)hist_len + 2);
        LSH_HIP(hipMemcpyAsync(h.data(), dh.get(), sizeof(unsigned long long) * h.size(), hipMemcpyDeviceToHost, s));
        LSH_HIP(hipStreamSynchronize(s));
        const int64_t buckets = (int64_t)1 << idx->sortBits;
        const int64_t nonempty = (int64_t)h[hist_len];
        for (int i = 0; i < hist_len; ++i) hist[i] = (int64_t)h[i];
        if (hist_len > 0) hist[0] += buckets - nonempty;  // empty buckets
        summary4[0] = buckets;
        summary4[1] = nonempty;
        summary4[2] = (int64_t)N;
        summary4[3] = (int64_t)h[hist_len + 1];
        return;
    }
    const uint32_t T = idx->tableSize;
    const uint32_t* off = idx->boff.get() + (size_t)table * (T + 1);
    DevBuf<unsigned long long> dh((size_t)hist_len + 2);
    LSH_HIP(hipMemsetAsync(dh.get(), 0, sizeof(unsigned long long) * ((size_t)hist_len + 2), s));
    unsigned long long* summary = dh.get() + hist_len;
    const unsigned grid = blocks_for(T, kTPB * 4, 2048);
    if (hist_len > 0 && hist_len <= kLdsBins) {
        hipLaunchKernelGGL(bucket_occupancy_kernel<true>, dim3(grid), dim3(kTPB), sizeof(uint32_t) * hist_len, s, off,
                           T, hist_len, dh.get(), summary);
    } else {
        hipLaunchKernelGGL(bucket_occupancy_kernel<false>, dim3(grid), dim3(kTPB), 0, s, off, T, hist_len, dh.get(),
                           summary);
    }
    check_launch("bucket_occupancy_kernel");
    uint32_t ends[2] = {0, 0};
    LSH_HIP(hipMemcpyAsync(&ends[0], off, sizeof(uint32_t), hipMemcpyDeviceToHost, s));
    LSH_HIP(hipMemcpyAsync(&ends[1], off + T, sizeof(uint32_t), hipMemcpyDeviceToHost, s));
    std::vector<unsigned long long> h((size_t)hist_len + 2);
    LSH_HIP(hipMemcpyAsync(h.data(), dh.get(), sizeof(unsigned long long) * h.size(), hipMemcpyDeviceToHost, s));
    LSH_HIP(hipStreamSynchronize(s));
    for (int i = 0; i < hist_len; ++i) hist[i] = (int64_t)h[i];
    summary4[0] = T;
    summary4[1] = (int64_t)h[hist_len];
    summary4[2] = (int64_t)ends[1] - (int64_t)ends[0];
    summary4[3] = (int64_t)h[hist_len + 1];
}

}  // namespace lshgpu
