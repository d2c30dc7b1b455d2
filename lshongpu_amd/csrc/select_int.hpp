// select_int.hpp -- the K-th smallest of a set of integer-distance keys, by
// one workgroup (header-only device code of the integer-key top-K select).
//
// A histogram of the distances (the key's high word; the last bin also holds
// every larger distance) in LDS gives the bin T that holds the K-th key.  The
// keys of that bin, when at most kTieCap, are gathered in LDS and the K-th is
// the one whose rank among them (keys are distinct) is right; a larger bin
// is resolved by a radix select over its keys in place (8-bit digits of the
// ID, or of the whole key in the last bin).
#pragma once

#include "lshgpu_internal.hpp"

namespace lshgpu {

constexpr int kSelBins = 1280;  // one bin per distance up to 1278 (1024-bit sketches); the last holds the rest
constexpr int kTieCap = 512;      // keys of the boundary bin gathered in LDS (more: selected in place)
constexpr int kSelScratch = 8;    // LDS words int_kth_key uses for its shared scalars
// Wave 0 of int_kth_key scans the histogram in 64 equal slices.
static_assert(kSelBins % 64 == 0, "the histogram scan splits kSelBins bins over 64 lanes");

// The keys are visited through `each(f)`, which calls f(key) for every key
// the calling thread owns (every thread of the workgroup together covers the
// set once): from memory, or from registers loaded once.
struct KeysInMemory {
    const uint64_t* in;
    int n;
    template <class F>
    __device__ void operator()(F&& f) const {
        for (int i = threadIdx.x; i < n; i += blockDim.x) f(in[i]);
    }
};

// The rank-th smallest (1-based) key among the non-empty keys whose bin is T
// and whose bits above `shift` + 8 match prefix; shifts run from `top` down
// to 0.  Uses hist[0, 256).
template <class Each>
__device__ inline uint64_t radix_kth_in_bin(const Each& each, uint32_t T, uint32_t rank, int top, uint64_t prefix,
                                            uint32_t* hist, uint32_t* s_dig, uint32_t* s_rank) {
    uint64_t mask = top == 24 ? 0xFFFFFFFF00000000ull : 0ull;
    for (int shift = top; shift >= 0; shift -= 8) {
        for (int i = threadIdx.x; i < 256; i += blockDim.x) hist[i] = 0;
        __syncthreads();
        each([&](uint64_t key) {
            if (key == kEmptyKey || min((uint32_t)(key >> 32), (uint32_t)(kSelBins - 1)) != T) return;
            if ((key & mask) == prefix) atomicAdd(&hist[(uint32_t)(key >> shift) & 255u], 1u);
        });
        __syncthreads();
        if (threadIdx.x < 64) {
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
                    *s_dig = 4u * lane + e;
                    *s_rank = rank - run;
                }
                run += c[e];
            }
        }
        __syncthreads();
        prefix |= (uint64_t)*s_dig << shift;
        mask |= 0xFFull << shift;
        rank = *s_rank;
        __syncthreads();
    }
    return prefix;
}

// The caller's hook of int_kth_key: nothing.
struct NoBoundaryPrep {
    __device__ void operator()(uint32_t) const {}
};

// The K-th smallest non-empty key of the set `each` visits (K >= 1), or
// kEmptyKey when there are fewer than K.  LDS: hist[kSelBins],
// ties[kTieCap], sv[kSelScratch].  Every thread of the workgroup must call
// it; it ends with a barrier, so the caller may reuse that LDS at once.
// prep(T) runs in every thread once the boundary bin T is known (T =
// kSelBins: fewer than K keys), before the keys of that bin are compared:
// the caller may then rewrite the low words of its keys in bin T (the
// order of equal distances), which `each` must then visit.
template <class Each, class Prep = NoBoundaryPrep>
__device__ inline uint64_t int_kth_key(const Each& each, uint32_t K, uint32_t* hist, uint64_t* ties, uint32_t* sv,
                                       const Prep& prep = Prep{}) {
    uint32_t& s_T = sv[0];
    uint32_t& s_below = sv[1];
    uint32_t& s_nties = sv[4];
    for (int i = threadIdx.x; i < kSelBins; i += blockDim.x) hist[i] = 0;
    if (threadIdx.x == 0) s_nties = 0;
    __syncthreads();
    each([&](uint64_t key) {
        if (key != kEmptyKey) atomicAdd(&hist[min((uint32_t)(key >> 32), (uint32_t)(kSelBins - 1))], 1u);
    });
    __syncthreads();
    if (threadIdx.x < 64) {
        // Wave 0 finds the first bin at which the running count reaches K.
        const int lane = threadIdx.x;
        constexpr int per = kSelBins / 64;
        uint32_t local = 0;
        for (int b = 0; b < per; ++b) local += hist[lane * per + b];
        uint32_t incl = local;
#pragma unroll
        for (int o = 1; o < 64; o <<= 1) {
            const uint32_t t = __shfl_up(incl, o);
            if (lane >= o) incl += t;
        }
        const uint32_t excl = incl - local;
        const uint32_t all = __shfl(incl, 63);
        const unsigned long long has = __ballot(incl >= K && excl < K);
        if (all < K) {
            if (lane == 0) {
                s_T = kSelBins;  // fewer than K keys in total: keep them all
                s_below = all;
            }
        } else {
            const int owner = __ffsll((long long)has) - 1;
            if (lane == owner) {
                uint32_t c = excl;
                int b = lane * per;
                while (c + hist[b] < K) c += hist[b++];
                s_T = (uint32_t)b;
                s_below = c;
            }
        }
    }
    __syncthreads();
    const uint32_t T = s_T;
    prep(T);
    uint64_t kth = kEmptyKey;
    if (T < (uint32_t)kSelBins) {
        const bool last = T == (uint32_t)(kSelBins - 1);
        const uint32_t nt = hist[T];  // keys in the boundary bin
        const uint32_t below = s_below;
        __syncthreads();
        if (nt <= (uint32_t)kTieCap) {
            // Gather them; the K-th is the key with at most want = K - below
            // - 1 smaller keys and more than want smaller or equal (0 <=
            // want < nt, so exactly one key value qualifies).
            each([&](uint64_t key) {
                if (key != kEmptyKey && min((uint32_t)(key >> 32), (uint32_t)(kSelBins - 1)) == T)
                    ties[atomicAdd(&s_nties, 1u)] = key;
            });
            __syncthreads();
            const uint32_t want = K - below - 1u;
            for (uint32_t t = threadIdx.x; t < nt; t += blockDim.x) {
                const uint64_t key = ties[t];
                uint32_t lt = 0, le = 0;
                for (uint32_t u = 0; u < nt; ++u) {
                    const uint64_t x = ties[u];
                    lt += x < key ? 1u : 0u;
                    le += x <= key ? 1u : 0u;
                }
                if (lt <= want && want < le) {
                    // Exactly one thread finds it; the key goes to two scalar words.
                    sv[5] = (uint32_t)key;
                    sv[6] = (uint32_t)(key >> 32);
                }
            }
            __syncthreads();
            kth = ((uint64_t)sv[6] << 32) | sv[5];
        } else {
            kth = radix_kth_in_bin(each, T, K - below, last ? 56 : 24, last ? 0ull : (uint64_t)T << 32, hist, &sv[2],
                                   &sv[3]);
        }
    }
    __syncthreads();
    return kth;
}

}  // namespace lshgpu
