// device_common.hpp -- device helpers shared by the scan, select and rerank
// kernels (header-only so every translation unit inlines them).
#pragma once

#include "lshgpu_internal.hpp"

namespace lshgpu {

// Appends `key` for the lanes with `pass` to query row `qrow` of the top-K
// state: one atomic per wave reserves the slots, each lane writes at its rank
// among the passing lanes.  Slots past newcap are counted but not written
// (the select kernel records the overflow).
__device__ inline void append_key(const TopKView& v, int qrow, bool pass, uint64_t key) {
    const unsigned long long pm = __ballot(pass);
    if (pm == 0) return;
    const int leader = __ffsll((long long)pm) - 1;
    const int lane = __lane_id();
    uint32_t base = 0;
    if (lane == leader) base = atomicAdd(&v.cnt[qrow], (uint32_t)__popcll(pm));
    base = __shfl(base, leader);
    if (pass) {
        const uint32_t rank = (uint32_t)__popcll(pm & ((1ull << lane) - 1ull));
        const uint32_t slot = base + rank;
        if (slot < (uint32_t)v.newcap) v.keys[(size_t)qrow * v.stride + v.K + slot] = key;
    }
}

// FP64 distance with the fixed operation order shared with the oracle
// (distanceFunctions.cuh:10-71 in double precision): sequential fused
// multiply-adds of the float inputs promoted to double, correctly rounded
// square root and division.
__device__ inline double dist64_dev(int kind, const float* __restrict__ q,
                                    const float* __restrict__ x, int d) {
    if (kind == kNegCos || kind == kAngular) {
        double dot = 0, nq = 0, nx = 0;
        for (int j = 0; j < d; ++j) {
            const double a = q[j], b = x[j];
            dot = __builtin_fma(a, b, dot);
            nq = __builtin_fma(a, a, nq);
            nx = __builtin_fma(b, b, nx);
        }
        const double c = dot / (__builtin_sqrt(nq) * __builtin_sqrt(nx));
        return kind == kNegCos ? -c : 1.0 - c;
    } else if (kind == kGenJaccard) {
        double smin = 0, smax = 0;
        for (int j = 0; j < d; ++j) {
            const double a = q[j], b = x[j];
            smin = smin + ((b < a) ? b : a);
            smax = smax + ((a < b) ? b : a);
        }
        return 1.0 - smin / smax;
    } else {
        double s = 0;
        for (int j = 0; j < d; ++j) {
            const double t = (double)q[j] - (double)x[j];
            s = __builtin_fma(t, t, s);
        }
        return __builtin_sqrt(s);
    }
}

// dist64_dev with the query's squared norm nq (angular kinds) given, and the
// row read with 16-byte loads when d % 4 == 0 and x is 16-byte aligned; the
// operations and their order are those of dist64_dev, so the result is the
// same.
__device__ inline double dist64_row(int kind, const float* __restrict__ q, const float* __restrict__ x, int d,
                                    double nq) {
    if ((d & 3) != 0 || (reinterpret_cast<uintptr_t>(x) & 15) != 0 || (reinterpret_cast<uintptr_t>(q) & 15) != 0)
        return dist64_dev(kind, q, x, d);
    const float4* __restrict__ x4 = reinterpret_cast<const float4*>(x);
    const float4* __restrict__ q4 = reinterpret_cast<const float4*>(q);
    if (kind == kNegCos || kind == kAngular) {
        double dot = 0, nx = 0;
#pragma unroll 5
        for (int j = 0; j < (d >> 2); ++j) {
            const float4 a = q4[j], b = x4[j];
            dot = __builtin_fma((double)a.x, (double)b.x, dot);
            nx = __builtin_fma((double)b.x, (double)b.x, nx);
            dot = __builtin_fma((double)a.y, (double)b.y, dot);
            nx = __builtin_fma((double)b.y, (double)b.y, nx);
            dot = __builtin_fma((double)a.z, (double)b.z, dot);
            nx = __builtin_fma((double)b.z, (double)b.z, nx);
            dot = __builtin_fma((double)a.w, (double)b.w, dot);
            nx = __builtin_fma((double)b.w, (double)b.w, nx);
        }
        const double c = dot / (__builtin_sqrt(nq) * __builtin_sqrt(nx));
        return kind == kNegCos ? -c : 1.0 - c;
    } else if (kind == kGenJaccard) {
        double smin = 0, smax = 0;
#pragma unroll 5
        for (int j = 0; j < (d >> 2); ++j) {
            const float4 a4 = q4[j], b4 = x4[j];
            const float av[4] = {a4.x, a4.y, a4.z, a4.w}, bv[4] = {b4.x, b4.y, b4.z, b4.w};
#pragma unroll
            for (int e = 0; e < 4; ++e) {
                const double a = av[e], b = bv[e];
                smin = smin + ((b < a) ? b : a);
                smax = smax + ((a < b) ? b : a);
            }
        }
        return 1.0 - smin / smax;
    } else {
        double s2 = 0;
#pragma unroll 5
        for (int j = 0; j < (d >> 2); ++j) {
            const float4 a4 = q4[j], b4 = x4[j];
            const float av[4] = {a4.x, a4.y, a4.z, a4.w}, bv[4] = {b4.x, b4.y, b4.z, b4.w};
#pragma unroll
            for (int e = 0; e < 4; ++e) {
                const double t = (double)av[e] - (double)bv[e];
                s2 = __builtin_fma(t, t, s2);
            }
        }
        return __builtin_sqrt(s2);
    }
}

// Squared norm of q in FP64, summed in index order as dist64_dev does.
__device__ inline double norm2_64(const float* __restrict__ q, int d) {
    double nq = 0;
    for (int j = 0; j < d; ++j) nq = __builtin_fma((double)q[j], (double)q[j], nq);
    return nq;
}

// FP32 distance as the reference's device code writes it
// (distanceFunctions.cuh:23-52; kNegCos per optimizedLinearScan.cu:45-57).
__device__ inline float dist32_dev(int kind, const float* __restrict__ q,
                                   const float* __restrict__ x, int d, float q_norm) {
    if (kind == kNegCos || kind == kAngular) {
        float dot = 0.0f, nx = 0.0f;
        for (int j = 0; j < d; ++j) {
            dot = __builtin_fmaf(q[j], x[j], dot);
            nx = __builtin_fmaf(x[j], x[j], nx);
        }
        const float c = dot / (q_norm * __builtin_sqrtf(nx));
        return kind == kNegCos ? -c : 1.0f - c;
    } else if (kind == kGenJaccard) {
        float smin = 0.0f, smax = 0.0f;
        for (int j = 0; j < d; ++j) {
            smin += fminf(x[j], q[j]);
            smax += fmaxf(x[j], q[j]);
        }
        return 1.0f - smin / smax;
    } else {
        float s = 0.0f;
        for (int j = 0; j < d; ++j) {
            const float t = q[j] - x[j];
            s = __builtin_fmaf(t, t, s);
        }
        return s;  // squared; kEuclid ranks by the square
    }
}

// Partial sums of one lane of a group that splits a row over its lanes:
// (dot, |x|^2) for the angular kinds, (sum min, sum max) for generalized
// Jaccard, (sum of squares, -) for L2.  x is the data value, y the query's.
__device__ inline void group_accumulate(int kind, float x, float y, float& a, float& b) {
    if (kind == kNegCos || kind == kAngular) {
        a = __builtin_fmaf(x, y, a);
        b = __builtin_fmaf(x, x, b);
    } else if (kind == kGenJaccard) {
        a += fminf(x, y);
        b += fmaxf(x, y);
    } else {
        const float t = y - x;
        a = __builtin_fmaf(t, t, a);
    }
}

// FP32 distance from the reduced partial sums (q_norm = ||q|| in FP32).
__device__ inline float group_distance(int kind, float a, float b, float q_norm) {
    if (kind == kNegCos || kind == kAngular) {
        const float c = a / (q_norm * __builtin_sqrtf(b));
        return kind == kNegCos ? -c : 1.0f - c;
    } else if (kind == kGenJaccard) {
        return 1.0f - a / b;
    }
    return a;  // squared; kEuclid ranks by the square
}

__device__ inline void cmpx_u64(uint64_t* a, int i, int j, bool asc) {
    const uint64_t x = a[i], y = a[j];
    if ((x > y) == asc) {
        a[i] = y;
        a[j] = x;
    }
}

// Ascending bitonic sort of n (power of two) keys in LDS by the whole block.
__device__ inline void bitonic_sort_lds(uint64_t* a, int n) {
    for (int size = 2; size <= n; size <<= 1) {
        for (int stride = size >> 1; stride > 0; stride >>= 1) {
            __syncthreads();
            for (int t = threadIdx.x; t < (n >> 1); t += blockDim.x) {
                const int i = 2 * t - (t & (stride - 1));
                cmpx_u64(a, i, i + stride, (i & size) == 0);
            }
        }
    }
    __syncthreads();
}

// Ascending sort of (key, id) pairs (n a power of two) in LDS.
__device__ inline void bitonic_sort_pairs(uint64_t* k, int32_t* id, int n) {
    for (int size = 2; size <= n; size <<= 1) {
        for (int stride = size >> 1; stride > 0; stride >>= 1) {
            __syncthreads();
            for (int t = threadIdx.x; t < (n >> 1); t += blockDim.x) {
                const int i = 2 * t - (t & (stride - 1));
                const int j = i + stride;
                const bool asc = (i & size) == 0;
                const uint64_t ki = k[i], kj = k[j];
                const int32_t ii = id[i], ij = id[j];
                const bool gt = ki > kj || (ki == kj && ii > ij);
                if (gt == asc) {
                    k[i] = kj;
                    k[j] = ki;
                    id[i] = ij;
                    id[j] = ii;
                }
            }
        }
    }
    __syncthreads();
}

}  // namespace lshgpu
