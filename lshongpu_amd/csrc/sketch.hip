// sketch.hip -- sketch and key builders plus the data transforms they need.
//
// SimHash (simHash.cu:22-47, simHash.cuh:20-46) and SimHash-JL
// (simHashJL.cuh:34-51): one lane per point; the hyperplanes are stored
// transposed ([word][dim][bit]) so that, for a fixed dimension, the 32
// hyperplane values a wave needs are wave-uniform and arrive as scalar
// operands; 32 independent fused multiply-add chains per lane run in
// dimension order, the same rounding sequence as the reference's
// `dot += x*r` (contracted to fmaf by nvcc), so the bits match it exactly.
//
// Cross-polytope (crossPolyTope.cuh:25-85): one lane per point, the padded
// vector lives in VGPRs and the Walsh-Hadamard butterflies are register
// moves (the reference keeps it in global scratch).
#include "device_common.hpp"

namespace lshgpu {

namespace {

constexpr int TPB = 256;

// SimHash of 64 rows per workgroup: the rows are staged in LDS with
// coalesced reads (row stride d + 1, so that the lanes' reads of one
// dimension spread over the banks), and each of the 4 waves computes one
// word (blockIdx.y * 4 + wave) of all 64 rows, one row per lane.  The
// hyperplane values are wave-uniform (scalar operands).
constexpr int kStageRows = 64;
__global__ __launch_bounds__(256) void simhash_staged_kernel(const float* __restrict__ x, int64_t n, int d,
                                                             const float* __restrict__ pT, int words, int hb,
                                                             uint32_t* __restrict__ out, int ostride) {
    extern __shared__ float xs[];  // [kStageRows][d + 1]
    const int64_t i0 = (int64_t)blockIdx.x * kStageRows;
    const int rows = (int)min<int64_t>(kStageRows, n - i0);
    const int ld = d + 1;
    for (int e = threadIdx.x; e < rows * d; e += 256) {
        const int r = e / d, j = e - r * d;
        xs[r * ld + j] = x[(size_t)i0 * d + e];
    }
    __syncthreads();
    const int lane = threadIdx.x & 63;
    // The wave's word, made provably uniform so the hyperplane reads are scalar.
    const int w = __builtin_amdgcn_readfirstlane(blockIdx.y * 4 + (threadIdx.x >> 6));
    if (w >= words || lane >= rows) return;
    const float* __restrict__ xr = xs + lane * ld;
    float acc[32];
#pragma unroll
    for (int b = 0; b < 32; ++b) acc[b] = 0.0f;
    const float* __restrict__ p = pT + (size_t)w * d * hb;
    if (hb == 32) {
        // Full words: the 32 hyperplane values of a dimension are one
        // contiguous uniform block (wide scalar loads, no per-bit tests).
        // Two register sets alternate: the block and row value of dimension
        // j + 1 load while dimension j's multiply-adds run.  Scalar loads
        // complete out of order, so any use waits for every load in flight:
        // each set's first use (which takes that wait) is fenced before the
        // other set's loads are issued, and its other 31 multiply-adds then
        // cover their latency.
        float pa[32], pb[32];
#pragma unroll
        for (int b = 0; b < 32; ++b) pa[b] = p[b];
        float xa = xr[0], xb;
        int j = 0;
        for (; j + 1 < d; j += 2) {
            acc[0] = __builtin_fmaf(xa, pa[0], acc[0]);
            __builtin_amdgcn_sched_barrier(0);
            xb = xr[j + 1];
#pragma unroll
            for (int b = 0; b < 32; ++b) pb[b] = p[(size_t)(j + 1) * 32 + b];
            __builtin_amdgcn_sched_barrier(0);
#pragma unroll
            for (int b = 1; b < 32; ++b) acc[b] = __builtin_fmaf(xa, pa[b], acc[b]);
            acc[0] = __builtin_fmaf(xb, pb[0], acc[0]);
            __builtin_amdgcn_sched_barrier(0);
            const int jn = min(j + 2, d - 1);  // past d - 1 the set is loaded but unused
            xa = xr[jn];
#pragma unroll
            for (int b = 0; b < 32; ++b) pa[b] = p[(size_t)jn * 32 + b];
            __builtin_amdgcn_sched_barrier(0);
#pragma unroll
            for (int b = 1; b < 32; ++b) acc[b] = __builtin_fmaf(xb, pb[b], acc[b]);
        }
        if (j < d) {  // odd d: the last dimension is in the first set
#pragma unroll
            for (int b = 0; b < 32; ++b) acc[b] = __builtin_fmaf(xa, pa[b], acc[b]);
        }
    } else {
        for (int j = 0; j < d; ++j) {
            const float xj = xr[j];
            const float* __restrict__ pj = p + (size_t)j * hb;
#pragma unroll
            for (int b = 0; b < 32; ++b)
                if (b < hb) acc[b] = __builtin_fmaf(xj, pj[b], acc[b]);
        }
    }
    uint32_t sk = 0;
#pragma unroll
    for (int b = 0; b < 32; ++b)
        if (b < hb) sk |= (acc[b] >= 0.0f ? 1u : 0u) << b;
    out[(size_t)(i0 + lane) * ostride + w] = sk;
}

// One lane per (point, word); blockIdx.y is the word, so the hyperplane
// values a wave reads are uniform across it.  Used when a staged block of
// rows does not fit the LDS.
__global__ __launch_bounds__(TPB) void simhash_kernel(const float* __restrict__ x, int64_t n, int d,
                                                      const float* __restrict__ pT, int words, int hb,
                                                      uint32_t* __restrict__ out, int ostride) {
    const int64_t i = (int64_t)blockIdx.x * TPB + threadIdx.x;
    const int w = blockIdx.y;
    if (i >= n) return;
    const float* __restrict__ row = x + (size_t)i * d;
    float acc[32];
#pragma unroll
    for (int b = 0; b < 32; ++b) acc[b] = 0.0f;
    const float* __restrict__ p = pT + (size_t)w * d * hb;
    for (int j = 0; j < d; ++j) {
        const float xj = row[j];
        const float* __restrict__ pj = p + (size_t)j * hb;
#pragma unroll
        for (int b = 0; b < 32; ++b)
            if (b < hb) acc[b] = __builtin_fmaf(xj, pj[b], acc[b]);
    }
    uint32_t s = 0;
#pragma unroll
    for (int b = 0; b < 32; ++b)
        if (b < hb) s |= (acc[b] >= 0.0f ? 1u : 0u) << b;
    out[(size_t)i * ostride + w] = s;
}

// pT is [groups of 32 sketch values][d][32].
__global__ __launch_bounds__(TPB) void jl_kernel(const float* __restrict__ x, int64_t n, int d,
                                                 const float* __restrict__ pT, int sdim,
                                                 float* __restrict__ out) {
    const int64_t i = (int64_t)blockIdx.x * TPB + threadIdx.x;
    if (i >= n) return;
    const float* __restrict__ row = x + (size_t)i * d;
    const int groups = (sdim + 31) / 32;
    for (int g = 0; g < groups; ++g) {
        float acc[32];
#pragma unroll
        for (int b = 0; b < 32; ++b) acc[b] = 0.0f;
        const float* __restrict__ p = pT + (size_t)g * d * 32;
        for (int j = 0; j < d; ++j) {
            const float xj = row[j];
#pragma unroll
            for (int b = 0; b < 32; ++b) acc[b] = __builtin_fmaf(xj, p[(size_t)j * 32 + b], acc[b]);
        }
#pragma unroll
        for (int b = 0; b < 32; ++b)
            if (g * 32 + b < sdim) out[(size_t)i * sdim + g * 32 + b] = acc[b];
    }
}

// processingUtils.cuh:48-67: magnitude as a fused multiply-add chain, IEEE
// square root and division.
__global__ void normalize_rows_kernel(float* __restrict__ x, int64_t n, int d) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    float* r = x + (size_t)i * d;
    float m = 0.0f;
    for (int j = 0; j < d; ++j) m = __builtin_fmaf(r[j], r[j], m);
    m = __builtin_sqrtf(m);
    for (int j = 0; j < d; ++j) r[j] = r[j] / m;
}

// The same, 64 rows per wave staged through LDS (row stride d + 1, so the
// per-row chains of the 64 lanes read distinct banks): the loads and stores
// are coalesced, and each lane's chain reads LDS instead of strided memory.
__global__ __launch_bounds__(64) void normalize_rows_lds_kernel(float* __restrict__ x, int64_t n, int d) {
    extern __shared__ float t[];
    const int64_t r0 = (int64_t)blockIdx.x * 64;
    const int rows = (int)min<int64_t>(64, n - r0);
    const int S = d + 1;
    float* __restrict__ g = x + (size_t)r0 * d;
    for (int e = threadIdx.x; e < rows * d; e += 64) {
        const int r = e / d;
        t[r * S + (e - r * d)] = g[e];
    }
    __syncthreads();
    if ((int)threadIdx.x < rows) {
        float* r = t + threadIdx.x * S;
        float m = 0.0f;
        for (int j = 0; j < d; ++j) m = __builtin_fmaf(r[j], r[j], m);
        m = __builtin_sqrtf(m);
        for (int j = 0; j < d; ++j) r[j] = r[j] / m;
    }
    __syncthreads();
    for (int e = threadIdx.x; e < rows * d; e += 64) {
        const int r = e / d;
        g[e] = t[r * S + (e - r * d)];
    }
}

// processingUtils.cuh:12-31: per-dimension min of (int)floor(x), starting
// from 0, reduced per workgroup and then with one atomic per dimension.
__global__ void gj_mins_kernel(const float* __restrict__ x, int64_t n, int d, int rows_per_block,
                               int* __restrict__ mins) {
    const int64_t r0 = (int64_t)blockIdx.x * rows_per_block;
    const int64_t r1 = min(n, r0 + rows_per_block);
    for (int j = threadIdx.x; j < d; j += blockDim.x) {
        int m = 0;
        for (int64_t r = r0; r < r1; ++r) m = min(m, (int)floorf(x[(size_t)r * d + j]));
        atomicMin(&mins[j], m);
    }
}

// processingUtils.cuh:33-46: x += abs(min).
__global__ void gj_shift_kernel(float* __restrict__ x, int64_t n, int d, const int* __restrict__ mins) {
    const size_t total = (size_t)n * d;
    for (size_t t = (size_t)blockIdx.x * blockDim.x + threadIdx.x; t < total;
         t += (size_t)gridDim.x * blockDim.x) {
        const int j = (int)(t % d);
        x[t] += (float)abs(mins[j]);
    }
}

template <int P>
__global__ __launch_bounds__(64) void crosspoly_kernel(const float* __restrict__ x, int64_t n, int d,
                                                       int tables, const int32_t* __restrict__ signs,
                                                       uint16_t* __restrict__ keys) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const float* __restrict__ row = x + (size_t)i * d;
    for (int t = 0; t < tables; ++t) {
        float v[P];
#pragma unroll
        for (int j = 0; j < P; ++j) v[j] = j < d ? row[j] : 0.0f;
#pragma unroll
        for (int r = 0; r < 3; ++r) {
            const int32_t* __restrict__ sg = signs + ((size_t)t * 3 + r) * P;
#pragma unroll
            for (int j = 0; j < P; ++j) v[j] *= (float)sg[j];
#pragma unroll
            for (int o = P / 2; o > 0; o >>= 1) {
#pragma unroll
                for (int j = 0; j < P; ++j) {
                    if ((j & o) == 0) {
                        const float a = v[j], b = v[j + o];
                        v[j] = a + b;
                        v[j + o] = a - b;
                    }
                }
            }
        }
        int index = 0;
        float best = -2147483648.0f;  // (float)INT_MIN
#pragma unroll
        for (int j = 0; j < P; ++j) {
            const bool neg = v[j] < 0.0f;
            const float m = neg ? fmaxf(best, -v[j]) : fmaxf(best, v[j]);
            if (m > best) {
                index = neg ? j : j + P;
                best = m;
            }
        }
        keys[(size_t)i * tables + t] = (uint16_t)index;
    }
}

unsigned blocks_for(int64_t n, int tpb) { return (unsigned)((n + tpb - 1) / tpb); }

}  // namespace

void launch_simhash(const float* x, int64_t n, int d, const float* planesT, int words, int hashBits,
                    uint32_t* out, int out_stride, hipStream_t s) {
    if (n <= 0) return;
    LSH_REQUIRE(hashBits >= 1 && hashBits <= 32, LSH_ERR_INVALID, "hashBits must be in 1..32");
    LSH_REQUIRE(words <= 65535, LSH_ERR_UNSUPPORTED, "at most 65535 sketch words");
    const size_t lds = sizeof(float) * kStageRows * (size_t)(d + 1);
    if (lds <= 64 * 1024) {
        hipLaunchKernelGGL(simhash_staged_kernel, dim3(blocks_for(n, kStageRows), (words + 3) / 4), dim3(256), lds, s,
                           x, n, d, planesT, words, hashBits, out, out_stride);
        check_launch("simhash_staged_kernel");
        return;
    }
    hipLaunchKernelGGL(simhash_kernel, dim3(blocks_for(n, TPB), words), dim3(TPB), 0, s, x, n, d, planesT,
                       words, hashBits, out, out_stride);
    check_launch("simhash_kernel");
}

void launch_jl(const float* x, int64_t n, int d, const float* planesT, int sdim, float* out,
               hipStream_t s) {
    if (n <= 0) return;
    hipLaunchKernelGGL(jl_kernel, dim3(blocks_for(n, TPB)), dim3(TPB), 0, s, x, n, d, planesT, sdim, out);
    check_launch("jl_kernel");
}

void launch_normalize_rows(float* x, int64_t n, int d, hipStream_t s) {
    if (n <= 0) return;
    if (d <= 255) {
        const size_t lds = sizeof(float) * 64 * (size_t)(d + 1);
        hipLaunchKernelGGL(normalize_rows_lds_kernel, dim3((unsigned)((n + 63) / 64)), dim3(64), lds, s, x, n, d);
        check_launch("normalize_rows_lds_kernel");
        return;
    }
    hipLaunchKernelGGL(normalize_rows_kernel, dim3(blocks_for(n, 256)), dim3(256), 0, s, x, n, d);
    check_launch("normalize_rows_kernel");
}

void launch_gj_mins(const float* x, int64_t n, int d, int* mins, hipStream_t s) {
    if (n <= 0) return;
    const int rpb = 256;
    hipLaunchKernelGGL(gj_mins_kernel, dim3(blocks_for(n, rpb)), dim3(128), 0, s, x, n, d, rpb, mins);
    check_launch("gj_mins_kernel");
}

void launch_gj_shift(float* x, int64_t n, int d, const int* mins, hipStream_t s) {
    if (n <= 0) return;
    const size_t total = (size_t)n * d;
    const unsigned b = (unsigned)std::min<size_t>((total + 255) / 256, 65535 * 4);
    hipLaunchKernelGGL(gj_shift_kernel, dim3(b), dim3(256), 0, s, x, n, d, mins);
    check_launch("gj_shift_kernel");
}

void launch_crosspoly(const float* x, int64_t n, int d, int tables, const int32_t* signs,
                      uint16_t* keys, hipStream_t s) {
    if (n <= 0) return;
    int P = 1;
    while (P < d) P *= 2;
    const dim3 g(blocks_for(n, 64)), b(64);
    switch (P) {
        case 1: hipLaunchKernelGGL(crosspoly_kernel<1>, g, b, 0, s, x, n, d, tables, signs, keys); break;
        case 2: hipLaunchKernelGGL(crosspoly_kernel<2>, g, b, 0, s, x, n, d, tables, signs, keys); break;
        case 4: hipLaunchKernelGGL(crosspoly_kernel<4>, g, b, 0, s, x, n, d, tables, signs, keys); break;
        case 8: hipLaunchKernelGGL(crosspoly_kernel<8>, g, b, 0, s, x, n, d, tables, signs, keys); break;
        case 16: hipLaunchKernelGGL(crosspoly_kernel<16>, g, b, 0, s, x, n, d, tables, signs, keys); break;
        case 32: hipLaunchKernelGGL(crosspoly_kernel<32>, g, b, 0, s, x, n, d, tables, signs, keys); break;
        case 64: hipLaunchKernelGGL(crosspoly_kernel<64>, g, b, 0, s, x, n, d, tables, signs, keys); break;
        case 128: hipLaunchKernelGGL(crosspoly_kernel<128>, g, b, 0, s, x, n, d, tables, signs, keys); break;
        case 256: hipLaunchKernelGGL(crosspoly_kernel<256>, g, b, 0, s, x, n, d, tables, signs, keys); break;
        default: fail(LSH_ERR_UNSUPPORTED, "cross-polytope hashing supports d <= 256");
    }
    check_launch("crosspoly_kernel");
}

}  // namespace lshgpu
