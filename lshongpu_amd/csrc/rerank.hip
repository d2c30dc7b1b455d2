// rerank.hip -- exact rerank of a sketch-distance candidate set.
//
// candidateSetScanner.cuh:10-32 recomputes the exact distance of every slot
// of a lane's queue and the final startSort/write of
// sketchedDistanceScanners.cuh:87-99 keeps the k best.  Here one workgroup
// owns one query: the C candidates get their FP32 distance (gathered rows),
// a bitonic sort in LDS orders them by (FP32 distance, ID), the best K' are
// recomputed in FP64 and re-sorted by (FP64 distance, ID).  When the FP32
// gap between the k-th and the K'-th candidate is not larger than twice the
// FP32 error bound, every candidate is recomputed in FP64 instead, so the
// returned IDs are always the FP64 order of the candidate set.
#include "device_common.hpp"

namespace lshgpu {

namespace {

__global__ __launch_bounds__(256) void rerank_kernel(const float* __restrict__ data,
                                                     const float* __restrict__ queries, int d,
                                                     int kind, const uint64_t* __restrict__ keys,
                                                     int key_stride, const int32_t* __restrict__ cand,
                                                     int C, int Cp, int k, int kprime, double eps,
                                                     lsh_point_t* __restrict__ out) {
    extern __shared__ uint64_t sm[];
    uint64_t* k32 = sm;                                        // Cp
    uint64_t* k64 = sm + Cp;                                   // Cp
    int32_t* ids = reinterpret_cast<int32_t*>(sm + 2 * Cp);    // Cp
    __shared__ int s_ok;
    const int q = blockIdx.x;
    const float* __restrict__ qv = queries + (size_t)q * d;

    float qn = 0.0f;
    if (kind == kNegCos || kind == kAngular) {
        for (int j = 0; j < d; ++j) qn = __builtin_fmaf(qv[j], qv[j], qn);
        qn = __builtin_sqrtf(qn);
    }
    // FP32 distances: a group of 16 lanes reads one candidate row with
    // coalesced float4 loads and reduces its partial sums with shuffles.
    const int gl = threadIdx.x & 15;
    const int ngroups = blockDim.x >> 4;
    const bool vec4 = (d & 3) == 0;
    for (int t = threadIdx.x >> 4; t < Cp; t += ngroups) {
        int32_t id = -1;
        if (t < C) {
            if (cand) {
                id = cand[(size_t)q * C + t];
            } else {
                const uint64_t key = keys[(size_t)q * key_stride + t];
                id = key == kEmptyKey ? -1 : (int32_t)(uint32_t)key;
            }
        }
        float a = 0.0f, b = 0.0f;
        if (id >= 0) {
            const float* __restrict__ x = data + (size_t)id * d;
            if (vec4) {
                for (int j = gl * 4; j < d; j += 64) {
                    const float4 xv = *reinterpret_cast<const float4*>(x + j);
                    const float4 yv = *reinterpret_cast<const float4*>(qv + j);
                    group_accumulate(kind, xv.x, yv.x, a, b);
                    group_accumulate(kind, xv.y, yv.y, a, b);
                    group_accumulate(kind, xv.z, yv.z, a, b);
                    group_accumulate(kind, xv.w, yv.w, a, b);
                }
            } else {
                for (int j = gl; j < d; j += 16) group_accumulate(kind, x[j], qv[j], a, b);
            }
        }
#pragma unroll
        for (int o = 8; o > 0; o >>= 1) {
            a += __shfl_xor(a, o, 16);
            b += __shfl_xor(b, o, 16);
        }
        if (gl == 0) {
            const float dist = group_distance(kind, a, b, qn);
            k32[t] = id < 0 ? kEmptyKey : make_key(f32_key(dist), (uint32_t)id);
        }
    }
    bitonic_sort_lds(k32, Cp);

    const int kp = min(kprime, Cp);
    for (int t = threadIdx.x; t < kp; t += blockDim.x) {
        const uint64_t key = k32[t];
        if (key == kEmptyKey) {
            k64[t] = kEmptyKey;
            ids[t] = 0x7FFFFFFF;
        } else {
            const int32_t id = (int32_t)(uint32_t)key;
            k64[t] = f64_key(dist64_dev(kind, qv, data + (size_t)id * d, d));
            ids[t] = id;
        }
    }
    if (threadIdx.x == 0) {
        bool ok = true;
        if (kp < Cp) {
            const uint64_t kk = k32[k - 1], kK = k32[kp - 1];
            if (kK != kEmptyKey && kk != kEmptyKey) {
                const double vk = key_f32((uint32_t)(kk >> 32));
                const double vK = key_f32((uint32_t)(kK >> 32));
                if (kind == kEuclid)
                    ok = vK * (1.0 - eps) > vk * (1.0 + eps);
                else
                    ok = vK - vk > 2.0 * eps;
            }
        }
        s_ok = ok;
    }
    __syncthreads();
    int n = kp;
    if (!s_ok) {
        // FP64 for every candidate.
        for (int t = threadIdx.x + kp; t < Cp; t += blockDim.x) {
            const uint64_t key = k32[t];
            if (key == kEmptyKey) {
                k64[t] = kEmptyKey;
                ids[t] = 0x7FFFFFFF;
            } else {
                const int32_t id = (int32_t)(uint32_t)key;
                k64[t] = f64_key(dist64_dev(kind, qv, data + (size_t)id * d, d));
                ids[t] = id;
            }
        }
        n = Cp;
    }
    bitonic_sort_pairs(k64, ids, n);
    const float pad = kind == kNegCos ? 2.0f : kPadDistance;
    for (int t = threadIdx.x; t < k; t += blockDim.x) {
        lsh_point_t p;
        if (t < n && ids[t] != 0x7FFFFFFF) {
            p.ID = ids[t];
            p.distance = (float)key_f64(k64[t]);
        } else {
            p.ID = -1;
            p.distance = pad;
        }
        out[(size_t)q * k + t] = p;
    }
}

}  // namespace

void launch_rerank(const float* data, const float* queries, int d, int kind, const uint64_t* keys,
                   int key_stride, const int32_t* cand, int C, int Q, int k, int kprime, double eps,
                   lsh_point_t* out, hipStream_t s) {
    if (Q <= 0) return;
    const int Cp = pow2_ceil(std::max(C, 2));
    LSH_REQUIRE(Cp <= kMaxTopK, LSH_ERR_INVALID, "candidate set larger than 4096");
    const int kp = pow2_ceil(std::max(kprime, k));
    const size_t lds = (size_t)Cp * (2 * sizeof(uint64_t) + sizeof(int32_t));
    set_dynamic_lds((const void*)rerank_kernel, lds);
    hipLaunchKernelGGL(rerank_kernel, dim3(Q), dim3(256), lds, s, data, queries, d, kind, keys, key_stride,
                       cand, C, Cp, k, kp, eps, out);
    check_launch("rerank_kernel");
}

}  // namespace lshgpu
