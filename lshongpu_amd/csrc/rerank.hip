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
    // Each group takes U rows per step so that U row reads are in flight.
    constexpr int U = 4;
    for (int t0 = threadIdx.x >> 4; t0 < Cp; t0 += ngroups * U) {
        int32_t id[U];
        float a[U], b[U];
#pragma unroll
        for (int u = 0; u < U; ++u) {
            const int t = t0 + u * ngroups;
            id[u] = -1;
            if (t < C) {
                if (cand) {
                    id[u] = cand[(size_t)q * C + t];
                } else {
                    const uint64_t key = keys[(size_t)q * key_stride + t];
                    id[u] = key == kEmptyKey ? -1 : (int32_t)(uint32_t)key;
                }
            }
            a[u] = 0.0f;
            b[u] = 0.0f;
        }
        if (vec4) {
            for (int j = gl * 4; j < d; j += 64) {
                const float4 yv = *reinterpret_cast<const float4*>(qv + j);
                float4 xv[U];
#pragma unroll
                for (int u = 0; u < U; ++u)
                    xv[u] = id[u] >= 0 ? *reinterpret_cast<const float4*>(data + (size_t)id[u] * d + j)
                                       : make_float4(0.0f, 0.0f, 0.0f, 0.0f);
#pragma unroll
                for (int u = 0; u < U; ++u) {
                    group_accumulate(kind, xv[u].x, yv.x, a[u], b[u]);
                    group_accumulate(kind, xv[u].y, yv.y, a[u], b[u]);
                    group_accumulate(kind, xv[u].z, yv.z, a[u], b[u]);
                    group_accumulate(kind, xv[u].w, yv.w, a[u], b[u]);
                }
            }
        } else {
            for (int j = gl; j < d; j += 16) {
                const float y = qv[j];
                float xs[U];
#pragma unroll
                for (int u = 0; u < U; ++u) xs[u] = id[u] >= 0 ? data[(size_t)id[u] * d + j] : 0.0f;
#pragma unroll
                for (int u = 0; u < U; ++u) group_accumulate(kind, xs[u], y, a[u], b[u]);
            }
        }
#pragma unroll
        for (int u = 0; u < U; ++u) {
#pragma unroll
            for (int o = 8; o > 0; o >>= 1) {
                a[u] += __shfl_xor(a[u], o, 16);
                b[u] += __shfl_xor(b[u], o, 16);
            }
            const int t = t0 + u * ngroups;
            if (gl == 0 && t < Cp) {
                const float dist = group_distance(kind, a[u], b[u], qn);
                k32[t] = id[u] < 0 ? kEmptyKey : make_key(f32_key(dist), (uint32_t)id[u]);
            }
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
