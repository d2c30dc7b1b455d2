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

__device__ inline uint64_t shfl_xor_u64(uint64_t x, int m) {
    const uint32_t lo = (uint32_t)__shfl_xor((int)(uint32_t)x, m);
    const uint32_t hi = (uint32_t)__shfl_xor((int)(uint32_t)(x >> 32), m);
    return ((uint64_t)hi << 32) | lo;
}

// Ascending bitonic sort of one value per lane across a wave.
__device__ inline uint64_t wave_sort_u64(uint64_t x) {
    const int lane = threadIdx.x & 63;
    for (int size = 2; size <= 64; size <<= 1)
        for (int stride = size >> 1; stride > 0; stride >>= 1) {
            const uint64_t y = shfl_xor_u64(x, stride);
            const bool up = (lane & size) == 0, lower = (lane & stride) == 0;
            x = (lower == up) ? (x < y ? x : y) : (x < y ? y : x);
        }
    return x;
}

// The kk smallest of the n (a power of two, distinct or kEmptyKey) keys, in
// ascending order, into out[0..kk): a radix select of the kk-th smallest
// distance key (four 8-bit digits of the high word), every key at or below
// it collected, and those sorted (in registers when they fit a wave).
// Returns the number collected (>= kk).  out has room for n keys.
__device__ int select_smallest(const uint64_t* keys, int n, int kk, uint64_t* out, uint32_t* hist,
                               uint32_t* sv) {
    const int tid = threadIdx.x, bs = blockDim.x;
    uint32_t prefix = 0, mask = 0, rank = (uint32_t)kk;
    for (int shift = 24; shift >= 0; shift -= 8) {
        for (int i = tid; i < 256; i += bs) hist[i] = 0;
        __syncthreads();
        for (int t = tid; t < n; t += bs) {
            const uint32_t dk = (uint32_t)(keys[t] >> 32);
            if ((dk & mask) == prefix) atomicAdd(&hist[(dk >> shift) & 255u], 1u);
        }
        __syncthreads();
        if (tid < 64) {
            uint32_t c[4], sum = 0;
#pragma unroll
            for (int e = 0; e < 4; ++e) {
                c[e] = hist[4 * tid + e];
                sum += c[e];
            }
            uint32_t inc = sum;
#pragma unroll
            for (int o = 1; o < 64; o <<= 1) {
                const uint32_t x = __shfl_up(inc, o);
                if (tid >= o) inc += x;
            }
            uint32_t run = inc - sum;
#pragma unroll
            for (int e = 0; e < 4; ++e) {
                if (run < rank && run + c[e] >= rank) {
                    sv[0] = 4u * tid + e;
                    sv[1] = rank - run;
                }
                run += c[e];
            }
        }
        __syncthreads();
        prefix |= sv[0] << shift;
        mask |= 0xFFu << shift;
        rank = sv[1];
        __syncthreads();
    }
    if (tid == 0) sv[2] = 0;
    __syncthreads();
    for (int t = tid; t < n; t += bs) {
        const uint64_t key = keys[t];
        if ((uint32_t)(key >> 32) <= prefix) out[atomicAdd(&sv[2], 1u)] = key;
    }
    __syncthreads();
    const int m = (int)sv[2];
    if (m <= 64) {
        if (tid < 64) {
            const uint64_t x = wave_sort_u64(tid < m ? out[tid] : kEmptyKey);
            if (tid < n) out[tid] = x;
        }
    } else {
        int m2 = 1;
        while (m2 < m) m2 <<= 1;
        for (int t = m + tid; t < m2; t += bs) out[t] = kEmptyKey;
        bitonic_sort_lds(out, m2);
    }
    __syncthreads();
    return m;
}

// Sum over the 16 lanes of a DPP row, in every lane of it: rotations by 8,
// 4, 2 and 1 within the row (row_ror), one VALU add each.
__device__ inline float row_sum16(float x) {
    x += __builtin_bit_cast(float, __builtin_amdgcn_update_dpp(0, __builtin_bit_cast(int, x), 0x128, 0xF, 0xF, false));
    x += __builtin_bit_cast(float, __builtin_amdgcn_update_dpp(0, __builtin_bit_cast(int, x), 0x124, 0xF, 0xF, false));
    x += __builtin_bit_cast(float, __builtin_amdgcn_update_dpp(0, __builtin_bit_cast(int, x), 0x122, 0xF, 0xF, false));
    x += __builtin_bit_cast(float, __builtin_amdgcn_update_dpp(0, __builtin_bit_cast(int, x), 0x121, 0xF, 0xF, false));
    return x;
}

// group_accumulate / group_distance with the distance kind fixed at compile
// time (the same operations).
template <int KIND>
__device__ inline void accumulate_k(float x, float y, float& a, float& b) {
    if (KIND == kNegCos || KIND == kAngular) {
        a = __builtin_fmaf(x, y, a);
        b = __builtin_fmaf(x, x, b);
    } else if (KIND == kGenJaccard) {
        a += fminf(x, y);
        b += fmaxf(x, y);
    } else {
        const float t = y - x;
        a = __builtin_fmaf(t, t, a);
    }
}

template <int KIND>
__device__ inline float distance_k(float a, float b, float q_norm) {
    if (KIND == kNegCos || KIND == kAngular) {
        const float c = a / (q_norm * __builtin_sqrtf(b));
        return KIND == kNegCos ? -c : 1.0f - c;
    } else if (KIND == kGenJaccard) {
        return 1.0f - a / b;
    }
    return a;
}

template <int KIND>
__global__ __launch_bounds__(256) void rerank_kernel(const float* __restrict__ data,
                                                     const float* __restrict__ queries, int d,
                                                     const uint64_t* __restrict__ keys,
                                                     int key_stride, const int32_t* __restrict__ cand,
                                                     int C, int Cp, int k, int kprime, double eps,
                                                     lsh_point_t* __restrict__ out,
                                                     const int32_t* __restrict__ qorder, int Q, int per,
                                                     const int32_t* __restrict__ rowmap) {
    extern __shared__ uint64_t sm[];
    uint64_t* k32 = sm;                                        // Cp
    uint64_t* k64 = sm + Cp;                                   // Cp
    int32_t* ids = reinterpret_cast<int32_t*>(sm + 2 * Cp);    // Cp
    int32_t* rows = ids + Cp;                                  // Cp: data row of each candidate
    __shared__ int s_ok;
    __shared__ uint32_t rs_hist[256];
    __shared__ uint32_t rs_sv[4];
    // With an order, XCD x (workgroups x, x + 8, ...) takes the x-th eighth
    // of it in sequence.
    int q = blockIdx.x;
    if (qorder) {
        const int pos = (int)(blockIdx.x & 7) * per + (int)(blockIdx.x >> 3);
        if (pos >= Q) return;
        q = qorder[pos];
    }
    const float* __restrict__ qv = queries + (size_t)q * d;

    float qn = 0.0f;
    double nq64 = 0.0;
    if (KIND == kNegCos || KIND == kAngular) {
        for (int j = 0; j < d; ++j) qn = __builtin_fmaf(qv[j], qv[j], qn);
        qn = __builtin_sqrtf(qn);
        nq64 = norm2_64(qv, d);
    }
    // Candidate IDs first (to LDS), so the row reads below do not wait on
    // the key reads.
    for (int t = threadIdx.x; t < Cp; t += blockDim.x) {
        int32_t id = -1;
        if (t < C) {
            if (cand) {
                id = cand[(size_t)q * C + t];
            } else {
                const uint64_t key = keys[(size_t)q * key_stride + t];
                id = key == kEmptyKey ? -1 : (int32_t)(uint32_t)key;
            }
        }
        ids[t] = id;
        rows[t] = id < 0 ? -1 : (rowmap ? rowmap[id] : id);
    }
    __syncthreads();
    // FP32 distances: a group of 16 lanes reads one candidate row with
    // coalesced float4 loads; the group's partial sums are added with DPP
    // rotations inside the 16-lane row, and lane u of the group keeps row u's
    // sums in LDS (k64 serves as scratch: two floats per candidate).  The
    // distances themselves are then taken one candidate per lane.
    const int gl = threadIdx.x & 15;
    const int ngroups = blockDim.x >> 4;
    const bool vec4 = (d & 3) == 0;
    float2* sums = reinterpret_cast<float2*>(k64);
    // Each group takes U rows per step so that U row reads are in flight.
    constexpr int U = 6;
    for (int t0 = threadIdx.x >> 4; t0 < Cp; t0 += ngroups * U) {
        const float* rp[U];
        float a[U], b[U];
#pragma unroll
        for (int u = 0; u < U; ++u) {
            const int t = t0 + u * ngroups;
            const int r = t < Cp ? rows[t] : -1;
            rp[u] = data + (size_t)(r < 0 ? 0 : r) * d;  // row 0 stands in; its sums are not used
            a[u] = 0.0f;
            b[u] = 0.0f;
        }
        if (vec4) {
            for (int j = gl * 4; j < d; j += 64) {
                const float4 yv = *reinterpret_cast<const float4*>(qv + j);
                float4 xv[U];
#pragma unroll
                for (int u = 0; u < U; ++u) xv[u] = *reinterpret_cast<const float4*>(rp[u] + j);
#pragma unroll
                for (int u = 0; u < U; ++u) {
                    accumulate_k<KIND>(xv[u].x, yv.x, a[u], b[u]);
                    accumulate_k<KIND>(xv[u].y, yv.y, a[u], b[u]);
                    accumulate_k<KIND>(xv[u].z, yv.z, a[u], b[u]);
                    accumulate_k<KIND>(xv[u].w, yv.w, a[u], b[u]);
                }
            }
        } else {
            for (int j = gl; j < d; j += 16) {
                const float y = qv[j];
                float xs[U];
#pragma unroll
                for (int u = 0; u < U; ++u) xs[u] = rp[u][j];
#pragma unroll
                for (int u = 0; u < U; ++u) accumulate_k<KIND>(xs[u], y, a[u], b[u]);
            }
        }
#pragma unroll
        for (int u = 0; u < U; ++u) {
            a[u] = row_sum16(a[u]);
            b[u] = row_sum16(b[u]);
        }
        float as = a[0], bs = b[0];
#pragma unroll
        for (int u = 1; u < U; ++u)
            if (gl == u) {
                as = a[u];
                bs = b[u];
            }
        const int t = t0 + gl * ngroups;
        if (gl < U && t < Cp) sums[t] = make_float2(as, bs);
    }
    __syncthreads();
    for (int t = threadIdx.x; t < Cp; t += blockDim.x) {
        const int32_t idt = ids[t];
        const float2 ab = sums[t];
        k32[t] = idt < 0 ? kEmptyKey : make_key(f32_key(distance_k<KIND>(ab.x, ab.y, qn)), (uint32_t)idt);
    }
    __syncthreads();
    // The kp best by FP32 distance, sorted, in k64 (k32 keeps every key).
    const int kp = min(kprime, Cp);
    select_smallest(k32, Cp, kp, k64, rs_hist, rs_sv);
    if (threadIdx.x == 0) {
        bool ok = true;
        if (kp < Cp) {
            const uint64_t kk = k64[k - 1], kK = k64[kp - 1];
            if (kK != kEmptyKey && kk != kEmptyKey) {
                const double vk = key_f32((uint32_t)(kk >> 32));
                const double vK = key_f32((uint32_t)(kK >> 32));
                if (KIND == kEuclid)
                    ok = vK * (1.0 - eps) > vk * (1.0 + eps);
                else
                    ok = vK - vk > 2.0 * eps;
            }
        }
        s_ok = ok;
    }
    __syncthreads();
    int n = kp;
    if (s_ok) {
        // FP64 for the kp best, in place.
        for (int t = threadIdx.x; t < kp; t += blockDim.x) {
            const uint64_t key = k64[t];
            if (key == kEmptyKey) {
                ids[t] = 0x7FFFFFFF;
            } else {
                const int32_t id = (int32_t)(uint32_t)key;
                k64[t] = f64_key(dist64_row(KIND, qv, data + (size_t)(rowmap ? rowmap[id] : id) * d, d, nq64));
                ids[t] = id;
            }
        }
    } else {
        // FP64 for every candidate.
        for (int t = threadIdx.x; t < Cp; t += blockDim.x) {
            const uint64_t key = k32[t];
            if (key == kEmptyKey) {
                k64[t] = kEmptyKey;
                ids[t] = 0x7FFFFFFF;
            } else {
                const int32_t id = (int32_t)(uint32_t)key;
                k64[t] = f64_key(dist64_row(KIND, qv, data + (size_t)(rowmap ? rowmap[id] : id) * d, d, nq64));
                ids[t] = id;
            }
        }
        n = Cp;
    }
    bitonic_sort_pairs(k64, ids, n);
    const float pad = KIND == kNegCos ? 2.0f : kPadDistance;
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
                   lsh_point_t* out, hipStream_t s, const int32_t* qorder, const int32_t* rowmap) {
    if (Q <= 0) return;
    const int Cp = pow2_ceil(std::max(C, 2));
    LSH_REQUIRE(Cp <= kMaxTopK, LSH_ERR_INVALID, "candidate set larger than 4096");
    const int kp = pow2_ceil(std::max(kprime, k));
    const size_t lds = (size_t)Cp * (2 * sizeof(uint64_t) + 2 * sizeof(int32_t));
    const int per = (Q + 7) / 8;
    auto go = [&](auto kernel) {
        set_dynamic_lds((const void*)kernel, lds);
        hipLaunchKernelGGL(kernel, dim3(qorder ? 8 * per : Q), dim3(256), lds, s, data, queries, d, keys, key_stride,
                           cand, C, Cp, k, kp, eps, out, qorder, Q, per, rowmap);
    };
    switch (kind) {
        case kNegCos: go(rerank_kernel<kNegCos>); break;
        case kAngular: go(rerank_kernel<kAngular>); break;
        case kGenJaccard: go(rerank_kernel<kGenJaccard>); break;
        case kEuclid: go(rerank_kernel<kEuclid>); break;
        default: fail(LSH_ERR_INTERNAL, "rerank: unknown distance kind");
    }
    check_launch("rerank_kernel");
}

}  // namespace lshgpu
