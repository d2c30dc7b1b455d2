// scan_exact.hip -- exact-distance scan (FP32) and its FP64 finalisation.
//
// Replaces the reference's warp-per-query scan (memOptimizedLinearScan.cu:20-106,
// optimizedLinearScan.cu:22-91) with a tiled scan: each lane owns one data
// point and keeps a chunk of its row in VGPRs; a wave walks a group of 32
// queries whose values are wave-uniform, so they arrive as scalar-register
// operands of the fused multiply-adds.  Query groups vary fastest across the
// grid, so the blocks in flight share their data rows through L2.  Each
// (query, point) distance is compared with the query's current threshold
// (the K-th best key) and the few that pass are appended with one atomic per
// wave; select.hip folds them into the best list between rounds.
#include "device_common.hpp"

namespace lshgpu {

namespace {

constexpr int QG = 32;   // queries per group
constexpr int DC = 32;   // dimensions per register chunk
constexpr int TPB = 256;

template <int KIND, bool VEC4>
__global__ __launch_bounds__(TPB) void exact_scan_kernel(ExactScanArgs a, TopKView v) {
    const int g = blockIdx.x;
    const int qbase = g * QG;
    const int nq = min(QG, a.Q - qbase);
    const int64_t i = a.p0 + (int64_t)blockIdx.y * TPB + threadIdx.x;
    const bool active = i < a.p1;
    const int d = a.d;
    const int64_t ri = active ? i : a.p0;
    const float* __restrict__ row = a.data + (size_t)ri * d;
    const float* __restrict__ qg = a.qT + (size_t)g * d * QG;

    float acc[QG];
    float acc2[QG];
#pragma unroll
    for (int q = 0; q < QG; ++q) {
        acc[q] = 0.0f;
        acc2[q] = 0.0f;
    }

    for (int j0 = 0; j0 < d; j0 += DC) {
        const int jn = min(DC, d - j0);
        float xr[DC];
        if (VEC4 && jn == DC) {
#pragma unroll
            for (int c = 0; c < DC / 4; ++c) {
                const float4 t = *reinterpret_cast<const float4*>(row + j0 + 4 * c);
                xr[4 * c + 0] = t.x;
                xr[4 * c + 1] = t.y;
                xr[4 * c + 2] = t.z;
                xr[4 * c + 3] = t.w;
            }
        } else {
#pragma unroll
            for (int jj = 0; jj < DC; ++jj) xr[jj] = jj < jn ? row[j0 + jj] : 0.0f;
        }
#pragma unroll
        for (int jj = 0; jj < DC; ++jj) {
            if (jj < jn) {
                const float* __restrict__ qv = qg + (size_t)(j0 + jj) * QG;
                const float x = xr[jj];
#pragma unroll
                for (int q = 0; q < QG; ++q) {
                    const float y = qv[q];
                    if constexpr (KIND == kNegCos || KIND == kAngular) {
                        acc[q] = __builtin_fmaf(x, y, acc[q]);
                    } else if constexpr (KIND == kGenJaccard) {
                        acc[q] += fminf(x, y);
                        acc2[q] += fmaxf(x, y);
                    } else if constexpr (KIND == kSqL2) {
                        // distanceFunctions.cuh:87-99: pow(q - x, 2) then a
                        // separate add (the square is exact in the promoted
                        // type and rounds to fl(t*t)); the JL candidate set is
                        // ranked by these FP32 sums, so they are not fused.
                        const float t = y - x;
                        const float sq = t * t;
                        acc[q] = acc[q] + sq;
                    } else {
                        // Euclidean ranking only (an FP64 recheck follows).
                        const float t = y - x;
                        acc[q] = __builtin_fmaf(t, t, acc[q]);
                    }
                }
            }
        }
    }

    float xinv = 0.0f;
    if constexpr (KIND == kNegCos || KIND == kAngular) xinv = a.data_inv[ri];

#pragma unroll
    for (int q = 0; q < QG; ++q) {
        if (q < nq) {
            const int qrow = qbase + q;
            float dist;
            if constexpr (KIND == kAngular) {
                const float c = acc[q] * xinv;
                dist = __builtin_fmaf(-c, a.q_inv[qrow], 1.0f);
            } else if constexpr (KIND == kNegCos) {
                dist = -((acc[q] * xinv) * a.q_inv[qrow]);
            } else if constexpr (KIND == kGenJaccard) {
                dist = 1.0f - acc[q] / acc2[q];
            } else {
                dist = acc[q];  // squared L2 (kEuclid ranks by the square)
            }
            const uint64_t thr = v.thr[qrow];
            const float thr_f = key_f32((uint32_t)(thr >> 32));
            const bool coarse = active && !(dist > thr_f);
            if (__ballot(coarse)) {
                const uint64_t key = make_key(f32_key(dist), (uint32_t)i);
                append_key(v, qrow, coarse && key < thr, key);
            }
        }
    }
}

template <int KIND>
void launch_kind(const ExactScanArgs& a, const TopKView& v, hipStream_t s) {
    const int groups = (a.Q + QG - 1) / QG;
    const int64_t npts = a.p1 - a.p0;
    const bool vec4 = (a.d % 4) == 0 && (reinterpret_cast<uintptr_t>(a.data) % 16) == 0;
    // grid.y is limited to 65535 blocks; split long ranges.
    const int64_t max_pts = (int64_t)65535 * TPB;
    for (int64_t off = 0; off < npts; off += max_pts) {
        ExactScanArgs b = a;
        b.p0 = a.p0 + off;
        b.p1 = std::min(a.p1, b.p0 + max_pts);
        const unsigned by = (unsigned)((b.p1 - b.p0 + TPB - 1) / TPB);
        if (vec4)
            hipLaunchKernelGGL((exact_scan_kernel<KIND, true>), dim3(groups, by), dim3(TPB), 0, s, b, v);
        else
            hipLaunchKernelGGL((exact_scan_kernel<KIND, false>), dim3(groups, by), dim3(TPB), 0, s, b, v);
    }
    check_launch("exact_scan_kernel");
}

// ---------------------------------------------------------------- helpers

__global__ void transpose_queries_kernel(const float* __restrict__ q, int Q, int d,
                                         float* __restrict__ qT) {
    // qT[g][j][c] = q[(g*QG + c)*d + j], zero for rows past Q.
    const int groups = (Q + QG - 1) / QG;
    const size_t total = (size_t)groups * d * QG;
    for (size_t t = (size_t)blockIdx.x * blockDim.x + threadIdx.x; t < total;
         t += (size_t)gridDim.x * blockDim.x) {
        const int c = (int)(t % QG);
        const int j = (int)((t / QG) % d);
        const int g = (int)(t / ((size_t)QG * d));
        const int r = g * QG + c;
        qT[t] = r < Q ? q[(size_t)r * d + j] : 0.0f;
    }
}

__global__ void row_inv_norms_kernel(const float* __restrict__ x, int64_t n, int d,
                                     float* __restrict__ inv) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const float* r = x + (size_t)i * d;
    float m = 0.0f;
    for (int j = 0; j < d; ++j) m = __builtin_fmaf(r[j], r[j], m);
    inv[i] = 1.0f / __builtin_sqrtf(m);
}

}  // namespace

namespace {

__global__ __launch_bounds__(256) void exact_finalize_kernel(const float* __restrict__ data,
                                                             const float* __restrict__ queries, int d,
                                                             int kind, TopKView v, int k, double eps,
                                                             lsh_point_t* __restrict__ out,
                                                             uint32_t* __restrict__ flags) {
    extern __shared__ uint64_t sm[];
    const int K = v.K;
    uint64_t* dk = sm;
    int32_t* ids = reinterpret_cast<int32_t*>(sm + K);
    const int q = blockIdx.x;
    const uint64_t* row = v.keys + (size_t)q * v.stride;
    const float* qv = queries + (size_t)q * d;
    for (int t = threadIdx.x; t < K; t += blockDim.x) {
        const uint64_t key = row[t];
        if (key == kEmptyKey) {
            dk[t] = kEmptyKey;
            ids[t] = 0x7FFFFFFF;
        } else {
            const int32_t id = (int32_t)(uint32_t)key;
            dk[t] = f64_key(dist64_dev(kind, qv, data + (size_t)id * d, d));
            ids[t] = id;
        }
    }
    bitonic_sort_pairs(dk, ids, K);
    if (threadIdx.x == 0) {
        bool ok = true;
        const uint64_t kk = row[k - 1], kK = row[K - 1];
        if (kK != kEmptyKey && kk != kEmptyKey) {
            const double vk = key_f32((uint32_t)(kk >> 32));
            const double vK = key_f32((uint32_t)(kK >> 32));
            if (kind == kEuclid)
                ok = vK * (1.0 - eps) > vk * (1.0 + eps);
            else
                ok = vK - vk > 2.0 * eps;
        }
        flags[q] = ok ? 0u : 1u;
    }
    const float pad = kind == kNegCos ? 2.0f : kPadDistance;
    for (int t = threadIdx.x; t < k; t += blockDim.x) {
        lsh_point_t p;
        if (t < K && ids[t] != 0x7FFFFFFF) {
            p.ID = ids[t];
            p.distance = (float)key_f64(dk[t]);
        } else {
            p.ID = -1;
            p.distance = pad;
        }
        out[(size_t)q * k + t] = p;
    }
}

}  // namespace

int exact_scan_query_group() { return QG; }

void launch_exact_scan(const ExactScanArgs& a, const TopKView& v, hipStream_t s) {
    if (a.p1 <= a.p0 || a.Q <= 0) return;
    switch (a.kind) {
        case kNegCos: launch_kind<kNegCos>(a, v, s); break;
        case kAngular: launch_kind<kAngular>(a, v, s); break;
        case kGenJaccard: launch_kind<kGenJaccard>(a, v, s); break;
        case kEuclid: launch_kind<kEuclid>(a, v, s); break;
        case kSqL2: launch_kind<kSqL2>(a, v, s); break;
        default: fail(LSH_ERR_INVALID, "unknown distance kind");
    }
}

void launch_transpose_queries(const float* q, int Q, int d, float* qT, hipStream_t s) {
    const int groups = (Q + QG - 1) / QG;
    const size_t total = (size_t)groups * d * QG;
    const unsigned blocks = (unsigned)std::min<size_t>((total + 255) / 256, 65535);
    hipLaunchKernelGGL(transpose_queries_kernel, dim3(std::max(blocks, 1u)), dim3(256), 0, s, q, Q, d, qT);
    check_launch("transpose_queries_kernel");
}

void launch_row_inv_norms(const float* x, int64_t n, int d, float* inv, hipStream_t s) {
    if (n <= 0) return;
    hipLaunchKernelGGL(row_inv_norms_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, x, n, d, inv);
    check_launch("row_inv_norms_kernel");
}

void launch_exact_finalize(const float* data, const float* queries, int d, int kind,
                           const TopKView& v, int k, double eps, lsh_point_t* out,
                           uint32_t* flags, hipStream_t s) {
    const size_t lds = (size_t)v.K * (sizeof(uint64_t) + sizeof(int32_t));
    set_dynamic_lds((const void*)exact_finalize_kernel, lds);
    hipLaunchKernelGGL(exact_finalize_kernel, dim3(v.Q), dim3(256), lds, s, data, queries, d, kind, v,
                       k, eps, out, flags);
    check_launch("exact_finalize_kernel");
}

double fp32_error_bound(int kind, int d) {
    const double u = 1.0 / 16777216.0;  // 2^-24
    switch (kind) {
        case kNegCos:
        case kAngular: return (2.0 * d + 16.0) * u;
        case kGenJaccard: return (2.0 * d + 8.0) * u;
        case kEuclid: return (d + 8.0) * u;
        default: return 0.0;
    }
}

}  // namespace lshgpu
