// scan_exact_mfma.hip -- the exact angular scan on the FP32 matrix cores.
//
// The same rounds as exact_scan_kernel (each launch scans points [p0, p1)
// for every query and appends the pairs that beat the query's threshold),
// with the dot products on v_mfma_f32_32x32x2f32: a wave multiplies a tile of
// 32 points (the A operand) with its group of 32 queries (the B operand,
// held in registers for the whole launch), 2 dimensions per instruction.
// Points and queries are stored in operand order -- tile t, step s, lane l
// holds dimension 2s + l/32 of row 32t + l%32, zero past d -- so each step's
// operand is one coalesced 256-byte load; the next tile's operands load while
// the current one is multiplied.  Accumulator register r of lane l holds the
// dot product of point 32t + 8(r/4) + 4(l/32) + r%4 with query l%32 of the
// group.  Four waves of a workgroup take four query groups over the same
// point tiles, so a tile's loads are shared through L1/L2.
//
// The products and sums of the matrix cores are not the sequential FMA chain
// of the vector kernel; the FP64 finalisation takes the error bound
// (5d + 32) 2^-24 for these distances -- each product and each addition
// within 2^-23 relative, in any order (4d 2^-24 of ||q|| ||x||), plus the FP32
// inverse norms and the final products (d + 8) 2^-24 -- so the returned
// results are the same exact (FP64, ID) top-k.  Should the matrix cores
// flush denormal products, the absolute error d 2^-126 that adds stays under
// 2^-24 ||q|| ||x|| while ||q|| ||x|| >= 2^-100 (the FP32 norms of the vector
// kernel lose their range below about 2^-63 per row already).
#include "device_common.hpp"

namespace lshgpu {

namespace {

typedef float f32x16 __attribute__((ext_vector_type(16)));

constexpr int kTPB = 256;

struct ExactMfmaArgs {
    const float* dT;      // [tiles][SP][64]
    const float* dinv;    // [tiles*32] 1/||x|| (zero for padding rows)
    const float* qB;      // [groups][SP][64]
    const float* qinv;    // [Q]
    int Q, groups;
    int64_t p0, p1;       // points scanned
    int64_t t0, t1;       // their tiles
};

template <int KIND, int SP>
__global__ __launch_bounds__(kTPB, 2) void exact_mfma_kernel(ExactMfmaArgs a, TopKView v) {
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63, kh = lane >> 5, l32 = lane & 31;
    const int g = blockIdx.y * 4 + wave;
    if (g >= a.groups) return;  // no barriers below: idle waves leave
    const int q = g * 32 + l32;
    const bool qvalid = q < a.Q;
    float bq[SP];
    const float* __restrict__ qsrc = a.qB + (size_t)g * SP * 64 + lane;
#pragma unroll
    for (int s = 0; s < SP; ++s) bq[s] = qsrc[s * 64];
    const uint64_t thr = qvalid ? v.thr[q] : 0ull;
    const float thr_f = key_f32((uint32_t)(thr >> 32));
    const float qi = qvalid ? a.qinv[q] : 0.0f;
    const size_t row0 = (size_t)(qvalid ? q : 0) * v.stride + v.K;

    float an[SP];
    int64_t t = a.t0 + blockIdx.x;
    if (t < a.t1) {
        const float* __restrict__ src = a.dT + (size_t)t * SP * 64 + lane;
#pragma unroll
        for (int s = 0; s < SP; ++s) an[s] = src[s * 64];
    }
    for (; t < a.t1; t += gridDim.x) {
        float ac[SP];
#pragma unroll
        for (int s = 0; s < SP; ++s) ac[s] = an[s];
        const int64_t tn = t + gridDim.x;
        if (tn < a.t1) {
            const float* __restrict__ src = a.dT + (size_t)tn * SP * 64 + lane;
#pragma unroll
            for (int s = 0; s < SP; ++s) an[s] = src[s * 64];
        }
        f32x16 acc = (f32x16){};
#pragma unroll
        for (int s = 0; s < SP; ++s) acc = __builtin_amdgcn_mfma_f32_32x32x2f32(ac[s], bq[s], acc, 0, 0, 0);
        // Inverse norms of this lane's 16 points, four consecutive rows at a time.
        float xinv[16];
#pragma unroll
        for (int r4 = 0; r4 < 4; ++r4) {
            const float4 xv = *reinterpret_cast<const float4*>(a.dinv + (size_t)t * 32 + r4 * 8 + kh * 4);
            xinv[4 * r4 + 0] = xv.x;
            xinv[4 * r4 + 1] = xv.y;
            xinv[4 * r4 + 2] = xv.z;
            xinv[4 * r4 + 3] = xv.w;
        }
        uint32_t pass = 0;
        uint64_t keys[16];
#pragma unroll
        for (int r = 0; r < 16; ++r) {
            const int64_t pt = t * 32 + (r >> 2) * 8 + kh * 4 + (r & 3);
            const float c = acc[r] * xinv[r];
            const float dist = KIND == kAngular ? __builtin_fmaf(-c, qi, 1.0f) : -(c * qi);
            keys[r] = make_key(f32_key(dist), (uint32_t)pt);
            const bool ok = qvalid && pt >= a.p0 && pt < a.p1 && !(dist > thr_f) && keys[r] < thr;
            pass |= (ok ? 1u : 0u) << r;
        }
        if (pass) {
            uint32_t slot = atomicAdd(&v.cnt[q], (uint32_t)__popc(pass));
#pragma unroll
            for (int r = 0; r < 16; ++r)
                if ((pass >> r) & 1u) {
                    if (slot < (uint32_t)v.newcap) v.keys[row0 + slot] = keys[r];
                    ++slot;
                }
        }
    }
}

__global__ void tile_rows_kernel(const float* __restrict__ x, int64_t n, int d, int SP, int64_t tiles,
                                 float* __restrict__ out) {
    const size_t total = (size_t)tiles * SP * 64;
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < total; i += (size_t)gridDim.x * blockDim.x) {
        const int l = (int)(i & 63);
        const size_t ts = i >> 6;
        const int s = (int)(ts % SP);
        const int64_t t = (int64_t)(ts / SP);
        const int64_t r = t * 32 + (l & 31);
        const int j = 2 * s + (l >> 5);
        out[i] = (r < n && j < d) ? x[(size_t)r * d + j] : 0.0f;
    }
}

template <int KIND>
void launch_sp(const ExactMfmaArgs& a, int SP, dim3 grid, const TopKView& v, hipStream_t s) {
    switch (SP) {
#define LSHGPU_EM_CASE(n) \
    case n: hipLaunchKernelGGL((exact_mfma_kernel<KIND, n>), grid, dim3(kTPB), 0, s, a, v); break;
        LSHGPU_EM_CASE(4) LSHGPU_EM_CASE(8) LSHGPU_EM_CASE(12) LSHGPU_EM_CASE(16)
        LSHGPU_EM_CASE(20) LSHGPU_EM_CASE(24) LSHGPU_EM_CASE(28) LSHGPU_EM_CASE(32)
        LSHGPU_EM_CASE(36) LSHGPU_EM_CASE(40) LSHGPU_EM_CASE(44) LSHGPU_EM_CASE(48)
        LSHGPU_EM_CASE(52) LSHGPU_EM_CASE(56) LSHGPU_EM_CASE(60) LSHGPU_EM_CASE(64)
#undef LSHGPU_EM_CASE
        default: fail(LSH_ERR_INTERNAL, "exact matrix-core scan: unsupported padded dimension");
    }
}

}  // namespace

// Steps of 2 dimensions, padded to a multiple of 4 (8 dimensions).
int exact_mfma_steps(int d) { return (d + 7) / 8 * 4; }

bool exact_mfma_supported(int kind, int d) { return (kind == kAngular || kind == kNegCos) && d >= 1 && d <= 128; }

void launch_tile_rows(const float* x, int64_t n, int d, float* out, hipStream_t s) {
    const int SP = exact_mfma_steps(d);
    const int64_t tiles = (n + 31) / 32;
    const size_t total = (size_t)tiles * SP * 64;
    if (total == 0) return;
    const unsigned blocks = (unsigned)std::min<size_t>((total + 255) / 256, 65535 * 16);
    hipLaunchKernelGGL(tile_rows_kernel, dim3(blocks), dim3(256), 0, s, x, n, d, SP, tiles, out);
    check_launch("tile_rows_kernel");
}

void launch_exact_mfma(const float* dT, const float* dinv, int d, const float* qB, const float* qinv, int Q,
                       int64_t p0, int64_t p1, int kind, const TopKView& v, hipStream_t s) {
    if (p1 <= p0 || Q <= 0) return;
    ExactMfmaArgs a{};
    a.dT = dT;
    a.dinv = dinv;
    a.qB = qB;
    a.qinv = qinv;
    a.Q = Q;
    a.groups = (Q + 31) / 32;
    a.p0 = p0;
    a.p1 = p1;
    a.t0 = p0 / 32;
    a.t1 = (p1 + 31) / 32;
    const int SP = exact_mfma_steps(d);
    const unsigned gy = (unsigned)((a.groups + 3) / 4);
    const int64_t tiles = a.t1 - a.t0;
    // Tiles are taken grid-stride; about 2048 workgroups fill the 256 CUs.
    const unsigned gx = (unsigned)std::max<int64_t>(1, std::min<int64_t>(tiles, (2048 + gy - 1) / gy));
    if (kind == kAngular)
        launch_sp<kAngular>(a, SP, dim3(gx, gy), v, s);
    else
        launch_sp<kNegCos>(a, SP, dim3(gx, gy), v, s);
    check_launch("exact_mfma_kernel");
}

double mfma_error_bound(int d) { return (5.0 * d + 32.0) / 16777216.0; }

}  // namespace lshgpu
