// scan_exact_mfma.hip -- the exact angular scan on the FP32 matrix cores.
//
// The same rounds as exact_scan_kernel (each launch scans points [p0, p1)
// for every query and appends the pairs that beat the query's threshold),
// with the dot products on v_mfma_f32_32x32x2f32: a wave multiplies a tile of
// 32 points (the A operand, in registers) with its group of 32 queries (the
// B operand, held in LDS for the whole launch), 2 dimensions per instruction.
// Points and queries are stored in operand order -- tile t, step s, lane l
// holds dimension 2s + l/32 of row 32t + l%32, zero past d (padded to a
// multiple of 4 dimensions) -- so each step's operand is one coalesced
// 256-byte load; the next tile's operands load while the current one is
// multiplied.  Accumulator register r of lane l holds the
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
#include <cmath>

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
    int kind;             // kAngular or kNegCos
    int64_t p0, p1;       // points scanned
    int64_t t0, t1;       // their tiles
};

// One tile: 32 points x this wave's 32 queries.  A/X are the tile's operands
// and inverse norms (loaded a tile ahead).  Distances are tested against the
// query's FP32 threshold first; only the pairs that pass build their keys
// (distance, point) for the exact comparison and are appended.
template <int SP>
__device__ __forceinline__ void exact_mfma_tile(const float (&A)[SP], const float (&X)[16], const float* lb,
                                                uint32_t boff,
                                                int64_t t, bool angular, uint32_t vmask_full, float qi,
                                                float thr_f, uint64_t thr, int q, size_t row0, int kh,
                                                const ExactMfmaArgs& a, const TopKView& v) {
    // The B offset is made opaque here so that the compiler reads B from LDS
    // per tile instead of hoisting all of it into registers.
    asm volatile("" : "+v"(boff));
    const float* bq = lb + boff;
    f32x16 acc = (f32x16){};
#pragma unroll
    for (int s = 0; s < SP; ++s) acc = __builtin_amdgcn_mfma_f32_32x32x2f32(A[s], bq[s * 64], acc, 0, 0, 0);
    // Rows of the tile inside [p0, p1): all of them except in the first and
    // last tile of the range.
    uint32_t vm = vmask_full;
    const int64_t base = t * 32;
    if (vm && (base < a.p0 || base + 32 > a.p1)) {
        vm = 0;
#pragma unroll
        for (int r = 0; r < 16; ++r) {
            const int64_t pt = base + (r >> 2) * 8 + kh * 4 + (r & 3);
            vm |= (pt >= a.p0 && pt < a.p1 ? 1u : 0u) << r;
        }
    }
    uint32_t pass = 0;
    if (angular) {
#pragma unroll
        for (int r = 0; r < 16; ++r) {
            acc[r] = __builtin_fmaf(-(acc[r] * X[r]), qi, 1.0f);
            pass |= (!(acc[r] > thr_f) ? 1u : 0u) << r;
        }
    } else {
#pragma unroll
        for (int r = 0; r < 16; ++r) {
            acc[r] = -((acc[r] * X[r]) * qi);
            pass |= (!(acc[r] > thr_f) ? 1u : 0u) << r;
        }
    }
    pass &= vm;
    if (pass) {
        uint32_t fine = 0;
#pragma unroll
        for (int r = 0; r < 16; ++r) {
            const uint32_t pt = (uint32_t)(base + (r >> 2) * 8 + kh * 4 + (r & 3));
            fine |= (((pass >> r) & 1u) && make_key(f32_key(acc[r]), pt) < thr ? 1u : 0u) << r;
        }
        if (fine) {
            uint32_t slot = atomicAdd(&v.cnt[q], (uint32_t)__popc(fine));
#pragma unroll
            for (int r = 0; r < 16; ++r)
                if ((fine >> r) & 1u) {
                    const uint32_t pt = (uint32_t)(base + (r >> 2) * 8 + kh * 4 + (r & 3));
                    if (slot < (uint32_t)v.newcap) v.keys[row0 + slot] = make_key(f32_key(acc[r]), pt);
                    ++slot;
                }
        }
    }
}

template <int SP>
__device__ __forceinline__ void exact_mfma_load(float (&A)[SP], float (&X)[16], int64_t t, int kh, int lane,
                                                const ExactMfmaArgs& a) {
    // The inverse norms go first: the tile's epilogue waits for them while
    // the following tile's operands are still in flight.
#pragma unroll
    for (int r4 = 0; r4 < 4; ++r4) {
        const float4 xv = *reinterpret_cast<const float4*>(a.dinv + (size_t)t * 32 + r4 * 8 + kh * 4);
        X[4 * r4 + 0] = xv.x;
        X[4 * r4 + 1] = xv.y;
        X[4 * r4 + 2] = xv.z;
        X[4 * r4 + 3] = xv.w;
    }
    const float* __restrict__ src = a.dT + (size_t)t * SP * 64 + lane;
#pragma unroll
    for (int s = 0; s < SP; ++s) A[s] = src[s * 64];
}

template <int SP>
__global__ __launch_bounds__(kTPB, 2) void exact_mfma_kernel(ExactMfmaArgs a, TopKView v) {
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63, kh = lane >> 5, l32 = lane & 31;
    const int g = blockIdx.y * 4 + wave;
    if (g >= a.groups) return;  // no barriers below: idle waves leave
    const int q = g * 32 + l32;
    const bool qvalid = q < a.Q;
    // The wave's B operands live in LDS (registers hold two tiles of A): each
    // lane reads back only what it wrote itself, so no barrier is needed.
    __shared__ float lb[4 * SP * 64];
    const uint32_t boff = (uint32_t)(wave * SP * 64 + lane);
    const float* __restrict__ qsrc = a.qB + (size_t)g * SP * 64 + lane;
#pragma unroll
    for (int s = 0; s < SP; ++s) lb[boff + s * 64] = qsrc[s * 64];
    const uint64_t thr = qvalid ? v.thr[q] : 0ull;
    const float thr_f = key_f32((uint32_t)(thr >> 32));
    const float qi = qvalid ? a.qinv[q] : 0.0f;
    const size_t row0 = (size_t)(qvalid ? q : 0) * v.stride + v.K;
    const uint32_t vfull = qvalid ? 0xFFFFu : 0u;
    const bool angular = a.kind == kAngular;

    // Two operand buffers in turn: the next tile loads while this one is
    // multiplied.
    float a0[SP], a1[SP], x0[16], x1[16];
    const int64_t G = gridDim.x;
    int64_t t = a.t0 + blockIdx.x;
    if (t < a.t1) exact_mfma_load<SP>(a0, x0, t, kh, lane, a);
    while (t < a.t1) {
        const int64_t tn = t + G;
        if (tn < a.t1) exact_mfma_load<SP>(a1, x1, tn, kh, lane, a);
        exact_mfma_tile<SP>(a0, x0, lb, boff, t, angular, vfull, qi, thr_f, thr, q, row0, kh, a, v);
        if (tn >= a.t1) break;
        t = tn + G;
        if (t < a.t1) exact_mfma_load<SP>(a0, x0, t, kh, lane, a);
        exact_mfma_tile<SP>(a1, x1, lb, boff, tn, angular, vfull, qi, thr_f, thr, q, row0, kh, a, v);
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

void launch_sp(const ExactMfmaArgs& a, int SP, dim3 grid, const TopKView& v, hipStream_t s) {
    switch (SP) {
#define LSHGPU_EM_CASE(n) \
    case n: hipLaunchKernelGGL((exact_mfma_kernel<n>), grid, dim3(kTPB), 0, s, a, v); break;
        LSHGPU_EM_CASE(2) LSHGPU_EM_CASE(4) LSHGPU_EM_CASE(6) LSHGPU_EM_CASE(8)
        LSHGPU_EM_CASE(10) LSHGPU_EM_CASE(12) LSHGPU_EM_CASE(14) LSHGPU_EM_CASE(16)
        LSHGPU_EM_CASE(18) LSHGPU_EM_CASE(20) LSHGPU_EM_CASE(22) LSHGPU_EM_CASE(24)
        LSHGPU_EM_CASE(26) LSHGPU_EM_CASE(28) LSHGPU_EM_CASE(30) LSHGPU_EM_CASE(32)
        LSHGPU_EM_CASE(34) LSHGPU_EM_CASE(36) LSHGPU_EM_CASE(38) LSHGPU_EM_CASE(40)
        LSHGPU_EM_CASE(42) LSHGPU_EM_CASE(44) LSHGPU_EM_CASE(46) LSHGPU_EM_CASE(48)
        LSHGPU_EM_CASE(50) LSHGPU_EM_CASE(52) LSHGPU_EM_CASE(54) LSHGPU_EM_CASE(56)
        LSHGPU_EM_CASE(58) LSHGPU_EM_CASE(60) LSHGPU_EM_CASE(62) LSHGPU_EM_CASE(64)
#undef LSHGPU_EM_CASE
        default: fail(LSH_ERR_INTERNAL, "exact matrix-core scan: unsupported padded dimension");
    }
}

// Workgroups of exact_mfma_kernel resident at once: two a CU (its
// registers allow two waves a SIMD).
int64_t wg_slots() {
    static const int64_t slots = [] {
        int dev = 0, cus = 0;
        if (hipGetDevice(&dev) != hipSuccess ||
            hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess || cus <= 0)
            cus = 256;
        return (int64_t)2 * cus;
    }();
    return slots;
}

}  // namespace

// Steps of 2 dimensions, padded to an even count (4 dimensions).
int exact_mfma_steps(int d) { return (d + 3) / 4 * 2; }

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
    const int64_t gy = (a.groups + 3) / 4;
    const int64_t tiles = a.t1 - a.t0;
    // Each workgroup keeps its query block and takes tiles grid-stride, so
    // every workgroup of a launch does about the same work: the launch lasts
    // as many workgroup lifetimes as it has waves of workgroups over the
    // machine's slots.  Pick the tile split gx whose grid fills its last wave
    // best, times the balance of tiles per workgroup.
    const int64_t slots = wg_slots();
    int64_t gx = 1;
    double best = -1.0;
    const int64_t gx_max = std::max<int64_t>(1, std::min<int64_t>(tiles, (8 * slots + gy - 1) / gy));
    for (int64_t x = 1; x <= gx_max; ++x) {
        const int64_t wgs = x * gy;
        const double fill = (double)wgs / (double)(slots * ((wgs + slots - 1) / slots));
        const double per = (double)tiles / (double)x;
        const double eff = fill * per / std::ceil(per);
        if (eff > best + 1e-9) {
            best = eff;
            gx = x;
        }
    }
    a.kind = kind;
    launch_sp(a, SP, dim3((unsigned)gx, (unsigned)gy), v, s);
    check_launch("exact_mfma_kernel");
}

double mfma_error_bound(int d) { return (5.0 * d + 32.0) / 16777216.0; }

}  // namespace lshgpu
