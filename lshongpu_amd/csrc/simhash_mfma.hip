// simhash_mfma.hip -- SimHash sketch bits on the bf16 matrix cores, with the
// bits whose projection is too close to zero recomputed in FP32.
//
// The bit of hyperplane j for row x is (y32 >= 0), y32 the sequential FP32
// fused multiply-add chain of simHash.cu:22-47 over dimensions in index
// order.  Here y' = x.p is formed from bfloat16 splits of both operands,
// x = x_hi + x_lo + dx and p = p_hi + p_lo + dp with |dx| <= 2^-16 |x| and
// |dp| <= 2^-16 |p|, as x_hi.p_hi + x_lo.p_hi + x_hi.p_lo: three
// v_mfma_f32_32x32x16_bf16 per 16 dimensions with exact products and FP32
// sums.  Per term the dropped parts (x_lo.p_lo, dx.p, x.dp and smaller) are
// at most 4 * 2^-16 |x_i p_i|; the matrix-core sums of 3*dpad products add at
// most 3*dpad * 2^-23 of sum |x_i p_i| in any order, and the FP32 chain's own
// error is at most d * 2^-24 of it.  With sum |x_i p_i| <= ||x|| ||p||:
//   |y' - y32| <= E = ||x|| ||p_j|| (4 * 2^-16 + (3 dpad + d + 8) 2^-23).
// When |y'| > E, y32 has the sign of y' and is not zero, so the bit is
// (y' >= 0); otherwise (and for NaN) the bit is written as (y' >= 0) and the
// (row, hyperplane) pair is listed, and a second kernel recomputes y32 of
// every listed pair by the FP32 chain, one pair a thread, and flips the bits
// that differ (a list past its capacity is rechecked in place instead).  The
// sketch is then bit-identical to the FP32 definition.  At d = 100 with unit
// rows and N(0,1) hyperplanes E is about 1e-3 and fewer than 0.1% of the
// bits are listed.
//
// Layout: a workgroup of 4 waves takes 128 rows (32 per wave, the A
// operand: lane l holds row l%32, dimensions 16s + 8(l/32) + [0,8) of step s,
// in registers for every word) and a range of words.  The hyperplanes of word
// w are the B operand, stored once per index in operand order (lane l: plane
// 32w + l%32, the same dimensions), so each lane's 16-byte loads of a step
// are one coalesced kilobyte per wave.  Accumulator register r of lane l
// holds plane 32w + l%32 of row 8(r/4) + 4(l/32) + r%4, so a 64-lane ballot
// of the bits of register r is the word w of two rows.
#include "device_common.hpp"

namespace lshgpu {

namespace {

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float f32x16 __attribute__((ext_vector_type(16)));

constexpr int kTPB = 256;
constexpr int kRows = 128;  // rows per workgroup
constexpr int kKS = 8;      // 16-dimension steps in registers (d <= 128)
constexpr int kMaxWpw = 64; // words per workgroup (LDS staging of the output words)
constexpr int kWaveFix = 256;  // pairs a wave lists in LDS before rechecking in place
constexpr int kBPre = 2 * kKS * 64 / kTPB;  // 16-byte B-operand chunks per thread

struct SmArgs {
    const float* x;        // [n][d]
    int64_t n;
    int d, KS, words, wpw;
    const bf16x8* ph;      // [words][KS][64] operand order, high parts
    const bf16x8* pl;      // low parts
    const float* pnorm;    // [words*32] ||p_j||, rounded up
    const float* planesT;  // [words][d][32] FP32 hyperplanes
    float c;
    uint32_t* out;
    int ostride;
    uint64_t* fix;         // listed pairs: row << 22 | word << 6 | bit << 1 | provisional bit
    uint32_t* fixn;        // their count (may pass fix_cap; those past it are rechecked in place)
    uint32_t fix_cap;
};

__device__ inline float fp32_chain(const float* __restrict__ xv, const float* __restrict__ pv, int d) {
    float t = 0.0f;
    for (int i = 0; i < d; ++i) t = __builtin_fmaf(xv[i], pv[(size_t)i * 32], t);
    return t;
}

// The same chain over a contiguous hyperplane row, its loads issued 16 at a
// time ahead of the multiply-adds.
__device__ inline float fp32_chain_rows(const float* __restrict__ xv, const float* __restrict__ pv, int d) {
    float t = 0.0f;
    int i = 0;
    for (; i + 16 <= d; i += 16) {
        float xa[16], pa[16];
#pragma unroll
        for (int u = 0; u < 16; ++u) {
            xa[u] = xv[i + u];
            pa[u] = pv[i + u];
        }
#pragma unroll
        for (int u = 0; u < 16; ++u) t = __builtin_fmaf(xa[u], pa[u], t);
    }
    for (; i < d; ++i) t = __builtin_fmaf(xv[i], pv[i], t);
    return t;
}

template <int KS>
__global__ __launch_bounds__(kTPB) void simhash_mfma_kernel(SmArgs a) {
    __shared__ float xn[kTPB / 64][32];
    __shared__ uint32_t wst[kTPB / 64][32][kMaxWpw + 1];
    __shared__ uint64_t lfix[kTPB / 64][kWaveFix];  // each wave's listed pairs
    __shared__ bf16x8 bst[2 * kKS * 64];             // the current word's B operands
    __shared__ float bpn[32];                        // and its hyperplane norms
    const int tid = threadIdx.x, wave = tid >> 6, lane = tid & 63, kh = lane >> 5, l32 = lane & 31;
    const int64_t p0 = (int64_t)blockIdx.x * kRows + wave * 32;
    const int64_t pt = p0 + l32;
    const bool valid = pt < a.n;
    const int w0 = blockIdx.y * a.wpw, w1 = min(a.words, w0 + a.wpw);

    bf16x8 xh[KS], xl[KS];
    float ss = 0.0f;
    const float* __restrict__ xr = a.x + (size_t)(valid ? pt : 0) * a.d;
    float xv[KS][8];
    if ((a.d & 3) == 0) {
        // 16-byte loads: rows of a multiple of 4 floats are 16-byte aligned
#pragma unroll
        for (int ks = 0; ks < KS; ++ks)
#pragma unroll
            for (int h = 0; h < 2; ++h) {
                const int i = ks * 16 + kh * 8 + 4 * h;
                const float4 v = (valid && i < a.d) ? *reinterpret_cast<const float4*>(xr + i)
                                                    : make_float4(0.f, 0.f, 0.f, 0.f);
                xv[ks][4 * h + 0] = v.x;
                xv[ks][4 * h + 1] = v.y;
                xv[ks][4 * h + 2] = v.z;
                xv[ks][4 * h + 3] = v.w;
            }
    } else {
#pragma unroll
        for (int ks = 0; ks < KS; ++ks)
#pragma unroll
            for (int e = 0; e < 8; ++e) {
                const int i = ks * 16 + kh * 8 + e;
                xv[ks][e] = (valid && i < a.d) ? xr[i] : 0.0f;
            }
    }
#pragma unroll
    for (int ks = 0; ks < KS; ++ks)
#pragma unroll
        for (int e = 0; e < 8; ++e) {
            const float v = xv[ks][e];
            ss = __builtin_fmaf(v, v, ss);
            const __bf16 hi = (__bf16)v;
            xh[ks][e] = hi;
            xl[ks][e] = (__bf16)(v - (float)hi);
        }
    ss += __shfl_xor(ss, 32);
    if (kh == 0) xn[wave][l32] = __builtin_sqrtf(ss) * (1.0f + 1.0f / 1024.0f);
    __syncthreads();
    // This lane's row among the 32, and which register and half hold its word.
    const int my_r = (l32 >> 3) * 4 + (l32 & 3);
    const bool my_hi = (l32 >> 2) & 1;
    uint32_t nfix = 0;  // pairs this wave listed (wave-uniform)
    // The word's B operands (high parts, then low parts: 2*KS*64 chunks of
    // 16 bytes) are shared by the 4 waves: staged in LDS, the next word's
    // loaded into registers while the current one is multiplied.
    constexpr int nchunk = 2 * KS * 64;
    auto fetch = [&](int w, uint4* pre) {
#pragma unroll
        for (int u = 0; u < kBPre; ++u) {
            const int c = tid + u * kTPB;
            const int half = c >= KS * 64;
            const int cc = half ? c - KS * 64 : c;
            const bf16x8* src = half ? a.pl : a.ph;
            pre[u] = c < nchunk ? *reinterpret_cast<const uint4*>(src + (size_t)w * KS * 64 + cc)
                                : make_uint4(0, 0, 0, 0);
        }
    };
    uint4 pre[kBPre];
    float pre_pn = 0.0f;
    if (w0 < w1) {
        fetch(w0, pre);
        if (tid < 32) pre_pn = a.pnorm[w0 * 32 + tid];
    }
    for (int w = w0; w < w1; ++w) {
        __syncthreads();  // every wave is done with the previous word's operands
#pragma unroll
        for (int u = 0; u < kBPre; ++u) {
            const int c = tid + u * kTPB;
            if (c < nchunk) reinterpret_cast<uint4*>(bst)[c] = pre[u];
        }
        if (tid < 32) bpn[tid] = pre_pn;
        __syncthreads();
        if (w + 1 < w1) {
            fetch(w + 1, pre);
            if (tid < 32) pre_pn = a.pnorm[(w + 1) * 32 + tid];
        }
        bf16x8 bh[KS], bl[KS];
#pragma unroll
        for (int ks = 0; ks < KS; ++ks) {
            bh[ks] = bst[ks * 64 + lane];
            bl[ks] = bst[(KS + ks) * 64 + lane];
        }
        f32x16 acc = (f32x16){};
#pragma unroll
        for (int ks = 0; ks < KS; ++ks) {
            acc = __builtin_amdgcn_mfma_f32_32x32x16_bf16(xh[ks], bh[ks], acc, 0, 0, 0);
            acc = __builtin_amdgcn_mfma_f32_32x32x16_bf16(xl[ks], bh[ks], acc, 0, 0, 0);
            acc = __builtin_amdgcn_mfma_f32_32x32x16_bf16(xh[ks], bl[ks], acc, 0, 0, 0);
        }
        const float pn = bpn[l32] * a.c;
        uint32_t myword = 0;
#pragma unroll
        for (int r = 0; r < 16; ++r) {
            const int row = (r >> 2) * 8 + kh * 4 + (r & 3);
            float y = acc[r];
            const bool need = !(__builtin_fabsf(y) > xn[wave][row] * pn) && p0 + row < a.n;
            const uint64_t nm = __ballot(need);
            if (nm) {
                // Too close to zero for the bound (or NaN): listed in LDS for
                // the FP32 chain (in place once the wave's list is full).
                if (need) {
                    const uint32_t idx = nfix + (uint32_t)__popcll(nm & ((1ull << lane) - 1ull));
                    if (idx < (uint32_t)kWaveFix) {
                        lfix[wave][idx] = ((uint64_t)(p0 + row) << 22) | ((uint64_t)w << 6) |
                                          ((uint64_t)l32 << 1) | (y >= 0.0f ? 1u : 0u);
                    } else {
                        y = fp32_chain(a.x + (size_t)(p0 + row) * a.d, a.planesT + (size_t)w * a.d * 32 + l32, a.d);
                    }
                }
                nfix += (uint32_t)__popcll(nm);
            }
            const uint64_t m = __ballot(y >= 0.0f);
            if (r == my_r) myword = my_hi ? (uint32_t)(m >> 32) : (uint32_t)m;
        }
        if (kh == 0) wst[wave][l32][w - w0] = myword;
    }
    // The wave's listed pairs go to the global list with one atomic; those
    // past its capacity are settled here, in the staged words.
    const uint32_t nl = min(nfix, (uint32_t)kWaveFix);
    if (nl > 0) {
        uint32_t base = 0;
        if (lane == 0) base = atomicAdd(a.fixn, nl);
        base = __shfl(base, 0);
        for (uint32_t i = lane; i < nl; i += 64) {
            const uint64_t e = lfix[wave][i];
            if (base + i < a.fix_cap) {
                a.fix[base + i] = e;
            } else {
                const int64_t row = (int64_t)(e >> 22);
                const int w = (int)((e >> 6) & 0xFFFFu), b = (int)((e >> 1) & 31u);
                const float y = fp32_chain(a.x + (size_t)row * a.d, a.planesT + (size_t)w * a.d * 32 + b, a.d);
                if ((y >= 0.0f ? 1u : 0u) != (uint32_t)(e & 1u)) atomicXor(&wst[wave][row - p0][w - w0], 1u << b);
            }
        }
    }
    __syncthreads();
    const int nw = w1 - w0;
    for (int i = lane; i < 32 * nw; i += 64) {
        const int row = i / nw, wi = i - row * nw;
        if (p0 + row < a.n) a.out[(size_t)(p0 + row) * a.ostride + w0 + wi] = wst[wave][row][wi];
    }
}

// One listed (row, hyperplane) pair a thread: its FP32 chain, and the bit
// flipped when it differs from the provisional one.
__global__ void simhash_fix_kernel(const float* __restrict__ x, int d, const float* __restrict__ prow,
                                   const uint64_t* __restrict__ fix, const uint32_t* __restrict__ fixn,
                                   uint32_t cap, uint32_t* __restrict__ out, int ostride) {
    const uint32_t n = min(*fixn, cap);
    for (uint32_t t = blockIdx.x * blockDim.x + threadIdx.x; t < n; t += gridDim.x * blockDim.x) {
        const uint64_t e = fix[t];
        const int64_t row = (int64_t)(e >> 22);
        const int w = (int)((e >> 6) & 0xFFFFu), b = (int)((e >> 1) & 31u);
        const uint32_t prov = (uint32_t)(e & 1u);
        const float y = fp32_chain_rows(x + (size_t)row * d, prow + ((size_t)w * 32 + b) * d, d);
        const uint32_t bit = y >= 0.0f ? 1u : 0u;
        if (bit != prov) atomicXor(&out[(size_t)row * ostride + w], 1u << b);
    }
}

uint16_t bf16_rne(float f) {
    const uint32_t u = __builtin_bit_cast(uint32_t, f);
    return (uint16_t)((u + 0x7FFFu + ((u >> 16) & 1u)) >> 16);
}
float bf16_value(uint16_t h) { return __builtin_bit_cast(float, (uint32_t)h << 16); }

}  // namespace

bool simhash_mfma_supported(int d, int hashBits) { return hashBits == 32 && d >= 1 && d <= 128; }

void SimhashMfma::build(const float* planes, int d_, int words_, const float* planesT_dev, hipStream_t s) {
    d = d_;
    words = words_;
    KS = (d + 15) / 16;
    planesT = planesT_dev;
    // Operand order: [w][ks][lane][e] = plane 32w + lane%32, dimension
    // 16ks + 8(lane/32) + e (zero past d).
    std::vector<uint16_t> hi((size_t)words * KS * 64 * 8, 0), lo(hi.size(), 0);
    std::vector<float> pn((size_t)words * 32);
    for (int w = 0; w < words; ++w)
        for (int ks = 0; ks < KS; ++ks)
            for (int l = 0; l < 64; ++l)
                for (int e = 0; e < 8; ++e) {
                    const int i = ks * 16 + (l >> 5) * 8 + e;
                    if (i >= d) continue;
                    const float v = planes[((size_t)w * 32 + (l & 31)) * d + i];
                    const uint16_t h = bf16_rne(v);
                    const size_t o = (((size_t)w * KS + ks) * 64 + l) * 8 + e;
                    hi[o] = h;
                    lo[o] = bf16_rne(v - bf16_value(h));
                }
    for (size_t j = 0; j < pn.size(); ++j) {
        double n2 = 0.0;
        for (int i = 0; i < d; ++i) {
            const double v = planes[j * d + i];
            n2 += v * v;
        }
        pn[j] = std::nextafter((float)(std::sqrt(n2) * (1.0 + 1e-6)), 3.0e38f);
    }
    ph.alloc(hi.size());
    pl.alloc(lo.size());
    pnorm.alloc(pn.size());
    prow.alloc((size_t)words * 32 * d);
    LSH_HIP(hipMemcpyAsync(prow.get(), planes, sizeof(float) * prow.size(), hipMemcpyHostToDevice, s));
    LSH_HIP(hipMemcpyAsync(ph.get(), hi.data(), sizeof(uint16_t) * hi.size(), hipMemcpyHostToDevice, s));
    LSH_HIP(hipMemcpyAsync(pl.get(), lo.data(), sizeof(uint16_t) * lo.size(), hipMemcpyHostToDevice, s));
    LSH_HIP(hipMemcpyAsync(pnorm.get(), pn.data(), sizeof(float) * pn.size(), hipMemcpyHostToDevice, s));
    LSH_HIP(hipStreamSynchronize(s));
}

void SimhashMfma::sketch(const float* x, int64_t n, uint32_t* out, int out_stride, hipStream_t s) const {
    if (n <= 0) return;
    SmArgs a{};
    a.x = x;
    a.n = n;
    a.d = d;
    a.KS = KS;
    a.words = words;
    a.ph = reinterpret_cast<const bf16x8*>(ph.get());
    a.pl = reinterpret_cast<const bf16x8*>(pl.get());
    a.pnorm = pnorm.get();
    a.planesT = planesT;
    a.c = (float)(4.0 * std::ldexp(1.0, -16) + (3.0 * 16 * KS + d + 8) * std::ldexp(1.0, -23));
    a.out = out;
    a.ostride = out_stride;
    // The recheck list: 1/256 of the bits (at least 4096 pairs), grow-only.
    const size_t want = std::max<size_t>(4096, (size_t)n * words * 32 / 256);
    if (want > fix.size()) fix.alloc(want);
    if (!fixn.get()) fixn.alloc(1);
    a.fix = fix.get();
    a.fixn = fixn.get();
    a.fix_cap = (uint32_t)std::min<size_t>(fix.size(), 0xFFFFFFFFu);
    launch_fill_i32(reinterpret_cast<int32_t*>(fixn.get()), 1, 0, s);
    const unsigned gx = (unsigned)((n + kRows - 1) / kRows);
    // Enough workgroups for the 256 CUs: small batches split the words.
    int groups = (words + kMaxWpw - 1) / kMaxWpw;
    groups = std::max(groups, (int)std::min<int64_t>(words, (1024 + gx - 1) / gx));
    a.wpw = (words + groups - 1) / groups;
    const unsigned gy = (unsigned)((words + a.wpw - 1) / a.wpw);
    switch (KS) {
#define LSHGPU_SM_CASE(k) \
    case k: hipLaunchKernelGGL(simhash_mfma_kernel<k>, dim3(gx, gy), dim3(kTPB), 0, s, a); break;
        LSHGPU_SM_CASE(1) LSHGPU_SM_CASE(2) LSHGPU_SM_CASE(3) LSHGPU_SM_CASE(4)
        LSHGPU_SM_CASE(5) LSHGPU_SM_CASE(6) LSHGPU_SM_CASE(7) LSHGPU_SM_CASE(8)
#undef LSHGPU_SM_CASE
        default: fail(LSH_ERR_INTERNAL, "matrix-core SimHash: d must be <= 128");
    }
    check_launch("simhash_mfma_kernel");
    const unsigned fg = (unsigned)std::min<size_t>(1024, (want + 255) / 256);
    hipLaunchKernelGGL(simhash_fix_kernel, dim3(fg), dim3(256), 0, s, x, d, prow.get(), a.fix, a.fixn, a.fix_cap, out,
                       out_stride);
    check_launch("simhash_fix_kernel");
}

}  // namespace lshgpu
