// crosspoly_dense.hip -- dense-rotation cross-polytope keys on the matrix
// cores (keysImplementation 8).
//
// Hash h of a row x: y = G_h x with G_h a P x d Gaussian rotation, and the
// index of the largest |y_j| encoded as in crossPolyTope.cuh:64-76 (j when
// y_j < 0, else j + P); a table's key packs K such hashes into
// log2(2P)-bit fields.  The keys are defined by FP32 arithmetic (sequential
// fused multiply-adds over i in index order); the matrix cores compute y
// faster and an error bound decides which hashes need that FP32 arithmetic:
//
// - G's entries are bfloat16 values (the Gaussian draws rounded to 8
//   significant bits), so they enter the products exactly;
// - x is split into x_hi = bf16(x) and x_lo = bf16(x - x_hi), so
//   |x - x_hi - x_lo| <= 2^-16 |x|, and y' = G x_hi + G x_lo is two
//   v_mfma_f32_32x32x16_bf16 per 16 dimensions (exact products, FP32 sums);
// - |y'_j - y32_j| <= E = ||x|| max_j ||G_hj|| (2^-16 + (2 dpad + d + 4) 2^-23)
//   covers the split, the matrix-core sums (at most a 2^-23 relative error per
//   addition, in any order) and the FP32 chain's own error (Cauchy-Schwarz
//   bounds sum |g_i x_i| by ||g|| ||x||);
// - with M = max |y'_j|, a coordinate with |y'_j| < M - 2E has
//   |y32_j| < M - E <= |y32_m| and cannot be the FP32 argmax.  When exactly
//   one coordinate is left and M > E, it is the argmax and its sign is that
//   of y'.  Otherwise the lanes recompute the remaining coordinates in FP32
//   from the rotation in LDS and the row in memory and apply the reference's
//   rule to them.  The keys are then bit-identical to the FP32 definition.
//
// Layout: a workgroup of 4 waves takes 128 rows, 32 per wave; the wave's
// rows are the B operand (lane l: row l%32, dimensions 8*(l/32)+[0,8) of each
// 16-dimension step), held in registers for every hash.  The rotation rows
// are the A operand and are staged hash by hash into LDS (row stride dpad + 8
// bfloat16, so the 16-byte operand reads of a 16-lane group hit all 64
// banks); the next hash's rotation is loaded into registers while the current
// one is multiplied.  Accumulator register r of lane l holds coordinate
// 32*mb + 8*(r/4) + 4*(l/32) + r%4 of row l%32, so the two lanes l and l^32
// hold the 128 coordinates of one row.  The keys of the workgroup's rows
// collect in LDS and are written out once, [n][L] and (for the bucket sort)
// [L][n].
#include "device_common.hpp"

namespace lshgpu {

namespace {

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float f32x16 __attribute__((ext_vector_type(16)));

constexpr int kTPB = 256;
constexpr int kRows = 128;  // rows per workgroup
constexpr int kKS = 8;      // 16-dimension steps held in registers (d <= 128)
constexpr int kPre = 8;     // 16-byte chunks of the next rotation per thread (P * dpad <= 128 * 128)

struct DenseArgs {
    const float* x;        // [n][d], unit rows
    int64_t n;
    int d, dpad, P, H, K, L, b;
    const uint16_t* G;     // [H][P][dpad] bfloat16 bits, zero past d
    const float* gmax;     // [H] max row norm of each rotation (rounded up)
    float c;               // error factor of E
    uint32_t* keys_nl;     // [n][L] or null
    uint32_t* keys_ln;     // [L][n] or null
    int tpw;               // tables per workgroup (grid.y = ceil(L / tpw) table groups)
};

__device__ inline float bf16_bits_to_f32(uint16_t h) { return __builtin_bit_cast(float, (uint32_t)h << 16); }

template <int PB>
__global__ __launch_bounds__(kTPB) void crosspoly_dense_kernel(DenseArgs a) {
    extern __shared__ __align__(16) uint8_t lds_raw[];
    const int RS = a.dpad + 8;  // LDS row stride (bfloat16)
    uint16_t* g_lds = reinterpret_cast<uint16_t*>(lds_raw);
    uint32_t* key_lds = reinterpret_cast<uint32_t*>(lds_raw + (size_t)PB * 32 * RS * 2);
    const int tid = threadIdx.x, wave = tid >> 6, lane = tid & 63, kh = lane >> 5;
    const int64_t row0 = (int64_t)blockIdx.x * kRows;
    // This workgroup's tables [t0, t1): hashes [t0*K, t1*K).
    const int t0 = blockIdx.y * a.tpw, t1 = min(a.L, t0 + a.tpw), nt = t1 - t0;
    const int h0 = t0 * a.K, h1 = t1 * a.K;
    const int64_t pt = row0 + wave * 32 + (lane & 31);
    const bool valid = pt < a.n;
    const int KS = a.dpad / 16;
    const int chunks_row = a.dpad / 8;
    const int chunks = a.P * chunks_row;

    // Rotation rows past P (a P below 32) stay zero; keys start at zero.
    for (int i = tid; i < PB * 32 * RS; i += kTPB)
        if (i / RS >= a.P) g_lds[i] = 0;
    for (int i = tid; i < kRows * nt; i += kTPB) key_lds[i] = 0u;

    // The wave's rows as B operands, split into bfloat16 high and low parts.
    bf16x8 xh[kKS], xl[kKS];
    float ss = 0.0f;
    const float* __restrict__ xr = a.x + (size_t)(valid ? pt : 0) * a.d;
#pragma unroll
    for (int ks = 0; ks < kKS; ++ks) {
        if (ks < KS) {
#pragma unroll
            for (int e = 0; e < 8; ++e) {
                const int i = ks * 16 + kh * 8 + e;
                const float v = (valid && i < a.d) ? xr[i] : 0.0f;
                ss = __builtin_fmaf(v, v, ss);
                const __bf16 hi = (__bf16)v;
                xh[ks][e] = hi;
                xl[ks][e] = (__bf16)(v - (float)hi);
            }
        }
    }
    ss += __shfl_xor(ss, 32);
    const float xnorm = __builtin_sqrtf(ss) * (1.0f + 1.0f / 1024.0f);
    // A NaN in the row makes every FP32 chain NaN (0 * NaN is NaN too), so no
    // coordinate ever wins the reference's comparison and the hash is P; such
    // rows (a zero row after normalisation) skip the recheck.
    const bool xnan = ss != ss;

    // Rotation of hash 0 into LDS.
    auto stage = [&](const uint4* pre) {
#pragma unroll
        for (int u = 0; u < kPre; ++u) {
            const int c = tid + u * kTPB;
            if (c < chunks) {
                const int r = c / chunks_row, col = (c - r * chunks_row) * 8;
                *reinterpret_cast<uint4*>(g_lds + (size_t)r * RS + col) = pre[u];
            }
        }
    };
    auto fetch = [&](int h, uint4* pre) {
        const uint4* __restrict__ src = reinterpret_cast<const uint4*>(a.G + (size_t)h * a.P * a.dpad);
#pragma unroll
        for (int u = 0; u < kPre; ++u) {
            const int c = tid + u * kTPB;
            pre[u] = c < chunks ? src[c] : make_uint4(0, 0, 0, 0);
        }
    };
    uint4 pre[kPre];
    fetch(h0, pre);
    stage(pre);
    __syncthreads();

    const int myrow = wave * 32 + (lane & 31);
    for (int h = h0; h < h1; ++h) {
        if (h + 1 < h1) fetch(h + 1, pre);
        f32x16 acc[PB];
#pragma unroll
        for (int mb = 0; mb < PB; ++mb) acc[mb] = (f32x16){};
#pragma unroll
        for (int ks = 0; ks < kKS; ++ks) {
            if (ks < KS) {
#pragma unroll
                for (int mb = 0; mb < PB; ++mb) {
                    const bf16x8 g = *reinterpret_cast<const bf16x8*>(g_lds + (size_t)(mb * 32 + (lane & 31)) * RS +
                                                                       ks * 16 + kh * 8);
                    acc[mb] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(g, xh[ks], acc[mb], 0, 0, 0);
                    acc[mb] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(g, xl[ks], acc[mb], 0, 0, 0);
                }
            }
        }
        // Largest |y'| of the row (both lanes of the pair), the candidates
        // within 2E of it, and the certain hash when there is one.
        const float E = xnorm * a.gmax[h] * a.c + 1e-30f;
        float best = -1.0f, bval = 0.0f;
        int bj = 0;
        bool nonfin = false;
#pragma unroll
        for (int mb = 0; mb < PB; ++mb)
#pragma unroll
            for (int r = 0; r < 16; ++r) {
                const int j = mb * 32 + (r >> 2) * 8 + kh * 4 + (r & 3);
                if (j < a.P) {
                    const float v = acc[mb][r], av = __builtin_fabsf(v);
                    nonfin |= !(av <= 3.0e38f);
                    if (av > best) {
                        best = av;
                        bval = v;
                        bj = j;
                    }
                }
            }
        const float M = fmaxf(best, __shfl_xor(best, 32));
        const float thr = M - 2.0f * E;
        int cnt = 0;
#pragma unroll
        for (int mb = 0; mb < PB; ++mb)
#pragma unroll
            for (int r = 0; r < 16; ++r) {
                const int j = mb * 32 + (r >> 2) * 8 + kh * 4 + (r & 3);
                if (j < a.P && __builtin_fabsf(acc[mb][r]) >= thr) ++cnt;
            }
        cnt += __shfl_xor(cnt, 32);
        nonfin |= __shfl_xor((int)nonfin, 32) != 0;
        const bool amb = nonfin || cnt != 1 || !(M > E);
        uint32_t hash;
        if (!amb) {
            const int mine = best == M ? (bval < 0.0f ? bj : bj + a.P) : -1;
            hash = (uint32_t)max(mine, __shfl_xor(mine, 32));
        } else {
            // FP32 recheck of this lane's remaining coordinates (all of them
            // when a value is not finite), in ascending j, by the reference's
            // comparison; then the pair's better one (the lower j on a tie).
            float rb = (float)INT32_MIN;
            int ridx = -1, rj = 0x7FFFFFFF;
            if (valid && !xnan) {
#pragma unroll
                for (int mb = 0; mb < PB; ++mb)
#pragma unroll
                    for (int r = 0; r < 16; ++r) {
                        const int j = mb * 32 + (r >> 2) * 8 + kh * 4 + (r & 3);
                        if (j < a.P && (nonfin || __builtin_fabsf(acc[mb][r]) >= thr)) {
                            const uint16_t* __restrict__ gr = g_lds + (size_t)j * RS;
                            float y = 0.0f;
                            for (int i = 0; i < a.d; ++i) y = __builtin_fmaf(bf16_bits_to_f32(gr[i]), xr[i], y);
                            const bool neg = y < 0.0f;
                            const float m = neg ? fmaxf(rb, -y) : fmaxf(rb, y);
                            if (m > rb) {
                                ridx = neg ? j : j + a.P;
                                rb = m;
                                rj = j;
                            }
                        }
                    }
            }
            const float orb = __shfl_xor(rb, 32);
            const int oidx = __shfl_xor(ridx, 32), orj = __shfl_xor(rj, 32);
            if (oidx >= 0 && (ridx < 0 || orb > rb || (orb == rb && orj < rj))) ridx = oidx;
            hash = ridx < 0 ? (uint32_t)a.P : (uint32_t)ridx;
        }
        if (kh == 0) {
            const int t = h / a.K, k = h - t * a.K;
            key_lds[myrow * nt + (t - t0)] |= hash << (k * a.b);
        }
        __syncthreads();  // every wave is done with this rotation
        if (h + 1 < h1) {
            stage(pre);
            __syncthreads();
        }
    }
    const int nrows = (int)min<int64_t>(kRows, a.n - row0);
    if (a.keys_nl)
        for (int i = tid; i < nrows * nt; i += kTPB) {
            const int r = i / nt, t = i - r * nt;
            a.keys_nl[(size_t)(row0 + r) * a.L + t0 + t] = key_lds[i];
        }
    if (a.keys_ln)
        for (int i = tid; i < kRows * nt; i += kTPB) {
            const int t = i / kRows, r = i - t * kRows;
            if (r < nrows) a.keys_ln[(size_t)(t0 + t) * a.n + row0 + r] = key_lds[r * nt + t];
        }
}

}  // namespace

int dense_key_field_bits(int d) {
    int P = 1;
    while (P < d) P *= 2;
    int b = 0;
    while ((1 << b) < 2 * P) ++b;
    return b;
}

void DenseRotations::build(int d_, int tables, int K_, hipStream_t s) {
    LSH_REQUIRE(d_ >= 1 && d_ <= 128, LSH_ERR_UNSUPPORTED, "dense cross-polytope keys support d <= 128");
    d = d_;
    K = K_;
    L = tables;
    P = 1;
    while (P < d) P *= 2;
    b = dense_key_field_bits(d);
    dpad = (d + 15) / 16 * 16;
    H = L * K;
    // randomVectorGenerator.cpp:32-53 stream, rounded to bfloat16 (ties to even)
    const std::vector<float> g = normal_stream((int64_t)H * P * d, 1.0);
    std::vector<uint16_t> bits((size_t)H * P * dpad, 0);
    std::vector<float> gm((size_t)H, 0.0f);
    for (int h = 0; h < H; ++h) {
        double mx = 0.0;
        for (int j = 0; j < P; ++j) {
            double n2 = 0.0;
            for (int i = 0; i < d; ++i) {
                const uint32_t u = __builtin_bit_cast(uint32_t, g[((size_t)h * P + j) * d + i]);
                const uint32_t r = (u + 0x7FFFu + ((u >> 16) & 1u)) >> 16;
                bits[((size_t)h * P + j) * dpad + i] = (uint16_t)r;
                const double v = (double)__builtin_bit_cast(float, r << 16);
                n2 += v * v;
            }
            mx = std::max(mx, std::sqrt(n2));
        }
        gm[h] = std::nextafter((float)(mx * (1.0 + 1e-6)), 3.0e38f);
    }
    G.alloc(bits.size());
    gmax.alloc(gm.size());
    LSH_HIP(hipMemcpyAsync(G.get(), bits.data(), sizeof(uint16_t) * bits.size(), hipMemcpyHostToDevice, s));
    LSH_HIP(hipMemcpyAsync(gmax.get(), gm.data(), sizeof(float) * gm.size(), hipMemcpyHostToDevice, s));
    LSH_HIP(hipStreamSynchronize(s));
}

void DenseRotations::keys(const float* x, int64_t n, uint32_t* keys_nl, uint32_t* keys_ln, hipStream_t s) const {
    if (n <= 0) return;
    DenseArgs a{};
    a.x = x;
    a.n = n;
    a.d = d;
    a.dpad = dpad;
    a.P = P;
    a.H = H;
    a.K = K;
    a.L = L;
    a.b = b;
    a.G = G.get();
    a.gmax = gmax.get();
    a.c = (float)(std::ldexp(1.0, -16) + (2.0 * dpad + d + 4) * std::ldexp(1.0, -23));
    a.keys_nl = keys_nl;
    a.keys_ln = keys_ln;
    const int PB = (P + 31) / 32;
    const size_t lds = (size_t)PB * 32 * (dpad + 8) * 2 + (size_t)kRows * L * 4;
    LSH_REQUIRE(lds <= 160 * 1024, LSH_ERR_UNSUPPORTED, "too many tables for the dense cross-polytope kernel");
    const unsigned grid = (unsigned)((n + kRows - 1) / kRows);
    // Enough workgroups for the 256 CUs: small batches split the tables.
    const int groups = (int)std::min<int64_t>(L, std::max<int64_t>(1, (1024 + grid - 1) / grid));
    a.tpw = (L + groups - 1) / groups;
    const unsigned gy = (unsigned)((L + a.tpw - 1) / a.tpw);
    const void* fn = PB == 1 ? (const void*)crosspoly_dense_kernel<1>
                             : PB == 2 ? (const void*)crosspoly_dense_kernel<2> : (const void*)crosspoly_dense_kernel<4>;
    set_dynamic_lds(fn, lds);
    if (PB == 1)
        hipLaunchKernelGGL(crosspoly_dense_kernel<1>, dim3(grid, gy), dim3(kTPB), lds, s, a);
    else if (PB == 2)
        hipLaunchKernelGGL(crosspoly_dense_kernel<2>, dim3(grid, gy), dim3(kTPB), lds, s, a);
    else
        hipLaunchKernelGGL(crosspoly_dense_kernel<4>, dim3(grid, gy), dim3(kTPB), lds, s, a);
    check_launch("crosspoly_dense_kernel");
}

}  // namespace lshgpu
