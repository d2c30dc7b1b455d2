// minhash.hip -- weighted MinHash sketches (weightedMinHash.cuh:41-166).
//
// After the shift + normalisation transform, every dimension j owns the
// interval [start_j, start_j + bound_j) of [0, M) with bound_j = max(0,
// ceil(max x_j)) (weightedMinHash.cuh:41-90).  Hash h draws r = M * U_t for
// t = 0, 1, ... until r <= start_c + x_c, with c the dimension owning
// floor(r) (isGreen, weightedMinHash.cuh:98-106); its value is the number of
// red draws (sketchDataOriginal, :144-166).  The one-bit variant stores
// bitmap[count] per bit (sketchDataOneBit, :109-140).  Hash seeds are
// i*1234 + 92138 (:252-262).
//
// The reference draws U_t from cuRAND XORWOW (curand_init(seed, 0, 10000)),
// whose stream is not reproducible here.  This build draws U_t from
// Philox4x32-10 (Salmon et al., SC'11): key (seed, 0x5EED5EED), counter
// (t/4, 0, 0, 0), word t%4, U = (word >> 8) * 2^-24 in [0, 1).  A draw
// sequence depends only on the hash, so a wave of 64 points walks one
// sequence together: the draws are wave-uniform (precomputed per hash into a
// table of R draws, Philox beyond), the point rows sit in LDS, and the wave
// stops when every lane has seen its first green draw.  A lane that sees no
// green draw in kMaxDraws draws (a zero or NaN row) records kMaxDraws.
#include "device_common.hpp"

namespace lshgpu {

namespace {

constexpr int kMaxDraws = 1 << 16;

__device__ inline uint32_t philox_word(uint32_t seed, uint32_t t) {
    uint32_t c0 = t >> 2, c1 = 0, c2 = 0, c3 = 0;
    uint32_t k0 = seed, k1 = 0x5EED5EEDu;
#pragma unroll
    for (int r = 0; r < 10; ++r) {
        const uint32_t hi0 = __umulhi(0xD2511F53u, c0), lo0 = 0xD2511F53u * c0;
        const uint32_t hi1 = __umulhi(0xCD9E8D57u, c2), lo1 = 0xCD9E8D57u * c2;
        const uint32_t n0 = hi1 ^ c1 ^ k0;
        const uint32_t n2 = hi0 ^ c3 ^ k1;
        c1 = lo1;
        c3 = lo0;
        c0 = n0;
        c2 = n2;
        k0 += 0x9E3779B9u;
        k1 += 0xBB67AE85u;
    }
    switch (t & 3) {
        case 0: return c0;
        case 1: return c1;
        case 2: return c2;
        default: return c3;
    }
}

__device__ inline float wmh_draw(uint32_t seed, uint32_t t, float Mf) {
    return Mf * ((float)(philox_word(seed, t) >> 8) * (1.0f / 16777216.0f));
}

__global__ void wmh_bounds_kernel(const float* __restrict__ x, int64_t n, int d, int rows_per_block,
                                  int* __restrict__ bounds) {
    const int64_t r0 = (int64_t)blockIdx.x * rows_per_block;
    const int64_t r1 = min(n, r0 + rows_per_block);
    for (int j = threadIdx.x; j < d; j += blockDim.x) {
        int m = 0;
        for (int64_t r = r0; r < r1; ++r) m = max(m, (int)ceilf(x[(size_t)r * d + j]));
        atomicMax(&bounds[j], m);
    }
}

// draws[h][t], owner_of_draw[h][t] for t < R.
__global__ void wmh_draws_kernel(int nhash, int R, int M, const int* __restrict__ owner,
                                 float* __restrict__ draws, int* __restrict__ downer) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= (int64_t)nhash * R) return;
    const int h = (int)(i / R), t = (int)(i % R);
    const uint32_t seed = (uint32_t)(h * 1234 + 92138);
    const float r = wmh_draw(seed, (uint32_t)t, (float)M);
    draws[i] = r;
    downer[i] = owner[min((int)r, M - 1)];
}

// Block: 64 points (one wave) x `per` hashes (blockIdx.y selects the hashes
// per*y .. per*y+per-1).  mode 0 (per = 32): out[i][h] = red count; mode 1
// (per = hbits): bit b of out[i][y] = bitmap[count of hash y*hbits + b %
// bitmap_len] (sketchDataOneBit, weightedMinHash.cuh:115-136).
__global__ __launch_bounds__(64) void wmh_sketch_kernel(const float* __restrict__ x, int64_t n, int d,
                                                        const int* __restrict__ start,
                                                        const int* __restrict__ owner, int M, int nhash,
                                                        int mode, int per, const uint8_t* __restrict__ bitmap,
                                                        int bitmap_len, const float* __restrict__ draws,
                                                        const int* __restrict__ downer, int R,
                                                        uint32_t* __restrict__ out, int ostride) {
    extern __shared__ float rows[];  // [64][d + 1]
    const int lane = threadIdx.x;
    const int64_t i0 = (int64_t)blockIdx.x * 64;
    const int ld = d + 1;
    for (int e = lane; e < 64 * d; e += 64) {
        const int p = e / d, j = e % d;
        rows[p * ld + j] = i0 + p < n ? x[(size_t)(i0 + p) * d + j] : 0.0f;
    }
    __syncthreads();
    const bool valid = i0 + lane < n;
    const float* row = rows + lane * ld;
    const float Mf = (float)M;
    uint32_t word = 0;
    for (int hb = 0; hb < per; ++hb) {
        const int h = blockIdx.y * per + hb;
        if (h >= nhash) break;
        const uint32_t seed = (uint32_t)(h * 1234 + 92138);
        int count = (valid && M > 0) ? -1 : 0;
        for (int t = 0; t < kMaxDraws; ++t) {
            if (!__ballot(count < 0)) break;
            float r;
            int c;
            if (t < R) {
                r = draws[(size_t)h * R + t];
                c = downer[(size_t)h * R + t];
            } else {
                r = wmh_draw(seed, (uint32_t)t, Mf);
                c = owner[min((int)r, M - 1)];
            }
            if (count < 0 && r <= (float)start[c] + row[c]) count = t;
        }
        if (count < 0) count = kMaxDraws;
        if (mode == 0) {
            if (valid) out[(size_t)(i0 + lane) * ostride + h] = (uint32_t)count;
        } else {
            word |= (uint32_t)bitmap[(uint32_t)count % (uint32_t)bitmap_len] << hb;
        }
    }
    if (mode == 1 && valid) out[(size_t)(i0 + lane) * ostride + blockIdx.y] = word;
}

}  // namespace

int wmh_max_draws() { return kMaxDraws; }

void launch_wmh_bounds(const float* x, int64_t n, int d, int* bounds, hipStream_t s) {
    if (n <= 0) return;
    const int rpb = 256;
    hipLaunchKernelGGL(wmh_bounds_kernel, dim3((unsigned)((n + rpb - 1) / rpb)), dim3(128), 0, s, x, n, d, rpb,
                       bounds);
    check_launch("wmh_bounds_kernel");
}

void launch_wmh_draws(int nhash, int R, int M, const int* owner, float* draws, int* downer, hipStream_t s) {
    const int64_t total = (int64_t)nhash * R;
    if (total <= 0 || M <= 0) return;
    hipLaunchKernelGGL(wmh_draws_kernel, dim3((unsigned)((total + 255) / 256)), dim3(256), 0, s, nhash, R, M,
                       owner, draws, downer);
    check_launch("wmh_draws_kernel");
}

void launch_wmh_sketch(const float* x, int64_t n, int d, const int* owner, const int* start, int M, int nhash,
                       int mode, int hbits, const uint8_t* bitmap, int bitmap_len, const float* draws,
                       const int* draw_owner, int R, uint32_t* out, int out_stride, hipStream_t s) {
    if (n <= 0 || nhash <= 0) return;
    LSH_REQUIRE(mode == 0 || (hbits >= 1 && hbits <= 32 && nhash % hbits == 0), LSH_ERR_INVALID,
                "one-bit MinHash needs 1..32 hashes per word");
    const size_t lds = (size_t)64 * (d + 1) * sizeof(float);
    set_dynamic_lds((const void*)wmh_sketch_kernel, lds);
    const int per = mode == 1 ? hbits : 32;
    const unsigned gy = (unsigned)((nhash + per - 1) / per);
    LSH_REQUIRE(gy <= 65535, LSH_ERR_UNSUPPORTED, "too many MinHash hashes");
    hipLaunchKernelGGL(wmh_sketch_kernel, dim3((unsigned)((n + 63) / 64), gy), dim3(64), lds, s, x, n, d, start,
                       owner, M, nhash, mode, per, bitmap, bitmap_len, draws, draw_owner, R, out, out_stride);
    check_launch("wmh_sketch_kernel");
}

}  // namespace lshgpu
