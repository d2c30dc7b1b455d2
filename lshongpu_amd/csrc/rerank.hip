// rerank.hip -- exact rerank of a sketch-distance candidate set.
//
// candidateSetScanner.cuh:10-32 recomputes the exact distance of every slot
// of a lane's queue and the final startSort/write of
// sketchedDistanceScanners.cuh:87-99 keeps the k best.  Here one workgroup
// owns one query.  The C candidates get an approximate distance with a
// proven error bound eps: FP32 over the FP32 rows, or, for the angular
// kinds, FP32 arithmetic over an FP16 image of the unit rows.  A radix
// select finds the k-th smallest approximate distance v_k; every candidate
// whose approximate distance is within the band that eps leaves above v_k
// (v_k + 2 eps, or the relative band for the Euclidean kind) is recomputed
// in FP64 and the set is sorted by (FP64 distance, ID).  A candidate outside
// the band is farther, in exact arithmetic, than each of the k approximate
// best, so the k written are always the FP64 order of the whole candidate
// set.
#include "device_common.hpp"

#include <cstdlib>
#include <type_traits>

namespace lshgpu {

namespace {

// Rows of the FP16 image hold x / ||x|| scaled by 2^14, so that components
// down to 2^-28 of the row norm stay out of the FP16 subnormal range.
constexpr float kHalfScale = 16384.0f;

typedef _Float16 half8 __attribute__((ext_vector_type(8)));

__device__ inline uint64_t shfl_xor_u64(uint64_t x, int m) {
    const uint32_t lo = (uint32_t)__shfl_xor((int)(uint32_t)x, m);
    const uint32_t hi = (uint32_t)__shfl_xor((int)(uint32_t)(x >> 32), m);
    return ((uint64_t)hi << 32) | lo;
}

// Ascending sort of one (key, id) pair per lane across a wave.
__device__ inline void wave_sort_pairs(uint64_t& k, int32_t& id) {
    const int lane = threadIdx.x & 63;
    for (int size = 2; size <= 64; size <<= 1)
        for (int stride = size >> 1; stride > 0; stride >>= 1) {
            const uint64_t ko = shfl_xor_u64(k, stride);
            const int32_t io = __shfl_xor(id, stride);
            const bool up = (lane & size) == 0, lower = (lane & stride) == 0;
            const bool mine_smaller = k < ko || (k == ko && id < io);
            // The lower lane of an ascending pair keeps the smaller pair.
            if (mine_smaller != (lower == up)) {
                k = ko;
                id = io;
            }
        }
}

// High word of the kk-th smallest (1-based) of the n keys (kk <= n).  The
// high words' minimum and maximum fix the bits they share; 8-bit digits are
// then histogrammed from the first bit that differs down, until the bin
// holding the kk-th has at most 256 keys, whose high words are gathered (in
// ties) and ranked against each other.  sv: 8 words, hist: 256 words, ties:
// min(n, 256) words, all LDS; every thread must call it.
__device__ uint32_t kth_high_word(const uint64_t* keys, int n, int kk, uint32_t* hist, uint32_t* sv,
                                  uint32_t* ties) {
    const int tid = threadIdx.x, bs = blockDim.x, lane = tid & 63;
    if (tid == 0) {
        sv[0] = 0xFFFFFFFFu;
        sv[1] = 0u;
        sv[4] = 0u;
    }
    __syncthreads();
    uint32_t mn = 0xFFFFFFFFu, mx = 0u;
    for (int t = tid; t < n; t += bs) {
        const uint32_t h = (uint32_t)(keys[t] >> 32);
        mn = min(mn, h);
        mx = max(mx, h);
    }
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) {
        mn = min(mn, (uint32_t)__shfl_xor((int)mn, o));
        mx = max(mx, (uint32_t)__shfl_xor((int)mx, o));
    }
    if (lane == 0) {
        atomicMin(&sv[0], mn);
        atomicMax(&sv[1], mx);
    }
    __syncthreads();
    const uint32_t lo = sv[0], hi = sv[1];
    if (lo == hi) return lo;
    // The first differing bit is 31 - clz(lo ^ hi); the digit [shift, shift
    // + 8) holds it, and every key shares lo's bits above the digit.
    int shift = max(0, 24 - (int)__builtin_clz(lo ^ hi));
    uint32_t mask = shift + 8 >= 32 ? 0u : ~0u << (shift + 8);
    uint32_t prefix = lo & mask;
    uint32_t rank = (uint32_t)kk;
    uint32_t nt;
    while (true) {
        for (int i = tid; i < 256; i += bs) hist[i] = 0;
        __syncthreads();
        for (int t = tid; t < n; t += bs) {
            const uint32_t h = (uint32_t)(keys[t] >> 32);
            if ((h & mask) == prefix) atomicAdd(&hist[(h >> shift) & 255u], 1u);
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
                    sv[2] = 4u * tid + e;
                    sv[3] = rank - run;
                    sv[5] = c[e];
                }
                run += c[e];
            }
        }
        __syncthreads();
        const uint32_t dig = sv[2];
        rank = sv[3];
        nt = sv[5];
        prefix |= dig << shift;
        mask |= 255u << shift;
        __syncthreads();
        if (shift == 0) return prefix;  // every bit fixed
        if (nt <= 256u) break;
        shift = max(0, shift - 8);
    }
    // The nt keys of the bin: the one with at most rank - 1 smaller high
    // words and more than rank - 1 smaller or equal ones.
    for (int t = tid; t < n; t += bs) {
        const uint32_t h = (uint32_t)(keys[t] >> 32);
        if ((h & mask) == prefix) ties[atomicAdd(&sv[4], 1u)] = h;
    }
    __syncthreads();
    const uint32_t want = rank - 1u;
    for (uint32_t i = tid; i < nt; i += bs) {
        const uint32_t x = ties[i];
        uint32_t lt = 0, le = 0;
        for (uint32_t j = 0; j < nt; ++j) {
            const uint32_t y = ties[j];
            lt += y < x ? 1u : 0u;
            le += y <= x ? 1u : 0u;
        }
        if (lt <= want && want < le) sv[6] = x;
    }
    __syncthreads();
    return sv[6];
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

template <int CTRL>
__device__ inline float dpp_f(float x) {
    return __builtin_bit_cast(float, __builtin_amdgcn_update_dpp(0, __builtin_bit_cast(int, x), CTRL, 0xF, 0xF, false));
}

// The sums over the 16 lanes of a DPP row of four values a0..a3 at once: lane
// l of the row returns the sum of a_{l / 4}.  The halves of the row trade
// the pairs they drop (rotation by 8), the lanes of each half then trade
// the value they drop with their mirror in the half (l <-> 7 - l), and each
// quad, holding one value summed over four lanes that cover all 16, sums
// it with two quad permutations: 11 VALU operations for the four sums
// instead of 16 for four rotation chains.
__device__ inline float row_sums4(float a0, float a1, float a2, float a3, int l) {
    const bool hi = l >= 8;
    const float b0 = (hi ? a2 : a0) + dpp_f<0x128>(hi ? a0 : a2);  // row_ror:8
    const float b1 = (hi ? a3 : a1) + dpp_f<0x128>(hi ? a1 : a3);
    const bool up = (l & 7) >= 4;
    float c = (up ? b1 : b0) + dpp_f<0x141>(up ? b0 : b1);  // row_half_mirror
    c += dpp_f<0xB1>(c);                                    // quad_perm [1,0,3,2]
    c += dpp_f<0x4E>(c);                                    // quad_perm [2,3,0,1]
    return c;
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

// Approximate distances of candidates [0, Cp) into sums[] (two floats per
// candidate), FP32 rows: a group of 16 lanes reads one candidate row with
// coalesced float4 loads, U rows per group in flight; the partial sums are
// added with DPP rotations and lane u of the group keeps row u's sums.
template <int KIND, int U>
__device__ inline void sums_f32(const float* __restrict__ data, const float* __restrict__ qv, int d,
                                const int32_t* rows, int Cp, float2* sums) {
    const int gl = threadIdx.x & 15;
    const int ngroups = blockDim.x >> 4;
    const bool vec4 = (d & 3) == 0;
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
}

// Whether sums_f16<4> takes rows[] as offsets into the FP16 image, in
// 16-byte pieces (row * hs / 8; the image is built only while these fit 32
// bits), rather than as row numbers.
__device__ inline bool rows_as_offsets(int hs, int Cp) { return hs <= 128 && Cp >= 4; }

// The FP16 dot products of sums_f16 for hs <= 128 and Cp >= 4: four
// consecutive candidates per 16-lane group and step (Cp, a power of two, is
// a multiple of 4).  rows[] holds each candidate's row offset in 16-byte
// pieces (rows_as_offsets), read in two 8-byte LDS reads; a lane's row piece
// (jp) is one add from there, addressed in bytes from the image base while
// the image is under 4 GiB (B32: 32-bit offsets from a uniform base), else
// as a 64-bit piece index.  Lanes past the row (act false) issue no loads
// and add zeros: the data path that returns loaded bytes to the lanes
// bounds this loop, and rows of hs <= 104 halves leave at least three of
// the 16 lanes idle.  Lanes 4u..4u+3 end with candidate u's sum.  The first
// product is a fused multiply-add onto zero: the product's own rounding,
// in one mixed-precision instruction.
template <bool B32>
__device__ inline void rows4_sums(const _Float16* __restrict__ hdata, const int32_t* rows, int Cp, uint32_t jp,
                                  bool act, float4 y0, float4 y1, float2* sums) {
    const int gl = threadIdx.x & 15;
    const int ngroups = blockDim.x >> 4;
    const char* __restrict__ bytes = reinterpret_cast<const char*>(hdata);
    const half8* __restrict__ base = reinterpret_cast<const half8*>(hdata) + jp;
    for (int g4 = 4 * (threadIdx.x >> 4); g4 < Cp; g4 += 4 * ngroups) {
        const uint2 r01 = *reinterpret_cast<const uint2*>(rows + g4);
        const uint2 r23 = *reinterpret_cast<const uint2*>(rows + g4 + 2);
        const uint32_t r[4] = {r01.x, r01.y, r23.x, r23.y};
        float a[4] = {0.0f, 0.0f, 0.0f, 0.0f};
        if (act) {
            half8 xv[4];
#pragma unroll
            for (int u = 0; u < 4; ++u) {
                if (B32) xv[u] = *reinterpret_cast<const half8*>(bytes + ((r[u] + jp) << 4));
                else xv[u] = base[r[u]];
            }
#pragma unroll
            for (int u = 0; u < 4; ++u) {
                a[u] = __builtin_fmaf((float)xv[u][0], y0.x, 0.0f);
                a[u] = __builtin_fmaf((float)xv[u][1], y0.y, a[u]);
                a[u] = __builtin_fmaf((float)xv[u][2], y0.z, a[u]);
                a[u] = __builtin_fmaf((float)xv[u][3], y0.w, a[u]);
                a[u] = __builtin_fmaf((float)xv[u][4], y1.x, a[u]);
                a[u] = __builtin_fmaf((float)xv[u][5], y1.y, a[u]);
                a[u] = __builtin_fmaf((float)xv[u][6], y1.z, a[u]);
                a[u] = __builtin_fmaf((float)xv[u][7], y1.w, a[u]);
            }
        }
        const float as = row_sums4(a[0], a[1], a[2], a[3], gl);
        if ((gl & 3) == 0) sums[g4 + (gl >> 2)] = make_float2(as, 0.0f);
    }
}

// The same over the FP16 image (rows of hs halves, hs a multiple of 8, zero
// padded): one 16-byte load per lane covers 8 components, the query comes
// from LDS (qs, zero padded to hs), and only the dot product is summed.
// With hs <= 128 a row is one load per lane and the query stays in
// registers; the U loads of a step are issued before any is used.
template <int U>
__device__ inline void sums_f16(const _Float16* __restrict__ hdata, int hs, const float* qs, const int32_t* rows,
                                int Cp, float2* sums, bool img32) {
    const int gl = threadIdx.x & 15;
    const int ngroups = blockDim.x >> 4;
    if (hs <= 128) {
        const bool act = gl * 8 < hs;
        const int j = act ? gl * 8 : 0;
        float4 y0 = *reinterpret_cast<const float4*>(qs + j);
        float4 y1 = *reinterpret_cast<const float4*>(qs + j + 4);
        if (!act) y0 = y1 = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
        if (U == 4 && Cp >= 4) {
            if (img32) rows4_sums<true>(hdata, rows, Cp, (uint32_t)j >> 3, act, y0, y1, sums);
            else rows4_sums<false>(hdata, rows, Cp, (uint32_t)j >> 3, act, y0, y1, sums);
            return;
        }
        // Any other U, or fewer than four candidates: rows[] holds rows.
        for (int t0 = threadIdx.x >> 4; t0 < Cp; t0 += ngroups * U) {
            half8 xv[U];
#pragma unroll
            for (int u = 0; u < U; ++u) {
                const int t = t0 + u * ngroups;
                const int r = t < Cp ? rows[t] : -1;
                xv[u] = *reinterpret_cast<const half8*>(hdata + (size_t)(r < 0 ? 0 : r) * hs + j);
            }
            float a[U];
#pragma unroll
            for (int u = 0; u < U; ++u) {
                a[u] = (float)xv[u][0] * y0.x;
                a[u] = __builtin_fmaf((float)xv[u][1], y0.y, a[u]);
                a[u] = __builtin_fmaf((float)xv[u][2], y0.z, a[u]);
                a[u] = __builtin_fmaf((float)xv[u][3], y0.w, a[u]);
                a[u] = __builtin_fmaf((float)xv[u][4], y1.x, a[u]);
                a[u] = __builtin_fmaf((float)xv[u][5], y1.y, a[u]);
                a[u] = __builtin_fmaf((float)xv[u][6], y1.z, a[u]);
                a[u] = __builtin_fmaf((float)xv[u][7], y1.w, a[u]);
            }
            if constexpr (U == 4) {
                // Lanes 4u..4u+3 of the group hold row u's sum.
                const float as = row_sums4(a[0], a[1], a[2], a[3], gl);
                const int t = t0 + (gl >> 2) * ngroups;
                if ((gl & 3) == 0 && t < Cp) sums[t] = make_float2(as, 0.0f);
            } else {
#pragma unroll
                for (int u = 0; u < U; ++u) a[u] = row_sum16(a[u]);
                float as = a[0];
#pragma unroll
                for (int u = 1; u < U; ++u)
                    if (gl == u) as = a[u];
                const int t = t0 + gl * ngroups;
                if (gl < U && t < Cp) sums[t] = make_float2(as, 0.0f);
            }
        }
        return;
    }
    for (int t0 = threadIdx.x >> 4; t0 < Cp; t0 += ngroups * U) {
        const _Float16* rp[U];
        float a[U];
#pragma unroll
        for (int u = 0; u < U; ++u) {
            const int t = t0 + u * ngroups;
            const int r = t < Cp ? rows[t] : -1;
            rp[u] = hdata + (size_t)(r < 0 ? 0 : r) * hs;
            a[u] = 0.0f;
        }
        for (int j = gl * 8; j < hs; j += 128) {
            const float4 y0 = *reinterpret_cast<const float4*>(qs + j);
            const float4 y1 = *reinterpret_cast<const float4*>(qs + j + 4);
            half8 xv[U];
#pragma unroll
            for (int u = 0; u < U; ++u) xv[u] = *reinterpret_cast<const half8*>(rp[u] + j);
#pragma unroll
            for (int u = 0; u < U; ++u) {
                a[u] = __builtin_fmaf((float)xv[u][0], y0.x, a[u]);
                a[u] = __builtin_fmaf((float)xv[u][1], y0.y, a[u]);
                a[u] = __builtin_fmaf((float)xv[u][2], y0.z, a[u]);
                a[u] = __builtin_fmaf((float)xv[u][3], y0.w, a[u]);
                a[u] = __builtin_fmaf((float)xv[u][4], y1.x, a[u]);
                a[u] = __builtin_fmaf((float)xv[u][5], y1.y, a[u]);
                a[u] = __builtin_fmaf((float)xv[u][6], y1.z, a[u]);
                a[u] = __builtin_fmaf((float)xv[u][7], y1.w, a[u]);
            }
        }
#pragma unroll
        for (int u = 0; u < U; ++u) a[u] = row_sum16(a[u]);
        float as = a[0];
#pragma unroll
        for (int u = 1; u < U; ++u)
            if (gl == u) as = a[u];
        const int t = t0 + gl * ngroups;
        if (gl < U && t < Cp) sums[t] = make_float2(as, 0.0f);
    }
}

// HALF: the approximate distances come from hdata (angular kinds only).
// BAND: the kernel stops once the band is known and writes it, (row << 32 |
// ID) per member, to band[pos][0, m) with bcnt[pos] = m, pos being the
// query's place in qorder (its index without one); rerank_recheck_kernel
// finishes the query.
template <int KIND, bool HALF, bool BAND>
__global__ __launch_bounds__(256) void rerank_kernel(const float* __restrict__ data,
                                                     const _Float16* __restrict__ hdata, int hs,
                                                     const float* __restrict__ queries, int d,
                                                     const uint64_t* __restrict__ keys,
                                                     int key_stride, const int32_t* __restrict__ cand,
                                                     int C, int Cp, int k, double eps,
                                                     lsh_point_t* __restrict__ out,
                                                     const int32_t* __restrict__ qorder, int Q, int per,
                                                     const int32_t* __restrict__ rowmap,
                                                     uint64_t* __restrict__ band, uint32_t* __restrict__ bcnt,
                                                     const int32_t* __restrict__ rows_in, int rows_stride,
                                                     int img32, const int32_t* __restrict__ fused_ids) {
    extern __shared__ uint64_t sm[];
    // LDS, in bytes from sm: k64 [0, 8Cp) (absent with BAND: it aliases
    // k32), ids [8Cp, 12Cp), k32 [12Cp, 20Cp), rows [20Cp, 24Cp), qs [24Cp,
    // 24Cp + 4hs), each shifted down by 8Cp with BAND.  k32 and rows are
    // adjacent: once the band is known they serve as one 12Cp-byte staging
    // buffer for its rows.  The approximate sums (float2) alias k64, or k32
    // with BAND (each key replaces its own candidate's sums).
    uint64_t* k64 = sm;                                              // Cp
    int32_t* ids = reinterpret_cast<int32_t*>(sm + (BAND ? 0 : Cp)); // Cp
    uint64_t* k32 = reinterpret_cast<uint64_t*>(ids + Cp);           // Cp
    if (BAND) k64 = k32;
    int32_t* rows = reinterpret_cast<int32_t*>(k32 + Cp);            // Cp: data row of each candidate
    float* qs = reinterpret_cast<float*>(rows + Cp);                 // hs (HALF only)
    __shared__ uint32_t rs_hist[256];
    __shared__ uint32_t rs_sv[8];
    // With an order, XCD x (workgroups x, x + 8, ...) takes the x-th eighth
    // of it in sequence.
    int pos = blockIdx.x, q = pos;  // pos: place in the list of queries (band row)
    if (qorder) {
        pos = (int)(blockIdx.x & 7) * per + (int)(blockIdx.x >> 3);
        if (pos >= Q) return;
        q = qorder[pos];
    }
    const float* __restrict__ qv = queries + (size_t)q * d;

    // Candidate IDs and rows first (to LDS), so the row reads below do not
    // wait on the key reads: each thread issues the loads of four candidates
    // at a time (keys, then their row-map entries, or with rows_in the rows
    // beside the keys) before it uses any.
    if (HALF)
        for (int j = threadIdx.x; j < hs; j += blockDim.x) qs[j] = j < d ? qv[j] : 0.0f;
    // Fused keys (BAND only): (distance << 16 | q, row), q the candidate's
    // 16-bit approximate distance from the fused cell scan; no row is read
    // and the IDs come from fused_ids for the band alone.
    const bool fused = BAND && HALF && fused_ids != nullptr;
    // With offsets in rows[], a candidate's row is its offset / pieces.
    const bool roff = HALF && rows_as_offsets(hs, Cp) && !fused;
    const uint32_t pieces = (uint32_t)(hs >> 3);
    auto row_of = [&](int t) -> uint32_t { return roff ? (uint32_t)rows[t] / pieces : (uint32_t)rows[t]; };
    constexpr int B = 4;
    // The approximate sums (or, with fused keys, distances) go to `sums`,
    // which aliases k64 (k32 with BAND) and is filled below.
    float2* sums = reinterpret_cast<float2*>(k64);
    for (int t0 = threadIdx.x; t0 < Cp; t0 += B * blockDim.x) {
        int32_t id[B], r[B];
        float fd[B];
#pragma unroll
        for (int u = 0; u < B; ++u) {
            const int t = t0 + u * (int)blockDim.x;
            id[u] = -1;
            r[u] = -1;
            fd[u] = 0.0f;
            if (t < C) {
                if (rows_in) r[u] = rows_in[(size_t)q * rows_stride + t];
                if (cand) {
                    id[u] = cand[(size_t)q * C + t];
                } else {
                    const uint64_t key = keys[(size_t)q * key_stride + t];
                    if (fused) {
                        // ID unknown (0 marks a candidate); the row and q from the key.
                        id[u] = key == kEmptyKey ? -1 : 0;
                        r[u] = (int32_t)(uint32_t)key;
                        fd[u] = (float)((uint32_t)(key >> 32) & 0xFFFFu) * 0x1p-15f - (KIND == kNegCos ? 1.0f : 0.0f);
                    } else {
                        id[u] = key == kEmptyKey ? -1 : (int32_t)(uint32_t)key;
                    }
                }
            }
        }
        if (!fused) {
#pragma unroll
            for (int u = 0; u < B; ++u)
                r[u] = id[u] < 0 ? -1 : (rows_in ? r[u] : (rowmap ? rowmap[id[u]] : id[u]));
        }
#pragma unroll
        for (int u = 0; u < B; ++u) {
            const int t = t0 + u * (int)blockDim.x;
            if (t < Cp) {
                ids[t] = id[u];
                // An empty slot reads row 0 in the filter; its sum is not used.
                rows[t] = roff ? (int32_t)((uint32_t)max(r[u], 0) * pieces) : r[u];
                if (fused) sums[t] = make_float2(fd[u], 0.0f);
            }
        }
    }
    __syncthreads();
    // The query norm in FP32 and in FP64 for the band.  Over the FP16 image
    // each wave sums the LDS copy (its zero padding adds exact zeros) with
    // one multiply-add chain per lane and a butterfly; every lane of every
    // wave ends with the same value (each step adds the same two operands in
    // both lanes of a pair), and the error bound holds for any order.  The
    // FP32 rows keep the sequential chain.
    float qn = 0.0f;
    double nq64 = 0.0;
    if (KIND == kNegCos || KIND == kAngular) {
        if (HALF) {
            for (int j = threadIdx.x & 63; j < hs; j += 64) qn = __builtin_fmaf(qs[j], qs[j], qn);
#pragma unroll
            for (int o = 32; o > 0; o >>= 1) qn += __shfl_xor(qn, o);
        } else {
            for (int j = 0; j < d; ++j) qn = __builtin_fmaf(qv[j], qv[j], qn);
        }
        qn = __builtin_sqrtf(qn);
        if (!BAND) nq64 = norm2_64(qv, d);
    }
    // Approximate distances (k64 serves as scratch for the sums; fused keys
    // brought theirs).
    if (HALF && !fused)
        sums_f16<4>(hdata, hs, qs, rows, Cp, sums, img32 != 0);
    else if (!HALF)
        sums_f32<KIND, 6>(data, qv, d, rows, Cp, sums);
    __syncthreads();
    // Queries of extreme norm leave the FP16 bound's range assumptions;
    // every candidate is then rechecked.
    const bool all = HALF && !(qn >= 0x1p-60f && qn <= 0x1p60f);
    // The image's scale and the query norm as one reciprocal (one rounding
    // more than a division per candidate; the bound allows for it).
    const float inv = HALF ? (1.0f / kHalfScale) / qn : 0.0f;
    for (int t = threadIdx.x; t < Cp; t += blockDim.x) {
        const int32_t idt = ids[t];
        const float2 ab = sums[t];
        float dist;
        if (fused) {
            dist = ab.x;
        } else if (HALF) {
            const float c = ab.x * inv;
            dist = KIND == kNegCos ? -c : 1.0f - c;
        } else {
            dist = distance_k<KIND>(ab.x, ab.y, qn);
        }
        // Fused keys carry the row in place of the ID until the band.
        k32[t] = idt < 0 ? kEmptyKey : make_key(f32_key(dist), fused ? (uint32_t)rows[t] : (uint32_t)idt);
    }
    __syncthreads();
    // The band: approximate distance at most ub (as a key's high word).
    uint32_t ub = 0xFFFFFFFFu;
    if (k < Cp && !all) {
        // ids is free until the band is collected: it holds the ties.
        const uint32_t kth = kth_high_word(k32, Cp, k, rs_hist, rs_sv, reinterpret_cast<uint32_t*>(ids));
        const double vk = (double)key_f32(kth);
        const double t = KIND == kEuclid ? vk * (1.0 + eps) / (1.0 - eps) : vk + 2.0 * eps;
        if (kth != 0xFFFFFFFFu && __builtin_isfinite(t) && t < 3.0e38) {
            // Rounded up: consecutive floats have consecutive keys.
            const float tf = (float)t;
            ub = f32_key(tf) + ((double)tf < t ? 1u : 0u);
        }
    }
    // Collect the band: ids[slot] = candidate index (the candidate IDs live
    // on in k32's low words).
    if (threadIdx.x == 0) rs_sv[2] = 0;
    __syncthreads();
    for (int t = threadIdx.x; t < Cp; t += blockDim.x) {
        const uint64_t key = k32[t];
        if (key != kEmptyKey && (uint32_t)(key >> 32) <= ub) ids[atomicAdd(&rs_sv[2], 1u)] = t;
    }
    __syncthreads();
    const int m = (int)rs_sv[2];
    if (BAND) {
        uint64_t* __restrict__ bq = band + (size_t)pos * Cp;
        for (int s = threadIdx.x; s < m; s += blockDim.x) {
            const int t = ids[s];
            const uint32_t idb = fused ? (uint32_t)fused_ids[rows[t]] : (uint32_t)k32[t];
            bq[s] = ((uint64_t)row_of(t) << 32) | idb;
        }
        if (threadIdx.x == 0) bcnt[pos] = (uint32_t)m;
        return;
    }
    // FP64 for the band.  FP64 keys carry no ID: the sort compares IDs on
    // equal keys.  Each slot's (row, ID) is parked in k64 first; k32 and
    // rows are then free and hold chunks of the band's rows, read with
    // coalesced loads, so that each FP64 sum reads its row from LDS.
    for (int s = threadIdx.x; s < m; s += blockDim.x) {
        const int t = ids[s];
        k64[s] = ((uint64_t)row_of(t) << 32) | (uint32_t)k32[t];
    }
    __syncthreads();
    // Staging: k32 and rows, 12 Cp bytes from byte 12 Cp (16-byte aligned
    // for Cp >= 4); R rows of 4d bytes per chunk, R * 4d <= 12 Cp.
    float* stage = reinterpret_cast<float*>(k32);
    const int R = ((d & 3) == 0 && Cp >= 4) ? (3 * Cp) / d : 0;
    if (R > 0) {
        const int d4 = d >> 2;
        for (int c0 = 0; c0 < m; c0 += R) {
            const int nr = min(R, m - c0);
            for (int i = threadIdx.x; i < nr * d4; i += blockDim.x) {
                const int rr = i / d4, j = i - rr * d4;
                const uint32_t row = (uint32_t)(k64[c0 + rr] >> 32);
                reinterpret_cast<float4*>(stage)[i] = reinterpret_cast<const float4*>(data + (size_t)row * d)[j];
            }
            __syncthreads();
            for (int s = c0 + threadIdx.x; s < c0 + nr; s += blockDim.x) {
                const int32_t id = (int32_t)(uint32_t)k64[s];
                k64[s] = f64_key(dist64_row(KIND, qv, stage + (size_t)(s - c0) * d, d, nq64));
                ids[s] = id;
            }
            __syncthreads();
        }
    } else {
        for (int s = threadIdx.x; s < m; s += blockDim.x) {
            const uint64_t v = k64[s];
            k64[s] = f64_key(dist64_row(KIND, qv, data + (size_t)(uint32_t)(v >> 32) * d, d, nq64));
            ids[s] = (int32_t)(uint32_t)v;
        }
        __syncthreads();
    }
    int n = 1;
    while (n < m) n <<= 1;
    if (n <= 64) {
        if (threadIdx.x < 64) {
            uint64_t kk = (int)threadIdx.x < m ? k64[threadIdx.x] : kEmptyKey;
            int32_t ii = (int)threadIdx.x < m ? ids[threadIdx.x] : 0x7FFFFFFF;
            wave_sort_pairs(kk, ii);
            // n <= Cp: the first n lanes hold the m entries and padding.
            if ((int)threadIdx.x < n) {
                k64[threadIdx.x] = kk;
                ids[threadIdx.x] = ii;
            }
        }
        __syncthreads();
    } else {
        for (int t = m + threadIdx.x; t < n; t += blockDim.x) {
            k64[t] = kEmptyKey;
            ids[t] = 0x7FFFFFFF;
        }
        bitonic_sort_pairs(k64, ids, n);
    }
    const float pad = KIND == kNegCos ? 2.0f : kPadDistance;
    for (int t = threadIdx.x; t < k; t += blockDim.x) {
        lsh_point_t p;
        if (t < m) {
            p.ID = ids[t];
            p.distance = (float)key_f64(k64[t]);
        } else {
            p.ID = -1;
            p.distance = pad;
        }
        out[(size_t)q * k + t] = p;
    }
}

// Smaller of two (key, ID) pairs.
__device__ inline bool pair_less(uint64_t ka, int32_t ia, uint64_t kb, int32_t ib) {
    return ka < kb || (ka == kb && ia < ib);
}

// Ascending bitonic merge of a bitonic sequence held one pair per lane.
__device__ inline void wave_bitonic_clean(uint64_t& k, int32_t& id) {
    const int lane = threadIdx.x & 63;
#pragma unroll
    for (int stride = 32; stride > 0; stride >>= 1) {
        const uint64_t ko = shfl_xor_u64(k, stride);
        const int32_t io = __shfl_xor(id, stride);
        const bool lower = (lane & stride) == 0;
        const bool other_less = pair_less(ko, io, k, id);
        if (lower == other_less) {
            k = ko;
            id = io;
        }
    }
}

// The band's FP64 recheck, one wave per query (k <= 64): 64 members at a
// time, each lane computes one member's FP64 distance (the oracle's
// operation order), the wave sorts them and merges them into the running 64
// best held one per lane (min against the reversed chunk, then a bitonic
// clean); the first k are written.
template <int KIND>
__global__ __launch_bounds__(64) void rerank_recheck_kernel(const float* __restrict__ data,
                                                            const float* __restrict__ queries, int d,
                                                            const uint64_t* __restrict__ band,
                                                            const uint32_t* __restrict__ bcnt, int Cp, int k,
                                                            lsh_point_t* __restrict__ out,
                                                            const int32_t* __restrict__ qlist,
                                                            const int32_t* __restrict__ band_ids) {
    const int lane = threadIdx.x;
    const int pos = blockIdx.x;
    const int q = qlist ? qlist[pos] : pos;
    const float* __restrict__ qv = queries + (size_t)q * d;
    const double nq64 = (KIND == kNegCos || KIND == kAngular) ? norm2_64(qv, d) : 0.0;
    const int m = (int)bcnt[pos];
    const uint64_t* __restrict__ bq = band + (size_t)pos * Cp;
    uint64_t bk = kEmptyKey;  // running best, ascending across lanes
    int32_t bi = 0x7FFFFFFF;
    for (int c0 = 0; c0 < m; c0 += 64) {
        uint64_t ck = kEmptyKey;
        int32_t ci = 0x7FFFFFFF;
        if (c0 + lane < m) {
            const uint64_t v = bq[c0 + lane];
            ci = band_ids ? band_ids[(uint32_t)(v >> 32)] : (int32_t)(uint32_t)v;
            ck = f64_key(dist64_row(KIND, qv, data + (size_t)(uint32_t)(v >> 32) * d, d, nq64));
        }
        wave_sort_pairs(ck, ci);
        // The 64 smallest of both: min(best[i], chunk[63 - i]) is bitonic.
        const uint64_t rk = shfl_xor_u64(ck, 63);
        const int32_t ri = __shfl_xor(ci, 63);
        if (pair_less(rk, ri, bk, bi)) {
            bk = rk;
            bi = ri;
        }
        wave_bitonic_clean(bk, bi);
    }
    const float pad = KIND == kNegCos ? 2.0f : kPadDistance;
    if (lane < k) {
        lsh_point_t p;
        if (bi != 0x7FFFFFFF) {
            p.ID = bi;
            p.distance = (float)key_f64(bk);
        } else {
            p.ID = -1;
            p.distance = pad;
        }
        out[(size_t)q * k + lane] = p;
    }
}

// One 16-lane group per row: out row = (float)(x * 2^14 / ||x||) rounded to
// FP16, zero padded to hs; padding positions (ids < 0) are zero rows.  A
// real row whose FP64 norm is zero or not finite sets *bad.
__global__ void unit_rows_half_kernel(const float* __restrict__ rows, const int32_t* __restrict__ ids,
                                      int64_t npos, int d, int hs, _Float16* __restrict__ out,
                                      uint32_t* __restrict__ bad) {
    const int gl = threadIdx.x & 15;
    const int64_t ngroups = (int64_t)gridDim.x * (blockDim.x >> 4);
    for (int64_t r = (int64_t)blockIdx.x * (blockDim.x >> 4) + (threadIdx.x >> 4); r < npos; r += ngroups) {
        const float* x = rows + (size_t)r * d;
        const bool real = ids[r] >= 0;
        double s = 0.0;
        for (int j = gl; j < d; j += 16) s = __builtin_fma((double)x[j], (double)x[j], s);
        for (int o = 8; o > 0; o >>= 1) s += __shfl_xor(s, o, 16);
        const bool ok = real && s > 0.0 && s <= 1.0e300;
        if (gl == 0 && real && !ok) atomicOr(bad, 1u);
        const double scale = ok ? (double)kHalfScale / __builtin_sqrt(s) : 0.0;
        for (int j = gl; j < hs; j += 16)
            out[(size_t)r * hs + j] = (_Float16)(j < d ? (float)((double)x[j] * scale) : 0.0f);
    }
}

// Per query (one wave each): the unit query scaled by 2^14, c = q * (2^14 /
// ||q||) in FP32, split into FP16 parts hi = (half)c and lo = (half)(c -
// hi) (c - hi is exact in FP32), 16 ksteps halves each, zero past d; a
// query whose norm is zero, not finite or outside [2^-60, 2^60] gets zeros
// (the filter rechecks all its candidates).
__global__ __launch_bounds__(256) void query_half_kernel(const float* __restrict__ q, int Q, int d, int ksteps,
                                                         _Float16* __restrict__ out, uint32_t* __restrict__ bad) {
    const int qi = (int)blockIdx.x * 4 + (int)(threadIdx.x >> 6);
    const int lane = threadIdx.x & 63;
    if (qi >= Q) return;
    const float* __restrict__ x = q + (size_t)qi * d;
    float s = 0.0f;
    for (int j = lane; j < d; j += 64) s = __builtin_fmaf(x[j], x[j], s);
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) s += __shfl_xor(s, o);
    const float n = __builtin_sqrtf(s);
    const bool ok = n >= 0x1p-60f && n <= 0x1p60f;
    const float r = ok ? kHalfScale / n : 0.0f;
    if (lane == 0) bad[qi] = ok ? 0u : 1u;
    const int kp = 16 * ksteps;
    _Float16* __restrict__ o = out + (size_t)qi * 2 * kp;
    for (int j = lane; j < kp; j += 64) {
        const float c = j < d ? x[j] * r : 0.0f;
        const _Float16 hi = (_Float16)c;
        o[j] = hi;
        o[kp + j] = (_Float16)(c - (float)hi);
    }
}

}  // namespace

double half_error_bound(int d) {
    // Per component, the FP16 image differs from 2^14 x / ||x|| by at most
    // (2^-11 + 2^-23) of its magnitude (two roundings, FP64 -> FP32 -> FP16)
    // plus 2^-14 (a subnormal flushed), i.e. 2^-28 of the unit row.  Over
    // the dot product with q that is (2^-11 + 2^-23) + sqrt(d) 2^-28 of
    // ||q||; the FP32 sums, norm and division add at most the FP32 bound,
    // and the filter's reciprocal (2^-14 / ||q||, then one product per
    // candidate in place of a division) one rounding more.
    const double u = 1.0 / 16777216.0;
    return 1.01 * (1.0 / 2048.0 + 1.0 / 8388608.0) + std::sqrt((double)d) * (1.0 / 268435456.0) +
           (2.0 * d + 17.0) * u;
}

int half_row_stride(int d) { return (d + 7) / 8 * 8; }

int fused_k_steps(int d) {
    const int ks = (d + 15) / 16;
    return ks >= 1 && ks <= 8 ? ks : 0;
}

double fused_error_bound(int d) {
    // The image's part as in half_error_bound.  The query's FP16 parts
    // differ from q 2^14 / ||q|| by the FP32 norm, quotient and product
    // (a relative (d/2 + 3) 2^-24, mostly a common factor), the split's
    // residue (2^-22 of each component) and, should the matrix cores flush
    // FP16 subnormals, parts below 2^-14 (2^-28 of the unit query each, for
    // the high and the low part).  The 2 x 16 ksteps exact FP16 products are
    // summed in FP32 in the matrix cores' order: at most 32 ksteps 2^-24 of
    // the sum of their magnitudes, at most the product of the norms; then
    // 1 - c rounds once.  Last, the 16-bit key floors (1 - c) 2^15: the
    // filter takes the midpoint of that step, 2^-16 either way.
    const double u = 1.0 / 16777216.0;
    const int ks = fused_k_steps(d);
    return 1.01 * (1.0 / 2048.0 + 1.0 / 8388608.0) + 3.0 * std::sqrt((double)d) * (1.0 / 268435456.0) +
           (32.0 * ks + 0.5 * d + 8.0) * u + 1.0 / 4194304.0 + 1.0 / 65536.0;
}

int fused_band_delta(int d) { return (int)std::floor(2.0 * fused_error_bound(d) * 32768.0); }

void launch_query_half(const float* q, int Q, int d, int ksteps, _Float16* out, hipStream_t s) {
    if (Q <= 0) return;
    LSH_REQUIRE(ksteps >= 1 && ksteps <= 8 && 16 * ksteps >= d, LSH_ERR_INTERNAL, "query parts: bad K-steps");
    hipLaunchKernelGGL(query_half_kernel, dim3((unsigned)((Q + 3) / 4)), dim3(256), 0, s, q, Q, d, ksteps, out,
                       query_half_flags(out, Q, ksteps));
    check_launch("query_half_kernel");
}

bool build_unit_rows_half(const float* rows, const int32_t* ids, int64_t npos, int d, DevBuf<_Float16>& out,
                          hipStream_t s) {
    const int hs = half_row_stride(d);
    // The filter addresses the image by 32-bit offsets in 16-byte pieces.
    if ((uint64_t)npos * (uint64_t)(hs / 8) > 0xFFFFFFFFull) {
        out.release();
        return false;
    }
    out.alloc((size_t)npos * hs);
    DevBuf<uint32_t> bad(1);
    LSH_HIP(hipMemsetAsync(bad.get(), 0, sizeof(uint32_t), s));
    if (npos > 0) {
        const unsigned blocks = (unsigned)std::min<int64_t>(65535, (npos + 15) / 16);
        hipLaunchKernelGGL(unit_rows_half_kernel, dim3(blocks), dim3(256), 0, s, rows, ids, npos, d, hs, out.get(),
                           bad.get());
        check_launch("unit_rows_half_kernel");
    }
    uint32_t h = 0;
    LSH_HIP(hipMemcpyAsync(&h, bad.get(), sizeof(uint32_t), hipMemcpyDeviceToHost, s));
    LSH_HIP(hipStreamSynchronize(s));
    if (h) out.release();
    return h == 0;
}

bool launch_rerank(const float* data, const float* queries, int d, int kind, const uint64_t* keys,
                   int key_stride, const int32_t* cand, int C, int Q, int k, double eps,
                   lsh_point_t* out, hipStream_t s, const int32_t* qorder, const int32_t* rowmap,
                   const _Float16* hdata, RerankBand* rb, bool recheck_too, const int32_t* rows_in,
                   int rows_stride, size_t hdata_elems, const int32_t* fused_ids) {
    if (Q <= 0) return false;
    LSH_REQUIRE(!rows_in || (!cand && rows_stride >= C), LSH_ERR_INTERNAL, "rerank: bad candidate rows");
    const int Cp = pow2_ceil(std::max(C, 2));
    LSH_REQUIRE(Cp <= kMaxTopK, LSH_ERR_INVALID, "candidate set larger than 4096");
    LSH_REQUIRE(!hdata || kind == kAngular || kind == kNegCos, LSH_ERR_INTERNAL,
                "rerank: the FP16 image serves the angular kinds only");
    LSH_REQUIRE(!fused_ids || (hdata && rb && k <= 64 && !cand && !rows_in), LSH_ERR_INTERNAL,
                "rerank: fused keys need the FP16 filter's split form");
    const int hs = hdata ? half_row_stride(d) : 0;
    // An image under 4 GiB is addressed by 32-bit byte offsets
    // (LSHGPU_RERANK_OFF64=1 takes the 64-bit path of larger images).
    const char* off64 = std::getenv("LSHGPU_RERANK_OFF64");
    const int img32 = hdata && hdata_elems > 0 && hdata_elems * sizeof(_Float16) <= 0xFFFFFFFFull &&
                              !(off64 && off64[0] == '1')
                          ? 1
                          : 0;
    // Split into the filter and the recheck when a band workspace is given
    // and k fits one wave.
    const bool split = rb != nullptr && k <= 64;
    uint64_t* band = nullptr;
    uint32_t* bcnt = nullptr;
    if (split) {
        rb->band.ensure((size_t)Q * Cp);
        rb->cnt.ensure((size_t)Q);
        band = rb->band.get();
        bcnt = rb->cnt.get();
    }
    const size_t lds = (size_t)Cp * ((split ? 1 : 2) * sizeof(uint64_t) + 2 * sizeof(int32_t)) +
                       (size_t)hs * sizeof(float);
    const int per = (Q + 7) / 8;
    auto go = [&](auto kernel) {
        set_dynamic_lds((const void*)kernel, lds);
        hipLaunchKernelGGL(kernel, dim3(qorder ? 8 * per : Q), dim3(256), lds, s, data, hdata, hs, queries, d, keys,
                           key_stride, cand, C, Cp, k, eps, out, qorder, Q, per, rowmap, band, bcnt, rows_in,
                           rows_stride, img32, fused_ids);
    };
    auto pick = [&](auto band_tag) {
        constexpr bool B = decltype(band_tag)::value;
        if (hdata) {
            if (kind == kAngular) go(rerank_kernel<kAngular, true, B>);
            else go(rerank_kernel<kNegCos, true, B>);
        } else {
            switch (kind) {
                case kNegCos: go(rerank_kernel<kNegCos, false, B>); break;
                case kAngular: go(rerank_kernel<kAngular, false, B>); break;
                case kGenJaccard: go(rerank_kernel<kGenJaccard, false, B>); break;
                case kEuclid: go(rerank_kernel<kEuclid, false, B>); break;
                default: fail(LSH_ERR_INTERNAL, "rerank: unknown distance kind");
            }
        }
    };
    if (split) pick(std::true_type{});
    else pick(std::false_type{});
    check_launch("rerank_kernel");
    if (split && recheck_too) launch_rerank_recheck(data, queries, d, kind, C, Q, k, out, s, qorder, *rb);
    return split;
}

void launch_rerank_recheck(const float* data, const float* queries, int d, int kind, int C, int Q, int k,
                           lsh_point_t* out, hipStream_t s, const int32_t* qorder, RerankBand& rb,
                           const int32_t* band_ids) {
    if (Q <= 0) return;
    const int Cp = pow2_ceil(std::max(C, 2));
    LSH_REQUIRE(k <= 64 && rb.band.size() >= (size_t)Q * Cp && rb.cnt.size() >= (size_t)Q, LSH_ERR_INTERNAL,
                "rerank recheck: no band for these queries");
    auto recheck = [&](auto kernel) {
        hipLaunchKernelGGL(kernel, dim3(Q), dim3(64), 0, s, data, queries, d, rb.band.get(), rb.cnt.get(), Cp, k,
                           out, qorder, band_ids);
    };
    switch (kind) {
        case kNegCos: recheck(rerank_recheck_kernel<kNegCos>); break;
        case kAngular: recheck(rerank_recheck_kernel<kAngular>); break;
        case kGenJaccard: recheck(rerank_recheck_kernel<kGenJaccard>); break;
        case kEuclid: recheck(rerank_recheck_kernel<kEuclid>); break;
        default: fail(LSH_ERR_INTERNAL, "rerank: unknown distance kind");
    }
    check_launch("rerank_recheck_kernel");
}

}  // namespace lshgpu
