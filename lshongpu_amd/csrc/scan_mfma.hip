// scan_mfma.hip -- Hamming scan on the CDNA4 matrix cores.
//
// A sketch bit b becomes the FP4 (e2m1) value s = +1 (b = 0) or -1 (b = 1);
// for two sketches of B bits, sum_k s_a[k] s_b[k] = B - 2 * Hamming(a, b),
// and the FP32 accumulation of these +-1 products is exact.  So the
// Hamming distances of a 32-query x 32-point tile are one chain of
// v_mfma_scale_f32_32x32x64_f8f6f4 (FP4 x FP4, unit scales) over the
// sketch, 64 bits per instruction.
//
// Operand layout: for K-step s (bits 64s .. 64s+63), lane l supplies row
// l%32 and the 32 bits of word 2s + l/32 of that row, one nibble per bit,
// in the low four registers of the operand.  A (queries) and B (points) use
// the same mapping, so the dot product pairs equal bit positions.
//
// Workgroup: 4 waves; each wave keeps the expanded fragments of MQ = 2 query
// tiles (64 queries) in VGPRs for the whole sketch, so a workgroup covers 256
// queries.  The workgroup walks point tiles of 32 points: the packed words
// of a tile (32 x W words, one coalesced read) are expanded once into LDS in
// fragment order and every wave reads its B fragments from there.  The
// epilogue compares each accumulator with the query's threshold (one float
// compare: dot >= B - 2*threshold); the rare survivors are checked against
// the exact 64-bit threshold and queued in an LDS pool with per-query LDS
// counters, and the pool is flushed to the global top-K state with one
// atomic per query, so no global round trip sits on the tile loop.  Ranges
// whose candidates are still dense (the first rounds) use the VALU scan.
#include "device_common.hpp"

namespace lshgpu {

namespace {

typedef int v4i __attribute__((ext_vector_type(4)));
typedef int v8i __attribute__((ext_vector_type(8)));
typedef float v16f __attribute__((ext_vector_type(16)));

constexpr int MQ = 2;                     // query tiles per wave
constexpr int WG_QUERIES = 4 * MQ * 32;   // queries per workgroup
constexpr int TPB = 256;

// 8 bits -> 8 FP4 nibbles: bit i -> nibble i = 0x2 (+1.0) or 0xA (-1.0).
__device__ inline uint32_t spread8(uint32_t x) {
    x &= 0xFFu;
    x = (x | (x << 12)) & 0x000F000Fu;
    x = (x | (x << 6)) & 0x03030303u;
    x = (x | (x << 3)) & 0x11111111u;
    return (x << 3) | 0x22222222u;
}

__device__ inline v4i expand_word(uint32_t w) {
    v4i r;
    r[0] = (int)spread8(w);
    r[1] = (int)spread8(w >> 8);
    r[2] = (int)spread8(w >> 16);
    r[3] = (int)spread8(w >> 24);
    return r;
}

// FP4 operands occupy the low four registers; the upper half is not read.
__device__ inline v8i widen(v4i a) { return __builtin_shufflevector(a, a, 0, 1, 2, 3, -1, -1, -1, -1); }

constexpr int POOL = 2048;  // queued candidates per workgroup

// KS = sketch words / 2 (K-steps of 64 bits).
template <int KS>
__global__ __launch_bounds__(TPB, 2) void hamming_mfma_kernel(const uint32_t* __restrict__ ds,
                                                              const uint32_t* __restrict__ qs, int Q,
                                                              int64_t p0, int64_t p1, int tiles_per_wg,
                                                              TopKView v) {
    constexpr int W = 2 * KS;
    constexpr float BITS = 32.0f * W;
    __shared__ v4i bfrag[KS * 64];          // expanded B fragments of the current tile
    __shared__ float spill[4][MQ * 16][64]; // per-wave accumulators for the append path
    __shared__ float tfs[WG_QUERIES];       // coarse thresholds: dot >= tfs[q] may pass
    __shared__ uint64_t lthr[WG_QUERIES];   // exact thresholds (keys)
    __shared__ uint32_t lcnt[WG_QUERIES];   // queued candidates per query since the last flush
    __shared__ uint32_t lbase[WG_QUERIES];  // global slot of the first of them
    __shared__ uint64_t pkey[POOL];         // queued keys
    __shared__ uint16_t pq[POOL];           // query (within the workgroup) of each queued key
    __shared__ uint32_t npool;

    const int lane = threadIdx.x & 63;
    const int wave = threadIdx.x >> 6;
    const int qblock = blockIdx.x * WG_QUERIES;

    // Resident A fragments of this wave's query tiles.
    v4i afrag[MQ][KS];
#pragma unroll
    for (int t = 0; t < MQ; ++t) {
        const int qt = qblock + (wave * MQ + t) * 32;
        const uint32_t* __restrict__ qr = qs + (size_t)(qt + (lane & 31)) * W + (lane >> 5);
#pragma unroll
        for (int s = 0; s < KS; ++s) afrag[t][s] = expand_word(qr[2 * s]);
    }
    {
        const int qi = qblock + (int)threadIdx.x;
        float f = __builtin_inff();  // rows past Q never pass
        uint64_t th = 0;
        if (qi < Q) {
            th = v.thr[qi];
            const uint32_t hi = (uint32_t)(th >> 32);
            f = hi == 0xFFFFFFFFu ? -__builtin_inff() : BITS - 2.0f * (float)hi;
        }
        tfs[threadIdx.x] = f;
        lthr[threadIdx.x] = th;
        lcnt[threadIdx.x] = 0;
        if (threadIdx.x == 0) npool = 0;
    }

    // Moves the queued candidates to the global top-K state (all threads):
    // per-query ranks from LDS atomics, one global atomic per query, scatter.
    auto flush = [&]() {
        const uint32_t n = min(npool, (uint32_t)POOL);
        uint32_t rank[POOL / TPB];
#pragma unroll
        for (int j = 0; j < POOL / TPB; ++j) {
            const uint32_t e = threadIdx.x + j * TPB;
            rank[j] = e < n ? atomicAdd(&lcnt[pq[e]], 1u) : 0u;
        }
        __syncthreads();
        const uint32_t c = lcnt[threadIdx.x];
        if (c) lbase[threadIdx.x] = atomicAdd(&v.cnt[qblock + threadIdx.x], c);
        __syncthreads();
#pragma unroll
        for (int j = 0; j < POOL / TPB; ++j) {
            const uint32_t e = threadIdx.x + j * TPB;
            if (e < n) {
                const uint32_t q = pq[e];
                const uint32_t slot = lbase[q] + rank[j];
                if (slot < (uint32_t)v.newcap) v.keys[(size_t)(qblock + q) * v.stride + v.K + slot] = pkey[e];
            }
        }
        __syncthreads();
        lcnt[threadIdx.x] = 0;
        if (threadIdx.x == 0) npool = 0;
        __syncthreads();
    };

    const int64_t ntiles = (p1 - p0 + 31) / 32;
    const int64_t t0 = (int64_t)blockIdx.y * tiles_per_wg;
    const int64_t t1 = min(ntiles, t0 + tiles_per_wg);
    // Staging: wave w fills the fragment rows of K-steps w, w+4, ...; lane l
    // of row s holds word 2s + l/32 of point l%32, so the LDS writes of a wave
    // are 64 consecutive 16-byte slots.  The words of the next tile are
    // loaded into registers while the current tile is multiplied.
    constexpr int SPW = (KS + 3) / 4;  // K-steps staged per wave
    const int sp = lane & 31, sw = lane >> 5;
    uint32_t next[SPW];
    auto load_tile = [&](int64_t tile) {
        const int64_t pi = p0 + tile * 32 + sp;
#pragma unroll
        for (int j = 0; j < SPW; ++j) {
            const int st = wave + 4 * j;
            next[j] = (st < KS && pi < p1) ? ds[(size_t)pi * W + 2 * st + sw] : 0u;
        }
    };
    if (t0 < t1) load_tile(t0);
    for (int64_t tile = t0; tile < t1; ++tile) {
        const int64_t pbase = p0 + tile * 32;
        __syncthreads();  // the previous tile's reads of bfrag and its queueing are complete
        const uint32_t queued = npool;
#pragma unroll
        for (int j = 0; j < SPW; ++j) {
            const int st = wave + 4 * j;
            if (st < KS) bfrag[st * 64 + lane] = expand_word(next[j]);
        }
        if (tile + 1 < t1) load_tile(tile + 1);
        __syncthreads();
        if (queued > POOL / 2) flush();

        v16f acc[MQ];
#pragma unroll
        for (int t = 0; t < MQ; ++t) acc[t] = (v16f){};
        // Four K-steps per group; the scheduling fence keeps the B fragments
        // of one group live at a time.
#pragma unroll
        for (int s0 = 0; s0 < KS; s0 += 4) {
            v4i b[4];
#pragma unroll
            for (int i = 0; i < 4; ++i) b[i] = bfrag[(s0 + i) * 64 + lane];
#pragma unroll
            for (int i = 0; i < 4; ++i)
#pragma unroll
                for (int t = 0; t < MQ; ++t)
                    acc[t] = __builtin_amdgcn_mfma_scale_f32_32x32x64_f8f6f4(
                        widen(afrag[t][s0 + i]), widen(b[i]), acc[t], 4, 4, 0, 0x7F7F7F7F, 0, 0x7F7F7F7F);
            __builtin_amdgcn_sched_barrier(0);
        }

        const int64_t pi = pbase + (lane & 31);
        const bool pvalid = pi < p1;
        // Bit t*16+r of `cand`: accumulator (t, r) of this lane may beat its
        // query's threshold.
        uint32_t cand = 0;
#pragma unroll
        for (int t = 0; t < MQ; ++t)
#pragma unroll
            for (int r = 0; r < 16; ++r)
                cand |= (acc[t][r] >= tfs[(wave * MQ + t) * 32 + (r >> 2) * 8 + (lane >> 5) * 4 + (r & 3)] ? 1u : 0u)
                        << (t * 16 + r);
        if (!pvalid) cand = 0;
        if (__ballot(cand != 0)) {
            // Rare: visit the accumulator registers some lane flagged (their
            // union over the wave), in wave-uniform order through LDS.
#pragma unroll
            for (int t = 0; t < MQ; ++t)
#pragma unroll
                for (int r = 0; r < 16; ++r) spill[wave][t * 16 + r][lane] = acc[t][r];
            uint32_t any = cand;
#pragma unroll
            for (int o = 1; o < 64; o <<= 1) any |= (uint32_t)__shfl_xor((int)any, o);
            // Exact check against the 64-bit threshold, then one pool
            // reservation for the whole wave.
            uint32_t pass = 0;
            for (uint32_t m = any; m; m &= m - 1) {
                const int bit = __ffs((int)m) - 1;
                const int qrel = (wave * MQ + (bit >> 4)) * 32 + ((bit & 15) >> 2) * 8 + (lane >> 5) * 4 + (bit & 3);
                const uint32_t h = (uint32_t)(((int)BITS - (int)spill[wave][bit][lane]) >> 1);
                if (((cand >> bit) & 1u) && make_key(h, (uint32_t)pi) < lthr[qrel]) pass |= 1u << bit;
            }
            const uint32_t mine = (uint32_t)__popc(pass);
            uint32_t incl = mine;
#pragma unroll
            for (int o = 1; o < 64; o <<= 1) {
                const uint32_t t = __shfl_up(incl, o);
                if (lane >= o) incl += t;
            }
            const uint32_t wtotal = __shfl(incl, 63);
            uint32_t base = 0;
            if (lane == 0 && wtotal) base = atomicAdd(&npool, wtotal);
            base = __shfl(base, 0);
            uint32_t e = base + incl - mine;
            for (uint32_t m = pass; m; m &= m - 1, ++e) {
                const int bit = __ffs((int)m) - 1;
                const int qrel = (wave * MQ + (bit >> 4)) * 32 + ((bit & 15) >> 2) * 8 + (lane >> 5) * 4 + (bit & 3);
                if (e < (uint32_t)POOL) {
                    const uint32_t h = (uint32_t)(((int)BITS - (int)spill[wave][bit][lane]) >> 1);
                    pkey[e] = make_key(h, (uint32_t)pi);
                    pq[e] = (uint16_t)qrel;
                } else {
                    v.ovf[qblock + qrel] = 1u;  // dropped: the query is redone
                }
            }
        }
    }
    __syncthreads();
    flush();
}

template <int KS>
void launch_ks(const uint32_t* ds, const uint32_t* qs, int Q, int64_t p0, int64_t p1, const TopKView& v,
               hipStream_t s) {
    const int qblocks = (Q + WG_QUERIES - 1) / WG_QUERIES;
    const int64_t ntiles = (p1 - p0 + 31) / 32;
    // Enough workgroups to fill 256 CUs twice over, at least 4 tiles each.
    int64_t slices = std::max<int64_t>(1, std::min<int64_t>(ntiles / 4, (2048 + qblocks - 1) / qblocks));
    slices = std::min<int64_t>(slices, 65535);
    const int tiles_per_wg = (int)((ntiles + slices - 1) / slices);
    slices = (ntiles + tiles_per_wg - 1) / tiles_per_wg;
    hipLaunchKernelGGL((hamming_mfma_kernel<KS>), dim3(qblocks, (unsigned)slices), dim3(TPB), 0, s, ds, qs, Q,
                       p0, p1, tiles_per_wg, v);
    check_launch("hamming_mfma_kernel");
}

}  // namespace

int hamming_mfma_query_rows() { return WG_QUERIES; }

bool hamming_mfma_supported(int words) { return words == 8 || words == 16 || words == 24 || words == 32; }

void launch_hamming_mfma(const uint32_t* ds, const uint32_t* qs, int words, int Q, int64_t p0, int64_t p1,
                         const TopKView& v, hipStream_t s) {
    if (p1 <= p0 || Q <= 0) return;
    switch (words) {
        case 8: launch_ks<4>(ds, qs, Q, p0, p1, v, s); break;
        case 16: launch_ks<8>(ds, qs, Q, p0, p1, v, s); break;
        case 24: launch_ks<12>(ds, qs, Q, p0, p1, v, s); break;
        case 32: launch_ks<16>(ds, qs, Q, p0, p1, v, s); break;
        default: fail(LSH_ERR_INTERNAL, "MFMA Hamming scan needs 8, 16, 24 or 32 words");
    }
}

}  // namespace lshgpu
