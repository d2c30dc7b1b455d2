// scan_mfma.hip -- Hamming scan on the CDNA4 matrix cores.
//
// A sketch bit b becomes the FP4 (e2m1) value s = +1 (b = 0) or -1 (b = 1);
// for two sketches of B bits, sum_k s_a[k] s_b[k] = B - 2 * Hamming(a, b),
// and the FP32 accumulation of these +-1 products is exact.  So the
// Hamming distances of a 32-point x 32-query tile are one chain of
// v_mfma_scale_f32_32x32x64_f8f6f4 (FP4 x FP4, unit scales) over the
// sketch, 64 bits per instruction.
//
// Operand layout: for K-step s (bits 64s .. 64s+63), lane l supplies row
// l%32 and the 32 bits of word 2s + l/32 of that row, one nibble per bit,
// in the low four registers of the operand.  Both operands use the same
// mapping, so the dot product pairs equal bit positions.  With the points as
// A and the queries as B, accumulator register r of lane l holds point row
// (r/4)*8 + (l/32)*4 + r%4 and query l%32: a lane sees one query, whose
// threshold it keeps in a register.
//
// The data sketches are kept in HBM already expanded to FP4 in operand
// order, [tile of 32 points][K-step][lane][16 bytes] (512 B per point at
// 1024 bits, 4x the packed sketch), so staging a tile is a copy: each wave
// moves a quarter of it with direct global-to-LDS loads (16 bytes a lane).
// A workgroup holds 4 waves of 32 queries each (their B operands resident
// in VGPRs) and walks a slice of point tiles through two LDS buffers: the
// loads of the next tile are in flight while the current one is multiplied,
// with one barrier per tile.  Workgroups are numbered so that the query
// blocks of one slice run on the same XCD (blocks are dealt to the 8 XCDs
// round-robin), which keeps one copy of the slice in that XCD's L2.
//
// Epilogue: the 16 accumulators of a lane are compared with its query's
// threshold (dot > B - 2*h_thr, i.e. Hamming < h_thr); the lane masks stay in
// scalar registers and only the registers with hits append (Hamming, ID) keys
// to the wave's LDS queue, which is flushed to the global top-K state with
// one atomic per query (or, when full, spills to the global slots).  Ranges
// whose candidates are still dense (the first rounds) use the VALU scan.
#include "fp4.hpp"

namespace lshgpu {

namespace {

using fp4::expand_word;
using fp4::v16f;
using fp4::v4i;
using fp4::v8i;
using fp4::widen;

constexpr int MQ = 1;                      // query tiles per wave
constexpr int WAVES = 4;
constexpr int WAVE_QUERIES = MQ * 32;      // at most 64: one query per lane in the flush
constexpr int WG_QUERIES = WAVES * WAVE_QUERIES;
constexpr int TPB = WAVES * 64;
constexpr int POOLW = 256;                 // queued candidates per wave
constexpr int TPS = 1;                     // point tiles per staging step
constexpr int OCC = 3;                     // workgroups per CU (registers and LDS)

__device__ inline void wave_sync() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// KS = sketch words / 2 (K-steps of 64 bits).
template <int KS>
__global__ __launch_bounds__(TPB, OCC) void hamming_mfma_kernel(const v4i* __restrict__ frag,
                                                              const uint32_t* __restrict__ qs, int Q,
                                                              int64_t p0, int64_t p1, int64_t tile_begin,
                                                              int64_t tile_end, int tiles_per_wg, int qblocks,
                                                              int slices, TopKView v) {
    constexpr int W = 2 * KS;
    constexpr int BITS = 32 * W;
    constexpr int SPW = (KS + WAVES - 1) / WAVES;  // K-steps staged per wave
    __shared__ v4i bufA[TPS * KS * 64];            // point tiles in flight (A operands); two
    __shared__ v4i bufB[TPS * KS * 64];            // arrays so reads of one overlap loads of the other
    __shared__ float tf[WG_QUERIES];               // pass iff dot > tf (Hamming < threshold distance)
    __shared__ uint32_t lcnt[WAVES][WAVE_QUERIES];
    __shared__ uint32_t lbase[WAVES][WAVE_QUERIES];
    __shared__ uint64_t pkey[WAVES][POOLW];
    __shared__ uint8_t pq[WAVES][POOLW];

    // XCD-aware numbering: block L runs on XCD L % 8; the blocks of one XCD
    // take the query blocks of slices xcd, xcd + 8, ...
    const int L = blockIdx.x;
    const int j = L >> 3;
    const int qb = j % qblocks;
    const int slice = (L & 7) + 8 * (j / qblocks);
    if (slice >= slices) return;
    const int lane = threadIdx.x & 63;
    const int wave = threadIdx.x >> 6;
    const int qwave = qb * WG_QUERIES + wave * WAVE_QUERIES;

    if (lane < WAVE_QUERIES) {
        const int q = qwave + lane;
        float f = __builtin_inff();  // rows past Q never pass
        if (q < Q) {
            const uint32_t h = (uint32_t)(v.thr[q] >> 32);
            f = h == 0xFFFFFFFFu ? -__builtin_inff() : (float)BITS - 2.0f * (float)h;
        }
        tf[wave * WAVE_QUERIES + lane] = f;
        lcnt[wave][lane] = 0;
    }
    // Query fragments (B operands), resident for the whole launch.
    v4i qfrag[MQ][KS];
#pragma unroll
    for (int t = 0; t < MQ; ++t) {
        const uint32_t* __restrict__ qr = qs + (size_t)(qwave + t * 32 + (lane & 31)) * W + (lane >> 5);
#pragma unroll
        for (int s = 0; s < KS; ++s) qfrag[t][s] = expand_word(qr[2 * s]);
    }
    __syncthreads();
    // Lane l holds column l%32 (a query) of each 32x32 product; register r
    // holds row (r/4)*8 + (l/32)*4 + r%4 (a point of the tile).
    float thr[MQ];
#pragma unroll
    for (int t = 0; t < MQ; ++t) thr[t] = tf[wave * WAVE_QUERIES + t * 32 + (lane & 31)];
    const int rowbase = (lane >> 5) * 4;

    // The wave's queue holds `queued` entries; it is moved to the global
    // top-K state with LDS ranks per query and one global atomic per query.
    uint32_t queued = 0;
    auto flush = [&]() {
        wave_sync();
        const uint32_t n = min(queued, (uint32_t)POOLW);
        uint32_t rank[POOLW / 64];
#pragma unroll
        for (int k = 0; k < POOLW / 64; ++k) {
            const uint32_t e = lane + 64 * k;
            rank[k] = e < n ? atomicAdd(&lcnt[wave][pq[wave][e]], 1u) : 0u;
        }
        wave_sync();
        if (lane < WAVE_QUERIES) {
            const uint32_t c = lcnt[wave][lane];
            if (c) lbase[wave][lane] = atomicAdd(&v.cnt[qwave + lane], c);
        }
        wave_sync();
#pragma unroll
        for (int k = 0; k < POOLW / 64; ++k) {
            const uint32_t e = lane + 64 * k;
            if (e < n) {
                const uint32_t q = pq[wave][e];
                const uint32_t slot = lbase[wave][q] + rank[k];
                if (slot < (uint32_t)v.newcap) v.keys[(size_t)(qwave + q) * v.stride + v.K + slot] = pkey[wave][e];
            }
        }
        if (lane < WAVE_QUERIES) lcnt[wave][lane] = 0;
        queued = 0;
        wave_sync();
    };

    const int64_t t0 = tile_begin + (int64_t)slice * tiles_per_wg;
    const int64_t t1 = min(tile_end, t0 + tiles_per_wg);
    // Wave w stages K-steps w, w+4, ... of a point tile with direct
    // global-to-LDS loads (16 bytes per lane, the operand layout).
    // Staging step u holds tiles t0 + u*TPS .. + TPS-1 (those below t1).
    auto stage = [&](int64_t u, v4i* buf) {
#pragma unroll
        for (int i = 0; i < TPS; ++i) {
            const int64_t tile = t0 + u * TPS + i;
            if (tile >= t1) break;
#pragma unroll
            for (int k = 0; k < SPW; ++k) {
                const int st = wave + WAVES * k;
                if (st < KS)
                    __builtin_amdgcn_global_load_lds(
                        (const void*)(frag + ((size_t)tile * KS + st) * 64 + lane),
                        (__attribute__((address_space(3))) void*)&buf[(i * KS + st) * 64], 16, 0, 0);
            }
        }
    };
    // One point tile: products from the staged operands, then the epilogue.
    auto process = [&](int64_t tile, const v4i* buf) {
        v16f acc[MQ];
#pragma unroll
        for (int t = 0; t < MQ; ++t) acc[t] = (v16f){};
#pragma unroll
        for (int s0 = 0; s0 < KS; s0 += 4) {
            v4i a[4];
#pragma unroll
            for (int i = 0; i < 4; ++i) a[i] = buf[(s0 + i) * 64 + lane];
#pragma unroll
            for (int i = 0; i < 4; ++i)
#pragma unroll
                for (int t = 0; t < MQ; ++t)
                    acc[t] = __builtin_amdgcn_mfma_scale_f32_32x32x64_f8f6f4(
                        widen(a[i]), widen(qfrag[t][s0 + i]), acc[t], 4, 4, 0, 0x7F7F7F7F, 0, 0x7F7F7F7F);
            __builtin_amdgcn_sched_barrier(0);
        }

        // A pair passes when its Hamming distance is below the query's
        // threshold distance.  Equal distances never pass: every key of the
        // best list has a smaller ID than the points of this range.  Each
        // register is one compare and one ballot; the rare passing lanes
        // take slots of the wave's queue in lane order.
        const uint32_t pbase = (uint32_t)(tile * 32);
        auto epilogue = [&](uint32_t vmask) {
            // All compares first (their lane masks stay in scalar
            // registers), then a scalar walk over the registers with hits.
            uint64_t m[MQ][16];
            uint64_t any = 0;
#pragma unroll
            for (int t = 0; t < MQ; ++t)
#pragma unroll
                for (int r = 0; r < 16; ++r) {
                    m[t][r] = __ballot(acc[t][r] > thr[t] && ((vmask >> r) & 1u));
                    any |= m[t][r];
                }
            if (!any) return;
#pragma unroll
            for (int t = 0; t < MQ; ++t)
#pragma unroll
                for (int r = 0; r < 16; ++r) {
                    const uint64_t mr = m[t][r];
                    if (!mr) continue;
                    if ((mr >> lane) & 1ull) {
                        const uint32_t e = queued + (uint32_t)__popcll(mr & ((1ull << lane) - 1ull));
                        const int qrel = t * 32 + (lane & 31);
                        const uint32_t pi = pbase + (uint32_t)((r >> 2) * 8 + rowbase + (r & 3));
                        const uint64_t key = make_key((uint32_t)((BITS - (int)acc[t][r]) >> 1), pi);
                        if (e < (uint32_t)POOLW) {
                            pkey[wave][e] = key;
                            pq[wave][e] = (uint8_t)qrel;
                        } else {
                            // Queue full (a dense tile): append straight to the
                            // query's global slots; the select kernel counts
                            // any excess as an overflow.
                            const uint32_t slot = atomicAdd(&v.cnt[qwave + qrel], 1u);
                            if (slot < (uint32_t)v.newcap) v.keys[(size_t)(qwave + qrel) * v.stride + v.K + slot] = key;
                        }
                    }
                    queued += (uint32_t)__popcll(mr);
                }
        };
        if ((int64_t)pbase >= p0 && (int64_t)pbase + 32 <= p1) {
            epilogue(0xFFFFu);
        } else {
            // Rows of a tile cut by the range: bit r set when row
            // (r/4)*8 + rowbase + r%4 lies in [p0, p1).
            uint32_t vm = 0;
#pragma unroll
            for (int r = 0; r < 16; ++r) {
                const int64_t pi = (int64_t)pbase + (r >> 2) * 8 + rowbase + (r & 3);
                vm |= (pi >= p0 && pi < p1 ? 1u : 0u) << r;
            }
            asm volatile("" : "+v"(vm));  // keep the mask; do not re-derive it per register
            epilogue(vm);
        }
        if (queued > POOLW / 2) flush();
    };

    // Staging steps alternate between the two buffers: after the barrier
    // that publishes step u, the loads of step u+1 go to the other buffer and
    // overlap the products of step u.
    const int64_t steps = (t1 - t0 + TPS - 1) / TPS;
    auto process_step = [&](int64_t u, const v4i* buf) {
#pragma unroll
        for (int i = 0; i < TPS; ++i) {
            const int64_t tile = t0 + u * TPS + i;
            if (tile < t1) process(tile, buf + i * KS * 64);
        }
    };
    if (steps > 0) stage(0, bufA);
    for (int64_t u = 0; u < steps; u += 2) {
        asm volatile("s_waitcnt vmcnt(0)" ::: "memory");
        __syncthreads();  // step u is in bufA; bufB is free
        if (u + 1 < steps) stage(u + 1, bufB);
        process_step(u, bufA);
        if (u + 1 >= steps) break;
        asm volatile("s_waitcnt vmcnt(0)" ::: "memory");
        __syncthreads();  // step u+1 is in bufB; bufA is free
        if (u + 2 < steps) stage(u + 2, bufA);
        process_step(u + 1, bufB);
    }
    flush();
}

int wg_slots() {
    // OCC workgroups of this kernel fit a CU (LDS and registers).
    static const int slots = [] {
        int dev = 0, cus = 0;
        if (hipGetDevice(&dev) != hipSuccess ||
            hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess || cus <= 0)
            cus = 256;
        return OCC * cus;
    }();
    return slots;
}

template <int KS>
void launch_ks(const v4i* frag, const uint32_t* qs, int Q, int64_t p0, int64_t p1, const TopKView& v,
               hipStream_t s) {
    const int qblocks = (Q + WG_QUERIES - 1) / WG_QUERIES;
    const int64_t tile_begin = p0 / 32, tile_end = (p1 + 31) / 32;
    const int64_t ntiles = tile_end - tile_begin;
    // At least ~2048 workgroups of at least 2*TPS tiles each; among the
    // group counts up to twice the minimum, the one whose workgroups fill
    // the last wave of the machine best.
    const int64_t gmax = std::max<int64_t>(1, ntiles / (2 * TPS * 8));
    const int64_t gmin = std::min<int64_t>(gmax, std::max<int64_t>(1, (2048 + 8 * qblocks - 1) / (8 * qblocks)));
    int64_t groups = gmin;
    double best = -1.0;
    const int slots = wg_slots();
    for (int64_t g = gmin; g <= std::min<int64_t>(gmax, 2 * gmin); ++g) {
        const int64_t b = 8 * (int64_t)qblocks * g;
        const double eff = (double)b / (double)(((b + slots - 1) / slots) * slots);
        if (eff > best + 1e-9) {
            best = eff;
            groups = g;
        }
    }
    int64_t slices = 8 * groups;
    int64_t per = (ntiles + slices - 1) / slices;
    per = (per + TPS - 1) / TPS * TPS;
    slices = (ntiles + per - 1) / per;
    groups = (slices + 7) / 8;
    const int64_t blocks = 8 * (int64_t)qblocks * groups;
    LSH_REQUIRE(blocks < (int64_t)1 << 31, LSH_ERR_UNSUPPORTED, "matrix-core scan grid too large");
    hipLaunchKernelGGL((hamming_mfma_kernel<KS>), dim3((unsigned)blocks), dim3(TPB), 0, s, frag, qs, Q, p0, p1,
                       tile_begin, tile_end, (int)per, qblocks, (int)slices, v);
    check_launch("hamming_mfma_kernel");
}

__global__ void expand_fp4_kernel(const uint32_t* __restrict__ ds, int64_t N, int wpad, int KS, int64_t total,
                                  v4i* __restrict__ frag) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= total) return;
    const int lane = (int)(i & 63);
    const int s = (int)((i >> 6) % KS);
    const int64_t T = (i >> 6) / KS;
    const int64_t row = T * 32 + (lane & 31);
    const uint32_t x = row < N ? ds[(size_t)row * wpad + 2 * s + (lane >> 5)] : 0u;
    frag[i] = expand_word(x);
}

}  // namespace

int hamming_mfma_query_rows() { return WG_QUERIES; }

bool hamming_mfma_supported(int words) { return words == 8 || words == 16 || words == 24 || words == 32; }

size_t hamming_fp4_words(int64_t N, int words) {
    const int64_t tiles = (N + 31) / 32 + 2;  // the scan reads one tile past a range
    return (size_t)tiles * (size_t)(words / 2) * 64 * 4;
}

void launch_expand_fp4(const uint32_t* ds, int64_t N, int words, uint32_t* frag, hipStream_t s) {
    const int KS = words / 2;
    const int64_t total = (int64_t)hamming_fp4_words(N, words) / 4;
    hipLaunchKernelGGL(expand_fp4_kernel, dim3((unsigned)((total + 255) / 256)), dim3(256), 0, s, ds, N, words, KS,
                       total, reinterpret_cast<v4i*>(frag));
    check_launch("expand_fp4_kernel");
}

void launch_hamming_mfma(const uint32_t* frag, const uint32_t* qs, int words, int Q, int64_t p0, int64_t p1,
                         const TopKView& v, hipStream_t s) {
    if (p1 <= p0 || Q <= 0) return;
    LSH_REQUIRE(frag, LSH_ERR_INTERNAL, "matrix-core scan needs the FP4 image");
    const v4i* f = reinterpret_cast<const v4i*>(frag);
    switch (words) {
        case 8: launch_ks<4>(f, qs, Q, p0, p1, v, s); break;
        case 16: launch_ks<8>(f, qs, Q, p0, p1, v, s); break;
        case 24: launch_ks<12>(f, qs, Q, p0, p1, v, s); break;
        case 32: launch_ks<16>(f, qs, Q, p0, p1, v, s); break;
        default: fail(LSH_ERR_INTERNAL, "MFMA Hamming scan needs 8, 16, 24 or 32 words");
    }
}

}  // namespace lshgpu
