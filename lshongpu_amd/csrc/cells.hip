// cells.hip -- exact pruning of the matrix-core Hamming pass with pivot cells.
//
// Build: P pivots (every N/P-th data row) split the data into cells: each
// point belongs to its nearest pivot by Hamming distance (ties to the lower
// pivot).  The points of a cell are laid out by ascending distance to the
// pivot, in 32-point tiles (the last tile of a cell padded), and each tile
// records the smallest and largest distance of its points to the pivot.
// The FP4 image of the data is written again in this order.
//
// Search: Hamming distance is a metric, so for a query q, a pivot p and a
// point x, H(q, x) >= |H(q, p) - H(x, p)|.  A tile whose points lie at
// [lo, hi] from its pivot is at least max(H(q, p) - hi, lo - H(q, p), 0) from
// q; when that bound is not below q's threshold distance h (the pass keeps
// pairs with H < h) no point of the tile can be kept, and the tile is
// skipped.  A search takes the distances of all queries to all pivots (one
// matrix-core product), then:
//   1. Threshold.  Every point x of a tile t of cell j satisfies
//      H(q, x) <= H(q, p_j) + hi_t, so the K-th smallest of these bounds over
//      the tiles of the query's nearest cells (those whose pivots are nearest
//      and hold at least K points, plus a margin) bounds the K-th nearest
//      distance from above: T = that bound + 1 is a valid threshold.
//   2. Pass.  Per cell, the queries whose bound on the cell is below their
//      threshold are listed and cut into chunks of 32; one wave per chunk
//      walks the cell's tiles, tests each tile's bound against the chunk's
//      queries, and multiplies the tiles that can hold a hit, appending the
//      pairs below the threshold.
// Every pair below the threshold is appended, so the K best of the appended
// pairs are the exact top-K, as after a pass over all points.
#include <cstdio>
#include <cstdlib>

#include "engine.hpp"
#include "fp4.hpp"

namespace lshgpu {

namespace {

using fp4::expand_word;
using fp4::v16f;
using fp4::v4i;

constexpr int kKeyShift = 11;  // cell key = pivot << 11 | distance (distance <= 1024)
constexpr int kMaxPivots = 8192;  // keys below 2^24 (the counting sort's table limit)
constexpr uint32_t kDistMask = (1u << kKeyShift) - 1u;

// LSHGPU_DEBUG_CELLS=1 prints the radii, thresholds and list sizes of each
// cell search to stderr.
bool debug_cells() {
    static const bool on = [] {
        const char* e = std::getenv("LSHGPU_DEBUG_CELLS");
        return e && e[0] == '1';
    }();
    return on;
}

unsigned blocks_for(size_t n, int per_block) {
    return (unsigned)std::max<size_t>(1, (n + per_block - 1) / per_block);
}

// Nearest pivot of each row of `rows` (packed [n][W]) among the P pivots
// of the FP4 image pfrag: best[row] = pivot << 11 | distance, ties to the
// lower pivot.  One wave per 32 rows.
template <int KS>
__global__ __launch_bounds__(256) void pivot_assign_kernel(const uint32_t* __restrict__ rows, int64_t n,
                                                           const v4i* __restrict__ pfrag, int P,
                                                           uint32_t* __restrict__ best) {
    constexpr int W = 2 * KS;
    constexpr int BITS = 32 * W;
    const int lane = threadIdx.x & 63;
    const int64_t row0 = ((int64_t)blockIdx.x * 4 + (threadIdx.x >> 6)) * 32;
    if (row0 >= n) return;
    const int64_t row = row0 + (lane & 31);
    const bool live = row < n;
    v4i b[KS];
#pragma unroll
    for (int s = 0; s < KS; ++s) b[s] = expand_word(live ? rows[(size_t)row * W + 2 * s + (lane >> 5)] : 0u);
    float bdot = -__builtin_inff();
    int bp = 0;
    const int ptiles = (P + 31) / 32;
    for (int pt = 0; pt < ptiles; ++pt) {
        v16f acc = {};
#pragma unroll
        for (int s = 0; s < KS; ++s) acc = fp4::mma(pfrag[((size_t)pt * KS + s) * 64 + lane], b[s], acc);
#pragma unroll
        for (int r = 0; r < 16; ++r) {
            const int p = pt * 32 + fp4::acc_row(r, lane);
            if (p < P && acc[r] > bdot) {
                bdot = acc[r];
                bp = p;
            }
        }
    }
    const float od = __shfl_xor(bdot, 32);
    const int op = __shfl_xor(bp, 32);
    if (od > bdot || (od == bdot && op < bp)) {
        bdot = od;
        bp = op;
    }
    if (lane < 32 && live) best[row] = ((uint32_t)bp << kKeyShift) | (uint32_t)((BITS - (int)bdot) >> 1);
}

// D[q][p] (row stride Pp, a multiple of 32) = Hamming distance of query q
// to pivot p.  Each wave takes 32 queries (blockIdx.x, wave), their FP4
// operands expanded once in registers, and the workgroup a group of kGroup
// pivot tiles (blockIdx.y): each tile's FP4 image (the A operand, KS KB) is
// copied into LDS once for the four waves with direct global-to-LDS loads,
// the next tile while the current one is multiplied.  A 32x32 block of
// distances goes through LDS so that every lane stores 32 contiguous bytes
// of one query's row.
constexpr int kGroup = 8;
template <int KS>
__global__ __launch_bounds__(256) void pivot_matrix_kernel(const uint32_t* __restrict__ qs, int Q,
                                                           const v4i* __restrict__ pfrag, int P, int Pp,
                                                           uint16_t* __restrict__ D, uint32_t* __restrict__ zero,
                                                           int nzero, TopKView tv) {
    // The first workgroup also zeroes zero[0, nzero) (a search's counters),
    // and the whole grid clears the fresh top-K view tv when it has one (tv.Q
    // > 0: thresholds empty, counts and overflow flags zero).
    if (blockIdx.x == 0 && blockIdx.y == 0)
        for (int i = threadIdx.x; i < nzero; i += blockDim.x) zero[i] = 0u;
    {
        const int nthreads = gridDim.x * gridDim.y * blockDim.x;
        for (int i = (blockIdx.y * gridDim.x + blockIdx.x) * blockDim.x + threadIdx.x; i < tv.Q; i += nthreads) {
            tv.thr[i] = kEmptyKey;
            tv.cnt[i] = 0u;
            tv.ovf[i] = 0u;
        }
    }
    constexpr int W = 2 * KS;
    constexpr int BITS = 32 * W;
    __shared__ uint32_t tb[4][32][33];
    __shared__ v4i pa[2][KS * 64];  // a pivot tile's A operand, two buffers
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int q0 = (blockIdx.x * 4 + wave) * 32;
    // A wave past the queries still stages its share of the tiles and meets
    // the barriers; it multiplies and stores nothing.
    const bool live = q0 < Q;
    const int q = q0 + (lane & 31);
    v4i b[KS];
#pragma unroll
    for (int s = 0; s < KS; ++s) b[s] = expand_word(q < Q ? qs[(size_t)q * W + 2 * s + (lane >> 5)] : 0u);
    const int ptiles = (P + 31) / 32;
    const int pt0 = blockIdx.y * kGroup;
    const int pt1 = min(ptiles, pt0 + kGroup);
    uint32_t(*t)[33] = tb[wave];
    // Wave w copies K-steps w, w + 4, ... of pivot tile pt (1 KB each).
    auto stage = [&](int pt, int buf) {
        for (int st = wave; st < KS; st += 4)
            __builtin_amdgcn_global_load_lds((const void*)(pfrag + ((size_t)pt * KS + st) * 64 + lane),
                                             (__attribute__((address_space(3))) void*)&pa[buf][st * 64], 16, 0, 0);
    };
    if (pt0 < pt1) stage(pt0, 0);
    for (int pt = pt0; pt < pt1; ++pt) {
        const int buf = (pt - pt0) & 1;
        // This wave's copies of tile pt have landed; the barrier makes every
        // wave's visible and frees the other buffer (read for tile pt - 1).
        asm volatile("s_waitcnt vmcnt(0)" ::: "memory");
        __syncthreads();
        if (pt + 1 < pt1) stage(pt + 1, buf ^ 1);
        if (!live) continue;
        v16f acc = {};
#pragma unroll
        for (int s = 0; s < KS; ++s) acc = fp4::mma(pa[buf][s * 64 + lane], b[s], acc);
#pragma unroll
        for (int r = 0; r < 16; ++r) t[lane & 31][fp4::acc_row(r, lane)] = (uint32_t)((BITS - (int)acc[r]) >> 1);
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_ACQ_REL, "wavefront");
        const int half = lane >> 5;
        uint32_t w[8];
#pragma unroll
        for (int i = 0; i < 8; ++i)
            w[i] = t[lane & 31][half * 16 + 2 * i] | (t[lane & 31][half * 16 + 2 * i + 1] << 16);
        if (q < Q) {
            uint4* dst = reinterpret_cast<uint4*>(D + (size_t)q * Pp + pt * 32 + half * 16);
            dst[0] = make_uint4(w[0], w[1], w[2], w[3]);
            dst[1] = make_uint4(w[4], w[5], w[6], w[7]);
        }
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_ACQ_REL, "wavefront");
    }
}

// Sorted position r (perm from the counting sort by cell key) -> slot in the
// cell layout.
__global__ void cell_place_kernel(const int32_t* __restrict__ perm, const uint32_t* __restrict__ ckey, int64_t N,
                                  const uint32_t* __restrict__ cstart, const uint32_t* __restrict__ ctile,
                                  int32_t* __restrict__ ids, uint16_t* __restrict__ pdist) {
    for (int64_t r = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; r < N; r += (int64_t)gridDim.x * blockDim.x) {
        const int32_t i = perm[r];
        const uint32_t key = ckey[i];
        const uint32_t j = key >> kKeyShift;
        const int64_t pos = (int64_t)ctile[j] * 32 + (r - (int64_t)cstart[j]);
        ids[pos] = i;
        pdist[pos] = (uint16_t)(key & kDistMask);
    }
}

// Per tile: min | max << 11 | rows << 22 of the valid rows (a prefix).
__global__ void cell_tile_info_kernel(const int32_t* __restrict__ ids, const uint16_t* __restrict__ pdist,
                                      int64_t ntiles, uint32_t* __restrict__ tinfo) {
    for (int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; t < ntiles;
         t += (int64_t)gridDim.x * blockDim.x) {
        uint32_t lo = kDistMask, hi = 0, nv = 0;
        for (int i = 0; i < 32; ++i) {
            if (ids[t * 32 + i] < 0) continue;
            const uint32_t d = pdist[t * 32 + i];
            lo = d < lo ? d : lo;
            hi = d > hi ? d : hi;
            ++nv;
        }
        tinfo[t] = nv ? (lo | (hi << kKeyShift) | (nv << 22)) : 0u;
    }
}

// Per cell: distance range of its points to the pivot (first and last point
// of the cell in layout order); an empty cell gets min > max.  crec packs
// (size, min | max << 16) for the threshold kernel's one-load reads.
__global__ void cell_range_kernel(const uint32_t* __restrict__ cstart, const uint32_t* __restrict__ ctile,
                                  const uint16_t* __restrict__ pdist, int P, uint16_t* __restrict__ cmin,
                                  uint16_t* __restrict__ cmax, uint2* __restrict__ crec) {
    for (int j = blockIdx.x * blockDim.x + threadIdx.x; j < P; j += gridDim.x * blockDim.x) {
        const uint32_t n = cstart[j + 1] - cstart[j];
        uint32_t lo = 1, hi = 0;
        if (n != 0) {
            lo = pdist[(size_t)ctile[j] * 32];
            hi = pdist[(size_t)ctile[j] * 32 + n - 1];
        }
        cmin[j] = (uint16_t)lo;
        cmax[j] = (uint16_t)hi;
        crec[j] = make_uint2(n, lo | (hi << 16));
    }
}

__global__ void cell_starts_kernel(const uint32_t* __restrict__ off, int P, uint32_t* __restrict__ cstart) {
    for (int j = blockIdx.x * blockDim.x + threadIdx.x; j <= P; j += gridDim.x * blockDim.x)
        cstart[j] = off[(size_t)j << kKeyShift];
}

// Packed sketches in cell order: position pos takes data row ids[pos], its
// even words first, then its odd words (zero for padding, which the tile
// row counts exclude), so the matrix-core lane that supplies word 2s + h of
// a row reads the 4*W/2 contiguous bytes of half h.
__global__ void pack_cells_kernel(const uint32_t* __restrict__ ds, int wpad, const int32_t* __restrict__ ids,
                                  int64_t npos, uint32_t* __restrict__ sk) {
    const int64_t total = npos * wpad;
    const int KS = wpad / 2;
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < total;
         i += (int64_t)gridDim.x * blockDim.x) {
        const int64_t pos = i / wpad;
        const int w = (int)(i - pos * wpad);
        const int src = w < KS ? 2 * w : 2 * (w - KS) + 1;
        const int32_t row = ids[pos];
        sk[i] = row >= 0 ? ds[(size_t)row * wpad + src] : 0u;
    }
}

// Float rows in cell order (zero rows for padding) and the position of
// every data row.
__global__ void cell_rows_kernel(const float* __restrict__ src, int d, const int32_t* __restrict__ ids,
                                 int64_t npos, float* __restrict__ dst, int32_t* __restrict__ pos_of) {
    const int64_t total = npos * d;
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < total;
         i += (int64_t)gridDim.x * blockDim.x) {
        const int64_t pos = i / d;
        const int c = (int)(i - pos * d);
        const int32_t row = ids[pos];
        dst[i] = row >= 0 ? src[(size_t)row * d + c] : 0.0f;
        if (c == 0 && row >= 0) pos_of[row] = (int32_t)pos;
    }
}

__device__ inline uint32_t lower_bound_to(uint32_t dq, uint32_t lo, uint32_t hi) {
    return dq > hi ? dq - hi : (lo > dq ? lo - dq : 0u);
}

// Smallest bin b with sum(hist[0..b]) >= K (nbins when the total is below
// K), by the first wave of the workgroup.
__device__ int kth_bin(const uint32_t* hist, int nbins, uint32_t K) {
    const int lane = threadIdx.x & 63;
    uint32_t carry = 0;
    for (int base = 0; base < nbins; base += 64) {
        const int b = base + lane;
        uint32_t inc = b < nbins ? hist[b] : 0u;
#pragma unroll
        for (int o = 1; o < 64; o <<= 1) {
            const uint32_t x = __shfl_up(inc, o);
            if (lane >= o) inc += x;
        }
        const uint64_t hit = __ballot(b < nbins && carry + inc >= K);
        if (hit) return base + __ffsll((unsigned long long)hit) - 1;
        carry += __shfl(inc, 63);
    }
    return nbins;
}

// Slot of this lane among the wave's lanes with `pass`, counted from a
// shared counter (one atomic per wave).  Every lane of the wave must call it.
__device__ inline uint32_t wave_slot(bool pass, uint32_t* counter) {
    const unsigned long long m = __ballot(pass);
    if (m == 0) return 0;
    const int lane = threadIdx.x & 63;
    const int leader = __ffsll((long long)m) - 1;
    uint32_t base = 0;
    if (lane == leader) base = atomicAdd(counter, (uint32_t)__popcll(m));
    base = __shfl(base, leader);
    return base + (uint32_t)__popcll(m & ((1ull << lane) - 1ull));
}

// Exclusive prefix of v over the NT threads of the workgroup; total = the
// sum.  ws holds NT / 64 words.  Every thread must call it.
template <int NT>
__device__ inline uint32_t block_exclusive_scan(uint32_t v, uint32_t* ws, uint32_t& total) {
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    uint32_t inc = v;
#pragma unroll
    for (int o = 1; o < 64; o <<= 1) {
        const uint32_t x = __shfl_up(inc, o);
        if (lane >= o) inc += x;
    }
    if (lane == 63) ws[w] = inc;
    __syncthreads();
    uint32_t pre = 0, tot = 0;
#pragma unroll
    for (int i = 0; i < NT / 64; ++i) {
        const uint32_t x = ws[i];
        pre += i < w ? x : 0u;
        tot += x;
    }
    __syncthreads();
    total = tot;
    return pre + inc - v;
}

// kth_bin by the whole workgroup: each thread sums a run of bins, a block
// scan places the K-th count in one run, whose thread walks it.  *out
// (shared) receives the bin; every thread must call it.
template <int NT>
__device__ inline int block_kth_bin(const uint32_t* hist, int nbins, uint32_t K, uint32_t* ws, int* out) {
    const int per = (nbins + NT - 1) / NT;
    const int b0 = (int)threadIdx.x * per, b1 = min(b0 + per, nbins);
    uint32_t sum = 0;
    for (int b = b0; b < b1; ++b) sum += hist[b];
    uint32_t total;
    const uint32_t pre = block_exclusive_scan<NT>(sum, ws, total);
    if (pre < K && pre + sum >= K) {
        uint32_t run = pre;
        for (int b = b0; b < b1; ++b) {
            run += hist[b];
            if (run >= K) {
                *out = b;
                break;
            }
        }
    }
    if (threadIdx.x == 0 && total < K) *out = nbins;
    __syncthreads();
    return *out;
}

// Threshold of each query (one workgroup per query):
//  - its nearest non-empty cell (kept for the query order);
//  - the radius R at which the nearest cells hold K points, from a histogram
//    of cell sizes over the 256 distances above the nearest cell's (all
//    cells when that window holds fewer than K points); the cells within R
//    are counted into cnt1 for the first pass;
//  - an upper bound on the K-th distance: the K-th smallest D + max (each
//    cell's points lie within its max of the pivot, so within D + max of
//    the query), weighted by cell size, over the cells within R +
//    kRadiusMargin: thr = (bound + 1, 0);
//  - per query, the lists of its first-pass cells (qcells1, n1) and of the
//    cells whose bound is below that threshold (qcells2, n2; rows of P).
// Thread i of the NT-thread workgroup owns cells i, i + NT, ... (CPT of
// them, P <= NT CPT): their distance and packed record are loaded once, in
// one batch, and every step reads them from registers.
constexpr int kRadiusMargin = 16;
constexpr int kWindow = 256;
constexpr int kHistBins = 128;  // first-pass histogram bins per query (8 bits each)
constexpr int kBoundBins = 512;  // bins of the D + max bound histogram (the last holds every larger bound)
template <int CPT, int NT>
__global__ __launch_bounds__(NT) void cell_threshold_kernel(const uint16_t* __restrict__ D, int Q, int P, int Pp,
                                                            const uint2* __restrict__ crec, int bits, int K,
                                                            uint64_t* __restrict__ thr, uint16_t* __restrict__ R,
                                                            uint32_t* __restrict__ nearest,
                                                            uint32_t* __restrict__ cnt1,
                                                            int32_t* __restrict__ qcells1, uint32_t* __restrict__ n1,
                                                            int32_t* __restrict__ qcells2,
                                                            uint32_t* __restrict__ n2,
                                                            uint32_t* __restrict__ hist) {
    __shared__ uint32_t h1[kWindow];
    __shared__ uint32_t h2[kBoundBins];
    __shared__ uint32_t ws[NT / 64];
    __shared__ int s_bin;
    __shared__ uint32_t best, s_n1, s_n2;
    const int q = blockIdx.x;
    const int nb2 = min(2 * bits + 1, kBoundBins);  // bound histogram bins; larger bounds share the last
    const uint16_t* __restrict__ drow = D + (size_t)q * Pp;
    uint32_t dq[CPT], sz[CPT], rg[CPT];
#pragma unroll
    for (int i = 0; i < CPT; ++i) {
        const int j = (int)threadIdx.x + NT * i;
        const bool in = j < P;
        const int jj = in ? j : 0;  // every load in bounds (P >= 1); values masked
        const uint32_t a = drow[jj];
        const uint2 r = crec[jj];
        dq[i] = in ? a : 0u;
        sz[i] = in ? r.x : 0u;
        rg[i] = in ? r.y : 1u;  // min 1 > max 0: empty
    }
    for (int b = threadIdx.x; b < kBoundBins; b += NT) {
        if (b < kWindow) h1[b] = 0;
        h2[b] = 0;
    }
    // The query's first-pass histogram starts at zero (the first pass runs
    // after this kernel).
    for (int b = threadIdx.x; b < kHistBins; b += NT) hist[(size_t)q * kHistBins + b] = 0u;
    if (threadIdx.x == 0) {
        best = 0xFFFFFFFFu;
        s_n1 = 0;
        s_n2 = 0;
    }
    __syncthreads();
    uint32_t mine = 0xFFFFFFFFu;  // (distance, pivot) of the nearest non-empty cell
#pragma unroll
    for (int i = 0; i < CPT; ++i)
        if (sz[i]) mine = min(mine, (dq[i] << 13) | (uint32_t)(threadIdx.x + NT * i));
    atomicMin(&best, mine);
    __syncthreads();
    const uint32_t dmin = best == 0xFFFFFFFFu ? 0u : best >> 13;
    if (threadIdx.x == 0) nearest[q] = best == 0xFFFFFFFFu ? 0u : (best & 0x1FFFu);
#pragma unroll
    for (int i = 0; i < CPT; ++i) {
        const uint32_t dd = dq[i] - dmin;
        if (sz[i] && dd < (uint32_t)kWindow) atomicAdd(&h1[dd], sz[i]);
    }
    __syncthreads();
    const int r1 = block_kth_bin<NT>(h1, kWindow, (uint32_t)K, ws, &s_bin);
    const int rr = r1 < kWindow ? (int)dmin + r1 : bits;
    if (threadIdx.x == 0) R[q] = (uint16_t)min(rr, bits);
    const int rm = rr + kRadiusMargin;
    const int r0 = rr;
    int32_t* __restrict__ q1 = qcells1 + (size_t)q * P;
#pragma unroll
    for (int i = 0; i < CPT; ++i) {
        const int j = (int)threadIdx.x + NT * i;
        const bool pass = sz[i] != 0 && (int)dq[i] <= rm;
        const bool first = pass && (int)dq[i] <= r0;
        const uint32_t s1 = wave_slot(first, &s_n1);
        if (first) {
            atomicAdd(&cnt1[j], 1u);
            q1[s1] = j;
        }
        if (pass) atomicAdd(&h2[min(dq[i] + (rg[i] >> 16), (uint32_t)(nb2 - 1))], sz[i]);
    }
    __syncthreads();
    int ub = block_kth_bin<NT>(h2, nb2, (uint32_t)K, ws, &s_bin);
    // A K-th bound in the clamped last bin is only known to be at least its
    // index: the loosest bound, 2 bits (every pair is within it), is used.
    if (ub == nb2 - 1 && nb2 < 2 * bits + 1) ub = 2 * bits;
    else if (ub == nb2) ub = 2 * bits + 1;
    if (threadIdx.x == 0) thr[q] = ub > 2 * bits ? kEmptyKey : make_key((uint32_t)ub + 1u, 0u);
    // The cells the second pass can need: the final threshold is at most this
    // one, so the cells whose bound is below it are a superset.
    const uint32_t h = ub > 2 * bits ? 0xFFFFFFFFu : (uint32_t)ub + 1u;
    int32_t* __restrict__ q2 = qcells2 + (size_t)q * P;
#pragma unroll
    for (int i = 0; i < CPT; ++i) {
        const int j = (int)threadIdx.x + NT * i;
        const uint32_t lo = rg[i] & 0xFFFFu, hi = rg[i] >> 16;
        const bool keep = lo <= hi && lower_bound_to(dq[i], lo, hi) < h;
        const uint32_t s2 = wave_slot(keep, &s_n2);
        if (keep) q2[s2] = j;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        n1[q] = s_n1;
        n2[q] = s_n2;
    }
}

// The queries grouped by nearest cell (a counting sort by one workgroup,
// counts in LDS): order[] lists the queries of cell 0, then of cell 1, ...;
// the order inside a cell is not fixed.
// One 1024-thread workgroup; cnt is LDS of P words.
__device__ void order_by_cell(const uint32_t* __restrict__ nearest, int Q, int P, int32_t* __restrict__ order,
                              uint32_t* cnt) {
    __shared__ uint32_t ws[16];
    for (int j = threadIdx.x; j < P; j += 1024) cnt[j] = 0;
    __syncthreads();
    for (int i = threadIdx.x; i < Q; i += 1024) atomicAdd(&cnt[min(nearest[i], (uint32_t)(P - 1))], 1u);
    __syncthreads();
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    uint32_t carry = 0;
    for (int c0 = 0; c0 < P; c0 += 1024) {
        const int j = c0 + (int)threadIdx.x;
        const uint32_t a = j < P ? cnt[j] : 0u;
        uint32_t inc = a;
#pragma unroll
        for (int o = 1; o < 64; o <<= 1) {
            const uint32_t x = __shfl_up(inc, o);
            if (lane >= o) inc += x;
        }
        if (lane == 63) ws[w] = inc;
        __syncthreads();
        uint32_t pre = 0, tot = 0;
#pragma unroll
        for (int i = 0; i < 16; ++i) {
            pre += i < w ? ws[i] : 0u;
            tot += ws[i];
        }
        if (j < P) cnt[j] = carry + pre + inc - a;
        carry += tot;
        __syncthreads();
    }
    for (int i = threadIdx.x; i < Q; i += 1024)
        order[atomicAdd(&cnt[min(nearest[i], (uint32_t)(P - 1))], 1u)] = i;
}

// A chunk is up to 32 queries of one cell over up to kChunkTiles of its
// tiles, so that no wave walks a long cell alone.
constexpr int kChunkTiles = 8;
__device__ inline uint32_t cell_parts(const uint32_t* ctile, int j) {
    return (ctile[j + 1] - ctile[j] + kChunkTiles - 1) / kChunkTiles;
}

// coff = exclusive prefix of cnt, choff = exclusive prefix of the cells'
// chunks (ceil(cnt/32) query groups x cell_parts tile ranges), both with a
// final total entry; choff[P+1] and *work_out = the tile products those
// chunks can make (query groups x tiles, summed); *dense (may be null) = 1
// when that exceeds dense_limit; cursor (= cnt for the placing pass) zeroed;
// cmap[c] = the cell of chunk c.  By one 1024-thread workgroup.
__device__ void cell_offsets(uint32_t* __restrict__ cnt, int P, const uint32_t* __restrict__ ctile,
                             uint32_t* __restrict__ coff, uint32_t* __restrict__ choff, uint32_t* __restrict__ cmap,
                             uint32_t* __restrict__ work_out, uint32_t* __restrict__ dense, uint32_t dense_limit) {
    __shared__ uint32_t ws[3][16];
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    uint32_t carry_a = 0, carry_b = 0, carry_w = 0;
    for (int c0 = 0; c0 < P; c0 += 1024) {
        const int j = c0 + threadIdx.x;
        const uint32_t a = j < P ? cnt[j] : 0u;
        const uint32_t groups = (a + 31) / 32;
        const uint32_t b = groups * (j < P ? cell_parts(ctile, j) : 0u);
        uint32_t wk = groups ? groups * (ctile[j + 1] - ctile[j]) : 0u;
#pragma unroll
        for (int o = 32; o > 0; o >>= 1) wk += __shfl_xor(wk, o);
        if (lane == 0) ws[2][w] = wk;
        uint32_t ia = a, ib = b;
#pragma unroll
        for (int o = 1; o < 64; o <<= 1) {
            const uint32_t ta = __shfl_up(ia, o), tb = __shfl_up(ib, o);
            if (lane >= o) {
                ia += ta;
                ib += tb;
            }
        }
        if (lane == 63) {
            ws[0][w] = ia;
            ws[1][w] = ib;
        }
        __syncthreads();
        uint32_t pa = 0, pb = 0, ta = 0, tb = 0, tw = 0;
#pragma unroll
        for (int q = 0; q < 16; ++q) {
            pa += q < w ? ws[0][q] : 0u;
            pb += q < w ? ws[1][q] : 0u;
            ta += ws[0][q];
            tb += ws[1][q];
            tw += ws[2][q];
        }
        if (j < P) {
            const uint32_t c0 = carry_b + pb + ib - b;
            coff[j] = carry_a + pa + ia - a;
            choff[j] = c0;
            cnt[j] = 0;
            for (uint32_t c = 0; c < b; ++c) cmap[c0 + c] = (uint32_t)j;
        }
        carry_a += ta;
        carry_b += tb;
        carry_w += tw;
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        coff[P] = carry_a;
        choff[P] = carry_b;
        choff[P + 1] = carry_w;
        *work_out = carry_w;
        if (dense) *dense = carry_w > dense_limit ? 1u : 0u;
    }
}

__global__ __launch_bounds__(1024) void cell_offsets_kernel(uint32_t* __restrict__ cnt, int P,
                                                            const uint32_t* __restrict__ ctile,
                                                            uint32_t* __restrict__ coff,
                                                            uint32_t* __restrict__ choff,
                                                            uint32_t* __restrict__ cmap,
                                                            uint32_t* __restrict__ work_out,
                                                            uint32_t* __restrict__ dense, uint32_t dense_limit) {
    cell_offsets(cnt, P, ctile, coff, choff, cmap, work_out, dense, dense_limit);
}

// The first pass's offsets (workgroup 0) and, in the same launch, the query
// order by nearest cell (workgroup 1, when launched): both need only the
// threshold kernel's output.
__global__ __launch_bounds__(1024) void offsets_and_order_kernel(uint32_t* __restrict__ cnt, int P,
                                                                 const uint32_t* __restrict__ ctile,
                                                                 uint32_t* __restrict__ coff,
                                                                 uint32_t* __restrict__ choff,
                                                                 uint32_t* __restrict__ cmap,
                                                                 uint32_t* __restrict__ work_out,
                                                                 const uint32_t* __restrict__ nearest, int Q,
                                                                 int32_t* __restrict__ order) {
    extern __shared__ uint32_t ocnt[];  // P, for the order
    if (blockIdx.x == 0) cell_offsets(cnt, P, ctile, coff, choff, cmap, work_out, nullptr, 0u);
    else order_by_cell(nearest, Q, P, order, ocnt);
}

// Coarse distance histogram of the exact first pass: bin = distance / 8,
// the last bin also holding the distances past it.  In LDS two 16-bit
// counters share a word (a chunk adds at most 32 pairs per query per tile,
// so no counter reaches 2^16); the workgroup flushes them to the global
// histogram at the end of its chunk.
static_assert(32 * kChunkTiles < 65536, "a chunk's pair counts must fit the 16-bit histogram counters");
__device__ inline int hist_bin(int d) { return min(d >> 3, kHistBins - 1); }

// Fused appending pass: at most kMaxFuseSteps K-steps of 16 halves (rows of
// up to 128 halves), and the 16-bit approximate angular distance of a pair
// from its product sum p (unit rows and unit query, both scaled by 2^14):
// floor((1 - p 2^-28) 2^15), clamped to [0, 65535].
constexpr int kMaxFuseSteps = 8;
typedef _Float16 h8 __attribute__((ext_vector_type(8)));
__device__ inline uint32_t fused_q16(float p) {
    const float t = (1.0f - p * 0x1p-28f) * 32768.0f;
    return t <= 0.0f ? 0u : (t >= 65535.0f ? 65535u : (uint32_t)t);
}

// One workgroup per chunk of up to 32 queries of one cell (chunks taken in a
// grid-stride loop; their count is on the device).  The chunk's query
// sketches are expanded once into LDS as the B operand; the tiles that the
// bounds leave are dealt to the four waves in turn, each wave loading the
// packed point sketches of its next tile while it multiplies the current
// one (the A operand expanded through a byte table in LDS).  kHist: counts
// the distances of every pair in hist[q][bin] (the first pass); otherwise
// appends the pairs below each query's threshold to v.  Bounded to 96
// registers, so that five workgroups (20 waves) share a CU.
//
// kFuse (appending pass, angular index with an FP16 image of the unit rows,
// v.fused_keys): each tile's pairs also get their approximate angular
// distance on the matrix cores, the tile's FP16 rows (chalf, hs halves a
// row, zero padded) against the queries' FP16 high and low parts (q16,
// query_half_kernel: 2 x 16 ksteps halves a query, staged in dynamic LDS
// as the B operands), and an appended key's high word becomes (distance <<
// 16 | q), q = floor((1 - c) * 2^15) clamped to 16 bits, with c = the
// product sum * 2^-28: the rerank's filter then reads the distance from the
// key instead of gathering the row.  The tile's rows are read once per
// chunk of up to 32 queries.  Bounded to 128 registers (four workgroups a
// CU).
template <int KS, bool kHist, bool kFuse>
__global__ __launch_bounds__(256, kFuse ? 4 : 5) void cell_scan_kernel(const uint32_t* __restrict__ sk,
                                                        const int32_t* __restrict__ ids,
                                                        const uint32_t* __restrict__ tinfo,
                                                        const uint32_t* __restrict__ ctile,
                                                        const uint32_t* __restrict__ coff,
                                                        const uint32_t* __restrict__ choff,
                                                        const uint32_t* __restrict__ cmap, int P, int Pp,
                                                        const int32_t* __restrict__ list,
                                                        const uint16_t* __restrict__ D,
                                                        const uint32_t* __restrict__ qs, TopKView v,
                                                        uint32_t* __restrict__ hist,
                                                        const uint32_t* __restrict__ skip,
                                                        const _Float16* __restrict__ chalf, int hs,
                                                        const _Float16* __restrict__ q16, int ksteps) {
    constexpr int W = 2 * KS;
    constexpr int BITS = 32 * W;
    constexpr int HW = kHistBins / 2;  // LDS words per query
    constexpr int LH = kHist ? 32 * HW : 1;
    constexpr int NV = KS / 4;         // uint4 loads per lane per tile
    static_assert(kChunkTiles <= 64, "a chunk's tiles must fit one 64-bit mask");
    __shared__ v4i sb[KS * 64];        // B operand: step s of lane l at [s * 64 + l]
    __shared__ uint32_t lhist[LH];
    __shared__ uint32_t lut[256];      // byte -> its 8 FP4 nibbles
    // kFuse: [2][ksteps][64] FP16 B operands (high parts, then low parts).
    extern __shared__ h8 qb[];
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63, col = lane & 31;
    lut[threadIdx.x] = fp4::spread8(threadIdx.x);  // read after the chunk's first barrier
    const uint32_t nchunks = (skip && *skip) ? 0u : choff[P];
    for (uint32_t c = blockIdx.x; c < nchunks; c += gridDim.x) {
        const uint32_t j = cmap[c];
        const uint32_t parts = cell_parts(ctile, j);
        const uint32_t local = c - choff[j];
        const uint32_t q0 = coff[j] + 32 * (local / parts);
        const uint32_t m = min(32u, coff[j + 1] - q0);
        const int q = (uint32_t)col < m ? list[q0 + col] : -1;
        uint32_t h = 0;
        float tf = __builtin_inff();  // idle columns never pass
        uint32_t dq = 0;
        if (q >= 0) {
            h = kHist ? 0xFFFFFFFFu : (uint32_t)(v.thr[q] >> 32);
            tf = h == 0xFFFFFFFFu ? -__builtin_inff() : (float)BITS - 2.0f * (float)h;
            dq = D[(size_t)q * Pp + j];
        }
        for (int st = wave; st < KS; st += 4)
            sb[st * 64 + lane] = fp4::expand_word(q >= 0 ? qs[(size_t)q * W + 2 * st + (lane >> 5)] : 0u);
        if (kFuse)
            for (int st = wave; st < 2 * ksteps; st += 4) {
                // Step st % ksteps of part st / ksteps: halves 16 (st % ksteps)
                // + 8 (lane / 32) .. + 7 of the part.
                h8 z = {};
                if (q >= 0)
                    z = *reinterpret_cast<const h8*>(q16 + (size_t)q * (32 * ksteps) + (st / ksteps) * (16 * ksteps) +
                                                     16 * (st % ksteps) + 8 * (lane >> 5));
                qb[st * 64 + lane] = z;
            }
        if (kHist)
            for (int e = threadIdx.x; e < LH; e += 256) lhist[e] = 0;
        const uint32_t t0 = ctile[j] + (local % parts) * kChunkTiles;
        const uint32_t nt = min((uint32_t)kChunkTiles, ctile[j + 1] - t0);
        const uint32_t myinfo = (uint32_t)lane < nt ? tinfo[t0 + lane] : 0u;
        uint64_t needm = nt == 64 ? ~0ull : ((1ull << nt) - 1ull);
        if (!kHist) {
            uint64_t nm = 0;
            for (uint32_t tt = 0; tt < nt; ++tt) {
                const uint32_t info = (uint32_t)__shfl((int)myinfo, (int)tt);
                const uint32_t tlo = info & kDistMask, thi = (info >> kKeyShift) & kDistMask;
                if (__ballot(q >= 0 && lower_bound_to(dq, tlo, thi) < h)) nm |= 1ull << tt;
            }
            needm = nm;
        }
        // This wave's share: the needed tiles of rank wave, wave + 4, ...
        uint64_t mine = 0;
        {
            uint64_t mm = needm;
            for (int r = 0; mm; ++r) {
                const uint64_t low = mm & (~mm + 1ull);
                if ((r & 3) == wave) mine |= low;
                mm ^= low;
            }
        }
        __syncthreads();  // sb, lhist and lut ready
        // Lane l's half of point row l%32 of tile t: W/2 words (and, when
        // appending, the ID of that point).
        auto load = [&](uint32_t t, uint4* x, int32_t& id) {
            const uint4* src = reinterpret_cast<const uint4*>(sk + ((size_t)t * 32 + col) * W + (lane >> 5) * KS);
    #pragma unroll
            for (int i = 0; i < NV; ++i) x[i] = src[i];
            if (!kHist && !v.row_keys) id = ids[(size_t)t * 32 + col];
        };
        // Appends of one tile: slots from the count atomic's old value `ret`
        // (lanes < 32; `other` = the upper half's hits), IDs from the tile's
        // ID lanes, or with v.row_keys the points' positions in the layout
        // (the tile's first is pos0) in their place.
        // dist: the 16 distances, two 16-bit halves per word; dqs (kFuse):
        // their 16-bit approximate angular distances, packed the same way.
        auto emit = [&](const uint32_t* dist, const uint32_t* dqs, uint32_t pass, uint32_t ret, uint32_t other,
                        int32_t tid, uint32_t pos0) {
            uint32_t base = (uint32_t)__shfl((int)ret, col);
            if (lane >= 32) base += other;
            const size_t row0 = (size_t)q * v.stride + v.K;
    #pragma unroll
            for (int r = 0; r < 16; ++r) {
                if (!__ballot((pass >> r) & 1u)) continue;
                const int32_t id = v.row_keys ? (int32_t)(pos0 + (uint32_t)fp4::acc_row(r, lane))
                                              : __shfl(tid, fp4::acc_row(r, lane));
                if ((pass >> r) & 1u) {
                    uint32_t hi = (dist[r >> 1] >> (16 * (r & 1))) & 0xFFFFu;
                    if (kFuse) hi = (hi << 16) | ((dqs[r >> 1] >> (16 * (r & 1))) & 0xFFFFu);
                    if (base < (uint32_t)v.newcap) v.keys[row0 + base] = make_key(hi, (uint32_t)id);
                    ++base;
                }
            }
        };
        if (mine) {
            uint4 nx[NV];
            int32_t nid = 0;
            int nt_next = __builtin_ctzll(mine);
            load(t0 + nt_next, nx, nid);
            while (true) {
                const int tt = nt_next;
                uint4 cur[NV];
    #pragma unroll
                for (int i = 0; i < NV; ++i) cur[i] = nx[i];
                const int32_t cid = nid;
                mine &= mine - 1ull;
                if (mine) {
                    nt_next = __builtin_ctzll(mine);
                    load(t0 + nt_next, nx, nid);
                }
                const uint32_t nv = (uint32_t)__shfl((int)myinfo, tt) >> 22;
                const uint32_t* cw = reinterpret_cast<const uint32_t*>(cur);
                // The B offset is made opaque so that the operands are read
                // from LDS per tile rather than hoisted into registers.
                uint32_t boff = (uint32_t)lane;
                asm volatile("" : "+v"(boff));
                v16f acc = {};
    #pragma unroll
                for (int st = 0; st < KS; ++st) {
                    const uint32_t w = cw[st];
                    v4i a;
                    a[0] = (int)lut[w & 0xFFu];
                    a[1] = (int)lut[(w >> 8) & 0xFFu];
                    a[2] = (int)lut[(w >> 16) & 0xFFu];
                    a[3] = (int)lut[w >> 24];
                    acc = fp4::mma(a, sb[boff + st * 64], acc);
                }
                uint32_t pass = 0;
    #pragma unroll
                for (int r = 0; r < 16; ++r)
                    pass |= (acc[r] > tf && (uint32_t)fp4::acc_row(r, lane) < nv ? 1u : 0u) << r;
                if (kHist) {
    #pragma unroll
                    for (int r = 0; r < 16; ++r)
                        if ((pass >> r) & 1u) {
                            const int bin = hist_bin((BITS - (int)acc[r]) >> 1);
                            atomicAdd(&lhist[col * HW + (bin >> 1)], 1u << (16 * (bin & 1)));
                        }
                } else {
                    const uint32_t mine_n = (uint32_t)__popc(pass);
                    const uint32_t other = (uint32_t)__shfl_xor((int)mine_n, 32);
                    uint32_t ret = 0;
                    if (lane < 32 && mine_n + other > 0) ret = atomicAdd(&v.cnt[q], mine_n + other);
                    uint32_t dist[8];
    #pragma unroll
                    for (int r = 0; r < 16; r += 2)
                        dist[r >> 1] = (uint32_t)((BITS - (int)acc[r]) >> 1) |
                                       ((uint32_t)((BITS - (int)acc[r + 1]) >> 1) << 16);
                    uint32_t dqs[kFuse ? 8 : 1];
                    if (kFuse) {
                        // This lane's pieces of tile row col: halves 16 st + 8
                        // (lane / 32) .. + 7, zero past the row.
                        const _Float16* __restrict__ rp =
                            chalf + ((size_t)(t0 + (uint32_t)tt) * 32 + col) * hs + 8 * (lane >> 5);
                        // Two groups of four steps, each group's loads issued together.
                        v16f ac = {};
    #pragma unroll
                        for (int g = 0; g < kMaxFuseSteps; g += 4) {
                            h8 ah[4];
    #pragma unroll
                            for (int i = 0; i < 4; ++i) {
                                const int st = g + i;
                                h8 z = {};
                                if (st < ksteps && 16 * st + 8 * (lane >> 5) < hs)
                                    z = *reinterpret_cast<const h8*>(rp + 16 * st);
                                ah[i] = z;
                            }
    #pragma unroll
                            for (int i = 0; i < 4; ++i) {
                                const int st = g + i;
                                if (st < ksteps) {
                                    ac = __builtin_amdgcn_mfma_f32_32x32x16_f16(ah[i], qb[st * 64 + lane], ac, 0, 0,
                                                                                0);
                                    ac = __builtin_amdgcn_mfma_f32_32x32x16_f16(ah[i], qb[(ksteps + st) * 64 + lane],
                                                                                ac, 0, 0, 0);
                                }
                            }
                        }
    #pragma unroll
                        for (int r = 0; r < 16; r += 2)
                            dqs[r >> 1] = fused_q16(ac[r]) | (fused_q16(ac[r + 1]) << 16);
                    }
                    // emit shuffles across the wave: every lane calls it, or none.
                    if (__ballot(pass != 0)) emit(dist, dqs, pass, ret, other, cid, (t0 + (uint32_t)tt) * 32u);
                }
                if (!mine) break;
            }
        }
        __syncthreads();  // every wave is done with sb and lhist
        if (kHist) {
            for (int e = threadIdx.x; e < LH; e += 256) {
                const uint32_t x = lhist[e];
                const int cq = e / HW;
                if (x && (uint32_t)cq < m) {
                    uint32_t* row = hist + (size_t)list[q0 + cq] * kHistBins + 2 * (e % HW);
                    if (x & 0xFFFFu) atomicAdd(&row[0], x & 0xFFFFu);
                    if (x >> 16) atomicAdd(&row[1], x >> 16);
                }
            }
            __syncthreads();  // the flush has read lhist before the next chunk clears it
        }
    }
}

// thr[q] = min(thr[q], (T + 1, 0)) with T the upper edge of the bin at which
// the query's first-pass counts reach K (the last bin's edge: bits); then the
// query's candidate cells (qcells2) whose bound is below the final threshold
// are kept (compacted in place, n2 updated) and counted into cnt2 for the
// second pass.  One wave per query.
// One wave per query: the threshold from the first pass's histogram, then
// the query's second-pass cell list compacted to the cells whose bound is
// below it, counted into cnt2.
__global__ __launch_bounds__(256) void hist_threshold_kernel(const uint32_t* __restrict__ hist, int Q, int bits,
                                                             int K, uint64_t* __restrict__ thr,
                                                             const uint16_t* __restrict__ D, int P, int Pp,
                                                             const uint16_t* __restrict__ cmin,
                                                             const uint16_t* __restrict__ cmax,
                                                             uint32_t* __restrict__ cnt2,
                                                             int32_t* __restrict__ qcells2,
                                                             uint32_t* __restrict__ n2) {
    const int q = blockIdx.x * 4 + (threadIdx.x >> 6);
    if (q >= Q) return;
    const int lane = threadIdx.x & 63;
    const int b = kth_bin(hist + (size_t)q * kHistBins, kHistBins, (uint32_t)K);
    uint64_t t = thr[q];
    if (b < kHistBins) {
        const uint32_t T = b == kHistBins - 1 ? (uint32_t)bits : (uint32_t)(8 * b + 7);
        const uint64_t th = make_key(T + 1u, 0u);
        t = th < t ? th : t;
    }
    if (lane == 0) thr[q] = t;
    const uint32_t h = (uint32_t)(t >> 32);
    const uint16_t* __restrict__ row = D + (size_t)q * Pp;
    int32_t* __restrict__ cells = qcells2 + (size_t)q * P;
    const uint32_t n = n2[q];
    // U blocks of 64 cells at a time: every list load, then every gather, is
    // issued before the first use.  The list is compacted in place, in
    // order: a write never lands past an entry already read.
    constexpr int U = 4;
    uint32_t kept = 0;
    for (uint32_t i0 = 0; i0 < n; i0 += 64 * U) {
        int32_t j[U];
#pragma unroll
        for (int u = 0; u < U; ++u) {
            const uint32_t i = i0 + 64 * u + lane;
            j[u] = i < n ? cells[i] : -1;
        }
        uint32_t dv[U], lo[U], hi[U];
#pragma unroll
        for (int u = 0; u < U; ++u) {
            const int jj = j[u] >= 0 ? j[u] : 0;
            dv[u] = row[jj];
            lo[u] = cmin[jj];
            hi[u] = cmax[jj];
        }
#pragma unroll
        for (int u = 0; u < U; ++u) {
            const bool keep = j[u] >= 0 && lower_bound_to(dv[u], lo[u], hi[u]) < h;
            const uint64_t m = __ballot(keep);
            if (keep) {
                atomicAdd(&cnt2[j[u]], 1u);
                cells[kept + __popcll(m & ((1ull << lane) - 1ull))] = j[u];
            }
            kept += (uint32_t)__popcll(m);
        }
    }
    if (lane == 0) n2[q] = kept;
}

// Per-cell query lists from the per-query cell lists: query q goes to slot
// cursor[j]++ of cell j's list for each of its cells.  One wave per query;
// a set skip flag leaves the lists empty.
__global__ __launch_bounds__(256) void place_from_cells_kernel(const int32_t* __restrict__ qcells,
                                                               const uint32_t* __restrict__ nq, int Q, int P,
                                                               uint32_t* __restrict__ cursor,
                                                               const uint32_t* __restrict__ coff,
                                                               int32_t* __restrict__ list,
                                                               const uint32_t* __restrict__ skip) {
    const int q = blockIdx.x * 4 + (threadIdx.x >> 6);
    if (q >= Q || (skip && *skip)) return;
    const int lane = threadIdx.x & 63;
    const uint32_t n = nq[q];
    for (uint32_t i = lane; i < n; i += 64) {
        const int32_t j = qcells[(size_t)q * P + i];
        list[coff[j] + atomicAdd(&cursor[j], 1u)] = q;
    }
}

template <int KS>
void launch_pivots_ks(const uint32_t* rows, int64_t n, const uint32_t* pfrag, int P, int Pp,
                      uint32_t* best, uint16_t* D, uint32_t* zero, int nzero, const TopKView* tv, hipStream_t s) {
    if (best) {
        const v4i* f = reinterpret_cast<const v4i*>(pfrag);
        hipLaunchKernelGGL((pivot_assign_kernel<KS>), dim3(blocks_for((size_t)((n + 31) / 32), 4)), dim3(256), 0, s,
                           rows, n, f, P, best);
        check_launch("pivot_assign_kernel");
    } else {
        const unsigned gy = (unsigned)(((P + 31) / 32 + kGroup - 1) / kGroup);
        hipLaunchKernelGGL((pivot_matrix_kernel<KS>), dim3(blocks_for((size_t)((n + 31) / 32), 4), gy), dim3(256), 0,
                           s, rows, (int)n, reinterpret_cast<const v4i*>(pfrag), P, Pp, D, zero, nzero,
                           tv ? *tv : TopKView{});
        check_launch("pivot_matrix_kernel");
    }
}

// best != nullptr: nearest pivot of each row; otherwise the distance matrix
// D, whose
// launch also zeroes zero[0, nzero) and clears the fresh view *tv (when
// given).
void pivots(int words, const uint32_t* rows, int64_t n, const uint32_t* pfrag, int P, int Pp,
            uint32_t* best, uint16_t* D, hipStream_t s, uint32_t* zero = nullptr, int nzero = 0,
            const TopKView* tv = nullptr) {
    switch (words) {
        case 8: launch_pivots_ks<4>(rows, n, pfrag, P, Pp, best, D, zero, nzero, tv, s); break;
        case 16: launch_pivots_ks<8>(rows, n, pfrag, P, Pp, best, D, zero, nzero, tv, s); break;
        case 24: launch_pivots_ks<12>(rows, n, pfrag, P, Pp, best, D, zero, nzero, tv, s); break;
        case 32: launch_pivots_ks<16>(rows, n, pfrag, P, Pp, best, D, zero, nzero, tv, s); break;
        default: fail(LSH_ERR_INTERNAL, "pivot cells need 8, 16, 24 or 32 sketch words");
    }
}

template <int KS>
void launch_cell_scan_ks(const CellIndex& c, unsigned grid, const uint32_t* coff, const uint32_t* choff,
                         const uint32_t* cmap, int Pp, const int32_t* list, const uint16_t* D, const uint32_t* qs,
                         const TopKView& v, uint32_t* hist, const uint32_t* skip, const CellFuse* fz, hipStream_t s) {
    auto go = [&](auto kernel, size_t lds, const _Float16* chalf, int hs, const _Float16* q16, int ksteps) {
        hipLaunchKernelGGL(kernel, dim3(grid), dim3(256), lds, s, c.sk.get(), c.ids.get(), c.tinfo.get(),
                           c.ctile.get(), coff, choff, cmap, c.P, Pp, list, D, qs, v, hist, skip, chalf, hs, q16,
                           ksteps);
    };
    if (hist) {
        go(cell_scan_kernel<KS, true, false>, 0, nullptr, 0, nullptr, 0);
    } else if (fz) {
        LSH_REQUIRE(fz->ksteps >= 1 && fz->ksteps <= kMaxFuseSteps && 16 * fz->ksteps >= fz->hs && v.row_keys &&
                        v.fused_keys,
                    LSH_ERR_INTERNAL, "cell scan: fused distances need rows of at most 128 halves and row keys");
        go(cell_scan_kernel<KS, false, true>, sizeof(h8) * 2 * 64 * (size_t)fz->ksteps, fz->chalf, fz->hs, fz->q16,
           fz->ksteps);
    } else {
        go(cell_scan_kernel<KS, false, false>, 0, nullptr, 0, nullptr, 0);
    }
    check_launch("cell_scan_kernel");
}

// The scan over a grid of `grid` workgroups that loops over the chunks.
void cell_scan(const CellIndex& c, unsigned grid, const uint32_t* coff, const uint32_t* choff, uint32_t* cmap,
               int Pp, const int32_t* list, const uint16_t* D, const uint32_t* qs, const TopKView& v, uint32_t* hist,
               const uint32_t* skip, const CellFuse* fz, hipStream_t s) {
    switch (c.words) {
        case 8: launch_cell_scan_ks<4>(c, grid, coff, choff, cmap, Pp, list, D, qs, v, hist, skip, fz, s); break;
        case 16: launch_cell_scan_ks<8>(c, grid, coff, choff, cmap, Pp, list, D, qs, v, hist, skip, fz, s); break;
        case 24: launch_cell_scan_ks<12>(c, grid, coff, choff, cmap, Pp, list, D, qs, v, hist, skip, fz, s); break;
        case 32: launch_cell_scan_ks<16>(c, grid, coff, choff, cmap, Pp, list, D, qs, v, hist, skip, fz, s); break;
        default: fail(LSH_ERR_INTERNAL, "pivot cells need 8, 16, 24 or 32 sketch words");
    }
}



}  // namespace

// About kPointsPerPivot points per cell before the outlier rounds.
constexpr int64_t kPointsPerPivot = 768;
int cell_pivots_for(int64_t N) {
    if (N < kCellMinPoints) return 0;
    return (int)std::min<int64_t>(kMaxPivots / 2, std::max<int64_t>(64, N / kPointsPerPivot));
}

void CellIndex::build(const uint32_t* ds, int64_t N, int wpad, hipStream_t s) {
    P = 0;
    int np = cell_pivots_for(N);
    if (np == 0 || !hamming_mfma_supported(wpad)) return;
    LSH_REQUIRE(N < ((int64_t)1 << 31), LSH_ERR_UNSUPPORTED, "pivot cells need N < 2^31");
    words = wpad;
    // Pivots: every (N/np)-th row (np at most half of kMaxPivots).  A group
    // of points that drew no pivot of its own ends up far from the pivot it
    // joins and widens that cell's range.  So a point is "far" beyond the
    // smaller of twice the median and three times the 10th percentile of the
    // distances to the pivots (and beyond bits/16); far points seed more
    // pivots, and the points are assigned again, until at most one point in
    // a thousand is far (or eight rounds, or kMaxPivots).  A round's seeds
    // come from a stride over the far points (four candidates per seed it
    // may add, up to half the pivots a round), thinned greedily: a candidate
    // within the far distance of a seed already taken this round belongs to
    // the same group and is skipped, so each seed covers a group of its own.
    // The median keeps well-seeded data from drawing pivots for the tails of
    // its groups; the low percentile keeps the rule working when most points
    // start without a near pivot (the median is then far too).
    const int bits = 32 * wpad;
    std::vector<int32_t> prow((size_t)np);
    const int64_t step = N / np;
    for (int i = 0; i < np; ++i) prow[i] = (int32_t)(i * step);
    DevBuf<uint32_t> ckey((size_t)N);
    for (int round = 0;; ++round) {
        const int n = (int)prow.size();
        DevBuf<int32_t> didx((size_t)n);
        LSH_HIP(hipMemcpyAsync(didx.get(), prow.data(), sizeof(int32_t) * n, hipMemcpyHostToDevice, s));
        DevBuf<uint32_t> psk((size_t)n * wpad);
        launch_gather_rows_u32(ds, wpad, didx.get(), n, psk.get(), s);
        pfp4.alloc(hamming_fp4_words(n, wpad));
        launch_expand_fp4(psk.get(), n, wpad, pfp4.get(), s);
        // Nearest pivot of every point.
        pivots(wpad, ds, N, pfp4.get(), n, 0, ckey.get(), nullptr, s);
        if (round == 8 || n >= kMaxPivots) break;
        std::vector<uint32_t> hk((size_t)N);
        LSH_HIP(hipMemcpyAsync(hk.data(), ckey.get(), sizeof(uint32_t) * N, hipMemcpyDeviceToHost, s));
        LSH_HIP(hipStreamSynchronize(s));
        std::vector<int64_t> dh((size_t)bits + 1, 0);
        for (uint32_t k : hk) ++dh[k & kDistMask];
        int64_t acc = 0;
        int p10 = -1, med = -1;
        for (int b = 0; b <= bits && med < 0; ++b) {
            acc += dh[b];
            if (p10 < 0 && acc * 10 >= N) p10 = b;
            if (acc * 2 >= N) med = b;
        }
        const int far_at = std::max(std::min(2 * med, 3 * p10), bits / 16);
        std::vector<int32_t> add;
        for (int64_t i = 0; i < N; ++i)
            if ((int)(hk[i] & kDistMask) > far_at) add.push_back((int32_t)i);
        if ((int64_t)add.size() * 1000 <= N) break;
        const size_t take = std::min<size_t>({add.size(), (size_t)(kMaxPivots - n), (size_t)std::max(1, n / 2)});
        const size_t ncand = std::min(add.size(), 4 * take);
        std::vector<int32_t> cand(ncand);
        for (size_t i = 0; i < ncand; ++i) cand[i] = add[i * add.size() / ncand];
        DevBuf<int32_t> cidx(ncand);
        DevBuf<uint32_t> csk(ncand * wpad);
        LSH_HIP(hipMemcpyAsync(cidx.get(), cand.data(), sizeof(int32_t) * ncand, hipMemcpyHostToDevice, s));
        launch_gather_rows_u32(ds, wpad, cidx.get(), (int)ncand, csk.get(), s);
        std::vector<uint64_t> hsk(ncand * wpad / 2);
        LSH_HIP(hipMemcpyAsync(hsk.data(), csk.get(), sizeof(uint32_t) * ncand * wpad, hipMemcpyDeviceToHost, s));
        LSH_HIP(hipStreamSynchronize(s));
        const int w2 = wpad / 2;  // wpad is a multiple of 8
        std::vector<size_t> kept;
        for (size_t i = 0; i < ncand && kept.size() < take; ++i) {
            const uint64_t* a = &hsk[i * w2];
            bool fresh = true;
            for (size_t j : kept) {
                const uint64_t* b = &hsk[j * w2];
                int h = 0;
                for (int w = 0; w < w2 && h <= far_at; ++w) h += __builtin_popcountll(a[w] ^ b[w]);
                if (h <= far_at) {
                    fresh = false;
                    break;
                }
            }
            if (fresh) kept.push_back(i);
        }
        for (size_t i : kept) prow.push_back(cand[i]);
    }
    np = (int)prow.size();
    {
        // Pivots renumbered along a greedy nearest-neighbour chain (each next
        // pivot the nearest not yet taken), then the points assigned again:
        // cells with consecutive numbers lie near each other, so queries that
        // run side by side in nearest-cell order (and the cells' rows, laid
        // out in cell order) share candidate cells in L2.  Any assignment to
        // the nearest pivot keeps the search exact.
        DevBuf<int32_t> didx((size_t)np);
        DevBuf<uint32_t> psk((size_t)np * wpad);
        LSH_HIP(hipMemcpyAsync(didx.get(), prow.data(), sizeof(int32_t) * np, hipMemcpyHostToDevice, s));
        launch_gather_rows_u32(ds, wpad, didx.get(), np, psk.get(), s);
        std::vector<uint64_t> hp((size_t)np * wpad / 2);
        LSH_HIP(hipMemcpyAsync(hp.data(), psk.get(), sizeof(uint32_t) * (size_t)np * wpad, hipMemcpyDeviceToHost, s));
        LSH_HIP(hipStreamSynchronize(s));
        const int w2 = wpad / 2;
        std::vector<int32_t> chain;
        chain.reserve((size_t)np);
        std::vector<char> taken((size_t)np, 0);
        int cur = 0;
        for (int step = 0; step < np; ++step) {
            chain.push_back(prow[cur]);
            taken[cur] = 1;
            int best = -1, bd = 1 << 30;
            const uint64_t* a = &hp[(size_t)cur * w2];
            for (int j = 0; j < np; ++j) {
                if (taken[j]) continue;
                const uint64_t* b = &hp[(size_t)j * w2];
                int h = 0;
                for (int w = 0; w < w2 && h < bd; ++w) h += __builtin_popcountll(a[w] ^ b[w]);
                if (h < bd) {
                    bd = h;
                    best = j;
                }
            }
            if (best < 0) break;
            cur = best;
        }
        prow.swap(chain);
        LSH_HIP(hipMemcpyAsync(didx.get(), prow.data(), sizeof(int32_t) * np, hipMemcpyHostToDevice, s));
        launch_gather_rows_u32(ds, wpad, didx.get(), np, psk.get(), s);
        pfp4.alloc(hamming_fp4_words(np, wpad));
        launch_expand_fp4(psk.get(), np, wpad, pfp4.get(), s);
        pivots(wpad, ds, N, pfp4.get(), np, 0, ckey.get(), nullptr, s);
        LSH_HIP(hipStreamSynchronize(s));
    }

    // Counting sort by (pivot, distance).
    const uint32_t T = (uint32_t)np << kKeyShift;
    DevBuf<uint32_t> off((size_t)T + 1), cursor((size_t)T + 1), perm((size_t)N);
    launch_bucket_build(ckey.get(), N, 1, T, off.get(), cursor.get(), perm.get(), s);
    DevBuf<uint32_t> cstart((size_t)np + 1);
    hipLaunchKernelGGL(cell_starts_kernel, dim3(blocks_for((size_t)np + 1, 256)), dim3(256), 0, s, off.get(), np,
                       cstart.get());
    check_launch("cell_starts_kernel");
    std::vector<uint32_t> hs((size_t)np + 1), ht((size_t)np + 1);
    LSH_HIP(hipMemcpyAsync(hs.data(), cstart.get(), sizeof(uint32_t) * hs.size(), hipMemcpyDeviceToHost, s));
    LSH_HIP(hipStreamSynchronize(s));
    ht[0] = 0;
    for (int j = 0; j < np; ++j) ht[j + 1] = ht[j] + (hs[j + 1] - hs[j] + 31) / 32;
    ntiles = ht[np];
    nparts = 0;
    for (int j = 0; j < np; ++j) nparts += (ht[j + 1] - ht[j] + kChunkTiles - 1) / kChunkTiles;
    ctile.alloc((size_t)np + 1);
    LSH_HIP(hipMemcpyAsync(ctile.get(), ht.data(), sizeof(uint32_t) * ht.size(), hipMemcpyHostToDevice, s));
    std::vector<uint32_t> hz((size_t)np);
    for (int j = 0; j < np; ++j) hz[j] = hs[j + 1] - hs[j];
    csize.alloc((size_t)np);
    LSH_HIP(hipMemcpyAsync(csize.get(), hz.data(), sizeof(uint32_t) * hz.size(), hipMemcpyHostToDevice, s));

    const int64_t npos = ntiles * 32;
    ids.alloc((size_t)npos);
    launch_fill_i32(ids.get(), (size_t)npos, -1, s);
    DevBuf<uint16_t> pdist((size_t)npos);
    hipLaunchKernelGGL(cell_place_kernel, dim3(blocks_for((size_t)N, 256)), dim3(256), 0, s,
                       reinterpret_cast<const int32_t*>(perm.get()), ckey.get(), N, cstart.get(), ctile.get(),
                       ids.get(), pdist.get());
    check_launch("cell_place_kernel");
    tinfo.alloc((size_t)ntiles);
    hipLaunchKernelGGL(cell_tile_info_kernel, dim3(blocks_for((size_t)ntiles, 256)), dim3(256), 0, s, ids.get(),
                       pdist.get(), ntiles, tinfo.get());
    check_launch("cell_tile_info_kernel");
    cmin.alloc((size_t)np);
    cmax.alloc((size_t)np);
    crec.alloc((size_t)np);
    hipLaunchKernelGGL(cell_range_kernel, dim3(blocks_for((size_t)np, 256)), dim3(256), 0, s, cstart.get(),
                       ctile.get(), pdist.get(), np, cmin.get(), cmax.get(), crec.get());
    check_launch("cell_range_kernel");
    sk.alloc((size_t)npos * wpad);
    hipLaunchKernelGGL(pack_cells_kernel, dim3(blocks_for((size_t)npos * wpad, 256)), dim3(256), 0, s, ds, wpad,
                       ids.get(), npos, sk.get());
    check_launch("pack_cells_kernel");
    LSH_HIP(hipStreamSynchronize(s));
    P = np;
}

void cell_rows(const CellIndex& c, const float* data, int64_t N, int d, DevBuf<float>& rows, DevBuf<int32_t>& pos,
               hipStream_t s) {
    const int64_t npos = c.ntiles * 32;
    rows.alloc((size_t)npos * d);
    pos.alloc((size_t)N);
    hipLaunchKernelGGL(cell_rows_kernel, dim3((unsigned)std::min<size_t>(65535u * 16u, blocks_for((size_t)npos * d, 256))),
                       dim3(256), 0, s, data, d, c.ids.get(), npos, rows.get(), pos.get());
    check_launch("cell_rows_kernel");
}

size_t CellIndex::bytes() const {
    return pfp4.bytes() + sk.bytes() + ids.bytes() + tinfo.bytes() + ctile.bytes() + csize.bytes() +
           cmin.bytes() + cmax.bytes() + crec.bytes();
}

namespace {

// Elements of the four scratch slots a cell search of Q queries takes:
// D (u16), meta (u32), the lists and chunk maps (i32), the histograms and
// per-query cell lists (u32).  The layouts are described in cell_search.
struct CellScratchSizes {
    size_t d, meta, list, hist;
};
CellScratchSizes cell_scratch_sizes(const CellIndex& c, int Q) {
    const size_t P = (size_t)c.P, Pp = (P + 31) / 32 * 32, q = (size_t)Q;
    const size_t qgroups = (q + 31) / 32;
    return {q * Pp, 4 * (P + 1) + 1 + q + 4, q * P + qgroups * (size_t)c.nparts,
            q * kHistBins + 2 * q * P + 2 * q + (q + 1) / 2};
}

}  // namespace

void cell_search_reserve(const CellIndex& c, int Q, Scratch& sc, int depth) {
    if (!c.built() || Q <= 0) return;
    const CellScratchSizes z = cell_scratch_sizes(c, Q);
    (void)sc.get<uint16_t>(depth, 0, z.d);
    (void)sc.get<uint32_t>(depth, 1, z.meta);
    (void)sc.get<int32_t>(depth, 2, z.list);
    (void)sc.get<uint32_t>(depth, 3, z.hist);
}

const uint32_t* cell_search(const CellIndex& c, const uint32_t* qs, int Q, int K, const TopKView& v, Scratch& sc,
                            int depth, hipStream_t s, int32_t* qorder, const CellFuse* fuse, PhaseTimer* pt,
                            int phase) {
    if (!c.built() || Q <= 0) return nullptr;
    const int P = c.P;
    const int Pp = (P + 31) / 32 * 32;
    const int bits = 32 * c.words;
    const CellScratchSizes z = cell_scratch_sizes(c, Q);
    uint16_t* D = sc.get<uint16_t>(depth, 0, z.d);
    // meta: cnt1, cnt2 [P+1 each], stat [4] = {second pass too dense,
    // first-pass tile products, second-pass tile products}, coff [P+1],
    // choff [P+2], nearest cell [Q].  The first 2 (P+1) + 4 words start at
    // zero (cleared by the pivot-matrix launch).
    uint32_t* meta = sc.get<uint32_t>(depth, 1, z.meta);
    uint32_t* cnt1 = meta;
    uint32_t* cnt2 = meta + (P + 1);
    uint32_t* stat = meta + 2 * (P + 1);
    uint32_t* coff = stat + 4;
    uint32_t* choff = coff + (P + 1);
    uint32_t* nearest = choff + (P + 2);
    const int nzero = 2 * (P + 1) + 4;
    // hist [Q][kHistBins], the per-query cell lists [Q][P] x 2 and their
    // lengths [Q] x 2, the radii [Q] (u16).
    const size_t hn = (size_t)Q * kHistBins, qp = (size_t)Q * P;
    uint32_t* hist = sc.get<uint32_t>(depth, 3, z.hist);
    int32_t* qcells1 = reinterpret_cast<int32_t*>(hist + hn);
    int32_t* qcells2 = qcells1 + qp;
    uint32_t* n1 = reinterpret_cast<uint32_t*>(qcells2 + qp);
    uint32_t* n2 = n1 + Q;
    uint16_t* R = reinterpret_cast<uint16_t*>(n2 + Q);
    // Lists of at most every (query, cell) pair, and chunk maps of at most
    // every (query group, tile range) pair: no count has to come back to the
    // host before the scans.
    const size_t qgroups = ((size_t)Q + 31) / 32;
    const size_t cap_list = (size_t)Q * P, cap_chunks = qgroups * (size_t)c.nparts;
    int32_t* list = sc.get<int32_t>(depth, 2, z.list);
    uint32_t* cmap = reinterpret_cast<uint32_t*>(list + cap_list);
    const unsigned grid = (unsigned)std::max<size_t>(1, std::min<size_t>(4096, cap_chunks));
    // The second pass is "too dense" when it would multiply more than half
    // of the tiles of a full pass (query tiles x point tiles).
    const uint32_t dense_limit = (uint32_t)std::min<double>(4e9, 0.5 * (double)qgroups * (double)c.ntiles);

    LSH_REQUIRE(v.fresh && v.Q == Q, LSH_ERR_INTERNAL, "cell_search takes a fresh top-K view of its queries");
    pivots(c.words, qs, Q, c.pfp4.get(), P, Pp, nullptr, D, s, meta, nzero, &v);
    // Threshold bound from the tile ranges, the first-pass radii and their
    // cell counts, and the nearest cell of each query.
    auto threshold = [&](auto kernel, int nt) {
        hipLaunchKernelGGL(kernel, dim3((unsigned)Q), dim3(nt), 0, s, D, Q, P, Pp, c.crec.get(), bits, K, v.thr, R,
                           nearest, cnt1, qcells1, n1, qcells2, n2, hist);
    };
    LSH_REQUIRE(P >= 1 && P <= 256 * 32, LSH_ERR_INTERNAL, "cell_search: pivot count outside the threshold kernel's range");
    // Up to 2,048 pivots two waves own a query (one wave holding 32 cells a
    // lane needs 151 registers, 3 waves a SIMD, and was slower); beyond, a
    // 256-thread workgroup.
    if (P <= 128 * 8) threshold(cell_threshold_kernel<8, 128>, 128);
    else if (P <= 128 * 16) threshold(cell_threshold_kernel<16, 128>, 128);
    else if (P <= 256 * 10) threshold(cell_threshold_kernel<10, 256>, 256);
    else if (P <= 256 * 10) threshold(cell_threshold_kernel<10, 256>, 256);
    else if (P <= 256 * 12) threshold(cell_threshold_kernel<12, 256>, 256);
    else if (P <= 256 * 16) threshold(cell_threshold_kernel<16, 256>, 256);
    else threshold(cell_threshold_kernel<32, 256>, 256);
    check_launch("cell_threshold_kernel");
    // First pass: exact distances over the nearest cells, K-th by histogram,
    // then the second pass's cell counts under the final threshold.  Its
    // offsets are computed in the launch that also orders the queries by
    // nearest cell (a counting sort): neighbouring queries then share most
    // candidates, which the rerank can reuse.
    const unsigned qwaves = (unsigned)((Q + 3) / 4);
    hipLaunchKernelGGL(offsets_and_order_kernel, dim3(qorder ? 2 : 1), dim3(1024),
                       qorder ? sizeof(uint32_t) * (size_t)P : 0, s, cnt1, P, c.ctile.get(), coff, choff, cmap,
                       stat + 1, nearest, Q, qorder);
    check_launch("offsets_and_order_kernel");
    hipLaunchKernelGGL(place_from_cells_kernel, dim3(qwaves), dim3(256), 0, s, qcells1, n1, Q, P, cnt1, coff, list,
                       nullptr);
    check_launch("place_from_cells_kernel");
    cell_scan(c, grid, coff, choff, cmap, Pp, list, D, qs, v, hist, nullptr, nullptr, s);
    hipLaunchKernelGGL(hist_threshold_kernel, dim3(qwaves), dim3(256), 0, s, hist, Q, bits, K, v.thr, D, P, Pp,
                       c.cmin.get(), c.cmax.get(), cnt2, qcells2, n2);
    check_launch("hist_threshold_kernel");
    // Second pass: the tiles the bounds leave (skipped when too dense).
    hipLaunchKernelGGL(cell_offsets_kernel, dim3(1), dim3(1024), 0, s, cnt2, P, c.ctile.get(), coff, choff, cmap,
                       stat + 2, stat, dense_limit);
    check_launch("cell_offsets_kernel");
    hipLaunchKernelGGL(place_from_cells_kernel, dim3(qwaves), dim3(256), 0, s, qcells2, n2, Q, P, cnt2, coff, list,
                       stat);
    check_launch("place_from_cells_kernel");
    if (pt) pt->begin(phase, s);
    cell_scan(c, grid, coff, choff, cmap, Pp, list, D, qs, v, nullptr, stat, fuse, s);
    if (pt) pt->end(phase, s);
    hipStreamCaptureStatus capturing = hipStreamCaptureStatusNone;
    if (debug_cells()) LSH_HIP(hipStreamIsCapturing(s, &capturing));
    if (debug_cells() && capturing == hipStreamCaptureStatusNone) {
        std::vector<uint64_t> ht((size_t)Q);
        uint32_t hs[3], pairs2 = 0, chunks2 = 0;
        LSH_HIP(hipMemcpyAsync(ht.data(), v.thr, sizeof(uint64_t) * Q, hipMemcpyDeviceToHost, s));
        LSH_HIP(hipMemcpyAsync(hs, stat, sizeof(hs), hipMemcpyDeviceToHost, s));
        LSH_HIP(hipMemcpyAsync(&pairs2, coff + P, sizeof(uint32_t), hipMemcpyDeviceToHost, s));
        LSH_HIP(hipMemcpyAsync(&chunks2, choff + P, sizeof(uint32_t), hipMemcpyDeviceToHost, s));
        LSH_HIP(hipStreamSynchronize(s));
        double mt = 0;
        for (int q = 0; q < Q; ++q) mt += ht[q] == kEmptyKey ? 2.0 * bits : (double)(ht[q] >> 32);
        std::fprintf(stderr,
                     "cells: P=%d tiles=%lld Q=%d K=%d mean T=%.1f tile products %u + %u%s; second pass: %u "
                     "(query, cell) pairs in %u chunks\n",
                     P, (long long)c.ntiles, Q, K, mt / Q, hs[1], hs[2], hs[0] ? " (too dense)" : "", pairs2, chunks2);
    }
    return stat;
}

}  // namespace lshgpu
