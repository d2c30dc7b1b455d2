// scan_sketch.hip -- integer sketch-distance scans (the hot kernel of the
// sketch paths).
//
// Replaces scanHammingDistance / scanJaccardDistance
// (sketchedDistanceScanners.cuh:14-100, 194-273), which give each query one
// warp that strides the packed sketches with a 128-byte lane stride and feeds
// a per-lane buffer into a warp sort.  Here each lane owns one data point and
// streams its sketch through VGPRs 8 words at a time; a wave walks a group of
// 64 queries whose sketch words are wave-uniform (scalar loads), so the
// distance of a (query, point) pair is W XORs plus W bit-count-and-adds
// (v_bcnt_u32_b32 accumulates).  Query groups vary fastest across the grid so
// the blocks in flight share data rows through L2.  A pair is kept only when
// its (distance, ID) key beats the query's current threshold; the survivors
// are appended with one atomic per wave.
//
//   kind 0: Hamming distance over `words` packed words (implementations 3
//           and 5; jaccardOneBitDistanceFunc orders the same way)
//   kind 1: number of unequal hash values (implementation 4:
//           jaccardDistanceFunc = 1 - equal/H orders the same way)
#include "device_common.hpp"

namespace lshgpu {

namespace {

constexpr int QGS = 64;  // queries per group
constexpr int CH = 8;    // sketch words per register chunk (rows are padded to a multiple of 8)
constexpr int TPB = 256;

template <int KIND, bool RAGGED>
__global__ __launch_bounds__(TPB) void sketch_scan_kernel(SketchScanArgs a, TopKView v) {
    const int g = blockIdx.x;
    const int qbase = g * QGS;
    const int nq = min(QGS, a.Q - qbase);
    const int64_t i = a.p0 + (int64_t)blockIdx.y * TPB + threadIdx.x;
    const bool active = i < a.p1;
    const int W = a.words;  // multiple of CH
    const int DW = RAGGED ? a.dwords : W;
    const int64_t ri = active ? i : a.p0;
    const uint32_t* __restrict__ row = a.data + (size_t)ri * DW;
    const uint32_t* __restrict__ qg = a.qs + (size_t)qbase * W;

    uint32_t h[QGS];
#pragma unroll
    for (int q = 0; q < QGS; ++q) h[q] = 0;

#pragma clang loop unroll(disable) vectorize(disable) interleave(disable)
    for (int w0 = 0; w0 < W; w0 += CH) {
        uint32_t xr[CH];
        if constexpr (RAGGED) {
            // words past the row read as zero, as the zero-padded query words
#pragma unroll
            for (int c = 0; c < CH; ++c) xr[c] = w0 + c < DW ? row[w0 + c] : 0u;
        } else {
#pragma unroll
            for (int c = 0; c < CH / 4; ++c) {
                const uint4 t = *reinterpret_cast<const uint4*>(row + w0 + 4 * c);
                xr[4 * c + 0] = t.x;
                xr[4 * c + 1] = t.y;
                xr[4 * c + 2] = t.z;
                xr[4 * c + 3] = t.w;
            }
        }
#pragma unroll
        for (int q = 0; q < QGS; ++q) {
            const uint32_t* __restrict__ qv = qg + (size_t)q * W + w0;
#pragma unroll
            for (int ww = 0; ww < CH; ++ww) {
                if constexpr (KIND == 0)
                    h[q] += __builtin_popcount(xr[ww] ^ qv[ww]);
                else
                    h[q] += (xr[ww] != qv[ww]) ? 1u : 0u;
            }
        }
    }

#pragma unroll
    for (int q = 0; q < QGS; ++q) {
        if (q < nq) {
            const int qrow = qbase + q;
            const uint64_t thr = v.thr[qrow];
            const uint32_t thr_hi = (uint32_t)(thr >> 32);
            const bool coarse = active && h[q] <= thr_hi;
            if (__ballot(coarse)) {
                const uint64_t key = make_key(h[q], (uint32_t)i);
                append_key(v, qrow, coarse && key < thr, key);
            }
        }
    }
}

template <int KIND>
void launch_kind(const SketchScanArgs& a, const TopKView& v, hipStream_t s) {
    const int groups = (a.Q + QGS - 1) / QGS;
    const int64_t npts = a.p1 - a.p0;
    LSH_REQUIRE(a.words % CH == 0 && (reinterpret_cast<uintptr_t>(a.qs) % 16) == 0 && a.dwords >= 0 &&
                    a.dwords <= a.words,
                LSH_ERR_INTERNAL, "query sketch rows must be padded to 8 words and 16-byte aligned");
    const bool ragged = (a.dwords != 0 && a.dwords != a.words) || (reinterpret_cast<uintptr_t>(a.data) % 16) != 0;
    SketchScanArgs base = a;
    if (base.dwords == 0) base.dwords = a.words;
    const int64_t max_pts = (int64_t)65535 * TPB;
    for (int64_t off = 0; off < npts; off += max_pts) {
        SketchScanArgs b = base;
        b.p0 = a.p0 + off;
        b.p1 = std::min(a.p1, b.p0 + max_pts);
        const unsigned by = (unsigned)((b.p1 - b.p0 + TPB - 1) / TPB);
        if (ragged)
            hipLaunchKernelGGL((sketch_scan_kernel<KIND, true>), dim3(groups, by), dim3(TPB), 0, s, b, v);
        else
            hipLaunchKernelGGL((sketch_scan_kernel<KIND, false>), dim3(groups, by), dim3(TPB), 0, s, b, v);
    }
    check_launch("sketch_scan_kernel");
}

}  // namespace

void launch_sketch_scan(const SketchScanArgs& a, const TopKView& v, hipStream_t s) {
    if (a.p1 <= a.p0 || a.Q <= 0) return;
    if (a.kind == 0)
        launch_kind<0>(a, v, s);
    else
        launch_kind<1>(a, v, s);
}

}  // namespace lshgpu
