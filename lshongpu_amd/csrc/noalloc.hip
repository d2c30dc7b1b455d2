// noalloc.hip -- Hamming top-C over device sketches with no device
// allocation: the working memory (padded query sketches and the per-query
// top-K state) lives in one file-scope device array, `scratch`, of
// kNoAllocScratchBytes per GPU, which every call on that GPU shares.
//
// The search is the exact one of lsh_hamming_topc: each round scans at most
// the append capacity of points (so no query's appends can overflow and no
// round is redone), folds the appends into each query's best list, and the
// last step orders the lists and writes (ID, distance) pairs.  Nothing waits
// on the host, so the launch sequence may be captured in a graph.
#include <mutex>

#include "device_common.hpp"

namespace lshgpu {

namespace {

constexpr size_t kNoAllocScratchBytes = size_t(32) << 20;  // 32 MiB per GPU
constexpr int kMaxDevices = 64;

static __device__ __align__(256) uint8_t scratch[kNoAllocScratchBytes];

// Device address of `scratch` on each device (the array's name is a host
// stand-in; only these addresses may reach kernels and copies).
uint8_t* g_scratch[kMaxDevices] = {};
std::mutex g_scratch_mu;

uint8_t* scratch_on(int dev) {
    LSH_REQUIRE(dev >= 0 && dev < kMaxDevices, LSH_ERR_INVALID, "device ordinal out of range");
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    if (!g_scratch[dev]) {
        void* p = nullptr;
        LSH_HIP(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
        LSH_REQUIRE(p != nullptr, LSH_ERR_DEVICE, "no device address for the scratch array");
        g_scratch[dev] = static_cast<uint8_t*>(p);
    }
    return g_scratch[dev];
}

size_t align256(size_t b) { return (b + 255) / 256 * 256; }

__global__ void topc_extract_kernel(const uint64_t* __restrict__ keys, int stride, int Q, int C,
                                    int32_t* __restrict__ ids, int32_t* __restrict__ dists) {
    const size_t total = (size_t)Q * C;
    for (size_t t = (size_t)blockIdx.x * blockDim.x + threadIdx.x; t < total; t += (size_t)gridDim.x * blockDim.x) {
        const size_t q = t / C;
        const uint64_t key = keys[q * stride + (t - q * C)];
        const bool empty = key == kEmptyKey;
        ids[t] = empty ? -1 : (int32_t)(uint32_t)key;
        if (dists) dists[t] = empty ? -1 : (int32_t)(key >> 32);
    }
}

}  // namespace

void hamming_topc_device_noalloc(const uint32_t* ds, int64_t N, const uint32_t* qs, int64_t Q, int words, int C,
                                 int32_t* ids, int32_t* dists, hipStream_t s) {
    LSH_REQUIRE(N >= 0 && Q >= 0 && words > 0 && C >= 1, LSH_ERR_INVALID, "bad Hamming top-C arguments");
    LSH_REQUIRE(N < 0x7FFFFFFF && Q < 0x7FFFFFFF, LSH_ERR_INVALID, "N_data and N_query must be below 2^31");
    const int K = C <= kMaxIntTopK ? C : pow2_ceil(C);
    LSH_REQUIRE(K <= kMaxTopK, LSH_ERR_UNSUPPORTED, "C must be <= 4096");
    if (Q == 0) return;
    LSH_REQUIRE(ds || N == 0, LSH_ERR_INVALID, "data sketches are null");
    LSH_REQUIRE(qs && ids, LSH_ERR_INVALID, "query sketches or ids are null");
    int dev = 0;
    LSH_HIP(hipGetDevice(&dev));
    uint8_t* base = scratch_on(dev);

    // Layout: padded query sketches, then the top-K state.
    const int wpad = (words + 7) / 8 * 8;
    const int64_t Qpad = (Q + 255) / 256 * 256;
    const int newcap = newcap_for(K);
    const int stride = K + newcap;
    const size_t b_q = align256(sizeof(uint32_t) * (size_t)Qpad * wpad);
    const size_t b_keys = align256(sizeof(uint64_t) * (size_t)Q * stride);
    const size_t b_cnt = align256(sizeof(uint32_t) * (size_t)Q);
    const size_t b_thr = align256(sizeof(uint64_t) * (size_t)Q);
    const size_t b_ovf = align256(sizeof(uint32_t) * (size_t)Q);
    const size_t need = b_q + b_keys + b_cnt + b_thr + b_ovf;
    LSH_REQUIRE(need <= kNoAllocScratchBytes, LSH_ERR_NOMEM,
                "N_query x C too large for the allocation-free scratch: " + std::to_string(need) + " bytes needed, " +
                    std::to_string(kNoAllocScratchBytes) + " available");
    uint32_t* qpad = reinterpret_cast<uint32_t*>(base);
    uint8_t* p = base + b_q;
    TopKView v{};
    v.keys = reinterpret_cast<uint64_t*>(p);
    p += b_keys;
    v.cnt = reinterpret_cast<uint32_t*>(p);
    p += b_cnt;
    v.thr = reinterpret_cast<uint64_t*>(p);
    p += b_thr;
    v.ovf = reinterpret_cast<uint32_t*>(p);
    v.K = K;
    v.newcap = newcap;
    v.stride = stride;
    v.Q = (int)Q;
    v.int_keys = K <= kMaxIntTopK ? 1 : 0;
    v.fresh = 0;

    // Every call starts from cleared state: the previous call's contents stay
    // in the array.
    launch_pad_rows_u32(qs, Q, words, Qpad, wpad, qpad, s);
    launch_topk_clear(v, s);
    for (int64_t p0 = 0; p0 < N; p0 += newcap) {
        SketchScanArgs a{};
        a.data = ds;
        a.qs = qpad;
        a.words = wpad;
        a.dwords = words;
        a.Q = (int)Q;
        a.p0 = p0;
        a.p1 = std::min<int64_t>(N, p0 + newcap);
        a.kind = 0;
        launch_sketch_scan(a, v, s);
        launch_select_merge(v, s);
    }
    launch_sort_rows(v, s);
    const size_t total = (size_t)Q * C;
    const unsigned grid = (unsigned)std::min<size_t>((total + 255) / 256, 65535 * 8);
    hipLaunchKernelGGL(topc_extract_kernel, dim3(grid), dim3(256), 0, s, v.keys, v.stride, (int)Q, C, ids, dists);
    check_launch("topc_extract_kernel");
}

}  // namespace lshgpu
