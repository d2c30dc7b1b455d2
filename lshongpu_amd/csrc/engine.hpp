// engine.hpp -- the resident index and its search paths.
#pragma once

#include "lshgpu_internal.hpp"

struct lsh_index {
    // Resolved configuration.
    int impl = 2;
    int dfunc = 1;
    int d = 0;
    int64_t N = 0;
    int sketchDim = 0;
    int framework = 0;
    int keysImpl = 0;
    int bucketKeyBits = 0;
    int tables = 0;
    bool withSketch = false;
    int candidates = 4096;
    int margin = 32;
    int device = 0;

    hipStream_t stream = nullptr;
    bool profiling = false;
    double phase_ms[8] = {0, 0, 0, 0, 0, 0, 0, 0};

    // Working copy of the data after the path's preprocessing.
    lshgpu::DevBuf<float> data;
    lshgpu::DevBuf<float> data_inv;  // 1/||x|| for the angular kinds
    bool gj_transform = false;       // shift + normalise (generalized Jaccard, weighted MinHash)
    bool normalize = false;          // unit rows (SimHash-JL, cross-polytope keys)
    lshgpu::DevBuf<int> mins;        // shift of the generalized-Jaccard transform

    // Sketches.
    int words = 0;   // uint32 words per point
    int wpad = 0;    // words padded to a multiple of 8
    lshgpu::DevBuf<uint32_t> dsk;    // [N][wpad]
    lshgpu::DevBuf<float> planesT;   // SimHash [words][d][32], JL [groups][d][32]
    int sdim = 0;                    // JL sketch length
    lshgpu::DevBuf<float> djl;       // [N][sdim]

    // Weighted MinHash.
    int wmh_M = 0;
    int wmh_mode = 0;                // 0 counts (impl 4), 1 one-bit words (impl 5)
    lshgpu::DevBuf<int> wmh_owner;   // [M] component of each unit of mass
    lshgpu::DevBuf<int> wmh_start;   // [d] first unit of each component
    lshgpu::DevBuf<uint8_t> wmh_bitmap;
    int wmh_bitmap_len = 0;
    lshgpu::DevBuf<float> wmh_draws;     // [hashes][R] draw values r = M*U
    lshgpu::DevBuf<int> wmh_draw_owner;  // [hashes][R] owner component of each draw
    int wmh_R = 0;

    // Per-search scratch (grow-only).
    lshgpu::DevBuf<float> q, qT, q_inv, qjl, qjlT;
    lshgpu::DevBuf<uint32_t> qsk;
    lshgpu::DevBuf<lsh_point_t> out;
    lshgpu::DevBuf<uint32_t> flags;
    lshgpu::TopKState st;
    int64_t scan_launches = 0;
    int64_t mfma_launches = 0;   // matrix-core Hamming scan launches of the last search
    int64_t mfma_points = 0;     // data points those launches covered

    size_t device_bytes() const;
    void build_wmh_tables(const float* x, int64_t n, const float* extra, int64_t n_extra, hipStream_t s);
    void wmh_sketch_rows(const float* x, int64_t n, uint32_t* out, hipStream_t s);
    void build(const float* src, bool on_device, const float* extra_host, int64_t n_extra);
    void search(const float* queries, int Q, int k, bool q_on_device, lsh_point_t* results,
                bool r_on_device, hipStream_t s);
};

namespace lshgpu {

// One-shot helpers behind the reference-shaped entry points.
int calculate_k(int k);

// Sketch-stage entry points (host in, host out).
void simhash_host(const float* x, int64_t n, int d, int words, int hashBits, const float* planes,
                  uint32_t* out, int device);
void jl_host(const float* x, int64_t n, int d, int sdim, const float* planes, float* out, int device);
void gj_transform_host(float* data, int64_t N, float* queries, int64_t Q, int d, int device);
void hamming_topc_host(const uint32_t* ds, int64_t N, const uint32_t* qs, int64_t Q, int words, int C,
                       int32_t* ids, int32_t* dists, int device);
void wmh_host(const float* x, int64_t n, int d, int sketchDim, int one_bit, uint32_t* out, int device);
void crosspoly_host(const float* x, int64_t n, int d, int tables, const int32_t* signs, uint16_t* keys,
                    int device);

}  // namespace lshgpu
