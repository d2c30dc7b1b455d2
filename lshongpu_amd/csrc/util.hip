// util.hip -- small data-movement kernels used by the host engine.
#include "device_common.hpp"

namespace lshgpu {

namespace {

template <class T>
__global__ void gather_rows_kernel(const T* __restrict__ src, int cols, const int32_t* __restrict__ idx,
                                   int m, T* __restrict__ dst) {
    const size_t total = (size_t)m * cols;
    for (size_t t = (size_t)blockIdx.x * blockDim.x + threadIdx.x; t < total;
         t += (size_t)gridDim.x * blockDim.x) {
        const int r = (int)(t / cols);
        const int c = (int)(t % cols);
        dst[t] = src[(size_t)idx[r] * cols + c];
    }
}

template <class T>
__global__ void scatter_rows_kernel(const T* __restrict__ src, int src_stride, int cols,
                                    const int32_t* __restrict__ idx, int m, T* __restrict__ dst,
                                    int dst_stride) {
    const size_t total = (size_t)m * cols;
    for (size_t t = (size_t)blockIdx.x * blockDim.x + threadIdx.x; t < total;
         t += (size_t)gridDim.x * blockDim.x) {
        const int r = (int)(t / cols);
        const int c = (int)(t % cols);
        dst[(size_t)idx[r] * dst_stride + c] = src[(size_t)r * src_stride + c];
    }
}

// Copies rows of `cols` words into rows of `pcols` words (pcols >= cols),
// zero-filling the padding and the rows in [n, npad).
__global__ void pad_rows_u32_kernel(const uint32_t* __restrict__ src, int64_t n, int cols,
                                    int64_t npad, int pcols, uint32_t* __restrict__ dst) {
    const size_t total = (size_t)npad * pcols;
    for (size_t t = (size_t)blockIdx.x * blockDim.x + threadIdx.x; t < total;
         t += (size_t)gridDim.x * blockDim.x) {
        const int64_t r = (int64_t)(t / pcols);
        const int c = (int)(t % pcols);
        dst[t] = (r < n && c < cols) ? src[(size_t)r * cols + c] : 0u;
    }
}

// rows[src + 1 + i] = rows[src] for i < n (rows of `cols` floats).
__global__ void repeat_row_f32_kernel(float* __restrict__ rows, int cols, int src, int n) {
    const size_t total = (size_t)n * cols;
    for (size_t t = (size_t)blockIdx.x * blockDim.x + threadIdx.x; t < total;
         t += (size_t)gridDim.x * blockDim.x)
        rows[(size_t)(src + 1) * cols + t] = rows[(size_t)src * cols + t % cols];
}

__global__ void fill_i32_kernel(int32_t* p, size_t n, int32_t v) {
    for (size_t t = (size_t)blockIdx.x * blockDim.x + threadIdx.x; t < n;
         t += (size_t)gridDim.x * blockDim.x)
        p[t] = v;
}

unsigned grid_of(size_t total) {
    return (unsigned)std::max<size_t>(1, std::min<size_t>((total + 255) / 256, 65535 * 4));
}

}  // namespace

void launch_gather_rows_f32(const float* src, int cols, const int32_t* idx, int m, float* dst,
                            hipStream_t s) {
    if (m <= 0) return;
    hipLaunchKernelGGL(gather_rows_kernel<float>, dim3(grid_of((size_t)m * cols)), dim3(256), 0, s, src,
                       cols, idx, m, dst);
    check_launch("gather_rows_kernel<float>");
}

void launch_repeat_row_f32(float* rows, int cols, int src, int n, hipStream_t s) {
    if (n <= 0 || cols <= 0) return;
    hipLaunchKernelGGL(repeat_row_f32_kernel, dim3(grid_of((size_t)n * cols)), dim3(256), 0, s, rows, cols, src, n);
    check_launch("repeat_row_f32_kernel");
}

void launch_gather_rows_u32(const uint32_t* src, int cols, const int32_t* idx, int m, uint32_t* dst,
                            hipStream_t s) {
    if (m <= 0) return;
    hipLaunchKernelGGL(gather_rows_kernel<uint32_t>, dim3(grid_of((size_t)m * cols)), dim3(256), 0, s,
                       src, cols, idx, m, dst);
    check_launch("gather_rows_kernel<uint32_t>");
}

void launch_scatter_rows_u64(const uint64_t* src, int src_stride, int cols, const int32_t* idx, int m,
                             uint64_t* dst, int dst_stride, hipStream_t s) {
    if (m <= 0) return;
    hipLaunchKernelGGL(scatter_rows_kernel<uint64_t>, dim3(grid_of((size_t)m * cols)), dim3(256), 0, s,
                       src, src_stride, cols, idx, m, dst, dst_stride);
    check_launch("scatter_rows_kernel<uint64_t>");
}

void launch_scatter_rows_pts(const lsh_point_t* src, int cols, const int32_t* idx, int m,
                             lsh_point_t* dst, hipStream_t s) {
    if (m <= 0) return;
    hipLaunchKernelGGL(scatter_rows_kernel<lsh_point_t>, dim3(grid_of((size_t)m * cols)), dim3(256), 0,
                       s, src, cols, cols, idx, m, dst, cols);
    check_launch("scatter_rows_kernel<lsh_point_t>");
}

void launch_pad_rows_u32(const uint32_t* src, int64_t n, int cols, int64_t npad, int pcols,
                         uint32_t* dst, hipStream_t s) {
    if (npad <= 0) return;
    hipLaunchKernelGGL(pad_rows_u32_kernel, dim3(grid_of((size_t)npad * pcols)), dim3(256), 0, s, src, n,
                       cols, npad, pcols, dst);
    check_launch("pad_rows_u32_kernel");
}

void launch_fill_i32(int32_t* p, size_t n, int32_t v, hipStream_t s) {
    if (n == 0) return;
    hipLaunchKernelGGL(fill_i32_kernel, dim3(grid_of(n)), dim3(256), 0, s, p, n, v);
    check_launch("fill_i32_kernel");
}

}  // namespace lshgpu
