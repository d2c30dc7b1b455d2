// fp4.hpp -- sketch bits as FP4 (e2m1) matrix-core operands.
//
// A sketch bit b becomes s = +1 (b = 0, nibble 0x2) or -1 (b = 1, nibble
// 0xA); for two sketches of B bits, sum_k s_a[k] s_b[k] = B - 2 * Hamming,
// exact in FP32.  For K-step s (bits 64s .. 64s+63) lane l supplies row l%32
// and the 32 bits of word 2s + l/32 of that row, one nibble per bit, in the
// low four registers of the operand.  Accumulator register r of lane l of a
// 32x32 product then holds row (r/4)*8 + (l/32)*4 + r%4 of the A operand and
// row l%32 of the B operand.
#pragma once

#include "device_common.hpp"

namespace lshgpu {
namespace fp4 {

typedef int v4i __attribute__((ext_vector_type(4)));
typedef int v8i __attribute__((ext_vector_type(8)));
typedef float v16f __attribute__((ext_vector_type(16)));

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

__device__ inline v16f mma(v4i a, v4i b, v16f acc) {
    return __builtin_amdgcn_mfma_scale_f32_32x32x64_f8f6f4(widen(a), widen(b), acc, 4, 4, 0, 0x7F7F7F7F, 0,
                                                           0x7F7F7F7F);
}

// Row of the A operand held by accumulator register r of lane l.
__device__ inline int acc_row(int r, int lane) { return (r >> 2) * 8 + (lane >> 5) * 4 + (r & 3); }

}  // namespace fp4
}  // namespace lshgpu
