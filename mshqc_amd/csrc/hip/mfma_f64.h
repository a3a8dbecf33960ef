// FP64 matrix-core primitive of gfx950: v_mfma_f64_16x16x4_f64.
//
// One wavefront (64 lanes) computes D(16x16) = A(16x4) * B(4x16) + C.
//   A operand: lane l supplies A[row = l % 16][k = l / 16]
//   B operand: lane l supplies B[k = l / 16][col = l % 16]
//   C/D:       lane l, element r (0..3) holds D[row = l / 16 + 4 * r][col = l % 16]
// (rows interleaved across the four 16-lane groups; measured on gfx950 with
// one-hot operands, and exercised by the GEMM and pair-kernel parity tests).
#pragma once
#include <hip/hip_runtime.h>

typedef double v4d __attribute__((ext_vector_type(4)));

__device__ __forceinline__ v4d mfma_f64(double a, double b, v4d c) {
    return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
}

__device__ __forceinline__ int mfma_row(int lane, int r) { return (lane >> 4) + 4 * r; }
__device__ __forceinline__ int mfma_col(int lane) { return lane & 15; }
