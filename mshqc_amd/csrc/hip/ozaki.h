// Split-BF16 arithmetic on gfx950 BF16 matrix cores (v_mfma_f32_16x16x32_bf16,
// v_mfma_f32_32x32x16_bf16).
//
// Each FP64 operand element x of a block (one row over one block of the
// contraction index) is written as
//     x = 2^{e-8} * sum_{p < S} s_p,   s_p = m_p * 2^{-9p}  (+ truncation < 2^{e-9S})
// where 2^e bounds |x| over the block and every m_p is an integer with
// |m_p| <= 256, so each s_p is exact in BF16 (8 significant bits; integers
// up to 256 are exact).  m_0 = rint(256 x / 2^e) leaves a residual of at most
// 1/2 in its unit, so each later slice takes 9 bits (m_p = rint(512 r)), not
// 8: at four slices the truncation then sits near 2^-36 of the block scale
// instead of 2^-32, with the same number of products.  A slice product
// s_p s_q = m_p m_q 2^{-9(p+q)} is exact.  Products are summed on the FP32
// MFMA accumulator in three tiers by class c = p+q; for blocks of up to 160
// products per output (the pair kernel's 128- or 160-wide Q blocks):
//   tier 0 (c = 0): integers <= 2^16 each, sums <= 160 * 2^16 < 2^24    -> exact
//   tier 1 (c = 1): multiples of 2^-9, <= 2^7 each, sums below 2^24
//                   units in practice (2^24.3 at worst, where a last
//                   bit of 2^-33 of the block scale may round)
// (with 256-wide blocks, tier-0 sums up to 2^24 measurably lost their last
// bit on gfx950 MFMAs, so the pair kernel's blocks stay at 128, where the
// coherent self-term sums reach that range; the three-slice gradient GEMM,
// whose tier-0 sums are random in sign and whose truncation sits near
// 2^-26, uses 512-wide k blocks (the transforms 256) and two tiers, c = 0
// and c >= 1)
//   tier 2 (c >= 2): below 2^-2 per product; FP32 rounding there sits near
//                    2^-42 of the tier-0 scale, under the truncation error.
// After every block the tiers are added in FP64 and scaled by 2^{eA+eB-16}.
// Products with p + q <= S-1 are kept (S(S+1)/2 MFMAs per output tile
// instead of one FP64 MFMA).
//
// MFMA 16x16x32 bf16 lane layout (measured on gfx950 with one-hot operands):
//   A: lane l holds A[row = l%16][k = 8*(l/16) + 0..7]
//   B: lane l holds B[k = 8*(l/16) + 0..7][col = l%16]
//   D: lane l, element r holds D[row = 4*(l/16) + r][col = l%16]
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>

typedef float v4f __attribute__((ext_vector_type(4)));
typedef float v16f __attribute__((ext_vector_type(16)));
typedef __bf16 v8bf __attribute__((ext_vector_type(8)));
typedef short v8s __attribute__((ext_vector_type(8)));

__device__ __forceinline__ v4f mfma_bf16(const v8s& a, const v8s& b, v4f c) {
    return __builtin_amdgcn_mfma_f32_16x16x32_bf16(__builtin_bit_cast(v8bf, a),
                                                   __builtin_bit_cast(v8bf, b), c, 0, 0, 0);
}

__device__ __forceinline__ int bf16_row(int lane, int r) { return 4 * (lane >> 4) + r; }

// Exponent e with |x| < 2^e for the block maximum `mx` (0 for an all-zero
// block), kept inside the int8 range of the exponent tables:
//   - a block below 2^-128 takes e = -127 (|x| < 2^-127 still holds, so the
//     slicing stays exact, with fewer significant bits in the slices);
//   - a block whose maximum is 2^127 or more, or not finite, returns
//     EXP_OVERFLOW: slice_value() then writes NaN slices, so the split-BF16
//     result is NaN (as the FP64 path overflows on such inputs) instead of a
//     silently wrapped scale.
constexpr int EXP_OVERFLOW = 128;

__device__ __forceinline__ int block_exponent(double mx) {
    if (!(mx > 0.0)) return 0;
    if (!(mx < 0x1p127)) return EXP_OVERFLOW;
    int e;
    frexp(mx, &e);
    return e < -127 ? -127 : e;
}

// The exponent as stored in the int8 tables (an overflowed block stores 127;
// its slices are NaN).
__device__ __forceinline__ int8_t exp_byte(int e) { return (int8_t)(e > 127 ? 127 : e); }

// 2^k as a double for |k| < 1000.
__device__ __forceinline__ double pow2(int k) {
    return __longlong_as_double((long long)(k + 1023) << 52);
}

// BF16 bit pattern of a float carrying at most 8 significant bits (exact).
__device__ __forceinline__ uint16_t bf16_bits_exact(float m) {
    return (uint16_t)(__float_as_uint(m) >> 16);
}

// Slice x (|x| < 2^e) into S BF16 values s_p = m_p 2^{-9p} (bit patterns).
template <int S>
__device__ __forceinline__ void slice_value(double x, int e, uint16_t (&out)[S]) {
    if (e == EXP_OVERFLOW) {
#pragma unroll
        for (int p = 0; p < S; ++p) out[p] = 0x7FC0;  // BF16 quiet NaN
        return;
    }
    double r = x * pow2(-e);  // |r| < 1 (exact scaling)
    float sc = 1.0f;
#pragma unroll
    for (int p = 0; p < S; ++p) {
        const double t = r * (p == 0 ? 256.0 : 512.0);  // |t| <= 256
        const double m = rint(t);
        out[p] = bf16_bits_exact((float)m * sc);
        r = t - m;  // exact: |r| <= 1/2
        sc *= (1.0f / 512.0f);
    }
}

// Accumulator tier of a slice product p, q.
__host__ __device__ constexpr int ozaki_tier(int p, int q) { return p + q < 2 ? p + q : 2; }

// Number of kept products for S slices.
__host__ __device__ constexpr int ozaki_products(int S) { return S * (S + 1) / 2; }
