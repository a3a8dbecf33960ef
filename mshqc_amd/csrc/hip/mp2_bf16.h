// Split-BF16 form of the MP2 kernels (interfaces used by the MP2 driver).
#pragma once
#include "common.h"
#include "pair_kernel.h"

// Largest supported slice count per operand.
constexpr int BF16_MAX_SLICES = 6;

// Device-resident BF16 slices of the B operands of one MP2 problem.
struct Bf16Operands {
    int slices = 0;
    int Q = 0, qblocks = 0;
    // Left and right operands: slices [k][Q][o][v] as bf16 bits and per
    // (q-block, i, a) power-of-two exponents.
    const uint16_t* sl = nullptr;
    const int8_t* xl = nullptr;
    const uint16_t* sr = nullptr;
    const int8_t* xr = nullptr;
};

// Slice Bl (and Br when not null) into BF16 planes.
int bf16_prepare(mshqc_hip_ctx* ctx, int slices, const double* Bl, int ol, int vl,
                 const double* Br, int orr, int vr, int Q, Bf16Operands* out);
// Pair-tile kernel on BF16 MFMA with the shared FP64 epilogue.
int bf16_launch_pair(mshqc_hip_ctx* ctx, int mode, const PairArgs& a, const Bf16Operands& ops);
// C += A * op(X) on BF16 MFMA with split operands (same contract as gemm_f64).
int bf16_gemm(mshqc_hip_ctx* ctx, int slices, const GemmF64& g);
