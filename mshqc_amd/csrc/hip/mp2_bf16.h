// Split-BF16 form of the MP2 kernels (interfaces used by the MP2 driver).
#pragma once
#include "common.h"
#include "pair_kernel.h"

// Largest supported slice count per operand.
constexpr int BF16_MAX_SLICES = 6;
// Slices of the gradient contractions: 3 keep the relative gradient error
// near 1e-7, below the 1e-6 bar, at 6 instead of up to 21 BF16 products.
constexpr int BF16_GRAD_SLICES = 3;
// Largest slice count of the split-BF16 AO->MO transform.
constexpr int BF16_TRANSFORM_SLICES = 4;

// Device-resident BF16 slices of the B operands of one MP2 problem.
struct Bf16Operands {
    int slices = 0;   // slices of the (ia|jb) operands
    int gslices = 0;  // slices of the gradient-GEMM operands
    int Q = 0, Qp = 0, qblocks = 0;
    int spb = 8;      // pair-kernel stages (of 16 Q) per exponent block, 8 or 10
    int ol = 0, vl = 0, orr = 0, vr = 0;
    int vEl = 0, vEr = 0;  // row stride of the pair-kernel exponent tables
    // Pair-kernel operands: slices [Qp/16][ceil(o*v/32) + 1][S][32][16] (bf16 bits) and
    // exponents [Qp/(16 spb)][o][vE], left and right (right = left in the
    // symmetric modes).
    const uint16_t* sl = nullptr;
    const int8_t* xl = nullptr;
    const uint16_t* sr = nullptr;
    const int8_t* xr = nullptr;
    // Gradient-GEMM operands, per occupied block of bo orbitals: block b
    // holds columns (i, a) of its orbitals as one compact run of nb*v values,
    // padded once to kw = round_up(nb*v, 256): slices [S][Q][kw] at
    // g + b * S*Q*kwb and exponents [kw/256][Qx] (Qx = Q rounded up to 128)
    // at ge + b * Qx*kwb/256, with kwb the kw of a full block.  Left and
    // right.
    int bol = 0, bor = 0;
    long kwbl = 0, kwbr = 0;
    const uint16_t* gl = nullptr;
    const int8_t* gel = nullptr;
    const uint16_t* gr = nullptr;
    const int8_t* ger = nullptr;
    // FP64 self terms sum_Q Bl^Q_ia Br^Q_ia, [nself_o][nself_v] (null when
    // not requested); see PairArgs::gself.
    const double* gself = nullptr;
    int nself_o = 0, nself_v = 0;
};

// Slice Bl (and Br when not null) into BF16 planes; `grad` adds the GEMM
// layout for occupied blocks of bo_l (left) and bo_r (right) orbitals.
// `self_terms` also forms the FP64 self terms sum_Q Bl_ia Br_ia (Br = Bl
// when null) for i < min(ol, orr), a < min(vl, vr).
int bf16_prepare(mshqc_hip_ctx* ctx, int slices, const double* Bl, int ol, int vl,
                 const double* Br, int orr, int vr, int Q, bool grad, int bo_l, int bo_r,
                 bool self_terms, Bf16Operands* out);
// AO->MO transform B^Q = C_occ^T L^Q C_vir (B: [Q][o][v]) on split-BF16
// GEMMs with S slices per operand, occupied first, in chunks of Q.
int bf16_transform(mshqc_hip_ctx* ctx, const double* L, int Q, int n, const double* Cocc, int o,
                   const double* Cvir, int v, int S, double* B);
// Pair-tile kernel on BF16 MFMA with the shared FP64 epilogue.
int bf16_launch_pair(mshqc_hip_ctx* ctx, int mode, const PairArgs& a, const Bf16Operands& ops);
// Gradient contractions of one occupied block pair from the FP64 amplitude
// block Z (rows (i,a) of I, columns (j,b) of J):
//   dBl[q][(i,a)] += sum Br[q][(j,b)] Z[(i,a)][(j,b)]
//   dBr[q][(j,b)] += sum Bl[q][(i,a)] Z[(i,a)][(j,b)]   (when `second`)
int bf16_grad_contract(mshqc_hip_ctx* ctx, const Bf16Operands& ops, const double* Z, long ldz,
                       int I0, int nI, int J0, int nJ, double* dBl, double* dBr, bool second);
