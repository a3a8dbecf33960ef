// Internal interfaces of the MP2 path: the shared pair-tile epilogue and the
// kernel launchers of each precision.
#pragma once
#include "common.h"
#include "mfma_f64.h"
#include "pair_kernel.h"

// Splits of the operands for the split-BF16 form (see mp2_bf16.hip).
struct SlicedB;

int launch_pair_f64(mshqc_hip_ctx* ctx, int mode, const PairArgs& a);

// Wave-wide sum over the lanes selected by the xor masks in [lo, hi).
__device__ __forceinline__ double xor_sum(double v, int lo_mask, int hi_mask) {
    for (int m = lo_mask; m < hi_mask; m <<= 1) v += __shfl_xor(v, m, 64);
    return v;
}

// Fused denominator epilogue for one 64x64 tile held by 4 waves in the
// v_mfma 16x16 accumulator layout (wave tile 32x32 = 2x2 subtiles).
// Writes the weighted tile partials (energy, and for gradients the dE/dD
// sums by row/column) and, when requested, the gradient amplitudes z or t2.
// ROWMAP selects the accumulator row layout: 0 = FP64 MFMA (row = l/16 + 4r),
// 1 = FP32 MFMA of the BF16 path (row = 4(l/16) + r).  Waves with
// active == false only take part in the barrier (waves 0..3 own the tile).
template <int MODE, int ROWMAP = 0>
__device__ __forceinline__ void pair_epilogue(const PairArgs& p, int i, int j, double w, int A0,
                                              int B0, bool mirror, int lane, int wave, int wm,
                                              int wn, const v4d (&accG)[2][2],
                                              const v4d (&accH)[2][2], double (*red_row)[64],
                                              double (*red_col)[64], double* red_e, double* red_s,
                                              long work, bool active = true) {
#pragma clang fp contract(off)
    const bool grad = p.s_part != nullptr;
    auto rowof = [](int l, int r) { return ROWMAP ? 4 * (l >> 4) + r : (l >> 4) + 4 * r; };
    if (!active) {
        __syncthreads();
        return;
    }
    const double ei = p.eol[i] + p.eor[j];
    double e = 0.0, s = 0.0;
    double rowq[2][4], colq[2];
#pragma unroll
    for (int a = 0; a < 2; ++a) {
        colq[a] = 0.0;
#pragma unroll
        for (int r = 0; r < 4; ++r) rowq[a][r] = 0.0;
    }
    double* zrow_base = nullptr;
    if (MODE == MODE_T2 || MODE == MODE_T2OS) zrow_base = p.Z + ((long)i * p.orr + j) * p.vl * p.vr;
#pragma unroll
    for (int sa = 0; sa < 2; ++sa)
#pragma unroll
        for (int tb = 0; tb < 2; ++tb)
#pragma unroll
            for (int r = 0; r < 4; ++r) {
                const int a = A0 + wm + 16 * sa + rowof(lane, r);
                const int b = B0 + wn + 16 * tb + mfma_col(lane);
                if (a >= p.vl || b >= p.vr) continue;
                const double D = ei - p.evl[a] - p.evr[b];
                const double G = accG[sa][tb][r];
                const double H = pair_mode_sym(MODE) ? accH[sa][tb][r] : 0.0;
                double e1, m1, z1;
                pair_terms<MODE>(G, H, D, e1, m1, z1);
                double ee = e1, q = m1, e2 = 0.0, m2 = 0.0, z2 = 0.0;
                if (mirror) {
                    pair_terms<MODE>(H, G, D, e2, m2, z2);
                    ee += e2;
                    q += m2;
                }
                e += ee;
                if (grad) {
                    s += q;
                    rowq[sa][r] += q;
                    colq[tb] += q;
                }
                if (MODE == MODE_T2 || MODE == MODE_T2OS) {
                    const bool ok = fabs(D) >= 1e-12;
                    zrow_base[(long)a * p.vr + b] = ok ? G / D : 0.0;
                    if (mirror) zrow_base[(long)b * p.vr + a] = ok ? H / D : 0.0;
                } else if (p.Z) {
                    const long r1 = (long)(i - p.zi0) * p.vl, c1 = (long)(j - p.zj0) * p.vr;
                    p.Z[(r1 + a) * p.ldz + c1 + b] = z1;
                    if (mirror) p.Z[(r1 + b) * p.ldz + c1 + a] = z2;
                    if (p.ztrans && i != j) {
                        const long r2 = (long)(j - p.zi0) * p.vl, c2 = (long)(i - p.zj0) * p.vr;
                        p.Z[(r2 + b) * p.ldz + c2 + a] = z1;
                        if (mirror) p.Z[(r2 + a) * p.ldz + c2 + b] = z2;
                    }
                }
            }
    // Deterministic tile reductions: fixed shuffle trees, then fixed-order
    // sums of the per-wave results in LDS.
    e = xor_sum(e, 1, 64);
    if (grad) {
        s = xor_sum(s, 1, 64);
#pragma unroll
        for (int sa = 0; sa < 2; ++sa)
#pragma unroll
            for (int r = 0; r < 4; ++r) {
                const double v = xor_sum(rowq[sa][r], 1, 16);
                if ((lane & 15) == 0) red_row[wave & 1][wm + 16 * sa + rowof(lane, r)] = v;
            }
#pragma unroll
        for (int tb = 0; tb < 2; ++tb) {
            const double v = xor_sum(colq[tb], 16, 64);
            if ((lane >> 4) == 0) red_col[wave >> 1][wn + 16 * tb + mfma_col(lane)] = v;
        }
    }
    if (lane == 0) {
        red_e[wave] = e;
        red_s[wave] = s;
    }
    __syncthreads();
    const int tid = threadIdx.x;
    if (tid == 0) p.e_part[work] = w * (((red_e[0] + red_e[1]) + red_e[2]) + red_e[3]);
    if (grad) {
        if (tid == 0) p.s_part[work] = w * (((red_s[0] + red_s[1]) + red_s[2]) + red_s[3]);
        if (tid < 64) p.vrow[work * 64 + tid] = w * (red_row[0][tid] + red_row[1][tid]);
        else if (tid < 128)
            p.vcol[work * 64 + tid - 64] = w * (red_col[0][tid - 64] + red_col[1][tid - 64]);
    }
}
