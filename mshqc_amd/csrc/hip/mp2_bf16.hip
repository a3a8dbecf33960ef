// Split-BF16 form of the MP2 kernels on gfx950 BF16 matrix cores.
//
//   slice_pair_kernel   B[Q][o][v] -> BF16 slices [S][o*v][Qp] with one exponent
//                       per (i,a) and 128-wide Q block (per-Q-block scaling)
//   pair_bf16_kernel    (ia|jb) tiles from the slices, exact FP32 class sums per
//                       Q block, FP64 across blocks, then the shared FP64
//                       denominator epilogue (same one as the FP64 kernel)
//   slice_rows_kernel / slice_cols_kernel
//                       row- and column-blocked slices of the gradient GEMM
//                       operands (B and the amplitude block Z, Z^T), one
//                       exponent per row and 128-wide k block
//   gemm_bf16_nt_kernel C[m][n] += sum_k A[m][k] X[n][k] from slices
// The arithmetic contract is described in ozaki.h.
#include <algorithm>

#include "mp2_bf16.h"
#include "mp2_internal.h"
#include "ozaki.h"
#include "tiles.h"

namespace {

constexpr int QB = 128;   // Q block: exponent block and exact-FP32 span of the pair kernel
constexpr int KC = 32;    // contraction chunk (one MFMA K step)
constexpr int LROW = 64;  // LDS bytes per 32-element BF16 row
constexpr int GB = 128;   // exponent block of the gradient-GEMM operands along k
constexpr int GT = 128;   // gradient-GEMM workgroup tile (rows and columns)

// LDS byte offset of 16-byte segment `sg` of panel row `row`: the segment is
// XOR-swizzled by (row >> 1) & 3, which makes the 16-lane fragment reads and
// the 4-lanes-per-row stores free of bank conflicts (measured on gfx950).
__device__ __forceinline__ int lds_off(int row, int sg) { return row * LROW + ((sg ^ ((row >> 1) & 3)) << 4); }

inline int ceil_div(long a, long b) { return (int)((a + b - 1) / b); }

// ------------------------------------------------------------- slicing

// B[Q][ov] (row-major, rows = q) -> Ps[S][ov][Qp], Pe[ov][Qb].  One workgroup
// per 64 columns x 128 q; the tile is transposed through LDS so both the
// FP64 reads and the BF16 writes are contiguous.
template <int S>
__global__ __launch_bounds__(256) void slice_pair_kernel(const double* __restrict__ B, int Q, int ov,
                                                         int Qp, int Qb, uint16_t* __restrict__ Ps,
                                                         int8_t* __restrict__ Pe) {
    __shared__ double T[QB][65];
    __shared__ int ex[64];
    const int r0 = blockIdx.x * 64, q0 = blockIdx.y * QB, tid = threadIdx.x;
    {
        const int c = tid & 63;
        for (int k = 0; k < QB / 4; ++k) {
            const int qq = (tid >> 6) + 4 * k;
            double x = 0.0;
            if (r0 + c < ov && q0 + qq < Q) x = B[(long)(q0 + qq) * ov + r0 + c];
            T[qq][c] = x;
        }
    }
    __syncthreads();
    if (tid < 64) {
        double mx = 0.0;
        for (int qq = 0; qq < QB; ++qq) mx = fmax(mx, fabs(T[qq][tid]));
        ex[tid] = block_exponent(mx);
    }
    __syncthreads();
    const int r = tid >> 2, seg = tid & 3;  // row r, q in [32 seg, 32 seg + 32)
    if (r0 + r >= ov) return;
    const int e = ex[r];
    const long slice = (long)ov * Qp;
    uint16_t* dst = Ps + (long)(r0 + r) * Qp + q0 + 32 * seg;
    for (int q8 = 0; q8 < 32; q8 += 8) {
        uint16_t out[S][8];
#pragma unroll
        for (int u = 0; u < 8; ++u) {
            uint16_t m[S];
            slice_value<S>(T[32 * seg + q8 + u][r], e, m);
#pragma unroll
            for (int p = 0; p < S; ++p) out[p][u] = m[p];
        }
#pragma unroll
        for (int p = 0; p < S; ++p)
            *reinterpret_cast<uint4*>(dst + p * slice + q8) = *reinterpret_cast<uint4*>(out[p]);
    }
    if (seg == 0) Pe[(long)(r0 + r) * Qb + blockIdx.y] = (int8_t)e;
}

// src[R][ld] with column index seg*v + b (seg < nseg, b < v)  ->
// dst[S][R][nseg*vp] (column seg*vp + b, zero padded), ex[R][nseg*(vp/GB)].
// One wave per (row, seg, GB-wide b tile); each lane slices two elements.
template <int S>
__global__ __launch_bounds__(256) void slice_rows_kernel(const double* __restrict__ src, long ld, int R,
                                                         int nseg, int v, int vp,
                                                         uint16_t* __restrict__ dst,
                                                         int8_t* __restrict__ ex) {
    const int ntv = vp / GB;
    const long job = (long)blockIdx.x * 4 + (threadIdx.x >> 6);
    const long njobs = (long)R * nseg * ntv;
    if (job >= njobs) return;
    const int lane = threadIdx.x & 63;
    const int row = (int)(job / (nseg * ntv));
    const int rem = (int)(job % (nseg * ntv));
    const int seg = rem / ntv, tile = rem % ntv;
    const int b0 = tile * GB + lane, b1 = b0 + 64;
    const double* s = src + (long)row * ld + (long)seg * v;
    const double x0 = b0 < v ? s[b0] : 0.0;
    const double x1 = b1 < v ? s[b1] : 0.0;
    double mx = fmax(fabs(x0), fabs(x1));
    for (int m = 1; m < 64; m <<= 1) mx = fmax(mx, __shfl_xor(mx, m, 64));
    const int e = block_exponent(mx);
    uint16_t s0[S], s1[S];
    slice_value<S>(x0, e, s0);
    slice_value<S>(x1, e, s1);
    const long kw = (long)nseg * vp;
    const long slice = (long)R * kw;
    uint16_t* d = dst + (long)row * kw + (long)seg * vp + tile * GB + lane;
#pragma unroll
    for (int p = 0; p < S; ++p) {
        d[p * slice] = s0[p];
        d[p * slice + 64] = s1[p];
    }
    if (lane == 0) ex[(long)row * nseg * ntv + rem] = (int8_t)e;
}

// Transposed slicing of Z[(i',a)][(j',b)] (rows i'*v + a, i' < nI; ld = ldz;
// nc = nJ*vr columns):  dst[S][nc][nI*vp] with row n = column of Z and
// column i'*vp + a;  ex[nc][nI*(vp/GB)].  A workgroup transposes a tile of
// 32 columns x GB values of a through LDS.
template <int S>
__global__ __launch_bounds__(256) void slice_cols_kernel(const double* __restrict__ Z, long ldz, int nI,
                                                         int v, int vp, int nc,
                                                         uint16_t* __restrict__ dst,
                                                         int8_t* __restrict__ ex) {
    __shared__ double T[GB][33];
    __shared__ int exs[32];
    const int ntv = vp / GB;
    const int c0 = blockIdx.x * 32;
    const int ip = blockIdx.y / ntv, at = blockIdx.y % ntv, a0 = at * GB;
    const int tid = threadIdx.x;
    {
        const int cc = tid & 31;
        for (int k = 0; k < GB / 8; ++k) {
            const int aa = (tid >> 5) + 8 * k;
            double x = 0.0;
            if (a0 + aa < v && c0 + cc < nc) x = Z[((long)ip * v + a0 + aa) * ldz + c0 + cc];
            T[aa][cc] = x;
        }
    }
    __syncthreads();
    {
        // 8 lanes per column, 16 values each, then a shuffle maximum.
        const int cc = tid >> 3, part = tid & 7;
        double mx = 0.0;
        for (int k = 0; k < GB / 8; ++k) mx = fmax(mx, fabs(T[part * (GB / 8) + k][cc]));
        for (int m = 1; m < 8; m <<= 1) mx = fmax(mx, __shfl_xor(mx, m, 64));
        if (part == 0) exs[cc] = block_exponent(mx);
    }
    __syncthreads();
    const long kw = (long)nI * vp, slice = (long)nc * kw;
    // thread -> (column cc, 16 consecutive a)
    const int cc = tid >> 3, part = tid & 7;
    if (c0 + cc >= nc) return;
    const int e = exs[cc];
    uint16_t* d = dst + (long)(c0 + cc) * kw + (long)ip * vp + a0 + 16 * part;
    for (int u8 = 0; u8 < 16; u8 += 8) {
        uint16_t out[S][8];
#pragma unroll
        for (int u = 0; u < 8; ++u) {
            uint16_t m[S];
            slice_value<S>(T[16 * part + u8 + u][cc], e, m);
#pragma unroll
            for (int p = 0; p < S; ++p) out[p][u] = m[p];
        }
#pragma unroll
        for (int p = 0; p < S; ++p)
            *reinterpret_cast<uint4*>(d + p * slice + u8) = *reinterpret_cast<uint4*>(out[p]);
    }
    if (part == 0) ex[(long)(c0 + cc) * nI * ntv + ip * ntv + at] = (int8_t)e;
}

// ------------------------------------------------ shared MFMA helpers

// 16-byte LDS fragment read: row `row` of a [64 rows][32 k] BF16 panel, the
// 8 elements k = 8*seg .. 8*seg + 7.
__device__ __forceinline__ v8s lds_frag(const char* panel, int row, int seg) {
    return *reinterpret_cast<const v8s*>(panel + lds_off(row, seg));
}

// Class sums of one KC chunk for a 32x32 wave tile: rows from panel pa,
// columns from panel pb (each [S] slices of [64][32] at stride `sstride`).
template <int S>
__device__ __forceinline__ void chunk_mfma(const char* pa, const char* pb, long sstride, int wm,
                                           int wn, int lane, v4f (&acc)[3][2][2]) {
    // All S column slices stay in registers; row slices are loaded one at a
    // time (slice p meets column slices q < S - p).
    v8s b[S][2];
    const int rr = lane & 15, sg = lane >> 4;
#pragma unroll
    for (int q = 0; q < S; ++q)
#pragma unroll
        for (int t = 0; t < 2; ++t) b[q][t] = lds_frag(pb + q * sstride, wn + 16 * t + rr, sg);
#pragma unroll
    for (int p = 0; p < S; ++p) {
        v8s a[2];
#pragma unroll
        for (int s = 0; s < 2; ++s) a[s] = lds_frag(pa + p * sstride, wm + 16 * s + rr, sg);
#pragma unroll
        for (int q = 0; q < S - p; ++q)
#pragma unroll
            for (int s = 0; s < 2; ++s)
#pragma unroll
                for (int t = 0; t < 2; ++t)
                    acc[ozaki_tier(p, q)][s][t] = mfma_bf16(a[s], b[q][t], acc[ozaki_tier(p, q)][s][t]);
    }
}

// acc64 += (tier0 + tier1 + tier2) * 2^(eA(row) + eB(col) - 16); tiers = 0.
__device__ __forceinline__ void flush(v4f (&acc)[3][2][2], v4d (&acc64)[2][2], const int (&eA)[2][4],
                                      const int (&eB)[2]) {
#pragma unroll
    for (int s = 0; s < 2; ++s)
#pragma unroll
        for (int t = 0; t < 2; ++t)
#pragma unroll
            for (int r = 0; r < 4; ++r) {
                const double sum = ((double)acc[0][s][t][r] + (double)acc[1][s][t][r]) +
                                   (double)acc[2][s][t][r];
                acc64[s][t][r] += sum * pow2(eA[s][r] + eB[t] - 16);
            }
#pragma unroll
    for (int c = 0; c < 3; ++c)
#pragma unroll
        for (int s = 0; s < 2; ++s)
#pragma unroll
            for (int t = 0; t < 2; ++t) acc[c][s][t] = v4f{0.f, 0.f, 0.f, 0.f};
}

// ----------------------------------------------------------- pair kernel

template <int MODE, int S>
__global__ __launch_bounds__(MODE == MODE_OS ? 256 : 512) void pair_bf16_kernel(
    PairArgs p, const uint16_t* __restrict__ Psl, const int8_t* __restrict__ Pel,
    const uint16_t* __restrict__ Psr, const int8_t* __restrict__ Per, int Qp, int Qb) {
    constexpr bool SYM = MODE != MODE_OS;
    constexpr int NP = SYM ? 4 : 2;
    constexpr int NT = SYM ? 512 : 256;
    constexpr long PSTRIDE = (long)S * 64 * LROW;  // bytes per panel
    extern __shared__ __align__(16) char lds[];
    __shared__ double red_row[2][64], red_col[2][64], red_e[4], red_s[4];

    const long work = xcd_work_index();
    if (work >= p.nwork) return;
    const int pslot = (int)(work / p.ntiles), tslot = (int)(work % p.ntiles);
    const int2 ij = p.pairs[pslot];
    const int i = ij.x, j = ij.y;
    const double w = p.pw[pslot];
    const int2 tb = p.tiles[tslot];
    const int A0 = tb.x * 64, B0 = tb.y * 64;
    const bool mirror = SYM && (tb.x != tb.y);

    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int wq = wave & 3, isH = wave >> 2;
    const int wm = (wq >> 1) * 32, wn = (wq & 1) * 32;

    // Panel 0: Bl_i[A] (G rows), 1: Br_j[B] (G cols), 2: Br_j[A] (H rows),
    // 3: Bl_i[B] (H cols).  Selected with branches, not an indexed array,
    // so nothing lands in scratch.
    const uint16_t* pb0 = Psl + ((long)i * p.vl + A0) * Qp;
    const uint16_t* pb1 = Psr + ((long)j * p.vr + B0) * Qp;
    const uint16_t* pb2 = Psr + ((long)j * p.vr + A0) * Qp;
    const uint16_t* pb3 = Psl + ((long)i * p.vl + B0) * Qp;
    const long sl_l = (long)p.ol * p.vl * Qp, sl_r = (long)p.orr * p.vr * Qp;
    const int lim0 = p.vl - A0, lim1 = p.vr - B0, lim2 = p.vr - A0, lim3 = p.vl - B0;

    constexpr int NSEG = NP * S * 256;  // 16-byte segments per chunk
    constexpr int PER = (NSEG + NT - 1) / NT;
    uint4 reg[PER];
    auto fetch = [&](int q0) {
#pragma unroll
        for (int u = 0; u < PER; ++u) {
            const int g = tid + u * NT;
            uint4 val = make_uint4(0, 0, 0, 0);
            if (g < NSEG) {
                const int P = g / (S * 256), sl = (g / 256) % S, row = (g & 255) >> 2, sg = g & 3;
                const uint16_t* base = P == 0 ? pb0 : P == 1 ? pb1 : P == 2 ? pb2 : pb3;
                const long pst = (P == 0 || P == 3) ? sl_l : sl_r;
                const int lim = P == 0 ? lim0 : P == 1 ? lim1 : P == 2 ? lim2 : lim3;
                if (row < lim)
                    val = *reinterpret_cast<const uint4*>(base + sl * pst + (long)row * Qp + q0 + 8 * sg);
            }
            reg[u] = val;
        }
    };
    auto stash = [&]() {
#pragma unroll
        for (int u = 0; u < PER; ++u) {
            const int g = tid + u * NT;
            if (g < NSEG) {
                const int P = g / (S * 256), sl = (g / 256) % S, row = (g & 255) >> 2, sg = g & 3;
                *reinterpret_cast<uint4*>(lds + P * PSTRIDE + (long)sl * 64 * LROW + lds_off(row, sg)) =
                    reg[u];
            }
        }
    };

    v4f acc[3][2][2];
    v4d acc64[2][2];
#pragma unroll
    for (int s = 0; s < 2; ++s)
#pragma unroll
        for (int t = 0; t < 2; ++t) {
            acc64[s][t] = v4d{0.0, 0.0, 0.0, 0.0};
#pragma unroll
            for (int c = 0; c < 3; ++c) acc[c][s][t] = v4f{0.f, 0.f, 0.f, 0.f};
        }
    // This wave's operand panels and exponent rows.
    const char* pa = lds + (isH ? 2 : 0) * PSTRIDE;
    const char* pb = lds + (isH ? 3 : 1) * PSTRIDE;
    const int8_t* exA = isH ? Per + ((long)j * p.vr) * Qb : Pel + ((long)i * p.vl) * Qb;
    const int8_t* exB = isH ? Pel + ((long)i * p.vl) * Qb : Per + ((long)j * p.vr) * Qb;
    const int limA = isH ? p.vr : p.vl, limB = isH ? p.vl : p.vr;

    // Register prefetch of the next chunk overlaps the loads with the MFMAs.
    constexpr bool PREFETCH = true;
    const int nchunk = Qp / KC;
    if (PREFETCH) fetch(0);
    for (int ch = 0; ch < nchunk; ++ch) {
        __syncthreads();
        if (!PREFETCH) fetch(ch * KC);
        stash();
        __syncthreads();
        if (PREFETCH && ch + 1 < nchunk) fetch((ch + 1) * KC);
        chunk_mfma<S>(pa, pb, (long)64 * LROW, wm, wn, lane, acc);
        if ((ch + 1) % (QB / KC) == 0) {
            const int qb = ch / (QB / KC);
            int eA[2][4], eB[2];
#pragma unroll
            for (int s = 0; s < 2; ++s)
#pragma unroll
                for (int r = 0; r < 4; ++r) {
                    const int a = A0 + wm + 16 * s + bf16_row(lane, r);
                    eA[s][r] = a < limA ? exA[(long)a * Qb + qb] : 0;
                }
#pragma unroll
            for (int t = 0; t < 2; ++t) {
                const int b = B0 + wn + 16 * t + (lane & 15);
                eB[t] = b < limB ? exB[(long)b * Qb + qb] : 0;
            }
            flush(acc, acc64, eA, eB);
        }
    }

    v4d accH[2][2];
    if (SYM) {
        // H waves hand their tile to the G waves through LDS.
        __syncthreads();
        double* Hs = reinterpret_cast<double*>(lds);  // [64][65]
        if (isH) {
#pragma unroll
            for (int s = 0; s < 2; ++s)
#pragma unroll
                for (int t = 0; t < 2; ++t)
#pragma unroll
                    for (int r = 0; r < 4; ++r)
                        Hs[(wm + 16 * s + bf16_row(lane, r)) * 65 + wn + 16 * t + (lane & 15)] =
                            acc64[s][t][r];
        }
        __syncthreads();
        if (!isH) {
#pragma unroll
            for (int s = 0; s < 2; ++s)
#pragma unroll
                for (int t = 0; t < 2; ++t)
#pragma unroll
                    for (int r = 0; r < 4; ++r)
                        accH[s][t][r] =
                            Hs[(wm + 16 * s + bf16_row(lane, r)) * 65 + wn + 16 * t + (lane & 15)];
        }
    } else {
#pragma unroll
        for (int s = 0; s < 2; ++s)
#pragma unroll
            for (int t = 0; t < 2; ++t) accH[s][t] = v4d{0.0, 0.0, 0.0, 0.0};
    }
    pair_epilogue<MODE, 1>(p, i, j, w, A0, B0, mirror, lane, wq, wm, wn, acc64, accH, red_row,
                           red_col, red_e, red_s, work, !isH);
}

// ----------------------------------------------------------- GEMM (NT)

struct SlicedMat {
    const uint16_t* s;  // [S] planes, plane stride `plane`
    long plane, ld;     // elements
    const int8_t* e;    // exponents [row][kb], row stride `lde`
    long lde;
    int rows;
};

// Wait for this wave's outstanding vector-memory operations (the LDS DMA
// loads), then a workgroup barrier.
__device__ __forceinline__ void dma_barrier() {
    __builtin_amdgcn_s_waitcnt(0x0F70);  // vmcnt(0)
    __syncthreads();
}

// Direct global -> LDS load of 16 BF16 panel rows x 64 bytes by one wave.
// The instruction writes lane l's 16 bytes at dst + 16*l, i.e. panel row
// l/4 of the 16, physical segment l&3; lane_seg() is the logical segment
// that lds_off() places there.  `dst` must be wave-uniform.
__device__ __forceinline__ int lane_seg(int lane) { return (lane & 3) ^ ((lane >> 3) & 3); }

__device__ __forceinline__ void dma16(const uint16_t* src, char* dst) {
    __builtin_amdgcn_global_load_lds((const void*)src, (__attribute__((address_space(3))) void*)dst, 16, 0, 0);
}

// C[m][n] += sum_k A[m][k] X[n][k] for the gradient contractions; K is a
// multiple of GB.  128 x 128 workgroup tile, four waves of 64 x 64, one
// 32-wide k step per stage; stages are loaded straight into a double-buffered
// LDS by the DMA path (one barrier per stage, no staging registers).  The
// gradient needs ~1e-7 relative accuracy, so products go to two FP32 tiers:
// c = 0 (exact integers) and c >= 1 (rounding near 2^-31 of the block
// scale); both are flushed to FP64 with the block exponents once per GB-wide
// k block.
template <int S>
__global__ __launch_bounds__(256) void gemm_bf16_nt_kernel(SlicedMat A, SlicedMat X, int K, double* __restrict__ C,
                                                           long ldc) {
    constexpr long PANEL = (long)GT * LROW;   // bytes of one slice panel (128 rows x 32 k)
    constexpr long STAGE = 2L * S * PANEL;    // both operands, all slices
    constexpr int PER = 2 * S * (GT / 16) / 4; // 16-row DMA loads per wave and stage
    extern __shared__ __align__(16) char lds[];
    int mt, ntile;
    if (!grouped_tile((A.rows + GT - 1) / GT, (X.rows + GT - 1) / GT, mt, ntile)) return;
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int wm = (wave >> 1) * 64, wn = (wave & 1) * 64;
    const int m0 = mt * GT, n0 = ntile * GT;
    // DMA load u of this wave covers operand P, slice sl and panel rows
    // 16*rb .. 16*rb+15 (id = wave + 4u enumerates (P, sl, rb)); lanes of rows
    // past the matrix load nothing, and those LDS rows only feed discarded
    // outputs.
    const int wv = __builtin_amdgcn_readfirstlane(wave);
    const int lrow = lane >> 2;
    const long laneA = (long)lrow * A.ld + 8 * lane_seg(lane);
    const long laneX = (long)lrow * X.ld + 8 * lane_seg(lane);
    auto issue = [&](int k0, int buf) {
#pragma unroll
        for (int u = 0; u < PER; ++u) {
            const int id = wv + 4 * u;
            const int P = id / (8 * S), sl = (id / 8) % S, rb = id % 8;
            const SlicedMat& M = P ? X : A;
            const int row = (P ? n0 : m0) + 16 * rb;
            if (row + lrow < M.rows)
                dma16(M.s + sl * M.plane + (long)row * M.ld + k0 + (P ? laneX : laneA),
                      lds + buf * STAGE + (P * S + sl) * PANEL + 1024 * rb);
        }
    };
    v4f acc[2][4][4];
    v4d acc64[4][4];
#pragma unroll
    for (int s = 0; s < 4; ++s)
#pragma unroll
        for (int t = 0; t < 4; ++t) {
            acc64[s][t] = v4d{0.0, 0.0, 0.0, 0.0};
            acc[0][s][t] = acc[1][s][t] = v4f{0.f, 0.f, 0.f, 0.f};
        }
    const int rr = lane & 15, fs = lane >> 4;
    const int nkb = K / GB;
    constexpr int CPB = GB / KC;  // chunks per exponent block (even: buffer = chunk parity)
    issue(0, 0);
    dma_barrier();
    for (int kb = 0; kb < nkb; ++kb) {
        // Exponents of this block, fetched early; used by the flush.
        int eA[4][4], eB[4];
#pragma unroll
        for (int t = 0; t < 4; ++t) {
            const int n = n0 + wn + 16 * t + rr;
            eB[t] = n < X.rows ? X.e[(long)n * X.lde + kb] : 0;
        }
#pragma unroll
        for (int s = 0; s < 4; ++s)
#pragma unroll
            for (int r = 0; r < 4; ++r) {
                const int m = m0 + wm + 16 * s + 4 * fs + r;
                eA[s][r] = m < A.rows ? A.e[(long)m * A.lde + kb] : 0;
            }
#pragma unroll
        for (int c = 0; c < CPB; ++c) {
            const int buf = c & 1;
            if (kb * CPB + c + 1 < nkb * CPB) issue((kb * CPB + c + 1) * KC, buf ^ 1);
            const char* pa = lds + buf * STAGE;
            const char* pb = pa + S * PANEL;
            v8s b[S][4];
#pragma unroll
            for (int q = 0; q < S; ++q)
#pragma unroll
                for (int t = 0; t < 4; ++t) b[q][t] = lds_frag(pb + q * PANEL, wn + 16 * t + rr, fs);
#pragma unroll
            for (int p = 0; p < S; ++p) {
                v8s a[4];
#pragma unroll
                for (int s = 0; s < 4; ++s) a[s] = lds_frag(pa + p * PANEL, wm + 16 * s + rr, fs);
#pragma unroll
                for (int q = 0; q < S - p; ++q)
#pragma unroll
                    for (int s = 0; s < 4; ++s)
#pragma unroll
                        for (int t = 0; t < 4; ++t)
                            acc[p + q ? 1 : 0][s][t] = mfma_bf16(a[s], b[q][t], acc[p + q ? 1 : 0][s][t]);
            }
            dma_barrier();
        }
#pragma unroll
        for (int s = 0; s < 4; ++s)
#pragma unroll
            for (int t = 0; t < 4; ++t) {
#pragma unroll
                for (int r = 0; r < 4; ++r) {
                    const double sum = (double)acc[0][s][t][r] + (double)acc[1][s][t][r];
                    acc64[s][t][r] += sum * pow2(eA[s][r] + eB[t] - 16);
                }
                acc[0][s][t] = acc[1][s][t] = v4f{0.f, 0.f, 0.f, 0.f};
            }
    }
#pragma unroll
    for (int s = 0; s < 4; ++s)
#pragma unroll
        for (int r = 0; r < 4; ++r) {
            const int m = m0 + wm + 16 * s + 4 * fs + r;
            if (m >= A.rows) continue;
#pragma unroll
            for (int t = 0; t < 4; ++t) {
                const int n = n0 + wn + 16 * t + rr;
                if (n < X.rows) C[(long)m * ldc + n] += acc64[s][t][r];
            }
        }
}

// ------------------------------------------------------------ dispatch

template <template <int> class F, typename... Args>
int by_slices(int S, Args&&... args) {
    switch (S) {
        case 2: return F<2>::run(args...);
        case 3: return F<3>::run(args...);
        case 4: return F<4>::run(args...);
        case 5: return F<5>::run(args...);
        case 6: return F<6>::run(args...);
        default: return -1;
    }
}

// Gradient-GEMM operands carry at most BF16_GRAD_SLICES slices.
template <template <int> class F, typename... Args>
int by_gslices(int S, Args&&... args) {
    static_assert(BF16_GRAD_SLICES == 3, "instantiate the gradient slice counts");
    switch (S) {
        case 2: return F<2>::run(args...);
        case 3: return F<3>::run(args...);
        default: return -1;
    }
}

int round_up(int x, int m) { return (x + m - 1) / m * m; }

// Dynamic LDS of the gradient GEMM: two stages of both operands' slices.
constexpr size_t gemm_lds(int S) { return (size_t)2 * 2 * S * GT * LROW; }

template <int S>
struct SlicePair {
    static int run(mshqc_hip_ctx* ctx, const double* B, int Q, int ov, int Qp, int Qb, uint16_t* Ps,
                   int8_t* Pe) {
        dim3 grid(ceil_div(ov, 64), Qb);
        slice_pair_kernel<S><<<grid, 256, 0, ctx->stream>>>(B, Q, ov, Qp, Qb, Ps, Pe);
        LAUNCH_CHECK(ctx);
        return 0;
    }
};

template <int S>
struct SliceRows {
    static int run(mshqc_hip_ctx* ctx, const double* src, long ld, int R, int nseg, int v, int vp,
                   uint16_t* dst, int8_t* ex) {
        const long jobs = (long)R * nseg * (vp / GB);
        slice_rows_kernel<S><<<ceil_div(jobs, 4), 256, 0, ctx->stream>>>(src, ld, R, nseg, v, vp, dst, ex);
        LAUNCH_CHECK(ctx);
        return 0;
    }
};

template <int S>
struct SliceCols {
    static int run(mshqc_hip_ctx* ctx, const double* Z, long ldz, int nI, int v, int vp, int nc,
                   uint16_t* dst, int8_t* ex) {
        dim3 grid(ceil_div(nc, 32), nI * (vp / GB));
        slice_cols_kernel<S><<<grid, 256, 0, ctx->stream>>>(Z, ldz, nI, v, vp, nc, dst, ex);
        LAUNCH_CHECK(ctx);
        return 0;
    }
};

template <int S>
struct PairLaunch {
    static int run(mshqc_hip_ctx* ctx, int mode, const PairArgs& a, const Bf16Operands& o) {
        const long nwg = ((a.nwork + 7) / 8) * 8;
        dim3 grid((unsigned)nwg);
        // The symmetric kernel reuses its panels for the 64 x 65 FP64 H tile.
        const size_t lds_sym = std::max((size_t)4 * S * 64 * LROW, (size_t)64 * 65 * 8);
        const size_t lds_os = (size_t)2 * S * 64 * LROW;
        const uint16_t* sr = o.sr ? o.sr : o.sl;
        const int8_t* xr = o.sr ? o.xr : o.xl;
        switch (mode) {
            case MODE_RMP2:
                pair_bf16_kernel<MODE_RMP2, S><<<grid, 512, lds_sym, ctx->stream>>>(a, o.sl, o.xl, sr, xr, o.Qp, o.qblocks);
                break;
            case MODE_SS:
                pair_bf16_kernel<MODE_SS, S><<<grid, 512, lds_sym, ctx->stream>>>(a, o.sl, o.xl, sr, xr, o.Qp, o.qblocks);
                break;
            case MODE_OS:
                pair_bf16_kernel<MODE_OS, S><<<grid, 256, lds_os, ctx->stream>>>(a, o.sl, o.xl, sr, xr, o.Qp, o.qblocks);
                break;
            default:
                return set_error(ctx, MSHQC_HIP_ERR_ARG, "split-BF16 pair kernel: unsupported mode %d", mode);
        }
        LAUNCH_CHECK(ctx);
        return 0;
    }
};

template <int S>
struct GemmLaunch {
    static int run(mshqc_hip_ctx* ctx, const SlicedMat& A, const SlicedMat& X, int K, double* C, long ldc) {
        if (K % GB != 0)
            return set_error(ctx, MSHQC_HIP_ERR_ARG, "split-BF16 GEMM: K=%d is not a multiple of %d", K, GB);
        dim3 grid(grouped_grid(ceil_div(A.rows, GT), ceil_div(X.rows, GT)));
        gemm_bf16_nt_kernel<S><<<grid, 256, gemm_lds(S), ctx->stream>>>(A, X, K, C, ldc);
        LAUNCH_CHECK(ctx);
        return 0;
    }
};

template <int S>
struct AllowLds {
    static int run(mshqc_hip_ctx* ctx) {
        // Dynamic LDS above 64 KiB must be opted into per kernel.
        const int big = std::max(4 * S * 64 * LROW, 64 * 65 * 8);
        hipFuncSetAttribute((const void*)pair_bf16_kernel<MODE_RMP2, S>,
                            hipFuncAttributeMaxDynamicSharedMemorySize, big);
        hipFuncSetAttribute((const void*)pair_bf16_kernel<MODE_SS, S>,
                            hipFuncAttributeMaxDynamicSharedMemorySize, big);
        hipFuncSetAttribute((const void*)pair_bf16_kernel<MODE_OS, S>,
                            hipFuncAttributeMaxDynamicSharedMemorySize, big / 2);
        hipGetLastError();
        (void)ctx;
        return 0;
    }
};

template <int S>
struct AllowLdsGemm {
    static int run(mshqc_hip_ctx* ctx) {
        hipFuncSetAttribute((const void*)gemm_bf16_nt_kernel<S>,
                            hipFuncAttributeMaxDynamicSharedMemorySize, (int)gemm_lds(S));
        hipGetLastError();
        (void)ctx;
        return 0;
    }
};

}  // namespace

int bf16_prepare(mshqc_hip_ctx* ctx, int slices, const double* Bl, int ol, int vl, const double* Br,
                 int orr, int vr, int Q, bool grad, Bf16Operands* out) {
    const int S = slices;
    by_slices<AllowLds>(S, ctx);
    Bf16Operands o;
    o.slices = S;
    o.gslices = std::min(S, BF16_GRAD_SLICES);
    by_gslices<AllowLdsGemm>(o.gslices, ctx);
    o.Q = Q;
    o.Qp = round_up(Q, QB);
    o.qblocks = o.Qp / QB;
    const long ovl = (long)ol * vl, ovr = Br ? (long)orr * vr : 0;
    // Pair-kernel slices.
    const size_t bl = (size_t)S * ovl * o.Qp * 2, br = (size_t)S * ovr * o.Qp * 2;
    const size_t el = (size_t)ovl * o.qblocks, er = (size_t)ovr * o.qblocks;
    char* w = static_cast<char*>(workspace(ctx, WS_SLICE, bl + br + el + er + 1024));
    if (!w) return MSHQC_HIP_ERR_ALLOC;
    o.sl = reinterpret_cast<uint16_t*>(w);
    o.sr = Br ? reinterpret_cast<uint16_t*>(w + bl) : nullptr;
    o.xl = reinterpret_cast<int8_t*>(w + bl + br);
    o.xr = Br ? reinterpret_cast<int8_t*>(w + bl + br + el) : nullptr;
    int rc = by_slices<SlicePair>(S, ctx, Bl, Q, (int)ovl, o.Qp, o.qblocks, const_cast<uint16_t*>(o.sl),
                                  const_cast<int8_t*>(o.xl));
    if (rc) return rc;
    if (Br) {
        rc = by_slices<SlicePair>(S, ctx, Br, Q, (int)ovr, o.Qp, o.qblocks, const_cast<uint16_t*>(o.sr),
                                  const_cast<int8_t*>(o.xr));
        if (rc) return rc;
    }
    if (grad) {
        // GEMM-operand slices of B: rows q, columns (i, a) blocked by 64 along a.
        o.vlp = round_up(vl, GB);
        o.vrp = round_up(vr, GB);
        const int SG = o.gslices;
        const size_t gl = (size_t)SG * Q * ol * o.vlp * 2, gr = Br ? (size_t)SG * Q * orr * o.vrp * 2 : 0;
        const size_t gel = (size_t)Q * ol * (o.vlp / GB), ger = Br ? (size_t)Q * orr * (o.vrp / GB) : 0;
        char* g = static_cast<char*>(workspace(ctx, WS_SLICE2, gl + gr + gel + ger + 1024));
        if (!g) return MSHQC_HIP_ERR_ALLOC;
        o.gl = reinterpret_cast<uint16_t*>(g);
        o.gr = Br ? reinterpret_cast<uint16_t*>(g + gl) : o.gl;
        o.gel = reinterpret_cast<int8_t*>(g + gl + gr);
        o.ger = Br ? reinterpret_cast<int8_t*>(g + gl + gr + gel) : o.gel;
        rc = by_gslices<SliceRows>(SG, ctx, Bl, (long)ovl, Q, ol, vl, o.vlp, const_cast<uint16_t*>(o.gl),
                                  const_cast<int8_t*>(o.gel));
        if (rc) return rc;
        if (Br) {
            rc = by_gslices<SliceRows>(SG, ctx, Br, (long)ovr, Q, orr, vr, o.vrp, const_cast<uint16_t*>(o.gr),
                                      const_cast<int8_t*>(o.ger));
            if (rc) return rc;
        }
    }
    o.ol = ol;
    o.vl = vl;
    o.orr = Br ? orr : ol;
    o.vr = Br ? vr : vl;
    *out = o;
    return MSHQC_HIP_OK;
}

int bf16_launch_pair(mshqc_hip_ctx* ctx, int mode, const PairArgs& a, const Bf16Operands& ops) {
    if (a.nwork <= 0) return MSHQC_HIP_OK;
    return by_slices<PairLaunch>(ops.slices, ctx, mode, a, ops);
}

int bf16_grad_contract(mshqc_hip_ctx* ctx, const Bf16Operands& o, const double* Z, long ldz, int I0,
                       int nI, int J0, int nJ, double* dBl, double* dBr, bool second) {
    const int S = o.gslices, Q = o.Q;
    const int vl = o.vl, vr = o.vr, vlp = o.vlp, vrp = o.vrp;
    int rc;
    if (dBl) {
        // dBl[q][(i,a)] += sum_{(j,b)} Br[q][(j,b)] Z[(i,a)][(j,b)]
        const long R = (long)nI * vl, kw = (long)nJ * vrp;
        const size_t zb = (size_t)S * R * kw * 2, ze = (size_t)R * nJ * (vrp / GB);
        char* w = static_cast<char*>(workspace(ctx, WS_ZS, zb + ze + 256));
        if (!w) return MSHQC_HIP_ERR_ALLOC;
        uint16_t* zs = reinterpret_cast<uint16_t*>(w);
        int8_t* zx = reinterpret_cast<int8_t*>(w + zb);
        rc = by_gslices<SliceRows>(S, ctx, Z, ldz, (int)R, nJ, vr, vrp, zs, zx);
        if (rc) return rc;
        SlicedMat A{o.gr + (long)J0 * vrp, (long)Q * o.orr * vrp, (long)o.orr * vrp,
                    o.ger + (long)J0 * (vrp / GB), (long)o.orr * (vrp / GB), Q};
        SlicedMat X{zs, (long)R * kw, kw, zx, (long)nJ * (vrp / GB), (int)R};
        rc = by_gslices<GemmLaunch>(S, ctx, A, X, (int)kw, dBl + (long)I0 * vl, (long)o.ol * vl);
        if (rc) return rc;
    }
    if (second && dBr) {
        // dBr[q][(j,b)] += sum_{(i,a)} Bl[q][(i,a)] Z[(i,a)][(j,b)]
        const long R = (long)nJ * vr, kw = (long)nI * vlp;
        const size_t zb = (size_t)S * R * kw * 2, ze = (size_t)R * nI * (vlp / GB);
        char* w = static_cast<char*>(workspace(ctx, WS_ZTS, zb + ze + 256));
        if (!w) return MSHQC_HIP_ERR_ALLOC;
        uint16_t* zs = reinterpret_cast<uint16_t*>(w);
        int8_t* zx = reinterpret_cast<int8_t*>(w + zb);
        rc = by_gslices<SliceCols>(S, ctx, Z, ldz, nI, vl, vlp, (int)R, zs, zx);
        if (rc) return rc;
        SlicedMat A{o.gl + (long)I0 * vlp, (long)Q * o.ol * vlp, (long)o.ol * vlp,
                    o.gel + (long)I0 * (vlp / GB), (long)o.ol * (vlp / GB), Q};
        SlicedMat X{zs, (long)R * kw, kw, zx, (long)nI * (vlp / GB), (int)R};
        rc = by_gslices<GemmLaunch>(S, ctx, A, X, (int)kw, dBr + (long)J0 * vr, (long)o.orr * vr);
        if (rc) return rc;
    }
    return MSHQC_HIP_OK;
}
