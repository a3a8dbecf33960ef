// Split-BF16 form of the MP2 kernels on gfx950 BF16 matrix cores.
//
//   slice_pair_kernel   B[Q][o*v] -> BF16 slices [Qp/16][ceil(o*v/32) + 1][S][32][16] (each
//                       16-wide Q step of 32 rows is one contiguous KiB) and
//                       one exponent per (i, a) and 128-wide Q block
//   pair_bf16_kernel    (ia|jb) tiles from the slices on 32x32x16 MFMAs, exact
//                       FP32 class sums per Q block, FP64 across blocks, then
//                       the FP64 denominator epilogue
//   slice_rows_kernel / slice_cols_kernel
//                       row- and column-blocked slices of the gradient GEMM
//                       operands (B and the amplitude block Z, Z^T), one
//                       exponent per row and 256-wide k block
//   gemm_bf16_nt_kernel C[m][n] += sum_k A[m][k] X[n][k] from slices
// The arithmetic contract is described in ozaki.h.
#include <algorithm>
#include <type_traits>

#include "dma.h"
#include "mp2_bf16.h"
#include "mp2_internal.h"
#include "ozaki.h"
#include "tiles.h"

namespace {

constexpr int KS = 16;    // Q per pair-kernel stage (one 32x32x16 MFMA K step)
// Q block of the (ia|jb) kernel (exponent block and exact-FP32 span): SPB
// stages of KS, 8 or 10, whichever pads Q to fewer stages (Q = 1600, 2400,
// 4000, 5600 are whole 160-blocks; 128-blocks left 2-6 stages of padding).
// Tier-0 class sums stay below 160 * 2^16 < 2^24; at 256 Q (sums up to
// 2^24) the energy and gradient errors measured on gfx950 grew ~3x and ~12x,
// so the MFMA accumulation is not exact in that last bit.
// Five slices keep 8: their 10-stage symmetric kernels do not fit the
// register file without spilling.
constexpr int SPB_SHORT = 8, SPB_LONG = 10;
constexpr bool spb_long_ok(int S) { return S != 5; }
int pair_stages_per_block(int Q, int S) {
    const int st = (Q + KS - 1) / KS;
    auto padded = [st](int spb) { return (st + spb - 1) / spb * spb; };
    return spb_long_ok(S) && padded(SPB_LONG) < padded(SPB_SHORT) ? SPB_LONG : SPB_SHORT;
}
constexpr int RB = 32;    // LDS bytes per panel row, slice and pair-kernel stage
constexpr int KC = 32;    // contraction chunk of the gradient GEMM (one 16x16x32 MFMA K step)
constexpr int LROW = 64;  // LDS bytes per 32-element BF16 row
// Exponent block widths of the BF16 GEMM operands along k (the GEMM flushes
// its FP32 class sums to FP64 once per block).  The gradient contractions
// use 512: their tier-0 sums are sums of products of unrelated operands
// (B against amplitudes), random in sign, so they stay far below 2^24, where
// an MFMA accumulation can lose a last bit; the FP64 flushes halve against
// 256-wide blocks (4% of the GEMM time, gradient error 3.5e-8 -> 4.8e-8 at
// o=100/v=1100/Q=4000).  The transforms keep 256, which pads their AO index
// less (n = 1200: 1280 instead of 1536).
constexpr int GB_GRAD = 512;
constexpr int GB_TRANSFORM = 256;
constexpr int GT = 128;   // gradient-GEMM workgroup tile (rows and columns)
constexpr int GW = 8;     // gradient-GEMM waves per workgroup (two per SIMD, 64 x 32 each)

// Slots of the pair kernel's LDS ring of slices (12 KiB per slice in the
// symmetric modes, 8 KiB in the opposite-spin mode) within 160 KiB: two whole
// stages plus as many slices of a third as fit (all S where possible).
constexpr int ring_slots(int S, bool sym) {
    const int fit = (160 * 1024) / ((sym ? 12 : 8) * 1024);
    const int third = fit - 2 * S < S ? fit - 2 * S : S;
    return 2 * S + third;
}

// LDS byte offset of 16-byte segment `sg` of panel row `row`: the segment is
// XOR-swizzled by (row >> 1) & 3, which makes the 16-lane fragment reads and
// the 4-lanes-per-row stores free of bank conflicts (measured on gfx950).
__device__ __forceinline__ int lds_off(int row, int sg) { return row * LROW + ((sg ^ ((row >> 1) & 3)) << 4); }

// LDS byte offset of 16-byte segment `seg` (0, 1) of a 32-byte pair-kernel
// panel row: the segment is swapped on bit 2 xor bit 3 of the row, which
// keeps the 32-lane fragment reads free of bank conflicts for 32 or 64 banks.
__device__ __forceinline__ int lds_off32(int row, int seg) {
    return row * RB + ((seg ^ (((row >> 2) ^ (row >> 3)) & 1)) << 4);
}

inline int ceil_div(long a, long b) { return (int)((a + b - 1) / b); }

// ------------------------------------------------------------- slicing

// B[Q][o*v] (row-major, rows = q) -> Ps[Qp/16][ceil(o*v/32) + 1][S][32][16] and
// Pe[Qp/QB][o][vE] (exponent of row (i, a) over one QB-wide Q block; the
// caller zeroes Pe, so the padding a >= v reads 0).  One workgroup per 32
// rows x QB q; the tile is transposed through LDS (dynamic, QB x 33 doubles)
// so both the FP64 reads and the BF16 writes are contiguous.  Rows past o*v,
// including the whole extra 32-row block, hold zero slices: the pair kernel
// feeds that block to the MFMAs of padded virtual indices, whose outputs it
// discards.  Zero operands draw less power than the next orbital's rows did,
// and the pair kernel runs power-limited (1.8-1.9 GHz).
constexpr size_t slice_pair_lds(int QB) { return (size_t)QB * 33 * sizeof(double); }

template <int S, int QB>
__global__ __launch_bounds__(256) void slice_pair_kernel(const double* __restrict__ B, int Q, int ov, int v,
                                                         int vE, int o, int Qp, uint16_t* __restrict__ Ps,
                                                         int8_t* __restrict__ Pe) {
    extern __shared__ __align__(16) char lds_raw[];
    double (*T)[33] = reinterpret_cast<double (*)[33]>(lds_raw);
    __shared__ int ex[32];
    const int r0 = blockIdx.x * 32, q0 = blockIdx.y * QB, tid = threadIdx.x;
    {
        const int c = tid & 31;
        for (int k = 0; k < QB / 8; ++k) {
            const int qq = (tid >> 5) + 8 * k;
            double x = 0.0;
            if (r0 + c < ov && q0 + qq < Q) x = B[(long)(q0 + qq) * ov + r0 + c];
            T[qq][c] = x;
        }
    }
    __syncthreads();
    {
        // 8 lanes per row, QB/8 values each, then a shuffle maximum.
        const int row = tid >> 3, part = tid & 7;
        double mx = 0.0;
        for (int k = 0; k < QB / 8; ++k) mx = fmax(mx, fabs(T[part * (QB / 8) + k][row]));
        for (int m = 1; m < 8; m <<= 1) mx = fmax(mx, __shfl_xor(mx, m, 64));
        if (part == 0) {
            const int e = block_exponent(mx);
            ex[row] = e;
            const int r = r0 + row;
            if (r < ov) Pe[((long)blockIdx.y * o + r / v) * vE + r % v] = exp_byte(e);
        }
    }
    __syncthreads();
    // Item (q16 block kq, row, half): 8 consecutive q of one row; consecutive
    // threads write consecutive 16-byte pieces of the 1 KiB (32 rows x 32 B)
    // block of one 16-wide Q step.
    for (int idx = tid; idx < (QB / KS) * 32 * 2; idx += 256) {
        const int kq = idx >> 6, row = (idx & 63) >> 1, half = idx & 1;
        const int e = ex[row];
        uint16_t out[S][8];
#pragma unroll
        for (int u = 0; u < 8; ++u) {
            uint16_t m[S];
            slice_value<S>(T[kq * KS + half * 8 + u][row], e, m);
#pragma unroll
            for (int p = 0; p < S; ++p) out[p][u] = m[p];
        }
        uint16_t* dst = Ps + (((long)(q0 / KS + kq) * gridDim.x + blockIdx.x) * S * 32 + row) * KS + half * 8;
#pragma unroll
        for (int p = 0; p < S; ++p)
            *reinterpret_cast<uint4*>(dst + p * 32 * KS) = *reinterpret_cast<uint4*>(out[p]);
    }
}

// Self terms out[i * nv + a] = sum_q Bl[q][i][a] Br[q][i][a] in FP64 (q in
// ascending order, fused multiply-adds), i < no, a < nv.  One thread per
// (i, a): consecutive threads read consecutive addresses of each q plane;
// eight independent partial chains keep several loads in flight and are
// added in a fixed order.
__global__ __launch_bounds__(256) void self_terms_kernel(const double* __restrict__ Bl, const double* __restrict__ Br,
                                                         int Q, long ovl, long ovr, int vl, int vr, int no, int nv,
                                                         double* __restrict__ out) {
    const long idx = (long)blockIdx.x * 256 + threadIdx.x;
    if (idx >= (long)no * nv) return;
    const int i = (int)(idx / nv), a = (int)(idx % nv);
    const double* pl = Bl + (long)i * vl + a;
    const double* pr = Br + (long)i * vr + a;
    double acc[8] = {0, 0, 0, 0, 0, 0, 0, 0};
    int q = 0;
    for (; q + 8 <= Q; q += 8)
#pragma unroll
        for (int u = 0; u < 8; ++u) acc[u] = fma(pl[(long)(q + u) * ovl], pr[(long)(q + u) * ovr], acc[u]);
    for (int u = 0; q < Q; ++q, ++u) acc[u] = fma(pl[(long)q * ovl], pr[(long)q * ovr], acc[u]);
    out[idx] = ((acc[0] + acc[1]) + (acc[2] + acc[3])) + ((acc[4] + acc[5]) + (acc[6] + acc[7]));
}

// src[R][ld] with column index seg*v + b (seg < nseg, b < v)  ->
// dst[S][R][nseg*vp] (column seg*vp + b, zero padded), exponents
// ex[nseg*(vp/GB)][ldx] (k-block major, ldx >= R rows: four consecutive rows
// are one dword).  One wave per (row, seg, GB-wide b tile); each lane slices
// GB/64 elements.
template <int S, int GB>
__global__ __launch_bounds__(256) void slice_rows_kernel(const double* __restrict__ src, long ld, int R,
                                                         int nseg, int v, int vp,
                                                         uint16_t* __restrict__ dst,
                                                         int8_t* __restrict__ ex, long ldx) {
    const int ntv = vp / GB;
    const long job = (long)blockIdx.x * 4 + (threadIdx.x >> 6);
    const long njobs = (long)R * nseg * ntv;
    if (job >= njobs) return;
    const int lane = threadIdx.x & 63;
    const int row = (int)(job / (nseg * ntv));
    const int rem = (int)(job % (nseg * ntv));
    const int seg = rem / ntv, tile = rem % ntv;
    constexpr int NE = GB / 64;
    const double* s = src + (long)row * ld + (long)seg * v;
    double x[NE];
    double mx = 0.0;
#pragma unroll
    for (int u = 0; u < NE; ++u) {
        const int b = tile * GB + 64 * u + lane;
        x[u] = b < v ? s[b] : 0.0;
        mx = fmax(mx, fabs(x[u]));
    }
    for (int m = 1; m < 64; m <<= 1) mx = fmax(mx, __shfl_xor(mx, m, 64));
    const int e = block_exponent(mx);
    const long kw = (long)nseg * vp;
    const long slice = (long)R * kw;
    uint16_t* d = dst + (long)row * kw + (long)seg * vp + tile * GB + lane;
#pragma unroll
    for (int u = 0; u < NE; ++u) {
        uint16_t sv[S];
        slice_value<S>(x[u], e, sv);
#pragma unroll
        for (int p = 0; p < S; ++p) d[p * slice + 64 * u] = sv[p];
    }
    if (lane == 0) ex[(long)rem * ldx + row] = exp_byte(e);
}

// Transposed slicing of Z[(i',a)][(j',b)] (rows i'*v + a, i' < nI; ld = ldz;
// nc = nJ*vr columns):  dst[S][nc][nI*vp] with row n = column of Z and
// column i'*vp + a;  exponents ex[nI*(vp/GB)][ldx] (k-block major, ldx >= nc).
// A workgroup transposes a tile of sc_cols(GB) columns x GB values of a through
// LDS.
// (16 columns for 256-wide blocks, 8 for 512: the LDS tile stays 37 KiB or less)
constexpr int sc_cols(int GB) { return GB >= 512 ? 8 : 16; }

template <int S, int GB>
__global__ __launch_bounds__(256) void slice_cols_kernel(const double* __restrict__ Z, long ldz, int nI,
                                                         int v, int vp, int nc,
                                                         uint16_t* __restrict__ dst,
                                                         int8_t* __restrict__ ex, long ldx) {
    constexpr int NC = sc_cols(GB), PARTS = 256 / NC, PER = GB / PARTS;  // PER values of a per thread
    static_assert(PER % 8 == 0, "whole 16-byte stores");
    __shared__ double T[GB][NC + 1];
    __shared__ int exs[NC];
    const int ntv = vp / GB;
    // Work item w = (a block, column tile), column tiles fastest.  The
    // dispatcher deals workgroups to the 8 XCDs in turn; remapped so that
    // each XCD takes a contiguous range of w, the workgroups that read the
    // two halves of a 128-byte line of Z (adjacent column tiles) run on one
    // XCD and share its L2.  The grid is a multiple of 8; items past the end
    // return (whole workgroups, before any barrier).
    const int nct = (nc + NC - 1) / NC;
    const long w = (long)(blockIdx.x & 7) * (gridDim.x >> 3) + (blockIdx.x >> 3);
    if (w >= (long)nct * nI * ntv) return;
    const int c0 = (int)(w % nct) * NC;
    const int rb = (int)(w / nct);
    const int ip = rb / ntv, at = rb % ntv, a0 = at * GB;
    const int tid = threadIdx.x;
    {
        const int cc = tid % NC;
        for (int k = 0; k < GB / PARTS; ++k) {
            const int aa = tid / NC + PARTS * k;
            double x = 0.0;
            if (a0 + aa < v && c0 + cc < nc) x = Z[((long)ip * v + a0 + aa) * ldz + c0 + cc];
            // columns swizzled by row group: the PARTS lanes reading one
            // column below (rows PER apart) then hit distinct banks
            T[aa][cc ^ ((aa / PER) & (NC - 1))] = x;
        }
    }
    __syncthreads();
    // thread -> (column cc, PER consecutive a)
    const int cc = tid / PARTS, part = tid % PARTS;
    {
        // PARTS lanes per column, then a shuffle maximum.
        double mx = 0.0;
        for (int k = 0; k < PER; ++k) mx = fmax(mx, fabs(T[part * PER + k][cc ^ (part & (NC - 1))]));
        for (int m = 1; m < PARTS; m <<= 1) mx = fmax(mx, __shfl_xor(mx, m, 64));
        if (part == 0) exs[cc] = block_exponent(mx);
    }
    __syncthreads();
    const long kw = (long)nI * vp, slice = (long)nc * kw;
    if (c0 + cc >= nc) return;
    const int e = exs[cc];
    uint16_t* d = dst + (long)(c0 + cc) * kw + (long)ip * vp + a0 + PER * part;
    for (int u8 = 0; u8 < PER; u8 += 8) {
        uint16_t out[S][8];
#pragma unroll
        for (int u = 0; u < 8; ++u) {
            uint16_t m[S];
            slice_value<S>(T[PER * part + u8 + u][cc ^ (part & (NC - 1))], e, m);
#pragma unroll
            for (int p = 0; p < S; ++p) out[p][u] = m[p];
        }
#pragma unroll
        for (int p = 0; p < S; ++p)
            *reinterpret_cast<uint4*>(d + p * slice + u8) = *reinterpret_cast<uint4*>(out[p]);
    }
    if (part == 0) ex[(long)(ip * ntv + at) * ldx + c0 + cc] = exp_byte(e);
}

// ------------------------------------------------ shared MFMA helpers

// 16-byte LDS fragment read: row `row` of a [rows][32 k] BF16 panel, the
// 8 elements k = 8*seg .. 8*seg + 7.
__device__ __forceinline__ v8s lds_frag(const char* panel, int row, int seg) {
    return *reinterpret_cast<const v8s*>(panel + lds_off(row, seg));
}


// Logical 16-byte segment a lane fetches in a 16-row x 64-byte DMA block
// (lane l lands at row l/4, physical segment l&3 of lds_off()).
__device__ __forceinline__ int lane_seg(int lane) { return (lane & 3) ^ ((lane >> 3) & 3); }

// ----------------------------------------------------------- pair kernel
//
// Work unit: one occupied pair (i, j) and one unit of 64 x 64 tiles.
//   Symmetric modes: unit (A, B2) = rows a in [64A, 64A+64) x columns b in
//   [128 B2, 128 B2 + 128), i.e. tiles (A, 2B2) and (A, 2B2+1); a tile (A, B)
//   counts only when A <= B (A < B with the mirrored element).  Waves 0, 1
//   form G = Bl_i^T Br_j on the two 64-column halves, waves 2, 3 the mirrored
//   H[a][b] = sum_Q Br^Q_ja Bl^Q_ib on the same halves; H is handed to the G
//   waves through LDS for the epilogue.
//   Symmetric modes, diagonal block k (encoded as unit (-1 - k, k)): rows and
//   columns [128k, 128k + 128), tiles (2k + x, 2k + y), one wave per tile, G
//   only; it replaces units (2k, k) and (2k + 1, k), whose tile (2k + 1, 2k)
//   lies below the diagonal and whose diagonal tiles repeat their own mirror.
//   The mirrors are all inside the block and go through LDS in the epilogue;
//   its stages hold 8 blocks (0-3 Bl_i, 4-7 Br_j), two per wave.
//   Opposite-spin mode: unit (A2, B2) = 128 x 128, tiles (2A2 + x, 2B2 + y),
//   one wave per tile, G only.
// Stage = 16 Q of every panel row and slice, streamed by DMA into LDS
// (double-buffered in the symmetric modes, triple-buffered in the opposite-
// spin mode).  Panel rows (32-row blocks): symmetric: 0-1 Bl_i[A],
// 2-5 Br_j[B], 6-7 Br_j[A], 8-11 Bl_i[B]; opposite spin: 0-3 Bl_i[A],
// 4-7 Br_j[B].  Each wave streams 3 (2) of the blocks for every slice.
//
// v_mfma_f32_32x32x16_bf16 lane layout (measured on gfx950 with one-hot
// operands): A lane l holds row l%32, k = 8(l/32)..+7; B lane l holds
// k = 8(l/32)..+7, column l%32; D lane l, element r holds row
// 8(r/4) + 4(l/32) + r%4, column l%32.

__device__ __forceinline__ v16f mfma32_bf16(const v8s& a, const v8s& b, v16f c) {
    return __builtin_amdgcn_mfma_f32_32x32x16_bf16(__builtin_bit_cast(v8bf, a), __builtin_bit_cast(v8bf, b), c, 0,
                                                  0, 0);
}

// Fused denominator epilogue for one 64 x 64 tile held by one wave in the
// 32x32 accumulator layout (2 x 2 subtiles): G (and H in the symmetric
// modes) at row a = a0 + 32s + 8u + 4h + c, column b = b0 + 32t + l%32,
// h = l/32.  Writes the weighted tile partials to `slot` and, when requested,
// the amplitudes z (gradient) or t2.
// The orbital energies and the output come in as restrict parameters: the
// amplitude stores then do not order the later energy loads behind them.
template <int MODE>
__device__ __forceinline__ void pair_epilogue32_r(const PairArgs& p, const double* __restrict__ evl,
                                                  const double* __restrict__ evr, double* __restrict__ Z,
                                                  int i, int j, double w, int a0, int b0, bool mirror, int lane,
                                                  const v4d (&G)[2][2][4], const v4d (&H)[2][2][4], long slot) {
#pragma clang fp contract(off)
    const bool grad = p.s_part != nullptr;
    const int h = lane >> 5, col = lane & 31;
    const double ei = p.eol[i] + p.eor[j];
    double e = 0.0, sm = 0.0;
    double rowq[2][4][4], colq[2];
#pragma unroll
    for (int t = 0; t < 2; ++t) colq[t] = 0.0;
#pragma unroll
    for (int s = 0; s < 2; ++s)
#pragma unroll
        for (int u = 0; u < 4; ++u)
#pragma unroll
            for (int c = 0; c < 4; ++c) rowq[s][u][c] = 0.0;
    double* t2base = nullptr;
    if (MODE == MODE_T2) t2base = Z + ((long)i * p.orr + j) * p.vl * p.vr;
    const long zr = (long)(i - p.zi0) * p.vl, zc = (long)(j - p.zj0) * p.vr;
    const long zr2 = (long)(j - p.zi0) * p.vl, zc2 = (long)(i - p.zj0) * p.vr;
    // The lane's virtual energies, loaded up front with clamped indices so
    // the loads issue together instead of one masked load-and-wait each.
    double evbv[2], eva[2][4][4];
#pragma unroll
    for (int t = 0; t < 2; ++t) evbv[t] = evr[min(b0 + 32 * t + col, p.vr - 1)];
#pragma unroll
    for (int s = 0; s < 2; ++s)
#pragma unroll
        for (int u = 0; u < 4; ++u)
#pragma unroll
            for (int c = 0; c < 4; ++c) eva[s][u][c] = evl[min(a0 + 32 * s + 8 * u + 4 * h + c, p.vl - 1)];
#pragma unroll
    for (int t = 0; t < 2; ++t) {
        const int b = b0 + 32 * t + col;
        const bool bin = b < p.vr;
        const double evb = evbv[t];
#pragma unroll
        for (int s = 0; s < 2; ++s)
#pragma unroll
            for (int u = 0; u < 4; ++u) {
#pragma unroll
                for (int c = 0; c < 4; ++c) {
                    const int a = a0 + 32 * s + 8 * u + 4 * h + c;
                    if (!bin || a >= p.vl) continue;
                    const double D = ei - eva[s][u][c] - evb;
                    const double g = G[s][t][u][c];
                    const double hh = (MODE == MODE_OS) ? 0.0 : H[s][t][u][c];
                    double e1, m1, z1;
                    pair_terms<MODE>(g, hh, D, e1, m1, z1);
                    double q = m1, e2 = 0.0, m2 = 0.0, z2 = 0.0;
                    e += e1;
                    if (mirror) {
                        pair_terms<MODE>(hh, g, D, e2, m2, z2);
                        e += e2;
                        q += m2;
                    }
                    if (grad) {
                        sm += q;
                        rowq[s][u][c] += q;
                        colq[t] += q;
                    }
                    if (MODE == MODE_T2) {
                        const bool ok = fabs(D) >= 1e-12;
                        t2base[(long)a * p.vr + b] = ok ? g / D : 0.0;
                        if (mirror) t2base[(long)b * p.vr + a] = ok ? hh / D : 0.0;
                    } else if (Z) {
                        Z[(zr + a) * p.ldz + zc + b] = z1;
                        if (mirror) Z[(zr + b) * p.ldz + zc + a] = z2;
                        if (p.ztrans && i != j) {
                            Z[(zr2 + b) * p.ldz + zc2 + a] = z1;
                            if (mirror) Z[(zr2 + a) * p.ldz + zc2 + b] = z2;
                        }
                    }
                }
            }
    }
    // Deterministic reductions: fixed in-lane order, then fixed shuffle trees.
    e = xor_sum(e, 1, 64);
    if (lane == 0) p.e_part[slot] = w * e;
    if (!grad) return;
    sm = xor_sum(sm, 1, 64);
    if (lane == 0) p.s_part[slot] = w * sm;
#pragma unroll
    for (int s = 0; s < 2; ++s)
#pragma unroll
        for (int u = 0; u < 4; ++u)
#pragma unroll
            for (int c = 0; c < 4; ++c) {
                const double v = xor_sum(rowq[s][u][c], 1, 32);
                if (col == 0) p.vrow[slot * 64 + 32 * s + 8 * u + 4 * h + c] = w * v;
            }
#pragma unroll
    for (int t = 0; t < 2; ++t) {
        const double v = xor_sum(colq[t], 32, 64);
        if (h == 0) p.vcol[slot * 64 + 32 * t + col] = w * v;
    }
}

template <int MODE>
__device__ __forceinline__ void pair_epilogue32(const PairArgs& p, int i, int j, double w, int a0, int b0,
                                                bool mirror, int lane, const v4d (&G)[2][2][4],
                                                const v4d (&H)[2][2][4], long slot) {
    pair_epilogue32_r<MODE>(p, p.evl, p.evr, p.Z, i, j, w, a0, b0, mirror, lane, G, H, slot);
}

template <int MODE, int S, int SPB>
__global__ __launch_bounds__(256) void pair_bf16_kernel(PairArgs p, const uint16_t* __restrict__ Psl,
                                                        const int8_t* __restrict__ Pel,
                                                        const uint16_t* __restrict__ Psr,
                                                        const int8_t* __restrict__ Per, int Qp, int vEl,
                                                        int vEr) {
    constexpr bool SYM = MODE != MODE_OS;
    constexpr int NBLK = SYM ? 12 : 8;         // 32-row panel blocks per slice and stage
    constexpr int BPW = NBLK / 4;              // blocks streamed per wave
    constexpr long SLICE = (long)NBLK * 1024;  // bytes per slice and stage
    // LDS ring of per-slice slots: (stage st, slice sl) lives in slot
    // (S st + sl) mod NSLOT, NSLOT = 2S + PF.  During stage st a wave loads
    // slices PF..S-1 of stage st+1, then slices 0..PF-1 of stage st+2; those
    // land in exactly the S slots stage st-1 used, free after its closing
    // barrier, and the barrier closing stage st waits only until the stage
    // st+1 slices have landed (the stage st+2 ones may stay in flight).
    // PF = S (three whole stages) where LDS allows; the symmetric units'
    // 12 KiB slices fit 13 slots, so PF = 3 at S = 5.
    constexpr int NSLOT = ring_slots(S, SYM);
    constexpr int PF = NSLOT - 2 * S;  // slices of stage st+2 loaded during stage st
    extern __shared__ __align__(16) char lds[];

    const long work = xcd_work_index();
    if (work >= (long)p.npairs * p.nunits) return;
    const int pslot = (int)(work / p.nunits), uslot = (int)(work % p.nunits);
    const int2 ij = p.pairs[pslot];
    const int i = ij.x, j = ij.y;
    const double w = p.pw[pslot];
    const int2 un = p.units[uslot];
    const bool diag = SYM && un.x < 0;            // diagonal 128 x 128 block
    const int R0 = diag ? un.y * 128 : SYM ? un.x * 64 : un.x * 128;  // first row (left virtual index)
    const int C0 = un.y * 128;                    // first column (right virtual index)
    const int tid = threadIdx.x, lane = tid & 63;
    const int wv = __builtin_amdgcn_readfirstlane(tid >> 6);

    // DMA sources: lane l of a 32-row block fetches row l/2, the logical
    // segment that lds_off32() stores at its physical position l&1.  Rows
    // past the operand are clamped to its last row (their outputs are
    // discarded).
    const long ovl = (long)p.ol * p.vl, ovr = (long)p.orr * p.vr;
    const int lr = lane >> 1;
    const int lseg = (lane & 1) ^ (((lr >> 2) ^ (lr >> 3)) & 1);
    bool rleft[BPW];    // source operand (wave-uniform): Bl (true) or Br
    unsigned roff[BPW]; // the lane's byte offset in a stage block (slice 0)
    int rdst[BPW];
#pragma unroll
    for (int k = 0; k < BPW; ++k) {
        // a diagonal block streams its 8 blocks two per wave (the third load
        // repeats the first: same source, same place)
        const int blk = diag ? (k < 2 ? 2 * wv + k : 2 * wv) : BPW * wv + k;
        rdst[k] = blk * 1024;
        bool left;
        int va;  // virtual index of the block's first row in its orbital
        if (SYM && !diag) {
            left = blk < 2 || blk >= 8;
            va = blk < 2 ? R0 + 32 * blk : blk < 6 ? C0 + 32 * (blk - 2) : blk < 8 ? R0 + 32 * (blk - 6)
                                                                                   : C0 + 32 * (blk - 8);
        } else {
            // diagonal block (symmetric) or opposite-spin unit: 0-3 Bl_i
            // rows, 4-7 Br_j rows
            left = blk < 4;
            va = blk < 4 ? R0 + 32 * blk : C0 + 32 * (blk - 4);
        }
        const long rows = left ? ovl : ovr;
        const long r0 = (long)(left ? i : j) * (left ? p.vl : p.vr) + va;
        // Rows past the orbital's virtual count (padding of the last tile)
        // read the all-zero 32-row block after the operand's last, so the
        // MFMAs whose outputs are discarded run on zeros.
        const long row = va + lr < (left ? p.vl : p.vr) ? r0 + lr : (rows + 31) / 32 * 32 + lr;
        rleft[k] = left;
        roff[k] = (unsigned)(2 * (((row >> 5) * S * 32 + (row & 31)) * KS + 8 * lseg));
    }
    // Slices are [Qp/16][ceil(o*v/32) + 1][S][32][16]: the S slices of one
    // 32-row block of a 16-Q stage are adjacent KiBs.  A load addresses its
    // stage by the resource base, its row by the lane offset and its slice by
    // a constant scalar offset.
    const long sstepl = (long)S * 32 * KS * ((ovl + 31) / 32 + 1), sstepr = (long)S * 32 * KS * ((ovr + 31) / 32 + 1);
    // One slice of a stage; a stage's loads are spread over its MFMA groups
    // (a burst of them stalls instruction issue behind the address unit).
    const int nst = Qp / KS;
    // Ring state, as LDS byte offsets: cur[sl] holds slice sl of the current
    // stage st, lo[k] slice k < PF of stage st+1 (loaded during stage st-1),
    // fslot[g] the slots stage st-1 read, which stage st's load items g fill
    // (items g < S-PF: slice PF+g of stage st+1; the rest: slice g-(S-PF) of
    // stage st+2).  Each stage permutes these 2S+PF offsets; inside the
    // unrolled stages that is register renaming, where the modulo form costs
    // ~60 scalar instructions a stage.  The symmetric S = 6 kernel has no
    // registers to spare for the state and keeps the modulo form.
    constexpr bool RREG = !(SYM && S >= 6);
    int cur[S], lo[PF > 0 ? PF : 1], fslot[S];
#pragma unroll
    for (int k = 0; k < S; ++k) {
        cur[k] = k * (int)SLICE;
        fslot[k] = (S + PF + k) * (int)SLICE;
    }
#pragma unroll
    for (int k = 0; k < PF; ++k) lo[k] = (S + k) * (int)SLICE;
    auto mslot = [&](int st, int sl) { return (int)(((S * st + sl) % NSLOT) * SLICE); };
    // offset of slice q of stage st, and the target of load item g in stage st
    auto rd_off = [&](int st, int q) { return RREG ? cur[q] : mslot(st, q); };
    auto ld_off = [&](int st, int g) {
        return RREG ? fslot[g] : g < S - PF ? mslot(st + 1, PF + g) : mslot(st + 2, g - (S - PF));
    };
    // Every wave issues exactly BPW loads per slice on every path (rows are
    // clamped; a diagonal block's third load repeats its first, same bytes
    // to the same place; past the last stage the source stage is clamped
    // into slots no one reads), so the waits below count loads exactly and
    // the compiler's own wait analysis sees one straight path.
    auto issue_slice = [&](int st, int sl, char* dst) {
        const long sq = min(st, nst - 1);
        const uint16_t* bl = Psl + sq * sstepl;
        const uint16_t* br = Psr + sq * sstepr;
#pragma unroll
        for (int k = 0; k < BPW; ++k) {
            // Buffer loads (their slice offsets below 4 KiB in the scalar
            // operand) in the symmetric kernels up to four slices.  The
            // five- and six-slice ones keep the global form, since the
            // resources leave them too few scalar registers to build without
            // scratch; so does the opposite-spin kernel, where the global
            // form measured 1.3% faster.
            if constexpr (!SYM || S >= 5)
                dma16(reinterpret_cast<const char*>(rleft[k] ? bl : br) + roff[k] + 1024 * sl, dst + rdst[k]);
            else
                dma16b(rleft[k] ? bl : br, roff[k] + 4096u * (sl >> 2), 1024u * (sl & 3), dst + rdst[k]);
        }
    };
    // Wait until at most the PF newest slices' loads of this wave are in
    // flight, then a workgroup barrier.
    auto ring_barrier = [&]() { dma_barrier<PF * BPW>(); };

    // This wave's operand rows in a stage (LDS panel rows) and exponent rows.
    int prow, pcol;          // first LDS panel row of the row / column operand
    int arow, bcol;          // first virtual index of this wave's rows / columns
    const int8_t *erow, *ecol;
    int vErow, vEcol;
    if (SYM && !diag) {
        const int isH = wv >> 1, half = wv & 1;
        prow = isH ? 192 : 0;
        pcol = (isH ? 256 : 64) + 64 * half;
        arow = R0;
        bcol = C0 + 64 * half;
        erow = isH ? Per + (long)j * vEr : Pel + (long)i * vEl;
        ecol = isH ? Pel + (long)i * vEl : Per + (long)j * vEr;
        vErow = isH ? p.orr * vEr : p.ol * vEl;
        vEcol = isH ? p.ol * vEl : p.orr * vEr;
    } else {
        const int wm = wv >> 1, wn = wv & 1;
        prow = 64 * wm;
        pcol = 128 + 64 * wn;
        arow = R0 + 64 * wm;
        bcol = C0 + 64 * wn;
        erow = Pel + (long)i * vEl;
        ecol = Per + (long)j * vEr;
        vErow = p.ol * vEl;
        vEcol = p.orr * vEr;
    }
    const int h = lane >> 5, fr = lane & 31;

    v16f acc[3][2][2];
#pragma unroll
    for (int c = 0; c < 3; ++c)
#pragma unroll
        for (int s = 0; s < 2; ++s)
#pragma unroll
            for (int t = 0; t < 2; ++t)
#pragma unroll
                for (int r = 0; r < 16; ++r) acc[c][s][t][r] = 0.f;
    v4d acc64[2][2][4];
#pragma unroll
    for (int s = 0; s < 2; ++s)
#pragma unroll
        for (int t = 0; t < 2; ++t)
#pragma unroll
            for (int u = 0; u < 4; ++u) acc64[s][t][u] = v4d{0.0, 0.0, 0.0, 0.0};

    constexpr int QB = SPB * KS;  // Q per exponent block
#pragma unroll
    for (int sl = 0; sl < S; ++sl) issue_slice(0, sl, lds + cur[sl]);
#pragma unroll
    for (int sl = 0; sl < PF; ++sl) issue_slice(1, sl, lds + lo[sl]);
    ring_barrier();  // stage 0 has landed
    for (int qb = 0; qb < Qp / QB; ++qb) {
        // Exponents of this Q block, fetched early and without branches
        // (rows past the operand read a clamped row; those outputs are
        // discarded).  Row exponents come four rows to a dword: the tables
        // are [Qp/QB][o][vE] with vE a multiple of 128.
        int erw[2][4], ecl[2];
#pragma unroll
        for (int s = 0; s < 2; ++s)
#pragma unroll
            for (int u = 0; u < 4; ++u)
                erw[s][u] = *reinterpret_cast<const int*>(erow + (long)qb * vErow + arow + 32 * s + 8 * u + 4 * h);
#pragma unroll
        for (int t = 0; t < 2; ++t) ecl[t] = ecol[(long)qb * vEcol + bcol + 32 * t + fr];
#pragma unroll
        for (int c8 = 0; c8 < SPB; ++c8) {
            const int st = qb * SPB + c8;
            // Group pp's row fragments are read first and the column
            // fragments in q order (group 0 uses bf[q] in that order), so
            // the first MFMAs wait for only a few reads; the next group's row
            // fragments are read while the current group's MFMAs run.  (The
            // MFMA order, and so every FP32 rounding, is unchanged.)
            auto rd_col = [&](v8s(&f)[2], int q) {
#pragma unroll
                for (int t = 0; t < 2; ++t)
                    f[t] = *reinterpret_cast<const v8s*>(lds + rd_off(st, q) + lds_off32(pcol + 32 * t + fr, h));
            };
            auto rd_row = [&](v8s(&f)[2], int pp) {
#pragma unroll
                for (int s = 0; s < 2; ++s)
                    f[s] = *reinterpret_cast<const v8s*>(lds + rd_off(st, pp) + lds_off32(prow + 32 * s + fr, h));
            };
            v8s bf[S][2], af[2], an[2];
            rd_row(af, 0);
#pragma unroll
            for (int q = 0; q < S; ++q) rd_col(bf[q], q);
#pragma unroll
            for (int pp = 0; pp < S; ++pp) {
                // load item pp of this stage: slice PF + pp of stage st+1,
                // then slices 0.. of stage st+2
                if (pp < S - PF) issue_slice(st + 1, PF + pp, lds + ld_off(st, pp));
                else issue_slice(st + 2, pp - (S - PF), lds + ld_off(st, pp));
                if (pp + 1 < S) rd_row(an, pp + 1);
#pragma unroll
                for (int q = 0; q < S - pp; ++q)
#pragma unroll
                    for (int s = 0; s < 2; ++s)
#pragma unroll
                        for (int t = 0; t < 2; ++t)
                            acc[ozaki_tier(pp, q)][s][t] = mfma32_bf16(af[s], bf[q][t], acc[ozaki_tier(pp, q)][s][t]);
                if (pp + 1 < S) {
                    af[0] = an[0];
                    af[1] = an[1];
                }
            }
            // stage st+1 must have landed; stage st+2's slices may stay in flight
            ring_barrier();
            // advance the ring: stage st+1's slices are lo[] then the first
            // S-PF slots just filled; the rest now hold stage st+2's low
            // slices; stage st's slots are free
            if (RREG) {
                int nc[S], nl[PF > 0 ? PF : 1];
#pragma unroll
                for (int k = 0; k < S; ++k) nc[k] = k < PF ? lo[k] : fslot[k - PF];
#pragma unroll
                for (int k = 0; k < PF; ++k) nl[k] = fslot[S - PF + k];
#pragma unroll
                for (int k = 0; k < S; ++k) {
                    fslot[k] = cur[k];
                    cur[k] = nc[k];
                }
#pragma unroll
                for (int k = 0; k < PF; ++k) lo[k] = nl[k];
            }
        }
        // FP64 flush of the QB-wide block: (tier0 + tier1) + tier2, scaled by
        // 2^(e_row + e_col - 16).
#pragma unroll
        for (int s = 0; s < 2; ++s)
#pragma unroll
            for (int t = 0; t < 2; ++t)
#pragma unroll
                for (int r = 0; r < 16; ++r) {
                    const int er = (int)(signed char)(erw[s][r >> 2] >> (8 * (r & 3)));
                    const double sum =
                        ((double)acc[0][s][t][r] + (double)acc[1][s][t][r]) + (double)acc[2][s][t][r];
                    acc64[s][t][r >> 2][r & 3] += sum * pow2(er + ecl[t] - 16);
                    acc[0][s][t][r] = acc[1][s][t][r] = acc[2][s][t][r] = 0.f;
                }
    }
    // The clamped loads of the last two stages land before the LDS is
    // reused (or handed to the next workgroup).
    dma_barrier<0>();

    // Self terms: on a diagonal tile (arow = bcol) of a pair i = j the
    // elements a = b take the FP64 sums sum_Q Bl_ia Br_ia.  Lane (h, fr)
    // holds element a = b = arow + 32t + fr of subtile (t, t) when
    // h = (fr >> 2) & 1, at u = fr >> 3, c = fr & 3.  They are set here, in
    // the block accumulators, so the mirrored tiles exchanged through LDS
    // carry them too and the epilogue needs no extra registers.
    if (p.gself != nullptr && i == j && arow == bcol && i < p.nself_o) {
        const bool mine = h == ((fr >> 2) & 1);
#pragma unroll
        for (int t = 0; t < 2; ++t) {
            const int a = arow + 32 * t + fr;
            const double gs = p.gself[(long)i * p.nself_v + min(a, p.nself_v - 1)];
            const bool use = mine && a < p.nself_v;
#pragma unroll
            for (int u = 0; u < 4; ++u)
#pragma unroll
                for (int c = 0; c < 4; ++c)
                    if (use && u == (fr >> 3) && c == (fr & 3)) acc64[t][t][u][c] = gs;
        }
    }

    v4d accH[2][2][4];
    if (SYM && diag) {
        // Diagonal block: wave 2 holds tile (2k+1, 2k), the mirror of wave 1's
        // tile (2k, 2k+1); waves 0 and 3 hold diagonal tiles, their own mirrors.
        // Tiles go through LDS as [64][65] doubles (row, column) and the
        // mirror is read transposed.
        double* T = reinterpret_cast<double*>(lds);
        constexpr int TP = 65;
        const int put = wv == 2 ? 0 : wv == 0 ? 1 : wv == 3 ? 2 : -1;
        if (put >= 0) {
#pragma unroll
            for (int s = 0; s < 2; ++s)
#pragma unroll
                for (int t = 0; t < 2; ++t)
#pragma unroll
                    for (int u = 0; u < 4; ++u)
#pragma unroll
                        for (int c = 0; c < 4; ++c)
                            T[(put * 64 + 32 * s + 8 * u + 4 * h + c) * TP + 32 * t + fr] = acc64[s][t][u][c];
        }
        __syncthreads();
        if (wv == 2) return;
        const int get = wv == 1 ? 0 : wv == 0 ? 1 : 2;
#pragma unroll
        for (int s = 0; s < 2; ++s)
#pragma unroll
            for (int t = 0; t < 2; ++t)
#pragma unroll
                for (int u = 0; u < 4; ++u)
#pragma unroll
                    for (int c = 0; c < 4; ++c)
                        accH[s][t][u][c] = T[(get * 64 + 32 * t + fr) * TP + 32 * s + 8 * u + 4 * h + c];
        const int A = 2 * un.y + (wv >> 1), B = 2 * un.y + (wv & 1);
        if (A >= p.ntA || B >= p.ntB) return;
        const long tslot = (long)A * p.ntA - (long)A * (A - 1) / 2 + (B - A);
        pair_epilogue32<MODE>(p, i, j, w, 64 * A, 64 * B, A < B, lane, acc64, accH,
                              (long)pslot * p.ntiles + tslot);
    } else if (SYM) {
        // H waves hand their tile to the G waves through LDS (same lane map).
        double* Hs = reinterpret_cast<double*>(lds);
        const int isH = wv >> 1, half = wv & 1;
        if (isH) {
#pragma unroll
            for (int s = 0; s < 2; ++s)
#pragma unroll
                for (int t = 0; t < 2; ++t)
#pragma unroll
                    for (int u = 0; u < 4; ++u)
#pragma unroll
                        for (int c = 0; c < 4; ++c)
                            Hs[((((half * 2 + s) * 2 + t) * 4 + u) * 4 + c) * 64 + lane] = acc64[s][t][u][c];
        }
        __syncthreads();
        if (isH) return;
#pragma unroll
        for (int s = 0; s < 2; ++s)
#pragma unroll
            for (int t = 0; t < 2; ++t)
#pragma unroll
                for (int u = 0; u < 4; ++u)
#pragma unroll
                    for (int c = 0; c < 4; ++c)
                        accH[s][t][u][c] = Hs[((((half * 2 + s) * 2 + t) * 4 + u) * 4 + c) * 64 + lane];
        const int A = un.x, B = 2 * un.y + half;
        if (B >= p.ntB || A > B) return;
        const long tslot = (long)A * p.ntA - (long)A * (A - 1) / 2 + (B - A);
        pair_epilogue32<MODE>(p, i, j, w, 64 * A, 64 * B, A < B, lane, acc64, accH,
                              (long)pslot * p.ntiles + tslot);
    } else {
#pragma unroll
        for (int s = 0; s < 2; ++s)
#pragma unroll
            for (int t = 0; t < 2; ++t)
#pragma unroll
                for (int u = 0; u < 4; ++u) accH[s][t][u] = v4d{0.0, 0.0, 0.0, 0.0};
        const int A = 2 * un.x + (wv >> 1), B = 2 * un.y + (wv & 1);
        if (A >= p.ntA || B >= p.ntB) return;
        pair_epilogue32<MODE>(p, i, j, w, 64 * A, 64 * B, false, lane, acc64, accH,
                              (long)pslot * p.ntiles + (long)A * p.ntB + B);
    }
}

// ----------------------------------------------------------- GEMM (NT)

struct SlicedMat {
    const uint16_t* s;  // [S] planes, plane stride `plane`
    long plane, ld;     // elements
    const int8_t* e;    // exponents [kb][row], k-block stride `lde` (a multiple of GT >= rows)
    long lde;
    int rows;
};

// C[m][n] += sum_k A[m][k] X[n][k] for the gradient contractions; K is a
// multiple of GB.  128 x 128 workgroup tile, eight waves of 64 x 32 (two
// per SIMD, so one wave's MFMAs fill the other's load, LDS and barrier
// gaps), one 32-wide k step per stage; stages are loaded straight into a
// triple-buffered LDS by the DMA path, two stages ahead (one barrier per
// stage, no staging registers).  The
// gradient needs ~1e-7 relative accuracy, so products go to two FP32 tiers:
// c = 0 (integers, exact up to the last bit above 2^23) and c >= 1 (rounding
// near 2^-33 of the block scale); both are flushed to FP64 with the block
// exponents once per GB-wide k block.
// Output row m of C is stored at C + m * ldc, or, when mdiv > 0, at
// C + (m / mdiv) * ld1 + (m % mdiv) * ldc (the transform writes its rows
// (i, q) into the [q][i][a] layout of B).
struct OutRows {
    long ldc;
    int mdiv;
    long ld1;
};

// Stages of the GEMM's LDS ring: three (loads two stages ahead) while they
// fit 160 KiB, else two (loads one stage ahead).
constexpr int gemm_nbuf(int S) { return 3 * 2 * S * GT * LROW <= 160 * 1024 ? 3 : 2; }

template <int S, int GB>
__global__ __launch_bounds__(GW * 64) void gemm_bf16_nt_kernel(SlicedMat A, SlicedMat X, int K,
                                                               double* __restrict__ C, OutRows out) {
    constexpr long PANEL = (long)GT * LROW;   // bytes of one slice panel (128 rows x 32 k)
    constexpr long STAGE = 2L * S * PANEL;    // both operands, all slices
    constexpr int NBUF = gemm_nbuf(S);
    constexpr int AHEAD = NBUF - 1;           // stages loaded ahead of the one in use
    constexpr int PER = 2 * S * (GT / 16) / GW; // 16-row DMA loads per wave and stage
    constexpr int WAITN = (AHEAD - 1) * PER;  // loads left in flight at a stage barrier
    static_assert(PER == 2 * S, "each wave loads one 16-row block of every operand slice");
    extern __shared__ __align__(16) char lds[];
    int mt, ntile;
    if (!grouped_tile((A.rows + GT - 1) / GT, (X.rows + GT - 1) / GT, mt, ntile)) return;
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    // eight waves, two per SIMD: wave w owns rows wm .. wm+63, columns
    // wn .. wn+31 of the 128 x 128 tile
    const int wm = (wave >> 2) * 64, wn = (wave & 3) * 32;
    const int m0 = mt * GT, n0 = ntile * GT;
    // DMA load u of wave w covers operand P = u / S, slice sl = u % S and
    // panel rows 16w .. 16w+15.  Rows past the matrix are clamped to its last
    // row, and chunks past the end to the last chunk (those LDS bytes only
    // feed discarded outputs or are never read): every wave issues exactly
    // PER loads per chunk on every path, which the vmcnt(PER) chunk barrier
    // relies on.
    const int wv = __builtin_amdgcn_readfirstlane(wave);
    const int lrow = lane >> 2;
    const int lseg8 = 8 * lane_seg(lane);
    const int nkb = K / GB;
    constexpr int CPB = GB / KC;  // chunks per exponent block
    const int nchunk = nkb * CPB;
    // Loads u in [2g, 2g+2) of a chunk; a chunk's loads are spread over its
    // S MFMA groups.
    // Buffer loads: the wave's first (clamped) row in the resource base, the
    // lane's row and segment in its byte offset (below 16 rows).
    int rb[2];
    unsigned voff[2];
#pragma unroll
    for (int Pn = 0; Pn < 2; ++Pn) {
        const SlicedMat& M = Pn ? X : A;
        const int r0 = (Pn ? n0 : m0) + 16 * wv;
        rb[Pn] = min(r0, M.rows - 1);
        voff[Pn] = (unsigned)(2 * ((long)(min(r0 + lrow, M.rows - 1) - rb[Pn]) * M.ld + lseg8));
    }
    auto issue_part = [&](int ch, int g) {
        const int k0 = min(ch, nchunk - 1) * KC;
        char* dst = lds + (ch % NBUF) * STAGE;
#pragma unroll
        for (int u = 2 * g; u < 2 * g + 2; ++u) {
            const int Pn = u / S, sl = u % S;
            const SlicedMat& M = Pn ? X : A;
            dma16b(M.s + sl * M.plane + (long)rb[Pn] * M.ld + k0, voff[Pn], 0, dst + (Pn * S + sl) * PANEL + 1024 * wv);
        }
    };
    v4f acc[2][4][2];
    v4d acc64[4][2];
#pragma unroll
    for (int s = 0; s < 4; ++s)
#pragma unroll
        for (int t = 0; t < 2; ++t) {
            acc64[s][t] = v4d{0.0, 0.0, 0.0, 0.0};
            acc[0][s][t] = acc[1][s][t] = v4f{0.f, 0.f, 0.f, 0.f};
        }
    const int rr = lane & 15, fs = lane >> 4;
#pragma unroll
    for (int a = 0; a < AHEAD; ++a)
#pragma unroll
        for (int g = 0; g < S; ++g) issue_part(a, g);
    dma_barrier<WAITN>();
    for (int kb = 0; kb < nkb; ++kb) {
        // Exponents of this block, fetched early and without branches (rows
        // past the matrix read the last row's: their outputs are discarded),
        // so the loads issue back to back and their wait lands at the flush.
        // The tables are padded to whole 128-row tiles, so no row is clamped;
        // the four rows 4fs .. 4fs+3 of a 16-row block are one dword.
        int eA[4], eB[2];
#pragma unroll
        for (int t = 0; t < 2; ++t) eB[t] = X.e[(long)kb * X.lde + n0 + wn + 16 * t + rr];
#pragma unroll
        for (int s = 0; s < 4; ++s)
            eA[s] = *reinterpret_cast<const int*>(A.e + (long)kb * A.lde + m0 + wm + 16 * s + 4 * fs);
#pragma unroll
        for (int c = 0; c < CPB; ++c) {
            const int ch = kb * CPB + c;
            const char* pa = lds + (ch % NBUF) * STAGE;
            const char* pb = pa + S * PANEL;
            // Row fragments first, column fragments in q order (group 0 uses
            // them in that order); the other wave on the SIMD covers the
            // read latency of each group's row fragments.
            v8s b[S][2];
#pragma unroll
            for (int q = 0; q < S; ++q)
#pragma unroll
                for (int t = 0; t < 2; ++t) b[q][t] = lds_frag(pb + q * PANEL, wn + 16 * t + rr, fs);
#pragma unroll
            for (int p = 0; p < S; ++p) {
                issue_part(ch + AHEAD, p);
                v8s a[4];
#pragma unroll
                for (int s = 0; s < 4; ++s) a[s] = lds_frag(pa + p * PANEL, wm + 16 * s + rr, fs);
#pragma unroll
                for (int q = 0; q < S - p; ++q)
#pragma unroll
                    for (int s = 0; s < 4; ++s)
#pragma unroll
                        for (int t = 0; t < 2; ++t)
                            acc[p + q ? 1 : 0][s][t] = mfma_bf16(a[s], b[q][t], acc[p + q ? 1 : 0][s][t]);
            }
            // Chunk ch+1 must have landed; with three stages chunk ch+2 may
            // stay in flight.
            dma_barrier<WAITN>();
        }
#pragma unroll
        for (int s = 0; s < 4; ++s)
#pragma unroll
            for (int t = 0; t < 2; ++t) {
#pragma unroll
                for (int r = 0; r < 4; ++r) {
                    const int er = (int)(signed char)(eA[s] >> (8 * r));
                    const double sum = (double)acc[0][s][t][r] + (double)acc[1][s][t][r];
                    acc64[s][t][r] += sum * pow2(er + eB[t] - 16);
                }
                acc[0][s][t] = acc[1][s][t] = v4f{0.f, 0.f, 0.f, 0.f};
            }
    }
    // The clamped loads of the last two chunks land before the LDS is
    // handed to the next workgroup.
    dma_barrier<0>();
#pragma unroll
    for (int s = 0; s < 4; ++s)
#pragma unroll
        for (int r = 0; r < 4; ++r) {
            const int m = m0 + wm + 16 * s + 4 * fs + r;
            if (m >= A.rows) continue;
            const long row = out.mdiv > 0 ? (long)(m / out.mdiv) * out.ld1 + (long)(m % out.mdiv) * out.ldc
                                          : (long)m * out.ldc;
#pragma unroll
            for (int t = 0; t < 2; ++t) {
                const int n = n0 + wn + 16 * t + rr;
                if (n < X.rows) C[row + n] += acc64[s][t][r];
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

// The AO->MO transform takes 2 to BF16_TRANSFORM_SLICES slices.
template <template <int> class F, typename... Args>
int by_tslices(int S, Args&&... args) {
    static_assert(BF16_TRANSFORM_SLICES == 4, "instantiate the transform slice counts");
    switch (S) {
        case 2: return F<2>::run(args...);
        case 3: return F<3>::run(args...);
        case 4: return F<4>::run(args...);
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

// Row pitch (elements) of the gradient GEMM operands' slices of n columns:
// whole GB_GRAD-wide exponent blocks, plus one more (zeros) when the pitch
// would be a multiple of 8 KiB.  Rows a power of two apart (n = 32400 pads to
// 32768: 64 KiB) put a 128-row panel's rows on the same memory channels.
int grad_pitch(int n) {
    const int k = round_up(n, GB_GRAD);
    return (2L * k) % 8192 == 0 ? k + GB_GRAD : k;
}

// Dynamic LDS of the BF16 GEMM: its ring of stages of both operands' slices.
constexpr size_t gemm_lds(int S) { return (size_t)gemm_nbuf(S) * 2 * S * GT * LROW; }

template <int S>
struct SlicePair {
    template <int QB>
    static void launch(mshqc_hip_ctx* ctx, const double* B, int Q, int o, int v, int vE, int Qp, uint16_t* Ps,
                       int8_t* Pe) {
        const int ov = o * v;
        dim3 grid(ceil_div(ov, 32) + 1, Qp / QB);
        slice_pair_kernel<S, QB><<<grid, 256, slice_pair_lds(QB), ctx->stream>>>(B, Q, ov, v, vE, o, Qp, Ps, Pe);
    }
    static int run(mshqc_hip_ctx* ctx, int spb, const double* B, int Q, int o, int v, int vE, int Qp,
                   uint16_t* Ps, int8_t* Pe) {
        if (spb == SPB_LONG) launch<SPB_LONG * KS>(ctx, B, Q, o, v, vE, Qp, Ps, Pe);
        else launch<SPB_SHORT * KS>(ctx, B, Q, o, v, vE, Qp, Ps, Pe);
        LAUNCH_CHECK(ctx);
        return 0;
    }
};

// The functors below take the exponent block width gb (GB_GRAD or
// GB_TRANSFORM) at run time and launch the matching instantiation.
template <int S>
struct SliceRows {
    template <int GB>
    static void launch(mshqc_hip_ctx* ctx, const double* src, long ld, int R, int nseg, int v, int vp,
                       uint16_t* dst, int8_t* ex, long ldx) {
        const long jobs = (long)R * nseg * (vp / GB);
        slice_rows_kernel<S, GB><<<ceil_div(jobs, 4), 256, 0, ctx->stream>>>(src, ld, R, nseg, v, vp, dst, ex, ldx);
    }
    static int run(mshqc_hip_ctx* ctx, int gb, const double* src, long ld, int R, int nseg, int v, int vp,
                   uint16_t* dst, int8_t* ex, long ldx) {
        if (gb == GB_GRAD) launch<GB_GRAD>(ctx, src, ld, R, nseg, v, vp, dst, ex, ldx);
        else launch<GB_TRANSFORM>(ctx, src, ld, R, nseg, v, vp, dst, ex, ldx);
        LAUNCH_CHECK(ctx);
        return 0;
    }
};

template <int S>
struct SliceCols {
    template <int GB>
    static void launch(mshqc_hip_ctx* ctx, const double* Z, long ldz, int nI, int v, int vp, int nc,
                       uint16_t* dst, int8_t* ex, long ldx) {
        const long items = (long)ceil_div(nc, sc_cols(GB)) * nI * (vp / GB);
        slice_cols_kernel<S, GB><<<(unsigned)((items + 7) / 8 * 8), 256, 0, ctx->stream>>>(Z, ldz, nI, v, vp, nc,
                                                                                     dst, ex, ldx);
    }
    static int run(mshqc_hip_ctx* ctx, int gb, const double* Z, long ldz, int nI, int v, int vp, int nc,
                   uint16_t* dst, int8_t* ex, long ldx) {
        // one workgroup per (a block, column tile), 1-D grid
        if ((long)ceil_div(nc, sc_cols(gb)) * nI * (vp / gb) > 0x7ffffff8L)
            return set_error(ctx, MSHQC_HIP_ERR_ARG, "split-BF16 transposed slicing: grid too large");
        if (gb == GB_GRAD) launch<GB_GRAD>(ctx, Z, ldz, nI, v, vp, nc, dst, ex, ldx);
        else launch<GB_TRANSFORM>(ctx, Z, ldz, nI, v, vp, nc, dst, ex, ldx);
        LAUNCH_CHECK(ctx);
        return 0;
    }
};

// Dynamic LDS of the pair kernel: the ring of slice slots (12 KiB in the
// symmetric modes, 8 KiB in the opposite-spin mode), reused in the symmetric
// modes by the epilogue exchanges (two H tiles, or three [64][65] tiles of a
// diagonal block).
constexpr size_t pair_lds(int S, bool sym) {
    const size_t ring = (size_t)ring_slots(S, sym) * (sym ? 12 : 8) * 1024;
    return sym ? std::max(ring, (size_t)3 * 64 * 65 * 8) : ring;
}

template <int S>
struct PairLaunch {
    static int run(mshqc_hip_ctx* ctx, int mode, const PairArgs& a, const Bf16Operands& o) {
        const long nwork = (long)a.npairs * a.nunits;
        const long nwg = ((nwork + 7) / 8) * 8;
        if (nwg > 0x7fffffffL) return set_error(ctx, MSHQC_HIP_ERR_ARG, "split-BF16 pair kernel: grid too large");
        dim3 grid((unsigned)nwg);
        const uint16_t* sr = o.sr ? o.sr : o.sl;
        const int8_t* xr = o.sr ? o.xr : o.xl;
        if (mode != MODE_RMP2 && mode != MODE_SS && mode != MODE_OS)
            return set_error(ctx, MSHQC_HIP_ERR_ARG, "split-BF16 pair kernel: unsupported mode %d", mode);
        auto go = [&](auto spb_c) {
            constexpr int SPB = decltype(spb_c)::value;
            if (mode == MODE_RMP2)
                pair_bf16_kernel<MODE_RMP2, S, SPB><<<grid, 256, pair_lds(S, true), ctx->stream>>>(
                    a, o.sl, o.xl, sr, xr, o.Qp, o.vEl, o.vEr);
            else if (mode == MODE_SS)
                pair_bf16_kernel<MODE_SS, S, SPB><<<grid, 256, pair_lds(S, true), ctx->stream>>>(
                    a, o.sl, o.xl, sr, xr, o.Qp, o.vEl, o.vEr);
            else
                pair_bf16_kernel<MODE_OS, S, SPB><<<grid, 256, pair_lds(S, false), ctx->stream>>>(
                    a, o.sl, o.xl, sr, xr, o.Qp, o.vEl, o.vEr);
        };
        if constexpr (spb_long_ok(S)) {
            if (o.spb == SPB_LONG) {
                go(std::integral_constant<int, SPB_LONG>{});
                LAUNCH_CHECK(ctx);
                return 0;
            }
        }
        go(std::integral_constant<int, SPB_SHORT>{});
        LAUNCH_CHECK(ctx);
        return 0;
    }
};

template <int S>
struct GemmLaunch {
    static int run(mshqc_hip_ctx* ctx, int gb, const SlicedMat& A, const SlicedMat& X, int K, double* C, long ldc,
                   int mdiv = 0, long ld1 = 0) {
        if (K % gb != 0)
            return set_error(ctx, MSHQC_HIP_ERR_ARG, "split-BF16 GEMM: K=%d is not a multiple of %d", K, gb);
        dim3 grid(grouped_grid(ceil_div(A.rows, GT), ceil_div(X.rows, GT)));
        const OutRows out{ldc, mdiv, ld1};
        if (gb == GB_GRAD)
            gemm_bf16_nt_kernel<S, GB_GRAD><<<grid, GW * 64, gemm_lds(S), ctx->stream>>>(A, X, K, C, out);
        else
            gemm_bf16_nt_kernel<S, GB_TRANSFORM><<<grid, GW * 64, gemm_lds(S), ctx->stream>>>(A, X, K, C, out);
        LAUNCH_CHECK(ctx);
        return 0;
    }
};

template <int S>
struct AllowLds {
    template <int SPB>
    static void allow() {
        hipFuncSetAttribute((const void*)pair_bf16_kernel<MODE_RMP2, S, SPB>,
                            hipFuncAttributeMaxDynamicSharedMemorySize, (int)pair_lds(S, true));
        hipFuncSetAttribute((const void*)pair_bf16_kernel<MODE_SS, S, SPB>,
                            hipFuncAttributeMaxDynamicSharedMemorySize, (int)pair_lds(S, true));
        hipFuncSetAttribute((const void*)pair_bf16_kernel<MODE_OS, S, SPB>,
                            hipFuncAttributeMaxDynamicSharedMemorySize, (int)pair_lds(S, false));
    }
    static int run(mshqc_hip_ctx* ctx) {
        // Dynamic LDS above 64 KiB must be opted into per kernel.
        hipFuncSetAttribute((const void*)slice_pair_kernel<S, SPB_SHORT * KS>,
                            hipFuncAttributeMaxDynamicSharedMemorySize, (int)slice_pair_lds(SPB_SHORT * KS));
        hipFuncSetAttribute((const void*)slice_pair_kernel<S, SPB_LONG * KS>,
                            hipFuncAttributeMaxDynamicSharedMemorySize, (int)slice_pair_lds(SPB_LONG * KS));
        allow<SPB_SHORT>();
        if constexpr (spb_long_ok(S)) allow<SPB_LONG>();
        hipGetLastError();
        (void)ctx;
        return 0;
    }
};

template <int S>
struct AllowLdsGemm {
    static int run(mshqc_hip_ctx* ctx) {
        hipFuncSetAttribute((const void*)gemm_bf16_nt_kernel<S, GB_GRAD>,
                            hipFuncAttributeMaxDynamicSharedMemorySize, (int)gemm_lds(S));
        hipFuncSetAttribute((const void*)gemm_bf16_nt_kernel<S, GB_TRANSFORM>,
                            hipFuncAttributeMaxDynamicSharedMemorySize, (int)gemm_lds(S));
        hipGetLastError();
        (void)ctx;
        return 0;
    }
};

}  // namespace

int bf16_prepare(mshqc_hip_ctx* ctx, int slices, const double* Bl, int ol, int vl, const double* Br,
                 int orr, int vr, int Q, bool grad, int bo_l, int bo_r, bool self_terms, Bf16Operands* out) {
    const int S = slices;
    by_slices<AllowLds>(S, ctx);
    Bf16Operands o;
    o.slices = S;
    o.gslices = std::min(S, BF16_GRAD_SLICES);
    by_gslices<AllowLdsGemm>(o.gslices, ctx);
    o.Q = Q;
    o.spb = pair_stages_per_block(Q, S);
    const int QB = o.spb * KS;
    o.Qp = round_up(Q, QB);
    o.qblocks = o.Qp / QB;
    const long ovl = (long)ol * vl, ovr = Br ? (long)orr * vr : 0;
    // The pair kernel's buffer loads address one 16-Q stage block
    // ([ceil(o*v/32) + 1][S][32][16]) by 32-bit byte offsets.
    if ((long)S * (std::max(ovl, ovr) + 63) >= (1L << 27))
        return set_error(ctx, MSHQC_HIP_ERR_ARG,
                         "split-BF16: a stage block of %d slices x o*v = %ld rows exceeds the pair kernel's 4 GiB", S,
                         std::max(ovl, ovr));
    // Pair-kernel slices [Qp/16][ceil(o*v/32) + 1][S][32][16] and exponent tables
    // [Qp/QB][o][vE], vE = v rounded up to 128 (zero padding).
    o.vEl = round_up(vl, 128);
    o.vEr = Br ? round_up(vr, 128) : o.vEl;
    const size_t bl = (size_t)S * ((ovl + 31) / 32 * 32 + 32) * o.Qp * 2, br = (size_t)S * ((ovr + 31) / 32 * 32 + 32) * o.Qp * 2;
    const size_t el = (size_t)o.qblocks * ol * o.vEl, er = Br ? (size_t)o.qblocks * orr * o.vEr : 0;
    // self terms after the exponent tables, 256-byte aligned
    const int nso = self_terms ? (Br ? std::min(ol, orr) : ol) : 0;
    const int nsv = self_terms ? (Br ? std::min(vl, vr) : vl) : 0;
    const size_t gs_off = (bl + br + el + er + 255) / 256 * 256;
    const size_t gs_bytes = sizeof(double) * (size_t)nso * nsv;
    char* w = static_cast<char*>(workspace(ctx, WS_SLICE, gs_off + gs_bytes + 1024));
    if (!w) return MSHQC_HIP_ERR_ALLOC;
    o.sl = reinterpret_cast<uint16_t*>(w);
    o.sr = Br ? reinterpret_cast<uint16_t*>(w + bl) : nullptr;
    o.xl = reinterpret_cast<int8_t*>(w + bl + br);
    o.xr = Br ? reinterpret_cast<int8_t*>(w + bl + br + el) : nullptr;
    HIP_TRY(ctx, hipMemsetAsync(w + bl + br, 0, el + er, ctx->stream));
    if (nso > 0 && nsv > 0) {
        double* gs = reinterpret_cast<double*>(w + gs_off);
        const long n = (long)nso * nsv;
        self_terms_kernel<<<ceil_div(n, 256), 256, 0, ctx->stream>>>(Bl, Br ? Br : Bl, Q, ovl, Br ? ovr : ovl, vl,
                                                                   Br ? vr : vl, nso, nsv, gs);
        LAUNCH_CHECK(ctx);
        o.gself = gs;
        o.nself_o = nso;
        o.nself_v = nsv;
    }
    int rc = by_slices<SlicePair>(S, ctx, o.spb, Bl, Q, ol, vl, o.vEl, o.Qp, const_cast<uint16_t*>(o.sl),
                                  const_cast<int8_t*>(o.xl));
    if (rc) return rc;
    if (Br) {
        rc = by_slices<SlicePair>(S, ctx, o.spb, Br, Q, orr, vr, o.vEr, o.Qp, const_cast<uint16_t*>(o.sr),
                                  const_cast<int8_t*>(o.xr));
        if (rc) return rc;
    }
    if (grad) {
        // GEMM-operand slices of B per occupied block: rows q, the block's
        // columns (i, a) as one run, GB-wide exponent blocks along it.
        const int SG = o.gslices;
        o.bol = std::max(1, std::min(bo_l, ol));
        o.bor = Br ? std::max(1, std::min(bo_r, orr)) : o.bol;
        o.kwbl = grad_pitch(o.bol * vl);
        o.kwbr = Br ? grad_pitch(o.bor * vr) : o.kwbl;
        const int nbl = ceil_div(ol, o.bol), nbr = Br ? ceil_div(orr, o.bor) : 0;
        const size_t gl = (size_t)nbl * SG * Q * o.kwbl * 2, gr = (size_t)nbr * SG * Q * o.kwbr * 2;
        const long Qx = round_up(Q, GT);  // exponent tables: [kb][Qx] per block
        const size_t gel = (size_t)nbl * Qx * (o.kwbl / GB_GRAD), ger = (size_t)nbr * Qx * (o.kwbr / GB_GRAD);
        char* g = static_cast<char*>(workspace(ctx, WS_SLICE2, gl + gr + gel + ger + 1024));
        if (!g) return MSHQC_HIP_ERR_ALLOC;
        o.gl = reinterpret_cast<uint16_t*>(g);
        o.gr = Br ? reinterpret_cast<uint16_t*>(g + gl) : o.gl;
        o.gel = reinterpret_cast<int8_t*>(g + gl + gr);
        o.ger = Br ? reinterpret_cast<int8_t*>(g + gl + gr + gel) : o.gel;
        for (int side = 0; side < (Br ? 2 : 1); ++side) {
            const double* src = side ? Br : Bl;
            const int oo = side ? orr : ol, vv = side ? vr : vl, bo = side ? o.bor : o.bol;
            const long kwb = side ? o.kwbr : o.kwbl;
            uint16_t* dst = const_cast<uint16_t*>(side ? o.gr : o.gl);
            int8_t* ex = const_cast<int8_t*>(side ? o.ger : o.gel);
            for (int b = 0, i0 = 0; i0 < oo; ++b, i0 += bo) {
                const int nb = std::min(bo, oo - i0);
                rc = by_gslices<SliceRows>(SG, ctx, GB_GRAD, src + (long)i0 * vv, (long)oo * vv, Q, 1, nb * vv,
                                          grad_pitch(nb * vv), dst + (size_t)b * SG * Q * kwb,
                                          ex + (size_t)b * Qx * (kwb / GB_GRAD), Qx);
                if (rc) return rc;
            }
        }
    }
    o.ol = ol;
    o.vl = vl;
    o.orr = Br ? orr : ol;
    o.vr = Br ? vr : vl;
    *out = o;
    return MSHQC_HIP_OK;
}

int bf16_transform(mshqc_hip_ctx* ctx, const double* L, int Q, int n, const double* Cocc, int o,
                   const double* Cvir, int v, int S, double* B) {
    if (S < 2 || S > BF16_TRANSFORM_SLICES)
        return set_error(ctx, MSHQC_HIP_ERR_ARG, "split-BF16 transform: %d slices (2 to %d)", S,
                         BF16_TRANSFORM_SLICES);
    by_tslices<AllowLdsGemm>(S, ctx);
    auto al = [](size_t b) { return (b + 255) / 256 * 256; };
    // The half transform T_q = (L^q C_occ)^T runs in FP64: it reads FP64 L
    // once, where slicing L would read it and write its slices, then read
    // them again (three passes over the largest operand).  The virtual
    // product B_q = T_q C_vir runs on split-BF16 slices of T and C_vir, one
    // exponent per row and GB-wide k block, the AO index padded to whole blocks.
    const int GB = GB_TRANSFORM;
    const long kw = round_up(n, GB), nkb = kw / GB;
    const long vx = round_up(v, GT);
    // Q chunk: about 1.5 GiB of T and its slices at a time.
    const long per_q = (long)o * n * 8 + (long)S * o * kw * 2;
    const int qc = (int)std::max(1L, std::min((long)Q, (long)(1.5 * (1L << 30)) / per_q));
    const long rt = (long)o * qc, rtx = round_up(rt, GT);  // T rows (q, i) of a chunk
    // workspace: C_vir^T slices, then per chunk T in FP64 and its slices,
    // each slice set with its exponent table
    size_t off = 0;
    auto take = [&](size_t bytes) { size_t at = off; off += al(bytes); return at; };
    const size_t o_vs = take((size_t)S * v * kw * 2), o_vx = take((size_t)nkb * vx);
    const size_t o_t = take(sizeof(double) * (size_t)rt * n);
    const size_t o_ts = take((size_t)S * rt * kw * 2), o_tx = take((size_t)nkb * rtx);
    char* w = static_cast<char*>(workspace(ctx, WS_ZS, off + 256));
    if (!w) return MSHQC_HIP_ERR_ALLOC;
    auto u16 = [&](size_t at) { return reinterpret_cast<uint16_t*>(w + at); };
    auto i8 = [&](size_t at) { return reinterpret_cast<int8_t*>(w + at); };
    double* T = reinterpret_cast<double*>(w + o_t);
    // C_vir^T: rows a, k = mu (transposed slicing of the n x v block)
    int rc = by_tslices<SliceCols>(S, ctx, GB, Cvir, (long)v, 1, n, (int)kw, v, u16(o_vs), i8(o_vx), vx);
    if (rc) return rc;
    const SlicedMat Vt{u16(o_vs), (long)v * kw, kw, i8(o_vx), vx, v};
    const long nn = (long)n * n;
    for (int q0 = 0; q0 < Q; q0 += qc) {
        const int nq = std::min(qc, Q - q0);
        const long nt = (long)o * nq;
        // T[q][i][mu] = (L^q C_occ)[mu][i]: the FP64 half transforms of the
        // chunk as one tall GEMM storing each q's block transposed
        GemmF64 g1;
        g1.A = L + (long)q0 * nn;
        g1.lda = n;
        g1.sA = nn;
        g1.X = Cocc;
        g1.ldx = o;
        g1.C = T;
        g1.ldc = n;
        g1.sC = (long)o * n;
        g1.trans_c = true;
        g1.M = n;
        g1.N = o;
        g1.K = n;
        g1.batch = nq;
        {
            ProfScope ps(ctx, PROF_GEMM);
            if ((rc = gemm_f64(ctx, g1))) return rc;
        }
        // T rows (q, i) = row q*o + i, k = mu
        rc = by_tslices<SliceRows>(S, ctx, GB, T, (long)n, (int)nt, 1, n, (int)kw, u16(o_ts), i8(o_tx), rtx);
        if (rc) return rc;
        // B[q][i][a] = sum_mu T[(q, i)][mu] C_vir[mu][a]: GEMM row q*o + i is
        // row i of B_q
        double* Bq = B + (long)q0 * o * v;
        if ((rc = fill_zero(ctx, Bq, nt * v))) return rc;
        const SlicedMat Ts{u16(o_ts), nt * kw, kw, i8(o_tx), rtx, (int)nt};
        {
            ProfScope ps(ctx, PROF_GEMM);
            rc = by_tslices<GemmLaunch>(S, ctx, GB, Ts, Vt, (int)kw, Bq, (long)v, 0, 0L);
        }
        if (rc) return rc;
    }
    return MSHQC_HIP_OK;
}

int bf16_launch_pair(mshqc_hip_ctx* ctx, int mode, const PairArgs& a, const Bf16Operands& ops) {
    if (a.nwork <= 0) return MSHQC_HIP_OK;
    return by_slices<PairLaunch>(ops.slices, ctx, mode, a, ops);
}

int bf16_grad_contract(mshqc_hip_ctx* ctx, const Bf16Operands& o, const double* Z, long ldz, int I0,
                       int nI, int J0, int nJ, double* dBl, double* dBr, bool second) {
    const int S = o.gslices, Q = o.Q;
    const int vl = o.vl, vr = o.vr;
    if (I0 % o.bol != 0 || J0 % o.bor != 0 || nI > o.bol || nJ > o.bor)
        return set_error(ctx, MSHQC_HIP_ERR_ARG, "split-BF16 gradient: block (%d+%d, %d+%d) off the %d/%d grid",
                         I0, nI, J0, nJ, o.bol, o.bor);
    int rc;
    if (dBl) {
        // dBl[q][(i,a)] += sum_{(j,b)} Br[q][(j,b)] Z[(i,a)][(j,b)]
        const long R = (long)nI * vl, kw = grad_pitch(nJ * vr), Rx = round_up(R, GT), Qx = round_up(Q, GT);
        const size_t zb = (size_t)S * R * kw * 2, ze = (size_t)Rx * (kw / GB_GRAD);
        char* w = static_cast<char*>(workspace(ctx, WS_ZS, zb + ze + 256));
        if (!w) return MSHQC_HIP_ERR_ALLOC;
        uint16_t* zs = reinterpret_cast<uint16_t*>(w);
        int8_t* zx = reinterpret_cast<int8_t*>(w + zb);
        rc = by_gslices<SliceRows>(S, ctx, GB_GRAD, Z, ldz, (int)R, 1, nJ * vr, (int)kw, zs, zx, Rx);
        if (rc) return rc;
        const long b = J0 / o.bor;
        SlicedMat A{o.gr + (size_t)b * S * Q * o.kwbr, (long)Q * kw, kw, o.ger + (size_t)b * Qx * (o.kwbr / GB_GRAD),
                    Qx, Q};
        SlicedMat X{zs, (long)R * kw, kw, zx, Rx, (int)R};
        rc = by_gslices<GemmLaunch>(S, ctx, GB_GRAD, A, X, (int)kw, dBl + (long)I0 * vl, (long)o.ol * vl, 0, 0L);
        if (rc) return rc;
    }
    if (second && dBr) {
        // dBr[q][(j,b)] += sum_{(i,a)} Bl[q][(i,a)] Z[(i,a)][(j,b)]
        const long R = (long)nJ * vr, kw = grad_pitch(nI * vl), Rx = round_up(R, GT), Qx = round_up(Q, GT);
        const size_t zb = (size_t)S * R * kw * 2, ze = (size_t)Rx * (kw / GB_GRAD);
        char* w = static_cast<char*>(workspace(ctx, WS_ZTS, zb + ze + 256));
        if (!w) return MSHQC_HIP_ERR_ALLOC;
        uint16_t* zs = reinterpret_cast<uint16_t*>(w);
        int8_t* zx = reinterpret_cast<int8_t*>(w + zb);
        rc = by_gslices<SliceCols>(S, ctx, GB_GRAD, Z, ldz, 1, nI * vl, (int)kw, (int)R, zs, zx, Rx);
        if (rc) return rc;
        const long b = I0 / o.bol;
        SlicedMat A{o.gl + (size_t)b * S * Q * o.kwbl, (long)Q * kw, kw, o.gel + (size_t)b * Qx * (o.kwbl / GB_GRAD),
                    Qx, Q};
        SlicedMat X{zs, (long)R * kw, kw, zx, Rx, (int)R};
        rc = by_gslices<GemmLaunch>(S, ctx, GB_GRAD, A, X, (int)kw, dBr + (long)J0 * vr, (long)o.orr * vr, 0, 0L);
        if (rc) return rc;
    }
    return MSHQC_HIP_OK;
}
