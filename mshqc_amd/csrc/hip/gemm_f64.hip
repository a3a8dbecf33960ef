// FP64 MFMA GEMM engine (gfx950, v_mfma_f64_16x16x4_f64) and deterministic
// reductions.  Used by the J/K build, the AO->MO transforms, the gradient
// contractions of the MP2 backward, the OMP2 and MP3 contractions.
//
// Two kernels:
//   gemm_f64_dma_kernel  operand panels streamed straight from memory into
//                        an LDS ring by direct-to-LDS loads, three stages
//                        ahead; 128 x 128 tiles of eight 64 x 32 waves (two
//                        per SIMD) or 64 x 64 tiles of four 32 x 32 waves.
//                        Needs 16-byte aligned operands whose 16-byte
//                        segments stay inside one row (see dma_eligible).
//   gemm_f64_kernel      64 x 64 tiles, panels staged through registers with
//                        8-byte loads; any alignment and shape.
#include <algorithm>

#include "common.h"
#include "dma.h"
#include "mfma_f64.h"
#include "tiles.h"

namespace {

constexpr int TM = 64, TN = 64, KC = 32, NT = 256;
constexpr int PR = TM * KC / NT;  // panel elements per thread and step

// Panel loader: S[kk][c] = M_b(row0 + c, k0 + kk) for a matrix whose element
// (r, k) sits at base[r * ld + k] (KCONTIG) or base[k * ld + r] (!KCONTIG).
template <bool KCONTIG>
struct Panel {
    static __device__ __forceinline__ void load(const double* __restrict__ base, long ld, int rows,
                                                int kmax, int row0, int k0, int tid, double (&reg)[PR]) {
#pragma unroll
        for (int r = 0; r < PR; ++r) {
            int kk, c;
            if (KCONTIG) {
                kk = tid % KC;
                c = tid / KC + (NT / KC) * r;
            } else {
                c = tid & 63;
                kk = (tid >> 6) + 4 * r;
            }
            const int gr = row0 + c, gk = k0 + kk;
            double x = 0.0;
            if (gr < rows && gk < kmax)
                x = KCONTIG ? base[(long)gr * ld + gk] : base[(long)gk * ld + gr];
            reg[r] = x;
        }
    }
    static __device__ __forceinline__ void store(double (*S)[TM + 1], int tid, const double (&reg)[PR]) {
#pragma unroll
        for (int r = 0; r < PR; ++r) {
            int kk, c;
            if (KCONTIG) {
                kk = tid % KC;
                c = tid / KC + (NT / KC) * r;
            } else {
                c = tid & 63;
                kk = (tid >> 6) + 4 * r;
            }
            S[kk][c] = reg[r];
        }
    }
};

// A_b(m,k): transA -> A[k*lda + m] (m contiguous)  else A[m*lda + k] (k contiguous)
// X_b(k,n) as a panel over n: transX -> X[n*ldx + k] (k contiguous) else X[k*ldx + n]
template <bool TA, bool TX>
__global__ __launch_bounds__(NT) void gemm_f64_kernel(GemmF64 g, double* __restrict__ part, int bper) {
    __shared__ double As[KC][TM + 1];
    __shared__ double Xs[KC][TN + 1];
    int mt, ntile;
    if (!grouped_tile((g.M + TM - 1) / TM, (g.N + TN - 1) / TN, mt, ntile)) return;
    if (g.lower_only && ntile > mt) return;
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int wm = (wave >> 1) * 32, wn = (wave & 1) * 32;
    const int m0 = mt * TM, n0 = ntile * TN;
    int b_lo, b_hi;
    if (g.reduce_batch) {
        b_lo = blockIdx.z * bper;
        b_hi = min(g.batch, b_lo + bper);
    } else {
        b_lo = blockIdx.z;
        b_hi = b_lo + 1;
    }
    v4d acc[2][2];
#pragma unroll
    for (int s = 0; s < 2; ++s)
#pragma unroll
        for (int t = 0; t < 2; ++t) acc[s][t] = v4d{0.0, 0.0, 0.0, 0.0};

    const int nk = (g.K + KC - 1) / KC;
    const long total = (long)(b_hi - b_lo) * nk;
    double ra[PR], rx[PR];
    auto fetch = [&](long step) {
        const int b = b_lo + (int)(step / nk);
        const int k0 = (int)(step % nk) * KC;
        const int bo = b / g.inner, bi = b - bo * g.inner;
        const double* Ab = g.A + (long)bo * g.sA + (long)bi * g.sA2;
        const double* Xb = g.X + (long)bo * g.sX + (long)bi * g.sX2;
        Panel<!TA>::load(Ab, g.lda, g.M, g.K, m0, k0, tid, ra);
        Panel<TX>::load(Xb, g.ldx, g.N, g.K, n0, k0, tid, rx);
    };
    if (total > 0) fetch(0);
    for (long step = 0; step < total; ++step) {
        __syncthreads();
        Panel<!TA>::store(As, tid, ra);
        Panel<TX>::store(Xs, tid, rx);
        __syncthreads();
        if (step + 1 < total) fetch(step + 1);
#pragma unroll
        for (int k4 = 0; k4 < KC / 4; ++k4) {
            const int kr = k4 * 4 + (lane >> 4);
            double a[2], x[2];
#pragma unroll
            for (int s = 0; s < 2; ++s) a[s] = As[kr][wm + 16 * s + (lane & 15)];
#pragma unroll
            for (int t = 0; t < 2; ++t) x[t] = Xs[kr][wn + 16 * t + (lane & 15)];
#pragma unroll
            for (int s = 0; s < 2; ++s)
#pragma unroll
                for (int t = 0; t < 2; ++t) acc[s][t] = mfma_f64(a[s], x[t], acc[s][t]);
        }
    }
    // Epilogue.
    double* Cb;
    long ldc;
    bool direct = !g.reduce_batch || gridDim.z == 1;
    if (direct) {
        Cb = g.C + (g.reduce_batch ? 0 : (long)blockIdx.z * g.sC);
        ldc = g.ldc;
    } else {
        Cb = part + (long)blockIdx.z * g.M * g.N;
        ldc = g.N;
    }
#pragma unroll
    for (int s = 0; s < 2; ++s)
#pragma unroll
        for (int t = 0; t < 2; ++t)
#pragma unroll
            for (int r = 0; r < 4; ++r) {
                const int row = m0 + wm + 16 * s + mfma_row(lane, r);
                const int col = n0 + wn + 16 * t + mfma_col(lane);
                if (row < g.M && col < g.N) {
                    double* p = g.trans_c ? Cb + (long)(row / g.tblock) * g.sC + (long)col * ldc + row % g.tblock
                                          : Cb + (long)row * ldc + col;
                    if (direct) {
                        double val = g.alpha * acc[s][t][r];
                        *p = g.accumulate ? *p + val : val;
                    } else {
                        *p = acc[s][t][r];
                    }
                }
            }
}

// ------------------------------------------------------------ DMA engine
//
// A stage holds DK (8 or 16) values of the contraction index for every panel
// row of both operands.  Panels live in LDS as 16-byte segments (two
// doubles), written by 1-KiB direct-to-LDS loads (lane l's segment lands at
// 16 l; the swizzles below are applied by choosing which segment each lane
// fetches):
//   layout K  (operand stored k-contiguous; panel rows = its m or n index,
//              SG = DK/2 segments per row): segment (r, kp), kp = k/2 < SG,
//              at 16 * (r*SG + (kp ^ ((r / (16/SG)) & (SG-1))));
//   layout MN (operand stored m- or n-contiguous; panel rows = k, R*8 bytes
//              per row): segment (k, mp), mp = m/2 < R/2, at
//              k*R*8 + 16 * (mp ^ (8 * (k & 1))).
// A 16-lane quarter of a fragment read then touches 16 distinct 4-bank groups
// (all 64 banks) in either layout.
//
// Each 16-byte read feeds two MFMAs.  The contraction index a lane supplies is
// free as long as both operands agree: in k step t (0, 1) of the 8-wide group
// u, lane quarter g supplies k = 8u + 2g + t.  A layout-K read of (row, kp =
// 4u + g) gives that row's two k; a layout-MN read of (k row 8u + 2g + t,
// pair mp) gives two adjacent m (n), which feed MFMA tiles 2s and 2s+1: tile
// 2s holds rows m = 32s + 2 rho, tile 2s+1 rows 32s + 2 rho + 1.
template <bool TA, bool TX, int BM, int BN, int WM, int WN, int NBUF, int DK>
struct DmaCfg {
    static_assert(DK == 8 || DK == 16, "stage depth");
    static constexpr int SG = DK / 2;                 // segments per layout-K panel row
    static constexpr int NW = (BM / WM) * (BN / WN);  // waves
    static constexpr int NTH = NW * 64;
    static constexpr bool AK = !TA;                   // A stored k-contiguous
    static constexpr bool XK = TX;                    // X stored k-contiguous
    static constexpr int PA = BM * DK * 8, PX = BN * DK * 8;    // panel bytes per stage
    static constexpr int STAGE = PA + PX;
    static constexpr int LA = PA / 1024, LX = PX / 1024;        // 1-KiB loads per stage
    static constexpr int PER = (LA + LX) / NW;                  // loads per wave and stage
    static constexpr int AHEAD = NBUF - 1;                      // stages loaded ahead
    static constexpr int WAITN = (AHEAD - 1) * PER;             // loads left in flight at a stage barrier
    static constexpr size_t LDS = (size_t)NBUF * STAGE;
    static_assert((LA + LX) % NW == 0, "every wave issues the same loads");
    static_assert(WM % 32 == 0 && WN % 32 == 0, "layout-MN fragments pair 16-row tiles");
    static_assert(BM % 8 == 0 && BN % 8 == 0 && BM >= 32 && BN >= 32, "panel shape");
};

// Byte offset of the segment holding panel row r (layout K, SG segments per
// row) at k pair kp.
template <int SG>
__device__ __forceinline__ int seg_k(int r, int kp) {
    return (r * SG + (kp ^ ((r / (16 / SG)) & (SG - 1)))) << 4;
}
// Byte offset of the segment holding k row k, pair mp of an R-row layout-MN panel.
template <int R>
__device__ __forceinline__ int seg_mn(int k, int mp) { return k * (R * 8) + ((mp ^ ((k & 1) << 3)) << 4); }

typedef double v2d __attribute__((ext_vector_type(2)));

template <bool TA, bool TX, int BM, int BN, int WM, int WN, int NBUF, int DK>
__global__ __launch_bounds__((BM / WM) * (BN / WN) * 64) void gemm_f64_dma_kernel(
    GemmF64 g, double* __restrict__ part, long fchunk, int zsplit) {
    using C = DmaCfg<TA, TX, BM, BN, WM, WN, NBUF, DK>;
    constexpr int SG = C::SG;
    constexpr int MS = WM / 16, NS = WN / 16;  // MFMA tiles per wave
    extern __shared__ __align__(16) char lds[];
    int mt, ntile;
    if (!grouped_tile((g.M + BM - 1) / BM, (g.N + BN - 1) / BN, mt, ntile)) return;
    if (g.lower_only && ntile > mt) return;
    const int tid = threadIdx.x, lane = tid & 63;
    const int wv = __builtin_amdgcn_readfirstlane(tid >> 6);
    const int wm = (wv / (BN / WN)) * WM, wn = (wv % (BN / WN)) * WN;
    const int m0 = mt * BM, n0 = ntile * BN;
    // The contraction runs over a range [f0, f1) of the flattened index
    // f = b * K + k: with zsplit, workgroup layer z takes the z-th chunk of
    // fchunk values of all batch entries (its partial output is summed by
    // the caller); without, layer z is batch entry z.  DK values per stage;
    // only the last stage can be partial.
    const int K = g.K;
    long f0, f1;
    if (zsplit) {
        f0 = (long)blockIdx.z * fchunk;
        f1 = min((long)g.batch * K, f0 + fchunk);
    } else {
        f0 = (long)blockIdx.z * K;
        f1 = f0 + K;
    }
    const int b_lo = (int)(f0 / K), k_lo = (int)(f0 - (long)b_lo * K);
    const int nb = (int)((f1 - 1) / K) - b_lo + 1;  // entries touched
    const long flat = f1 - f0;
    const int total = (int)((flat + DK - 1) / DK);  // stages

    // This wave's loads: u = wv*PER + k < LA are A loads, the rest X loads.
    // Per lane and load: a pointer to its segment's row (layout K) or pair
    // (layout MN) in the current batch entry, the segment's k in that entry
    // (kk), the entries passed (bc) and the inner index (bi).  Rows past the
    // matrix are clamped to its last row (layout K) or last pair (layout
    // MN); those LDS rows only feed discarded outputs.  K is even for a
    // layout-K operand, so a segment's two k never straddle two entries.
    const double* ptr[C::PER];
    int kk[C::PER], bc[C::PER], bi[C::PER], dst[C::PER];
    bool isA[C::PER], kl[C::PER];
    const int bo0 = b_lo / g.inner, bi0 = b_lo - bo0 * g.inner;
#pragma unroll
    for (int k = 0; k < C::PER; ++k) {
        const int u = wv * C::PER + k;
        const bool a = u < C::LA;
        const int uu = a ? u : u - C::LA;
        const bool klay = a ? C::AK : C::XK;
        const int R = a ? BM : BN, row0 = a ? m0 : n0, rows = a ? g.M : g.N;
        const long ld = a ? g.lda : g.ldx;
        isA[k] = a;
        kl[k] = klay;
        dst[k] = (a ? 0 : C::PA) + uu * 1024;
        const double* base = a ? g.A + (long)bo0 * g.sA + (long)bi0 * g.sA2
                               : g.X + (long)bo0 * g.sX + (long)bi0 * g.sX2;
        if (klay) {
            const int r = (64 / SG) * uu + lane / SG;
            const int kp = (lane % SG) ^ ((r / (16 / SG)) & (SG - 1));
            ptr[k] = base + (long)min(row0 + r, rows - 1) * ld;
            kk[k] = k_lo + 2 * kp;
        } else {
            const int idx = uu * 64 + lane, segr = R / 2;
            const int kr = idx / segr, p = idx % segr;
            const int mp = p ^ ((kr & 1) << 3);
            ptr[k] = base + min(row0 + 2 * mp, rows - 2);
            kk[k] = k_lo + kr;
        }
        bc[k] = 0;
        bi[k] = bi0;
    }
    // Moves load k's position to the next batch entry while it lies past K
    // (never past the last entry: a position beyond the workgroup's range
    // stays in its last entry, clamped below, and lands only in entries the
    // last stage's zeroing covers).
    auto settle = [&](int k) {
        const long s2 = isA[k] ? g.sA2 : g.sX2, s1 = isA[k] ? g.sA : g.sX;
        while (kk[k] >= K && bc[k] + 1 < nb) {
            kk[k] -= K;
            ++bc[k];
            if (++bi[k] == g.inner) {
                bi[k] = 0;
                ptr[k] += s1 - (long)(g.inner - 1) * s2;
            } else {
                ptr[k] += s2;
            }
        }
    };
#pragma unroll
    for (int k = 0; k < C::PER; ++k) settle(k);
    // Issue the next stage (stages are issued in order; past the end the
    // positions stay clamped in the last entry, loaded into slots no one
    // reads, so every wave issues exactly PER loads per stage on every path).
    int st_issue = 0;
    auto issue = [&]() {
        char* slot = lds + (st_issue % NBUF) * C::STAGE;
        ++st_issue;
#pragma unroll
        for (int k = 0; k < C::PER; ++k) {
            const long ld = isA[k] ? g.lda : g.ldx;
            const double* src = kl[k] ? ptr[k] + min(kk[k], K - 2) : ptr[k] + (long)min(kk[k], K - 1) * ld;
            dma16(src, slot + dst[k]);
            kk[k] += DK;
            settle(k);
        }
    };
    // Zero the k >= kv entries of both panels of a slot (the last, partial
    // stage), so neither operand's clamped values reach the sums.
    auto zero_tail = [&](char* slot, int kv) {
        for (int e = tid; e < (BM + BN) * DK; e += C::NTH) {
            const bool a = e < BM * DK;
            const int ee = a ? e : e - BM * DK;
            char* pan = slot + (a ? 0 : C::PA);
            const bool klay = a ? C::AK : C::XK;
            int k, boff;
            if (klay) {
                const int r = ee / DK;
                k = ee % DK;
                boff = seg_k<SG>(r, k >> 1) + 8 * (k & 1);
            } else {
                const int R = a ? BM : BN;
                k = ee / R;
                const int m = ee % R;
                boff = (a ? seg_mn<BM>(k, m >> 1) : seg_mn<BN>(k, m >> 1)) + 8 * (m & 1);
            }
            if (k >= kv) *reinterpret_cast<double*>(pan + boff) = 0.0;
        }
    };

    v4d acc[MS][NS];
#pragma unroll
    for (int s = 0; s < MS; ++s)
#pragma unroll
        for (int c = 0; c < NS; ++c) acc[s][c] = v4d{0.0, 0.0, 0.0, 0.0};
    const int q4 = lane >> 4, rho = lane & 15;

    if (total > 0) {
#pragma unroll
        for (int a = 0; a < C::AHEAD; ++a) issue();
        dma_barrier<C::WAITN>();  // stage 0 has landed
    }
    for (int st = 0; st < total; ++st) {
        char* pa = lds + (st % NBUF) * C::STAGE;
        const char* px = pa + C::PA;
        if (st == total - 1) {
            const int kv = (int)(flat - (long)st * DK);
            if (kv < DK) {
                zero_tail(pa, kv);
                dma_barrier<C::WAITN>();  // zeros written before anyone reads
            }
        }
        issue();
#pragma unroll
        for (int u = 0; u < DK / 8; ++u) {
            double af[MS][2], xf[NS][2];
            if constexpr (C::AK) {
#pragma unroll
                for (int s = 0; s < MS; ++s) {
                    const v2d v = *reinterpret_cast<const v2d*>(pa + seg_k<SG>(wm + 16 * s + rho, 4 * u + q4));
                    af[s][0] = v.x;
                    af[s][1] = v.y;
                }
            } else {
#pragma unroll
                for (int t = 0; t < 2; ++t)
#pragma unroll
                    for (int s2 = 0; s2 < MS / 2; ++s2) {
                        const v2d v = *reinterpret_cast<const v2d*>(
                            pa + seg_mn<BM>(8 * u + 2 * q4 + t, wm / 2 + 16 * s2 + rho));
                        af[2 * s2][t] = v.x;
                        af[2 * s2 + 1][t] = v.y;
                    }
            }
            if constexpr (C::XK) {
#pragma unroll
                for (int c = 0; c < NS; ++c) {
                    const v2d v = *reinterpret_cast<const v2d*>(px + seg_k<SG>(wn + 16 * c + rho, 4 * u + q4));
                    xf[c][0] = v.x;
                    xf[c][1] = v.y;
                }
            } else {
#pragma unroll
                for (int t = 0; t < 2; ++t)
#pragma unroll
                    for (int c2 = 0; c2 < NS / 2; ++c2) {
                        const v2d v = *reinterpret_cast<const v2d*>(
                            px + seg_mn<BN>(8 * u + 2 * q4 + t, wn / 2 + 16 * c2 + rho));
                        xf[2 * c2][t] = v.x;
                        xf[2 * c2 + 1][t] = v.y;
                    }
            }
#pragma unroll
            for (int t = 0; t < 2; ++t)
#pragma unroll
                for (int s = 0; s < MS; ++s)
#pragma unroll
                    for (int c = 0; c < NS; ++c) acc[s][c] = mfma_f64(af[s][t], xf[c][t], acc[s][c]);
        }
        dma_barrier<C::WAITN>();  // stage st+1 has landed
    }
    // The clamped loads past the last stage land before the LDS is handed to
    // the next workgroup.
    dma_barrier<0>();

    double* Cb;
    long ldc;
    const bool direct = !zsplit || gridDim.z == 1;
    if (direct) {
        Cb = g.C + (zsplit ? 0 : (long)blockIdx.z * g.sC);
        ldc = g.ldc;
    } else {
        Cb = part + (long)blockIdx.z * g.M * g.N;
        ldc = g.N;
    }
#pragma unroll
    for (int s = 0; s < MS; ++s)
#pragma unroll
        for (int c = 0; c < NS; ++c)
#pragma unroll
            for (int r = 0; r < 4; ++r) {
                const int rr = mfma_row(lane, r), cc = mfma_col(lane);
                const int row = m0 + wm + (C::AK ? 16 * s + rr : 32 * (s >> 1) + 2 * rr + (s & 1));
                const int col = n0 + wn + (C::XK ? 16 * c + cc : 32 * (c >> 1) + 2 * cc + (c & 1));
                if (row < g.M && col < g.N) {
                    double* p = g.trans_c ? Cb + (long)(row / g.tblock) * g.sC + (long)col * ldc + row % g.tblock
                                          : Cb + (long)row * ldc + col;
                    if (direct) {
                        const double val = g.alpha * acc[s][c][r];
                        *p = g.accumulate ? *p + val : val;
                    } else {
                        *p = acc[s][c][r];
                    }
                }
            }
}

__global__ void split_reduce_kernel(const double* __restrict__ part, int splits, int M, int N,
                                    double* __restrict__ C, long ldc, double alpha, int accumulate,
                                    int lower_only) {
    const long idx = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (idx >= (long)M * N) return;
    const int m = (int)(idx / N), n = (int)(idx % N);
    if (lower_only && n / TN > m / TM) return;  // a tile no split formed
    double s = 0.0;
    for (int z = 0; z < splits; ++z) s += part[(long)z * M * N + idx];
    double* p = C + (long)m * ldc + n;
    *p = accumulate ? *p + alpha * s : alpha * s;
}

__global__ __launch_bounds__(1024) void reduce_sum_kernel(const double* __restrict__ x, long n,
                                                          double* __restrict__ out, int accumulate) {
    __shared__ double sh[1024];
    double s = 0.0, c = 0.0;
    for (long i = threadIdx.x; i < n; i += 1024) {
        const double v = x[i], t = s + v;
        c += (fabs(s) >= fabs(v)) ? (s - t) + v : (v - t) + s;
        s = t;
    }
    sh[threadIdx.x] = s + c;
    __syncthreads();
    for (int w = 512; w > 0; w >>= 1) {
        if (threadIdx.x < w) sh[threadIdx.x] += sh[threadIdx.x + w];
        __syncthreads();
    }
    if (threadIdx.x == 0) *out = accumulate ? *out + sh[0] : sh[0];
}

// The DMA kernel reads 16-byte segments: two k of one row (layout K) or two
// m (n) of one k row (layout MN).  Every segment must be 16-byte aligned and
// lie inside its row: aligned bases, even leading dimensions and batch
// strides, an even K for a layout-K operand and an even M (N) for a
// layout-MN one.
bool dma_eligible(const GemmF64& g) {
    auto al16 = [](const void* p) { return (reinterpret_cast<uintptr_t>(p) & 15) == 0; };
    if (!al16(g.A) || !al16(g.X)) return false;
    if ((g.lda | g.sA | g.sA2 | g.ldx | g.sX | g.sX2) & 1) return false;
    const bool aK = !g.transA, xK = g.transX;
    if (aK ? (g.K & 1) : (g.M & 1)) return false;
    if (xK ? (g.K & 1) : (g.N & 1)) return false;
    return true;
}

template <int BM, int BN, int WM, int WN, int NBUF, int DK, bool TA, bool TX>
int launch_dma_t(mshqc_hip_ctx* ctx, const GemmF64& g, double* part, long fchunk, int zsplit, unsigned gz) {
    using C = DmaCfg<TA, TX, BM, BN, WM, WN, NBUF, DK>;
    auto* fn = gemm_f64_dma_kernel<TA, TX, BM, BN, WM, WN, NBUF, DK>;
    // dynamic LDS above 64 KiB must be opted into per kernel
    hipFuncSetAttribute((const void*)fn, hipFuncAttributeMaxDynamicSharedMemorySize, (int)C::LDS);
    dim3 grid(grouped_grid((g.M + BM - 1) / BM, (g.N + BN - 1) / BN), 1, gz);
    fn<<<grid, C::NTH, C::LDS, ctx->stream>>>(g, part, fchunk, zsplit);
    LAUNCH_CHECK(ctx);
    return MSHQC_HIP_OK;
}

template <int BM, int BN, int WM, int WN, int NBUF, int DK>
int launch_dma(mshqc_hip_ctx* ctx, const GemmF64& g, double* part, long fchunk, int zsplit, unsigned gz) {
    if (g.transA)
        return g.transX ? launch_dma_t<BM, BN, WM, WN, NBUF, DK, true, true>(ctx, g, part, fchunk, zsplit, gz)
                        : launch_dma_t<BM, BN, WM, WN, NBUF, DK, true, false>(ctx, g, part, fchunk, zsplit, gz);
    return g.transX ? launch_dma_t<BM, BN, WM, WN, NBUF, DK, false, true>(ctx, g, part, fchunk, zsplit, gz)
                    : launch_dma_t<BM, BN, WM, WN, NBUF, DK, false, false>(ctx, g, part, fchunk, zsplit, gz);
}

}  // namespace

// Kernel choice of the DMA engine for one GEMM.
enum DmaMode { DMA_NONE, DMA_64, DMA_128_DEEP, DMA_128, DMA_SKINNY };

// Rows of the tall tiles for outputs at most 64 wide (eight 64 x 32 waves,
// 120 KiB of LDS, one workgroup per CU).
constexpr int SKINNY_BM = 256;

int gemm_f64(mshqc_hip_ctx* ctx, const GemmF64& g0) {
    if (g0.M <= 0 || g0.N <= 0 || g0.batch <= 0) return MSHQC_HIP_OK;
    if (g0.K <= 0) return set_error(ctx, MSHQC_HIP_ERR_ARG, "gemm_f64: K must be positive");
    if (g0.inner < 1 || (g0.inner > 1 && !g0.reduce_batch))
        return set_error(ctx, MSHQC_HIP_ERR_ARG, "gemm_f64: two-level batches need reduce_batch");
    if (g0.trans_c && (g0.reduce_batch || g0.lower_only))
        return set_error(ctx, MSHQC_HIP_ERR_ARG, "gemm_f64: a transposed C needs a plain batch");
    // A plain batch whose entries' A (k-contiguous) and C rows follow one
    // another, with one shared X, is one GEMM over all the entries' rows:
    // no per-entry padding of M to the tile.  A transposed C keeps its entry
    // stride: row r of the flat GEMM is row r % M of entry r / M.
    GemmF64 g = g0;
    g.tblock = g0.M;
    if (!g.reduce_batch && g.batch > 1 && g.inner == 1 && !g.transA && !g.lower_only && g.sX == 0 &&
        g.sA == (long)g.M * g.lda && (g.trans_c || g.sC == (long)g.M * g.ldc) &&
        (long)g.M * g.batch <= 0x7fffffffL) {
        g.M *= g.batch;
        g.batch = 1;
        g.sA = 0;
        if (!g.trans_c) g.sC = 0;
    }
    // Kernel choice (timed on gfx950 across the J/K, transform, MP2, OMP2 and
    // MP3 shapes):
    //   - the register-staged kernel where the DMA alignment rules do not hold;
    //   - 128 x 128 DMA tiles (two workgroups per CU) where M and N are both
    //     512 or more, or both 256 or more and the tiles of all plain batch
    //     entries fill one round of the chip; 16-deep stages one ahead, or
    //     8-deep stages three ahead when A is k-contiguous, X n-contiguous
    //     and K at most 1024 (there the deeper prefetch of the short
    //     contractions wins over the longer rows of 16-deep ones);
    //   - 256 x 64 DMA tiles (one workgroup per CU) for outputs at most 64
    //     wide over at least 64 tiles of rows (the occupied-orbital
    //     half-transforms L^q C: X is read once per 256 rows, not per 64);
    //   - 64 x 64 DMA tiles (three workgroups per CU) otherwise.
    int mode = DMA_NONE;
    if (dma_eligible(g)) {
        const long t128 = (long)((g.M + 127) / 128) * ((g.N + 127) / 128) * (g.reduce_batch ? 1 : g.batch);
        const bool big = (g.M >= 512 && g.N >= 512) || (g.M >= 256 && g.N >= 256 && t128 >= 512);
        if (!big) mode = DMA_64;
        else mode = (!g.transA && !g.transX && g.K <= 1024) ? DMA_128_DEEP : DMA_128;
        // tall and skinny (N <= 64): tall tiles read the shared X panel once
        // per SKINNY_BM rows instead of once per 64
        const long rows = (long)g.M * (g.reduce_batch ? 1 : g.batch);
        if (g.N <= 64 && rows >= 64L * SKINNY_BM) mode = DMA_SKINNY;
        // up to 128 wide over many rows: 128 x 128 tiles read each A row once
        // (64 x 64 tiles read it once per 64 columns)
        else if (g.N <= 128 && rows >= 64L * 128 && mode == DMA_64)
            mode = (!g.transA && !g.transX && g.K <= 1024) ? DMA_128_DEEP : DMA_128;
    }
    const bool big = mode == DMA_128 || mode == DMA_128_DEEP;
    // Work split over workgroup layers (z): batch entries for a plain batched
    // GEMM; for a batch-reduced GEMM (and, in the DMA kernels, a single GEMM
    // with too few tiles to fill the chip) chunks of the contraction whose
    // partial outputs are then summed in layer order (deterministic).
    int splits = 1, bper = 1;
    long fchunk = 0;
    bool zsplit = g.reduce_batch;
    if (mode == DMA_NONE) {
        if (g.reduce_batch) {
            splits = (int)std::max(1, std::min(g.splits, g.batch));
            bper = (g.batch + splits - 1) / splits;
            splits = (g.batch + bper - 1) / bper;
        }
    } else if (g.reduce_batch || g.batch == 1) {
        // enough layers for about four rounds of workgroups over 256 CUs,
        // each layer at least 32 stages of 16
        const int TMm = mode == DMA_SKINNY ? SKINNY_BM : big ? 128 : 64, TNn = mode == DMA_SKINNY ? 64 : TMm;
        const long tm = (g.M + TMm - 1) / TMm, tn = (g.N + TNn - 1) / TNn;
        const long tiles = g.lower_only ? tm * (tm + 1) / 2 : tm * tn;
        const long slots = 256L * (mode == DMA_SKINNY ? (SKINNY_BM > 128 ? 1 : 2) : big ? 2 : 3);
        const long F = (long)(g.reduce_batch ? g.batch : 1) * g.K;
        // (at most four whole rounds: rounding up left a few workgroups
        // alone in a fifth round, e.g. 15 x 137 = 4 x 512 + 7)
        const long want = std::max(1L, std::min(std::max(1L, 4 * slots / tiles), F / (16 * 32)));
        if (g.reduce_batch || (want > 1 && !g.trans_c)) {
            zsplit = true;
            fchunk = (F + want - 1) / want;
            fchunk = (fchunk + 15) / 16 * 16;  // whole stages of 16 (even, as layout-K pairs need)
            splits = (int)((F + fchunk - 1) / fchunk);
        }
    }
    double* part = nullptr;
    if (zsplit && splits > 1) {
        part = static_cast<double*>(workspace(ctx, WS_TMP2, sizeof(double) * (size_t)splits * g.M * g.N));
        if (!part) return MSHQC_HIP_ERR_ALLOC;
    }
    const unsigned gz = zsplit ? splits : g.batch;
    if (gz > 65535u) return set_error(ctx, MSHQC_HIP_ERR_ARG, "gemm_f64: batch too large");
    if (zsplit && fchunk == 0) fchunk = (long)g.batch * g.K;  // one layer
    int rc = MSHQC_HIP_OK;
    const int zs = zsplit ? 1 : 0;
    if (mode == DMA_128_DEEP) {
        rc = launch_dma<128, 128, 64, 32, 4, 8>(ctx, g, part, fchunk, zs, gz);
    } else if (mode == DMA_128) {
        rc = launch_dma<128, 128, 64, 32, 2, 16>(ctx, g, part, fchunk, zs, gz);
    } else if (mode == DMA_64) {
        rc = launch_dma<64, 64, 32, 32, 3, 16>(ctx, g, part, fchunk, zs, gz);
    } else if (mode == DMA_SKINNY) {
        rc = launch_dma<SKINNY_BM, 64, 64, 32, 3, 16>(ctx, g, part, fchunk, zs, gz);
    } else {
        const unsigned gx = grouped_grid((g.M + TM - 1) / TM, (g.N + TN - 1) / TN);
        dim3 grid(gx, 1, gz);
        if (g.transA) {
            if (g.transX) gemm_f64_kernel<true, true><<<grid, NT, 0, ctx->stream>>>(g, part, bper);
            else gemm_f64_kernel<true, false><<<grid, NT, 0, ctx->stream>>>(g, part, bper);
        } else {
            if (g.transX) gemm_f64_kernel<false, true><<<grid, NT, 0, ctx->stream>>>(g, part, bper);
            else gemm_f64_kernel<false, false><<<grid, NT, 0, ctx->stream>>>(g, part, bper);
        }
        LAUNCH_CHECK(ctx);
    }
    if (rc) return rc;
    if (part) {
        const long tot = (long)g.M * g.N;
        split_reduce_kernel<<<(unsigned)((tot + 255) / 256), 256, 0, ctx->stream>>>(
            part, splits, g.M, g.N, g.C, g.ldc, g.alpha, g.accumulate ? 1 : 0, g.lower_only ? 1 : 0);
        LAUNCH_CHECK(ctx);
    }
    return MSHQC_HIP_OK;
}

// Tile (ti, tj) of 64 x 64 lies in the block pair (ti / tpb, tj / tpb);
// where the column block exceeds the row block it is overwritten by the
// transpose of tile (tj, ti), through LDS.
constexpr int MIR_TILE = 64;

static __global__ __launch_bounds__(MIR_TILE * 8) void mirror_upper_kernel(double* __restrict__ M, long ld,
                                                                            long n, long tpb) {
    const long ti = blockIdx.y, tj = blockIdx.x;
    if (tj / tpb <= ti / tpb) return;
    __shared__ double t[MIR_TILE][MIR_TILE + 1];
    const int x = threadIdx.x, y0 = threadIdx.y;
    for (int y = y0; y < MIR_TILE; y += 8) {  // tile (tj, ti)
        const long r = tj * MIR_TILE + y, c = ti * MIR_TILE + x;
        if (r < n && c < n) t[y][x] = M[r * ld + c];
    }
    __syncthreads();
    for (int y = y0; y < MIR_TILE; y += 8) {  // element (ti T + y, tj T + x) = (tj T + x, ti T + y)
        const long r = ti * MIR_TILE + y, c = tj * MIR_TILE + x;
        if (r < n && c < n) M[r * ld + c] = t[x][y];
    }
}

int mirror_upper(mshqc_hip_ctx* ctx, double* M, long ld, long n, long tpb) {
    if (n <= 0) return MSHQC_HIP_OK;
    const long nt = (n + MIR_TILE - 1) / MIR_TILE;
    if (nt > 65535) return set_error(ctx, MSHQC_HIP_ERR_ARG, "mirror_upper: matrix too large");
    mirror_upper_kernel<<<dim3((unsigned)nt, (unsigned)nt), dim3(MIR_TILE, 8), 0, ctx->stream>>>(M, ld, n, tpb);
    LAUNCH_CHECK(ctx);
    return MSHQC_HIP_OK;
}

static __global__ void fill_zero_kernel(double* __restrict__ x, long count) {
    const long k = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (k < count) x[k] = 0.0;
}

int fill_zero(mshqc_hip_ctx* ctx, double* x, long count) {
    if (count <= 0) return MSHQC_HIP_OK;
    fill_zero_kernel<<<(unsigned)((count + 255) / 256), 256, 0, ctx->stream>>>(x, count);
    LAUNCH_CHECK(ctx);
    return MSHQC_HIP_OK;
}

int reduce_sum(mshqc_hip_ctx* ctx, const double* x, long n, double* out, bool accumulate) {
    reduce_sum_kernel<<<1, 1024, 0, ctx->stream>>>(x, n, out, accumulate ? 1 : 0);
    LAUNCH_CHECK(ctx);
    return MSHQC_HIP_OK;
}
