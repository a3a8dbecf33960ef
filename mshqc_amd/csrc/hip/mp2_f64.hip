// FP64 pair-block kernel: (ia|jb) on v_mfma_f64_16x16x4_f64 with the
// denominator epilogue fused (energy, dE/deps partials, gradient amplitudes).
//
// Two forms of the same work unit (one occupied pair, one 64 x 64 virtual
// tile, four waves of 32 x 32 for G and H):
//   pair_f64_dma_kernel  the four operand panels of each stage of Q stream
//                        from memory into an LDS ring by direct-to-LDS
//                        loads, stages ahead; needs 16-byte aligned operands
//                        with even virtual counts (then every 16-byte
//                        segment, two adjacent virtual indices, stays inside
//                        one orbital's row);
//   pair_f64_kernel      panels staged through registers; any shape.
#include "common.h"
#include "dma.h"
#include "mfma_f64.h"
#include "mp2_internal.h"
#include "pair_kernel.h"

namespace {

constexpr int TILE = 64, KC = 16, NT = 256, LDSP = TILE + 1;

template <int MODE>
__global__ __launch_bounds__(NT) void pair_f64_kernel(PairArgs p) {
    constexpr bool SYM = pair_mode_sym(MODE);
    constexpr int NP = SYM ? 4 : 2;
    __shared__ double S[NP][KC][LDSP];
    __shared__ double red_row[2][TILE], red_col[2][TILE], red_e[4], red_s[4];

    const long work = xcd_work_index();
    if (work >= p.nwork) return;
    const int pslot = (int)(work / p.ntiles), tslot = (int)(work % p.ntiles);
    const int2 ij = p.pairs[pslot];
    const int i = ij.x, j = ij.y;
    const double w = p.pw[pslot];
    const int2 tb = p.tiles[tslot];
    const int A0 = tb.x * TILE, B0 = tb.y * TILE;
    const bool mirror = SYM && (tb.x != tb.y);

    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int wm = (wave >> 1) * 32, wn = (wave & 1) * 32;

    // Panel sources: row base pointers at q = 0; panel columns are contiguous.
    const long sl = (long)p.ol * p.vl, sr = (long)p.orr * p.vr;
    const double* src[4];
    int lim[4];
    src[0] = p.Bl + (long)i * p.vl + A0;  lim[0] = p.vl - A0;  // G rows: Bl_i[A]
    src[1] = p.Br + (long)j * p.vr + B0;  lim[1] = p.vr - B0;  // G cols: Br_j[B]
    if (SYM) {
        src[2] = p.Br + (long)j * p.vr + A0;  lim[2] = p.vr - A0;  // H rows: Br_j[A]
        src[3] = p.Bl + (long)i * p.vl + B0;  lim[3] = p.vl - B0;  // H cols: Bl_i[B]
    }
    const long sstride[4] = {sl, sr, sr, sl};

    v4d accG[2][2], accH[2][2];
#pragma unroll
    for (int s = 0; s < 2; ++s)
#pragma unroll
        for (int t = 0; t < 2; ++t) {
            accG[s][t] = v4d{0.0, 0.0, 0.0, 0.0};
            accH[s][t] = v4d{0.0, 0.0, 0.0, 0.0};
        }

    const int c = tid & 63, kk0 = tid >> 6;
    double reg[NP][4];
    auto fetch = [&](int q0) {
#pragma unroll
        for (int P = 0; P < NP; ++P)
#pragma unroll
            for (int r = 0; r < 4; ++r) {
                const int q = q0 + kk0 + 4 * r;
                reg[P][r] = (c < lim[P] && q < p.Q) ? src[P][(long)q * sstride[P] + c] : 0.0;
            }
    };
    const int nchunk = (p.Q + KC - 1) / KC;
    fetch(0);
    for (int ch = 0; ch < nchunk; ++ch) {
        __syncthreads();
#pragma unroll
        for (int P = 0; P < NP; ++P)
#pragma unroll
            for (int r = 0; r < 4; ++r) S[P][kk0 + 4 * r][c] = reg[P][r];
        __syncthreads();
        if (ch + 1 < nchunk) fetch((ch + 1) * KC);
#pragma unroll
        for (int k4 = 0; k4 < KC / 4; ++k4) {
            const int kr = k4 * 4 + (lane >> 4), cc = lane & 15;
            double a[2], b[2];
#pragma unroll
            for (int s = 0; s < 2; ++s) a[s] = S[0][kr][wm + 16 * s + cc];
#pragma unroll
            for (int t = 0; t < 2; ++t) b[t] = S[1][kr][wn + 16 * t + cc];
#pragma unroll
            for (int s = 0; s < 2; ++s)
#pragma unroll
                for (int t = 0; t < 2; ++t) accG[s][t] = mfma_f64(a[s], b[t], accG[s][t]);
            if (SYM) {
#pragma unroll
                for (int s = 0; s < 2; ++s) a[s] = S[2][kr][wm + 16 * s + cc];
#pragma unroll
                for (int t = 0; t < 2; ++t) b[t] = S[3][kr][wn + 16 * t + cc];
#pragma unroll
                for (int s = 0; s < 2; ++s)
#pragma unroll
                    for (int t = 0; t < 2; ++t) accH[s][t] = mfma_f64(a[s], b[t], accH[s][t]);
            }
        }
    }

    pair_epilogue<MODE>(p, i, j, w, A0, B0, mirror, lane, wave, wm, wn, accG, accH, red_row,
                        red_col, red_e, red_s, work);
}

// Direct-to-LDS form.  A stage holds DK values of Q for each panel (64
// virtual columns: Bl_i[A], Br_j[B], and in the symmetric modes Br_j[A],
// Bl_i[B]).  Panel element (q, c) sits in 16-byte segment (q, c/2) at byte
// q*512 + 16 * ((c/2) ^ (8 * (q & 1))) + 8 * (c & 1): the 16 lanes of a
// fragment read touch 8 segments of one q row, the next 16 lanes the other
// half-bank group of the next row, so reads are free of bank conflicts.
// Columns past the operand's virtual count read zeros (their outputs are
// discarded); rows q >= Q of the last stage are zeroed.

// 16 bytes of zeros (never written): the DMA source of the padded columns.
__device__ __align__(16) double pair_zero_seg[2];

__device__ __forceinline__ int pf_off(int q, int c) {
    return q * 512 + ((((c >> 1) ^ ((q & 1) << 3))) << 4) + ((c & 1) << 3);
}

template <int MODE, int DK, int NBUF, int WPE>
__global__ __launch_bounds__(NT) __attribute__((amdgpu_waves_per_eu(WPE))) void pair_f64_dma_kernel(PairArgs p) {
    constexpr bool SYM = pair_mode_sym(MODE);
    constexpr int NP = SYM ? 4 : 2;
    constexpr int PANEL = DK * TILE * 8;  // bytes per panel and stage
    constexpr int STAGE = NP * PANEL;
    constexpr int LPP = PANEL / 1024;     // 1-KiB loads per panel and stage
    constexpr int PER = NP * LPP / 4;     // loads per wave and stage
    constexpr int AHEAD = NBUF - 1;
    constexpr int WAITN = (AHEAD - 1) * PER;
    static_assert((NP * LPP) % 4 == 0 && DK % 4 == 0, "every wave issues the same loads");
    extern __shared__ __align__(16) char lds[];
    __shared__ double red_row[2][TILE], red_col[2][TILE], red_e[4], red_s[4];

    const long work = xcd_work_index();
    if (work >= p.nwork) return;
    const int pslot = (int)(work / p.ntiles), tslot = (int)(work % p.ntiles);
    const int2 ij = p.pairs[pslot];
    const int i = ij.x, j = ij.y;
    const double w = p.pw[pslot];
    const int2 tb = p.tiles[tslot];
    const int A0 = tb.x * TILE, B0 = tb.y * TILE;
    const bool mirror = SYM && (tb.x != tb.y);

    const int tid = threadIdx.x, lane = tid & 63;
    const int wave = __builtin_amdgcn_readfirstlane(tid >> 6);
    const int wm = (wave >> 1) * 32, wn = (wave & 1) * 32;

    // This wave's loads: u = wave*PER + k is load u % LPP of panel u / LPP;
    // one load covers two q rows (lane l: row l/32, segment l%32 in its
    // swizzled position).
    const double* col_ptr[PER];
    long qstride[PER];
    int qrow[PER], dst[PER];
#pragma unroll
    for (int k = 0; k < PER; ++k) {
        const int u = wave * PER + k, P = u / LPP, uu = u % LPP;
        const int qr = 2 * uu + (lane >> 5);
        const int cp = (lane & 31) ^ ((qr & 1) << 3);
        const bool left = P == 0 || P == 3;
        const int v = left ? p.vl : p.vr;
        const int c0 = (P == 0 || P == 2) ? A0 : B0;
        const double* base = left ? p.Bl + (long)i * p.vl : p.Br + (long)j * p.vr;
        // Columns past the virtual count (v is even, so a segment is whole
        // or wholly past) read a fixed 16-byte zero segment at every q: their
        // outputs are discarded, and zero operands draw less power than real
        // columns would.
        const bool pad = c0 + 2 * cp >= v;
        col_ptr[k] = pad ? pair_zero_seg : base + c0 + 2 * cp;
        qstride[k] = pad ? 0 : left ? (long)p.ol * p.vl : (long)p.orr * p.vr;
        qrow[k] = qr;
        dst[k] = P * PANEL + uu * 1024;
    }
    const int nst = (p.Q + DK - 1) / DK;
    // Stage st (clamped to the last Q row: past the end it loads valid rows
    // into a slot no one reads, so every wave issues exactly PER loads per
    // stage on every path).
    auto issue = [&](int st) {
        char* slot = lds + (st % NBUF) * STAGE;
#pragma unroll
        for (int k = 0; k < PER; ++k) {
            const int q = min(st * DK + qrow[k], p.Q - 1);
            dma16(col_ptr[k] + (long)q * qstride[k], slot + dst[k]);
        }
    };

    v4d accG[2][2], accH[2][2];
#pragma unroll
    for (int s = 0; s < 2; ++s)
#pragma unroll
        for (int t = 0; t < 2; ++t) {
            accG[s][t] = v4d{0.0, 0.0, 0.0, 0.0};
            accH[s][t] = v4d{0.0, 0.0, 0.0, 0.0};
        }
#pragma unroll
    for (int a = 0; a < AHEAD; ++a) issue(a);
    dma_barrier<WAITN>();  // stage 0 has landed
    for (int st = 0; st < nst; ++st) {
        const char* slot = lds + (st % NBUF) * STAGE;
        if (st == nst - 1) {
            const int kv = p.Q - st * DK;
            if (kv < DK) {
                for (int e = tid; e < NP * DK * TILE; e += NT) {
                    const int P = e / (DK * TILE), q = (e / TILE) % DK, c = e % TILE;
                    if (q >= kv) *reinterpret_cast<double*>(lds + (st % NBUF) * STAGE + P * PANEL + pf_off(q, c)) = 0.0;
                }
                dma_barrier<WAITN>();  // zeros written before anyone reads
            }
        }
        issue(st + AHEAD);
#pragma unroll
        for (int k4 = 0; k4 < DK / 4; ++k4) {
            const int kr = k4 * 4 + (lane >> 4), cc = lane & 15;
            double a[2], b[2];
#pragma unroll
            for (int s = 0; s < 2; ++s) a[s] = *reinterpret_cast<const double*>(slot + pf_off(kr, wm + 16 * s + cc));
#pragma unroll
            for (int t = 0; t < 2; ++t)
                b[t] = *reinterpret_cast<const double*>(slot + PANEL + pf_off(kr, wn + 16 * t + cc));
#pragma unroll
            for (int s = 0; s < 2; ++s)
#pragma unroll
                for (int t = 0; t < 2; ++t) accG[s][t] = mfma_f64(a[s], b[t], accG[s][t]);
            if (SYM) {
#pragma unroll
                for (int s = 0; s < 2; ++s)
                    a[s] = *reinterpret_cast<const double*>(slot + 2 * PANEL + pf_off(kr, wm + 16 * s + cc));
#pragma unroll
                for (int t = 0; t < 2; ++t)
                    b[t] = *reinterpret_cast<const double*>(slot + 3 * PANEL + pf_off(kr, wn + 16 * t + cc));
#pragma unroll
                for (int s = 0; s < 2; ++s)
#pragma unroll
                    for (int t = 0; t < 2; ++t) accH[s][t] = mfma_f64(a[s], b[t], accH[s][t]);
            }
        }
        dma_barrier<WAITN>();  // stage st+1 has landed
    }
    // The clamped loads past the last stage land before the epilogue and
    // before the LDS is handed to the next workgroup.
    dma_barrier<0>();
    pair_epilogue<MODE>(p, i, j, w, A0, B0, mirror, lane, wave, wm, wn, accG, accH, red_row,
                        red_col, red_e, red_s, work);
}

template <int DK, int NBUF>
constexpr size_t pair_dma_lds(bool sym) { return (size_t)NBUF * (sym ? 4 : 2) * DK * TILE * 8; }

// The direct-to-LDS form reads 16-byte segments of two adjacent virtual
// indices of one orbital: aligned operands and even virtual counts.
bool pair_dma_eligible(const PairArgs& a) {
    auto al16 = [](const void* q) { return (reinterpret_cast<uintptr_t>(q) & 15) == 0; };
    return al16(a.Bl) && al16(a.Br) && a.vl % 2 == 0 && a.vr % 2 == 0;
}

template <int MODE, int DK, int NBUF, int WPE>
void launch_pair_dma(mshqc_hip_ctx* ctx, dim3 grid, const PairArgs& a) {
    constexpr size_t L = pair_dma_lds<DK, NBUF>(pair_mode_sym(MODE));
    hipFuncSetAttribute((const void*)pair_f64_dma_kernel<MODE, DK, NBUF, WPE>,
                        hipFuncAttributeMaxDynamicSharedMemorySize, (int)L);
    pair_f64_dma_kernel<MODE, DK, NBUF, WPE><<<grid, NT, L, ctx->stream>>>(a);
}

template <int DK, int NBUF, int WPE>
int dispatch_pair_dma(mshqc_hip_ctx* ctx, int mode, dim3 grid, const PairArgs& a) {
    switch (mode) {
        case MODE_RMP2: launch_pair_dma<MODE_RMP2, DK, NBUF, WPE>(ctx, grid, a); break;
        case MODE_SS: launch_pair_dma<MODE_SS, DK, NBUF, WPE>(ctx, grid, a); break;
        case MODE_OS: launch_pair_dma<MODE_OS, DK, NBUF, WPE>(ctx, grid, a); break;
        case MODE_T2: launch_pair_dma<MODE_T2, DK, NBUF, WPE>(ctx, grid, a); break;
        case MODE_T2OS: launch_pair_dma<MODE_T2OS, DK, NBUF, WPE>(ctx, grid, a); break;
        default: return set_error(ctx, MSHQC_HIP_ERR_ARG, "pair kernel: bad mode %d", mode);
    }
    LAUNCH_CHECK(ctx);
    return MSHQC_HIP_OK;
}

}  // namespace

int launch_pair_f64(mshqc_hip_ctx* ctx, int mode, const PairArgs& a) {
    if (a.nwork <= 0) return MSHQC_HIP_OK;
    const long nwg = ((a.nwork + 7) / 8) * 8;
    if (nwg > 0x7fffffffL) return set_error(ctx, MSHQC_HIP_ERR_ARG, "pair kernel: too much work");
    dim3 grid((unsigned)nwg);
    // 12 values of Q per stage, two stages, three waves per SIMD (48 KiB
    // of LDS and at most 170 registers per wave); timed on gfx950 against
    // 4- to 24-deep stages, two to six stages and two to four waves per SIMD.
    if (pair_dma_eligible(a)) return dispatch_pair_dma<12, 2, 3>(ctx, mode, grid, a);
    switch (mode) {
        case MODE_RMP2: pair_f64_kernel<MODE_RMP2><<<grid, NT, 0, ctx->stream>>>(a); break;
        case MODE_SS: pair_f64_kernel<MODE_SS><<<grid, NT, 0, ctx->stream>>>(a); break;
        case MODE_OS: pair_f64_kernel<MODE_OS><<<grid, NT, 0, ctx->stream>>>(a); break;
        case MODE_T2: pair_f64_kernel<MODE_T2><<<grid, NT, 0, ctx->stream>>>(a); break;
        case MODE_T2OS: pair_f64_kernel<MODE_T2OS><<<grid, NT, 0, ctx->stream>>>(a); break;
        default: return set_error(ctx, MSHQC_HIP_ERR_ARG, "pair kernel: bad mode %d", mode);
    }
    LAUNCH_CHECK(ctx);
    return MSHQC_HIP_OK;
}
