// FP64 pair-block kernel: (ia|jb) on v_mfma_f64_16x16x4_f64 with the
// denominator epilogue fused (energy, dE/deps partials, gradient amplitudes).
#include "common.h"
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

}  // namespace

int launch_pair_f64(mshqc_hip_ctx* ctx, int mode, const PairArgs& a) {
    if (a.nwork <= 0) return MSHQC_HIP_OK;
    const long nwg = ((a.nwork + 7) / 8) * 8;
    if (nwg > 0x7fffffffL) return set_error(ctx, MSHQC_HIP_ERR_ARG, "pair kernel: too much work");
    dim3 grid((unsigned)nwg);
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
