// FP64 MFMA GEMM engine (gfx950, v_mfma_f64_16x16x4_f64) and deterministic
// reductions.  Used by the J/K build, the AO->MO transforms and the
// gradient contractions of the MP2 backward.
#include "common.h"
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
                    double* p = Cb + (long)row * ldc + col;
                    if (direct) {
                        double val = g.alpha * acc[s][t][r];
                        *p = g.accumulate ? *p + val : val;
                    } else {
                        *p = acc[s][t][r];
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

}  // namespace

int gemm_f64(mshqc_hip_ctx* ctx, const GemmF64& g) {
    if (g.M <= 0 || g.N <= 0 || g.batch <= 0) return MSHQC_HIP_OK;
    if (g.K <= 0) return set_error(ctx, MSHQC_HIP_ERR_ARG, "gemm_f64: K must be positive");
    if (g.inner < 1 || (g.inner > 1 && !g.reduce_batch))
        return set_error(ctx, MSHQC_HIP_ERR_ARG, "gemm_f64: two-level batches need reduce_batch");
    const unsigned gx = grouped_grid((g.M + TM - 1) / TM, (g.N + TN - 1) / TN);
    int splits = 1, bper = 1;
    double* part = nullptr;
    if (g.reduce_batch) {
        splits = g.splits < 1 ? 1 : (g.splits > g.batch ? g.batch : g.splits);
        bper = (g.batch + splits - 1) / splits;
        splits = (g.batch + bper - 1) / bper;
        if (splits > 1) {
            part = static_cast<double*>(
                workspace(ctx, WS_TMP2, sizeof(double) * (size_t)splits * g.M * g.N));
            if (!part) return MSHQC_HIP_ERR_ALLOC;
        }
    }
    const unsigned gz = g.reduce_batch ? splits : g.batch;
    if (gz > 65535u) return set_error(ctx, MSHQC_HIP_ERR_ARG, "gemm_f64: batch too large");
    dim3 grid(gx, 1, gz);
    if (g.transA) {
        if (g.transX) gemm_f64_kernel<true, true><<<grid, NT, 0, ctx->stream>>>(g, part, bper);
        else gemm_f64_kernel<true, false><<<grid, NT, 0, ctx->stream>>>(g, part, bper);
    } else {
        if (g.transX) gemm_f64_kernel<false, true><<<grid, NT, 0, ctx->stream>>>(g, part, bper);
        else gemm_f64_kernel<false, false><<<grid, NT, 0, ctx->stream>>>(g, part, bper);
    }
    LAUNCH_CHECK(ctx);
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
