// Fock-build pieces (J, K) and AO->MO transforms from Cholesky vectors.
//
//   J   : two streaming passes over the mu <= nu half of L (HBM-bound; L^Q
//         and J are symmetric): gamma_Q = <L^Q, P_tot> with wavefront-shuffle
//         reductions, then J = sum_Q gamma_Q L^Q, mirrored.
//   K   : FP64 MFMA GEMMs, batched over Q (general density: T = P L^Q then
//         K += L^Q T; factored density: X^Q = L^Q Y then K += X^Q X^Q^T).
//   B   : occupied-first transform X^Q = L^Q C_occ, B^Q = X^Q^T C_vir.
#include <algorithm>
#include <cmath>

#include "common.h"
#include "mp2_bf16.h"

namespace {

// One workgroup per Q: out[q] = the squared Frobenius norm of the symmetric
// L^Q from its upper triangle (off-diagonal terms twice).  Wave w takes rows
// m = w, w+4, ...; each row's upper part [m, n) is one contiguous run read 64
// lanes at a time.  Fixed orders throughout.
__global__ __launch_bounds__(256) void norm2_upper_kernel(const double* __restrict__ L, int n,
                                                          double* __restrict__ out) {
    const long nn = (long)n * n;
    const double* Lq = L + (long)blockIdx.x * nn;
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    double s = 0.0;
    for (int m = wave; m < n; m += 4) {
        const double* lr = Lq + (long)m * n;
        double d = 0.0;
#pragma unroll 4
        for (int v = m + lane; v < n; v += 64) d += (v == m ? 1.0 : 2.0) * lr[v] * lr[v];
        s += d;
    }
    for (int k = 32; k > 0; k >>= 1) s += __shfl_xor(s, k, 64);
    __shared__ double sh[4];
    if (lane == 0) sh[wave] = s;
    __syncthreads();
    if (threadIdx.x == 0) out[blockIdx.x] = ((sh[0] + sh[1]) + sh[2]) + sh[3];
}

// J[m][v] = J[v][m] = sum_q gamma[q] L[q][m][v] for m <= v: row m is
// blockIdx.y, two consecutive v per thread; pairs wholly below the diagonal
// return at once, so only the upper half of L is read.  The second element
// of a pair at the row's end (odd n) re-reads the first, in bounds.
__global__ __launch_bounds__(256) void j_accumulate_upper_kernel(const double* __restrict__ L, int Q, int n,
                                                                 const double* __restrict__ gam,
                                                                 double* __restrict__ J) {
    const int m = blockIdx.y;
    const int v0 = 2 * (blockIdx.x * 256 + threadIdx.x);
    if (v0 >= n || v0 + 1 < m) return;
    const long nn = (long)n * n;
    const bool has1 = v0 + 1 < n;
    const int off1 = has1 ? 1 : 0;
    double s0 = 0.0, s1 = 0.0;
    const double* p = L + (long)m * n + v0;
    // unrolled so several q's loads are in flight; the sums stay in q order
#pragma unroll 8
    for (int q = 0; q < Q; ++q) {
        const double g = gam[q];
        s0 += g * p[(long)q * nn];
        s1 += g * p[(long)q * nn + off1];
    }
    if (v0 >= m) {
        J[(long)m * n + v0] = s0;
        J[(long)v0 * n + m] = s0;
    }
    if (has1) {
        J[(long)m * n + v0 + 1] = s1;
        J[(long)(v0 + 1) * n + m] = s1;
    }
}

// gam[q] = sum_{m < n, i < k} w_i X[q][m][i] Y[m][i] for a factored density
// P = sum_i w_i y_i y_i^T (w_i = (i < wa) + (i < wb)), from X_q = L^q Y:
// <L^q, P> = sum_i w_i y_i^T L^q y_i.  One workgroup per q: wave w takes rows
// m = w, w+4, ..., its lanes the columns i of a row; each thread sums in
// order, then a fixed tree.
__global__ __launch_bounds__(256) void gamma_factor_kernel(const double* __restrict__ X, int n, int k,
                                                           const double* __restrict__ Y, long ldy, int wa,
                                                           int wb, double* __restrict__ gam) {
    const double* Xq = X + (long)blockIdx.x * n * k;
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    double s = 0.0;
    for (int m = wave; m < n; m += 4)
        for (int i = lane; i < k; i += 64)
            s += (double)((i < wa) + (i < wb)) * (Xq[(long)m * k + i] * Y[(long)m * ldy + i]);
    __shared__ double sh[256];
    sh[threadIdx.x] = s;
    __syncthreads();
    for (int h = 128; h > 0; h >>= 1) {
        if (threadIdx.x < h) sh[threadIdx.x] += sh[threadIdx.x + h];
        __syncthreads();
    }
    if (threadIdx.x == 0) gam[blockIdx.x] = sh[0];
}

// gam[q] (+)= tr(T[q]) for the n x n blocks T[q], in a fixed order.
__global__ __launch_bounds__(256) void trace_kernel(const double* __restrict__ T, int n,
                                                    double* __restrict__ gam, int accumulate) {
    const double* Tq = T + (long)blockIdx.x * n * n;
    double s = 0.0;
    for (int m = threadIdx.x; m < n; m += 256) s += Tq[(long)m * (n + 1)];
    __shared__ double sh[256];
    sh[threadIdx.x] = s;
    __syncthreads();
    for (int h = 128; h > 0; h >>= 1) {
        if (threadIdx.x < h) sh[threadIdx.x] += sh[threadIdx.x + h];
        __syncthreads();
    }
    if (threadIdx.x == 0) gam[blockIdx.x] = accumulate ? gam[blockIdx.x] + sh[0] : sh[0];
}

// y += x elementwise.
__global__ void add_kernel(double* __restrict__ y, const double* __restrict__ x, long count) {
    const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < count) y[i] += x[i];
}


// Zero T_q for screened vectors (norm2[q] <= thr^2): their K contribution
// then vanishes exactly.
__global__ void mask_zero_kernel(double* __restrict__ T, const double* __restrict__ norm2, int q0,
                                 int nq, long nn, double thr2) {
    const long idx = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (idx >= (long)nq * nn) return;
    const int q = (int)(idx / nn);
    if (!(norm2[q0 + q] > thr2)) T[idx] = 0.0;
}

// J = sum_q gam[q] L^q from the upper half of each L^q.
int j_from_gamma(mshqc_hip_ctx* ctx, const double* L, int Q, int n, const double* gam, double* J) {
    ProfScope ps(ctx, PROF_J);
    dim3 grid((unsigned)((n + 511) / 512), (unsigned)n);
    j_accumulate_upper_kernel<<<grid, 256, 0, ctx->stream>>>(L, Q, n, gam, J);
    LAUNCH_CHECK(ctx);
    return MSHQC_HIP_OK;
}

// Largest Q chunk whose n x m temporaries fit in `cap` doubles.
int q_chunk(int Q, long per_q, size_t cap) {
    long c = (long)(cap / (size_t)std::max(per_q, 1L));
    return (int)std::max(1L, std::min((long)Q, c));
}

constexpr size_t kTmpCap = size_t(1) << 28;  // doubles (2 GiB) per temporary

int gemm_splits(int M, int N, int batch, bool lower_only = false) {
    const long tm = (M + 63) / 64, tn = (N + 63) / 64;
    const long tiles = lower_only ? tm * (tm + 1) / 2 : tm * tn;  // tiles that do work
    long s = (2048 + tiles - 1) / tiles;
    return (int)std::max(1L, std::min((long)batch, s));
}

}  // namespace

extern "C" {

int mshqc_hip_jk(mshqc_hip_ctx* ctx, const double* L, int Q, int n, const double* P, int nspin,
                 int screen, double* J, double* K) {
    int rc = ctx_enter(ctx);
    if (rc) return rc;
    CHECK_ARG(ctx, L && P && J && K, "jk: null pointer");
    CHECK_ARG(ctx, Q > 0 && n > 0 && nspin >= 1 && nspin <= 2, "jk: bad dimensions");
    hipEventRecord(ctx->ev0, ctx->stream);
    if ((rc = jk_general(ctx, L, Q, n, P, nspin, screen, J, K))) return rc;
    hipEventRecord(ctx->ev1, ctx->stream);
    if ((rc = ctx_finish(ctx))) return rc;
    hipEventElapsedTime(&ctx->last_ms, ctx->ev0, ctx->ev1);
    return MSHQC_HIP_OK;
}

}  // extern "C"

int jk_general(mshqc_hip_ctx* ctx, const double* L, int Q, int n, const double* P, int nspin, int screen,
               double* J, double* K) {
    int rc;
    const long nn = (long)n * n;
    const int qc = q_chunk(Q, nn, kTmpCap);
    double* tmp = static_cast<double*>(
        workspace(ctx, WS_TMP, sizeof(double) * ((size_t)qc * nn + 2 * (size_t)Q + 64)));
    if (!tmp) return MSHQC_HIP_ERR_ALLOC;
    double* T = tmp;
    double* gam = tmp + (size_t)qc * nn;
    double* norm2 = gam + Q;
    // J's densities gamma_q = <L^q, P_tot> = sum_s tr(L^q P_s) are the traces
    // of the exchange's first products T'_q = L^q P_s (taken before the
    // screening), so J reads L once instead of twice
    if (screen) {
        norm2_upper_kernel<<<Q, 256, 0, ctx->stream>>>(L, n, norm2);
        LAUNCH_CHECK(ctx);
    }
    for (int s = 0; s < nspin; ++s) {
        double* Ks = K + (long)s * nn;
        for (int q0 = 0; q0 < Q; q0 += qc) {
            const int nq = std::min(qc, Q - q0);
            // T_q = P_s L^q = (L^q P_s)^T (both symmetric).  The rows of all
            // q's L^q form one k-contiguous operand, so the products run as one
            // (nq*n) x n x n GEMM; it stores T'_q = L^q P_s = T_q^T, which the
            // second product reads as a k-contiguous X
            GemmF64 g1;
            g1.A = L + (long)q0 * nn;
            g1.lda = n;
            g1.sA = nn;
            g1.X = P + (long)s * nn;
            g1.ldx = n;
            g1.sX = 0;
            g1.C = T;
            g1.ldc = n;
            g1.sC = nn;
            g1.M = g1.N = g1.K = n;
            g1.batch = nq;
            { ProfScope ps(ctx, PROF_GEMM); rc = gemm_f64(ctx, g1); }
            if (rc) return rc;
            {
                ProfScope ps(ctx, PROF_J);
                trace_kernel<<<nq, 256, 0, ctx->stream>>>(T, n, gam + q0, s > 0 ? 1 : 0);
                LAUNCH_CHECK(ctx);
            }
            if (screen) {
                const long tot = (long)nq * nn;
                mask_zero_kernel<<<(unsigned)((tot + 255) / 256), 256, 0, ctx->stream>>>(
                    T, norm2, q0, nq, nn, 1e-16);
                LAUNCH_CHECK(ctx);
            }
            // K_s (+)= sum_q L^q T_q
            GemmF64 g2;
            g2.A = L + (long)q0 * nn;
            g2.lda = n;
            g2.sA = nn;
            g2.X = T;
            g2.ldx = n;
            g2.sX = nn;
            g2.transX = true;  // T_q[k][n] = T'_q[n][k]
            g2.C = Ks;
            g2.ldc = n;
            g2.M = g2.N = g2.K = n;
            g2.batch = nq;
            g2.reduce_batch = true;
            g2.splits = gemm_splits(n, n, nq, true);
            g2.accumulate = q0 > 0;
            g2.lower_only = true;  // K_s = sum_q L^q P_s L^q is symmetric
            { ProfScope ps(ctx, PROF_GEMM); rc = gemm_f64(ctx, g2); }
            if (rc) return rc;
        }
        if ((rc = mirror_upper(ctx, Ks, n, n, 1))) return rc;
    }
    return j_from_gamma(ctx, L, Q, n, gam, J);
}

extern "C" {

int mshqc_hip_jk_factor(mshqc_hip_ctx* ctx, const double* L, int Q, int n, const double* Y, int m,
                        double* J, double* K) {
    int rc = ctx_enter(ctx);
    if (rc) return rc;
    CHECK_ARG(ctx, L && Y && J && K, "jk_factor: null pointer");
    CHECK_ARG(ctx, Q > 0 && n > 0 && m > 0, "jk_factor: bad dimensions");
    hipEventRecord(ctx->ev0, ctx->stream);
    const long nn = (long)n * n, nm = (long)n * m;
    const int qc = q_chunk(Q, nm, kTmpCap);
    double* tmp = static_cast<double*>(
        workspace(ctx, WS_TMP, sizeof(double) * ((size_t)qc * nm + (size_t)Q + 64)));
    if (!tmp) return MSHQC_HIP_ERR_ALLOC;
    double* X = tmp;
    double* gam = tmp + (size_t)qc * nm;
    // J's densities gamma_q = <L^q, Y Y^T> = <L^q Y, Y> come from the X_q of
    // the exchange, so J reads L once instead of twice
    for (int q0 = 0; q0 < Q; q0 += qc) {
        const int nq = std::min(qc, Q - q0);
        // X_q = L^q Y
        GemmF64 g1;
        g1.A = L + (long)q0 * nn;
        g1.lda = n;
        g1.sA = nn;
        g1.X = Y;
        g1.ldx = m;
        g1.sX = 0;
        g1.C = X;
        g1.ldc = m;
        g1.sC = nm;
        g1.M = n;
        g1.N = m;
        g1.K = n;
        g1.batch = nq;
        { ProfScope ps(ctx, PROF_GEMM); rc = gemm_f64(ctx, g1); }
            if (rc) return rc;
        {
            ProfScope ps(ctx, PROF_J);
            gamma_factor_kernel<<<nq, 256, 0, ctx->stream>>>(X, n, m, Y, m, m, 0, gam + q0);
            LAUNCH_CHECK(ctx);
        }
        // K (+)= sum_q X_q X_q^T
        GemmF64 g2;
        g2.A = X;
        g2.lda = m;
        g2.sA = nm;
        g2.X = X;
        g2.ldx = m;
        g2.sX = nm;
        g2.transX = true;
        g2.C = K;
        g2.ldc = n;
        g2.M = g2.N = n;
        g2.K = m;
        g2.batch = nq;
        g2.reduce_batch = true;
        g2.splits = gemm_splits(n, n, nq, true);
        g2.accumulate = q0 > 0;
        g2.lower_only = true;  // symmetric
        { ProfScope ps(ctx, PROF_GEMM); rc = gemm_f64(ctx, g2); }
            if (rc) return rc;
    }
    if ((rc = j_from_gamma(ctx, L, Q, n, gam, J))) return rc;
    if ((rc = mirror_upper(ctx, K, n, n, 1))) return rc;
    hipEventRecord(ctx->ev1, ctx->stream);
    if ((rc = ctx_finish(ctx))) return rc;
    hipEventElapsedTime(&ctx->last_ms, ctx->ev0, ctx->ev1);
    return MSHQC_HIP_OK;
}

int mshqc_hip_jk_rohf(mshqc_hip_ctx* ctx, const double* L, int Q, int n, const double* C, int ldc,
                      int na, int nb, double* J, double* Ka, double* Kb) {
    // cholesky_rohf.cc:89-147.  C_beta = C_alpha there, so the beta occupied
    // block is the first nb columns of the alpha one and X^Q = L^Q C_{0..m}
    // (m = max(na, nb)) feeds both exchange matrices.
    int rc = ctx_enter(ctx);
    if (rc) return rc;
    CHECK_ARG(ctx, L && C && J && Ka && Kb, "jk_rohf: null pointer");
    CHECK_ARG(ctx, Q > 0 && n > 0 && na >= 0 && nb >= 0, "jk_rohf: bad dimensions");
    const int m = std::max(na, nb);
    CHECK_ARG(ctx, ldc >= std::max(m, 1), "jk_rohf: ldc smaller than max(na, nb)");
    hipEventRecord(ctx->ev0, ctx->stream);
    const long nn = (long)n * n, nm = (long)n * std::max(m, 1);
    if (m == 0) {
        if ((rc = fill_zero(ctx, J, nn)) || (rc = fill_zero(ctx, Ka, nn)) || (rc = fill_zero(ctx, Kb, nn)))
            return rc;
    } else {
        const int qc = q_chunk(Q, nm, kTmpCap);
        double* tmp = static_cast<double*>(
            workspace(ctx, WS_TMP, sizeof(double) * ((size_t)qc * nm + (size_t)Q + 64)));
        if (!tmp) return MSHQC_HIP_ERR_ALLOC;
        double* X = tmp;
        double* gam = tmp + (size_t)qc * nm;
        // J's densities gamma_q = <L^q, C_a C_a^T + C_b C_b^T> = sum_i w_i
        // (L^q C)[:, i] . C[:, i] with w_i = (i < na) + (i < nb), from the X_q
        // of the exchange, so J reads L once instead of twice
        if (na == 0 && (rc = fill_zero(ctx, Ka, nn))) return rc;
        if (nb == 0 && (rc = fill_zero(ctx, Kb, nn))) return rc;
        const int ms = std::min(na, nb);
        double* Ksmall = na <= nb ? Ka : Kb;
        double* Kbig = na <= nb ? Kb : Ka;
        for (int q0 = 0; q0 < Q; q0 += qc) {
            const int nq = std::min(qc, Q - q0);
            // X_q = L^q C_{0..m}
            GemmF64 g1;
            g1.A = L + (long)q0 * nn;
            g1.lda = n;
            g1.sA = nn;
            g1.X = C;
            g1.ldx = ldc;
            g1.sX = 0;
            g1.C = X;
            g1.ldc = m;
            g1.sC = nm;
            g1.M = n;
            g1.N = m;
            g1.K = n;
            g1.batch = nq;
            { ProfScope ps(ctx, PROF_GEMM); rc = gemm_f64(ctx, g1); }
            if (rc) return rc;
            {
                ProfScope ps(ctx, PROF_J);
                gamma_factor_kernel<<<nq, 256, 0, ctx->stream>>>(X, n, m, C, ldc, na, nb, gam + q0);
                LAUNCH_CHECK(ctx);
            }
            // K_small (+)= sum_q X_q[:, 0..ms) X_q[:, 0..ms)^T over the smaller
            // spin's columns, and the larger spin's exchange only over the
            // columns between: K_big = K_small + sum_q X_q[:, ms..m) X_q[:, ms..m)^T
            for (int part = 0; part < 2; ++part) {
                const int c0 = part == 0 ? 0 : ms, c1 = part == 0 ? ms : m;
                if (c1 == c0) continue;
                GemmF64 g2;
                g2.A = X + c0;
                g2.lda = m;
                g2.sA = nm;
                g2.X = X + c0;
                g2.ldx = m;
                g2.sX = nm;
                g2.transX = true;
                g2.C = part == 0 ? Ksmall : Kbig;
                g2.ldc = n;
                g2.M = g2.N = n;
                g2.K = c1 - c0;
                g2.batch = nq;
                g2.reduce_batch = true;
                g2.splits = gemm_splits(n, n, nq, true);
                g2.accumulate = q0 > 0;
                g2.lower_only = true;  // symmetric
                { ProfScope ps(ctx, PROF_GEMM); rc = gemm_f64(ctx, g2); }
                if (rc) return rc;
            }
        }
        if ((rc = j_from_gamma(ctx, L, Q, n, gam, J))) return rc;
        if (ms > 0) {
            if (m > ms) {
                add_kernel<<<(unsigned)((nn + 255) / 256), 256, 0, ctx->stream>>>(Kbig, Ksmall, nn);
                LAUNCH_CHECK(ctx);
            } else {
                HIP_TRY(ctx, hipMemcpyAsync(Kbig, Ksmall, sizeof(double) * nn, hipMemcpyDeviceToDevice,
                                              ctx->stream));
            }
        }
    }
    if (na > 0 && (rc = mirror_upper(ctx, Ka, n, n, 1))) return rc;
    if (nb > 0 && (rc = mirror_upper(ctx, Kb, n, n, 1))) return rc;
    hipEventRecord(ctx->ev1, ctx->stream);
    if ((rc = ctx_finish(ctx))) return rc;
    hipEventElapsedTime(&ctx->last_ms, ctx->ev0, ctx->ev1);
    return MSHQC_HIP_OK;
}

int mshqc_hip_half_transform(mshqc_hip_ctx* ctx, const double* L, int Q, int n, const double* C,
                             int m, double* X) {
    int rc = ctx_enter(ctx);
    if (rc) return rc;
    CHECK_ARG(ctx, L && C && X, "half_transform: null pointer");
    CHECK_ARG(ctx, Q > 0 && n > 0 && m > 0, "half_transform: bad dimensions");
    hipEventRecord(ctx->ev0, ctx->stream);
    GemmF64 g;
    g.A = L;
    g.lda = n;
    g.sA = (long)n * n;
    g.X = C;
    g.ldx = m;
    g.sX = 0;
    g.C = X;
    g.ldc = m;
    g.sC = (long)n * m;
    g.M = n;
    g.N = m;
    g.K = n;
    g.batch = Q;
    { ProfScope ps(ctx, PROF_GEMM); rc = gemm_f64(ctx, g); }
    if (rc) return rc;
    hipEventRecord(ctx->ev1, ctx->stream);
    if ((rc = ctx_finish(ctx))) return rc;
    hipEventElapsedTime(&ctx->last_ms, ctx->ev0, ctx->ev1);
    return MSHQC_HIP_OK;
}

int mshqc_hip_transform(mshqc_hip_ctx* ctx, const double* L, int Q, int n, const double* Cocc,
                        int o, const double* Cvir, int v, int precision, double* B) {
    int rc = ctx_enter(ctx);
    if (rc) return rc;
    CHECK_ARG(ctx, L && Cocc && Cvir && B, "transform: null pointer");
    CHECK_ARG(ctx, Q > 0 && n > 0 && o > 0 && v > 0, "transform: bad dimensions");
    if (precision != MSHQC_PREC_FP64 && (precision < 2 || precision > BF16_TRANSFORM_SLICES))
        return set_error(ctx, MSHQC_HIP_ERR_ARG,
                         "transform: precision must be 0 (FP64) or a slice count in [2, %d], got %d",
                         BF16_TRANSFORM_SLICES, precision);
    hipEventRecord(ctx->ev0, ctx->stream);
    if (precision != MSHQC_PREC_FP64) {
        if ((rc = bf16_transform(ctx, L, Q, n, Cocc, o, Cvir, v, precision, B))) return rc;
        hipEventRecord(ctx->ev1, ctx->stream);
        if ((rc = ctx_finish(ctx))) return rc;
        hipEventElapsedTime(&ctx->last_ms, ctx->ev0, ctx->ev1);
        return MSHQC_HIP_OK;
    }
    const long nn = (long)n * n, no = (long)n * o;
    const int qc = q_chunk(Q, no, kTmpCap);
    double* X = static_cast<double*>(workspace(ctx, WS_TMP, sizeof(double) * (size_t)qc * no));
    if (!X) return MSHQC_HIP_ERR_ALLOC;
    for (int q0 = 0; q0 < Q; q0 += qc) {
        const int nq = std::min(qc, Q - q0);
        // X_q = L^q C_occ (n x o), stored transposed (o x n): the rows of all
        // q's X_q^T are then one k-contiguous operand of the second product,
        // which runs as one (nq*o) x v x n GEMM
        GemmF64 g1;
        g1.A = L + (long)q0 * nn;
        g1.lda = n;
        g1.sA = nn;
        g1.X = Cocc;
        g1.ldx = o;
        g1.C = X;
        g1.ldc = n;
        g1.sC = no;
        g1.trans_c = true;
        g1.M = n;
        g1.N = o;
        g1.K = n;
        g1.batch = nq;
        { ProfScope ps(ctx, PROF_GEMM); rc = gemm_f64(ctx, g1); }
            if (rc) return rc;
        GemmF64 g2;  // B_q = X_q^T C_vir  (o x v)
        g2.A = X;
        g2.lda = n;
        g2.sA = no;
        g2.X = Cvir;
        g2.ldx = v;
        g2.C = B + (long)q0 * o * v;
        g2.ldc = v;
        g2.sC = (long)o * v;
        g2.M = o;
        g2.N = v;
        g2.K = n;
        g2.batch = nq;
        { ProfScope ps(ctx, PROF_GEMM); rc = gemm_f64(ctx, g2); }
            if (rc) return rc;
    }
    hipEventRecord(ctx->ev1, ctx->stream);
    if ((rc = ctx_finish(ctx))) return rc;
    hipEventElapsedTime(&ctx->last_ms, ctx->ev0, ctx->ev1);
    return MSHQC_HIP_OK;
}

}  // extern "C"
