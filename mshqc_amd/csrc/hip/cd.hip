// Pivoted incomplete Cholesky of the ERI matrix on the device, FP64
// (cholesky_eri.cc:40-203).
//
// The ERI matrix M[(ij)][(kl)] = (ij|kl) (n^2 x n^2, row-major) stays in HBM.
// Each step is two kernels:
//   pivot : one workgroup scans the diagonal D for its first maximum;
//           stops the decomposition when it is below the threshold or the
//           output is full
//   vector: L_new(ij) = (M[ij][p] - sum_k L_k(ij) L_k(p)) / sqrt(D_p),
//           D(ij) -= L_new(ij)^2, clamped at 0
// with the same operation order as the reference's scalar loop and no fused
// multiply-adds (contraction off; IEEE division and square root), so the
// vectors and the pivot sequence are those of the FP64 host algorithm.
// Steps are queued in batches; a small device state word makes queued steps
// after the stop into no-ops, so the host waits once per batch.
#include <algorithm>
#include <cmath>

#include "common.h"

namespace {

constexpr int PIV_THREADS = 1024;
constexpr int VEC_THREADS = 256;
constexpr int STEPS_PER_BATCH = 64;

struct CdState {
    int nvec;     // committed vectors
    int pending;  // a vector was written by the last step and is not yet counted
    int done;     // 1: below threshold, 2: output full
    int pivot;
    double dpiv;  // D at the pivot when it was chosen
    double dmax;  // largest remaining diagonal at the stop
};

__global__ void cd_init_kernel(const double* __restrict__ M, long np, double* __restrict__ D,
                               CdState* st) {
    const long k = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (k < np) D[k] = M[k * np + k];
    if (k == 0) {
        st->nvec = 0;
        st->pending = 0;
        st->done = 0;
        st->pivot = 0;
        st->dpiv = 0.0;
        st->dmax = 0.0;
    }
}

// First maximum of D (strictly greater wins; ties keep the lower index), as
// cholesky_eri.cc:155-167 scans it.
__global__ __launch_bounds__(PIV_THREADS) void cd_pivot_kernel(const double* __restrict__ D, long np,
                                                               double thr, int cap, const double* __restrict__ L,
                                                               double* __restrict__ Lp, CdState* st) {
    __shared__ double sv[PIV_THREADS];
    __shared__ long si[PIV_THREADS];
    __shared__ int skip, gather, nv_sh;
    if (threadIdx.x == 0) {
        if (st->pending) {
            st->nvec += 1;
            st->pending = 0;
        }
        skip = st->done;
    }
    __syncthreads();
    if (skip) return;
    double best = -INFINITY;
    long bi = np;
    for (long k = threadIdx.x; k < np; k += PIV_THREADS) {
        const double d = D[k];
        if (d > best) {
            best = d;
            bi = k;
        }
    }
    sv[threadIdx.x] = best;
    si[threadIdx.x] = bi;
    __syncthreads();
    for (int w = PIV_THREADS / 2; w > 0; w >>= 1) {
        if (threadIdx.x < w) {
            const double o = sv[threadIdx.x + w];
            const long oi = si[threadIdx.x + w];
            if (o > sv[threadIdx.x] || (o == sv[threadIdx.x] && oi < si[threadIdx.x])) {
                sv[threadIdx.x] = o;
                si[threadIdx.x] = oi;
            }
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        const double dp = sv[0];
        st->dmax = dp;
        gather = 0;
        nv_sh = st->nvec;
        if (st->nvec >= cap) {
            st->done = 2;  // the reference's loop bound: stop unconverged
        } else if (dp < thr) {
            st->done = 1;
        } else {
            st->pivot = (int)si[0];
            st->dpiv = dp;
            st->pending = 1;
            gather = 1;
        }
    }
    __syncthreads();
    // gather the pivot's entries of the earlier vectors: Lp[k] = L_k(p)
    if (gather) {
        const long p = si[0];
        for (int k = threadIdx.x; k < nv_sh; k += PIV_THREADS) Lp[k] = L[(long)k * np + p];
    }
}

// New vector nvec and the diagonal update (cholesky_eri.cc:107-113, 169-203).
// Loads run UNR vectors ahead of the (sequential, k-ordered) differences.
constexpr int UNR = 16;

__global__ __launch_bounds__(VEC_THREADS) void cd_vector_kernel(const double* __restrict__ M, long np,
                                                                double* __restrict__ L, double* __restrict__ D,
                                                                const double* __restrict__ Lp,
                                                                const CdState* st) {
#pragma clang fp contract(off)
    if (st->done || !st->pending) return;
    const int nv = st->nvec;
    const long p = st->pivot;
    const double nf = sqrt(st->dpiv);
    const long ij = (long)blockIdx.x * VEC_THREADS + threadIdx.x;
    if (ij >= np) return;
    double val = M[ij * np + p];
    const double* Lc = L + ij;
    int k = 0;
    for (; k + UNR <= nv; k += UNR) {
        double a[UNR], b[UNR];
#pragma unroll
        for (int u = 0; u < UNR; ++u) {
            a[u] = Lc[(long)(k + u) * np];
            b[u] = Lp[k + u];
        }
#pragma unroll
        for (int u = 0; u < UNR; ++u) val = val - a[u] * b[u];
    }
    for (; k < nv; ++k) val = val - Lc[(long)k * np] * Lp[k];
    if (nf > 1e-12) val = val / nf;
    L[(long)nv * np + ij] = val;
    double d = D[ij] - val * val;
    if (d < 0.0) d = 0.0;
    D[ij] = d;
}

}  // namespace

extern "C" int mshqc_hip_cholesky(mshqc_hip_ctx* ctx, const double* eri, int n, double threshold,
                                  int max_vectors, double* L, int* n_vectors, int* converged,
                                  double* max_residual) {
    int rc = ctx_enter(ctx);
    if (rc) return rc;
    CHECK_ARG(ctx, eri && L && n_vectors, "cholesky: null pointer");
    CHECK_ARG(ctx, n > 0 && max_vectors > 0 && threshold > 0.0, "cholesky: bad arguments");
    hipEventRecord(ctx->ev0, ctx->stream);
    const long np = (long)n * n;
    double* D = static_cast<double*>(workspace(ctx, WS_TMP, sizeof(double) * (size_t)np));
    CdState* st = static_cast<CdState*>(workspace(ctx, WS_MISC, sizeof(CdState)));
    double* Lp = static_cast<double*>(workspace(ctx, WS_TMP2, sizeof(double) * (size_t)max_vectors));
    if (!D || !st || !Lp) return MSHQC_HIP_ERR_ALLOC;
    cd_init_kernel<<<(unsigned)((np + 255) / 256), 256, 0, ctx->stream>>>(eri, np, D, st);
    LAUNCH_CHECK(ctx);
    const unsigned vgrid = (unsigned)((np + VEC_THREADS - 1) / VEC_THREADS);
    CdState h{};
    for (;;) {
        for (int s = 0; s < STEPS_PER_BATCH; ++s) {
            cd_pivot_kernel<<<1, PIV_THREADS, 0, ctx->stream>>>(D, np, threshold, max_vectors, L, Lp, st);
            LAUNCH_CHECK(ctx);
            cd_vector_kernel<<<vgrid, VEC_THREADS, 0, ctx->stream>>>(eri, np, L, D, Lp, st);
            LAUNCH_CHECK(ctx);
        }
        HIP_TRY(ctx, hipMemcpyAsync(&h, st, sizeof(CdState), hipMemcpyDeviceToHost, ctx->stream));
        HIP_TRY(ctx, hipStreamSynchronize(ctx->stream));
        if (h.done) break;
    }
    *n_vectors = h.nvec + h.pending;
    if (converged) *converged = h.done == 1;
    if (max_residual) *max_residual = h.dmax;
    hipEventRecord(ctx->ev1, ctx->stream);
    if ((rc = ctx_finish(ctx))) return rc;
    hipEventElapsedTime(&ctx->last_ms, ctx->ev0, ctx->ev1);
    return MSHQC_HIP_OK;
}
