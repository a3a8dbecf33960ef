// Pivoted incomplete Cholesky of the ERI matrix on the device, FP64
// (cholesky_eri.cc:40-203).
//
// The ERI matrix M[(ij)][(kl)] = (ij|kl) (n^2 x n^2, row-major) stays in HBM.
// After one pivot scan, each step is one kernel:
//   L_new(ij) = (M[ij][p] - sum_k L_k(ij) L_k(p)) / sqrt(D_p),
//   D(ij) -= L_new(ij)^2, clamped at 0;
//   each workgroup writes the first maximum of its updated D entries, and
//   the last workgroup to finish (an arrival counter) reduces those in
//   index order to the next pivot, stops the decomposition when the maximum
//   is below the threshold or the output is full, and gathers the pivot's
//   entries of the earlier vectors
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
    unsigned arrived;  // workgroups of the current step that have finished
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
        st->arrived = 0;
    }
}

// Commit the pending vector, then pick the next pivot from the first
// maximum dp at index pi: stop at the output bound or below the threshold
// (cholesky_eri.cc:155-167 and the loop bounds around them).  One thread.
__device__ __forceinline__ bool decide_pivot(double dp, long pi, double thr, int cap, CdState* st) {
    if (st->pending) {
        st->nvec += 1;
        st->pending = 0;
    }
    st->dmax = dp;
    if (st->nvec >= cap) {
        st->done = 2;  // the reference's loop bound: stop unconverged
        return false;
    }
    if (dp < thr) {
        st->done = 1;
        return false;
    }
    st->pivot = (int)pi;
    st->dpiv = dp;
    st->pending = 1;
    return true;
}

// First maximum over a workgroup of (v, i) pairs, strictly greater wins and
// ties keep the lower index; the result is in sv[0], si[0].
template <int NT>
__device__ __forceinline__ void block_first_max(double v, long i, double* sv, long* si) {
    sv[threadIdx.x] = v;
    si[threadIdx.x] = i;
    __syncthreads();
    for (int w = NT / 2; w > 0; w >>= 1) {
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
}

// The first pivot: the first maximum of the initial diagonal (strictly
// greater wins; ties keep the lower index), as cholesky_eri.cc:155-167 scans
// it.  Later pivots come from cd_step_kernel.
__global__ __launch_bounds__(PIV_THREADS) void cd_pivot_kernel(const double* __restrict__ D, long np,
                                                               double thr, int cap, CdState* st) {
    __shared__ double sv[PIV_THREADS];
    __shared__ long si[PIV_THREADS];
    double best = -INFINITY;
    long bi = np;
    for (long k = threadIdx.x; k < np; k += PIV_THREADS) {
        const double d = D[k];
        if (d > best) {
            best = d;
            bi = k;
        }
    }
    block_first_max<PIV_THREADS>(best, bi, sv, si);
    if (threadIdx.x == 0) decide_pivot(sv[0], si[0], thr, cap, st);
}

// New vector nvec and the diagonal update (cholesky_eri.cc:107-113, 169-203).
// Loads run UNR vectors ahead of the (sequential, k-ordered) differences.
constexpr int UNR = 16;

// One step: new vector nvec and the diagonal update (cholesky_eri.cc:107-113,
// 169-203; loads run UNR vectors ahead of the sequential, k-ordered
// differences), then the next pivot from the updated diagonal by the last
// workgroup.
// bmax/bidx hold one first maximum per workgroup.
__global__ __launch_bounds__(VEC_THREADS) void cd_step_kernel(const double* __restrict__ M, long np,
                                                              double* __restrict__ L, double* __restrict__ D,
                                                              double* __restrict__ Lp, CdState* st, double thr,
                                                              int cap, double* __restrict__ bmax,
                                                              long* __restrict__ bidx) {
#pragma clang fp contract(off)
    __shared__ double sv[VEC_THREADS];
    __shared__ long si[VEC_THREADS];
    __shared__ int last, gather, nv_sh;
    if (st->done || !st->pending) return;
    const int nv = st->nvec;
    const long p = st->pivot;
    const double nf = sqrt(st->dpiv);
    const long ij = (long)blockIdx.x * VEC_THREADS + threadIdx.x;
    double d = -INFINITY;
    if (ij < np) {
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
        d = D[ij] - val * val;
        if (d < 0.0) d = 0.0;
        D[ij] = d;
    }
    block_first_max<VEC_THREADS>(d, ij < np ? ij : np, sv, si);
    if (threadIdx.x == 0) {
        bmax[blockIdx.x] = sv[0];
        bidx[blockIdx.x] = si[0];
        __threadfence();
        last = atomicAdd(&st->arrived, 1u) == gridDim.x - 1;
    }
    __syncthreads();
    if (!last) return;
    // The last workgroup: every other one has written its vector entries,
    // diagonal and maximum (each fenced before its arrival).
    __threadfence();
    double best = -INFINITY;
    long bi = np;
    for (int b = threadIdx.x; b < (int)gridDim.x; b += VEC_THREADS) {
        const double v = bmax[b];
        if (v > best) {
            best = v;
            bi = bidx[b];
        }
    }
    __syncthreads();
    block_first_max<VEC_THREADS>(best, bi, sv, si);
    if (threadIdx.x == 0) {
        st->arrived = 0;
        gather = decide_pivot(sv[0], si[0], thr, cap, st) ? 1 : 0;
        nv_sh = st->nvec;
    }
    __syncthreads();
    if (gather) {
        const long pn = si[0];
        for (int k = threadIdx.x; k < nv_sh; k += VEC_THREADS) Lp[k] = L[(long)k * np + pn];
    }
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
    const unsigned vgrid = (unsigned)((np + VEC_THREADS - 1) / VEC_THREADS);
    char* blk = static_cast<char*>(workspace(ctx, WS_PART, (sizeof(double) + sizeof(long)) * (size_t)vgrid + 256));
    if (!D || !st || !Lp || !blk) return MSHQC_HIP_ERR_ALLOC;
    double* bmax = reinterpret_cast<double*>(blk);
    long* bidx = reinterpret_cast<long*>(blk + sizeof(double) * (size_t)vgrid);
    cd_init_kernel<<<(unsigned)((np + 255) / 256), 256, 0, ctx->stream>>>(eri, np, D, st);
    LAUNCH_CHECK(ctx);
    cd_pivot_kernel<<<1, PIV_THREADS, 0, ctx->stream>>>(D, np, threshold, max_vectors, st);
    LAUNCH_CHECK(ctx);
    CdState h{};
    for (;;) {
        for (int s = 0; s < STEPS_PER_BATCH; ++s) {
            cd_step_kernel<<<vgrid, VEC_THREADS, 0, ctx->stream>>>(eri, np, L, D, Lp, st, threshold, max_vectors,
                                                                   bmax, bidx);
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
