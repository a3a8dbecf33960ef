// CD-RMP3 second-order amplitudes and energy (cholesky_rmp3.cc:79-200) from
// the MO Cholesky blocks B (ov), Loo (oo), Lvv (vv) and the first-order
// amplitudes t1, as FP64 MFMA GEMMs:
//
//   particle-particle: (ac|bd) = sum_Q Lvv_ac Lvv_bd is built for a block of
//     a at a time as W[(a c)][(b d)] (GEMM over Q), permuted to
//     W'[(a b)][(c d)], and contracted with t1(ijcd) - t1(ijdc):
//     R[(ij)][(ab)] += 1/2 T-[(ij)][(cd)] W'[(ab)][(cd)]
//   hole-hole: G4[(ki)][(lj)] = sum_Q Loo_ki Loo_lj, U[(ij)][(kl)] =
//     G4[(ki)][(lj)] - G4[(kj)][(li)], R += 1/2 U t1
//   rings: G2[(ia)][(jb)] = (ia|jb), Y[(jb)][(kc)] = 2 (jb|kc) - (kb|jc),
//     Ts[(ia)][(kc)] = t1(ikac), R1 = Ts Y^T; the two ring terms of the
//     reference are R1[(ia)][(jb)] and R1[(jb)][(ia)]
//   t2_2 = (R - rings) / D, E3 = sum [2 (ia|jb) - (ib|ja)] t2_2
#include <algorithm>
#include <cmath>

#include "common.h"

namespace {

constexpr int EW = 256;  // threads of the elementwise kernels

__global__ void tminus_kernel(const double* __restrict__ t1, long oo, int v, double* __restrict__ Tm) {
    const long total = oo * v * v;
    for (long t = (long)blockIdx.x * EW + threadIdx.x; t < total; t += (long)gridDim.x * EW) {
        const int d = (int)(t % v);
        const long r = t / v;
        const int c = (int)(r % v);
        const long ij = r / v;
        Tm[t] = t1[t] - t1[(ij * v + d) * v + c];
    }
}

// W'[((a*v + b)*v + c)*v + d] = W[((a*v + c)*v + b)*v + d], a < na
__global__ void permute_w_kernel(const double* __restrict__ W, int na, int v, double* __restrict__ Wp) {
    const long total = (long)na * v * v * v;
    for (long t = (long)blockIdx.x * EW + threadIdx.x; t < total; t += (long)gridDim.x * EW) {
        const int d = (int)(t % v);
        long r = t / v;
        const int c = (int)(r % v);
        r /= v;
        const int b = (int)(r % v);
        const long a = r / v;
        Wp[t] = W[((a * v + c) * v + b) * v + d];
    }
}

// U[(i o + j) o^2 + (k o + l)] = G4[(k o + i) o^2 + (l o + j)] - G4[(k o + j) o^2 + (l o + i)]
__global__ void u_kernel(const double* __restrict__ G4, int o, double* __restrict__ U) {
    const long oo = (long)o * o, total = oo * oo;
    for (long t = (long)blockIdx.x * EW + threadIdx.x; t < total; t += (long)gridDim.x * EW) {
        const int l = (int)(t % o);
        long r = t / o;
        const int k = (int)(r % o);
        r /= o;
        const int j = (int)(r % o);
        const int i = (int)(r / o);
        U[t] = G4[((long)k * o + i) * oo + (long)l * o + j] - G4[((long)k * o + j) * oo + (long)l * o + i];
    }
}

// Y[(j v + b) ov + (k v + c)] = 2 G2[(j v + b) ov + (k v + c)] - G2[(k v + b) ov + (j v + c)]
__global__ void y_kernel(const double* __restrict__ G2, int o, int v, double* __restrict__ Y) {
    const long ov = (long)o * v, total = ov * ov;
    for (long t = (long)blockIdx.x * EW + threadIdx.x; t < total; t += (long)gridDim.x * EW) {
        const int c = (int)(t % v);
        long r = t / v;
        const int k = (int)(r % o);
        r /= o;
        const int b = (int)(r % v);
        const int j = (int)(r / v);
        Y[t] = 2.0 * G2[t] - G2[((long)k * v + b) * ov + (long)j * v + c];
    }
}

// Ts[((i v + a) o + k) v + c] = t1[((i o + k) v + a) v + c]
__global__ void ts_kernel(const double* __restrict__ t1, int o, int v, double* __restrict__ Ts) {
    const long total = (long)o * o * v * v;
    for (long t = (long)blockIdx.x * EW + threadIdx.x; t < total; t += (long)gridDim.x * EW) {
        const int c = (int)(t % v);
        long r = t / v;
        const int k = (int)(r % o);
        r /= o;
        const int a = (int)(r % v);
        const int i = (int)(r / v);
        Ts[t] = t1[(((long)i * o + k) * v + a) * v + c];
    }
}

// t2_2 and the energy partials (cholesky_rmp3.cc:159-178).
__global__ __launch_bounds__(EW) void rmp3_final_kernel(const double* __restrict__ R, const double* __restrict__ R1,
                                                       const double* __restrict__ G2, int o, int v,
                                                       const double* __restrict__ eo, const double* __restrict__ ev,
                                                       double* __restrict__ t22, double* __restrict__ part) {
    const long ov = (long)o * v, total = (long)o * o * v * v;
    double e = 0.0;
    for (long t = (long)blockIdx.x * EW + threadIdx.x; t < total; t += (long)gridDim.x * EW) {
        const int b = (int)(t % v);
        long r = t / v;
        const int a = (int)(r % v);
        r /= v;
        const int j = (int)(r % o);
        const int i = (int)(r / o);
        const long ia = (long)i * v + a, jb = (long)j * v + b;
        const double D = eo[i] + eo[j] - ev[a] - ev[b];
        double val = 0.0;
        if (fabs(D) >= 1e-12) val = (R[t] - R1[ia * ov + jb] - R1[jb * ov + ia]) / D;
        if (t22) t22[t] = val;
        e += (2.0 * G2[ia * ov + jb] - G2[((long)i * v + b) * ov + (long)j * v + a]) * val;
    }
    for (int m = 32; m > 0; m >>= 1) e += __shfl_xor(e, m, 64);
    __shared__ double sh[EW / 64];
    if ((threadIdx.x & 63) == 0) sh[threadIdx.x >> 6] = e;
    __syncthreads();
    if (threadIdx.x == 0) part[blockIdx.x] = ((sh[0] + sh[1]) + sh[2]) + sh[3];
}

unsigned ew_grid(long total) { return (unsigned)std::min<long>(8192, std::max<long>(1, (total + EW - 1) / EW)); }

}  // namespace

extern "C" int mshqc_hip_rmp3(mshqc_hip_ctx* ctx, const double* B, const double* Loo, const double* Lvv, int Q,
                              int o, int v, const double* t1, const double* eps_o, const double* eps_v,
                              double* E3, double* t2_2) {
    int rc = ctx_enter(ctx);
    if (rc) return rc;
    CHECK_ARG(ctx, E3 != nullptr, "rmp3: null energy pointer");
    CHECK_ARG(ctx, Q > 0 && o >= 0 && v >= 0, "rmp3: bad dimensions");
    *E3 = 0.0;
    if (o == 0 || v == 0) return MSHQC_HIP_OK;
    CHECK_ARG(ctx, B && Loo && Lvv && t1 && eps_o && eps_v, "rmp3: null pointer");
    hipEventRecord(ctx->ev0, ctx->stream);
    const long oo = (long)o * o, vv = (long)v * v, ov = (long)o * v, oovv = oo * vv;
    const int na = (int)std::max<long>(1, std::min<long>(v, (long)(ctx->zblock_elems / (size_t)(vv * v))));
    double* R = static_cast<double*>(workspace(ctx, WS_MP3_A, sizeof(double) * oovv));
    double* Tm = static_cast<double*>(workspace(ctx, WS_MP3_B, sizeof(double) * oovv));
    double* G2 = static_cast<double*>(workspace(ctx, WS_MP3_C, sizeof(double) * oovv));
    double* Y = static_cast<double*>(workspace(ctx, WS_MP3_D, sizeof(double) * oovv));
    double* R1 = static_cast<double*>(workspace(ctx, WS_OMP2_V, sizeof(double) * oovv));
    double* Wb = static_cast<double*>(workspace(ctx, WS_OMP2_T, sizeof(double) * 2 * (size_t)na * vv * v));
    double* G4 = static_cast<double*>(workspace(ctx, WS_TMP, sizeof(double) * 2 * (size_t)oo * oo));
    const unsigned fgrid = ew_grid(oovv);
    double* part = static_cast<double*>(workspace(ctx, WS_PART, sizeof(double) * fgrid));
    double* slot = static_cast<double*>(workspace(ctx, WS_MISC, sizeof(double) * 2));
    if (!R || !Tm || !G2 || !Y || !R1 || !Wb || !G4 || !part || !slot) return MSHQC_HIP_ERR_ALLOC;
    double* Wp = Wb + (size_t)na * vv * v;
    double* U = G4 + (size_t)oo * oo;
    if ((rc = fill_zero(ctx, R, oovv))) return rc;

    // particle-particle ladder (cholesky_rmp3.cc:120-127)
    tminus_kernel<<<ew_grid(oovv), EW, 0, ctx->stream>>>(t1, oo, v, Tm);
    LAUNCH_CHECK(ctx);
    for (int a0 = 0; a0 < v; a0 += na) {
        const int nb = std::min(na, v - a0);
        GemmF64 gw;  // W[(a c)][(b d)] = sum_Q Lvv[Q][a0 + a][c] Lvv[Q][b][d]
        gw.A = Lvv + (long)a0 * v;
        gw.lda = vv;
        gw.transA = true;
        gw.X = Lvv;
        gw.ldx = vv;
        gw.C = Wb;
        gw.ldc = vv;
        gw.M = nb * v;
        gw.N = (int)vv;
        gw.K = Q;
        { ProfScope ps(ctx, PROF_GEMM); rc = gemm_f64(ctx, gw); }
        if (rc) return rc;
        permute_w_kernel<<<ew_grid((long)nb * vv * v), EW, 0, ctx->stream>>>(Wb, nb, v, Wp);
        LAUNCH_CHECK(ctx);
        // R[(ij)][(a0 + a, b)] += 1/2 sum_cd T-[(ij)][(cd)] W'[(ab)][(cd)], the
        // v^2-long contraction cut into v chunks of one c each (batch-reduced)
        GemmF64 gp;
        gp.A = Tm;
        gp.lda = vv;
        gp.sA = v;
        gp.X = Wp;
        gp.ldx = vv;
        gp.sX = v;
        gp.transX = true;
        gp.C = R + (long)a0 * v;
        gp.ldc = vv;
        gp.M = (int)oo;
        gp.N = nb * v;
        gp.K = v;
        gp.batch = v;
        gp.reduce_batch = true;
        gp.splits = (int)std::min<long>(std::min<long>(v, 256), std::max<long>(1, (1L << 25) / (oo * nb * v)));
        gp.alpha = 0.5;
        gp.accumulate = true;
        { ProfScope ps(ctx, PROF_GEMM); rc = gemm_f64(ctx, gp); }
        if (rc) return rc;
    }

    // hole-hole ladder (:130-137)
    {
        GemmF64 g4;
        g4.A = Loo;
        g4.lda = oo;
        g4.transA = true;
        g4.X = Loo;
        g4.ldx = oo;
        g4.C = G4;
        g4.ldc = oo;
        g4.M = g4.N = (int)oo;
        g4.K = Q;
        { ProfScope ps(ctx, PROF_GEMM); rc = gemm_f64(ctx, g4); }
        if (rc) return rc;
        u_kernel<<<ew_grid(oo * oo), EW, 0, ctx->stream>>>(G4, o, U);
        LAUNCH_CHECK(ctx);
        GemmF64 gh;
        gh.A = U;
        gh.lda = oo;
        gh.X = t1;
        gh.ldx = vv;
        gh.C = R;
        gh.ldc = vv;
        gh.M = (int)oo;
        gh.N = (int)vv;
        gh.K = (int)oo;
        gh.alpha = 0.5;
        gh.accumulate = true;
        { ProfScope ps(ctx, PROF_GEMM); rc = gemm_f64(ctx, gh); }
        if (rc) return rc;
    }

    // rings (:140-157)
    {
        GemmF64 g2;
        g2.A = B;
        g2.lda = ov;
        g2.transA = true;
        g2.X = B;
        g2.ldx = ov;
        g2.C = G2;
        g2.ldc = ov;
        g2.M = g2.N = (int)ov;
        g2.K = Q;
        { ProfScope ps(ctx, PROF_GEMM); rc = gemm_f64(ctx, g2); }
        if (rc) return rc;
        y_kernel<<<ew_grid(oovv), EW, 0, ctx->stream>>>(G2, o, v, Y);
        LAUNCH_CHECK(ctx);
        ts_kernel<<<ew_grid(oovv), EW, 0, ctx->stream>>>(t1, o, v, Tm);  // T- no longer needed
        LAUNCH_CHECK(ctx);
        GemmF64 gr;  // R1[(ia)][(jb)] = sum_kc Ts[(ia)][(kc)] Y[(jb)][(kc)]
        gr.A = Tm;
        gr.lda = ov;
        gr.X = Y;
        gr.ldx = ov;
        gr.transX = true;
        gr.C = R1;
        gr.ldc = ov;
        gr.M = gr.N = (int)ov;
        gr.K = (int)ov;
        { ProfScope ps(ctx, PROF_GEMM); rc = gemm_f64(ctx, gr); }
        if (rc) return rc;
    }

    rmp3_final_kernel<<<fgrid, EW, 0, ctx->stream>>>(R, R1, G2, o, v, eps_o, eps_v, t2_2, part);
    LAUNCH_CHECK(ctx);
    if ((rc = reduce_sum(ctx, part, fgrid, slot, false))) return rc;
    HIP_TRY(ctx, hipMemcpyAsync(E3, slot, sizeof(double), hipMemcpyDeviceToHost, ctx->stream));
    hipEventRecord(ctx->ev1, ctx->stream);
    if ((rc = ctx_finish(ctx))) return rc;
    hipEventElapsedTime(&ctx->last_ms, ctx->ev0, ctx->ev1);
    return MSHQC_HIP_OK;
}
