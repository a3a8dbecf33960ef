// CD-OMP2 amplitudes, amplitude energy and one-particle density blocks
// (cholesky_omp2.cc:94-326) from the MO Cholesky vectors B^Q_ia of each spin.
//
// The reference stores the three amplitude tensors t_aa, t_bb, t_ab whole.
// Here each spin block (X, Y) is walked in blocks of its second occupied
// index k:
//   V[(i a)][(k b)] = (ia|kb) = sum_Q B_X^Q_ia B_Y^Q_kb     FP64 MFMA GEMM
//   T[k][i][a][b]   = amplitude (antisymmetrised for X = Y)  elementwise
//   G_oo^X += s_o sum_k T_k T_k^T,  G_vv^X += s_v sum_{k,i} T_ki T_ki^T
//   (opposite spin also G_vv^Y += sum_{k,i} T_ki^T T_ki and G_oo^Y)
// so memory is bounded by two blocks, not by o^2 v^2 per spin pair.
#include <algorithm>
#include <cmath>

#include "common.h"
#include "mfma_f64.h"

namespace {

constexpr int AMP_THREADS = 256;
constexpr int AMP_MAX_GROUPS = 4096;

// Amplitudes of one k block.  V row (i, a), column (kk, b) holds (ia|kb),
// k = k0 + kk.  same != 0: t = ((ia|kb) - (ib|ka)) / D, zero when i == k or
// a == b (cholesky_omp2.cc:115-137); else t = (ia|kb) / D (:161-180).
// |D| < 1e-12 leaves t = 0.  part[g] = sum of t^2 D over workgroup g's
// elements (fixed grid-stride assignment, so the sum is reproducible).
// full != 0: V is the whole symmetric (ia|kb) matrix (same spin, one block),
// so (ib|ka) is read as its mirror (ka|ib), contiguous along b.
__global__ __launch_bounds__(AMP_THREADS) void omp2_amp_kernel(
    const double* __restrict__ V, long ldv, int oX, int vX, int vY, int k0, int nk,
    const double* __restrict__ eoX, const double* __restrict__ evX, const double* __restrict__ eoY,
    const double* __restrict__ evY, int same, int full, double* __restrict__ T, double* __restrict__ part) {
    const long total = (long)nk * oX * vX * vY;
    double e = 0.0;
    for (long t = (long)blockIdx.x * AMP_THREADS + threadIdx.x; t < total;
         t += (long)gridDim.x * AMP_THREADS) {
        const int b = (int)(t % vY);
        long r = t / vY;
        const int a = (int)(r % vX);
        r /= vX;
        const int i = (int)(r % oX);
        const int kk = (int)(r / oX);
        const int k = k0 + kk;
        const double D = eoX[i] + eoY[k] - evX[a] - evY[b];
        const double g = V[((long)i * vX + a) * ldv + (long)kk * vY + b];
        double val = 0.0;
        if (fabs(D) >= 1e-12) {
            if (!same) {
                val = g / D;
            } else if (i != k && a != b) {
                const double gx = full ? V[((long)k * vX + a) * ldv + (long)i * vY + b]
                                       : V[((long)i * vX + b) * ldv + (long)kk * vY + a];
                val = (g - gx) / D;
            }
        }
        T[t] = val;
        e += val * val * D;
    }
    for (int m = 32; m > 0; m >>= 1) e += __shfl_xor(e, m, 64);
    __shared__ double sh[AMP_THREADS / 64];
    if ((threadIdx.x & 63) == 0) sh[threadIdx.x >> 6] = e;
    __syncthreads();
    if (threadIdx.x == 0) part[blockIdx.x] = ((sh[0] + sh[1]) + sh[2]) + sh[3];
}

// Small symmetric products for the G_oo contractions (o x o outputs, o <= 64):
// part[wg][m][n] = sum over workgroup wg's batch entries b of
// sum_k A_b[m][k] A_b[n][k], A_b[m][k] = A[bo sA + bi sA2 + m lda + k]
// (b = bo inner + bi).  Each wave keeps the whole G x G grid of 16 x 16
// FP64 MFMA tiles and walks every fourth entry; one loaded fragment serves
// as both MFMA operands.  The four waves are summed through LDS in wave
// order, so the partials are reproducible.
constexpr int SY_THREADS = 256;

template <int G>
__global__ __launch_bounds__(SY_THREADS) void omp2_syrk_small_kernel(const double* __restrict__ A, long lda,
                                                                     long sA, long sA2, int inner, int M,
                                                                     int K, int batch, int bper,
                                                                     double* __restrict__ part) {
    __shared__ double red[64 * 64];
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int r = lane & 15, kq = lane >> 4;
    v4d acc[G][G];
#pragma unroll
    for (int g = 0; g < G; ++g)
#pragma unroll
        for (int h = 0; h < G; ++h) acc[g][h] = v4d{0.0, 0.0, 0.0, 0.0};
    const int b_lo = blockIdx.x * bper, b_hi = min(batch, b_lo + bper);
    for (int b = b_lo + wave; b < b_hi; b += 4) {
        const int bo = b / inner, bi = b - bo * inner;
        const double* Ab = A + (long)bo * sA + (long)bi * sA2;
        // the next step's fragments load while this step's MFMAs run
        auto fetch = [&](int k0, double(&f)[G]) {
            const int k = k0 + kq;
#pragma unroll
            for (int g = 0; g < G; ++g) {
                const int m = 16 * g + r;
                f[g] = (m < M && k < K) ? Ab[(long)m * lda + k] : 0.0;
            }
        };
        double f[G], fn[G];
        fetch(0, f);
        for (int k0 = 0; k0 < K; k0 += 4) {
            fetch(k0 + 4, fn);
#pragma unroll
            for (int g = 0; g < G; ++g)
#pragma unroll
                for (int h = 0; h < G; ++h) acc[g][h] = mfma_f64(f[g], f[h], acc[g][h]);
#pragma unroll
            for (int g = 0; g < G; ++g) f[g] = fn[g];
        }
    }
    for (int w = 0; w < 4; ++w) {
        if (wave == w) {
#pragma unroll
            for (int g = 0; g < G; ++g)
#pragma unroll
                for (int h = 0; h < G; ++h)
#pragma unroll
                    for (int q = 0; q < 4; ++q) {
                        const int idx = (16 * g + mfma_row(lane, q)) * 64 + 16 * h + mfma_col(lane);
                        red[idx] = w == 0 ? acc[g][h][q] : red[idx] + acc[g][h][q];
                    }
        }
        __syncthreads();
    }
    double* out = part + (long)blockIdx.x * M * M;
    for (int e = threadIdx.x; e < M * M; e += SY_THREADS) out[e] = red[(e / M) * 64 + e % M];
}

// C[m][n] (+)= alpha * sum_z part[z][m][n]: one workgroup per element,
// thread t sums z = t, t + 256, ... and a fixed tree adds the threads, so
// the order is the same on every call.
__global__ __launch_bounds__(256) void omp2_sum_parts_kernel(const double* __restrict__ part, int nz, int M,
                                                             double* __restrict__ C, long ldc, double alpha) {
    __shared__ double sh[256];
    const int e = blockIdx.x;
    double s = 0.0;
    for (int z = threadIdx.x; z < nz; z += 256) s += part[(long)z * M * M + e];
    sh[threadIdx.x] = s;
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
        if (threadIdx.x < w) sh[threadIdx.x] += sh[threadIdx.x + w];
        __syncthreads();
    }
    if (threadIdx.x == 0) C[(long)(e / M) * ldc + e % M] += alpha * sh[0];
}

// C (M x M, accumulated) += alpha * sum_b A_b A_b^T for M <= 64 (else the
// generic batch-reduced GEMM).
int small_syrk(mshqc_hip_ctx* ctx, const double* A, long lda, long sA, long sA2, int inner, int M, int K,
               int batch, double alpha, double* C, long ldc) {
    if (M <= 0 || batch <= 0) return MSHQC_HIP_OK;
    if (M > 64) {
        GemmF64 g;
        g.A = A;
        g.lda = lda;
        g.sA = sA;
        g.sA2 = sA2;
        g.X = A;
        g.ldx = lda;
        g.sX = sA;
        g.sX2 = sA2;
        g.inner = inner;
        g.transX = true;
        g.C = C;
        g.ldc = ldc;
        g.M = g.N = M;
        g.K = K;
        g.batch = batch;
        g.reduce_batch = true;
        g.splits = std::min<long>(batch, std::max<long>(1, (1L << 25) / ((long)M * M)));
        g.splits = std::min(g.splits, 1024);
        g.alpha = alpha;
        g.accumulate = true;
        return gemm_f64(ctx, g);
    }
    const int nwg = std::min(1024, (batch + 3) / 4);
    const int bper = (batch + nwg - 1) / nwg;
    const int nz = (batch + bper - 1) / bper;
    double* part = static_cast<double*>(workspace(ctx, WS_TMP2, sizeof(double) * (size_t)nz * M * M));
    if (!part) return MSHQC_HIP_ERR_ALLOC;
    const int G = (M + 15) / 16;
    switch (G) {
        case 1: omp2_syrk_small_kernel<1><<<nz, SY_THREADS, 0, ctx->stream>>>(A, lda, sA, sA2, inner, M, K, batch, bper, part); break;
        case 2: omp2_syrk_small_kernel<2><<<nz, SY_THREADS, 0, ctx->stream>>>(A, lda, sA, sA2, inner, M, K, batch, bper, part); break;
        case 3: omp2_syrk_small_kernel<3><<<nz, SY_THREADS, 0, ctx->stream>>>(A, lda, sA, sA2, inner, M, K, batch, bper, part); break;
        default: omp2_syrk_small_kernel<4><<<nz, SY_THREADS, 0, ctx->stream>>>(A, lda, sA, sA2, inner, M, K, batch, bper, part); break;
    }
    LAUNCH_CHECK(ctx);
    omp2_sum_parts_kernel<<<M * M, 256, 0, ctx->stream>>>(part, nz, M, C, ldc, alpha);
    LAUNCH_CHECK(ctx);
    return MSHQC_HIP_OK;
}

constexpr int SYM_TILE = 64;  // tile of the shared mirror (mirror_upper)

// P[p] = T[k][i] (blocks of vv doubles) for the pairs i < k of an o-orbital
// block, p = k (k - 1) / 2 + i.
__global__ __launch_bounds__(256) void omp2_pack_lower_kernel(const double* __restrict__ T, int o, long vv,
                                                              double* __restrict__ P) {
    const long npair = (long)o * (o - 1) / 2;
    for (long p = blockIdx.x; p < npair; p += gridDim.x) {
        // k from p: the largest k with k (k - 1) / 2 <= p
        int k = (int)((1.0 + sqrt(1.0 + 8.0 * (double)p)) / 2.0);
        while ((long)k * (k - 1) / 2 > p) --k;
        while ((long)(k + 1) * k / 2 <= p) ++k;
        const int i = (int)(p - (long)k * (k - 1) / 2);
        const double* src = T + ((long)k * o + i) * vv;
        double* dst = P + p * vv;
        for (long e = threadIdx.x; e < vv; e += 256) dst[e] = src[e];
    }
}

struct SpinPair {
    const double* BX;  // [Q][oX][vX]
    const double* BY;  // [Q][oY][vY]
    int oX, vX, oY, vY;
    const double *eoX, *evX, *eoY, *evY;
    bool same;
};

struct Want {
    double* GooX = nullptr;  // [oX][oX], scaled by -1/2 (same) or -1
    double* GvvX = nullptr;  // [vX][vX], scaled by 1/2 (same) or 1
    double* GvvY = nullptr;  // [vY][vY] (opposite spin only)
    double* GooY = nullptr;  // [oY][oY] (opposite spin, single k block only)
    double* e_slot = nullptr;
};

// Batch layers of a batch-reduced GEMM: at most 1024, and partials
// ([splits][M][N]) within 2^25 doubles.
int red_splits(int batch, long M, long N) {
    const long cap = 1L << 25;
    const long s = std::min<long>(batch, std::max<long>(1, cap / std::max<long>(M * N, 1)));
    return (int)std::min<long>(s, 1024);
}

// Largest k block whose V and T blocks (oX vX vY doubles per k) fit `cap`.
int k_block(const SpinPair& s, size_t cap) {
    const size_t per_k = (size_t)s.oX * s.vX * s.vY;
    return (int)std::max<size_t>(1, std::min<size_t>((size_t)s.oY, cap / std::max<size_t>(per_k, 1)));
}

int spin_pair(mshqc_hip_ctx* ctx, int Q, const SpinPair& s, const Want& w, int kb) {
    const long per_k = (long)s.oX * s.vX * s.vY;
    double* V = static_cast<double*>(workspace(ctx, WS_OMP2_V, sizeof(double) * (size_t)kb * per_k));
    double* T = static_cast<double*>(workspace(ctx, WS_OMP2_T, sizeof(double) * (size_t)kb * per_k));
    double* part = static_cast<double*>(workspace(ctx, WS_PART, sizeof(double) * AMP_MAX_GROUPS));
    if (!V || !T || !part) return MSHQC_HIP_ERR_ALLOC;
    const double so = s.same ? -0.5 : -1.0, sv = s.same ? 0.5 : 1.0;
    int rc;
    for (int k0 = 0; k0 < s.oY; k0 += kb) {
        const int nk = std::min(kb, s.oY - k0);
        const long ldv = (long)nk * s.vY;
        // V = B_X^T B_Y[:, k0:k0+nk, :] over Q, as blocks of rows r0.. and
        // columns c0..
        auto v_block = [&](long r0, long rn, long c0, long cn) {
            GemmF64 gv;
            gv.A = s.BX + r0;
            gv.lda = (long)s.oX * s.vX;
            gv.transA = true;
            gv.X = s.BY + (long)k0 * s.vY + c0;
            gv.ldx = (long)s.oY * s.vY;
            gv.C = V + r0 * ldv + c0;
            gv.ldc = ldv;
            gv.M = (int)rn;
            gv.N = (int)cn;
            gv.K = Q;
            ProfScope ps(ctx, PROF_PAIR);
            return gemm_f64(ctx, gv);
        };
        const long nov = (long)s.oX * s.vX;
        if (s.same && nk == s.oY) {
            // Same spin with every k in one block: V is symmetric.  The
            // blocks on and below the diagonal are formed (each element the
            // same ordered sum over Q as in the full product) and the rest
            // mirrored.
            const long rb = (nov + 4 * SYM_TILE - 1) / (4 * SYM_TILE) * SYM_TILE;
            for (long r0 = 0; r0 < nov; r0 += rb)
                for (long c0 = 0; c0 <= r0; c0 += rb)
                    if ((rc = v_block(r0, std::min(rb, nov - r0), c0, std::min(rb, nov - c0)))) return rc;
            if ((rc = mirror_upper(ctx, V, ldv, nov, rb / SYM_TILE))) return rc;
        } else if ((rc = v_block(0, nov, 0, ldv))) {
            return rc;
        }
        const long total = (long)nk * per_k;
        const int groups = (int)std::min<long>(AMP_MAX_GROUPS, (total + AMP_THREADS - 1) / AMP_THREADS);
        omp2_amp_kernel<<<groups, AMP_THREADS, 0, ctx->stream>>>(V, ldv, s.oX, s.vX, s.vY, k0, nk, s.eoX,
                                                                 s.evX, s.eoY, s.evY, s.same ? 1 : 0,
                                                                 (s.same && nk == s.oY) ? 1 : 0, T, part);
        LAUNCH_CHECK(ctx);
        if (w.e_slot && (rc = reduce_sum(ctx, part, groups, w.e_slot, true))) return rc;
        const long vv = (long)s.vX * s.vY;
        if (w.GooX) {  // sum_k T_k T_k^T, T_k = [oX][vX vY], cut into one chunk per a
            { ProfScope ps(ctx, PROF_GEMM); rc = small_syrk(ctx, T, vv, per_k, s.vY, s.vX, s.oX, s.vY, nk * s.vX, so, w.GooX, s.oX); }
            if (rc) return rc;
        }
        if (w.GvvX) {  // sum_{k,i} T_ki T_ki^T, T_ki = [vX][vY]
            // Same spin with every k in this block: T_ik = -T_ki and T_kk = 0,
            // so the sum is twice that over i < k, whose blocks are packed
            // into V (free once the amplitudes are formed).
            const bool half = s.same && nk == s.oY;
            const double* src = T;
            long nb = (long)nk * s.oX;
            if (half) {
                nb = (long)s.oX * (s.oX - 1) / 2;
                if (nb > 0) {
                    omp2_pack_lower_kernel<<<(unsigned)std::min<long>(nb, 65535), 256, 0, ctx->stream>>>(
                        T, s.oX, vv, V);
                    LAUNCH_CHECK(ctx);
                }
                src = V;
            }
            GemmF64 g;
            g.A = src;
            g.lda = s.vY;
            g.sA = vv;
            g.X = src;
            g.ldx = s.vY;
            g.sX = vv;
            g.transX = true;
            g.C = w.GvvX;
            g.ldc = s.vX;
            g.M = g.N = s.vX;
            g.K = s.vY;
            g.batch = (int)nb;
            g.reduce_batch = true;
            g.splits = red_splits(g.batch, s.vX, s.vX);
            g.alpha = half ? 2.0 * sv : sv;
            g.accumulate = true;
            { ProfScope ps(ctx, PROF_GEMM); rc = gemm_f64(ctx, g); }
            if (rc) return rc;
        }
        if (w.GvvY) {  // sum_{k,i} T_ki^T T_ki
            GemmF64 g;
            g.A = T;
            g.lda = s.vY;
            g.sA = vv;
            g.transA = true;
            g.X = T;
            g.ldx = s.vY;
            g.sX = vv;
            g.C = w.GvvY;
            g.ldc = s.vY;
            g.M = g.N = s.vY;
            g.K = s.vX;
            g.batch = nk * s.oX;
            g.reduce_batch = true;
            g.splits = red_splits(nk * s.oX, s.vY, s.vY);
            g.alpha = 1.0;
            g.accumulate = true;
            { ProfScope ps(ctx, PROF_GEMM); rc = gemm_f64(ctx, g); }
            if (rc) return rc;
        }
        if (w.GooY) {  // one block holds every k: -sum_{i,a} T_(.ia) T_(.ia)^T over b
            { ProfScope ps(ctx, PROF_GEMM); rc = small_syrk(ctx, T, per_k, vv, s.vY, s.vX, nk, s.vY, s.oX * s.vX, -1.0, w.GooY, s.oY); }
            if (rc) return rc;
        }
    }
    return MSHQC_HIP_OK;
}

__global__ void omp2_energy_out_kernel(const double* __restrict__ slots, double* __restrict__ out) {
    if (threadIdx.x == 0) {
        out[0] = 0.25 * slots[0];
        out[1] = 0.25 * slots[1];
        out[2] = slots[2];
    }
}

}  // namespace

extern "C" int mshqc_hip_omp2_opdm(mshqc_hip_ctx* ctx, const double* Ba, const double* Bb, int Q, int oa,
                                   int va, int ob, int vb, const double* eps_oa, const double* eps_va,
                                   const double* eps_ob, const double* eps_vb, double* E3,
                                   double* Goo_a, double* Gvv_a, double* Goo_b, double* Gvv_b) {
    int rc = ctx_enter(ctx);
    if (rc) return rc;
    CHECK_ARG(ctx, E3 != nullptr, "omp2_opdm: null energy pointer");
    CHECK_ARG(ctx, Q > 0 && oa >= 0 && va >= 0 && ob >= 0 && vb >= 0, "omp2_opdm: bad dimensions");
    CHECK_ARG(ctx, (oa == 0 || (Goo_a && eps_oa)) && (va == 0 || (Gvv_a && eps_va)) &&
                       (ob == 0 || (Goo_b && eps_ob)) && (vb == 0 || (Gvv_b && eps_vb)),
              "omp2_opdm: null output or orbital-energy pointer");
    CHECK_ARG(ctx, (oa * va == 0 || Ba) && (ob * vb == 0 || Bb), "omp2_opdm: null B");
    hipEventRecord(ctx->ev0, ctx->stream);
    double* slots = static_cast<double*>(workspace(ctx, WS_MISC, sizeof(double) * 8));
    if (!slots) return MSHQC_HIP_ERR_ALLOC;
    if ((rc = fill_zero(ctx, slots, 8))) return rc;
    if ((rc = fill_zero(ctx, Goo_a, (long)oa * oa)) || (rc = fill_zero(ctx, Gvv_a, (long)va * va)) ||
        (rc = fill_zero(ctx, Goo_b, (long)ob * ob)) || (rc = fill_zero(ctx, Gvv_b, (long)vb * vb)))
        return rc;
    const size_t cap = ctx->zblock_elems;
    // same spin: alpha-alpha, beta-beta (cholesky_omp2.cc:115-158, 232-326)
    const SpinPair aa{Ba, Ba, oa, va, oa, va, eps_oa, eps_va, eps_oa, eps_va, true};
    const SpinPair bb{Bb, Bb, ob, vb, ob, vb, eps_ob, eps_vb, eps_ob, eps_vb, true};
    for (int s = 0; s < 2; ++s) {
        const SpinPair& sp = s == 0 ? aa : bb;
        if (sp.oX < 2 || sp.vX < 2) continue;  // no i < j, a < b pairs
        Want w;
        w.GooX = s == 0 ? Goo_a : Goo_b;
        w.GvvX = s == 0 ? Gvv_a : Gvv_b;
        w.e_slot = slots + s;
        if ((rc = spin_pair(ctx, Q, sp, w, k_block(sp, cap)))) return rc;
    }
    // opposite spin (cholesky_omp2.cc:161-180 and the t_ab terms of the OPDM)
    if (oa > 0 && va > 0 && ob > 0 && vb > 0) {
        const SpinPair ab{Ba, Bb, oa, va, ob, vb, eps_oa, eps_va, eps_ob, eps_vb, false};
        const int kb = k_block(ab, cap);
        Want w;
        w.GooX = Goo_a;
        w.GvvX = Gvv_a;
        w.GvvY = Gvv_b;
        w.GooY = kb >= ob ? Goo_b : nullptr;
        w.e_slot = slots + 2;
        if ((rc = spin_pair(ctx, Q, ab, w, kb))) return rc;
        if (kb < ob) {
            // G_oo^b needs every beta k at once: walk t_ab again with the
            // spins swapped, blocked over the alpha index
            const SpinPair ba{Bb, Ba, ob, vb, oa, va, eps_ob, eps_vb, eps_oa, eps_va, false};
            Want wb;
            wb.GooX = Goo_b;
            if ((rc = spin_pair(ctx, Q, ba, wb, k_block(ba, cap)))) return rc;
        }
    }
    omp2_energy_out_kernel<<<1, 64, 0, ctx->stream>>>(slots, slots + 4);
    LAUNCH_CHECK(ctx);
    HIP_TRY(ctx, hipMemcpyAsync(E3, slots + 4, sizeof(double) * 3, hipMemcpyDeviceToHost, ctx->stream));
    hipEventRecord(ctx->ev1, ctx->stream);
    if ((rc = ctx_finish(ctx))) return rc;
    hipEventElapsedTime(&ctx->last_ms, ctx->ev0, ctx->ev1);
    return MSHQC_HIP_OK;
}

// Two-particle term of the orbital gradient for one spin: with G = dE/dB
// [Q][o][v] and the same orbitals' Loo [Q][o][o], Lvv [Q][v][v],
//   w2(c, k) = sum_Q [ sum_a Lvv^Q(c, a) G^Q(k, a) - sum_i G^Q(i, c) Loo^Q(i, k) ]
// (the derivative of E through B^Q_ia = C_i^T L^Q C_a when C_k gains
// kappa_ck C_c and C_c loses kappa_ck C_k).  Two batch-reduced FP64 GEMMs
// over Q; the split partials are summed in a fixed order.
extern "C" int mshqc_hip_omp2_grad2p(mshqc_hip_ctx* ctx, const double* G, const double* Loo,
                                     const double* Lvv, int Q, int o, int v, double* w2) {
    int rc = ctx_enter(ctx);
    if (rc) return rc;
    CHECK_ARG(ctx, Q > 0 && o >= 0 && v >= 0, "omp2_grad2p: bad dimensions");
    CHECK_ARG(ctx, (long)o * v == 0 || (G && Loo && Lvv && w2), "omp2_grad2p: null pointer");
    if ((long)o * v == 0) return MSHQC_HIP_OK;
    hipEventRecord(ctx->ev0, ctx->stream);
    GemmF64 g;  // w2 = sum_Q Lvv_Q G_Q^T
    g.A = Lvv;
    g.lda = v;
    g.sA = (long)v * v;
    g.X = G;
    g.ldx = v;
    g.sX = (long)o * v;
    g.transX = true;
    g.C = w2;
    g.ldc = o;
    g.M = v;
    g.N = o;
    g.K = v;
    g.batch = Q;
    g.reduce_batch = true;
    g.splits = red_splits(Q, v, o);
    {
        ProfScope ps(ctx, PROF_GEMM);
        rc = gemm_f64(ctx, g);
    }
    if (rc) return rc;
    GemmF64 h;  // w2 -= sum_Q G_Q^T Loo_Q
    h.A = G;
    h.lda = v;
    h.sA = (long)o * v;
    h.transA = true;
    h.X = Loo;
    h.ldx = o;
    h.sX = (long)o * o;
    h.C = w2;
    h.ldc = o;
    h.M = v;
    h.N = o;
    h.K = o;
    h.batch = Q;
    h.reduce_batch = true;
    h.splits = red_splits(Q, v, o);
    h.alpha = -1.0;
    h.accumulate = true;
    {
        ProfScope ps(ctx, PROF_GEMM);
        rc = gemm_f64(ctx, h);
    }
    if (rc) return rc;
    hipEventRecord(ctx->ev1, ctx->stream);
    if ((rc = ctx_finish(ctx))) return rc;
    hipEventElapsedTime(&ctx->last_ms, ctx->ev0, ctx->ev1);
    return MSHQC_HIP_OK;
}

// ------------------------------------------------- the orbital step, on the device
//
// Per spin the orbitals are held as C_occ [n][o] and C_vir [n][v] (columns =
// orbitals), the density P [n][n], the MO-basis generalised Fock F [n][n].
namespace {

__global__ void omp2_fock_ao_kernel(const double* __restrict__ H, const double* __restrict__ J,
                                    const double* __restrict__ K, long nn, double* __restrict__ F) {
    const long k = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (k < nn) F[k] = (H[k] + J[k]) - K[k];
}

// A = -kappa of the reference's step (cholesky_omp2.cc:375-425) from
// w = 2 F[o+a][i] (+ w2[a][i]) and F's diagonal; A is zero elsewhere (the
// caller clears it).  One thread per (a, i).
__global__ void omp2_minus_kappa_kernel(int n, int o, const double* __restrict__ F, const double* __restrict__ w2,
                                        double* __restrict__ A) {
    const int v = n - o;
    const long idx = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (idx >= (long)v * o) return;
    const int a = (int)(idx / o), i = (int)(idx % o);
    double w = 2.0 * F[(long)(o + a) * n + i];
    if (w2) w += w2[idx];
    double diff = F[(long)(o + a) * n + o + a] - F[(long)i * n + i] + 0.2;
    if (fabs(diff) < 1e-2) diff = diff >= 0 ? 1e-2 : -1e-2;
    const double val = (w / diff) * 0.5;
    A[(long)(o + a) * n + i] = -val;
    A[(long)i * n + o + a] = val;
}

// out[0] = max_j sum_i |A[i][j]| (one workgroup, fixed order).
__global__ __launch_bounds__(256) void omp2_norm1_kernel(int n, const double* __restrict__ A, double* out) {
    double m = 0.0;
    for (int j = threadIdx.x; j < n; j += 256) {
        double c = 0.0;
        for (int i = 0; i < n; ++i) c += fabs(A[(long)i * n + j]);
        m = fmax(m, c);
    }
    for (int k = 32; k > 0; k >>= 1) m = fmax(m, __shfl_xor(m, k, 64));
    __shared__ double sh[4];
    if ((threadIdx.x & 63) == 0) sh[threadIdx.x >> 6] = m;
    __syncthreads();
    if (threadIdx.x == 0) out[0] = fmax(fmax(sh[0], sh[1]), fmax(sh[2], sh[3]));
}

// X = s X + t I (n x n).
__global__ void omp2_scale_shift_kernel(int n, double s, double t, double* __restrict__ X) {
    const long k = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= (long)n * n) return;
    const int i = (int)(k / n), j = (int)(k % n);
    X[k] = s * X[k] + (i == j ? t : 0.0);
}

// out[0] = sum_{a,i} (2 F[o+a][i] + w2[a][i])^2 (one workgroup, fixed order).
__global__ __launch_bounds__(256) void omp2_grad_norm_kernel(int n, int o, const double* __restrict__ F,
                                                             const double* __restrict__ w2, double* out) {
    const long nvo = (long)(n - o) * o;
    double s = 0.0;
    for (long idx = threadIdx.x; idx < nvo; idx += 256) {
        const int a = (int)(idx / o), i = (int)(idx % o);
        double w = 2.0 * F[(long)(o + a) * n + i];
        if (w2) w += w2[idx];
        s += w * w;
    }
    for (int k = 32; k > 0; k >>= 1) s += __shfl_xor(s, k, 64);
    __shared__ double sh[4];
    if ((threadIdx.x & 63) == 0) sh[threadIdx.x >> 6] = s;
    __syncthreads();
    if (threadIdx.x == 0) out[0] = ((sh[0] + sh[1]) + sh[2]) + sh[3];
}

inline unsigned blocks_of(long n) { return (unsigned)((n + 255) / 256); }

// C[M][N] (+)= op(A) op(X) with row strides; skipped when a dimension is 0.
int small_gemm(mshqc_hip_ctx* ctx, const double* A, long lda, bool tA, const double* X, long ldx, bool tX,
               double* C, long ldc, int M, int N, int K, bool acc) {
    if (M <= 0 || N <= 0) return MSHQC_HIP_OK;
    if (K <= 0) {  // an empty contraction: C = 0 (or unchanged when accumulating)
        if (acc) return MSHQC_HIP_OK;
        for (int m = 0; m < M; ++m) {
            int rc = fill_zero(ctx, C + (long)m * ldc, N);
            if (rc) return rc;
        }
        return MSHQC_HIP_OK;
    }
    GemmF64 g;
    g.A = A;
    g.lda = lda;
    g.transA = tA;
    g.X = X;
    g.ldx = ldx;
    g.transX = tX;
    g.C = C;
    g.ldc = ldc;
    g.M = M;
    g.N = N;
    g.K = K;
    g.accumulate = acc;
    return gemm_f64(ctx, g);
}

// E = expm(A) for an n x n A (overwritten: A is scaled in place), with the
// host expm's algorithm (core.cc): scale by 2^-sq to 1-norm <= 1/2, Taylor
// to the first degree m whose next term is below 2^-60, Horner, square sq
// times.  tmp holds 2 n^2 doubles; norm1 one double.
int device_expm(mshqc_hip_ctx* ctx, int n, double* A, double* E, double* tmp, double* norm1) {
    const long nn = (long)n * n;
    omp2_norm1_kernel<<<1, 256, 0, ctx->stream>>>(n, A, norm1);
    LAUNCH_CHECK(ctx);
    double nrm = 0.0;
    HIP_TRY(ctx, hipMemcpyAsync(&nrm, norm1, sizeof(double), hipMemcpyDeviceToHost, ctx->stream));
    HIP_TRY(ctx, hipStreamSynchronize(ctx->stream));
    int sq = 0;
    while (nrm > 0.5) {
        nrm *= 0.5;
        ++sq;
    }
    int m = 1;
    double term = nrm;
    while (m < 30 && term * nrm / (m + 1) >= std::ldexp(1.0, -60)) {
        term *= nrm / (m + 1);
        ++m;
    }
    omp2_scale_shift_kernel<<<blocks_of(nn), 256, 0, ctx->stream>>>(n, std::ldexp(1.0, -sq), 0.0, A);
    LAUNCH_CHECK(ctx);
    // Horner: E = I + X/1 (I + X/2 (I + ... (I + X/m)))
    double* cur = tmp;
    double* nxt = tmp + nn;
    int rc = fill_zero(ctx, cur, nn);
    if (rc) return rc;
    omp2_scale_shift_kernel<<<blocks_of(nn), 256, 0, ctx->stream>>>(n, 0.0, 1.0, cur);
    LAUNCH_CHECK(ctx);
    for (int k = m; k >= 1; --k) {
        if ((rc = small_gemm(ctx, A, n, false, cur, n, false, nxt, n, n, n, n, false))) return rc;
        omp2_scale_shift_kernel<<<blocks_of(nn), 256, 0, ctx->stream>>>(n, 1.0 / k, 1.0, nxt);
        LAUNCH_CHECK(ctx);
        std::swap(cur, nxt);
    }
    for (int k = 0; k < sq; ++k) {
        if ((rc = small_gemm(ctx, cur, n, false, cur, n, false, nxt, n, n, n, n, false))) return rc;
        std::swap(cur, nxt);
    }
    HIP_TRY(ctx, hipMemcpyAsync(E, cur, sizeof(double) * nn, hipMemcpyDeviceToDevice, ctx->stream));
    return MSHQC_HIP_OK;
}

// HIP events around a call's kernels (the ctx's last_ms), then the stream
// synchronisation that turns asynchronous faults into a status.
struct StepTimer {
    mshqc_hip_ctx* ctx;
    explicit StepTimer(mshqc_hip_ctx* c) : ctx(c) { hipEventRecord(ctx->ev0, ctx->stream); }
    int stop() {
        hipEventRecord(ctx->ev1, ctx->stream);
        int rc = ctx_finish(ctx);
        if (rc) return rc;
        hipEventElapsedTime(&ctx->last_ms, ctx->ev0, ctx->ev1);
        hipGetLastError();
        return MSHQC_HIP_OK;
    }
};

int omp2_step_checks(mshqc_hip_ctx* ctx, int n, int o, const double* Co, const double* Cv, const char* who) {
    if (n <= 0 || o < 0 || o > n)
        return set_error(ctx, MSHQC_HIP_ERR_ARG, "%s: need n > 0 and 0 <= o <= n (n=%d o=%d)", who, n, o);
    if ((o > 0 && !Co) || (o < n && !Cv)) return set_error(ctx, MSHQC_HIP_ERR_ARG, "%s: null orbital block", who);
    return MSHQC_HIP_OK;
}

}  // namespace

extern "C" int mshqc_hip_omp2_fock(mshqc_hip_ctx* ctx, const double* L, int Q, int n, const double* H,
                                   const double* Co_a, const double* Cv_a, int na, const double* Co_b,
                                   const double* Cv_b, int nb, const double* P_a, const double* P_b,
                                   const double* Goo_a, const double* Gvv_a, const double* Goo_b,
                                   const double* Gvv_b, double* F_a, double* F_b) {
    int rc = ctx_enter(ctx);
    if (rc) return rc;
    CHECK_ARG(ctx, L && H && P_a && P_b && F_a && F_b && Q > 0, "omp2_fock: null pointer or Q <= 0");
    if ((rc = omp2_step_checks(ctx, n, na, Co_a, Cv_a, "omp2_fock (alpha)"))) return rc;
    if ((rc = omp2_step_checks(ctx, n, nb, Co_b, Cv_b, "omp2_fock (beta)"))) return rc;
    const int va = n - na, vb = n - nb;
    CHECK_ARG(ctx, (na == 0 || Goo_a) && (va == 0 || Gvv_a) && (nb == 0 || Goo_b) && (vb == 0 || Gvv_b),
              "omp2_fock: null density block");
    StepTimer tm(ctx);
    const long nn = (long)n * n;
    // Pt [2][n][n] | J | K [2] | F_ao | T (n x n)
    double* w = static_cast<double*>(workspace(ctx, WS_OMP2_V, sizeof(double) * (size_t)7 * nn));
    if (!w) return MSHQC_HIP_ERR_ALLOC;
    double *Pt = w, *J = w + 2 * nn, *K = w + 3 * nn, *Fao = w + 5 * nn, *T = w + 6 * nn;
    const double* Co[2] = {Co_a, Co_b};
    const double* Cv[2] = {Cv_a, Cv_b};
    const double* P[2] = {P_a, P_b};
    const double* Goo[2] = {Goo_a, Goo_b};
    const double* Gvv[2] = {Gvv_a, Gvv_b};
    double* F[2] = {F_a, F_b};
    const int no[2] = {na, nb};
    // P_tot = P + C_occ G_oo C_occ^T + C_vir G_vv C_vir^T (cholesky_omp2.cc:451-459)
    for (int s = 0; s < 2; ++s) {
        const int o = no[s], v = n - o;
        double* Ps = Pt + s * nn;
        HIP_TRY(ctx, hipMemcpyAsync(Ps, P[s], sizeof(double) * nn, hipMemcpyDeviceToDevice, ctx->stream));
        if ((rc = small_gemm(ctx, Co[s], o, false, Goo[s], o, false, T, o, n, o, o, false))) return rc;
        if ((rc = small_gemm(ctx, T, o, false, Co[s], o, true, Ps, n, n, n, o, true))) return rc;
        if ((rc = small_gemm(ctx, Cv[s], v, false, Gvv[s], v, false, T, v, n, v, v, false))) return rc;
        if ((rc = small_gemm(ctx, T, v, false, Cv[s], v, true, Ps, n, n, n, v, true))) return rc;
    }
    // H + J[Pt_a + Pt_b] - K[Pt_s] (cholesky_omp2.cc:331-369; one J for both spins)
    if ((rc = jk_general(ctx, L, Q, n, Pt, 2, 0, J, K))) return rc;
    for (int s = 0; s < 2; ++s) {
        const int o = no[s], v = n - o;
        omp2_fock_ao_kernel<<<blocks_of(nn), 256, 0, ctx->stream>>>(H, J, K + s * nn, nn, Fao);
        LAUNCH_CHECK(ctx);
        // F_mo = C^T F_ao C by blocks: [F_oo F_ov; F_vo F_vv]
        double* Xo = T;                    // F_ao C_occ [n][o]
        double* Xv = T + (long)n * o;      // F_ao C_vir [n][v]
        if ((rc = small_gemm(ctx, Fao, n, false, Co[s], o, false, Xo, o, n, o, n, false))) return rc;
        if ((rc = small_gemm(ctx, Fao, n, false, Cv[s], v, false, Xv, v, n, v, n, false))) return rc;
        double* Fs = F[s];
        if ((rc = small_gemm(ctx, Co[s], o, true, Xo, o, false, Fs, n, o, o, n, false))) return rc;
        if ((rc = small_gemm(ctx, Co[s], o, true, Xv, v, false, Fs + o, n, o, v, n, false))) return rc;
        if ((rc = small_gemm(ctx, Cv[s], v, true, Xo, o, false, Fs + (long)o * n, n, v, o, n, false))) return rc;
        if ((rc = small_gemm(ctx, Cv[s], v, true, Xv, v, false, Fs + (long)o * n + o, n, v, v, n, false))) return rc;
    }
    return tm.stop();
}

extern "C" int mshqc_hip_omp2_grad_norm(mshqc_hip_ctx* ctx, int n, int o, const double* F, const double* w2,
                                        double* g2) {
    int rc = ctx_enter(ctx);
    if (rc) return rc;
    CHECK_ARG(ctx, F && g2 && n > 0 && o >= 0 && o <= n, "omp2_grad_norm: bad arguments");
    double* slot = static_cast<double*>(workspace(ctx, WS_MISC, sizeof(double) * 8));
    if (!slot) return MSHQC_HIP_ERR_ALLOC;
    omp2_grad_norm_kernel<<<1, 256, 0, ctx->stream>>>(n, o, F, w2, slot);
    LAUNCH_CHECK(ctx);
    HIP_TRY(ctx, hipMemcpyAsync(g2, slot, sizeof(double), hipMemcpyDeviceToHost, ctx->stream));
    return ctx_finish(ctx);
}

extern "C" int mshqc_hip_omp2_rotate(mshqc_hip_ctx* ctx, int n, int o, const double* F, const double* w2,
                                     double* Co, double* Cv, double* P) {
    int rc = ctx_enter(ctx);
    if (rc) return rc;
    CHECK_ARG(ctx, F && P, "omp2_rotate: null pointer");
    if ((rc = omp2_step_checks(ctx, n, o, Co, Cv, "omp2_rotate"))) return rc;
    const int v = n - o;
    if (o == 0 || v == 0) return MSHQC_HIP_OK;  // no occupied-virtual rotation exists
    StepTimer tm(ctx);
    const long nn = (long)n * n;
    // A | U | 2 n^2 expm temporaries | new C_occ, C_vir | norm
    double* w = static_cast<double*>(workspace(ctx, WS_OMP2_T, sizeof(double) * (size_t)(5 * nn + 8)));
    if (!w) return MSHQC_HIP_ERR_ALLOC;
    double *A = w, *U = w + nn, *tmp = w + 2 * nn, *Cn = w + 4 * nn, *nrm = w + 5 * nn;
    if ((rc = fill_zero(ctx, A, nn))) return rc;
    omp2_minus_kappa_kernel<<<blocks_of((long)v * o), 256, 0, ctx->stream>>>(n, o, F, w2, A);
    LAUNCH_CHECK(ctx);
    if ((rc = device_expm(ctx, n, A, U, tmp, nrm))) return rc;
    // [C_occ | C_vir] <- [C_occ | C_vir] U, by blocks of U
    double* Con = Cn;
    double* Cvn = Cn + (long)n * o;
    if ((rc = small_gemm(ctx, Co, o, false, U, n, false, Con, o, n, o, o, false))) return rc;
    if ((rc = small_gemm(ctx, Cv, v, false, U + (long)o * n, n, false, Con, o, n, o, v, true))) return rc;
    if ((rc = small_gemm(ctx, Co, o, false, U + o, n, false, Cvn, v, n, v, o, false))) return rc;
    if ((rc = small_gemm(ctx, Cv, v, false, U + (long)o * n + o, n, false, Cvn, v, n, v, v, true))) return rc;
    HIP_TRY(ctx, hipMemcpyAsync(Co, Con, sizeof(double) * (size_t)n * o, hipMemcpyDeviceToDevice, ctx->stream));
    HIP_TRY(ctx, hipMemcpyAsync(Cv, Cvn, sizeof(double) * (size_t)n * v, hipMemcpyDeviceToDevice, ctx->stream));
    // P = C_occ C_occ^T
    if ((rc = small_gemm(ctx, Co, o, false, Co, o, true, P, n, n, n, o, false))) return rc;
    return tm.stop();
}
