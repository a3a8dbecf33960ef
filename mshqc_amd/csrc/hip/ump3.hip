// CD-UMP3 energy components (cholesky_ump3.cc:160-637) from the MO Cholesky
// blocks of each spin: B [Q][o][v], Loo [Q][o][o], Lvv [Q][v][v].
//
// The reference forms its integral blocks as products of column-major
// flattened L blocks (:65-97) and indexes some of them as if row-major:
// for the o x v blocks that reads L(r mod o, r div o) at r = i*v + a.  The
// first-order amplitudes and the beta ring term use that reading; here it
// is one gather of B into S (scrambled) before the GEMMs, so every later
// step is an ordinary product:
//   t_aa(ijab) = [S(ia|jb) - S(ib|ja)] / D,  t_ab(ijab) = S(ia|jb) / D
//   same spin:  1/2 sum_kl [(kl|ij) - (kl|ji)] t(klab)       (|.| > 1e-12)
//               + sum_cd (ab|cd) t(ijcd)   (alpha; beta: 1/2 [(ab|cd) - (ab|dc)])
//               + P(ij)P(ab) sum_kc [(kj|bc) - (kc|jb)] t(ikac)
//                 ((kc|jb) from B for alpha, from S for beta)
//               E = 1/4 sum t new / D
//   opposite:   sum_kl (ki|lj) t(klab) (|.| > 1e-12) + sum_cd (ac|bd) t(ijcd)
//               + sum_kc (ka|ic) t(kjcb) + sum_ld (lb|jd) t(ilad),  E = sum t new / D
// The (ab|cd) terms contract through the Cholesky factors (o^2 v^2 Q work);
// the interleaved (ac|bd) is built in blocks of a, as in CD-RMP3.
#include <algorithm>
#include <cmath>
#include <vector>

#include "common.h"

namespace {

constexpr int EW = 256;

unsigned ew_grid(long total) { return (unsigned)std::min<long>(8192, std::max<long>(1, (total + EW - 1) / EW)); }

// dst[t] = src[map(t)] over t < total, for a gather given as a functor.
template <class Map>
__global__ void gather_kernel(const double* __restrict__ src, double* __restrict__ dst, long total, Map map) {
    for (long t = (long)blockIdx.x * EW + threadIdx.x; t < total; t += (long)gridDim.x * EW) dst[t] = src[map(t)];
}

// S[q][r] = B[q][(r mod o) v + r div o]
struct ScrambleMap {
    long ov;
    int o, v;
    __device__ long operator()(long t) const {
        const long q = t / ov, r = t % ov;
        return q * ov + (r % o) * v + r / o;
    }
};

// [(i a)][(k c)] <- G[(k a)][(i c)] over an (o v)^2 matrix
struct SwapOccMap {
    int o, v;
    __device__ long operator()(long t) const {
        const long ov = (long)o * v;
        const int c = (int)(t % v);
        long r = t / v;
        const int k = (int)(r % o);
        r /= o;
        const int a = (int)(r % v);
        const int i = (int)(r / v);
        return ((long)k * v + a) * ov + (long)i * v + c;
    }
};

// Ts[((i v1 + a) o2 + k) v2 + c] = t[((i o2 + k) v1 + a) v2 + c] for t [o1][o2][v1][v2]
struct TsMap {
    int o2, v1, v2;
    __device__ long operator()(long t) const {
        const int c = (int)(t % v2);
        long r = t / v2;
        const int k = (int)(r % o2);
        r /= o2;
        const int a = (int)(r % v1);
        const long i = r / v1;
        return ((i * o2 + k) * v1 + a) * v2 + c;
    }
};

// W'[((a v2 + b) v1 + c) v2 + d] = W[((a v1 + c) v2 + b) v2 + d]
struct PermWMap {
    int v1, v2;
    __device__ long operator()(long t) const {
        const int d = (int)(t % v2);
        long r = t / v2;
        const int c = (int)(r % v1);
        r /= v1;
        const int b = (int)(r % v2);
        const long a = r / v2;
        return ((a * v1 + c) * v2 + b) * v2 + d;
    }
};

template <class Map>
int gather(mshqc_hip_ctx* ctx, const double* src, double* dst, long total, Map map) {
    if (total <= 0) return MSHQC_HIP_OK;
    gather_kernel<Map><<<ew_grid(total), EW, 0, ctx->stream>>>(src, dst, total, map);
    LAUNCH_CHECK(ctx);
    return MSHQC_HIP_OK;
}

// First-order amplitudes from a (scrambled) pair-integral matrix
// G[(i a)][(j b)] (row length ldg): same spin t = (G[ia,jb] - G[ib,ja]) / D,
// opposite spin t = G[ia,jb] / D; |D| <= 1e-12 gives 0 (:223-300).
__global__ void amp_kernel(const double* __restrict__ G, long ldg, int o1, int v1, int o2, int v2, int same,
                           const double* __restrict__ e1o, const double* __restrict__ e1v,
                           const double* __restrict__ e2o, const double* __restrict__ e2v, double* __restrict__ t) {
    const long total = (long)o1 * o2 * v1 * v2;
    for (long x = (long)blockIdx.x * EW + threadIdx.x; x < total; x += (long)gridDim.x * EW) {
        const int b = (int)(x % v2);
        long r = x / v2;
        const int a = (int)(r % v1);
        r /= v1;
        const int j = (int)(r % o2);
        const int i = (int)(r / o2);
        double num = G[((long)i * v1 + a) * ldg + (long)j * v2 + b];
        if (same) num -= G[((long)i * v1 + b) * ldg + (long)j * v2 + a];
        const double D = e1o[i] + e2o[j] - e1v[a] - e2v[b];
        t[x] = fabs(D) > 1e-12 ? num / D : 0.0;
    }
}

// Same spin: V[(i o + j) o^2 + k o + l] = (kl|ij) - (kl|ji) from O4[(k l)][(i j)],
// zero when |.| <= 1e-12 (:323-333, :466-469).
__global__ void vss_kernel(const double* __restrict__ O4, int o, double* __restrict__ V) {
    const long oo = (long)o * o, total = oo * oo;
    for (long t = (long)blockIdx.x * EW + threadIdx.x; t < total; t += (long)gridDim.x * EW) {
        const long ij = t / oo, kl = t % oo;
        const int i = (int)(ij / o), j = (int)(ij % o);
        const double x = O4[kl * oo + ij] - O4[kl * oo + (long)j * o + i];
        V[t] = fabs(x) > 1e-12 ? x : 0.0;
    }
}

// Opposite spin: U[(i ob + j)][(k ob + l)] = (ki|lj) from O4[(k i)][(l j)], zero
// when |.| <= 1e-12 (:556-566).
__global__ void uos_kernel(const double* __restrict__ O4, int oa, int ob, double* __restrict__ U) {
    const long n = (long)oa * ob, total = n * n;
    for (long t = (long)blockIdx.x * EW + threadIdx.x; t < total; t += (long)gridDim.x * EW) {
        const long ij = t / n, kl = t % n;
        const int i = (int)(ij / ob), j = (int)(ij % ob), k = (int)(kl / ob), l = (int)(kl % ob);
        const double x = O4[((long)k * oa + i) * ((long)ob * ob) + (long)l * ob + j];
        U[t] = fabs(x) > 1e-12 ? x : 0.0;
    }
}

// A[q][c][d] = L[q][c][d] - L[q][d][c]
__global__ void asym_kernel(const double* __restrict__ L, long Qvv, int v, double* __restrict__ A) {
    const long vv = (long)v * v;
    for (long t = (long)blockIdx.x * EW + threadIdx.x; t < Qvv; t += (long)gridDim.x * EW) {
        const long q = t / vv, cd = t % vv;
        const int c = (int)(cd / v), d = (int)(cd % v);
        A[t] = L[t] - L[q * vv + (long)d * v + c];
    }
}

// M[(j v + b) ov + k v + c] = OV[(k o + j) v^2 + b v + c] - G2[(k v + c) ov + j v + b]
__global__ void mring_kernel(const double* __restrict__ OV, const double* __restrict__ G2, int o, int v,
                             double* __restrict__ M) {
    const long ov = (long)o * v, vv = (long)v * v, total = ov * ov;
    for (long t = (long)blockIdx.x * EW + threadIdx.x; t < total; t += (long)gridDim.x * EW) {
        const int c = (int)(t % v);
        long r = t / v;
        const int k = (int)(r % o);
        r /= o;
        const int b = (int)(r % v);
        const int j = (int)(r / v);
        M[t] = OV[((long)k * o + j) * vv + (long)b * v + c] - G2[((long)k * v + c) * ov + (long)j * v + b];
    }
}

// Energy partials.  Same spin: new = R[(ij)(ab)] + R1[ia,jb] - R1[ib,ja] -
// R1[ja,ib] + R1[jb,ia], e += 1/4 t new / D; opposite spin: new = R + R1[ia,jb],
// e += t new / D.
__global__ __launch_bounds__(EW) void ump3_final_kernel(const double* __restrict__ R, const double* __restrict__ R1,
                                                       const double* __restrict__ t, int o1, int v1, int o2, int v2,
                                                       int same, const double* __restrict__ e1o,
                                                       const double* __restrict__ e1v,
                                                       const double* __restrict__ e2o,
                                                       const double* __restrict__ e2v, double* __restrict__ part) {
    const long ld = (long)o2 * v2, total = (long)o1 * o2 * v1 * v2;
    double e = 0.0;
    for (long x = (long)blockIdx.x * EW + threadIdx.x; x < total; x += (long)gridDim.x * EW) {
        const int b = (int)(x % v2);
        long r = x / v2;
        const int a = (int)(r % v1);
        r /= v1;
        const int j = (int)(r % o2);
        const int i = (int)(r / o2);
        const long ia = (long)i * v1 + a, jb = (long)j * v2 + b;
        double nw = R[x] + R1[ia * ld + jb];
        if (same) {
            const long ib = (long)i * v2 + b, ja = (long)j * v1 + a;
            nw += -R1[ib * ld + ja] - R1[ja * ld + ib] + R1[jb * ld + ia];
        }
        const double D = e1o[i] + e2o[j] - e1v[a] - e2v[b];
        if (fabs(D) > 1e-12) e += (same ? 0.25 : 1.0) * t[x] * (nw / D);
    }
    for (int m = 32; m > 0; m >>= 1) e += __shfl_xor(e, m, 64);
    __shared__ double sh[EW / 64];
    if ((threadIdx.x & 63) == 0) sh[threadIdx.x >> 6] = e;
    __syncthreads();
    if (threadIdx.x == 0) part[blockIdx.x] = ((sh[0] + sh[1]) + sh[2]) + sh[3];
}

// C (+)= alpha op(A) op(X), row-major, one GEMM.
int mm(mshqc_hip_ctx* ctx, const double* A, long lda, bool tA, const double* X, long ldx, bool tX, double* C,
       long ldc, long M, long N, long K, double alpha = 1.0, bool acc = false) {
    if (M <= 0 || N <= 0 || K <= 0) return MSHQC_HIP_OK;
    GemmF64 g;
    g.A = A;
    g.lda = lda;
    g.transA = tA;
    g.X = X;
    g.ldx = ldx;
    g.transX = tX;
    g.C = C;
    g.ldc = ldc;
    g.M = (int)M;
    g.N = (int)N;
    g.K = (int)K;
    g.alpha = alpha;
    g.accumulate = acc;
    ProfScope ps(ctx, PROF_GEMM);
    return gemm_f64(ctx, g);
}

// Bump allocator over one workspace slot.
struct Arena {
    double* base = nullptr;
    size_t used = 0;
    double* take(size_t n) {
        double* p = base + used;
        used += (n + 31) / 32 * 32;
        return p;
    }
};

struct Spin {
    const double *B, *Loo, *Lvv, *eo, *ev;
    int o, v;
};

int energy_sum(mshqc_hip_ctx* ctx, const double* R, const double* R1, const double* t, const Spin& x,
               const Spin& y, bool same, double* part, double* slot) {
    const long total = (long)x.o * y.o * x.v * y.v;
    const unsigned grid = ew_grid(total);
    ump3_final_kernel<<<grid, EW, 0, ctx->stream>>>(R, R1, t, x.o, x.v, y.o, y.v, same ? 1 : 0, x.eo, x.ev, y.eo,
                                                    y.ev, part);
    LAUNCH_CHECK(ctx);
    return reduce_sum(ctx, part, grid, slot, false);
}

// One same-spin component (compute_mp3_ss_alpha :308-456 / _beta :462-531).
int same_spin(mshqc_hip_ctx* ctx, Arena ar, int Q, const Spin& s, const double* Gs, const double* Gu, bool beta,
              double* part, double* slot) {
    const int o = s.o, v = s.v;
    const long oo = (long)o * o, vv = (long)v * v, ov = (long)o * v, oovv = oo * vv;
    double* t = ar.take(oovv);
    double* R = ar.take(oovv);
    double* Ts = ar.take(oovv);
    double* M = ar.take(ov * ov);
    double* R1 = ar.take(ov * ov);
    double* OV = ar.take(oovv);
    double* O4 = ar.take(oo * oo);
    double* V = ar.take(oo * oo);
    double* X1 = ar.take(oo * Q);
    double* La = beta ? ar.take((size_t)Q * vv) : nullptr;
    int rc;
    amp_kernel<<<ew_grid(oovv), EW, 0, ctx->stream>>>(Gs, ov, o, v, o, v, 1, s.eo, s.ev, s.eo, s.ev, t);
    LAUNCH_CHECK(ctx);
    // hole-hole ladder
    if ((rc = mm(ctx, s.Loo, oo, true, s.Loo, oo, false, O4, oo, oo, oo, Q))) return rc;
    vss_kernel<<<ew_grid(oo * oo), EW, 0, ctx->stream>>>(O4, o, V);
    LAUNCH_CHECK(ctx);
    if ((rc = mm(ctx, V, oo, false, t, vv, false, R, vv, oo, vv, oo, 0.5, false))) return rc;
    // (ab|cd) ladder through the factors: X1[(ij)][q] = sum_cd t[(ij)(cd)] L_q[cd]
    const double* Lc = s.Lvv;
    if (beta) {
        asym_kernel<<<ew_grid((long)Q * vv), EW, 0, ctx->stream>>>(s.Lvv, (long)Q * vv, v, La);
        LAUNCH_CHECK(ctx);
        Lc = La;
    }
    if ((rc = mm(ctx, t, vv, false, Lc, vv, true, X1, Q, oo, Q, vv))) return rc;
    if ((rc = mm(ctx, X1, Q, false, s.Lvv, vv, false, R, vv, oo, vv, Q, beta ? 0.5 : 1.0, true))) return rc;
    // ring: M[(jb)][(kc)] = (kj|bc) - (kc|jb), R1 = Ts M^T
    if ((rc = mm(ctx, s.Loo, oo, true, s.Lvv, vv, false, OV, vv, oo, vv, Q))) return rc;
    mring_kernel<<<ew_grid(ov * ov), EW, 0, ctx->stream>>>(OV, beta ? Gs : Gu, o, v, M);
    LAUNCH_CHECK(ctx);
    if ((rc = gather(ctx, t, Ts, oovv, TsMap{o, v, v}))) return rc;
    if ((rc = mm(ctx, Ts, ov, false, M, ov, true, R1, ov, ov, ov, ov))) return rc;
    return energy_sum(ctx, R, R1, t, s, s, true, part, slot);
}

// The opposite-spin component (compute_mp3_os :537-637).
int opposite_spin(mshqc_hip_ctx* ctx, Arena ar, int Q, const Spin& a, const Spin& b, const double* Gab,
                  const double* Gua, const double* Gub, size_t wcap, double* part, double* slot) {
    const int oa = a.o, va = a.v, ob = b.o, vb = b.v;
    const long nij = (long)oa * ob, nab = (long)va * vb, ova = (long)oa * va, ovb = (long)ob * vb;
    double* t = ar.take(nij * nab);
    double* R = ar.take(nij * nab);
    double* T12 = ar.take(nij * nab);
    double* Ya = ar.take(ova * ova);
    double* Yb = ar.take(ovb * ovb);
    double* R1 = ar.take(ova * ovb);
    double* O4 = ar.take(nij * nij);
    double* U = ar.take(nij * nij);
    // (ac|bd) blocks: na values of a, each va * vb^2 doubles
    const long wrow = (long)va * vb * vb;
    const int na = (int)std::max<long>(1, std::min<long>(va, (long)(wcap / (size_t)wrow)));
    double* Wb = ar.take((size_t)na * wrow);
    double* Wp = ar.take((size_t)na * wrow);
    int rc;
    amp_kernel<<<ew_grid(nij * nab), EW, 0, ctx->stream>>>(Gab, ovb, oa, va, ob, vb, 0, a.eo, a.ev, b.eo, b.ev, t);
    LAUNCH_CHECK(ctx);
    // hole-hole: O4[(k i)][(l j)] = sum_q Looa[ki] Loob[lj]
    if ((rc = mm(ctx, a.Loo, (long)oa * oa, true, b.Loo, (long)ob * ob, false, O4, (long)ob * ob, (long)oa * oa,
                 (long)ob * ob, Q)))
        return rc;
    uos_kernel<<<ew_grid(nij * nij), EW, 0, ctx->stream>>>(O4, oa, ob, U);
    LAUNCH_CHECK(ctx);
    if ((rc = mm(ctx, U, nij, false, t, nab, false, R, nab, nij, nab, nij, 1.0, false))) return rc;
    // (ac|bd) ladder, a in blocks of na
    for (int a0 = 0; a0 < va; a0 += na) {
        const int nb = std::min(na, va - a0);
        if ((rc = mm(ctx, a.Lvv + (long)a0 * va, (long)va * va, true, b.Lvv, (long)vb * vb, false, Wb,
                     (long)vb * vb, (long)nb * va, (long)vb * vb, Q)))
            return rc;
        if ((rc = gather(ctx, Wb, Wp, (long)nb * wrow, PermWMap{va, vb}))) return rc;
        if ((rc = mm(ctx, t, nab, false, Wp, nab, true, R + (long)a0 * vb, nab, nij, (long)nb * vb, nab, 1.0,
                     true)))
            return rc;
    }
    // rings: (ka|ic) t(kjcb) + (lb|jd) t(ilad)
    if ((rc = gather(ctx, t, T12, nij * nab, TsMap{ob, va, vb}))) return rc;
    if ((rc = gather(ctx, Gua, Ya, ova * ova, SwapOccMap{oa, va}))) return rc;
    if ((rc = gather(ctx, Gub, Yb, ovb * ovb, SwapOccMap{ob, vb}))) return rc;
    if ((rc = mm(ctx, Ya, ova, false, T12, ovb, false, R1, ovb, ova, ovb, ova))) return rc;
    if ((rc = mm(ctx, T12, ovb, false, Yb, ovb, true, R1, ovb, ova, ovb, ovb, 1.0, true))) return rc;
    return energy_sum(ctx, R, R1, t, a, b, false, part, slot);
}

}  // namespace

extern "C" int mshqc_hip_ump3(mshqc_hip_ctx* ctx, const double* Ba, const double* Bb, const double* Looa,
                              const double* Loob, const double* Lvva, const double* Lvvb, int Q, int oa, int va,
                              int ob, int vb, const double* eps_oa, const double* eps_va, const double* eps_ob,
                              const double* eps_vb, double* E) {
    int rc = ctx_enter(ctx);
    if (rc) return rc;
    CHECK_ARG(ctx, E != nullptr, "ump3: null energy pointer");
    CHECK_ARG(ctx, Q > 0 && oa >= 0 && va >= 0 && ob >= 0 && vb >= 0, "ump3: bad dimensions");
    E[0] = E[1] = E[2] = 0.0;
    const bool ha = oa > 0 && va > 0, hb = ob > 0 && vb > 0;
    CHECK_ARG(ctx, (!ha || (Ba && Looa && Lvva && eps_oa && eps_va)) && (!hb || (Bb && Loob && Lvvb && eps_ob && eps_vb)),
              "ump3: null pointer");
    if (!ha && !hb) return MSHQC_HIP_OK;
    hipEventRecord(ctx->ev0, ctx->stream);
    const long ova = (long)oa * va, ovb = (long)ob * vb;
    const size_t wcap = ctx->zblock_elems;
    // arena sizing: scrambled blocks and pair matrices, then the largest phase
    auto ss_need = [&](int o, int v) -> size_t {
        const size_t oo = (size_t)o * o, vv = (size_t)v * v, ovv = (size_t)o * v;
        return 4 * oo * vv + 2 * ovv * ovv + 2 * oo * oo + oo * Q + (size_t)Q * vv + 10 * 32;
    };
    size_t phase = 0;
    if (ha) phase = std::max(phase, ss_need(oa, va));
    if (hb) phase = std::max(phase, ss_need(ob, vb));
    if (ha && hb) {
        const size_t nij = (size_t)oa * ob, nab = (size_t)va * vb;
        const size_t wrow = (size_t)va * vb * vb;
        const size_t na = std::max<size_t>(1, std::min<size_t>(va, wcap / wrow));
        phase = std::max(phase, 3 * nij * nab + (size_t)ova * ova + (size_t)ovb * ovb + (size_t)ova * ovb +
                                    2 * nij * nij + 2 * na * wrow + 10 * 32);
    }
    const size_t fixed = (size_t)Q * (ova + ovb) + (size_t)ova * ova * 2 + (size_t)ovb * ovb * 2 +
                         (size_t)ova * ovb + 8 * 32;
    double* base = static_cast<double*>(workspace(ctx, WS_MP3_A, sizeof(double) * (fixed + phase)));
    double* part = static_cast<double*>(workspace(ctx, WS_PART, sizeof(double) * 8192));
    double* slots = static_cast<double*>(workspace(ctx, WS_MISC, sizeof(double) * 4));
    if (!base || !part || !slots) return MSHQC_HIP_ERR_ALLOC;
    Arena ar{base, 0};
    double* Sa = ha ? ar.take((size_t)Q * ova) : nullptr;
    double* Sb = hb ? ar.take((size_t)Q * ovb) : nullptr;
    double* Gsa = ha ? ar.take((size_t)ova * ova) : nullptr;  // scrambled (ia|jb), alpha
    double* Gua = ha ? ar.take((size_t)ova * ova) : nullptr;  // (ia|jb), alpha
    double* Gsb = hb ? ar.take((size_t)ovb * ovb) : nullptr;
    double* Gub = hb ? ar.take((size_t)ovb * ovb) : nullptr;
    double* Gab = (ha && hb) ? ar.take((size_t)ova * ovb) : nullptr;
    if (ha) {
        if ((rc = gather(ctx, Ba, Sa, (long)Q * ova, ScrambleMap{ova, oa, va}))) return rc;
        if ((rc = mm(ctx, Sa, ova, true, Sa, ova, false, Gsa, ova, ova, ova, Q))) return rc;
        if ((rc = mm(ctx, Ba, ova, true, Ba, ova, false, Gua, ova, ova, ova, Q))) return rc;
    }
    if (hb) {
        if ((rc = gather(ctx, Bb, Sb, (long)Q * ovb, ScrambleMap{ovb, ob, vb}))) return rc;
        if ((rc = mm(ctx, Sb, ovb, true, Sb, ovb, false, Gsb, ovb, ovb, ovb, Q))) return rc;
        if ((rc = mm(ctx, Bb, ovb, true, Bb, ovb, false, Gub, ovb, ovb, ovb, Q))) return rc;
    }
    if (ha && hb && (rc = mm(ctx, Sa, ova, true, Sb, ovb, false, Gab, ovb, ova, ovb, Q))) return rc;
    const Spin sa{Ba, Looa, Lvva, eps_oa, eps_va, oa, va}, sb{Bb, Loob, Lvvb, eps_ob, eps_vb, ob, vb};
    Arena ph{base + ar.used, 0};
    if (ha && (rc = same_spin(ctx, ph, Q, sa, Gsa, Gua, false, part, slots + 0))) return rc;
    if (hb && (rc = same_spin(ctx, ph, Q, sb, Gsb, Gub, true, part, slots + 1))) return rc;
    if (ha && hb && (rc = opposite_spin(ctx, ph, Q, sa, sb, Gab, Gua, Gub, wcap, part, slots + 2))) return rc;
    double h[3] = {0, 0, 0};
    HIP_TRY(ctx, hipMemcpyAsync(h, slots, sizeof(double) * 3, hipMemcpyDeviceToHost, ctx->stream));
    hipEventRecord(ctx->ev1, ctx->stream);
    if ((rc = ctx_finish(ctx))) return rc;
    hipEventElapsedTime(&ctx->last_ms, ctx->ev0, ctx->ev1);
    E[0] = ha ? h[0] : 0.0;
    E[1] = hb ? h[1] : 0.0;
    E[2] = (ha && hb) ? h[2] : 0.0;
    return MSHQC_HIP_OK;
}
