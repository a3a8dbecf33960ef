/*
 * cd_oracle.c -- CPU FP64 restatement of mshqc's Cholesky-vector hot path.
 *
 * TEST INFRASTRUCTURE ONLY.  This file is the checker for the HIP library in
 * mshqc_amd/csrc: only tests/, __graft_entry__.smoke() and bench.py's CPU
 * baseline leg may load it.  Nothing in the product links or calls it.
 *
 * Each function restates the arithmetic of one reference routine (cited per
 * function as file:line into syahrulhidayat/mshqc).  Loop order follows the
 * reference where it matters for semantics (pivot choice, skip branches);
 * floating-point sums use Neumaier compensated summation in a fixed order so
 * the result does not depend on the OpenMP thread count (the reference's
 * `reduction(+:e_corr)` does).  That makes the oracle closer to the exact value
 * than the reference itself, never further.
 *
 * Layouts (all row-major, C order):
 *   L  [Q][n][n]      Cholesky vectors, L^Q_{mu nu} at Q*n*n + mu*n + nu
 *   B  [Q][o][v]      B^Q_{ia} at Q*o*v + i*v + a
 *   C  [n][m]         MO coefficient block, C_{mu p} at mu*m + p
 *   eri[n][n][n][n]   (ij|kl) at ((i*n+j)*n+k)*n+l
 *   t2 [o][o][v][v]   t_{ij}^{ab}
 *
 * Parity status of every function is listed at the end of this header block:
 *   orc_rmp2_energy        pinned: Q=1/o=1 closed forms, Q-rotation invariance,
 *                          E = sum t^2 D identity, reference build (oracle/_ref)
 *   orc_rmp2_grad          pinned: central finite differences, <B,dE/dB> = 4E,
 *                          sum_p dE/deps_p = 0, sum_p eps_p dE/deps_p = -E,
 *                          OMP2 OPDM diagonals
 *   orc_ump2_energy/grad   pinned: UMP2(alpha=beta) = RMP2, finite differences,
 *                          reference build (oracle/_ref)
 *   orc_jk                 pinned: brute-force 4-index contraction at tiny n
 *   orc_transform          pinned: numpy einsum at tiny n
 *   orc_cholesky           pinned: exact reconstruction at thr -> 0, reference
 *                          build (oracle/_ref)
 *   orc_rmp2_energy_pairs_rank1  (timing baseline) pinned: equals
 *                          orc_rmp2_energy summed over all ordered pairs
 */
#include <math.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

/* ---------------------------------------------------------------- helpers */

typedef struct { double s, c; } nsum_t;

static inline void nsum_add(nsum_t* a, double x) {
    double t = a->s + x;
    if (fabs(a->s) >= fabs(x)) a->c += (a->s - t) + x;
    else a->c += (x - t) + a->s;
    a->s = t;
}
static inline double nsum_val(const nsum_t* a) { return a->s + a->c; }

/* G[a*vb + b] = sum_Q Bl[Q][i][a] * Br[Q][j][b]: the rank-1 updates over Q of
 * cholesky_rmp2.cc:125-132, done per 64x64 block of G so the block stays in
 * L1.  Every element still accumulates Q = 0, 1, ... in order, so the result
 * is identical to the unblocked loop. */
static void pair_block(const double* Bl, int ol, int vl, int i,
                       const double* Br, int orr, int vr, int j,
                       int Q, double* G) {
    enum { BS = 64 };
    memset(G, 0, sizeof(double) * (size_t)vl * vr);
    for (int a0 = 0; a0 < vl; a0 += BS) {
        const int a1 = a0 + BS < vl ? a0 + BS : vl;
        for (int b0 = 0; b0 < vr; b0 += BS) {
            const int b1 = b0 + BS < vr ? b0 + BS : vr;
            for (int q = 0; q < Q; ++q) {
                const double* bi = Bl + ((size_t)q * ol + i) * vl;
                const double* bj = Br + ((size_t)q * orr + j) * vr;
                int a = a0;
                for (; a + 4 <= a1; a += 4) {
                    const double x0 = bi[a], x1 = bi[a + 1], x2 = bi[a + 2], x3 = bi[a + 3];
                    double* g0 = G + (size_t)a * vr;
                    double* g1 = g0 + vr;
                    double* g2 = g1 + vr;
                    double* g3 = g2 + vr;
                    for (int b = b0; b < b1; ++b) {
                        const double y = bj[b];
                        g0[b] += x0 * y;
                        g1[b] += x1 * y;
                        g2[b] += x2 * y;
                        g3[b] += x3 * y;
                    }
                }
                for (; a < a1; ++a) {
                    const double x = bi[a];
                    double* g = G + (size_t)a * vr;
                    for (int b = b0; b < b1; ++b) g[b] += x * bj[b];
                }
            }
        }
    }
}

/* ----------------------------------------------------------------- RMP2 */

/*
 * CD-RMP2 correlation energy.  Follows cholesky_rmp2.cc:118-155: for every
 * ordered pair (i,j), G_ab = sum_Q B^Q_ia B^Q_jb, D = e_i + e_j - e_a - e_b,
 * skip |D| < 1e-12, E += (2 G_ab - G_ba) * G_ab / D, t2(i,j,a,b) = G_ab / D.
 * `rows` (nullable) restricts i to a subset (CPU-baseline sampling); the return
 * is then the partial sum over those rows.
 */
double orc_rmp2_energy(const double* B, int Q, int o, int v,
                       const double* eo, const double* ev,
                       const int* rows, int nrows, double* t2) {
    if (!rows) nrows = o;
    double* pe = (double*)calloc((size_t)nrows * o, sizeof(double));
#pragma omp parallel
    {
        double* G = (double*)malloc(sizeof(double) * (size_t)v * v);
#pragma omp for collapse(2) schedule(dynamic)
        for (int r = 0; r < nrows; ++r)
            for (int j = 0; j < o; ++j) {
                int i = rows ? rows[r] : r;
                pair_block(B, o, v, i, B, o, v, j, Q, G);
                nsum_t acc = {0.0, 0.0};
                for (int a = 0; a < v; ++a)
                    for (int b = 0; b < v; ++b) {
                        double D = eo[i] + eo[j] - ev[a] - ev[b];
                        if (fabs(D) < 1e-12) continue;
                        double g = G[(size_t)a * v + b], gx = G[(size_t)b * v + a];
                        double t = g / D;
                        if (t2) t2[(((size_t)i * o + j) * v + a) * v + b] = t;
                        nsum_add(&acc, (2.0 * g - gx) * t);
                    }
                pe[(size_t)r * o + j] = nsum_val(&acc);
            }
        free(G);
    }
    nsum_t tot = {0.0, 0.0};
    for (size_t k = 0; k < (size_t)nrows * o; ++k) nsum_add(&tot, pe[k]);
    free(pe);
    return nsum_val(&tot);
}

/*
 * Energy plus gradient of the CD-RMP2 energy (no reference routine; the
 * derivative of the cholesky_rmp2.cc:118-155 energy).  With t = G/D and
 * T = 2t - t^T for each ordered pair (i,j):
 *   dE/dB^Q_ia  = 4 sum_{j,b} T^{ij}_ab B^Q_jb          (= 4 Gamma^Q_ia)
 *   dE/deps_i   = -2 sum_{j,a,b} t^{ij}_ab T^{ij}_ab
 *   dE/deps_a   = +2 sum_{i,j,b} t^{ij}_ab T^{ij}_ab
 * Pairs with |D| < 1e-12 contribute nothing (the reference's skip).
 * Threads own rows i, so dEdB is written without races or atomics.
 * `rows` (nullable) restricts i; dEdB/deo/dev then hold partial sums.
 */
double orc_rmp2_grad(const double* B, int Q, int o, int v,
                     const double* eo, const double* ev,
                     const int* rows, int nrows,
                     double* dEdB, double* deo, double* dev) {
    if (!rows) nrows = o;
    double* pe = (double*)calloc((size_t)nrows * o, sizeof(double));
    double* ps = (double*)calloc((size_t)nrows * o, sizeof(double));
    double* pv = (double*)calloc((size_t)nrows * v, sizeof(double));
    if (dEdB) memset(dEdB, 0, sizeof(double) * (size_t)Q * o * v);
#pragma omp parallel
    {
        double* G = (double*)malloc(sizeof(double) * (size_t)v * v);
        double* T = (double*)malloc(sizeof(double) * (size_t)v * v);
#pragma omp for schedule(dynamic)
        for (int r = 0; r < nrows; ++r) {
            int i = rows ? rows[r] : r;
            double* pvr = pv + (size_t)r * v;
            for (int j = 0; j < o; ++j) {
                pair_block(B, o, v, i, B, o, v, j, Q, G);
                nsum_t e = {0.0, 0.0}, s = {0.0, 0.0};
                for (int a = 0; a < v; ++a) {
                    nsum_t ra = {0.0, 0.0};
                    for (int b = 0; b < v; ++b) {
                        double D = eo[i] + eo[j] - ev[a] - ev[b];
                        double g = G[(size_t)a * v + b], gx = G[(size_t)b * v + a];
                        double Tab = 0.0;
                        if (fabs(D) >= 1e-12) {
                            double t = g / D;
                            Tab = (2.0 * g - gx) / D;
                            nsum_add(&e, g * Tab);
                            nsum_add(&s, t * Tab);
                            nsum_add(&ra, t * Tab);
                        }
                        T[(size_t)a * v + b] = Tab;
                    }
                    pvr[a] += nsum_val(&ra);
                }
                /* column sums: sum_b t_ba T_ba goes to eps_a */
                for (int a = 0; a < v; ++a) {
                    nsum_t ca = {0.0, 0.0};
                    for (int b = 0; b < v; ++b) {
                        double D = eo[i] + eo[j] - ev[a] - ev[b];
                        if (fabs(D) < 1e-12) continue;
                        double gba = G[(size_t)b * v + a];
                        nsum_add(&ca, (gba / D) * T[(size_t)b * v + a]);
                    }
                    pvr[a] += nsum_val(&ca);
                }
                pe[(size_t)r * o + j] = nsum_val(&e);
                ps[(size_t)r * o + j] = nsum_val(&s);
                if (dEdB) {
                    for (int q = 0; q < Q; ++q) {
                        const double* bj = B + ((size_t)q * o + j) * v;
                        double* gi = dEdB + ((size_t)q * o + i) * v;
                        for (int a = 0; a < v; ++a) {
                            const double* Ta = T + (size_t)a * v;
                            double acc = 0.0;
                            for (int b = 0; b < v; ++b) acc += Ta[b] * bj[b];
                            gi[a] += 4.0 * acc;
                        }
                    }
                }
            }
        }
        free(G);
        free(T);
    }
    nsum_t tot = {0.0, 0.0};
    for (size_t k = 0; k < (size_t)nrows * o; ++k) nsum_add(&tot, pe[k]);
    if (deo) {
        /* eps_k enters D^{kj} and D^{ik}: each ordered pair (i,j) adds -s_ij
         * to both of its occupied indices; with s_ij = s_ji the total is
         * dE/deps_k = -2 sum_j s_kj. */
        for (int k = 0; k < o; ++k) deo[k] = 0.0;
        for (int r = 0; r < nrows; ++r) {
            int i = rows ? rows[r] : r;
            for (int j = 0; j < o; ++j) {
                double s = ps[(size_t)r * o + j];
                deo[i] -= s;
                deo[j] -= s;
            }
        }
    }
    if (dev) {
        for (int a = 0; a < v; ++a) {
            nsum_t acc = {0.0, 0.0};
            for (int r = 0; r < nrows; ++r) nsum_add(&acc, pv[(size_t)r * v + a]);
            dev[a] = nsum_val(&acc);
        }
    }
    free(pe);
    free(ps);
    free(pv);
    return nsum_val(&tot);
}

/* ----------------------------------------------------------------- UMP2 */

/*
 * Same-spin energy, cholesky_ump2.cc:299-358 (and :364-411 for beta):
 * E_ss = 1/4 sum_{ijab} (g_ijab - g_ijba)^2 / D over all ordered i,j,a,b,
 * skip |D| < 1e-12.  Optional gradient (derivative of that expression):
 *   dE/dB^Q_ia = 2 sum_{j,b} u^{ij}_ab B^Q_jb,  u = (G - G^T)/D
 *   dE/deps_i  = -1/2 sum_{j,a,b} (u^{ij}_ab)^2,  dE/deps_a = +1/2 sum_{i,j,b} (u^{ij}_ab)^2
 */
static double ss_block(const double* B, int Q, int o, int v,
                       const double* eo, const double* ev,
                       double* dEdB, double* deo, double* dev) {
    double* pe = (double*)calloc((size_t)o * o, sizeof(double));
    double* ps = (double*)calloc((size_t)o * o, sizeof(double));
    double* pv = (double*)calloc((size_t)o * v, sizeof(double));
    if (dEdB) memset(dEdB, 0, sizeof(double) * (size_t)Q * o * v);
#pragma omp parallel
    {
        double* G = (double*)malloc(sizeof(double) * (size_t)v * v);
        double* U = (double*)malloc(sizeof(double) * (size_t)v * v);
#pragma omp for schedule(dynamic)
        for (int i = 0; i < o; ++i) {
            for (int j = 0; j < o; ++j) {
                pair_block(B, o, v, i, B, o, v, j, Q, G);
                nsum_t e = {0.0, 0.0}, s = {0.0, 0.0};
                for (int a = 0; a < v; ++a) {
                    nsum_t ra = {0.0, 0.0};
                    for (int b = 0; b < v; ++b) {
                        double D = eo[i] + eo[j] - ev[a] - ev[b];
                        double u = 0.0;
                        if (fabs(D) >= 1e-12) {
                            double A = G[(size_t)a * v + b] - G[(size_t)b * v + a];
                            u = A / D;
                            nsum_add(&e, 0.25 * A * u);
                            nsum_add(&s, u * u);
                            nsum_add(&ra, u * u);
                        }
                        U[(size_t)a * v + b] = u;
                    }
                    pv[(size_t)i * v + a] += nsum_val(&ra);
                }
                pe[(size_t)i * o + j] = nsum_val(&e);
                ps[(size_t)i * o + j] = nsum_val(&s);
                if (dEdB) {
                    for (int q = 0; q < Q; ++q) {
                        const double* bj = B + ((size_t)q * o + j) * v;
                        double* gi = dEdB + ((size_t)q * o + i) * v;
                        for (int a = 0; a < v; ++a) {
                            const double* Ua = U + (size_t)a * v;
                            double acc = 0.0;
                            for (int b = 0; b < v; ++b) acc += Ua[b] * bj[b];
                            gi[a] += 2.0 * acc;
                        }
                    }
                }
            }
        }
        free(G);
        free(U);
    }
    nsum_t tot = {0.0, 0.0};
    for (size_t k = 0; k < (size_t)o * o; ++k) nsum_add(&tot, pe[k]);
    if (deo) {
        for (int k = 0; k < o; ++k) deo[k] = 0.0;
        for (int i = 0; i < o; ++i)
            for (int j = 0; j < o; ++j) {
                /* u^{ij} enters D with eps_i and eps_j; the 1/4 prefactor times
                 * d(A^2/D)/dD = -u^2 gives -1/4 u^2 per occupied index. */
                deo[i] -= 0.25 * ps[(size_t)i * o + j];
                deo[j] -= 0.25 * ps[(size_t)i * o + j];
            }
    }
    if (dev) {
        /* Row sums over b give eps_a; by the a<->b antisymmetry of u the column
         * sums are equal, so the 1/4 * 2 = 1/2 prefactor applies to row sums. */
        for (int a = 0; a < v; ++a) {
            nsum_t acc = {0.0, 0.0};
            for (int i = 0; i < o; ++i) nsum_add(&acc, pv[(size_t)i * v + a]);
            dev[a] = 0.5 * nsum_val(&acc);
        }
    }
    free(pe);
    free(ps);
    free(pv);
    return nsum_val(&tot);
}

/*
 * Opposite-spin energy, cholesky_ump2.cc:417-465:
 * E_os = sum_{i in a, j in b, a, b} g^2 / D,  g = sum_Q Ba^Q_ia Bb^Q_jb.
 * Gradient: dE/dBa^Q_ia = 2 sum_{jb} t_ab Bb^Q_jb, dE/dBb^Q_jb = 2 sum_{ia} t_ab Ba^Q_ia,
 *           dE/deps = -/+ sum t^2 over the other three indices.
 */
static double os_block(const double* Ba, int oa, int va, const double* Bb, int ob, int vb,
                       int Q, const double* eoa, const double* eva,
                       const double* eob, const double* evb,
                       double* dBa, double* dBb, double* deoa, double* deva,
                       double* deob, double* devb) {
    double* pe = (double*)calloc((size_t)oa * ob, sizeof(double));
    double* ps = (double*)calloc((size_t)oa * ob, sizeof(double));
    double* pva = (double*)calloc((size_t)oa * va, sizeof(double));
    double* pvb = (double*)calloc((size_t)oa * vb, sizeof(double));
    double* Tst = NULL;
    if (dBa) memset(dBa, 0, sizeof(double) * (size_t)Q * oa * va);
    if (dBb) {
        memset(dBb, 0, sizeof(double) * (size_t)Q * ob * vb);
        Tst = (double*)malloc(sizeof(double) * (size_t)oa * ob * va * vb);
    }
#pragma omp parallel
    {
        double* G = (double*)malloc(sizeof(double) * (size_t)va * vb);
        double* T = (double*)malloc(sizeof(double) * (size_t)va * vb);
#pragma omp for schedule(dynamic)
        for (int i = 0; i < oa; ++i) {
            for (int j = 0; j < ob; ++j) {
                pair_block(Ba, oa, va, i, Bb, ob, vb, j, Q, G);
                nsum_t e = {0.0, 0.0}, s = {0.0, 0.0};
                for (int a = 0; a < va; ++a) {
                    nsum_t ra = {0.0, 0.0};
                    for (int b = 0; b < vb; ++b) {
                        double D = eoa[i] + eob[j] - eva[a] - evb[b];
                        double t = 0.0;
                        if (fabs(D) >= 1e-12) {
                            double g = G[(size_t)a * vb + b];
                            t = g / D;
                            nsum_add(&e, g * t);
                            nsum_add(&s, t * t);
                            nsum_add(&ra, t * t);
                            pvb[(size_t)i * vb + b] += t * t;
                        }
                        T[(size_t)a * vb + b] = t;
                    }
                    pva[(size_t)i * va + a] += nsum_val(&ra);
                }
                pe[(size_t)i * ob + j] = nsum_val(&e);
                ps[(size_t)i * ob + j] = nsum_val(&s);
                if (Tst) memcpy(Tst + ((size_t)i * ob + j) * va * vb, T, sizeof(double) * (size_t)va * vb);
                if (dBa) {
                    for (int q = 0; q < Q; ++q) {
                        const double* bj = Bb + ((size_t)q * ob + j) * vb;
                        double* gi = dBa + ((size_t)q * oa + i) * va;
                        for (int a = 0; a < va; ++a) {
                            const double* Ta = T + (size_t)a * vb;
                            double acc = 0.0;
                            for (int b = 0; b < vb; ++b) acc += Ta[b] * bj[b];
                            gi[a] += 2.0 * acc;
                        }
                    }
                }
            }
        }
        free(G);
        free(T);
    }
    if (dBb) {
        /* beta side owned by j so each thread writes its own rows */
#pragma omp parallel for schedule(dynamic)
        for (int j = 0; j < ob; ++j)
            for (int i = 0; i < oa; ++i) {
                const double* T = Tst + ((size_t)i * ob + j) * va * vb;
                for (int q = 0; q < Q; ++q) {
                    const double* bi = Ba + ((size_t)q * oa + i) * va;
                    double* gj = dBb + ((size_t)q * ob + j) * vb;
                    for (int a = 0; a < va; ++a) {
                        const double x = 2.0 * bi[a];
                        const double* Ta = T + (size_t)a * vb;
                        for (int b = 0; b < vb; ++b) gj[b] += Ta[b] * x;
                    }
                }
            }
        free(Tst);
    }
    nsum_t tot = {0.0, 0.0};
    for (size_t k = 0; k < (size_t)oa * ob; ++k) nsum_add(&tot, pe[k]);
    if (deoa) {
        for (int i = 0; i < oa; ++i) {
            nsum_t acc = {0.0, 0.0};
            for (int j = 0; j < ob; ++j) nsum_add(&acc, ps[(size_t)i * ob + j]);
            deoa[i] = -nsum_val(&acc);
        }
    }
    if (deob) {
        for (int j = 0; j < ob; ++j) {
            nsum_t acc = {0.0, 0.0};
            for (int i = 0; i < oa; ++i) nsum_add(&acc, ps[(size_t)i * ob + j]);
            deob[j] = -nsum_val(&acc);
        }
    }
    if (deva) {
        for (int a = 0; a < va; ++a) {
            nsum_t acc = {0.0, 0.0};
            for (int i = 0; i < oa; ++i) nsum_add(&acc, pva[(size_t)i * va + a]);
            deva[a] = nsum_val(&acc);
        }
    }
    if (devb) {
        for (int b = 0; b < vb; ++b) {
            nsum_t acc = {0.0, 0.0};
            for (int i = 0; i < oa; ++i) nsum_add(&acc, pvb[(size_t)i * vb + b]);
            devb[b] = nsum_val(&acc);
        }
    }
    free(pe);
    free(ps);
    free(pva);
    free(pvb);
    return nsum_val(&tot);
}

/* E[0] = E_ss(aa), E[1] = E_ss(bb), E[2] = E_os; gradients nullable.
 * deps arrays receive the sum of the ss and os contributions per spin. */
void orc_ump2(const double* Ba, const double* Bb, int Q, int oa, int va, int ob, int vb,
              const double* eoa, const double* eva, const double* eob, const double* evb,
              double* E, double* dBa, double* dBb,
              double* deoa, double* deva, double* deob, double* devb) {
    int grad = dBa != NULL;
    double* tBa = grad ? (double*)malloc(sizeof(double) * (size_t)Q * oa * va) : NULL;
    double* tBb = grad ? (double*)malloc(sizeof(double) * (size_t)Q * ob * vb) : NULL;
    double* t1 = grad ? (double*)malloc(sizeof(double) * (size_t)(oa + va + ob + vb) * 2) : NULL;
    double *so_a = t1, *sv_a = t1 ? t1 + oa : NULL, *so_b = t1 ? t1 + oa + va : NULL,
           *sv_b = t1 ? t1 + oa + va + ob : NULL;
    double* t2 = t1 ? t1 + (oa + va + ob + vb) : NULL;
    E[0] = ss_block(Ba, Q, oa, va, eoa, eva, dBa, so_a, sv_a);
    E[1] = ss_block(Bb, Q, ob, vb, eob, evb, dBb, so_b, sv_b);
    E[2] = os_block(Ba, oa, va, Bb, ob, vb, Q, eoa, eva, eob, evb, tBa, tBb,
                    t2, t2 ? t2 + oa : NULL, t2 ? t2 + oa + va : NULL, t2 ? t2 + oa + va + ob : NULL);
    if (grad) {
        for (size_t k = 0; k < (size_t)Q * oa * va; ++k) dBa[k] += tBa[k];
        for (size_t k = 0; k < (size_t)Q * ob * vb; ++k) dBb[k] += tBb[k];
        for (int k = 0; k < oa; ++k) deoa[k] = so_a[k] + t2[k];
        for (int k = 0; k < va; ++k) deva[k] = sv_a[k] + t2[oa + k];
        for (int k = 0; k < ob; ++k) deob[k] = so_b[k] + t2[oa + va + k];
        for (int k = 0; k < vb; ++k) devb[k] = sv_b[k] + t2[oa + va + ob + k];
        free(tBa);
        free(tBb);
        free(t1);
    }
}

/* ------------------------------------------------------------ J/K build */

/*
 * CD Fock pieces, cholesky_rhf.cc:108-136 (RHF, one density, optional
 * screening ||L^Q||_F <= 1e-8 skips K) and cholesky_uhf.cc:81-102 (UHF: J from
 * P_a + P_b, one K per spin, no screening).
 *   J = sum_Q tr(L^Q P_tot) L^Q,   K_s = sum_Q L^Q P_s L^Q
 * P: [nspin][n][n]; K: [nspin][n][n].  Threads own Q ranges and the partial
 * matrices are merged in thread-index order (fixed for a fixed thread count).
 */
void orc_jk(const double* L, int Q, int n, const double* P, int nspin, int screen,
            double* J, double* K) {
    size_t nn = (size_t)n * n;
    double* Pt = (double*)calloc(nn, sizeof(double));
    for (int s = 0; s < nspin; ++s)
        for (size_t k = 0; k < nn; ++k) Pt[k] += P[s * nn + k];
    /* gamma_Q = <L^Q, P_tot> first (fixed order), then J */
    double* gam = (double*)calloc((size_t)Q, sizeof(double));
#pragma omp parallel for schedule(static)
    for (int q = 0; q < Q; ++q) {
        nsum_t acc = {0.0, 0.0};
        const double* Lq = L + q * nn;
        for (size_t k = 0; k < nn; ++k) nsum_add(&acc, Lq[k] * Pt[k]);
        gam[q] = nsum_val(&acc);
    }
#pragma omp parallel for schedule(static)
    for (long k = 0; k < (long)nn; ++k) {
        nsum_t acc = {0.0, 0.0};
        for (int q = 0; q < Q; ++q) nsum_add(&acc, gam[q] * L[q * nn + k]);
        J[k] = nsum_val(&acc);
    }
    memset(K, 0, sizeof(double) * nn * nspin);
    int nth = 1;
#ifdef _OPENMP
    nth = omp_get_max_threads();
#endif
    double* Kp = (double*)calloc(nn * nspin * nth, sizeof(double));
#pragma omp parallel
    {
        int tid = 0;
#ifdef _OPENMP
        tid = omp_get_thread_num();
#endif
        double* T = (double*)malloc(sizeof(double) * nn);
        double* Kl = Kp + (size_t)tid * nn * nspin;
#pragma omp for schedule(static)
        for (int q = 0; q < Q; ++q) {
            const double* Lq = L + q * nn;
            if (screen) {
                double f = 0.0;
                for (size_t k = 0; k < nn; ++k) f += Lq[k] * Lq[k];
                if (!(sqrt(f) > 1e-8)) continue;
            }
            for (int s = 0; s < nspin; ++s) {
                const double* Ps = P + s * nn;
                /* T = P_s L^Q, then K_s += L^Q T */
                for (int m = 0; m < n; ++m)
                    for (int c = 0; c < n; ++c) {
                        double acc = 0.0;
                        for (int k = 0; k < n; ++k) acc += Ps[(size_t)m * n + k] * Lq[(size_t)k * n + c];
                        T[(size_t)m * n + c] = acc;
                    }
                double* Ks = Kl + s * nn;
                for (int m = 0; m < n; ++m)
                    for (int c = 0; c < n; ++c) {
                        double acc = 0.0;
                        for (int k = 0; k < n; ++k) acc += Lq[(size_t)m * n + k] * T[(size_t)k * n + c];
                        Ks[(size_t)m * n + c] += acc;
                    }
            }
        }
        free(T);
    }
    for (int t = 0; t < nth; ++t)
        for (size_t k = 0; k < nn * nspin; ++k) K[k] += Kp[(size_t)t * nn * nspin + k];
    free(Kp);
    free(gam);
    free(Pt);
}

/* ----------------------------------------------------------- transforms */

/* B^Q = C_occ^T (L^Q C_vir): cholesky_rmp2.cc:80-89 (virtual-first, as the
 * reference does) and cholesky_ump2.cc:265-284 (same product per spin). */
void orc_transform(const double* L, int Q, int n, const double* Co, int o,
                   const double* Cv, int v, double* B) {
    size_t nn = (size_t)n * n;
#pragma omp parallel
    {
        double* T = (double*)malloc(sizeof(double) * (size_t)n * v);
#pragma omp for schedule(dynamic)
        for (int q = 0; q < Q; ++q) {
            const double* Lq = L + q * nn;
            for (int m = 0; m < n; ++m)
                for (int a = 0; a < v; ++a) {
                    double acc = 0.0;
                    for (int k = 0; k < n; ++k) acc += Lq[(size_t)m * n + k] * Cv[(size_t)k * v + a];
                    T[(size_t)m * v + a] = acc;
                }
            double* Bq = B + (size_t)q * o * v;
            for (int i = 0; i < o; ++i)
                for (int a = 0; a < v; ++a) {
                    double acc = 0.0;
                    for (int m = 0; m < n; ++m) acc += Co[(size_t)m * o + i] * T[(size_t)m * v + a];
                    Bq[(size_t)i * v + a] = acc;
                }
        }
        free(T);
    }
}

/* Half transform X^Q = L^Q C  ([Q][n][m]); the occupied factor of K. */
void orc_half_transform(const double* L, int Q, int n, const double* C, int m, double* X) {
    size_t nn = (size_t)n * n;
#pragma omp parallel for schedule(static)
    for (int q = 0; q < Q; ++q) {
        const double* Lq = L + q * nn;
        double* Xq = X + (size_t)q * n * m;
        for (int r = 0; r < n; ++r)
            for (int c = 0; c < m; ++c) {
                double acc = 0.0;
                for (int k = 0; k < n; ++k) acc += Lq[(size_t)r * n + k] * C[(size_t)k * m + c];
                Xq[(size_t)r * m + c] = acc;
            }
    }
}

/* ------------------------------------------------- pivoted Cholesky (a1) */

/*
 * Pivoted incomplete Cholesky of the n^2 x n^2 ERI matrix, following
 * cholesky_eri.cc:40-203: diagonal D(ij) = (ij|ij) over ordered pairs
 * ij = i*n + j; pivot = first index of the maximum (strict '>' scan,
 * :155-167); stop when D(pivot) < thr; new vector
 * L(ij) = ((ij|p) - sum_k L_k(ij) L_k(p)) / sqrt(D(p)) (divide only when
 * sqrt(D(p)) > 1e-12, :195-200); D -= L^2 and clamp negatives to 0
 * (:107-110).  Returns the number of vectors (<= maxvec), L as [nvec][n*n].
 */
int orc_cholesky(const double* eri, int n, double thr, int maxvec, double* L) {
    int np = n * n;
    double* D = (double*)malloc(sizeof(double) * np);
    for (int i = 0; i < n; ++i)
        for (int j = 0; j < n; ++j) {
            int ij = i * n + j;
            D[ij] = eri[(size_t)ij * np + ij];
        }
    double dmax = D[0];
    for (int k = 1; k < np; ++k)
        if (D[k] > dmax) dmax = D[k];
    if (dmax < thr) {
        free(D);
        return 0;
    }
    int nv = 0;
    int limit = maxvec < np ? maxvec : np;
    while (nv < limit) {
        int p = 0;
        double mv = D[0];
        for (int k = 1; k < np; ++k)
            if (D[k] > mv) {
                mv = D[k];
                p = k;
            }
        if (mv < thr) break;
        double* Ln = L + (size_t)nv * np;
        for (int ij = 0; ij < np; ++ij) {
            double val = eri[(size_t)ij * np + p];
            for (int k = 0; k < nv; ++k) val -= L[(size_t)k * np + ij] * L[(size_t)k * np + p];
            Ln[ij] = val;
        }
        double nf = sqrt(D[p]);
        if (nf > 1e-12)
            for (int ij = 0; ij < np; ++ij) Ln[ij] /= nf;
        for (int ij = 0; ij < np; ++ij) {
            D[ij] -= Ln[ij] * Ln[ij];
            if (D[ij] < 0.0) D[ij] = 0.0;
        }
        ++nv;
    }
    free(D);
    return nv;
}

int orc_num_threads(void) {
#ifdef _OPENMP
    return omp_get_max_threads();
#else
    return 1;
#endif
}

/*
 * The per-pair work of orc_rmp2_grad for an explicit list of ordered pairs
 * (CPU-baseline sampling): G, amplitudes, energy and the dEdB row update of
 * each pair.  Pairs must have distinct first indices (each thread then owns
 * its rows of dEdB).  Returns the sum of the listed pair energies.
 */
double orc_rmp2_grad_pairs(const double* B, int Q, int o, int v,
                           const double* eo, const double* ev,
                           const int* pairs, int npairs, double* dEdB) {
    double* pe = (double*)calloc((size_t)npairs, sizeof(double));
#pragma omp parallel
    {
        double* G = (double*)malloc(sizeof(double) * (size_t)v * v);
        double* T = (double*)malloc(sizeof(double) * (size_t)v * v);
#pragma omp for schedule(dynamic)
        for (int p = 0; p < npairs; ++p) {
            const int i = pairs[2 * p], j = pairs[2 * p + 1];
            pair_block(B, o, v, i, B, o, v, j, Q, G);
            nsum_t e = {0.0, 0.0};
            for (int a = 0; a < v; ++a)
                for (int b = 0; b < v; ++b) {
                    double D = eo[i] + eo[j] - ev[a] - ev[b];
                    double g = G[(size_t)a * v + b], gx = G[(size_t)b * v + a];
                    double Tab = fabs(D) >= 1e-12 ? (2.0 * g - gx) / D : 0.0;
                    nsum_add(&e, g * Tab);
                    T[(size_t)a * v + b] = Tab;
                }
            pe[p] = nsum_val(&e);
            for (int q = 0; q < Q; ++q) {
                const double* bj = B + ((size_t)q * o + j) * v;
                double* gi = dEdB + ((size_t)q * o + i) * v;
                for (int a = 0; a < v; ++a) {
                    const double* Ta = T + (size_t)a * v;
                    double acc = 0.0;
                    for (int b = 0; b < v; ++b) acc += Ta[b] * bj[b];
                    gi[a] += 4.0 * acc;
                }
            }
        }
        free(G);
        free(T);
    }
    nsum_t tot = {0.0, 0.0};
    for (int p = 0; p < npairs; ++p) nsum_add(&tot, pe[p]);
    free(pe);
    return nsum_val(&tot);
}

/*
 * The reference's energy loop as it is written (cholesky_rmp2.cc:118-155),
 * for the CPU timing baseline: for each listed ordered pair (i, j) the v x v
 * block Int_dir (column-major, as Eigen holds it) is built by Q rank-1
 * updates Int_dir += B^k_i B^k_j^T, column by column, then the epilogue adds
 * (2 (ia|jb) - (ib|ja)) t_ij^ab with the |D| < 1e-12 skip.  The t2 store of
 * the reference is left out (at o = 100, v = 1100 it is 97 GB).  OpenMP over
 * the listed pairs with a reduction, as the reference's collapse(2) loop.
 * Returns the sum of the listed pair energies.
 */
double orc_rmp2_energy_pairs_rank1(const double* B, int Q, int o, int v,
                                   const double* eo, const double* ev,
                                   const int* pairs, int npairs) {
    double e_corr = 0.0;
#pragma omp parallel reduction(+ : e_corr)
    {
        double* G = (double*)malloc(sizeof(double) * (size_t)v * v);
#pragma omp for schedule(dynamic)
        for (int p = 0; p < npairs; ++p) {
            const int i = pairs[2 * p], j = pairs[2 * p + 1];
            memset(G, 0, sizeof(double) * (size_t)v * v);
            for (int k = 0; k < Q; ++k) {
                const double* x = B + ((size_t)k * o + i) * v;
                const double* y = B + ((size_t)k * o + j) * v;
                for (int b = 0; b < v; ++b) {
                    double* col = G + (size_t)b * v;
                    const double yb = y[b];
                    for (int a = 0; a < v; ++a) col[a] += x[a] * yb;
                }
            }
            for (int a = 0; a < v; ++a)
                for (int b = 0; b < v; ++b) {
                    const double denom = eo[i] + eo[j] - ev[a] - ev[b];
                    if (fabs(denom) < 1e-12) continue;
                    const double iajb = G[a + (size_t)v * b], ibja = G[b + (size_t)v * a];
                    const double t = iajb / denom;
                    e_corr += (2.0 * iajb - ibja) * t;
                }
        }
        free(G);
    }
    return e_corr;
}
