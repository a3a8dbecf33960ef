/*
 * mshqc_hip.h -- C ABI of the MI355X (gfx950) Cholesky-vector hot path.
 *
 * One shared library, libmshqc_hip.so, replaces the OpenMP/Eigen loops that
 * syahrulhidayat/mshqc runs inside its SCF and MP2 classes.  Every entry point
 * takes plain pointers and sizes; no C++ or torch types cross this boundary.
 *
 * Conventions
 *   - All array arguments are DEVICE pointers (hipMalloc'd, e.g. through
 *     mshqc_hip_alloc, or torch tensors' data_ptr()), C order, FP64:
 *       L  [Q][n][n]   Cholesky vectors L^Q_{mu nu} (symmetric in mu,nu)
 *       P  [nspin][n][n]  densities;  Y [n][m] density factor (P = Y Y^T)
 *       C  [n][m]      MO coefficient block (columns = orbitals)
 *       B  [Q][o][v]   B^Q_{ia};  dEdB has the same layout
 *       t2 [o][o][v][v]
 *     Scalar energies are written to HOST pointers.
 *   - Calls are synchronous: work is queued on the context's stream and the
 *     stream is synchronised before return.  A context is not re-entrant;
 *     use one context per host thread.
 *   - Every call returns MSHQC_HIP_OK (0) or an error code; the message is
 *     available from mshqc_hip_last_error().  No C++ exception crosses the ABI.
 *   - Reductions run in a fixed order: results are bit-reproducible for a
 *     given device, shape and precision.
 *   - precision: MSHQC_PREC_FP64 (0) computes in FP64 on FP64 MFMA.  A value
 *     k >= 2 selects the split-BF16 form with k BF16 slices per operand
 *     (k = 2 is the hi+lo three-product form); slice products with
 *     p + q <= k - 1 are kept, the two leading product classes accumulated
 *     exactly in FP32 within 128-wide Q (or contraction) blocks, and the
 *     blocks summed in FP64.  Denominators and every energy/gradient sum
 *     stay FP64.
 */
#ifndef MSHQC_HIP_H
#define MSHQC_HIP_H

#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

#define MSHQC_HIP_OK 0
#define MSHQC_HIP_ERR_ARG 1
#define MSHQC_HIP_ERR_DEVICE 2
#define MSHQC_HIP_ERR_ALLOC 3
#define MSHQC_HIP_ERR_UNSUPPORTED 4

#define MSHQC_PREC_FP64 0

typedef struct mshqc_hip_ctx mshqc_hip_ctx;

/* ------------------------------------------------------------ context */

/* Create a context on `device` with its own stream. */
int mshqc_hip_ctx_create(int device, mshqc_hip_ctx** out);
/* Destroy a context, its stream and its workspace.  NULL is accepted. */
int mshqc_hip_ctx_destroy(mshqc_hip_ctx* ctx);
/* Message of the last failed call on `ctx` ("" if none). */
const char* mshqc_hip_last_error(const mshqc_hip_ctx* ctx);
/* Queue subsequent work on `stream` (a hipStream_t); NULL restores the
 * context's own stream. */
int mshqc_hip_set_stream(mshqc_hip_ctx* ctx, void* stream);
/* Device time, in ms, of the last compute call (HIP events around its kernels). */
int mshqc_hip_last_ms(const mshqc_hip_ctx* ctx, float* ms);
/* Cap, in doubles, of the per-occupied-block amplitude workspace used by the
 * gradient calls and by mshqc_hip_omp2_opdm, and of the (ac|bd) blocks of the
 * MP3 calls (default 2^30 = 8 GiB; 0 restores the default).  Smaller caps mean
 * more, smaller blocks; results do not depend on it beyond rounding. */
int mshqc_hip_set_zblock(mshqc_hip_ctx* ctx, size_t elems);
/* Per-kernel-class timing (HIP events around each launch, on the stream the
 * kernel runs on).  kind: 0 = pair-block (ia|jb) kernels, 1 = GEMM kernels
 * (gradient contractions, K, transforms), 2 = J streaming kernels.
 * mshqc_hip_profile(ctx, 1) clears and enables the counters, 0 disables. */
int mshqc_hip_profile(mshqc_hip_ctx* ctx, int enable);
int mshqc_hip_profile_get(const mshqc_hip_ctx* ctx, int kind, long* launches, double* total_ms);
/* Library build tag, e.g. "mshqc_hip gfx950 <date>". */
const char* mshqc_hip_version(void);

/* ------------------------------------------------------------- memory */

int mshqc_hip_alloc(mshqc_hip_ctx* ctx, size_t bytes, void** dptr);
int mshqc_hip_free(mshqc_hip_ctx* ctx, void* dptr);
int mshqc_hip_memcpy_h2d(mshqc_hip_ctx* ctx, void* dst, const void* src, size_t bytes);
int mshqc_hip_memcpy_d2h(mshqc_hip_ctx* ctx, void* dst, const void* src, size_t bytes);

/* ----------------------------------------------------- Cholesky vectors */

/*
 * Pivoted incomplete Cholesky of the ERI matrix, on the device in FP64.
 * Replaces CholeskyERI::decompose (src/integrals/cholesky_eri.cc:40-203):
 *   diagonal D(ij) = (ij|ij); repeatedly take the first maximum p of D, stop
 *   when D_p < threshold or max_vectors vectors exist;
 *   L_new(ij) = ((ij|p) - sum_k L_k(ij) L_k(p)) / sqrt(D_p),
 *   D(ij) -= L_new(ij)^2, negative values clamped to 0.
 * eri: DEVICE [n*n][n*n], element (ij|kl) at (i*n+j)*n*n + k*n+l.
 * L:   DEVICE [max_vectors][n*n], vectors in pivot order.
 * Operation order and rounding follow the reference's scalar loop (no fused
 * multiply-adds), so vectors and pivots equal the FP64 host algorithm's.
 * Host outputs: *n_vectors; *converged = 1 when the threshold was reached
 * (0 when max_vectors stopped it, as the reference's n*n bound does);
 * *max_residual = the largest remaining diagonal.  converged and
 * max_residual may be NULL.
 */
int mshqc_hip_cholesky(mshqc_hip_ctx* ctx, const double* eri, int n, double threshold,
                       int max_vectors, double* L, int* n_vectors, int* converged,
                       double* max_residual);

/* --------------------------------------------------------- Fock pieces */

/*
 * Coulomb and exchange from Cholesky vectors with general densities.
 * Replaces CholeskyRHF::build_fock (src/scf/cholesky_rhf.cc:99-137) and
 * CholeskyUHF::build_fock_cholesky (src/scf/cholesky_uhf.cc:77-103):
 *   J      = sum_Q tr(L^Q P_tot) L^Q,   P_tot = sum_s P_s
 *   K[s]   = sum_Q L^Q P_s L^Q
 * screen != 0 skips K for vectors with ||L^Q||_F <= 1e-8 (the RHF
 * `screen_exchange` option, cholesky_rhf.cc:120-126).  The caller assembles
 * F = H + J - K/2 (RHF) or F_s = H + J - K[s] (UHF).
 */
int mshqc_hip_jk(mshqc_hip_ctx* ctx, const double* L, int Q, int n,
                 const double* P, int nspin, int screen, double* J, double* K);

/*
 * Coulomb and exchange for a factored density P = Y Y^T (Y is n x m):
 *   J = sum_Q tr(L^Q P) L^Q,   K = sum_Q (L^Q Y)(L^Q Y)^T
 * The occupied-factor exchange of CholeskyROHF (src/scf/cholesky_rohf.cc:
 * 125-133); with Y = sqrt(2) C_occ it equals the RHF K of an idempotent
 * density (cholesky_rhf.cc:90-94,127).
 */
int mshqc_hip_jk_factor(mshqc_hip_ctx* ctx, const double* L, int Q, int n,
                        const double* Y, int m, double* J, double* K);

/*
 * Restricted open-shell Coulomb and exchange: one spatial orbital set C
 * (row-major, n x ldc), alpha electrons in columns [0, na), beta in [0, nb):
 *   J  = sum_Q tr(L^Q (Pa + Pb)) L^Q,     Pa = C_a C_a^T, Pb = C_b C_b^T
 *   Ka = sum_Q (L^Q C_a)(L^Q C_a)^T,      Kb = sum_Q (L^Q C_b)(L^Q C_b)^T
 * Replaces CholeskyROHF::build_fock_cholesky (src/scf/cholesky_rohf.cc:
 * 89-147), whose C_beta equals C_alpha on every iteration (:77-83, :235-236);
 * the caller assembles F_a = H + J - Ka, F_b = H + J - Kb.  One half
 * transform L^Q C_{0..max(na,nb)} serves both spins.  na = 0 or nb = 0 give
 * zero matrices for that spin.
 */
int mshqc_hip_jk_rohf(mshqc_hip_ctx* ctx, const double* L, int Q, int n,
                      const double* C, int ldc, int na, int nb, double* J,
                      double* Ka, double* Kb);

/* ---------------------------------------------------------- transforms */

/* Half transform X^Q = L^Q C  (X: [Q][n][m]). */
int mshqc_hip_half_transform(mshqc_hip_ctx* ctx, const double* L, int Q, int n,
                             const double* C, int m, double* X);

/*
 * Full transform B^Q = C_occ^T L^Q C_vir  (B: [Q][o][v]).
 * Replaces CholeskyRMP2::transform_vectors (src/mp2/cholesky_rmp2.cc:68-92)
 * and CholeskyUMP2::transform_cholesky_vectors (src/mp2/cholesky_ump2.cc:
 * 244-293, called once per spin).  Computed occupied-first.
 * precision 0: FP64 MFMA.  precision k in [2, 4]: split-BF16 GEMMs with k
 * slices of L, C_occ, C_vir and the half-transformed T^Q = L^Q C_occ, one
 * exponent per row and 256-wide block of the contracted AO index, the
 * first product class exact in FP32 per block, blocks summed in FP64; the
 * relative error of B is near 2^-(9k - 1) of its block scale (k = 4: about
 * 1e-10).  L is symmetric, so T = C_occ^T L^Q reads L^Q row-wise.
 */
int mshqc_hip_transform(mshqc_hip_ctx* ctx, const double* L, int Q, int n,
                        const double* Cocc, int o, const double* Cvir, int v,
                        int precision, double* B);

/* ---------------------------------------------------------------- RMP2 */

/*
 * CD-RMP2 correlation energy, replacing the pair loop of
 * CholeskyRMP2::compute (src/mp2/cholesky_rmp2.cc:118-155):
 *   E = sum_{ij} sum_{ab} (ia|jb) [2 (ia|jb) - (ib|ja)] / D_ij^ab,
 *   (ia|jb) = sum_Q B^Q_ia B^Q_jb,  D = e_i + e_j - e_a - e_b,
 * terms with |D| < 1e-12 skipped.  (ia|jb) never leaves the chip.
 */
int mshqc_hip_rmp2_energy(mshqc_hip_ctx* ctx, const double* B, int Q, int o, int v,
                          const double* eps_o, const double* eps_v, int precision,
                          double* E);

/*
 * Energy and gradient (no reference counterpart: the backward of the energy
 * above).  With t = (ia|jb)/D and T = 2t - t^T:
 *   dEdB^Q_ia   = 4 Gamma^Q_ia,  Gamma^Q_ia = sum_{jb} T_ij^ab B^Q_jb
 *   dE_deps_o_i = -2 sum_{jab} t T,   dE_deps_v_a = +2 sum_{ijb} t T
 * Amplitudes are recomputed per occupied pair block, never stored whole.
 * Any of dEdB / dE_deps_o / dE_deps_v may be NULL.
 */
int mshqc_hip_rmp2_grad(mshqc_hip_ctx* ctx, const double* B, int Q, int o, int v,
                        const double* eps_o, const double* eps_v, int precision,
                        double* E, double* dEdB, double* dE_deps_o, double* dE_deps_v);

/*
 * The FP64 CD-RMP2 energy of mshqc_hip_rmp2_energy (precision 0), allocating
 * nothing.  It has no counterpart in the reference: it is for callers that
 * cannot allocate device memory (code under HIP graph capture, tight serving
 * loops).  Same kernels and summation order, so the same energy bit for bit.
 *   - Scratch: a fixed 64 MiB device array per GPU, held until the process
 *     ends, holding the pair and tile lists and one partial per (pair,
 *     64 x 64 tile): about 8 (t + 2p + p t) bytes for p = o(o+1)/2 pairs and
 *     t = T(T+1)/2 tiles, T = ceil(v/64).  Sizes beyond it are rejected with
 *     MSHQC_HIP_ERR_ARG (o=100, v=1100 needs 7 MiB).
 *   - E is a DEVICE pointer to one double, overwritten.
 *   - Asynchronous: the kernels are queued on the context's stream and the
 *     call returns without synchronising; E is ready when that stream is.
 *   - Every call on one GPU uses that GPU's one scratch array, so calls that
 *     overlap on one GPU (two streams, two contexts, two threads) race.  Each
 *     call writes all of the scratch it reads, so nothing is cleared between
 *     calls.
 */
int mshqc_hip_rmp2_energy_noalloc(mshqc_hip_ctx* ctx, const double* B, int Q, int o, int v,
                                  const double* eps_o, const double* eps_v, double* E);

/*
 * Amplitudes t2[i][j][a][b] = (ia|jb)/D (0 where |D| < 1e-12): the t2 tensor
 * CholeskyRMP2 stores for CholeskyRMP3 (cholesky_rmp2.cc:145-148,160).
 */
int mshqc_hip_rmp2_t2(mshqc_hip_ctx* ctx, const double* B, int Q, int o, int v,
                      const double* eps_o, const double* eps_v, double* t2);

/* ---------------------------------------------------------------- UMP2 */

/*
 * CD-UMP2 energy components, replacing CholeskyUMP2::compute_ss_alpha,
 * compute_ss_beta and compute_os (src/mp2/cholesky_ump2.cc:299-465):
 *   E[0] = 1/4 sum_{ijab in aa} (g_ijab - g_ijba)^2 / D
 *   E[1] = the same for beta,  E[2] = sum_{i in a, j in b, ab} g^2 / D
 * A spin may have no occupied or no virtual orbitals (o = 0 or v = 0, e.g.
 * the beta spin of the H atom); its blocks then contribute 0, as the
 * reference's loops over empty ranges do, and its (empty) arrays may be
 * NULL.
 */
int mshqc_hip_ump2_energy(mshqc_hip_ctx* ctx, const double* Ba, const double* Bb,
                          int Q, int oa, int va, int ob, int vb,
                          const double* eps_oa, const double* eps_va,
                          const double* eps_ob, const double* eps_vb,
                          int precision, double* E);

/* Energy components plus gradients with respect to Ba, Bb and the orbital
 * energies of each spin (any gradient pointer may be NULL). */
int mshqc_hip_ump2_grad(mshqc_hip_ctx* ctx, const double* Ba, const double* Bb,
                        int Q, int oa, int va, int ob, int vb,
                        const double* eps_oa, const double* eps_va,
                        const double* eps_ob, const double* eps_vb,
                        int precision, double* E,
                        double* dEdBa, double* dEdBb,
                        double* dE_deps_oa, double* dE_deps_va,
                        double* dE_deps_ob, double* dE_deps_vb);

/*
 * UMP2 amplitudes of every spin block, replacing
 * CholeskyUMP2::compute_t2_amplitudes (src/mp2/cholesky_ump2.cc:471-556):
 *   t2_aa[i][j][a][b] = g_ijab / (e_i + e_j - e_a - e_b)   (alpha, all i, j)
 *   t2_bb likewise for beta,
 *   t2_ab[i][j][a][b] = sum_Q Ba^Q_ia Bb^Q_jb / D  (i alpha, j beta)
 * with g_ijab = sum_Q B^Q_ia B^Q_jb of the same spin and 0 where
 * |D| < 1e-12; not antisymmetrised, as the reference stores them.  DEVICE
 * outputs [oa][oa][va][va], [ob][ob][vb][vb], [oa][ob][va][vb]; any may be
 * NULL, and an empty block is skipped.  FP64.
 */
int mshqc_hip_ump2_t2(mshqc_hip_ctx* ctx, const double* Ba, const double* Bb,
                      int Q, int oa, int va, int ob, int vb,
                      const double* eps_oa, const double* eps_va,
                      const double* eps_ob, const double* eps_vb,
                      double* t2_aa, double* t2_bb, double* t2_ab);

/* ---------------------------------------------------------------- OMP2 */

/*
 * One CD-OMP2 iteration's amplitude work, replacing
 * CholeskyOMP2::compute_t2_amplitudes, compute_mp2_energy_from_t2 and
 * build_opdm_alpha/beta (src/mp2/cholesky_omp2.cc:94-326).  With
 * g_ijab = sum_Q B^Q_ia B^Q_jb and D the orbital-energy denominator
 * (|D| < 1e-12 gives t = 0):
 *   t_aa(ijab) = (g_ijab - g_ijba)/D (i != j, a != b; else 0), t_bb alike,
 *   t_ab(ijab) = g_ijab/D  (i alpha, j beta)
 *   E[0] = sum_{i<j, a<b} t_aa^2 D,  E[1] likewise for beta,
 *   E[2] = sum t_ab^2 D                                   (E: HOST, 3)
 *   Goo_a(i,j) = -(1/2 sum_kab t_aa(ikab) t_aa(jkab) + sum_kab t_ab(ikab) t_ab(jkab))
 *   Gvv_a(a,b) =  1/2 sum_ijc t_aa(ijac) t_aa(ijbc) + sum_ijc t_ab(ijac) t_ab(ijbc)
 *   Goo_b, Gvv_b with the roles of the spins exchanged in t_ab.
 * G blocks are DEVICE arrays [o][o] and [v][v].  The amplitudes are formed
 * block by block (at most the mshqc_hip_set_zblock cap per block) and never
 * stored whole.
 */
int mshqc_hip_omp2_opdm(mshqc_hip_ctx* ctx, const double* Ba, const double* Bb,
                        int Q, int oa, int va, int ob, int vb,
                        const double* eps_oa, const double* eps_va,
                        const double* eps_ob, const double* eps_vb, double* E,
                        double* Goo_a, double* Gvv_a, double* Goo_b, double* Gvv_b);

/*
 * Two-particle term of the CD-OMP2 orbital gradient for one spin, the term
 * the reference's w_ai = 2 F_vo leaves out (build_gfock_from_density and its
 * use at src/mp2/cholesky_omp2.cc:331-369,462-467 form F from the OPDM
 * only).  With G = dE/dB (DEVICE [Q][o][v], e.g. from mshqc_hip_ump2_grad
 * in FP64) and the same orbitals' Loo (DEVICE [Q][o][o]) and Lvv (DEVICE
 * [Q][v][v]):
 *   w2(c,k) = sum_Q [ sum_a Lvv^Q(c,a) G^Q(k,a) - sum_i G^Q(i,c) Loo^Q(i,k) ]
 * = dE_MP2/dkappa_ck at fixed denominators, for the rotation that adds
 * kappa_ck C_c to occupied orbital k and subtracts kappa_ck C_k from
 * virtual c (the convention in which the reference's 2 F_vo is the SCF
 * energy's gradient).  w2 is a DEVICE array [v][o], overwritten.
 */
int mshqc_hip_omp2_grad2p(mshqc_hip_ctx* ctx, const double* G, const double* Loo,
                          const double* Lvv, int Q, int o, int v, double* w2);

/*
 * One CD-OMP2 iteration's orbital work on the device, replacing
 * build_gfock_from_density, the gradient and the rotation of
 * CholeskyOMP2::compute (src/mp2/cholesky_omp2.cc:331-369, 375-425,
 * 451-490).  Per spin the orbitals are DEVICE blocks C_occ [n][o] and
 * C_vir [n][v] (columns = orbitals; NULL for an empty block), the density
 * P [n][n] and the MO-basis generalised Fock F [n][n].
 *
 * mshqc_hip_omp2_fock: with G_oo, G_vv from mshqc_hip_omp2_opdm,
 *   P_tot,s = P_s + C_occ G_oo C_occ^T + C_vir G_vv C_vir^T,
 *   F_s = C^T (H + J[P_tot,a + P_tot,b] - K[P_tot,s]) C,  C = [C_occ | C_vir].
 * mshqc_hip_omp2_grad_norm: *g2 (HOST) = sum_ai (2 F[o+a][i] + w2[a][i])^2,
 *   w2 [v][o] the optional two-particle term (NULL for none).
 * mshqc_hip_omp2_rotate: kappa_ai = 0.5 w_ai / d, w = 2 F_vo (+ w2),
 *   d = F_aa - F_ii + 0.2 raised to |d| >= 1e-2 (sign kept); C <- C exp(-kappa)
 *   (in place, the exponential by scaling and squaring of a Taylor series);
 *   P <- C_occ C_occ^T.  o = 0 or o = n leaves everything unchanged.
 */
int mshqc_hip_omp2_fock(mshqc_hip_ctx* ctx, const double* L, int Q, int n, const double* H,
                        const double* Cocc_a, const double* Cvir_a, int na,
                        const double* Cocc_b, const double* Cvir_b, int nb,
                        const double* P_a, const double* P_b,
                        const double* Goo_a, const double* Gvv_a,
                        const double* Goo_b, const double* Gvv_b,
                        double* F_a, double* F_b);
int mshqc_hip_omp2_grad_norm(mshqc_hip_ctx* ctx, int n, int o, const double* F, const double* w2,
                             double* g2);
int mshqc_hip_omp2_rotate(mshqc_hip_ctx* ctx, int n, int o, const double* F, const double* w2,
                          double* Cocc, double* Cvir, double* P);

/* ---------------------------------------------------------------- RMP3 */

/*
 * CD-RMP3 second-order amplitudes and energy, replacing the t2(2) and E(3)
 * loops of foundation::CholeskyRMP3::compute (src/mp3/cholesky_rmp3.cc:
 * 100-178).  Inputs: B [Q][o][v], Loo [Q][o][o], Lvv [Q][v][v] (each
 * C_left^T L^Q C_right, e.g. from mshqc_hip_transform), the first-order
 * amplitudes t1 [o][o][v][v] (the CD-RMP2 t2) and the orbital energies.
 *   D t2_2(ijab) = 1/2 sum_cd [(ac|bd) - (ad|bc)] t1(ijcd)
 *                + 1/2 sum_kl [(ki|lj) - (kj|li)] t1(klab)
 *                - sum_kc [2(jb|kc) - (kb|jc)] t1(ikac)
 *                - sum_kc [2(ia|kc) - (ka|ic)] t1(jkbc)
 * (|D| < 1e-12 gives 0);  *E3 (HOST) = sum_ijab [2(ia|jb) - (ib|ja)] t2_2.
 * t2_2 (DEVICE [o][o][v][v]) may be NULL.  Workspace: about 5 o^2 v^2 doubles
 * plus (ac|bd) blocks of at most twice the mshqc_hip_set_zblock cap.
 */
int mshqc_hip_rmp3(mshqc_hip_ctx* ctx, const double* B, const double* Loo,
                   const double* Lvv, int Q, int o, int v, const double* t1,
                   const double* eps_o, const double* eps_v, double* E3, double* t2_2);

/*
 * CD-UMP3 energy components E = {ss_aa, ss_bb, os} (HOST, 3), replacing
 * CholeskyUMP3::compute_t2_first_order and compute_mp3_ss_alpha / _beta /
 * _os (src/mp3/cholesky_ump3.cc:212-637), from per-spin B [Q][o][v],
 * Loo [Q][o][o], Lvv [Q][v][v] and orbital energies.  The reference builds
 * its first-order amplitudes (and the beta ring term) from its o x v blocks
 * read in column-major order as if row-major, i.e. from
 * B~[q][i*v + a] = B[q][(r mod o)*v + r div o] at r = i*v + a; this call
 * reproduces that.  A spin with no occupied or no virtual orbitals
 * contributes 0.  Workspace: a few (o v)^2 and o^2 v^2 blocks plus (ac|bd)
 * blocks bounded by the mshqc_hip_set_zblock cap.
 */
int mshqc_hip_ump3(mshqc_hip_ctx* ctx, const double* Ba, const double* Bb,
                   const double* Looa, const double* Loob, const double* Lvva,
                   const double* Lvvb, int Q, int oa, int va, int ob, int vb,
                   const double* eps_oa, const double* eps_va,
                   const double* eps_ob, const double* eps_vb, double* E);

#ifdef __cplusplus
}
#endif

#endif /* MSHQC_HIP_H */
