// Pair-block (ia|jb) kernel shared by the FP64 and split-BF16 MP2 paths:
// mode definitions, argument block and the fused denominator epilogue.
//
// A workgroup owns one occupied pair (i,j) and one 64 x 64 virtual tile (A,B).
// It forms
//     G[a][b] = sum_Q Bl^Q_ia Br^Q_jb         (a in A, b in B)
// and, in the symmetric modes, the mirrored block
//     H[a][b] = sum_Q Br^Q_ja Bl^Q_ib = G[b][a]
// so the epilogue sees (ia|jb) and (ib|ja) in the same lane.  Only tiles with
// A <= B are launched in the symmetric modes; a tile with A < B accounts for
// both (a,b) and (b,a).  Nothing of G leaves registers.
#pragma once
#include <hip/hip_runtime.h>

enum PairMode {
    MODE_RMP2 = 0,  // E += G(2G - H)/D     (cholesky_rmp2.cc:134-151)
    MODE_SS = 1,    // E += 1/4 (G - H)^2/D (cholesky_ump2.cc:315-350)
    MODE_OS = 2,    // E += G^2/D           (cholesky_ump2.cc:434-457)
    MODE_T2 = 3,    // writes t2[i][j][a][b] = G/D (cholesky_rmp2.cc:145-148)
    MODE_T2OS = 4,  // the same for two operands, all tiles (A, B), no mirror:
                    // t2_ab of CholeskyUMP2::compute_t2_amplitudes
                    // (cholesky_ump2.cc:533-552)
};

// Symmetric modes take the tiles A <= B with the mirrored H block; MODE_OS
// and MODE_T2OS take every tile (A, B) and no mirror.
__host__ __device__ constexpr bool pair_mode_sym(int mode) { return mode != MODE_OS && mode != MODE_T2OS; }

struct PairArgs {
    const double* Bl;  // [Q][ol][vl]
    const double* Br;  // [Q][or][vr]
    int ol, vl, orr, vr, Q;
    const double *eol, *evl, *eor, *evr;  // orbital energies (occupied/virtual, left/right)
    const int2* pairs;                    // (i, j) per pair slot
    const double* pw;                     // weight per pair slot
    int npairs;
    const int2* tiles;                    // (A, B) tile indices
    int ntiles;
    long nwork;                           // npairs * ntiles
    // Work units of the split-BF16 kernel (groups of 64 x 64 tiles; the
    // partials above stay per tile).  ntA, ntB: 64-wide tiles along vl, vr.
    const int2* units;
    int nunits;
    int ntA, ntB;
    double* e_part;                       // [nwork] weighted energy partials
    // Gradient outputs (null for energy only).
    double* Z;       // amplitude block, row (i - zi0) * vl + a, col (j - zj0) * vr + b
    long ldz;
    int zi0, zj0;
    int ztrans;      // diagonal block of a symmetric mode: pairs i < j also
                     // write the (j, i) entries, z^{ji}_{ba} = z^{ij}_{ab}
    double* s_part;  // [nwork]      weighted sum of m over the tile
    double* vrow;    // [nwork][64]  weighted row sums of m (left virtual index)
    double* vcol;    // [nwork][64]  weighted column sums of m (right virtual index)
    // Self terms (split-BF16 only; null in FP64): gself[i * nself_v + a] =
    // sum_Q Bl^Q_ia Br^Q_ia in FP64 for i < nself_o, a < nself_v, used for the
    // elements (ia|ia) of the pairs i = j.  With Bl = Br these sums are
    // positive definite, so their truncated slice products err in one
    // direction (the dropped p = q products are squares) instead of
    // averaging out as every other element's do.
    const double* gself;
    int nself_o, nself_v;
};

// Per-element epilogue.  Given G = (ia|jb), H = (ib|ja) and D, returns the
// energy term `e`, the dE/dD weight `m` (dE/dD = -m summed over both index
// pairs of the element) and the gradient amplitude `z`
// (dE/dB^Q_ia = sum_jb z^{ij}_ab B^Q_jb over ordered pairs), for the direct
// element (a,b); the mirrored element (b,a) is the same call with G and H
// swapped.
template <int MODE>
__device__ __forceinline__ void pair_terms(double G, double H, double D, double& e, double& m,
                                           double& z) {
    // No FMA contraction: the FP64 and split-BF16 kernels then round the
    // epilogue identically, and so does the CPU oracle.
#pragma clang fp contract(off)
    if (fabs(D) < 1e-12) {  // the reference's skip branch
        e = m = z = 0.0;
        return;
    }
    if (MODE == MODE_RMP2 || MODE == MODE_T2) {
        const double T = (2.0 * G - H) / D;  // 2t_ab - t_ba
        e = G * T;
        m = (G / D) * T;
        z = 4.0 * T;
    } else if (MODE == MODE_SS) {
        const double u = (G - H) / D;
        e = 0.25 * (G - H) * u;
        m = 0.25 * u * u;
        z = 2.0 * u;
    } else {  // MODE_OS, MODE_T2OS
        const double t = G / D;
        e = G * t;
        m = t * t;
        z = 2.0 * t;
    }
}

// Linear work index with an XCD-aware remap: the dispatcher hands workgroup
// ids to the 8 XCDs round-robin, so consecutive work items (the tiles of one
// occupied pair, which share Bl_i and Br_j) are placed on the same XCD and
// its L2.  gridDim.x is a multiple of 8.
__device__ __forceinline__ long xcd_work_index() {
    const long per = gridDim.x >> 3;
    return (long)(blockIdx.x & 7) * per + (blockIdx.x >> 3);
}
