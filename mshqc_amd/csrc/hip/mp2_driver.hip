// MP2 drivers: pair/tile scheduling, deterministic reductions, the
// gradient contractions, and the RMP2/UMP2 entry points of the C ABI.
#include <algorithm>
#include <cmath>
#include <cstring>

#include "mp2_internal.h"
#include "mp2_bf16.h"

namespace {

constexpr int TILE = 64;

// s_pair[p] = sum_t s_part[p * ntiles + t]  (fixed order)
__global__ void pair_sum_kernel(const double* __restrict__ s_part, int npairs, int ntiles,
                                double* __restrict__ s_pair) {
    const int pidx = blockIdx.x * blockDim.x + threadIdx.x;
    if (pidx >= npairs) return;
    double acc = 0.0;
    for (int t = 0; t < ntiles; ++t) acc += s_part[(long)pidx * ntiles + t];
    s_pair[pidx] = acc;
}

// dE/deps_occ: each pair (i,j) adds -s to eps_i (left) and eps_j (right).
__global__ void occ_reduce_kernel(const int2* __restrict__ pairs, const double* __restrict__ s_pair,
                                  int npairs, double* deo_l, int ol, double* deo_r, int orr,
                                  int sym) {
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (sym) {
        if (k >= ol) return;
        double acc = 0.0;
        for (int q = 0; q < npairs; ++q) {
            const int2 ij = pairs[q];
            if (ij.x == k) acc += s_pair[q];
            if (ij.y == k) acc += s_pair[q];
        }
        deo_l[k] -= acc;
    } else {
        if (k < ol && deo_l) {
            double acc = 0.0;
            for (int q = 0; q < npairs; ++q)
                if (pairs[q].x == k) acc += s_pair[q];
            deo_l[k] -= acc;
        }
        if (k < orr && deo_r) {
            double acc = 0.0;
            for (int q = 0; q < npairs; ++q)
                if (pairs[q].y == k) acc += s_pair[q];
            deo_r[k] -= acc;
        }
    }
}

// dE/deps_vir: row sums feed the left virtual index, column sums the right.
// Only the tiles that touch the index's 64-block are visited: in the
// symmetric modes the tile list holds (A,B), A <= B, in row-major order; in
// the opposite-spin mode it holds all (A,B) with B fastest.  A workgroup
// covers 16 virtual indices; its 16 thread rows split the pairs into fixed
// chunks whose sums are combined in chunk order (deterministic).
constexpr int VR_IDX = 16, VR_CHUNKS = 16;
__global__ __launch_bounds__(VR_IDX * VR_CHUNKS) void vir_reduce_kernel(
    int ntA, int ntB, int npairs, const double* __restrict__ vrow, const double* __restrict__ vcol,
    double* dev_l, int vl, double* dev_r, int vr, int sym) {
    __shared__ double sl[VR_CHUNKS][VR_IDX], sr[VR_CHUNKS][VR_IDX];
    const int li = threadIdx.x % VR_IDX, c = threadIdx.x / VR_IDX;
    const int a = blockIdx.x * VR_IDX + li;
    const int blk = a / TILE, off = a % TILE;
    const int ntiles = sym ? ntA * (ntA + 1) / 2 : ntA * ntB;
    auto sym_index = [ntA](int A, int B) { return A * ntA - A * (A - 1) / 2 + (B - A); };
    const int per = (npairs + VR_CHUNKS - 1) / VR_CHUNKS;
    const int q0 = c * per, q1 = min(npairs, q0 + per);
    double accl = 0.0, accr = 0.0;
    if (sym) {
        if (a < vl)
            for (int q = q0; q < q1; ++q) {
                const long base = (long)q * ntiles;
                for (int B = blk; B < ntA; ++B) accl += vrow[(base + sym_index(blk, B)) * TILE + off];
                for (int A = 0; A <= blk; ++A) accl += vcol[(base + sym_index(A, blk)) * TILE + off];
            }
    } else {
        if (a < vl && dev_l)
            for (int q = q0; q < q1; ++q)
                for (int B = 0; B < ntB; ++B)
                    accl += vrow[((long)q * ntiles + (long)blk * ntB + B) * TILE + off];
        if (a < vr && dev_r)
            for (int q = q0; q < q1; ++q)
                for (int A = 0; A < ntA; ++A)
                    accr += vcol[((long)q * ntiles + (long)A * ntB + blk) * TILE + off];
    }
    sl[c][li] = accl;
    sr[c][li] = accr;
    __syncthreads();
    if (c != 0) return;
    double tl = 0.0, tr = 0.0;
    for (int k = 0; k < VR_CHUNKS; ++k) {
        tl += sl[k][li];
        tr += sr[k][li];
    }
    if (sym) {
        if (a < vl) dev_l[a] += tl;
    } else {
        if (a < vl && dev_l) dev_l[a] += tl;
        if (a < vr && dev_r) dev_r[a] += tr;
    }
}

struct Mp2Problem {
    int mode = MODE_RMP2;
    const double *Bl = nullptr, *Br = nullptr;
    int ol = 0, vl = 0, orr = 0, vr = 0, Q = 0;
    const double *eol = nullptr, *evl = nullptr, *eor = nullptr, *evr = nullptr;
    int precision = 0;
    double* e_slot = nullptr;  // device scalar, accumulated into
    bool grad = false;
    double *dBl = nullptr, *dBr = nullptr;
    double *deol = nullptr, *devl = nullptr, *deor = nullptr, *devr = nullptr;
    double* t2 = nullptr;      // MODE_T2 / MODE_T2OS output
};

inline int ceil_div(int a, int b) { return (a + b - 1) / b; }

// Runs the pair kernel over all (pair, tile) work of a problem, either in one
// launch (energy, t2) or per occupied block pair with the gradient GEMMs.
int run_pairs(mshqc_hip_ctx* ctx, const Mp2Problem& pr) {
    const bool sym = pair_mode_sym(pr.mode);
    const bool t2mode = pr.mode == MODE_T2 || pr.mode == MODE_T2OS;
    const int ntA = ceil_div(pr.vl, TILE), ntB = ceil_div(pr.vr, TILE);
    std::vector<int2> tiles;
    for (int A = 0; A < ntA; ++A)
        for (int B = sym ? A : 0; B < ntB; ++B) tiles.push_back(make_int2(A, B));
    const int ntiles = (int)tiles.size();
    // Work units of the split-BF16 kernel: symmetric modes (A, B2) = tiles
    // (A, 2B2), (A, 2B2+1) with A < 2B2, and for each k the diagonal block
    // (-1 - k, k) = tiles (2k .. 2k+1)^2 in place of units (2k, k) and
    // (2k+1, k); opposite spin (A2, B2) = 2 x 2 tiles.
    std::vector<int2> units;
    if (sym) {
        for (int A = 0; A < ntA; ++A)
            for (int B2 = 0; B2 < (ntA + 1) / 2; ++B2) {
                if (A == 2 * B2) units.push_back(make_int2(-1 - B2, B2));
                else if (A < 2 * B2) units.push_back(make_int2(A, B2));
            }
    } else {
        for (int A2 = 0; A2 < (ntA + 1) / 2; ++A2)
            for (int B2 = 0; B2 < (ntB + 1) / 2; ++B2) units.push_back(make_int2(A2, B2));
    }

    // Pair lists: one segment per launch.
    struct Seg {
        int off, n, I0, nI, J0, nJ;
    };
    std::vector<int2> pairs;
    std::vector<double> wts;
    std::vector<Seg> segs;
    int bo_l = pr.ol, bo_r = pr.orr;
    if (!pr.grad) {
        Seg sg{0, 0, 0, pr.ol, 0, pr.orr};
        for (int i = 0; i < pr.ol; ++i)
            for (int j = 0; j < pr.orr; ++j) {
                if (t2mode) {
                    pairs.push_back(make_int2(i, j));
                    wts.push_back(1.0);
                } else if (sym) {
                    if (j < i || (pr.mode == MODE_SS && j == i)) continue;
                    pairs.push_back(make_int2(i, j));
                    wts.push_back(j == i ? 1.0 : 2.0);
                } else {
                    pairs.push_back(make_int2(i, j));
                    wts.push_back(1.0);
                }
            }
        sg.n = (int)pairs.size();
        segs.push_back(sg);
    } else {
        const int vmax = std::max(pr.vl, pr.vr);
        const double side = std::sqrt((double)ctx->zblock_elems);
        const int bo = std::max(1, (int)std::floor(side / vmax));
        // as few blocks as a cap of bo orbitals allows, each only as large as
        // that count needs (o = 100, bo = 29: four blocks of 25, not 29 x 3 + 13)
        auto balanced = [bo](int o) {
            if (o <= 0) return 0;
            const int nb = ceil_div(o, std::min(bo, o));
            return ceil_div(o, nb);
        };
        bo_l = balanced(pr.ol);
        bo_r = balanced(pr.orr);
        if (sym) bo_r = bo_l;
        for (int I0 = 0; I0 < pr.ol; I0 += bo_l) {
            const int nI = std::min(bo_l, pr.ol - I0);
            for (int J0 = sym ? I0 : 0; J0 < pr.orr; J0 += bo_r) {
                const int nJ = std::min(bo_r, pr.orr - J0);
                Seg sg{(int)pairs.size(), 0, I0, nI, J0, nJ};
                // A diagonal block of a symmetric mode takes each unordered
                // pair once (weight 2 off the diagonal); the kernel writes
                // both its Z entries.
                for (int i = I0; i < I0 + nI; ++i)
                    for (int j = J0; j < J0 + nJ; ++j) {
                        if (sym && I0 == J0 && j < i) continue;
                        pairs.push_back(make_int2(i, j));
                        wts.push_back(sym && i != j ? 2.0 : 1.0);
                    }
                sg.n = (int)pairs.size() - sg.off;
                segs.push_back(sg);
            }
        }
    }
    const int npairs_all = (int)pairs.size();
    if (npairs_all == 0 || ntiles == 0) return MSHQC_HIP_OK;
    int maxseg = 0;
    for (auto& s : segs) maxseg = std::max(maxseg, s.n);

    // Upload tiles | units | pairs | weights, 256-byte aligned pieces.
    const size_t tb = sizeof(int2) * tiles.size(), ub = sizeof(int2) * units.size(),
                 pb = sizeof(int2) * pairs.size(), wb = sizeof(double) * wts.size();
    const size_t o_units = (tb + 255) / 256 * 256;
    const size_t o_pairs = o_units + (ub + 255) / 256 * 256, o_w = o_pairs + (pb + 255) / 256 * 256;
    const size_t o_spair = o_w + (wb + 255) / 256 * 256;
    const size_t misc_bytes = o_spair + sizeof(double) * (size_t)maxseg + 256;
    char* misc = static_cast<char*>(workspace(ctx, WS_MISC, misc_bytes));
    if (!misc) return MSHQC_HIP_ERR_ALLOC;
    std::vector<char> host(o_spair, 0);
    std::memcpy(host.data(), tiles.data(), tb);
    std::memcpy(host.data() + o_units, units.data(), ub);
    std::memcpy(host.data() + o_pairs, pairs.data(), pb);
    std::memcpy(host.data() + o_w, wts.data(), wb);
    HIP_TRY(ctx, hipMemcpyAsync(misc, host.data(), o_spair, hipMemcpyHostToDevice, ctx->stream));
    HIP_TRY(ctx, hipStreamSynchronize(ctx->stream));
    const int2* d_tiles = reinterpret_cast<const int2*>(misc);
    const int2* d_units = reinterpret_cast<const int2*>(misc + o_units);
    const int2* d_pairs = reinterpret_cast<const int2*>(misc + o_pairs);
    const double* d_w = reinterpret_cast<const double*>(misc + o_w);
    double* d_spair = reinterpret_cast<double*>(misc + o_spair);

    const long maxwork = (long)maxseg * ntiles;
    const size_t part_doubles = (size_t)maxwork * (pr.grad ? 2 : 1);
    double* part = static_cast<double*>(workspace(ctx, WS_PART, sizeof(double) * part_doubles));
    if (!part) return MSHQC_HIP_ERR_ALLOC;
    double* vpart = nullptr;
    double* Z = nullptr;
    long ldz = 0;
    if (pr.grad) {
        vpart = static_cast<double*>(
            workspace(ctx, WS_PART2, sizeof(double) * (size_t)maxwork * 2 * TILE));
        if (!vpart) return MSHQC_HIP_ERR_ALLOC;
        if (pr.dBl || pr.dBr) {
            ldz = (long)bo_r * pr.vr;
            Z = static_cast<double*>(
                workspace(ctx, WS_Z, sizeof(double) * (size_t)bo_l * pr.vl * ldz));
            if (!Z) return MSHQC_HIP_ERR_ALLOC;
        }
    }

    // Split-BF16 operand preparation (once per problem).
    Bf16Operands bops;
    if (pr.precision != 0) {
        int rc = bf16_prepare(ctx, pr.precision, pr.Bl, pr.ol, pr.vl,
                              sym ? nullptr : pr.Br, pr.orr, pr.vr, pr.Q,
                              pr.grad && (pr.dBl || pr.dBr), bo_l, bo_r,
                              pr.mode == MODE_RMP2 || pr.mode == MODE_OS, &bops);
        if (rc) return rc;
    }

    for (const Seg& sg : segs) {
        PairArgs a{};
        a.Bl = pr.Bl;
        a.Br = sym ? pr.Bl : pr.Br;
        a.ol = pr.ol;
        a.vl = pr.vl;
        a.orr = pr.orr;
        a.vr = pr.vr;
        a.Q = pr.Q;
        a.eol = pr.eol;
        a.evl = pr.evl;
        a.eor = pr.eor;
        a.evr = pr.evr;
        a.pairs = d_pairs + sg.off;
        a.pw = d_w + sg.off;
        a.npairs = sg.n;
        a.tiles = d_tiles;
        a.ntiles = ntiles;
        a.units = d_units;
        a.nunits = (int)units.size();
        a.ntA = ntA;
        a.ntB = ntB;
        a.nwork = (long)sg.n * ntiles;
        a.e_part = part;
        if (pr.precision != 0) {
            a.gself = bops.gself;
            a.nself_o = bops.nself_o;
            a.nself_v = bops.nself_v;
        }
        if (t2mode) {
            a.Z = pr.t2;
        } else if (pr.grad) {
            a.s_part = part + maxwork;
            a.vrow = vpart;
            a.vcol = vpart + (size_t)maxwork * TILE;
            a.Z = Z;
            a.ldz = ldz;
            a.zi0 = sg.I0;
            a.zj0 = sg.J0;
            a.ztrans = (sym && sg.I0 == sg.J0) ? 1 : 0;
        }
        int rc;
        {
            ProfScope ps(ctx, PROF_PAIR);
            rc = pr.precision == 0 ? launch_pair_f64(ctx, pr.mode, a)
                                   : bf16_launch_pair(ctx, pr.mode, a, bops);
        }
        if (rc) return rc;
        if (!t2mode) {
            rc = reduce_sum(ctx, part, a.nwork, pr.e_slot, true);
            if (rc) return rc;
        }
        if (!pr.grad) continue;
        if (pr.deol || pr.deor) {
            pair_sum_kernel<<<ceil_div(sg.n, 128), 128, 0, ctx->stream>>>(a.s_part, sg.n, ntiles,
                                                                            d_spair);
            LAUNCH_CHECK(ctx);
            const int nk = std::max(pr.ol, pr.orr);
            occ_reduce_kernel<<<ceil_div(nk, 128), 128, 0, ctx->stream>>>(
                a.pairs, d_spair, sg.n, pr.deol, pr.deol ? pr.ol : 0, pr.deor,
                pr.deor ? pr.orr : 0, sym ? 1 : 0);
            LAUNCH_CHECK(ctx);
        }
        if (pr.devl || pr.devr) {
            const int nv = std::max(pr.vl, pr.vr);
            vir_reduce_kernel<<<ceil_div(nv, VR_IDX), VR_IDX * VR_CHUNKS, 0, ctx->stream>>>(
                ntA, ntB, sg.n, a.vrow, a.vcol, pr.devl, pr.devl ? pr.vl : 0, pr.devr,
                pr.devr ? pr.vr : 0, sym ? 1 : 0);
            LAUNCH_CHECK(ctx);
        }
        if (!Z) continue;
        const bool second = !sym || sg.I0 != sg.J0;
        if (pr.precision != 0) {
            ProfScope ps(ctx, PROF_GEMM);
            rc = bf16_grad_contract(ctx, bops, Z, ldz, sg.I0, sg.nI, sg.J0, sg.nJ, pr.dBl,
                                    pr.dBr, second);
            if (rc) return rc;
            continue;
        }
        // dBl[q][(i,a) in I] += sum_{(j,b) in J} Br[q][(j,b)] Z[(i,a)][(j,b)]
        if (pr.dBl) {
            GemmF64 g;
            g.A = (sym ? pr.Bl : pr.Br) + (long)sg.J0 * pr.vr;
            g.lda = (long)pr.orr * pr.vr;
            g.transA = false;
            g.X = Z;
            g.ldx = ldz;
            g.transX = true;
            g.C = pr.dBl + (long)sg.I0 * pr.vl;
            g.ldc = (long)pr.ol * pr.vl;
            g.M = pr.Q;
            g.N = sg.nI * pr.vl;
            g.K = sg.nJ * pr.vr;
            g.accumulate = true;
            {
                ProfScope ps(ctx, PROF_GEMM);
                rc = gemm_f64(ctx, g);
            }
            if (rc) return rc;
        }
        // dBr[q][(j,b) in J] += sum_{(i,a) in I} Bl[q][(i,a)] Z[(i,a)][(j,b)]
        if (second && pr.dBr) {
            GemmF64 g;
            g.A = pr.Bl + (long)sg.I0 * pr.vl;
            g.lda = (long)pr.ol * pr.vl;
            g.transA = false;
            g.X = Z;
            g.ldx = ldz;
            g.transX = false;
            g.C = pr.dBr + (long)sg.J0 * pr.vr;
            g.ldc = (long)pr.orr * pr.vr;
            g.M = pr.Q;
            g.N = sg.nJ * pr.vr;
            g.K = sg.nI * pr.vl;
            g.accumulate = true;
            {
                ProfScope ps(ctx, PROF_GEMM);
                rc = gemm_f64(ctx, g);
            }
            if (rc) return rc;
        }
    }
    return MSHQC_HIP_OK;
}

int check_precision(mshqc_hip_ctx* ctx, int precision) {
    if (precision == MSHQC_PREC_FP64) return MSHQC_HIP_OK;
    if (precision >= 2 && precision <= BF16_MAX_SLICES) return MSHQC_HIP_OK;
    return set_error(ctx, MSHQC_HIP_ERR_ARG,
                     "precision must be 0 (FP64) or a split-BF16 slice count in [2, %d], got %d",
                     BF16_MAX_SLICES, precision);
}

int check_dims(mshqc_hip_ctx* ctx, const double* B, int Q, int o, int v, const double* eo,
               const double* ev, const char* who) {
    if (!B || !eo || !ev)
        return set_error(ctx, MSHQC_HIP_ERR_ARG, "%s: null input pointer", who);
    if (Q <= 0 || o <= 0 || v <= 0)
        return set_error(ctx, MSHQC_HIP_ERR_ARG, "%s: dimensions must be positive (Q=%d o=%d v=%d)",
                         who, Q, o, v);
    if ((long)o * v > (1L << 30) || (long)Q * o * v > (1L << 40))
        return set_error(ctx, MSHQC_HIP_ERR_ARG, "%s: dimensions too large", who);
    return MSHQC_HIP_OK;
}

// One spin's inputs of a UMP2 call: a spin may have no occupied or no
// virtual orbitals (the H atom has no beta electrons), and then contributes
// nothing; pointers of empty arrays may be null.
int check_spin(mshqc_hip_ctx* ctx, const double* B, int Q, int o, int v, const double* eo,
               const double* ev, const char* who) {
    if (Q <= 0 || o < 0 || v < 0)
        return set_error(ctx, MSHQC_HIP_ERR_ARG, "%s: need Q > 0 and o, v >= 0 (Q=%d o=%d v=%d)", who, Q, o, v);
    if ((o > 0 && v > 0 && !B) || (o > 0 && !eo) || (v > 0 && !ev))
        return set_error(ctx, MSHQC_HIP_ERR_ARG, "%s: null input pointer", who);
    if ((long)o * v > (1L << 30) || (long)Q * o * v > (1L << 40))
        return set_error(ctx, MSHQC_HIP_ERR_ARG, "%s: dimensions too large", who);
    return MSHQC_HIP_OK;
}

struct Timer {
    mshqc_hip_ctx* ctx;
    explicit Timer(mshqc_hip_ctx* c) : ctx(c) { hipEventRecord(ctx->ev0, ctx->stream); }
    int stop() {
        hipEventRecord(ctx->ev1, ctx->stream);
        int rc = ctx_finish(ctx);
        if (rc) return rc;
        float ms = 0.f;
        if (hipEventElapsedTime(&ms, ctx->ev0, ctx->ev1) == hipSuccess) ctx->last_ms = ms;
        hipGetLastError();
        return MSHQC_HIP_OK;
    }
};

double* energy_slots(mshqc_hip_ctx* ctx, int n) {
    // Kept apart from WS_MISC, which run_pairs may regrow.
    return static_cast<double*>(workspace(ctx, WS_TMP, sizeof(double) * (size_t)n));
}

}  // namespace

extern "C" {

int mshqc_hip_rmp2_energy(mshqc_hip_ctx* ctx, const double* B, int Q, int o, int v,
                          const double* eps_o, const double* eps_v, int precision, double* E) {
    int rc = ctx_enter(ctx);
    if (rc) return rc;
    if ((rc = check_dims(ctx, B, Q, o, v, eps_o, eps_v, "rmp2_energy"))) return rc;
    if ((rc = check_precision(ctx, precision))) return rc;
    CHECK_ARG(ctx, E != nullptr, "rmp2_energy: null E");
    Timer tm(ctx);
    double* slot = energy_slots(ctx, 1);
    if (!slot) return MSHQC_HIP_ERR_ALLOC;
    HIP_TRY(ctx, hipMemsetAsync(slot, 0, sizeof(double), ctx->stream));
    Mp2Problem pr;
    pr.mode = MODE_RMP2;
    pr.Bl = pr.Br = B;
    pr.ol = pr.orr = o;
    pr.vl = pr.vr = v;
    pr.Q = Q;
    pr.eol = pr.eor = eps_o;
    pr.evl = pr.evr = eps_v;
    pr.precision = precision;
    pr.e_slot = slot;
    if ((rc = run_pairs(ctx, pr))) return rc;
    if ((rc = tm.stop())) return rc;
    HIP_TRY(ctx, hipMemcpy(E, slot, sizeof(double), hipMemcpyDeviceToHost));
    return MSHQC_HIP_OK;
}

int mshqc_hip_rmp2_grad(mshqc_hip_ctx* ctx, const double* B, int Q, int o, int v,
                        const double* eps_o, const double* eps_v, int precision, double* E,
                        double* dEdB, double* dE_deps_o, double* dE_deps_v) {
    int rc = ctx_enter(ctx);
    if (rc) return rc;
    if ((rc = check_dims(ctx, B, Q, o, v, eps_o, eps_v, "rmp2_grad"))) return rc;
    if ((rc = check_precision(ctx, precision))) return rc;
    CHECK_ARG(ctx, E != nullptr, "rmp2_grad: null E");
    Timer tm(ctx);
    double* slot = energy_slots(ctx, 1);
    if (!slot) return MSHQC_HIP_ERR_ALLOC;
    HIP_TRY(ctx, hipMemsetAsync(slot, 0, sizeof(double), ctx->stream));
    if (dEdB) HIP_TRY(ctx, hipMemsetAsync(dEdB, 0, sizeof(double) * (size_t)Q * o * v, ctx->stream));
    if (dE_deps_o) HIP_TRY(ctx, hipMemsetAsync(dE_deps_o, 0, sizeof(double) * o, ctx->stream));
    if (dE_deps_v) HIP_TRY(ctx, hipMemsetAsync(dE_deps_v, 0, sizeof(double) * v, ctx->stream));
    Mp2Problem pr;
    pr.mode = MODE_RMP2;
    pr.Bl = pr.Br = B;
    pr.ol = pr.orr = o;
    pr.vl = pr.vr = v;
    pr.Q = Q;
    pr.eol = pr.eor = eps_o;
    pr.evl = pr.evr = eps_v;
    pr.precision = precision;
    pr.e_slot = slot;
    pr.grad = true;
    pr.dBl = pr.dBr = dEdB;
    pr.deol = pr.deor = dE_deps_o;
    pr.devl = pr.devr = dE_deps_v;
    if ((rc = run_pairs(ctx, pr))) return rc;
    if ((rc = tm.stop())) return rc;
    HIP_TRY(ctx, hipMemcpy(E, slot, sizeof(double), hipMemcpyDeviceToHost));
    return MSHQC_HIP_OK;
}

int mshqc_hip_ump2_t2(mshqc_hip_ctx* ctx, const double* Ba, const double* Bb, int Q, int oa, int va, int ob,
                      int vb, const double* eps_oa, const double* eps_va, const double* eps_ob,
                      const double* eps_vb, double* t2_aa, double* t2_bb, double* t2_ab) {
    int rc = ctx_enter(ctx);
    if (rc) return rc;
    if ((rc = check_spin(ctx, Ba, Q, oa, va, eps_oa, eps_va, "ump2_t2 (alpha)"))) return rc;
    if ((rc = check_spin(ctx, Bb, Q, ob, vb, eps_ob, eps_vb, "ump2_t2 (beta)"))) return rc;
    Timer tm(ctx);
    // (Bl, Br, ol, vl, orr, vr, eps ..., out) of the three blocks
    struct Blk {
        const double *Bl, *Br;
        int ol, vl, orr, vr;
        const double *eol, *evl, *eor, *evr;
        double* out;
        int mode;
    } blks[3] = {{Ba, Ba, oa, va, oa, va, eps_oa, eps_va, eps_oa, eps_va, t2_aa, MODE_T2},
                 {Bb, Bb, ob, vb, ob, vb, eps_ob, eps_vb, eps_ob, eps_vb, t2_bb, MODE_T2},
                 {Ba, Bb, oa, va, ob, vb, eps_oa, eps_va, eps_ob, eps_vb, t2_ab, MODE_T2OS}};
    for (const Blk& b : blks) {
        if (!b.out || b.ol == 0 || b.vl == 0 || b.orr == 0 || b.vr == 0) continue;
        Mp2Problem pr;
        pr.mode = b.mode;
        pr.Bl = b.Bl;
        pr.Br = b.Br;
        pr.ol = b.ol;
        pr.vl = b.vl;
        pr.orr = b.orr;
        pr.vr = b.vr;
        pr.Q = Q;
        pr.eol = b.eol;
        pr.evl = b.evl;
        pr.eor = b.eor;
        pr.evr = b.evr;
        pr.t2 = b.out;
        if ((rc = run_pairs(ctx, pr))) return rc;
    }
    return tm.stop();
}

int mshqc_hip_rmp2_t2(mshqc_hip_ctx* ctx, const double* B, int Q, int o, int v,
                      const double* eps_o, const double* eps_v, double* t2) {
    int rc = ctx_enter(ctx);
    if (rc) return rc;
    if ((rc = check_dims(ctx, B, Q, o, v, eps_o, eps_v, "rmp2_t2"))) return rc;
    CHECK_ARG(ctx, t2 != nullptr, "rmp2_t2: null t2");
    Timer tm(ctx);
    Mp2Problem pr;
    pr.mode = MODE_T2;
    pr.Bl = pr.Br = B;
    pr.ol = pr.orr = o;
    pr.vl = pr.vr = v;
    pr.Q = Q;
    pr.eol = pr.eor = eps_o;
    pr.evl = pr.evr = eps_v;
    pr.t2 = t2;
    if ((rc = run_pairs(ctx, pr))) return rc;
    return tm.stop();
}

static int ump2_common(mshqc_hip_ctx* ctx, const double* Ba, const double* Bb, int Q, int oa,
                       int va, int ob, int vb, const double* eoa, const double* eva,
                       const double* eob, const double* evb, int precision, double* E,
                       bool grad, double* dBa, double* dBb, double* deoa, double* deva,
                       double* deob, double* devb) {
    int rc = ctx_enter(ctx);
    if (rc) return rc;
    if ((rc = check_spin(ctx, Ba, Q, oa, va, eoa, eva, "ump2 (alpha)"))) return rc;
    if ((rc = check_spin(ctx, Bb, Q, ob, vb, eob, evb, "ump2 (beta)"))) return rc;
    if ((rc = check_precision(ctx, precision))) return rc;
    CHECK_ARG(ctx, E != nullptr, "ump2: null E");
    Timer tm(ctx);
    double* slot = energy_slots(ctx, 3);
    if (!slot) return MSHQC_HIP_ERR_ALLOC;
    HIP_TRY(ctx, hipMemsetAsync(slot, 0, 3 * sizeof(double), ctx->stream));
    if (grad) {
        // zero-length outputs of an empty spin are not touched
        const size_t na = (size_t)Q * oa * va, nb = (size_t)Q * ob * vb;
        if (dBa && na) HIP_TRY(ctx, hipMemsetAsync(dBa, 0, sizeof(double) * na, ctx->stream));
        if (dBb && nb) HIP_TRY(ctx, hipMemsetAsync(dBb, 0, sizeof(double) * nb, ctx->stream));
        if (deoa && oa) HIP_TRY(ctx, hipMemsetAsync(deoa, 0, sizeof(double) * oa, ctx->stream));
        if (deva && va) HIP_TRY(ctx, hipMemsetAsync(deva, 0, sizeof(double) * va, ctx->stream));
        if (deob && ob) HIP_TRY(ctx, hipMemsetAsync(deob, 0, sizeof(double) * ob, ctx->stream));
        if (devb && vb) HIP_TRY(ctx, hipMemsetAsync(devb, 0, sizeof(double) * vb, ctx->stream));
    }
    for (int spin = 0; spin < 2; ++spin) {
        Mp2Problem pr;
        pr.mode = MODE_SS;
        pr.Bl = pr.Br = spin ? Bb : Ba;
        pr.ol = pr.orr = spin ? ob : oa;
        pr.vl = pr.vr = spin ? vb : va;
        pr.Q = Q;
        pr.eol = pr.eor = spin ? eob : eoa;
        pr.evl = pr.evr = spin ? evb : eva;
        pr.precision = precision;
        pr.e_slot = slot + spin;
        pr.grad = grad;
        pr.dBl = pr.dBr = spin ? dBb : dBa;
        pr.deol = pr.deor = spin ? deob : deoa;
        pr.devl = pr.devr = spin ? devb : deva;
        // same-spin pairs need two occupied and one virtual orbital
        if (pr.ol >= 2 && pr.vl >= 1 && (rc = run_pairs(ctx, pr))) return rc;
    }
    Mp2Problem pr;
    pr.mode = MODE_OS;
    pr.Bl = Ba;
    pr.Br = Bb;
    pr.ol = oa;
    pr.vl = va;
    pr.orr = ob;
    pr.vr = vb;
    pr.Q = Q;
    pr.eol = eoa;
    pr.evl = eva;
    pr.eor = eob;
    pr.evr = evb;
    pr.precision = precision;
    pr.e_slot = slot + 2;
    pr.grad = grad;
    pr.dBl = dBa;
    pr.dBr = dBb;
    pr.deol = deoa;
    pr.devl = deva;
    pr.deor = deob;
    pr.devr = devb;
    if (oa > 0 && va > 0 && ob > 0 && vb > 0 && (rc = run_pairs(ctx, pr))) return rc;
    if ((rc = tm.stop())) return rc;
    HIP_TRY(ctx, hipMemcpy(E, slot, 3 * sizeof(double), hipMemcpyDeviceToHost));
    return MSHQC_HIP_OK;
}

int mshqc_hip_ump2_energy(mshqc_hip_ctx* ctx, const double* Ba, const double* Bb, int Q, int oa,
                          int va, int ob, int vb, const double* eps_oa, const double* eps_va,
                          const double* eps_ob, const double* eps_vb, int precision, double* E) {
    return ump2_common(ctx, Ba, Bb, Q, oa, va, ob, vb, eps_oa, eps_va, eps_ob, eps_vb, precision,
                       E, false, nullptr, nullptr, nullptr, nullptr, nullptr, nullptr);
}

int mshqc_hip_ump2_grad(mshqc_hip_ctx* ctx, const double* Ba, const double* Bb, int Q, int oa,
                        int va, int ob, int vb, const double* eps_oa, const double* eps_va,
                        const double* eps_ob, const double* eps_vb, int precision, double* E,
                        double* dEdBa, double* dEdBb, double* dE_deps_oa, double* dE_deps_va,
                        double* dE_deps_ob, double* dE_deps_vb) {
    return ump2_common(ctx, Ba, Bb, Q, oa, va, ob, vb, eps_oa, eps_va, eps_ob, eps_vb, precision,
                       E, true, dEdBa, dEdBb, dE_deps_oa, dE_deps_va, dE_deps_ob, dE_deps_vb);
}

}  // extern "C"
