// Cholesky-vector drop-in classes (host side).  SCF control flow, damping,
// DIIS placement and result packing follow the reference drivers cited per
// function; the Fock pieces and the MP2 pair work are C-ABI calls.
#include "cd.h"

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <iostream>
#include <mutex>
#include <string>

namespace mshqc {

// ================================================================= device

namespace device {

mshqc_hip_ctx* ctx() {
    static std::once_flag once;
    static mshqc_hip_ctx* c = nullptr;
    static int rc = 0;
    std::call_once(once, [] { rc = mshqc_hip_ctx_create(0, &c); });
    if (rc != 0 || !c)
        throw std::runtime_error("mshqc_hip: cannot create a device context (code " + std::to_string(rc) +
                                 "); the hot path needs an MI355X (gfx950) and has no CPU fallback");
    return c;
}

void check(int rc, const char* what) {
    if (rc != 0)
        throw std::runtime_error(std::string(what) + " failed (code " + std::to_string(rc) +
                                 "): " + mshqc_hip_last_error(ctx()));
}

Buffer::Buffer(size_t count) : n_(count) {
    void* p = nullptr;
    check(mshqc_hip_alloc(ctx(), count * sizeof(double), &p), "device alloc");
    p_ = static_cast<double*>(p);
}

Buffer::~Buffer() {
    if (p_) mshqc_hip_free(ctx(), p_);
}

void Buffer::upload(const double* host, size_t count) {
    check(mshqc_hip_memcpy_h2d(ctx(), p_, host, count * sizeof(double)), "upload");
}

void Buffer::download(double* host, size_t count) const {
    check(mshqc_hip_memcpy_d2h(ctx(), host, p_, count * sizeof(double)), "download");
}

}  // namespace device

namespace {

std::unique_ptr<device::Buffer> to_device(const Mat& M) {
    auto b = std::make_unique<device::Buffer>(M.size());
    b->upload(M.data(), M.size());
    return b;
}

std::unique_ptr<device::Buffer> to_device(const Vec& v) {
    auto b = std::make_unique<device::Buffer>(v.size());
    b->upload(v.data(), v.size());
    return b;
}

// X = S^{-1/2} by symmetric orthogonalisation.
Mat inverse_sqrt(const Mat& S) {
    Vec w;
    Mat U;
    eigh(S, w, U);
    const int n = S.rows;
    Mat T(n, n);
    for (int i = 0; i < n; ++i)
        for (int k = 0; k < n; ++k) T(i, k) = U(i, k) / std::sqrt(w[k]);
    return matmul_nt(T, U);
}

double seconds_since(std::chrono::high_resolution_clock::time_point t0) {
    return std::chrono::duration<double>(std::chrono::high_resolution_clock::now() - t0).count();
}

}  // namespace

// ============================================================ CholeskyERI

namespace integrals {

CholeskyERI::CholeskyERI(double threshold) : threshold_(threshold) {
    if (threshold_ <= 0.0) throw std::invalid_argument("Cholesky threshold must be positive");
}

void CholeskyERI::clear() {
    L_.clear();
    dev_.reset();
}

void CholeskyERI::add_vector(const Vec& v) {
    if (n_basis_ == 0) n_basis_ = (int)std::lround(std::sqrt((double)v.size()));
    if ((size_t)n_basis_ * n_basis_ != v.size())
        throw std::invalid_argument("add_vector: length must be n_basis^2");
    L_.push_back(v);
    dev_.reset();
}

void CholeskyERI::set_vectors(const std::vector<Vec>& vectors, int n_basis) {
    for (const auto& v : vectors)
        if (v.size() != (size_t)n_basis * n_basis)
            throw std::invalid_argument("set_vectors: every vector needs n_basis^2 entries");
    L_ = vectors;
    n_basis_ = n_basis;
    decomposed_ = true;
    dev_.reset();
}

double CholeskyERI::compression_ratio() const {
    if (L_.empty() || n_basis_ == 0) return 0.0;
    const double full = std::pow((double)n_basis_, 4);
    return full / ((double)L_.size() * n_basis_ * n_basis_);
}

CholeskyDecompositionResult CholeskyERI::decompose(const ERITensor& eri) {
    // cholesky_eri.cc:40-149 (driver), :155-167 (pivot), :169-203 (new vector)
    clear();
    n_basis_ = eri.n;
    const int n = eri.n, np = n * n;
    Vec D(np);
    for (int ij = 0; ij < np; ++ij) D[ij] = eri.v[(size_t)ij * np + ij];
    CholeskyDecompositionResult res;
    res.n_basis = n;
    res.threshold = threshold_;
    double dmax = np ? D[0] : 0.0;
    for (int k = 1; k < np; ++k) dmax = std::max(dmax, D[k]);
    if (np == 0 || dmax < threshold_) {
        decomposed_ = true;
        res.converged = true;
        return res;
    }
    int iter = 0;
    while (iter < np) {
        int p = 0;
        double mv = D[0];
        for (int k = 1; k < np; ++k)
            if (D[k] > mv) {
                mv = D[k];
                p = k;
            }
        if (mv < threshold_) break;
        Vec Ln(np);
        const int nv = (int)L_.size();
        Vec Lp(nv);
        for (int k = 0; k < nv; ++k) Lp[k] = L_[k][p];
#pragma omp parallel for if (np > 4096)
        for (int ij = 0; ij < np; ++ij) {
            double val = eri.v[(size_t)ij * np + p];
            for (int k = 0; k < nv; ++k) val -= L_[k][ij] * Lp[k];
            Ln[ij] = val;
        }
        const double nf = std::sqrt(D[p]);
        if (nf > 1e-12)
            for (auto& x : Ln) x /= nf;
        for (int ij = 0; ij < np; ++ij) {
            D[ij] -= Ln[ij] * Ln[ij];
            if (D[ij] < 0.0) D[ij] = 0.0;
        }
        L_.push_back(std::move(Ln));
        ++iter;
    }
    decomposed_ = true;
    res.n_vectors = (int)L_.size();
    res.compression_ratio = compression_ratio();
    res.converged = iter < np;
    return res;
}

CholeskyDecompositionResult CholeskyERI::decompose_gpu(const ERITensor& eri) {
    // cholesky_eri.cc:40-203 on the device
    clear();
    n_basis_ = eri.n;
    const int n = eri.n;
    const long np = (long)n * n;
    CholeskyDecompositionResult res;
    res.n_basis = n;
    res.threshold = threshold_;
    if (np == 0) {
        decomposed_ = true;
        res.converged = true;
        return res;
    }
    // the reference's bound is n^2 vectors; hold that many when the buffer is
    // at most the size of the ERI matrix itself, which is always the case
    const int cap = (int)np;
    auto dM = std::make_unique<device::Buffer>((size_t)np * np);
    dM->upload(eri.v.data(), (size_t)np * np);
    device::Buffer dL((size_t)cap * np);
    int nv = 0, conv = 0;
    double resid = 0.0;
    device::check(mshqc_hip_cholesky(device::ctx(), dM->get(), n, threshold_, cap, dL.get(), &nv, &conv, &resid),
                  "mshqc_hip_cholesky");
    dM.reset();
    Vec host((size_t)nv * np);
    if (nv > 0) dL.download(host.data(), host.size());
    L_.assign((size_t)nv, Vec());
    for (int q = 0; q < nv; ++q) L_[q].assign(host.begin() + (size_t)q * np, host.begin() + (size_t)(q + 1) * np);
    decomposed_ = true;
    res.n_vectors = nv;
    res.compression_ratio = compression_ratio();
    res.converged = conv != 0;
    return res;
}

double CholeskyERI::reconstruct(int i, int j, int a, int b) const {
    if (!decomposed_) throw std::runtime_error("CholeskyERI::reconstruct: not decomposed");
    const int ij = i * n_basis_ + j, ab = a * n_basis_ + b;
    double s = 0.0;
    for (const auto& L : L_) s += L[ij] * L[ab];
    return s;
}

std::shared_ptr<device::Buffer> CholeskyERI::device_buffer() const {
    device_vectors();
    return dev_;
}

const double* CholeskyERI::device_vectors() const {
    if (L_.empty()) throw std::runtime_error("CholeskyERI: no vectors (decompose or set them first)");
    if (!dev_) {
        const size_t nn = (size_t)n_basis_ * n_basis_;
        Vec packed(nn * L_.size());
        for (size_t q = 0; q < L_.size(); ++q) std::copy(L_[q].begin(), L_[q].end(), packed.begin() + q * nn);
        auto b = std::make_shared<device::Buffer>(packed.size());
        b->upload(packed.data(), packed.size());
        dev_ = b;
    }
    return dev_->get();
}

}  // namespace integrals

// ================================================================= CD-RHF

CholeskyRHF::CholeskyRHF(const Molecule& mol, const BasisSet& basis,
                         std::shared_ptr<IntegralEngine> integrals, const CholeskyRHFConfig& config)
    : mol_(mol), basis_(basis), integrals_(std::move(integrals)), config_(config),
      diis_(config.diis_max_vectors) {
    // cholesky_rhf.cc:17-39
    nbasis_ = basis.n_basis_functions();
    const int ne = mol.n_electrons();
    if (ne % 2 != 0)
        throw std::runtime_error("CholeskyRHF: Molecule must be closed-shell (even electrons).");
    n_occ_ = ne / 2;
    internal_ = std::make_unique<integrals::CholeskyERI>(config_.cholesky_threshold);
    internal_->decompose_gpu(integrals_->compute_eri());
    chol_ = internal_.get();
}

CholeskyRHF::CholeskyRHF(const Molecule& mol, const BasisSet& basis,
                         std::shared_ptr<IntegralEngine> integrals, const CholeskyRHFConfig& config,
                         const integrals::CholeskyERI& existing_cholesky)
    : mol_(mol), basis_(basis), integrals_(std::move(integrals)), config_(config),
      chol_(&existing_cholesky), diis_(config.diis_max_vectors) {
    // cholesky_rhf.cc:44-63
    nbasis_ = basis.n_basis_functions();
    const int ne = mol.n_electrons();
    if (ne % 2 != 0)
        throw std::runtime_error("CholeskyRHF: Molecule must be closed-shell (even electrons).");
    n_occ_ = ne / 2;
}

void CholeskyRHF::init_integrals() {
    // cholesky_rhf.cc:69-78
    S_ = integrals_->compute_overlap();
    H_ = add(integrals_->compute_kinetic(), integrals_->compute_nuclear());
    X_ = inverse_sqrt(S_);
}

void CholeskyRHF::initial_guess() {
    // cholesky_rhf.cc:80-88: core-Hamiltonian guess
    Mat Hp = matmul(matmul_tn(X_, H_), X_);
    Mat U;
    eigh(Hp, eps_, U);
    C_ = matmul(X_, U);
    P_ = build_density();
    energy_ = 0.0;
}

Mat CholeskyRHF::build_density() const {
    Mat Co = left_cols(C_, n_occ_);
    return scale(matmul_nt(Co, Co), 2.0);
}

void CholeskyRHF::build_fock() {
    // cholesky_rhf.cc:99-137: F = H + J - K/2 with J, K from the Cholesky
    // vectors (screen_exchange drops vectors with ||L||_F <= 1e-8 from K).
    if (chol_->n_basis() != nbasis_)
        throw std::runtime_error("CholeskyRHF: Cholesky vectors do not match the basis");
    const int n = nbasis_, Q = chol_->n_vectors();
    auto dP = to_device(P_);
    device::Buffer dJ((size_t)n * n), dK((size_t)n * n);
    device::check(mshqc_hip_jk(device::ctx(), chol_->device_vectors(), Q, n, dP->get(), 1,
                               config_.screen_exchange ? 1 : 0, dJ.get(), dK.get()),
                  "mshqc_hip_jk");
    Mat J(n, n), K(n, n);
    dJ.download(J.data(), J.size());
    dK.download(K.data(), K.size());
    F_ = add(add(H_, J), K, -0.5);
}

void CholeskyRHF::solve_fock() {
    Mat Fp = matmul(matmul_tn(X_, F_), X_);
    Mat U;
    eigh(Fp, eps_, U);
    C_ = matmul(X_, U);
}

SCFResult CholeskyRHF::compute() {
    // cholesky_rhf.cc:157-254
    init_integrals();
    initial_guess();
    const double e_nuc = mol_.nuclear_repulsion_energy();
    bool converged = false;
    int iter = 0;
    double energy_old = 0.0;
    energy_ = 0.0;
    Mat P_prev = P_;
    for (iter = 1; iter <= config_.max_iterations; ++iter) {
        energy_old = energy_;
        P_prev = P_;
        build_fock();
        // The reference's DIIS block can only run once two vectors are stored,
        // but vectors are only stored inside it, so it never runs
        // (cholesky_rhf.cc:188-196); kept for the same behaviour.
        if (iter > 1 && diis_.can_extrapolate()) {
            Mat err = add(matmul(matmul(F_, P_), S_), matmul(matmul(S_, P_), F_), -1.0);
            double nrm = std::sqrt(dot(err, err));
            if (nrm < config_.diis_threshold) {
                diis_.add_iteration(F_, err);
                if (diis_.can_extrapolate()) F_ = diis_.extrapolate();
            }
        }
        solve_fock();
        Mat P_new = build_density();
        const double alpha = 0.7;  // density damping (cholesky_rhf.cc:206-210)
        P_ = add(scale(P_new, alpha), P_prev, 1.0 - alpha);
        const double e_elec = 0.5 * trace_product(P_, add(H_, F_));
        energy_ = e_elec + e_nuc;
        const double de = iter > 1 ? energy_ - energy_old : energy_;
        if (config_.print_level > 0)
            std::printf("Iter %3d E(Tot) = %.10f dE = %.4e\n", iter, energy_, de);
        if (std::fabs(de) < config_.energy_threshold && iter > 1) {
            converged = true;
            break;
        }
    }
    SCFResult res;
    res.energy_total = energy_;
    res.energy_nuclear = e_nuc;
    res.energy_electronic = energy_ - e_nuc;
    res.converged = converged;
    res.iterations = iter;
    res.C_alpha = res.C_beta = C_;
    res.P_alpha = res.P_beta = scale(P_, 0.5);
    res.F_alpha = res.F_beta = F_;
    res.orbital_energies_alpha = res.orbital_energies_beta = eps_;
    res.n_occ_alpha = res.n_occ_beta = n_occ_;
    return res;
}

// ================================================================= CD-UHF

CholeskyUHF::CholeskyUHF(const Molecule& mol, const BasisSet& basis,
                         std::shared_ptr<IntegralEngine> integrals, int n_alpha, int n_beta,
                         const CholeskyUHFConfig& config)
    : mol_(mol), basis_(basis), integrals_(std::move(integrals)), config_(config), n_alpha_(n_alpha),
      n_beta_(n_beta), nbasis_(basis.n_basis_functions()), chol_(config.cholesky_threshold),
      diis_a_(config.diis_max_vectors), diis_b_(config.diis_max_vectors) {}

void CholeskyUHF::set_cholesky_vectors(const std::vector<Vec>& vectors) {
    // cholesky_uhf.cc:26-30
    chol_.set_vectors(vectors, nbasis_);
    external_ = true;
}

void CholeskyUHF::init_integrals_and_cholesky() {
    // cholesky_uhf.cc:32-63
    S_ = integrals_->compute_overlap();
    H_ = add(integrals_->compute_kinetic(), integrals_->compute_nuclear());
    X_ = inverse_sqrt(S_);
    if (!external_ && chol_.n_vectors() == 0) chol_.decompose_gpu(integrals_->compute_eri());
}

void CholeskyUHF::solve_fock(const Mat& F, Mat& C, Vec& eps) const {
    Mat Fp = matmul(matmul_tn(X_, F), X_);
    Mat U;
    eigh(Fp, eps, U);
    C = matmul(X_, U);
}

Mat CholeskyUHF::build_density(const Mat& C, int n_occ) const {
    if (n_occ == 0) return Mat(nbasis_, nbasis_);
    Mat Co = left_cols(C, n_occ);
    return matmul_nt(Co, Co);
}

void CholeskyUHF::initial_guess() {
    // cholesky_uhf.cc:68-75: core Hamiltonian with -0.5 on the diagonal
    Mat Hb = H_;
    for (int i = 0; i < nbasis_; ++i) Hb(i, i) -= 0.5;
    solve_fock(Hb, Ca_, eps_a_);
    solve_fock(Hb, Cb_, eps_b_);
    Pa_ = build_density(Ca_, n_alpha_);
    Pb_ = build_density(Cb_, n_beta_);
}

void CholeskyUHF::build_fock_cholesky() {
    // cholesky_uhf.cc:77-103: F_s = H + J[Pa + Pb] - K[P_s]
    const int n = nbasis_, Q = chol_.n_vectors();
    Mat P2(2 * n, n);
    std::copy(Pa_.a.begin(), Pa_.a.end(), P2.a.begin());
    std::copy(Pb_.a.begin(), Pb_.a.end(), P2.a.begin() + (size_t)n * n);
    auto dP = to_device(P2);
    device::Buffer dJ((size_t)n * n), dK((size_t)2 * n * n);
    device::check(mshqc_hip_jk(device::ctx(), chol_.device_vectors(), Q, n, dP->get(), 2, 0, dJ.get(),
                               dK.get()),
                  "mshqc_hip_jk");
    Mat J(n, n), K(2 * n, n);
    dJ.download(J.data(), J.size());
    dK.download(K.data(), K.size());
    Mat Ka(n, n), Kb(n, n);
    std::copy(K.a.begin(), K.a.begin() + (size_t)n * n, Ka.a.begin());
    std::copy(K.a.begin() + (size_t)n * n, K.a.end(), Kb.a.begin());
    Mat HJ = add(H_, J);
    Fa_ = add(HJ, Ka, -1.0);
    Fb_ = add(HJ, Kb, -1.0);
}

double CholeskyUHF::compute_energy() const {
    return 0.5 * (trace_product(Pa_, add(H_, Fa_)) + trace_product(Pb_, add(H_, Fb_)));
}

SCFResult CholeskyUHF::compute() {
    // cholesky_uhf.cc:121-159
    init_integrals_and_cholesky();
    initial_guess();
    int iter = 0;
    bool converged = false;
    for (iter = 1; iter <= config_.max_iterations; ++iter) {
        energy_old_ = energy_;
        build_fock_cholesky();
        Mat ea = add(matmul(matmul(Fa_, Pa_), S_), matmul(matmul(S_, Pa_), Fa_), -1.0);
        Mat eb = add(matmul(matmul(Fb_, Pb_), S_), matmul(matmul(S_, Pb_), Fb_), -1.0);
        diis_a_.add_iteration(Fa_, ea);
        diis_b_.add_iteration(Fb_, eb);
        if (iter > 2) {
            Fa_ = diis_a_.extrapolate();
            Fb_ = diis_b_.extrapolate();
        }
        solve_fock(Fa_, Ca_, eps_a_);
        solve_fock(Fb_, Cb_, eps_b_);
        Mat Pa_new = build_density(Ca_, n_alpha_);
        Mat Pb_new = build_density(Cb_, n_beta_);
        if (iter > 1) {
            Pa_ = add(scale(Pa_new, 0.7), Pa_, 0.3);
            Pb_ = add(scale(Pb_new, 0.7), Pb_, 0.3);
        } else {
            Pa_ = Pa_new;
            Pb_ = Pb_new;
        }
        energy_ = compute_energy();
        if (config_.print_level > 0) std::printf("UHF iter %3d E(elec) = %.10f\n", iter, energy_);
        if (std::fabs(energy_ - energy_old_) < config_.energy_threshold) {
            converged = true;
            break;
        }
    }
    SCFResult res;
    res.energy_nuclear = mol_.nuclear_repulsion_energy();
    res.energy_total = energy_ + res.energy_nuclear;
    res.energy_electronic = energy_;
    res.iterations = iter;
    res.converged = converged;
    res.n_occ_alpha = n_alpha_;
    res.n_occ_beta = n_beta_;
    res.C_alpha = Ca_;
    res.C_beta = Cb_;
    res.orbital_energies_alpha = eps_a_;
    res.orbital_energies_beta = eps_b_;
    res.P_alpha = Pa_;
    res.P_beta = Pb_;
    res.F_alpha = Fa_;
    res.F_beta = Fb_;
    return res;
}

double CholeskyUHF::compute_s_squared() const {
    // cholesky_uhf.cc:161-166
    const double sz = 0.5 * (n_alpha_ - n_beta_);
    Mat Sab = matmul(matmul_tn(left_cols(Ca_, n_alpha_), S_), left_cols(Cb_, n_beta_));
    return sz * (sz + 1.0) + n_beta_ - dot(Sab, Sab);
}

// ================================================================ CD-ROHF

CholeskyROHF::CholeskyROHF(const Molecule& mol, const BasisSet& basis,
                           std::shared_ptr<IntegralEngine> integrals, int n_alpha, int n_beta,
                           const CholeskyROHFConfig& config)
    : mol_(mol), basis_(basis), integrals_(std::move(integrals)), config_(config),
      nbasis_(basis.n_basis_functions()), n_alpha_(n_alpha), n_beta_(n_beta),
      chol_(config.cholesky_threshold) {
    // cholesky_rohf.cc:19-29: the decomposition waits for compute()
    if (n_alpha < 0 || n_beta < 0 || n_alpha > nbasis_ || n_beta > nbasis_)
        throw std::runtime_error("CholeskyROHF: occupation numbers out of range");
}

CholeskyROHF::CholeskyROHF(const Molecule& mol, const BasisSet& basis,
                           std::shared_ptr<IntegralEngine> integrals, int n_alpha, int n_beta,
                           const CholeskyROHFConfig& config, const integrals::CholeskyERI& existing_cholesky)
    : mol_(mol), basis_(basis), integrals_(std::move(integrals)), config_(config),
      nbasis_(basis.n_basis_functions()), n_alpha_(n_alpha), n_beta_(n_beta), chol_(existing_cholesky) {
    // cholesky_rohf.cc:34-51: a copy of the caller's vectors (the device
    // copy, once uploaded, is shared with it)
    if (n_alpha < 0 || n_beta < 0 || n_alpha > nbasis_ || n_beta > nbasis_)
        throw std::runtime_error("CholeskyROHF: occupation numbers out of range");
    if (config_.print_level > 0)
        std::printf("\n[CholeskyROHF] Initializing with REUSED vectors.\n  Vectors provided: %d\n",
                    chol_.n_vectors());
}

void CholeskyROHF::init_integrals() {
    // cholesky_rohf.cc:56-75
    S_ = integrals_->compute_overlap();
    H_ = add(integrals_->compute_kinetic(), integrals_->compute_nuclear());
    if (chol_.n_vectors() == 0) {
        if (config_.print_level > 0) std::printf("[CholeskyROHF] Decomposing Integrals...\n");
        chol_.decompose_gpu(integrals_->compute_eri());
    }
    if (chol_.n_basis() != nbasis_)
        throw std::runtime_error("CholeskyROHF: Cholesky vectors do not match the basis");
    X_ = inverse_sqrt(S_);
}

void CholeskyROHF::solve_fock(const Mat& F) {
    Mat Fp = matmul(matmul_tn(X_, F), X_);
    Mat U;
    eigh(Fp, eps_, U);
    C_ = matmul(X_, U);
}

Mat CholeskyROHF::build_density(int n_occ) const {
    if (n_occ == 0) return Mat(nbasis_, nbasis_);
    Mat Co = left_cols(C_, n_occ);
    return matmul_nt(Co, Co);
}

void CholeskyROHF::initial_guess() {
    // cholesky_rohf.cc:77-83: core-Hamiltonian guess; the reference solves it
    // once per spin, which gives the same orbitals twice
    solve_fock(H_);
    Pa_ = build_density(n_alpha_);
    Pb_ = build_density(n_beta_);
}

void CholeskyROHF::build_fock_cholesky() {
    // cholesky_rohf.cc:89-147: F_s = H + J[Pa + Pb] - K_s with the exchange
    // from the occupied columns of the shared orbitals
    const int n = nbasis_, Q = chol_.n_vectors();
    Fa_ = H_;
    Fb_ = H_;
    if (n_alpha_ == 0 && n_beta_ == 0) return;
    const int m = std::max(n_alpha_, n_beta_);
    auto dC = to_device(left_cols(C_, m));
    device::Buffer dJ((size_t)n * n), dKa((size_t)n * n), dKb((size_t)n * n);
    device::check(mshqc_hip_jk_rohf(device::ctx(), chol_.device_vectors(), Q, n, dC->get(), m, n_alpha_,
                                    n_beta_, dJ.get(), dKa.get(), dKb.get()),
                  "mshqc_hip_jk_rohf");
    Mat J(n, n), Ka(n, n), Kb(n, n);
    dJ.download(J.data(), J.size());
    dKa.download(Ka.data(), Ka.size());
    dKb.download(Kb.data(), Kb.size());
    Fa_ = add(add(H_, J), Ka, -1.0);
    Fb_ = add(add(H_, J), Kb, -1.0);
}

double CholeskyROHF::compute_energy() const {
    // cholesky_rohf.cc:159-165 (electronic energy)
    return 0.5 * dot(Pa_, add(H_, Fa_)) + 0.5 * dot(Pb_, add(H_, Fb_));
}

SCFResult CholeskyROHF::compute() {
    // cholesky_rohf.cc:184-276
    if (config_.print_level > 0)
        std::printf("\n%s\n  Cholesky-ROHF SCF Calculation\n%s\n", std::string(60, '=').c_str(),
                    std::string(60, '=').c_str());
    init_integrals();
    initial_guess();
    DIIS diis_a(config_.diis_max_vectors), diis_b(config_.diis_max_vectors);
    const double e_nuc = mol_.nuclear_repulsion_energy();
    const double na = n_alpha_, nb = n_beta_;
    int iter = 0;
    bool converged = false;
    while (iter < config_.max_iterations) {
        ++iter;
        energy_old_ = energy_;
        build_fock_cholesky();
        energy_ = compute_energy();
        Mat err_a = add(matmul(matmul(Fa_, Pa_), S_), matmul(matmul(S_, Pa_), Fa_), -1.0);
        Mat err_b = add(matmul(matmul(Fb_, Pb_), S_), matmul(matmul(S_, Pb_), Fb_), -1.0);
        if (std::sqrt(dot(err_a, err_a)) < config_.diis_threshold) {
            diis_a.add_iteration(Fa_, err_a);
            if (diis_a.can_extrapolate()) Fa_ = diis_a.extrapolate();
        }
        if (std::sqrt(dot(err_b, err_b)) < config_.diis_threshold) {
            diis_b.add_iteration(Fb_, err_b);
            if (diis_b.can_extrapolate()) Fb_ = diis_b.extrapolate();
        }
        // occupation-weighted average (cholesky_rohf.cc:228-232)
        Mat F_eff = scale(add(scale(Fa_, na), Fb_, nb), 1.0 / (na + nb));
        solve_fock(F_eff);
        Pa_ = build_density(n_alpha_);
        Pb_ = build_density(n_beta_);
        const double de = energy_ - energy_old_;
        if (config_.print_level > 0)
            std::printf("Iter %3d E = %.10f dE = %.4e\n", iter, energy_ + e_nuc, de);
        if (std::fabs(de) < config_.energy_threshold && iter > 1) {
            converged = true;
            break;
        }
    }
    SCFResult res;
    res.energy_nuclear = e_nuc;
    res.energy_electronic = energy_;
    res.energy_total = energy_ + e_nuc;
    res.C_alpha = res.C_beta = C_;
    res.P_alpha = Pa_;
    res.P_beta = Pb_;
    res.orbital_energies_alpha = res.orbital_energies_beta = eps_;
    res.converged = converged;
    res.iterations = iter;
    res.n_occ_alpha = n_alpha_;
    res.n_occ_beta = n_beta_;
    return res;
}

// ================================================================ CD-RMP2

namespace {

// B [Q][o][v] on the device from the AO vectors and an SCF coefficient block;
// precision k >= 2 runs the transform on split-BF16 GEMMs with min(k, 4)
// slices (the MP2 classes' split-BF16 form), 0 in FP64.
std::unique_ptr<device::Buffer> transform_device(const integrals::CholeskyERI& chol, const Mat& C, int o,
                                                 int precision = MSHQC_PREC_FP64) {
    const int n = C.rows, v = C.cols - o, Q = chol.n_vectors();
    if (chol.n_basis() != n) throw std::runtime_error("Cholesky vectors do not match the MO basis");
    auto dCo = to_device(left_cols(C, o));
    auto dCv = to_device(right_cols(C, v));
    auto B = std::make_unique<device::Buffer>((size_t)Q * o * v);
    device::check(mshqc_hip_transform(device::ctx(), chol.device_vectors(), Q, n, dCo->get(), o, dCv->get(),
                                      v, precision > 4 ? 4 : precision, B->get()),
                  "mshqc_hip_transform");
    return B;
}

}  // namespace

CholeskyRMP2::CholeskyRMP2(const Molecule&, const BasisSet& basis, std::shared_ptr<IntegralEngine> integrals,
                           const SCFResult& rhf_result, const CholeskyRMP2Config& config)
    : rhf_(rhf_result), config_(config) {
    // cholesky_rmp2.cc:19-41
    nbf_ = basis.n_basis_functions();
    nocc_ = rhf_.n_occ_alpha;
    nvirt_ = nbf_ - nocc_;
    internal_ = std::make_unique<integrals::CholeskyERI>(config_.cholesky_threshold);
    internal_->decompose_gpu(integrals->compute_eri());
    chol_ = internal_.get();
}

CholeskyRMP2::CholeskyRMP2(const Molecule&, const BasisSet& basis, std::shared_ptr<IntegralEngine>,
                           const SCFResult& rhf_result, const CholeskyRMP2Config& config,
                           const integrals::CholeskyERI& existing_cholesky)
    : rhf_(rhf_result), config_(config), chol_(&existing_cholesky) {
    // cholesky_rmp2.cc:46-63
    nbf_ = basis.n_basis_functions();
    nocc_ = rhf_.n_occ_alpha;
    nvirt_ = nbf_ - nocc_;
}

foundation::CholeskyRMP2Result CholeskyRMP2::compute() {
    // cholesky_rmp2.cc:97-179
    if (nocc_ <= 0 || nvirt_ <= 0) throw std::runtime_error("CholeskyRMP2: need occupied and virtual orbitals");
    const int o = nocc_, v = nvirt_, Q = chol_->n_vectors();
    auto t0 = std::chrono::high_resolution_clock::now();
    auto B = transform_device(*chol_, rhf_.C_alpha, o, config_.precision);
    foundation::CholeskyRMP2Result r;
    r.time_transform_s = seconds_since(t0);
    const Vec& eps = rhf_.orbital_energies_alpha;
    auto deo = to_device(Vec(eps.begin(), eps.begin() + o));
    auto dev = to_device(Vec(eps.begin() + o, eps.begin() + o + v));
    t0 = std::chrono::high_resolution_clock::now();
    double E = 0.0;
    device::check(mshqc_hip_rmp2_energy(device::ctx(), B->get(), Q, o, v, deo->get(), dev->get(),
                                        config_.precision, &E),
                  "mshqc_hip_rmp2_energy");
    r.time_energy_s = seconds_since(t0);
    const size_t nt2 = (size_t)o * o * v * v;
    const bool store = config_.store_t2 == 1 || (config_.store_t2 == -1 && nt2 <= (size_t(1) << 27));
    if (store) {
        device::Buffer dt2(nt2);
        device::check(mshqc_hip_rmp2_t2(device::ctx(), B->get(), Q, o, v, deo->get(), dev->get(), dt2.get()),
                      "mshqc_hip_rmp2_t2");
        r.t2.resize(nt2);
        dt2.download(r.t2.data(), nt2);
    }
    r.e_rhf = rhf_.energy_total;
    r.e_corr = E;
    r.e_total = r.e_rhf + E;
    r.n_occ = o;
    r.n_virt = v;
    r.n_chol_vectors = Q;
    r.chol_device = chol_->device_buffer();
    r.n_basis = nbf_;
    if (config_.print_level > 0)
        std::printf("  RHF Energy:    %16.10f Ha\n  Cholesky RMP2: %16.10f Ha\n  Total Energy:  %16.10f Ha\n",
                    r.e_rhf, r.e_corr, r.e_total);
    return r;
}

// ================================================================ CD-RMP3

namespace foundation {

CholeskyRMP3::CholeskyRMP3(const SCFResult& rhf_result, const CholeskyRMP2Result& crmp2_result,
                           const BasisSet& basis)
    : rhf_(rhf_result), crmp2_(crmp2_result) {
    // cholesky_rmp3.cc:18-29
    nbf_ = basis.n_basis_functions();
    nocc_ = crmp2_.n_occ;
    nvirt_ = crmp2_.n_virt;
    n_chol_vectors_ = crmp2_.n_chol_vectors;
}

RMP3Result CholeskyRMP3::compute() {
    // cholesky_rmp3.cc:79-200: the three MO blocks of the vectors, t2(2) and
    // E(3) on the device
    if (!crmp2_.chol_device || crmp2_.n_basis != nbf_)
        throw std::runtime_error("CholeskyRMP3: the CholeskyRMP2Result carries no Cholesky vectors for this basis");
    const int n = nbf_, o = nocc_, v = nvirt_, Q = n_chol_vectors_;
    const double* dL = crmp2_.chol_device->get();
    const Mat& C = rhf_.C_alpha;
    auto dCo = to_device(left_cols(C, o));
    auto dCv = to_device(right_cols(C, v));
    const Vec& eps = rhf_.orbital_energies_alpha;
    auto deo = to_device(Vec(eps.begin(), eps.begin() + o));
    auto dev = to_device(Vec(eps.begin() + o, eps.begin() + o + v));
    device::Buffer B((size_t)Q * o * v), Loo((size_t)Q * o * o), Lvv((size_t)Q * v * v);
    device::check(mshqc_hip_transform(device::ctx(), dL, Q, n, dCo->get(), o, dCv->get(), v, MSHQC_PREC_FP64, B.get()),
                  "mshqc_hip_transform");
    device::check(mshqc_hip_transform(device::ctx(), dL, Q, n, dCo->get(), o, dCo->get(), o, MSHQC_PREC_FP64, Loo.get()),
                  "mshqc_hip_transform");
    device::check(mshqc_hip_transform(device::ctx(), dL, Q, n, dCv->get(), v, dCv->get(), v, MSHQC_PREC_FP64, Lvv.get()),
                  "mshqc_hip_transform");
    // first-order amplitudes: the RMP2 result's when it kept them
    const size_t nt2 = (size_t)o * o * v * v;
    device::Buffer t1(nt2);
    if (crmp2_.t2.size() == nt2)
        t1.upload(crmp2_.t2.data(), nt2);
    else
        device::check(mshqc_hip_rmp2_t2(device::ctx(), B.get(), Q, o, v, deo->get(), dev->get(), t1.get()),
                      "mshqc_hip_rmp2_t2");
    const bool keep = nt2 <= (size_t(1) << 27);
    std::unique_ptr<device::Buffer> t22;
    if (keep) t22 = std::make_unique<device::Buffer>(nt2);
    double e3 = 0.0;
    device::check(mshqc_hip_rmp3(device::ctx(), B.get(), Loo.get(), Lvv.get(), Q, o, v, t1.get(), deo->get(),
                                 dev->get(), &e3, keep ? t22->get() : nullptr),
                  "mshqc_hip_rmp3");
    RMP3Result r;
    r.e_rhf = crmp2_.e_rhf;
    r.e_mp2 = crmp2_.e_corr;
    r.e_mp3 = e3;
    r.e_corr_total = r.e_mp2 + e3;
    r.e_total = r.e_rhf + r.e_corr_total;
    r.n_occ = o;
    r.n_virt = v;
    r.t2_1 = crmp2_.t2;
    if (keep) {
        r.t2_2.resize(nt2);
        t22->download(r.t2_2.data(), nt2);
    }
    std::printf("[INFO] MP2 Correlation : %.8f Ha\n[INFO] MP3 Correction  : %.8f Ha\n"
                "[INFO] Total Correlation: %.8f Ha\n[INFO] Total Energy    : %.8f Ha\n",
                r.e_mp2, r.e_mp3, r.e_corr_total, r.e_total);
    return r;
}

}  // namespace foundation

// ================================================================ CD-UMP2

CholeskyUMP2::CholeskyUMP2(const SCFResult& uhf_result, const BasisSet& basis,
                           std::shared_ptr<IntegralEngine> integrals, const CholeskyUMP2Config& config,
                           const integrals::CholeskyERI& cholesky_from_uhf)
    : uhf_(uhf_result), basis_(basis), integrals_(std::move(integrals)), config_(config),
      chol_(cholesky_from_uhf) {
    // cholesky_ump2.cc:37-62
    nbf_ = basis.n_basis_functions();
    na_ = uhf_.n_occ_alpha;
    nb_ = uhf_.n_occ_beta;
    va_ = nbf_ - na_;
    vb_ = nbf_ - nb_;
}

CholeskyUMP2Result CholeskyUMP2::compute() {
    // cholesky_ump2.cc:68-174
    CholeskyUMP2Result r;
    auto t0 = std::chrono::high_resolution_clock::now();
    if (chol_.n_vectors() == 0) {
        chol_.decompose_gpu(integrals_->compute_eri());
        r.time_cholesky_s = seconds_since(t0);
    }
    const int Q = chol_.n_vectors();
    r.n_cholesky_vectors = Q;
    r.compression_ratio = chol_.compression_ratio();
    t0 = std::chrono::high_resolution_clock::now();
    // a spin without occupied or virtual orbitals has no B (and contributes
    // 0, as the reference's loops over its empty ranges do)
    std::unique_ptr<device::Buffer> Ba, Bb;
    if (na_ > 0 && va_ > 0) Ba = transform_device(chol_, uhf_.C_alpha, na_, config_.precision);
    if (nb_ > 0 && vb_ > 0) Bb = transform_device(chol_, uhf_.C_beta, nb_, config_.precision);
    r.time_transform_s = seconds_since(t0);
    const Vec &ea = uhf_.orbital_energies_alpha, &eb = uhf_.orbital_energies_beta;
    auto deoa = to_device(Vec(ea.begin(), ea.begin() + na_));
    auto deva = to_device(Vec(ea.begin() + na_, ea.end()));
    auto deob = to_device(Vec(eb.begin(), eb.begin() + nb_));
    auto devb = to_device(Vec(eb.begin() + nb_, eb.end()));
    t0 = std::chrono::high_resolution_clock::now();
    double E[3] = {0, 0, 0};
    device::check(mshqc_hip_ump2_energy(device::ctx(), Ba ? Ba->get() : nullptr, Bb ? Bb->get() : nullptr, Q,
                                        na_, va_, nb_, vb_, deoa->get(),
                                        deva->get(), deob->get(), devb->get(), config_.precision, E),
                  "mshqc_hip_ump2_energy");
    r.time_energy_s = seconds_since(t0);
    r.e_corr_ss_aa = E[0];
    r.e_corr_ss_bb = E[1];
    r.e_corr_os = E[2];
    r.e_corr_total = E[0] + E[1] + E[2];
    r.e_total = uhf_.energy_total + r.e_corr_total;
    r.memory_mb = (double)Q * ((double)na_ * va_ + (double)nb_ * vb_) * 8.0 / (1024.0 * 1024.0);
    if (config_.print_level > 0)
        std::printf("SS(aa) %.10f  SS(bb) %.10f  OS %.10f  corr %.10f  total %.10f\n", r.e_corr_ss_aa,
                    r.e_corr_ss_bb, r.e_corr_os, r.e_corr_total, r.e_total);
    return r;
}

void CholeskyUMP2::compute_t2_amplitudes() {
    // cholesky_ump2.cc:471-556 (the reference reuses the vectors its compute()
    // transformed; here they are transformed again, on the device)
    if (config_.print_level > 0) std::printf("\nComputing T2 amplitudes (memory-intensive)...\n");
    if (chol_.n_vectors() == 0) chol_.decompose_gpu(integrals_->compute_eri());
    const int Q = chol_.n_vectors();
    std::unique_ptr<device::Buffer> Ba, Bb;
    if (na_ > 0 && va_ > 0) Ba = transform_device(chol_, uhf_.C_alpha, na_);
    if (nb_ > 0 && vb_ > 0) Bb = transform_device(chol_, uhf_.C_beta, nb_);
    const Vec &ea = uhf_.orbital_energies_alpha, &eb = uhf_.orbital_energies_beta;
    auto deoa = to_device(Vec(ea.begin(), ea.begin() + na_));
    auto deva = to_device(Vec(ea.begin() + na_, ea.end()));
    auto deob = to_device(Vec(eb.begin(), eb.begin() + nb_));
    auto devb = to_device(Vec(eb.begin() + nb_, eb.end()));
    const size_t naa = (size_t)na_ * na_ * va_ * va_, nbb = (size_t)nb_ * nb_ * vb_ * vb_,
                 nab = (size_t)na_ * nb_ * va_ * vb_;
    device::Buffer taa(naa), tbb(nbb), tab(nab);
    device::check(mshqc_hip_ump2_t2(device::ctx(), Ba ? Ba->get() : nullptr, Bb ? Bb->get() : nullptr, Q, na_, va_,
                                    nb_, vb_, deoa->get(), deva->get(), deob->get(), devb->get(),
                                    naa ? taa.get() : nullptr, nbb ? tbb.get() : nullptr,
                                    nab ? tab.get() : nullptr),
                  "mshqc_hip_ump2_t2");
    // blocks with an empty index range stay zero, as the reference's setZero()
    t2_aa_.assign(naa, 0.0);
    t2_bb_.assign(nbb, 0.0);
    t2_ab_.assign(nab, 0.0);
    if (naa && Ba) taa.download(t2_aa_.data(), naa);
    if (nbb && Bb) tbb.download(t2_bb_.data(), nbb);
    if (nab && Ba && Bb) tab.download(t2_ab_.data(), nab);
    if (config_.print_level > 0) std::printf("T2 amplitudes computed and stored.\n");
}

// ================================================================ CD-UMP3

CholeskyUMP3::CholeskyUMP3(const SCFResult& uhf_result, const BasisSet& basis,
                           std::shared_ptr<IntegralEngine> integrals, const CholeskyUMP3Config& config)
    : uhf_(uhf_result), basis_(basis), integrals_(std::move(integrals)), config_(config),
      chol_(config.cholesky_threshold) {
    // cholesky_ump3.cc:26-39
    nbf_ = basis.n_basis_functions();
    na_ = uhf_.n_occ_alpha;
    nb_ = uhf_.n_occ_beta;
    va_ = nbf_ - na_;
    vb_ = nbf_ - nb_;
}

CholeskyUMP3::CholeskyUMP3(const CholeskyUMP2& ump2_solver, const CholeskyUMP3Config& config)
    : uhf_(ump2_solver.uhf_), basis_(ump2_solver.basis_), integrals_(ump2_solver.integrals_), config_(config),
      chol_(ump2_solver.get_cholesky()) {
    // cholesky_ump3.cc:41-56: reuses the UMP2 solver's vectors
    nbf_ = ump2_solver.nbf_;
    na_ = ump2_solver.na_;
    nb_ = ump2_solver.nb_;
    va_ = ump2_solver.va_;
    vb_ = ump2_solver.vb_;
}

// Device MO blocks of both spins (cholesky_ump3.cc:119-150): B = C_occ^T L
// C_vir, Loo, Lvv and the orbital energies, for the vector count they were
// made from.
struct CholeskyUMP3::MoBlocks {
    struct Spin {
        std::unique_ptr<device::Buffer> B, Loo, Lvv, eo, ev;
    } s[2];
    int Q = 0;
};

void CholeskyUMP3::initialize_cholesky() {
    // cholesky_ump3.cc:108-117
    if (chol_.n_vectors() == 0) chol_.decompose_gpu(integrals_->compute_eri());
    transform_cholesky_vectors();
}

void CholeskyUMP3::transform_cholesky_vectors() {
    // cholesky_ump3.cc:119-150
    if (chol_.n_vectors() == 0) throw std::runtime_error("CholeskyUMP3: no Cholesky vectors to transform");
    if (chol_.n_basis() != nbf_) throw std::runtime_error("CholeskyUMP3: Cholesky vectors do not match the basis");
    const int n = nbf_, Q = chol_.n_vectors();
    const double* dL = chol_.device_vectors();
    auto blk = std::make_shared<MoBlocks>();
    blk->Q = Q;
    for (int s = 0; s < 2; ++s) {
        const Mat& C = s == 0 ? uhf_.C_alpha : uhf_.C_beta;
        const Vec& e = s == 0 ? uhf_.orbital_energies_alpha : uhf_.orbital_energies_beta;
        const int o = s == 0 ? na_ : nb_, v = s == 0 ? va_ : vb_;
        if (o == 0 || v == 0) continue;
        auto dCo = to_device(left_cols(C, o));
        auto dCv = to_device(right_cols(C, v));
        MoBlocks::Spin& b = blk->s[s];
        b.B = std::make_unique<device::Buffer>((size_t)Q * o * v);
        b.Loo = std::make_unique<device::Buffer>((size_t)Q * o * o);
        b.Lvv = std::make_unique<device::Buffer>((size_t)Q * v * v);
        device::check(mshqc_hip_transform(device::ctx(), dL, Q, n, dCo->get(), o, dCv->get(), v, MSHQC_PREC_FP64, b.B->get()),
                      "mshqc_hip_transform");
        device::check(mshqc_hip_transform(device::ctx(), dL, Q, n, dCo->get(), o, dCo->get(), o, MSHQC_PREC_FP64, b.Loo->get()),
                      "mshqc_hip_transform");
        device::check(mshqc_hip_transform(device::ctx(), dL, Q, n, dCv->get(), v, dCv->get(), v, MSHQC_PREC_FP64, b.Lvv->get()),
                      "mshqc_hip_transform");
        b.eo = to_device(Vec(e.begin(), e.begin() + o));
        b.ev = to_device(Vec(e.begin() + o, e.begin() + o + v));
    }
    blocks_ = std::move(blk);
}

CholeskyUMP3Result CholeskyUMP3::compute() {
    // cholesky_ump3.cc:160-205
    CholeskyUMP3Result r;
    auto t_start = std::chrono::high_resolution_clock::now();
    if (config_.print_level > 0)
        std::printf("\n========================================\n  Cholesky UMP3 (Block-Contraction Optimized)\n"
                    "========================================\n");
    if (chol_.n_vectors() == 0) chol_.decompose_gpu(integrals_->compute_eri());
    if (!blocks_ || blocks_->Q != chol_.n_vectors()) transform_cholesky_vectors();
    const int Q = chol_.n_vectors();
    r.n_cholesky_vectors = Q;
    r.compression_ratio = chol_.compression_ratio();
    const auto& blk = blocks_->s;
    auto ptr = [](const std::unique_ptr<device::Buffer>& x) -> const double* { return x ? x->get() : nullptr; };
    auto t1 = std::chrono::high_resolution_clock::now();
    double E[3] = {0, 0, 0};
    device::check(mshqc_hip_ump3(device::ctx(), ptr(blk[0].B), ptr(blk[1].B), ptr(blk[0].Loo), ptr(blk[1].Loo),
                                 ptr(blk[0].Lvv), ptr(blk[1].Lvv), Q, na_, va_, nb_, vb_, ptr(blk[0].eo),
                                 ptr(blk[0].ev), ptr(blk[1].eo), ptr(blk[1].ev), E),
                  "mshqc_hip_ump3");
    r.time_mp3_s = seconds_since(t1);
    r.e_mp3_ss_aa = E[0];
    r.e_mp3_ss_bb = E[1];
    r.e_mp3_os = E[2];
    r.e_mp3_total = r.e_mp3_ss_aa + r.e_mp3_ss_bb + r.e_mp3_os;
    r.e_corr_total = r.e_mp2_total + r.e_mp3_total;
    r.time_total_s = seconds_since(t_start);
    if (config_.print_level > 0)
        std::printf("MP3 Corr (aa): %.8f Ha\nMP3 Corr (bb): %.8f Ha\nMP3 Corr (ab): %.8f Ha\nTotal MP3 Corr: %.8f Ha\n",
                    r.e_mp3_ss_aa, r.e_mp3_ss_bb, r.e_mp3_os, r.e_mp3_total);
    return r;
}

// ================================================================ CD-OMP2

CholeskyOMP2::CholeskyOMP2(const Molecule& mol, const BasisSet& basis, std::shared_ptr<IntegralEngine> integrals,
                           const SCFResult& scf_guess, const CholeskyOMP2Config& config)
    : mol_(mol), basis_(basis), integrals_(std::move(integrals)), scf_(scf_guess), config_(config) {
    // cholesky_omp2.cc:18-41: decompose with the configured threshold
    nbf_ = scf_.C_alpha.rows;
    na_ = scf_.n_occ_alpha;
    nb_ = scf_.n_occ_beta;
    va_ = nbf_ - na_;
    vb_ = nbf_ - nb_;
    if (config_.print_level > 0)
        std::printf("[CholeskyOMP2] Decomposing integrals (Thresh=%g)...\n", config_.cholesky_threshold);
    internal_ = std::make_unique<integrals::CholeskyERI>(config_.cholesky_threshold);
    internal_->decompose_gpu(integrals_->compute_eri());
    chol_ = internal_.get();
}

CholeskyOMP2::CholeskyOMP2(const Molecule& mol, const BasisSet& basis, std::shared_ptr<IntegralEngine> integrals,
                           const SCFResult& scf_guess, const CholeskyOMP2Config& config,
                           const integrals::CholeskyERI& cholesky_vectors)
    : mol_(mol), basis_(basis), integrals_(std::move(integrals)), scf_(scf_guess), config_(config),
      chol_(&cholesky_vectors) {
    // cholesky_omp2.cc:46-66; the reference exits the process on empty
    // vectors, this throws
    nbf_ = scf_.C_alpha.rows;
    na_ = scf_.n_occ_alpha;
    nb_ = scf_.n_occ_beta;
    va_ = nbf_ - na_;
    vb_ = nbf_ - nb_;
    if (chol_->n_vectors() == 0)
        throw std::runtime_error("CholeskyOMP2 initialized with empty reused vectors");
}

void CholeskyOMP2::two_particle_terms(const double* Ba, const double* Bb, const double* eoa, const double* eva,
                                      const double* eob, const double* evb, bool same_orbitals,
                                      double* const (&Co)[2], double* const (&Cv)[2],
                                      std::unique_ptr<device::Buffer> (&w2)[2]) const {
    // dE/dB per spin in FP64 (the energy is the UMP2 expression with the
    // SCF denominators), then w2 = sum_Q Lvv G^T - G^T Loo per spin, all on
    // the device
    const int n = nbf_, Q = chol_->n_vectors();
    std::unique_ptr<device::Buffer> dBa, dBb;
    if (na_ * va_ > 0) dBa = std::make_unique<device::Buffer>((size_t)Q * na_ * va_);
    if (nb_ * vb_ > 0) dBb = std::make_unique<device::Buffer>((size_t)Q * nb_ * vb_);
    double E[3];
    device::check(mshqc_hip_ump2_grad(device::ctx(), Ba, Bb, Q, na_, va_, nb_, vb_, eoa, eva,
                                      eob, evb, MSHQC_PREC_FP64, E, dBa ? dBa->get() : nullptr,
                                      dBb ? dBb->get() : nullptr, nullptr, nullptr, nullptr, nullptr),
                  "mshqc_hip_ump2_grad");
    std::unique_ptr<device::Buffer> Loo, Lvv;
    for (int s = 0; s < 2; ++s) {
        const int no = s == 0 ? na_ : nb_, nv = s == 0 ? va_ : vb_;
        w2[s].reset();
        if (no * nv == 0) continue;
        if (s == 0 || !same_orbitals) {
            Loo = std::make_unique<device::Buffer>((size_t)Q * no * no);
            Lvv = std::make_unique<device::Buffer>((size_t)Q * nv * nv);
            device::check(mshqc_hip_transform(device::ctx(), chol_->device_vectors(), Q, n, Co[s], no, Co[s], no,
                                              MSHQC_PREC_FP64, Loo->get()),
                          "mshqc_hip_transform");
            device::check(mshqc_hip_transform(device::ctx(), chol_->device_vectors(), Q, n, Cv[s], nv, Cv[s], nv,
                                              MSHQC_PREC_FP64, Lvv->get()),
                          "mshqc_hip_transform");
        }
        w2[s] = std::make_unique<device::Buffer>((size_t)nv * no);
        device::check(mshqc_hip_omp2_grad2p(device::ctx(), (s == 0 ? dBa : dBb)->get(), Loo->get(), Lvv->get(), Q,
                                            no, nv, w2[s]->get()),
                      "mshqc_hip_omp2_grad2p");
    }
}

OMP2Result CholeskyOMP2::compute() {
    // cholesky_omp2.cc:427-513, the iteration on the device: amplitudes, OPDM,
    // generalised Fock, gradient and rotation; the host keeps the energies
    // and the gradient norm for the convergence test
    const int n = nbf_, Q = chol_->n_vectors();
    if (chol_->n_basis() != n) throw std::runtime_error("CholeskyOMP2: Cholesky vectors do not match the basis");
    if (config_.print_level > 0)
        std::printf("\n========================================\n     Cholesky Orbital-Optimized MP2\n"
                    "     Vectors: %d\n========================================\n",
                    Q);
    const Mat H = add(integrals_->compute_kinetic(), integrals_->compute_nuclear());
    const Vec &ea = scf_.orbital_energies_alpha, &eb = scf_.orbital_energies_beta;
    // the denominators keep the SCF orbital energies on every iteration
    auto deoa = to_device(Vec(ea.begin(), ea.begin() + na_));
    auto deva = to_device(Vec(ea.begin() + na_, ea.end()));
    auto deob = to_device(Vec(eb.begin(), eb.begin() + nb_));
    auto devb = to_device(Vec(eb.begin() + nb_, eb.end()));
    device::Buffer dGoa((size_t)std::max(na_ * na_, 1)), dGva((size_t)std::max(va_ * va_, 1)),
        dGob((size_t)std::max(nb_ * nb_, 1)), dGvb((size_t)std::max(vb_ * vb_, 1));
    auto dH = to_device(H);
    // per spin: C_occ [n][o], C_vir [n][v], the density (the SCF result's on
    // the first iteration, :451-459) and the MO-basis generalised Fock
    std::unique_ptr<device::Buffer> oCo[2], oCv[2], oP[2], oF[2], w2[2];
    double *Co[2] = {nullptr, nullptr}, *Cv[2] = {nullptr, nullptr}, *P[2], *F[2];
    for (int s = 0; s < 2; ++s) {
        const int no = s == 0 ? na_ : nb_, nv = s == 0 ? va_ : vb_;
        const Mat& C = s == 0 ? scf_.C_alpha : scf_.C_beta;
        if (no > 0) Co[s] = (oCo[s] = to_device(left_cols(C, no)))->get();
        if (nv > 0) Cv[s] = (oCv[s] = to_device(right_cols(C, nv)))->get();
        P[s] = (oP[s] = to_device(s == 0 ? scf_.P_alpha : scf_.P_beta))->get();
        F[s] = (oF[s] = std::make_unique<device::Buffer>((size_t)n * n))->get();
    }
    // beta follows alpha once the alpha-only rotation of na == nb has run
    // (:484-490), or from the start when the SCF orbitals are the same
    bool same_orbitals = na_ == nb_ && scf_.C_beta.a == scf_.C_alpha.a;
    double e_total_prev = scf_.energy_total, e_mp2 = 0.0, e_ss = 0.0, e_os = 0.0;
    bool converged = false;
    int iter = 0;
    while (iter < config_.max_iterations) {
        ++iter;
        // amplitudes, their energy and the OPDM blocks (:94-326)
        std::unique_ptr<device::Buffer> Ba, Bb;
        if (na_ > 0 && va_ > 0) {
            Ba = std::make_unique<device::Buffer>((size_t)Q * na_ * va_);
            device::check(mshqc_hip_transform(device::ctx(), chol_->device_vectors(), Q, n, Co[0], na_, Cv[0], va_,
                                              MSHQC_PREC_FP64, Ba->get()),
                          "mshqc_hip_transform");
        }
        if (nb_ > 0 && vb_ > 0 && !same_orbitals) {
            Bb = std::make_unique<device::Buffer>((size_t)Q * nb_ * vb_);
            device::check(mshqc_hip_transform(device::ctx(), chol_->device_vectors(), Q, n, Co[1], nb_, Cv[1], vb_,
                                              MSHQC_PREC_FP64, Bb->get()),
                          "mshqc_hip_transform");
        }
        const double* pBb = same_orbitals ? (Ba ? Ba->get() : nullptr) : (Bb ? Bb->get() : nullptr);
        double E[3] = {0, 0, 0};
        device::check(mshqc_hip_omp2_opdm(device::ctx(), Ba ? Ba->get() : nullptr, pBb, Q, na_, va_, nb_, vb_,
                                          deoa->get(), deva->get(), deob->get(), devb->get(), E, dGoa.get(),
                                          dGva.get(), dGob.get(), dGvb.get()),
                      "mshqc_hip_omp2_opdm");
        e_ss = E[0] + E[1];
        e_os = E[2];
        e_mp2 = (E[0] + E[1]) + E[2];
        const double e_total_curr = scf_.energy_total + e_mp2;
        // P + C G C^T, then the generalised Fock H + J[Pt_a + Pt_b] - K[Pt_s]
        // in the MO basis per spin (:331-369, :451-467)
        device::check(mshqc_hip_omp2_fock(device::ctx(), chol_->device_vectors(), Q, n, dH->get(), Co[0], Cv[0],
                                          na_, Co[1], Cv[1], nb_, P[0], P[1], dGoa.get(), dGva.get(), dGob.get(),
                                          dGvb.get(), F[0], F[1]),
                      "mshqc_hip_omp2_fock");
        if (config_.two_particle_gradient)
            two_particle_terms(Ba ? Ba->get() : nullptr, pBb, deoa->get(), deva->get(), deob->get(), devb->get(),
                               same_orbitals, Co, Cv, w2);
        double g2 = 0.0;
        for (int s = 0; s < 2; ++s) {
            double gs = 0.0;
            device::check(mshqc_hip_omp2_grad_norm(device::ctx(), n, s == 0 ? na_ : nb_, F[s],
                                                   w2[s] ? w2[s]->get() : nullptr, &gs),
                          "mshqc_hip_omp2_grad_norm");
            g2 += gs;
        }
        const double g_norm = std::sqrt(g2);
        if (config_.print_level > 0)
            std::printf("%4d    %.8f    %.8f    %.2e\n", iter, e_total_curr, e_mp2, g_norm);
        const double de = std::fabs(e_total_curr - e_total_prev);
        if (g_norm < config_.gradient_threshold && de < config_.energy_threshold) {
            converged = true;
            break;
        }
        // damped, level-shifted rotation C <- C exp(-kappa) (:375-425)
        device::check(mshqc_hip_omp2_rotate(device::ctx(), n, na_, F[0], w2[0] ? w2[0]->get() : nullptr, Co[0],
                                            Cv[0], P[0]),
                      "mshqc_hip_omp2_rotate");
        if (na_ != nb_) {
            device::check(mshqc_hip_omp2_rotate(device::ctx(), n, nb_, F[1], w2[1] ? w2[1]->get() : nullptr, Co[1],
                                                Cv[1], P[1]),
                          "mshqc_hip_omp2_rotate");
        } else {
            Co[1] = Co[0];
            Cv[1] = Cv[0];
            P[1] = P[0];
            same_orbitals = true;
        }
        e_total_prev = e_total_curr;
    }
    // the final orbitals back on the host (columns: occupied, then virtual)
    for (int s = 0; s < 2; ++s) {
        const int no = s == 0 ? na_ : nb_, nv = s == 0 ? va_ : vb_;
        Mat& C = s == 0 ? scf_.C_alpha : scf_.C_beta;
        Mat Mo(n, no), Mv(n, nv);
        if (no > 0) device::check(mshqc_hip_memcpy_d2h(device::ctx(), Mo.data(), Co[s], sizeof(double) * Mo.size()),
                                  "download");
        if (nv > 0) device::check(mshqc_hip_memcpy_d2h(device::ctx(), Mv.data(), Cv[s], sizeof(double) * Mv.size()),
                                  "download");
        for (int mu = 0; mu < n; ++mu) {
            for (int i = 0; i < no; ++i) C(mu, i) = Mo(mu, i);
            for (int a = 0; a < nv; ++a) C(mu, no + a) = Mv(mu, a);
        }
        device::check(mshqc_hip_memcpy_d2h(device::ctx(), (s == 0 ? scf_.P_alpha : scf_.P_beta).data(), P[s],
                                           sizeof(double) * (size_t)n * n),
                      "download");
    }
    OMP2Result res;
    res.energy_scf = scf_.energy_total;
    res.energy_total = e_total_prev;  // the reference keeps the previous iteration's total on exit
    res.energy_mp2_corr = e_mp2;
    res.energy_mp2_ss = e_ss;
    res.energy_mp2_os = e_os;
    res.converged = converged;
    res.iterations = iter;
    res.n_occ_alpha = na_;
    res.n_occ_beta = nb_;
    res.n_virt_alpha = va_;
    res.n_virt_beta = vb_;
    res.C_alpha = scf_.C_alpha;
    res.C_beta = scf_.C_beta;
    res.orbital_energies_alpha = scf_.orbital_energies_alpha;
    res.orbital_energies_beta = scf_.orbital_energies_beta;
    return res;
}

}  // namespace mshqc
