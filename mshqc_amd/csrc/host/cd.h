// Cholesky-vector drop-in classes: CholeskyERI, CholeskyRHF, CholeskyUHF,
// CholeskyROHF, CholeskyRMP2, CholeskyUMP2, CholeskyOMP2.  Public names, constructor arguments and
// defaults follow syahrulhidayat/mshqc; the J/K build, the AO->MO transforms
// and the (ia|jb)/denominator reductions run through libmshqc_hip.so.
#pragma once

#include <chrono>
#include <memory>

#include "core.h"
#include "mshqc_hip.h"

namespace mshqc {

// ----------------------------------------------------------------- device

namespace device {
// Process-wide context on device 0 (created on first use).
mshqc_hip_ctx* ctx();
// Throws std::runtime_error with the library's message when rc != 0.
void check(int rc, const char* what);

// Owning device buffer of doubles.
class Buffer {
public:
    Buffer() = default;
    explicit Buffer(size_t count);
    ~Buffer();
    Buffer(const Buffer&) = delete;
    Buffer& operator=(const Buffer&) = delete;
    double* get() const { return p_; }
    size_t count() const { return n_; }
    void upload(const double* host, size_t count);
    void download(double* host, size_t count) const;

private:
    double* p_ = nullptr;
    size_t n_ = 0;
};
}  // namespace device

// ----------------------------------------------------------- CholeskyERI

struct CholeskyDecompositionResult {
    int n_vectors = 0;
    int n_basis = 0;
    double threshold = 0.0;
    double max_error = 0.0;
    double rms_error = 0.0;
    double compression_ratio = 0.0;
    bool converged = false;
};

namespace integrals {

// Pivoted incomplete Cholesky of the ERI matrix in FP64 on the host
// (cholesky_eri.cc:40-203 semantics: diagonal over all n^2 ordered pairs
// i*n+j, first maximum as pivot, stop when it falls below the threshold,
// negative updated diagonals clamped to zero).  The vectors live on the host
// and are uploaded to the device once, on first use by a Fock or MP2 call.
class CholeskyERI {
public:
    explicit CholeskyERI(double threshold = 1e-6);
    CholeskyDecompositionResult decompose(const ERITensor& eri);
    // The same decomposition on the device (mshqc_hip_cholesky): identical
    // vectors and pivots; the ERI matrix is uploaded once and the vectors come
    // back to the host.
    CholeskyDecompositionResult decompose_gpu(const ERITensor& eri);
    double reconstruct(int i, int j, int a, int b) const;
    int n_vectors() const { return (int)L_.size(); }
    int n_basis() const { return n_basis_; }
    double threshold() const { return threshold_; }
    bool is_decomposed() const { return decomposed_; }
    double compression_ratio() const;
    size_t storage_bytes() const { return L_.size() * (size_t)n_basis_ * n_basis_ * sizeof(double); }
    const std::vector<Vec>& get_L_vectors() const { return L_; }
    void clear();
    // Appends one vector of n*n values (index i*n+j) and counts it.
    void add_vector(const Vec& v);
    // Replaces all vectors (each n*n, index i*n+j).
    void set_vectors(const std::vector<Vec>& vectors, int n_basis);
    // Device copy [Q][n][n] (uploaded on first call).
    const double* device_vectors() const;
    // The same device copy as a shared handle, for results that outlive this object.
    std::shared_ptr<device::Buffer> device_buffer() const;

private:
    std::vector<Vec> L_;
    int n_basis_ = 0;
    double threshold_ = 1e-6;
    bool decomposed_ = false;
    mutable std::shared_ptr<device::Buffer> dev_;
};

}  // namespace integrals

// ------------------------------------------------------------- CD-RHF

struct CholeskyRHFConfig : public SCFConfig {
    double cholesky_threshold = 1e-6;
    bool screen_exchange = true;
    CholeskyRHFConfig() { print_level = 1; }
};

class CholeskyRHF {
public:
    CholeskyRHF(const Molecule& mol, const BasisSet& basis, std::shared_ptr<IntegralEngine> integrals,
                const CholeskyRHFConfig& config);
    CholeskyRHF(const Molecule& mol, const BasisSet& basis, std::shared_ptr<IntegralEngine> integrals,
                const CholeskyRHFConfig& config, const integrals::CholeskyERI& existing_cholesky);
    SCFResult compute();
    double energy() const { return energy_; }

private:
    Molecule mol_;
    BasisSet basis_;
    std::shared_ptr<IntegralEngine> integrals_;
    CholeskyRHFConfig config_;
    std::unique_ptr<integrals::CholeskyERI> internal_;
    const integrals::CholeskyERI* chol_ = nullptr;
    DIIS diis_;
    int nbasis_ = 0, n_occ_ = 0;
    Mat S_, H_, X_, F_, P_, C_;
    Vec eps_;
    double energy_ = 0.0;

    void init_integrals();
    void initial_guess();
    Mat build_density() const;
    void build_fock();
    void solve_fock();
};

// ------------------------------------------------------------- CD-UHF

struct CholeskyUHFConfig {
    double energy_threshold = 1e-8;
    double cholesky_threshold = 1e-6;
    int max_iterations = 100;
    int diis_max_vectors = 6;
    int print_level = 1;
};

class CholeskyUHF {
public:
    CholeskyUHF(const Molecule& mol, const BasisSet& basis, std::shared_ptr<IntegralEngine> integrals,
                int n_alpha, int n_beta, const CholeskyUHFConfig& config);
    SCFResult compute();
    void set_cholesky_vectors(const std::vector<Vec>& vectors);
    double compute_s_squared() const;

private:
    Molecule mol_;
    BasisSet basis_;
    std::shared_ptr<IntegralEngine> integrals_;
    CholeskyUHFConfig config_;
    int n_alpha_, n_beta_, nbasis_;
    integrals::CholeskyERI chol_;
    bool external_ = false;
    Mat S_, H_, X_, Fa_, Fb_, Pa_, Pb_, Ca_, Cb_;
    Vec eps_a_, eps_b_;
    double energy_ = 0.0, energy_old_ = 0.0;
    DIIS diis_a_, diis_b_;

    void init_integrals_and_cholesky();
    void initial_guess();
    void build_fock_cholesky();
    void solve_fock(const Mat& F, Mat& C, Vec& eps) const;
    Mat build_density(const Mat& C, int n_occ) const;
    double compute_energy() const;
};

// ------------------------------------------------------------ CD-ROHF

struct CholeskyROHFConfig : public SCFConfig {
    double cholesky_threshold = 1e-6;
    bool screen_exchange = true;  // accepted for parity; the reference's ROHF Fock build ignores it
};

// cholesky_rohf.h / cholesky_rohf.cc: alpha and beta Fock matrices from one
// shared orbital set, orbitals from the occupation-weighted average
// (n_a F_a + n_b F_b) / (n_a + n_b), per-spin DIIS below diis_threshold.
class CholeskyROHF {
public:
    CholeskyROHF(const Molecule& mol, const BasisSet& basis, std::shared_ptr<IntegralEngine> integrals,
                 int n_alpha, int n_beta, const CholeskyROHFConfig& config = CholeskyROHFConfig());
    CholeskyROHF(const Molecule& mol, const BasisSet& basis, std::shared_ptr<IntegralEngine> integrals,
                 int n_alpha, int n_beta, const CholeskyROHFConfig& config,
                 const integrals::CholeskyERI& existing_cholesky);
    SCFResult compute();

private:
    Molecule mol_;
    BasisSet basis_;
    std::shared_ptr<IntegralEngine> integrals_;
    CholeskyROHFConfig config_;
    int nbasis_, n_alpha_, n_beta_;
    integrals::CholeskyERI chol_;
    Mat S_, H_, X_, C_, Pa_, Pb_, Fa_, Fb_;
    Vec eps_;
    double energy_ = 0.0, energy_old_ = 0.0;

    void init_integrals();
    void initial_guess();
    void build_fock_cholesky();
    void solve_fock(const Mat& F);
    Mat build_density(int n_occ) const;
    double compute_energy() const;
};

// ------------------------------------------------------------- CD-RMP2

namespace foundation {
struct CholeskyRMP2Result {
    double e_rhf = 0.0, e_corr = 0.0, e_total = 0.0;
    // t2[i][j][a][b] (flattened, o*o*v*v) when stored; empty otherwise.
    Vec t2;
    // The AO Cholesky vectors for consumers such as CholeskyRMP3: the
    // reference copies them to the host here (cholesky_rmp2.cc:167); this
    // result shares the device copy instead.
    std::shared_ptr<device::Buffer> chol_device;
    int n_basis = 0;
    int n_chol_vectors = 0;
    int n_occ = 0, n_virt = 0;
    double time_transform_s = 0.0, time_energy_s = 0.0;
};
}  // namespace foundation

struct CholeskyRMP2Config {
    double cholesky_threshold = 1e-6;
    int print_level = 1;
    // 0 = FP64; k >= 2 = split-BF16 with k slices (see mshqc_hip.h).
    int precision = 0;
    // Amplitudes kept for later use (the reference always stores them):
    // 1 = always, 0 = never, -1 = when o^2 v^2 <= 2^27.
    int store_t2 = -1;
};

class CholeskyRMP2 {
public:
    CholeskyRMP2(const Molecule& mol, const BasisSet& basis, std::shared_ptr<IntegralEngine> integrals,
                 const SCFResult& rhf_result, const CholeskyRMP2Config& config);
    CholeskyRMP2(const Molecule& mol, const BasisSet& basis, std::shared_ptr<IntegralEngine> integrals,
                 const SCFResult& rhf_result, const CholeskyRMP2Config& config,
                 const integrals::CholeskyERI& existing_cholesky);
    foundation::CholeskyRMP2Result compute();

private:
    SCFResult rhf_;
    CholeskyRMP2Config config_;
    std::unique_ptr<integrals::CholeskyERI> internal_;
    const integrals::CholeskyERI* chol_ = nullptr;
    int nbf_, nocc_, nvirt_;
};

// ------------------------------------------------------------- CD-RMP3

namespace foundation {
// rmp3.h:52-65 of the reference (t2_1, t2_2 flattened [o][o][v][v]; t2_2
// is kept when o^2 v^2 <= 2^27).
struct RMP3Result {
    double e_rhf = 0.0, e_mp2 = 0.0, e_mp3 = 0.0, e_corr_total = 0.0, e_total = 0.0;
    int n_occ = 0, n_virt = 0;
    Vec t2_1, t2_2;
};

// cholesky_rmp3.h / cholesky_rmp3.cc: second-order amplitudes and the MP3
// energy from the CD-RMP2 result's vectors and amplitudes.
class CholeskyRMP3 {
public:
    CholeskyRMP3(const SCFResult& rhf_result, const CholeskyRMP2Result& crmp2_result, const BasisSet& basis);
    RMP3Result compute();

private:
    SCFResult rhf_;
    CholeskyRMP2Result crmp2_;
    int nbf_, nocc_, nvirt_, n_chol_vectors_;
};
}  // namespace foundation

// ------------------------------------------------------------- CD-UMP2

struct CholeskyUMP2Config {
    double cholesky_threshold = 1e-6;
    bool use_on_the_fly = false;
    bool validate_energy = false;
    int print_level = 1;
    int precision = 0;
};

struct CholeskyUMP2Result {
    double e_corr_ss_aa = 0.0, e_corr_ss_bb = 0.0, e_corr_os = 0.0, e_corr_total = 0.0, e_total = 0.0;
    int n_cholesky_vectors = 0;
    double compression_ratio = 0.0, memory_mb = 0.0;
    double time_cholesky_s = 0.0, time_transform_s = 0.0, time_energy_s = 0.0;
};

class CholeskyUMP2 {
public:
    CholeskyUMP2(const SCFResult& uhf_result, const BasisSet& basis,
                 std::shared_ptr<IntegralEngine> integrals, const CholeskyUMP2Config& config,
                 const integrals::CholeskyERI& cholesky_from_uhf);
    CholeskyUMP2Result compute();
    // cholesky_ump2.cc:471-556: stores t2_aa, t2_bb, t2_ab (i, j, a, b) =
    // g_ijab / D for every ordered pair (memory o^2 v^2 per block).  The
    // vectors are decomposed first when there are none.
    void compute_t2_amplitudes();
    const integrals::CholeskyERI& get_cholesky() const { return chol_; }
    // The stored amplitudes, flattened [o][o'][v][v'] (empty before
    // compute_t2_amplitudes); not part of the reference's Python surface,
    // which keeps them private.
    const Vec& t2_aa() const { return t2_aa_; }
    const Vec& t2_bb() const { return t2_bb_; }
    const Vec& t2_ab() const { return t2_ab_; }
    int n_occ_alpha() const { return na_; }
    int n_occ_beta() const { return nb_; }
    int n_virt_alpha() const { return va_; }
    int n_virt_beta() const { return vb_; }

private:
    friend class CholeskyUMP3;
    Vec t2_aa_, t2_bb_, t2_ab_;
    SCFResult uhf_;
    BasisSet basis_;
    std::shared_ptr<IntegralEngine> integrals_;
    CholeskyUMP2Config config_;
    integrals::CholeskyERI chol_;  // shares the device copy with the source
    int nbf_, na_, nb_, va_, vb_;
};

// ------------------------------------------------------------- CD-UMP3

struct CholeskyUMP3Config {
    double cholesky_threshold = 1e-6;
    bool use_ump2_result = true;
    bool store_intermediates = false;
    bool validate_energy = false;
    int print_level = 1;
};

// cholesky_ump3.h of the reference; like its compute(), e_mp2_* stay 0 and
// e_total is not filled.
struct CholeskyUMP3Result {
    double e_mp2_ss_aa = 0.0, e_mp2_ss_bb = 0.0, e_mp2_os = 0.0, e_mp2_total = 0.0;
    double e_mp3_ss_aa = 0.0, e_mp3_ss_bb = 0.0, e_mp3_os = 0.0, e_mp3_total = 0.0;
    double e_corr_total = 0.0, e_total = 0.0;
    int n_cholesky_vectors = 0;
    double compression_ratio = 0.0, memory_mb = 0.0;
    double time_mp2_s = 0.0, time_mp3_s = 0.0, time_total_s = 0.0;
};

// cholesky_ump3.cc: MP3 energy components from the UHF orbitals and the
// Cholesky vectors (decomposed here, or reused from a CholeskyUMP2 solver).
class CholeskyUMP3 {
public:
    CholeskyUMP3(const SCFResult& uhf_result, const BasisSet& basis, std::shared_ptr<IntegralEngine> integrals,
                 const CholeskyUMP3Config& config = CholeskyUMP3Config());
    CholeskyUMP3(const CholeskyUMP2& ump2_solver, const CholeskyUMP3Config& config = CholeskyUMP3Config());
    CholeskyUMP3Result compute();
    // cholesky_ump3.cc:108-117: decompose when there are no vectors, then
    // transform_cholesky_vectors().
    void initialize_cholesky();
    // cholesky_ump3.cc:119-150: the per-spin MO blocks B = C_occ^T L C_vir,
    // Loo and Lvv, kept on the device for compute().
    void transform_cholesky_vectors();

private:
    struct MoBlocks;
    std::shared_ptr<MoBlocks> blocks_;
    SCFResult uhf_;
    BasisSet basis_;
    std::shared_ptr<IntegralEngine> integrals_;
    CholeskyUMP3Config config_;
    integrals::CholeskyERI chol_;
    int nbf_, na_, nb_, va_, vb_;
};

// ------------------------------------------------------------- CD-OMP2

// mp2.h:52-72 of the reference.  energy_mp2_ss / energy_mp2_os, which the
// reference's CholeskyOMP2 leaves unset, carry the same-spin and
// opposite-spin parts of the last iteration's energy here.
struct OMP2Result {
    double energy_scf = 0.0, energy_mp2_ss = 0.0, energy_mp2_os = 0.0, energy_mp2_corr = 0.0,
           energy_total = 0.0;
    int n_occ_alpha = 0, n_occ_beta = 0, n_virt_alpha = 0, n_virt_beta = 0;
    Vec orbital_energies_alpha, orbital_energies_beta;
    Mat C_alpha, C_beta;
    bool converged = false;
    int iterations = 0;
};

struct CholeskyOMP2Config {
    int max_iterations = 50;
    double energy_threshold = 1e-6;
    double gradient_threshold = 1e-5;
    double cholesky_threshold = 1e-6;
    int print_level = 1;
    // Not in the reference (default off, which is the reference's loop):
    // add the two-particle term dE_MP2/dkappa (mshqc_hip_omp2_grad2p) to
    // each spin's w_ai = 2 F_vo, in the gradient norm and in the rotation.
    bool two_particle_gradient = false;
};

// cholesky_omp2.h / cholesky_omp2.cc: MP2 amplitudes and one-particle density
// from the rotated orbitals each iteration (denominators from the SCF
// orbital energies throughout), generalised Fock from P + C G C^T, a damped
// level-shifted rotation C <- C exp(-kappa).
class CholeskyOMP2 {
public:
    CholeskyOMP2(const Molecule& mol, const BasisSet& basis, std::shared_ptr<IntegralEngine> integrals,
                 const SCFResult& scf_guess, const CholeskyOMP2Config& config = CholeskyOMP2Config());
    CholeskyOMP2(const Molecule& mol, const BasisSet& basis, std::shared_ptr<IntegralEngine> integrals,
                 const SCFResult& scf_guess, const CholeskyOMP2Config& config,
                 const integrals::CholeskyERI& cholesky_vectors);
    OMP2Result compute();

private:
    Molecule mol_;
    BasisSet basis_;
    std::shared_ptr<IntegralEngine> integrals_;
    SCFResult scf_;
    CholeskyOMP2Config config_;
    std::unique_ptr<integrals::CholeskyERI> internal_;
    const integrals::CholeskyERI* chol_ = nullptr;
    int nbf_, na_, nb_, va_, vb_;

    // w2[s] = the two-particle term of spin s, on the device ([v][o]), from
    // device B, orbital energies and orbital blocks
    void two_particle_terms(const double* Ba, const double* Bb, const double* eoa, const double* eva,
                            const double* eob, const double* evb, bool same_orbitals, double* const (&Co)[2],
                            double* const (&Cv)[2], std::unique_ptr<device::Buffer> (&w2)[2]) const;
};

}  // namespace mshqc
