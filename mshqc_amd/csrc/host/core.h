// Host-side classes of the drop-in (C++17): small dense linear algebra,
// Molecule, BasisSet (.gbs), the s/p integral engine, the ERI tensor,
// CholeskyERI, SCF containers and DIIS.  Class and member names follow
// syahrulhidayat/mshqc so callers keep their code; the arithmetic that the
// reference runs through OpenMP/Eigen on the hot path goes through the C ABI
// of libmshqc_hip.so instead (see device.h).
#pragma once

#include <array>
#include <cstddef>
#include <memory>
#include <stdexcept>
#include <string>
#include <vector>

namespace mshqc {

// ------------------------------------------------------------ linear algebra

// Dense row-major matrix of doubles (element (r,c) at r*cols + c).
struct Mat {
    int rows = 0, cols = 0;
    std::vector<double> a;
    Mat() = default;
    Mat(int r, int c, double v = 0.0) : rows(r), cols(c), a((size_t)r * c, v) {}
    double& operator()(int r, int c) { return a[(size_t)r * cols + c]; }
    double operator()(int r, int c) const { return a[(size_t)r * cols + c]; }
    double* data() { return a.data(); }
    const double* data() const { return a.data(); }
    size_t size() const { return a.size(); }
    static Mat identity(int n);
};
using Vec = std::vector<double>;

Mat matmul(const Mat& A, const Mat& B);          // A B
Mat matmul_tn(const Mat& A, const Mat& B);       // A^T B
Mat matmul_nt(const Mat& A, const Mat& B);       // A B^T
Mat transpose(const Mat& A);
Mat add(const Mat& A, const Mat& B, double beta = 1.0);  // A + beta B
Mat scale(const Mat& A, double s);
double trace_product(const Mat& A, const Mat& B);        // tr(A B) = sum_ij A_ij B_ji
double dot(const Mat& A, const Mat& B);                  // sum_ij A_ij B_ij
Mat left_cols(const Mat& A, int k);
Mat right_cols(const Mat& A, int k);
// Symmetric eigendecomposition (Householder + implicit QL): ascending eigenvalues w,
// eigenvectors as the columns of V.
void eigh(const Mat& A, Vec& w, Mat& V);
// Solve A x = b by LU with full pivoting, as the DIIS of the reference does
// (diis.cc, Eigen fullPivLu().solve): pivots searched column by column,
// first largest magnitude wins; pivots at or below n * eps * (largest pivot)
// count as zero and their unknowns are set to 0, so a singular system (e.g.
// vanishing DIIS error vectors) still returns a solution.
Vec solve(Mat A, Vec b);
// exp(A) by scaling and squaring with a Taylor polynomial: A is scaled by
// 2^-s so that ||A||_1 <= 1/2, the series is summed (Horner) to the first
// degree whose remainder bound is below 2^-60, then squared s times.
Mat expm(const Mat& A);

// ----------------------------------------------------------------- molecule

struct Atom {
    int atomic_number = 0;
    double x = 0, y = 0, z = 0;  // bohr
    Atom() = default;
    Atom(int Z, double x_, double y_, double z_) : atomic_number(Z), x(x_), y(y_), z(z_) {}
    std::array<double, 3> position() const { return {x, y, z}; }
};

class Molecule {
public:
    Molecule() = default;
    Molecule(int charge, int multiplicity) : charge_(charge), multiplicity_(multiplicity) {}
    void add_atom(int Z, double x, double y, double z) { atoms_.emplace_back(Z, x, y, z); }
    void add_atom(const Atom& a) { atoms_.push_back(a); }
    int n_atoms() const { return (int)atoms_.size(); }
    const Atom& atom(int i) const;
    const std::vector<Atom>& atoms() const { return atoms_; }
    int total_nuclear_charge() const;
    int n_electrons() const { return total_nuclear_charge() - charge_; }
    int charge() const { return charge_; }
    void set_charge(int q) { charge_ = q; }
    int multiplicity() const { return multiplicity_; }
    void set_multiplicity(int m) { multiplicity_ = m; }
    double nuclear_repulsion_energy() const;

private:
    std::vector<Atom> atoms_;
    int charge_ = 0, multiplicity_ = 1;
};

// -------------------------------------------------------------------- basis

enum class AngularMomentum { S = 0, P = 1, D = 2, F = 3, G = 4, H = 5 };

struct GaussianPrimitive {
    double exponent = 0, coefficient = 0;
    GaussianPrimitive() = default;
    GaussianPrimitive(double e, double c) : exponent(e), coefficient(c) {}
};

class Shell {
public:
    Shell(AngularMomentum am, int center, const std::array<double, 3>& pos)
        : am_(am), center_(center), position_(pos) {}
    void add_primitive(double exponent, double coeff) { primitives_.emplace_back(exponent, coeff); }
    AngularMomentum angular_momentum() const { return am_; }
    int l() const { return (int)am_; }
    int center() const { return center_; }
    const std::array<double, 3>& position() const { return position_; }
    const std::vector<GaussianPrimitive>& primitives() const { return primitives_; }
    bool is_spherical() const { return spherical_; }
    void set_spherical(bool s) { spherical_ = s; }
    int n_functions() const;
    // Coefficient convention of the reference (basis.cc:55-106): primitive
    // norms folded into the coefficients, then the contraction rescaled by
    // 1/sqrt(sum_ij c_i c_j (2 sqrt(a_i a_j)/(a_i+a_j))^(L+3/2)) computed on
    // the already-scaled coefficients.  The resulting coefficients multiply
    // unnormalised primitives x^l y^m z^n exp(-a r^2).
    void normalize();

private:
    AngularMomentum am_;
    int center_;
    std::array<double, 3> position_;
    std::vector<GaussianPrimitive> primitives_;
    bool spherical_ = true;
};

class BasisSet {
public:
    BasisSet() = default;
    BasisSet(const std::string& basis_name, const Molecule& mol,
             const std::string& basis_dir = "data/basis");
    bool read_gbs(const std::string& file, const Molecule& mol);
    void add_shell(const Shell& s) { shells_.push_back(s); }
    int n_shells() const { return (int)shells_.size(); }
    int n_basis_functions() const;
    const Shell& shell(int i) const;
    const std::vector<Shell>& shells() const { return shells_; }
    const std::string& name() const { return name_; }

private:
    std::string name_ = "unknown";
    bool spherical_ = true;
    std::vector<Shell> shells_;
};

// -------------------------------------------------------------- ERI tensor

// Full AO ERI (ij|kl), stored row-major: ((i*n + j)*n + k)*n + l.
struct ERITensor {
    int n = 0;
    std::vector<double> v;
    ERITensor() = default;
    explicit ERITensor(int n_) : n(n_), v((size_t)n_ * n_ * n_ * n_, 0.0) {}
    double& operator()(int i, int j, int k, int l) {
        return v[(((size_t)i * n + j) * n + k) * n + l];
    }
    double operator()(int i, int j, int k, int l) const {
        return v[(((size_t)i * n + j) * n + k) * n + l];
    }
    size_t size() const { return v.size(); }
};

// ------------------------------------------------------- integral engine

// One- and two-electron integrals over contracted Cartesian Gaussians by the
// McMurchie-Davidson scheme (Hermite expansion + Boys function).  Spherical
// shells are supported up to L = 1, where pure and Cartesian functions
// coincide; the reference's integrals come from libint2, which is not part
// of this build.
class IntegralEngine {
public:
    IntegralEngine(const Molecule& mol, const BasisSet& basis);
    Mat compute_overlap() const;
    Mat compute_kinetic() const;
    Mat compute_nuclear() const;
    ERITensor compute_eri() const;
    // integrals.cc:976-1032 of the reference: the diagonal (pq|pq) and one
    // column (pq|rs) over the pairs p >= q in row order (index p(p+1)/2 + q),
    // both from the full ERI.  An out-of-range pivot index leaves (r, s) at
    // (0, 0), as the reference's decoding loop does.
    Vec compute_eri_diagonal() const;
    Vec compute_eri_column(int pivot_index) const;
    int n_basis() const { return nbf_; }

private:
    Molecule mol_;
    BasisSet basis_;
    int nbf_ = 0;
    std::vector<int> offset_;  // first function of each shell
};

// --------------------------------------------------------------- SCF data

struct SCFConfig {
    int max_iterations = 100;
    double energy_threshold = 1e-8;
    double density_threshold = 1e-6;
    double diis_threshold = 1e-2;
    int diis_max_vectors = 8;
    int print_level = 1;
    double level_shift = 0.0;
};

struct SCFResult {
    double energy_electronic = 0, energy_nuclear = 0, energy_total = 0;
    Vec orbital_energies_alpha, orbital_energies_beta;
    Mat C_alpha, C_beta, P_alpha, P_beta, F_alpha, F_beta;
    int iterations = 0;
    bool converged = false;
    double gradient_norm = 0;
    int n_occ_alpha = 0, n_occ_beta = 0;
};

// Pulay DIIS (same bookkeeping as the reference's diis.cc:36-108).
class DIIS {
public:
    explicit DIIS(int max_vectors = 8);
    void add_iteration(const Mat& F, const Mat& error);
    Mat extrapolate() const;
    void clear();
    size_t size() const { return focks_.size(); }
    bool can_extrapolate() const { return size() >= 2; }

private:
    int max_vectors_;
    std::vector<Mat> focks_, errors_;
};

}  // namespace mshqc
