// Host data layer: dense linear algebra, Molecule, BasisSet, the
// McMurchie-Davidson integral engine and DIIS.
#include "core.h"

#include <algorithm>
#include <cctype>
#include <cmath>
#include <limits>
#include <fstream>
#include <sstream>

namespace mshqc {

// ================================================================ algebra

Mat Mat::identity(int n) {
    Mat I(n, n);
    for (int i = 0; i < n; ++i) I(i, i) = 1.0;
    return I;
}

Mat matmul(const Mat& A, const Mat& B) {
    if (A.cols != B.rows) throw std::invalid_argument("matmul: shape mismatch");
    Mat C(A.rows, B.cols);
#pragma omp parallel for if ((long)A.rows * B.cols * A.cols > 1000000)
    for (int i = 0; i < A.rows; ++i)
        for (int k = 0; k < A.cols; ++k) {
            const double a = A(i, k);
            const double* b = &B.a[(size_t)k * B.cols];
            double* c = &C.a[(size_t)i * C.cols];
            for (int j = 0; j < B.cols; ++j) c[j] += a * b[j];
        }
    return C;
}

Mat transpose(const Mat& A) {
    Mat T(A.cols, A.rows);
    for (int i = 0; i < A.rows; ++i)
        for (int j = 0; j < A.cols; ++j) T(j, i) = A(i, j);
    return T;
}

Mat matmul_tn(const Mat& A, const Mat& B) { return matmul(transpose(A), B); }
Mat matmul_nt(const Mat& A, const Mat& B) { return matmul(A, transpose(B)); }

Mat add(const Mat& A, const Mat& B, double beta) {
    if (A.rows != B.rows || A.cols != B.cols) throw std::invalid_argument("add: shape mismatch");
    Mat C = A;
    for (size_t k = 0; k < C.a.size(); ++k) C.a[k] += beta * B.a[k];
    return C;
}

Mat scale(const Mat& A, double s) {
    Mat C = A;
    for (auto& x : C.a) x *= s;
    return C;
}

double trace_product(const Mat& A, const Mat& B) {
    double t = 0.0;
    for (int i = 0; i < A.rows; ++i)
        for (int k = 0; k < A.cols; ++k) t += A(i, k) * B(k, i);
    return t;
}

double dot(const Mat& A, const Mat& B) {
    double t = 0.0;
    for (size_t k = 0; k < A.a.size(); ++k) t += A.a[k] * B.a[k];
    return t;
}

Mat left_cols(const Mat& A, int k) {
    Mat R(A.rows, k);
    for (int i = 0; i < A.rows; ++i)
        for (int j = 0; j < k; ++j) R(i, j) = A(i, j);
    return R;
}

Mat right_cols(const Mat& A, int k) {
    Mat R(A.rows, k);
    const int off = A.cols - k;
    for (int i = 0; i < A.rows; ++i)
        for (int j = 0; j < k; ++j) R(i, j) = A(i, off + j);
    return R;
}

// Symmetric eigendecomposition in two steps (Householder, then the implicit
// QL iteration of Wilkinson and Reinsch on the tridiagonal form), O(n^3) with
// small constants, where the cyclic Jacobi sweeps it replaces took seconds at
// n = 400.
//  1. Householder: for k = 0 .. n-3 the reflector H_k = I - beta v v^T
//     (v zero above row k+1) maps column k below the subdiagonal to zero;
//     A <- H_k A H_k is applied to the trailing block as a rank-2 update, and
//     Q = H_0 H_1 ... H_{n-3} is accumulated backwards, so A = Q T Q^T.
//  2. QL: T's eigenvalues by implicitly shifted QL sweeps (shift from the
//     leading 2 x 2 block, Givens rotations chasing the bulge), the same
//     rotations applied to the rows of Z = Q^T, so the eigenvectors are the
//     rows of Z (contiguous updates).
// Eigenvalues ascending, eigenvectors the columns of V.
void eigh(const Mat& Ain, Vec& w, Mat& V) {
    const int n = Ain.rows;
    if (Ain.cols != n) throw std::invalid_argument("eigh: matrix must be square");
    w.assign(n, 0.0);
    V = Mat(n, n);
    if (n == 0) return;
    Mat A = Ain;
    // symmetrise against round-off in the input
    for (int i = 0; i < n; ++i)
        for (int j = i + 1; j < n; ++j) A(i, j) = A(j, i) = 0.5 * (A(i, j) + A(j, i));
    std::vector<std::vector<double>> refl(n > 2 ? n - 2 : 0);  // reflector v_k (rows k+1 ..)
    std::vector<double> beta(n > 2 ? n - 2 : 0, 0.0);
    std::vector<double> p(n), q(n);
    for (int k = 0; k + 2 < n; ++k) {
        const int m = n - k - 1;  // length of v
        std::vector<double>& v = refl[k];
        v.assign(m, 0.0);
        double sc = 0.0;
        for (int i = 0; i < m; ++i) sc = std::max(sc, std::fabs(A(k + 1 + i, k)));
        if (sc == 0.0) continue;  // column already reduced: H_k = I
        double nrm2 = 0.0;
        for (int i = 0; i < m; ++i) {
            v[i] = A(k + 1 + i, k) / sc;
            nrm2 += v[i] * v[i];
        }
        const double nrm = std::sqrt(nrm2);
        const double alpha = v[0] >= 0.0 ? -nrm : nrm;  // H x = alpha e_1 (scaled)
        v[0] -= alpha;
        const double vv = nrm2 - 2.0 * alpha * (v[0] + alpha) + alpha * alpha;  // |x - alpha e1|^2
        if (vv == 0.0) continue;
        const double bk = 2.0 / vv;
        beta[k] = bk;
        // column k and row k of the reduced matrix
        A(k + 1, k) = A(k, k + 1) = alpha * sc;
        for (int i = 1; i < m; ++i) A(k + 1 + i, k) = A(k, k + 1 + i) = 0.0;
        // trailing block B (rows/cols k+1..): B <- B - v q^T - q v^T with
        // p = beta B v, q = p - (beta/2)(v^T p) v
        for (int i = 0; i < m; ++i) {
            double t = 0.0;
            const double* row = &A(k + 1 + i, k + 1);
            for (int j = 0; j < m; ++j) t += row[j] * v[j];
            p[i] = bk * t;
        }
        double vp = 0.0;
        for (int i = 0; i < m; ++i) vp += v[i] * p[i];
        const double K = 0.5 * bk * vp;
        for (int i = 0; i < m; ++i) q[i] = p[i] - K * v[i];
        for (int i = 0; i < m; ++i) {
            double* row = &A(k + 1 + i, k + 1);
            const double vi = v[i], qi = q[i];
            for (int j = 0; j < m; ++j) row[j] -= vi * q[j] + qi * v[j];
        }
    }
    std::vector<double> d(n), e(n, 0.0);
    for (int i = 0; i < n; ++i) d[i] = A(i, i);
    for (int i = 0; i + 1 < n; ++i) e[i] = A(i + 1, i);
    // Z = Q^T = H_{n-3} ... H_0 (the H_k are symmetric): rows of Z are
    // Q's columns.  Built as Z = I, then Z <- Z H_k for k descending applied
    // to rows: Z^T = H_0 ... H_{n-3} accumulated from the right end.
    Mat Z = Mat::identity(n);  // holds Q row-major: Z(i, j) = Q(i, j)
    for (int k = n - 3; k >= 0; --k) {
        if (beta[k] == 0.0) continue;
        const int m = n - k - 1;
        const std::vector<double>& v = refl[k];
        // Q[k+1.., k+1..] <- H_k Q[k+1.., k+1..]
        std::fill(p.begin(), p.end(), 0.0);
        for (int i = 0; i < m; ++i) {
            const double vi = v[i];
            const double* row = &Z(k + 1 + i, k + 1);
            for (int j = 0; j < m; ++j) p[j] += vi * row[j];
        }
        for (int i = 0; i < m; ++i) {
            const double f = beta[k] * v[i];
            double* row = &Z(k + 1 + i, k + 1);
            for (int j = 0; j < m; ++j) row[j] -= f * p[j];
        }
    }
    Mat E = transpose(Z);  // E(r, :) = column r of Q: the rotations act on rows
    // implicit QL on (d, e)
    for (int l = 0; l < n; ++l) {
        for (int iter = 0;; ++iter) {
            int m = l;
            for (; m + 1 < n; ++m) {
                const double dd = std::fabs(d[m]) + std::fabs(d[m + 1]);
                if (std::fabs(e[m]) <= std::numeric_limits<double>::epsilon() * dd) break;
            }
            if (m == l) break;
            if (iter == 60) throw std::runtime_error("eigh: QL iteration did not converge");
            // shift from the 2 x 2 block at l
            double g = (d[l + 1] - d[l]) / (2.0 * e[l]);
            double r = std::hypot(g, 1.0);
            g = d[m] - d[l] + e[l] / (g + (g >= 0.0 ? r : -r));
            double s = 1.0, c = 1.0, pp = 0.0;
            int i = m - 1;
            for (; i >= l; --i) {
                double f = s * e[i];
                const double b = c * e[i];
                r = std::hypot(f, g);
                e[i + 1] = r;
                if (r == 0.0) {  // split: recover and restart
                    d[i + 1] -= pp;
                    e[m] = 0.0;
                    break;
                }
                s = f / r;
                c = g / r;
                g = d[i + 1] - pp;
                r = (d[i] - g) * s + 2.0 * c * b;
                pp = s * r;
                d[i + 1] = g + pp;
                g = c * r - b;
                // rotate eigenvector rows i, i+1
                double* zi = &E(i, 0);
                double* zj = &E(i + 1, 0);
                for (int t = 0; t < n; ++t) {
                    const double a0 = zi[t], a1 = zj[t];
                    zj[t] = s * a0 + c * a1;
                    zi[t] = c * a0 - s * a1;
                }
            }
            if (r == 0.0 && i >= l) continue;
            d[l] -= pp;
            e[l] = g;
            e[m] = 0.0;
        }
    }
    std::vector<int> idx(n);
    for (int i = 0; i < n; ++i) idx[i] = i;
    std::stable_sort(idx.begin(), idx.end(), [&](int a2, int b2) { return d[a2] < d[b2]; });
    for (int k = 0; k < n; ++k) {
        w[k] = d[idx[k]];
        const double* src = &E(idx[k], 0);
        for (int r = 0; r < n; ++r) V(r, k) = src[r];
    }
}

Vec solve(Mat A, Vec b) {
    const int n = A.rows;
    std::vector<int> colperm(n);
    for (int i = 0; i < n; ++i) colperm[i] = i;
    int nz = n;  // pivots before the first all-zero corner
    double maxpivot = 0.0;
    for (int k = 0; k < n; ++k) {
        int pr = k, pc = k;
        double best = -1.0;
        for (int j = k; j < n; ++j)
            for (int i = k; i < n; ++i)
                if (std::fabs(A(i, j)) > best) best = std::fabs(A(i, j)), pr = i, pc = j;
        if (best == 0.0) {
            nz = k;
            break;
        }
        maxpivot = std::max(maxpivot, best);
        if (pr != k) {
            for (int j = 0; j < n; ++j) std::swap(A(k, j), A(pr, j));
            std::swap(b[k], b[pr]);
        }
        if (pc != k) {
            for (int i = 0; i < n; ++i) std::swap(A(i, k), A(i, pc));
            std::swap(colperm[k], colperm[pc]);
        }
        for (int i = k + 1; i < n; ++i) {
            const double f = A(i, k) / A(k, k);
            for (int j = k + 1; j < n; ++j) A(i, j) -= f * A(k, j);
            A(i, k) = f;
            b[i] -= f * b[k];
        }
    }
    const double thr = maxpivot * std::numeric_limits<double>::epsilon() * n;
    int rank = 0;
    for (int k = 0; k < nz; ++k) rank += std::fabs(A(k, k)) > thr;
    Vec y(n, 0.0);
    for (int i = rank - 1; i >= 0; --i) {
        double s = b[i];
        for (int j = i + 1; j < rank; ++j) s -= A(i, j) * y[j];
        y[i] = s / A(i, i);
    }
    Vec x(n, 0.0);
    for (int i = 0; i < n; ++i) x[colperm[i]] = y[i];
    return x;
}

Mat expm(const Mat& A) {
    const int n = A.rows;
    if (A.cols != n) throw std::invalid_argument("expm: matrix must be square");
    double norm1 = 0.0;
    for (int j = 0; j < n; ++j) {
        double c = 0.0;
        for (int i = 0; i < n; ++i) c += std::fabs(A(i, j));
        norm1 = std::max(norm1, c);
    }
    int sq = 0;
    while (norm1 > 0.5) {
        norm1 *= 0.5;
        ++sq;
    }
    const Mat X = scale(A, std::ldexp(1.0, -sq));
    // smallest m with ||X||^(m+1)/(m+1)! < 2^-60
    int m = 1;
    double term = norm1;
    while (m < 30 && term * norm1 / (m + 1) >= std::ldexp(1.0, -60)) {
        term *= norm1 / (m + 1);
        ++m;
    }
    // Horner: E = I + X/1 (I + X/2 (I + ... (I + X/m)))
    Mat E = Mat::identity(n);
    for (int k = m; k >= 1; --k) {
        Mat XE = scale(matmul(X, E), 1.0 / k);
        for (int i = 0; i < n; ++i) XE(i, i) += 1.0;
        E = XE;
    }
    for (int k = 0; k < sq; ++k) E = matmul(E, E);
    return E;
}

// ================================================================ molecule

const Atom& Molecule::atom(int i) const {
    if (i < 0 || i >= n_atoms()) throw std::out_of_range("Molecule::atom: index out of range");
    return atoms_[i];
}

int Molecule::total_nuclear_charge() const {
    int z = 0;
    for (const auto& a : atoms_) z += a.atomic_number;
    return z;
}

double Molecule::nuclear_repulsion_energy() const {
    double e = 0.0;
    for (size_t i = 0; i < atoms_.size(); ++i)
        for (size_t j = i + 1; j < atoms_.size(); ++j) {
            const double dx = atoms_[i].x - atoms_[j].x, dy = atoms_[i].y - atoms_[j].y,
                         dz = atoms_[i].z - atoms_[j].z;
            e += atoms_[i].atomic_number * atoms_[j].atomic_number / std::sqrt(dx * dx + dy * dy + dz * dz);
        }
    return e;
}

// =================================================================== basis

namespace {

const char* kSymbols[] = {"",   "H",  "He", "Li", "Be", "B",  "C",  "N",  "O",  "F",
                          "Ne", "Na", "Mg", "Al", "Si", "P",  "S",  "Cl", "Ar", "K",
                          "Ca", "Sc", "Ti", "V",  "Cr", "Mn", "Fe", "Co", "Ni", "Cu",
                          "Zn", "Ga", "Ge", "As", "Se", "Br", "Kr"};

double double_factorial(int n) {
    double r = 1.0;
    for (int k = n; k > 1; k -= 2) r *= k;
    return r;
}

std::string upper(std::string s) {
    for (auto& c : s) c = (char)std::toupper((unsigned char)c);
    return s;
}

double parse_number(std::string s) {
    for (auto& c : s)
        if (c == 'D' || c == 'd') c = 'E';
    return std::stod(s);
}

int am_of(char c) {
    const std::string L = "SPDFGH";
    const auto k = L.find((char)std::toupper((unsigned char)c));
    return k == std::string::npos ? -1 : (int)k;
}

}  // namespace

int Shell::n_functions() const {
    const int L = l();
    return spherical_ ? 2 * L + 1 : (L + 1) * (L + 2) / 2;
}

void Shell::normalize() {
    const int L = l();
    for (auto& p : primitives_) {
        const double N = std::pow(2.0 * p.exponent / M_PI, 0.75) * std::pow(4.0 * p.exponent, L / 2.0) /
                         std::sqrt(double_factorial(2 * L - 1));
        p.coefficient *= N;
    }
    double sum = 0.0;
    for (const auto& a : primitives_)
        for (const auto& b : primitives_) {
            const double ratio = 2.0 * std::sqrt(a.exponent * b.exponent) / (a.exponent + b.exponent);
            sum += a.coefficient * b.coefficient * std::pow(ratio, L + 1.5);
        }
    if (sum > 1e-14) {
        const double k = 1.0 / std::sqrt(sum);
        for (auto& p : primitives_) p.coefficient *= k;
    }
}

BasisSet::BasisSet(const std::string& basis_name, const Molecule& mol, const std::string& basis_dir)
    : name_(basis_name) {
    std::string lower = basis_name;
    for (auto& c : lower) c = (char)std::tolower((unsigned char)c);
    const std::string file = basis_dir + "/" + lower + ".gbs";
    if (!read_gbs(file, mol)) throw std::runtime_error("Failed to read basis set: " + file);
}

bool BasisSet::read_gbs(const std::string& file, const Molecule& mol) {
    std::ifstream f(file);
    if (!f) return false;
    std::vector<std::string> lines;
    std::string line;
    while (std::getline(f, line)) lines.push_back(line);
    if (lines.empty()) return false;
    spherical_ = lines[0].find("cartesian") == std::string::npos;
    shells_.clear();
    for (int ia = 0; ia < mol.n_atoms(); ++ia) {
        const Atom& at = mol.atom(ia);
        if (at.atomic_number <= 0 || at.atomic_number > 36)
            throw std::runtime_error("read_gbs: unsupported element Z=" + std::to_string(at.atomic_number));
        const std::string sym = upper(kSymbols[at.atomic_number]);
        // Locate "SYM 0" right after a "****" line.
        size_t k = 0;
        bool found = false;
        for (; k + 1 < lines.size(); ++k) {
            if (lines[k].find("****") == std::string::npos) continue;
            std::istringstream is(lines[k + 1]);
            std::string s;
            if ((is >> s) && upper(s) == sym) {
                found = true;
                k += 2;
                break;
            }
        }
        if (!found) throw std::runtime_error("read_gbs: element " + sym + " not in " + file);
        const std::array<double, 3> pos = at.position();
        while (k < lines.size() && lines[k].find("****") == std::string::npos) {
            std::istringstream is(lines[k]);
            std::string type;
            int nprim = 0;
            if (!(is >> type >> nprim)) {
                ++k;
                continue;
            }
            type = upper(type);
            ++k;
            if (type == "SP") {
                Shell s(AngularMomentum::S, ia, pos), p(AngularMomentum::P, ia, pos);
                s.set_spherical(spherical_);
                p.set_spherical(spherical_);
                for (int i = 0; i < nprim && k < lines.size(); ++i, ++k) {
                    std::istringstream ps(lines[k]);
                    std::string e, cs, cp;
                    ps >> e >> cs >> cp;
                    s.add_primitive(parse_number(e), parse_number(cs));
                    p.add_primitive(parse_number(e), parse_number(cp));
                }
                shells_.push_back(s);
                shells_.push_back(p);
            } else {
                const int L = am_of(type[0]);
                if (L < 0 || type.size() != 1) throw std::runtime_error("read_gbs: bad shell type " + type);
                Shell s((AngularMomentum)L, ia, pos);
                s.set_spherical(spherical_);
                for (int i = 0; i < nprim && k < lines.size(); ++i, ++k) {
                    std::istringstream ps(lines[k]);
                    std::string e, c;
                    ps >> e >> c;
                    s.add_primitive(parse_number(e), parse_number(c));
                }
                shells_.push_back(s);
            }
        }
    }
    for (auto& s : shells_) s.normalize();
    return true;
}

int BasisSet::n_basis_functions() const {
    int n = 0;
    for (const auto& s : shells_) n += s.n_functions();
    return n;
}

const Shell& BasisSet::shell(int i) const {
    if (i < 0 || i >= n_shells()) throw std::out_of_range("BasisSet::shell: index out of range");
    return shells_[i];
}

// ========================================================= integral engine

namespace {

constexpr int LMAX = 4;                 // per shell (s .. g)
constexpr int TMAX = 4 * LMAX;          // Hermite index bound in an ERI

int n_cart(int L) { return (L + 1) * (L + 2) / 2; }

// Cartesian exponents of component c of a shell with angular momentum L, in
// the order x^L, x^(L-1) y, x^(L-1) z, ..., z^L (x, y, z for L = 1; xx, xy,
// xz, yy, yz, zz for L = 2).
void cart(int L, int c, int& lx, int& ly, int& lz) {
    lx = ly = lz = 0;
    int idx = 0;
    for (int a = L; a >= 0; --a)
        for (int b = L - a; b >= 0; --b) {
            if (idx == c) {
                lx = a;
                ly = b;
                lz = L - a - b;
                return;
            }
            ++idx;
        }
}

int cart_index(int L, int lx, int ly) {
    int idx = 0;
    for (int a = L; a > lx; --a) idx += L - a + 1;
    return idx + (L - lx - ly);
}

double binom(int n, int k) {
    if (k < 0 || k > n) return 0.0;
    double r = 1.0;
    for (int i = 1; i <= k; ++i) r = r * (n - k + i) / i;
    return r;
}

// Boys function F_n(T), n = 0..nmax: series for the top order, downward
// recursion below (F_n = (2T F_{n+1} + e^-T) / (2n+1)); the asymptotic form
// (2n-1)!! / 2^(n+1) sqrt(pi / T^(2n+1)) once the neglected e^-T terms are
// below ~1e-14 of F_n, which needs larger T for larger n.
void boys(int nmax, double T, double* F) {
    if (T < 40.0 + 2.0 * nmax) {
        const double e = std::exp(-T);
        double term = 1.0 / (2 * nmax + 1), sum = term;
        for (int k = 1; k < 600; ++k) {
            term *= 2.0 * T / (2 * nmax + 2 * k + 1);
            sum += term;
            if (term < 1e-17 * sum) break;
        }
        F[nmax] = e * sum;
        for (int n = nmax - 1; n >= 0; --n) F[n] = (2.0 * T * F[n + 1] + e) / (2 * n + 1);
    } else {
        F[0] = 0.5 * std::sqrt(M_PI / T);
        for (int n = 1; n <= nmax; ++n) F[n] = F[n - 1] * (2 * n - 1) / (2.0 * T);
    }
}

// 1-D Hermite expansion coefficients E[i][j][t] for exponents a, b, centres
// separated by Xab = A - B.
struct Herm {
    double E[LMAX + 3][LMAX + 3][2 * LMAX + 5];
};

void hermite(double a, double b, double Xab, int imax, int jmax, Herm& h) {
    const double p = a + b, mu = a * b / p;
    const double Xpa = -b * Xab / p, Xpb = a * Xab / p;
    for (int i = 0; i <= imax; ++i)
        for (int j = 0; j <= jmax; ++j)
            for (int t = 0; t <= i + j + 1; ++t) h.E[i][j][t] = 0.0;
    h.E[0][0][0] = std::exp(-mu * Xab * Xab);
    for (int i = 0; i <= imax; ++i)
        for (int j = 0; j <= jmax; ++j) {
            if (i == 0 && j == 0) continue;
            for (int t = 0; t <= i + j; ++t) {
                double v = 0.0;
                if (i > 0) {
                    if (t > 0) v += h.E[i - 1][j][t - 1] / (2 * p);
                    v += Xpa * h.E[i - 1][j][t];
                    if (t + 1 <= i - 1 + j) v += (t + 1) * h.E[i - 1][j][t + 1];
                } else {
                    if (t > 0) v += h.E[i][j - 1][t - 1] / (2 * p);
                    v += Xpb * h.E[i][j - 1][t];
                    if (t + 1 <= i + j - 1) v += (t + 1) * h.E[i][j - 1][t + 1];
                }
                h.E[i][j][t] = v;
            }
        }
}

// Hermite Coulomb integrals R_{tuv}^{(0)}(p, PC) for t+u+v <= nmax.
struct RTab {
    double R[TMAX + 1][TMAX + 1][TMAX + 1];
};

void hermite_coulomb(int nmax, double p, double X, double Y, double Z, RTab& out) {
    static thread_local double Rn[TMAX + 1][TMAX + 1][TMAX + 1][TMAX + 1];
    double F[TMAX + 1];
    boys(nmax, p * (X * X + Y * Y + Z * Z), F);
    double pw = 1.0;
    for (int n = 0; n <= nmax; ++n) {
        Rn[n][0][0][0] = pw * F[n];
        pw *= -2.0 * p;
    }
    // Build R^n_{tuv} for t+u+v <= nmax - n, top-down in n.
    for (int n = nmax - 1; n >= 0; --n) {
        const int lim = nmax - n;
        for (int t = 0; t <= lim; ++t)
            for (int u = 0; u <= lim - t; ++u)
                for (int v = 0; v <= lim - t - u; ++v) {
                    if (t + u + v == 0) continue;
                    double r;
                    if (t > 0) {
                        r = X * Rn[n + 1][t - 1][u][v];
                        if (t > 1) r += (t - 1) * Rn[n + 1][t - 2][u][v];
                    } else if (u > 0) {
                        r = Y * Rn[n + 1][t][u - 1][v];
                        if (u > 1) r += (u - 1) * Rn[n + 1][t][u - 2][v];
                    } else {
                        r = Z * Rn[n + 1][t][u][v - 1];
                        if (v > 1) r += (v - 1) * Rn[n + 1][t][u][v - 2];
                    }
                    Rn[n][t][u][v] = r;
                }
    }
    for (int t = 0; t <= nmax; ++t)
        for (int u = 0; u <= nmax - t; ++u)
            for (int v = 0; v <= nmax - t - u; ++v) out.R[t][u][v] = Rn[0][t][u][v];
}

// Real solid harmonics S_lm (l >= 2; m = -l..l, the sine type for m < 0) over
// a shell's Cartesian components, from the closed form in Helgaker,
// Jorgensen and Olsen, "Molecular Electronic-Structure Theory", eq. 6.4.47-50,
// each row scaled so that S_lm is normalised when every Cartesian component
// carries the normalisation of x^l (the shells' coefficient convention): for
// one exponent, <x^a y^b z^c | x^a' y^b' z^c'> is proportional to
// (a+a'-1)!! (b+b'-1)!! (c+c'-1)!!.  p shells stay Cartesian (x, y, z).
struct SphTab {
    std::vector<double> T[LMAX + 1];  // [(2l+1)][n_cart(l)]
    SphTab() {
        for (int l = 2; l <= LMAX; ++l) {
            const int nc = n_cart(l);
            T[l].assign((size_t)(2 * l + 1) * nc, 0.0);
            for (int m = -l; m <= l; ++m) {
                double* row = &T[l][(size_t)(m + l) * nc];
                const int am = std::abs(m), km = m < 0 ? 1 : 0;
                for (int t = 0; t <= (l - am) / 2; ++t)
                    for (int u = 0; u <= t; ++u)
                        for (int k = km; k <= am; k += 2) {
                            const int ex = 2 * t + am - 2 * u - k, ey = 2 * u + k;
                            const double sgn = ((t + (k - km) / 2) & 1) ? -1.0 : 1.0;
                            row[cart_index(l, ex, ey)] += sgn * std::pow(0.25, t) * binom(l, t) *
                                                          binom(l - t, am + t) * binom(t, u) * binom(am, k);
                        }
                double nrm = 0.0;
                for (int a = 0; a < nc; ++a)
                    for (int b = 0; b < nc; ++b) {
                        int ax, ay, az, bx, by, bz;
                        cart(l, a, ax, ay, az);
                        cart(l, b, bx, by, bz);
                        if (row[a] == 0.0 || row[b] == 0.0 || ((ax + bx) | (ay + by) | (az + bz)) & 1) continue;
                        nrm += row[a] * row[b] * double_factorial(ax + bx - 1) * double_factorial(ay + by - 1) *
                               double_factorial(az + bz - 1);
                    }
                const double s = 1.0 / std::sqrt(nrm / double_factorial(2 * l - 1));
                for (int a = 0; a < nc; ++a) row[a] *= s;
            }
        }
    }
};

const SphTab& sph_tab() {
    static const SphTab t;
    return t;
}

// Cartesian -> shell-function matrix of a shell ([n_functions][n_cart]), or
// null when the shell's functions are its Cartesian components.
const double* sph_rows(const Shell& s) {
    return (s.is_spherical() && s.l() >= 2) ? sph_tab().T[s.l()].data() : nullptr;
}

// out[f][rest] = sum_c T[f][c] in[c][rest] along one index of a row-major
// block [outer][nc][rest] (T null: a copy).
void to_shell_index(const double* T, int nf, int nc, long outer, long rest, const double* in, double* out) {
    for (long o = 0; o < outer; ++o)
        for (int f = 0; f < nf; ++f) {
            double* dst = out + (o * nf + f) * rest;
            for (long r = 0; r < rest; ++r) dst[r] = 0.0;
            for (int c = 0; c < nc; ++c) {
                const double w = T ? T[(size_t)f * nc + c] : (f == c ? 1.0 : 0.0);
                if (w == 0.0) continue;
                const double* src = in + (o * nc + c) * rest;
                for (long r = 0; r < rest; ++r) dst[r] += w * src[r];
            }
        }
}

void check_shells(const BasisSet& b) {
    for (const auto& s : b.shells())
        if (s.l() > LMAX)
            throw std::runtime_error(
                "IntegralEngine: shells above L=4 (g) are not supported by this build's integral engine");
}

}  // namespace

IntegralEngine::IntegralEngine(const Molecule& mol, const BasisSet& basis) : mol_(mol), basis_(basis) {
    check_shells(basis_);
    nbf_ = 0;
    for (const auto& s : basis_.shells()) {
        offset_.push_back(nbf_);
        nbf_ += s.n_functions();
    }
}

namespace {
enum OneKind { OVERLAP, KINETIC, NUCLEAR };

Mat one_body(const BasisSet& basis, const std::vector<int>& offset, int nbf, const Molecule& mol,
             OneKind kind) {
    Mat M(nbf, nbf);
    const auto& sh = basis.shells();
    const int ns = (int)sh.size();
    for (int s1 = 0; s1 < ns; ++s1)
        for (int s2 = 0; s2 <= s1; ++s2) {
            const Shell &A = sh[s1], &B = sh[s2];
            const int LA = A.l(), LB = B.l(), ncA = n_cart(LA), ncB = n_cart(LB);
            const auto &PA = A.position(), &PB = B.position();
            std::vector<double> blk((size_t)ncA * ncB, 0.0);
            for (const auto& pa : A.primitives())
                for (const auto& pb : B.primitives()) {
                    const double a = pa.exponent, b = pb.exponent, p = a + b;
                    Herm hx, hy, hz;
                    hermite(a, b, PA[0] - PB[0], LA, LB + 2, hx);
                    hermite(a, b, PA[1] - PB[1], LA, LB + 2, hy);
                    hermite(a, b, PA[2] - PB[2], LA, LB + 2, hz);
                    const double cc = pa.coefficient * pb.coefficient;
                    const double s3 = std::pow(M_PI / p, 1.5);
                    RTab R;
                    for (int ca = 0; ca < ncA; ++ca)
                        for (int cb = 0; cb < ncB; ++cb) {
                            int ax, ay, az, bx, by, bz;
                            cart(LA, ca, ax, ay, az);
                            cart(LB, cb, bx, by, bz);
                            double val = 0.0;
                            if (kind == OVERLAP) {
                                val = hx.E[ax][bx][0] * hy.E[ay][by][0] * hz.E[az][bz][0] * s3;
                            } else if (kind == KINETIC) {
                                auto k1 = [&](const Herm& h, int i, int j) {
                                    double t = b * (2 * j + 1) * h.E[i][j][0] - 2 * b * b * h.E[i][j + 2][0];
                                    if (j >= 2) t -= 0.5 * j * (j - 1) * h.E[i][j - 2][0];
                                    return t;
                                };
                                val = (k1(hx, ax, bx) * hy.E[ay][by][0] * hz.E[az][bz][0] +
                                       hx.E[ax][bx][0] * k1(hy, ay, by) * hz.E[az][bz][0] +
                                       hx.E[ax][bx][0] * hy.E[ay][by][0] * k1(hz, az, bz)) * s3;
                            } else {
                                const double Px = (a * PA[0] + b * PB[0]) / p, Py = (a * PA[1] + b * PB[1]) / p,
                                             Pz = (a * PA[2] + b * PB[2]) / p;
                                for (const auto& at : mol.atoms()) {
                                    const int nmax = ax + bx + ay + by + az + bz;
                                    hermite_coulomb(nmax, p, Px - at.x, Py - at.y, Pz - at.z, R);
                                    double s = 0.0;
                                    for (int t = 0; t <= ax + bx; ++t)
                                        for (int u = 0; u <= ay + by; ++u)
                                            for (int v = 0; v <= az + bz; ++v)
                                                s += hx.E[ax][bx][t] * hy.E[ay][by][u] * hz.E[az][bz][v] *
                                                     R.R[t][u][v];
                                    val -= at.atomic_number * 2.0 * M_PI / p * s;
                                }
                            }
                            blk[(size_t)ca * ncB + cb] += cc * val;
                        }
                }
            // Cartesian block -> shell functions, index A then index B
            const int na = A.n_functions(), nb = B.n_functions();
            std::vector<double> t1((size_t)na * ncB), t2((size_t)nb * na);
            to_shell_index(sph_rows(A), na, ncA, 1, ncB, blk.data(), t1.data());
            // transpose to [ncB][na], transform B, giving [nb][na]
            std::vector<double> tt((size_t)ncB * na);
            for (int f = 0; f < na; ++f)
                for (int cb = 0; cb < ncB; ++cb) tt[(size_t)cb * na + f] = t1[(size_t)f * ncB + cb];
            to_shell_index(sph_rows(B), nb, ncB, 1, na, tt.data(), t2.data());
            for (int fa = 0; fa < na; ++fa)
                for (int fb = 0; fb < nb; ++fb) {
                    const double v = t2[(size_t)fb * na + fa];
                    M(offset[s1] + fa, offset[s2] + fb) = v;
                    M(offset[s2] + fb, offset[s1] + fa) = v;
                }
        }
    return M;
}

// Hermite coefficients of one primitive pair's Cartesian products: for
// product (a, b) the nonzero E^a_t E^b_u E^c_v, t+u+v indexed in a table of
// the pair's Hermite functions.
struct PairHerm {
    int nh = 0;                    // Hermite functions t+u+v <= La+Lb
    std::vector<int> tuv;          // [nh][3]
    std::vector<int> start;        // [ncart pairs + 1] into idx / coef
    std::vector<int> idx;
    std::vector<double> coef;
};

int hidx(int L, int t, int u, int v) {
    // flat index of (t, u, v), t+u+v <= L, ordered by t, then u, then v
    int k = 0;
    for (int tt = 0; tt < t; ++tt) k += (L - tt + 1) * (L - tt + 2) / 2;
    for (int uu = 0; uu < u; ++uu) k += L - t - uu + 1;
    return k + v;
}

void pair_herm(const Shell& A, const Shell& B, const Herm& ex, const Herm& ey, const Herm& ez, PairHerm& ph) {
    const int LA = A.l(), LB = B.l(), L = LA + LB, ncA = n_cart(LA), ncB = n_cart(LB);
    ph.nh = (L + 1) * (L + 2) * (L + 3) / 6;
    ph.tuv.resize((size_t)3 * ph.nh);
    for (int t = 0; t <= L; ++t)
        for (int u = 0; u <= L - t; ++u)
            for (int v = 0; v <= L - t - u; ++v) {
                const int k = hidx(L, t, u, v);
                ph.tuv[3 * k] = t;
                ph.tuv[3 * k + 1] = u;
                ph.tuv[3 * k + 2] = v;
            }
    ph.start.assign((size_t)ncA * ncB + 1, 0);
    ph.idx.clear();
    ph.coef.clear();
    for (int ca = 0; ca < ncA; ++ca)
        for (int cb = 0; cb < ncB; ++cb) {
            int ax, ay, az, bx, by, bz;
            cart(LA, ca, ax, ay, az);
            cart(LB, cb, bx, by, bz);
            for (int t = 0; t <= ax + bx; ++t)
                for (int u = 0; u <= ay + by; ++u)
                    for (int v = 0; v <= az + bz; ++v) {
                        const double e = ex.E[ax][bx][t] * ey.E[ay][by][u] * ez.E[az][bz][v];
                        if (e == 0.0) continue;
                        ph.idx.push_back(hidx(L, t, u, v));
                        ph.coef.push_back(e);
                    }
            ph.start[(size_t)ca * ncB + cb + 1] = (int)ph.idx.size();
        }
}
}  // namespace

Mat IntegralEngine::compute_overlap() const { return one_body(basis_, offset_, nbf_, mol_, OVERLAP); }
Mat IntegralEngine::compute_kinetic() const { return one_body(basis_, offset_, nbf_, mol_, KINETIC); }
Mat IntegralEngine::compute_nuclear() const { return one_body(basis_, offset_, nbf_, mol_, NUCLEAR); }

Vec IntegralEngine::compute_eri_diagonal() const {
    const ERITensor eri = compute_eri();
    Vec d;
    d.reserve((size_t)nbf_ * (nbf_ + 1) / 2);
    for (int i = 0; i < nbf_; ++i)
        for (int j = 0; j <= i; ++j) d.push_back(eri(i, j, i, j));
    return d;
}

Vec IntegralEngine::compute_eri_column(int pivot_index) const {
    int r = 0, s = 0;
    for (int i = 0, idx = 0; i < nbf_; ++i)
        for (int j = 0; j <= i; ++j, ++idx)
            if (idx == pivot_index) {
                r = i;
                s = j;
            }
    const ERITensor eri = compute_eri();
    Vec col;
    col.reserve((size_t)nbf_ * (nbf_ + 1) / 2);
    for (int p = 0; p < nbf_; ++p)
        for (int q = 0; q <= p; ++q) col.push_back(eri(p, q, r, s));
    return col;
}

ERITensor IntegralEngine::compute_eri() const {
    ERITensor eri(nbf_);
    const auto& sh = basis_.shells();
    const int ns = (int)sh.size();
    // Shell pairs (s1 >= s2), then quartets with pair1 >= pair2 (8-fold symmetry).
    std::vector<std::pair<int, int>> pairs;
    for (int s1 = 0; s1 < ns; ++s1)
        for (int s2 = 0; s2 <= s1; ++s2) pairs.emplace_back(s1, s2);
    const long np = (long)pairs.size();
#pragma omp parallel for schedule(dynamic)
    for (long p1 = 0; p1 < np; ++p1) {
        PairHerm bra, ket;
        std::vector<double> W, buf, tmp;
        for (long p2 = 0; p2 <= p1; ++p2) {
            const int sa = pairs[p1].first, sb = pairs[p1].second, sc = pairs[p2].first,
                      sd = pairs[p2].second;
            const Shell &A = sh[sa], &B = sh[sb], &C = sh[sc], &D = sh[sd];
            const int ncA = n_cart(A.l()), ncB = n_cart(B.l()), ncC = n_cart(C.l()), ncD = n_cart(D.l());
            const long nab = (long)ncA * ncB, ncd = (long)ncC * ncD;
            buf.assign((size_t)(nab * ncd), 0.0);
            const auto &RA = A.position(), &RB = B.position(), &RC = C.position(), &RD = D.position();
            const int Lab = A.l() + B.l(), nmax = Lab + C.l() + D.l();
            for (const auto& pa : A.primitives())
                for (const auto& pb : B.primitives()) {
                    const double a = pa.exponent, b = pb.exponent, p = a + b;
                    Herm ex, ey, ez;
                    hermite(a, b, RA[0] - RB[0], A.l(), B.l(), ex);
                    hermite(a, b, RA[1] - RB[1], A.l(), B.l(), ey);
                    hermite(a, b, RA[2] - RB[2], A.l(), B.l(), ez);
                    pair_herm(A, B, ex, ey, ez, bra);
                    const double P[3] = {(a * RA[0] + b * RB[0]) / p, (a * RA[1] + b * RB[1]) / p,
                                         (a * RA[2] + b * RB[2]) / p};
                    for (const auto& pc : C.primitives())
                        for (const auto& pd : D.primitives()) {
                            const double c = pc.exponent, d = pd.exponent, q = c + d;
                            Herm fx, fy, fz;
                            hermite(c, d, RC[0] - RD[0], C.l(), D.l(), fx);
                            hermite(c, d, RC[1] - RD[1], C.l(), D.l(), fy);
                            hermite(c, d, RC[2] - RD[2], C.l(), D.l(), fz);
                            pair_herm(C, D, fx, fy, fz, ket);
                            const double Qc[3] = {(c * RC[0] + d * RD[0]) / q, (c * RC[1] + d * RD[1]) / q,
                                                  (c * RC[2] + d * RD[2]) / q};
                            const double alpha = p * q / (p + q);
                            const double pref = 2.0 * std::pow(M_PI, 2.5) / (p * q * std::sqrt(p + q)) *
                                                pa.coefficient * pb.coefficient * pc.coefficient *
                                                pd.coefficient;
                            RTab R;
                            hermite_coulomb(nmax, alpha, P[0] - Qc[0], P[1] - Qc[1], P[2] - Qc[2], R);
                            // W[cd][h] = sum_{ket Hermite k} (-1)^|k| E^cd_k R[h + k] for every bra
                            // Hermite function h, then buf[ab][cd] += pref sum_h E^ab_h W[cd][h]
                            W.assign((size_t)(ncd * bra.nh), 0.0);
                            for (long cd = 0; cd < ncd; ++cd) {
                                double* w = &W[(size_t)(cd * bra.nh)];
                                for (int e = ket.start[cd]; e < ket.start[cd + 1]; ++e) {
                                    const int k = ket.idx[e];
                                    const int tt = ket.tuv[3 * k], uu = ket.tuv[3 * k + 1], vv = ket.tuv[3 * k + 2];
                                    const double f = ((tt + uu + vv) & 1) ? -ket.coef[e] : ket.coef[e];
                                    for (int h = 0; h < bra.nh; ++h)
                                        w[h] += f * R.R[bra.tuv[3 * h] + tt][bra.tuv[3 * h + 1] + uu]
                                                       [bra.tuv[3 * h + 2] + vv];
                                }
                            }
                            for (long ab = 0; ab < nab; ++ab) {
                                double* out = &buf[(size_t)(ab * ncd)];
                                for (int e = bra.start[ab]; e < bra.start[ab + 1]; ++e) {
                                    const double g = pref * bra.coef[e];
                                    const double* w = &W[bra.idx[e]];
                                    for (long cd = 0; cd < ncd; ++cd) out[cd] += g * w[cd * bra.nh];
                                }
                            }
                        }
                }
            // Cartesian quartet -> shell functions, one index at a time
            const int na = A.n_functions(), nb = B.n_functions(), nc = C.n_functions(), nd = D.n_functions();
            tmp.resize(buf.size());
            to_shell_index(sph_rows(A), na, ncA, 1, (long)ncB * ncC * ncD, buf.data(), tmp.data());
            to_shell_index(sph_rows(B), nb, ncB, na, (long)ncC * ncD, tmp.data(), buf.data());
            to_shell_index(sph_rows(C), nc, ncC, (long)na * nb, ncD, buf.data(), tmp.data());
            to_shell_index(sph_rows(D), nd, ncD, (long)na * nb * nc, 1, tmp.data(), buf.data());
            for (int ia = 0; ia < na; ++ia)
                for (int ib = 0; ib < nb; ++ib)
                    for (int ic = 0; ic < nc; ++ic)
                        for (int id = 0; id < nd; ++id) {
                            const int i = offset_[sa] + ia, j = offset_[sb] + ib, k = offset_[sc] + ic,
                                      l = offset_[sd] + id;
                            const double v = buf[(((size_t)ia * nb + ib) * nc + ic) * nd + id];
                            eri(i, j, k, l) = v;
                            eri(j, i, k, l) = v;
                            eri(i, j, l, k) = v;
                            eri(j, i, l, k) = v;
                            eri(k, l, i, j) = v;
                            eri(l, k, i, j) = v;
                            eri(k, l, j, i) = v;
                            eri(l, k, j, i) = v;
                        }
        }
    }
    return eri;
}

// ==================================================================== DIIS

DIIS::DIIS(int max_vectors) : max_vectors_(max_vectors) {
    if (max_vectors < 2) throw std::invalid_argument("DIIS requires at least 2 vectors");
}

void DIIS::add_iteration(const Mat& F, const Mat& error) {
    focks_.push_back(F);
    errors_.push_back(error);
    if ((int)focks_.size() > max_vectors_) {
        focks_.erase(focks_.begin());
        errors_.erase(errors_.begin());
    }
}

Mat DIIS::extrapolate() const {
    if (!can_extrapolate()) throw std::runtime_error("DIIS: not enough vectors");
    const int n = (int)focks_.size();
    Mat A(n + 1, n + 1);
    for (int i = 0; i < n; ++i)
        for (int j = 0; j < n; ++j) A(i, j) = dot(errors_[i], errors_[j]);
    for (int i = 0; i < n; ++i) A(i, n) = A(n, i) = -1.0;
    Vec b(n + 1, 0.0);
    b[n] = -1.0;
    const Vec x = solve(A, b);
    Mat F(focks_[0].rows, focks_[0].cols);
    for (int i = 0; i < n; ++i)
        for (size_t k = 0; k < F.a.size(); ++k) F.a[k] += x[i] * focks_[i].a[k];
    return F;
}

void DIIS::clear() {
    focks_.clear();
    errors_.clear();
}

}  // namespace mshqc
