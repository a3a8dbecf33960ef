// Driver around the REFERENCE's own hot-path code (built from its sources by
// oracle/ref.mk into oracle/_ref/) -- TEST INFRASTRUCTURE ONLY.
//
// Reads one binary problem file and runs, with the reference's classes:
//   CholeskyERI::decompose (src/integrals/cholesky_eri.cc)
//   CholeskyRMP2::compute  (src/mp2/cholesky_rmp2.cc), reuse-vectors constructor
//   CholeskyUMP2::compute  (src/mp2/cholesky_ump2.cc)
//   CholeskyRMP3::compute  (src/mp3/cholesky_rmp3.cc), fed the RMP2 result
//   CholeskyUMP3::compute  (src/mp3/cholesky_ump3.cc), from the UMP2 solver
// and prints the results as text.  Input layout (little-endian):
//   int32 n, int32 n_occ_a, int32 n_occ_b; double threshold;
//   double eri[n^4] ((i*n+j)*n+k)*n+l;
//   double Ca[n*n], Cb[n*n] (row-major, columns = orbitals);
//   double eps_a[n], eps_b[n]
// Output lines: "nvec Q", "L <Q*n*n values>", "rmp2 e_corr", "t2 <o*o*v*v>",
//   "ump2 ss_aa ss_bb os", "rmp3 e_mp3", "t22 <o*o*v*v>", "ump3 aa bb os".
#include <cstdio>
#include <cstdint>
#include <iostream>
#include <memory>
#include <vector>

#include "mshqc/basis.h"
#include "mshqc/cholesky_rmp2.h"
#include "mshqc/cholesky_rmp3.h"
#include "mshqc/cholesky_ump3.h"
#include "mshqc/cholesky_ump2.h"
#include "mshqc/integrals/cholesky_eri.h"
#include "mshqc/molecule.h"
#include "mshqc/scf.h"

using namespace mshqc;

int main(int argc, char** argv) {
    if (argc < 2) {
        std::fprintf(stderr, "usage: ref_driver problem.bin\n");
        return 2;
    }
    FILE* f = std::fopen(argv[1], "rb");
    if (!f) return 3;
    int32_t hdr[3];
    double thr;
    if (std::fread(hdr, 4, 3, f) != 3 || std::fread(&thr, 8, 1, f) != 1) return 4;
    const int n = hdr[0], na = hdr[1], nb = hdr[2];
    std::vector<double> eri((size_t)n * n * n * n), Ca(n * n), Cb(n * n), ea(n), eb(n);
    if (std::fread(eri.data(), 8, eri.size(), f) != eri.size()) return 5;
    if (std::fread(Ca.data(), 8, Ca.size(), f) != Ca.size()) return 5;
    if (std::fread(Cb.data(), 8, Cb.size(), f) != Cb.size()) return 5;
    if (std::fread(ea.data(), 8, n, f) != (size_t)n || std::fread(eb.data(), 8, n, f) != (size_t)n) return 5;
    std::fclose(f);

    Eigen::Tensor<double, 4> T(n, n, n, n);
    for (int i = 0; i < n; ++i)
        for (int j = 0; j < n; ++j)
            for (int k = 0; k < n; ++k)
                for (int l = 0; l < n; ++l) T(i, j, k, l) = eri[(((size_t)i * n + j) * n + k) * n + l];

    // The reference prints its progress on stdout: keep it on stderr.
    std::streambuf* keep = std::cout.rdbuf(std::cerr.rdbuf());
    integrals::CholeskyERI chol(thr);
    chol.decompose(T);

    BasisSet basis;
    for (int k = 0; k < n; ++k) basis.add_shell(Shell(AngularMomentum::S, 0, {0.0, 0.0, 0.0}));
    Molecule mol;
    SCFResult scf;
    scf.energy_total = 0.0;
    scf.n_occ_alpha = na;
    scf.n_occ_beta = nb;
    scf.C_alpha = Eigen::MatrixXd(n, n);
    scf.C_beta = Eigen::MatrixXd(n, n);
    for (int r = 0; r < n; ++r)
        for (int c = 0; c < n; ++c) {
            scf.C_alpha(r, c) = Ca[r * n + c];
            scf.C_beta(r, c) = Cb[r * n + c];
        }
    scf.orbital_energies_alpha = Eigen::Map<Eigen::VectorXd>(ea.data(), n);
    scf.orbital_energies_beta = Eigen::Map<Eigen::VectorXd>(eb.data(), n);

    CholeskyRMP2Config rc;
    rc.print_level = 0;
    CholeskyRMP2 rmp2(mol, basis, nullptr, scf, rc, chol);
    auto rr = rmp2.compute();

    CholeskyUMP2Config uc;
    uc.print_level = 0;
    CholeskyUMP2 ump2(scf, basis, nullptr, uc, chol);
    auto ur = ump2.compute();

    foundation::CholeskyRMP3 rmp3(scf, rr, basis);
    auto r3 = rmp3.compute();

    CholeskyUMP3Config u3c;
    u3c.print_level = 0;
    CholeskyUMP3 ump3(ump2, u3c);
    auto u3 = ump3.compute();
    std::cout.rdbuf(keep);

    std::printf("nvec %d\n", chol.n_vectors());
    std::printf("L");
    for (const auto& v : chol.get_L_vectors())
        for (int k = 0; k < v.size(); ++k) std::printf(" %.17g", v(k));
    std::printf("\nrmp2 %.17g\n", rr.e_corr);
    const int o = na, v = n - na;
    std::printf("t2");
    for (int i = 0; i < o; ++i)
        for (int j = 0; j < o; ++j)
            for (int a = 0; a < v; ++a)
                for (int b = 0; b < v; ++b) std::printf(" %.17g", rr.t2(i, j, a, b));
    std::printf("\nump2 %.17g %.17g %.17g\n", ur.e_corr_ss_aa, ur.e_corr_ss_bb, ur.e_corr_os);
    std::printf("rmp3 %.17g\n", r3.e_mp3);
    std::printf("t22");
    for (int i = 0; i < o; ++i)
        for (int j = 0; j < o; ++j)
            for (int a = 0; a < v; ++a)
                for (int b = 0; b < v; ++b) std::printf(" %.17g", r3.t2_2(i, j, a, b));
    std::printf("\nump3 %.17g %.17g %.17g\n", u3.e_mp3_ss_aa, u3.e_mp3_ss_bb, u3.e_mp3_os);
    return 0;
}
