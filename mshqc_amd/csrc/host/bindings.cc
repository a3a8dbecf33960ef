// pybind11 module `_mshqc`: the Python surface of the Cholesky-vector path
// with the class names, constructor argument names and defaults of the
// reference bindings (python/bindings.cc:73-592).  Matrices cross as numpy
// arrays.  Additions for synthetic inputs: ERITensor(numpy), CholeskyERI.
// set_vectors / add_vector, precision and store_t2 options.
#include <pybind11/numpy.h>
#include <pybind11/pybind11.h>
#include <pybind11/stl.h>

#include "cd.h"

namespace py = pybind11;
using namespace mshqc;
using mshqc::integrals::CholeskyERI;

namespace {

using arr = py::array_t<double, py::array::c_style | py::array::forcecast>;

py::array_t<double> mat_to_np(const Mat& M) {
    py::array_t<double> a({(py::ssize_t)M.rows, (py::ssize_t)M.cols});
    std::copy(M.a.begin(), M.a.end(), a.mutable_data());
    return a;
}

Mat np_to_mat(const arr& a) {
    if (a.ndim() != 2) throw std::invalid_argument("expected a 2-D array");
    Mat M((int)a.shape(0), (int)a.shape(1));
    std::copy(a.data(), a.data() + M.size(), M.a.begin());
    return M;
}

// A flattened [d0][d1][d2][d3] amplitude block as a 4-D array (None when empty).
py::object t2_array(const Vec& t, int d0, int d1, int d2, int d3) {
    if (t.empty()) return py::none();
    py::array_t<double> a({d0, d1, d2, d3});
    std::copy(t.begin(), t.end(), a.mutable_data());
    return std::move(a);
}

py::array_t<double> vec_to_np(const Vec& v) {
    py::array_t<double> a((py::ssize_t)v.size());
    std::copy(v.begin(), v.end(), a.mutable_data());
    return a;
}

Vec np_to_vec(const arr& a) { return Vec(a.data(), a.data() + a.size()); }

std::vector<Vec> to_vectors(py::object obj, int& n_basis) {
    std::vector<Vec> out;
    if (py::isinstance<py::array>(obj)) {
        arr a = obj.cast<arr>();
        if (a.ndim() == 3) {
            if (a.shape(1) != a.shape(2)) throw std::invalid_argument("L must be [Q, n, n]");
            n_basis = (int)a.shape(1);
            const size_t nn = (size_t)n_basis * n_basis;
            for (py::ssize_t q = 0; q < a.shape(0); ++q)
                out.emplace_back(a.data() + q * nn, a.data() + (q + 1) * nn);
            return out;
        }
        if (a.ndim() != 2) throw std::invalid_argument("L must be [Q, n, n], [Q, n*n] or a list");
        const size_t len = (size_t)a.shape(1);
        for (py::ssize_t q = 0; q < a.shape(0); ++q) out.emplace_back(a.data() + q * len, a.data() + (q + 1) * len);
    } else {
        for (auto item : obj) {
            arr a = item.cast<arr>();
            out.emplace_back(a.data(), a.data() + a.size());
        }
    }
    if (!out.empty()) n_basis = (int)std::lround(std::sqrt((double)out[0].size()));
    return out;
}

template <class C, class Field>
void mat_prop(py::class_<C>& cls, const char* name, Field C::*f) {
    cls.def_property(
        name, [f](const C& self) { return mat_to_np(self.*f); },
        [f](C& self, const arr& a) { self.*f = np_to_mat(a); });
}

template <class C>
void vec_prop(py::class_<C>& cls, const char* name, Vec C::*f) {
    cls.def_property(
        name, [f](const C& self) { return vec_to_np(self.*f); },
        [f](C& self, const arr& a) { self.*f = np_to_vec(a); });
}

}  // namespace

PYBIND11_MODULE(_mshqc, m) {
    m.doc() = "MI355X-native Cholesky-vector path of mshqc (CD-RHF/UHF J/K, CD-RMP2/UMP2)";

    // ------------------------------------------------------------- ERI
    py::class_<ERITensor>(m, "ERITensor")
        .def(py::init([](const arr& a) {
                 if (a.ndim() != 4 || a.shape(0) != a.shape(1) || a.shape(0) != a.shape(2) ||
                     a.shape(0) != a.shape(3))
                     throw std::invalid_argument("ERITensor needs an [n, n, n, n] array");
                 ERITensor t((int)a.shape(0));
                 std::copy(a.data(), a.data() + t.size(), t.v.begin());
                 return t;
             }),
             py::arg("values"))
        .def("size", [](const ERITensor& t) { return (long)t.size(); })
        .def("dimension", [](const ERITensor& t, int) { return t.n; })
        .def("shape", [](const ERITensor& t) { return std::make_tuple(t.n, t.n, t.n, t.n); })
        .def("numpy",
             [](const ERITensor& t) {
                 py::array_t<double> a({t.n, t.n, t.n, t.n});
                 std::copy(t.v.begin(), t.v.end(), a.mutable_data());
                 return a;
             })
        .def("__repr__", [](const ERITensor& t) {
            const std::string n = std::to_string(t.n);
            return "<mshqc.ERITensor shape=(" + n + ", " + n + ", " + n + ", " + n + ")>";
        });

    // ----------------------------------------------- molecule and basis
    py::class_<Atom>(m, "Atom")
        .def(py::init<int, double, double, double>(), py::arg("atomic_number"), py::arg("x"), py::arg("y"),
             py::arg("z"))
        .def_readwrite("atomic_number", &Atom::atomic_number)
        .def_readwrite("x", &Atom::x)
        .def_readwrite("y", &Atom::y)
        .def_readwrite("z", &Atom::z)
        .def("position", &Atom::position);

    py::class_<Molecule>(m, "Molecule")
        .def(py::init<>())
        .def(py::init<int, int>(), py::arg("charge"), py::arg("multiplicity"))
        .def("add_atom", py::overload_cast<int, double, double, double>(&Molecule::add_atom), py::arg("Z"),
             py::arg("x"), py::arg("y"), py::arg("z"))
        .def("add_atom", py::overload_cast<const Atom&>(&Molecule::add_atom), py::arg("atom"))
        .def("n_atoms", &Molecule::n_atoms)
        .def("atom", &Molecule::atom, py::arg("i"))
        .def("atoms", &Molecule::atoms)
        .def("total_nuclear_charge", &Molecule::total_nuclear_charge)
        .def("n_electrons", &Molecule::n_electrons)
        .def("charge", &Molecule::charge)
        .def("set_charge", &Molecule::set_charge, py::arg("q"))
        .def("multiplicity", &Molecule::multiplicity)
        .def("set_multiplicity", &Molecule::set_multiplicity, py::arg("m"))
        .def("nuclear_repulsion_energy", &Molecule::nuclear_repulsion_energy);

    py::enum_<AngularMomentum>(m, "AngularMomentum")
        .value("S", AngularMomentum::S)
        .value("P", AngularMomentum::P)
        .value("D", AngularMomentum::D)
        .value("F", AngularMomentum::F)
        .value("G", AngularMomentum::G)
        .value("H", AngularMomentum::H);

    py::class_<GaussianPrimitive>(m, "GaussianPrimitive")
        .def(py::init<double, double>(), py::arg("exponent"), py::arg("coefficient"))
        .def_readwrite("exponent", &GaussianPrimitive::exponent)
        .def_readwrite("coefficient", &GaussianPrimitive::coefficient);

    py::class_<Shell>(m, "Shell")
        .def(py::init<AngularMomentum, int, const std::array<double, 3>&>(), py::arg("am"), py::arg("center"),
             py::arg("center_pos"))
        .def("add_primitive", &Shell::add_primitive, py::arg("exponent"), py::arg("coefficient"))
        .def("normalize", &Shell::normalize)
        .def("n_functions", &Shell::n_functions)
        .def_property_readonly("angular_momentum", &Shell::angular_momentum)
        .def_property_readonly("center", &Shell::center)
        .def_property_readonly("primitives", &Shell::primitives);

    py::class_<BasisSet>(m, "BasisSet")
        .def(py::init<>())
        .def(py::init<const std::string&, const Molecule&, const std::string&>(), py::arg("basis_name"),
             py::arg("mol"), py::arg("basis_dir") = "data/basis")
        .def("read_gbs", &BasisSet::read_gbs, py::arg("basis_file"), py::arg("mol"))
        .def("add_shell", &BasisSet::add_shell, py::arg("shell"))
        .def("n_shells", &BasisSet::n_shells)
        .def("n_basis_functions", &BasisSet::n_basis_functions)
        .def("shell", &BasisSet::shell, py::arg("i"))
        .def("name", &BasisSet::name);

    py::class_<IntegralEngine, std::shared_ptr<IntegralEngine>>(m, "IntegralEngine")
        .def(py::init<const Molecule&, const BasisSet&>(), py::arg("mol"), py::arg("basis"))
        .def("compute_overlap", [](const IntegralEngine& e) { return mat_to_np(e.compute_overlap()); })
        .def("compute_kinetic", [](const IntegralEngine& e) { return mat_to_np(e.compute_kinetic()); })
        .def("compute_nuclear", [](const IntegralEngine& e) { return mat_to_np(e.compute_nuclear()); })
        .def("compute_eri", &IntegralEngine::compute_eri)
        .def("compute_eri_diagonal", [](const IntegralEngine& e) { return vec_to_np(e.compute_eri_diagonal()); },
             "Compute diagonal ERI elements (pq|pq)")
        .def("compute_eri_column",
             [](const IntegralEngine& e, int pivot_index) { return vec_to_np(e.compute_eri_column(pivot_index)); },
             "Compute specific ERI column (pq|rs) for a pivot index", py::arg("pivot_index"));

    // ---------------------------------------------------- CholeskyERI
    py::class_<CholeskyDecompositionResult>(m, "CholeskyDecompositionResult")
        .def_readonly("n_vectors", &CholeskyDecompositionResult::n_vectors)
        .def_readonly("n_basis", &CholeskyDecompositionResult::n_basis)
        .def_readonly("threshold", &CholeskyDecompositionResult::threshold)
        .def_readonly("compression_ratio", &CholeskyDecompositionResult::compression_ratio)
        .def_readonly("converged", &CholeskyDecompositionResult::converged);

    py::class_<CholeskyERI>(m, "CholeskyERI")
        .def(py::init<double>(), py::arg("threshold") = 1e-6)
        .def("decompose", &CholeskyERI::decompose)
        .def("decompose_gpu", &CholeskyERI::decompose_gpu)
        .def("get_L_vectors",
             [](const CholeskyERI& c) {
                 py::list out;
                 for (const auto& v : c.get_L_vectors()) out.append(vec_to_np(v));
                 return out;
             })
        .def("reconstruct", &CholeskyERI::reconstruct)
        .def("n_vectors", &CholeskyERI::n_vectors)
        .def("n_basis", &CholeskyERI::n_basis)
        .def("threshold", &CholeskyERI::threshold)
        .def("compression_ratio", &CholeskyERI::compression_ratio)
        .def(
            "set_vectors",
            [](CholeskyERI& c, py::object L, int n_basis) {
                int nb = n_basis;
                auto v = to_vectors(L, nb);
                c.set_vectors(v, n_basis > 0 ? n_basis : nb);
            },
            py::arg("vectors"), py::arg("n_basis") = 0)
        .def("add_vector", [](CholeskyERI& c, const arr& v) { c.add_vector(np_to_vec(v)); }, py::arg("vector"));

    // ------------------------------------------------------------ SCF
    py::class_<SCFConfig>(m, "SCFConfig")
        .def(py::init<>())
        .def_readwrite("max_iterations", &SCFConfig::max_iterations)
        .def_readwrite("energy_threshold", &SCFConfig::energy_threshold)
        .def_readwrite("density_threshold", &SCFConfig::density_threshold)
        .def_readwrite("diis_threshold", &SCFConfig::diis_threshold)
        .def_readwrite("diis_max_vectors", &SCFConfig::diis_max_vectors)
        .def_readwrite("print_level", &SCFConfig::print_level)
        .def_readwrite("level_shift", &SCFConfig::level_shift);

    py::class_<SCFResult> scf(m, "SCFResult");
    scf.def(py::init<>())
        .def_readwrite("energy_electronic", &SCFResult::energy_electronic)
        .def_readwrite("energy_nuclear", &SCFResult::energy_nuclear)
        .def_readwrite("energy_total", &SCFResult::energy_total)
        .def_readwrite("iterations", &SCFResult::iterations)
        .def_readwrite("converged", &SCFResult::converged)
        .def_readwrite("gradient_norm", &SCFResult::gradient_norm)
        .def_readwrite("n_occ_alpha", &SCFResult::n_occ_alpha)
        .def_readwrite("n_occ_beta", &SCFResult::n_occ_beta);
    vec_prop(scf, "orbital_energies_alpha", &SCFResult::orbital_energies_alpha);
    vec_prop(scf, "orbital_energies_beta", &SCFResult::orbital_energies_beta);
    mat_prop(scf, "C_alpha", &SCFResult::C_alpha);
    mat_prop(scf, "C_beta", &SCFResult::C_beta);
    mat_prop(scf, "P_alpha", &SCFResult::P_alpha);
    mat_prop(scf, "P_beta", &SCFResult::P_beta);
    mat_prop(scf, "F_alpha", &SCFResult::F_alpha);
    mat_prop(scf, "F_beta", &SCFResult::F_beta);

    py::class_<CholeskyRHFConfig, SCFConfig>(m, "CholeskyRHFConfig")
        .def(py::init<>())
        .def_readwrite("cholesky_threshold", &CholeskyRHFConfig::cholesky_threshold)
        .def_readwrite("screen_exchange", &CholeskyRHFConfig::screen_exchange);

    py::class_<CholeskyRHF>(m, "CholeskyRHF")
        .def(py::init<const Molecule&, const BasisSet&, std::shared_ptr<IntegralEngine>,
                      const CholeskyRHFConfig&>(),
             py::arg("molecule"), py::arg("basis"), py::arg("integrals"), py::arg("config") = CholeskyRHFConfig())
        .def(py::init<const Molecule&, const BasisSet&, std::shared_ptr<IntegralEngine>, const CholeskyRHFConfig&,
                      const CholeskyERI&>(),
             py::arg("molecule"), py::arg("basis"), py::arg("integrals"), py::arg("config"),
             py::arg("existing_cholesky"), py::keep_alive<1, 6>())
        .def("compute", &CholeskyRHF::compute)
        .def("energy", &CholeskyRHF::energy);

    py::class_<CholeskyUHFConfig>(m, "CholeskyUHFConfig")
        .def(py::init<>())
        .def_readwrite("cholesky_threshold", &CholeskyUHFConfig::cholesky_threshold)
        .def_readwrite("print_level", &CholeskyUHFConfig::print_level)
        .def_readwrite("max_iterations", &CholeskyUHFConfig::max_iterations)
        .def_readwrite("energy_threshold", &CholeskyUHFConfig::energy_threshold)
        .def_readwrite("diis_max_vectors", &CholeskyUHFConfig::diis_max_vectors);

    py::class_<CholeskyUHF>(m, "CholeskyUHF")
        .def(py::init<const Molecule&, const BasisSet&, std::shared_ptr<IntegralEngine>, int, int,
                      const CholeskyUHFConfig&>(),
             py::arg("molecule"), py::arg("basis"), py::arg("integrals"), py::arg("n_alpha"), py::arg("n_beta"),
             py::arg("config") = CholeskyUHFConfig())
        .def("compute", &CholeskyUHF::compute)
        .def(
            "set_cholesky_vectors",
            [](CholeskyUHF& u, py::object L) {
                int nb = 0;
                u.set_cholesky_vectors(to_vectors(L, nb));
            },
            py::arg("vectors"))
        .def("compute_s_squared", &CholeskyUHF::compute_s_squared);

    // ---------------------------------------------------------- ROHF
    // bindings.cc:249-275 of the reference
    py::class_<CholeskyROHFConfig, SCFConfig>(m, "CholeskyROHFConfig")
        .def(py::init<>())
        .def_readwrite("cholesky_threshold", &CholeskyROHFConfig::cholesky_threshold)
        .def_readwrite("screen_exchange", &CholeskyROHFConfig::screen_exchange);

    py::class_<CholeskyROHF>(m, "CholeskyROHF")
        .def(py::init<const Molecule&, const BasisSet&, std::shared_ptr<IntegralEngine>, int, int,
                      const CholeskyROHFConfig&>(),
             py::arg("molecule"), py::arg("basis"), py::arg("integrals"), py::arg("n_alpha"), py::arg("n_beta"),
             py::arg("config") = CholeskyROHFConfig())
        .def(py::init<const Molecule&, const BasisSet&, std::shared_ptr<IntegralEngine>, int, int,
                      const CholeskyROHFConfig&, const integrals::CholeskyERI&>(),
             py::arg("molecule"), py::arg("basis"), py::arg("integrals"), py::arg("n_alpha"), py::arg("n_beta"),
             py::arg("config"), py::arg("existing_cholesky"))
        .def("compute", &CholeskyROHF::compute);

    // ---------------------------------------------------------- UMP3
    // (no Python binding in the reference; names follow cholesky_ump3.h)
    py::class_<CholeskyUMP3Config>(m, "CholeskyUMP3Config")
        .def(py::init<>())
        .def_readwrite("cholesky_threshold", &CholeskyUMP3Config::cholesky_threshold)
        .def_readwrite("use_ump2_result", &CholeskyUMP3Config::use_ump2_result)
        .def_readwrite("store_intermediates", &CholeskyUMP3Config::store_intermediates)
        .def_readwrite("validate_energy", &CholeskyUMP3Config::validate_energy)
        .def_readwrite("print_level", &CholeskyUMP3Config::print_level);
    py::class_<CholeskyUMP3Result>(m, "CholeskyUMP3Result")
        .def_readonly("e_mp2_total", &CholeskyUMP3Result::e_mp2_total)
        .def_readonly("e_mp3_ss_aa", &CholeskyUMP3Result::e_mp3_ss_aa)
        .def_readonly("e_mp3_ss_bb", &CholeskyUMP3Result::e_mp3_ss_bb)
        .def_readonly("e_mp3_os", &CholeskyUMP3Result::e_mp3_os)
        .def_readonly("e_mp3_total", &CholeskyUMP3Result::e_mp3_total)
        .def_readonly("e_corr_total", &CholeskyUMP3Result::e_corr_total)
        .def_readonly("e_total", &CholeskyUMP3Result::e_total)
        .def_readonly("n_cholesky_vectors", &CholeskyUMP3Result::n_cholesky_vectors)
        .def_readonly("time_mp3_s", &CholeskyUMP3Result::time_mp3_s);
    py::class_<CholeskyUMP3>(m, "CholeskyUMP3")
        .def(py::init<const SCFResult&, const BasisSet&, std::shared_ptr<IntegralEngine>, const CholeskyUMP3Config&>(),
             py::arg("uhf_result"), py::arg("basis"), py::arg("integrals"), py::arg("config") = CholeskyUMP3Config())
        .def(py::init<const CholeskyUMP2&, const CholeskyUMP3Config&>(), py::arg("ump2_solver"),
             py::arg("config") = CholeskyUMP3Config())
        .def("compute", &CholeskyUMP3::compute)
        .def("initialize_cholesky", &CholeskyUMP3::initialize_cholesky)
        .def("transform_cholesky_vectors", &CholeskyUMP3::transform_cholesky_vectors);

    // ---------------------------------------------------------- OMP2
    // bindings.cc:311-331 (OMP2Result) and 495-519 of the reference
    py::class_<OMP2Result> omp2r(m, "OMP2Result");
    omp2r.def(py::init<>())
        .def_readwrite("energy_scf", &OMP2Result::energy_scf)
        .def_readwrite("energy_mp2_ss", &OMP2Result::energy_mp2_ss)
        .def_readwrite("energy_mp2_os", &OMP2Result::energy_mp2_os)
        .def_readwrite("energy_mp2_corr", &OMP2Result::energy_mp2_corr)
        .def_readwrite("energy_total", &OMP2Result::energy_total)
        .def_readwrite("converged", &OMP2Result::converged)
        .def_readwrite("iterations", &OMP2Result::iterations)
        .def_readwrite("n_occ_alpha", &OMP2Result::n_occ_alpha)
        .def_readwrite("n_occ_beta", &OMP2Result::n_occ_beta)
        .def_readwrite("n_virt_alpha", &OMP2Result::n_virt_alpha)
        .def_readwrite("n_virt_beta", &OMP2Result::n_virt_beta);
    mat_prop(omp2r, "C_alpha", &OMP2Result::C_alpha);
    mat_prop(omp2r, "C_beta", &OMP2Result::C_beta);
    vec_prop(omp2r, "orbital_energies_alpha", &OMP2Result::orbital_energies_alpha);
    vec_prop(omp2r, "orbital_energies_beta", &OMP2Result::orbital_energies_beta);

    py::class_<CholeskyOMP2Config>(m, "CholeskyOMP2Config")
        .def(py::init<>())
        .def_readwrite("max_iterations", &CholeskyOMP2Config::max_iterations)
        .def_readwrite("energy_threshold", &CholeskyOMP2Config::energy_threshold)
        .def_readwrite("gradient_threshold", &CholeskyOMP2Config::gradient_threshold)
        .def_readwrite("cholesky_threshold", &CholeskyOMP2Config::cholesky_threshold)
        .def_readwrite("print_level", &CholeskyOMP2Config::print_level)
        .def_readwrite("two_particle_gradient", &CholeskyOMP2Config::two_particle_gradient);

    py::class_<CholeskyOMP2>(m, "CholeskyOMP2")
        .def(py::init<const Molecule&, const BasisSet&, std::shared_ptr<IntegralEngine>, const SCFResult&,
                      const CholeskyOMP2Config&>(),
             py::arg("molecule"), py::arg("basis"), py::arg("integrals"), py::arg("scf_guess"),
             py::arg("config") = CholeskyOMP2Config())
        .def(py::init<const Molecule&, const BasisSet&, std::shared_ptr<IntegralEngine>, const SCFResult&,
                      const CholeskyOMP2Config&, const CholeskyERI&>(),
             py::arg("molecule"), py::arg("basis"), py::arg("integrals"), py::arg("scf_guess"), py::arg("config"),
             py::arg("existing_cholesky"), py::keep_alive<1, 7>())
        .def("compute", &CholeskyOMP2::compute);

    // ----------------------------------------------------------- MP2
    py::class_<CholeskyRMP2Config>(m, "CholeskyRMP2Config")
        .def(py::init<>())
        .def_readwrite("cholesky_threshold", &CholeskyRMP2Config::cholesky_threshold)
        .def_readwrite("print_level", &CholeskyRMP2Config::print_level)
        .def_readwrite("precision", &CholeskyRMP2Config::precision)
        .def_readwrite("store_t2", &CholeskyRMP2Config::store_t2);

    py::class_<foundation::CholeskyRMP2Result>(m, "CholeskyRMP2Result")
        .def_readonly("e_rhf", &foundation::CholeskyRMP2Result::e_rhf)
        .def_readonly("e_corr", &foundation::CholeskyRMP2Result::e_corr)
        .def_readonly("e_total", &foundation::CholeskyRMP2Result::e_total)
        .def_readonly("n_chol_vectors", &foundation::CholeskyRMP2Result::n_chol_vectors)
        .def_readonly("n_occ", &foundation::CholeskyRMP2Result::n_occ)
        .def_readonly("n_virt", &foundation::CholeskyRMP2Result::n_virt)
        .def_readonly("time_transform_s", &foundation::CholeskyRMP2Result::time_transform_s)
        .def_readonly("time_energy_s", &foundation::CholeskyRMP2Result::time_energy_s)
        .def_property_readonly("t2", [](const foundation::CholeskyRMP2Result& r) -> py::object {
            if (r.t2.empty()) return py::none();
            py::array_t<double> a({r.n_occ, r.n_occ, r.n_virt, r.n_virt});
            std::copy(r.t2.begin(), r.t2.end(), a.mutable_data());
            return std::move(a);
        });

    // bindings.cc:395-399 (RMP3Result) and 594-600 (CholeskyRMP3) of the reference
    py::class_<foundation::RMP3Result>(m, "RMP3Result")
        .def(py::init<>())
        .def_readwrite("e_rhf", &foundation::RMP3Result::e_rhf)
        .def_readwrite("e_mp2", &foundation::RMP3Result::e_mp2)
        .def_readwrite("e_mp3", &foundation::RMP3Result::e_mp3)
        .def_readwrite("e_corr_total", &foundation::RMP3Result::e_corr_total)
        .def_readwrite("e_total", &foundation::RMP3Result::e_total)
        .def_readonly("n_occ", &foundation::RMP3Result::n_occ)
        .def_readonly("n_virt", &foundation::RMP3Result::n_virt)
        .def_property_readonly("t2_2", [](const foundation::RMP3Result& r) -> py::object {
            if (r.t2_2.empty()) return py::none();
            py::array_t<double> a({r.n_occ, r.n_occ, r.n_virt, r.n_virt});
            std::copy(r.t2_2.begin(), r.t2_2.end(), a.mutable_data());
            return std::move(a);
        });

    py::class_<foundation::CholeskyRMP3>(m, "CholeskyRMP3")
        .def(py::init<const SCFResult&, const foundation::CholeskyRMP2Result&, const BasisSet&>(),
             py::arg("rhf_result"), py::arg("crmp2_result"), py::arg("basis"))
        .def("compute", &foundation::CholeskyRMP3::compute);

    py::class_<CholeskyRMP2>(m, "CholeskyRMP2")
        .def(py::init<const Molecule&, const BasisSet&, std::shared_ptr<IntegralEngine>, const SCFResult&,
                      const CholeskyRMP2Config&>(),
             py::arg("molecule"), py::arg("basis"), py::arg("integrals"), py::arg("rhf_result"),
             py::arg("config") = CholeskyRMP2Config())
        .def(py::init<const Molecule&, const BasisSet&, std::shared_ptr<IntegralEngine>, const SCFResult&,
                      const CholeskyRMP2Config&, const CholeskyERI&>(),
             py::arg("molecule"), py::arg("basis"), py::arg("integrals"), py::arg("rhf_result"), py::arg("config"),
             py::arg("existing_cholesky"), py::keep_alive<1, 7>())
        .def("compute", &CholeskyRMP2::compute);

    py::class_<CholeskyUMP2Config>(m, "CholeskyUMP2Config")
        .def(py::init<>())
        .def_readwrite("cholesky_threshold", &CholeskyUMP2Config::cholesky_threshold)
        .def_readwrite("use_on_the_fly", &CholeskyUMP2Config::use_on_the_fly)
        .def_readwrite("validate_energy", &CholeskyUMP2Config::validate_energy)
        .def_readwrite("print_level", &CholeskyUMP2Config::print_level)
        .def_readwrite("precision", &CholeskyUMP2Config::precision);

    py::class_<CholeskyUMP2Result>(m, "CholeskyUMP2Result")
        .def_readonly("e_corr_ss_aa", &CholeskyUMP2Result::e_corr_ss_aa)
        .def_readonly("e_corr_ss_bb", &CholeskyUMP2Result::e_corr_ss_bb)
        .def_readonly("e_corr_os", &CholeskyUMP2Result::e_corr_os)
        .def_readonly("e_corr_total", &CholeskyUMP2Result::e_corr_total)
        .def_readonly("e_total", &CholeskyUMP2Result::e_total)
        .def_readonly("n_cholesky_vectors", &CholeskyUMP2Result::n_cholesky_vectors)
        .def_readonly("compression_ratio", &CholeskyUMP2Result::compression_ratio)
        .def_readonly("memory_mb", &CholeskyUMP2Result::memory_mb)
        .def_readonly("time_cholesky_s", &CholeskyUMP2Result::time_cholesky_s)
        .def_readonly("time_transform_s", &CholeskyUMP2Result::time_transform_s)
        .def_readonly("time_energy_s", &CholeskyUMP2Result::time_energy_s);

    py::class_<CholeskyUMP2>(m, "CholeskyUMP2")
        .def(py::init<const SCFResult&, const BasisSet&, std::shared_ptr<IntegralEngine>,
                      const CholeskyUMP2Config&, const CholeskyERI&>())
        .def("compute", &CholeskyUMP2::compute, "Compute UMP2 energy using Cholesky vectors")
        .def("compute_t2_amplitudes", &CholeskyUMP2::compute_t2_amplitudes)
        .def("get_cholesky", &CholeskyUMP2::get_cholesky, py::return_value_policy::reference_internal)
        // additions: the stored amplitudes (None before compute_t2_amplitudes)
        .def_property_readonly("t2_aa", [](const CholeskyUMP2& u) -> py::object {
            return t2_array(u.t2_aa(), u.n_occ_alpha(), u.n_occ_alpha(), u.n_virt_alpha(), u.n_virt_alpha());
        })
        .def_property_readonly("t2_bb", [](const CholeskyUMP2& u) -> py::object {
            return t2_array(u.t2_bb(), u.n_occ_beta(), u.n_occ_beta(), u.n_virt_beta(), u.n_virt_beta());
        })
        .def_property_readonly("t2_ab", [](const CholeskyUMP2& u) -> py::object {
            return t2_array(u.t2_ab(), u.n_occ_alpha(), u.n_occ_beta(), u.n_virt_alpha(), u.n_virt_beta());
        });

    m.def("linalg_eigh", [](const arr& A) {
        Vec w;
        Mat V;
        eigh(np_to_mat(A), w, V);
        return py::make_tuple(vec_to_np(w), mat_to_np(V));
    });
    m.def("linalg_expm", [](const arr& A) { return mat_to_np(expm(np_to_mat(A))); });
    m.def("linalg_solve", [](const arr& A, const arr& b) {
        const Mat M = np_to_mat(A);
        auto bb = b.unchecked<1>();
        if (M.rows != M.cols || bb.shape(0) != M.rows) throw std::invalid_argument("linalg_solve: shape mismatch");
        Vec v(M.rows);
        for (int i = 0; i < M.rows; ++i) v[i] = bb(i);
        return vec_to_np(solve(M, v));
    });
}
