// fm_engine.hpp -- C++ drop-in for the engine's assembler and solver entry
// points, running on one MI355X through the C ABI of fm_api.h.
//
// The functions keep the names, argument meaning and error behaviour of the
// reference (namespace atlas::fem, FEM/adaptive_fem_engine.hpp and
// FEM/nonlinear_solver.hpp) so host code can call atlas::fem::gpu::X beside
// the CPU path atlas::fem::X.  They are templates over the mesh, problem and
// CSR types and only use the members the reference declares:
//   Mesh : nodes[i].x/.y/.z, elements[e].nodes[0..3], n_dofs         (FEM/fem_types.hpp:58-205)
//   Prob : E, nu                                                      (FEM/fem_types.hpp:241-256)
//   CSR  : n_rows, row_ptr, col_idx, val (std::vector)                (FEM/fem_types.hpp:304-320)
// so the reference's own MeshTopology / ProblemDescriptor / SparseCSR work
// unchanged.  Header-only; link with libflexmesh.so.
#pragma once

#include <cmath>
#include <cstdint>
#include <cstring>
#include <stdexcept>
#include <string>
#include <tuple>
#include <vector>

#include "flexmesh/fm_api.h"

namespace atlas::fem::gpu {

// One device context per thread (device 0 unless select_device() is called
// first).  The mesh is recognised by content: a hash of the connectivity and
// one of the coordinates.  A new topology is uploaded with fm_set_mesh (new
// pattern, colouring and schedules); new coordinates on the same topology
// (nodes moved in place, as the adapt loop's smoothing does) with
// fm_set_coords, which keeps them.
class Device {
public:
    static Device& get() {
        thread_local Device d;
        return d;
    }
    static void select_device(int dev) { get().device_ = dev; }

    fm_ctx* ctx() {
        if (!ctx_) {
            const int rc = fm_create(device_, &ctx_);
            if (rc != FM_OK) throw std::runtime_error("flexmesh: fm_create failed (" + std::to_string(rc) + ")");
        }
        return ctx_;
    }

    template <class Mesh>
    void upload(const Mesh& mesh) {
        const size_t nn = mesh.nodes.size(), ne = mesh.elements.size();
        std::vector<uint32_t> conn(4 * ne);
        for (size_t e = 0; e < ne; ++e)
            for (int a = 0; a < 4; ++a) conn[a * ne + e] = (uint32_t)mesh.elements[e].nodes[a];
        std::vector<double> X(3 * nn);
        for (size_t n = 0; n < nn; ++n) {
            X[n] = mesh.nodes[n].x;
            X[nn + n] = mesh.nodes[n].y;
            X[2 * nn + n] = mesh.nodes[n].z;
        }
        const uint64_t ht = hash(conn.data(), conn.size() * sizeof(uint32_t), nn * 0x9E3779B97F4A7C15ull + ne);
        const uint64_t hx = hash(X.data(), X.size() * sizeof(double), 0x243F6A8885A308D3ull);
        if (have_mesh_ && ht == topo_hash_ && nn == nn_ && ne == ne_) {
            if (hx != coord_hash_) {
                check(fm_set_coords(ctx(), (uint32_t)nn, X.data()), "fm_set_coords");
                coord_hash_ = hx;
            }
            return;
        }
        check(fm_set_mesh(ctx(), (uint32_t)nn, X.data(), (uint32_t)ne, conn.data()), "fm_set_mesh");
        have_mesh_ = true;
        topo_hash_ = ht;
        coord_hash_ = hx;
        nn_ = nn;
        ne_ = ne;
        E_ = nu_ = std::nan("");
    }

    // one (E, nu) per problem (ProblemDescriptor.E/.nu): the Lame parameters of
    // NeoHookeanParameters::mu/lambda (FEM/constitutive_models.hpp:35-41)
    void material(double E, double nu) {
        if (E == E_ && nu == nu_) return;
        const double mu = E / (2.0 * (1.0 + nu));
        const double lam = E * nu / ((1.0 + nu) * (1.0 - 2.0 * nu));
        std::vector<double> m(ne_, mu), l(ne_, lam);
        check(fm_set_material(ctx(), m.data(), l.data()), "fm_set_material");
        E_ = E;
        nu_ = nu;
    }

    void check(int rc, const char* what) {
        if (rc != FM_OK)
            throw std::runtime_error(std::string("flexmesh: ") + what + ": " + fm_last_error(ctx_));
    }

    ~Device() {
        if (ctx_) fm_destroy(ctx_);
    }

private:
    Device() = default;
    // 64-bit multiply-xorshift hash over 8-byte words (and the tail bytes)
    static uint64_t hash(const void* data, size_t bytes, uint64_t h) {
        const unsigned char* p = static_cast<const unsigned char*>(data);
        size_t i = 0;
        for (; i + 8 <= bytes; i += 8) {
            uint64_t w;
            std::memcpy(&w, p + i, 8);
            h = (h ^ w) * 0xBF58476D1CE4E5B9ull;
            h ^= h >> 31;
        }
        for (; i < bytes; ++i) h = (h ^ p[i]) * 0x94D049BB133111EBull;
        return h ^ (h >> 29);
    }
    fm_ctx* ctx_ = nullptr;
    int device_ = 0;
    bool have_mesh_ = false;
    uint64_t topo_hash_ = 0, coord_hash_ = 0;
    size_t nn_ = 0, ne_ = 0;
    double E_ = std::nan(""), nu_ = std::nan("");
};

// build_csr_sparsity_pattern_rigorous (FEM/adaptive_fem_engine.hpp:1069-1120),
// with the intended column rule (row 3n+d lists 3m..3m+2 for sorted m in nbr(n)).
template <class Mesh, class CSR>
void build_csr_sparsity_pattern_rigorous(const Mesh& mesh, CSR& K) {
    Device& d = Device::get();
    d.upload(mesh);
    uint64_t nnz = 0;
    d.check(fm_build_pattern(d.ctx(), nullptr, nullptr, &nnz), "fm_build_pattern");
    K.n_rows = mesh.n_dofs;
    K.row_ptr.assign(K.n_rows + 1, 0);
    K.col_idx.assign(nnz, 0);
    K.val.assign(nnz, 0.0);
    d.check(fm_build_pattern(d.ctx(), K.row_ptr.data(), K.col_idx.data(), &nnz), "fm_build_pattern");
}

// assemble_global_threadlocal (FEM/adaptive_fem_engine.hpp:1126-1233): K.val
// overwritten on K's (pre-built) pattern, f_int resized and overwritten.
template <class Mesh, class Prob, class CSR>
void assemble_global_threadlocal(const Mesh& mesh, const Prob& prob, const double* u, CSR& K,
                                 std::vector<double>& f_int) {
    Device& d = Device::get();
    d.upload(mesh);
    d.material(prob.E, prob.nu);
    uint64_t nnz = 0;
    d.check(fm_build_pattern(d.ctx(), nullptr, nullptr, &nnz), "fm_build_pattern");
    if (K.col_idx.size() != nnz || K.row_ptr.size() != (size_t)mesh.n_dofs + 1)
        throw std::runtime_error("flexmesh: K does not hold the pattern of build_csr_sparsity_pattern_rigorous");
    K.val.resize(nnz);
    f_int.assign(mesh.n_dofs, 0.0);
    d.check(fm_assemble(d.ctx(), u, f_int.data(), K.val.data(), FM_FP64), "fm_assemble");
}

// assemble_global_residual_only (FEM/adaptive_fem_engine.hpp:990-1059): f_int
// resized and overwritten; no boundary condition or other state is touched.
template <class Mesh, class Prob>
void assemble_global_residual_only(const Mesh& mesh, const Prob& prob, const double* u,
                                   std::vector<double>& f_int) noexcept {
    try {
        Device& d = Device::get();
        d.upload(mesh);
        d.material(prob.E, prob.nu);
        f_int.assign(mesh.n_dofs, 0.0);
        d.check(fm_assemble_residual_only(d.ctx(), u, f_int.data()), "fm_assemble_residual_only");
    } catch (...) {
        f_int.assign(mesh.n_dofs, std::nan(""));
    }
}

// pcg_solve_ilu0 (FEM/adaptive_fem_engine.hpp:1283-1341): factorizes A, runs
// PCG from x with the stopping rule sqrt(r.z / b.b) < tol; returns the
// iteration count (max_iter = not converged, 0 = trivially converged), -1 on a
// device error (the reference is noexcept).  A call with the pattern of the
// previous call uploads the values only (the sort and the sweep schedules are
// kept by fm_set_matrix).
template <class CSR>
[[nodiscard]] int pcg_solve_ilu0(const CSR& A, const double* b, double* x, uint32_t N, double tol = 1e-10,
                                 int max_iter = 1000) noexcept {
    try {
        Device& d = Device::get();
        if (N != A.n_rows) return -1;
        d.check(fm_set_matrix(d.ctx(), A.n_rows, A.row_ptr.data(), A.col_idx.data(), A.val.data()), "fm_set_matrix");
        int iters = 0;
        double rel = 0.0;
        d.check(fm_pcg(d.ctx(), FM_FP64, b, x, tol, max_iter, &iters, &rel), "fm_pcg");
        return iters;
    } catch (...) {
        return -1;
    }
}

// MeshColoring (FEM/adaptive_fem_engine.hpp:1549-1602)
struct MeshColoring {
    std::vector<uint32_t> element_colors;
    uint32_t n_colors{1};

    template <class Mesh>
    static MeshColoring compute(const Mesh& mesh) {
        Device& d = Device::get();
        d.upload(mesh);
        MeshColoring c;
        c.element_colors.assign(mesh.elements.size(), 0);
        d.check(fm_color(d.ctx(), c.element_colors.data(), &c.n_colors), "fm_color");
        return c;
    }
};

// StressExtractor (FEM/stress_extraction.hpp:31-241) over fm_element_fields:
// extract_element_fields fills the reference's ElementFieldData in element
// order from each element's exact P1 gradient F (the reference's
// compute_deformation_gradient_at_element uses constant shape gradients that
// ignore the geometry).  The strain energy is the energy whose derivative is
// the stress (lam/2 ln^2 J); the error indicators come from the caller's
// estimator, as in the reference (:205-206, :237-240).
template <class Material>
class StressExtractor {
public:
    struct ElementFieldData {
        std::vector<double> von_mises, pressure, max_principal_stress, strain_energy, max_stretch, det_F,
            error_indicators;
    };
    explicit StressExtractor(const Material& mat) : material(mat) {}
    Material material;

    template <class Mesh, class Solution, class Estimator>
    ElementFieldData extract_element_fields(const Mesh& mesh, const Solution& u, const Estimator& estimator) const {
        Device& d = Device::get();
        d.upload(mesh);
        d.material(material.params.E, material.params.nu);
        if (u.size() != 3 * mesh.nodes.size()) throw std::runtime_error("flexmesh: u is not 3 * n_nodes long");
        const size_t ne = mesh.elements.size();
        ElementFieldData data;
        for (auto* v : {&data.von_mises, &data.pressure, &data.max_principal_stress, &data.strain_energy,
                        &data.max_stretch, &data.det_F, &data.error_indicators})
            v->assign(ne, 0.0);
        fm_element_fields_out o{};
        o.det_F = data.det_F.data();
        o.von_mises = data.von_mises.data();
        o.pressure = data.pressure.data();
        o.max_principal = data.max_principal_stress.data();
        o.energy = data.strain_energy.data();
        o.max_stretch = data.max_stretch.data();
        d.check(fm_element_fields(d.ctx(), u.data(), &o), "fm_element_fields");
        std::vector<double> elem_errors;
        estimator.estimate(elem_errors);
        for (size_t e = 0; e < ne && e < elem_errors.size(); ++e) data.error_indicators[e] = elem_errors[e];
        return data;
    }
};

// ZZErrorEstimator (FEM/error_estimator.hpp:106-239) over fm_zz_estimate.  The
// reference reads the element states of the mesh; the GPU twin takes the
// displacement they come from (the states' F is recomputed exactly, without
// the det F = 1 projection).  estimate() returns the global estimate and fills
// the element indicators; estimate_with_bounds() adds the reference's bounds.
template <class Mesh>
class ZZErrorEstimator {
public:
    static constexpr double C_EFFICIENCY_LINEAR_ELEMENTS = 1.0;
    static constexpr double C_RELIABILITY_LINEAR_ELEMENTS = 2.0;
    struct EstimateResult {
        double global_estimate{0.0};
        double lower_bound{0.0};
        double upper_bound{0.0};
        double effective_mean_ratio{0.0};
        bool meets_reliability{false};
        std::vector<double> elem_errors;
    };
    ZZErrorEstimator(const Mesh& mesh, const std::vector<double>& u) : mesh_(mesh), u_(u) {}

    EstimateResult estimate_with_bounds(std::vector<double>& elem_errors) const {
        Device& d = Device::get();
        d.upload(mesh_);
        if (u_.size() != 3 * mesh_.nodes.size()) throw std::runtime_error("flexmesh: u is not 3 * n_nodes long");
        elem_errors.assign(mesh_.elements.size(), 0.0);
        EstimateResult r;
        d.check(fm_zz_estimate(d.ctx(), u_.data(), elem_errors.data(), nullptr, nullptr, &r.global_estimate),
                "fm_zz_estimate");
        r.elem_errors = elem_errors;
        r.lower_bound = r.global_estimate / C_RELIABILITY_LINEAR_ELEMENTS;
        r.upper_bound = r.global_estimate * C_EFFICIENCY_LINEAR_ELEMENTS;
        r.meets_reliability = true;
        return r;
    }
    double estimate(std::vector<double>& elem_errors) const { return estimate_with_bounds(elem_errors).global_estimate; }

private:
    const Mesh& mesh_;
    const std::vector<double>& u_;
};

// NewtonSolverConfig / NewtonSolverResult (FEM/nonlinear_solver.hpp:23-64).
// Every member of the reference's configuration is honoured except verbose
// (nothing is printed): the absolute and relative residual tests, the energy
// test |E_k - E_(k-1)| < energy_convergence_tol after an iteration, the
// backtracking line search from line_search_alpha with factor
// line_search_beta, and reuse_tangent (the step's first tangent and its
// ILU(0) factor for every iteration).
struct NewtonSolverConfig {
    int max_iterations = 50;
    double tol_absolute = 1e-8;
    double tol_relative = 1e-6;
    int max_line_search = 10;
    double line_search_alpha = 1.0;
    double line_search_beta = 0.5;
    bool reuse_tangent = false;
    bool verbose = true;
    double energy_convergence_tol = 1e-2;
    // device extensions
    int precision = FM_FP64;
    double pcg_tol = 1e-10;
    int pcg_max = 300;
};

// residual_history[0] = ||r_0||, then one entry per iteration (and one after
// the load-step predictor solve when the prescribed values changed);
// step_size_history has one entry fewer, energy_history one per residual entry
// (the potential energy sum_e V_e psi_e - f_ext . u).
struct NewtonSolverResult {
    bool converged = false;
    int num_iterations = 0;
    double final_residual_norm = 0;
    double final_residual_relative = 0;
    double final_energy_change = 0;
    std::vector<double> residual_history;
    std::vector<double> step_size_history;
    std::vector<double> energy_history;
    int pcg_iterations = 0;
};

// NonlinearNewtonSolver (FEM/nonlinear_solver.hpp:68-238): constructed from
// the mesh, a material with params.E/params.nu and BoundaryConditions
// (dirichlet_bcs (node, dof, value), neumann_bcs (node, fx, fy, fz)); solve()
// brings u to equilibrium at load.load_factor, with the Dirichlet values and
// the Neumann forces scaled by it.  A trial step is accepted when
// ||r(u + alpha du)|| <= (1 - 0.1 alpha) ||r(u)|| (:163-165); when no trial is
// accepted the last (smallest) step is taken and the iteration continues, as
// the engine's armijo_line_search_true does (FEM/adaptive_fem_engine.hpp:933-985).
template <class Mesh, class Material, class BCs>
class NonlinearNewtonSolver {
public:
    NonlinearNewtonSolver(const Mesh& mesh, const Material& mat, const BCs& bcs)
        : mesh_(mesh), material_(mat), bcs_(bcs) {}

    template <class Solution, class LoadStep>
    NewtonSolverResult solve(Solution& u, const LoadStep& load,
                             const NewtonSolverConfig& config = NewtonSolverConfig()) const {
        Device& d = Device::get();
        d.upload(mesh_);
        d.material(material_.params.E, material_.params.nu);
        if (u.size() != (size_t)mesh_.n_dofs) throw std::runtime_error("flexmesh: u is not n_dofs long");
        std::vector<uint32_t> dofs;
        std::vector<double> vals;
        for (const auto& bc : bcs_.dirichlet_bcs) {
            dofs.push_back(3u * (uint32_t)std::get<0>(bc) + (uint32_t)std::get<1>(bc));
            vals.push_back(std::get<2>(bc));
        }
        d.check(fm_set_bcs(d.ctx(), (uint32_t)dofs.size(), dofs.data(), vals.data()), "fm_set_bcs");
        std::vector<double> f(mesh_.n_dofs, 0.0);
        for (const auto& nb : bcs_.neumann_bcs) {
            const int n = std::get<0>(nb);
            f[3 * n] += std::get<1>(nb);
            f[3 * n + 1] += std::get<2>(nb);
            f[3 * n + 2] += std::get<3>(nb);
        }
        d.check(fm_set_load(d.ctx(), bcs_.neumann_bcs.empty() ? nullptr : f.data()), "fm_set_load");
        fm_newton_cfg cfg;
        fm_newton_cfg_default(&cfg);
        cfg.prec = config.precision;
        cfg.newton_rtol = config.tol_relative;
        cfg.tol_absolute = config.tol_absolute;
        cfg.max_newton = config.max_iterations;
        cfg.ls_max = config.max_line_search;
        cfg.ls_alpha0 = config.line_search_alpha;
        cfg.ls_beta = config.line_search_beta;
        cfg.armijo_c = 0.1;
        cfg.reuse_tangent = config.reuse_tangent ? 1 : 0;
        cfg.energy_tol = config.energy_convergence_tol;
        cfg.record_energy = 1;
        cfg.pcg_tol = config.pcg_tol;
        cfg.pcg_max = config.pcg_max;
        fm_newton_stats st{};
        const int rc = fm_newton_solve_to(d.ctx(), load.load_factor, u.data(), &cfg, &st);
        if (rc != FM_OK && rc != FM_ENOCONV) d.check(rc, "fm_newton_solve_to");
        NewtonSolverResult r;
        r.converged = st.converged != 0;
        r.num_iterations = st.newton_iters;
        r.final_residual_relative = st.rel_res;
        r.final_residual_norm = st.rel_res * st.r0;
        r.pcg_iterations = st.pcg_iters;
        uint32_t len = 0;
        d.check(fm_newton_history(d.ctx(), 0, nullptr, nullptr, nullptr, &len), "fm_newton_history");
        r.residual_history.assign(len, 0.0);
        r.energy_history.assign(len, 0.0);
        r.step_size_history.assign(len > 0 ? len - 1 : 0, 0.0);
        d.check(fm_newton_history(d.ctx(), len, r.residual_history.data(),
                                  r.step_size_history.empty() ? nullptr : r.step_size_history.data(),
                                  r.energy_history.data(), &len),
                "fm_newton_history");
        if (r.converged && !r.energy_history.empty()) {
            const bool by_residual = r.final_residual_norm < config.tol_absolute ||
                                     r.final_residual_relative <= config.tol_relative;
            const size_t k = r.energy_history.size();
            // the reference reports E_last - E_first after the residual tests and
            // |E_k - E_(k-1)| after the energy test (:130-131, :219)
            r.final_energy_change = by_residual || k < 2
                                        ? r.energy_history.back() - r.energy_history.front()
                                        : std::fabs(r.energy_history[k - 1] - r.energy_history[k - 2]);
        }
        return r;
    }

private:
    const Mesh& mesh_;
    const Material& material_;
    const BCs& bcs_;
};

}  // namespace atlas::fem::gpu
