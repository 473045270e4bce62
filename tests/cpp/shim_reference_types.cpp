// Compile-only check of the C++ drop-in: every template of fm_engine.hpp
// instantiated with the reference's own types (atlas::fem::MeshTopology,
// ProblemDescriptor, SparseCSR, NeoHookeanMaterial, BoundaryConditions,
// LoadStep, Solution, ZZErrorEstimator from FEM/fem_types.hpp,
// FEM/constitutive_models.hpp, FEM/error_estimator.hpp and
// FEM/nonlinear_solver.hpp) exactly as engine code would call the CPU
// functions.  tests/test_shim_reference_types.py compiles it against the
// reference headers where they lie; nothing is linked, run or shipped.
#include "fem/nonlinear_solver.hpp"
#include "fem/stress_extraction.hpp"
#include "flexmesh/fm_engine.hpp"

namespace F = atlas::fem;
namespace G = atlas::fem::gpu;

void drop_in(F::MeshTopology& mesh, const F::ProblemDescriptor& prob, F::Solution& u, const F::NeoHookeanMaterial& mat,
             const F::BoundaryConditions& bcs, const F::LoadStep& load) {
    F::SparseCSR K;
    G::build_csr_sparsity_pattern_rigorous(mesh, K);
    std::vector<double> f_int;
    G::assemble_global_threadlocal(mesh, prob, u.data(), K, f_int);
    G::assemble_global_residual_only(mesh, prob, u.data(), f_int);
    std::vector<double> x(K.n_rows, 0.0);
    const int its = G::pcg_solve_ilu0(K, f_int.data(), x.data(), K.n_rows, 1e-10, 300);
    (void)its;
    G::MeshColoring col = G::MeshColoring::compute(mesh);
    (void)col;

    G::NonlinearNewtonSolver solver(mesh, mat, bcs);  // class template argument deduction
    G::NewtonSolverConfig cfg;
    G::NewtonSolverResult r = solver.solve(u, load, cfg);
    // the reference's result type holds the same members, filled the same way
    F::NewtonSolverResult rr;
    rr.converged = r.converged;
    rr.num_iterations = r.num_iterations;
    rr.final_residual_norm = r.final_residual_norm;
    rr.final_residual_relative = r.final_residual_relative;
    rr.final_energy_change = r.final_energy_change;
    rr.residual_history = r.residual_history;
    rr.step_size_history = r.step_size_history;
    rr.energy_history = r.energy_history;

    F::ZZErrorEstimator zz_cpu(mesh);
    G::StressExtractor<F::NeoHookeanMaterial> sx(mat);
    auto fields = sx.extract_element_fields(mesh, u, zz_cpu);
    (void)fields;
    G::ZZErrorEstimator<F::MeshTopology> zz_gpu(mesh, u);
    std::vector<double> errs;
    auto est = zz_gpu.estimate_with_bounds(errs);
    (void)est;
}
