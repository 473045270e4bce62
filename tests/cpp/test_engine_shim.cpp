// Exercises include/flexmesh/fm_engine.hpp the way engine code would: with
// structs that carry the reference's member names (FEM/fem_types.hpp), checked
// against the FP64 CPU oracle.  Prints one line per check; exit code 0 = pass.
#include <algorithm>
#include <cmath>
#include <cstdio>
#include <tuple>
#include <vector>

#include "flexmesh/fm_engine.hpp"
#include "../../oracle/fm_oracle.h"

struct MeshNode { double x{0}, y{0}, z{0}; };
struct TetraElement { uint32_t nodes[4]; };
struct MeshTopology { std::vector<MeshNode> nodes; std::vector<TetraElement> elements; uint32_t n_dofs{0}; };
struct ProblemDescriptor { double E{1.0}; double nu{0.3}; };
struct SparseCSR { uint32_t n_rows{0}; std::vector<uint32_t> row_ptr, col_idx; std::vector<double> val; };
struct NeoHookeanParameters { double E = 1.0; double nu = 0.3; };
struct NeoHookeanMaterial { NeoHookeanParameters params; };
struct BoundaryConditions {
    std::vector<std::tuple<int, int, double>> dirichlet_bcs;
    std::vector<std::tuple<int, double, double, double>> neumann_bcs;
};
struct LoadStep { double load_factor{1.0}; };
struct ZZErrorEstimator {
    void estimate(std::vector<double>& e) const { e.assign(5, 0.25); }
};

static int failures = 0;
static void check(const char* name, bool ok, double v) {
    std::printf("%-46s %s (%.3e)\n", name, ok ? "ok" : "FAIL", v);
    if (!ok) ++failures;
}

static double rel(const std::vector<double>& a, const std::vector<double>& b) {
    double d = 0, m = 0;
    for (size_t i = 0; i < a.size(); ++i) {
        d = std::fmax(d, std::fabs(a[i] - b[i]));
        m = std::fmax(m, std::fabs(b[i]));
    }
    return d / (m > 0 ? m : 1.0);
}

int run();

int main() {
    std::setvbuf(stdout, nullptr, _IONBF, 0);
    try {
        return run();
    } catch (const std::exception& e) {
        std::printf("exception: %s\n", e.what());
        return 2;
    }
}

int run() {
    namespace G = atlas::fem::gpu;
    const int N = 4;
    const uint32_t nn = (N + 1) * (N + 1) * (N + 1), ne = 6 * N * N * N;
    std::vector<double> X(3 * nn);
    std::vector<uint32_t> conn(4 * ne);
    fm_mesh_box(N, N, N, 1, 1, 1, X.data(), conn.data());
    MeshTopology mesh;
    mesh.nodes.resize(nn);
    for (uint32_t n = 0; n < nn; ++n) mesh.nodes[n] = {X[n] + 0.01 * std::sin(7.0 * n), X[nn + n], X[2 * nn + n]};
    for (uint32_t n = 0; n < nn; ++n) X[n] = mesh.nodes[n].x;
    mesh.elements.resize(ne);
    for (uint32_t e = 0; e < ne; ++e)
        for (int a = 0; a < 4; ++a) mesh.elements[e].nodes[a] = conn[a * ne + e];
    mesh.n_dofs = 3 * nn;
    ProblemDescriptor prob;

    SparseCSR K;
    G::build_csr_sparsity_pattern_rigorous(mesh, K);
    std::vector<uint32_t> rp(3 * nn + 1);
    const uint64_t nnz = orc_pattern_rows(nn, ne, conn.data(), rp.data());
    std::vector<uint32_t> ci(nnz);
    orc_pattern_cols(nn, ne, conn.data(), rp.data(), ci.data());
    check("pattern bit-exact", K.row_ptr == rp && K.col_idx == ci, 0.0);

    auto col = G::MeshColoring::compute(mesh);
    std::vector<uint32_t> co(ne);
    const uint32_t nco = orc_color(nn, ne, conn.data(), co.data());
    check("colouring bit-exact", col.element_colors == co && col.n_colors == nco, (double)nco);

    std::vector<double> u(3 * nn);
    for (uint32_t i = 0; i < 3 * nn; ++i) u[i] = 0.01 * std::cos(0.37 * i);
    std::vector<double> f;
    G::assemble_global_threadlocal(mesh, prob, u.data(), K, f);
    const double mu = 1.0 / 2.6, lam = 0.3 / (1.3 * 0.4);
    std::vector<double> mv(ne, mu), lv(ne, lam), Ko(nnz), fo(3 * nn);
    orc_assemble(nn, ne, X.data(), conn.data(), mv.data(), lv.data(), co.data(), nco, u.data(), rp.data(), ci.data(),
                 Ko.data(), fo.data());
    check("assemble_global_threadlocal K vs oracle", rel(K.val, Ko) < 1e-12, rel(K.val, Ko));
    check("assemble_global_threadlocal f vs oracle", rel(f, fo) < 1e-12, rel(f, fo));
    std::vector<double> fr;
    G::assemble_global_residual_only(mesh, prob, u.data(), fr);
    check("assemble_global_residual_only vs oracle", rel(fr, fo) < 1e-12, rel(fr, fo));

    // pcg_solve_ilu0 on the 2-D Poisson matrix of the reference's validation runner
    const int g = 24;
    SparseCSR A;
    A.n_rows = g * g;
    A.row_ptr.push_back(0);
    const double h = 1.0 / (g + 1), ih2 = 1.0 / (h * h);
    for (int i = 0; i < g; ++i)
        for (int j = 0; j < g; ++j) {
            const int r = i * g + j;
            A.col_idx.push_back(r);
            A.val.push_back(4 * ih2);
            if (i > 0) { A.col_idx.push_back(r - g); A.val.push_back(-ih2); }
            if (i + 1 < g) { A.col_idx.push_back(r + g); A.val.push_back(-ih2); }
            if (j > 0) { A.col_idx.push_back(r - 1); A.val.push_back(-ih2); }
            if (j + 1 < g) { A.col_idx.push_back(r + 1); A.val.push_back(-ih2); }
            A.row_ptr.push_back((uint32_t)A.col_idx.size());
        }
    std::vector<double> b(g * g), x(g * g, 0.0);
    for (int r = 0; r < g * g; ++r) b[r] = std::sin(0.1 * r);
    const int its = G::pcg_solve_ilu0(A, b.data(), x.data(), A.n_rows, 1e-10, 4000);
    // the oracle needs sorted rows: sort a copy
    SparseCSR S = A;
    for (uint32_t r = 0; r < S.n_rows; ++r) {
        std::vector<std::pair<uint32_t, double>> row;
        for (uint32_t k = A.row_ptr[r]; k < A.row_ptr[r + 1]; ++k) row.push_back({A.col_idx[k], A.val[k]});
        std::sort(row.begin(), row.end());
        for (uint32_t k = A.row_ptr[r]; k < A.row_ptr[r + 1]; ++k) {
            S.col_idx[k] = row[k - A.row_ptr[r]].first;
            S.val[k] = row[k - A.row_ptr[r]].second;
        }
    }
    std::vector<double> xo(g * g, 0.0);
    double relo = 0;
    const int itso = orc_pcg(S.n_rows, S.row_ptr.data(), S.col_idx.data(), S.val.data(), b.data(), xo.data(), 1e-10,
                             4000, &relo);
    check("pcg_solve_ilu0 iterations vs oracle (+-1)", std::abs(its - itso) <= 1, (double)its);
    check("pcg_solve_ilu0 solution vs oracle", rel(x, xo) < 1e-9, rel(x, xo));
    // the same pattern again with new values: the kept plan must give the new solution
    SparseCSR A2 = A;
    for (uint32_t r = 0; r < A2.n_rows; ++r)
        for (uint32_t k = A2.row_ptr[r]; k < A2.row_ptr[r + 1]; ++k)
            if (A2.col_idx[k] == r) A2.val[k] = 6 * ih2;
    for (uint32_t r = 0; r < S.n_rows; ++r)
        for (uint32_t k = S.row_ptr[r]; k < S.row_ptr[r + 1]; ++k)
            if (S.col_idx[k] == r) S.val[k] = 6 * ih2;
    std::fill(x.begin(), x.end(), 0.0);
    std::fill(xo.begin(), xo.end(), 0.0);
    const int its2 = G::pcg_solve_ilu0(A2, b.data(), x.data(), A2.n_rows, 1e-10, 4000);
    const int itso2 = orc_pcg(S.n_rows, S.row_ptr.data(), S.col_idx.data(), S.val.data(), b.data(), xo.data(), 1e-10,
                              4000, &relo);
    check("pcg_solve_ilu0 same pattern, new values vs oracle", std::abs(its2 - itso2) <= 1 && rel(x, xo) < 1e-9,
          rel(x, xo));

    // nodes moved in place (same mesh object, same topology): the device copy follows
    for (uint32_t n = 0; n < nn; ++n) mesh.nodes[n].y += 0.004 * std::cos(3.0 * n);
    for (uint32_t n = 0; n < nn; ++n) X[nn + n] = mesh.nodes[n].y;
    G::assemble_global_threadlocal(mesh, prob, u.data(), K, f);
    orc_assemble(nn, ne, X.data(), conn.data(), mv.data(), lv.data(), co.data(), nco, u.data(), rp.data(), ci.data(),
                 Ko.data(), fo.data());
    check("nodes moved in place: K, f follow the coordinates", rel(K.val, Ko) < 1e-12 && rel(f, fo) < 1e-12,
          std::fmax(rel(K.val, Ko), rel(f, fo)));

    // NonlinearNewtonSolver: homogeneous uniaxial stretch with sliding faces has
    // the closed form F = diag(l1, l2, l2), mu (l2 - 1/l2) + lam ln(l1 l2^2)/l2 = 0
    MeshTopology cube;
    cube.nodes.resize(nn);
    std::vector<double> Xc(3 * nn);
    fm_mesh_box(N, N, N, 1, 1, 1, Xc.data(), conn.data());
    for (uint32_t n = 0; n < nn; ++n) cube.nodes[n] = {Xc[n], Xc[nn + n], Xc[2 * nn + n]};
    cube.elements = mesh.elements;
    cube.n_dofs = 3 * nn;
    BoundaryConditions bcs;
    const double delta = 0.1;
    for (uint32_t n = 0; n < nn; ++n) {
        if (cube.nodes[n].x == 0.0) bcs.dirichlet_bcs.emplace_back(n, 0, 0.0);
        if (cube.nodes[n].x == 1.0) bcs.dirichlet_bcs.emplace_back(n, 0, delta);
        if (cube.nodes[n].y == 0.0) bcs.dirichlet_bcs.emplace_back(n, 1, 0.0);
        if (cube.nodes[n].z == 0.0) bcs.dirichlet_bcs.emplace_back(n, 2, 0.0);
    }
    NeoHookeanMaterial mat;
    G::NonlinearNewtonSolver<MeshTopology, NeoHookeanMaterial, BoundaryConditions> solver(cube, mat, bcs);
    std::vector<double> uc(3 * nn, 0.0);
    G::NewtonSolverConfig cfg;
    cfg.tol_relative = 1e-12;
    cfg.tol_absolute = 0.0;
    cfg.energy_convergence_tol = 0.0;
    cfg.pcg_tol = 1e-13;
    cfg.pcg_max = 2000;
    LoadStep half{0.5}, full{1.0};
    auto r1 = solver.solve(uc, half, cfg);
    auto r2 = solver.solve(uc, full, cfg);
    {
        // histories in the reference's shapes (NewtonSolverResult::print indexes all three)
        const size_t L = r2.residual_history.size();
        bool ok = L >= 2 && r2.step_size_history.size() == L - 1 && r2.energy_history.size() == L &&
                  std::fabs(r2.residual_history.back() - r2.final_residual_norm) <= 1e-12 * r2.residual_history[0] &&
                  std::fabs(r2.final_residual_norm / r2.residual_history[0] - r2.final_residual_relative) <= 1e-12;
        for (size_t i = 0; i < L; ++i) ok = ok && std::isfinite(r2.energy_history[i]);
        for (double a : r2.step_size_history) ok = ok && a > 0.0 && a <= 1.0;
        // no external force: the energy is the strain energy, positive at the stretched state
        ok = ok && r2.energy_history.back() > 0.0;
        check("NonlinearNewtonSolver histories (lengths, last = final)", ok, (double)L);
    }
    {
        // tol_absolute: a solve from the converged state stops at once (:127)
        G::NewtonSolverConfig c2 = cfg;
        c2.tol_absolute = 1e-8;
        std::vector<double> ucc = uc;
        auto r3 = solver.solve(ucc, full, c2);
        check("tol_absolute: converged state, 0 iterations", r3.converged && r3.num_iterations == 0 &&
                                                              r3.residual_history.size() == 1,
              r3.final_residual_norm);
    }
    {
        // the energy test (:215-227) with the reference's default 1e-2: stops once |dE| < 1e-2
        G::NewtonSolverConfig c3;
        c3.tol_absolute = 0.0;
        c3.tol_relative = 1e-14;
        std::vector<double> u0(3 * nn, 0.0);
        auto r4 = solver.solve(u0, half, c3);
        const size_t k = r4.energy_history.size();
        const double dE = k >= 2 ? std::fabs(r4.energy_history[k - 1] - r4.energy_history[k - 2]) : 1.0;
        check("energy_convergence_tol stops Newton", r4.converged && k >= 2 && dE < 1e-2 &&
                                                    r4.final_residual_relative > 1e-14 &&
                                                    std::fabs(r4.final_energy_change - dE) <= 1e-15,
              dE);
    }
    {
        // reuse_tangent: the first tangent for every iteration still reaches the same state
        G::NewtonSolverConfig c4 = cfg;
        c4.reuse_tangent = true;
        c4.tol_relative = 1e-10;
        c4.max_iterations = 200;
        std::vector<double> u0(3 * nn, 0.0);
        auto r5 = solver.solve(u0, half, c4);
        auto r5b = solver.solve(u0, full, c4);
        check("reuse_tangent converges to the same state", r5.converged && r5b.converged &&
                                                          r5b.num_iterations >= r2.num_iterations &&
                                                          rel(u0, uc) < 1e-8,
              rel(u0, uc));
    }
    {
        bool threw = false;
        G::NewtonSolverConfig c5 = cfg;
        c5.line_search_beta = 1.5;
        std::vector<double> u0(3 * nn, 0.0);
        try {
            (void)solver.solve(u0, half, c5);
        } catch (const std::exception&) {
            threw = true;
        }
        check("line_search_beta outside (0, 1) is rejected", threw, 0.0);
    }
    double lo = 0.3, hi = 1.5;
    auto fl = [&](double l2) { return mu * (l2 - 1 / l2) + lam * std::log((1 + delta) * l2 * l2) / l2; };
    for (int it = 0; it < 200; ++it) {
        const double m = 0.5 * (lo + hi);
        if (fl(lo) * fl(m) <= 0) hi = m; else lo = m;
    }
    const double l2 = 0.5 * (lo + hi);
    double err = 0;
    for (uint32_t n = 0; n < nn; ++n) {
        err = std::fmax(err, std::fabs(uc[3 * n] - delta * cube.nodes[n].x));
        err = std::fmax(err, std::fabs(uc[3 * n + 1] - (l2 - 1) * cube.nodes[n].y));
        err = std::fmax(err, std::fabs(uc[3 * n + 2] - (l2 - 1) * cube.nodes[n].z));
    }
    check("NonlinearNewtonSolver converged (two load steps)", r1.converged && r2.converged, r2.final_residual_relative);
    check("NonlinearNewtonSolver vs closed-form uniaxial", err < 1e-10, err);

    // StressExtractor at the converged uniaxial state against the oracle
    G::StressExtractor<NeoHookeanMaterial> sx(mat);
    ZZErrorEstimator zz;
    auto fd = sx.extract_element_fields(cube, uc, zz);
    const uint32_t nec = (uint32_t)cube.elements.size();
    std::vector<double> vm(nec), pr(nec), s1(nec), psi(nec), st(nec), dj(nec), ml(nec, mu), ll(nec, lam);
    orc_element_fields(nn, Xc.data(), nec, conn.data(), ml.data(), ll.data(), uc.data(), dj.data(), nullptr,
                       vm.data(), pr.data(), s1.data(), psi.data(), st.data(), nullptr);
    const double fe = std::fmax(std::fmax(rel(fd.von_mises, vm), rel(fd.pressure, pr)),
                                std::fmax(std::fmax(rel(fd.max_principal_stress, s1), rel(fd.strain_energy, psi)),
                                          std::fmax(rel(fd.max_stretch, st), rel(fd.det_F, dj))));
    check("StressExtractor fields vs oracle", fe < 1e-12, fe);
    std::vector<double> u2(uc);  // a non-uniform field, so the indicators are not rounding noise
    for (size_t i = 0; i < u2.size(); ++i) u2[i] += 1e-3 * std::sin(0.37 * (double)i);
    G::ZZErrorEstimator<MeshTopology> zzg(cube, u2);
    std::vector<double> ezz;
    auto zr = zzg.estimate_with_bounds(ezz);
    std::vector<double> ezo(nec);
    const double gzo = orc_zz_estimate(nn, Xc.data(), nec, conn.data(), u2.data(), ezo.data(), nullptr, nullptr);
    double ezmax = 0.0;
    for (double v : ezo) ezmax = std::fmax(ezmax, v);
    double ezd = 0.0;
    for (uint32_t e = 0; e < nec; ++e) ezd = std::fmax(ezd, std::fabs(ezz[e] - ezo[e]));
    check("ZZErrorEstimator vs oracle", std::fabs(zr.global_estimate - gzo) <= 1e-9 * gzo + 1e-300 &&
                                            ezd <= 1e-9 * ezmax + 1e-300 && zr.lower_bound == zr.global_estimate / 2,
          ezd);
    check("StressExtractor error indicators from the estimator",
          fd.error_indicators.size() == nec && fd.error_indicators[4] == 0.25 && fd.error_indicators[5] == 0.0, 0.0);
    std::printf("%s\n", failures ? "FAILED" : "ALL OK");
    return failures ? 1 : 0;
}
