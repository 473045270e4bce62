// ref_driver.cpp -- C entry points over the compiled reference, used only to
// write golden vectors (tests/golden/make_golden.py) and to cross-check the
// oracle in this container.  TEST INFRASTRUCTURE ONLY.
//
// Built by oracle/build_ref.sh from the reference sources where they lie;
// the output goes to oracle/_ref/ and is never committed.  Every function
// below calls the reference's own code unchanged.
#include "fem/adaptive_fem_engine.hpp"
#include "fem/constitutive_models.hpp"
#include "fem/stress_extraction.hpp"

#include <chrono>
#include <cstring>

using namespace atlas::fem;

namespace {
MeshTopology mesh_from(uint32_t nn, const double* X, uint32_t ne, const uint32_t* conn) {
    MeshTopology m;
    m.nodes.resize(nn);
    for (uint32_t n = 0; n < nn; ++n) {
        m.nodes[n].x = X[n];
        m.nodes[n].y = X[nn + n];
        m.nodes[n].z = X[2 * nn + n];
        m.nodes[n].id = n;
    }
    m.elements.resize(ne);
    m.states.resize(ne);
    for (uint32_t e = 0; e < ne; ++e) {
        for (int a = 0; a < 4; ++a) m.elements[e].nodes[a] = conn[a * (size_t)ne + e];
        m.elements[e].id = e;
        m.elements[e].volume = 1.0;
    }
    m.n_dofs = 3 * nn;
    return m;
}

SparseCSR csr_from(uint32_t n, const uint32_t* row_ptr, const uint32_t* col_idx, const double* val) {
    SparseCSR A;
    A.n_rows = n;
    A.row_ptr.assign(row_ptr, row_ptr + n + 1);
    A.col_idx.assign(col_idx, col_idx + row_ptr[n]);
    A.val.assign(val, val + row_ptr[n]);
    return A;
}
}  // namespace

extern "C" {

// NeoHookeanMaterial::compute_pk1_stress (FEM/constitutive_models.hpp:61-113)
void ref_pk1(const double F[9], double E, double nu, double P[9]) {
    NeoHookeanParameters p;
    p.E = E;
    p.nu = nu;
    NeoHookeanMaterial m(p);
    m.compute_pk1_stress(F, P);
}

// NeoHookeanMaterial::compute_cauchy_stress (FEM/constitutive_models.hpp:116-148)
void ref_cauchy(const double F[9], double E, double nu, double sigma[9]) {
    NeoHookeanParameters p;
    p.E = E;
    p.nu = nu;
    NeoHookeanMaterial m(p);
    m.compute_cauchy_stress(F, sigma);
}

// StressExtractor::compute_von_mises (FEM/stress_extraction.hpp:143-157)
double ref_von_mises(const double sigma[9]) {
    NeoHookeanParameters p;
    StressExtractor x{NeoHookeanMaterial(p)};
    return x.compute_von_mises(sigma);
}

// generate_unit_cube_mesh (FEM/adaptive_fem_engine.hpp:2219-2294)
uint32_t ref_cube_mesh(int N, double* X, uint32_t* conn) {
    MeshTopology m = generate_unit_cube_mesh(N);
    const uint32_t nn = (uint32_t)m.n_nodes(), ne = (uint32_t)m.n_elements();
    for (uint32_t n = 0; n < nn; ++n) {
        X[n] = m.nodes[n].x;
        X[nn + n] = m.nodes[n].y;
        X[2 * nn + n] = m.nodes[n].z;
    }
    for (uint32_t e = 0; e < ne; ++e)
        for (int a = 0; a < 4; ++a) conn[a * (size_t)ne + e] = m.elements[e].nodes[a];
    return ne;
}

// build_csr_sparsity_pattern_rigorous (:1069-1120); returns nnz, fills row_ptr (n_dofs+1)
// and col_idx (nnz, may be NULL)
uint32_t ref_csr(uint32_t nn, const double* X, uint32_t ne, const uint32_t* conn, uint32_t* row_ptr,
                 uint32_t* col_idx) {
    MeshTopology m = mesh_from(nn, X, ne, conn);
    SparseCSR K;
    build_csr_sparsity_pattern_rigorous(m, K);
    std::memcpy(row_ptr, K.row_ptr.data(), K.row_ptr.size() * sizeof(uint32_t));
    if (col_idx) std::memcpy(col_idx, K.col_idx.data(), K.col_idx.size() * sizeof(uint32_t));
    return K.row_ptr.back();
}

// MeshColoring::compute (:1549-1602)
uint32_t ref_color(uint32_t nn, const double* X, uint32_t ne, const uint32_t* conn, uint32_t* colors) {
    MeshTopology m = mesh_from(nn, X, ne, conn);
    MeshColoring c = MeshColoring::compute(m);
    std::memcpy(colors, c.element_colors.data(), ne * sizeof(uint32_t));
    return c.n_colors;
}

// ILU0Preconditioner::factorize (:305-470)
void ref_ilu0(uint32_t n, const uint32_t* row_ptr, const uint32_t* col_idx, const double* val, double* L_val,
              double* U_val, double* diag) {
    SparseCSR A = csr_from(n, row_ptr, col_idx, val);
    auto f = ILU0Preconditioner::factorize(A);
    std::memcpy(L_val, f.L_val.data(), f.L_val.size() * sizeof(double));
    std::memcpy(U_val, f.U_val.data(), f.U_val.size() * sizeof(double));
    std::memcpy(diag, f.diag.data(), n * sizeof(double));
}

// ILU0Preconditioner::solve (:473-504) on a fresh factorization
void ref_ilu0_solve(uint32_t n, const uint32_t* row_ptr, const uint32_t* col_idx, const double* val, const double* x,
                    double* y) {
    SparseCSR A = csr_from(n, row_ptr, col_idx, val);
    auto f = ILU0Preconditioner::factorize(A);
    ILU0Preconditioner::solve(f, x, y, n);
}

// pcg_solve_ilu0 (:1283-1341)
int ref_pcg(uint32_t n, const uint32_t* row_ptr, const uint32_t* col_idx, const double* val, const double* b,
            double* x, double tol, int max_iter) {
    SparseCSR A = csr_from(n, row_ptr, col_idx, val);
    return pcg_solve_ilu0(A, b, x, n, tol, max_iter);
}

// CheckpointFile::save (FEM/adaptive_fem_engine.hpp:1347-1398) of a mesh with
// the displacement u (3 nn), which also goes into the node records' u, v, w.
// Element records: the reference's tet_volume and tet_quality
// (FEM/mesh_adaptation.hpp:110-146) and jacobian_det = 6 volume, as
// generate_unit_cube_mesh sets them (:2266-2268).  States: F of each element
// (the P1 gradient, F_data[i*3+j] = F_ij as update_element_states_rigorous
// fills it, :2755-2758, without its projection onto det F = 1) and the Mandel
// stress of neo_hookean_response (:79-129) with J = |det F|; the plastic fields
// stay zero.  Returns 0 on success.
int ref_checkpoint_save(const char* path, uint32_t nn, const double* X, const double* u, uint32_t ne,
                        const uint32_t* conn, double E, double nu, uint32_t load_step, uint32_t adapt_iter) {
    MeshTopology m = mesh_from(nn, X, ne, conn);
    for (uint32_t n = 0; n < nn; ++n) {
        m.nodes[n].u = u[3 * n];
        m.nodes[n].v = u[3 * n + 1];
        m.nodes[n].w = u[3 * n + 2];
    }
    for (uint32_t e = 0; e < ne; ++e) {
        TetraElement& el = m.elements[e];
        const MeshNode &n0 = m.nodes[el.nodes[0]], &n1 = m.nodes[el.nodes[1]], &n2 = m.nodes[el.nodes[2]],
                       &n3 = m.nodes[el.nodes[3]];
        el.volume = tet_volume(n0, n1, n2, n3);
        el.quality_metric = tet_quality(n0, n1, n2, n3);
        el.jacobian_det = el.volume * 6.0;
        double J[9];
        const MeshNode* nd[4] = {&n0, &n1, &n2, &n3};
        for (int d = 0; d < 3; ++d)
            for (int a = 1; a < 4; ++a) J[d * 3 + a - 1] = nd[a]->coord(d) - nd[0]->coord(d);
        double Ji[9];
        matrix_3x3_inv(J, Ji);
        const double dxi[4][3] = {{-1, -1, -1}, {1, 0, 0}, {0, 1, 0}, {0, 0, 1}};
        double F[9] = {1, 0, 0, 0, 1, 0, 0, 0, 1};
        for (int a = 0; a < 4; ++a) {
            double g[3];
            for (int i = 0; i < 3; ++i) g[i] = Ji[0 * 3 + i] * dxi[a][0] + Ji[1 * 3 + i] * dxi[a][1] + Ji[2 * 3 + i] * dxi[a][2];
            for (int i = 0; i < 3; ++i)
                for (int j = 0; j < 3; ++j) F[i * 3 + j] += u[3 * el.nodes[a] + i] * g[j];
        }
        const double detF = F[0] * (F[4] * F[8] - F[5] * F[7]) - F[1] * (F[3] * F[8] - F[5] * F[6]) +
                            F[2] * (F[3] * F[7] - F[4] * F[6]);
        ElementState& st = m.states[e];
        st.initialize();
        for (int k = 0; k < 9; ++k) st.F_data[k] = F[k];
        neo_hookean_response(st.F_data, E, nu, st.stress, std::fabs(detF));
    }
    std::vector<double> uu(u, u + 3 * (size_t)nn);
    return CheckpointFile::save(path, m, uu, load_step, adapt_iter) ? 0 : 1;
}

// Wall seconds (best of reps) of assemble_global_threadlocal
// (FEM/adaptive_fem_engine.hpp:1126-1233), or of assemble_global_residual_only
// (:990-1059) when residual_only, on the mesh with one (E, nu); the pattern
// comes from build_csr_sparsity_pattern_rigorous (:1069-1120), built once.
double ref_time_assemble(uint32_t nn, const double* X, uint32_t ne, const uint32_t* conn, double E, double nu,
                         const double* u, int reps, int residual_only) {
    MeshTopology m = mesh_from(nn, X, ne, conn);
    ProblemDescriptor prob;
    prob.E = E;
    prob.nu = nu;
    SparseCSR K;
    build_csr_sparsity_pattern_rigorous(m, K);
    std::vector<double> f;
    double best = 1e300;
    for (int r = 0; r < reps; ++r) {
        const auto t0 = std::chrono::steady_clock::now();
        if (residual_only)
            assemble_global_residual_only(m, prob, u, f);
        else
            assemble_global_threadlocal(m, prob, u, K, f);
        const double dt = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
        best = dt < best ? dt : best;
    }
    return best;
}

// Wall seconds (best of reps) of SparseCSR::matvec (FEM/fem_types.hpp:310-319)
double ref_time_spmv(uint32_t n, const uint32_t* row_ptr, const uint32_t* col_idx, const double* val,
                     const double* x, double* y, int reps) {
    SparseCSR A = csr_from(n, row_ptr, col_idx, val);
    double best = 1e300;
    for (int r = 0; r < reps; ++r) {
        const auto t0 = std::chrono::steady_clock::now();
        A.matvec(x, y);
        const double dt = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
        best = dt < best ? dt : best;
    }
    return best;
}

// exponential_map / matrix_logarithm (CORE/lie_operator.cpp:94-275), column-major 3x3
void ref_expm(const double A[9], double out[9]) {
    atlas::Matrix3x3 M;
    for (int i = 0; i < 9; ++i) M.data[i] = A[i];
    atlas::Matrix3x3 R = atlas::exponential_map(M);
    for (int i = 0; i < 9; ++i) out[i] = R.data[i];
}

void ref_logm(const double A[9], double out[9]) {
    atlas::Matrix3x3 M;
    for (int i = 0; i < 9; ++i) M.data[i] = A[i];
    atlas::Matrix3x3 R = atlas::matrix_logarithm(M);
    for (int i = 0; i < 9; ++i) out[i] = R.data[i];
}

}  // extern "C"
