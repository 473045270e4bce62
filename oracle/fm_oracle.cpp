// fm_oracle.cpp -- FP64 CPU oracle for the FlexMesh quasi-static Newton hot path.
//
// TEST INFRASTRUCTURE ONLY (the checker and the CPU baseline; never part of the
// product).  Each function restates one reference algorithm and cites the
// reference file:line it follows.  Where the reference is defective the
// correction is stated next to the citation (defect list: CSR column fill,
// inconsistent tangent, O(n^2) ILU scan, non-conforming mesher, O(n^2)
// coloring adjacency).
//
// Parity pinning of every function here is done in tests/ (closed forms,
// finite differences, brute force on tiny inputs, and golden vectors written
// by the compiled reference in oracle/_ref).
#include "fm_oracle.h"

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstring>
#include <vector>
#ifdef _OPENMP
#include <omp.h>
#endif

namespace {

int g_threads = 1;

double now_s() {
    using namespace std::chrono;
    return duration<double>(steady_clock::now().time_since_epoch()).count();
}

// Kuhn (Freudenthal) split of a hex into 6 tets that all share the 0-6 diagonal.
// Corner numbering follows the reference hex_off table
// (FEM/adaptive_fem_engine.hpp:2245-2248); the reference's own split
// (:2240-2243) is non-conforming, so it is replaced here.  Vertex order is
// chosen so that every tet has positive orientation on an undistorted grid.
const int kKuhn[6][4] = {{0, 1, 2, 6}, {0, 5, 1, 6}, {0, 2, 3, 6},
                         {0, 3, 7, 6}, {0, 4, 5, 6}, {0, 7, 4, 6}};
const int kHexOff[8][3] = {{0, 0, 0}, {1, 0, 0}, {1, 1, 0}, {0, 1, 0},
                           {0, 0, 1}, {1, 0, 1}, {1, 1, 1}, {0, 1, 1}};

inline double det3(const double A[9]) {
    // same cofactor expansion as the reference (FEM/adaptive_fem_engine.hpp:53-55)
    return A[0] * (A[4] * A[8] - A[7] * A[5]) - A[3] * (A[1] * A[8] - A[7] * A[2]) +
           A[6] * (A[1] * A[5] - A[4] * A[2]);
}

// node neighbours (sorted, including the node itself) through element adjacency
void node_neighbours(uint32_t nn, uint32_t ne, const uint32_t* conn,
                     std::vector<uint32_t>& ptr, std::vector<uint32_t>& nbr) {
    std::vector<uint32_t> eptr(nn + 1, 0), elist;
    for (int a = 0; a < 4; ++a)
        for (uint32_t e = 0; e < ne; ++e) eptr[conn[a * (size_t)ne + e] + 1]++;
    for (uint32_t n = 0; n < nn; ++n) eptr[n + 1] += eptr[n];
    elist.resize(eptr[nn]);
    std::vector<uint32_t> fill(eptr.begin(), eptr.end() - 1);
    for (uint32_t e = 0; e < ne; ++e)
        for (int a = 0; a < 4; ++a) elist[fill[conn[a * (size_t)ne + e]]++] = e;
    ptr.assign(nn + 1, 0);
    std::vector<std::vector<uint32_t>> tmp(nn);
    for (uint32_t n = 0; n < nn; ++n) {
        auto& v = tmp[n];
        for (uint32_t k = eptr[n]; k < eptr[n + 1]; ++k)
            for (int a = 0; a < 4; ++a) v.push_back(conn[a * (size_t)ne + elist[k]]);
        std::sort(v.begin(), v.end());
        v.erase(std::unique(v.begin(), v.end()), v.end());
        ptr[n + 1] = ptr[n] + (uint32_t)v.size();
    }
    nbr.resize(ptr[nn]);
    for (uint32_t n = 0; n < nn; ++n) std::copy(tmp[n].begin(), tmp[n].end(), nbr.begin() + ptr[n]);
}

// position of column c in a sorted CSR row, or -1
inline int64_t find_col(const uint32_t* row_ptr, const uint32_t* col_idx, uint32_t r, uint32_t c) {
    const uint32_t* b = col_idx + row_ptr[r];
    const uint32_t* e = col_idx + row_ptr[r + 1];
    const uint32_t* it = std::lower_bound(b, e, c);
    if (it == e || *it != c) return -1;
    return (int64_t)(it - col_idx);
}

struct ElemGeo {
    double g[4][3];
    double vol;
};

// Shape-function gradients, restating FEM/adaptive_fem_engine.hpp:527-562.
bool elem_geometry(const double Xe[12], ElemGeo& G) {
    double J[9];
    for (int d = 0; d < 3; ++d) {
        J[d * 3 + 0] = Xe[1 * 3 + d] - Xe[0 * 3 + d];
        J[d * 3 + 1] = Xe[2 * 3 + d] - Xe[0 * 3 + d];
        J[d * 3 + 2] = Xe[3 * 3 + d] - Xe[0 * 3 + d];
    }
    double Jinv[9];
    orc_inv3(J, Jinv);
    const double detJ = std::fabs(det3(J));
    G.vol = detJ / 6.0;
    if (G.vol < 1e-30) return false;  // :545-549
    static const double dxi[4][3] = {{-1, -1, -1}, {1, 0, 0}, {0, 1, 0}, {0, 0, 1}};
    for (int a = 0; a < 4; ++a)
        for (int i = 0; i < 3; ++i) {
            double s = 0.0;
            for (int j = 0; j < 3; ++j) s += Jinv[j * 3 + i] * dxi[a][j];
            G.g[a][i] = s;
        }
    return true;
}

struct ElemKin {
    double F[9], FiT[9];  // F and F^{-T}, row-major
    double lnJ;
    bool clamped;
};

// F = I + sum_a u_a (x) g_a (FEM/adaptive_fem_engine.hpp:565-575), J = |det F|
// (= sqrt|det C|, FEM/constitutive_models.hpp:79-83) and ln(max(J,0.01)) (:96).
void kin_from_F(ElemKin& K) {
    const double* F = K.F;
    const double det = det3(F);
    const double J = std::fabs(det);
    if (J < 1e-15) {
        // reference inverse guard on C (|det C| < 1e-30 -> C^{-1} = 1e30 I,
        // FEM/adaptive_fem_engine.hpp:56-60) carried to F^{-T} = F C^{-1}
        for (int k = 0; k < 9; ++k) K.FiT[k] = 1e30 * F[k];
    } else {
        const double id = 1.0 / det;
        // F^{-T}_ij = cof(F)_ij / det
        K.FiT[0] = (F[4] * F[8] - F[5] * F[7]) * id;
        K.FiT[1] = (F[5] * F[6] - F[3] * F[8]) * id;
        K.FiT[2] = (F[3] * F[7] - F[4] * F[6]) * id;
        K.FiT[3] = (F[2] * F[7] - F[1] * F[8]) * id;
        K.FiT[4] = (F[0] * F[8] - F[2] * F[6]) * id;
        K.FiT[5] = (F[1] * F[6] - F[0] * F[7]) * id;
        K.FiT[6] = (F[1] * F[5] - F[2] * F[4]) * id;
        K.FiT[7] = (F[2] * F[3] - F[0] * F[5]) * id;
        K.FiT[8] = (F[0] * F[4] - F[1] * F[3]) * id;
    }
    K.clamped = J < 0.01;
    K.lnJ = std::log(std::max(J, 0.01));
}

void elem_kinematics(const ElemGeo& G, const double ue[12], ElemKin& K) {
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j) {
            double s = (i == j) ? 1.0 : 0.0;
            for (int a = 0; a < 4; ++a) s += ue[a * 3 + i] * G.g[a][j];
            K.F[i * 3 + j] = s;
        }
    kin_from_F(K);
}

// Element internal force f_a = V P g_a with P = mu F + (lam lnJ - mu) F^{-T}
// (the law of FEM/constitutive_models.hpp:61-113 written without C^{-1}), and
// the exact tangent K_ab,ik = V [mu (g_a.g_b) d_ik + lam' a_a,i a_b,k + c a_a,k a_b,i]
// with a = F^{-T} g, c = mu - lam lnJ, lam' = 0 where ln J is clamped.  The
// reference tangent (FEM/adaptive_fem_engine.hpp:594-692) is not the derivative
// of its own force and is not reproduced.
void elem_force_tangent(const ElemGeo& G, const ElemKin& K, double mu, double lam, double fe[12],
                        double* Ke) {
    double a[4][3], Fg[4][3];
    for (int n = 0; n < 4; ++n)
        for (int i = 0; i < 3; ++i) {
            a[n][i] = K.FiT[i * 3 + 0] * G.g[n][0] + K.FiT[i * 3 + 1] * G.g[n][1] + K.FiT[i * 3 + 2] * G.g[n][2];
            Fg[n][i] = K.F[i * 3 + 0] * G.g[n][0] + K.F[i * 3 + 1] * G.g[n][1] + K.F[i * 3 + 2] * G.g[n][2];
        }
    const double V = G.vol;
    const double s = lam * K.lnJ - mu;
    for (int n = 0; n < 4; ++n)
        for (int i = 0; i < 3; ++i) fe[n * 3 + i] = V * (mu * Fg[n][i] + s * a[n][i]);
    if (!Ke) return;
    const double lp = K.clamped ? 0.0 : lam;
    const double c = mu - lam * K.lnJ;
    for (int p = 0; p < 4; ++p)
        for (int q = 0; q < 4; ++q) {
            const double gg = G.g[p][0] * G.g[q][0] + G.g[p][1] * G.g[q][1] + G.g[p][2] * G.g[q][2];
            for (int i = 0; i < 3; ++i)
                for (int k = 0; k < 3; ++k) {
                    double v = lp * a[p][i] * a[q][k] + c * a[p][k] * a[q][i];
                    if (i == k) v += mu * gg;
                    Ke[(p * 3 + i) * 12 + (q * 3 + k)] = V * v;
                }
        }
}

void gather_elem(const double* X, uint32_t nn, const uint32_t* conn, uint32_t ne, uint32_t e,
                 const double* u, double Xe[12], double ue[12], uint32_t nodes[4]) {
    for (int a = 0; a < 4; ++a) {
        const uint32_t n = conn[a * (size_t)ne + e];
        nodes[a] = n;
        for (int d = 0; d < 3; ++d) {
            Xe[a * 3 + d] = X[d * (size_t)nn + n];
            ue[a * 3 + d] = u ? u[3 * (size_t)n + d] : 0.0;
        }
    }
}

double dot(uint32_t n, const double* a, const double* b) {
    double s = 0.0;
    if (g_threads > 1) {
#pragma omp parallel for num_threads(g_threads) reduction(+ : s) schedule(static)
        for (int64_t i = 0; i < (int64_t)n; ++i) s += a[i] * b[i];
    } else {
        for (uint32_t i = 0; i < n; ++i) s += a[i] * b[i];
    }
    return s;
}

// colour buckets in colour order (element index ascending inside a colour)
void colour_buckets(uint32_t ne, const uint32_t* colors, uint32_t nc, std::vector<uint32_t>& ptr,
                    std::vector<uint32_t>& list) {
    ptr.assign(nc + 1, 0);
    for (uint32_t e = 0; e < ne; ++e) ptr[colors[e] + 1]++;
    for (uint32_t c = 0; c < nc; ++c) ptr[c + 1] += ptr[c];
    list.resize(ne);
    std::vector<uint32_t> fill(ptr.begin(), ptr.end() - 1);
    for (uint32_t e = 0; e < ne; ++e) list[fill[colors[e]]++] = e;
}

}  // namespace

extern "C" {

void orc_set_threads(int n) { g_threads = n < 1 ? 1 : n; }
int orc_get_threads(void) { return g_threads; }

// Restates generate_unit_cube_mesh (FEM/adaptive_fem_engine.hpp:2219-2294):
// node id k*NN*NN + j*NN + i (:2236-2238), node loop order k, j, i
// (:2227-2229), hex loop k, j, i, 6 tets per hex (:2250-2273); generalised to a
// box of nx*ny*nz hexes and extent Lx*Ly*Lz.
void orc_box_mesh(int nx, int ny, int nz, double Lx, double Ly, double Lz, double* X, uint32_t* conn) {
    const uint32_t NX = nx + 1, NY = ny + 1, NZ = nz + 1;
    const size_t nn = (size_t)NX * NY * NZ;
    const size_t ne = 6ull * nx * ny * nz;
    const double hx = Lx / nx, hy = Ly / ny, hz = Lz / nz;
    for (uint32_t k = 0; k < NZ; ++k)
        for (uint32_t j = 0; j < NY; ++j)
            for (uint32_t i = 0; i < NX; ++i) {
                const size_t n = (size_t)k * NX * NY + (size_t)j * NX + i;
                X[n] = i * hx;
                X[nn + n] = j * hy;
                X[2 * nn + n] = k * hz;
            }
    size_t e = 0;
    for (int k = 0; k < nz; ++k)
        for (int j = 0; j < ny; ++j)
            for (int i = 0; i < nx; ++i) {
                uint32_t hex[8];
                for (int c = 0; c < 8; ++c)
                    hex[c] = (uint32_t)((k + kHexOff[c][2]) * (size_t)NX * NY + (j + kHexOff[c][1]) * (size_t)NX +
                                        (i + kHexOff[c][0]));
                for (int t = 0; t < 6; ++t, ++e)
                    for (int v = 0; v < 4; ++v) conn[v * ne + e] = hex[kKuhn[t][v]];
            }
}

// Restates matrix_3x3_inv (FEM/adaptive_fem_engine.hpp:52-71) including its
// |det| < 1e-30 guard.
void orc_inv3(const double A[9], double Ainv[9]) {
    const double det = det3(A);
    if (std::fabs(det) < 1e-30) {
        std::memset(Ainv, 0, 9 * sizeof(double));
        Ainv[0] = Ainv[4] = Ainv[8] = 1e30;
        return;
    }
    const double id = 1.0 / det;
    Ainv[0] = (A[4] * A[8] - A[7] * A[5]) * id;
    Ainv[1] = -(A[1] * A[8] - A[7] * A[2]) * id;
    Ainv[2] = (A[1] * A[5] - A[4] * A[2]) * id;
    Ainv[3] = -(A[3] * A[8] - A[6] * A[5]) * id;
    Ainv[4] = (A[0] * A[8] - A[6] * A[2]) * id;
    Ainv[5] = -(A[0] * A[5] - A[3] * A[2]) * id;
    Ainv[6] = (A[3] * A[7] - A[6] * A[4]) * id;
    Ainv[7] = -(A[0] * A[7] - A[6] * A[1]) * id;
    Ainv[8] = (A[0] * A[4] - A[3] * A[1]) * id;
}

// Literal restatement of NeoHookeanMaterial::compute_pk1_stress
// (FEM/constitutive_models.hpp:61-113): C = F^T F, J = sqrt|det C|,
// S = mu (I - C^{-1}) + lam ln(max(J,0.01)) C^{-1}, P = F S.
void orc_pk1_refform(const double F[9], double mu, double lam, double P[9]) {
    double C[9] = {0};
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j)
            for (int k = 0; k < 3; ++k) C[i * 3 + j] += F[k * 3 + i] * F[k * 3 + j];
    const double det_C = C[0] * (C[4] * C[8] - C[7] * C[5]) - C[3] * (C[1] * C[8] - C[7] * C[2]) +
                         C[6] * (C[1] * C[5] - C[4] * C[2]);
    const double J = std::sqrt(std::fabs(det_C));
    double Ci[9];
    orc_inv3(C, Ci);
    const double lnJ = std::log(std::max(J, 0.01));
    double S[9];
    for (int i = 0; i < 9; ++i) {
        const double delta = (i % 4 == 0) ? 1.0 : 0.0;
        S[i] = mu * (delta - Ci[i]) + lam * lnJ * Ci[i];
    }
    std::memset(P, 0, 9 * sizeof(double));
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j)
            for (int k = 0; k < 3; ++k) P[i * 3 + j] += F[i * 3 + k] * S[k * 3 + j];
}

// The same law in the closed form used by the element kernels:
// P = mu F + (lam ln J - mu) F^{-T}, since F C^{-1} = F^{-T}.
void orc_pk1(const double F[9], double mu, double lam, double P[9]) {
    ElemKin K;
    std::memcpy(K.F, F, sizeof K.F);
    kin_from_F(K);
    const double s = lam * K.lnJ - mu;
    for (int k = 0; k < 9; ++k) P[k] = mu * F[k] + s * K.FiT[k];
}

// psi = mu/2 (I1 - 3) - mu ln J + lam/2 ln^2 J: the energy whose derivative is
// the stress above.  FEM/constitutive_models.hpp:156-178 uses 0.25 lam ln^2 J,
// which is inconsistent with its own stress; the consistent lam/2 is used.
double orc_energy(const double F[9], double mu, double lam) {
    double I1 = 0.0;
    for (int k = 0; k < 9; ++k) I1 += F[k] * F[k];
    const double J = std::fabs(det3(F));
    const double lnJ = std::log(std::max(J, 0.01));
    return 0.5 * mu * (I1 - 3.0) - mu * lnJ + 0.5 * lam * lnJ * lnJ;
}

int orc_element(const double Xe[12], const double ue[12], double mu, double lam, double fe[12], double Ke[144]) {
    ElemGeo G;
    if (!elem_geometry(Xe, G)) {
        std::memset(fe, 0, 12 * sizeof(double));
        if (Ke) std::memset(Ke, 0, 144 * sizeof(double));
        return 1;
    }
    ElemKin K;
    elem_kinematics(G, ue, K);
    elem_force_tangent(G, K, mu, lam, fe, Ke);
    return 0;
}

// Row structure of build_csr_sparsity_pattern_rigorous
// (FEM/adaptive_fem_engine.hpp:1069-1102): node adjacency through elements as
// a sorted set (:1077-1086); row 3n+d holds 3*|nbr(n)| entries (:1091-1102).
uint64_t orc_pattern_rows(uint32_t nn, uint32_t ne, const uint32_t* conn, uint32_t* row_ptr) {
    std::vector<uint32_t> ptr, nbr;
    node_neighbours(nn, ne, conn, ptr, nbr);
    row_ptr[0] = 0;
    uint64_t acc = 0;
    for (uint32_t n = 0; n < nn; ++n) {
        const uint32_t deg = ptr[n + 1] - ptr[n];
        for (int d = 0; d < 3; ++d) {
            acc += 3ull * deg;
            row_ptr[3 * n + d + 1] = (uint32_t)acc;
        }
    }
    return acc;
}

// Column fill with the intended rule: row 3n+d lists 3m, 3m+1, 3m+2 for each m
// in sorted nbr(n).  The reference fill (:1105-1119) writes the 3x3 block of
// each neighbour sequentially across the node's three rows, which leaves
// duplicate and missing columns; that defect is corrected here.
void orc_pattern_cols(uint32_t nn, uint32_t ne, const uint32_t* conn, const uint32_t* row_ptr, uint32_t* col_idx) {
    std::vector<uint32_t> ptr, nbr;
    node_neighbours(nn, ne, conn, ptr, nbr);
    for (uint32_t n = 0; n < nn; ++n)
        for (int d = 0; d < 3; ++d) {
            uint32_t k = row_ptr[3 * n + d];
            for (uint32_t q = ptr[n]; q < ptr[n + 1]; ++q)
                for (int e = 0; e < 3; ++e) col_idx[k++] = 3 * nbr[q] + e;
        }
}

// MeshColoring::compute (FEM/adaptive_fem_engine.hpp:1549-1602): elements in
// index order take the smallest colour not used by an earlier element sharing
// a node.  The O(n_e^2) all-pairs adjacency (:1559-1581) is replaced by the
// node->element lists; the assignment rule (:1584-1598) is unchanged.
uint32_t orc_color(uint32_t nn, uint32_t ne, const uint32_t* conn, uint32_t* colors) {
    std::vector<uint32_t> eptr(nn + 1, 0), elist;
    for (int a = 0; a < 4; ++a)
        for (uint32_t e = 0; e < ne; ++e) eptr[conn[a * (size_t)ne + e] + 1]++;
    for (uint32_t n = 0; n < nn; ++n) eptr[n + 1] += eptr[n];
    elist.resize(eptr[nn]);
    std::vector<uint32_t> fill(eptr.begin(), eptr.end() - 1);
    for (uint32_t e = 0; e < ne; ++e)
        for (int a = 0; a < 4; ++a) elist[fill[conn[a * (size_t)ne + e]]++] = e;
    std::vector<uint32_t> stamp;
    uint32_t ncol = 0;
    for (uint32_t e = 0; e < ne; ++e) {
        for (int a = 0; a < 4; ++a) {
            const uint32_t n = conn[a * (size_t)ne + e];
            for (uint32_t k = eptr[n]; k < eptr[n + 1]; ++k) {
                const uint32_t f = elist[k];
                if (f >= e) continue;
                const uint32_t c = colors[f];
                if (c >= stamp.size()) stamp.resize(c + 1, UINT32_MAX);
                stamp[c] = e;
            }
        }
        uint32_t c = 0;
        while (c < stamp.size() && stamp[c] == e) ++c;
        colors[e] = c;
        ncol = std::max(ncol, c + 1);
    }
    return ncol;
}

// assemble_global_threadlocal (FEM/adaptive_fem_engine.hpp:1126-1233) and
// assemble_global_residual_only (:990-1059): zero K and f, evaluate every
// element, add its 12x12 block at the CSR positions of its node pairs and its
// force at its nodes.  Elements are visited colour by colour (index order
// inside a colour), so each K and f entry receives its contributions in a
// fixed order that does not depend on the thread count.
void orc_assemble(uint32_t nn, uint32_t ne, const double* X, const uint32_t* conn, const double* mu,
                  const double* lam, const uint32_t* colors, uint32_t n_colors, const double* u,
                  const uint32_t* row_ptr, const uint32_t* col_idx, double* K, double* f_int) {
    const size_t n = 3ull * nn;
    std::memset(f_int, 0, n * sizeof(double));
    if (K) std::memset(K, 0, (size_t)row_ptr[n] * sizeof(double));
    std::vector<uint32_t> cptr, clist;
    colour_buckets(ne, colors, n_colors, cptr, clist);
    for (uint32_t c = 0; c < n_colors; ++c) {
#pragma omp parallel for num_threads(g_threads) schedule(static) if (g_threads > 1)
        for (int64_t t = cptr[c]; t < (int64_t)cptr[c + 1]; ++t) {
            const uint32_t e = clist[t];
            double Xe[12], ue[12], fe[12], Ke[144];
            uint32_t nodes[4];
            gather_elem(X, nn, conn, ne, e, u, Xe, ue, nodes);
            if (orc_element(Xe, ue, mu[e], lam[e], fe, K ? Ke : nullptr)) continue;
            for (int a = 0; a < 4; ++a)
                for (int i = 0; i < 3; ++i) f_int[3 * (size_t)nodes[a] + i] += fe[a * 3 + i];
            if (!K) continue;
            for (int a = 0; a < 4; ++a)
                for (int i = 0; i < 3; ++i) {
                    const uint32_t r = 3 * nodes[a] + i;
                    for (int b = 0; b < 4; ++b) {
                        const int64_t pos = find_col(row_ptr, col_idx, r, 3 * nodes[b]);
                        for (int k = 0; k < 3; ++k) K[pos + k] += Ke[(a * 3 + i) * 12 + b * 3 + k];
                    }
                }
        }
    }
}

// Dirichlet elimination as NonlinearNewtonSolver applies it
// (FEM/nonlinear_solver.hpp:388-396: zero row and column, unit diagonal),
// in place of the reference engine's 1e12 penalty (:710-730).
void orc_eliminate(uint32_t n, const uint32_t* row_ptr, const uint32_t* col_idx, double* K, const uint8_t* is_dir) {
#pragma omp parallel for num_threads(g_threads) schedule(static) if (g_threads > 1)
    for (int64_t r = 0; r < (int64_t)n; ++r)
        for (uint32_t k = row_ptr[r]; k < row_ptr[r + 1]; ++k) {
            const uint32_t c = col_idx[k];
            if (is_dir[r] || is_dir[c]) K[k] = (c == (uint32_t)r) ? 1.0 : 0.0;
        }
}

// SparseCSR::matvec (FEM/fem_types.hpp:310-319).
void orc_spmv(uint32_t n, const uint32_t* row_ptr, const uint32_t* col_idx, const double* val, const double* x,
              double* y) {
#pragma omp parallel for num_threads(g_threads) schedule(static) if (g_threads > 1)
    for (int64_t i = 0; i < (int64_t)n; ++i) {
        double s = 0.0;
        for (uint32_t k = row_ptr[i]; k < row_ptr[i + 1]; ++k) s += val[k] * x[col_idx[k]];
        y[i] = s;
    }
}

// ILU0Preconditioner::factorize (FEM/adaptive_fem_engine.hpp:305-470).  Its
// pivot-row search (:327-343) always selects the lowest unprocessed row (every
// L dependency of row i lies below i), so the factorization is natural-order
// Doolittle ILU(0).  For each row i, in ascending column order k < i:
//   L(i,k) = (A_ik - sum_{l<k} L(i,l) U(l,k)) / u_kk          (:427-466)
//   U(i,j) = A_ij - sum_{l<i} L(i,l) U(l,j), j > i             (:392-425)
//   u_ii   = A_ii - sum_{k<i} L(i,k) U(k,i), then the pivot rule
//            |u_ii| < 1e-12 -> u_ii +/- 1e-8                     (:358-388)
// with every sum subtracted term by term in ascending l, which is the order a
// row-wise IKJ sweep produces.  Columns must be sorted within each row.
void orc_ilu0(uint32_t n, const uint32_t* row_ptr, const uint32_t* col_idx, const double* val, double* LU,
              double* diag) {
    std::vector<double> w(n, 0.0);
    std::vector<uint32_t> mark(n, UINT32_MAX);
    std::vector<uint32_t> dpos(n, UINT32_MAX);
    for (uint32_t i = 0; i < n; ++i) {
        const uint32_t b = row_ptr[i], e = row_ptr[i + 1];
        for (uint32_t k = b; k < e; ++k) {
            w[col_idx[k]] = val[k];
            mark[col_idx[k]] = i;
        }
        for (uint32_t kk = b; kk < e; ++kk) {
            const uint32_t k = col_idx[kk];
            if (k >= i) break;
            const double l = w[k] / diag[k];
            w[k] = l;
            // U part of row k (columns > k)
            for (uint32_t t = dpos[k] + 1; t < row_ptr[k + 1]; ++t) {
                const uint32_t j = col_idx[t];
                if (mark[j] == i) w[j] -= l * LU[t];
            }
        }
        double uii = (mark[i] == i) ? w[i] : 0.0;
        if (std::fabs(uii) < 1e-12) uii += (uii >= 0 ? 1e-8 : -1e-8);
        diag[i] = uii;
        for (uint32_t k = b; k < e; ++k) {
            const uint32_t c = col_idx[k];
            LU[k] = (c == i) ? uii : w[c];
            if (c == i) dpos[i] = k;
        }
        if (dpos[i] == UINT32_MAX) {
            // no stored diagonal: upper part starts at the first column > i
            uint32_t k = b;
            while (k < e && col_idx[k] < i) ++k;
            dpos[i] = k - 1;
        }
    }
}

// ILU0Preconditioner::solve (FEM/adaptive_fem_engine.hpp:473-504): forward
// z_i = x_i - sum_{j<i} L(i,j) z_j, then backward
// y_i = (z_i - sum_{j>i} U(i,j) y_j) / u_ii, ascending j inside each sum.
void orc_ilu0_apply(uint32_t n, const uint32_t* row_ptr, const uint32_t* col_idx, const double* LU,
                    const double* diag, const double* x, double* y) {
    std::vector<double> z(n);
    for (uint32_t i = 0; i < n; ++i) {
        double s = x[i];
        for (uint32_t k = row_ptr[i]; k < row_ptr[i + 1] && col_idx[k] < i; ++k) s -= LU[k] * z[col_idx[k]];
        z[i] = s;
    }
    for (int64_t i = (int64_t)n - 1; i >= 0; --i) {
        double s = z[i];
        for (uint32_t k = row_ptr[i]; k < row_ptr[i + 1]; ++k) {
            const uint32_t c = col_idx[k];
            if (c > (uint32_t)i) s -= LU[k] * y[c];
        }
        y[i] = s / diag[i];
    }
}

// pcg_solve_ilu0 (FEM/adaptive_fem_engine.hpp:1283-1341), same operation order
// and stopping rule sqrt(r.z / b.b) < tol.  relres_out receives sqrt(r.z/b.b)
// at exit (or ||r0|| for the early exit).
int orc_pcg(uint32_t n, const uint32_t* row_ptr, const uint32_t* col_idx, const double* val, const double* b,
            double* x, double tol, int max_iter, double* relres_out) {
    const uint64_t nnz = row_ptr[n];
    std::vector<double> LU(nnz), diag(n);
    orc_ilu0(n, row_ptr, col_idx, val, LU.data(), diag.data());
    std::vector<double> r(n), p(n), Ap(n), z(n);
    orc_spmv(n, row_ptr, col_idx, val, x, Ap.data());
    double r_norm_sq = 0.0;
    for (uint32_t i = 0; i < n; ++i) {
        r[i] = b[i] - Ap[i];
        r_norm_sq += r[i] * r[i];
    }
    if (std::sqrt(r_norm_sq) < tol * 1e-3) {
        if (relres_out) *relres_out = std::sqrt(r_norm_sq);
        return 0;
    }
    orc_ilu0_apply(n, row_ptr, col_idx, LU.data(), diag.data(), r.data(), z.data());
    double rz = 0.0;
    for (uint32_t i = 0; i < n; ++i) {
        p[i] = z[i];
        rz += r[i] * z[i];
    }
    double bb = 0.0;
    for (uint32_t i = 0; i < n; ++i) bb += b[i] * b[i];
    double rel = std::sqrt(rz / std::max(bb, 1e-300));
    for (int it = 0; it < max_iter; ++it) {
        orc_spmv(n, row_ptr, col_idx, val, p.data(), Ap.data());
        const double pAp = dot(n, p.data(), Ap.data());
        if (std::fabs(pAp) < 1e-60) break;
        const double alpha = rz / pAp;
#pragma omp parallel for num_threads(g_threads) schedule(static) if (g_threads > 1)
        for (int64_t i = 0; i < (int64_t)n; ++i) {
            x[i] += alpha * p[i];
            r[i] -= alpha * Ap[i];
        }
        orc_ilu0_apply(n, row_ptr, col_idx, LU.data(), diag.data(), r.data(), z.data());
        const double rz_new = dot(n, r.data(), z.data());
        rel = std::sqrt(rz_new / std::max(bb, 1e-300));
        if (rel < tol) {
            if (relres_out) *relres_out = rel;
            return it + 1;
        }
        const double beta = rz_new / rz;
#pragma omp parallel for num_threads(g_threads) schedule(static) if (g_threads > 1)
        for (int64_t i = 0; i < (int64_t)n; ++i) p[i] = z[i] + beta * p[i];
        rz = rz_new;
    }
    if (relres_out) *relres_out = rel;
    return max_iter;
}

}  // extern "C"

namespace {

struct NewtonWork {
    std::vector<double> K, f, R, du, rhs, ut;
    std::vector<uint8_t> is_dir;
};

double residual(const orc_problem* P, double lf, const double* u, const uint8_t* is_dir, std::vector<double>& f,
                std::vector<double>& R) {
    const size_t n = 3ull * P->n_nodes;
    orc_assemble(P->n_nodes, P->n_elems, P->X, P->conn, P->mu, P->lam, P->colors, P->n_colors, u, P->row_ptr,
                 P->col_idx, nullptr, f.data());
    for (size_t i = 0; i < n; ++i) R[i] = is_dir[i] ? 0.0 : (P->f_ext ? lf * P->f_ext[i] : 0.0) - f[i];
    return std::sqrt(dot((uint32_t)n, R.data(), R.data()));
}

}  // namespace

extern "C" {

// Potential energy: sum over elements of V_e psi(F_e) (orc_energy) minus the
// work lf f_ext.u; its gradient is -R, so a Newton direction with R.du > 0 is
// a descent direction.  Elements in index order (serial) for a fixed sum.
double orc_total_energy(const orc_problem* P, double lf, const double* u) {
    const uint32_t nn = P->n_nodes, ne = P->n_elems;
    double e_int = 0.0;
    for (uint32_t e = 0; e < ne; ++e) {
        double Xe[12], ue[12];
        uint32_t nodes[4];
        gather_elem(P->X, nn, P->conn, ne, e, u, Xe, ue, nodes);
        ElemGeo G;
        if (!elem_geometry(Xe, G)) continue;
        ElemKin K;
        elem_kinematics(G, ue, K);
        e_int += G.vol * orc_energy(K.F, P->mu[e], P->lam[e]);
    }
    double w = 0.0;
    if (P->f_ext)
        for (size_t i = 0; i < 3ull * nn; ++i) w += P->f_ext[i] * u[i];
    return e_int - lf * w;
}

// armijo_line_search_true (FEM/adaptive_fem_engine.hpp:933-985): alpha = 1,
// 1/2, ... for at most ls_max trials; accept the first alpha with
// ||R(u + alpha du)|| <= (1 - c alpha) ||R(u)|| + 1e-14; otherwise return the
// last halved alpha.  With ls_mode = 1 a trial is also accepted when the
// energy satisfies Pi(u + alpha du) <= Pi(u) - c alpha R.du (R.du > 0): the
// residual norm may grow on the way into the quadratic basin (slender bodies
// under traction) while the energy decreases.  u_out receives u + alpha du.
double orc_armijo(const orc_problem* P, double lf, const double* u, const double* du, double R_norm, double* u_out,
                  double* R_norm_out) {
    const size_t n = 3ull * P->n_nodes;
    std::vector<uint8_t> is_dir(n, 0);
    for (uint32_t k = 0; k < P->n_dir; ++k) is_dir[P->dir_dofs[k]] = 1;
    std::vector<double> f(n), R(n);
    double pi0 = 0.0, slope = 0.0;
    if (P->ls_mode == 1) {
        pi0 = orc_total_energy(P, lf, u);
        residual(P, lf, u, is_dir.data(), f, R);
        for (size_t i = 0; i < n; ++i) slope += R[i] * du[i];
    }
    double alpha = 1.0;
    for (int bt = 0; bt < P->ls_max; ++bt) {
        for (size_t i = 0; i < n; ++i) u_out[i] = u[i] + alpha * du[i];
        const double Rt = residual(P, lf, u_out, is_dir.data(), f, R);
        if (R_norm_out) *R_norm_out = Rt;
        if (Rt <= (1.0 - P->armijo_c * alpha) * R_norm + 1e-14) return alpha;
        if (P->ls_mode == 1 && slope > 0.0 &&
            orc_total_energy(P, lf, u_out) <= pi0 - P->armijo_c * alpha * slope)
            return alpha;
        alpha *= 0.5;
    }
    for (size_t i = 0; i < n; ++i) u_out[i] = u[i] + alpha * du[i];
    if (R_norm_out) *R_norm_out = residual(P, lf, u_out, is_dir.data(), f, R);
    return alpha;
}

// One load step s (1-based) of the Newton loop of AdaptiveFEMEngine::run
// (FEM/adaptive_fem_engine.hpp:2456-2592) with the corrections documented in
// the design notes: load factor lf = s / n_steps (:2457); Dirichlet values are
// lf * dir_vals and eliminated (FEM/nonlinear_solver.hpp:329-334, 388-396);
// convergence is ||R_k|| <= rtol ||R_0|| (nonlinear_solver.hpp:124-127) with
// R_0 evaluated at the previous state with this step's prescribed values set;
// with lift = 1 the step starts with a predictor solve K du = R - K[:,D] dD;
// each Newton iteration solves K du = R by pcg_solve_ilu0 (:2577-2578) and
// takes an Armijo step (:2581-2586).
int orc_newton_step(const orc_problem* P, uint32_t step, double* u, orc_step_stats* st) {
    const double t_begin = now_s();
    const uint32_t nn = P->n_nodes;
    const size_t n = 3ull * nn;
    const uint64_t nnz = P->row_ptr[n];
    const double lf = (double)step / (double)P->n_steps;
    orc_step_stats S;
    std::memset(&S, 0, sizeof S);
    NewtonWork W;
    W.K.resize(nnz);
    W.f.resize(n);
    W.R.resize(n);
    W.du.resize(n);
    W.rhs.resize(n);
    W.ut.resize(n);
    W.is_dir.assign(n, 0);
    for (uint32_t k = 0; k < P->n_dir; ++k) W.is_dir[P->dir_dofs[k]] = 1;

    // prescribed increment
    std::vector<double> dD(n, 0.0);
    bool any_inc = false;
    for (uint32_t k = 0; k < P->n_dir; ++k) {
        const uint32_t d = P->dir_dofs[k];
        dD[d] = lf * P->dir_vals[k] - u[d];
        if (dD[d] != 0.0) any_inc = true;
    }
    // R_0 at the previous state with this step's prescribed values
    for (size_t i = 0; i < n; ++i) W.ut[i] = u[i] + dD[i];
    double t0 = now_s();
    S.r0 = residual(P, lf, W.ut.data(), W.is_dir.data(), W.f, W.R);
    S.t_residual += now_s() - t0;
    S.residuals++;

    if (P->lift && any_inc) {
        t0 = now_s();
        orc_assemble(nn, P->n_elems, P->X, P->conn, P->mu, P->lam, P->colors, P->n_colors, u, P->row_ptr,
                     P->col_idx, W.K.data(), W.f.data());
        S.t_assemble += now_s() - t0;
        S.assemblies++;
        orc_spmv((uint32_t)n, P->row_ptr, P->col_idx, W.K.data(), dD.data(), W.du.data());
        for (size_t i = 0; i < n; ++i)
            W.rhs[i] = W.is_dir[i] ? 0.0 : (P->f_ext ? lf * P->f_ext[i] : 0.0) - W.f[i] - W.du[i];
        orc_eliminate((uint32_t)n, P->row_ptr, P->col_idx, W.K.data(), W.is_dir.data());
        std::fill(W.du.begin(), W.du.end(), 0.0);
        t0 = now_s();
        S.pcg_iters += orc_pcg((uint32_t)n, P->row_ptr, P->col_idx, W.K.data(), W.rhs.data(), W.du.data(),
                               P->pcg_tol, P->pcg_max, nullptr);
        S.t_pcg += now_s() - t0;
        S.factorizations++;
        for (size_t i = 0; i < n; ++i) u[i] += W.is_dir[i] ? dD[i] : W.du[i];
    } else {
        for (size_t i = 0; i < n; ++i) u[i] += dD[i];
    }

    t0 = now_s();
    double rn = residual(P, lf, u, W.is_dir.data(), W.f, W.R);
    S.t_residual += now_s() - t0;
    S.residuals++;
    const double target = P->newton_rtol * S.r0;
    int it = 0;
    while (!(rn <= target) && it < P->max_newton) {
        t0 = now_s();
        orc_assemble(nn, P->n_elems, P->X, P->conn, P->mu, P->lam, P->colors, P->n_colors, u, P->row_ptr,
                     P->col_idx, W.K.data(), W.f.data());
        S.t_assemble += now_s() - t0;
        S.assemblies++;
        orc_eliminate((uint32_t)n, P->row_ptr, P->col_idx, W.K.data(), W.is_dir.data());
        std::fill(W.du.begin(), W.du.end(), 0.0);
        t0 = now_s();
        S.pcg_iters += orc_pcg((uint32_t)n, P->row_ptr, P->col_idx, W.K.data(), W.R.data(), W.du.data(),
                               P->pcg_tol, P->pcg_max, nullptr);
        S.t_pcg += now_s() - t0;
        S.factorizations++;
        t0 = now_s();
        double rnew = 0.0;
        orc_armijo(P, lf, u, W.du.data(), rn, W.ut.data(), &rnew);
        S.t_residual += now_s() - t0;
        std::memcpy(u, W.ut.data(), n * sizeof(double));
        rn = residual(P, lf, u, W.is_dir.data(), W.f, W.R);
        S.residuals++;
        ++it;
    }
    S.newton_iters = it;
    S.converged = rn <= target;
    S.rel_res = S.r0 > 0 ? rn / S.r0 : 0.0;
    S.t_total = now_s() - t_begin;
    if (st) *st = S;
    return S.converged ? 0 : 1;
}

// ---------------------------------------------------------------------------
// a20: FP64 rotation part.  The reference takes exp(alpha skew(log F))
// (FEM/state_transport.hpp:201-249) with series log/exp that lose accuracy at
// large angles (CORE/lie_operator.cpp:94-275).  Here the rotation part is the
// polar factor R of F (scaled Newton iteration, Higham 1986), its logarithm is
// the angle-axis vector and the exponential is Rodrigues' formula.
// ---------------------------------------------------------------------------
int orc_polar(const double F[9], double R[9]) {
    double X[9];
    std::memcpy(X, F, sizeof X);
    for (int it = 0; it < 60; ++it) {
        double Xi[9];
        const double det = det3(X);
        if (std::fabs(det) < 1e-300) return 1;
        orc_inv3(X, Xi);
        double nx = 0, ni = 0;
        for (int k = 0; k < 9; ++k) {
            nx += X[k] * X[k];
            ni += Xi[k] * Xi[k];
        }
        const double zeta = (it < 8) ? std::sqrt(std::sqrt(ni / nx)) : 1.0;
        double Xn[9], diff = 0.0;
        for (int i = 0; i < 3; ++i)
            for (int j = 0; j < 3; ++j) {
                Xn[i * 3 + j] = 0.5 * (zeta * X[i * 3 + j] + Xi[j * 3 + i] / zeta);
                diff += (Xn[i * 3 + j] - X[i * 3 + j]) * (Xn[i * 3 + j] - X[i * 3 + j]);
            }
        std::memcpy(X, Xn, sizeof X);
        if (it >= 8 && diff < 1e-30) break;
    }
    std::memcpy(R, X, sizeof X);
    return det3(R) > 0 ? 0 : 2;
}

void orc_rot_log(const double R[9], double w[3]) {
    const double c = 0.5 * (R[0] + R[4] + R[8] - 1.0);
    const double v[3] = {0.5 * (R[7] - R[5]), 0.5 * (R[2] - R[6]), 0.5 * (R[3] - R[1])};
    const double s = std::sqrt(v[0] * v[0] + v[1] * v[1] + v[2] * v[2]);
    const double th = std::atan2(s, c);
    if (th < 1e-4) {
        const double t2 = th * th;
        const double f = 1.0 + t2 / 6.0 + 7.0 * t2 * t2 / 360.0;  // theta / sin(theta)
        for (int i = 0; i < 3; ++i) w[i] = f * v[i];
        return;
    }
    if (c > -0.9) {
        const double f = th / s;
        for (int i = 0; i < 3; ++i) w[i] = f * v[i];
        return;
    }
    // near pi: axis from the symmetric part (R + R^T)/2 - c I = (1 - c) n n^T
    double B[9];
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j) B[i * 3 + j] = 0.5 * (R[i * 3 + j] + R[j * 3 + i]) - (i == j ? c : 0.0);
    int m = 0;
    if (B[4] > B[m * 4]) m = 1;
    if (B[8] > B[m * 4]) m = 2;
    double nrm = 0.0, nv[3];
    for (int i = 0; i < 3; ++i) {
        nv[i] = B[i * 3 + m];
        nrm += nv[i] * nv[i];
    }
    nrm = std::sqrt(nrm);
    double sgn = nv[0] * v[0] + nv[1] * v[1] + nv[2] * v[2];
    sgn = (sgn < 0) ? -1.0 : 1.0;
    for (int i = 0; i < 3; ++i) w[i] = sgn * th * nv[i] / nrm;
}

void orc_rot_exp(const double w[3], double R[9]) {
    const double th2 = w[0] * w[0] + w[1] * w[1] + w[2] * w[2];
    const double th = std::sqrt(th2);
    double a, b;
    if (th < 1e-6) {
        a = 1.0 - th2 / 6.0;
        b = 0.5 - th2 / 24.0;
    } else {
        a = std::sin(th) / th;
        const double h = std::sin(0.5 * th) / (0.5 * th);
        b = 0.5 * h * h;  // (1 - cos th) / th^2 without cancellation
    }
    const double K[9] = {0, -w[2], w[1], w[2], 0, -w[0], -w[1], w[0], 0};
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j) {
            double k2 = 0.0;
            for (int k = 0; k < 3; ++k) k2 += K[i * 3 + k] * K[k * 3 + j];
            R[i * 3 + j] = (i == j ? 1.0 : 0.0) + a * K[i * 3 + j] + b * k2;
        }
}

// a22: adjoint of the final-load equilibrium R(u, p) = s f_ext - f_int(u; mu, lam)
// with u_D = s dir_vals.  K lambda = dL/du on free DOFs (K symmetric), then
// dL/dmu_e = -lambda_e . V (F - F^{-T}) g, dL/dlam_e = -lambda_e . V lnJ F^{-T} g,
// dL/ds = lambda . (f_ext - K[:,D] u_D) + dL/du_D . u_D.
int orc_adjoint(const orc_problem* P, const double* u, const double* dLdu, double tol, int max_iter,
                double* dL_dmu, double* dL_dlam, double* dL_dload, double* lam_out) {
    const uint32_t nn = P->n_nodes, ne = P->n_elems;
    const size_t n = 3ull * nn;
    const uint64_t nnz = P->row_ptr[n];
    std::vector<double> K(nnz), f(n), g(n), lam(n, 0.0), uD(n, 0.0), KuD(n);
    std::vector<uint8_t> is_dir(n, 0);
    for (uint32_t k = 0; k < P->n_dir; ++k) {
        is_dir[P->dir_dofs[k]] = 1;
        uD[P->dir_dofs[k]] = P->dir_vals[k];
    }
    orc_assemble(nn, ne, P->X, P->conn, P->mu, P->lam, P->colors, P->n_colors, u, P->row_ptr, P->col_idx,
                 K.data(), f.data());
    orc_spmv((uint32_t)n, P->row_ptr, P->col_idx, K.data(), uD.data(), KuD.data());
    orc_eliminate((uint32_t)n, P->row_ptr, P->col_idx, K.data(), is_dir.data());
    for (size_t i = 0; i < n; ++i) g[i] = is_dir[i] ? 0.0 : dLdu[i];
    const int its = orc_pcg((uint32_t)n, P->row_ptr, P->col_idx, K.data(), g.data(), lam.data(), tol, max_iter,
                            nullptr);
    if (lam_out) std::memcpy(lam_out, lam.data(), n * sizeof(double));
    double dl = 0.0;
    for (size_t i = 0; i < n; ++i) {
        if (is_dir[i])
            dl += dLdu[i] * uD[i];
        else
            dl += lam[i] * ((P->f_ext ? P->f_ext[i] : 0.0) - KuD[i]);
    }
    if (dL_dload) *dL_dload = dl;
#pragma omp parallel for num_threads(g_threads) schedule(static) if (g_threads > 1)
    for (int64_t e = 0; e < (int64_t)ne; ++e) {
        double Xe[12], ue[12];
        uint32_t nodes[4];
        gather_elem(P->X, nn, P->conn, ne, (uint32_t)e, u, Xe, ue, nodes);
        ElemGeo G;
        if (!elem_geometry(Xe, G)) {
            dL_dmu[e] = 0.0;
            dL_dlam[e] = 0.0;
            continue;
        }
        ElemKin Kk;
        elem_kinematics(G, ue, Kk);
        double dm = 0.0, dlm = 0.0;
        for (int a = 0; a < 4; ++a)
            for (int i = 0; i < 3; ++i) {
                double Fg = 0.0, Ag = 0.0;
                for (int J = 0; J < 3; ++J) {
                    Fg += Kk.F[i * 3 + J] * G.g[a][J];
                    Ag += Kk.FiT[i * 3 + J] * G.g[a][J];
                }
                const double le = lam[3 * (size_t)nodes[a] + i];
                dm += le * G.vol * (Fg - Ag);
                dlm += le * G.vol * Kk.lnJ * Ag;
            }
        dL_dmu[e] = -dm;
        dL_dlam[e] = -dlm;
    }
    return its;
}

// ---------------------------------------------------------------------------
// Element fields of a displacement: the downstream consumer of the converged
// state.  F follows update_element_states_rigorous
// (FEM/adaptive_fem_engine.hpp:2716-2774) with its defect D9 corrected: F is
// stored as computed (no projection onto det F = 1, which is wrong for a
// compressible law) and J = det F is evaluated, not passed uninitialised.
// StressExtractor::compute_deformation_gradient_at_element
// (FEM/stress_extraction.hpp:39-91) uses constant shape gradients that ignore
// the geometry; the exact P1 gradient of the element kernels is used instead.
// ---------------------------------------------------------------------------

// Eigenvalues of a symmetric 3x3 matrix (row-major; the upper triangle is
// read), descending, by cyclic Jacobi rotations (Golub & Van Loan, Alg. 8.5.3)
// until the off-diagonal part is below 1e-18 of the matrix norm (the diagonal
// then holds the eigenvalues to rounding) or 60 sweeps.
void orc_sym_eig3(const double A_in[9], double w[3]) {
    double A[3][3];
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j) A[i][j] = A_in[std::min(i, j) * 3 + std::max(i, j)];
    for (int sweep = 0; sweep < 60; ++sweep) {
        const double off = A[0][1] * A[0][1] + A[0][2] * A[0][2] + A[1][2] * A[1][2];
        const double dg = A[0][0] * A[0][0] + A[1][1] * A[1][1] + A[2][2] * A[2][2];
        if (off <= 1e-36 * (dg + 2.0 * off)) break;
        for (int p = 0; p < 2; ++p)
            for (int q = p + 1; q < 3; ++q) {
                if (A[p][q] == 0.0) continue;
                const double theta = (A[q][q] - A[p][p]) / (2.0 * A[p][q]);
                const double t = (theta >= 0 ? 1.0 : -1.0) / (std::fabs(theta) + std::sqrt(theta * theta + 1.0));
                const double c = 1.0 / std::sqrt(t * t + 1.0), sn = t * c;
                for (int k = 0; k < 3; ++k) {  // A <- A J (columns p, q)
                    const double akp = A[k][p], akq = A[k][q];
                    A[k][p] = c * akp - sn * akq;
                    A[k][q] = sn * akp + c * akq;
                }
                for (int k = 0; k < 3; ++k) {  // A <- J^T A (rows p, q)
                    const double apk = A[p][k], aqk = A[q][k];
                    A[p][k] = c * apk - sn * aqk;
                    A[q][k] = sn * apk + c * aqk;
                }
            }
    }
    w[0] = A[0][0];
    w[1] = A[1][1];
    w[2] = A[2][2];
    std::sort(w, w + 3, [](double a, double b) { return a > b; });
}

// NeoHookeanMaterial::compute_cauchy_stress (FEM/constitutive_models.hpp:116-148):
// sigma = P F^T / max(|det F|, 0.01), with P the stress law of orc_pk1.
void orc_cauchy(const double F[9], double mu, double lam, double sigma[9]) {
    double P[9];
    orc_pk1(F, mu, lam, P);
    const double inv_J = 1.0 / std::max(std::fabs(det3(F)), 0.01);
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j) {
            double t = 0.0;
            for (int k = 0; k < 3; ++k) t += P[i * 3 + k] * F[j * 3 + k];
            sigma[i * 3 + j] = inv_J * t;
        }
}

// StressExtractor::compute_von_mises (FEM/stress_extraction.hpp:143-157):
// sqrt(3/2 dev:dev) with the off-diagonal terms of the upper triangle doubled.
double orc_von_mises(const double s[9]) {
    const double p = (s[0] + s[4] + s[8]) / 3.0;
    const double d0 = s[0] - p, d4 = s[4] - p, d8 = s[8] - p;
    const double n2 = d0 * d0 + d4 * d4 + d8 * d8 + 2.0 * (s[1] * s[1] + s[2] * s[2] + s[5] * s[5]);
    return std::sqrt(1.5 * n2);
}

// StressExtractor::extract_element_fields (FEM/stress_extraction.hpp:187-240)
// per element: det F (:229-231), Cauchy stress as (xx, yy, zz, yz, xz, xy)
// from the upper triangle, von Mises, pressure tr(sigma)/3 (:226), largest
// principal stress (:234-236), strain energy (orc_energy), maximum stretch
// sqrt(max(eig_max(F^T F), 1)) (:160-176), and F.  The reference's power
// iteration (:98-140) converges to the eigenvalue of largest magnitude from a
// fixed start vector; the exact algebraically largest eigenvalue is used.
// Degenerate elements (volume < 1e-30, skipped by :2750) give zeros.  Outputs
// are per element in element order; any pointer may be null.
void orc_element_fields(uint32_t nn, const double* X, uint32_t ne, const uint32_t* conn, const double* mu,
                        const double* lam, const double* u, double* det_F, double* cauchy, double* vm,
                        double* pressure, double* s1, double* psi, double* stretch, double* Fout) {
#pragma omp parallel for num_threads(g_threads) schedule(static)
    for (int64_t ei = 0; ei < (int64_t)ne; ++ei) {
        const uint32_t e = (uint32_t)ei;
        double Xe[12], ue[12];
        uint32_t nodes[4];
        gather_elem(X, nn, conn, ne, e, u, Xe, ue, nodes);
        ElemGeo G;
        double F[9] = {0}, sig[9] = {0}, f[8] = {0};
        if (elem_geometry(Xe, G)) {
            ElemKin K;
            elem_kinematics(G, ue, K);
            std::memcpy(F, K.F, sizeof F);
            orc_cauchy(F, mu[e], lam[e], sig);
            double w[3], Cm[9] = {0};
            f[0] = det3(F);
            f[1] = orc_von_mises(sig);
            f[2] = (sig[0] + sig[4] + sig[8]) / 3.0;
            orc_sym_eig3(sig, w);
            f[3] = w[0];
            f[4] = orc_energy(F, mu[e], lam[e]);
            for (int i = 0; i < 3; ++i)
                for (int j = 0; j < 3; ++j)
                    for (int k = 0; k < 3; ++k) Cm[i * 3 + j] += F[k * 3 + i] * F[k * 3 + j];
            orc_sym_eig3(Cm, w);
            f[5] = std::sqrt(std::max(w[0], 1.0));
        }
        if (det_F) det_F[e] = f[0];
        if (vm) vm[e] = f[1];
        if (pressure) pressure[e] = f[2];
        if (s1) s1[e] = f[3];
        if (psi) psi[e] = f[4];
        if (stretch) stretch[e] = f[5];
        if (cauchy) {
            const double c6[6] = {sig[0], sig[4], sig[8], sig[5], sig[2], sig[1]};
            for (int k = 0; k < 6; ++k) cauchy[(size_t)k * ne + e] = c6[k];
        }
        if (Fout)
            for (int k = 0; k < 9; ++k) Fout[(size_t)k * ne + e] = F[k];
    }
}

// ---------------------------------------------------------------------------
// ZZErrorEstimator::estimate_with_bounds (FEM/error_estimator.hpp:128-232).
// Raw element "gradient" g_e = (F11 - 1, F22 - 1, F33 - 1) of the element
// state (:141-146), volume-weighted node averages (:148-164), an SPR linear fit
// [1, x, y, z] over each node's patch of at most 32 incident elements in
// element order (:30-99, :166-196), element indicators
// sqrt(sum over the 4 nodes of |recovered - average|^2 / 4) cbrt(V_e) and the
// global sqrt(sum V_e sum/4) (:198-214).  Corrections: F is the exact,
// unprojected gradient (see orc_element_fields); the fit uses coordinates relative to the
// node, so its constant term is the recovered value at the node (the
// reference fits in global coordinates and takes the value at the origin,
// contrary to its own comment at :94-96); a patch whose normal equations are
// singular to 1e-10 of their diagonal (fewer than four non-coplanar
// centroids, e.g. at corners, where the reference's absolute 1e-30 pivot test
// divides by rounding noise) takes the constant fit, the patch mean.
// ---------------------------------------------------------------------------
namespace {
// SPR fit of one patch: m samples (centroid offsets d, raw gradients r);
// returns the value of the fit at offset 0
void spr_fit(int m, const double (*d)[3], const double (*r)[3], double g[3]) {
    if (m == 0) {
        g[0] = g[1] = g[2] = 0.0;
        return;
    }
    double M[4][7] = {};
    for (int e = 0; e < m; ++e) {
        const double b[4] = {1.0, d[e][0], d[e][1], d[e][2]};
        for (int j = 0; j < 4; ++j) {
            for (int k = 0; k < 4; ++k) M[j][k] += b[j] * b[k];
            for (int c = 0; c < 3; ++c) M[j][4 + c] += b[j] * r[e][c];
        }
    }
    double diag[4];
    for (int j = 0; j < 4; ++j) diag[j] = M[j][j];
    bool singular = false;
    // elimination without pivoting, as :64-82 (the normal matrix is SPD)
    for (int p = 0; p < 4 && !singular; ++p) {
        if (!(std::fabs(M[p][p]) > 1e-10 * diag[p])) {
            singular = true;
            break;
        }
        const double ip = 1.0 / M[p][p];
        for (int rr = p + 1; rr < 4; ++rr) {
            const double f = M[rr][p] * ip;
            for (int c = p; c < 7; ++c) M[rr][c] -= f * M[p][c];
        }
    }
    if (singular) {
        for (int c = 0; c < 3; ++c) {
            double s = 0.0;
            for (int e = 0; e < m; ++e) s += r[e][c];
            g[c] = s / m;
        }
        return;
    }
    double sol[4][3];
    for (int p = 3; p >= 0; --p)
        for (int c = 0; c < 3; ++c) {
            double rhs = M[p][4 + c];
            for (int k = p + 1; k < 4; ++k) rhs -= M[p][k] * sol[k][c];
            sol[p][c] = rhs / M[p][p];
        }
    for (int c = 0; c < 3; ++c) g[c] = sol[0][c];
}
}  // namespace

double orc_zz_estimate(uint32_t nn, const double* X, uint32_t ne, const uint32_t* conn, const double* u,
                       double* elem_err, double* raw_node, double* recovered) {
    std::vector<double> g(3 * (size_t)ne, 0.0), vol(ne, 0.0), cen(3 * (size_t)ne, 0.0);
#pragma omp parallel for num_threads(g_threads) schedule(static)
    for (int64_t ei = 0; ei < (int64_t)ne; ++ei) {
        const uint32_t e = (uint32_t)ei;
        double Xe[12], ue[12];
        uint32_t nodes[4];
        gather_elem(X, nn, conn, ne, e, u, Xe, ue, nodes);
        for (int a = 0; a < 4; ++a)
            for (int c = 0; c < 3; ++c) cen[3 * (size_t)e + c] += 0.25 * Xe[a * 3 + c];
        ElemGeo G;
        const bool ok = elem_geometry(Xe, G);
        double J[9];
        for (int c = 0; c < 3; ++c) {
            J[c * 3 + 0] = Xe[3 + c] - Xe[c];
            J[c * 3 + 1] = Xe[6 + c] - Xe[c];
            J[c * 3 + 2] = Xe[9 + c] - Xe[c];
        }
        vol[e] = std::fabs(det3(J)) / 6.0;
        if (ok) {
            ElemKin K;
            elem_kinematics(G, ue, K);
            for (int c = 0; c < 3; ++c) g[3 * (size_t)e + c] = K.F[c * 4] - 1.0;
        }
    }
    // node -> elements in element order (generate_unit_cube_mesh :2279-2291)
    std::vector<uint32_t> ptr(nn + 1, 0), data(4 * (size_t)ne);
    for (uint32_t e = 0; e < ne; ++e)
        for (int a = 0; a < 4; ++a) ptr[conn[(size_t)a * ne + e] + 1]++;
    for (uint32_t v = 0; v < nn; ++v) ptr[v + 1] += ptr[v];
    {
        std::vector<uint32_t> fill(ptr.begin(), ptr.end() - 1);
        for (uint32_t e = 0; e < ne; ++e)
            for (int a = 0; a < 4; ++a) data[fill[conn[(size_t)a * ne + e]]++] = e;
    }
    std::vector<double> raw(3 * (size_t)nn, 0.0), rec(3 * (size_t)nn, 0.0);
#pragma omp parallel for num_threads(g_threads) schedule(static)
    for (int64_t vi = 0; vi < (int64_t)nn; ++vi) {
        const uint32_t v = (uint32_t)vi;
        double s[3] = {0, 0, 0}, sv = 0.0;
        for (uint32_t p = ptr[v]; p < ptr[v + 1]; ++p) {
            const uint32_t e = data[p];
            for (int c = 0; c < 3; ++c) s[c] += g[3 * (size_t)e + c] * vol[e];
            sv += vol[e];
        }
        for (int c = 0; c < 3; ++c) raw[3 * (size_t)v + c] = sv > 1e-30 ? s[c] / sv : s[c];
        double d[32][3], r[32][3];
        int m = 0;
        for (uint32_t p = ptr[v]; p < ptr[v + 1] && m < 32; ++p, ++m) {
            const uint32_t e = data[p];
            for (int c = 0; c < 3; ++c) {
                d[m][c] = cen[3 * (size_t)e + c] - X[(size_t)c * nn + v];
                r[m][c] = g[3 * (size_t)e + c];
            }
        }
        spr_fit(m, d, r, &rec[3 * (size_t)v]);
    }
    double gsq = 0.0;
    for (uint32_t e = 0; e < ne; ++e) {
        double sq = 0.0;
        for (int a = 0; a < 4; ++a) {
            const uint32_t v = conn[(size_t)a * ne + e];
            for (int c = 0; c < 3; ++c) {
                const double df = rec[3 * (size_t)v + c] - raw[3 * (size_t)v + c];
                sq += df * df;
            }
        }
        if (elem_err) elem_err[e] = std::sqrt(sq / 4.0) * std::cbrt(vol[e]);
        gsq += vol[e] * sq / 4.0;
    }
    if (raw_node) std::memcpy(raw_node, raw.data(), raw.size() * sizeof(double));
    if (recovered) std::memcpy(recovered, rec.data(), rec.size() * sizeof(double));
    return std::sqrt(gsq);
}

}  // extern "C"
