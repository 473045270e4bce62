// fm_internal.hpp -- context state and kernel launchers of the MI355X Newton path.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <string>
#include <vector>

#include "flexmesh/fm_api.h"
#include "fm_plan.hpp"

namespace fm {

// ---------------------------------------------------------------------------
// device buffer (owned by the context)
// ---------------------------------------------------------------------------
template <typename T>
struct DevBuf {
    T* p = nullptr;
    size_t n = 0;
    DevBuf() = default;
    DevBuf(const DevBuf&) = delete;
    DevBuf& operator=(const DevBuf&) = delete;
    ~DevBuf() { release(); }
    hipError_t alloc(size_t count) {
        if (count == n && p) return hipSuccess;
        release();
        if (count == 0) return hipSuccess;
        hipError_t e = hipMalloc(reinterpret_cast<void**>(&p), count * sizeof(T));
        if (e == hipSuccess) n = count;
        return e;
    }
    void release() {
        if (p) (void)hipFree(p);
        p = nullptr;
        n = 0;
    }
    size_t bytes() const { return n * sizeof(T); }
};

// Scalars of one PCG solve, kept on the device so that no host round trip is
// needed between iterations.  Kernels read `done` and return at once when set.
constexpr int kReduceBlocks = 64;  // most blocks of one reduce (each sums a contiguous chunk of partials)
constexpr int kPcgHistCap = 4096;  // iterations whose alpha and beta a PCG solve records
constexpr int kLanczosCap = 48;    // Lanczos vectors a harvesting PCG solve keeps

struct PcgState {
    double rz, pAp, alpha, beta, bb, rz_new, relres, tol;
    double slot[4];
    double gal[48];  // Gram matrix, right-hand side and a norm of a Galerkin start (reduce op 5)
    double rz_coarse;  // r.(Z E^{-1} Z^T r) of the additive coarse correction, added to r.z (0 without one)
    double rz_cross;   // r.(Z c_old) with the previous coefficients: the coarse part of the flexible beta's z_old term
    int done;     // 1 once converged, broken down or out of iterations
    int status;   // 0 running, 1 converged, 2 breakdown (|pAp| < 1e-60), 3 max_it
    int iters;
    int max_it;
    uint32_t ticket[4];  // work-queue counters of the triangular-solve kernels
    uint32_t red_count;  // blocks of a multi-block reduce that have stored their sum (0 between reduces)
    double red_stage[kReduceBlocks];
    double* hist;        // alpha and beta of every iteration (hist_cap iterations at most), or null
    uint32_t hist_cap;
};

// Level schedule of the natural-order ILU(0) dependence graph on groups of B
// rows ("nodes": B = 3 for the 3-DOF mesh pattern, B = 1 otherwise).
struct Schedule {
    int B = 1;
    uint32_t n_nodes = 0;
    int slots = 1;           // ceil(max row length / 64)
    int lpn = 16;            // lanes per node in the triangular solves
    uint32_t levels_fwd = 0, levels_bwd = 0;
    // level-ordered node queues; each level is padded with 0xFFFFFFFF entries
    // to a multiple of the nodes one wave takes, so that a wave never holds
    // two nodes of which one depends on the other
    uint32_t q_fwd = 0, q_bwd = 0;
    DevBuf<uint32_t> fwd_order, bwd_order;
    DevBuf<uint16_t> nlow;   // number of lower column blocks per node
};

// Chain-bundle schedule of the triangular sweeps of a 3-row node-blocked
// matrix.  A chain is a maximal run of consecutive nodes in which each node
// depends on its predecessor (the x-lines of a box mesh in natural order); a
// bundle is 64/lpn chains that one wave advances together, one node per chain
// per step, in a skewed wavefront: dependencies inside the bundle pass through
// an LDS ring, dependencies on other bundles through the ready-by-value
// protocol.  Bundles are processed in a topological order of the bundle graph.
struct ChainSweep {
    uint32_t nb = 0;                 // bundles, in processing order
    uint32_t nq = 0;                 // wave-steps over all bundles
    DevBuf<uint32_t> bstart, bsteps; // per bundle: first wave-step, number of steps
    DevBuf<uint32_t> far;            // nq * kRecWords: step records (fm_plan.hpp)
                                     // (kDepNone: idle), dependency words 0 and 1, zero
    DevBuf<float> pk32;              // packed factor values of this sweep
    DevBuf<double> pk64;
};

struct ChainSchedule {
    bool ok = false;
    int lpn = 4;                     // lanes per node (the sweep kernels use one quad per node)
    uint32_t chains = 0, stride = 0, gj = 0, gk = 0;
    ChainSweep fwd, bwd;
};

struct Matrix {
    uint32_t n = 0;
    uint64_t nnz = 0;
    std::vector<uint32_t> h_row_ptr, h_col_idx;
    DevBuf<uint32_t> row_ptr, col_idx;
    DevBuf<double> val64;
    DevBuf<float> val32;
    DevBuf<double> lu64;
    DevBuf<float> lu32;
    bool lu_valid[2] = {false, false};
    bool val_valid[2] = {false, false};
    bool pk_valid[2] = {false, false};  // sweep payload packed from the current factor
    bool lu32_rounded = false;          // lu32 is the FP64 factor rounded (FM_MIXED_PC), not an FP32 factorization
    Schedule sched;
    ChainSchedule chain;
    // column nodes of each node row block (3-row node-blocked matrices): the
    // node-blocked SpMV reads one index per 3x3 block
    bool node_spmv = false;
    DevBuf<uint32_t> nptr, ncol;
    std::vector<uint32_t> perm;  // sorted position k -> caller's storage position (external matrices)
    bool ready = false;
};

// matrix value arrays hold this many elements past nnz: the LDS-staged SpMV
// reads whole 16-byte chunks, the last of which may end past the last value
constexpr size_t kValPad = 4;

enum { PREC64 = 0, PREC32 = 1, PRECPC = 2 };  // PRECPC: FP64 Krylov, FP32 ILU(0) preconditioner

struct Ctx {
    int device = 0;
    hipStream_t stream = nullptr;
    std::string err;

    // mesh (host copies are kept for the host-side set-up passes)
    uint32_t nn = 0, ne = 0;
    std::vector<double> hX;
    std::vector<uint32_t> hconn;
    std::vector<double> hmu, hlam;
    DevBuf<double> X;  // 3*nn SoA

    // colouring and colour-ordered element arrays
    bool have_colors = false;
    uint32_t n_colors = 0;
    std::vector<uint32_t> hcolors, color_ptr, hperm;
    DevBuf<uint32_t> conn_c;  // 4*ne SoA, colour order
    DevBuf<uint32_t> eperm;   // colour-order index -> element id
    DevBuf<uint32_t> eiperm;  // element id -> colour-order index
    DevBuf<double> mu_c, lam_c;
    DevBuf<uint8_t> erank_c;  // 16*ne SoA: rank of node b in nbr(node a)
    // node -> incident elements in colour order, (t << 2) | local index (the
    // node-gather assembly; used when every node is within its limits)
    DevBuf<uint32_t> n2e_ptr, n2e;
    bool node_asm = false;
    bool have_n2e = false;  // node -> incident element lists (colour order) on the device
    // converged states for the extrapolating predictor (lift = 2): the device
    // state u at load factor sec_lf[0] (sec_n >= 1) and the older ones,
    // sec_old[i] at sec_lf[i + 1] (sec_n >= i + 2), newest first, as many as
    // the extrapolation's degree uses; cleared by anything that changes the
    // problem or u
    // Galerkin-start basis per PCG precision ([0] FP64, [1] FP32): V is
    // n x cs_cols (leading dimension n) with a ring of the latest twelve
    // normalised Newton corrections of this load path in columns 0..11
    // (gal_n valid, gal_next the slot written next; cleared with the history
    // below) and the polynomial fields of degree <= cs_deg from column 12 on
    // (cs_nmono monomials per component; rebuilt after a change of the mesh,
    // the coordinates or the constrained DOFs); W holds K V and the
    // right-hand side
    DevBuf<double> cs_V64, cs_W64;
    DevBuf<float> cs_V32, cs_W32;
    int cs_cols[2] = {0, 0}, cs_nmono[2] = {0, 0};
    bool cs_poly_ok[2] = {false, false};
    DevBuf<double> cs_partials, cs_G, cs_c;
    DevBuf<int> cs_col;
    // the scaled node coordinates of the polynomial fields (built with them)
    // and the additive coarse correction of the running PCG: E^{-1} over the
    // 3 ad_nmono polynomial fields (ad_nmono = 0: none), its moment partials
    // and coefficients
    DevBuf<float4> cs_xn;
    DevBuf<double> ad_Einv, ad_part, ad_coef;
    int ad_nmono = 0;
    // E^{-1} in ad_Einv belongs to the last Galerkin start's tangent, over
    // lag_nmono monomials per component in precision lag_prec (0: none); a
    // lagged preconditioner reuses it within the load step
    int lag_nmono = 0, lag_prec = -1;
    int gal_n[2] = {0, 0}, gal_next[2] = {0, 0};
    // Ritz vectors of the last harvested solve in the basis columns after the
    // corrections (nritz valid per precision), and the Lanczos vectors
    // z_j / sqrt(r_j.z_j) of the running solve (n x lz_cap, PCG precision)
    int nritz[2] = {0, 0};
    DevBuf<float> lz32;
    DevBuf<double> lz64;
    int harvest = 0;  // Ritz vectors the next pcg_run's Lanczos vectors are kept for (0: none)
    DevBuf<double> ritz_S;  // the Ritz combination coefficients (kLanczosCap x kRitz)
    static constexpr int kExtrapMax = 4;  // highest degree of the extrapolation
    DevBuf<double> sec_old[kExtrapMax], sec_start;
    double sec_lf[kExtrapMax + 1] = {0.0, 0.0, 0.0, 0.0, 0.0};
    // du/dlf at load factor slope_lf (slope_ok): the predictor solve's
    // increment from the unloaded state over its load factor, Hermite data
    // of the extrapolation while fewer states are known than its degree needs
    DevBuf<double> slope;
    double slope_lf = 0.0;
    bool slope_ok = false;
    // alpha, beta of the last pcg_run (fm_pcg_history)
    DevBuf<double> pcg_hist;
    uint32_t pcg_hist_len = 0;
    // load steps whose first Newton iteration used the current tangent factor
    // (the factor_steps option); fact_gen counts the tangent's factorizations
    int fact_age = 0;
    uint64_t fact_gen = 0;
    // the polynomial fields' block Z^T K Z of the last fused Galerkin start
    // (host, row-major, 3 gpp_nmono square, precision gpp_prec) and the factor
    // generation whose tangent it was formed with: a step that keeps that
    // factor keeps this block too
    std::vector<double> gpp;
    int gpp_nmono = 0, gpp_prec = -1;
    uint64_t gpp_gen = 0;
    bool gpp_ok = false;
    // ad_Einv is the inverse of that block (generation einv_gen)
    uint64_t einv_gen = 0;
    bool einv_ok = false;
    int sec_n = 0;
    bool have_material = false;

    // pattern (node adjacency + tangent matrix)
    bool have_pattern = false;
    std::vector<uint32_t> nbr_ptr, nbr;
    Matrix tangent;
    Matrix ext;
    uint64_t ext_value_uploads = 0;  // fm_set_matrix calls that kept the previous pattern and plan
    Matrix* cur = nullptr;
    bool tangent_elim = false;

    // boundary conditions and load
    std::vector<uint32_t> dir_dofs;
    std::vector<double> dir_vals;
    DevBuf<uint8_t> isdir;
    DevBuf<uint32_t> d_dir_dofs;
    DevBuf<double> d_dir_vals;
    DevBuf<double> f_ext;
    bool have_fext = false;

    // FP64 state and work vectors
    DevBuf<double> u, f, R, du, ut, dD, lift, hv;
    bool u_valid = false;
    // per-iteration history of the last Newton solve (host scalars)
    std::vector<double> hist_res, hist_alpha, hist_energy;
    // PCG work vectors
    DevBuf<double> b64, x64, r64, p64, Ap64, z64, y64;
    DevBuf<float> b32, x32, r32, p32, Ap32, z32, y32;

    // scalars and reductions
    DevBuf<double> partials;
    DevBuf<PcgState> st;
    PcgState* h_st = nullptr;  // pinned, 3 entries: [0] general use, [1..2] the PCG batch checks
    hipEvent_t batch_ev[2] = {nullptr, nullptr};
    DevBuf<int> d_err;

    // profiling
    bool prof = false;
    int prof_every = 1;               // time one launch in prof_every of each class
    uint64_t prof_seen[FM_K_COUNT] = {0};  // launches of each class while profiling
    double ms[FM_K_COUNT] = {0};
    uint64_t launches[FM_K_COUNT] = {0};
    std::vector<hipEvent_t> ev_pool;
    std::vector<std::pair<int, std::pair<hipEvent_t, hipEvent_t>>> ev_pending;
};

// ---------------------------------------------------------------------------
// host passes (fm_setup.cpp)
// ---------------------------------------------------------------------------
void node_adjacency(uint32_t nn, uint32_t ne, const uint32_t* conn, std::vector<uint32_t>& ptr,
                    std::vector<uint32_t>& nbr);
uint32_t greedy_colors(uint32_t nn, uint32_t ne, const uint32_t* conn, uint32_t* colors);
// level schedule over a sorted CSR grouped in B rows; returns false if the
// matrix lacks a stored diagonal or exceeds the supported row length
bool build_schedule(const std::vector<uint32_t>& rp, const std::vector<uint32_t>& ci, int B, Schedule& S,
                    hipStream_t stream, std::string& err);
int detect_block(const std::vector<uint32_t>& rp, const std::vector<uint32_t>& ci);
// chain-bundle schedule (plan_chain_schedule, then upload); returns false
// with the reason in `why` when the matrix does not fit it (the
// level-scheduled sweeps are used then), and also on an upload failure (err set)
bool build_chain_schedule(const std::vector<uint32_t>& rp, const std::vector<uint32_t>& ci, int lpn,
                          ChainSchedule& C, hipStream_t stream, std::string& why, std::string& err);

// ---------------------------------------------------------------------------
// kernel launchers (fm_kernels.hip); all enqueue on `s`
// ---------------------------------------------------------------------------
struct ElemArgs {
    const double* X;
    uint32_t nn;
    const uint32_t* conn;  // colour order, stride ne
    uint32_t ne;
    const double* mu;
    const double* lam;
    const uint8_t* erank;
    const double* u;       // 3nn or nullptr
};

// scratch of fm_residual_noalloc (a static device array, fm_kernels.hip):
// kNoAllocPartials block partials (one per 256 rows, so at most
// 256 * kNoAllocPartials rows) and one PcgState
constexpr uint32_t kNoAllocPartials = 32768;
struct NoAllocScratch {
    double* partials;
    PcgState* state;
};
// looks the array's address up for `device`, which must be current; once per device
hipError_t noalloc_scratch_init(int device);
// the address looked up for `device` (null members before noalloc_scratch_init)
NoAllocScratch noalloc_scratch(int device);

// error of the last failed persistent-kernel launch (cleared when read)
hipError_t take_launch_error();
void record_launch_error(hipError_t e);
// grid of a persistent kernel with `block` threads per block: min(want, half
// of what the device keeps resident at once), so every block runs from the
// start
uint32_t persistent_grid(const void* kern, uint32_t want, int block = 64);
void launch_fill_u8(hipStream_t s, uint8_t* p, size_t n, uint8_t v);
void launch_fill_f64(hipStream_t s, double* p, size_t n, double v);
void launch_fill_f32(hipStream_t s, float* p, size_t n, float v);
void launch_expand_pattern(hipStream_t s, uint32_t nn, const uint32_t* nbr_ptr, const uint32_t* nbr,
                           const uint32_t* row_ptr, uint32_t* col_idx);
void launch_assemble_color(hipStream_t s, int prec, bool with_K, const ElemArgs& a, uint32_t begin, uint32_t end,
                           double* f, void* K, const uint32_t* row_ptr, const uint8_t* isdir, const double* dD,
                           double* lift);
// node-gather tangent assembly limits: incident elements and column blocks per node
constexpr uint32_t kAsmMaxElems = 32, kAsmMaxBlk = 32;
void launch_assemble_node(hipStream_t s, int prec, const ElemArgs& a, const uint32_t* n2e_ptr, const uint32_t* n2e,
                          double* f, void* K, const uint32_t* row_ptr, const uint32_t* col_idx, const uint8_t* isdir,
                          const double* dD, double* lift);
void launch_set_dir_diag(hipStream_t s, int prec, uint32_t n_dir, const uint32_t* dofs, const uint32_t* row_ptr,
                         const uint16_t* nlow, void* K);
void launch_eliminate(hipStream_t s, int prec, uint32_t n, const uint32_t* row_ptr, const uint32_t* col_idx,
                      const uint8_t* isdir, void* K);
// R = load f_ext - f (constrained rows zeroed); block partials of R.R into partials; returns #partials
uint32_t launch_residual_finish(hipStream_t s, uint32_t n, double load, const double* f_ext, const double* f,
                                const uint8_t* isdir, double* R, double* partials);
// partials of a.b; returns #partials
uint32_t launch_dot64(hipStream_t s, uint32_t n, const double* a, const double* b, double* partials);
// reduce partials into st: op 0 = store in slot[k]; 1 = pAp/alpha; 2 = rz_new/convergence; 3 = init rz;
// 4 = as 2 with the flexible beta (the r.z_old term stored in slot[k] by op 0); 5 = store in gal[k]
void launch_reduce(hipStream_t s, const double* partials, uint32_t np, PcgState* st, int op, int k);
void launch_spmv(hipStream_t s, int prec, const Matrix& A, const void* x, void* y, bool dot, double* partials,
                 void* sent0, void* sent1, const PcgState* st);
// Y[:, j] = A X[:, j], j < nv (leading dimension ld), with the matrix read
// once for all vectors; false (nothing launched) where the LDS-staged
// node-blocked SpMV does not apply
bool launch_spmv_multi(hipStream_t s, int prec, const Matrix& A, const void* X, void* Y, uint32_t ld, int nv);
uint32_t spmv_blocks(const Matrix& A, int prec);  // partials written by launch_spmv with dot
void launch_ilu_factor(hipStream_t s, int prec, const Matrix& A, const void* val, void* lu, PcgState* st,
                       int* err);
void launch_ilu_fwd(hipStream_t s, int prec, const Matrix& A, const void* lu, const void* rhs, void* r, void* x,
                    const void* p, const void* Ap, void* z, PcgState* st, bool upd, int* err);
uint32_t launch_ilu_bwd(hipStream_t s, int prec, const Matrix& A, const void* lu, const void* z, const void* r,
                        void* y, double* partials, PcgState* st, bool check_done, int* err);
// chain-bundle sweeps (fm_sweep.hip): repack of the factor after each
// factorization, then forward (z = L^{-1} rhs, or with the fused PCG update
// when upd) and backward (y = U^{-1} z with the fused r.y partials; returns
// the number of partials)
int chain_pack(hipStream_t s, int prec, Matrix& A);
void launch_chain_fwd(hipStream_t s, int prec, const Matrix& A, const void* rhs, void* z, PcgState* st,
                      bool check_done, int* err);
uint32_t launch_chain_bwd(hipStream_t s, int prec, const Matrix& A, const void* z, const void* r, void* y,
                          double* partials, PcgState* st, bool check_done, int* err);
void launch_fill_sentinel(hipStream_t s, int prec, void* a, void* b, size_t n);
void launch_pcg_p_update(hipStream_t s, int prec, uint32_t n, const void* y, void* p, const PcgState* st);
// dst = (double)src and block partials of r.dst (and of r.dst_before when
// old_partials != null); returns #partials
uint32_t launch_f32_to_f64_dot(hipStream_t s, uint32_t n, const float* src, const double* r, double* dst,
                               double* partials, double* old_partials, const PcgState* st);
void launch_pcg_xr_update(hipStream_t s, int prec, uint32_t n, const void* p, const void* Ap, void* x, void* r,
                          const PcgState* st);
// r = b - Ap; x untouched; partials of r.r
uint32_t launch_pcg_init_r(hipStream_t s, int prec, uint32_t n, const void* b, const void* Ap, void* r,
                           double* partials);
void launch_copy(hipStream_t s, int prec, uint32_t n, const void* src, void* dst);
// dst32 = (float)(src64 * scale)
void launch_f64_to_f32(hipStream_t s, uint32_t n, const double* src, double scale, float* dst);
// dst64 = (double)src32 * scale
void launch_f32_to_f64(hipStream_t s, uint32_t n, const float* src, double scale, double* dst);
// y = a + alpha * b (FP64)
// y = alpha a
void launch_scale(hipStream_t s, uint32_t n, const double* a, double alpha, double* y);
void launch_axpy_to(hipStream_t s, uint32_t n, const double* a, double alpha, const double* b, double* y);
// u[dofs[k]] = scale * vals[k] (values), and dD[dofs[k]] = scale*vals[k] - u_old[dofs[k]]
void launch_dir_increment(hipStream_t s, uint32_t n_dir, const uint32_t* dofs, const double* vals, double scale,
                          const double* u, double* dD);
// rhs = R - lift (free rows), 0 on constrained rows
void launch_sub_masked(hipStream_t s, uint32_t n, const double* R, const double* lift, const uint8_t* isdir,
                       double* out);
// u += du on free rows, u += dD on constrained rows
void launch_apply_predictor(hipStream_t s, uint32_t n, const double* du, const double* dD, const uint8_t* isdir,
                            double* u);
// block partials of sum_e V_e psi(F_e); returns #partials
uint32_t launch_energy_elem(hipStream_t s, const ElemArgs& a, double* partials);
void launch_adjoint_elem(hipStream_t s, const ElemArgs& a, const uint32_t* eperm, const double* lam_vec,
                         double* dmu, double* dlam);
void launch_rot_batch(hipStream_t s, uint32_t n, const double* F, double alpha, double* R, double* w, double* Q);
void launch_elem_rot(hipStream_t s, const ElemArgs& a, const uint32_t* eiperm, double* w);
// per-element fields, kFieldCount x ne SoA in element order: det F, Cauchy
// stress (6), von Mises, pressure, largest principal stress, strain energy,
// maximum stretch, F (9)
constexpr int kFieldCount = 21;
void launch_element_fields(hipStream_t s, const ElemArgs& a, const uint32_t* eiperm, double* out);
// out = sum over i < k of w[i] x[i] (k <= 5)
struct Lincomb {
    const double* x[5];
    double w[5];
    int k;
};
void launch_lincomb(hipStream_t s, uint32_t n, const Lincomb& c, double* out);
// Galerkin start of a linear solve (precision prec: PREC32 or PREC64 vectors)
// block partials of a.b with FP64 accumulation; returns #partials
uint32_t launch_dot(hipStream_t s, int prec, uint32_t n, const void* a, const void* b, double* partials);
// out = sum_{i < m} c[i] v[i] (m <= 4)
void launch_combine(hipStream_t s, int prec, uint32_t n, int m, const double* c, const void* const* v, void* out);
// out = in / sqrt(*nrm2) (*nrm2 read on the device)
void launch_normalize(hipStream_t s, int prec, uint32_t n, const double* nrm2, const void* in, void* out);
// xn[v]: node coordinates mapped by (X - center) * scale (FP32) and, in .w,
// the bit mask of the node's constrained components
void launch_coarse_coords(hipStream_t s, uint32_t nn, const double* X, const uint8_t* isdir, const double* center,
                          const double* scale, float4* xn);
// columns col0 + c ldm + k (k < nmono, c < 3) of out (leading dimension
// ld = 3 nn): monomial k (1, x, y, z, x^2, y^2, z^2, xy, yz, xz, cubics,
// quartics) of the scaled coordinates xn in displacement component c, zero
// on constrained DOFs
void launch_poly_basis(hipStream_t s, int prec, uint32_t nn, const float4* xn, int nmono, int ldm, void* out,
                       uint32_t ld, int col0);
// additive coarse correction: moments g = Z^T r (partials:
// coarse_moment_blocks() * 3 nmono doubles), coef = Einv g (Einv 3 nmono
// square, row-major) and st->rz_coarse = g.coef; a no-op once st->done
uint32_t coarse_moment_blocks();
void launch_coarse_correct(hipStream_t s, int prec, uint32_t n, int nmono, const float4* xn, const void* r,
                           const double* Einv, double* partials, double* coef, PcgState* st);
// x += alpha p, r -= alpha Ap with the moment partials of the new r in the
// same pass; launch_coarse_solve then turns them into the coefficients (coef:
// 3 nmono FP64 values followed by the same in FP32, room for 6 nmono doubles)
void launch_pcg_xr_update_moments(hipStream_t s, int prec, uint32_t n, int nmono, const void* p, const void* Ap,
                                  void* x, void* r, const float4* xn, double* partials, const PcgState* st);
void launch_coarse_solve(hipStream_t s, int nmono, const double* partials, const double* Einv, double* coef,
                         PcgState* st, bool cross = false);  // cross: also st->rz_cross (the flexible beta of pcg_run_pc)
// block partials (coarse_moment_blocks() x 3 nmono) of the moments Z^T r alone
void launch_coarse_moments(hipStream_t s, int prec, uint32_t n, int nmono, const float4* xn, const void* r,
                           double* partials, const PcgState* st);
// p = y + Z coef + beta p (first: p = y + Z coef); lz (optional, lz_cap
// columns of n): the iteration's Lanczos vector z / sqrt(r.z), z = y + Z coef
void launch_pcg_p_update_coarse(hipStream_t s, int prec, uint32_t n, int nmono, const void* y, void* p,
                                const float4* xn, const double* coef, bool first, const PcgState* st,
                                void* lz = nullptr, uint32_t lz_cap = 0);
// Y[:, q] = sum_j S[j k + q] Z[:, j], q < k <= 8, j < m (S row-major, FP64)
void launch_basis_combine_multi(hipStream_t s, int prec, uint32_t n, int m, int k, const double* S, const void* Z,
                                void* Y);
// G (m x mw, row-major, FP64) = V^T W for m and mw (both <= 128) columns of
// leading dimension n, in a fixed order; partials: gram_blocks(m, mw) * m * mw
// doubles
uint32_t gram_blocks(int m, int mw);
// G (3 nmono x (3 nmono + kPolyGramStored), row-major, FP64): the Gram rows
// of the polynomial fields Z over the node-blocked matrix A without forming
// A Z: Z^T A Z, then Z^T S_j for the kPolyGramStored stored columns S
// (leading dimension A.n; the basis products and the right-hand side);
// partials: poly_gram_blocks() * 3 nmono * (3 nmono + kPolyGramStored)
uint32_t poly_gram_blocks();
// out[p] = sum over b < nb of partials[b np + p], in a fixed order
void launch_partial_sums(hipStream_t s, uint32_t nb, uint32_t np, const double* partials, double* out);
constexpr int kPolyGramStored = 21;
void launch_poly_gram(hipStream_t s, int prec, const Matrix& A, const float4* xn, int nmono, const void* S,
                      double* partials, double* G);
void launch_gram(hipStream_t s, int prec, uint32_t n, int m, int mw, const void* V, const void* W, double* partials,
                 double* G);
// x = sum_q c[q] V[:, col[q]] (device arrays col, c of k entries)
void launch_basis_combine(hipStream_t s, int prec, uint32_t n, int k, const int* col, const double* c, const void* V,
                          void* x);
// ZZ estimate: zz is kZzRows x ne scratch (colour order), raw and rec 3 nn,
// err ne (element order); returns the number of partials of sum V sq / 4
constexpr int kZzRows = 7;
uint32_t launch_zz(hipStream_t s, const ElemArgs& a, const uint32_t* eiperm, const uint32_t* n2e_ptr,
                   const uint32_t* n2e, double* zz, double* raw, double* rec, double* err, double* partials);

}  // namespace fm
