// fm_api.cpp -- C ABI and host drivers (PCG, Newton, adjoint) of the MI355X
// quasi-static Newton path.  Host code orchestrates device-resident data; the
// only host round trips inside a solve are the scalar checks (convergence,
// line search), read through pinned memory.
#include <omp.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <numeric>
#include <utility>

#include "fm_internal.hpp"

using namespace fm;

struct fm_ctx : public fm::Ctx {};

namespace {

constexpr int kBatch = 8;  // PCG iterations enqueued between host checks

#define FM_HIP(call)                                                            \
    do {                                                                        \
        hipError_t e_ = (call);                                                 \
        if (e_ != hipSuccess) {                                                 \
            ctx->err = std::string(#call) + ": " + hipGetErrorString(e_);       \
            return FM_EHIP;                                                     \
        }                                                                       \
    } while (0)

#define FM_TRY(expr)          \
    do {                      \
        int rc_ = (expr);     \
        if (rc_ != FM_OK) return rc_; \
    } while (0)

int fail(fm_ctx* ctx, int code, const std::string& msg) {
    if (ctx) ctx->err = msg;
    return code;
}

double now_s() {
    using namespace std::chrono;
    return duration<double>(steady_clock::now().time_since_epoch()).count();
}

// ---- profiling: HIP events around launches of one kernel class --------------
void prof_flush(fm_ctx* ctx) {
    if (ctx->ev_pending.empty()) return;
    (void)hipEventSynchronize(ctx->ev_pending.back().second.second);
    for (auto& pe : ctx->ev_pending) {
        float ms = 0.f;
        if (hipEventElapsedTime(&ms, pe.second.first, pe.second.second) == hipSuccess) ctx->ms[pe.first] += ms;
        ctx->launches[pe.first] += 1;
        ctx->ev_pool.push_back(pe.second.first);
        ctx->ev_pool.push_back(pe.second.second);
    }
    ctx->ev_pending.clear();
}

struct ProfScope {
    fm_ctx* ctx;
    int k;
    hipEvent_t a = nullptr, b = nullptr;
    ProfScope(fm_ctx* c, int kind) : ctx(c), k(kind) {
        if (!ctx->prof || kind < 0) return;  // kind < 0: never timed (creates no events)
        // sampled profiling: the events put a bubble of a few microseconds
        // into the stream, so only every prof_every-th launch of the classes
        // launched once per PCG iteration is timed
        const bool per_iteration = kind == FM_K_SPMV || kind == FM_K_ILU_FWD || kind == FM_K_ILU_BWD ||
                                   kind == FM_K_PCG_OTHER || kind == FM_K_COARSE;
        if (ctx->prof_seen[kind]++ % (uint64_t)(per_iteration ? ctx->prof_every : 1) != 0) return;
        if (ctx->ev_pool.size() < 2) {
            if (ctx->ev_pending.size() >= 1024) prof_flush(ctx);
            if (ctx->ev_pool.size() < 2) {
                hipEvent_t e1, e2;
                if (hipEventCreate(&e1) != hipSuccess || hipEventCreate(&e2) != hipSuccess) return;
                ctx->ev_pool.push_back(e1);
                ctx->ev_pool.push_back(e2);
            }
        }
        a = ctx->ev_pool.back();
        ctx->ev_pool.pop_back();
        b = ctx->ev_pool.back();
        ctx->ev_pool.pop_back();
        (void)hipEventRecord(a, ctx->stream);
    }
    ~ProfScope() {
        if (!a) return;
        (void)hipEventRecord(b, ctx->stream);
        ctx->ev_pending.push_back({k, {a, b}});
    }
};

int check_launch(fm_ctx* ctx, const char* what) {
    hipError_t e = take_launch_error();
    if (e == hipSuccess) e = hipGetLastError();
    if (e != hipSuccess) return fail(ctx, FM_EHIP, std::string(what) + ": " + hipGetErrorString(e));
    return FM_OK;
}

int sync(fm_ctx* ctx) {
    FM_HIP(hipStreamSynchronize(ctx->stream));
    if (ctx->prof) prof_flush(ctx);
    return FM_OK;
}

int check_device_error(fm_ctx* ctx) {
    int h = 0;
    FM_HIP(hipMemcpyAsync(&h, ctx->d_err.p, sizeof(int), hipMemcpyDeviceToHost, ctx->stream));
    FM_HIP(hipStreamSynchronize(ctx->stream));
    if (h) {
        FM_HIP(hipMemsetAsync(ctx->d_err.p, 0, sizeof(int), ctx->stream));
        return fail(ctx, FM_EDEVICE, "a device-side wait exceeded its spin limit (dependence schedule broken)");
    }
    return FM_OK;
}

// ---- set-up ---------------------------------------------------------------------
// chain-bundle sweeps for 3-row node-blocked matrices that fit them
// (FM_SWEEP=level keeps the level-scheduled sweeps)
int chain_setup(fm_ctx* ctx, Matrix& M, const std::vector<uint32_t>& rp, const std::vector<uint32_t>& ci) {
    M.chain.ok = false;
    static const bool level_only = [] {
        const char* e = std::getenv("FM_SWEEP");
        return e && std::string(e) == "level";
    }();
    if (level_only || M.sched.B != 3) return FM_OK;
    std::string why, err;
    if (!build_chain_schedule(rp, ci, 4, M.chain, ctx->stream, why, err) && !err.empty())
        return fail(ctx, FM_EDEVICE, err);
    return FM_OK;
}

// column-node lists of a 3-row node-blocked matrix for the node-blocked SpMV
// (rows 3v hold the node's columns; rp/ci may be the compact view)
int node_cols_setup(fm_ctx* ctx, Matrix& M, const std::vector<uint32_t>& rp, const std::vector<uint32_t>& ci) {
    M.node_spmv = false;
    if (M.sched.B != 3) return FM_OK;
    const uint32_t nn = (uint32_t)((rp.size() - 1) / 3);
    std::vector<uint32_t> nptr(nn + 1, 0), ncol;
    ncol.reserve(rp[3 * nn] / 3);
    for (uint32_t v = 0; v < nn; ++v) {
        for (uint32_t k = rp[3 * v]; k < rp[3 * v + 1]; k += 3) ncol.push_back(ci[k] / 3);
        nptr[v + 1] = (uint32_t)ncol.size();
    }
    FM_HIP(M.nptr.alloc(nn + 1));
    FM_HIP(M.ncol.alloc(ncol.size()));
    FM_HIP(hipMemcpy(M.nptr.p, nptr.data(), (nn + 1) * 4ull, hipMemcpyHostToDevice));
    FM_HIP(hipMemcpy(M.ncol.p, ncol.data(), ncol.size() * 4, hipMemcpyHostToDevice));
    M.node_spmv = true;
    return FM_OK;
}

int ensure_pattern(fm_ctx* ctx);

int ensure_colors(fm_ctx* ctx) {
    if (ctx->have_colors) return FM_OK;
    if (!ctx->nn) return fail(ctx, FM_ESTATE, "no mesh");
    FM_TRY(ensure_pattern(ctx));
    const uint32_t ne = ctx->ne;
    ctx->hcolors.assign(ne, 0);
    const uint32_t nc = greedy_colors(ctx->nn, ne, ctx->hconn.data(), ctx->hcolors.data());
    if (nc == 0 && ne > 0) return fail(ctx, FM_EINVAL, "mesh needs more than 256 colours");
    ctx->n_colors = nc;
    ctx->color_ptr.assign(nc + 1, 0);
    for (uint32_t e = 0; e < ne; ++e) ctx->color_ptr[ctx->hcolors[e] + 1]++;
    for (uint32_t c = 0; c < nc; ++c) ctx->color_ptr[c + 1] += ctx->color_ptr[c];
    ctx->hperm.resize(ne);
    {
        std::vector<uint32_t> fill(ctx->color_ptr.begin(), ctx->color_ptr.end() - 1);
        for (uint32_t e = 0; e < ne; ++e) ctx->hperm[fill[ctx->hcolors[e]]++] = e;
    }
    // colour-ordered SoA element arrays and the node-pair rank table
    std::vector<uint32_t> connc((size_t)4 * ne);
    std::vector<uint8_t> rank((size_t)16 * ne);
    const auto& np = ctx->nbr_ptr;
    const auto& nb = ctx->nbr;
#pragma omp parallel for schedule(static)
    for (int64_t t = 0; t < (int64_t)ne; ++t) {
        const uint32_t e = ctx->hperm[t];
        uint32_t nd[4];
        for (int a = 0; a < 4; ++a) nd[a] = connc[(size_t)a * ne + t] = ctx->hconn[(size_t)a * ne + e];
        for (int a = 0; a < 4; ++a)
            for (int b = 0; b < 4; ++b) {
                const uint32_t* lo = nb.data() + np[nd[a]];
                const uint32_t* hi = nb.data() + np[nd[a] + 1];
                rank[(size_t)(a * 4 + b) * ne + t] = (uint8_t)(std::lower_bound(lo, hi, nd[b]) - lo);
            }
    }
    FM_HIP(ctx->conn_c.alloc((size_t)4 * ne));
    FM_HIP(ctx->eperm.alloc(ne));
    FM_HIP(ctx->erank_c.alloc((size_t)16 * ne));
    FM_HIP(hipMemcpy(ctx->conn_c.p, connc.data(), connc.size() * 4, hipMemcpyHostToDevice));
    FM_HIP(hipMemcpy(ctx->eperm.p, ctx->hperm.data(), (size_t)ne * 4, hipMemcpyHostToDevice));
    {
        std::vector<uint32_t> iperm(ne);
        for (uint32_t t = 0; t < ne; ++t) iperm[ctx->hperm[t]] = t;
        FM_HIP(ctx->eiperm.alloc(ne));
        FM_HIP(hipMemcpy(ctx->eiperm.p, iperm.data(), (size_t)ne * 4, hipMemcpyHostToDevice));
    }
    FM_HIP(hipMemcpy(ctx->erank_c.p, rank.data(), rank.size(), hipMemcpyHostToDevice));
    // node -> incident elements, in colour order (the order of the colour launches)
    {
        const uint32_t nn = ctx->nn;
        std::vector<uint32_t> ptr(nn + 1, 0), ent((size_t)4 * ne);
        for (size_t k = 0; k < (size_t)4 * ne; ++k) ptr[connc[k] + 1]++;
        for (uint32_t v = 0; v < nn; ++v) ptr[v + 1] += ptr[v];
        std::vector<uint32_t> fill(ptr.begin(), ptr.end() - 1);
        for (uint32_t t = 0; t < ne; ++t)
            for (uint32_t a = 0; a < 4; ++a) ent[fill[connc[(size_t)a * ne + t]]++] = (t << 2) | a;
        // the lists are also the patches of the ZZ estimator (entries need t < 2^30)
        ctx->have_n2e = ne < (1u << 30);
        bool ok = ctx->have_n2e;
        for (uint32_t v = 0; v < nn && ok; ++v)
            ok = ptr[v + 1] - ptr[v] <= kAsmMaxElems && np[v + 1] - np[v] <= kAsmMaxBlk;
        static const bool color_only = [] {
            const char* e = std::getenv("FM_ASM");
            return e && std::string(e) == "color";
        }();
        ctx->node_asm = ok && !color_only;
        if (ctx->have_n2e) {
            FM_HIP(ctx->n2e_ptr.alloc(nn + 1));
            FM_HIP(ctx->n2e.alloc(ent.size()));
            FM_HIP(hipMemcpy(ctx->n2e_ptr.p, ptr.data(), (size_t)(nn + 1) * 4, hipMemcpyHostToDevice));
            FM_HIP(hipMemcpy(ctx->n2e.p, ent.data(), ent.size() * 4, hipMemcpyHostToDevice));
        }
    }
    ctx->have_colors = true;
    return FM_OK;
}

int upload_material(fm_ctx* ctx) {
    const uint32_t ne = ctx->ne;
    std::vector<double> mc(ne), lc(ne);
    for (uint32_t t = 0; t < ne; ++t) {
        mc[t] = ctx->hmu[ctx->hperm[t]];
        lc[t] = ctx->hlam[ctx->hperm[t]];
    }
    FM_HIP(ctx->mu_c.alloc(ne));
    FM_HIP(ctx->lam_c.alloc(ne));
    FM_HIP(hipMemcpy(ctx->mu_c.p, mc.data(), (size_t)ne * 8, hipMemcpyHostToDevice));
    FM_HIP(hipMemcpy(ctx->lam_c.p, lc.data(), (size_t)ne * 8, hipMemcpyHostToDevice));
    ctx->have_material = true;
    return FM_OK;
}

int ensure_pattern(fm_ctx* ctx) {
    if (ctx->have_pattern) return FM_OK;
    if (!ctx->nn) return fail(ctx, FM_ESTATE, "no mesh");
    node_adjacency(ctx->nn, ctx->ne, ctx->hconn.data(), ctx->nbr_ptr, ctx->nbr);
    for (uint32_t n = 0; n < ctx->nn; ++n)
        if (ctx->nbr_ptr[n + 1] - ctx->nbr_ptr[n] > 85)
            return fail(ctx, FM_EINVAL, "a node has more than 85 neighbours (row length above 256)");
    Matrix& M = ctx->tangent;
    const uint32_t n = 3 * ctx->nn;
    M.n = n;
    M.h_row_ptr.assign(n + 1, 0);
    uint64_t acc = 0;
    for (uint32_t v = 0; v < ctx->nn; ++v) {
        const uint32_t deg = ctx->nbr_ptr[v + 1] - ctx->nbr_ptr[v];
        for (int d = 0; d < 3; ++d) {
            acc += 3ull * deg;
            if (acc > 0xFFFFFFFFull) return fail(ctx, FM_EINVAL, "nnz exceeds uint32 CSR indexing");
            M.h_row_ptr[3 * v + d + 1] = (uint32_t)acc;
        }
    }
    M.nnz = acc;
    FM_HIP(M.row_ptr.alloc(n + 1));
    FM_HIP(M.col_idx.alloc(M.nnz));
    FM_HIP(hipMemcpy(M.row_ptr.p, M.h_row_ptr.data(), (size_t)(n + 1) * 4, hipMemcpyHostToDevice));
    {
        DevBuf<uint32_t> dptr, dnbr;
        FM_HIP(dptr.alloc(ctx->nn + 1));
        FM_HIP(dnbr.alloc(ctx->nbr.size()));
        FM_HIP(hipMemcpy(dptr.p, ctx->nbr_ptr.data(), (size_t)(ctx->nn + 1) * 4, hipMemcpyHostToDevice));
        FM_HIP(hipMemcpy(dnbr.p, ctx->nbr.data(), ctx->nbr.size() * 4, hipMemcpyHostToDevice));
        launch_expand_pattern(ctx->stream, ctx->nn, dptr.p, dnbr.p, M.row_ptr.p, M.col_idx.p);
        FM_TRY(check_launch(ctx, "expand_pattern"));
        FM_HIP(hipStreamSynchronize(ctx->stream));
    }
    // level schedule: build_schedule reads rows 3v only, so a compact view
    // holding row 3v's columns (and empty rows 3v+1, 3v+2) is enough
    {
        std::vector<uint32_t> rp3(n + 1), ci3;
        ci3.reserve(3 * ctx->nbr.size());
        for (uint32_t v = 0; v < ctx->nn; ++v) {
            rp3[3 * v] = (uint32_t)ci3.size();
            for (uint32_t q = ctx->nbr_ptr[v]; q < ctx->nbr_ptr[v + 1]; ++q)
                for (int e = 0; e < 3; ++e) ci3.push_back(3 * ctx->nbr[q] + e);
            rp3[3 * v + 1] = rp3[3 * v + 2] = (uint32_t)ci3.size();
        }
        rp3[n] = (uint32_t)ci3.size();
        std::string err;
        if (!build_schedule(rp3, ci3, 3, M.sched, ctx->stream, err)) return fail(ctx, FM_EINVAL, err);
        FM_TRY(chain_setup(ctx, M, rp3, ci3));
        FM_TRY(node_cols_setup(ctx, M, rp3, ci3));
    }
    M.val_valid[0] = M.val_valid[1] = false;
    M.lu_valid[0] = M.lu_valid[1] = false;
    M.ready = true;
    ctx->have_pattern = true;
    return FM_OK;
}

ElemArgs elem_args(fm_ctx* ctx, const double* u) {
    ElemArgs a;
    a.X = ctx->X.p;
    a.nn = ctx->nn;
    a.conn = ctx->conn_c.p;
    a.ne = ctx->ne;
    // no material yet (the ZZ estimate needs none): the kernels read zeros
    a.mu = ctx->have_material ? ctx->mu_c.p : nullptr;
    a.lam = ctx->have_material ? ctx->lam_c.p : nullptr;
    a.erank = ctx->erank_c.p;
    a.u = u;
    return a;
}

int ensure_vectors(fm_ctx* ctx) {
    const size_t n = 3ull * ctx->nn;
    for (auto* b : {&ctx->u, &ctx->f, &ctx->R, &ctx->du, &ctx->ut, &ctx->dD, &ctx->lift, &ctx->hv}) FM_HIP(b->alloc(n));
    FM_HIP(ctx->isdir.alloc(n));
    return FM_OK;
}

// mesh-derived state: pattern, colours, node -> element lists, work vectors
int ensure_mesh(fm_ctx* ctx) {
    if (!ctx->nn) return fail(ctx, FM_ESTATE, "no mesh (fm_set_mesh)");
    FM_TRY(ensure_pattern(ctx));
    FM_TRY(ensure_colors(ctx));
    FM_TRY(ensure_vectors(ctx));
    return FM_OK;
}

int ensure_ready(fm_ctx* ctx) {
    FM_TRY(ensure_mesh(ctx));
    if (!ctx->have_material) return fail(ctx, FM_ESTATE, "no material (fm_set_material)");
    return FM_OK;
}

int ensure_partials(fm_ctx* ctx, size_t n) {
    if (ctx->partials.n < n) FM_HIP(ctx->partials.alloc(std::max<size_t>(n, 1024)));
    return FM_OK;
}

// f = f_int(u) (and the tangent with BC masking/lift when with_K)
int assemble(fm_ctx* ctx, const double* u, bool with_K, int prec, bool mask, const double* dD, int prof_kind) {
    Matrix& M = ctx->tangent;
    hipStream_t s = ctx->stream;
    const uint8_t* isdir = (mask && !ctx->dir_dofs.empty()) ? ctx->isdir.p : nullptr;
    if (with_K && ctx->node_asm) {
        // node-gather tangent assembly: every entry of f, lift and K is written once
        void* K;
        if (prec == PREC64) {
            FM_HIP(M.val64.alloc(M.nnz + kValPad));
            K = M.val64.p;
        } else {
            FM_HIP(M.val32.alloc(M.nnz + kValPad));
            K = M.val32.p;
        }
        {
            ProfScope ps(ctx, prof_kind);
            launch_assemble_node(s, prec, elem_args(ctx, u), ctx->n2e_ptr.p, ctx->n2e.p, ctx->f.p, K, M.row_ptr.p,
                                 M.col_idx.p, isdir, dD, ctx->lift.p);
        }
        FM_TRY(check_launch(ctx, "assemble"));
        if (isdir)
            launch_set_dir_diag(s, prec, (uint32_t)ctx->dir_dofs.size(), ctx->d_dir_dofs.p, M.row_ptr.p,
                                M.sched.nlow.p, K);
        M.val_valid[prec] = true;
        M.val_valid[1 - prec] = false;
        M.lu_valid[0] = M.lu_valid[1] = false;
        ctx->tangent_elim = isdir != nullptr;
        return FM_OK;
    }
    launch_fill_f64(s, ctx->f.p, ctx->f.n, 0.0);
    void* K = nullptr;
    if (with_K) {
        if (prec == PREC64) {
            FM_HIP(M.val64.alloc(M.nnz + kValPad));
            launch_fill_f64(s, M.val64.p, M.nnz, 0.0);
            K = M.val64.p;
        } else {
            FM_HIP(M.val32.alloc(M.nnz + kValPad));
            launch_fill_f32(s, M.val32.p, M.nnz, 0.f);
            K = M.val32.p;
        }
        if (dD) launch_fill_f64(s, ctx->lift.p, ctx->lift.n, 0.0);
    }
    ElemArgs a = elem_args(ctx, u);
    for (uint32_t c = 0; c < ctx->n_colors; ++c) {
        ProfScope ps(ctx, prof_kind);
        launch_assemble_color(s, prec, with_K, a, ctx->color_ptr[c], ctx->color_ptr[c + 1], ctx->f.p, K,
                              M.row_ptr.p, isdir, dD, ctx->lift.p);
    }
    FM_TRY(check_launch(ctx, "assemble"));
    if (with_K) {
        if (isdir)
            launch_set_dir_diag(s, prec, (uint32_t)ctx->dir_dofs.size(), ctx->d_dir_dofs.p, M.row_ptr.p,
                                M.sched.nlow.p, K);
        M.val_valid[prec] = true;
        M.val_valid[1 - prec] = false;
        M.lu_valid[0] = M.lu_valid[1] = false;
        ctx->tangent_elim = isdir != nullptr;
    }
    return FM_OK;
}

// R = load f_ext - f_int(u) (constrained rows zeroed); returns ||R|| in *norm
int residual(fm_ctx* ctx, const double* u, double load, double* norm) {
    FM_TRY(assemble(ctx, u, false, PREC64, false, nullptr, FM_K_RESIDUAL));
    const uint32_t n = 3 * ctx->nn;
    FM_TRY(ensure_partials(ctx, n / 256 + 1));
    const uint32_t np = launch_residual_finish(ctx->stream, n, load, ctx->have_fext ? ctx->f_ext.p : nullptr,
                                               ctx->f.p, ctx->dir_dofs.empty() ? nullptr : ctx->isdir.p, ctx->R.p,
                                               ctx->partials.p);
    launch_reduce(ctx->stream, ctx->partials.p, np, ctx->st.p, 0, 0);
    FM_HIP(hipMemcpyAsync(ctx->h_st, ctx->st.p, sizeof(PcgState), hipMemcpyDeviceToHost, ctx->stream));
    FM_TRY(sync(ctx));
    *norm = std::sqrt(ctx->h_st->slot[0]);
    return FM_OK;
}

// Potential energy sum_e V_e psi(F_e) - load f_ext.u (FP64, fixed-order
// block reductions) and, when du != nullptr, the slope R.du with the R of the
// last residual evaluation.
int energy(fm_ctx* ctx, const double* u, double load, const double* du, double* pi, double* slope) {
    hipStream_t s = ctx->stream;
    const uint32_t n = 3 * ctx->nn;
    FM_TRY(ensure_partials(ctx, std::max(n, ctx->ne) / 256 + 1));
    uint32_t np = launch_energy_elem(s, elem_args(ctx, u), ctx->partials.p);
    launch_reduce(s, ctx->partials.p, np, ctx->st.p, 0, 0);
    if (ctx->have_fext) {
        np = launch_dot64(s, n, ctx->f_ext.p, u, ctx->partials.p);
        launch_reduce(s, ctx->partials.p, np, ctx->st.p, 0, 1);
    }
    if (du) {
        np = launch_dot64(s, n, ctx->R.p, du, ctx->partials.p);
        launch_reduce(s, ctx->partials.p, np, ctx->st.p, 0, 2);
    }
    FM_HIP(hipMemcpyAsync(ctx->h_st, ctx->st.p, sizeof(PcgState), hipMemcpyDeviceToHost, s));
    FM_TRY(sync(ctx));
    *pi = ctx->h_st->slot[0] - (ctx->have_fext ? load * ctx->h_st->slot[1] : 0.0);
    if (slope) *slope = du ? ctx->h_st->slot[2] : 0.0;
    return FM_OK;
}

// ---- PCG ------------------------------------------------------------------------
template <typename T>
struct PcgVecs {
    T *b, *x, *r, *p, *Ap, *z, *y;
};

int pcg_vectors(fm_ctx* ctx, int prec, uint32_t n) {
    if (prec == PREC64) {
        for (auto* v : {&ctx->b64, &ctx->x64, &ctx->r64, &ctx->p64, &ctx->Ap64, &ctx->z64, &ctx->y64})
            FM_HIP(v->alloc(n));
    } else {
        for (auto* v : {&ctx->b32, &ctx->x32, &ctx->r32, &ctx->p32, &ctx->Ap32, &ctx->z32, &ctx->y32})
            FM_HIP(v->alloc(n));
    }
    return FM_OK;
}

void* vb(fm_ctx* c, int prec, int which) {
    if (prec == PREC64) {
        DevBuf<double>* v[] = {&c->b64, &c->x64, &c->r64, &c->p64, &c->Ap64, &c->z64, &c->y64};
        return v[which]->p;
    }
    DevBuf<float>* v[] = {&c->b32, &c->x32, &c->r32, &c->p32, &c->Ap32, &c->z32, &c->y32};
    return v[which]->p;
}
enum { VB = 0, VX, VR, VP, VAP, VZ, VY };

int pack_factor(fm_ctx* ctx, Matrix& A, int prec);

int factorize(fm_ctx* ctx, Matrix& A, int prec, bool pack = true) {
    const bool have = A.lu_valid[prec] && !(prec == PREC32 && A.lu32_rounded);
    if (!have) {
        if (!A.val_valid[prec]) return fail(ctx, FM_ESTATE, "matrix values not available in this precision");
        const void* val;
        void* lu;
        if (prec == PREC64) {
            FM_HIP(A.lu64.alloc(A.nnz));
            val = A.val64.p;
            lu = A.lu64.p;
        } else {
            FM_HIP(A.lu32.alloc(A.nnz));
            val = A.val32.p;
            lu = A.lu32.p;
        }
        // ticket[2] of the state is the factorization queue
        PcgState init{};
        *ctx->h_st = init;
        FM_HIP(hipMemcpyAsync(ctx->st.p, ctx->h_st, sizeof(PcgState), hipMemcpyHostToDevice, ctx->stream));
        {
            ProfScope ps(ctx, FM_K_ILU_FACTOR);
            launch_ilu_factor(ctx->stream, prec, A, val, lu, ctx->st.p, ctx->d_err.p);
        }
        FM_TRY(check_launch(ctx, "ilu_factor"));
        if (&A == &ctx->tangent) ++ctx->fact_gen;
        A.pk_valid[prec] = false;
        if (pack) FM_TRY(pack_factor(ctx, A, prec));
        FM_TRY(check_device_error(ctx));  // a pivot wait that hit its bound
        A.lu_valid[prec] = true;
        if (prec == PREC32) A.lu32_rounded = false;
    }
    if (pack) FM_TRY(pack_factor(ctx, A, prec));
    return FM_OK;
}

// sweep payload of the current factor in `prec`
int pack_factor(fm_ctx* ctx, Matrix& A, int prec) {
    if (!A.chain.ok || A.pk_valid[prec]) return FM_OK;
    {
        ProfScope ps(ctx, FM_K_ILU_FACTOR);
        if (chain_pack(ctx->stream, prec, A) != FM_OK) return fail(ctx, FM_EDEVICE, "allocation of the packed factor failed");
    }
    FM_TRY(check_launch(ctx, "chain_pack"));
    A.pk_valid[prec] = true;
    return FM_OK;
}

// Preconditioner of FM_MIXED_PC: the FP64 ILU(0) factor rounded to FP32 and
// applied in FP32.  A factorization carried out in FP32 loses digits in the
// Schur-complement updates of an ill-conditioned tangent (a slender beam),
// and the preconditioned iteration count then depends on the rounding of
// every other kernel; the rounded FP64 factor keeps it at the FP64 count.
int factorize_rounded(fm_ctx* ctx, Matrix& A) {
    if (!(A.lu_valid[PREC32] && A.lu32_rounded)) {
        FM_TRY(factorize(ctx, A, PREC64, false));
        FM_HIP(A.lu32.alloc(A.nnz));
        {
            ProfScope ps(ctx, FM_K_ILU_FACTOR);
            launch_f64_to_f32(ctx->stream, (uint32_t)A.nnz, A.lu64.p, 1.0, A.lu32.p);
        }
        FM_TRY(check_launch(ctx, "factor rounding"));
        A.lu_valid[PREC32] = true;
        A.lu32_rounded = true;
        A.pk_valid[PREC32] = false;
    }
    return pack_factor(ctx, A, PREC32);
}

// M^{-1} applied to vector r -> y (vectors of the PCG work set)
int precond_apply(fm_ctx* ctx, Matrix& A, int prec, const void* r, void* z, void* y, double* partials) {
    const void* lu = prec == PREC64 ? (const void*)A.lu64.p : (const void*)A.lu32.p;
    launch_fill_sentinel(ctx->stream, prec, z, y, A.n);
    {
        ProfScope ps(ctx, FM_K_ILU_FWD);
        if (A.chain.ok)
            launch_chain_fwd(ctx->stream, prec, A, r, z, ctx->st.p, false, ctx->d_err.p);
        else
            launch_ilu_fwd(ctx->stream, prec, A, lu, r, nullptr, nullptr, nullptr, nullptr, z, ctx->st.p, false,
                           ctx->d_err.p);
    }
    uint32_t np;
    {
        ProfScope ps(ctx, FM_K_ILU_BWD);
        np = A.chain.ok ? launch_chain_bwd(ctx->stream, prec, A, z, r, y, partials, ctx->st.p, false, ctx->d_err.p)
                        : launch_ilu_bwd(ctx->stream, prec, A, lu, z, r, y, partials, ctx->st.p, false, ctx->d_err.p);
    }
    return (int)np;
}

struct PcgResult {
    int iters;
    double relres;
    int status;
};

// pcg_solve_ilu0 (FEM/adaptive_fem_engine.hpp:1283-1341) on device vectors
// b -> x of the work set (x holds the initial guess).
// PCG iterations in batches of kBatch with one batch of lookahead: batch b+1
// is enqueued before the host waits for the state copied after batch b, so
// the device does not idle at a host check.  Iterations enqueued past
// convergence return at once (every kernel of the loop tests st->done).  On
// return the final state is in h_st[0].
// FM_TRACE_SOLVES=1 (read once per process): the Newton loop's linear
// solves and the PCG's relative residual at every batch on stderr
static bool trace_solves() {
    static const bool on = [] {
        const char* e = std::getenv("FM_TRACE_SOLVES");
        return e && *e == '1';
    }();
    return on;
}

template <typename F>
int pcg_batches(fm_ctx* ctx, int max_it, F&& enqueue_one) {
    const bool trace = trace_solves();
    hipStream_t s = ctx->stream;
    int enq = 0;
    for (int b = 0;; ++b) {
        for (int k = 0; k < kBatch && enq < max_it; ++k, ++enq) enqueue_one();
        FM_TRY(check_launch(ctx, "pcg iteration"));
        FM_HIP(hipMemcpyAsync(ctx->h_st + 1 + (b & 1), ctx->st.p, sizeof(PcgState), hipMemcpyDeviceToHost, s));
        FM_HIP(hipEventRecord(ctx->batch_ev[b & 1], s));
        if (b > 0) {
            FM_HIP(hipEventSynchronize(ctx->batch_ev[(b - 1) & 1]));
            const PcgState& h = ctx->h_st[1 + ((b - 1) & 1)];
            if (trace) std::fprintf(stderr, "[fm]   pcg %d relres %.3e\n", h.iters, h.relres);
            if (h.done) break;
        }
        if (enq >= max_it) break;
    }
    FM_HIP(hipMemcpyAsync(ctx->h_st, ctx->st.p, sizeof(PcgState), hipMemcpyDeviceToHost, s));
    return sync(ctx);
}

int pcg_run(fm_ctx* ctx, Matrix& A, int prec, double tol, int max_it, PcgResult* res) {
    hipStream_t s = ctx->stream;
    const uint32_t n = A.n;
    FM_TRY(factorize(ctx, A, prec));
    const uint32_t nbw = A.sched.q_bwd + 64;
    FM_TRY(ensure_partials(ctx, std::max<size_t>({(size_t)spmv_blocks(A, prec), (size_t)n / 256 + 1, (size_t)nbw})));
    void* b = vb(ctx, prec, VB);
    void* x = vb(ctx, prec, VX);
    void* r = vb(ctx, prec, VR);
    void* p = vb(ctx, prec, VP);
    void* Ap = vb(ctx, prec, VAP);
    void* z = vb(ctx, prec, VZ);
    void* y = vb(ctx, prec, VY);
    PcgState init{};
    init.tol = tol;
    init.max_it = max_it;
    // alpha and beta of every iteration, for fm_pcg_history
    FM_HIP(ctx->pcg_hist.alloc(2 * (size_t)kPcgHistCap));
    init.hist = ctx->pcg_hist.p;
    init.hist_cap = kPcgHistCap;
    ctx->pcg_hist_len = 0;
    *ctx->h_st = init;
    FM_HIP(hipMemcpyAsync(ctx->st.p, ctx->h_st, sizeof(PcgState), hipMemcpyHostToDevice, ctx->stream));
    // r = b - A x; ||r||^2 and b.b
    {
        ProfScope ps(ctx, FM_K_SPMV);
        launch_spmv(s, prec, A, x, Ap, false, nullptr, nullptr, nullptr, nullptr);
    }
    uint32_t np = launch_pcg_init_r(s, prec, n, b, Ap, r, ctx->partials.p);
    launch_reduce(s, ctx->partials.p, np, ctx->st.p, 0, 0);
    // b.b through the same kernel with a zero "Ap" (z is scratch here)
    if (prec == PREC64)
        launch_fill_f64(s, (double*)Ap, n, 0.0);
    else
        launch_fill_f32(s, (float*)Ap, n, 0.f);
    np = launch_pcg_init_r(s, prec, n, b, Ap, z, ctx->partials.p);
    launch_reduce(s, ctx->partials.p, np, ctx->st.p, 0, 1);
    FM_HIP(hipMemcpyAsync(ctx->h_st, ctx->st.p, sizeof(PcgState), hipMemcpyDeviceToHost, s));
    FM_TRY(sync(ctx));
    const double rr = ctx->h_st->slot[0], bb = ctx->h_st->slot[1];
    if (std::sqrt(rr) < tol * 1e-3) {
        res->iters = 0;
        res->relres = std::sqrt(rr);
        res->status = 1;
        return FM_OK;
    }
    ctx->h_st->bb = bb;
    ctx->h_st->tol = tol;
    ctx->h_st->max_it = max_it;
    FM_HIP(hipMemcpyAsync(ctx->st.p, ctx->h_st, sizeof(PcgState), hipMemcpyHostToDevice, s));
    // z = M^{-1} r (+ Z E^{-1} Z^T r); rz; p = z
    const int ad = ctx->ad_nmono;
    // the Lanczos vectors of this solve for its Ritz vectors (with the coarse level)
    void* lz = nullptr;
    uint32_t lz_cap = 0;
    if (ctx->harvest > 0 && ad) {
        if (prec == PREC64) {
            FM_HIP(ctx->lz64.alloc((size_t)n * kLanczosCap));
            lz = ctx->lz64.p;
        } else {
            FM_HIP(ctx->lz32.alloc((size_t)n * kLanczosCap));
            lz = ctx->lz32.p;
        }
        lz_cap = kLanczosCap;
    }
    if (ad) launch_coarse_correct(s, prec, n, ad, ctx->cs_xn.p, r, ctx->ad_Einv.p, ctx->ad_part.p, ctx->ad_coef.p, ctx->st.p);
    np = (uint32_t)precond_apply(ctx, A, prec, r, z, y, ctx->partials.p);
    launch_reduce(s, ctx->partials.p, np, ctx->st.p, 3, 0);
    if (ad)
        launch_pcg_p_update_coarse(s, prec, n, ad, y, p, ctx->cs_xn.p, ctx->ad_coef.p, true, ctx->st.p, lz, lz_cap);
    else
        launch_copy(s, prec, n, y, p);
    FM_TRY(check_launch(ctx, "pcg init"));
    const void* lu = prec == PREC64 ? (const void*)A.lu64.p : (const void*)A.lu32.p;
    FM_TRY(pcg_batches(ctx, max_it, [&] {
            {
                ProfScope ps(ctx, FM_K_SPMV);
                launch_spmv(s, prec, A, p, Ap, true, ctx->partials.p, z, y, ctx->st.p);
            }
            launch_reduce(s, ctx->partials.p, spmv_blocks(A, prec), ctx->st.p, 1, 0);
            {
                if (A.chain.ok) {
                    // the update runs as its own streaming kernel: fused, its
                    // stores would sit on the sweep's serial step path; with
                    // the coarse level it also forms the moments of the new r
                    if (ad) {
                        ProfScope pu(ctx, FM_K_COARSE);
                        launch_pcg_xr_update_moments(s, prec, n, ad, p, Ap, x, r, ctx->cs_xn.p, ctx->ad_part.p,
                                                     ctx->st.p);
                        launch_coarse_solve(s, ad, ctx->ad_part.p, ctx->ad_Einv.p, ctx->ad_coef.p, ctx->st.p);
                    } else {
                        ProfScope pu(ctx, FM_K_PCG_OTHER);
                        launch_pcg_xr_update(s, prec, n, p, Ap, x, r, ctx->st.p);
                    }
                    ProfScope ps(ctx, FM_K_ILU_FWD);
                    launch_chain_fwd(s, prec, A, r, z, ctx->st.p, true, ctx->d_err.p);
                } else {
                    ProfScope ps(ctx, FM_K_ILU_FWD);
                    launch_ilu_fwd(s, prec, A, lu, nullptr, r, x, p, Ap, z, ctx->st.p, true, ctx->d_err.p);
                }
            }
            uint32_t nbb;
            {
                ProfScope ps(ctx, FM_K_ILU_BWD);
                nbb = A.chain.ok
                          ? launch_chain_bwd(s, prec, A, z, r, y, ctx->partials.p, ctx->st.p, true, ctx->d_err.p)
                          : launch_ilu_bwd(s, prec, A, lu, z, r, y, ctx->partials.p, ctx->st.p, true, ctx->d_err.p);
            }
            if (ad && !A.chain.ok) {
                ProfScope pc(ctx, FM_K_COARSE);
                launch_coarse_correct(s, prec, n, ad, ctx->cs_xn.p, r, ctx->ad_Einv.p, ctx->ad_part.p,
                                      ctx->ad_coef.p, ctx->st.p);
            }
            launch_reduce(s, ctx->partials.p, nbb, ctx->st.p, 2, 0);
            if (ad)
                launch_pcg_p_update_coarse(s, prec, n, ad, y, p, ctx->cs_xn.p, ctx->ad_coef.p, false, ctx->st.p, lz,
                                           lz_cap);
            else
                launch_pcg_p_update(s, prec, n, y, p, ctx->st.p);
    }));
    FM_TRY(check_device_error(ctx));
    const PcgState& h = *ctx->h_st;
    res->status = h.status;
    res->relres = h.relres;
    res->iters = (h.status == 1) ? h.iters : max_it;
    ctx->pcg_hist_len = (uint32_t)std::min<int>(std::max(h.iters, 0), kPcgHistCap);
    return FM_OK;
}

// FP32 values derived on the device from the FP64 ones
int values_to_fp32(fm_ctx* ctx, Matrix& M) {
    FM_HIP(M.val32.alloc(M.nnz + kValPad));
    launch_f64_to_f32(ctx->stream, (uint32_t)M.nnz, M.val64.p, 1.0, M.val32.p);
    FM_TRY(check_launch(ctx, "f64_to_f32"));
    M.val_valid[PREC32] = true;
    return FM_OK;
}

// pcg_solve_ilu0 with FP64 vectors, SpMV and scalars, and the FP32 ILU(0)
// factor of the FP32-rounded matrix as preconditioner: z = M32^{-1} (float)r,
// r.z in FP64.  On an ill-conditioned matrix the FP32 application is a
// slightly variable preconditioner (its relative error grows with the
// condition number), so beta takes the flexible (Polak-Ribiere) form
// r_new.(z_new - z_old) / r_old.z_old, which equals the standard one for a
// fixed preconditioner.  On b64 -> x64 of the work set (x64: initial guess).
int pcg_run_pc(fm_ctx* ctx, Matrix& A, double tol, int max_it, PcgResult* res) {
    hipStream_t s = ctx->stream;
    const uint32_t n = A.n;
    FM_TRY(factorize_rounded(ctx, A));
    FM_TRY(pcg_vectors(ctx, PREC64, n));
    FM_TRY(pcg_vectors(ctx, PREC32, n));
    const uint32_t nbw = A.sched.q_bwd + 64;
    // partials, then the r.z_old partials of the flexible beta
    const size_t npart = std::max<size_t>({(size_t)spmv_blocks(A, PREC64), (size_t)n / 256 + 1, (size_t)nbw});
    FM_TRY(ensure_partials(ctx, npart + n / 256 + 1));
    double* old_part = ctx->partials.p + npart;
    double *b = ctx->b64.p, *x = ctx->x64.p, *r = ctx->r64.p, *p = ctx->p64.p, *Ap = ctx->Ap64.p, *y = ctx->y64.p;
    float *r32 = ctx->r32.p, *z32 = ctx->z32.p, *y32 = ctx->y32.p;
    const float* lu = A.lu32.p;
    PcgState init{};
    init.tol = tol;
    init.max_it = max_it;
    ctx->pcg_hist_len = 0;  // the flexible beta has no Lanczos relation: no coefficients recorded
    *ctx->h_st = init;
    FM_HIP(hipMemcpyAsync(ctx->st.p, ctx->h_st, sizeof(PcgState), hipMemcpyHostToDevice, s));
    {
        ProfScope ps(ctx, FM_K_SPMV);
        launch_spmv(s, PREC64, A, x, Ap, false, nullptr, nullptr, nullptr, nullptr);
    }
    uint32_t np = launch_pcg_init_r(s, PREC64, n, b, Ap, r, ctx->partials.p);
    launch_reduce(s, ctx->partials.p, np, ctx->st.p, 0, 0);
    launch_fill_f64(s, Ap, n, 0.0);
    np = launch_pcg_init_r(s, PREC64, n, b, Ap, y, ctx->partials.p);
    launch_reduce(s, ctx->partials.p, np, ctx->st.p, 0, 1);
    FM_HIP(hipMemcpyAsync(ctx->h_st, ctx->st.p, sizeof(PcgState), hipMemcpyDeviceToHost, s));
    FM_TRY(sync(ctx));
    const double rr = ctx->h_st->slot[0], bb = ctx->h_st->slot[1];
    if (std::sqrt(rr) < tol * 1e-3) {
        res->iters = 0;
        res->relres = std::sqrt(rr);
        res->status = 1;
        return FM_OK;
    }
    ctx->h_st->bb = bb;
    ctx->h_st->tol = tol;
    ctx->h_st->max_it = max_it;
    FM_HIP(hipMemcpyAsync(ctx->st.p, ctx->h_st, sizeof(PcgState), hipMemcpyHostToDevice, s));
    // y = M^{-1} r in FP32, cast back with the FP64 partials of r.y
    auto precond = [&](bool check_done) -> uint32_t {
        launch_f64_to_f32(s, n, r, 1.0, r32);
        launch_fill_sentinel(s, PREC32, z32, y32, n);
        {
            ProfScope ps(ctx, FM_K_ILU_FWD);
            if (A.chain.ok)
                launch_chain_fwd(s, PREC32, A, r32, z32, ctx->st.p, check_done, ctx->d_err.p);
            else
                launch_ilu_fwd(s, PREC32, A, lu, r32, nullptr, nullptr, nullptr, nullptr, z32, ctx->st.p, false,
                               ctx->d_err.p);
        }
        {
            ProfScope ps(ctx, FM_K_ILU_BWD);
            if (A.chain.ok)
                launch_chain_bwd(s, PREC32, A, z32, nullptr, y32, nullptr, ctx->st.p, check_done, ctx->d_err.p);
            else
                launch_ilu_bwd(s, PREC32, A, lu, z32, nullptr, y32, nullptr, ctx->st.p, check_done, ctx->d_err.p);
        }
        ProfScope pu(ctx, FM_K_PCG_OTHER);
        return launch_f32_to_f64_dot(s, n, y32, r, y, ctx->partials.p, check_done ? old_part : nullptr,
                                     check_done ? ctx->st.p : nullptr);
    };
    // with the additive coarse level (ctx->ad_nmono fields) z = y + Z c,
    // c = E^{-1} Z^T r, whose r.(Z c) the coarse solve adds to r.z
    const int ad = ctx->ad_nmono;
    if (ad) launch_coarse_correct(s, PREC64, n, ad, ctx->cs_xn.p, r, ctx->ad_Einv.p, ctx->ad_part.p, ctx->ad_coef.p, ctx->st.p);
    np = precond(false);
    launch_reduce(s, ctx->partials.p, np, ctx->st.p, 3, 0);
    if (ad)
        launch_pcg_p_update_coarse(s, PREC64, n, ad, y, p, ctx->cs_xn.p, ctx->ad_coef.p, true, ctx->st.p);
    else
        launch_copy(s, PREC64, n, y, p);
    FM_TRY(check_launch(ctx, "pcg init"));
    FM_TRY(pcg_batches(ctx, max_it, [&] {
            {
                ProfScope ps(ctx, FM_K_SPMV);
                launch_spmv(s, PREC64, A, p, Ap, true, ctx->partials.p, nullptr, nullptr, ctx->st.p);
            }
            launch_reduce(s, ctx->partials.p, spmv_blocks(A, PREC64), ctx->st.p, 1, 0);
            if (ad) {
                // the moments of the new r, then c and the coarse terms of r.z
                // and of the flexible beta (r_new.(Z c_old))
                ProfScope pu(ctx, FM_K_COARSE);
                launch_pcg_xr_update_moments(s, PREC64, n, ad, p, Ap, x, r, ctx->cs_xn.p, ctx->ad_part.p, ctx->st.p);
                launch_coarse_solve(s, ad, ctx->ad_part.p, ctx->ad_Einv.p, ctx->ad_coef.p, ctx->st.p, true);
            } else {
                ProfScope pu(ctx, FM_K_PCG_OTHER);
                launch_pcg_xr_update(s, PREC64, n, p, Ap, x, r, ctx->st.p);
            }
            np = precond(true);
            launch_reduce(s, old_part, np, ctx->st.p, 0, 3);
            launch_reduce(s, ctx->partials.p, np, ctx->st.p, 4, 3);
            if (ad)
                launch_pcg_p_update_coarse(s, PREC64, n, ad, y, p, ctx->cs_xn.p, ctx->ad_coef.p, false, ctx->st.p);
            else
                launch_pcg_p_update(s, PREC64, n, y, p, ctx->st.p);
    }));
    FM_TRY(check_device_error(ctx));
    const PcgState& h = *ctx->h_st;
    res->status = h.status;
    res->relres = h.relres;
    res->iters = (h.status == 1) ? h.iters : max_it;
    return FM_OK;
}

// Solve K du = rhs (FP64 device vector) with the tangent; du FP64.  The mixed
// solves, and the FP64 solve when `inexact`, work on the normalised right-hand
// side with a forcing-term tolerance.
// the share of the needed reduction a target-aware forcing term asks for: a
// solve achieves about half its forcing term's reduction (0.25 took 10% more
// PCG iterations over the 128^3 path, with the same worst final residual)
constexpr double kForcingShare = 0.5;

// Galerkin start of a PCG solve A x = b (b, x: the PCG vectors of precision
// p): x0 = V c with (V^T A V) c = V^T b over the chosen basis columns, i.e.
// the A-norm-best start in their span.  Columns are taken in order (the ring
// of recycled corrections, newest first, the Ritz vectors of the last solve,
// then the polynomial fields); one whose pivot in the Cholesky factor of
// V^T A V is not above 1e-10 of its diagonal adds nothing new and is left out.
constexpr int kCorr = 12;              // recycled corrections kept (basis columns 0..11)
constexpr int kRitz = 8;               // Ritz vectors of the last solve (columns 12..19)
constexpr int kRing = kCorr + kRitz;   // stored basis columns before the polynomial fields
static_assert(kRitz <= 8, "the combination kernel forms at most eight Ritz vectors");
constexpr int kMaxDegree = 4;
constexpr int kMonos[kMaxDegree + 1] = {0, 4, 10, 20, 35};  // monomials of total degree <= 0 (none), 1, 2, 3, 4
static_assert(kRing + 3 * kMonos[kMaxDegree] + 1 <= 128, "the Gram kernel takes at most 128 columns");

// the node coordinates mapped to [-1, 1] over the mesh's bounding box (FP32)
// and the constrained-component masks of the polynomial fields
int coarse_coords(fm_ctx* ctx) {
    double center[3], scale[3];
    for (int d = 0; d < 3; ++d) {
        const double* x = ctx->hX.data() + (size_t)d * ctx->nn;
        const auto mm = std::minmax_element(x, x + ctx->nn);
        center[d] = 0.5 * (*mm.first + *mm.second);
        scale[d] = *mm.second > *mm.first ? 2.0 / (*mm.second - *mm.first) : 1.0;
    }
    FM_HIP(ctx->cs_xn.alloc(ctx->nn));
    launch_coarse_coords(ctx->stream, ctx->nn, ctx->X.p, ctx->isdir.p, center, scale, ctx->cs_xn.p);
    return FM_OK;
}

// the basis matrices of precision p with room for the polynomial fields of
// degree deg (their columns rebuilt when missing)
int coarse_space(fm_ctx* ctx, int p, uint32_t n, int deg) {
    const int pi = p == PREC64 ? 0 : 1;
    const int nmono = kMonos[std::max(0, std::min(deg, kMaxDegree))];
    const int cols = kRing + 3 * nmono;
    if (cols > ctx->cs_cols[pi] || (pi == 0 ? ctx->cs_V64.n : ctx->cs_V32.n) < (size_t)n * cols) {
        const int c = std::max(cols, ctx->cs_cols[pi]);
        if (pi == 0) {
            FM_HIP(ctx->cs_V64.alloc((size_t)n * c));
            FM_HIP(ctx->cs_W64.alloc((size_t)n * (c + 1)));
        } else {
            FM_HIP(ctx->cs_V32.alloc((size_t)n * c));
            FM_HIP(ctx->cs_W32.alloc((size_t)n * (c + 1)));
        }
        ctx->cs_cols[pi] = c;
        ctx->cs_poly_ok[pi] = false;
        ctx->gal_n[pi] = 0;  // the ring's contents are gone
        ctx->nritz[pi] = 0;
    }
    if (nmono && (!ctx->cs_poly_ok[pi] || ctx->cs_nmono[pi] < nmono)) {
        FM_TRY(coarse_coords(ctx));
        const int build = std::min(std::max(nmono, ctx->cs_nmono[pi]), (ctx->cs_cols[pi] - kRing) / 3);
        void* V = pi == 0 ? (void*)ctx->cs_V64.p : (void*)ctx->cs_V32.p;
        launch_poly_basis(ctx->stream, p, ctx->nn, ctx->cs_xn.p, build, build, V, n, kRing);
        ctx->cs_nmono[pi] = build;
        ctx->cs_poly_ok[pi] = true;
    }
    return FM_OK;
}

// Cholesky factor of the symmetric matrix g over the columns cols (in
// order), leaving out a column whose pivot is not above 1e-10 of its
// diagonal; L (nc x nc, row-major) holds the kept columns' rows
template <typename G>
void pivoted_cholesky(const G& g, const int* cols, int nc, std::vector<double>& L, std::vector<int>& keep) {
    L.assign((size_t)nc * nc, 0.0);
    keep.clear();
    std::vector<double> row(nc);
    for (int q = 0; q < nc; ++q) {
        const int j = cols[q];
        const int k = (int)keep.size();
        for (int a = 0; a < k; ++a) {
            double v = g(keep[a], j);
            for (int t = 0; t < a; ++t) v -= L[(size_t)a * nc + t] * row[t];
            row[a] = v / L[(size_t)a * nc + a];
        }
        double d = g(j, j);
        for (int a = 0; a < k; ++a) d -= row[a] * row[a];
        if (!(d > 1e-10 * g(j, j))) continue;
        for (int a = 0; a < k; ++a) L[(size_t)k * nc + a] = row[a];
        L[(size_t)k * nc + k] = std::sqrt(d);
        keep.push_back(j);
    }
}

// additive: also E^{-1} over the polynomial fields for the additive coarse
// correction of the PCG that follows (ctx->ad_nmono)
int galerkin_start(fm_ctx* ctx, Matrix& A, int p, int ring_use, int deg, bool additive, bool reuse_poly = false,
                   int ritz_use = 0) {
    hipStream_t s = ctx->stream;
    const uint32_t n = A.n;
    const int pi = p == PREC64 ? 0 : 1;
    FM_TRY(coarse_space(ctx, p, n, deg));
    const int nmono = std::min(kMonos[std::max(0, std::min(deg, kMaxDegree))], ctx->cs_nmono[pi]);
    // the candidate columns, in order
    int cand[kRing + 3 * kMonos[kMaxDegree]], nc = 0;
    ring_use = std::min({ring_use, ctx->gal_n[pi], kCorr});
    ritz_use = std::min(ritz_use, ctx->nritz[pi]);
    for (int j = 0; j < ring_use; ++j) cand[nc++] = (ctx->gal_next[pi] - 1 - j + 2 * kCorr) % kCorr;
    for (int j = 0; j < ritz_use; ++j) cand[nc++] = kCorr + j;
    for (int c = 0; c < 3; ++c)
        for (int k = 0; k < nmono; ++k) cand[nc++] = kRing + c * ctx->cs_nmono[pi] + k;
    void* b = vb(ctx, p, VB);
    void* x = vb(ctx, p, VX);
    const size_t es = p == PREC64 ? 8 : 4;
    char* V = pi == 0 ? (char*)ctx->cs_V64.p : (char*)ctx->cs_V32.p;
    char* W = pi == 0 ? (char*)ctx->cs_W64.p : (char*)ctx->cs_W32.p;
    if (nc == 0) {
        if (p == PREC64)
            launch_fill_f64(s, (double*)x, n, 0.0);
        else
            launch_fill_f32(s, (float*)x, n, 0.f);
        return FM_OK;
    }
    // On a node-blocked matrix the polynomial fields' rows of the Gram matrix
    // come from one fused pass over the matrix (Z^T A Z, Z^T A V_ring and
    // Z^T b, without forming A Z), and the rows of the stored columns (the
    // recycled corrections and the Ritz vectors) from their products: G[ring]
    // over the kRing stored slots with W column kRing = b.  Otherwise every candidate column's
    // product is formed, and one Gram matrix covers the contiguous column
    // range [c0, c1) with W column c1 = b (FM_GALERKIN=generic selects this
    // path for every matrix; read once per process).
    static const bool generic = [] {
        const char* e = std::getenv("FM_GALERKIN");
        return e && std::string(e) == "generic";
    }();
    const bool fused = !generic && A.node_spmv && nmono > 0;
    int c0 = cand[0], c1 = cand[0] + 1;
    for (int q = 0; q < nc; ++q) {
        c0 = std::min(c0, cand[q]);
        c1 = std::max(c1, cand[q] + 1);
    }
    const int m = fused ? kRing : c1 - c0;
    const int mp = 3 * nmono, np = fused ? mp * (mp + kPolyGramStored) : 0;
    FM_HIP(ctx->cs_partials.alloc(std::max<size_t>(
        ctx->cs_partials.n, std::max((size_t)gram_blocks(m, m + 1) * m * (m + 1), (size_t)poly_gram_blocks() * np))));
    FM_HIP(ctx->cs_G.alloc(std::max<size_t>(ctx->cs_G.n, (size_t)m * (m + 1) + np)));
    static_assert(kPolyGramStored == kRing + 1, "the fused pass reads the ring products and b");
    // reuse_poly (a step that keeps an earlier step's factor): Z^T A Z of that
    // step's start stays; the moments of the stored columns (Z^T A v_ring,
    // Z^T b) come from the coarse level's moment kernel
    const bool reuse = fused && reuse_poly && ctx->gpp_ok && ctx->gpp_gen == ctx->fact_gen &&
                       ctx->gpp_nmono == nmono && ctx->gpp_prec == p;
    const uint32_t nmb = coarse_moment_blocks();
    if (reuse) {
        FM_HIP(ctx->cs_partials.alloc(std::max<size_t>(ctx->cs_partials.n, (size_t)nmb * mp)));
        FM_HIP(ctx->cs_G.alloc(std::max<size_t>(ctx->cs_G.n, (size_t)m * (m + 1) + (size_t)kPolyGramStored * mp)));
    }
    {
        ProfScope ps(ctx, FM_K_GALERKIN);
        // the ring candidates' products: one pass over the matrix for the
        // contiguous slot range that holds them where the multi-vector SpMV
        // applies (slots in the range that are not candidates get products
        // that are never read), one product per column otherwise
        int rlo = kRing, rhi = -1;
        for (int q = 0; q < nc; ++q)
            if (cand[q] < kRing) {
                rlo = std::min(rlo, cand[q]);
                rhi = std::max(rhi, cand[q]);
            }
        const bool multi = fused && rhi >= rlo &&
                           launch_spmv_multi(s, p, A, V + (size_t)rlo * n * es, W + (size_t)rlo * n * es, n,
                                             rhi - rlo + 1);
        for (int q = 0; q < nc; ++q)
            if (!fused || (cand[q] < kRing && !multi))
                launch_spmv(s, p, A, V + (size_t)cand[q] * n * es, W + (size_t)cand[q] * n * es, false, nullptr,
                            nullptr, nullptr, nullptr);
        const int bcol = fused ? kRing : c1;
        FM_HIP(hipMemcpyAsync(W + (size_t)bcol * n * es, b, n * es, hipMemcpyDeviceToDevice, s));
        if (fused) {
            // ring slots that are not candidates hold stale columns: their
            // entries are never read (the Ritz vectors are stored columns
            // too: the first solve after a reset has them without any
            // recycled correction)
            if (ring_use > 0 || ritz_use > 0) launch_gram(s, p, n, m, m + 1, V, W, ctx->cs_partials.p, ctx->cs_G.p);
            if (reuse) {
                // the moment kernels return at once on a finished PCG state
                PcgState init{};
                *ctx->h_st = init;
                FM_HIP(hipMemcpyAsync(ctx->st.p, ctx->h_st, sizeof(PcgState), hipMemcpyHostToDevice, s));
                for (int j = 0; j < kPolyGramStored; ++j) {
                    bool used = j == kRing;
                    for (int q = 0; q < nc && !used; ++q) used = cand[q] == j;
                    if (!used) continue;
                    launch_coarse_moments(s, p, n, nmono, ctx->cs_xn.p, W + (size_t)j * n * es, ctx->cs_partials.p,
                                          ctx->st.p);
                    launch_partial_sums(s, nmb, mp, ctx->cs_partials.p, ctx->cs_G.p + (size_t)m * (m + 1) + j * mp);
                }
            } else {
                launch_poly_gram(s, p, A, ctx->cs_xn.p, nmono, W, ctx->cs_partials.p,
                                 ctx->cs_G.p + (size_t)m * (m + 1));
            }
        } else {
            launch_gram(s, p, n, m, m + 1, V + (size_t)c0 * n * es, W + (size_t)c0 * n * es, ctx->cs_partials.p,
                        ctx->cs_G.p);
        }
    }
    std::vector<double> G((size_t)m * (m + 1) + np);
    const int ldp = mp + kPolyGramStored;
    if (reuse) {
        // Gp in the fused pass's layout: the kept block, then the moments
        std::vector<double> Gm((size_t)kPolyGramStored * mp);
        FM_HIP(hipMemcpyAsync(G.data(), ctx->cs_G.p, (size_t)m * (m + 1) * 8, hipMemcpyDeviceToHost, s));
        FM_HIP(hipMemcpyAsync(Gm.data(), ctx->cs_G.p + (size_t)m * (m + 1), Gm.size() * 8, hipMemcpyDeviceToHost, s));
        FM_TRY(sync(ctx));
        double* gp = G.data() + (size_t)m * (m + 1);
        for (int r = 0; r < mp; ++r) {
            for (int c = 0; c < mp; ++c) gp[(size_t)r * ldp + c] = ctx->gpp[(size_t)r * mp + c];
            for (int j = 0; j < kPolyGramStored; ++j) gp[(size_t)r * ldp + mp + j] = Gm[(size_t)j * mp + r];
        }
    } else {
        FM_HIP(hipMemcpyAsync(G.data(), ctx->cs_G.p, G.size() * 8, hipMemcpyDeviceToHost, s));
        FM_TRY(sync(ctx));
        if (fused) {
            // keep Z^T A Z for the steps that keep the factor this solve forms
            ctx->gpp.resize((size_t)mp * mp);
            const double* gp = G.data() + (size_t)m * (m + 1);
            for (int r = 0; r < mp; ++r)
                for (int c = 0; c < mp; ++c) ctx->gpp[(size_t)r * mp + c] = gp[(size_t)r * ldp + c];
            ctx->gpp_nmono = nmono;
            ctx->gpp_prec = p;
            ctx->gpp_gen = ctx->fact_gen + (A.lu_valid[p] ? 0 : 1);
            ctx->gpp_ok = true;
        }
    }
    const double* Gp = G.data() + (size_t)m * (m + 1);  // fused: mp x (mp + kPolyGramStored)
    auto prow = [&](int col) {  // row of polynomial V-column col in Gp
        const int o = col - kRing;
        return (o / ctx->cs_nmono[pi]) * nmono + o % ctx->cs_nmono[pi];
    };
    // symmetric part of V^T A V over V-column indices, and V^T b
    auto g = [&](int i, int j) -> double {
        if (!fused) {
            const int a = i - c0, c = j - c0;
            return 0.5 * (G[(size_t)a * (m + 1) + c] + G[(size_t)c * (m + 1) + a]);
        }
        if (i < kRing && j < kRing) return 0.5 * (G[(size_t)i * (m + 1) + j] + G[(size_t)j * (m + 1) + i]);
        if (i < kRing) return Gp[(size_t)prow(j) * ldp + mp + i];
        if (j < kRing) return Gp[(size_t)prow(i) * ldp + mp + j];
        const int a = prow(i), c = prow(j);
        return 0.5 * (Gp[(size_t)a * ldp + c] + Gp[(size_t)c * ldp + a]);
    };
    auto vtb = [&](int i) -> double {
        if (!fused) return G[(size_t)(i - c0) * (m + 1) + m];
        return i < kRing ? G[(size_t)i * (m + 1) + m] : Gp[(size_t)prow(i) * ldp + mp + kRing];
    };
    // Cholesky over the kept candidates, then the two triangular solves
    std::vector<double> L, y(nc), coef(nc);
    std::vector<int> keep;
    pivoted_cholesky(g, cand, nc, L, keep);
    const int k = (int)keep.size();
    for (int a = 0; a < k; ++a) {
        double v = vtb(keep[a]);
        for (int t = 0; t < a; ++t) v -= L[(size_t)a * nc + t] * y[t];
        y[a] = v / L[(size_t)a * nc + a];
    }
    for (int a = k - 1; a >= 0; --a) {
        double v = y[a];
        for (int t = a + 1; t < k; ++t) v -= L[(size_t)t * nc + a] * coef[t];
        coef[a] = v / L[(size_t)a * nc + a];
    }
    // x0: the kept ring columns combined from memory, the polynomial fields'
    // part evaluated from the node coordinates (coefficient c nmono + k, FP64
    // then FP32, the coarse kernels' layout)
    std::vector<int> rcol;
    std::vector<double> rcoef, pcoef(nmono ? 6 * nmono : 0, 0.0);
    for (int a = 0; a < k; ++a) {
        if (keep[a] < kRing) {
            rcol.push_back(keep[a]);
            rcoef.push_back(coef[a]);
        } else {
            const int col = keep[a] - kRing, c = col / ctx->cs_nmono[pi], q = col % ctx->cs_nmono[pi];
            pcoef[(size_t)c * nmono + q] = coef[a];
        }
    }
    for (int i = 0; i < 3 * nmono; ++i) reinterpret_cast<float*>(pcoef.data() + 3 * nmono)[i] = (float)pcoef[i];
    FM_HIP(ctx->cs_col.alloc(std::max<size_t>(ctx->cs_col.n, kRing + 3 * kMonos[kMaxDegree])));
    FM_HIP(ctx->cs_c.alloc(std::max<size_t>(ctx->cs_c.n, kRing + 6 * kMonos[kMaxDegree])));
    FM_HIP(hipMemcpyAsync(ctx->cs_col.p, rcol.data(), rcol.size() * sizeof(int), hipMemcpyHostToDevice, s));
    FM_HIP(hipMemcpyAsync(ctx->cs_c.p, rcoef.data(), rcoef.size() * 8, hipMemcpyHostToDevice, s));
    if (nmono) {
        FM_HIP(hipMemcpyAsync(ctx->cs_c.p + kRing, pcoef.data(), pcoef.size() * 8, hipMemcpyHostToDevice, s));
        void* z = vb(ctx, p, VZ);
        launch_basis_combine(s, p, n, (int)rcol.size(), ctx->cs_col.p, ctx->cs_c.p, V, z);
        // the coarse kernels return at once on a finished PCG state
        PcgState init{};
        *ctx->h_st = init;
        FM_HIP(hipMemcpyAsync(ctx->st.p, ctx->h_st, sizeof(PcgState), hipMemcpyHostToDevice, s));
        launch_pcg_p_update_coarse(s, p, n, nmono, z, x, ctx->cs_xn.p, ctx->cs_c.p + kRing, true, ctx->st.p);
    } else {
        launch_basis_combine(s, p, n, (int)rcol.size(), ctx->cs_col.p, ctx->cs_c.p, V, x);
    }
    // E^{-1} = (Z^T A Z)^{-1} over the polynomial fields (index c nmono + k),
    // from the same Gram entries; a left-out field gets a zero row and column
    const int ma = 3 * nmono;
    std::vector<double> Einv((size_t)ma * ma, 0.0);
    ctx->ad_nmono = 0;
    if (additive && ma > 0 && reuse && ctx->einv_ok && ctx->einv_gen == ctx->gpp_gen) {
        ctx->ad_nmono = nmono;  // the kept block's inverse is on the device already
    } else if (additive && ma > 0) {
        int pc[3 * kMonos[kMaxDegree]];
        for (int c = 0; c < 3; ++c)
            for (int q = 0; q < nmono; ++q) pc[c * nmono + q] = kRing + c * ctx->cs_nmono[pi] + q;
        std::vector<double> LE;
        std::vector<int> kept;
        pivoted_cholesky(g, pc, ma, LE, kept);
        const int ke = (int)kept.size();
        // inverse of the kept block: columns of L^{-T} L^{-1}, one unit vector at a time
        std::vector<double> e(ke), w(ke);
        std::vector<int> pos(ke);
        for (int a = 0; a < ke; ++a) pos[a] = (int)(std::find(pc, pc + ma, kept[a]) - pc);
        for (int col = 0; col < ke; ++col) {
            for (int a = 0; a < ke; ++a) {
                double v = a == col ? 1.0 : 0.0;
                for (int t = 0; t < a; ++t) v -= LE[(size_t)a * ma + t] * e[t];
                e[a] = v / LE[(size_t)a * ma + a];
            }
            for (int a = ke - 1; a >= 0; --a) {
                double v = e[a];
                for (int t = a + 1; t < ke; ++t) v -= LE[(size_t)t * ma + a] * w[t];
                w[a] = v / LE[(size_t)a * ma + a];
            }
            for (int a = 0; a < ke; ++a) Einv[(size_t)pos[a] * ma + pos[col]] = w[a];
        }
        FM_HIP(ctx->ad_Einv.alloc(std::max<size_t>(ctx->ad_Einv.n, (size_t)ma * ma)));
        FM_HIP(ctx->ad_coef.alloc(std::max<size_t>(ctx->ad_coef.n, (size_t)2 * ma)));
        FM_HIP(ctx->ad_part.alloc(std::max<size_t>(ctx->ad_part.n, (size_t)coarse_moment_blocks() * ma)));
        FM_HIP(hipMemcpyAsync(ctx->ad_Einv.p, Einv.data(), Einv.size() * 8, hipMemcpyHostToDevice, s));
        ctx->ad_nmono = nmono;
        ctx->einv_ok = fused && !reuse;
        ctx->einv_gen = ctx->gpp_gen;
    }
    ctx->lag_nmono = ctx->ad_nmono;
    ctx->lag_prec = p;
    FM_TRY(sync(ctx));  // the host arrays above are copied before they go out of scope
    return FM_OK;
}

// the start of a solve with a lagged preconditioner: x0 = Z E^{-1} Z^T b with
// the coarse operator E of the step's first tangent (the Galerkin start over
// the polynomial fields for that tangent), and the same E^{-1} as the additive
// coarse level of the PCG that follows
int lagged_start(fm_ctx* ctx, Matrix& A, int p) {
    hipStream_t s = ctx->stream;
    const uint32_t n = A.n;
    const int nm = ctx->lag_nmono;
    // the coarse kernels return at once on a finished PCG state
    PcgState init{};
    *ctx->h_st = init;
    FM_HIP(hipMemcpyAsync(ctx->st.p, ctx->h_st, sizeof(PcgState), hipMemcpyHostToDevice, s));
    void* b = vb(ctx, p, VB);
    void* x = vb(ctx, p, VX);
    void* z = vb(ctx, p, VZ);
    {
        ProfScope ps(ctx, FM_K_GALERKIN);
        launch_coarse_correct(s, p, n, nm, ctx->cs_xn.p, b, ctx->ad_Einv.p, ctx->ad_part.p, ctx->ad_coef.p,
                              ctx->st.p);
        if (p == PREC64)
            launch_fill_f64(s, (double*)z, n, 0.0);
        else
            launch_fill_f32(s, (float*)z, n, 0.f);
        launch_pcg_p_update_coarse(s, p, n, nm, z, x, ctx->cs_xn.p, ctx->ad_coef.p, true, ctx->st.p);
    }
    FM_TRY(check_launch(ctx, "lagged start"));
    ctx->ad_nmono = nm;
    return FM_OK;
}

// eigenpairs of the symmetric m x m matrix T (row-major, overwritten) by
// cyclic Jacobi rotations (Golub & Van Loan, Alg. 8.5.3): w the eigenvalues,
// Q (row-major) the eigenvectors in its columns
void jacobi_eigen(int m, std::vector<double>& T, std::vector<double>& w, std::vector<double>& Q) {
    Q.assign((size_t)m * m, 0.0);
    for (int i = 0; i < m; ++i) Q[(size_t)i * m + i] = 1.0;
    for (int sweep = 0; sweep < 100; ++sweep) {
        double off = 0.0, nrm = 0.0;
        for (int i = 0; i < m; ++i)
            for (int j = 0; j < m; ++j) {
                const double v = T[(size_t)i * m + j] * T[(size_t)i * m + j];
                nrm += v;
                if (i != j) off += v;
            }
        if (off <= 1e-30 * nrm) break;
        for (int p = 0; p < m; ++p)
            for (int q = p + 1; q < m; ++q) {
                const double apq = T[(size_t)p * m + q];
                if (apq == 0.0) continue;
                const double app = T[(size_t)p * m + p], aqq = T[(size_t)q * m + q];
                const double theta = (aqq - app) / (2.0 * apq);
                const double t = (theta >= 0 ? 1.0 : -1.0) / (std::fabs(theta) + std::sqrt(theta * theta + 1.0));
                const double c = 1.0 / std::sqrt(t * t + 1.0), sn = t * c;
                for (int k = 0; k < m; ++k) {  // columns p, q
                    const double akp = T[(size_t)k * m + p], akq = T[(size_t)k * m + q];
                    T[(size_t)k * m + p] = c * akp - sn * akq;
                    T[(size_t)k * m + q] = sn * akp + c * akq;
                }
                for (int k = 0; k < m; ++k) {  // rows p, q
                    const double apk = T[(size_t)p * m + k], aqk = T[(size_t)q * m + k];
                    T[(size_t)p * m + k] = c * apk - sn * aqk;
                    T[(size_t)q * m + k] = sn * apk + c * aqk;
                }
                for (int k = 0; k < m; ++k) {
                    const double qkp = Q[(size_t)k * m + p], qkq = Q[(size_t)k * m + q];
                    Q[(size_t)k * m + p] = c * qkp - sn * qkq;
                    Q[(size_t)k * m + q] = sn * qkp + c * qkq;
                }
            }
    }
    w.resize(m);
    for (int i = 0; i < m; ++i) w[i] = T[(size_t)i * m + i];
}

// The Ritz vectors of the last PCG solve (run with ctx->harvest set): with
// z_j the preconditioned residuals, the vectors z_j / sqrt(r_j.z_j) kept by
// the solve and the Lanczos matrix of its alpha, beta (T_jj = 1/alpha_j +
// beta_(j-1)/alpha_(j-1), T_j,j+1 = -sqrt(beta_j)/alpha_j) give P A V = V T,
// P the preconditioner; the k eigenvectors of T with the smallest eigenvalues
// combine V into approximations of the slowest modes of the preconditioned
// operator, stored in the basis columns after the recycled corrections.  A
// solve with fewer than 2k iterations keeps the previous ones.
int ritz_harvest(fm_ctx* ctx, uint32_t n, int p, int k) {
    const int pi = p == PREC64 ? 0 : 1;
    const int m = std::min<int>((int)ctx->pcg_hist_len, kLanczosCap);
    k = std::min(k, kRitz);
    if (k <= 0 || m < 2 * k) return FM_OK;
    std::vector<double> h(2 * (size_t)m);
    FM_HIP(hipMemcpyAsync(h.data(), ctx->pcg_hist.p, h.size() * 8, hipMemcpyDeviceToHost, ctx->stream));
    FM_TRY(sync(ctx));
    std::vector<double> T((size_t)m * m, 0.0), w, Q;
    for (int j = 0; j < m; ++j) {
        const double a = h[2 * j];
        T[(size_t)j * m + j] = 1.0 / a + (j > 0 ? h[2 * (j - 1) + 1] / h[2 * (j - 1)] : 0.0);
        if (j + 1 < m) {
            const double e = -std::sqrt(std::max(h[2 * j + 1], 0.0)) / a;
            T[(size_t)j * m + j + 1] = e;
            T[(size_t)(j + 1) * m + j] = e;
        }
    }
    jacobi_eigen(m, T, w, Q);
    std::vector<int> idx(m);
    std::iota(idx.begin(), idx.end(), 0);
    std::sort(idx.begin(), idx.end(), [&](int a, int b) { return w[a] < w[b]; });
    std::vector<double> S((size_t)m * k);
    for (int j = 0; j < m; ++j)
        for (int q = 0; q < k; ++q) S[(size_t)j * k + q] = Q[(size_t)j * m + idx[q]];
    FM_HIP(ctx->ritz_S.alloc(std::max<size_t>(ctx->ritz_S.n, (size_t)kLanczosCap * kRitz)));
    FM_HIP(hipMemcpyAsync(ctx->ritz_S.p, S.data(), S.size() * 8, hipMemcpyHostToDevice, ctx->stream));
    const size_t es = p == PREC64 ? 8 : 4;
    char* V = pi == 0 ? (char*)ctx->cs_V64.p : (char*)ctx->cs_V32.p;
    const void* Z = pi == 0 ? (const void*)ctx->lz64.p : (const void*)ctx->lz32.p;
    launch_basis_combine_multi(ctx->stream, p, n, m, k, ctx->ritz_S.p, Z, V + (size_t)kCorr * n * es);
    FM_TRY(check_launch(ctx, "ritz vectors"));
    FM_TRY(sync(ctx));  // S is copied before it goes out of scope
    ctx->nritz[pi] = k;
    return FM_OK;
}

// the solution of the last solve (the PCG vector x of precision p), normalised,
// becomes the newest ring column
int galerkin_push(fm_ctx* ctx, uint32_t n, int p, int deg) {
    hipStream_t s = ctx->stream;
    const int pi = p == PREC64 ? 0 : 1;
    FM_TRY(coarse_space(ctx, p, n, deg));
    const int slot = ctx->gal_next[pi];
    const size_t es = p == PREC64 ? 8 : 4;
    char* V = pi == 0 ? (char*)ctx->cs_V64.p : (char*)ctx->cs_V32.p;
    const void* x = vb(ctx, p, VX);
    FM_TRY(ensure_partials(ctx, n / 256 + 1));
    const uint32_t np = launch_dot(s, p, n, x, x, ctx->partials.p);
    launch_reduce(s, ctx->partials.p, np, ctx->st.p, 5, 47);
    launch_normalize(s, p, n, &ctx->st.p->gal[47], x, V + (size_t)slot * n * es);
    ctx->gal_next[pi] = (slot + 1) % kCorr;
    ctx->gal_n[pi] = std::min(ctx->gal_n[pi] + 1, kCorr);
    return FM_OK;
}

// FP32 and inexact FP64 solves: ring_use > 0 or deg > 0: the solve starts
// from the Galerkin start over that many of the latest recycled corrections
// and the polynomial fields of degree <= deg; push: its solution joins the
// recycled corrections
int linear_solve(fm_ctx* ctx, int prec, const double* rhs, double* du, double tol, int max_it, int* iters,
                 bool inexact = false, int* capped = nullptr, int ring_use = 0, int deg = 0, bool push = false,
                 bool additive = false, bool lag = false, bool reuse_poly = false, int ritz = 0,
                 int ritz_use = 0) {
    ctx->ad_nmono = 0;
    Matrix& A = ctx->tangent;
    ctx->cur = &A;
    const uint32_t n = A.n;
    PcgResult pr{};
    double nr = 1.0;
    if (prec != PREC64 || inexact) {
        // the inner solves work on the normalised right-hand side: the
        // reference's absolute early exit (||r|| < tol 1e-3) would otherwise
        // end a loose forcing-term solve at once on a small load-step residual
        FM_TRY(ensure_partials(ctx, n / 256 + 1));
        const uint32_t np = launch_dot64(ctx->stream, n, rhs, rhs, ctx->partials.p);
        launch_reduce(ctx->stream, ctx->partials.p, np, ctx->st.p, 0, 2);
        FM_HIP(hipMemcpyAsync(ctx->h_st, ctx->st.p, sizeof(PcgState), hipMemcpyDeviceToHost, ctx->stream));
        FM_TRY(sync(ctx));
        nr = std::sqrt(ctx->h_st->slot[2]);
        if (nr == 0.0) {
            launch_fill_f64(ctx->stream, du, n, 0.0);
            *iters = 0;
            return FM_OK;
        }
    }
    if (prec == PRECPC) {
        // FP64 Krylov vectors: the FP64 path's Galerkin start, recycled
        // corrections and coarse level (no Ritz vectors)
        FM_TRY(pcg_vectors(ctx, PREC64, n));
        launch_scale(ctx->stream, n, rhs, 1.0 / nr, ctx->b64.p);
        if (lag && additive && ring_use == 0 && ctx->lag_nmono > 0 && ctx->lag_prec == PREC64)
            FM_TRY(lagged_start(ctx, A, PREC64));
        else if ((ring_use > 0 && ctx->gal_n[0] > 0) || deg > 0)
            FM_TRY(galerkin_start(ctx, A, PREC64, ring_use, deg, additive, reuse_poly, 0));
        else
            launch_fill_f64(ctx->stream, ctx->x64.p, n, 0.0);
        const int prc = pcg_run_pc(ctx, A, tol, max_it, &pr);
        ctx->ad_nmono = 0;
        FM_TRY(prc);
        if (push) FM_TRY(galerkin_push(ctx, n, PREC64, deg));
        launch_scale(ctx->stream, n, ctx->x64.p, nr, du);
        *iters = pr.iters;
        if (capped && pr.status != 1) ++*capped;
        return FM_OK;
    }
    FM_TRY(pcg_vectors(ctx, prec, n));
    if (prec == PREC64 && inexact) {
        launch_scale(ctx->stream, n, rhs, 1.0 / nr, ctx->b64.p);
        if (lag && additive && ring_use == 0 && ctx->lag_nmono > 0 && ctx->lag_prec == PREC64)
            FM_TRY(lagged_start(ctx, A, PREC64));
        else if ((ring_use > 0 && ctx->gal_n[0] > 0) || deg > 0)
            FM_TRY(galerkin_start(ctx, A, PREC64, ring_use, deg, additive, reuse_poly, ritz_use));
        else
            launch_fill_f64(ctx->stream, ctx->x64.p, n, 0.0);
        ctx->harvest = additive ? ritz : 0;
        const int prc = pcg_run(ctx, A, PREC64, tol, max_it, &pr);
        const bool harvested = ctx->harvest > 0 && ctx->ad_nmono > 0;
        ctx->harvest = 0;
        ctx->ad_nmono = 0;
        FM_TRY(prc);
        if (harvested) FM_TRY(ritz_harvest(ctx, n, PREC64, ritz));
        if (push) FM_TRY(galerkin_push(ctx, n, PREC64, deg));
        launch_scale(ctx->stream, n, ctx->x64.p, nr, du);
    } else if (prec == PREC64) {
        FM_HIP(hipMemcpyAsync(ctx->b64.p, rhs, (size_t)n * 8, hipMemcpyDeviceToDevice, ctx->stream));
        launch_fill_f64(ctx->stream, ctx->x64.p, n, 0.0);
        FM_TRY(pcg_run(ctx, A, PREC64, tol, max_it, &pr));
        FM_HIP(hipMemcpyAsync(du, ctx->x64.p, (size_t)n * 8, hipMemcpyDeviceToDevice, ctx->stream));
    } else {
        // FP32 inner solve on the normalised right-hand side
        launch_f64_to_f32(ctx->stream, n, rhs, 1.0 / nr, ctx->b32.p);
        if (lag && additive && ring_use == 0 && ctx->lag_nmono > 0 && ctx->lag_prec == PREC32)
            FM_TRY(lagged_start(ctx, A, PREC32));
        else if ((ring_use > 0 && ctx->gal_n[1] > 0) || deg > 0)
            FM_TRY(galerkin_start(ctx, A, PREC32, ring_use, deg, additive, reuse_poly, ritz_use));
        else
            launch_fill_f32(ctx->stream, ctx->x32.p, n, 0.f);
        ctx->harvest = additive ? ritz : 0;
        const int prc = pcg_run(ctx, A, PREC32, tol, max_it, &pr);
        const bool harvested = ctx->harvest > 0 && ctx->ad_nmono > 0;
        ctx->harvest = 0;
        ctx->ad_nmono = 0;
        FM_TRY(prc);
        if (harvested) FM_TRY(ritz_harvest(ctx, n, PREC32, ritz));
        if (push) FM_TRY(galerkin_push(ctx, n, PREC32, deg));
        launch_f32_to_f64(ctx->stream, n, ctx->x32.p, nr, du);
    }
    *iters = pr.iters;
    if (capped && pr.status != 1) ++*capped;
    return FM_OK;
}

void fnv1a_update(uint64_t& h, const void* data, size_t n) {
    const uint8_t* p = (const uint8_t*)data;
    for (size_t i = 0; i < n; ++i) {
        h ^= p[i];
        h *= 1099511628211ull;
    }
}

// counter-based generator for the synthetic inputs (splitmix64 finaliser)
inline uint64_t mix64(uint64_t z) {
    z += 0x9E3779B97F4A7C15ull;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}
inline double uniform01(uint64_t seed, uint64_t stream, uint64_t i) {
    return (double)(mix64(mix64(seed ^ (stream * 0xD1B54A32D192ED03ull)) + i) >> 11) * 0x1.0p-53;
}

}  // namespace

// ===========================================================================
// C ABI
// ===========================================================================
extern "C" {

const char* fm_version(void) { return "flexmesh-mi355x 0.1 (gfx950)"; }

int fm_create(int device, fm_ctx** out) {
    if (!out) return FM_EINVAL;
    *out = nullptr;
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev == 0) return FM_EHIP;
    if (device < 0 || device >= ndev) return FM_EINVAL;
    fm_ctx* ctx = new fm_ctx();
    ctx->device = device;
    if (hipSetDevice(device) != hipSuccess || hipStreamCreateWithFlags(&ctx->stream, hipStreamNonBlocking) != hipSuccess ||
        ctx->st.alloc(1) != hipSuccess || ctx->d_err.alloc(1) != hipSuccess ||
        hipHostMalloc((void**)&ctx->h_st, 3 * sizeof(PcgState), hipHostMallocDefault) != hipSuccess ||
        hipEventCreateWithFlags(&ctx->batch_ev[0], hipEventDisableTiming) != hipSuccess ||
        hipEventCreateWithFlags(&ctx->batch_ev[1], hipEventDisableTiming) != hipSuccess) {
        delete ctx;
        return FM_EHIP;
    }
    int zero = 0;
    (void)hipMemcpy(ctx->d_err.p, &zero, sizeof(int), hipMemcpyHostToDevice);
    // the PCG state starts zeroed
    launch_fill_u8(ctx->stream, reinterpret_cast<uint8_t*>(ctx->st.p), sizeof(PcgState), 0);
    if (hipStreamSynchronize(ctx->stream) != hipSuccess) {
        fm_destroy(ctx);
        return FM_EHIP;
    }
    // the allocation-free residual's scratch address on this device
    const hipError_t se = noalloc_scratch_init(device);
    if (se != hipSuccess) {
        fm_destroy(ctx);
        return FM_EHIP;
    }
    *out = ctx;
    return FM_OK;
}

int fm_destroy(fm_ctx* ctx) {
    if (!ctx) return FM_OK;
    (void)hipSetDevice(ctx->device);
    (void)hipStreamSynchronize(ctx->stream);
    prof_flush(ctx);
    for (auto e : ctx->ev_pool) (void)hipEventDestroy(e);
    for (auto e : ctx->batch_ev)
        if (e) (void)hipEventDestroy(e);
    if (ctx->h_st) (void)hipHostFree(ctx->h_st);
    if (ctx->stream) (void)hipStreamDestroy(ctx->stream);
    delete ctx;
    return FM_OK;
}

const char* fm_last_error(const fm_ctx* ctx) { return ctx ? ctx->err.c_str() : "null context"; }

int fm_mesh_box(int nx, int ny, int nz, double Lx, double Ly, double Lz, double* X, uint32_t* conn) {
    if (nx < 1 || ny < 1 || nz < 1 || !X || !conn) return FM_EINVAL;
    // Kuhn split (all six tets share the 0-6 diagonal of the hex, corner
    // numbering of the reference hex_off table, FEM/adaptive_fem_engine.hpp:2245-2248);
    // vertex order gives positive orientation on the undistorted grid
    static const int kKuhn[6][4] = {{0, 1, 2, 6}, {0, 5, 1, 6}, {0, 2, 3, 6}, {0, 3, 7, 6}, {0, 4, 5, 6}, {0, 7, 4, 6}};
    static const int kOff[8][3] = {{0, 0, 0}, {1, 0, 0}, {1, 1, 0}, {0, 1, 0}, {0, 0, 1}, {1, 0, 1}, {1, 1, 1}, {0, 1, 1}};
    const uint64_t NX = nx + 1, NY = ny + 1, NZ = nz + 1;
    const uint64_t nn = NX * NY * NZ, ne = 6ull * nx * ny * nz;
    if (nn > 0xFFFFFFFFull || ne > 0xFFFFFFFFull) return FM_EINVAL;
    const double hx = Lx / nx, hy = Ly / ny, hz = Lz / nz;
#pragma omp parallel for schedule(static)
    for (int64_t k = 0; k < (int64_t)NZ; ++k)
        for (uint64_t j = 0; j < NY; ++j)
            for (uint64_t i = 0; i < NX; ++i) {
                const uint64_t v = k * NX * NY + j * NX + i;
                X[v] = i * hx;
                X[nn + v] = j * hy;
                X[2 * nn + v] = k * hz;
            }
#pragma omp parallel for schedule(static)
    for (int64_t k = 0; k < nz; ++k)
        for (int j = 0; j < ny; ++j)
            for (int i = 0; i < nx; ++i) {
                const uint64_t e0 = 6ull * (((uint64_t)k * ny + j) * nx + i);
                uint32_t hex[8];
                for (int c = 0; c < 8; ++c)
                    hex[c] = (uint32_t)((k + kOff[c][2]) * NX * NY + (j + kOff[c][1]) * NX + (i + kOff[c][0]));
                for (int t = 0; t < 6; ++t)
                    for (int a = 0; a < 4; ++a) conn[a * ne + e0 + t] = hex[kKuhn[t][a]];
            }
    return FM_OK;
}

int fm_mesh_jitter(uint32_t n_nodes, double* X, const uint8_t* is_interior, double amp, uint64_t seed) {
    if (!X || !is_interior) return FM_EINVAL;
#pragma omp parallel for schedule(static)
    for (int64_t v = 0; v < (int64_t)n_nodes; ++v) {
        if (!is_interior[v]) continue;
        for (int d = 0; d < 3; ++d) X[(size_t)d * n_nodes + v] += amp * (2.0 * uniform01(seed, 1 + d, v) - 1.0);
    }
    return FM_OK;
}

int fm_field_lognormal(uint32_t n, double median, double sigma, uint64_t seed, double* out) {
    if (!out) return FM_EINVAL;
#pragma omp parallel for schedule(static)
    for (int64_t e = 0; e < (int64_t)n; ++e) {
        const double u1 = 1.0 - uniform01(seed, 11, e);  // (0, 1]
        const double u2 = uniform01(seed, 12, e);
        const double z = std::sqrt(-2.0 * std::log(u1)) * std::cos(6.283185307179586 * u2);
        out[e] = median * std::exp(sigma * z);
    }
    return FM_OK;
}

int fm_set_mesh(fm_ctx* ctx, uint32_t n_nodes, const double* X, uint32_t n_elems, const uint32_t* conn) {
    if (!ctx || !X || !conn || !n_nodes || !n_elems) return fail(ctx, FM_EINVAL, "fm_set_mesh: bad arguments");
    ctx->sec_n = 0;
    ctx->slope_ok = false;
    ctx->gal_n[0] = ctx->gal_n[1] = 0;
    ctx->nritz[0] = ctx->nritz[1] = 0;
    ctx->cs_poly_ok[0] = ctx->cs_poly_ok[1] = false;  // the polynomial fields depend on X and the constraints
    if (3ull * n_nodes > 0xFFFFFFFFull) return fail(ctx, FM_EINVAL, "too many nodes");
    for (size_t k = 0; k < 4ull * n_elems; ++k)
        if (conn[k] >= n_nodes) return fail(ctx, FM_EINVAL, "connectivity refers to a node out of range");
    FM_HIP(hipSetDevice(ctx->device));
    ctx->nn = n_nodes;
    ctx->ne = n_elems;
    ctx->hX.assign(X, X + 3ull * n_nodes);
    ctx->hconn.assign(conn, conn + 4ull * n_elems);
    ctx->have_pattern = ctx->have_colors = ctx->have_material = false;
    ctx->hmu.clear();
    ctx->hlam.clear();
    ctx->u_valid = false;
    ctx->dir_dofs.clear();
    ctx->dir_vals.clear();
    ctx->have_fext = false;
    FM_HIP(ctx->X.alloc(3ull * n_nodes));
    FM_HIP(hipMemcpy(ctx->X.p, X, 3ull * n_nodes * 8, hipMemcpyHostToDevice));
    FM_TRY(ensure_vectors(ctx));
    launch_fill_u8(ctx->stream, ctx->isdir.p, ctx->isdir.n, 0);
    FM_HIP(hipStreamSynchronize(ctx->stream));
    return FM_OK;
}

int fm_set_coords(fm_ctx* ctx, uint32_t n_nodes, const double* X) {
    if (!ctx || !X) return fail(ctx, FM_EINVAL, "fm_set_coords: bad arguments");
    if (!ctx->nn) return fail(ctx, FM_ESTATE, "no mesh");
    if (n_nodes != ctx->nn) return fail(ctx, FM_EINVAL, "fm_set_coords: node count differs from the mesh");
    FM_HIP(hipSetDevice(ctx->device));
    ctx->hX.assign(X, X + 3ull * n_nodes);
    FM_HIP(hipMemcpy(ctx->X.p, X, 3ull * n_nodes * 8, hipMemcpyHostToDevice));
    ctx->sec_n = 0;
    ctx->slope_ok = false;
    ctx->gal_n[0] = ctx->gal_n[1] = 0;
    ctx->nritz[0] = ctx->nritz[1] = 0;
    ctx->cs_poly_ok[0] = ctx->cs_poly_ok[1] = false;  // the polynomial fields depend on X and the constraints
    Matrix& M = ctx->tangent;
    M.val_valid[0] = M.val_valid[1] = false;
    M.lu_valid[0] = M.lu_valid[1] = false;
    M.pk_valid[0] = M.pk_valid[1] = false;
    return FM_OK;
}

int fm_set_material(fm_ctx* ctx, const double* mu_e, const double* lam_e) {
    if (!ctx || !mu_e || !lam_e) return fail(ctx, FM_EINVAL, "fm_set_material: bad arguments");
    ctx->sec_n = 0;
    ctx->slope_ok = false;
    ctx->gal_n[0] = ctx->gal_n[1] = 0;
    ctx->nritz[0] = ctx->nritz[1] = 0;
    if (!ctx->nn) return fail(ctx, FM_ESTATE, "no mesh");
    ctx->hmu.assign(mu_e, mu_e + ctx->ne);
    ctx->hlam.assign(lam_e, lam_e + ctx->ne);
    FM_TRY(ensure_colors(ctx));
    return upload_material(ctx);
}

int fm_set_bcs(fm_ctx* ctx, uint32_t n_dir, const uint32_t* dofs, const double* vals) {
    if (!ctx || (n_dir && !dofs)) return fail(ctx, FM_EINVAL, "fm_set_bcs: bad arguments");
    ctx->sec_n = 0;
    ctx->slope_ok = false;
    ctx->gal_n[0] = ctx->gal_n[1] = 0;
    ctx->nritz[0] = ctx->nritz[1] = 0;
    ctx->cs_poly_ok[0] = ctx->cs_poly_ok[1] = false;  // the polynomial fields depend on X and the constraints
    if (!ctx->nn) return fail(ctx, FM_ESTATE, "no mesh");
    const uint32_t n = 3 * ctx->nn;
    std::vector<uint8_t> mask(n, 0);
    for (uint32_t k = 0; k < n_dir; ++k) {
        if (dofs[k] >= n) return fail(ctx, FM_EINVAL, "Dirichlet DOF out of range");
        if (mask[dofs[k]]) return fail(ctx, FM_EINVAL, "duplicate Dirichlet DOF");
        mask[dofs[k]] = 1;
    }
    ctx->dir_dofs.assign(dofs, dofs + n_dir);
    ctx->dir_vals.assign(n_dir, 0.0);
    if (vals) std::copy(vals, vals + n_dir, ctx->dir_vals.begin());
    FM_HIP(hipMemcpy(ctx->isdir.p, mask.data(), n, hipMemcpyHostToDevice));
    FM_HIP(ctx->d_dir_dofs.alloc(std::max<uint32_t>(n_dir, 1)));
    FM_HIP(ctx->d_dir_vals.alloc(std::max<uint32_t>(n_dir, 1)));
    if (n_dir) {
        FM_HIP(hipMemcpy(ctx->d_dir_dofs.p, dofs, n_dir * 4ull, hipMemcpyHostToDevice));
        FM_HIP(hipMemcpy(ctx->d_dir_vals.p, ctx->dir_vals.data(), n_dir * 8ull, hipMemcpyHostToDevice));
    }
    return FM_OK;
}

int fm_set_load(fm_ctx* ctx, const double* f_ext) {
    if (!ctx) return FM_EINVAL;
    ctx->sec_n = 0;
    ctx->slope_ok = false;
    ctx->gal_n[0] = ctx->gal_n[1] = 0;
    ctx->nritz[0] = ctx->nritz[1] = 0;
    if (!ctx->nn) return fail(ctx, FM_ESTATE, "no mesh");
    if (!f_ext) {
        ctx->have_fext = false;
        return FM_OK;
    }
    FM_HIP(ctx->f_ext.alloc(3ull * ctx->nn));
    FM_HIP(hipMemcpy(ctx->f_ext.p, f_ext, 3ull * ctx->nn * 8, hipMemcpyHostToDevice));
    ctx->have_fext = true;
    return FM_OK;
}

int fm_build_pattern(fm_ctx* ctx, uint32_t* row_ptr_out, uint32_t* col_idx_out, uint64_t* nnz_out) {
    if (!ctx) return FM_EINVAL;
    FM_TRY(ensure_pattern(ctx));
    Matrix& M = ctx->tangent;
    if (nnz_out) *nnz_out = M.nnz;
    if (row_ptr_out) std::copy(M.h_row_ptr.begin(), M.h_row_ptr.end(), row_ptr_out);
    if (col_idx_out) FM_HIP(hipMemcpy(col_idx_out, M.col_idx.p, M.nnz * 4, hipMemcpyDeviceToHost));
    return FM_OK;
}

int fm_color(fm_ctx* ctx, uint32_t* colors_out, uint32_t* n_colors_out) {
    if (!ctx) return FM_EINVAL;
    FM_TRY(ensure_colors(ctx));
    if (colors_out) std::copy(ctx->hcolors.begin(), ctx->hcolors.end(), colors_out);
    if (n_colors_out) *n_colors_out = ctx->n_colors;
    return FM_OK;
}

int fm_assemble(fm_ctx* ctx, const double* u, double* f_int_out, double* K_vals_out, int prec) {
    if (!ctx || (prec != FM_FP64 && prec != FM_MIXED)) return fail(ctx, FM_EINVAL, "fm_assemble: bad arguments");
    FM_TRY(ensure_ready(ctx));
    const size_t n = 3ull * ctx->nn;
    const double* du = nullptr;
    if (u) {
        FM_HIP(hipMemcpy(ctx->hv.p, u, n * 8, hipMemcpyHostToDevice));
        du = ctx->hv.p;
    }
    const int p = prec == FM_FP64 ? PREC64 : PREC32;
    FM_TRY(assemble(ctx, du, true, p, false, nullptr, FM_K_ASSEMBLE));
    FM_TRY(sync(ctx));
    if (f_int_out) FM_HIP(hipMemcpy(f_int_out, ctx->f.p, n * 8, hipMemcpyDeviceToHost));
    if (K_vals_out) {
        Matrix& M = ctx->tangent;
        if (p == PREC64) {
            FM_HIP(hipMemcpy(K_vals_out, M.val64.p, M.nnz * 8, hipMemcpyDeviceToHost));
        } else {
            std::vector<float> tmp(M.nnz);
            FM_HIP(hipMemcpy(tmp.data(), M.val32.p, M.nnz * 4, hipMemcpyDeviceToHost));
            std::copy(tmp.begin(), tmp.end(), K_vals_out);
        }
    }
    ctx->cur = &ctx->tangent;
    return FM_OK;
}

int fm_assemble_residual_only(fm_ctx* ctx, const double* u, double* f_int_out) {
    if (!ctx || !f_int_out) return fail(ctx, FM_EINVAL, "fm_assemble_residual_only: bad arguments");
    FM_TRY(ensure_ready(ctx));
    const size_t n = 3ull * ctx->nn;
    const double* du = nullptr;
    if (u) {
        FM_HIP(hipMemcpy(ctx->hv.p, u, n * 8, hipMemcpyHostToDevice));
        du = ctx->hv.p;
    }
    FM_TRY(assemble(ctx, du, false, PREC64, false, nullptr, FM_K_RESIDUAL));
    FM_TRY(sync(ctx));
    FM_HIP(hipMemcpy(f_int_out, ctx->f.p, n * 8, hipMemcpyDeviceToHost));
    return FM_OK;
}

int fm_residual(fm_ctx* ctx, const double* u, double load, double* R_out, double* norm_out) {
    if (!ctx) return FM_EINVAL;
    FM_TRY(ensure_ready(ctx));
    const size_t n = 3ull * ctx->nn;
    const double* du = nullptr;
    if (u) {
        FM_HIP(hipMemcpy(ctx->hv.p, u, n * 8, hipMemcpyHostToDevice));
        du = ctx->hv.p;
    }
    double nr = 0.0;
    FM_TRY(residual(ctx, du, load, &nr));
    if (R_out) FM_HIP(hipMemcpy(R_out, ctx->R.p, n * 8, hipMemcpyDeviceToHost));
    if (norm_out) *norm_out = nr;
    return FM_OK;
}

int fm_residual_noalloc(fm_ctx* ctx, const double* u, double load, double* R_out, double* norm_out) {
    if (!ctx) return FM_EINVAL;
    const uint64_t n = 3ull * ctx->nn;
    if (n > 256ull * kNoAllocPartials)
        return fail(ctx, FM_EINVAL,
                    "fm_residual_noalloc: " + std::to_string(n) + " DOFs exceed its scratch limit of " +
                        std::to_string(256ull * kNoAllocPartials) + " (use fm_residual)");
    // only buffers the set-up already made: fm_residual or fm_assemble makes them
    if (!ctx->nn || !ctx->have_pattern || !ctx->have_colors || !ctx->have_material || ctx->f.n != n ||
        ctx->R.n != n || ctx->hv.n != n)
        return fail(ctx, FM_ESTATE, "fm_residual_noalloc: the context is not set up (call fm_residual once first)");
    const NoAllocScratch sc = noalloc_scratch(ctx->device);
    if (!sc.partials) return fail(ctx, FM_EHIP, "fm_residual_noalloc: no scratch address for this device");
    FM_HIP(hipSetDevice(ctx->device));
    hipStream_t s = ctx->stream;
    const double* du = nullptr;
    if (u) {
        FM_HIP(hipMemcpyAsync(ctx->hv.p, u, n * 8, hipMemcpyHostToDevice, s));
        du = ctx->hv.p;
    }
    // the reduce state starts from zero on every call (the array keeps the last call's contents)
    launch_fill_u8(s, reinterpret_cast<uint8_t*>(sc.state), sizeof(PcgState), 0);
    FM_TRY(assemble(ctx, du, false, PREC64, false, nullptr, -1));
    const uint32_t np = launch_residual_finish(s, (uint32_t)n, load, ctx->have_fext ? ctx->f_ext.p : nullptr,
                                               ctx->f.p, ctx->dir_dofs.empty() ? nullptr : ctx->isdir.p, ctx->R.p,
                                               sc.partials);
    launch_reduce(s, sc.partials, np, sc.state, 0, 0);
    FM_HIP(hipMemcpyAsync(ctx->h_st, sc.state, sizeof(PcgState), hipMemcpyDeviceToHost, s));
    FM_TRY(check_launch(ctx, "residual_noalloc"));
    FM_HIP(hipStreamSynchronize(s));
    if (norm_out) *norm_out = std::sqrt(ctx->h_st->slot[0]);
    if (R_out) FM_HIP(hipMemcpy(R_out, ctx->R.p, n * 8, hipMemcpyDeviceToHost));
    return FM_OK;
}

int fm_set_matrix(fm_ctx* ctx, uint32_t n, const uint32_t* row_ptr, const uint32_t* col_idx, const double* vals) {
    if (!ctx || !n || !row_ptr || !col_idx || !vals) return fail(ctx, FM_EINVAL, "fm_set_matrix: bad arguments");
    FM_HIP(hipSetDevice(ctx->device));
    Matrix& M = ctx->ext;
    const uint64_t nnz = row_ptr[n];
    // the same pattern as the last call (same rows, same columns in the same
    // storage order): keep the sort, the schedules and the device pattern,
    // upload the values only
    bool same = M.ready && M.n == n && M.nnz == nnz && M.perm.size() == nnz &&
                std::equal(row_ptr, row_ptr + n + 1, M.h_row_ptr.begin());
    if (same) {
        int diff = 0;
#pragma omp parallel for schedule(static) reduction(| : diff)
        for (int64_t k = 0; k < (int64_t)nnz; ++k) diff |= col_idx[M.perm[k]] != M.h_col_idx[k];
        same = !diff;
    }
    if (same) {
        std::vector<double> sv(nnz);
        std::vector<float> sf(nnz);
#pragma omp parallel for schedule(static)
        for (int64_t k = 0; k < (int64_t)nnz; ++k) {
            sv[k] = vals[M.perm[k]];
            sf[k] = (float)sv[k];
        }
        FM_HIP(hipMemcpy(M.val64.p, sv.data(), nnz * 8, hipMemcpyHostToDevice));
        FM_HIP(hipMemcpy(M.val32.p, sf.data(), nnz * 4, hipMemcpyHostToDevice));
        M.val_valid[0] = M.val_valid[1] = true;
        M.lu_valid[0] = M.lu_valid[1] = false;
        M.pk_valid[0] = M.pk_valid[1] = false;
        ctx->cur = &M;
        ++ctx->ext_value_uploads;
        return FM_OK;
    }
    M.ready = false;
    M.n = n;
    M.nnz = nnz;
    M.h_row_ptr.assign(row_ptr, row_ptr + n + 1);
    M.h_col_idx.resize(nnz);
    M.perm.resize(nnz);
    std::vector<double> sv(nnz);
    for (uint32_t r = 0; r < n; ++r) {
        if (row_ptr[r + 1] < row_ptr[r]) return fail(ctx, FM_EINVAL, "row_ptr not monotone");
        const uint32_t b = row_ptr[r], e = row_ptr[r + 1];
        std::iota(M.perm.begin() + b, M.perm.begin() + e, b);
        std::sort(M.perm.begin() + b, M.perm.begin() + e, [&](uint32_t i, uint32_t j) { return col_idx[i] < col_idx[j]; });
        for (uint32_t k = b; k < e; ++k) {
            M.h_col_idx[k] = col_idx[M.perm[k]];
            if (M.h_col_idx[k] >= n) return fail(ctx, FM_EINVAL, "column index out of range");
            if (k > b && M.h_col_idx[k] == M.h_col_idx[k - 1]) return fail(ctx, FM_EINVAL, "duplicate column");
            sv[k] = vals[M.perm[k]];
        }
    }
    const int B = detect_block(M.h_row_ptr, M.h_col_idx);
    std::string err;
    if (!build_schedule(M.h_row_ptr, M.h_col_idx, B, M.sched, ctx->stream, err)) return fail(ctx, FM_EINVAL, err);
    FM_TRY(chain_setup(ctx, M, M.h_row_ptr, M.h_col_idx));
    FM_TRY(node_cols_setup(ctx, M, M.h_row_ptr, M.h_col_idx));
    FM_HIP(M.row_ptr.alloc(n + 1));
    FM_HIP(M.col_idx.alloc(nnz));
    FM_HIP(M.val64.alloc(nnz + kValPad));
    FM_HIP(M.val32.alloc(nnz + kValPad));
    FM_HIP(hipMemcpy(M.row_ptr.p, M.h_row_ptr.data(), (n + 1) * 4ull, hipMemcpyHostToDevice));
    FM_HIP(hipMemcpy(M.col_idx.p, M.h_col_idx.data(), nnz * 4, hipMemcpyHostToDevice));
    FM_HIP(hipMemcpy(M.val64.p, sv.data(), nnz * 8, hipMemcpyHostToDevice));
    std::vector<float> sf(sv.begin(), sv.end());
    FM_HIP(hipMemcpy(M.val32.p, sf.data(), nnz * 4, hipMemcpyHostToDevice));
    M.val_valid[0] = M.val_valid[1] = true;
    M.lu_valid[0] = M.lu_valid[1] = false;
    M.ready = true;
    ctx->cur = &M;
    return FM_OK;
}

int fm_use_tangent(fm_ctx* ctx, int eliminate) {
    if (!ctx) return FM_EINVAL;
    FM_TRY(ensure_pattern(ctx));
    Matrix& M = ctx->tangent;
    if (!M.val_valid[0] && !M.val_valid[1]) return fail(ctx, FM_ESTATE, "no assembled tangent (fm_assemble)");
    if (eliminate && !ctx->tangent_elim && !ctx->dir_dofs.empty()) {
        for (int p = 0; p < 2; ++p)
            if (M.val_valid[p])
                launch_eliminate(ctx->stream, p, M.n, M.row_ptr.p, M.col_idx.p, ctx->isdir.p,
                                 p == 0 ? (void*)M.val64.p : (void*)M.val32.p);
        M.lu_valid[0] = M.lu_valid[1] = false;
        ctx->tangent_elim = true;
        FM_TRY(sync(ctx));
    }
    ctx->cur = &M;
    return FM_OK;
}

static int matrix_prec(fm_ctx* ctx, int prec, Matrix** A, int* p) {
    if (!ctx->cur || !ctx->cur->ready) return fail(ctx, FM_ESTATE, "no matrix (fm_set_matrix / fm_use_tangent)");
    if (prec != FM_FP64 && prec != FM_MIXED) return fail(ctx, FM_EINVAL, "bad precision");
    *A = ctx->cur;
    *p = prec == FM_FP64 ? PREC64 : PREC32;
    if (!(*A)->val_valid[*p]) {
        // derive the other precision from the one present
        Matrix& M = **A;
        if (*p == PREC32 && M.val_valid[0]) {
            FM_TRY(values_to_fp32(ctx, M));
        } else {
            return fail(ctx, FM_ESTATE, "matrix values not available in FP64 (assembled in mixed precision)");
        }
    }
    return FM_OK;
}

static void scatter_perm(const Matrix& M, const double* sorted, double* out) {
    if (M.perm.empty()) {
        std::copy(sorted, sorted + M.nnz, out);
        return;
    }
    for (uint64_t k = 0; k < M.nnz; ++k) out[M.perm[k]] = sorted[k];
}

int fm_spmv(fm_ctx* ctx, int prec, const double* x, double* y) {
    if (!ctx || !x || !y) return fail(ctx, FM_EINVAL, "fm_spmv: bad arguments");
    Matrix* A;
    int p;
    FM_TRY(matrix_prec(ctx, prec, &A, &p));
    FM_TRY(pcg_vectors(ctx, p, A->n));
    if (p == PREC64) {
        FM_HIP(hipMemcpy(ctx->p64.p, x, A->n * 8ull, hipMemcpyHostToDevice));
    } else {
        std::vector<float> t(x, x + A->n);
        FM_HIP(hipMemcpy(ctx->p32.p, t.data(), A->n * 4ull, hipMemcpyHostToDevice));
    }
    {
        ProfScope ps(ctx, FM_K_SPMV);
        launch_spmv(ctx->stream, p, *A, vb(ctx, p, VP), vb(ctx, p, VAP), false, nullptr, nullptr, nullptr, nullptr);
    }
    FM_TRY(check_launch(ctx, "spmv"));
    FM_TRY(sync(ctx));
    if (p == PREC64) {
        FM_HIP(hipMemcpy(y, ctx->Ap64.p, A->n * 8ull, hipMemcpyDeviceToHost));
    } else {
        std::vector<float> t(A->n);
        FM_HIP(hipMemcpy(t.data(), ctx->Ap32.p, A->n * 4ull, hipMemcpyDeviceToHost));
        std::copy(t.begin(), t.end(), y);
    }
    return FM_OK;
}

int fm_ilu0_factor(fm_ctx* ctx, int prec, double* LU_out) {
    if (!ctx) return FM_EINVAL;
    Matrix* A;
    int p;
    FM_TRY(matrix_prec(ctx, prec, &A, &p));
    A->lu_valid[p] = false;
    FM_TRY(factorize(ctx, *A, p));
    FM_TRY(sync(ctx));
    if (LU_out) {
        std::vector<double> t(A->nnz);
        if (p == PREC64) {
            FM_HIP(hipMemcpy(t.data(), A->lu64.p, A->nnz * 8, hipMemcpyDeviceToHost));
        } else {
            std::vector<float> f(A->nnz);
            FM_HIP(hipMemcpy(f.data(), A->lu32.p, A->nnz * 4, hipMemcpyDeviceToHost));
            std::copy(f.begin(), f.end(), t.begin());
        }
        scatter_perm(*A, t.data(), LU_out);
    }
    return FM_OK;
}

int fm_ilu0_apply(fm_ctx* ctx, int prec, const double* r, double* z) {
    if (!ctx || !r || !z) return fail(ctx, FM_EINVAL, "fm_ilu0_apply: bad arguments");
    Matrix* A;
    int p;
    FM_TRY(matrix_prec(ctx, prec, &A, &p));
    FM_TRY(factorize(ctx, *A, p));
    FM_TRY(pcg_vectors(ctx, p, A->n));
    if (p == PREC64) {
        FM_HIP(hipMemcpy(ctx->r64.p, r, A->n * 8ull, hipMemcpyHostToDevice));
    } else {
        std::vector<float> t(r, r + A->n);
        FM_HIP(hipMemcpy(ctx->r32.p, t.data(), A->n * 4ull, hipMemcpyHostToDevice));
    }
    PcgState init{};
    *ctx->h_st = init;
    FM_HIP(hipMemcpyAsync(ctx->st.p, ctx->h_st, sizeof(PcgState), hipMemcpyHostToDevice, ctx->stream));
    FM_TRY(ensure_partials(ctx, A->sched.q_bwd + 64));
    precond_apply(ctx, *A, p, vb(ctx, p, VR), vb(ctx, p, VZ), vb(ctx, p, VY), nullptr);
    FM_TRY(check_launch(ctx, "ilu apply"));
    FM_TRY(sync(ctx));
    FM_TRY(check_device_error(ctx));
    if (p == PREC64) {
        FM_HIP(hipMemcpy(z, ctx->y64.p, A->n * 8ull, hipMemcpyDeviceToHost));
    } else {
        std::vector<float> t(A->n);
        FM_HIP(hipMemcpy(t.data(), ctx->y32.p, A->n * 4ull, hipMemcpyDeviceToHost));
        std::copy(t.begin(), t.end(), z);
    }
    return FM_OK;
}

int fm_pcg(fm_ctx* ctx, int prec, const double* b, double* x, double tol, int max_it, int* iters, double* relres) {
    if (!ctx || !b || !x || max_it < 0) return fail(ctx, FM_EINVAL, "fm_pcg: bad arguments");
    Matrix* A;
    int p;
    if (prec == FM_MIXED_PC) {
        if (!ctx->cur || !ctx->cur->ready) return fail(ctx, FM_ESTATE, "no matrix (fm_set_matrix / fm_use_tangent)");
        A = ctx->cur;
        if (!A->val_valid[PREC64]) return fail(ctx, FM_ESTATE, "FP64 matrix values needed for FM_MIXED_PC");
        A->lu_valid[PREC64] = A->lu_valid[PREC32] = false;  // pcg_solve_ilu0 factorizes on every call
        FM_TRY(pcg_vectors(ctx, PREC64, A->n));
        const size_t n = A->n;
        FM_HIP(hipMemcpy(ctx->b64.p, b, n * 8, hipMemcpyHostToDevice));
        FM_HIP(hipMemcpy(ctx->x64.p, x, n * 8, hipMemcpyHostToDevice));
        PcgResult pr{};
        FM_TRY(pcg_run_pc(ctx, *A, tol, max_it, &pr));
        FM_HIP(hipMemcpy(x, ctx->x64.p, n * 8, hipMemcpyDeviceToHost));
        if (iters) *iters = pr.iters;
        if (relres) *relres = pr.relres;
        return FM_OK;
    }
    FM_TRY(matrix_prec(ctx, prec, &A, &p));
    A->lu_valid[p] = false;  // pcg_solve_ilu0 factorizes on every call
    FM_TRY(pcg_vectors(ctx, p, A->n));
    const size_t n = A->n;
    if (p == PREC64) {
        FM_HIP(hipMemcpy(ctx->b64.p, b, n * 8, hipMemcpyHostToDevice));
        FM_HIP(hipMemcpy(ctx->x64.p, x, n * 8, hipMemcpyHostToDevice));
    } else {
        std::vector<float> tb(b, b + n), tx(x, x + n);
        FM_HIP(hipMemcpy(ctx->b32.p, tb.data(), n * 4, hipMemcpyHostToDevice));
        FM_HIP(hipMemcpy(ctx->x32.p, tx.data(), n * 4, hipMemcpyHostToDevice));
    }
    PcgResult pr{};
    FM_TRY(pcg_run(ctx, *A, p, tol, max_it, &pr));
    if (p == PREC64) {
        FM_HIP(hipMemcpy(x, ctx->x64.p, n * 8, hipMemcpyDeviceToHost));
    } else {
        std::vector<float> t(n);
        FM_HIP(hipMemcpy(t.data(), ctx->x32.p, n * 4, hipMemcpyDeviceToHost));
        std::copy(t.begin(), t.end(), x);
    }
    if (iters) *iters = pr.iters;
    if (relres) *relres = pr.relres;
    return FM_OK;
}

void fm_newton_cfg_default(fm_newton_cfg* c) {
    if (!c) return;
    c->n_steps = 1;
    c->prec = FM_FP64;
    c->newton_rtol = 1e-8;
    c->max_newton = 30;
    c->pcg_tol = 1e-10;  // newton_tol * 0.01 (FEM/adaptive_fem_engine.hpp:2577-2578)
    c->pcg_max = 300;
    // forcing term of the inner solves: at config 3 (128^3 hexes, 20 load
    // steps) 3e-3 takes 3 Newton iterations per step and 40,409 FP32 PCG
    // iterations in all, 1e-2 takes 4 and 44,336, 3e-2 diverges into the
    // line search (19 per step)
    c->inner_tol = 3e-3;
    c->ls_max = 10;
    c->armijo_c = 1e-4;
    c->lift = 1;
    c->inner_max = 1000;
    c->ls_mode = 0;
    c->tol_absolute = 0.0;
    c->ls_alpha0 = 1.0;
    c->ls_beta = 0.5;
    c->reuse_tangent = 0;
    c->energy_tol = 0.0;
    c->record_energy = 0;
    c->inexact = 0;
    c->recycle = 0;
    c->inner_tol_max = 0.0;
    c->coarse_degree = 0;
    c->coarse_additive = 0;
    c->lag_precond = 0;
    c->extrap_order = 0;
    c->slope_start = 0;
    c->factor_steps = 0;
    c->keep_coarse = 0;
    c->ritz = 0;
}

int fm_set_u(fm_ctx* ctx, const double* u) {
    if (!ctx) return FM_EINVAL;
    if (!ctx->nn) return fail(ctx, FM_ESTATE, "no mesh");
    FM_TRY(ensure_vectors(ctx));
    if (!u) {
        // the unloaded configuration: u = 0 is the equilibrium at load factor 0
        launch_fill_f64(ctx->stream, ctx->u.p, 3ull * ctx->nn, 0.0);
        FM_TRY(sync(ctx));
        ctx->u_valid = true;
        ctx->sec_n = 1;
        ctx->slope_ok = false;
        ctx->gal_n[0] = ctx->gal_n[1] = 0;
        ctx->nritz[0] = ctx->nritz[1] = 0;
        ctx->sec_lf[0] = 0.0;
        return FM_OK;
    }
    ctx->sec_n = 0;
    ctx->slope_ok = false;
    ctx->gal_n[0] = ctx->gal_n[1] = 0;
    ctx->nritz[0] = ctx->nritz[1] = 0;
    FM_HIP(hipMemcpy(ctx->u.p, u, 3ull * ctx->nn * 8, hipMemcpyHostToDevice));
    ctx->u_valid = true;
    return FM_OK;
}

int fm_newton_history(fm_ctx* ctx, uint32_t cap, double* residual, double* step_size, double* energy,
                      uint32_t* len) {
    if (!ctx) return FM_EINVAL;
    const size_t n = ctx->hist_res.size();
    for (size_t i = 0; i < n && i < cap; ++i) {
        if (residual) residual[i] = ctx->hist_res[i];
        if (energy) energy[i] = ctx->hist_energy[i];
        if (step_size && i + 1 < n) step_size[i] = ctx->hist_alpha[i];
    }
    if (len) *len = (uint32_t)n;
    return FM_OK;
}

int fm_pcg_history(fm_ctx* ctx, uint32_t cap, double* alpha, double* beta, uint32_t* len) {
    if (!ctx) return FM_EINVAL;
    const uint32_t n = std::min(cap, ctx->pcg_hist_len);
    if (n && (!alpha || !beta)) return FM_EINVAL;
    if (n) {
        std::vector<double> h(2 * (size_t)n);
        FM_HIP(hipMemcpy(h.data(), ctx->pcg_hist.p, h.size() * 8, hipMemcpyDeviceToHost));
        for (uint32_t i = 0; i < n; ++i) {
            alpha[i] = h[2 * i];
            beta[i] = h[2 * i + 1];
        }
    }
    if (len) *len = ctx->pcg_hist_len;
    return FM_OK;
}

int fm_coarse_moments(fm_ctx* ctx, int prec, int degree, const double* r, double* out) {
    if (!ctx || !r || !out) return FM_EINVAL;
    if (prec != FM_FP64 && prec != FM_MIXED) return fail(ctx, FM_EINVAL, "bad precision");
    if (degree < 1 || degree > kMaxDegree) return fail(ctx, FM_EINVAL, "degree must be 1..4");
    FM_TRY(ensure_ready(ctx));
    FM_TRY(ensure_vectors(ctx));
    hipStream_t s = ctx->stream;
    const int p = prec == FM_FP64 ? PREC64 : PREC32;
    const uint32_t n = 3 * ctx->nn;
    const int nmono = kMonos[degree], m = 3 * nmono;
    FM_TRY(coarse_coords(ctx));
    FM_TRY(pcg_vectors(ctx, p, n));
    FM_HIP(hipMemcpyAsync(ctx->hv.p, r, n * 8ull, hipMemcpyHostToDevice, s));
    void* rv = vb(ctx, p, VB);
    if (p == PREC64)
        FM_HIP(hipMemcpyAsync(rv, ctx->hv.p, n * 8ull, hipMemcpyDeviceToDevice, s));
    else
        launch_f64_to_f32(s, n, ctx->hv.p, 1.0, (float*)rv);
    const uint32_t nb = coarse_moment_blocks();
    FM_HIP(ctx->ad_part.alloc(std::max<size_t>(ctx->ad_part.n, (size_t)nb * m)));
    PcgState init{};
    *ctx->h_st = init;
    FM_HIP(hipMemcpyAsync(ctx->st.p, ctx->h_st, sizeof(PcgState), hipMemcpyHostToDevice, s));
    launch_coarse_moments(s, p, n, nmono, ctx->cs_xn.p, rv, ctx->ad_part.p, ctx->st.p);
    FM_TRY(check_launch(ctx, "coarse moments"));
    std::vector<double> part((size_t)nb * m);
    FM_HIP(hipMemcpyAsync(part.data(), ctx->ad_part.p, part.size() * 8, hipMemcpyDeviceToHost, s));
    FM_TRY(sync(ctx));
    for (int i = 0; i < m; ++i) {
        double v = 0.0;
        for (uint32_t b = 0; b < nb; ++b) v += part[(size_t)b * m + i];
        out[i] = v;
    }
    return FM_OK;
}

int fm_get_u(fm_ctx* ctx, double* u) {
    if (!ctx || !u) return FM_EINVAL;
    if (!ctx->u_valid) return fail(ctx, FM_ESTATE, "no device state");
    FM_HIP(hipMemcpy(u, ctx->u.p, 3ull * ctx->nn * 8, hipMemcpyDeviceToHost));
    return FM_OK;
}

// One load step of the Newton loop (same algorithm as the FP64 CPU oracle's
// orc_newton_step): prescribed values lf*dir_vals, R_0 at the previous state
// with them set, optional predictor solve with the prescribed increment,
// then Newton + Armijo until ||R|| <= rtol ||R_0||.
int fm_newton_step(fm_ctx* ctx, uint32_t step, double* u, const fm_newton_cfg* cfg_in, fm_newton_stats* stats) {
    if (!ctx) return FM_EINVAL;
    fm_newton_cfg cfg;
    if (cfg_in)
        cfg = *cfg_in;
    else
        fm_newton_cfg_default(&cfg);
    if (cfg.n_steps == 0 || step == 0 || step > cfg.n_steps) return fail(ctx, FM_EINVAL, "bad load step");
    return fm_newton_solve_to(ctx, (double)step / (double)cfg.n_steps, u, &cfg, stats);
}

// the predictor solve that gives the path's slope at the unloaded state
constexpr double kSlopeTol = 1e-5;

// FM_TRACE_SOLVES=1: one line per linear solve of the Newton loop on stderr
// (load factor, Newton iteration or -1 for the predictor solve, forcing
// term, PCG iterations, lagged preconditioner)
static void trace_solve(double lf, int it, double eta, int its, bool lag) {
    if (trace_solves()) std::fprintf(stderr, "[fm] lf %.4f newton %d eta %.3g pcg %d%s\n", lf, it, eta, its, lag ? " lagged" : "");
}

int fm_newton_solve_to(fm_ctx* ctx, double load_factor, double* u, const fm_newton_cfg* cfg_in,
                       fm_newton_stats* stats) {
    if (!ctx) return FM_EINVAL;
    fm_newton_cfg cfg;
    if (cfg_in)
        cfg = *cfg_in;
    else
        fm_newton_cfg_default(&cfg);
    if (!(load_factor >= 0.0) || !std::isfinite(load_factor)) return fail(ctx, FM_EINVAL, "bad load factor");
    if (cfg.prec != FM_FP64 && cfg.prec != FM_MIXED && cfg.prec != FM_MIXED_PC)
        return fail(ctx, FM_EINVAL, "bad precision");
    if (!(cfg.ls_beta > 0.0 && cfg.ls_beta < 1.0) || !(cfg.ls_alpha0 > 0.0) || !(cfg.tol_absolute >= 0.0) ||
        !(cfg.energy_tol >= 0.0))
        return fail(ctx, FM_EINVAL, "bad line-search or tolerance setting");
    FM_TRY(ensure_ready(ctx));
    const double t0 = now_s();
    hipStream_t s = ctx->stream;
    const uint32_t n = 3 * ctx->nn;
    if (u) {
        FM_HIP(hipMemcpyAsync(ctx->u.p, u, n * 8ull, hipMemcpyHostToDevice, s));
        ctx->u_valid = true;
        ctx->sec_n = 0;  // a state from outside: no known history
        ctx->slope_ok = false;
        ctx->gal_n[0] = ctx->gal_n[1] = 0;
        ctx->nritz[0] = ctx->nritz[1] = 0;
    }
    if (!ctx->u_valid) {
        // u = 0 is the equilibrium at load factor 0 (no load, no prescribed values)
        launch_fill_f64(s, ctx->u.p, n, 0.0);
        ctx->u_valid = true;
        ctx->sec_n = 1;
        ctx->slope_ok = false;
        ctx->gal_n[0] = ctx->gal_n[1] = 0;
        ctx->nritz[0] = ctx->nritz[1] = 0;
        ctx->sec_lf[0] = 0.0;
    }
    // linear-solve mode and the precision the tangent is assembled in
    const int kprec = cfg.prec == FM_FP64 ? PREC64 : (cfg.prec == FM_MIXED ? PREC32 : PRECPC);
    const int aprec = kprec == PREC32 ? PREC32 : PREC64;
    // the FP64 path's inner solves: the reference's pcg_tol / pcg_max, or the
    // mixed path's forcing term when inexact
    const bool inexact = kprec == PREC64 && cfg.inexact;
    const double lin_tol = kprec == PREC64 && !inexact ? cfg.pcg_tol : cfg.inner_tol;
    // the options of the inner solves: target-aware forcing, Galerkin start
    // (the forcing-term solves: mixed, FP32-preconditioned, inexact FP64)
    const bool inner = kprec != PREC64 || inexact;
    const bool adapt = inner && cfg.inner_tol_max > lin_tol;
    const bool recycle = inner && cfg.lift == 2 && cfg.recycle > 0;
    const int coarse = inner ? std::max(0, std::min(cfg.coarse_degree, kMaxDegree)) : 0;
    const bool additive = coarse > 0 && cfg.coarse_additive != 0;
    // Ritz vectors of each solve for the next step's first solve (with the coarse level)
    const int ritz = additive && (kprec == PREC32 || inexact) ? std::max(0, std::min(cfg.ritz, kRitz)) : 0;
    const int lin_cap = kprec == PREC64 && !inexact ? cfg.pcg_max : cfg.inner_max;
    const double lf = load_factor;
    fm_newton_stats S{};
    const bool want_energy = cfg.record_energy || cfg.energy_tol > 0.0;
    ctx->hist_res.clear();
    ctx->hist_alpha.clear();
    ctx->hist_energy.clear();
    // one history entry: ||R|| of `state`, the step that reached it, its energy
    auto record = [&](double rnorm, double alpha, const double* state) -> int {
        if (!ctx->hist_res.empty()) ctx->hist_alpha.push_back(alpha);
        ctx->hist_res.push_back(rnorm);
        double pi = std::nan("");
        if (want_energy) FM_TRY(energy(ctx, state, lf, nullptr, &pi, nullptr));
        ctx->hist_energy.push_back(pi);
        return FM_OK;
    };
    const uint32_t ndir = (uint32_t)ctx->dir_dofs.size();
    bool any_inc = false;
    for (double v : ctx->dir_vals) any_inc |= v != 0.0;
    launch_fill_f64(s, ctx->dD.p, n, 0.0);
    launch_dir_increment(s, ndir, ctx->d_dir_dofs.p, ctx->d_dir_vals.p, lf, ctx->u.p, ctx->dD.p);
    launch_axpy_to(s, n, ctx->u.p, 1.0, ctx->dD.p, ctx->ut.p);
    FM_TRY(residual(ctx, ctx->ut.p, lf, &S.r0));
    S.residuals++;
    FM_TRY(record(S.r0, 1.0, ctx->ut.p));
    // lift 2: the step starts from the Lagrange extrapolation to lf of the
    // previous converged states of this load path (of degree extrap_order
    // through that many + 1 states, default quadratic through three; lower
    // degrees while fewer are known), with the prescribed values set exactly
    // (no linear solve); the predictor solve below where there is no such
    // history; a start state that already passes the absolute test takes no
    // predictor
    const bool at_rest = S.r0 < cfg.tol_absolute;
    const int max_order = cfg.extrap_order > 0 ? std::min(cfg.extrap_order, fm_ctx::kExtrapMax) : 2;
    int order = 0;
    bool hermite = false;
    if (cfg.lift == 2 && !at_rest) {
        // the highest degree whose states are known at distinct load factors
        for (int o = std::min(max_order, ctx->sec_n - 1); o >= 1 && order == 0; --o) {
            bool distinct = true;
            for (int a = 0; a <= o; ++a)
                for (int b = a + 1; b <= o; ++b) distinct &= ctx->sec_lf[a] != ctx->sec_lf[b];
            if (distinct) order = o;
        }
        // slope_start: while the states alone fall short of the degree, all
        // m known states and the slope at one of them give degree m
        const int m = ctx->sec_n;
        if (cfg.slope_start && ctx->slope_ok && order < max_order && m >= 1 && m <= max_order) {
            bool has = false, distinct = true;
            for (int a = 0; a < m; ++a) {
                has |= ctx->sec_lf[a] == ctx->slope_lf;
                for (int b = a + 1; b < m; ++b) distinct &= ctx->sec_lf[a] != ctx->sec_lf[b];
            }
            if (has && distinct && !(m == 1 && ctx->sec_lf[0] == lf)) {
                order = m;
                hermite = true;
            }
        }
    }
    const bool extrapolate = order > 0;
    if (cfg.lift == 2) {
        FM_HIP(ctx->sec_start.alloc(n));
        for (int i = 0; i < std::max(max_order, 2); ++i) FM_HIP(ctx->sec_old[i].alloc(n));
        FM_HIP(hipMemcpyAsync(ctx->sec_start.p, ctx->u.p, n * 8ull, hipMemcpyDeviceToDevice, s));
    }
    if (extrapolate) {
        Lincomb lc{};
        lc.k = order + 1;
        if (!hermite) {
            // Lagrange weights of u (at sec_lf[0]) and sec_old[a - 1] (at sec_lf[a]) at lf
            for (int a = 0; a <= order; ++a) {
                double v = 1.0;
                for (int b = 0; b <= order; ++b)
                    if (b != a) v *= (lf - ctx->sec_lf[b]) / (ctx->sec_lf[a] - ctx->sec_lf[b]);
                lc.w[a] = v;
                lc.x[a] = a == 0 ? ctx->u.p : ctx->sec_old[a - 1].p;
            }
        } else {
            // Hermite weights: the m = order states and the slope at slope_lf
            // reproduce every polynomial of degree <= m, sum_a w_a l_a^k +
            // w_d k slope_lf^(k-1) = lf^k for k = 0..m (Gaussian elimination
            // with partial pivoting; the last unknown is the slope's weight)
            const int m = order, N = m + 1;
            double M[5][6];
            for (int k = 0; k < N; ++k) {
                for (int a = 0; a < m; ++a) M[k][a] = std::pow(ctx->sec_lf[a], k);
                M[k][m] = k == 0 ? 0.0 : k * std::pow(ctx->slope_lf, k - 1);
                M[k][N] = std::pow(lf, k);
            }
            for (int c = 0; c < N; ++c) {
                int piv = c;
                for (int r = c + 1; r < N; ++r)
                    if (std::fabs(M[r][c]) > std::fabs(M[piv][c])) piv = r;
                for (int j = 0; j <= N; ++j) std::swap(M[c][j], M[piv][j]);
                for (int r = 0; r < N; ++r)
                    if (r != c) {
                        const double f = M[r][c] / M[c][c];
                        for (int j = c; j <= N; ++j) M[r][j] -= f * M[c][j];
                    }
            }
            for (int a = 0; a < N; ++a) lc.w[a] = M[a][N] / M[a][a];
            for (int a = 0; a < m; ++a) lc.x[a] = a == 0 ? ctx->u.p : ctx->sec_old[a - 1].p;
            lc.x[m] = ctx->slope.p;
        }
        launch_lincomb(s, n, lc, ctx->ut.p);
        launch_fill_f64(s, ctx->dD.p, n, 0.0);
        launch_dir_increment(s, ndir, ctx->d_dir_dofs.p, ctx->d_dir_vals.p, lf, ctx->ut.p, ctx->dD.p);
        launch_axpy_to(s, n, ctx->ut.p, 1.0, ctx->dD.p, ctx->u.p);
    } else if (cfg.lift && any_inc && !at_rest) {
        FM_TRY(assemble(ctx, ctx->u.p, true, aprec, true, ctx->dD.p, FM_K_ASSEMBLE));
        S.assemblies++;
        FM_TRY(ensure_partials(ctx, n / 256 + 1));
        launch_residual_finish(s, n, lf, ctx->have_fext ? ctx->f_ext.p : nullptr, ctx->f.p, ctx->isdir.p, ctx->R.p,
                               ctx->partials.p);
        launch_sub_masked(s, n, ctx->R.p, ctx->lift.p, ctx->isdir.p, ctx->hv.p);
        int its = 0;
        // from the unloaded state the predictor solve's increment over lf is
        // the path's slope there (the linearised problem is linear in the
        // load factor): with slope_start it is solved to kSlopeTol and kept
        const bool want_slope = cfg.lift == 2 && cfg.slope_start && ctx->sec_n == 1 && ctx->sec_lf[0] == 0.0 && lf > 0.0;
        const double ptol = want_slope ? std::min(lin_tol, kSlopeTol) : lin_tol;
        FM_TRY(linear_solve(ctx, kprec, ctx->hv.p, ctx->du.p, ptol, lin_cap, &its, inexact, &S.pcg_cap_hits, 0,
                            coarse, false, additive, false, false, ritz, 0));
        S.pcg_iters += its;
        S.factorizations++;
        trace_solve(lf, -1, ptol, its, false);
        launch_apply_predictor(s, n, ctx->du.p, ctx->dD.p, ctx->isdir.p, ctx->u.p);
        if (want_slope) {
            FM_HIP(ctx->slope.alloc(n));
            Lincomb lc{};
            lc.k = 2;
            lc.x[0] = ctx->u.p;
            lc.w[0] = 1.0 / lf;
            lc.x[1] = ctx->sec_start.p;
            lc.w[1] = -1.0 / lf;
            launch_lincomb(s, n, lc, ctx->slope.p);
            ctx->slope_lf = 0.0;
            ctx->slope_ok = true;
        }
    } else {
        launch_axpy_to(s, n, ctx->u.p, 1.0, ctx->dD.p, ctx->u.p);
    }
    const bool predicted = extrapolate || (cfg.lift && any_inc && !at_rest);
    double rn = 0.0;
    FM_TRY(residual(ctx, ctx->u.p, lf, &rn));
    S.residuals++;
    if (predicted) FM_TRY(record(rn, 1.0, ctx->u.p));
    const double target = cfg.newton_rtol * S.r0;
    // NonlinearNewtonSolver's tests (FEM/nonlinear_solver.hpp:127, :215-227):
    // relative, absolute, and the energy change of the last iteration
    auto done = [&](double r, bool energy_stop) {
        return r <= target || r < cfg.tol_absolute || energy_stop;
    };
    int it = 0;
    double rn_prev = 0.0;
    bool energy_stop = false;
    while (!done(rn, energy_stop) && it < cfg.max_newton) {
        // from the quadratic extrapolation, later iterations keep the step's
        // first tangent and its factor while each iteration still cuts ||R||
        // by 100x or more (modified Newton; the FP64 residual decides
        // convergence, and a mixed iteration's reduction is set by the FP32
        // tangent rather than by this lag)
        // (reuse_tangent keeps it for every iteration of the step)
        // (not with the target-aware forcing: its second solve is the
        // expensive one, and on the first tangent it takes more PCG
        // iterations than the assembly and factorization it would save)
        const bool keep = it > 0 && (cfg.reuse_tangent || (order == 2 && !adapt && rn <= 1e-2 * rn_prev));
        // a lagged preconditioner: the new tangent with the first iteration's
        // factor (and coarse operator), which the assembly leaves in place
        const bool lag_on = it > 0 && !keep && cfg.lag_precond && inner;
        bool lag = false, reuse_poly = false;
        if (!keep) {
            const int fp = kprec == PREC32 ? PREC32 : PREC64;
            lag = lag_on && ctx->tangent.lu_valid[fp];
            // factor_steps: a step that starts from an extrapolation of degree
            // >= 2 keeps the factor of an earlier step while it has served
            // fewer than factor_steps steps
            const bool keep_factor = it == 0 && cfg.lag_precond && cfg.factor_steps > 1 && order >= 2 &&
                                     inner && ctx->tangent.lu_valid[fp] &&
                                     ctx->fact_age < cfg.factor_steps;
            FM_TRY(assemble(ctx, ctx->u.p, true, aprec, true, nullptr, FM_K_ASSEMBLE));
            S.assemblies++;
            if (lag || keep_factor) ctx->tangent.lu_valid[fp] = true;
            if (it == 0) ctx->fact_age = keep_factor ? ctx->fact_age + 1 : 1;
            lag |= keep_factor;
            reuse_poly = keep_factor;
        }
        int its = 0;
        // R of the current u is in ctx->R (last residual evaluation)
        FM_HIP(hipMemcpyAsync(ctx->hv.p, ctx->R.p, n * 8ull, hipMemcpyDeviceToDevice, s));
        // target-aware forcing term: a solve that needs only a small reduction
        // of ||R|| to reach the target stops early (it achieves about half the
        // forcing term's reduction, so 0.5 target / ||R|| leaves a margin of ~4)
        double eta = lin_tol;
        if (adapt) eta = std::min(cfg.inner_tol_max, std::max(lin_tol, kForcingShare * target / rn));
        // the first solve of a step starts from the Galerkin projection onto
        // the latest 2 recycle Newton corrections (both corrections of each of
        // the previous `recycle` steps in a steady run), and every correction
        // joins them
        const int start_from = recycle && it == 0 ? std::min(2 * cfg.recycle, kCorr) : 0;
        // the Ritz vectors of the previous solve join that start only beside
        // recycled corrections: at a path's first step they come from the
        // predictor solve at the unloaded state, and there they cost the
        // 6.4M-DOF cube a Newton iteration (3 solves, 502 PCG iterations
        // against 2 and 391)
        const int ritz_use = it == 0 && ctx->gal_n[kprec == PREC32 ? 1 : 0] > 0 ? ritz : 0;
        FM_TRY(linear_solve(ctx, kprec, ctx->hv.p, ctx->du.p, eta, lin_cap, &its, inexact, &S.pcg_cap_hits,
                            start_from, coarse, recycle, additive, lag, reuse_poly && cfg.keep_coarse, ritz,
                            ritz_use));
        S.pcg_iters += its;
        if (!keep && !lag) S.factorizations++;
        trace_solve(lf, it, eta, its, lag);
        // armijo_line_search_true (FEM/adaptive_fem_engine.hpp:933-985); with
        // ls_mode 1 a trial that fails the residual-norm test is also accepted
        // when the energy decreases by c alpha R.du (the residual norm may grow
        // on the way into the quadratic basin while the energy decreases)
        double alpha = cfg.ls_alpha0, rt = 0.0, pi0 = 0.0, slope = 0.0;
        bool accepted = false;
        if (cfg.ls_mode == 1) FM_TRY(energy(ctx, ctx->u.p, lf, ctx->du.p, &pi0, &slope));
        for (int bt = 0; bt < cfg.ls_max; ++bt) {
            launch_axpy_to(s, n, ctx->u.p, alpha, ctx->du.p, ctx->ut.p);
            FM_TRY(residual(ctx, ctx->ut.p, lf, &rt));
            S.residuals++;
            if (rt <= (1.0 - cfg.armijo_c * alpha) * rn + 1e-14) {
                accepted = true;
                break;
            }
            if (cfg.ls_mode == 1 && slope > 0.0) {
                double pit = 0.0;
                FM_TRY(energy(ctx, ctx->ut.p, lf, nullptr, &pit, nullptr));
                if (pit <= pi0 - cfg.armijo_c * alpha * slope) {
                    accepted = true;
                    break;
                }
            }
            alpha *= cfg.ls_beta;
        }
        if (!accepted) {
            launch_axpy_to(s, n, ctx->u.p, alpha, ctx->du.p, ctx->ut.p);
            FM_TRY(residual(ctx, ctx->ut.p, lf, &rt));
            S.residuals++;
        }
        const double e_prev = ctx->hist_energy.back();
        FM_TRY(record(rt, alpha, ctx->ut.p));
        if (cfg.energy_tol > 0.0) energy_stop = std::fabs(ctx->hist_energy.back() - e_prev) < cfg.energy_tol;
        std::swap(ctx->u.p, ctx->ut.p);
        rn_prev = rn;
        rn = rt;
        ++it;
    }
    FM_TRY(sync(ctx));
    S.newton_iters = it;
    S.converged = done(rn, energy_stop);
    if (cfg.lift == 2 && S.converged) {
        // the start state (the previous converged one) becomes sec_old[0], each
        // older one moves down, this step's converged state is the current
        // one; the states kept are those the degree uses (at least three)
        const int keep_n = std::max(max_order, 2) + 1;
        for (int i = keep_n - 2; i >= 1; --i) {
            std::swap(ctx->sec_old[i].p, ctx->sec_old[i - 1].p);
            std::swap(ctx->sec_old[i].n, ctx->sec_old[i - 1].n);
        }
        std::swap(ctx->sec_old[0].p, ctx->sec_start.p);
        std::swap(ctx->sec_old[0].n, ctx->sec_start.n);
        ctx->sec_n = ctx->sec_n >= 1 ? std::min(ctx->sec_n + 1, keep_n) : 1;
        for (int i = keep_n - 1; i >= 1; --i) ctx->sec_lf[i] = ctx->sec_lf[i - 1];
        ctx->sec_lf[0] = lf;
    } else {
        ctx->sec_n = 0;
        ctx->slope_ok = false;
        ctx->gal_n[0] = ctx->gal_n[1] = 0;
        ctx->nritz[0] = ctx->nritz[1] = 0;
    }
    S.rel_res = S.r0 > 0 ? rn / S.r0 : 0.0;
    S.seconds = now_s() - t0;
    if (stats) *stats = S;
    if (u) FM_HIP(hipMemcpy(u, ctx->u.p, n * 8ull, hipMemcpyDeviceToHost));
    return S.converged ? FM_OK : fail(ctx, FM_ENOCONV, "Newton did not reach the relative tolerance");
}

int fm_adjoint(fm_ctx* ctx, const double* dLdu, double tol, int max_it, double* dL_dmu, double* dL_dlam,
               double* dL_dload, int* iters) {
    if (!ctx || !dLdu) return fail(ctx, FM_EINVAL, "fm_adjoint: bad arguments");
    FM_TRY(ensure_ready(ctx));
    if (!ctx->u_valid) return fail(ctx, FM_ESTATE, "no converged state (fm_newton_step / fm_set_u)");
    hipStream_t s = ctx->stream;
    const uint32_t n = 3 * ctx->nn;
    const uint32_t ndir = (uint32_t)ctx->dir_dofs.size();
    // uD: Dirichlet values at load factor 1 (zero elsewhere), through dD
    launch_fill_f64(s, ctx->hv.p, n, 0.0);
    launch_fill_f64(s, ctx->dD.p, n, 0.0);
    launch_dir_increment(s, ndir, ctx->d_dir_dofs.p, ctx->d_dir_vals.p, 1.0, ctx->hv.p, ctx->dD.p);
    FM_TRY(assemble(ctx, ctx->u.p, true, PREC64, true, ctx->dD.p, FM_K_ASSEMBLE));
    std::vector<double> g(dLdu, dLdu + n);
    for (uint32_t d : ctx->dir_dofs) g[d] = 0.0;
    FM_HIP(hipMemcpyAsync(ctx->R.p, g.data(), n * 8ull, hipMemcpyHostToDevice, s));
    FM_TRY(sync(ctx));
    int its = 0;
    FM_TRY(linear_solve(ctx, PREC64, ctx->R.p, ctx->du.p, tol, max_it, &its));
    if (iters) *iters = its;
    double* dmu = nullptr;
    double* dlam = nullptr;
    DevBuf<double> bm, bl;
    if (dL_dmu) {
        FM_HIP(bm.alloc(ctx->ne));
        dmu = bm.p;
    }
    if (dL_dlam) {
        FM_HIP(bl.alloc(ctx->ne));
        dlam = bl.p;
    }
    if (dmu || dlam) {
        launch_adjoint_elem(s, elem_args(ctx, ctx->u.p), ctx->eperm.p, ctx->du.p, dmu, dlam);
        FM_TRY(check_launch(ctx, "adjoint"));
    }
    FM_TRY(sync(ctx));
    if (dL_dmu) FM_HIP(hipMemcpy(dL_dmu, bm.p, ctx->ne * 8ull, hipMemcpyDeviceToHost));
    if (dL_dlam) FM_HIP(hipMemcpy(dL_dlam, bl.p, ctx->ne * 8ull, hipMemcpyDeviceToHost));
    if (dL_dload) {
        std::vector<double> lam(n), lift(n), fe(n, 0.0);
        FM_HIP(hipMemcpy(lam.data(), ctx->du.p, n * 8ull, hipMemcpyDeviceToHost));
        FM_HIP(hipMemcpy(lift.data(), ctx->lift.p, n * 8ull, hipMemcpyDeviceToHost));
        if (ctx->have_fext) FM_HIP(hipMemcpy(fe.data(), ctx->f_ext.p, n * 8ull, hipMemcpyDeviceToHost));
        std::vector<uint8_t> isd(n, 0);
        std::vector<double> uD(n, 0.0);
        for (uint32_t k = 0; k < ndir; ++k) {
            isd[ctx->dir_dofs[k]] = 1;
            uD[ctx->dir_dofs[k]] = ctx->dir_vals[k];
        }
        double acc = 0.0;
        for (uint32_t i = 0; i < n; ++i) acc += isd[i] ? dLdu[i] * uD[i] : lam[i] * (fe[i] - lift[i]);
        *dL_dload = acc;
    }
    return FM_OK;
}

int fm_rot_logexp_batch(fm_ctx* ctx, uint32_t n, const double* F, double alpha, double* R, double* w, double* Q) {
    if (!ctx || !F) return fail(ctx, FM_EINVAL, "fm_rot_logexp_batch: bad arguments");
    if (!n) return FM_OK;
    DevBuf<double> dF, dR, dw, dQ;
    FM_HIP(dF.alloc(9ull * n));
    FM_HIP(hipMemcpy(dF.p, F, 9ull * n * 8, hipMemcpyHostToDevice));
    if (R) FM_HIP(dR.alloc(9ull * n));
    if (w) FM_HIP(dw.alloc(3ull * n));
    if (Q) FM_HIP(dQ.alloc(9ull * n));
    launch_rot_batch(ctx->stream, n, dF.p, alpha, dR.p, dw.p, dQ.p);
    FM_TRY(check_launch(ctx, "rot_batch"));
    FM_TRY(sync(ctx));
    if (R) FM_HIP(hipMemcpy(R, dR.p, 9ull * n * 8, hipMemcpyDeviceToHost));
    if (w) FM_HIP(hipMemcpy(w, dw.p, 3ull * n * 8, hipMemcpyDeviceToHost));
    if (Q) FM_HIP(hipMemcpy(Q, dQ.p, 9ull * n * 8, hipMemcpyDeviceToHost));
    return FM_OK;
}

int fm_element_rotations(fm_ctx* ctx, double* w) {
    if (!ctx || !w) return FM_EINVAL;
    FM_TRY(ensure_ready(ctx));
    if (!ctx->u_valid) return fail(ctx, FM_ESTATE, "no device state");
    DevBuf<double> dw;
    FM_HIP(dw.alloc(3ull * ctx->ne));
    launch_elem_rot(ctx->stream, elem_args(ctx, ctx->u.p), ctx->eiperm.p, dw.p);
    FM_TRY(check_launch(ctx, "elem_rot"));
    FM_TRY(sync(ctx));
    FM_HIP(hipMemcpy(w, dw.p, 3ull * ctx->ne * 8, hipMemcpyDeviceToHost));
    return FM_OK;
}

int fm_element_fields(fm_ctx* ctx, const double* u, const fm_element_fields_out* out) {
    if (!ctx || !out) return fail(ctx, FM_EINVAL, "fm_element_fields: bad arguments");
    FM_TRY(ensure_ready(ctx));
    const uint32_t n = 3 * ctx->nn, ne = ctx->ne;
    const double* ud = ctx->u.p;
    DevBuf<double> du;
    if (u) {
        FM_HIP(du.alloc(n));
        FM_HIP(hipMemcpyAsync(du.p, u, n * 8ull, hipMemcpyHostToDevice, ctx->stream));
        ud = du.p;
    } else if (!ctx->u_valid) {
        return fail(ctx, FM_ESTATE, "no device state (pass u)");
    }
    DevBuf<double> dout;
    FM_HIP(dout.alloc((size_t)kFieldCount * ne));
    {
        ProfScope ps(ctx, FM_K_FIELDS);
        launch_element_fields(ctx->stream, elem_args(ctx, ud), ctx->eiperm.p, dout.p);
    }
    FM_TRY(check_launch(ctx, "element_fields"));
    FM_TRY(sync(ctx));
    // rows of the device result: det F, sigma (6), von Mises, pressure, s1, psi, stretch, F (9)
    const struct {
        double* dst;
        int row, rows;
    } parts[] = {{out->det_F, 0, 1},      {out->cauchy, 1, 6}, {out->von_mises, 7, 1},  {out->pressure, 8, 1},
                 {out->max_principal, 9, 1}, {out->energy, 10, 1}, {out->max_stretch, 11, 1}, {out->F, 12, 9}};
    for (const auto& pt : parts)
        if (pt.dst)
            FM_HIP(hipMemcpy(pt.dst, dout.p + (size_t)pt.row * ne, (size_t)pt.rows * ne * 8, hipMemcpyDeviceToHost));
    return FM_OK;
}

int fm_zz_estimate(fm_ctx* ctx, const double* u, double* elem_errors, double* node_avg, double* recovered,
                   double* global_estimate) {
    if (!ctx) return FM_EINVAL;
    FM_TRY(ensure_mesh(ctx));
    if (!ctx->have_n2e) return fail(ctx, FM_EINVAL, "fm_zz_estimate: more than 2^30 elements");
    const uint32_t nn = ctx->nn, ne = ctx->ne;
    const double* ud = ctx->u.p;
    DevBuf<double> du;
    if (u) {
        FM_HIP(du.alloc(3ull * nn));
        FM_HIP(hipMemcpyAsync(du.p, u, 3ull * nn * 8, hipMemcpyHostToDevice, ctx->stream));
        ud = du.p;
    } else if (!ctx->u_valid) {
        return fail(ctx, FM_ESTATE, "no device state (pass u)");
    }
    DevBuf<double> zz, raw, rec, err;
    FM_HIP(zz.alloc((size_t)kZzRows * ne));
    FM_HIP(raw.alloc(3ull * nn));
    FM_HIP(rec.alloc(3ull * nn));
    FM_HIP(err.alloc(ne));
    FM_TRY(ensure_partials(ctx, ne / 256 + 1));
    uint32_t np;
    {
        ProfScope ps(ctx, FM_K_FIELDS);
        np = launch_zz(ctx->stream, elem_args(ctx, ud), ctx->eiperm.p, ctx->n2e_ptr.p, ctx->n2e.p, zz.p, raw.p,
                       rec.p, err.p, ctx->partials.p);
    }
    FM_TRY(check_launch(ctx, "zz_estimate"));
    std::vector<double> part(np);
    FM_HIP(hipMemcpyAsync(part.data(), ctx->partials.p, (size_t)np * 8, hipMemcpyDeviceToHost, ctx->stream));
    FM_TRY(sync(ctx));
    double gsq = 0.0;
    for (double p : part) gsq += p;
    if (global_estimate) *global_estimate = std::sqrt(gsq);
    if (elem_errors) FM_HIP(hipMemcpy(elem_errors, err.p, (size_t)ne * 8, hipMemcpyDeviceToHost));
    if (node_avg) FM_HIP(hipMemcpy(node_avg, raw.p, 3ull * nn * 8, hipMemcpyDeviceToHost));
    if (recovered) FM_HIP(hipMemcpy(recovered, rec.p, 3ull * nn * 8, hipMemcpyDeviceToHost));
    return FM_OK;
}

// ---- checkpoint ------------------------------------------------------------------
// File: "FMCKPT02", u32 version 2, u32 flags 0, u64 length L of the next part;
// then L bytes laid out as CheckpointFile::save writes them
// (FEM/adaptive_fem_engine.hpp:1347-1398): u32 load_step, adapt_iter, n_nodes,
// n_elements, n_dofs; per node x, y, z, u, v, w; per element 4 x u32 nodes,
// volume, jacobian_det, quality_metric; per element state F_data (9, F_ij at
// i*3+j), stress (6, Mandel: M11, M22, M33, sqrt2 M12, sqrt2 M13, sqrt2 M23 of
// M = C S), eps_plastic (6), kappa, gamma_p; u.  Then mu_e and lam_e per
// element, and the FNV-1a hash of every byte before it.
namespace {

struct CkptWriter {
    FILE* fp = nullptr;
    uint64_t h = 1469598103934665603ull;
    std::vector<uint8_t> buf;
    bool ok = true;
    void put(const void* p, size_t n) {
        fnv1a_update(h, p, n);
        const uint8_t* b = static_cast<const uint8_t*>(p);
        buf.insert(buf.end(), b, b + n);
        if (buf.size() >= (8u << 20)) flush();
    }
    void flush() {
        if (!buf.empty()) ok = ok && std::fwrite(buf.data(), 1, buf.size(), fp) == buf.size();
        buf.clear();
    }
};

constexpr size_t kCkptNodeBytes = 48, kCkptElemBytes = 40, kCkptStateBytes = 23 * 8;

// the reference's tet_signed_volume6 / tet_volume / tet_quality
// (FEM/mesh_adaptation.hpp:110-146) on element e of the SoA mesh
void elem_record(const std::vector<double>& X, uint32_t nn, const uint32_t* nd, double* vol, double* jac,
                 double* qual) {
    double P[4][3];
    for (int a = 0; a < 4; ++a)
        for (int d = 0; d < 3; ++d) P[a][d] = X[(size_t)d * nn + nd[a]];
    const double e1[3] = {P[1][0] - P[0][0], P[1][1] - P[0][1], P[1][2] - P[0][2]};
    const double e2[3] = {P[2][0] - P[0][0], P[2][1] - P[0][1], P[2][2] - P[0][2]};
    const double e3[3] = {P[3][0] - P[0][0], P[3][1] - P[0][1], P[3][2] - P[0][2]};
    const double v6 = std::fabs(e1[0] * (e2[1] * e3[2] - e2[2] * e3[1]) - e1[1] * (e2[0] * e3[2] - e2[2] * e3[0]) +
                                e1[2] * (e2[0] * e3[1] - e2[1] * e3[0]));
    *vol = v6 / 6.0;
    *jac = *vol * 6.0;
    double sq = 0.0;
    for (int a = 0; a < 4; ++a)
        for (int b = a + 1; b < 4; ++b)
            for (int d = 0; d < 3; ++d) sq += (P[a][d] - P[b][d]) * (P[a][d] - P[b][d]);
    *qual = sq < 1e-30 ? 0.0 : 12.0 * std::sqrt(2.0) * v6 / std::pow(sq, 1.5);
}

}  // namespace

int fm_checkpoint_save(fm_ctx* ctx, const char* path, uint32_t load_step) {
    if (!ctx || !path) return FM_EINVAL;
    if (!ctx->u_valid) return fail(ctx, FM_ESTATE, "no device state");
    if (!ctx->have_material) return fail(ctx, FM_ESTATE, "no material");
    const uint32_t nn = ctx->nn, ne = ctx->ne, nd = 3 * nn;
    std::vector<double> u(nd);
    FM_HIP(hipMemcpy(u.data(), ctx->u.p, nd * 8ull, hipMemcpyDeviceToHost));
    // F of every element (element order) from the element-fields kernel
    std::vector<double> F((size_t)9 * ne);
    {
        fm_element_fields_out o{};
        o.F = F.data();
        FM_TRY(fm_element_fields(ctx, nullptr, &o));
    }
    FILE* fp = std::fopen(path, "wb");
    if (!fp) return fail(ctx, FM_EINVAL, std::string("cannot open ") + path);
    CkptWriter w;
    w.fp = fp;
    const uint64_t image = 20 + (uint64_t)nn * kCkptNodeBytes + (uint64_t)ne * (kCkptElemBytes + kCkptStateBytes) +
                           (uint64_t)nd * 8;
    const uint32_t ver[2] = {2u, 0u};
    w.put("FMCKPT02", 8);
    w.put(ver, sizeof ver);
    w.put(&image, 8);
    const uint32_t hdr[5] = {load_step, 0u /*adapt_iter*/, nn, ne, nd};
    w.put(hdr, sizeof hdr);
    for (uint32_t v = 0; v < nn; ++v) {
        const double rec[6] = {ctx->hX[v], ctx->hX[nn + v], ctx->hX[2ull * nn + v], u[3ull * v], u[3ull * v + 1],
                               u[3ull * v + 2]};
        w.put(rec, sizeof rec);
    }
    std::vector<double> geo((size_t)3 * ne);
#pragma omp parallel for schedule(static)
    for (int64_t e = 0; e < (int64_t)ne; ++e) {
        const uint32_t c[4] = {ctx->hconn[e], ctx->hconn[(size_t)ne + e], ctx->hconn[2ull * ne + e],
                               ctx->hconn[3ull * ne + e]};
        elem_record(ctx->hX, nn, c, &geo[3 * e], &geo[3 * e + 1], &geo[3 * e + 2]);
    }
    for (uint32_t e = 0; e < ne; ++e) {
        const uint32_t c[4] = {ctx->hconn[e], ctx->hconn[(size_t)ne + e], ctx->hconn[2ull * ne + e],
                               ctx->hconn[3ull * ne + e]};
        w.put(c, sizeof c);
        w.put(&geo[3ull * e], 24);
    }
    // states: F, the Mandel stress M = C S = mu (C - I) + lam ln(max(J, 0.01)) I
    // (neo_hookean_response, FEM/adaptive_fem_engine.hpp:79-129, with J = |det F|),
    // zero plastic fields; degenerate elements keep the initial state (F = I)
    std::vector<double> rec((size_t)23 * ne, 0.0);
#pragma omp parallel for schedule(static)
    for (int64_t e = 0; e < (int64_t)ne; ++e) {
        double* r = &rec[23 * e];
        double f[9];
        for (int k = 0; k < 9; ++k) f[k] = F[(size_t)k * ne + e];
        if (geo[3 * e] < 1e-30) {
            for (int k = 0; k < 9; ++k) r[k] = (k % 4 == 0) ? 1.0 : 0.0;
            continue;
        }
        double C[9];
        for (int i = 0; i < 3; ++i)
            for (int j = 0; j < 3; ++j) C[i * 3 + j] = f[i] * f[j] + f[3 + i] * f[3 + j] + f[6 + i] * f[6 + j];
        const double J = std::fabs(f[0] * (f[4] * f[8] - f[5] * f[7]) - f[1] * (f[3] * f[8] - f[5] * f[6]) +
                                   f[2] * (f[3] * f[7] - f[4] * f[6]));
        const double mu = ctx->hmu[e], lam = ctx->hlam[e];
        const double p = lam * std::log(std::max(J, 0.01)) - mu;
        const double s2 = std::sqrt(2.0);
        for (int k = 0; k < 9; ++k) r[k] = f[k];
        r[9] = mu * C[0] + p;
        r[10] = mu * C[4] + p;
        r[11] = mu * C[8] + p;
        r[12] = s2 * mu * C[1];
        r[13] = s2 * mu * C[2];
        r[14] = s2 * mu * C[5];
    }
    w.put(rec.data(), rec.size() * 8);
    w.put(u.data(), nd * 8ull);
    w.put(ctx->hmu.data(), (size_t)ne * 8);
    w.put(ctx->hlam.data(), (size_t)ne * 8);
    w.flush();
    const uint64_t hh = w.h;
    bool ok = w.ok && std::fwrite(&hh, 1, 8, fp) == 8;
    ok = (std::fclose(fp) == 0) && ok;
    return ok ? FM_OK : fail(ctx, FM_EINVAL, "checkpoint write failed");
}

int fm_checkpoint_load(fm_ctx* ctx, const char* path, uint32_t* load_step) {
    if (!ctx || !path) return FM_EINVAL;
    FILE* fp = std::fopen(path, "rb");
    if (!fp) return fail(ctx, FM_EINVAL, std::string("cannot open ") + path);
    std::fseek(fp, 0, SEEK_END);
    const long size = std::ftell(fp);
    std::fseek(fp, 0, SEEK_SET);
    std::vector<uint8_t> buf(size > 0 ? (size_t)size : 0);
    const bool rd = size > 0 && std::fread(buf.data(), 1, buf.size(), fp) == buf.size();
    std::fclose(fp);
    if (!rd || buf.size() < 24 + 20 + 8) return fail(ctx, FM_EINVAL, "checkpoint truncated");
    if (std::memcmp(buf.data(), "FMCKPT02", 8) != 0) return fail(ctx, FM_EINVAL, "not a version-2 checkpoint file");
    uint64_t h = 1469598103934665603ull;
    fnv1a_update(h, buf.data(), buf.size() - 8);
    uint64_t stored;
    std::memcpy(&stored, buf.data() + buf.size() - 8, 8);
    if (h != stored) return fail(ctx, FM_EINVAL, "checkpoint checksum mismatch");
    uint64_t image;
    std::memcpy(&image, buf.data() + 16, 8);
    const uint8_t* img = buf.data() + 24;
    uint32_t hdr[5];
    std::memcpy(hdr, img, sizeof hdr);
    const uint32_t nn = hdr[2], ne = hdr[3], nd = hdr[4];
    const uint64_t expect_img = 20 + (uint64_t)nn * kCkptNodeBytes + (uint64_t)ne * (kCkptElemBytes + kCkptStateBytes) +
                                (uint64_t)nd * 8;
    if (nd != 3ull * nn || image != expect_img || buf.size() != 24 + image + 16ull * ne + 8)
        return fail(ctx, FM_EINVAL, "checkpoint sizes are inconsistent");
    // mesh and material of the file
    std::vector<double> X((size_t)3 * nn), u(nd), mu(ne), lam(ne);
    std::vector<uint32_t> conn((size_t)4 * ne);
    const uint8_t* pn = img + 20;
    for (uint32_t v = 0; v < nn; ++v) {
        double r[6];
        std::memcpy(r, pn + (size_t)v * kCkptNodeBytes, sizeof r);
        X[v] = r[0];
        X[nn + v] = r[1];
        X[2ull * nn + v] = r[2];
    }
    const uint8_t* pe = pn + (size_t)nn * kCkptNodeBytes;
    for (uint32_t e = 0; e < ne; ++e) {
        uint32_t c[4];
        std::memcpy(c, pe + (size_t)e * kCkptElemBytes, sizeof c);
        for (int a = 0; a < 4; ++a) {
            if (c[a] >= nn) return fail(ctx, FM_EINVAL, "checkpoint connectivity out of range");
            conn[(size_t)a * ne + e] = c[a];
        }
    }
    const uint8_t* pu = pe + (size_t)ne * (kCkptElemBytes + kCkptStateBytes);
    std::memcpy(u.data(), pu, (size_t)nd * 8);
    std::memcpy(mu.data(), pu + (size_t)nd * 8, (size_t)ne * 8);
    std::memcpy(lam.data(), pu + (size_t)nd * 8 + (size_t)ne * 8, (size_t)ne * 8);
    if (ctx->nn) {
        // a context with a mesh restores u only, and only onto the same mesh and material
        if (nn != ctx->nn || ne != ctx->ne) return fail(ctx, FM_EINVAL, "checkpoint is of a mesh of another size");
        if (std::memcmp(X.data(), ctx->hX.data(), X.size() * 8) != 0)
            return fail(ctx, FM_EINVAL, "checkpoint node coordinates differ from the mesh");
        if (std::memcmp(conn.data(), ctx->hconn.data(), conn.size() * 4) != 0)
            return fail(ctx, FM_EINVAL, "checkpoint connectivity differs from the mesh");
        if (ctx->have_material && (std::memcmp(mu.data(), ctx->hmu.data(), (size_t)ne * 8) != 0 ||
                                   std::memcmp(lam.data(), ctx->hlam.data(), (size_t)ne * 8) != 0))
            return fail(ctx, FM_EINVAL, "checkpoint material differs from the context's");
        if (!ctx->have_material) FM_TRY(fm_set_material(ctx, mu.data(), lam.data()));
    } else {
        // an empty context takes the mesh and the material from the file
        FM_TRY(fm_set_mesh(ctx, nn, X.data(), ne, conn.data()));
        FM_TRY(fm_set_material(ctx, mu.data(), lam.data()));
    }
    if (load_step) *load_step = hdr[0];
    return fm_set_u(ctx, u.data());
}

int fm_profile_enable(fm_ctx* ctx, int on) {
    if (!ctx) return FM_EINVAL;
    (void)hipStreamSynchronize(ctx->stream);
    prof_flush(ctx);
    ctx->prof = on != 0;
    ctx->prof_every = on > 1 ? on : 1;
    return FM_OK;
}

int fm_profile_reset(fm_ctx* ctx) {
    if (!ctx) return FM_EINVAL;
    (void)hipStreamSynchronize(ctx->stream);
    prof_flush(ctx);
    for (int k = 0; k < FM_K_COUNT; ++k) {
        ctx->ms[k] = 0;
        ctx->launches[k] = 0;
        ctx->prof_seen[k] = 0;
    }
    return FM_OK;
}

int fm_profile_get(fm_ctx* ctx, double* ms, uint64_t* launches) {
    if (!ctx) return FM_EINVAL;
    (void)hipStreamSynchronize(ctx->stream);
    prof_flush(ctx);
    // launches: every launch of the class while profiling was on; ms: the
    // mean of the timed ones times that count
    for (int k = 0; k < FM_K_COUNT; ++k) {
        const uint64_t seen = std::max<uint64_t>(ctx->prof_seen[k], ctx->launches[k]);
        if (ms) ms[k] = ctx->launches[k] ? ctx->ms[k] / (double)ctx->launches[k] * (double)seen : 0.0;
        if (launches) launches[k] = seen;
    }
    return FM_OK;
}

int fm_info(fm_ctx* ctx, uint64_t* out5) {
    if (!ctx || !out5) return FM_EINVAL;
    const Matrix* A = ctx->cur ? ctx->cur : &ctx->tangent;
    out5[0] = A->n;
    out5[1] = A->nnz;
    out5[2] = ctx->ne;
    out5[3] = ctx->nn;
    out5[4] = ctx->n_colors;
    return FM_OK;
}

int fm_sweep_info(fm_ctx* ctx, uint64_t* o) {
    if (!ctx || !o) return FM_EINVAL;
    const Matrix* A = ctx->cur ? ctx->cur : &ctx->tangent;
    const ChainSchedule& C = A->chain;
    o[0] = C.ok ? 1 : 0;
    o[1] = A->sched.levels_fwd;
    o[2] = A->sched.levels_bwd;
    o[3] = C.ok ? C.chains : 0;
    o[4] = C.ok ? C.fwd.nb : 0;
    o[5] = C.ok ? C.fwd.nq : 0;
    o[6] = C.ok ? (uint64_t)C.lpn : 0;
    o[7] = C.ok ? C.gj : 0;
    o[8] = C.ok ? C.gk : 0;
    o[9] = A->node_spmv ? 1 : 0;
    return FM_OK;
}

}  // extern "C"
