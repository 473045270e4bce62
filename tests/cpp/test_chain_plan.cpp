// CPU check of the chain-bundle sweep plan (no GPU): plans the ILU(0) sweeps
// of assembled tangents, replays the plan with the data flow of the sweep
// kernels (packed dependency blocks, the LDS ring with its overwrite timing,
// published values) and compares with the oracle's sequential ILU(0) apply.
// Every dependency word is checked against the CSR column it stands for, and
// the device form of each plan (the kernels' step records) decodes back to it.
#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <random>
#include <string>
#include <vector>

#include "../../atlas-flexmesh_amd/csrc/fm_plan.hpp"
#include "../../oracle/fm_oracle.h"

using fm::ChainPlan;
using fm::ChainPlanDir;

static int g_fail = 0;
#define CHECK(c, ...)                                                 \
    do {                                                              \
        if (!(c)) {                                                   \
            std::printf("FAIL %s:%d: ", __FILE__, __LINE__);          \
            std::printf(__VA_ARGS__);                                 \
            std::printf("\n");                                        \
            ++g_fail;                                                 \
            return false;                                             \
        }                                                             \
    } while (0)

struct Csr {
    uint32_t nn = 0;
    std::vector<uint32_t> rp, ci;
    std::vector<double> LU, diag;
};

// replay of one sweep: in -> out, in the order and with the storage the kernel uses
static bool replay(const Csr& A, const ChainPlan& P, const ChainPlanDir& W, bool bwd, const std::vector<double>& in,
                   std::vector<double>& out) {
    const int lpn = P.lpn, npw = 64 / lpn, bpl = fm::kChainMaxBlk / lpn;
    const uint32_t n = 3 * A.nn;
    out.assign(n, 0.0);
    std::vector<char> done(A.nn, 0);
    std::vector<double> ring((size_t)fm::kChainRing * npw * 3, NAN);
    std::vector<uint32_t> ring_node((size_t)fm::kChainRing * npw, fm::kDepNone);
    uint32_t processed = 0;
    for (uint32_t b = 0; b < W.nb; ++b) {
        for (uint32_t s = 0; s < W.bsteps[b]; ++s) {
            const uint64_t q = (uint64_t)W.bstart[b] + s;
            double res[64][3];
            uint32_t nodes[64];
            for (int g = 0; g < npw; ++g) {
                const uint32_t v = W.node[q * npw + g];
                nodes[g] = v;
                if (v == fm::kDepNone) continue;
                CHECK(v < A.nn && !done[v], "node %u out of range or repeated", v);
                const uint32_t r0 = A.rp[3 * v], nblk = (A.rp[3 * v + 1] - r0) / 3;
                uint32_t nl = 0;
                while (A.ci[r0 + 3 * nl] / 3 < v) ++nl;
                const uint32_t ndep = bwd ? nblk - nl - 1 : nl;
                double part[3] = {0, 0, 0};
                uint32_t seen = 0;
                for (int gl = 0; gl < lpn; ++gl)
                    for (int j = 0; j < bpl; ++j) {
                        const uint32_t k = gl + j * lpn;
                        const uint32_t dw = W.dep[(q * bpl + j) * 64 + g * lpn + gl];
                        if (k >= ndep) {
                            CHECK(dw == fm::kDepNone, "node %u block %u: word for a missing block", v, k);
                            continue;
                        }
                        CHECK(dw != fm::kDepNone, "node %u block %u: no word", v, k);
                        ++seen;
                        const uint32_t pos = bwd ? nl + 1 + k : k;
                        const uint32_t m = A.ci[r0 + 3 * pos] / 3;
                        double y[3];
                        if (dw & fm::kDepLocal) {
                            const uint32_t slot = dw & ~fm::kDepLocal;
                            CHECK(slot < ring_node.size(), "slot %u out of range", slot);
                            CHECK(ring_node[slot] == m, "node %u: LDS slot %u holds %u, wants %u", v, slot,
                                  ring_node[slot], m);
                            for (int c = 0; c < 3; ++c) y[c] = ring[slot * 3 + c];
                        } else {
                            CHECK(dw == m, "node %u: word %u for dependency %u", v, dw, m);
                            CHECK(done[m], "node %u: dependency %u not published yet", v, m);
                            for (int c = 0; c < 3; ++c) y[c] = out[3 * m + c];
                        }
                        for (int d = 0; d < 3; ++d)
                            for (int c = 0; c < 3; ++c) part[d] += A.LU[A.rp[3 * v + d] + 3 * pos + c] * y[c];
                    }
                CHECK(seen == ndep, "node %u: %u of %u dependency words", v, seen, ndep);
                const uint32_t o = 3 * nl;
                auto lu = [&](int d, int e) { return A.LU[A.rp[3 * v + d] + o + e]; };
                double w[3];
                if (!bwd) {
                    w[0] = in[3 * v] - part[0];
                    w[1] = in[3 * v + 1] - part[1] - lu(1, 0) * w[0];
                    w[2] = in[3 * v + 2] - part[2] - lu(2, 0) * w[0] - lu(2, 1) * w[1];
                } else {
                    w[2] = (in[3 * v + 2] - part[2]) / lu(2, 2);
                    w[1] = (in[3 * v + 1] - part[1] - lu(1, 2) * w[2]) / lu(1, 1);
                    w[0] = (in[3 * v] - part[0] - lu(0, 1) * w[1] - lu(0, 2) * w[2]) / lu(0, 0);
                }
                for (int d = 0; d < 3; ++d) res[g][d] = w[d];
            }
            // writes of the step follow all of its reads
            for (int g = 0; g < npw; ++g) {
                const uint32_t v = nodes[g];
                if (v == fm::kDepNone) continue;
                const size_t slot = (size_t)(s % fm::kChainRing) * npw + g;
                for (int d = 0; d < 3; ++d) {
                    out[3 * v + d] = res[g][d];
                    ring[slot * 3 + d] = res[g][d];
                }
                ring_node[slot] = v;
                done[v] = 1;
                ++processed;
            }
        }
    }
    CHECK(processed == A.nn, "%u of %u nodes processed", processed, A.nn);
    return true;
}

static Csr box_tangent(int nx, int ny, int nz, double Lx, uint32_t seed, bool permute) {
    Csr A;
    const uint32_t nn = (uint32_t)((nx + 1) * (ny + 1) * (nz + 1)), ne = 6u * nx * ny * nz;
    std::vector<double> X(3 * (size_t)nn);
    std::vector<uint32_t> conn(4 * (size_t)ne);
    orc_box_mesh(nx, ny, nz, Lx, 1.0, 1.0, X.data(), conn.data());
    std::mt19937 rng(seed);
    if (permute) {  // a numbering with no chain structure
        std::vector<uint32_t> perm(nn);
        for (uint32_t i = 0; i < nn; ++i) perm[i] = i;
        std::shuffle(perm.begin(), perm.end(), rng);
        std::vector<double> Xp(X.size());
        for (uint32_t i = 0; i < nn; ++i)
            for (int d = 0; d < 3; ++d) Xp[(size_t)d * nn + perm[i]] = X[(size_t)d * nn + i];
        X = Xp;
        for (auto& c : conn) c = perm[c];
    }
    A.nn = nn;
    A.rp.resize(3 * (size_t)nn + 1);
    const uint64_t nnz = orc_pattern_rows(nn, ne, conn.data(), A.rp.data());
    A.ci.resize(nnz);
    orc_pattern_cols(nn, ne, conn.data(), A.rp.data(), A.ci.data());
    std::vector<uint32_t> colors(ne);
    const uint32_t nc = orc_color(nn, ne, conn.data(), colors.data());
    std::vector<double> mu(ne), lam(ne), u(3 * (size_t)nn), f(3 * (size_t)nn), K(nnz);
    std::uniform_real_distribution<double> U(-0.02, 0.02), M(0.5, 2.0);
    for (uint32_t e = 0; e < ne; ++e) {
        mu[e] = M(rng);
        lam[e] = 2.0 * M(rng);
    }
    for (auto& x : u) x = U(rng);
    orc_assemble(nn, ne, X.data(), conn.data(), mu.data(), lam.data(), colors.data(), nc, u.data(), A.rp.data(),
                 A.ci.data(), K.data(), f.data());
    A.LU.resize(nnz);
    A.diag.resize(3 * (size_t)nn);
    orc_ilu0(3 * nn, A.rp.data(), A.ci.data(), K.data(), A.LU.data(), A.diag.data());
    return A;
}

// the device form of a plan direction (encode_chain_records) decodes back to
// the plan's nodes and dependency words exactly
static bool check_records(const ChainPlanDir& W, int lpn) {
    const int npw = 64 / lpn, bpl = fm::kChainMaxBlk / lpn;
    std::vector<uint32_t> far;
    fm::encode_chain_records(W, lpn, far);
    CHECK(far.size() == (size_t)W.nq * fm::kRecWords, "record size %zu", far.size());
    for (size_t q = 0; q < W.nq; ++q) {
        const uint32_t* f = &far[q * fm::kRecWords];
        for (int g = 0; g < npw; ++g) {
            const uint32_t x = f[128 + g];
            CHECK(x == fm::kOffNone || x % 12 == 0, "step %zu quad %d: offset %u", q, g, x);
            const uint32_t nd = x == fm::kOffNone ? fm::kDepNone : x / 12;
            CHECK(nd == W.node[q * npw + g], "step %zu quad %d: node %u, plan %u", q, g, nd, W.node[q * npw + g]);
        }
        for (int lane = 0; lane < 64; ++lane)
            for (int j = 0; j < bpl; ++j) {
                const uint32_t w = f[2 * lane + j];
                uint32_t dw;
                if ((int32_t)w >= 0) {
                    CHECK(w % 12 == 0, "step %zu lane %d: external offset %u", q, lane, w);
                    dw = w / 12;
                } else {
                    const uint32_t slot = w & 0xFFFFu;
                    CHECK(slot <= fm::kZeroSlot, "step %zu lane %d: slot %u", q, lane, slot);
                    dw = slot == fm::kZeroSlot ? fm::kDepNone : (fm::kDepLocal | slot);
                }
                const uint32_t want = W.dep[(q * bpl + j) * 64 + lane];
                CHECK(dw == want, "step %zu lane %d block %d: word %08x, plan %08x", q, lane, j, dw, want);
            }
    }
    return true;
}

static bool run_case(const char* name, int nx, int ny, int nz, double Lx, bool permute, int lpn, bool expect_plan) {
    const Csr A = box_tangent(nx, ny, nz, Lx, 17u + nx, permute);
    ChainPlan P;
    std::string why;
    const bool ok = fm::plan_chain_schedule(A.rp, A.ci, lpn, P, why);
    if (!ok) {
        std::printf("%-26s lpn %d: no plan (%s)\n", name, lpn, why.c_str());
        CHECK(!expect_plan, "%s: expected a plan", name);
        return true;
    }
    const uint32_t n = 3 * A.nn;
    std::vector<double> x(n), z, y, yref(n);
    std::mt19937 rng(5);
    std::normal_distribution<double> N(0.0, 1.0);
    for (auto& v : x) v = N(rng);
    if (!replay(A, P, P.fwd, false, x, z) || !replay(A, P, P.bwd, true, z, y)) return false;
    if (lpn == 4 && (!check_records(P.fwd, lpn) || !check_records(P.bwd, lpn))) return false;
    orc_ilu0_apply(n, A.rp.data(), A.ci.data(), A.LU.data(), A.diag.data(), x.data(), yref.data());
    double num = 0, den = 0;
    for (uint32_t i = 0; i < n; ++i) {
        num = std::max(num, std::fabs(y[i] - yref[i]));
        den = std::max(den, std::fabs(yref[i]));
    }
    const int npw = 64 / lpn;
    const double fill = (double)A.nn / ((double)P.fwd.nq * npw);
    std::printf("%-26s lpn %d: chains %u stride %u bundle %ux%u bundles %u steps %u/%u fill %.2f  rel %.2e\n", name,
                lpn, P.chains, P.stride, P.gj, P.gk, P.fwd.nb, P.fwd.nq, P.bwd.nq, fill, num / den);
    CHECK(num <= 1e-12 * den, "%s: replay differs from the sequential apply (%.3e)", name, num / den);
    return true;
}

int main() {
    for (int lpn : {4, 8}) {
        run_case("box 6x5x4", 6, 5, 4, 1.0, false, lpn, true);
        run_case("box 17x9x5", 17, 9, 5, 1.0, false, lpn, true);
        run_case("beam 24x3x3", 24, 3, 3, 8.0, false, lpn, true);
        run_case("line 10x1x1", 10, 1, 1, 1.0, false, lpn, true);
        run_case("cube 12x12x12", 12, 12, 12, 1.0, false, lpn, true);
        run_case("permuted 5x4x4", 5, 4, 4, 1.0, true, lpn, false);
    }
    std::printf(g_fail ? "FAILED (%d)\n" : "ALL OK\n", g_fail);
    return g_fail ? 1 : 0;
}
