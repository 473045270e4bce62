// fm_setup.cpp -- one-time host passes: node adjacency, greedy colouring and
// the level schedule of the ILU(0) dependence graph.  They run once per mesh
// (the pattern is fixed for the whole Newton solve), in O(nnz / 9).
#include <omp.h>

#include <algorithm>
#include <cstdlib>
#include <cstring>

#include "fm_internal.hpp"

namespace fm {

// Node adjacency through elements, sorted and including the node itself: the
// node-level form of the row structure of build_csr_sparsity_pattern_rigorous
// (FEM/adaptive_fem_engine.hpp:1077-1102).
void node_adjacency(uint32_t nn, uint32_t ne, const uint32_t* conn, std::vector<uint32_t>& ptr,
                    std::vector<uint32_t>& nbr) {
    std::vector<uint32_t> eptr(nn + 1, 0);
    for (int a = 0; a < 4; ++a)
        for (uint32_t e = 0; e < ne; ++e) eptr[conn[(size_t)a * ne + e] + 1]++;
    for (uint32_t n = 0; n < nn; ++n) eptr[n + 1] += eptr[n];
    std::vector<uint32_t> elist(eptr[nn]);
    {
        std::vector<uint32_t> fill(eptr.begin(), eptr.end() - 1);
        for (uint32_t e = 0; e < ne; ++e)
            for (int a = 0; a < 4; ++a) elist[fill[conn[(size_t)a * ne + e]]++] = e;
    }
    std::vector<uint32_t> deg(nn, 0);
    auto gather = [&](uint32_t n, std::vector<uint32_t>& v) {
        v.clear();
        for (uint32_t k = eptr[n]; k < eptr[n + 1]; ++k)
            for (int a = 0; a < 4; ++a) v.push_back(conn[(size_t)a * ne + elist[k]]);
        std::sort(v.begin(), v.end());
        v.erase(std::unique(v.begin(), v.end()), v.end());
    };
#pragma omp parallel
    {
        std::vector<uint32_t> v;
#pragma omp for schedule(static)
        for (int64_t n = 0; n < (int64_t)nn; ++n) {
            gather((uint32_t)n, v);
            deg[n] = (uint32_t)v.size();
        }
    }
    ptr.assign(nn + 1, 0);
    for (uint32_t n = 0; n < nn; ++n) ptr[n + 1] = ptr[n] + deg[n];
    nbr.resize(ptr[nn]);
#pragma omp parallel
    {
        std::vector<uint32_t> v;
#pragma omp for schedule(static)
        for (int64_t n = 0; n < (int64_t)nn; ++n) {
            gather((uint32_t)n, v);
            std::copy(v.begin(), v.end(), nbr.begin() + ptr[n]);
        }
    }
}

// MeshColoring::compute (FEM/adaptive_fem_engine.hpp:1549-1602): element e
// takes the smallest colour not held by an earlier element that shares a
// node.  The set of colours of earlier elements through node n is kept as a
// bit mask per node, so the pass is O(n_e) instead of the all-pairs O(n_e^2)
// adjacency of the reference; the assignment is identical.
uint32_t greedy_colors(uint32_t nn, uint32_t ne, const uint32_t* conn, uint32_t* colors) {
    constexpr int W = 4;  // up to 256 colours
    std::vector<uint64_t> mask((size_t)nn * W, 0);
    uint32_t ncol = 0;
    for (uint32_t e = 0; e < ne; ++e) {
        uint32_t nd[4];
        uint64_t used[W] = {0, 0, 0, 0};
        for (int a = 0; a < 4; ++a) {
            nd[a] = conn[(size_t)a * ne + e];
            for (int w = 0; w < W; ++w) used[w] |= mask[(size_t)nd[a] * W + w];
        }
        uint32_t c = UINT32_MAX;
        for (int w = 0; w < W; ++w)
            if (~used[w]) {
                c = 64 * w + __builtin_ctzll(~used[w]);
                break;
            }
        if (c == UINT32_MAX) return 0;  // more than 256 colours: caller reports it
        colors[e] = c;
        for (int a = 0; a < 4; ++a) mask[(size_t)nd[a] * W + c / 64] |= 1ull << (c % 64);
        ncol = std::max(ncol, c + 1);
    }
    return ncol;
}

// 3 when rows come in groups of three with one shared, sorted column set made
// of whole 3x3 blocks (the FE pattern); 1 otherwise.
int detect_block(const std::vector<uint32_t>& rp, const std::vector<uint32_t>& ci) {
    const size_t n = rp.size() - 1;
    if (n % 3 != 0 || n == 0) return 1;
    for (size_t node = 0; node < n / 3; ++node) {
        const uint32_t b0 = rp[3 * node], len = rp[3 * node + 1] - b0;
        if (len % 3 != 0) return 1;
        for (int d = 1; d < 3; ++d) {
            const uint32_t bd = rp[3 * node + d];
            if (rp[3 * node + d + 1] - bd != len) return 1;
            if (std::memcmp(&ci[b0], &ci[bd], len * sizeof(uint32_t)) != 0) return 1;
        }
        for (uint32_t j = 0; j < len; j += 3) {
            const uint32_t c = ci[b0 + j];
            if (c % 3 != 0 || ci[b0 + j + 1] != c + 1 || ci[b0 + j + 2] != c + 2) return 1;
            if (j > 0 && ci[b0 + j - 1] >= c) return 1;
        }
    }
    return 3;
}

// Level schedule of natural-order ILU(0) on groups of B rows.  level(node) =
// 1 + max level of its lower neighbours (forward) or upper neighbours
// (backward); nodes are queued by level, so every dependence of a queued node
// sits earlier in the queue.  Rows must have sorted columns and a stored
// diagonal.
bool build_schedule(const std::vector<uint32_t>& rp, const std::vector<uint32_t>& ci, int B, Schedule& S,
                    hipStream_t stream, std::string& err) {
    const uint32_t n = (uint32_t)(rp.size() - 1);
    const uint32_t nnode = n / B;
    std::vector<uint16_t> nlow(nnode);
    uint32_t maxlen = 0;
    uint64_t lower_total = 0;
    for (uint32_t node = 0; node < nnode; ++node) {
        const uint32_t r = B * node, b = rp[r], e = rp[r + 1];
        maxlen = std::max(maxlen, e - b);
        uint32_t k = b;
        while (k < e && ci[k] < r) ++k;
        if (k == e || ci[k] != r) {
            err = "matrix row " + std::to_string(r) + " has no stored diagonal";
            return false;
        }
        const uint32_t low = (k - b) / B;
        if (low > 65535) {
            err = "row has too many lower entries";
            return false;
        }
        nlow[node] = (uint16_t)low;
        lower_total += k - b;
    }
    if (maxlen > 256) {
        err = "row length " + std::to_string(maxlen) + " exceeds the 256 supported by the ILU(0) kernels";
        return false;
    }
    std::vector<uint32_t> lev(nnode, 0), blev(nnode, 0);
    uint32_t L = 0, LB = 0;
    for (uint32_t node = 0; node < nnode; ++node) {
        const uint32_t b = rp[B * node];
        uint32_t l = 0;
        for (uint32_t j = 0; j < (uint32_t)nlow[node] * B; j += B) l = std::max(l, lev[ci[b + j] / B] + 1);
        lev[node] = l;
        L = std::max(L, l + 1);
    }
    for (int64_t node = (int64_t)nnode - 1; node >= 0; --node) {
        const uint32_t b = rp[B * node], e = rp[B * node + 1];
        uint32_t l = 0;
        for (uint32_t j = b + (uint32_t)(nlow[node] + 1) * B; j < e; j += B) l = std::max(l, blev[ci[j] / B] + 1);
        blev[node] = l;
        LB = std::max(LB, l + 1);
    }
    S.lpn = (B == 3 || (nnode && (double)lower_total / nnode > 6.0)) ? 16 : 4;
    if (B == 3) {
        const char* e = std::getenv("FM_TRSV_LPN");  // lanes per node of the sweeps (16 or 32)
        if (e && std::atoi(e) == 32) S.lpn = 32;
    }
    const uint32_t npw = 64 / S.lpn;  // nodes per wave in the triangular solves
    auto queue = [npw](const std::vector<uint32_t>& level, uint32_t nlev, bool descending_ids) {
        std::vector<uint32_t> cnt(nlev + 1, 0);
        for (uint32_t l : level) cnt[l + 1]++;
        std::vector<uint32_t> start(nlev + 1, 0);
        for (uint32_t l = 0; l < nlev; ++l) start[l + 1] = start[l] + (cnt[l + 1] + npw - 1) / npw * npw;
        std::vector<uint32_t> q(start[nlev], 0xFFFFFFFFu);
        std::vector<uint32_t> fill(start.begin(), start.end() - 1);
        const uint32_t nn = (uint32_t)level.size();
        for (uint32_t k = 0; k < nn; ++k) {
            const uint32_t node = descending_ids ? nn - 1 - k : k;
            q[fill[level[node]]++] = node;
        }
        return q;
    };
    const std::vector<uint32_t> fwd = queue(lev, L, false);
    const std::vector<uint32_t> bwd = queue(blev, LB, true);
    S.B = B;
    S.n_nodes = nnode;
    S.slots = (int)((maxlen + 63) / 64);
    if (S.slots == 3) S.slots = 4;
    S.levels_fwd = L;
    S.levels_bwd = LB;
    S.q_fwd = (uint32_t)fwd.size();
    S.q_bwd = (uint32_t)bwd.size();
    if (S.fwd_order.alloc(S.q_fwd) != hipSuccess || S.bwd_order.alloc(S.q_bwd) != hipSuccess ||
        S.nlow.alloc(nnode) != hipSuccess) {
        err = "hipMalloc failed for the ILU schedule";
        return false;
    }
    if (hipMemcpyAsync(S.fwd_order.p, fwd.data(), S.q_fwd * 4ull, hipMemcpyHostToDevice, stream) != hipSuccess ||
        hipMemcpyAsync(S.bwd_order.p, bwd.data(), S.q_bwd * 4ull, hipMemcpyHostToDevice, stream) != hipSuccess ||
        hipMemcpyAsync(S.nlow.p, nlow.data(), nnode * 2, hipMemcpyHostToDevice, stream) != hipSuccess ||
        hipStreamSynchronize(stream) != hipSuccess) {
        err = "upload of the ILU schedule failed";
        return false;
    }
    return true;
}

// ---------------------------------------------------------------------------
// Chain-bundle schedule of the sweeps (B = 3).  Chains: node v continues the
// chain of v - 1 when v - 1 is one of its lower neighbours.  Chains are laid
// out on a grid of `stride` columns (the most frequent chain-to-chain
// dependence distance above one: the chains of one plane of a box mesh) and
// grouped into gj x gk bundles, gj * gk = 64 / lpn.  Within a bundle the
// step of a node is one more than the latest step of its chain predecessor
// and of its dependencies inside the bundle; bundles are ordered by their
// level in the bundle graph.  Every property the kernels rely on is checked
// here; a matrix that fails one keeps the level-scheduled sweeps.
// ---------------------------------------------------------------------------
namespace {

// dependency node blocks of v in one direction: lower blocks 0..nlow-1 or
// upper blocks nlow+1..nblk-1, in column order
inline void dep_blocks(const std::vector<uint32_t>& rp, const std::vector<uint32_t>& ci, uint32_t v, uint32_t nl,
                       bool bwd, uint32_t* out, uint32_t& cnt) {
    const uint32_t b = rp[3 * v], nblk = (rp[3 * v + 1] - b) / 3;
    cnt = 0;
    if (!bwd)
        for (uint32_t k = 0; k < nl; ++k) out[cnt++] = ci[b + 3 * k] / 3;
    else
        for (uint32_t k = nl + 1; k < nblk; ++k) out[cnt++] = ci[b + 3 * k] / 3;
}

bool chain_direction(const std::vector<uint32_t>& rp, const std::vector<uint32_t>& ci,
                     const std::vector<uint16_t>& nlow, const std::vector<uint32_t>& cid,
                     const std::vector<uint32_t>& bnd, const std::vector<uint32_t>& grp, uint32_t nbund, int lpn,
                     bool bwd, ChainPlanDir& H, std::string& why) {
    const uint32_t nn = (uint32_t)nlow.size();
    const int npw = 64 / lpn, bpl = kChainMaxBlk / lpn;
    // bundle graph: edge bundle(m) -> bundle(v) for a dependency m of v
    std::vector<uint64_t> edges;
    uint32_t deps[kChainMaxBlk + 1], cnt = 0;
    for (uint32_t v = 0; v < nn; ++v) {
        dep_blocks(rp, ci, v, nlow[v], bwd, deps, cnt);
        for (uint32_t k = 0; k < cnt; ++k)
            if (bnd[deps[k]] != bnd[v]) edges.push_back((uint64_t)bnd[deps[k]] << 32 | bnd[v]);
    }
    std::sort(edges.begin(), edges.end());
    edges.erase(std::unique(edges.begin(), edges.end()), edges.end());
    std::vector<uint32_t> indeg(nbund, 0), eptr(nbund + 1, 0), lev(nbund, 0);
    for (uint64_t e : edges) {
        eptr[(e >> 32) + 1]++;
        indeg[(uint32_t)e]++;
    }
    for (uint32_t b = 0; b < nbund; ++b) eptr[b + 1] += eptr[b];
    std::vector<uint32_t> stack;
    for (uint32_t b = 0; b < nbund; ++b)
        if (!indeg[b]) stack.push_back(b);
    uint32_t seen = 0;
    while (!stack.empty()) {
        const uint32_t b = stack.back();
        stack.pop_back();
        ++seen;
        for (uint32_t k = eptr[b]; k < eptr[b + 1]; ++k) {
            const uint32_t t = (uint32_t)edges[k];
            lev[t] = std::max(lev[t], lev[b] + 1);
            if (--indeg[t] == 0) stack.push_back(t);
        }
    }
    if (seen != nbund) {
        why = "bundle graph has a cycle";
        return false;
    }
    std::vector<uint32_t> order(nbund), pos(nbund);
    for (uint32_t b = 0; b < nbund; ++b) order[b] = b;
    std::sort(order.begin(), order.end(), [&](uint32_t a, uint32_t b) {
        return lev[a] != lev[b] ? lev[a] < lev[b] : a < b;
    });
    for (uint32_t i = 0; i < nbund; ++i) pos[order[i]] = i;
    // steps, in the direction's node order (a topological order of the node DAG)
    std::vector<uint32_t> st(nn, 0), nsteps(nbund, 0);
    for (uint32_t t = 0; t < nn; ++t) {
        const uint32_t v = bwd ? nn - 1 - t : t;
        uint32_t sv = 0;
        const uint32_t pred = bwd ? v + 1 : v - 1;
        if ((bwd ? v + 1 < nn : v > 0) && cid[pred] == cid[v]) sv = st[pred] + 1;
        dep_blocks(rp, ci, v, nlow[v], bwd, deps, cnt);
        for (uint32_t k = 0; k < cnt; ++k)
            if (bnd[deps[k]] == bnd[v]) sv = std::max(sv, st[deps[k]] + 1);
        st[v] = sv;
        nsteps[bnd[v]] = std::max(nsteps[bnd[v]], sv + 1);
    }
    H.bstart.assign(nbund, 0);
    H.bsteps.assign(nbund, 0);
    uint64_t q = 0;
    for (uint32_t i = 0; i < nbund; ++i) {
        H.bstart[i] = (uint32_t)q;
        H.bsteps[i] = nsteps[order[i]];
        q += nsteps[order[i]];
    }
    if (q >= (1ull << 31) / 64) {
        why = "schedule too large";
        return false;
    }
    H.nq = (uint32_t)q;
    H.node.assign((size_t)q * npw, kDepNone);
    H.dep.assign((size_t)q * bpl * 64, kDepNone);
    for (uint32_t v = 0; v < nn; ++v) {
        const uint32_t b = bnd[v], g = grp[v];
        const uint64_t qs = (uint64_t)H.bstart[pos[b]] + st[v];
        uint32_t& slot = H.node[qs * npw + g];
        if (slot != kDepNone) {
            why = "two nodes of one chain group share a step";
            return false;
        }
        slot = v;
        dep_blocks(rp, ci, v, nlow[v], bwd, deps, cnt);
        if (cnt > (uint32_t)kChainDepBlk) {
            why = "too many dependency blocks";
            return false;
        }
        for (uint32_t k = 0; k < cnt; ++k) {
            const uint32_t m = deps[k];
            uint32_t w;
            if (bnd[m] == b) {
                if (st[m] >= st[v]) {
                    why = "dependency inside a bundle is not earlier";
                    return false;
                }
                w = st[v] - st[m] < (uint32_t)kChainRing
                        ? kDepLocal | ((st[m] % kChainRing) * (uint32_t)npw + grp[m])
                        : m;  // far back in this bundle: already published
            } else {
                if (pos[bnd[m]] >= pos[b]) {
                    why = "dependency on a later bundle";
                    return false;
                }
                w = m;
            }
            const int lane = (int)g * lpn + (int)(k % lpn), j = (int)(k / lpn);
            H.dep[(qs * bpl + j) * 64 + lane] = w;
        }
    }
    return true;
}

}  // namespace

bool plan_chain_schedule(const std::vector<uint32_t>& rp, const std::vector<uint32_t>& ci, int lpn, ChainPlan& P,
                         std::string& why) {
    if (lpn != 4 && lpn != 8) {
        why = "lanes per node must be 4 or 8";
        return false;
    }
    const uint32_t n = (uint32_t)(rp.size() - 1);
    if (n % 3 != 0 || n == 0 || n / 3 >= kDepLocal) {
        why = "not a 3-row node-blocked matrix";
        return false;
    }
    const uint32_t nn = n / 3;
    std::vector<uint16_t> nlow(nn);
    for (uint32_t v = 0; v < nn; ++v) {
        const uint32_t b = rp[3 * v], e = rp[3 * v + 1];
        uint32_t k = b;
        while (k < e && ci[k] < 3 * v) ++k;
        if (k == e || ci[k] != 3 * v || (k - b) % 3 != 0) {
            why = "row without a stored diagonal block";
            return false;
        }
        const uint32_t nl = (k - b) / 3, nu = (e - b) / 3 - nl - 1;
        if (nl > (uint32_t)kChainDepBlk || nu > (uint32_t)kChainDepBlk) {
            why = "more than 7 dependency blocks in a row";
            return false;
        }
        nlow[v] = (uint16_t)nl;
    }
    // chains; the backward sweep walks the same chains in reverse, which
    // needs v + 1 to be an upper neighbour of v exactly when v is a lower
    // neighbour of v + 1 (true for a structurally symmetric pattern)
    std::vector<uint32_t> cid(nn);
    uint32_t nch = 0;
    for (uint32_t v = 0; v < nn; ++v) {
        const bool cont = v > 0 && nlow[v] > 0 && ci[rp[3 * v] + 3 * (nlow[v] - 1)] / 3 == v - 1;
        cid[v] = cont ? cid[v - 1] : nch++;
        if (v > 0) {
            const uint32_t b = rp[3 * (v - 1)], nblk = (rp[3 * (v - 1) + 1] - b) / 3;
            const bool up = (uint32_t)nlow[v - 1] + 1 < nblk && ci[b + 3 * (nlow[v - 1] + 1)] / 3 == v;
            if (up != cont) {
                why = "pattern is not structurally symmetric";
                return false;
            }
        }
    }
    // chain grid stride: most frequent chain distance above one
    std::vector<uint32_t> hist(nch + 1, 0);
    for (uint32_t v = 0; v < nn; ++v)
        for (uint32_t k = 0; k < nlow[v]; ++k) {
            const uint32_t d = cid[v] - cid[ci[rp[3 * v] + 3 * k] / 3];
            if (d >= 2) hist[d]++;
        }
    uint32_t stride = nch, best = 0;
    for (uint32_t d = 2; d <= nch; ++d)
        if (hist[d] > best) {
            best = hist[d];
            stride = d;
        }
    const uint32_t nj = stride, nk = (nch + stride - 1) / stride;
    const uint32_t npw = 64 / (uint32_t)lpn;
    uint32_t gj = npw, gk = 1, cost = UINT32_MAX;
    for (uint32_t a = npw; a >= 1; a /= 2) {
        const uint32_t c = (nj + a - 1) / a + (nk + npw / a - 1) / (npw / a);
        if (c < cost) {
            cost = c;
            gj = a;
            gk = npw / a;
        }
    }
    const uint32_t nbj = (nj + gj - 1) / gj, nbk = (nk + gk - 1) / gk, nbund = nbj * nbk;
    std::vector<uint32_t> bnd(nn), grp(nn);
    for (uint32_t v = 0; v < nn; ++v) {
        const uint32_t c = cid[v], j = c % stride, k = c / stride;
        bnd[v] = (j / gj) + nbj * (k / gk);
        grp[v] = (j % gj) + gj * (k % gk);
    }
    P.fwd.nb = P.bwd.nb = nbund;
    if (!chain_direction(rp, ci, nlow, cid, bnd, grp, nbund, lpn, false, P.fwd, why) ||
        !chain_direction(rp, ci, nlow, cid, bnd, grp, nbund, lpn, true, P.bwd, why))
        return false;
    P.lpn = lpn;
    P.chains = nch;
    P.stride = stride;
    P.gj = gj;
    P.gk = gk;
    return true;
}

void encode_chain_records(const ChainPlanDir& H, int lpn, std::vector<uint32_t>& far) {
    const int npw = 64 / lpn, bpl = kChainMaxBlk / lpn;
    far.assign((size_t)H.nq * kRecWords, 0);
    for (size_t q = 0; q < H.nq; ++q) {
        uint32_t* f = &far[q * kRecWords];
        for (int g = 0; g < npw; ++g) {
            const uint32_t nd = H.node[q * npw + g];
            f[128 + g] = nd == kDepNone ? kOffNone : 12u * nd;
        }
        for (int lane = 0; lane < 64; ++lane)
            for (int j = 0; j < bpl; ++j) {
                const uint32_t dw = H.dep[(q * bpl + j) * 64 + lane];
                f[2 * lane + j] = dw == kDepNone ? kDepLocal | kZeroSlot : (dw & kDepLocal) ? dw : 12u * dw;
            }
    }
}

bool build_chain_schedule(const std::vector<uint32_t>& rp, const std::vector<uint32_t>& ci, int lpn,
                          ChainSchedule& C, hipStream_t stream, std::string& why, std::string& err) {
    C.ok = false;
    // the sweeps address the published vector with 32-bit buffer offsets
    const size_t nn = rp.empty() ? 0 : (rp.size() - 1) / 3;
    if (nn * 3 * sizeof(double) >= (size_t)kChainMaxBytes) {
        why = "vector too large for 32-bit buffer offsets";
        return false;
    }
    ChainPlan P;
    if (!plan_chain_schedule(rp, ci, lpn, P, why)) return false;
    const uint32_t nbund = P.fwd.nb;
    std::vector<uint32_t> far;  // device form of the plan (encode_chain_records)
    auto upload = [&](const ChainPlanDir& H, ChainSweep& W) {
        W.nb = nbund;
        W.nq = H.nq;
        encode_chain_records(H, P.lpn, far);
        if (W.bstart.alloc(nbund) != hipSuccess || W.bsteps.alloc(nbund) != hipSuccess ||
            W.far.alloc(far.size()) != hipSuccess)
            return false;
        return hipMemcpyAsync(W.bstart.p, H.bstart.data(), nbund * 4ull, hipMemcpyHostToDevice, stream) ==
                   hipSuccess &&
               hipMemcpyAsync(W.bsteps.p, H.bsteps.data(), nbund * 4ull, hipMemcpyHostToDevice, stream) ==
                   hipSuccess &&
               hipMemcpy(W.far.p, far.data(), far.size() * 4, hipMemcpyHostToDevice) == hipSuccess;
    };
    if (P.lpn != 4 || !upload(P.fwd, C.fwd) || !upload(P.bwd, C.bwd) ||
        hipStreamSynchronize(stream) != hipSuccess) {
        err = "upload of the chain schedule failed";
        return false;
    }
    C.lpn = P.lpn;
    C.chains = P.chains;
    C.stride = P.stride;
    C.gj = P.gj;
    C.gk = P.gk;
    C.ok = true;
    return true;
}

}  // namespace fm
