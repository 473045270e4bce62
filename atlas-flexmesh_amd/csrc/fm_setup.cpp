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

}  // namespace fm
