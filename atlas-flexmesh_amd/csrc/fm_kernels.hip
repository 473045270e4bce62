// fm_kernels.hip -- hand-written gfx950 kernels of the quasi-static Newton path.
//
// All of this is HBM- or latency-bound integer/FP gather-scatter work: no
// dense contraction exists, so MFMA is not used.  Kernels are written for
// 64-wide wavefronts; grid-wide kernels remap blockIdx so that each of the
// eight XCDs (own L2 each) works on one contiguous slice of rows / elements.
#include "fm_device.hpp"
#include "fm_internal.hpp"

#include <algorithm>
#include <cmath>
#include <cstdlib>
#include <mutex>
#include <string>

namespace fm {
namespace {

constexpr uint32_t kNoNode = 0xFFFFFFFFu;   // padding entry of a level-ordered queue

__device__ __forceinline__ uint32_t xcd_remap(uint32_t bid, uint32_t nb) {
    // workgroups are handed to XCDs round-robin; give XCD x a contiguous chunk
    constexpr uint32_t NX = 8;
    if (nb < NX) return bid;
    const uint32_t q = nb / NX, r = nb % NX;
    const uint32_t x = bid % NX, loc = bid / NX;
    return (x < r ? x * (q + 1) : r * (q + 1) + (x - r) * q) + loc;
}

// fixed-order block sum (blockDim.x a multiple of 64); result valid in thread 0
__device__ __forceinline__ double block_sum(double v) {
    __shared__ double s_part[16];
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o);
    const int w = threadIdx.x >> 6;
    if ((threadIdx.x & 63) == 0) s_part[w] = v;
    __syncthreads();
    double r = 0.0;
    if (threadIdx.x == 0)
        for (int i = 0; i < (int)(blockDim.x >> 6); ++i) r += s_part[i];
    return r;
}
// the same sum, with wave w's partial in the caller's LDS at part[w * stride]
// (a slot the wave no longer reads), so the kernel needs no LDS beyond its own
__device__ __forceinline__ double block_sum(double v, double* part, int stride) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o);
    const int w = threadIdx.x >> 6;
    if ((threadIdx.x & 63) == 0) part[w * stride] = v;
    __syncthreads();
    double r = 0.0;
    if (threadIdx.x == 0)
        for (int i = 0; i < (int)(blockDim.x >> 6); ++i) r += part[i * stride];
    return r;
}

// ---------------------------------------------------------------------------
// element evaluation (FP64): geometry, kinematics, force and tangent factors
// ---------------------------------------------------------------------------
struct Elem {
    uint32_t node[4];
    double g[4][3];   // shape-function gradients
    double a[4][3];   // F^{-T} g
    double Fg[4][3];  // F g
    double V, lnJ, mu, lam;
    bool clamped, ok;
};

// a + b with the product that forms b already rounded: no fused multiply-add
// across the accumulation, so a sum of element terms has the same bits whether
// the terms are added straight into memory (colour launches) or first stored
// and then added (node-gather kernel)
template <typename T>
__device__ __forceinline__ T add_rounded(T a, T b) {
#pragma clang fp contract(off)
    return a + b;
}

__device__ __forceinline__ double det3(const double* A) {
    return A[0] * (A[4] * A[8] - A[7] * A[5]) - A[3] * (A[1] * A[8] - A[7] * A[2]) +
           A[6] * (A[1] * A[5] - A[4] * A[2]);
}

// F, ln J with the clamp J >= 0.01, F^{-T} (with the C^{-1} guard of the
// reference inverse carried to F^{-T} = F C^{-1} when |det F| < 1e-15)
__device__ __forceinline__ void kinematics(const double* F, double* FiT, double& lnJ, bool& clamped) {
    const double det = det3(F);
    const double J = fabs(det);
    if (J < 1e-15) {
        for (int k = 0; k < 9; ++k) FiT[k] = 1e30 * F[k];
    } else {
        const double id = 1.0 / det;
        FiT[0] = (F[4] * F[8] - F[5] * F[7]) * id;
        FiT[1] = (F[5] * F[6] - F[3] * F[8]) * id;
        FiT[2] = (F[3] * F[7] - F[4] * F[6]) * id;
        FiT[3] = (F[2] * F[7] - F[1] * F[8]) * id;
        FiT[4] = (F[0] * F[8] - F[2] * F[6]) * id;
        FiT[5] = (F[1] * F[6] - F[0] * F[7]) * id;
        FiT[6] = (F[1] * F[5] - F[2] * F[4]) * id;
        FiT[7] = (F[2] * F[3] - F[0] * F[5]) * id;
        FiT[8] = (F[0] * F[4] - F[1] * F[3]) * id;
    }
    clamped = J < 0.01;
    lnJ = log(fmax(J, 0.01));
}

__device__ __forceinline__ void elem_eval(const ElemArgs& A, uint32_t t, Elem& E, double* Fout = nullptr) {
#pragma unroll
    for (int q = 0; q < 4; ++q) E.node[q] = A.conn[(size_t)q * A.ne + t];
    double Xe[4][3];
#pragma unroll
    for (int q = 0; q < 4; ++q)
#pragma unroll
        for (int d = 0; d < 3; ++d) Xe[q][d] = A.X[(size_t)d * A.nn + E.node[q]];
    double J[9];
#pragma unroll
    for (int d = 0; d < 3; ++d) {
        J[d * 3 + 0] = Xe[1][d] - Xe[0][d];
        J[d * 3 + 1] = Xe[2][d] - Xe[0][d];
        J[d * 3 + 2] = Xe[3][d] - Xe[0][d];
    }
    const double det = det3(J);
    E.V = fabs(det) / 6.0;
    E.ok = E.V >= 1e-30;
    E.mu = A.mu ? A.mu[t] : 0.0;
    E.lam = A.lam ? A.lam[t] : 0.0;
    if (!E.ok) return;
    const double id = 1.0 / det;
    // J^{-1} rows (cofactor form of matrix_3x3_inv); g_q = J^{-T} dN_xi,q
    double Ji[9];
    Ji[0] = (J[4] * J[8] - J[7] * J[5]) * id;
    Ji[1] = -(J[1] * J[8] - J[7] * J[2]) * id;
    Ji[2] = (J[1] * J[5] - J[4] * J[2]) * id;
    Ji[3] = -(J[3] * J[8] - J[6] * J[5]) * id;
    Ji[4] = (J[0] * J[8] - J[6] * J[2]) * id;
    Ji[5] = -(J[0] * J[5] - J[3] * J[2]) * id;
    Ji[6] = (J[3] * J[7] - J[6] * J[4]) * id;
    Ji[7] = -(J[0] * J[7] - J[6] * J[1]) * id;
    Ji[8] = (J[0] * J[4] - J[3] * J[1]) * id;
#pragma unroll
    for (int i = 0; i < 3; ++i) {
        E.g[1][i] = Ji[0 * 3 + i];
        E.g[2][i] = Ji[1 * 3 + i];
        E.g[3][i] = Ji[2 * 3 + i];
        E.g[0][i] = -(Ji[0 * 3 + i] + Ji[1 * 3 + i] + Ji[2 * 3 + i]);
    }
    double F[9];
#pragma unroll
    for (int i = 0; i < 3; ++i)
#pragma unroll
        for (int j = 0; j < 3; ++j) F[i * 3 + j] = (i == j) ? 1.0 : 0.0;
    if (A.u) {
#pragma unroll
        for (int q = 0; q < 4; ++q) {
            const double* uq = A.u + 3 * (size_t)E.node[q];
            const double u0 = uq[0], u1 = uq[1], u2 = uq[2];
#pragma unroll
            for (int j = 0; j < 3; ++j) {
                F[0 * 3 + j] += u0 * E.g[q][j];
                F[1 * 3 + j] += u1 * E.g[q][j];
                F[2 * 3 + j] += u2 * E.g[q][j];
            }
        }
    }
    if (Fout)
        for (int k = 0; k < 9; ++k) Fout[k] = F[k];
    double FiT[9];
    kinematics(F, FiT, E.lnJ, E.clamped);
#pragma unroll
    for (int q = 0; q < 4; ++q)
#pragma unroll
        for (int i = 0; i < 3; ++i) {
            E.a[q][i] = FiT[i * 3 + 0] * E.g[q][0] + FiT[i * 3 + 1] * E.g[q][1] + FiT[i * 3 + 2] * E.g[q][2];
            E.Fg[q][i] = F[i * 3 + 0] * E.g[q][0] + F[i * 3 + 1] * E.g[q][1] + F[i * 3 + 2] * E.g[q][2];
        }
}

// Colour-by-colour element kernel: one element per lane computes f_e
// (FP64, scattered straight away) and the factors of its 12x12 tangent
//   K_ab,ik = V [mu (g_a.g_b) d_ik + lam' a_a,i a_b,k + c a_a,k a_b,i],
// which are staged in LDS; the wave then writes the 144 entries of each of
// its elements with consecutive lanes on consecutive CSR positions.  No two
// elements of one colour share a node, so no atomics are needed.
template <typename T, bool WITH_K>
__global__ __launch_bounds__(64) void assemble_color_kernel(ElemArgs A, uint32_t begin, uint32_t end,
                                                            double* __restrict__ f, T* __restrict__ K,
                                                            const uint32_t* __restrict__ row_ptr,
                                                            const uint8_t* __restrict__ isdir,
                                                            const double* __restrict__ dD,
                                                            double* __restrict__ lift) {
    const uint32_t blk = xcd_remap(blockIdx.x, gridDim.x);
    const uint32_t t = begin + blk * 64 + threadIdx.x;
    const int lane = threadIdx.x;
    Elem E;
    E.ok = false;
    if (t < end) elem_eval(A, t, E);
    if (E.ok) {
        const double s = E.lam * E.lnJ - E.mu;
#pragma unroll
        for (int q = 0; q < 4; ++q) {
            double* fq = f + 3 * (size_t)E.node[q];
#pragma unroll
            for (int i = 0; i < 3; ++i) fq[i] = add_rounded(fq[i], E.V * (E.mu * E.Fg[q][i] + s * E.a[q][i]));
        }
    }
    if constexpr (WITH_K) {
        __shared__ T sV[64], smu[64], slp[64], sc[64];
        __shared__ T sg[64][12], sa[64][12];
        __shared__ uint32_t srp[64][12];
        __shared__ uint8_t srank[64][16];
        __shared__ uint8_t sdir[64][12];
        __shared__ int sok[64];
        const double lp = E.clamped ? 0.0 : E.lam;
        const double c = E.mu - E.lam * E.lnJ;
        sok[lane] = E.ok ? 1 : 0;
        if (E.ok) {
            sV[lane] = (T)E.V;
            smu[lane] = (T)E.mu;
            slp[lane] = (T)lp;
            sc[lane] = (T)c;
#pragma unroll
            for (int q = 0; q < 4; ++q)
#pragma unroll
                for (int i = 0; i < 3; ++i) {
                    sg[lane][q * 3 + i] = (T)E.g[q][i];
                    sa[lane][q * 3 + i] = (T)E.a[q][i];
                    srp[lane][q * 3 + i] = row_ptr[3 * E.node[q] + i];
                    sdir[lane][q * 3 + i] = isdir ? isdir[3 * E.node[q] + i] : 0;
                }
#pragma unroll
            for (int k = 0; k < 16; ++k) srank[lane][k] = A.erank[(size_t)k * A.ne + t];
            if (dD) {
                // K_e dD_e with the un-eliminated element matrix (FP64): the
                // right-hand-side correction of a prescribed increment
                double dq[4][3];
                bool any = false;
#pragma unroll
                for (int q = 0; q < 4; ++q)
#pragma unroll
                    for (int i = 0; i < 3; ++i) {
                        dq[q][i] = dD[3 * (size_t)E.node[q] + i];
                        any |= dq[q][i] != 0.0;
                    }
                if (any) {
#pragma unroll
                    for (int q = 0; q < 4; ++q) {
                        double out[3] = {0, 0, 0};
#pragma unroll
                        for (int b = 0; b < 4; ++b) {
                            const double gg = E.g[q][0] * E.g[b][0] + E.g[q][1] * E.g[b][1] + E.g[q][2] * E.g[b][2];
                            const double ad = E.a[b][0] * dq[b][0] + E.a[b][1] * dq[b][1] + E.a[b][2] * dq[b][2];
                            const double qd = E.a[q][0] * dq[b][0] + E.a[q][1] * dq[b][1] + E.a[q][2] * dq[b][2];
#pragma unroll
                            for (int i = 0; i < 3; ++i) out[i] += E.mu * gg * dq[b][i] + lp * E.a[q][i] * ad + c * E.a[b][i] * qd;
                        }
                        double* lq = lift + 3 * (size_t)E.node[q];
#pragma unroll
                        for (int i = 0; i < 3; ++i) lq[i] = add_rounded(lq[i], E.V * out[i]);
                    }
                }
            }
        }
        __syncthreads();
        for (int el = 0; el < 64; ++el) {
            if (!sok[el]) continue;
            const T V = sV[el], mu = smu[el], lpe = slp[el], ce = sc[el];
            for (int j = lane; j < 144; j += 64) {
                // j = ((q*3 + i)*4 + b)*3 + k
                const int k = j % 3;
                const int b = (j / 3) & 3;
                const int qi = j / 12;
                const int q = qi / 3, i = qi % 3;
                if (sdir[el][qi] | sdir[el][b * 3 + k]) continue;
                const T gg = sg[el][q * 3 + 0] * sg[el][b * 3 + 0] + sg[el][q * 3 + 1] * sg[el][b * 3 + 1] +
                             sg[el][q * 3 + 2] * sg[el][b * 3 + 2];
                T v = lpe * sa[el][q * 3 + i] * sa[el][b * 3 + k] + ce * sa[el][q * 3 + k] * sa[el][b * 3 + i];
                if (i == k) v += mu * gg;
                T* Kp = K + srp[el][qi] + 3u * srank[el][q * 4 + b] + k;
                *Kp = add_rounded(*Kp, V * v);
            }
        }
    }
}

// Node-gather tangent assembly: two nodes per wave, one per half.  Lane l
// of a half evaluates its node's l-th incident element (colour order) and
// the contributions to the node's three rows: the force, the lift term and
// the 3x12 tangent rows
//   K_qb,ik = V [mu (g_q.g_b) d_ik + lam' a_q,i a_b,k + c a_q,k a_b,i]
// (the arithmetic of the colour kernel).  The half then adds them into an
// LDS copy of the node's CSR rows in colour order, the order in which the
// colour launches add them, and writes the rows once, contiguous: no fill
// of K or f and no read-back of K entries.
template <typename T>
__global__ __launch_bounds__(64) void assemble_node_kernel(ElemArgs A, const uint32_t* __restrict__ n2e_ptr,
                                                           const uint32_t* __restrict__ n2e, double* __restrict__ f,
                                                           T* __restrict__ K, const uint32_t* __restrict__ row_ptr,
                                                           const uint32_t* __restrict__ col_idx,
                                                           const uint8_t* __restrict__ isdir,
                                                           const double* __restrict__ dD, double* __restrict__ lift) {
    constexpr uint32_t H = kAsmMaxElems;  // lanes per node (a half wave)
    static_assert(2 * H == 64, "two nodes per wave");
    const uint32_t blk = xcd_remap(blockIdx.x, gridDim.x);
    const int h = threadIdx.x / H, hl = threadIdx.x % H;
    const uint32_t v = 2 * blk + h;
    const bool have = v < A.nn;
    const uint32_t e0 = have ? n2e_ptr[v] : 0, nev = have ? n2e_ptr[v + 1] - e0 : 0;  // <= H
    __shared__ T sK[2][H][36];
    __shared__ uint8_t srk[2][H][4];
    __shared__ double sf[2][H][3], sl[2][H][3];
    __shared__ T srow[2][3 * kAsmMaxBlk * 3];
    Elem E;
    E.ok = false;
    uint32_t t = 0;
    int q = 0;
    if ((uint32_t)hl < nev) {
        const uint32_t w = n2e[e0 + hl];
        t = w >> 2;
        q = (int)(w & 3u);
        elem_eval(A, t, E);
        // the node's own vertex q of the element, selected without a run-time
        // index into the element arrays (which would move them to scratch)
        double gq[3], aq[3], fgq[3];
#pragma unroll
        for (int i = 0; i < 3; ++i) {
            gq[i] = E.g[0][i];
            aq[i] = E.a[0][i];
            fgq[i] = E.Fg[0][i];
        }
#pragma unroll
        for (int qq = 1; qq < 4; ++qq)
            if (q == qq)
#pragma unroll
                for (int i = 0; i < 3; ++i) {
                    gq[i] = E.g[qq][i];
                    aq[i] = E.a[qq][i];
                    fgq[i] = E.Fg[qq][i];
                }
        const double s = E.lam * E.lnJ - E.mu;
#pragma unroll
        for (int i = 0; i < 3; ++i) sf[h][hl][i] = E.ok ? E.V * (E.mu * fgq[i] + s * aq[i]) : 0.0;
        const double lp = E.clamped ? 0.0 : E.lam;
        const double c = E.mu - E.lam * E.lnJ;
        const T V = (T)E.V, mu = (T)E.mu, lpe = (T)lp, ce = (T)c;
#pragma unroll
        for (int b = 0; b < 4; ++b) {
            srk[h][hl][b] = A.erank[(size_t)(q * 4 + b) * A.ne + t];
            const T gg = (T)gq[0] * (T)E.g[b][0] + (T)gq[1] * (T)E.g[b][1] + (T)gq[2] * (T)E.g[b][2];
#pragma unroll
            for (int i = 0; i < 3; ++i)
#pragma unroll
                for (int k = 0; k < 3; ++k) {
                    T val = lpe * (T)aq[i] * (T)E.a[b][k] + ce * (T)aq[k] * (T)E.a[b][i];
                    if (i == k) val += mu * gg;
                    sK[h][hl][(i * 4 + b) * 3 + k] = E.ok ? V * val : (T)0;
                }
        }
        double out[3] = {0, 0, 0};
        if (dD && E.ok) {
            // K_e dD_e with the un-eliminated element matrix (FP64), row block q
            double dq[4][3];
#pragma unroll
            for (int b = 0; b < 4; ++b)
#pragma unroll
                for (int i = 0; i < 3; ++i) dq[b][i] = dD[3 * (size_t)E.node[b] + i];
#pragma unroll
            for (int b = 0; b < 4; ++b) {
                const double gg = gq[0] * E.g[b][0] + gq[1] * E.g[b][1] + gq[2] * E.g[b][2];
                const double ad = E.a[b][0] * dq[b][0] + E.a[b][1] * dq[b][1] + E.a[b][2] * dq[b][2];
                const double qd = aq[0] * dq[b][0] + aq[1] * dq[b][1] + aq[2] * dq[b][2];
#pragma unroll
                for (int i = 0; i < 3; ++i) out[i] += E.mu * gg * dq[b][i] + lp * aq[i] * ad + c * E.a[b][i] * qd;
            }
        }
#pragma unroll
        for (int i = 0; i < 3; ++i) sl[h][hl][i] = E.V * out[i];
    }
    const uint32_t r0 = have ? row_ptr[3 * v] : 0, nbv = have ? (row_ptr[3 * v + 1] - r0) / 3 : 0;  // <= kAsmMaxBlk
    const uint32_t nrow = 3 * nbv, nent = 3 * nrow;
    for (uint32_t j = hl; j < nent; j += H) srow[h][j] = (T)0;
    __syncthreads();
    // force and lift: sums over the elements in colour order
    if (hl < 3 && have) {
        double sum = 0.0;
        for (uint32_t l = 0; l < nev; ++l) sum += sf[h][l][hl];
        f[3 * (size_t)v + hl] = sum;
    } else if (dD && hl >= 3 && hl < 6 && have) {
        double sum = 0.0;
        for (uint32_t l = 0; l < nev; ++l) sum += sl[h][l][hl - 3];
        lift[3 * (size_t)v + hl - 3] = sum;
    }
    // entry (i, b, k) of element l goes to row i, block rank_b, component k;
    // lane hl takes entries hl and hl + 32 (< 36) of each element, element by
    // element, so each row entry receives its contributions in colour order
    // (the wave's LDS accesses take effect in program order)
    {
        const int ea = hl, eb = hl + (int)H;
        const int ia = ea / 12, ba = (ea / 3) & 3, ka = ea % 3;
        const int ib = eb / 12, bb = (eb / 3) & 3, kb = eb % 3;
        for (uint32_t l = 0; l < nev; ++l) {
            srow[h][ia * nrow + 3u * srk[h][l][ba] + ka] += sK[h][l][ea];
            if (eb < 36) srow[h][ib * nrow + 3u * srk[h][l][bb] + kb] += sK[h][l][eb];
        }
    }
    __syncthreads();
    for (uint32_t j = hl; j < nent; j += H) {
        const uint32_t i = j / nrow, rem = j - i * nrow;
        const uint32_t pos = row_ptr[3 * v + i] + rem;
        T val = srow[h][j];
        if (isdir && (isdir[3 * v + i] || isdir[col_idx[pos]])) val = (T)0;
        K[pos] = val;
    }
}

template <typename T>
__global__ void set_dir_diag_kernel(uint32_t n_dir, const uint32_t* __restrict__ dofs,
                                    const uint32_t* __restrict__ row_ptr, const uint16_t* __restrict__ nlow,
                                    T* __restrict__ K) {
    const uint32_t k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= n_dir) return;
    const uint32_t r = dofs[k];
    const uint32_t node = r / 3, d = r % 3;
    K[row_ptr[r] + 3u * nlow[node] + d] = (T)1;
}

// Dirichlet rows and columns of an assembled matrix: kEliminateLanes
// consecutive lanes walk one row, so a wave reads whole runs of col_idx
// (one lane per row read each row's own cache lines: ~16x the bytes)
constexpr int kEliminateLanes = 32;
template <typename T>
__global__ __launch_bounds__(256) void eliminate_kernel(uint32_t n, const uint32_t* __restrict__ row_ptr,
                                                        const uint32_t* __restrict__ col_idx,
                                                        const uint8_t* __restrict__ isdir, T* __restrict__ K) {
    const uint32_t r = blockIdx.x * (256 / kEliminateLanes) + threadIdx.x / kEliminateLanes;
    if (r >= n) return;
    const bool dr = isdir[r];
    const uint32_t e = row_ptr[r + 1];
    for (uint32_t k = row_ptr[r] + threadIdx.x % kEliminateLanes; k < e; k += kEliminateLanes) {
        const uint32_t c = col_idx[k];
        if (dr || isdir[c]) K[k] = (c == r) ? (T)1 : (T)0;
    }
}

// ---------------------------------------------------------------------------
// CSR SpMV: LPR lanes per row, fixed-order lane reduction; optional fused
// x.y partial per block and sentinel preset of the solve outputs
// ---------------------------------------------------------------------------
template <typename T, int LPR, bool DOT>
__global__ __launch_bounds__(256) void spmv_kernel(uint32_t n, const uint32_t* __restrict__ rp,
                                                   const uint32_t* __restrict__ ci, const T* __restrict__ val,
                                                   const T* __restrict__ x, T* __restrict__ y,
                                                   double* __restrict__ partials, T* __restrict__ sent0,
                                                   T* __restrict__ sent1, const PcgState* __restrict__ st) {
    if (st && st->done) return;
    constexpr int RPB = 256 / LPR;
    const uint32_t blk = xcd_remap(blockIdx.x, gridDim.x);
    const uint32_t row = blk * RPB + threadIdx.x / LPR;
    const int l = threadIdx.x % LPR;
    T sum = 0;
    if (row < n) {
        // UF entries per lane are loaded before any x gather, so a wave keeps
        // UF times more bytes in flight (the kernel is bound by memory-level
        // parallelism, not by instruction count)
        constexpr int UF = 4;
        const uint32_t e = rp[row + 1];
        for (uint32_t k0 = rp[row] + l; k0 < e; k0 += LPR * UF) {
            T v[UF];
            uint32_t c[UF];
#pragma unroll
            for (int u = 0; u < UF; ++u) {
                const uint32_t k = k0 + u * LPR;
                v[u] = k < e ? __builtin_nontemporal_load(val + k) : (T)0;
                c[u] = k < e ? __builtin_nontemporal_load(ci + k) : row;
            }
#pragma unroll
            for (int u = 0; u < UF; ++u) sum += v[u] * x[c[u]];
        }
    }
#pragma unroll
    for (int o = LPR / 2; o > 0; o >>= 1) sum += __shfl_xor(sum, o);
    double part = 0.0;
    if (row < n && l == 0) {
        y[row] = sum;
        if (sent0) {
            sent0[row] = sentinel_value<T>();
            sent1[row] = sentinel_value<T>();
        }
        if (DOT) part = (double)sum * (double)x[row];
    }
    if (DOT) {
        const double bs = block_sum(part);
        if (threadIdx.x == 0) partials[blk] = bs;
    }
}

// CSR-stream SpMV: a block owns a tile of TR consecutive rows, streams the
// tile's contiguous nonzeros with every load of a pass issued before any
// gather (U per thread), stages the products in LDS and reduces one row per
// thread.  Tiles whose nonzeros exceed the LDS stage fall back to lanes-per-row.
constexpr int kTR = 64;          // rows per tile
constexpr int kStage = 64 * kTR; // products staged per tile

template <typename T, bool DOT>
__global__ __launch_bounds__(256) void spmv_stream_kernel(uint32_t n, const uint32_t* __restrict__ rp,
                                                          const uint32_t* __restrict__ ci, const T* __restrict__ val,
                                                          const T* __restrict__ x, T* __restrict__ y,
                                                          double* __restrict__ partials, T* __restrict__ sent0,
                                                          T* __restrict__ sent1, const PcgState* __restrict__ st) {
    if (st && st->done) return;
    __shared__ uint32_t srp[kTR + 1];
    __shared__ T sprod[kStage];
    const uint32_t tile = xcd_remap(blockIdx.x, gridDim.x);
    const uint32_t r0 = tile * kTR;
    const uint32_t nr = min((uint32_t)kTR, n - r0);
    const int tid = threadIdx.x;
    if (tid <= (int)nr) srp[tid] = rp[r0 + tid];
    __syncthreads();
    const uint32_t base = srp[0];
    const uint32_t cnt = srp[nr] - base;
    T sum = 0;
    if (cnt <= (uint32_t)kStage) {
        constexpr int U = 8;
        for (uint32_t k0 = tid; k0 < cnt; k0 += 256 * U) {
            T v[U];
            uint32_t c[U];
#pragma unroll
            for (int u = 0; u < U; ++u) {
                const uint32_t k = k0 + u * 256;
                v[u] = k < cnt ? __builtin_nontemporal_load(val + base + k) : (T)0;
                c[u] = k < cnt ? __builtin_nontemporal_load(ci + base + k) : 0u;
            }
#pragma unroll
            for (int u = 0; u < U; ++u) {
                const uint32_t k = k0 + u * 256;
                if (k < cnt) sprod[k] = v[u] * x[c[u]];
            }
        }
        __syncthreads();
        if (tid < (int)nr)
            for (uint32_t k = srp[tid] - base; k < srp[tid + 1] - base; ++k) sum += sprod[k];
    } else if (tid < (int)nr) {
        for (uint32_t k = srp[tid]; k < srp[tid + 1]; ++k) sum += val[k] * x[ci[k]];
    }
    double part = 0.0;
    if (tid < (int)nr) {
        const uint32_t row = r0 + tid;
        y[row] = sum;
        if (sent0) {
            sent0[row] = sentinel_value<T>();
            sent1[row] = sentinel_value<T>();
        }
        if (DOT) part = (double)sum * (double)x[row];
    }
    if (DOT) {
        const double bs = block_sum(part);
        if (tid == 0) partials[tile] = bs;
    }
}

// Node-blocked SpMV of a 3-row node-blocked matrix: one column-node index
// per 3x3 block (nptr/ncol) instead of one column index per entry, which is
// a third of the CSR bytes at 4-byte values.  16 lanes per node: lane k takes
// blocks k, k + 16, ...; each block is a 3-value gather of x and the three
// contiguous 3-value segments of the node's rows, which start at
// row_ptr[0] + 9 * nptr[node] (the three rows of a node have equal length).
// Each 16-lane group takes nodes_per_group nodes whose loads are all issued
// before any is used, so every wave has that many times the bytes in flight
// per round trip; lanes past a node's last block re-read a valid block with
// weight zero.  Fixed-order reduction (DPP within the 16-lane row).
constexpr int kNodeLanes = 16;
#ifndef FM_NPG32
#define FM_NPG32 2
#endif
#ifndef FM_NPG64
#define FM_NPG64 1
#endif
template <typename T>
constexpr int nodes_per_group() {
    return sizeof(T) == 4 ? FM_NPG32 : FM_NPG64;  // FP64: two nodes per group measured slower
}

template <int CTRL>
__device__ __forceinline__ float row_dpp(float v) {
    return __builtin_bit_cast(float, __builtin_amdgcn_update_dpp(0, __builtin_bit_cast(int, v), CTRL, 0xF, 0xF, false));
}
template <int CTRL>
__device__ __forceinline__ double row_dpp(double v) {
    const unsigned long long u = __builtin_bit_cast(unsigned long long, v);
    const unsigned lo = (unsigned)__builtin_amdgcn_update_dpp(0, (int)(unsigned)u, CTRL, 0xF, 0xF, false);
    const unsigned hi = (unsigned)__builtin_amdgcn_update_dpp(0, (int)(unsigned)(u >> 32), CTRL, 0xF, 0xF, false);
    return __builtin_bit_cast(double, (unsigned long long)lo | ((unsigned long long)hi << 32));
}
// sum over the 16 lanes of a DPP row, in every lane: rotations by 8 and 4,
// then quad permutes (xor 2, xor 1)
template <typename T>
__device__ __forceinline__ T row16_sum(T v) {
    v += row_dpp<0x128>(v);  // row_ror:8
    v += row_dpp<0x124>(v);  // row_ror:4
    v += row_dpp<0x4E>(v);   // quad_perm [2,3,0,1]
    v += row_dpp<0xB1>(v);   // quad_perm [1,0,3,2]
    return v;
}

#ifndef FM_SPMV_NT
#define FM_SPMV_NT 1
#endif
// matrix values and column nodes are streamed once per product
template <typename T>
__device__ __forceinline__ T spmv_stream_load(const T* p) {
    if constexpr (FM_SPMV_NT) return __builtin_nontemporal_load(p);
    return *p;
}

template <typename T, bool DOT>
__global__ __launch_bounds__(256) void spmv_node_kernel(uint32_t nn, const uint32_t* __restrict__ rp,
                                                        const uint32_t* __restrict__ nptr,
                                                        const uint32_t* __restrict__ ncol,
                                                        const T* __restrict__ val, const T* __restrict__ x,
                                                        T* __restrict__ y, double* __restrict__ partials,
                                                        T* __restrict__ sent0, T* __restrict__ sent1,
                                                        const PcgState* __restrict__ st) {
    if (st && st->done) return;
    constexpr int NPB = 256 / kNodeLanes, G = nodes_per_group<T>();
    const uint32_t blk = xcd_remap(blockIdx.x, gridDim.x);
    const int k = threadIdx.x % kNodeLanes;
    if constexpr (G == 1) {
        // one node per group: lanes past the node's last block skip the loop
        const uint32_t node = blk * NPB + threadIdx.x / kNodeLanes;
        T s0 = 0, s1 = 0, s2 = 0;
        if (node < nn) {
            const uint32_t b0 = nptr[node], nb = nptr[node + 1] - b0;
            const uint32_t r0 = rp[3 * node], r1 = rp[3 * node + 1], r2 = rp[3 * node + 2];
            for (uint32_t kk = k; kk < nb; kk += kNodeLanes) {
                const T* a0 = val + r0 + 3 * kk;
                const T* a1 = val + r1 + 3 * kk;
                const T* a2 = val + r2 + 3 * kk;
                T a[9];
#pragma unroll
                for (int c = 0; c < 3; ++c) {
                    a[c] = __builtin_nontemporal_load(a0 + c);
                    a[3 + c] = __builtin_nontemporal_load(a1 + c);
                    a[6 + c] = __builtin_nontemporal_load(a2 + c);
                }
                const T* xm = x + 3 * (size_t)__builtin_nontemporal_load(ncol + b0 + kk);
                const T x0 = xm[0], x1 = xm[1], x2 = xm[2];
                s0 += a[0] * x0 + a[1] * x1 + a[2] * x2;
                s1 += a[3] * x0 + a[4] * x1 + a[5] * x2;
                s2 += a[6] * x0 + a[7] * x1 + a[8] * x2;
            }
        }
#pragma unroll
        for (int o = kNodeLanes / 2; o > 0; o >>= 1) {
            s0 += __shfl_xor(s0, o);
            s1 += __shfl_xor(s1, o);
            s2 += __shfl_xor(s2, o);
        }
        double part = 0.0;
        if (node < nn && k < 3) {
            const uint32_t row = 3 * node + k;
            const T sum = k == 0 ? s0 : (k == 1 ? s1 : s2);
            y[row] = sum;
            if (sent0) {
                sent0[row] = sentinel_value<T>();
                sent1[row] = sentinel_value<T>();
            }
            if (DOT) part = (double)sum * (double)x[row];
        }
        if (DOT) {
            const double bs = block_sum(part);
            if (threadIdx.x == 0) partials[blk] = bs;
        }
        return;
    }
    const uint32_t rbase = rp[0];
    // three dependent rounds per wave: (1) the block ranges of its nodes and
    // the own-row x of the dot, (2) the column nodes, then the values, which
    // stay in flight while (3) the x gathers wait only for the column nodes
    // (loads complete in order).  Every load is unconditional, at a clamped
    // index: a load under a condition becomes a branch with a full wait.
    uint32_t node[G], b0[G], nb[G];
    T xown[G];
#pragma unroll
    for (int i = 0; i < G; ++i) {
        node[i] = (blk * G + i) * NPB + threadIdx.x / kNodeLanes;
        const uint32_t v = min(node[i], nn - 1);
        b0[i] = nptr[v];
        nb[i] = nptr[v + 1];
        if (DOT) xown[i] = x[3 * v + min(k, 2)];
    }
#pragma unroll
    for (int i = 0; i < G; ++i) nb[i] = node[i] < nn ? nb[i] - b0[i] : 0;
    // block k of each node (a valid block when k is past the end; weight zero)
    uint32_t kk[G], cn[G];
#pragma unroll
    for (int i = 0; i < G; ++i) {
        kk[i] = min((uint32_t)k, max(nb[i], 1u) - 1);
        cn[i] = spmv_stream_load(ncol + b0[i] + kk[i]);
    }
    T a[G][9];
#pragma unroll
    for (int i = 0; i < G; ++i) {
        const T* a0 = val + rbase + 9 * (size_t)b0[i] + 3 * kk[i];
        const T* a1 = a0 + 3 * nb[i];
        const T* a2 = a1 + 3 * nb[i];
#pragma unroll
        for (int c = 0; c < 3; ++c) {
            a[i][c] = spmv_stream_load(a0 + c);
            a[i][3 + c] = spmv_stream_load(a1 + c);
            a[i][6 + c] = spmv_stream_load(a2 + c);
        }
    }
    __builtin_amdgcn_sched_barrier(0);  // all value loads issue before the wait for the column nodes
    T s[G][3];
#pragma unroll
    for (int i = 0; i < G; ++i) {
        const T* xm = x + 3 * (size_t)cn[i];
        const T w = (uint32_t)k < nb[i] ? (T)1 : (T)0;
        const T x0 = xm[0], x1 = xm[1], x2 = xm[2];
        s[i][0] = (a[i][0] * x0 + a[i][1] * x1 + a[i][2] * x2) * w;
        s[i][1] = (a[i][3] * x0 + a[i][4] * x1 + a[i][5] * x2) * w;
        s[i][2] = (a[i][6] * x0 + a[i][7] * x1 + a[i][8] * x2) * w;
    }
    // nodes with more than 16 blocks
#pragma unroll
    for (int i = 0; i < G; ++i)
        for (uint32_t kk = k + kNodeLanes; kk < nb[i]; kk += kNodeLanes) {
            const T* a0 = val + rbase + 9 * (size_t)b0[i] + 3 * kk;
            const T* a1 = a0 + 3 * nb[i];
            const T* a2 = a1 + 3 * nb[i];
            const T* xm = x + 3 * (size_t)ncol[b0[i] + kk];
            const T x0 = xm[0], x1 = xm[1], x2 = xm[2];
            s[i][0] += a0[0] * x0 + a0[1] * x1 + a0[2] * x2;
            s[i][1] += a1[0] * x0 + a1[1] * x1 + a1[2] * x2;
            s[i][2] += a2[0] * x0 + a2[1] * x1 + a2[2] * x2;
        }
    double part = 0.0;
#pragma unroll
    for (int i = 0; i < G; ++i) {
        const T t0 = row16_sum(s[i][0]), t1 = row16_sum(s[i][1]), t2 = row16_sum(s[i][2]);
        if (node[i] < nn && k < 3) {
            const uint32_t row = 3 * node[i] + k;
            const T sum = k == 0 ? t0 : (k == 1 ? t1 : t2);
            y[row] = sum;
            if (sent0) {
                sent0[row] = sentinel_value<T>();
                sent1[row] = sentinel_value<T>();
            }
            if (DOT) part += (double)sum * (double)xown[i];
        }
    }
    if (DOT) {
        const double bs = block_sum(part);
        if (threadIdx.x == 0) partials[blk] = bs;
    }
}

// LDS-staged node SpMV.  A wave takes 4 G consecutive nodes, whose values
// are one contiguous run of the value array (node v's three rows start at
// row_ptr[0] + 9 nptr[v], node after node).  The wave streams the run into
// LDS with 16-byte loads of consecutive lanes, so every load instruction
// covers 1 KB of consecutive addresses and every cache line is read whole,
// once; lanes past the run reload its last chunk (the same line, no extra
// traffic).  The 12-byte row segments of the direct kernel touch several
// partial lines per instruction, and the L1 spends its cycles on pending
// misses.  Each 16-lane group then takes its node's blocks from LDS with the
// arithmetic of the direct kernel (the same bits).  A run longer than the LDS
// slice (more than ~17 blocks per node on average; a Kuhn box has at most 15)
// reads its values directly.  Five 16-byte rounds per lane leave 20 KB of LDS
// per block, so eight waves per SIMD stay resident: 285 us per FP32 product
// at 128^3 hexes against 295 us with six rounds and 301 us unstaged.
#ifndef FM_SPMV_LDS_ROUNDS
#define FM_SPMV_LDS_ROUNDS 5
#endif
template <typename T>
struct Vec16;
template <>
struct Vec16<float> {
    typedef float type __attribute__((ext_vector_type(4)));
};
template <>
struct Vec16<double> {
    typedef double type __attribute__((ext_vector_type(2)));
};

template <typename T, bool DOT>
__global__ __launch_bounds__(256) void spmv_node_lds_kernel(uint32_t nn, const uint32_t* __restrict__ rp,
                                                            const uint32_t* __restrict__ nptr,
                                                            const uint32_t* __restrict__ ncol,
                                                            const T* __restrict__ val, uint32_t nval,
                                                            const T* __restrict__ x, T* __restrict__ y,
                                                            double* __restrict__ partials, T* __restrict__ sent0,
                                                            T* __restrict__ sent1, const PcgState* __restrict__ st) {
    if (st && st->done) return;
    using V = typename Vec16<T>::type;
    constexpr int G = nodes_per_group<T>(), NPW = 4 * G, W = 16 / (int)sizeof(T);
    constexpr int ROUNDS = FM_SPMV_LDS_ROUNDS, CAP = ROUNDS * 64 * W;
    __shared__ V sv[4][ROUNDS * 64];
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63, q = lane >> 4, k = lane & 15;
    const uint32_t blk = xcd_remap(blockIdx.x, gridDim.x);
    const uint32_t v0 = (blk * 4 + wave) * NPW;  // first node of the wave
    const uint32_t rbase = rp[0];
    (void)nval;
    // Three dependent rounds: (1) the block ranges of the wave's nodes, lane j
    // holding nptr[v0 + j] (j <= NPW); (2) the column nodes, then the value
    // run; (3) the x gathers, which wait only for the column nodes (loads
    // complete in issue order) while the value run is still in flight.
    const uint32_t pl = nptr[min(v0 + min((uint32_t)lane, (uint32_t)NPW), nn)];
    uint32_t node[G];
    T xown[G];
#pragma unroll
    for (int i = 0; i < G; ++i) {
        node[i] = v0 + 4 * i + q;
        if (DOT) xown[i] = x[3 * min(node[i], nn - 1) + min(k, 2)];
    }
    // the wave's value run [r0, r1), 16-byte chunks from base
    const uint32_t r0 = rbase + 9 * (uint32_t)__builtin_amdgcn_readlane((int)pl, 0);
    const uint32_t r1 = rbase + 9 * (uint32_t)__builtin_amdgcn_readlane((int)pl, NPW);
    const uint32_t base = r0 & ~(uint32_t)(W - 1);
    const bool staged = r1 - base <= (uint32_t)CAP;  // wave-uniform
    const uint32_t last = r1 > base ? (r1 - 1) & ~(uint32_t)(W - 1) : base;  // chunk holding the run's last value
    // group q takes nodes v0 + 4 i + q; a node past the end gets no blocks
    // and reads block 0 of the column list and the run's first values
    uint32_t b0[G], off[G], nb[G], cn[G];
#pragma unroll
    for (int i = 0; i < G; ++i) {
        const uint32_t bb = __shfl(pl, 4 * i + q), e = __shfl(pl, 4 * i + q + 1);
        const bool ok = node[i] < nn;
        nb[i] = ok ? e - bb : 0;
        b0[i] = ok ? bb : 0;
        const uint32_t kk = min((uint32_t)k, max(nb[i], 1u) - 1);
        off[i] = ok ? rbase + 9 * bb + 3 * kk : base;
        cn[i] = spmv_stream_load(ncol + b0[i] + kk);
    }
    __builtin_amdgcn_sched_barrier(0);  // column nodes issue before the value run
    V t[ROUNDS];
#pragma unroll
    for (int r = 0; r < ROUNDS; ++r) {
        const uint32_t c = min(base + (uint32_t)(r * 64 + lane) * W, last);
        t[r] = __builtin_nontemporal_load(reinterpret_cast<const V*>(val + c));
    }
    __builtin_amdgcn_sched_barrier(0);  // x gathers issue after the value run
    T xg[G][3];
#pragma unroll
    for (int i = 0; i < G; ++i) {
        const T* xm = x + 3 * (size_t)cn[i];
        xg[i][0] = xm[0];
        xg[i][1] = xm[1];
        xg[i][2] = xm[2];
    }
    __builtin_amdgcn_sched_barrier(0);
#pragma unroll
    for (int r = 0; r < ROUNDS; ++r) sv[wave][r * 64 + lane] = t[r];
    __syncthreads();
    // The block products read the values through `av`.  The staged and the
    // unstaged run are two copies of the same code (the choice is
    // wave-uniform), so the staged copy reads LDS with ds_read: one pointer
    // selected between LDS and global memory compiles to generic (flat)
    // loads, which take the vector-memory path of the global loads (FP64
    // product at 128^3 hexes: 571 -> 496 us).
    T s[G][3];
    auto products = [&](auto av) {
#pragma unroll
        for (int i = 0; i < G; ++i) {
            const uint32_t o0 = off[i], o1 = o0 + 3 * nb[i], o2 = o1 + 3 * nb[i];
            T a[9];
#pragma unroll
            for (int c = 0; c < 3; ++c) {
                a[c] = av(o0 + c);
                a[3 + c] = av(o1 + c);
                a[6 + c] = av(o2 + c);
            }
            const T w = (uint32_t)k < nb[i] ? (T)1 : (T)0;
            const T x0 = xg[i][0], x1 = xg[i][1], x2 = xg[i][2];
            s[i][0] = (a[0] * x0 + a[1] * x1 + a[2] * x2) * w;
            s[i][1] = (a[3] * x0 + a[4] * x1 + a[5] * x2) * w;
            s[i][2] = (a[6] * x0 + a[7] * x1 + a[8] * x2) * w;
        }
        // nodes with more than 16 blocks
#pragma unroll
        for (int i = 0; i < G; ++i)
            for (uint32_t kb = k + kNodeLanes; kb < nb[i]; kb += kNodeLanes) {
                const uint32_t o0 = rbase + 9 * b0[i] + 3 * kb, o1 = o0 + 3 * nb[i], o2 = o1 + 3 * nb[i];
                const T* xm = x + 3 * (size_t)ncol[b0[i] + kb];
                const T x0 = xm[0], x1 = xm[1], x2 = xm[2];
                s[i][0] += av(o0) * x0 + av(o0 + 1) * x1 + av(o0 + 2) * x2;
                s[i][1] += av(o1) * x0 + av(o1 + 1) * x1 + av(o1 + 2) * x2;
                s[i][2] += av(o2) * x0 + av(o2 + 1) * x1 + av(o2 + 2) * x2;
            }
    };
    if (staged) {
        const T* lv = reinterpret_cast<const T*>(&sv[wave][0]);
        products([&](uint32_t o) { return lv[o - base]; });
    } else {
        products([&](uint32_t o) { return val[o]; });
    }
    double part = 0.0;
#pragma unroll
    for (int i = 0; i < G; ++i) {
        const T t0 = row16_sum(s[i][0]), t1 = row16_sum(s[i][1]), t2 = row16_sum(s[i][2]);
        if (node[i] < nn && k < 3) {
            const uint32_t row = 3 * node[i] + k;
            const T sum = k == 0 ? t0 : (k == 1 ? t1 : t2);
            y[row] = sum;
            if (sent0) {
                sent0[row] = sentinel_value<T>();
                sent1[row] = sentinel_value<T>();
            }
            if (DOT) part += (double)sum * (double)xown[i];
        }
    }
    if (DOT) {
        // each wave's partial goes to the start of its own slice, which it has
        // finished reading: 20 KB of LDS per block keeps 8 blocks per CU
        const double bs = block_sum(part, reinterpret_cast<double*>(&sv[0][0]), ROUNDS * 64 * 16 / 8);
        if (threadIdx.x == 0) partials[blk] = bs;
    }
}

// Y[:, j] = A X[:, j] for nv vectors of leading dimension ld: the LDS-staged
// node SpMV with the wave's value run staged once and used for every vector
// (the matrix is read once for all of them; each vector costs its x gathers
// and its stores).  Per vector the arithmetic is spmv_node_lds_kernel's, in
// the same order.
template <typename T>
__global__ __launch_bounds__(256) void spmv_node_multi_kernel(uint32_t nn, const uint32_t* __restrict__ rp,
                                                              const uint32_t* __restrict__ nptr,
                                                              const uint32_t* __restrict__ ncol,
                                                              const T* __restrict__ val, const T* __restrict__ X,
                                                              T* __restrict__ Y, uint32_t ld, int nv) {
    using V = typename Vec16<T>::type;
    constexpr int G = nodes_per_group<T>(), NPW = 4 * G, W = 16 / (int)sizeof(T);
    constexpr int ROUNDS = FM_SPMV_LDS_ROUNDS, CAP = ROUNDS * 64 * W;
    __shared__ V sv[4][ROUNDS * 64];
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63, q = lane >> 4, k = lane & 15;
    const uint32_t blk = xcd_remap(blockIdx.x, gridDim.x);
    const uint32_t v0 = (blk * 4 + wave) * NPW;
    const uint32_t rbase = rp[0];
    const uint32_t pl = nptr[min(v0 + min((uint32_t)lane, (uint32_t)NPW), nn)];
    uint32_t node[G];
#pragma unroll
    for (int i = 0; i < G; ++i) node[i] = v0 + 4 * i + q;
    const uint32_t r0 = rbase + 9 * (uint32_t)__builtin_amdgcn_readlane((int)pl, 0);
    const uint32_t r1 = rbase + 9 * (uint32_t)__builtin_amdgcn_readlane((int)pl, NPW);
    const uint32_t base = r0 & ~(uint32_t)(W - 1);
    const bool staged = r1 - base <= (uint32_t)CAP;  // wave-uniform
    const uint32_t last = r1 > base ? (r1 - 1) & ~(uint32_t)(W - 1) : base;
    uint32_t b0[G], off[G], nb[G], cn[G];
#pragma unroll
    for (int i = 0; i < G; ++i) {
        const uint32_t bb = __shfl(pl, 4 * i + q), e = __shfl(pl, 4 * i + q + 1);
        const bool ok = node[i] < nn;
        nb[i] = ok ? e - bb : 0;
        b0[i] = ok ? bb : 0;
        const uint32_t kk = min((uint32_t)k, max(nb[i], 1u) - 1);
        off[i] = ok ? rbase + 9 * bb + 3 * kk : base;
        cn[i] = ncol[b0[i] + kk];
    }
    if (staged) {
#pragma unroll
        for (int r = 0; r < ROUNDS; ++r) {
            const uint32_t c = min(base + (uint32_t)(r * 64 + lane) * W, last);
            sv[wave][r * 64 + lane] = __builtin_nontemporal_load(reinterpret_cast<const V*>(val + c));
        }
    }
    __syncthreads();
    const T* lv = reinterpret_cast<const T*>(&sv[wave][0]);
    for (int j = 0; j < nv; ++j) {
        const T* x = X + (size_t)j * ld;
        T s[G][3];
        auto products = [&](auto av) {
#pragma unroll
            for (int i = 0; i < G; ++i) {
                const uint32_t o0 = off[i], o1 = o0 + 3 * nb[i], o2 = o1 + 3 * nb[i];
                T a[9];
#pragma unroll
                for (int c = 0; c < 3; ++c) {
                    a[c] = av(o0 + c);
                    a[3 + c] = av(o1 + c);
                    a[6 + c] = av(o2 + c);
                }
                const T w = (uint32_t)k < nb[i] ? (T)1 : (T)0;
                const T* xm = x + 3 * (size_t)cn[i];
                const T x0 = xm[0], x1 = xm[1], x2 = xm[2];
                s[i][0] = (a[0] * x0 + a[1] * x1 + a[2] * x2) * w;
                s[i][1] = (a[3] * x0 + a[4] * x1 + a[5] * x2) * w;
                s[i][2] = (a[6] * x0 + a[7] * x1 + a[8] * x2) * w;
            }
#pragma unroll
            for (int i = 0; i < G; ++i)
                for (uint32_t kb = k + kNodeLanes; kb < nb[i]; kb += kNodeLanes) {
                    const uint32_t o0 = rbase + 9 * b0[i] + 3 * kb, o1 = o0 + 3 * nb[i], o2 = o1 + 3 * nb[i];
                    const T* xm = x + 3 * (size_t)ncol[b0[i] + kb];
                    const T x0 = xm[0], x1 = xm[1], x2 = xm[2];
                    s[i][0] += av(o0) * x0 + av(o0 + 1) * x1 + av(o0 + 2) * x2;
                    s[i][1] += av(o1) * x0 + av(o1 + 1) * x1 + av(o1 + 2) * x2;
                    s[i][2] += av(o2) * x0 + av(o2 + 1) * x1 + av(o2 + 2) * x2;
                }
        };
        if (staged)
            products([&](uint32_t o) { return lv[o - base]; });
        else
            products([&](uint32_t o) { return val[o]; });
        T* y = Y + (size_t)j * ld;
#pragma unroll
        for (int i = 0; i < G; ++i) {
            const T t0 = row16_sum(s[i][0]), t1 = row16_sum(s[i][1]), t2 = row16_sum(s[i][2]);
            if (node[i] < nn && k < 3) y[3 * node[i] + k] = k == 0 ? t0 : (k == 1 ? t1 : t2);
        }
    }
}

// ---------------------------------------------------------------------------
// natural-order ILU(0), node-blocked (B rows per node) and sync-free:
// nodes are taken from a level-ordered queue through a ticket counter, so a
// node only ever waits on nodes that a running or finished wave owns.
// ---------------------------------------------------------------------------
template <typename T, int S>
__device__ __forceinline__ T pick(const T (&a)[S], int sl) {
    T v = a[0];
#pragma unroll
    for (int s = 1; s < S; ++s)
        if (sl == s) v = a[s];
    return v;
}
template <typename T, int S>
__device__ __forceinline__ void put(T (&a)[S], int sl, T v) {
#pragma unroll
    for (int s = 0; s < S; ++s)
        if (sl == s) a[s] = v;
}

// One wave factorizes the B rows of one node (IKJ order, which subtracts the
// terms of every L and U entry in ascending pivot order, as the reference
// factorization does), reading the finished U rows of its lower neighbours.
template <typename T, int B, int S>
__device__ __forceinline__ void factor_node(uint32_t node, const uint32_t* __restrict__ rp,
                                            const uint32_t* __restrict__ ci, const uint16_t* __restrict__ nlow,
                                            const T* __restrict__ A, T* LU, int* err) {
    const int lane = threadIdx.x;
    const uint32_t r0 = B * node;
    uint32_t rs[B];
#pragma unroll
    for (int d = 0; d < B; ++d) rs[d] = rp[r0 + d];
    const uint32_t len = rp[r0 + 1] - rs[0];
    const uint32_t p = B * (uint32_t)nlow[node];
    T a[B][S];
    uint32_t col[S];
#pragma unroll
    for (int s = 0; s < S; ++s) {
        const uint32_t j = lane + 64 * s;
        const bool v = j < len;
        col[s] = v ? ci[rs[0] + j] : 0xFFFFFFFFu;
#pragma unroll
        for (int d = 0; d < B; ++d) a[d][s] = v ? A[rs[d] + j] : (T)0;
    }
    for (uint32_t q = 0; q < p; q += B) {
        const uint32_t m = ci[rs[0] + q] / B;
        const uint32_t rm = rp[B * m];
        const uint32_t lenm = rp[B * m + 1] - rm;
        const uint32_t pm = B * (uint32_t)nlow[m];
        int pos[S];
#pragma unroll
        for (int s = 0; s < S; ++s) {
            pos[s] = -1;
            const uint32_t j = lane + 64 * s;
            if (j < len && j > q) {
                uint32_t lo = pm, hi = lenm;
                const uint32_t c = col[s];
                while (lo < hi) {
                    const uint32_t mid = (lo + hi) >> 1;
                    if (ci[rm + mid] < c)
                        lo = mid + 1;
                    else
                        hi = mid;
                }
                if (lo < lenm && ci[rm + lo] == c) pos[s] = (int)lo;
            }
        }
#pragma unroll
        for (int e = 0; e < B; ++e) {
            const uint32_t k = B * m + e;
            const uint32_t rk = rp[k];
            const T ukk = await_val(&LU[rk + pm + e], err);
            const uint32_t jj = q + e;
            const int owner = jj & 63, sl = jj >> 6;
            T l[B];
#pragma unroll
            for (int d = 0; d < B; ++d) l[d] = __shfl(pick(a[d], sl), owner) / ukk;
            if (lane == owner)
#pragma unroll
                for (int d = 0; d < B; ++d) put(a[d], sl, l[d]);
#pragma unroll
            for (int s = 0; s < S; ++s) {
                if (pos[s] >= 0 && (uint32_t)(lane + 64 * s) > jj) {
                    const T ukj = await_val(&LU[rk + pos[s]], err);
#pragma unroll
                    for (int d = 0; d < B; ++d) a[d][s] -= l[d] * ukj;
                }
            }
        }
    }
    // diagonal block of the node itself
#pragma unroll
    for (int d = 0; d < B; ++d) {
#pragma unroll
        for (int e = 0; e < d; ++e) {
            const uint32_t jj = p + e;
            const int owner = jj & 63, sl = jj >> 6;
            const T ue = __shfl(pick(a[e], sl), owner);
            const T ld = __shfl(pick(a[d], sl), owner) / ue;
            if (lane == owner) put(a[d], sl, ld);
#pragma unroll
            for (int s = 0; s < S; ++s)
                if ((uint32_t)(lane + 64 * s) > jj && lane + 64 * s < (int)len) a[d][s] -= ld * a[e][s];
        }
        const uint32_t jd = p + d;
        if (lane == (int)(jd & 63)) {
            T u = pick(a[d], jd >> 6);
            if (fabs((double)u) < 1e-12) u += (u >= (T)0) ? (T)1e-8 : (T)-1e-8;
            put(a[d], jd >> 6, u);
        }
    }
#pragma unroll
    for (int s = 0; s < S; ++s) {
        const uint32_t j = lane + 64 * s;
        if (j < len)
#pragma unroll
            for (int d = 0; d < B; ++d) publish(&LU[rs[d] + j], a[d][s]);
    }
}

// value of lane l (a compile-time constant) in every lane, through a scalar register
__device__ __forceinline__ float lane_bcast(float v, int l) {
    return __builtin_bit_cast(float, __builtin_amdgcn_readlane(__builtin_bit_cast(int, v), l));
}
__device__ __forceinline__ double lane_bcast(double v, int l) {
    const uint64_t b = __builtin_bit_cast(uint64_t, v);
    const uint32_t lo = __builtin_amdgcn_readlane((int)(uint32_t)b, l),
                   hi = __builtin_amdgcn_readlane((int)(uint32_t)(b >> 32), l);
    return __builtin_bit_cast(double, ((uint64_t)hi << 32) | lo);
}

// B = 3: the same IKJ arithmetic as factor_node with the U rows of the lower
// neighbours staged in LDS.  The loads of all neighbours are issued together;
// each neighbour's values are re-polled (if still the sentinel) just before
// its elimination step, and the column positions come from a search of the
// neighbours' column lists in LDS.  A node costs a few memory round trips
// instead of one per search step and per U entry.  Returns false (nothing
// written) when the node exceeds the staging limits; the caller then runs
// factor_node.
constexpr uint32_t kFacLow = 8;           // lower neighbours (at most 7 in a Kuhn-split hex mesh)
constexpr uint32_t kFacUp = 10;           // blocks of a neighbour's row from its diagonal block on (at most 8 there)
constexpr uint32_t kFacColPad = 16;       // column list padded to a power of two for the search
constexpr uint32_t kFacSeg = 3 * kFacUp;  // staged entries per row
constexpr uint32_t kFacQ = 96;            // LDS stride of one neighbour: its 3 rows (two wave loads)
static_assert(3 * kFacSeg <= kFacQ && kFacQ <= 128, "a neighbour's rows fit two wave loads");
static_assert(kFacUp < kFacColPad && kFacLow * kFacColPad % 64 == 0, "padded column lists");
template <typename T, int S>
__device__ __forceinline__ bool factor_node3_staged(uint32_t node, const uint32_t* __restrict__ rp,
                                                    const uint32_t* __restrict__ ci,
                                                    const uint16_t* __restrict__ nlow, const T* __restrict__ A,
                                                    T* LU, int* err) {
    constexpr int B = 3;
    using U = typename SB<T>::U;
    constexpr int NIT = (int)(2 * kFacLow);  // item it: neighbour it / 2, half it % 2 of its padded stride
    static_assert(NIT <= 32, "pending mask holds one bit per item");
    __shared__ uint32_t s_col[kFacLow][kFacColPad];  // column node of each staged block; ~0 past the last
    __shared__ T s_u[kFacLow * kFacQ];           // [q][e * kFacSeg + j]: row 3 m_q + e, entry j after the diagonal offset
    const int lane = threadIdx.x;
    const uint32_t r0 = B * node;
    uint32_t rs[B];
#pragma unroll
    for (int d = 0; d < B; ++d) rs[d] = rp[r0 + d];
    const uint32_t len = rp[r0 + 1] - rs[0];
    const uint32_t nl = nlow[node], p = B * nl;
    if (len > 64u * S || nl > kFacLow) return false;
    // lane q < nl: lower neighbour m_q, its row starts, diagonal offset and staged blocks
    uint32_t rk0 = 0, rk1 = 0, rk2 = 0, pm = 0, nup = 0;
    if ((uint32_t)lane < nl) {
        const uint32_t m = ci[rs[0] + B * lane] / B;
        rk0 = rp[B * m];
        rk1 = rp[B * m + 1];
        rk2 = rp[B * m + 2];
        pm = B * (uint32_t)nlow[m];
        nup = (rk1 - rk0 - pm) / B;
    }
    if (__any((uint32_t)lane < nl && nup > kFacUp)) return false;
    T a[B][S];
    uint32_t col[S];
#pragma unroll
    for (int s = 0; s < S; ++s) {
        const uint32_t j = lane + 64 * s;
        const bool v = j < len;
        col[s] = v ? ci[rs[0] + j] : 0xFFFFFFFFu;
#pragma unroll
        for (int d = 0; d < B; ++d) a[d][s] = v ? A[rs[d] + j] : (T)0;
    }
    // column nodes of the staged blocks: loaded before the values, so waiting
    // for them does not wait for the values (loads complete in order)
#pragma unroll
    for (uint32_t k = 0; k < kFacLow * kFacColPad / 64; ++k) {
        const uint32_t c = lane + 64 * k, q = c / kFacColPad, b = c % kFacColPad;
        const uint32_t row = __shfl(rk0, (int)q), off = __shfl(pm, (int)q), nq = __shfl(nup, (int)q);
        (&s_col[0][0])[c] = (q < nl && b < nq) ? ci[row + off + B * b] / B : 0xFFFFFFFFu;
    }
    // global address of item it of this lane (neighbour constant per item: scalar reads)
    auto item_ptr = [&](int it, bool& ok) -> U* {
        const int q = it >> 1;
        const uint32_t idx = lane + 64u * (it & 1), e = idx / kFacSeg, j = idx % kFacSeg;
        const uint32_t x0 = __builtin_amdgcn_readlane(rk0, q), x1 = __builtin_amdgcn_readlane(rk1, q),
                       x2 = __builtin_amdgcn_readlane(rk2, q);
        const uint32_t off = __builtin_amdgcn_readlane(pm, q), nq = __builtin_amdgcn_readlane(nup, q);
        ok = (uint32_t)q < nl && e < B && j < B * nq;
        return reinterpret_cast<U*>(LU + (e == 0 ? x0 : e == 1 ? x1 : x2) + off + j);
    };
    U val[NIT];
    uint32_t pending = 0;
#pragma unroll
    for (int it = 0; it < NIT; ++it) {
        bool ok;
        U* ptr = item_ptr(it, ok);
        val[it] = SB<T>::S;
        if (ok) {
            val[it] = __hip_atomic_load(ptr, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            pending |= 1u << it;
        }
    }
    __syncthreads();
    // positions of this lane's columns in every neighbour's staged rows, while the values are in flight
    int pos[kFacLow][S];
#pragma unroll
    for (int q = 0; q < (int)kFacLow; ++q) {
        const uint32_t nq = __builtin_amdgcn_readlane(nup, q);
#pragma unroll
        for (int s = 0; s < S; ++s) {
            pos[q][s] = -1;
            const uint32_t j = lane + 64 * s;
            // branch-free lower bound over the padded list: the searches of all
            // neighbours are independent and overlap their LDS latencies
            const uint32_t cn = col[s] / B;
            uint32_t lo = 0;
#pragma unroll
            for (uint32_t step = kFacColPad / 2; step >= 1; step >>= 1)
                lo += s_col[q][lo + step - 1] < cn ? step : 0;
            const bool hit = lo < nq && s_col[q][lo] == cn;
            if ((uint32_t)q < nl && j < len && j > (uint32_t)(B * q) && hit) pos[q][s] = (int)(B * lo + col[s] % B);
        }
    }
    // IKJ over the lower neighbours in ascending order, as factor_node; each
    // neighbour is awaited just before its turn, so the elimination by early
    // neighbours overlaps the wait for late ones (bounded like await_val)
    uint32_t spins = 0;
#pragma unroll
    for (int q = 0; q < (int)kFacLow; ++q) {
        if ((uint32_t)q >= nl) continue;
        constexpr uint32_t none = 0;
        const uint32_t mine_q = 3u << (2 * q);
        for (;;) {
#pragma unroll
            for (int h = 0; h < 2; ++h) {
                const int it = 2 * q + h;
                if (((pending >> it) & 1u) && val[it] != SB<T>::S) {
                    s_u[q * kFacQ + lane + 64u * h] = __builtin_bit_cast(T, val[it]);
                    pending &= ~(1u << it);
                }
            }
            if (!__any((pending & mine_q) != none)) break;
            if (++spins > kSpinLimit) {
                if (pending & mine_q) atomicOr(err, 1);
                break;
            }
            if ((spins & 255u) == 0 && __hip_atomic_load(err, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT)) break;
            __builtin_amdgcn_s_sleep(FM_SPIN_SLEEP);
#pragma unroll
            for (int h = 0; h < 2; ++h) {
                const int it = 2 * q + h;
                bool ok;
                U* ptr = item_ptr(it, ok);
                if ((pending >> it) & 1u) val[it] = __hip_atomic_load(ptr, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            }
        }
        __syncthreads();
        const T* uq = s_u + q * kFacQ;
#pragma unroll
        for (int e = 0; e < B; ++e) {
            const T ukk = uq[e * kFacSeg + e];
            const int jj = B * q + e, owner = jj & 63, sl = jj >> 6;
            T l[B];
#pragma unroll
            for (int d = 0; d < B; ++d) l[d] = lane_bcast(pick(a[d], sl), owner) / ukk;
            if (lane == owner)
#pragma unroll
                for (int d = 0; d < B; ++d) put(a[d], sl, l[d]);
#pragma unroll
            for (int s = 0; s < S; ++s) {
                if (pos[q][s] >= 0 && lane + 64 * s > jj) {
                    const T ukj = uq[e * kFacSeg + pos[q][s]];
#pragma unroll
                    for (int d = 0; d < B; ++d) a[d][s] -= l[d] * ukj;
                }
            }
        }
    }
    // diagonal block of the node itself
#pragma unroll
    for (int d = 0; d < B; ++d) {
#pragma unroll
        for (int e = 0; e < d; ++e) {
            const uint32_t jj = p + e;
            const int owner = jj & 63, sl = jj >> 6;
            const T ue = __shfl(pick(a[e], sl), owner);
            const T ld = __shfl(pick(a[d], sl), owner) / ue;
            if (lane == owner) put(a[d], sl, ld);
#pragma unroll
            for (int s = 0; s < S; ++s)
                if ((uint32_t)(lane + 64 * s) > jj && lane + 64 * s < (int)len) a[d][s] -= ld * a[e][s];
        }
        const uint32_t jd = p + d;
        if (lane == (int)(jd & 63)) {
            T u = pick(a[d], jd >> 6);
            if (fabs((double)u) < 1e-12) u += (u >= (T)0) ? (T)1e-8 : (T)-1e-8;
            put(a[d], jd >> 6, u);
        }
    }
#pragma unroll
    for (int s = 0; s < S; ++s) {
        const uint32_t j = lane + 64 * s;
        if (j < len)
#pragma unroll
            for (int d = 0; d < B; ++d) publish(&LU[rs[d] + j], a[d][s]);
    }
    __syncthreads();  // the next node of this wave overwrites the stage
    return true;
}

// persistent, like the sweeps: wave w takes queue entries w, w+P, ...
template <typename T, int B, int S>
__global__ __launch_bounds__(64) void ilu_factor_kernel(uint32_t nq, const uint32_t* __restrict__ order,
                                                        const uint32_t* __restrict__ rp,
                                                        const uint32_t* __restrict__ ci,
                                                        const uint16_t* __restrict__ nlow, const T* __restrict__ A,
                                                        T* LU, uint32_t* ticket, int staged, int* err) {
    for (uint32_t t = blockIdx.x; t < nq; t += gridDim.x) {
        const uint32_t node = order[t];
        if (node == kNoNode) continue;
        if constexpr (B == 3 && S <= 2)
            if (staged && factor_node3_staged<T, S>(node, rp, ci, nlow, A, LU, err)) continue;
        factor_node<T, B, S>(node, rp, ci, nlow, A, LU, err);
    }
}

// Triangular sweeps.  Both are persistent with every wave resident from the
// start (launch_persistent): wave w works queue chunks w, w+P, w+2P, ... in order.
// The lowest unfinished chunk therefore always belongs to a running wave and
// depends only on lower, finished chunks, so the sweep always progresses; and
// the P waves in flight cover a window of consecutive chunks at the dependence
// frontier.  (A shared ticket counter was the bottleneck: one atomic per chunk
// on one address serializes at ~80 M/s.)  Everything a node needs that does not
// depend on other nodes (matrix values, columns, right-hand side, PCG vectors,
// diagonal-block entries) is loaded before the first wait.
constexpr int kMJ = 2;  // lower/upper entries per lane kept in registers (more fall back to a loop)

// forward sweep z = L^{-1} rhs for the NPW nodes of ticket t; with UPD the PCG
// update r -= alpha Ap, x += alpha p of the node's rows is fused in and the
// new r is the right-hand side
template <typename T, int B, int LPN, bool UPD>
__device__ __forceinline__ void fwd_ticket(uint32_t t, uint32_t nq, const uint32_t* __restrict__ order,
                                           const uint32_t* __restrict__ rp, const uint32_t* __restrict__ ci,
                                           const uint16_t* __restrict__ nlow, const T* __restrict__ LU,
                                           const T* __restrict__ rhs, T* __restrict__ r, T* __restrict__ x,
                                           const T* __restrict__ p, const T* __restrict__ Ap, T* z, T alpha,
                                           int* err) {
    constexpr int NPW = 64 / LPN;
    const int lane = threadIdx.x, gl = lane % LPN;
    const uint32_t idx = t * NPW + lane / LPN;
    const uint32_t entry = idx < nq ? order[idx] : kNoNode;
    const bool valid = entry != kNoNode;
    const uint32_t node = valid ? entry : 0;
    const uint32_t r0 = B * node;
    uint32_t rs[B];
    uint32_t pl = 0;
    uint32_t c[kMJ];
    T l[kMJ][B];
    T rv[B], apv[B], pv[B], xv[B], ld[B][B];
    if (valid) {
#pragma unroll
        for (int d = 0; d < B; ++d) rs[d] = rp[r0 + d];
        pl = B * (uint32_t)nlow[node];
#pragma unroll
        for (int m = 0; m < kMJ; ++m) {
            const uint32_t j = gl + m * LPN;
            if (j < pl) {
                c[m] = ci[rs[0] + j];
#pragma unroll
                for (int d = 0; d < B; ++d) l[m][d] = LU[rs[d] + j];
            }
        }
        if (gl == 0) {
#pragma unroll
            for (int d = 0; d < B; ++d) {
                if (UPD) {
                    rv[d] = r[r0 + d];
                    apv[d] = Ap[r0 + d];
                    pv[d] = p[r0 + d];
                    xv[d] = x[r0 + d];
                } else {
                    rv[d] = rhs[r0 + d];
                }
#pragma unroll
                for (int e = 0; e < d; ++e) ld[d][e] = LU[rs[d] + pl + e];
            }
        }
    }
    T acc[B];
#pragma unroll
    for (int d = 0; d < B; ++d) acc[d] = 0;
    if (valid) {
        const bool h0 = (uint32_t)gl < pl, h1 = (uint32_t)gl + LPN < pl;
        if (h0) {
            T z0, z1;
            await2(&z[c[0]], &z[h1 ? c[1] : c[0]], h1, z0, z1, err);
#pragma unroll
            for (int d = 0; d < B; ++d) acc[d] += l[0][d] * z0;
            if (h1)
#pragma unroll
                for (int d = 0; d < B; ++d) acc[d] += l[1][d] * z1;
        }
        for (uint32_t j = gl + kMJ * LPN; j < pl; j += LPN) {
            const T zc = await_val(&z[ci[rs[0] + j]], err);
#pragma unroll
            for (int d = 0; d < B; ++d) acc[d] += LU[rs[d] + j] * zc;
        }
    }
#pragma unroll
    for (int o = LPN / 2; o > 0; o >>= 1)
#pragma unroll
        for (int d = 0; d < B; ++d) acc[d] += __shfl_xor(acc[d], o);
    if (valid && gl == 0) {
        T v[B];
#pragma unroll
        for (int d = 0; d < B; ++d) {
            T rhs_d = rv[d];
            if (UPD) {
                rhs_d = rv[d] - alpha * apv[d];
                r[r0 + d] = rhs_d;
                x[r0 + d] = xv[d] + alpha * pv[d];
            }
            T s = rhs_d - acc[d];
#pragma unroll
            for (int e = 0; e < d; ++e) s -= ld[d][e] * v[e];
            v[d] = s;
        }
#pragma unroll
        for (int d = 0; d < B; ++d) publish(&z[r0 + d], v[d]);
    }
}

template <typename T, int B, int LPN, bool UPD>
__global__ __launch_bounds__(64) void ilu_fwd_kernel(uint32_t nq, const uint32_t* __restrict__ order,
                                                     const uint32_t* __restrict__ rp, const uint32_t* __restrict__ ci,
                                                     const uint16_t* __restrict__ nlow, const T* __restrict__ LU,
                                                     const T* __restrict__ rhs, T* __restrict__ r, T* __restrict__ x,
                                                     const T* __restrict__ p, const T* __restrict__ Ap, T* z,
                                                     PcgState* st, int* err) {
    if (UPD && st->done) return;
    const T alpha = UPD ? (T)st->alpha : (T)0;
    constexpr int NPW = 64 / LPN;
    const uint32_t ntk = nq / NPW;
    for (uint32_t t = blockIdx.x; t < ntk; t += gridDim.x)
        fwd_ticket<T, B, LPN, UPD>(t, nq, order, rp, ci, nlow, LU, rhs, r, x, p, Ap, z, alpha, err);
}

// backward sweep y = U^{-1} z with the diagonal u_ii at the diagonal
// positions; fused partial of r.y per ticket (fixed order, deterministic)
template <typename T, int B, int LPN>
__device__ __forceinline__ double bwd_ticket(uint32_t t, uint32_t nq, const uint32_t* __restrict__ order,
                                             const uint32_t* __restrict__ rp, const uint32_t* __restrict__ ci,
                                             const uint16_t* __restrict__ nlow, const T* __restrict__ LU,
                                             const T* __restrict__ z, const T* __restrict__ r, T* y, int* err) {
    constexpr int NPW = 64 / LPN;
    const int lane = threadIdx.x, gl = lane % LPN;
    const uint32_t idx = t * NPW + lane / LPN;
    const uint32_t entry = idx < nq ? order[idx] : kNoNode;
    const bool valid = entry != kNoNode;
    const uint32_t node = valid ? entry : 0;
    const uint32_t r0 = B * node;
    uint32_t rs[B];
    uint32_t us = 0, len = 0, pl = 0;
    uint32_t c[kMJ];
    T l[kMJ][B];
    T zv[B], rv[B], ud[B][B];
    if (valid) {
#pragma unroll
        for (int d = 0; d < B; ++d) rs[d] = rp[r0 + d];
        len = rp[r0 + 1] - rs[0];
        pl = B * (uint32_t)nlow[node];
        us = pl + B;
#pragma unroll
        for (int m = 0; m < kMJ; ++m) {
            const uint32_t j = us + gl + m * LPN;
            if (j < len) {
                c[m] = ci[rs[0] + j];
#pragma unroll
                for (int d = 0; d < B; ++d) l[m][d] = LU[rs[d] + j];
            }
        }
        if (gl == 0) {
#pragma unroll
            for (int d = 0; d < B; ++d) {
                zv[d] = z[r0 + d];
                rv[d] = r ? r[r0 + d] : (T)0;
#pragma unroll
                for (int e = d; e < B; ++e) ud[d][e] = LU[rs[d] + pl + e];
            }
        }
    }
    T acc[B];
#pragma unroll
    for (int d = 0; d < B; ++d) acc[d] = 0;
    if (valid) {
        const bool h0 = us + gl < len, h1 = us + gl + LPN < len;
        if (h0) {
            T y0, y1;
            await2(&y[c[0]], &y[h1 ? c[1] : c[0]], h1, y0, y1, err);
#pragma unroll
            for (int d = 0; d < B; ++d) acc[d] += l[0][d] * y0;
            if (h1)
#pragma unroll
                for (int d = 0; d < B; ++d) acc[d] += l[1][d] * y1;
        }
        for (uint32_t j = us + gl + kMJ * LPN; j < len; j += LPN) {
            const T yc = await_val(&y[ci[rs[0] + j]], err);
#pragma unroll
            for (int d = 0; d < B; ++d) acc[d] += LU[rs[d] + j] * yc;
        }
    }
#pragma unroll
    for (int o = LPN / 2; o > 0; o >>= 1)
#pragma unroll
        for (int d = 0; d < B; ++d) acc[d] += __shfl_xor(acc[d], o);
    double part = 0.0;
    if (valid && gl == 0) {
        T v[B];
#pragma unroll
        for (int d = B - 1; d >= 0; --d) {
            T s = zv[d] - acc[d];
#pragma unroll
            for (int e = d + 1; e < B; ++e) s -= ud[d][e] * v[e];
            v[d] = s / ud[d][d];
        }
#pragma unroll
        for (int d = 0; d < B; ++d) {
            publish(&y[r0 + d], v[d]);
            part += (double)rv[d] * (double)v[d];
        }
    }
    return part;
}

template <typename T, int B, int LPN>
__global__ __launch_bounds__(64) void ilu_bwd_kernel(uint32_t nq, const uint32_t* __restrict__ order,
                                                     const uint32_t* __restrict__ rp, const uint32_t* __restrict__ ci,
                                                     const uint16_t* __restrict__ nlow, const T* __restrict__ LU,
                                                     const T* __restrict__ z, const T* __restrict__ r, T* y,
                                                     double* __restrict__ partials, PcgState* st, int check_done,
                                                     int* err) {
    if (check_done && st->done) return;
    constexpr int NPW = 64 / LPN;
    const uint32_t ntk = nq / NPW;
    double acc = 0.0;  // this wave's chunks, in increasing order: a fixed summation order
    for (uint32_t t = blockIdx.x; t < ntk; t += gridDim.x)
        acc += bwd_ticket<T, B, LPN>(t, nq, order, rp, ci, nlow, LU, z, partials ? r : nullptr, y, err);
    if (partials) {
#pragma unroll
        for (int o = 32; o > 0; o >>= 1) acc += __shfl_xor(acc, o);
        if (threadIdx.x == 0) partials[blockIdx.x] = acc;
    }
}

// ---------------------------------------------------------------------------
// PCG scalar bookkeeping: one block, fixed-order sum of the partials
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(1024) void reduce_kernel(const double* __restrict__ partials, uint32_t np,
                                                      PcgState* st, int op, int k) {
    // Block b sums the contiguous chunk b of the partials with eight
    // independent accumulators per thread (loads in flight).  With several
    // blocks, each stores its sum, and the last block to finish (counted
    // with an atomic) adds the block sums in block order.  The result is
    // deterministic for a given partial count and grid.
    const uint32_t chunk = (np + gridDim.x - 1) / gridDim.x;
    const uint32_t lo = blockIdx.x * chunk, hi = min(np, lo + chunk);
    double acc[8] = {0, 0, 0, 0, 0, 0, 0, 0};
    const uint32_t stride = blockDim.x * 8;
    for (uint32_t i0 = lo + threadIdx.x; i0 < hi; i0 += stride)
#pragma unroll
        for (int u = 0; u < 8; ++u) {
            const uint32_t i = i0 + u * blockDim.x;
            if (i < hi) acc[u] += partials[i];
        }
    double s = ((acc[0] + acc[1]) + (acc[2] + acc[3])) + ((acc[4] + acc[5]) + (acc[6] + acc[7]));
    s = block_sum(s);
    if (gridDim.x > 1) {
        __shared__ int s_last;
        if (threadIdx.x == 0) {
            st->red_stage[blockIdx.x] = s;
            __threadfence();
            s_last = atomicAdd(&st->red_count, 1u) == gridDim.x - 1;
        }
        __syncthreads();
        if (!s_last) return;
        __threadfence();
        const double v = threadIdx.x < gridDim.x
                             ? __hip_atomic_load(&st->red_stage[threadIdx.x], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT)
                             : 0.0;
        __syncthreads();  // block_sum's shared slots are reused
        s = block_sum(v);
        if (threadIdx.x == 0) st->red_count = 0;
    }
    if (threadIdx.x != 0) return;
    switch (op) {
        case 0:
            st->slot[k] = s;
            break;
        case 1:  // p.Ap -> alpha (reference: break when |pAp| < 1e-60)
            if (st->done) return;
            st->pAp = s;
            st->ticket[0] = 0;
            st->ticket[1] = 0;
            if (fabs(s) < 1e-60) {
                st->done = 1;
                st->status = 2;
            } else {
                st->alpha = st->rz / s;
            }
            break;
        case 2: {  // r.z -> stopping rule sqrt(rz/bb) < tol, beta
            if (st->done) return;
            s += st->rz_coarse;  // r.(Z E^{-1} Z^T r) of the additive coarse correction (0 without)
            st->rz_new = s;
            st->iters += 1;
            const double rel = sqrt(s / fmax(st->bb, 1e-300));
            st->relres = rel;
            if (rel < st->tol) {
                st->done = 1;
                st->status = 1;
            } else if (st->iters >= st->max_it) {
                st->done = 1;
                st->status = 3;
            } else {
                st->beta = s / st->rz;
                st->rz = s;
            }
            if (st->hist && (uint32_t)(st->iters - 1) < st->hist_cap) {
                st->hist[2 * (st->iters - 1)] = st->alpha;
                st->hist[2 * (st->iters - 1) + 1] = st->done ? 0.0 : st->beta;
            }
            break;
        }
        case 3:
            s += st->rz_coarse;
            st->rz = s;
            st->relres = sqrt(s / fmax(st->bb, 1e-300));
            break;
        case 5:
            st->gal[k] = s;
            break;
        case 4: {  // as 2, with the flexible beta r_new.(z_new - z_old) / r_old.z_old (z_old term in slot[k])
            if (st->done) return;
            // with the additive coarse correction z = y + Z c: r.(Z c) and
            // r_new.(Z c_old) from the coarse solve (0 without one)
            s += st->rz_coarse;
            st->slot[k] += st->rz_cross;
            st->rz_new = s;
            st->iters += 1;
            const double rel = sqrt(s / fmax(st->bb, 1e-300));
            st->relres = rel;
            if (rel < st->tol) {
                st->done = 1;
                st->status = 1;
            } else if (st->iters >= st->max_it) {
                st->done = 1;
                st->status = 3;
            } else {
                st->beta = (s - st->slot[k]) / st->rz;
                st->rz = s;
            }
            break;
        }
    }
}

template <typename T>
__global__ void pcg_p_update_kernel(uint32_t n, const T* __restrict__ y, T* __restrict__ p,
                                    const PcgState* __restrict__ st) {
    if (st->done) return;
    const uint32_t i = xcd_remap(blockIdx.x, gridDim.x) * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const T beta = (T)st->beta;
    p[i] = y[i] + beta * p[i];
}

// x += alpha p, r -= alpha Ap (the update the level-scheduled forward sweep fuses)
template <typename T>
__global__ void pcg_xr_update_kernel(uint32_t n, const T* __restrict__ p, const T* __restrict__ Ap,
                                     T* __restrict__ x, T* __restrict__ r, const PcgState* __restrict__ st) {
    if (st->done) return;
    const uint32_t i = xcd_remap(blockIdx.x, gridDim.x) * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const T alpha = (T)st->alpha;
    x[i] = x[i] + alpha * p[i];
    r[i] = r[i] - alpha * Ap[i];
}

template <typename T>
__global__ __launch_bounds__(256) void pcg_init_r_kernel(uint32_t n, const T* __restrict__ b, const T* __restrict__ Ap,
                                                         T* __restrict__ r, double* __restrict__ partials) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    double part = 0.0;
    if (i < n) {
        const T v = b[i] - Ap[i];
        r[i] = v;
        part = (double)v * (double)v;
    }
    const double s = block_sum(part);
    if (threadIdx.x == 0) partials[blockIdx.x] = s;
}

template <typename T>
__global__ void copy_kernel(uint32_t n, const T* __restrict__ s, T* __restrict__ d) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) d[i] = s[i];
}

template <typename T>
__global__ void fill_kernel(T* __restrict__ p, size_t n, T v) {
    for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x)
        p[i] = v;
}

template <typename T>
__global__ void fill_sentinel_kernel(T* __restrict__ a, T* __restrict__ b, size_t n) {
    const T s = sentinel_value<T>();
    for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x) {
        a[i] = s;
        if (b) b[i] = s;
    }
}

__global__ void expand_pattern_kernel(uint32_t nn, const uint32_t* __restrict__ nbr_ptr,
                                      const uint32_t* __restrict__ nbr, const uint32_t* __restrict__ row_ptr,
                                      uint32_t* __restrict__ col_idx) {
    const uint32_t n = blockIdx.x * blockDim.x + threadIdx.x;
    if (n >= nn) return;
    const uint32_t b = nbr_ptr[n], e = nbr_ptr[n + 1];
    for (int d = 0; d < 3; ++d) {
        uint32_t k = row_ptr[3 * n + d];
        for (uint32_t q = b; q < e; ++q) {
            const uint32_t m = 3 * nbr[q];
            col_idx[k++] = m;
            col_idx[k++] = m + 1;
            col_idx[k++] = m + 2;
        }
    }
}

__global__ __launch_bounds__(256) void residual_finish_kernel(uint32_t n, double load,
                                                              const double* __restrict__ f_ext,
                                                              const double* __restrict__ f,
                                                              const uint8_t* __restrict__ isdir,
                                                              double* __restrict__ R, double* __restrict__ partials) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    double part = 0.0;
    if (i < n) {
        const double v = (isdir && isdir[i]) ? 0.0 : (f_ext ? load * f_ext[i] : 0.0) - f[i];
        R[i] = v;
        part = v * v;
    }
    const double s = block_sum(part);
    if (threadIdx.x == 0) partials[blockIdx.x] = s;
}

__global__ __launch_bounds__(256) void dot64_kernel(uint32_t n, const double* __restrict__ a,
                                                    const double* __restrict__ b, double* __restrict__ partials) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    const double part = (i < n) ? a[i] * b[i] : 0.0;
    const double s = block_sum(part);
    if (threadIdx.x == 0) partials[blockIdx.x] = s;
}

template <typename T>
__global__ __launch_bounds__(256) void dot_kernel(uint32_t n, const T* __restrict__ a, const T* __restrict__ b,
                                                  double* __restrict__ partials) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    const double part = (i < n) ? (double)a[i] * (double)b[i] : 0.0;
    const double s = block_sum(part);
    if (threadIdx.x == 0) partials[blockIdx.x] = s;
}

// out = c0 v0 + c1 v1 + c2 v2 + c3 v3 (the first m terms), in FP64, rounded once
struct Combine {
    double c[4];
    const void* v[4];
};
template <typename T>
__global__ void combine_kernel(uint32_t n, int m, Combine cb, T* __restrict__ out) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    double acc = 0.0;
    for (int j = 0; j < m; ++j) acc += cb.c[j] * (double)static_cast<const T*>(cb.v[j])[i];
    out[i] = (T)acc;
}

template <typename T>
__global__ void normalize_kernel(uint32_t n, const double* __restrict__ nrm2, const T* __restrict__ in,
                                 T* __restrict__ out) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const double s = *nrm2 > 0.0 ? 1.0 / sqrt(*nrm2) : 0.0;
    out[i] = (T)((double)in[i] * s);
}

// ---------------------------------------------------------------------------
// coarse space of the Galerkin start: polynomial displacement fields, the
// Gram matrix of a tall-skinny basis, and the start x0 = V c
// ---------------------------------------------------------------------------
constexpr int kMaxMono = 35;  // monomials of total degree <= 4 in three variables

// Monomial k of total degree <= 4 in (x, y, z): exponents packed as
// ex | ey << 3 | ez << 6, by degree (1; x, y, z; x^2, y^2, z^2, xy, yz, xz;
// the cubics; the quartics)
__host__ __device__ constexpr int mono_code(int k) {
    switch (k) {
        case 0: return 0;
        case 1: return 1;
        case 2: return 1 << 3;
        case 3: return 1 << 6;
        case 4: return 2;
        case 5: return 2 << 3;
        case 6: return 2 << 6;
        case 7: return 1 | 1 << 3;
        case 8: return 1 << 3 | 1 << 6;
        case 9: return 1 | 1 << 6;
        case 10: return 3;
        case 11: return 3 << 3;
        case 12: return 3 << 6;
        case 13: return 2 | 1 << 3;
        case 14: return 2 | 1 << 6;
        case 15: return 1 | 2 << 3;
        case 16: return 2 << 3 | 1 << 6;
        case 17: return 1 | 2 << 6;
        case 18: return 1 << 3 | 2 << 6;
        case 19: return 1 | 1 << 3 | 1 << 6;
        case 20: return 4;
        case 21: return 4 << 3;
        case 22: return 4 << 6;
        case 23: return 3 | 1 << 3;
        case 24: return 3 | 1 << 6;
        case 25: return 1 | 3 << 3;
        case 26: return 3 << 3 | 1 << 6;
        case 27: return 1 | 3 << 6;
        case 28: return 1 << 3 | 3 << 6;
        case 29: return 2 | 2 << 3;
        case 30: return 2 << 3 | 2 << 6;
        case 31: return 2 | 2 << 6;
        case 32: return 2 | 1 << 3 | 1 << 6;
        case 33: return 1 | 2 << 3 | 1 << 6;
        case 34: return 1 | 1 << 3 | 2 << 6;
        default: return 0;
    }
}
__constant__ int kMonoCodeDev[kMaxMono] = {
    mono_code(0),  mono_code(1),  mono_code(2),  mono_code(3),  mono_code(4),  mono_code(5),  mono_code(6),
    mono_code(7),  mono_code(8),  mono_code(9),  mono_code(10), mono_code(11), mono_code(12), mono_code(13),
    mono_code(14), mono_code(15), mono_code(16), mono_code(17), mono_code(18), mono_code(19), mono_code(20),
    mono_code(21), mono_code(22), mono_code(23), mono_code(24), mono_code(25), mono_code(26), mono_code(27),
    mono_code(28), mono_code(29), mono_code(30), mono_code(31), mono_code(32), mono_code(33), mono_code(34)};

// powers 1, x, .., x^4 and their product for one monomial: every kernel
// evaluates a monomial as (x^ex y^ey) z^ez from these, so the basis columns,
// the products with the matrix and the coarse correction see the same bits
template <typename T>
__device__ __forceinline__ void powers5(T a, T* q) {
    q[0] = (T)1;
    q[1] = a;
    q[2] = a * a;
    q[3] = q[2] * a;
    q[4] = q[2] * q[2];
}
template <typename T>
struct Powers {
    T x[5], y[5], z[5];
    __device__ __forceinline__ explicit Powers(float4 c) {
        powers5((T)c.x, x);
        powers5((T)c.y, y);
        powers5((T)c.z, z);
    }
    // compile-time exponents (unrolled loops)
    __device__ __forceinline__ T mono(int code) const { return (x[code & 7] * y[(code >> 3) & 7]) * z[code >> 6]; }
    // run-time exponents, selected without indexing the register arrays
    __device__ __forceinline__ static T pick(const T* q, int e) {
        return e == 0 ? q[0] : e == 1 ? q[1] : e == 2 ? q[2] : e == 3 ? q[3] : q[4];
    }
    __device__ __forceinline__ T mono_rt(int code) const {
        return (pick(x, code & 7) * pick(y, (code >> 3) & 7)) * pick(z, code >> 6);
    }
};

// xn[v] = ((X_v - center) * scale, mask): the node coordinates mapped to
// [-1, 1] over the bounding box, in FP32, and in .w the bit mask of the
// node's constrained components (bit c: component c)
__global__ void coarse_coords_kernel(uint32_t nn, const double* __restrict__ X, const uint8_t* __restrict__ isdir,
                                     double cx, double cy, double cz, double sx, double sy, double sz,
                                     float4* __restrict__ xn) {
    const uint32_t v = blockIdx.x * blockDim.x + threadIdx.x;
    if (v >= nn) return;
    int mask = 0;
    if (isdir)
        for (int c = 0; c < 3; ++c) mask |= isdir[3 * v + c] ? 1 << c : 0;
    xn[v] = make_float4((float)((X[v] - cx) * sx), (float)((X[nn + v] - cy) * sy), (float)((X[2 * nn + v] - cz) * sz),
                        (float)mask);
}

// column (col0 + c ldm + k), k < nmono, holds monomial k of the scaled node
// coordinates in displacement component c, zero in the other components and
// on constrained DOFs; one thread per node
template <typename T>
__global__ void poly_basis_kernel(uint32_t nn, const float4* __restrict__ xn, int nmono, int ldm,
                                  T* __restrict__ out, uint32_t ld, int col0) {
    const uint32_t v = blockIdx.x * blockDim.x + threadIdx.x;
    if (v >= nn) return;
    const float4 c4 = xn[v];
    const Powers<T> pw(c4);
    const int mask = (int)c4.w;
    for (int k = 0; k < nmono; ++k) {
        const T m = pw.mono_rt(kMonoCodeDev[k]);
        for (int c = 0; c < 3; ++c) {
            T* col = out + (size_t)(col0 + c * ldm + k) * ld + 3 * (size_t)v;
            for (int d = 0; d < 3; ++d) col[d] = (d == c && !(mask >> c & 1)) ? m : (T)0;
        }
    }
}

// Rows of the Galerkin start's Gram matrix that belong to the polynomial
// fields Z (3 NM columns, (k, c) at c NM + k), without forming K Z:
//   G[(k,c), (k',c')] = sum_v zeta_k,c(v) H_v[c][(k',c')],
//   H_v[c][(k',c')]   = sum_q K_vq[c][c'] zeta_k',c'(u_q)
// (zeta_k,c(v): monomial k at node v, zero where component c is constrained;
// K_vq: the 3x3 block of node row v and column node u_q), and the moments
// G[(k,c), 3 NM + j] = sum_v zeta_k,c(v) S_j[3v + c] of kPgStored stored
// vectors S_j (the ring's K V columns and the right-hand side).  A block
// takes tiles of TILE consecutive nodes, whose value runs are one contiguous
// range: the range and the powers of the column nodes' coordinates are
// staged in LDS with independent loads (a tile whose range exceeds the LDS
// slice reads them directly), then the tile's monomials and H rows (the
// row's products with every field, then the stored values), and each update
// thread adds the tile's contributions to its 5 x 7 tile of G (sums in T over
// the tile, FP64 over the tiles).  partials: one 3 NM x (3 NM + kPgStored)
// block per workgroup, row-major.
constexpr int kPgStored = 21;
constexpr int kPgBlocks = 512;
constexpr int kPgThreads = 384;
#ifndef FM_PG_TILE32
#define FM_PG_TILE32 16
#endif
#ifndef FM_PG_UNROLL
#define FM_PG_UNROLL 4
#endif
template <typename T>
constexpr int pg_tile() {
    return sizeof(T) == 4 ? FM_PG_TILE32 : 8;
}
template <typename T, int NM>
__global__ __launch_bounds__(kPgThreads) void poly_gram_kernel(uint32_t nn, const uint32_t* __restrict__ rp,
                                                               const uint32_t* __restrict__ nptr,
                                                               const uint32_t* __restrict__ ncol,
                                                               const T* __restrict__ val,
                                                               const float4* __restrict__ xn,
                                                               const T* __restrict__ S, uint32_t ld,
                                                               double* __restrict__ partials) {
    constexpr int TILE = pg_tile<T>();
    constexpr int CAPB = 20 * TILE;           // staged column blocks per tile
    constexpr int MP = 3 * NM;                // polynomial columns
    constexpr int NCOL = MP + kPgStored;      // columns of a G row
    constexpr int RG = (NM + 4) / 5;          // row groups of five per component
    constexpr int CG = (NCOL + 6) / 7;        // column groups of seven
    constexpr int NU = 3 * RG * CG;           // update threads
    static_assert(NU <= kPgThreads && TILE * NM <= 2 * kPgThreads, "tile work exceeds the block");
    __shared__ T s_phi[3][TILE][RG * 5];      // zeta_k,c at the tile's nodes (zero padded)
    __shared__ T s_h[TILE][3][CG * 7];        // H rows, then the stored values (zero padded)
    __shared__ T s_k[CAPB * 9];               // the tile's value range
    __shared__ T s_pw[CAPB][16];              // per column block: x^0..4, y^0..4, z^0..4, mask
    const int t = threadIdx.x;
    const uint32_t rbase = rp[0];
    // the update thread's tile of G: component c, rows 5 rg.., columns 7 cg..
    const int uc = t / (RG * CG), urg = (t % (RG * CG)) / CG, ucg = t % CG;
    double accd[5][7];
#pragma unroll
    for (int i = 0; i < 5; ++i)
#pragma unroll
        for (int j = 0; j < 7; ++j) accd[i][j] = 0.0;
    // the padding of both arrays stays zero
    for (int i = t; i < 3 * TILE * RG * 5; i += kPgThreads) (&s_phi[0][0][0])[i] = (T)0;
    for (int i = t; i < TILE * 3 * CG * 7; i += kPgThreads) (&s_h[0][0][0])[i] = (T)0;
    const uint32_t ntile = (nn + TILE - 1) / TILE;
    for (uint32_t tile = blockIdx.x; tile < ntile; tile += gridDim.x) {
        const uint32_t v0 = tile * TILE, v1 = min(v0 + TILE, nn);
        const uint32_t bs = nptr[v0], be = nptr[v1], nbt = be - bs;
        const bool staged = nbt <= (uint32_t)CAPB;  // block-uniform
        if (staged) {
            const T* src = val + rbase + 9 * (size_t)bs;
            for (uint32_t i = t; i < 9 * nbt; i += kPgThreads) s_k[i] = src[i];
            for (uint32_t q = t; q < nbt; q += kPgThreads) {
                const float4 c4 = xn[ncol[bs + q]];
                T p[5];
                powers5((T)c4.x, p);
#pragma unroll
                for (int e = 0; e < 5; ++e) s_pw[q][e] = p[e];
                powers5((T)c4.y, p);
#pragma unroll
                for (int e = 0; e < 5; ++e) s_pw[q][5 + e] = p[e];
                powers5((T)c4.z, p);
#pragma unroll
                for (int e = 0; e < 5; ++e) s_pw[q][10 + e] = p[e];
                s_pw[q][15] = (T)c4.w;
            }
        }
        // (2) the stored values of the tile's rows (zero past the last node)
        for (int w = t; w < kPgStored * TILE * 3; w += kPgThreads) {
            const int j = w / (TILE * 3), vr = w % (TILE * 3);
            const uint32_t row = 3 * v0 + vr;
            s_h[vr / 3][vr % 3][MP + j] = row < 3 * nn ? S[(size_t)j * ld + row] : (T)0;
        }
        __syncthreads();
        // (1) monomials and H rows: work item (v, k), two rounds
        for (int w = t; w < TILE * NM; w += kPgThreads) {
            const int v = w / NM, k = w % NM;
            const uint32_t node = v0 + v;
            T h[3][3];
#pragma unroll
            for (int r = 0; r < 3; ++r)
#pragma unroll
                for (int c = 0; c < 3; ++c) h[r][c] = (T)0;
            T ph = (T)0;
            int pmask = 7;
            if (node < nn) {
                const int code = kMonoCodeDev[k];
                const int ex = code & 7, ey = (code >> 3) & 7, ez = code >> 6;
                const float4 c4 = xn[node];
                ph = Powers<T>(c4).mono_rt(code);
                pmask = (int)c4.w;
                const uint32_t b0 = nptr[node], nb = nptr[node + 1] - b0;
                if (staged) {
                    const T* a = s_k + 9 * (b0 - bs);
#pragma unroll FM_PG_UNROLL
                    for (uint32_t q = 0; q < nb; ++q) {
                        const T* pw = s_pw[b0 - bs + q];
                        const T m = (pw[ex] * pw[5 + ey]) * pw[10 + ez];
                        const int um = (int)pw[15];
                        T f[3];
#pragma unroll
                        for (int c = 0; c < 3; ++c) f[c] = (um >> c & 1) ? (T)0 : m;
#pragma unroll
                        for (int r = 0; r < 3; ++r)
#pragma unroll
                            for (int c = 0; c < 3; ++c) h[r][c] += a[3 * nb * r + 3 * q + c] * f[c];
                    }
                } else {
                    const T* a = val + rbase + 9 * (size_t)b0;
                    for (uint32_t q = 0; q < nb; ++q) {
                        const float4 u4 = xn[ncol[b0 + q]];
                        const T m = Powers<T>(u4).mono_rt(code);
                        const int um = (int)u4.w;
                        T f[3];
#pragma unroll
                        for (int c = 0; c < 3; ++c) f[c] = (um >> c & 1) ? (T)0 : m;
#pragma unroll
                        for (int r = 0; r < 3; ++r)
#pragma unroll
                            for (int c = 0; c < 3; ++c) h[r][c] += a[3 * nb * r + 3 * q + c] * f[c];
                    }
                }
            }
#pragma unroll
            for (int c = 0; c < 3; ++c) s_phi[c][v][k] = (pmask >> c & 1) ? (T)0 : ph;
#pragma unroll
            for (int r = 0; r < 3; ++r)
#pragma unroll
                for (int c = 0; c < 3; ++c) s_h[v][r][c * NM + k] = h[r][c];
        }
        __syncthreads();
        // (3) the update threads add the tile to their 5 x 7 tiles of G
        if (t < NU) {
            T acc[5][7];
#pragma unroll
            for (int i = 0; i < 5; ++i)
#pragma unroll
                for (int j = 0; j < 7; ++j) acc[i][j] = (T)0;
#pragma unroll 4
            for (int v = 0; v < TILE; ++v) {
                T p[5], hv[7];
#pragma unroll
                for (int i = 0; i < 5; ++i) p[i] = s_phi[uc][v][5 * urg + i];
#pragma unroll
                for (int j = 0; j < 7; ++j) hv[j] = s_h[v][uc][7 * ucg + j];
#pragma unroll
                for (int i = 0; i < 5; ++i)
#pragma unroll
                    for (int j = 0; j < 7; ++j) acc[i][j] = fma(p[i], hv[j], acc[i][j]);
            }
#pragma unroll
            for (int i = 0; i < 5; ++i)
#pragma unroll
                for (int j = 0; j < 7; ++j) accd[i][j] += (double)acc[i][j];
        }
        __syncthreads();  // the next tile overwrites the staged arrays
    }
    if (t < NU) {
        double* out = partials + (size_t)blockIdx.x * MP * NCOL;
#pragma unroll
        for (int i = 0; i < 5; ++i)
#pragma unroll
            for (int j = 0; j < 7; ++j) {
                const int k = 5 * urg + i, col = 7 * ucg + j;
                if (k < NM && col < NCOL) out[(size_t)(uc * NM + k) * NCOL + col] = accd[i][j];
            }
    }
}

// ---------------------------------------------------------------------------
// additive coarse correction of the preconditioner: z = M^{-1} r + Z E^{-1}
// Z^T r with Z the polynomial fields (3 NM columns) and E = Z^T A Z
// ---------------------------------------------------------------------------
// blocks of the moment passes (and partials of the coarse solve): 256 cut
// the coarse kernels from 46 to 42 us per iteration at 128^3 hexes, but the
// changed rounding of the moments took 2,225 PCG iterations over the 20 load
// steps against 2,209
#ifndef FM_MOMENT_BLOCKS
#define FM_MOMENT_BLOCKS 512
#endif
constexpr int kMomentBlocks = FM_MOMENT_BLOCKS;

// DPP lane exchanges of a wave (GFX9 controls): the row-masked source value,
// zero in the rows the mask leaves out
template <int CTRL, int ROWS>
__device__ __forceinline__ float dpp_src(float v) {
    return __builtin_bit_cast(float, __builtin_amdgcn_update_dpp(0, __builtin_bit_cast(int, v), CTRL, ROWS, 0xF, false));
}
template <int CTRL, int ROWS>
__device__ __forceinline__ double dpp_src(double v) {
    const unsigned long long u = __builtin_bit_cast(unsigned long long, v);
    const unsigned lo = (unsigned)__builtin_amdgcn_update_dpp(0, (int)(unsigned)u, CTRL, ROWS, 0xF, false);
    const unsigned hi = (unsigned)__builtin_amdgcn_update_dpp(0, (int)(unsigned)(u >> 32), CTRL, ROWS, 0xF, false);
    return __builtin_bit_cast(double, (unsigned long long)lo | ((unsigned long long)hi << 32));
}
// the sum over the wave's 64 lanes, complete in lane 63: pairs and quads
// (quad permutes), rows of 16 (half-row and row mirrors), then row 0 into
// rows 1 and 3 (broadcast of lane 15) and rows 0 + 1 into rows 2 and 3
// (broadcast of lane 31); VALU only, no LDS round trip
template <typename T>
__device__ __forceinline__ T wave_sum_to_63(T v) {
    v += dpp_src<0xB1, 0xF>(v);   // quad_perm(1, 0, 3, 2)
    v += dpp_src<0x4E, 0xF>(v);   // quad_perm(2, 3, 0, 1)
    v += dpp_src<0x141, 0xF>(v);  // row_half_mirror
    v += dpp_src<0x140, 0xF>(v);  // row_mirror
    v += dpp_src<0x142, 0xA>(v);  // row_bcast:15
    v += dpp_src<0x143, 0xC>(v);  // row_bcast:31
    return v;
}

// The moments Z^T r (3 NM values) of one thread's nodes accumulate in T
// (FP32 products and sums over the ~16 nodes of a thread, then over the
// wave, on the FP32 path: a preconditioner component, as exact as the FP32
// sweeps beside it); the block sums are FP64, written as the block's
// partials.  The wave sums are DPP reductions (a butterfly of LDS lane
// exchanges waits for the exchange at every step).
template <typename T, int NM>
struct MomentAcc {
    T acc[3][NM];
    __device__ __forceinline__ MomentAcc() {
#pragma unroll
        for (int c = 0; c < 3; ++c)
#pragma unroll
            for (int k = 0; k < NM; ++k) acc[c][k] = (T)0;
    }
    // r: the node's three components, masked by the caller
    __device__ __forceinline__ void add(const Powers<T>& pw, const T (&rv)[3]) {
#pragma unroll
        for (int k = 0; k < NM; ++k) {
            const T m = pw.mono(mono_code(k));
#pragma unroll
            for (int c = 0; c < 3; ++c) acc[c][k] = fma(m, rv[c], acc[c][k]);
        }
    }
    __device__ __forceinline__ void store(double* __restrict__ partials) {
        __shared__ double s_w[4][3 * NM];
        const int w = threadIdx.x >> 6, lane = threadIdx.x & 63;
#pragma unroll
        for (int c = 0; c < 3; ++c)
#pragma unroll
            for (int k = 0; k < NM; ++k) acc[c][k] = wave_sum_to_63(acc[c][k]);
        if (lane == 63)
#pragma unroll
            for (int c = 0; c < 3; ++c)
#pragma unroll
                for (int k = 0; k < NM; ++k) s_w[w][c * NM + k] = (double)acc[c][k];
        __syncthreads();
        for (int i = threadIdx.x; i < 3 * NM; i += blockDim.x)
            partials[(size_t)blockIdx.x * 3 * NM + i] = (s_w[0][i] + s_w[1][i]) + (s_w[2][i] + s_w[3][i]);
    }
};

// block partials of Z^T r (3 NM moments, FP64), grid-stride over the nodes
template <typename T, int NM>
__global__ __launch_bounds__(256) void coarse_moments_kernel(uint32_t nn, const float4* __restrict__ xn,
                                                             const T* __restrict__ r, double* __restrict__ partials,
                                                             const PcgState* __restrict__ st) {
    if (st->done) return;
    MomentAcc<T, NM> M;
    for (uint32_t v = blockIdx.x * blockDim.x + threadIdx.x; v < nn; v += gridDim.x * blockDim.x) {
        const float4 c4 = xn[v];
        const Powers<T> pw(c4);
        const int mask = (int)c4.w;
        T rv[3];
#pragma unroll
        for (int c = 0; c < 3; ++c) rv[c] = (mask >> c & 1) ? (T)0 : r[3 * (size_t)v + c];
        M.add(pw, rv);
    }
    M.store(partials);
}

// the PCG update x += alpha p, r -= alpha Ap over whole nodes, with the block
// partials of the moments Z^T r of the updated r (the additive coarse
// correction's input) in the same pass.  A thread takes every G-th node
// (G = the grid's threads, ~16 nodes each); the next node's five loads are
// issued before the current node's moments are summed, so that two waves
// per SIMD keep the memory busy.
template <typename T, int NM>
__global__ __launch_bounds__(256) void pcg_xr_update_moments_kernel(uint32_t nn, const T* __restrict__ p,
                                                                    const T* __restrict__ Ap, T* __restrict__ x,
                                                                    T* __restrict__ r, const float4* __restrict__ xn,
                                                                    double* __restrict__ partials,
                                                                    const PcgState* __restrict__ st) {
    if (st->done) return;
    const T alpha = (T)st->alpha;
    MomentAcc<T, NM> M;
    const uint32_t G = gridDim.x * blockDim.x;
    struct In {
        float4 c4;
        T p[3], ap[3], x[3], r[3];
    };
    auto load = [&](uint32_t v, In& d) {
        const uint32_t w = min(v, nn - 1);  // past the end: a valid node, never stored
        const size_t i = 3 * (size_t)w;
        d.c4 = xn[w];
#pragma unroll
        for (int c = 0; c < 3; ++c) {
            d.p[c] = p[i + c];
            d.ap[c] = Ap[i + c];
            d.x[c] = x[i + c];
            d.r[c] = r[i + c];
        }
    };
    uint32_t v = blockIdx.x * blockDim.x + threadIdx.x;
    In cur;
    load(v, cur);
    for (; v < nn; v += G) {
        In nxt;
        load(v + G, nxt);
        const size_t i = 3 * (size_t)v;
        T rv[3];
#pragma unroll
        for (int c = 0; c < 3; ++c) {
            x[i + c] = cur.x[c] + alpha * cur.p[c];
            const T rn = cur.r[c] - alpha * cur.ap[c];
            r[i + c] = rn;
            rv[c] = rn;
        }
        const Powers<T> pw(cur.c4);
        const int mask = (int)cur.c4.w;
#pragma unroll
        for (int c = 0; c < 3; ++c) rv[c] = (mask >> c & 1) ? (T)0 : rv[c];
        M.add(pw, rv);
        cur = nxt;
    }
    M.store(partials);
}

// one block of 1024 threads: moments g = sum of the nb block partials,
// coefficients c = E^{-1} g (E^{-1} symmetric) and g.c (the coarse part of
// r.z) into the PCG state; coef holds the m FP64 coefficients followed by
// the same rounded to FP32 (the FP32 p update's).  Thread (s, i) takes
// moment i of every S-th partial, sixteen loads in flight, then entries i of
// every S-th row of E^{-1}; the S slices are added in slice order.
template <bool CROSS>
__global__ __launch_bounds__(1024) void coarse_solve_kernel(uint32_t nb, int m, const double* __restrict__ partials,
                                                            const double* __restrict__ Einv,
                                                            double* __restrict__ coef, PcgState* __restrict__ st) {
    if (st->done) return;
    __shared__ double s_part[1024];
    __shared__ double s_g[3 * kMaxMono];
    int mp = 16;
    while (mp < m) mp <<= 1;
    const int S = (int)blockDim.x / mp, i = threadIdx.x % mp, sl = threadIdx.x / mp;
    double a[4] = {0.0, 0.0, 0.0, 0.0};
    if (i < m) {
        uint32_t b = sl;
        for (; b + 15 * S < nb; b += 16 * S) {
            double t[16];
#pragma unroll
            for (int u = 0; u < 16; ++u) t[u] = partials[(size_t)(b + u * S) * m + i];
#pragma unroll
            for (int u = 0; u < 16; ++u) a[u & 3] += t[u];
        }
        for (; b < nb; b += S) a[0] += partials[(size_t)b * m + i];
    }
    s_part[threadIdx.x] = (a[0] + a[1]) + (a[2] + a[3]);
    __syncthreads();
    if (threadIdx.x < m) {
        double g = 0.0;
        for (int q = 0; q < S; ++q) g += s_part[q * mp + threadIdx.x];
        s_g[threadIdx.x] = g;
    }
    __syncthreads();
    double c = 0.0;
    if (i < m)
        for (int j = sl; j < m; j += S) c += Einv[(size_t)j * m + i] * s_g[j];
    s_part[threadIdx.x] = c;
    __syncthreads();
    double gc = 0.0, go = 0.0;
    if (threadIdx.x < m) {
        double ci = 0.0;
        for (int q = 0; q < S; ++q) ci += s_part[q * mp + threadIdx.x];
        // CROSS: the previous coefficients against the new moments (the
        // flexible beta's r_new.(Z c_old)); read before they are overwritten
        if (CROSS) go = coef[threadIdx.x] * s_g[threadIdx.x];
        coef[threadIdx.x] = ci;
        reinterpret_cast<float*>(coef + m)[threadIdx.x] = (float)ci;
        gc = ci * s_g[threadIdx.x];
    }
    __syncthreads();  // block_sum's slots are its own; the barrier orders the reads above
    const double t = block_sum(gc);
    if (threadIdx.x == 0) st->rz_coarse = t;
    if (CROSS) {
        __syncthreads();  // block_sum's shared scratch is reused
        const double to = block_sum(go);
        if (threadIdx.x == 0) st->rz_cross = to;
    }
}

// p = (y + Z c) + beta p, or p = y + Z c on the first (beta ignored); one
// thread per node, so the coefficient loads are the same for every lane
// (coefT: the coefficients in T, see coarse_solve_kernel).  lz (optional):
// the Lanczos vector z_j / sqrt(r_j.z_j) of this iteration j = st->iters
// (z = y + Z c, the preconditioned residual) goes to column j while j < lz_cap.
template <typename T, int NM>
__global__ void pcg_p_update_coarse_kernel(uint32_t nn, const T* __restrict__ y, T* __restrict__ p,
                                           const float4* __restrict__ xn, const T* __restrict__ coefT, int first,
                                           const PcgState* __restrict__ st, T* __restrict__ lz, uint32_t lz_cap) {
    if (st->done) return;
    const uint32_t v = xcd_remap(blockIdx.x, gridDim.x) * blockDim.x + threadIdx.x;
    if (v >= nn) return;
    const float4 c4 = xn[v];
    const Powers<T> pw(c4);
    const int mask = (int)c4.w;
    const T beta = first ? (T)0 : (T)st->beta;
    const uint32_t j = (uint32_t)st->iters;
    T* lzj = lz && j < lz_cap ? lz + (size_t)j * 3 * nn : nullptr;
    const T zs = lzj ? (T)(1.0 / sqrt(st->rz)) : (T)0;
#pragma unroll
    for (int c = 0; c < 3; ++c) {
        T zc = (T)0;
#pragma unroll
        for (int k = 0; k < NM; ++k) zc += coefT[c * NM + k] * pw.mono(mono_code(k));
        if (mask >> c & 1) zc = (T)0;
        const size_t i = 3 * (size_t)v + c;
        const T z = y[i] + zc;
        p[i] = first ? z : z + beta * p[i];
        if (lzj) lzj[i] = z * zs;
    }
}

// block partials of G = V^T W (V: m columns, W: mw columns, both with leading
// dimension n, m and mw <= 128), FP64 accumulation.  A block takes a
// contiguous row range in chunks of kGramRows rows, staged in LDS row by row
// (row r of V at sV[r mp ..], of W at sW[r mwp ..], padded to multiples of 4
// columns), so that the lanes of a wave read their four consecutive columns
// of W from consecutive 16-byte slots (conflict-free) and share V's (a
// broadcast).  Thread t owns the 4x4 tile (t / (mwp / 4), t % (mwp / 4)) of
// G; the next chunk's loads are in flight while the current one is summed.
// blockDim: gram_threads(m, mw), so that every thread stages at most
// kGramStage values per chunk.
constexpr int kGramRows = 32;
constexpr int kGramStage = 16;
template <typename T>
__global__ __launch_bounds__(1024) void gram_kernel(uint32_t n, int m, int mw, const T* __restrict__ V,
                                                    const T* __restrict__ W, double* __restrict__ partials) {
    extern __shared__ __align__(16) unsigned char gram_smem[];
    using V4 = typename Vec16<T>::type;  // 16 bytes: 4 floats or 2 doubles
    constexpr int PV = 16 / (int)sizeof(T);
    const int mp = (m + 3) & ~3, mwp = (mw + 3) & ~3;
    T* sV = reinterpret_cast<T*>(gram_smem);
    T* sW = sV + kGramRows * mp;
    const int tj_n = mwp / 4, ntile = (mp / 4) * tj_n;
    const int t = threadIdx.x;
    const bool active = t < ntile;
    const int i0 = active ? (t / tj_n) * 4 : 0, j0 = active ? (t % tj_n) * 4 : 0;
    const uint32_t per = ((n + gridDim.x - 1) / gridDim.x + kGramRows - 1) / kGramRows * kGramRows;
    const uint32_t r0 = blockIdx.x * per, r1 = min(n, r0 + per);
    const int nv = mp * kGramRows, total = (mp + mwp) * kGramRows;
    double acc[4][4];
#pragma unroll
    for (int a = 0; a < 4; ++a)
#pragma unroll
        for (int b = 0; b < 4; ++b) acc[a][b] = 0.0;
    T buf[kGramStage];
    auto load = [&](uint32_t rc) {
#pragma unroll
        for (int q = 0; q < kGramStage; ++q) {
            const int e = t + q * (int)blockDim.x;
            T v = (T)0;
            if (e < total) {
                const bool inv = e < nv;
                const int f = inv ? e : e - nv;
                const int j = f / kGramRows, r = f % kGramRows;
                if (rc + r < r1 && j < (inv ? m : mw)) v = (inv ? V : W)[(size_t)j * n + rc + r];
            }
            buf[q] = v;
        }
    };
    auto store = [&]() {
#pragma unroll
        for (int q = 0; q < kGramStage; ++q) {
            const int e = t + q * (int)blockDim.x;
            if (e < total) {
                const bool inv = e < nv;
                const int f = inv ? e : e - nv;
                const int j = f / kGramRows, r = f % kGramRows;
                if (inv)
                    sV[r * mp + j] = buf[q];
                else
                    sW[r * mwp + j] = buf[q];
            }
        }
    };
    if (r0 < r1) load(r0);
    for (uint32_t rc = r0; rc < r1; rc += kGramRows) {
        store();
        __syncthreads();
        if (rc + kGramRows < r1) load(rc + kGramRows);
        if (active) {
            for (int r = 0; r < kGramRows; ++r) {
                double a[4], b[4];
#pragma unroll
                for (int h = 0; h < 4 / PV; ++h) {
                    const V4 va = *reinterpret_cast<const V4*>(sV + r * mp + i0 + h * PV);
                    const V4 vb = *reinterpret_cast<const V4*>(sW + r * mwp + j0 + h * PV);
#pragma unroll
                    for (int u = 0; u < PV; ++u) {
                        a[h * PV + u] = (double)va[u];
                        b[h * PV + u] = (double)vb[u];
                    }
                }
#pragma unroll
                for (int u = 0; u < 4; ++u)
#pragma unroll
                    for (int w = 0; w < 4; ++w) acc[u][w] += a[u] * b[w];
            }
        }
        __syncthreads();
    }
    if (!active) return;
    double* out = partials + (size_t)blockIdx.x * m * mw;
    for (int u = 0; u < 4; ++u)
        for (int w = 0; w < 4; ++w)
            if (i0 + u < m && j0 + w < mw) out[(i0 + u) * mw + j0 + w] = acc[u][w];
}

// threads of a Gram block: one per 4x4 tile, and enough that each stages at
// most kGramStage values of a chunk
static int gram_threads(int m, int mw) {
    const int mp = (m + 3) & ~3, mwp = (mw + 3) & ~3;
    const int ntile = (mp / 4) * (mwp / 4);
    const int stage = ((mp + mwp) * kGramRows + kGramStage - 1) / kGramStage;
    const int t = std::max(ntile, stage);
    return std::min(1024, (t + 63) / 64 * 64);
}

// G[p] = sum over the blocks of partials[b][p]: eight slices of every
// eighth block each, added in slice order
__global__ __launch_bounds__(256) void gram_reduce_kernel(uint32_t nb, uint32_t np,
                                                          const double* __restrict__ partials,
                                                          double* __restrict__ G) {
    __shared__ double s_sl[8][32];
    const int e = threadIdx.x & 31, sl = threadIdx.x >> 5;
    const uint32_t p = blockIdx.x * 32 + e;
    double a[4] = {0.0, 0.0, 0.0, 0.0};
    if (p < np) {
        uint32_t b = sl;
        for (; b + 8 * 15 < nb; b += 8 * 16) {
            double t[16];
#pragma unroll
            for (int u = 0; u < 16; ++u) t[u] = partials[(size_t)(b + 8 * u) * np + p];
#pragma unroll
            for (int u = 0; u < 16; ++u) a[u & 3] += t[u];
        }
        for (; b < nb; b += 8) a[0] += partials[(size_t)b * np + p];
    }
    s_sl[sl][e] = (a[0] + a[1]) + (a[2] + a[3]);
    __syncthreads();
    if (sl == 0 && p < np) {
        double g = 0.0;
        for (int q = 0; q < 8; ++q) g += s_sl[q][e];
        G[p] = g;
    }
}

// x = sum_k c[k] V[:, col[k]] (FP64 accumulation)
template <typename T>
__global__ void basis_combine_kernel(uint32_t n, int k, const int* __restrict__ col, const double* __restrict__ c,
                                     const T* __restrict__ V, T* __restrict__ x) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    double acc = 0.0;
    for (int q = 0; q < k; ++q) acc += c[q] * (double)V[(size_t)col[q] * n + i];
    x[i] = (T)acc;
}

__global__ void f64_to_f32_kernel(uint32_t n, const double* __restrict__ s, double scale, float* __restrict__ d) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) d[i] = (float)(s[i] * scale);
}

__global__ void f32_to_f64_kernel(uint32_t n, const float* __restrict__ s, double scale, double* __restrict__ d) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) d[i] = (double)s[i] * scale;
}

// d = (double)s with block partials of r.d (new d) and, when old != null,
// of r.d with the d it overwrites (the FP64 r.z of a PCG whose
// preconditioner runs in FP32, and the r.z_old of its flexible beta)
__global__ __launch_bounds__(256) void f32_to_f64_dot_kernel(uint32_t n, const float* __restrict__ s,
                                                             const double* __restrict__ r, double* __restrict__ d,
                                                             double* __restrict__ partials,
                                                             double* __restrict__ old_partials,
                                                             const PcgState* __restrict__ st) {
    if (st && st->done) return;
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    double part = 0.0, part_old = 0.0;
    if (i < n) {
        const double v = (double)s[i];
        if (old_partials) part_old = r[i] * d[i];
        d[i] = v;
        part = r[i] * v;
    }
    const double bs = block_sum(part);
    if (threadIdx.x == 0) partials[blockIdx.x] = bs;
    if (old_partials) {
        __syncthreads();  // block_sum's shared scratch is reused
        const double bo = block_sum(part_old);
        if (threadIdx.x == 0) old_partials[blockIdx.x] = bo;
    }
}

__global__ void scale_kernel(uint32_t n, const double* __restrict__ a, double alpha, double* __restrict__ y) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) y[i] = alpha * a[i];
}

__global__ void axpy_to_kernel(uint32_t n, const double* __restrict__ a, double alpha, const double* __restrict__ b,
                               double* __restrict__ y) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) y[i] = a[i] + alpha * b[i];
}

__global__ void dir_increment_kernel(uint32_t n_dir, const uint32_t* __restrict__ dofs,
                                     const double* __restrict__ vals, double scale, const double* __restrict__ u,
                                     double* __restrict__ dD) {
    const uint32_t k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= n_dir) return;
    const uint32_t d = dofs[k];
    dD[d] = scale * (vals ? vals[k] : 0.0) - u[d];
}

__global__ void sub_masked_kernel(uint32_t n, const double* __restrict__ R, const double* __restrict__ lift,
                                  const uint8_t* __restrict__ isdir, double* __restrict__ out) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) out[i] = isdir[i] ? 0.0 : R[i] - lift[i];
}

__global__ void apply_predictor_kernel(uint32_t n, const double* __restrict__ du, const double* __restrict__ dD,
                                       const uint8_t* __restrict__ isdir, double* __restrict__ u) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) u[i] += isdir[i] ? dD[i] : du[i];
}

// dL/dmu_e = -lam_e . V (F g - F^{-T} g), dL/dlam_e = -lam_e . V lnJ F^{-T} g
__global__ void adjoint_elem_kernel(ElemArgs A, const uint32_t* __restrict__ eperm, const double* __restrict__ lv,
                                    double* __restrict__ dmu, double* __restrict__ dlam) {
    const uint32_t t = xcd_remap(blockIdx.x, gridDim.x) * blockDim.x + threadIdx.x;
    if (t >= A.ne) return;
    Elem E;
    elem_eval(A, t, E);
    double dm = 0.0, dl = 0.0;
    if (E.ok) {
#pragma unroll
        for (int q = 0; q < 4; ++q)
#pragma unroll
            for (int i = 0; i < 3; ++i) {
                const double l = lv[3 * (size_t)E.node[q] + i];
                dm += l * E.V * (E.Fg[q][i] - E.a[q][i]);
                dl += l * E.V * E.lnJ * E.a[q][i];
            }
    }
    const uint32_t e = eperm[t];
    if (dmu) dmu[e] = -dm;
    if (dlam) dlam[e] = -dl;
}

// Stored energy V_e psi(F_e), psi = mu/2 (tr F^T F - 3) - mu ln J + lam/2 ln^2 J
// with the same ln max(J, 0.01) clamp as the stress; block partials of the sum.
__global__ __launch_bounds__(256) void energy_elem_kernel(ElemArgs A, double* __restrict__ partials) {
    const uint32_t t = blockIdx.x * blockDim.x + threadIdx.x;
    double part = 0.0;
    if (t < A.ne) {
        Elem E;
        double F[9];
        elem_eval(A, t, E, F);
        if (E.ok) {
            double I1 = 0.0;
#pragma unroll
            for (int k = 0; k < 9; ++k) I1 += F[k] * F[k];
            const double lnJ = log(fmax(fabs(det3(F)), 0.01));
            part = E.V * (0.5 * E.mu * (I1 - 3.0) - E.mu * lnJ + 0.5 * E.lam * lnJ * lnJ);
        }
    }
    const double s = block_sum(part);
    if (threadIdx.x == 0) partials[blockIdx.x] = s;
}

// ---------------------------------------------------------------------------
// FP64 rotation part: polar factor by the scaled Newton iteration
// X <- (zeta X + X^{-T}/zeta)/2 (Higham 1986), angle-axis logarithm with the
// symmetric-part branch near pi, Rodrigues exponential with
// (1 - cos t)/t^2 = (sin(t/2)/(t/2))^2 / 2 to avoid cancellation.
// ---------------------------------------------------------------------------
__device__ int polar_dev(const double* F, double* R) {
    double X[9];
    for (int k = 0; k < 9; ++k) X[k] = F[k];
    for (int it = 0; it < 60; ++it) {
        const double det = det3(X);
        if (fabs(det) < 1e-300) return 1;
        const double id = 1.0 / det;
        double XiT[9];  // X^{-T} = cof(X) / det
        XiT[0] = (X[4] * X[8] - X[5] * X[7]) * id;
        XiT[1] = (X[5] * X[6] - X[3] * X[8]) * id;
        XiT[2] = (X[3] * X[7] - X[4] * X[6]) * id;
        XiT[3] = (X[2] * X[7] - X[1] * X[8]) * id;
        XiT[4] = (X[0] * X[8] - X[2] * X[6]) * id;
        XiT[5] = (X[1] * X[6] - X[0] * X[7]) * id;
        XiT[6] = (X[1] * X[5] - X[2] * X[4]) * id;
        XiT[7] = (X[2] * X[3] - X[0] * X[5]) * id;
        XiT[8] = (X[0] * X[4] - X[1] * X[3]) * id;
        double nx = 0.0, ni = 0.0;
        for (int k = 0; k < 9; ++k) {
            nx += X[k] * X[k];
            ni += XiT[k] * XiT[k];
        }
        const double zeta = (it < 8) ? sqrt(sqrt(ni / nx)) : 1.0;
        double diff = 0.0;
        for (int k = 0; k < 9; ++k) {
            const double xn = 0.5 * (zeta * X[k] + XiT[k] / zeta);
            diff += (xn - X[k]) * (xn - X[k]);
            X[k] = xn;
        }
        if (it >= 8 && diff < 1e-30) break;
    }
    for (int k = 0; k < 9; ++k) R[k] = X[k];
    return det3(R) > 0 ? 0 : 2;
}

__device__ void rot_log_dev(const double* R, double* w) {
    const double c = 0.5 * (R[0] + R[4] + R[8] - 1.0);
    const double v0 = 0.5 * (R[7] - R[5]), v1 = 0.5 * (R[2] - R[6]), v2 = 0.5 * (R[3] - R[1]);
    const double s = sqrt(v0 * v0 + v1 * v1 + v2 * v2);
    const double th = atan2(s, c);
    if (th < 1e-4) {
        const double t2 = th * th;
        const double f = 1.0 + t2 / 6.0 + 7.0 * t2 * t2 / 360.0;
        w[0] = f * v0;
        w[1] = f * v1;
        w[2] = f * v2;
        return;
    }
    if (c > -0.9) {
        const double f = th / s;
        w[0] = f * v0;
        w[1] = f * v1;
        w[2] = f * v2;
        return;
    }
    double B[9];
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j) B[i * 3 + j] = 0.5 * (R[i * 3 + j] + R[j * 3 + i]) - (i == j ? c : 0.0);
    int m = 0;
    if (B[4] > B[0]) m = 1;
    if (B[8] > B[m * 4]) m = 2;
    const double n0 = B[0 * 3 + m], n1 = B[1 * 3 + m], n2 = B[2 * 3 + m];
    const double nrm = sqrt(n0 * n0 + n1 * n1 + n2 * n2);
    const double sgn = (n0 * v0 + n1 * v1 + n2 * v2) < 0 ? -1.0 : 1.0;
    w[0] = sgn * th * n0 / nrm;
    w[1] = sgn * th * n1 / nrm;
    w[2] = sgn * th * n2 / nrm;
}

__device__ void rot_exp_dev(const double* w, double* R) {
    const double th2 = w[0] * w[0] + w[1] * w[1] + w[2] * w[2];
    const double th = sqrt(th2);
    double a, b;
    if (th < 1e-6) {
        a = 1.0 - th2 / 6.0;
        b = 0.5 - th2 / 24.0;
    } else {
        a = sin(th) / th;
        const double h = sin(0.5 * th) / (0.5 * th);
        b = 0.5 * h * h;
    }
    const double K[9] = {0, -w[2], w[1], w[2], 0, -w[0], -w[1], w[0], 0};
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j) {
            double k2 = 0.0;
            for (int k = 0; k < 3; ++k) k2 += K[i * 3 + k] * K[k * 3 + j];
            R[i * 3 + j] = (i == j ? 1.0 : 0.0) + a * K[i * 3 + j] + b * k2;
        }
}

__global__ void rot_batch_kernel(uint32_t n, const double* __restrict__ F, double alpha, double* __restrict__ Rout,
                                 double* __restrict__ wout, double* __restrict__ Qout) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    double Fi[9], R[9], w[3];
    for (int k = 0; k < 9; ++k) Fi[k] = F[9 * (size_t)i + k];
    polar_dev(Fi, R);
    rot_log_dev(R, w);
    if (Rout)
        for (int k = 0; k < 9; ++k) Rout[9 * (size_t)i + k] = R[k];
    if (wout)
        for (int k = 0; k < 3; ++k) wout[3 * (size_t)i + k] = w[k];
    if (Qout) {
        double aw[3] = {alpha * w[0], alpha * w[1], alpha * w[2]};
        double Q[9];
        rot_exp_dev(aw, Q);
        for (int k = 0; k < 9; ++k) Qout[9 * (size_t)i + k] = Q[k];
    }
}

// thread e takes element e (its colour-order index from eiperm): the node
// gathers of neighbouring elements share cache lines and the stores are contiguous
__global__ void elem_rot_kernel(ElemArgs A, const uint32_t* __restrict__ eiperm, double* __restrict__ wout) {
    const uint32_t e = blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= A.ne) return;
    Elem E;
    double F[9] = {1, 0, 0, 0, 1, 0, 0, 0, 1};
    elem_eval(A, eiperm[e], E, F);
    double R[9], w[3] = {0, 0, 0};
    if (E.ok && polar_dev(F, R) == 0) rot_log_dev(R, w);
    for (int k = 0; k < 3; ++k) wout[3 * (size_t)e + k] = w[k];
}

// ---------------------------------------------------------------------------
// Element fields of a displacement (FP64), the consumer of the converged
// state: StressExtractor::extract_element_fields (FEM/stress_extraction.hpp:
// 187-240) over the element state of update_element_states_rigorous
// (FEM/adaptive_fem_engine.hpp:2716-2794).  F is the exact P1 gradient, kept
// as computed (no projection onto det F = 1) with J = det F evaluated.
// ---------------------------------------------------------------------------
// largest eigenvalue of a symmetric 3x3 matrix (upper triangle a00 a01 a02 a11
// a12 a22) by cyclic Jacobi rotations (Golub & Van Loan, Alg. 8.5.3) until the
// off-diagonal part is below 1e-18 of the matrix norm: accurate to rounding
// also for repeated eigenvalues, where the trigonometric closed form loses
// half the digits
__device__ double sym3_max_eig(double a00, double a01, double a02, double a11, double a12, double a22) {
    double A[3][3] = {{a00, a01, a02}, {a01, a11, a12}, {a02, a12, a22}};
    for (int sweep = 0; sweep < 60; ++sweep) {
        const double off = A[0][1] * A[0][1] + A[0][2] * A[0][2] + A[1][2] * A[1][2];
        const double dg = A[0][0] * A[0][0] + A[1][1] * A[1][1] + A[2][2] * A[2][2];
        if (off <= 1e-36 * (dg + 2.0 * off)) break;
#pragma unroll
        for (int p = 0; p < 2; ++p)
#pragma unroll
            for (int q = p + 1; q < 3; ++q) {
                if (A[p][q] == 0.0) continue;
                const double th = (A[q][q] - A[p][p]) / (2.0 * A[p][q]);
                const double t = (th >= 0 ? 1.0 : -1.0) / (fabs(th) + sqrt(th * th + 1.0));
                const double c = 1.0 / sqrt(t * t + 1.0), sn = t * c;
#pragma unroll
                for (int k = 0; k < 3; ++k) {
                    const double akp = A[k][p], akq = A[k][q];
                    A[k][p] = c * akp - sn * akq;
                    A[k][q] = sn * akp + c * akq;
                }
#pragma unroll
                for (int k = 0; k < 3; ++k) {
                    const double apk = A[p][k], aqk = A[q][k];
                    A[p][k] = c * apk - sn * aqk;
                    A[q][k] = sn * apk + c * aqk;
                }
            }
    }
    return fmax(A[0][0], fmax(A[1][1], A[2][2]));
}

// out (kFieldCount x ne, SoA, element order): det F, Cauchy stress (xx, yy,
// zz, yz, xz, xy; sigma = P F^T / max(|det F|, 0.01), FEM/constitutive_models.hpp:
// 116-148), von Mises (FEM/stress_extraction.hpp:143-157), pressure tr/3,
// largest principal stress, strain energy psi, maximum stretch
// sqrt(max(eig_max(F^T F), 1)) (:160-176), F row-major.  Degenerate elements
// (volume < 1e-30) give zeros.  Thread e takes element e (colour-order index
// from eiperm), as elem_rot_kernel.
__global__ __launch_bounds__(128) void element_fields_kernel(ElemArgs A, const uint32_t* __restrict__ eiperm,
                                                             double* __restrict__ out) {
    const uint32_t e = blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= A.ne) return;
    Elem E;
    double F[9];
    elem_eval(A, eiperm[e], E, F);
    double f[kFieldCount];
#pragma unroll
    for (int k = 0; k < kFieldCount; ++k) f[k] = 0.0;
    if (E.ok) {
        double FiT[9], lnJ;
        bool cl;
        kinematics(F, FiT, lnJ, cl);
        double P[9];
        const double sP = E.lam * lnJ - E.mu;
#pragma unroll
        for (int k = 0; k < 9; ++k) P[k] = E.mu * F[k] + sP * FiT[k];
        const double det = det3(F);
        const double inv_J = 1.0 / fmax(fabs(det), 0.01);
        double sg[9];
#pragma unroll
        for (int i = 0; i < 3; ++i)
#pragma unroll
            for (int j = 0; j < 3; ++j)
                sg[i * 3 + j] = inv_J * (P[i * 3 + 0] * F[j * 3 + 0] + P[i * 3 + 1] * F[j * 3 + 1] +
                                         P[i * 3 + 2] * F[j * 3 + 2]);
        const double pr = (sg[0] + sg[4] + sg[8]) / 3.0;
        const double d0 = sg[0] - pr, d4 = sg[4] - pr, d8 = sg[8] - pr;
        double I1 = 0.0;
#pragma unroll
        for (int k = 0; k < 9; ++k) I1 += F[k] * F[k];
        double Cm[9];
#pragma unroll
        for (int i = 0; i < 3; ++i)
#pragma unroll
            for (int j = 0; j < 3; ++j)
                Cm[i * 3 + j] = F[0 * 3 + i] * F[0 * 3 + j] + F[1 * 3 + i] * F[1 * 3 + j] + F[2 * 3 + i] * F[2 * 3 + j];
        f[0] = det;
        f[1] = sg[0];
        f[2] = sg[4];
        f[3] = sg[8];
        f[4] = sg[5];
        f[5] = sg[2];
        f[6] = sg[1];
        f[7] = sqrt(1.5 * (d0 * d0 + d4 * d4 + d8 * d8 + 2.0 * (sg[1] * sg[1] + sg[2] * sg[2] + sg[5] * sg[5])));
        f[8] = pr;
        f[9] = sym3_max_eig(sg[0], sg[1], sg[2], sg[4], sg[5], sg[8]);
        f[10] = 0.5 * E.mu * (I1 - 3.0) - E.mu * lnJ + 0.5 * E.lam * lnJ * lnJ;
        f[11] = sqrt(fmax(sym3_max_eig(Cm[0], Cm[1], Cm[2], Cm[4], Cm[5], Cm[8]), 1.0));
#pragma unroll
        for (int k = 0; k < 9; ++k) f[12 + k] = F[k];
    }
#pragma unroll
    for (int k = 0; k < kFieldCount; ++k) out[(size_t)k * A.ne + e] = f[k];
}

// ---------------------------------------------------------------------------
// ZZ recovered-gradient error estimate (ZZErrorEstimator::estimate_with_bounds,
// FEM/error_estimator.hpp:128-232) in three passes: element samples, node
// patches, element indicators.
// ---------------------------------------------------------------------------
// pass 1, colour order t: raw sample g = (F11 - 1, F22 - 1, F33 - 1), the
// reference volume |det J|/6 and the centroid; zz is kZzRows x ne SoA in
// colour order, the order of the node -> element lists (within a colour,
// element ids ascend, so the patches of neighbouring nodes read nearby
// entries of every colour)
__global__ __launch_bounds__(128) void zz_elem_kernel(ElemArgs A, double* __restrict__ zz) {
    const uint32_t t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= A.ne) return;
    Elem E;
    double F[9] = {1, 0, 0, 0, 1, 0, 0, 0, 1};
    elem_eval(A, t, E, F);
    double c[3] = {0, 0, 0};
#pragma unroll
    for (int q = 0; q < 4; ++q)
#pragma unroll
        for (int d = 0; d < 3; ++d) c[d] += 0.25 * A.X[(size_t)d * A.nn + E.node[q]];
    const double v[kZzRows] = {E.ok ? F[0] - 1.0 : 0.0, E.ok ? F[4] - 1.0 : 0.0, E.ok ? F[8] - 1.0 : 0.0,
                               E.V, c[0], c[1], c[2]};
#pragma unroll
    for (int k = 0; k < kZzRows; ++k) zz[(size_t)k * A.ne + t] = v[k];
}

// pass 2, one node per lane: volume-weighted average of the incident samples
// (:148-164) and the SPR fit [1, dx, dy, dz] over at most 32 of them with
// offsets from the node, so the constant term is the value at the node
// (:30-99); normal equations eliminated without pivoting as :64-82, and the
// patch mean where they are singular to 1e-10 of their diagonal
__global__ __launch_bounds__(128) void zz_node_kernel(uint32_t nn, uint32_t ne, const double* __restrict__ X,
                                                      const uint32_t* __restrict__ n2e_ptr,
                                                      const uint32_t* __restrict__ n2e,
                                                      const double* __restrict__ zz, double* __restrict__ raw,
                                                      double* __restrict__ rec) {
    const uint32_t v = blockIdx.x * blockDim.x + threadIdx.x;
    if (v >= nn) return;
    const uint32_t p0 = n2e_ptr[v], p1 = n2e_ptr[v + 1];
    const double x0 = X[v], x1 = X[(size_t)nn + v], x2 = X[2 * (size_t)nn + v];
    double s[3] = {0, 0, 0}, sv = 0.0, mean[3] = {0, 0, 0};
    double M[4][7];
#pragma unroll
    for (int j = 0; j < 4; ++j)
#pragma unroll
        for (int k = 0; k < 7; ++k) M[j][k] = 0.0;
    // entries in chunks of four whose loads are all issued before any is used
    // (a plain loop waits a full round trip per entry); past the end, a chunk
    // re-reads the node's last entry with weight zero
    constexpr int CH = 4;
    for (uint32_t p = p0; p < p1; p += CH) {
        uint32_t t[CH];
#pragma unroll
        for (int k = 0; k < CH; ++k) t[k] = n2e[min(p + k, p1 - 1)] >> 2;
        double z[CH][kZzRows];
#pragma unroll
        for (int k = 0; k < CH; ++k)
#pragma unroll
            for (int r = 0; r < kZzRows; ++r) z[k][r] = zz[(size_t)r * ne + t[k]];
#pragma unroll
        for (int k = 0; k < CH; ++k) {
            if (p + k >= p1) break;
            const double vol = z[k][3];
            s[0] += z[k][0] * vol;
            s[1] += z[k][1] * vol;
            s[2] += z[k][2] * vol;
            sv += vol;
            if (p + k - p0 < 32) {
                const double b[4] = {1.0, z[k][4] - x0, z[k][5] - x1, z[k][6] - x2};
#pragma unroll
                for (int j = 0; j < 4; ++j) {
#pragma unroll
                    for (int q = 0; q < 4; ++q) M[j][q] += b[j] * b[q];
#pragma unroll
                    for (int c = 0; c < 3; ++c) M[j][4 + c] += b[j] * z[k][c];
                }
            }
        }
    }
    const double m = (double)min(p1 - p0, 32u);
#pragma unroll
    for (int c = 0; c < 3; ++c) mean[c] = m > 0 ? M[0][4 + c] / m : 0.0;
    double diag[4];
#pragma unroll
    for (int j = 0; j < 4; ++j) diag[j] = M[j][j];
    bool singular = m == 0;
#pragma unroll
    for (int pv = 0; pv < 4; ++pv) {
        singular = singular || !(fabs(M[pv][pv]) > 1e-10 * diag[pv]);
        const double ip = singular ? 0.0 : 1.0 / M[pv][pv];
#pragma unroll
        for (int rr = pv + 1; rr < 4; ++rr) {
            const double f = M[rr][pv] * ip;
#pragma unroll
            for (int c = pv; c < 7; ++c) M[rr][c] -= f * M[pv][c];
        }
    }
    double sol[4][3];
#pragma unroll
    for (int pv = 3; pv >= 0; --pv)
#pragma unroll
        for (int c = 0; c < 3; ++c) {
            double rhs = M[pv][4 + c];
#pragma unroll
            for (int k = pv + 1; k < 4; ++k) rhs -= M[pv][k] * sol[k][c];
            sol[pv][c] = singular ? 0.0 : rhs / M[pv][pv];
        }
#pragma unroll
    for (int c = 0; c < 3; ++c) {
        raw[3 * (size_t)v + c] = sv > 1e-30 ? s[c] / sv : s[c];
        rec[3 * (size_t)v + c] = singular ? mean[c] : sol[0][c];
    }
}

// pass 3, element order e: indicator sqrt(sum |rec - raw|^2 / 4) cbrt(V) and
// the block partial of V sum / 4 (:198-214)
__global__ __launch_bounds__(256) void zz_err_kernel(uint32_t ne, const uint32_t* __restrict__ conn_c,
                                                     const uint32_t* __restrict__ eiperm,
                                                     const double* __restrict__ zz, const double* __restrict__ raw,
                                                     const double* __restrict__ rec, double* __restrict__ err,
                                                     double* __restrict__ partials) {
    const uint32_t e = blockIdx.x * blockDim.x + threadIdx.x;
    double part = 0.0;
    if (e < ne) {
        const uint32_t t = eiperm[e];
        double sq = 0.0;
#pragma unroll
        for (int a = 0; a < 4; ++a) {
            const uint32_t v = conn_c[(size_t)a * ne + t];
#pragma unroll
            for (int c = 0; c < 3; ++c) {
                const double d = rec[3 * (size_t)v + c] - raw[3 * (size_t)v + c];
                sq += d * d;
            }
        }
        const double vol = zz[3 * (size_t)ne + t];
        err[e] = sqrt(sq / 4.0) * cbrt(vol);
        part = vol * sq / 4.0;
    }
    const double bs = block_sum(part);
    if (threadIdx.x == 0) partials[blockIdx.x] = bs;
}

__global__ void lincomb_kernel(uint32_t n, Lincomb c, double* __restrict__ out) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    double v = c.w[0] * c.x[0][i];
    for (int q = 1; q < c.k; ++q) v += c.w[q] * c.x[q][i];
    out[i] = v;
}

inline uint32_t cdiv(uint64_t a, uint64_t b) { return (uint32_t)((a + b - 1) / b); }

// waves of the persistent kernels: few for the sweeps (only the frontier
// should spin), more for the factorization (long per-node latency chains);
// FM_SOLVE_WAVES / FM_FACTOR_WAVES override
uint32_t env_waves(const char* name, uint32_t dflt) {
    const char* e = std::getenv(name);
    const long v = e ? std::atol(e) : 0;
    return (uint32_t)(v > 0 ? v : dflt);
}
thread_local hipError_t g_launch_err = hipSuccess;  // per calling thread, like its contexts

// Launch of a persistent kernel: at most `want` one-wave blocks, clamped to
// half of what the device keeps resident at once, so every block is running
// from the start (the dependence waits rely on it).  A plain launch is used,
// not hipLaunchCooperativeKernel: cooperative launches hold a device-wide
// resource and fail when another process on the same GPU uses them.  Should
// a block nevertheless not be resident, the bounded waits end the kernel with
// FM_EDEVICE instead of hanging.
template <typename... Args>
uint32_t launch_persistent(void (*kern)(Args...), uint32_t want, hipStream_t s, Args... args) {
    const uint32_t grid = persistent_grid(reinterpret_cast<const void*>(kern), want);
    kern<<<grid, 64, 0, s>>>(args...);
    record_launch_error(hipGetLastError());
    return grid;
}

uint32_t solve_waves() {
    static const uint32_t w = env_waves("FM_SOLVE_WAVES", 2048);
    return w;
}
uint32_t factor_waves() {
    static const uint32_t w = env_waves("FM_FACTOR_WAVES", 4096);
    return w;
}

}  // namespace

// ===========================================================================
// launchers
// ===========================================================================
// ---- scratch of the allocation-free residual (fm_residual_noalloc) ------------
// Block partials of R.R and the reduce state, in one file-scope device array:
// each GPU holds its own copy for the life of the process, shared by every
// call on that GPU.  Host code uses the address hipGetSymbolAddress returns
// for the current device, never the array's name.
constexpr size_t kNoAllocStateWords = (sizeof(PcgState) + 255) / 256 * 32;
static __device__ __align__(256) double scratch[kNoAllocPartials + kNoAllocStateWords];

namespace {
constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
double* g_scratch[kMaxDevices] = {};
}  // namespace

hipError_t noalloc_scratch_init(int device) {
    if (device < 0 || device >= kMaxDevices) return hipErrorInvalidDevice;
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (g_scratch[device]) return hipSuccess;
    int cur = -1;
    hipError_t e = hipGetDevice(&cur);
    if (e != hipSuccess) return e;
    if (cur != device) return hipErrorInvalidDevice;
    void* p = nullptr;
    e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess) return e;
    g_scratch[device] = static_cast<double*>(p);
    return hipSuccess;
}

NoAllocScratch noalloc_scratch(int device) {
    NoAllocScratch s{};
    if (device < 0 || device >= kMaxDevices) return s;
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (!g_scratch[device]) return s;
    s.partials = g_scratch[device];
    s.state = reinterpret_cast<PcgState*>(g_scratch[device] + kNoAllocPartials);
    return s;
}

hipError_t take_launch_error() {
    const hipError_t e = g_launch_err;
    g_launch_err = hipSuccess;
    return e;
}
void record_launch_error(hipError_t e) {
    if (e != hipSuccess) g_launch_err = e;
}
uint32_t persistent_grid(const void* kern, uint32_t want, int block) {
    static thread_local int dev = -1, cus = 0;
    int d = 0;
    (void)hipGetDevice(&d);
    if (d != dev) {
        dev = d;
        (void)hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, d);
    }
    int occ = 0;
    (void)hipOccupancyMaxActiveBlocksPerMultiprocessor(&occ, kern, block, 0);
    const uint32_t cap = (uint32_t)std::max(1, occ * cus / 2);
    return std::max(1u, std::min(want, cap));
}

void launch_fill_u8(hipStream_t s, uint8_t* p, size_t n, uint8_t v) {
    if (n) fill_kernel<uint8_t><<<std::min<uint64_t>(cdiv(n, 256), 65535), 256, 0, s>>>(p, n, v);
}
void launch_fill_f64(hipStream_t s, double* p, size_t n, double v) {
    if (n) fill_kernel<double><<<std::min<uint64_t>(cdiv(n, 256), 65535), 256, 0, s>>>(p, n, v);
}
void launch_fill_f32(hipStream_t s, float* p, size_t n, float v) {
    if (n) fill_kernel<float><<<std::min<uint64_t>(cdiv(n, 256), 65535), 256, 0, s>>>(p, n, v);
}

void launch_expand_pattern(hipStream_t s, uint32_t nn, const uint32_t* nbr_ptr, const uint32_t* nbr,
                           const uint32_t* row_ptr, uint32_t* col_idx) {
    expand_pattern_kernel<<<cdiv(nn, 256), 256, 0, s>>>(nn, nbr_ptr, nbr, row_ptr, col_idx);
}

void launch_assemble_color(hipStream_t s, int prec, bool with_K, const ElemArgs& a, uint32_t begin, uint32_t end,
                           double* f, void* K, const uint32_t* row_ptr, const uint8_t* isdir, const double* dD,
                           double* lift) {
    if (end <= begin) return;
    const uint32_t nb = cdiv(end - begin, 64);
    if (!with_K)
        assemble_color_kernel<double, false><<<nb, 64, 0, s>>>(a, begin, end, f, nullptr, row_ptr, isdir, nullptr,
                                                                nullptr);
    else if (prec == PREC64)
        assemble_color_kernel<double, true><<<nb, 64, 0, s>>>(a, begin, end, f, (double*)K, row_ptr, isdir, dD, lift);
    else
        assemble_color_kernel<float, true><<<nb, 64, 0, s>>>(a, begin, end, f, (float*)K, row_ptr, isdir, dD, lift);
}

void launch_assemble_node(hipStream_t s, int prec, const ElemArgs& a, const uint32_t* n2e_ptr, const uint32_t* n2e,
                          double* f, void* K, const uint32_t* row_ptr, const uint32_t* col_idx, const uint8_t* isdir,
                          const double* dD, double* lift) {
    if (!a.nn) return;
    const uint32_t nb = cdiv(a.nn, 2);
    if (prec == PREC64)
        assemble_node_kernel<double><<<nb, 64, 0, s>>>(a, n2e_ptr, n2e, f, (double*)K, row_ptr, col_idx, isdir, dD, lift);
    else
        assemble_node_kernel<float><<<nb, 64, 0, s>>>(a, n2e_ptr, n2e, f, (float*)K, row_ptr, col_idx, isdir, dD, lift);
}

void launch_set_dir_diag(hipStream_t s, int prec, uint32_t n_dir, const uint32_t* dofs, const uint32_t* row_ptr,
                         const uint16_t* nlow, void* K) {
    if (!n_dir) return;
    if (prec == PREC64)
        set_dir_diag_kernel<double><<<cdiv(n_dir, 256), 256, 0, s>>>(n_dir, dofs, row_ptr, nlow, (double*)K);
    else
        set_dir_diag_kernel<float><<<cdiv(n_dir, 256), 256, 0, s>>>(n_dir, dofs, row_ptr, nlow, (float*)K);
}

void launch_eliminate(hipStream_t s, int prec, uint32_t n, const uint32_t* row_ptr, const uint32_t* col_idx,
                      const uint8_t* isdir, void* K) {
    const uint32_t nb = cdiv(n, 256 / kEliminateLanes);
    if (prec == PREC64)
        eliminate_kernel<double><<<nb, 256, 0, s>>>(n, row_ptr, col_idx, isdir, (double*)K);
    else
        eliminate_kernel<float><<<nb, 256, 0, s>>>(n, row_ptr, col_idx, isdir, (float*)K);
}

uint32_t launch_residual_finish(hipStream_t s, uint32_t n, double load, const double* f_ext, const double* f,
                                const uint8_t* isdir, double* R, double* partials) {
    const uint32_t nb = cdiv(n, 256);
    residual_finish_kernel<<<nb, 256, 0, s>>>(n, load, f_ext, f, isdir, R, partials);
    return nb;
}

uint32_t launch_dot(hipStream_t s, int prec, uint32_t n, const void* a, const void* b, double* partials) {
    const uint32_t nb = cdiv(n, 256);
    if (prec == PREC64)
        dot_kernel<double><<<nb, 256, 0, s>>>(n, (const double*)a, (const double*)b, partials);
    else
        dot_kernel<float><<<nb, 256, 0, s>>>(n, (const float*)a, (const float*)b, partials);
    return nb;
}

void launch_combine(hipStream_t s, int prec, uint32_t n, int m, const double* c, const void* const* v, void* out) {
    Combine cb{};
    for (int j = 0; j < m && j < 4; ++j) {
        cb.c[j] = c[j];
        cb.v[j] = v[j];
    }
    m = std::min(m, 4);
    if (prec == PREC64)
        combine_kernel<double><<<cdiv(n, 256), 256, 0, s>>>(n, m, cb, (double*)out);
    else
        combine_kernel<float><<<cdiv(n, 256), 256, 0, s>>>(n, m, cb, (float*)out);
}

void launch_coarse_coords(hipStream_t s, uint32_t nn, const double* X, const uint8_t* isdir, const double* center,
                          const double* scale, float4* xn) {
    coarse_coords_kernel<<<cdiv(nn, 256), 256, 0, s>>>(nn, X, isdir, center[0], center[1], center[2], scale[0],
                                                        scale[1], scale[2], xn);
}

void launch_poly_basis(hipStream_t s, int prec, uint32_t nn, const float4* xn, int nmono, int ldm, void* out,
                       uint32_t ld, int col0) {
    if (prec == PREC64)
        poly_basis_kernel<double><<<cdiv(nn, 256), 256, 0, s>>>(nn, xn, nmono, ldm, (double*)out, ld, col0);
    else
        poly_basis_kernel<float><<<cdiv(nn, 256), 256, 0, s>>>(nn, xn, nmono, ldm, (float*)out, ld, col0);
}


uint32_t coarse_moment_blocks() { return kMomentBlocks; }

template <typename T>
static void coarse_correct_t(hipStream_t s, uint32_t n, int nmono, const float4* xn, const T* r, const double* Einv,
                             double* partials, double* coef, PcgState* st) {
    const uint32_t nn = n / 3;
    switch (nmono) {
        case 4: coarse_moments_kernel<T, 4><<<kMomentBlocks, 256, 0, s>>>(nn, xn, r, partials, st); break;
        case 10: coarse_moments_kernel<T, 10><<<kMomentBlocks, 256, 0, s>>>(nn, xn, r, partials, st); break;
        case 20: coarse_moments_kernel<T, 20><<<kMomentBlocks, 256, 0, s>>>(nn, xn, r, partials, st); break;
        default: coarse_moments_kernel<T, 35><<<kMomentBlocks, 256, 0, s>>>(nn, xn, r, partials, st); break;
    }
    coarse_solve_kernel<false><<<1, 1024, 0, s>>>(kMomentBlocks, 3 * nmono, partials, Einv, coef, st);
}

template <typename T>
static void coarse_moments_t(hipStream_t s, uint32_t n, int nmono, const float4* xn, const T* r, double* partials,
                             const PcgState* st) {
    const uint32_t nn = n / 3;
    switch (nmono) {
        case 4: coarse_moments_kernel<T, 4><<<kMomentBlocks, 256, 0, s>>>(nn, xn, r, partials, st); break;
        case 10: coarse_moments_kernel<T, 10><<<kMomentBlocks, 256, 0, s>>>(nn, xn, r, partials, st); break;
        case 20: coarse_moments_kernel<T, 20><<<kMomentBlocks, 256, 0, s>>>(nn, xn, r, partials, st); break;
        default: coarse_moments_kernel<T, 35><<<kMomentBlocks, 256, 0, s>>>(nn, xn, r, partials, st); break;
    }
}

void launch_coarse_moments(hipStream_t s, int prec, uint32_t n, int nmono, const float4* xn, const void* r,
                           double* partials, const PcgState* st) {
    if (prec == PREC64)
        coarse_moments_t<double>(s, n, nmono, xn, (const double*)r, partials, st);
    else
        coarse_moments_t<float>(s, n, nmono, xn, (const float*)r, partials, st);
}

template <typename T>
static void xr_update_moments_t(hipStream_t s, uint32_t n, int nmono, const T* p, const T* Ap, T* x, T* r,
                                const float4* xn, double* partials, const PcgState* st) {
    const uint32_t nn = n / 3;
    switch (nmono) {
        case 4: pcg_xr_update_moments_kernel<T, 4><<<kMomentBlocks, 256, 0, s>>>(nn, p, Ap, x, r, xn, partials, st); break;
        case 10: pcg_xr_update_moments_kernel<T, 10><<<kMomentBlocks, 256, 0, s>>>(nn, p, Ap, x, r, xn, partials, st); break;
        case 20: pcg_xr_update_moments_kernel<T, 20><<<kMomentBlocks, 256, 0, s>>>(nn, p, Ap, x, r, xn, partials, st); break;
        default: pcg_xr_update_moments_kernel<T, 35><<<kMomentBlocks, 256, 0, s>>>(nn, p, Ap, x, r, xn, partials, st); break;
    }
}

void launch_pcg_xr_update_moments(hipStream_t s, int prec, uint32_t n, int nmono, const void* p, const void* Ap,
                                  void* x, void* r, const float4* xn, double* partials, const PcgState* st) {
    if (prec == PREC64)
        xr_update_moments_t<double>(s, n, nmono, (const double*)p, (const double*)Ap, (double*)x, (double*)r, xn,
                                    partials, st);
    else
        xr_update_moments_t<float>(s, n, nmono, (const float*)p, (const float*)Ap, (float*)x, (float*)r, xn, partials,
                                   st);
}

void launch_coarse_solve(hipStream_t s, int nmono, const double* partials, const double* Einv, double* coef,
                         PcgState* st, bool cross) {
    if (cross)
        coarse_solve_kernel<true><<<1, 1024, 0, s>>>(kMomentBlocks, 3 * nmono, partials, Einv, coef, st);
    else
        coarse_solve_kernel<false><<<1, 1024, 0, s>>>(kMomentBlocks, 3 * nmono, partials, Einv, coef, st);
}

void launch_coarse_correct(hipStream_t s, int prec, uint32_t n, int nmono, const float4* xn, const void* r,
                           const double* Einv, double* partials, double* coef, PcgState* st) {
    if (prec == PREC64)
        coarse_correct_t<double>(s, n, nmono, xn, (const double*)r, Einv, partials, coef, st);
    else
        coarse_correct_t<float>(s, n, nmono, xn, (const float*)r, Einv, partials, coef, st);
}

template <typename T>
static void p_update_coarse_t(hipStream_t s, uint32_t n, int nmono, const T* y, T* p, const float4* xn,
                              const double* coef64, int first, const PcgState* st, T* lz, uint32_t lz_cap) {
    const uint32_t nn = n / 3, nb = cdiv(nn, 256);
    const T* coef = sizeof(T) == 8 ? reinterpret_cast<const T*>(coef64)
                                   : reinterpret_cast<const T*>(coef64 + 3 * nmono);
    switch (nmono) {
        case 4: pcg_p_update_coarse_kernel<T, 4><<<nb, 256, 0, s>>>(nn, y, p, xn, coef, first, st, lz, lz_cap); break;
        case 10: pcg_p_update_coarse_kernel<T, 10><<<nb, 256, 0, s>>>(nn, y, p, xn, coef, first, st, lz, lz_cap); break;
        case 20: pcg_p_update_coarse_kernel<T, 20><<<nb, 256, 0, s>>>(nn, y, p, xn, coef, first, st, lz, lz_cap); break;
        default: pcg_p_update_coarse_kernel<T, 35><<<nb, 256, 0, s>>>(nn, y, p, xn, coef, first, st, lz, lz_cap); break;
    }
}

void launch_pcg_p_update_coarse(hipStream_t s, int prec, uint32_t n, int nmono, const void* y, void* p,
                                const float4* xn, const double* coef, bool first, const PcgState* st, void* lz,
                                uint32_t lz_cap) {
    if (prec == PREC64)
        p_update_coarse_t<double>(s, n, nmono, (const double*)y, (double*)p, xn, coef, first ? 1 : 0, st,
                                  (double*)lz, lz_cap);
    else
        p_update_coarse_t<float>(s, n, nmono, (const float*)y, (float*)p, xn, coef, first ? 1 : 0, st,
                                 (float*)lz, lz_cap);
}

// Y[:, q] = sum over j < m of S[j k + q] Z[:, j] for q < k (k <= 8; Z, Y of
// leading dimension n; S row-major m x k, FP64): one thread per row, every
// column of Z read once
template <typename T>
__global__ void basis_combine_multi_kernel(uint32_t n, int m, int k, const double* __restrict__ S,
                                           const T* __restrict__ Z, T* __restrict__ Y) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    double acc[8] = {0, 0, 0, 0, 0, 0, 0, 0};
    for (int j = 0; j < m; ++j) {
        const double z = (double)Z[(size_t)j * n + i];
#pragma unroll
        for (int q = 0; q < 8; ++q)
            if (q < k) acc[q] += S[j * k + q] * z;
    }
#pragma unroll
    for (int q = 0; q < 8; ++q)
        if (q < k) Y[(size_t)q * n + i] = (T)acc[q];
}

void launch_basis_combine_multi(hipStream_t s, int prec, uint32_t n, int m, int k, const double* S, const void* Z,
                                void* Y) {
    if (prec == PREC64)
        basis_combine_multi_kernel<double><<<cdiv(n, 256), 256, 0, s>>>(n, m, k, S, (const double*)Z, (double*)Y);
    else
        basis_combine_multi_kernel<float><<<cdiv(n, 256), 256, 0, s>>>(n, m, k, S, (const float*)Z, (float*)Y);
}

// a Gram matrix of few columns has few update threads per block (one 4x4
// tile each): more blocks keep the chip's memory pipes busy
uint32_t gram_blocks(int m, int mw) { return m * mw <= 32 * 32 ? 2048 : 512; }
uint32_t poly_gram_blocks() { return kPgBlocks; }

template <typename T>
static void poly_gram_t(hipStream_t s, const Matrix& A, const float4* xn, int nmono, const T* val, const T* S,
                        double* partials) {
    const uint32_t nn = A.n / 3;
    switch (nmono) {
        case 4:
            poly_gram_kernel<T, 4><<<kPgBlocks, kPgThreads, 0, s>>>(nn, A.row_ptr.p, A.nptr.p, A.ncol.p, val, xn, S,
                                                                     A.n, partials);
            break;
        case 10:
            poly_gram_kernel<T, 10><<<kPgBlocks, kPgThreads, 0, s>>>(nn, A.row_ptr.p, A.nptr.p, A.ncol.p, val, xn, S,
                                                                      A.n, partials);
            break;
        case 20:
            poly_gram_kernel<T, 20><<<kPgBlocks, kPgThreads, 0, s>>>(nn, A.row_ptr.p, A.nptr.p, A.ncol.p, val, xn, S,
                                                                      A.n, partials);
            break;
        default:
            poly_gram_kernel<T, 35><<<kPgBlocks, kPgThreads, 0, s>>>(nn, A.row_ptr.p, A.nptr.p, A.ncol.p, val, xn, S,
                                                                      A.n, partials);
            break;
    }
}

void launch_partial_sums(hipStream_t s, uint32_t nb, uint32_t np, const double* partials, double* out) {
    gram_reduce_kernel<<<cdiv(np, 32), 256, 0, s>>>(nb, np, partials, out);
}

void launch_poly_gram(hipStream_t s, int prec, const Matrix& A, const float4* xn, int nmono, const void* S,
                      double* partials, double* G) {
    if (prec == PREC64)
        poly_gram_t<double>(s, A, xn, nmono, A.val64.p, (const double*)S, partials);
    else
        poly_gram_t<float>(s, A, xn, nmono, A.val32.p, (const float*)S, partials);
    const uint32_t np = (uint32_t)(3 * nmono * (3 * nmono + kPgStored));
    gram_reduce_kernel<<<cdiv(np, 32), 256, 0, s>>>(kPgBlocks, np, partials, G);
}

void launch_gram(hipStream_t s, int prec, uint32_t n, int m, int mw, const void* V, const void* W, double* partials,
                 double* G) {
    const uint32_t nb = gram_blocks(m, mw);
    const size_t es = prec == PREC64 ? 8 : 4;
    const size_t lds = (size_t)(((m + 3) & ~3) + ((mw + 3) & ~3)) * kGramRows * es;
    const int nt = gram_threads(m, mw);
    if (prec == PREC64)
        gram_kernel<double><<<nb, nt, lds, s>>>(n, m, mw, (const double*)V, (const double*)W, partials);
    else
        gram_kernel<float><<<nb, nt, lds, s>>>(n, m, mw, (const float*)V, (const float*)W, partials);
    const uint32_t np = (uint32_t)(m * mw);
    gram_reduce_kernel<<<cdiv(np, 32), 256, 0, s>>>(nb, np, partials, G);
}

void launch_basis_combine(hipStream_t s, int prec, uint32_t n, int k, const int* col, const double* c, const void* V,
                          void* x) {
    if (prec == PREC64)
        basis_combine_kernel<double><<<cdiv(n, 256), 256, 0, s>>>(n, k, col, c, (const double*)V, (double*)x);
    else
        basis_combine_kernel<float><<<cdiv(n, 256), 256, 0, s>>>(n, k, col, c, (const float*)V, (float*)x);
}

void launch_normalize(hipStream_t s, int prec, uint32_t n, const double* nrm2, const void* in, void* out) {
    if (prec == PREC64)
        normalize_kernel<double><<<cdiv(n, 256), 256, 0, s>>>(n, nrm2, (const double*)in, (double*)out);
    else
        normalize_kernel<float><<<cdiv(n, 256), 256, 0, s>>>(n, nrm2, (const float*)in, (float*)out);
}

uint32_t launch_dot64(hipStream_t s, uint32_t n, const double* a, const double* b, double* partials) {
    const uint32_t nb = cdiv(n, 256);
    dot64_kernel<<<nb, 256, 0, s>>>(n, a, b, partials);
    return nb;
}

void launch_reduce(hipStream_t s, const double* partials, uint32_t np, PcgState* st, int op, int k) {
    // one block per 8192 partials (one round of loads per thread), at most kReduceBlocks
    const uint32_t nb = std::max(1u, std::min<uint32_t>(kReduceBlocks, (np + 8191) / 8192));
    reduce_kernel<<<nb, 1024, 0, s>>>(partials, np, st, op, k);
}

static int spmv_lpr(const Matrix& A) {
    static const int forced = (int)env_waves("FM_SPMV_LPR", 0);
    if (forced == 4 || forced == 8 || forced == 16 || forced == 32) return forced;
    const double avg = A.n ? (double)A.nnz / A.n : 1.0;
    if (avg <= 6) return 4;
    if (avg <= 14) return 8;
    if (avg <= 40) return 16;
    if (avg <= 64) return 16;
    return 32;
}

static bool spmv_use_stream(const Matrix& A) {
    static const int forced = (int)env_waves("FM_SPMV_STREAM", 2);  // 1 on, anything else: automatic
    if (forced == 1) return true;
    if (forced == 3) return false;
    return A.n && (double)A.nnz / A.n <= 48.0;
}

// FM_SPMV_LDS=0 keeps the direct node kernel (default: LDS-staged)
static bool spmv_use_lds() {
    static const bool on = [] {
        const char* e = std::getenv("FM_SPMV_LDS");
        return !(e && e[0] == '0' && e[1] == 0);
    }();
    return on;
}

static bool spmv_use_node(const Matrix& A) {
    static const bool csr_only = [] {
        const char* e = std::getenv("FM_SPMV");
        return e && std::string(e) == "csr";
    }();
    return A.node_spmv && !csr_only;
}

uint32_t spmv_blocks(const Matrix& A, int prec) {
    if (spmv_use_node(A))
        return cdiv(A.n / 3, 256 / kNodeLanes * (prec == PREC64 ? nodes_per_group<double>() : nodes_per_group<float>()));
    return spmv_use_stream(A) ? cdiv(A.n, kTR) : cdiv(A.n, 256 / spmv_lpr(A));
}

template <typename T>
static void spmv_t(hipStream_t s, const Matrix& A, const T* val, const T* x, T* y, bool dot, double* partials,
                   T* s0, T* s1, const PcgState* st) {
    if (spmv_use_node(A)) {
        const uint32_t nn = A.n / 3, nb = cdiv(nn, 256 / kNodeLanes * nodes_per_group<T>());
        if (spmv_use_lds()) {
            const uint32_t nval = (uint32_t)A.nnz;
            if (dot)
                spmv_node_lds_kernel<T, true><<<nb, 256, 0, s>>>(nn, A.row_ptr.p, A.nptr.p, A.ncol.p, val, nval, x, y,
                                                                 partials, s0, s1, st);
            else
                spmv_node_lds_kernel<T, false><<<nb, 256, 0, s>>>(nn, A.row_ptr.p, A.nptr.p, A.ncol.p, val, nval, x, y,
                                                                  partials, s0, s1, st);
            return;
        }
        if (dot)
            spmv_node_kernel<T, true><<<nb, 256, 0, s>>>(nn, A.row_ptr.p, A.nptr.p, A.ncol.p, val, x, y, partials, s0,
                                                         s1, st);
        else
            spmv_node_kernel<T, false><<<nb, 256, 0, s>>>(nn, A.row_ptr.p, A.nptr.p, A.ncol.p, val, x, y, partials,
                                                          s0, s1, st);
        return;
    }
    if (spmv_use_stream(A)) {
        const uint32_t nb = cdiv(A.n, kTR);
        if (dot)
            spmv_stream_kernel<T, true><<<nb, 256, 0, s>>>(A.n, A.row_ptr.p, A.col_idx.p, val, x, y, partials, s0, s1,
                                                          st);
        else
            spmv_stream_kernel<T, false><<<nb, 256, 0, s>>>(A.n, A.row_ptr.p, A.col_idx.p, val, x, y, partials, s0,
                                                           s1, st);
        return;
    }
    const int lpr = spmv_lpr(A);
    const uint32_t nb = cdiv(A.n, 256 / lpr);
#define FM_SPMV(L)                                                                                         \
    if (dot)                                                                                               \
        spmv_kernel<T, L, true><<<nb, 256, 0, s>>>(A.n, A.row_ptr.p, A.col_idx.p, val, x, y, partials, s0, \
                                                   s1, st);                                                \
    else                                                                                                   \
        spmv_kernel<T, L, false><<<nb, 256, 0, s>>>(A.n, A.row_ptr.p, A.col_idx.p, val, x, y, partials, s0, s1, st);
    switch (lpr) {
        case 4:
            FM_SPMV(4);
            break;
        case 8:
            FM_SPMV(8);
            break;
        case 16:
            FM_SPMV(16);
            break;
        default:
            FM_SPMV(32);
            break;
    }
#undef FM_SPMV
}

bool launch_spmv_multi(hipStream_t s, int prec, const Matrix& A, const void* X, void* Y, uint32_t ld, int nv) {
    if (!spmv_use_node(A) || !spmv_use_lds() || nv <= 0) return false;
    const uint32_t nn = A.n / 3;
    if (prec == PREC64) {
        const uint32_t nb = cdiv(nn, 256 / kNodeLanes * nodes_per_group<double>());
        spmv_node_multi_kernel<double><<<nb, 256, 0, s>>>(nn, A.row_ptr.p, A.nptr.p, A.ncol.p, A.val64.p,
                                                          (const double*)X, (double*)Y, ld, nv);
    } else {
        const uint32_t nb = cdiv(nn, 256 / kNodeLanes * nodes_per_group<float>());
        spmv_node_multi_kernel<float><<<nb, 256, 0, s>>>(nn, A.row_ptr.p, A.nptr.p, A.ncol.p, A.val32.p,
                                                         (const float*)X, (float*)Y, ld, nv);
    }
    return true;
}

void launch_spmv(hipStream_t s, int prec, const Matrix& A, const void* x, void* y, bool dot, double* partials,
                 void* sent0, void* sent1, const PcgState* st) {
    if (prec == PREC64)
        spmv_t<double>(s, A, A.val64.p, (const double*)x, (double*)y, dot, partials, (double*)sent0, (double*)sent1,
                       st);
    else
        spmv_t<float>(s, A, A.val32.p, (const float*)x, (float*)y, dot, partials, (float*)sent0, (float*)sent1, st);
}

template <typename T, int B>
static void factor_t(hipStream_t s, const Matrix& A, const T* val, T* lu, PcgState* st, int* err) {
    const Schedule& S = A.sched;
    fill_sentinel_kernel<T><<<std::min<uint64_t>(cdiv(A.nnz, 256), 65535), 256, 0, s>>>(lu, nullptr, A.nnz);
    const uint32_t nb = std::min(S.q_fwd, factor_waves());
    const uint32_t q = S.q_fwd;
    // FM_FACTOR_STAGED=0 runs every node through the unstaged code (both give the same bits)
    const char* ev = std::getenv("FM_FACTOR_STAGED");
    const int staged = !(ev && ev[0] == '0');
    switch (S.slots) {
        case 1:
            launch_persistent(ilu_factor_kernel<T, B, 1>, nb, s, q, (const uint32_t*)S.fwd_order.p,
                        (const uint32_t*)A.row_ptr.p, (const uint32_t*)A.col_idx.p, (const uint16_t*)S.nlow.p, val,
                        lu, &st->ticket[2], staged, err);
            break;
        case 2:
            launch_persistent(ilu_factor_kernel<T, B, 2>, nb, s, q, (const uint32_t*)S.fwd_order.p,
                        (const uint32_t*)A.row_ptr.p, (const uint32_t*)A.col_idx.p, (const uint16_t*)S.nlow.p, val,
                        lu, &st->ticket[2], staged, err);
            break;
        default:
            launch_persistent(ilu_factor_kernel<T, B, 4>, nb, s, q, (const uint32_t*)S.fwd_order.p,
                        (const uint32_t*)A.row_ptr.p, (const uint32_t*)A.col_idx.p, (const uint16_t*)S.nlow.p, val,
                        lu, &st->ticket[2], staged, err);
            break;
    }
}

void launch_ilu_factor(hipStream_t s, int prec, const Matrix& A, const void* val, void* lu, PcgState* st, int* err) {
    if (prec == PREC64) {
        if (A.sched.B == 3)
            factor_t<double, 3>(s, A, (const double*)val, (double*)lu, st, err);
        else
            factor_t<double, 1>(s, A, (const double*)val, (double*)lu, st, err);
    } else {
        if (A.sched.B == 3)
            factor_t<float, 3>(s, A, (const float*)val, (float*)lu, st, err);
        else
            factor_t<float, 1>(s, A, (const float*)val, (float*)lu, st, err);
    }
}

template <typename T, int B, int LPN>
static void fwd_t(hipStream_t s, const Matrix& A, const T* lu, const T* rhs, T* r, T* x, const T* p, const T* Ap,
                  T* z, PcgState* st, bool upd, int* err) {
    const Schedule& S = A.sched;
    const uint32_t nb = std::min(cdiv(S.q_fwd, 64 / LPN), solve_waves());
    const uint32_t q = S.q_fwd;
    if (upd)
        launch_persistent(ilu_fwd_kernel<T, B, LPN, true>, nb, s, q, (const uint32_t*)S.fwd_order.p,
                    (const uint32_t*)A.row_ptr.p, (const uint32_t*)A.col_idx.p, (const uint16_t*)S.nlow.p, lu, rhs, r,
                    x, p, Ap, z, st, err);
    else
        launch_persistent(ilu_fwd_kernel<T, B, LPN, false>, nb, s, q, (const uint32_t*)S.fwd_order.p,
                    (const uint32_t*)A.row_ptr.p, (const uint32_t*)A.col_idx.p, (const uint16_t*)S.nlow.p, lu, rhs, r,
                    x, p, Ap, z, st, err);
}

template <typename T>
static void fwd_dispatch(hipStream_t s, const Matrix& A, const T* lu, const T* rhs, T* r, T* x, const T* p,
                         const T* Ap, T* z, PcgState* st, bool upd, int* err) {
    if (A.sched.B == 3) {
        if (A.sched.lpn == 32)
            fwd_t<T, 3, 32>(s, A, lu, rhs, r, x, p, Ap, z, st, upd, err);
        else if (A.sched.lpn == 16)
            fwd_t<T, 3, 16>(s, A, lu, rhs, r, x, p, Ap, z, st, upd, err);
        else
            fwd_t<T, 3, 4>(s, A, lu, rhs, r, x, p, Ap, z, st, upd, err);
    } else {
        if (A.sched.lpn == 16)
            fwd_t<T, 1, 16>(s, A, lu, rhs, r, x, p, Ap, z, st, upd, err);
        else
            fwd_t<T, 1, 4>(s, A, lu, rhs, r, x, p, Ap, z, st, upd, err);
    }
}

void launch_ilu_fwd(hipStream_t s, int prec, const Matrix& A, const void* lu, const void* rhs, void* r, void* x,
                    const void* p, const void* Ap, void* z, PcgState* st, bool upd, int* err) {
    if (prec == PREC64)
        fwd_dispatch<double>(s, A, (const double*)lu, (const double*)rhs, (double*)r, (double*)x, (const double*)p,
                             (const double*)Ap, (double*)z, st, upd, err);
    else
        fwd_dispatch<float>(s, A, (const float*)lu, (const float*)rhs, (float*)r, (float*)x, (const float*)p,
                            (const float*)Ap, (float*)z, st, upd, err);
}

template <typename T, int B, int LPN>
static uint32_t bwd_t(hipStream_t s, const Matrix& A, const T* lu, const T* z, const T* r, T* y, double* partials,
                      PcgState* st, bool check_done, int* err) {
    const Schedule& S = A.sched;
    const uint32_t ntk = cdiv(S.q_bwd, 64 / LPN);
    const uint32_t nb = std::min(ntk, solve_waves());
    return launch_persistent(ilu_bwd_kernel<T, B, LPN>, nb, s, S.q_bwd, (const uint32_t*)S.bwd_order.p,
                       (const uint32_t*)A.row_ptr.p, (const uint32_t*)A.col_idx.p, (const uint16_t*)S.nlow.p, lu, z, r,
                       y, partials, st, check_done ? 1 : 0, err);  // one partial per wave
}

template <typename T>
static uint32_t bwd_dispatch(hipStream_t s, const Matrix& A, const T* lu, const T* z, const T* r, T* y,
                             double* partials, PcgState* st, bool check_done, int* err) {
    if (A.sched.B == 3)
        return A.sched.lpn == 32   ? bwd_t<T, 3, 32>(s, A, lu, z, r, y, partials, st, check_done, err)
               : A.sched.lpn == 16 ? bwd_t<T, 3, 16>(s, A, lu, z, r, y, partials, st, check_done, err)
                                   : bwd_t<T, 3, 4>(s, A, lu, z, r, y, partials, st, check_done, err);
    return A.sched.lpn == 16 ? bwd_t<T, 1, 16>(s, A, lu, z, r, y, partials, st, check_done, err)
                             : bwd_t<T, 1, 4>(s, A, lu, z, r, y, partials, st, check_done, err);
}

uint32_t launch_ilu_bwd(hipStream_t s, int prec, const Matrix& A, const void* lu, const void* z, const void* r,
                        void* y, double* partials, PcgState* st, bool check_done, int* err) {
    if (prec == PREC64)
        return bwd_dispatch<double>(s, A, (const double*)lu, (const double*)z, (const double*)r, (double*)y, partials,
                                    st, check_done, err);
    return bwd_dispatch<float>(s, A, (const float*)lu, (const float*)z, (const float*)r, (float*)y, partials, st,
                               check_done, err);
}

void launch_fill_sentinel(hipStream_t s, int prec, void* a, void* b, size_t n) {
    const uint32_t nb = std::min<uint64_t>(cdiv(n, 256), 65535);
    if (prec == PREC64)
        fill_sentinel_kernel<double><<<nb, 256, 0, s>>>((double*)a, (double*)b, n);
    else
        fill_sentinel_kernel<float><<<nb, 256, 0, s>>>((float*)a, (float*)b, n);
}

void launch_pcg_p_update(hipStream_t s, int prec, uint32_t n, const void* y, void* p, const PcgState* st) {
    if (prec == PREC64)
        pcg_p_update_kernel<double><<<cdiv(n, 256), 256, 0, s>>>(n, (const double*)y, (double*)p, st);
    else
        pcg_p_update_kernel<float><<<cdiv(n, 256), 256, 0, s>>>(n, (const float*)y, (float*)p, st);
}

void launch_pcg_xr_update(hipStream_t s, int prec, uint32_t n, const void* p, const void* Ap, void* x, void* r,
                          const PcgState* st) {
    if (prec == PREC64)
        pcg_xr_update_kernel<double><<<cdiv(n, 256), 256, 0, s>>>(n, (const double*)p, (const double*)Ap,
                                                                  (double*)x, (double*)r, st);
    else
        pcg_xr_update_kernel<float><<<cdiv(n, 256), 256, 0, s>>>(n, (const float*)p, (const float*)Ap, (float*)x,
                                                                 (float*)r, st);
}

uint32_t launch_pcg_init_r(hipStream_t s, int prec, uint32_t n, const void* b, const void* Ap, void* r,
                           double* partials) {
    const uint32_t nb = cdiv(n, 256);
    if (prec == PREC64)
        pcg_init_r_kernel<double><<<nb, 256, 0, s>>>(n, (const double*)b, (const double*)Ap, (double*)r, partials);
    else
        pcg_init_r_kernel<float><<<nb, 256, 0, s>>>(n, (const float*)b, (const float*)Ap, (float*)r, partials);
    return nb;
}

void launch_copy(hipStream_t s, int prec, uint32_t n, const void* src, void* dst) {
    if (prec == PREC64)
        copy_kernel<double><<<cdiv(n, 256), 256, 0, s>>>(n, (const double*)src, (double*)dst);
    else
        copy_kernel<float><<<cdiv(n, 256), 256, 0, s>>>(n, (const float*)src, (float*)dst);
}

void launch_f64_to_f32(hipStream_t s, uint32_t n, const double* src, double scale, float* dst) {
    f64_to_f32_kernel<<<cdiv(n, 256), 256, 0, s>>>(n, src, scale, dst);
}
uint32_t launch_f32_to_f64_dot(hipStream_t s, uint32_t n, const float* src, const double* r, double* dst,
                               double* partials, double* old_partials, const PcgState* st) {
    const uint32_t nb = cdiv(n, 256);
    f32_to_f64_dot_kernel<<<nb, 256, 0, s>>>(n, src, r, dst, partials, old_partials, st);
    return nb;
}
void launch_f32_to_f64(hipStream_t s, uint32_t n, const float* src, double scale, double* dst) {
    f32_to_f64_kernel<<<cdiv(n, 256), 256, 0, s>>>(n, src, scale, dst);
}
void launch_scale(hipStream_t s, uint32_t n, const double* a, double alpha, double* y) {
    scale_kernel<<<cdiv(n, 256), 256, 0, s>>>(n, a, alpha, y);
}
void launch_axpy_to(hipStream_t s, uint32_t n, const double* a, double alpha, const double* b, double* y) {
    axpy_to_kernel<<<cdiv(n, 256), 256, 0, s>>>(n, a, alpha, b, y);
}
void launch_dir_increment(hipStream_t s, uint32_t n_dir, const uint32_t* dofs, const double* vals, double scale,
                          const double* u, double* dD) {
    if (n_dir) dir_increment_kernel<<<cdiv(n_dir, 256), 256, 0, s>>>(n_dir, dofs, vals, scale, u, dD);
}
void launch_sub_masked(hipStream_t s, uint32_t n, const double* R, const double* lift, const uint8_t* isdir,
                       double* out) {
    sub_masked_kernel<<<cdiv(n, 256), 256, 0, s>>>(n, R, lift, isdir, out);
}
void launch_apply_predictor(hipStream_t s, uint32_t n, const double* du, const double* dD, const uint8_t* isdir,
                            double* u) {
    apply_predictor_kernel<<<cdiv(n, 256), 256, 0, s>>>(n, du, dD, isdir, u);
}
void launch_adjoint_elem(hipStream_t s, const ElemArgs& a, const uint32_t* eperm, const double* lam_vec,
                         double* dmu, double* dlam) {
    adjoint_elem_kernel<<<cdiv(a.ne, 128), 128, 0, s>>>(a, eperm, lam_vec, dmu, dlam);
}
uint32_t launch_energy_elem(hipStream_t s, const ElemArgs& a, double* partials) {
    const uint32_t nb = cdiv(a.ne, 256);
    energy_elem_kernel<<<nb, 256, 0, s>>>(a, partials);
    return nb;
}
void launch_rot_batch(hipStream_t s, uint32_t n, const double* F, double alpha, double* R, double* w, double* Q) {
    if (n) rot_batch_kernel<<<cdiv(n, 128), 128, 0, s>>>(n, F, alpha, R, w, Q);
}
uint32_t launch_zz(hipStream_t s, const ElemArgs& a, const uint32_t* eiperm, const uint32_t* n2e_ptr,
                   const uint32_t* n2e, double* zz, double* raw, double* rec, double* err, double* partials) {
    if (!a.ne) return 0;
    zz_elem_kernel<<<cdiv(a.ne, 128), 128, 0, s>>>(a, zz);
    zz_node_kernel<<<cdiv(a.nn, 128), 128, 0, s>>>(a.nn, a.ne, a.X, n2e_ptr, n2e, zz, raw, rec);
    const uint32_t nb = cdiv(a.ne, 256);
    zz_err_kernel<<<nb, 256, 0, s>>>(a.ne, a.conn, eiperm, zz, raw, rec, err, partials);
    return nb;
}
void launch_lincomb(hipStream_t s, uint32_t n, const Lincomb& c, double* out) {
    if (n) lincomb_kernel<<<cdiv(n, 256), 256, 0, s>>>(n, c, out);
}
void launch_element_fields(hipStream_t s, const ElemArgs& a, const uint32_t* eiperm, double* out) {
    if (a.ne) element_fields_kernel<<<cdiv(a.ne, 128), 128, 0, s>>>(a, eiperm, out);
}
void launch_elem_rot(hipStream_t s, const ElemArgs& a, const uint32_t* eiperm, double* w) {
    elem_rot_kernel<<<cdiv(a.ne, 128), 128, 0, s>>>(a, eiperm, w);
}

}  // namespace fm
