// fm_sweep.hip -- chain-bundle ILU(0) triangular sweeps for gfx950.
//
// The level-scheduled sweeps pay one cross-CU visibility latency per level of
// the dependence DAG (n_x + n_y + n_z levels on a box mesh in natural order).
// Here one wave advances 16 chains together, one node per chain and step, in
// a skewed wavefront: dependencies along a chain and between the chains of
// one bundle are met through an LDS ring at instruction latency; only the
// dependencies between bundles cross CUs, through the ready-by-value protocol,
// and their loads are issued D steps early, so a producer that runs ahead
// costs no wait.
//
// A node belongs to one hardware quad (4 lanes).  Lane gl holds dependency
// blocks gl and gl + 4 (3x3 each); the three row sums are reduced across the
// quad and broadcast inside it with DPP quad permutes, which stay in the
// VALU (no LDS round trip) on the step's critical path.
//
// The factor values are repacked after each factorization into a step-major
// layout: the payload of one lane in one step is NC 16-byte chunks, stored so
// that each chunk index of a wave-step is one coalesced 1 KiB access.
//
// The workgroup is a single wave, so no barrier is needed between steps: LDS
// instructions of one wave execute in order, and the ring accesses use
// run-time indices, which the compiler keeps in program order.  Leaving out
// the barrier also keeps the prefetched global loads in flight.
//
// The step body has no branch around a global load or store: the vector
// memory counter is in-order, and where a load or store may be skipped the
// compiler has to wait for every outstanding load, which would drain the
// prefetch each step.  Loads therefore always issue (clamped steps), and the
// polls and publications of the result vector use buffer instructions whose
// offset is moved past the buffer for lanes with nothing to do: such a load
// returns zeros and such a store is dropped, without a memory access.  Lanes
// with nothing to store write a spare LDS slot.  The only branch left is the
// rare, wave-uniform re-poll of a dependency that was not yet published when
// it was polled early.
#include "fm_device.hpp"
#include "fm_internal.hpp"

#include <algorithm>
#include <cstdlib>

namespace fm {
namespace {

constexpr int kLPN = 4;                    // lanes per node: one quad
constexpr int kNPW = 64 / kLPN;            // chains per wave
constexpr int kBPL = kChainMaxBlk / kLPN;  // dependency blocks per lane
constexpr int kRingStride = 4;             // values per LDS ring slot (3 used)

template <typename T>
struct Chunk;
template <>
struct Chunk<float> {
    using V = float4;
    static constexpr int W = 4;
};
template <>
struct Chunk<double> {
    using V = double2;
    static constexpr int W = 2;
};

__device__ __forceinline__ void unpack(const float4& t, float* o) {
    o[0] = t.x;
    o[1] = t.y;
    o[2] = t.z;
    o[3] = t.w;
}
__device__ __forceinline__ void unpack(const double2& t, double* o) {
    o[0] = t.x;
    o[1] = t.y;
}

template <typename T>
struct Shape {
    static constexpr int W = Chunk<T>::W;
    static constexpr int DG = kBPL * 9;                // two diagonal-block slots follow the blocks
    static constexpr int NV = (DG + 2 + W - 1) / W * W;
    static constexpr int NC = NV / W;                  // 16-byte chunks per lane and step
};

// DPP quad permutes: xor 1, xor 2, broadcast of quad lane k
constexpr int qp(int a, int b, int c, int d) { return a | (b << 2) | (c << 4) | (d << 6); }
constexpr int kXor1 = qp(1, 0, 3, 2);
constexpr int kXor2 = qp(2, 3, 0, 1);
constexpr int kBc0 = qp(0, 0, 0, 0), kBc1 = qp(1, 1, 1, 1), kBc2 = qp(2, 2, 2, 2);

template <int CTRL>
__device__ __forceinline__ float quad(float v) {
    return __builtin_bit_cast(float, __builtin_amdgcn_update_dpp(0, __builtin_bit_cast(int, v), CTRL, 0xF, 0xF, false));
}
template <int CTRL>
__device__ __forceinline__ double quad(double v) {
    const unsigned long long u = __builtin_bit_cast(unsigned long long, v);
    const unsigned lo = (unsigned)__builtin_amdgcn_update_dpp(0, (int)(unsigned)u, CTRL, 0xF, 0xF, false);
    const unsigned hi = (unsigned)__builtin_amdgcn_update_dpp(0, (int)(unsigned)(u >> 32), CTRL, 0xF, 0xF, false);
    return __builtin_bit_cast(double, (unsigned long long)lo | ((unsigned long long)hi << 32));
}

// step data loaded 2D steps ahead: node of the lane's quad, dependency words
struct Far {
    uint32_t nd;
    uint32_t dw[kBPL];
};

// step data loaded D steps ahead
template <typename T>
struct Near {
    using S = Shape<T>;
    uint32_t nd;
    uint32_t dw[kBPL];
    T v[S::NV];      // dependency blocks (row-major 3x3 each), then two diagonal-block values
    T ext[kBPL][3];  // external dependency values polled early (may still be the sentinel)
    T vec[3];        // the node's three components of input vector gl (see SweepArgs)
};

template <typename T>
struct SweepArgs {
    uint32_t nb;
    const uint32_t* bstart;
    const uint32_t* bsteps;
    const uint4* far;  // per (wave-step, lane): node of the quad, dependency words 0 and 1
    const T* pk;
    // input vector read by quad lane gl: MODE 0 (forward, rhs): in[0] = rhs;
    // MODE 1 (forward with the fused PCG update): in = r, Ap, p, x;
    // MODE 2 (backward): in[0] = z, in[1] = r (for the r.y partials)
    const T* in[4];
    T* out;   // published result: z (forward) or y (backward)
    uint32_t out_bytes;  // its size (< kChainMaxBytes)
    T* r;     // MODE 1: r <- r - alpha Ap (written by quad lane 0)
    T* x;     // MODE 1: x <- x + alpha p (written by quad lane 3)
    T* sink;  // 4 values per lane and block: MODE 1 stores of lanes with nothing to do
    double* partials;
    PcgState* st;
    int check_done;
    int* err;
};

// Polls and publications go through a raw buffer resource over `out`.  A
// lane whose offset is kOob (past the buffer) loads zeros and its store is
// dropped, without a memory access: every lane issues the same instructions
// (no branch, so no waits at joins) and idle lanes add no traffic.  The
// cache policy is that of agent-scope atomics (sc1): the value is read from
// and written to the device-coherent level, never a stale CU or XCD copy.
constexpr uint32_t kOob = kChainMaxBytes;
constexpr int kSc1 = 16;               // buffer instruction cache-policy bits: sc1
constexpr int kRsrcWord3 = 0x00020000;  // raw buffer, 32-bit elements, no swizzle

template <typename T>
__device__ __forceinline__ __amdgpu_buffer_rsrc_t out_rsrc(const SweepArgs<T>& A) {
    return __builtin_amdgcn_make_buffer_rsrc(A.out, (short)0, (int)A.out_bytes, kRsrcWord3);
}

// LLVM's raw buffer intrinsics, declared directly: the clang builtins for
// loads wider than one dword return the first dword splatted in this
// toolchain
typedef float fm_f3 __attribute__((ext_vector_type(3)));
typedef double fm_d2 __attribute__((ext_vector_type(2)));
__device__ fm_f3 buf_load_f3(__amdgpu_buffer_rsrc_t, int, int, int) __asm("llvm.amdgcn.raw.ptr.buffer.load.v3f32");
__device__ fm_d2 buf_load_d2(__amdgpu_buffer_rsrc_t, int, int, int) __asm("llvm.amdgcn.raw.ptr.buffer.load.v2f64");
__device__ double buf_load_d1(__amdgpu_buffer_rsrc_t, int, int, int) __asm("llvm.amdgcn.raw.ptr.buffer.load.f64");
__device__ void buf_store_u32(uint32_t, __amdgpu_buffer_rsrc_t, int, int, int) __asm("llvm.amdgcn.raw.ptr.buffer.store.i32");
__device__ void buf_store_u64(unsigned long long, __amdgpu_buffer_rsrc_t, int, int,
                              int) __asm("llvm.amdgcn.raw.ptr.buffer.store.i64");

// the three components published for one node (byte offset `off`)
__device__ __forceinline__ void poll3(__amdgpu_buffer_rsrc_t r, uint32_t off, float (&v)[3]) {
    const fm_f3 w = buf_load_f3(r, (int)off, 0, kSc1);
    v[0] = w.x;
    v[1] = w.y;
    v[2] = w.z;
}
__device__ __forceinline__ void poll3(__amdgpu_buffer_rsrc_t r, uint32_t off, double (&v)[3]) {
    const fm_d2 a = buf_load_d2(r, (int)off, 0, kSc1);
    v[0] = a.x;
    v[1] = a.y;
    v[2] = buf_load_d1(r, (int)(off + 16), 0, kSc1);
}

// ready-by-value publication (the all-ones pattern is reserved for "not ready")
__device__ __forceinline__ void publish_at(__amdgpu_buffer_rsrc_t r, uint32_t off, float v) {
    uint32_t b = __builtin_bit_cast(uint32_t, v);
    if (b == SB<float>::S) b = SB<float>::Q;
    buf_store_u32(b, r, (int)off, 0, kSc1);
}
__device__ __forceinline__ void publish_at(__amdgpu_buffer_rsrc_t r, uint32_t off, double v) {
    unsigned long long b = __builtin_bit_cast(unsigned long long, v);
    if (b == SB<double>::S) b = SB<double>::Q;
    buf_store_u64(b, r, (int)off, 0, kSc1);
}

template <typename T>
__device__ __forceinline__ bool is_sentinel(T v) {
    return __builtin_bit_cast(typename SB<T>::U, v) == SB<T>::S;
}

template <typename T>
__device__ __forceinline__ void far_load(const SweepArgs<T>& A, uint32_t q, int lane, Far& F) {
    const uint4 w = A.far[(size_t)q * 64 + lane];
    F.nd = w.x;
    F.dw[0] = w.y;
    F.dw[1] = w.z;
}

template <typename T>
__device__ __forceinline__ T* sink_of(const SweepArgs<T>& A, int lane) {
    return A.sink + ((size_t)blockIdx.x * 64 + lane) * 4;
}

// every load issues: absent dependencies poll past the buffer (no access),
// idle quads read their zero payload and the first elements of the input
// vectors
template <typename T, int MODE>
__device__ __forceinline__ void near_load(const SweepArgs<T>& A, uint32_t q, int lane, const Far& F, Near<T>& N) {
    using S = Shape<T>;
    using V = typename Chunk<T>::V;
    N.nd = F.nd;
#pragma unroll
    for (int j = 0; j < kBPL; ++j) N.dw[j] = F.dw[j];
    const V* src = reinterpret_cast<const V*>(A.pk) + (size_t)q * S::NC * 64 + lane;
#pragma unroll
    for (int c = 0; c < S::NC; ++c) unpack(src[c * 64], &N.v[c * S::W]);
    const __amdgpu_buffer_rsrc_t ro = out_rsrc(A);
#pragma unroll
    for (int j = 0; j < kBPL; ++j) {
        const uint32_t dw = F.dw[j];  // kDepNone has the local bit set
        poll3(ro, (dw & kDepLocal) ? kOob : dw * (3 * (uint32_t)sizeof(T)), N.ext[j]);
    }
    const int gl = lane & (kLPN - 1);
    const int which = MODE == 1 ? gl : (MODE == 2 && gl == 1 ? 1 : 0);
    const T* v = A.in[which] + (F.nd == kDepNone ? 0 : 3 * (size_t)F.nd);
#pragma unroll
    for (int c = 0; c < 3; ++c) N.vec[c] = v[c];
}

// Re-poll of the external dependency values that were still the sentinel
// when polled early: all of them in flight together, so a stall costs one
// round trip rather than one per value.  Bounded like await_val.
template <typename T>
__device__ __forceinline__ void await_late(const SweepArgs<T>& A, const Near<T>& N, T (&y)[kBPL][3]) {
    const bool act = N.nd != kDepNone;
    uint32_t spins = 0;
    for (;;) {
        // the re-polls of both dependencies issue unconditionally (lanes
        // with nothing pending load past the buffer) and are waited for
        // together: one round trip per attempt, where loads under a
        // per-value branch would each end in a full wait
        __asm__ volatile("" ::: "memory");  // every attempt loads again
        T z[kBPL][3];
        bool need[kBPL][3];
#pragma unroll
        for (int j = 0; j < kBPL; ++j) {
            const uint32_t dw = N.dw[j];
            const bool ext = act && dw != kDepNone && !(dw & kDepLocal);
            bool any = false;
#pragma unroll
            for (int c = 0; c < 3; ++c) {
                need[j][c] = ext && is_sentinel(y[j][c]);
                any |= need[j][c];
            }
            poll3(out_rsrc(A), any ? dw * (3 * (uint32_t)sizeof(T)) : kOob, z[j]);
        }
        bool pending = false;
#pragma unroll
        for (int j = 0; j < kBPL; ++j)
#pragma unroll
            for (int c = 0; c < 3; ++c) {
                y[j][c] = need[j][c] ? z[j][c] : y[j][c];
                pending |= need[j][c] && is_sentinel(z[j][c]);
            }
        if (!__any(pending)) return;
        if (++spins > kSpinLimit) {
            if (pending) atomicOr(A.err, 1);
            return;
        }
        if ((spins & 255u) == 0 && __hip_atomic_load(A.err, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT)) return;
        __builtin_amdgcn_s_sleep(FM_SPIN_SLEEP);
    }
}

// One step of one quad: its node's row block against the dependency values,
// the 3x3 diagonal-block solve, publication of the three results.
template <typename T, int MODE>
__device__ __forceinline__ void process(const SweepArgs<T>& A, const Near<T>& N, uint32_t s, int lane, T* ring,
                                        T alpha, double& dot) {
    using S = Shape<T>;
    const int g = lane / kLPN, gl = lane & (kLPN - 1);
    const bool act = N.nd != kDepNone;
    T y[kBPL][3];
    bool late = false;
#pragma unroll
    for (int j = 0; j < kBPL; ++j) {
        const uint32_t dw = N.dw[j];
        const bool none = !act || dw == kDepNone, local = !none && (dw & kDepLocal);
        const T* src = ring + (size_t)(local ? (dw & ~kDepLocal) : 0) * kRingStride;
#pragma unroll
        for (int c = 0; c < 3; ++c) {
            const T l = src[c], e = N.ext[j][c];
            late |= !none && !local && is_sentinel(e);
            y[j][c] = none ? (T)0 : (local ? l : e);
        }
    }
    if (__any(late)) await_late<T>(A, N, y);  // a dependency was not yet published when polled early
    T part[3] = {(T)0, (T)0, (T)0};
#pragma unroll
    for (int j = 0; j < kBPL; ++j)
#pragma unroll
        for (int d = 0; d < 3; ++d)
#pragma unroll
            for (int c = 0; c < 3; ++c) part[d] += N.v[j * 9 + d * 3 + c] * y[j][c];
#pragma unroll
    for (int d = 0; d < 3; ++d) {
        part[d] += quad<kXor1>(part[d]);
        part[d] += quad<kXor2>(part[d]);
    }
    // right-hand side: quad lane 0 holds rhs / r / z, lane 1 Ap (MODE 1)
    T b[3];
#pragma unroll
    for (int d = 0; d < 3; ++d) {
        b[d] = quad<kBc0>(N.vec[d]);
        if (MODE == 1) b[d] -= alpha * quad<kBc1>(N.vec[d]);
    }
    const T d0 = N.v[S::DG], d1 = N.v[S::DG + 1];
    T v0, v1, v2;
    if (MODE != 2) {
        // unit lower diagonal block: l10, l20 in quad lane 0, l21 in lane 1
        const T l10 = quad<kBc0>(d0), l20 = quad<kBc0>(d1), l21 = quad<kBc1>(d0);
        v0 = b[0] - part[0];
        v1 = b[1] - part[1];
        v1 -= l10 * v0;
        v2 = b[2] - part[2];
        v2 -= l20 * v0;
        v2 -= l21 * v1;
    } else {
        // upper diagonal block with reciprocal pivots: (1/u00, u01) lane 0,
        // (u02, 1/u11) lane 1, (u12, 1/u22) lane 2
        const T r00 = quad<kBc0>(d0), u01 = quad<kBc0>(d1), u02 = quad<kBc1>(d0), r11 = quad<kBc1>(d1),
                u12 = quad<kBc2>(d0), r22 = quad<kBc2>(d1);
        v2 = (b[2] - part[2]) * r22;
        T t1 = b[1] - part[1];
        t1 -= u12 * v2;
        v1 = t1 * r11;
        T t0 = b[0] - part[0];
        t0 -= u01 * v1;
        t0 -= u02 * v2;
        v0 = t0 * r00;
    }
    const bool wr = act && gl < 3;
    const T vd = gl == 0 ? v0 : (gl == 1 ? v1 : v2);
    publish_at(out_rsrc(A), wr ? (3 * N.nd + gl) * (uint32_t)sizeof(T) : kOob, vd);
    ring[wr ? ((s % kChainRing) * kNPW + g) * kRingStride + gl : kChainRing * kNPW * kRingStride + lane] = vd;
    if (MODE == 1) {
        // lane 0 writes r - alpha Ap, lane 3 writes x + alpha p (p from lane 2)
        T pv[3];
#pragma unroll
        for (int d = 0; d < 3; ++d) pv[d] = quad<kBc2>(N.vec[d]);
        T* const sk = sink_of(A, lane);
        T* dst = !act ? sk : (gl == 0 ? A.r + 3 * (size_t)N.nd : (gl == 3 ? A.x + 3 * (size_t)N.nd : sk));
#pragma unroll
        for (int d = 0; d < 3; ++d) dst[d] = gl == 0 ? b[d] : N.vec[d] + alpha * pv[d];
    }
    if (MODE == 2 && act && gl == 1)
        dot += (double)N.vec[0] * (double)v0 + (double)N.vec[1] * (double)v1 + (double)N.vec[2] * (double)v2;
}

// Persistent: block b of the grid takes bundles b, b + P, ... in processing
// order, which is a topological order of the bundle graph, so the earliest
// unfinished bundle only ever waits on finished ones.
template <typename T, int D, int MODE>
__global__ __launch_bounds__(64) void chain_sweep_kernel(SweepArgs<T> A) {
    if (A.check_done && A.st->done) return;
    const T alpha = MODE == 1 ? (T)A.st->alpha : (T)0;
    // the ring, then one spare slot per lane for lanes with nothing to store
    __shared__ T ring[kChainRing * kNPW * kRingStride + 64];
    const int lane = threadIdx.x;
    double dot = 0.0;
    for (uint32_t b = blockIdx.x; b < A.nb; b += gridDim.x) {
        const uint32_t q0 = A.bstart[b], ns = A.bsteps[b], ql = q0 + ns - 1;
        // steps past the end re-read the last step (never processed)
        Far F[2 * D];
        Near<T> N[D];
#pragma unroll
        for (int u = 0; u < 2 * D; ++u) far_load(A, min(q0 + u, ql), lane, F[u]);
#pragma unroll
        for (int u = 0; u < D; ++u) near_load<T, MODE>(A, min(q0 + u, ql), lane, F[u], N[u]);
        // 2D steps per block, unguarded: every block issues the same memory
        // operations, so the waits the compiler derives inside the loop see
        // loads issued a full block earlier.  The first block is peeled (its
        // data was issued just before it), the partial last block follows
        // the loop with guards.
        auto block = [&](uint32_t s0) {
#pragma unroll
            for (int u = 0; u < 2 * D; ++u) {
                const uint32_t s = s0 + u;
                process<T, MODE>(A, N[u % D], s, lane, ring, alpha, dot);
                near_load<T, MODE>(A, min(q0 + s + D, ql), lane, F[(u + D) % (2 * D)], N[u % D]);
                far_load(A, min(q0 + s + 2 * D, ql), lane, F[u]);
            }
        };
        uint32_t s0 = 0;
        if (ns >= 2 * D) {
            block(0);
            for (s0 = 2 * D; s0 + 2 * D <= ns; s0 += 2 * D) block(s0);
        }
#pragma unroll
        for (int u = 0; u < 2 * D; ++u) {
            const uint32_t s = s0 + u;
            if (s < ns) {
                process<T, MODE>(A, N[u % D], s, lane, ring, alpha, dot);
                near_load<T, MODE>(A, min(q0 + s + D, ql), lane, F[(u + D) % (2 * D)], N[u % D]);
                far_load(A, min(q0 + s + 2 * D, ql), lane, F[u]);
            }
        }
    }
    if (MODE == 2 && A.partials) {
#pragma unroll
        for (int o = 32; o > 0; o >>= 1) dot += __shfl_xor(dot, o);
        if (lane == 0) A.partials[blockIdx.x] = dot;
    }
}

// Repack of the CSR factor into the step-major payload of one sweep: one
// thread per (wave-step, lane).  Forward: lower blocks k (node-block position
// k of the row), unit lower diagonal block (l10, l20 | l21).  Backward: upper
// blocks k (position nlow + 1 + k), upper diagonal block with reciprocal
// pivots, rounded from FP64 (1/u00, u01 | u02, 1/u11 | u12, 1/u22).
template <typename T, bool BWD>
__global__ __launch_bounds__(256) void chain_pack_kernel(uint32_t nq, const uint4* __restrict__ far,
                                                         const uint32_t* __restrict__ rp,
                                                         const uint16_t* __restrict__ nlow, const T* __restrict__ LU,
                                                         T* __restrict__ pk) {
    using S = Shape<T>;
    const uint64_t t = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= (uint64_t)nq * 64) return;
    const uint32_t q = (uint32_t)(t >> 6);
    const int lane = (int)(t & 63), gl = lane & (kLPN - 1);
    const uint32_t v = far[t].x;
    T val[S::NV];
#pragma unroll
    for (int k = 0; k < S::NV; ++k) val[k] = (T)0;
    if (v != kDepNone) {
        uint32_t rs[3];
#pragma unroll
        for (int d = 0; d < 3; ++d) rs[d] = rp[3 * v + d];
        const uint32_t nblk = (rp[3 * v + 1] - rs[0]) / 3, nl = nlow[v];
        const uint32_t nup = nblk - nl - 1;
#pragma unroll
        for (int j = 0; j < kBPL; ++j) {
            const uint32_t k = gl + j * kLPN;
            if (k >= (BWD ? nup : nl)) continue;
            const uint32_t pos = BWD ? nl + 1 + k : k;
#pragma unroll
            for (int d = 0; d < 3; ++d)
#pragma unroll
                for (int c = 0; c < 3; ++c) val[j * 9 + d * 3 + c] = LU[rs[d] + 3 * pos + c];
        }
        const uint32_t o = 3 * nl;
        if (!BWD) {
            if (gl == 0) {
                val[S::DG] = LU[rs[1] + o + 0];
                val[S::DG + 1] = LU[rs[2] + o + 0];
            } else if (gl == 1) {
                val[S::DG] = LU[rs[2] + o + 1];
            }
        } else {
            if (gl == 0) {
                val[S::DG] = (T)(1.0 / (double)LU[rs[0] + o + 0]);
                val[S::DG + 1] = LU[rs[0] + o + 1];
            } else if (gl == 1) {
                val[S::DG] = LU[rs[0] + o + 2];
                val[S::DG + 1] = (T)(1.0 / (double)LU[rs[1] + o + 1]);
            } else if (gl == 2) {
                val[S::DG] = LU[rs[1] + o + 2];
                val[S::DG + 1] = (T)(1.0 / (double)LU[rs[2] + o + 2]);
            }
        }
    }
    using V = typename Chunk<T>::V;
    V* dst = reinterpret_cast<V*>(pk) + (size_t)q * S::NC * 64 + lane;
#pragma unroll
    for (int c = 0; c < S::NC; ++c) {
        V w;
        T* wp = reinterpret_cast<T*>(&w);
#pragma unroll
        for (int k = 0; k < S::W; ++k) wp[k] = val[c * S::W + k];
        dst[c * 64] = w;
    }
}

inline uint32_t cdiv(uint64_t a, uint64_t b) { return (uint32_t)((a + b - 1) / b); }

// prefetch depth in steps
#ifndef FM_CHAIN_D32
#define FM_CHAIN_D32 2
#endif
#ifndef FM_CHAIN_D64
#define FM_CHAIN_D64 2
#endif
template <typename T>
constexpr int depth() {
    return sizeof(T) == 4 ? FM_CHAIN_D32 : FM_CHAIN_D64;
}

template <typename T, int MODE>
uint32_t launch_sweep(hipStream_t s, const ChainSchedule& C, const ChainSweep& W, const T* pk, SweepArgs<T> a) {
    a.sink = reinterpret_cast<T*>(C.sink.p);
    for (int k = 1; k < 4; ++k)
        if (!a.in[k]) a.in[k] = a.in[0];  // always read (value unused)
    if (!a.r) a.r = a.sink;
    if (!a.x) a.x = a.sink;
    a.nb = W.nb;
    a.bstart = W.bstart.p;
    a.bsteps = W.bsteps.p;
    a.far = reinterpret_cast<const uint4*>(W.far.p);
    a.pk = pk;
    auto kern = chain_sweep_kernel<T, depth<T>(), MODE>;
    const uint32_t grid = persistent_grid(reinterpret_cast<const void*>(kern), W.nb);
    kern<<<grid, 64, 0, s>>>(a);
    record_launch_error(hipGetLastError());
    return grid;
}

template <typename T>
void pack_t(hipStream_t s, const Matrix& A, const T* lu, T* pf, T* pb) {
    const ChainSchedule& C = A.chain;
    const uint16_t* nlow = A.sched.nlow.p;
    chain_pack_kernel<T, false><<<cdiv((uint64_t)C.fwd.nq * 64, 256), 256, 0, s>>>(
        C.fwd.nq, reinterpret_cast<const uint4*>(C.fwd.far.p), A.row_ptr.p, nlow, lu, pf);
    chain_pack_kernel<T, true><<<cdiv((uint64_t)C.bwd.nq * 64, 256), 256, 0, s>>>(
        C.bwd.nq, reinterpret_cast<const uint4*>(C.bwd.far.p), A.row_ptr.p, nlow, lu, pb);
    record_launch_error(hipGetLastError());
}

}  // namespace

int chain_pack(hipStream_t s, int prec, Matrix& A) {
    ChainSchedule& C = A.chain;
    if (prec == PREC64) {
        const size_t nv = (size_t)Shape<double>::NV * 64;
        if (C.fwd.pk64.alloc((size_t)C.fwd.nq * nv) != hipSuccess ||
            C.bwd.pk64.alloc((size_t)C.bwd.nq * nv) != hipSuccess)
            return FM_EDEVICE;
        pack_t<double>(s, A, A.lu64.p, C.fwd.pk64.p, C.bwd.pk64.p);
    } else {
        const size_t nv = (size_t)Shape<float>::NV * 64;
        if (C.fwd.pk32.alloc((size_t)C.fwd.nq * nv) != hipSuccess ||
            C.bwd.pk32.alloc((size_t)C.bwd.nq * nv) != hipSuccess)
            return FM_EDEVICE;
        pack_t<float>(s, A, A.lu32.p, C.fwd.pk32.p, C.bwd.pk32.p);
    }
    return FM_OK;
}

template <typename T>
static void chain_fwd_t(hipStream_t s, const Matrix& A, const void* rhs, void* r, void* x, const void* p,
                        const void* Ap, void* z, PcgState* st, bool upd, bool check_done, int* err) {
    const ChainSchedule& C = A.chain;
    SweepArgs<T> a{};
    if (upd) {
        a.in[0] = (const T*)r;
        a.in[1] = (const T*)Ap;
        a.in[2] = (const T*)p;
        a.in[3] = (const T*)x;
    } else {
        a.in[0] = (const T*)rhs;
    }
    a.out = (T*)z;
    a.out_bytes = (uint32_t)((size_t)A.n * sizeof(T));
    a.r = (T*)r;
    a.x = (T*)x;
    a.st = st;
    a.check_done = (upd || check_done) ? 1 : 0;
    a.err = err;
    const T* pk = sizeof(T) == 8 ? (const T*)C.fwd.pk64.p : (const T*)C.fwd.pk32.p;
    if (upd)
        launch_sweep<T, 1>(s, C, C.fwd, pk, a);
    else
        launch_sweep<T, 0>(s, C, C.fwd, pk, a);
}

void launch_chain_fwd(hipStream_t s, int prec, const Matrix& A, const void* rhs, void* r, void* x, const void* p,
                      const void* Ap, void* z, PcgState* st, bool upd, bool check_done, int* err) {
    if (prec == PREC64)
        chain_fwd_t<double>(s, A, rhs, r, x, p, Ap, z, st, upd, check_done, err);
    else
        chain_fwd_t<float>(s, A, rhs, r, x, p, Ap, z, st, upd, check_done, err);
}

template <typename T>
static uint32_t chain_bwd_t(hipStream_t s, const Matrix& A, const void* z, const void* r, void* y, double* partials,
                            PcgState* st, bool check_done, int* err) {
    const ChainSchedule& C = A.chain;
    SweepArgs<T> a{};
    a.in[0] = (const T*)z;
    a.in[1] = partials ? (const T*)r : nullptr;
    a.out = (T*)y;
    a.out_bytes = (uint32_t)((size_t)A.n * sizeof(T));
    a.partials = partials;
    a.st = st;
    a.check_done = check_done ? 1 : 0;
    a.err = err;
    const T* pk = sizeof(T) == 8 ? (const T*)C.bwd.pk64.p : (const T*)C.bwd.pk32.p;
    return launch_sweep<T, 2>(s, C, C.bwd, pk, a);
}

uint32_t launch_chain_bwd(hipStream_t s, int prec, const Matrix& A, const void* z, const void* r, void* y,
                          double* partials, PcgState* st, bool check_done, int* err) {
    if (prec == PREC64) return chain_bwd_t<double>(s, A, z, r, y, partials, st, check_done, err);
    return chain_bwd_t<float>(s, A, z, r, y, partials, st, check_done, err);
}

}  // namespace fm
