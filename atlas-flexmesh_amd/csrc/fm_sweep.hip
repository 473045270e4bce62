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
// blocks gl and gl + 4 (3x3 each); the slot of block 7 (lane 3) carries the
// node's diagonal block instead, so a node has at most 7 dependency blocks
// per direction.  The three row sums are reduced across the quad and the
// diagonal block is broadcast inside it with DPP quad permutes, which stay
// in the VALU (no LDS round trip) on the step's critical path.
//
// One wave issues about one instruction per four cycles, and a bundle's steps
// are serial, so the step's instruction count sets the sweep's pace.  The
// step therefore decides nothing per lane: the host encodes every lane's
// step as byte offsets and LDS slots (fm_plan.hpp), a dependency value is
// always (LDS slot) + (polled value), where a non-local block reads a slot
// that holds zeros and a non-external block polls past the end of the buffer
// (which loads zeros without a memory access), and the payload is addressed
// through per-bundle buffer descriptors with the step in a scalar offset.
//
// The factor values are repacked after each factorization into a step-major
// layout: the payload of one lane in one step is 18 values in 16-byte chunks
// (FP32: four and a half), stored so that each chunk index of a wave-step is
// one coalesced access of 64 lanes.
//
// The workgroup is a single wave, so no barrier is needed between steps: LDS
// instructions of one wave execute in order, and the ring accesses use
// run-time indices, which the compiler keeps in program order.  Leaving out
// the barrier also keeps the prefetched global loads in flight.
//
// The step body has no branch around a global load or store: the vector
// memory counter is in-order, and where a load or store may be skipped the
// compiler has to wait for every outstanding load, which would drain the
// prefetch each step.  Loads therefore always issue (clamped steps), lanes
// with nothing to poll or publish use an offset past the buffer (a dropped
// store), and lanes with nothing to store write a spare LDS slot.  The only
// branch left is the rare, wave-uniform re-poll of a dependency that was not
// yet published when it was polled early.
#include "fm_device.hpp"
#include "fm_internal.hpp"

#include <algorithm>
#include <cstdlib>

namespace fm {
// LLVM's raw buffer intrinsics, declared directly: the clang builtins for
// loads wider than one dword return the first dword splatted in this
// toolchain.  Outside the anonymous namespace: they have no definition.
typedef float fm_f2 __attribute__((ext_vector_type(2)));
typedef float fm_f3 __attribute__((ext_vector_type(3)));
typedef float fm_f4 __attribute__((ext_vector_type(4)));
typedef double fm_d2 __attribute__((ext_vector_type(2)));
typedef unsigned int fm_u2 __attribute__((ext_vector_type(2)));
typedef __amdgpu_buffer_rsrc_t Rsrc;
__device__ fm_f2 buf_load_f2(Rsrc, int, int, int) __asm("llvm.amdgcn.raw.ptr.buffer.load.v2f32");
__device__ fm_f3 buf_load_f3(Rsrc, int, int, int) __asm("llvm.amdgcn.raw.ptr.buffer.load.v3f32");
__device__ fm_f4 buf_load_f4(Rsrc, int, int, int) __asm("llvm.amdgcn.raw.ptr.buffer.load.v4f32");
__device__ fm_d2 buf_load_d2(Rsrc, int, int, int) __asm("llvm.amdgcn.raw.ptr.buffer.load.v2f64");
__device__ double buf_load_d1(Rsrc, int, int, int) __asm("llvm.amdgcn.raw.ptr.buffer.load.f64");
__device__ fm_u2 buf_load_u2(Rsrc, int, int, int) __asm("llvm.amdgcn.raw.ptr.buffer.load.v2i32");
__device__ uint32_t buf_load_u1(Rsrc, int, int, int) __asm("llvm.amdgcn.raw.ptr.buffer.load.i32");
__device__ void buf_store_u32(uint32_t, Rsrc, int, int, int) __asm("llvm.amdgcn.raw.ptr.buffer.store.i32");
__device__ void buf_store_u64(unsigned long long, Rsrc, int, int, int) __asm("llvm.amdgcn.raw.ptr.buffer.store.i64");

namespace {

constexpr int kLPN = 4;                    // lanes per node: one quad
constexpr int kNPW = 64 / kLPN;            // chains per wave
constexpr int kBPL = kChainMaxBlk / kLPN;  // dependency blocks per lane
constexpr int kRingStride = 4;             // values per LDS ring slot (3 used)
static_assert(kZeroSlot == (uint32_t)(kChainRing * kNPW), "the zero slot follows the ring");
static_assert(kBPL == 2, "the step record holds two dependency blocks per lane");

// payload of one lane and step: two 3x3 blocks; lane 3's second block slot
// holds the diagonal block (forward: l10, l20, l21; backward: 1/u00, u01,
// u02, 1/u11, u12, 1/u22)
template <typename T>
struct Shape {
    static constexpr int NV = kBPL * 9;
    static constexpr int DG = 9;                         // diagonal values in lane 3, from this index
    static constexpr int W = 16 / (int)sizeof(T);        // values per 16-byte chunk
    static constexpr int NC = NV / W;                    // full chunks per lane and step
    static constexpr int TAIL = NV - NC * W;             // values in the last, partial chunk
    static constexpr uint32_t STEP = 64u * NV * sizeof(T);  // bytes per wave-step
};

// DPP quad permutes: xor 1, xor 2, broadcast of quad lane k
constexpr int qp(int a, int b, int c, int d) { return a | (b << 2) | (c << 4) | (d << 6); }
constexpr int kXor1 = qp(1, 0, 3, 2);
constexpr int kXor2 = qp(2, 3, 0, 1);
constexpr int kBc3 = qp(3, 3, 3, 3);

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

// ---------------------------------------------------------------------------
// buffer accesses
// ---------------------------------------------------------------------------

// A lane whose offset is past the buffer (kOob, or kOffNone of the step
// record) loads zeros and its store is dropped, without a memory access.
// Polls and publications use the cache policy of agent-scope atomics (sc1):
// the value is read from and written to the device-coherent level, never a
// stale CU or XCD copy.
constexpr uint32_t kOob = kChainMaxBytes;
constexpr int kSc1 = 16;                // buffer instruction cache-policy bits: sc1
constexpr int kRsrcWord3 = 0x00020000;  // raw buffer, 32-bit elements, no swizzle
#ifndef FM_PK_AUX
#define FM_PK_AUX 2  // cache policy of the streamed step records and payload: nt (no reuse)
#endif

__device__ __forceinline__ Rsrc make_rsrc(const void* base, uint32_t bytes) {
    return __builtin_amdgcn_make_buffer_rsrc(const_cast<void*>(base), (short)0, (int)bytes, kRsrcWord3);
}

// step-record offsets count FP32 bytes
template <typename T>
__device__ __forceinline__ uint32_t scaled(uint32_t off) {
    return sizeof(T) == 4 ? off : off << 1;
}

// the three components of one node (byte offset `off`)
template <int AUX>
__device__ __forceinline__ void load3(Rsrc r, uint32_t off, float (&v)[3]) {
    const fm_f3 w = buf_load_f3(r, (int)off, 0, AUX);
    v[0] = w.x;
    v[1] = w.y;
    v[2] = w.z;
}
template <int AUX>
__device__ __forceinline__ void load3(Rsrc r, uint32_t off, double (&v)[3]) {
    const fm_d2 a = buf_load_d2(r, (int)off, 0, AUX);
    v[0] = a.x;
    v[1] = a.y;
    v[2] = buf_load_d1(r, (int)(off + 16), 0, AUX);
}

// ready-by-value publication (the all-ones pattern is reserved for "not ready")
__device__ __forceinline__ void publish_at(Rsrc r, uint32_t off, float v) {
    uint32_t b = __builtin_bit_cast(uint32_t, v);
    if (b == SB<float>::S) b = SB<float>::Q;
    buf_store_u32(b, r, (int)off, 0, kSc1);
}
__device__ __forceinline__ void publish_at(Rsrc r, uint32_t off, double v) {
    unsigned long long b = __builtin_bit_cast(unsigned long long, v);
    if (b == SB<double>::S) b = SB<double>::Q;
    buf_store_u64(b, r, (int)off, 0, kSc1);
}

template <typename T>
__device__ __forceinline__ bool is_sentinel(T v) {
    return __builtin_bit_cast(typename SB<T>::U, v) == SB<T>::S;
}
// high word of a value: all ones for the sentinel (and for no value a sweep
// publishes, except a negative NaN with that exact high word, which only
// costs an exact re-check)
__device__ __forceinline__ uint32_t hiword(float v) { return __builtin_bit_cast(uint32_t, v); }
__device__ __forceinline__ uint32_t hiword(double v) {
    return (uint32_t)(__builtin_bit_cast(unsigned long long, v) >> 32);
}

// ---------------------------------------------------------------------------
// kernel
// ---------------------------------------------------------------------------
template <typename T>
struct SweepArgs {
    uint32_t nb;
    const uint32_t* bstart;
    const uint32_t* bsteps;
    const uint32_t* far;  // step records (fm_plan.hpp), kRecWords words per wave-step
    const T* pk;
    const T* in0;  // forward: rhs; backward: z
    const T* in1;  // backward: r for the r.y partials (else in0)
    T* out;        // published result: z (forward) or y (backward)
    uint32_t out_bytes;  // size of in0, in1 and out (< kChainMaxBytes)
    double* partials;
    PcgState* st;
    int check_done;
    int* err;
};

// step record of a lane (2D steps ahead)
struct Far {
    uint32_t pub;        // offset of the quad's node triple
    uint32_t dep[kBPL];  // dependency blocks: external node offset or kDepLocal | ring slot
};
// an external dependency's offset (kOffNone for a local or absent block)
__device__ __forceinline__ uint32_t ext_off(uint32_t w) { return (int32_t)w >= 0 ? w : kOffNone; }
// a local dependency's ring slot (kZeroSlot for an external or absent block)
__device__ __forceinline__ uint32_t ring_slot(uint32_t w) { return (int32_t)w >= 0 ? kZeroSlot : (w & 0xFFFFu); }

// step data (D steps ahead)
template <typename T>
struct Near {
    using S = Shape<T>;
    Far f;
    T v[S::NV];  // dependency blocks (row-major 3x3 each), then two diagonal-block values
    T vec[3];    // the node's components of in0
    T vec1[3];   // backward: of in1
};

// external dependency values of a step, polled poll_ahead<T>() steps early (may
// still be the sentinel)
template <typename T>
struct Ext {
    T v[kBPL][3];
};
// The payload is prefetched D steps ahead (enough loads in flight to cover
// the memory latency), the external dependencies are polled one step ahead:
// a poll issued earlier finds the producer's value missing more often, and
// each miss costs a re-poll round trip (at the 128^3-hex box, FP32: 350/380 us
// forward/backward polling one step ahead, 416/425 us polling three ahead).
#ifndef FM_CHAIN_POLL32
#define FM_CHAIN_POLL32 1
#endif
#ifndef FM_CHAIN_POLL64
#define FM_CHAIN_POLL64 1
#endif
template <typename T>
constexpr int poll_ahead() {
    return sizeof(T) == 4 ? FM_CHAIN_POLL32 : FM_CHAIN_POLL64;
}

// descriptors of one bundle's step records and payload
struct BundleRsrc {
    Rsrc far, pk;
};

template <typename T>
__device__ __forceinline__ BundleRsrc bundle_rsrc(const SweepArgs<T>& A, uint32_t q0, uint32_t ns) {
    using S = Shape<T>;
    BundleRsrc R;
    R.far = make_rsrc(A.far + (size_t)q0 * kRecWords, ns * kRecWords * 4u);
    R.pk = make_rsrc(reinterpret_cast<const char*>(A.pk) + (size_t)q0 * S::STEP, ns * S::STEP);
    return R;
}

__device__ __forceinline__ void far_load(const BundleRsrc& R, uint32_t s, int lane, Far& F) {
    const fm_u2 w = buf_load_u2(R.far, lane * 8, (int)(s * kRecWords * 4), FM_PK_AUX);
    F.dep[0] = w.x;
    F.dep[1] = w.y;
    F.pub = buf_load_u1(R.far, 512 + (lane >> 2) * 4, (int)(s * kRecWords * 4), FM_PK_AUX);
}

// every load issues: idle quads read their zero payload, absent dependencies
// and idle quads' vectors load past the buffers (zeros, no access)
template <typename T, bool BWD>
__device__ __forceinline__ void near_load(const SweepArgs<T>& A, const BundleRsrc& R, uint32_t s, int lane,
                                          const Far& F, Near<T>& N) {
    using S = Shape<T>;
    N.f = F;
    const int soff = (int)(s * S::STEP);
#pragma unroll
    for (int c = 0; c < S::NC; ++c) {
        const int voff = lane * 16 + c * 1024;
        if constexpr (sizeof(T) == 4) {
            const fm_f4 w = buf_load_f4(R.pk, voff, soff, FM_PK_AUX);
            N.v[c * 4 + 0] = w.x;
            N.v[c * 4 + 1] = w.y;
            N.v[c * 4 + 2] = w.z;
            N.v[c * 4 + 3] = w.w;
        } else {
            const fm_d2 w = buf_load_d2(R.pk, voff, soff, FM_PK_AUX);
            N.v[c * 2 + 0] = w.x;
            N.v[c * 2 + 1] = w.y;
        }
    }
    static_assert(S::TAIL == 0 || (S::TAIL == 2 && sizeof(T) == 4), "payload tail layout");
    if constexpr (S::TAIL == 2 && sizeof(T) == 4) {
        const fm_f2 w = buf_load_f2(R.pk, lane * 8 + S::NC * 1024, soff, FM_PK_AUX);
        N.v[S::NC * 4 + 0] = w.x;
        N.v[S::NC * 4 + 1] = w.y;
    }
    load3<0>(make_rsrc(A.in0, A.out_bytes), scaled<T>(F.pub), N.vec);
    if (BWD) load3<0>(make_rsrc(A.in1, A.out_bytes), scaled<T>(F.pub), N.vec1);
}

template <typename T>
__device__ __forceinline__ void poll_ext(const SweepArgs<T>& A, const Far& F, Ext<T>& X) {
    const Rsrc ro = make_rsrc(A.out, A.out_bytes);
#pragma unroll
    for (int j = 0; j < kBPL; ++j) load3<kSc1>(ro, scaled<T>(ext_off(F.dep[j])), X.v[j]);
}

// Re-poll of the external dependency values that were still the sentinel
// when polled early.  The re-polls of both blocks issue unconditionally
// (lanes with nothing pending load past the buffer) and are waited for
// together: one round trip per attempt, where loads under a per-value branch
// would each end in a full wait.  Bounded like await_val.
template <typename T>
__device__ __forceinline__ void await_late(const SweepArgs<T>& A, const Far& f, T (&e)[kBPL][3]) {
    uint32_t spins = 0;
    for (;;) {
        __asm__ volatile("" ::: "memory");  // every attempt loads again
        T z[kBPL][3];
        bool need[kBPL][3];
#pragma unroll
        for (int j = 0; j < kBPL; ++j) {
            bool any = false;
#pragma unroll
            for (int c = 0; c < 3; ++c) {
                need[j][c] = is_sentinel(e[j][c]);  // values of absent blocks are zeros
                any |= need[j][c];
            }
            load3<kSc1>(make_rsrc(A.out, A.out_bytes), any ? scaled<T>(ext_off(f.dep[j])) : kOob, z[j]);
        }
        bool pending = false;
#pragma unroll
        for (int j = 0; j < kBPL; ++j)
#pragma unroll
            for (int c = 0; c < 3; ++c) {
                e[j][c] = need[j][c] ? z[j][c] : e[j][c];
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
// the publication of one step: lanes 0..2 of a quad store component gl of
// its node (an idle quad's offset is past the buffer)
template <typename T>
struct Pub {
    uint32_t off;
    T v;
};

template <typename T, bool BWD>
__device__ __forceinline__ Pub<T> process(const SweepArgs<T>& A, const Near<T>& N, const Ext<T>& X, uint32_t s,
                                          int lane, T* ring, double& dot) {
    using S = Shape<T>;
    const int g = lane / kLPN, gl = lane & (kLPN - 1);
    T e[kBPL][3];
    uint32_t hi = 0;
#pragma unroll
    for (int j = 0; j < kBPL; ++j)
#pragma unroll
        for (int c = 0; c < 3; ++c) {
            e[j][c] = X.v[j][c];
            hi = max(hi, hiword(e[j][c]));
        }
    // a dependency was not yet published when polled early
    if (__any(hi == 0xFFFFFFFFu)) await_late<T>(A, N.f, e);
    T y[kBPL][3];
#pragma unroll
    for (int j = 0; j < kBPL; ++j) {
        const T* src = ring + ring_slot(N.f.dep[j]) * kRingStride;
#pragma unroll
        for (int c = 0; c < 3; ++c) y[j][c] = src[c] + e[j][c];
    }
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
    // right-hand side: every lane of a quad holds its node's triple of in0
    // (loaded from the quad's node offset, or staged per quad by the loader)
    T b[3];
#pragma unroll
    for (int d = 0; d < 3; ++d) b[d] = N.vec[d];
    T v0, v1, v2;
    if (!BWD) {
        // unit lower diagonal block (l10, l20, l21 in quad lane 3)
        const T l10 = quad<kBc3>(N.v[S::DG]), l20 = quad<kBc3>(N.v[S::DG + 1]), l21 = quad<kBc3>(N.v[S::DG + 2]);
        v0 = b[0] - part[0];
        v1 = b[1] - part[1];
        v1 -= l10 * v0;
        v2 = b[2] - part[2];
        v2 -= l20 * v0;
        v2 -= l21 * v1;
    } else {
        // upper diagonal block with reciprocal pivots (quad lane 3)
        const T r00 = quad<kBc3>(N.v[S::DG]), u01 = quad<kBc3>(N.v[S::DG + 1]), u02 = quad<kBc3>(N.v[S::DG + 2]),
                r11 = quad<kBc3>(N.v[S::DG + 3]), u12 = quad<kBc3>(N.v[S::DG + 4]), r22 = quad<kBc3>(N.v[S::DG + 5]);
        v2 = (b[2] - part[2]) * r22;
        T t1 = b[1] - part[1];
        t1 -= u12 * v2;
        v1 = t1 * r11;
        T t0 = b[0] - part[0];
        t0 -= u01 * v1;
        t0 -= u02 * v2;
        v0 = t0 * r00;
    }
    // lanes 0..2 publish components 0..2 (an idle quad's offset is past the
    // buffer); lane 3 stores into its spare LDS slot
    const bool wr = gl < 3;
    const T vd = gl == 0 ? v0 : (gl == 1 ? v1 : v2);
    ring[wr ? ((s % kChainRing) * kNPW + g) * kRingStride + gl : (kZeroSlot + 1) * kRingStride + lane] = vd;
    // r.y partials (idle quads contribute zeros), with explicit fused
    // multiply-adds: the same bits in every kernel that inlines this step
    if (BWD && gl == 1)
        dot = __builtin_fma((double)N.vec1[2], (double)v2,
                            __builtin_fma((double)N.vec1[1], (double)v1, __builtin_fma((double)N.vec1[0], (double)v0, dot)));
    return Pub<T>{wr ? scaled<T>(N.f.pub) + gl * (uint32_t)sizeof(T) : kOob, vd};
}

// minimum waves per SIMD asked of the register allocator (1: no limit; the
// FP64 sweeps would spill at 3)
#ifndef FM_SWEEP_MINW
#define FM_SWEEP_MINW 1
#endif

// Persistent: block b of the grid takes bundles b, b + P, ... in processing
// order, which is a topological order of the bundle graph, so a block only
// ever waits on bundles of lower index.  The grid is one block per bundle
// whenever the device holds that many at once (every chain sweep at the box
// configurations): progress then needs only that blocks start in index order.
// (Bundles taken from an atomic ticket counter instead, which needs neither,
// cost ~100 us per sweep at 1,089 bundles: three serialised atomics per block
// on one address, and bundles placed on CUs at random.)
template <typename T, int D, bool BWD>
__global__ __launch_bounds__(64, FM_SWEEP_MINW) void chain_sweep_kernel(SweepArgs<T> A) {
    if (A.check_done && A.st->done) return;
    // the ring, the zero slot, then one spare slot per lane for lanes with nothing to store
    __shared__ T ring[(kZeroSlot + 1) * kRingStride + 64];
    const int lane = threadIdx.x;
    if (lane < kRingStride) ring[kZeroSlot * kRingStride + lane] = (T)0;
    for (uint32_t b = blockIdx.x; b < A.nb; b += gridDim.x) {
        double dot = 0.0;  // r.y of this bundle: one partial per bundle, a fixed order whichever block runs it
        // wave-uniform: the descriptors built from them must live in scalar registers
        const uint32_t q0 = __builtin_amdgcn_readfirstlane(A.bstart[b]);
        const uint32_t ns = __builtin_amdgcn_readfirstlane(A.bsteps[b]), sl = ns - 1;
        const BundleRsrc R = bundle_rsrc(A, q0, ns);
        // steps past the end re-read the last step (never processed)
        constexpr int DE = poll_ahead<T>();
        static_assert(DE <= D && (2 * D) % DE == 0, "poll distance must divide the unrolled block");
        Far F[2 * D];
        Near<T> N[D];
        Ext<T> X[DE];
#pragma unroll
        for (int u = 0; u < 2 * D; ++u) far_load(R, min((uint32_t)u, sl), lane, F[u]);
#pragma unroll
        for (int u = 0; u < DE; ++u) poll_ext(A, F[u], X[u]);
#pragma unroll
        for (int u = 0; u < D; ++u) near_load<T, BWD>(A, R, min((uint32_t)u, sl), lane, F[u], N[u]);
        // 2D steps per block, unguarded: every block issues the same memory
        // operations, so the waits the compiler derives inside the loop see
        // loads issued a full block earlier.  The first block is peeled (its
        // data was issued just before it), the partial last block follows
        // the loop with guards.  A step's value is stored after that step's
        // loads are issued: the memory counter the waits use drains in issue
        // order and counts stores too, so a store issued before a poll would
        // hold that poll's wait until the store is acknowledged.
        auto block = [&](uint32_t s0) {
#pragma unroll
            for (int u = 0; u < 2 * D; ++u) {
                const uint32_t s = s0 + u;
                const Pub<T> pb = process<T, BWD>(A, N[u % D], X[u % DE], s, lane, ring, dot);
                poll_ext(A, F[(u + DE) % (2 * D)], X[u % DE]);
                near_load<T, BWD>(A, R, min(s + D, sl), lane, F[(u + D) % (2 * D)], N[u % D]);
                far_load(R, min(s + 2 * D, sl), lane, F[u]);
                publish_at(make_rsrc(A.out, A.out_bytes), pb.off, pb.v);
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
                const Pub<T> pb = process<T, BWD>(A, N[u % D], X[u % DE], s, lane, ring, dot);
                poll_ext(A, F[(u + DE) % (2 * D)], X[u % DE]);
                near_load<T, BWD>(A, R, min(s + D, sl), lane, F[(u + D) % (2 * D)], N[u % D]);
                far_load(R, min(s + 2 * D, sl), lane, F[u]);
                publish_at(make_rsrc(A.out, A.out_bytes), pb.off, pb.v);
            }
        }
        if (BWD && A.partials) {
#pragma unroll
            for (int o = 32; o > 0; o >>= 1) dot += __shfl_xor(dot, o);
            if (lane == 0) A.partials[b] = dot;
        }
    }
}

// ---------------------------------------------------------------------------
// Two-wave variant: a loader wave streams the step records, the payload and
// the nodes' right-hand sides into an LDS ring of R steps; the solving wave
// polls, solves and publishes.  A wave's vector memory counter completes in
// issue order, so in the one-wave kernel a polled value is usable only once
// every payload load issued before the poll has landed: the step waits for
// the loaded latency of the payload stream.  Here the solving wave's counter
// holds only its polls and publications.  The waves hand over steps through
// two LDS counters (steps staged, steps consumed); LDS instructions complete
// in order per wave and a wave waits for its LDS writes before it bumps its
// counter, so a counter value read by the other wave covers the data.
// ---------------------------------------------------------------------------
template <typename T>
struct LdSlot {
    static constexpr uint32_t DEP = Shape<T>::STEP;            // payload first (global layout), then
    static constexpr uint32_t PUB = DEP + 64 * 8;               // two dependency words per lane,
    static constexpr uint32_t VEC = PUB + 64 * 4;               // the node word of each lane,
    static constexpr uint32_t VEC1 = VEC + kNPW * 4 * sizeof(T);  // in0 of each quad's node (4 values),
    static constexpr uint32_t BYTES = VEC1 + kNPW * 4 * sizeof(T);  // in1 (backward)
};

#ifndef FM_LD_SLOTS32
#define FM_LD_SLOTS32 4
#endif
#ifndef FM_LD_SLOTS64
#define FM_LD_SLOTS64 3
#endif
#ifndef FM_LD_DEPTH32
#define FM_LD_DEPTH32 3
#endif
#ifndef FM_LD_DEPTH64
#define FM_LD_DEPTH64 3
#endif
template <typename T>
constexpr int ld_slots() {
    return sizeof(T) == 4 ? FM_LD_SLOTS32 : FM_LD_SLOTS64;
}
template <typename T>
constexpr int ld_depth() {
    return sizeof(T) == 4 ? FM_LD_DEPTH32 : FM_LD_DEPTH64;
}
// minimum waves per SIMD asked of the register allocator: at most 128 VGPRs
// in FP32, 256 in FP64
#ifndef FM_LD_MINW32
#define FM_LD_MINW32 4
#endif
#ifndef FM_LD_MINW64
#define FM_LD_MINW64 2
#endif

__device__ __forceinline__ void lds_wait() { __asm__ volatile("s_waitcnt lgkmcnt(0)" ::: "memory"); }
// the hand-over counters, addressed as LDS (a volatile generic pointer would
// become a flat access, which also waits for the wave's global loads)
typedef __attribute__((address_space(3))) volatile uint32_t LdsCounter;
// spin on an LDS counter of the other wave (bounded like the global waits;
// the other wave always advances, so the bound is never reached)
__device__ __forceinline__ void lds_await(const LdsCounter* c, uint32_t at_least, int* err) {
    uint32_t spins = 0;
    while (*c < at_least) {
        if (++spins > kSpinLimit) {
            atomicOr(err, 1);
            break;
        }
        __builtin_amdgcn_s_sleep(1);
    }
    __asm__ volatile("" ::: "memory");
}

template <typename T, bool BWD>
__device__ __forceinline__ void slot_write(char* slot, int lane, const Near<T>& N) {
    using S = Shape<T>;
    using L = LdSlot<T>;
    using V4 = T __attribute__((ext_vector_type(4)));
#pragma unroll
    for (int c = 0; c < S::NC; ++c) {
        if constexpr (sizeof(T) == 4)
            *reinterpret_cast<fm_f4*>(slot + c * 1024 + lane * 16) =
                fm_f4{N.v[c * 4], N.v[c * 4 + 1], N.v[c * 4 + 2], N.v[c * 4 + 3]};
        else
            *reinterpret_cast<fm_d2*>(slot + c * 1024 + lane * 16) = fm_d2{N.v[c * 2], N.v[c * 2 + 1]};
    }
    if constexpr (S::TAIL == 2 && sizeof(T) == 4)
        *reinterpret_cast<fm_f2*>(slot + S::NC * 1024 + lane * 8) = fm_f2{N.v[S::NC * 4], N.v[S::NC * 4 + 1]};
    *reinterpret_cast<fm_u2*>(slot + L::DEP + lane * 8) = fm_u2{N.f.dep[0], N.f.dep[1]};
    *reinterpret_cast<uint32_t*>(slot + L::PUB + lane * 4) = N.f.pub;
    // the four lanes of a quad loaded the same node triple: lane 0 stores it
    if ((lane & (kLPN - 1)) == 0) {
        *reinterpret_cast<V4*>(slot + L::VEC + (lane / kLPN) * 4 * sizeof(T)) = V4{N.vec[0], N.vec[1], N.vec[2], (T)0};
        if (BWD)
            *reinterpret_cast<V4*>(slot + L::VEC1 + (lane / kLPN) * 4 * sizeof(T)) =
                V4{N.vec1[0], N.vec1[1], N.vec1[2], (T)0};
    }
}

template <typename T>
__device__ __forceinline__ void slot_read_far(const char* slot, int lane, Far& F) {
    using L = LdSlot<T>;
    const fm_u2 d = *reinterpret_cast<const fm_u2*>(slot + L::DEP + lane * 8);
    F.dep[0] = d.x;
    F.dep[1] = d.y;
    F.pub = *reinterpret_cast<const uint32_t*>(slot + L::PUB + lane * 4);
}

template <typename T, bool BWD>
__device__ __forceinline__ void slot_read(const char* slot, int lane, const Far& F, Near<T>& N) {
    using S = Shape<T>;
    using L = LdSlot<T>;
    using V4 = T __attribute__((ext_vector_type(4)));
    N.f = F;
#pragma unroll
    for (int c = 0; c < S::NC; ++c) {
        if constexpr (sizeof(T) == 4) {
            const fm_f4 w = *reinterpret_cast<const fm_f4*>(slot + c * 1024 + lane * 16);
            N.v[c * 4 + 0] = w.x;
            N.v[c * 4 + 1] = w.y;
            N.v[c * 4 + 2] = w.z;
            N.v[c * 4 + 3] = w.w;
        } else {
            const fm_d2 w = *reinterpret_cast<const fm_d2*>(slot + c * 1024 + lane * 16);
            N.v[c * 2 + 0] = w.x;
            N.v[c * 2 + 1] = w.y;
        }
    }
    if constexpr (S::TAIL == 2 && sizeof(T) == 4) {
        const fm_f2 w = *reinterpret_cast<const fm_f2*>(slot + S::NC * 1024 + lane * 8);
        N.v[S::NC * 4 + 0] = w.x;
        N.v[S::NC * 4 + 1] = w.y;
    }
    const V4 a = *reinterpret_cast<const V4*>(slot + L::VEC + (lane / kLPN) * 4 * sizeof(T));
    N.vec[0] = a.x;
    N.vec[1] = a.y;
    N.vec[2] = a.z;
    if (BWD) {
        const V4 c = *reinterpret_cast<const V4*>(slot + L::VEC1 + (lane / kLPN) * 4 * sizeof(T));
        N.vec1[0] = c.x;
        N.vec1[1] = c.y;
        N.vec1[2] = c.z;
    }
}

// loader wave: the one-wave kernel's prefetch pipeline (records 2P steps
// ahead, payload and right-hand side P steps ahead), each step then written
// to its LDS slot once the solving wave has consumed the slot's previous step
template <typename T, int P, int R, bool BWD>
__device__ __forceinline__ void sweep_loader(const SweepArgs<T>& A, uint32_t q0, uint32_t ns, int lane, char* slots,
                                             LdsCounter* ready, LdsCounter* done) {
    using L = LdSlot<T>;
    const uint32_t sl = ns - 1;
    const BundleRsrc Rs = bundle_rsrc(A, q0, ns);
    Far F[2 * P];
    Near<T> N[P];
#pragma unroll
    for (int u = 0; u < 2 * P; ++u) far_load(Rs, min((uint32_t)u, sl), lane, F[u]);
#pragma unroll
    for (int u = 0; u < P; ++u) near_load<T, BWD>(A, Rs, min((uint32_t)u, sl), lane, F[u], N[u]);
    auto step = [&](uint32_t j, int u) {
        if (j >= (uint32_t)R) lds_await(done, j - R + 1, A.err);
        slot_write<T, BWD>(slots + (j % R) * L::BYTES, lane, N[u % P]);
        lds_wait();
        if (lane == 0) *ready = j + 1;
        near_load<T, BWD>(A, Rs, min(j + P, sl), lane, F[(u + P) % (2 * P)], N[u % P]);
        far_load(Rs, min(j + 2 * P, sl), lane, F[u]);
    };
    uint32_t j0 = 0;
    if (ns >= 2 * P) {
#pragma unroll
        for (int u = 0; u < 2 * P; ++u) step(u, u);
        for (j0 = 2 * P; j0 + 2 * P <= ns; j0 += 2 * P)
#pragma unroll
            for (int u = 0; u < 2 * P; ++u) step(j0 + u, u);
    }
#pragma unroll
    for (int u = 0; u < 2 * P; ++u)
        if (j0 + u < ns) step(j0 + u, u);
}

// solving wave: the records of step s + 1 are read from LDS and polled while
// step s is solved
template <typename T, int R, bool BWD>
__device__ __forceinline__ void sweep_solver(const SweepArgs<T>& A, uint32_t ns, int lane, const char* slots,
                                             T* ring, LdsCounter* ready, LdsCounter* done, double& dot) {
    using L = LdSlot<T>;
    const Rsrc ro = make_rsrc(A.out, A.out_bytes);
    Far F[2];
    Ext<T> X[2];
    lds_await(ready, 1, A.err);
    slot_read_far<T>(slots, lane, F[0]);
    poll_ext(A, F[0], X[0]);
    auto step = [&](uint32_t s, int u) {
        const uint32_t s1 = min(s + 1, ns - 1);
        lds_await(ready, s1 + 1, A.err);
        slot_read_far<T>(slots + (s1 % R) * L::BYTES, lane, F[u ^ 1]);
        poll_ext(A, F[u ^ 1], X[u ^ 1]);
        Near<T> N;
        slot_read<T, BWD>(slots + (s % R) * L::BYTES, lane, F[u], N);
        const Pub<T> pb = process<T, BWD>(A, N, X[u], s, lane, ring, dot);
        publish_at(ro, pb.off, pb.v);
        lds_wait();
        if (lane == 0) *done = s + 1;
    };
    uint32_t s0 = 0;
    if (ns >= 2) {
        step(0, 0);
        step(1, 1);
        for (s0 = 2; s0 + 2 <= ns; s0 += 2) {
            step(s0, 0);
            step(s0 + 1, 1);
        }
    }
    if (s0 < ns) step(s0, 0);
}

template <typename T, int P, int R, bool BWD>
__global__ __launch_bounds__(128) __attribute__((amdgpu_waves_per_eu(sizeof(T) == 4 ? FM_LD_MINW32
                                                                                      : FM_LD_MINW64))) void
chain_sweep2_kernel(SweepArgs<T> A) {
    if (A.check_done && A.st->done) return;
    __shared__ T ring[(kZeroSlot + 1) * kRingStride + 64];
    __shared__ __align__(16) char slots[R * LdSlot<T>::BYTES];
    __shared__ uint32_t counters[2];  // steps staged by the loader, steps consumed by the solver
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    if (threadIdx.x < kRingStride) ring[kZeroSlot * kRingStride + threadIdx.x] = (T)0;
    for (uint32_t b = blockIdx.x; b < A.nb; b += gridDim.x) {
        if (threadIdx.x == 0) counters[0] = counters[1] = 0;
        __syncthreads();
        const uint32_t q0 = __builtin_amdgcn_readfirstlane(A.bstart[b]);
        const uint32_t ns = __builtin_amdgcn_readfirstlane(A.bsteps[b]);
        if (wave == 1) {
            sweep_loader<T, P, R, BWD>(A, q0, ns, lane, slots, (LdsCounter*)&counters[0], (LdsCounter*)&counters[1]);
        } else {
            double dot = 0.0;
            sweep_solver<T, R, BWD>(A, ns, lane, slots, ring, (LdsCounter*)&counters[0], (LdsCounter*)&counters[1],
                                    dot);
            if (BWD && A.partials) {
#pragma unroll
                for (int o = 32; o > 0; o >>= 1) dot += __shfl_xor(dot, o);
                if (lane == 0) A.partials[b] = dot;
            }
        }
        __syncthreads();
    }
}

// Repack of the CSR factor into the step-major payload of one sweep: one
// thread per (wave-step, lane).  Forward: lower blocks k (node-block position
// k of the row), unit lower diagonal block (l10, l20, l21).  Backward: upper
// blocks k (position nlow + 1 + k), upper diagonal block with reciprocal
// pivots, rounded from FP64 (1/u00, u01, u02, 1/u11, u12, 1/u22).  The
// diagonal block goes to lane 3's second block slot (block 7 never exists).
template <typename T, bool BWD>
__global__ __launch_bounds__(256) void chain_pack_kernel(uint32_t nq, const uint32_t* __restrict__ far,
                                                         const uint32_t* __restrict__ rp,
                                                         const uint16_t* __restrict__ nlow, const T* __restrict__ LU,
                                                         T* __restrict__ pk) {
    using S = Shape<T>;
    const uint64_t t = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= (uint64_t)nq * 64) return;
    const uint32_t q = (uint32_t)(t >> 6);
    const int lane = (int)(t & 63), gl = lane & (kLPN - 1);
    const uint32_t x = far[(size_t)q * kRecWords + 128 + lane / kLPN];
    T val[S::NV];
#pragma unroll
    for (int k = 0; k < S::NV; ++k) val[k] = (T)0;
    if (x != kOffNone) {
        const uint32_t v = x / 12u;
        uint32_t rs[3];
#pragma unroll
        for (int d = 0; d < 3; ++d) rs[d] = rp[3 * v + d];
        const uint32_t nblk = (rp[3 * v + 1] - rs[0]) / 3, nl = nlow[v];
        const uint32_t nup = nblk - nl - 1;
#pragma unroll
        for (int j = 0; j < kBPL; ++j) {
            const uint32_t k = gl + j * kLPN;  // block 7 never exists: at most kChainDepBlk
            if (k >= (BWD ? nup : nl)) continue;
            const uint32_t pos = BWD ? nl + 1 + k : k;
#pragma unroll
            for (int d = 0; d < 3; ++d)
#pragma unroll
                for (int c = 0; c < 3; ++c) val[j * 9 + d * 3 + c] = LU[rs[d] + 3 * pos + c];
        }
        const uint32_t o = 3 * nl;
        if (gl == 3) {
            if (!BWD) {
                val[S::DG] = LU[rs[1] + o + 0];
                val[S::DG + 1] = LU[rs[2] + o + 0];
                val[S::DG + 2] = LU[rs[2] + o + 1];
            } else {
                val[S::DG] = (T)(1.0 / (double)LU[rs[0] + o + 0]);
                val[S::DG + 1] = LU[rs[0] + o + 1];
                val[S::DG + 2] = LU[rs[0] + o + 2];
                val[S::DG + 3] = (T)(1.0 / (double)LU[rs[1] + o + 1]);
                val[S::DG + 4] = LU[rs[1] + o + 2];
                val[S::DG + 5] = (T)(1.0 / (double)LU[rs[2] + o + 2]);
            }
        }
    }
    char* step = reinterpret_cast<char*>(pk) + (size_t)q * S::STEP;
#pragma unroll
    for (int c = 0; c < S::NC; ++c)
#pragma unroll
        for (int k = 0; k < S::W; ++k)
            reinterpret_cast<T*>(step + c * 1024 + lane * 16)[k] = val[c * S::W + k];
#pragma unroll
    for (int k = 0; k < S::TAIL; ++k)
        reinterpret_cast<T*>(step + S::NC * 1024 + lane * S::TAIL * sizeof(T))[k] = val[S::NC * S::W + k];
}

inline uint32_t cdiv(uint64_t a, uint64_t b) { return (uint32_t)((a + b - 1) / b); }

// prefetch depth in steps
#ifndef FM_CHAIN_D32
#define FM_CHAIN_D32 3
#endif
#ifndef FM_CHAIN_D64
#define FM_CHAIN_D64 2
#endif
template <typename T>
constexpr int depth() {
    return sizeof(T) == 4 ? FM_CHAIN_D32 : FM_CHAIN_D64;
}

// The two-wave kernel serves the FP64 sweeps, the one-wave kernel the FP32
// sweeps (the faster of the two for each at the 128^3-hex box);
// FM_SWEEP_LOADER=0 or 1 selects one of them for both (read once per process)
template <typename T>
static bool sweep_loader_on() {
    static const int mode = [] {
        const char* e = std::getenv("FM_SWEEP_LOADER");
        if (e && (e[0] == '0' || e[0] == '1') && e[1] == 0) return e[0] - '0';
        return -1;
    }();
    return mode < 0 ? sizeof(T) == 8 : mode == 1;
}

template <typename T, bool BWD>
uint32_t launch_sweep(hipStream_t s, const Matrix& A, const ChainSweep& W, const T* pk, SweepArgs<T> a) {
    if (!a.in1) a.in1 = a.in0;  // always read (value unused)
    a.out_bytes = (uint32_t)((size_t)A.n * sizeof(T));
    a.nb = W.nb;
    a.bstart = W.bstart.p;
    a.bsteps = W.bsteps.p;
    a.far = W.far.p;
    a.pk = pk;
    if (sweep_loader_on<T>()) {
        // one block per bundle: a bundle only waits on bundles of lower
        // index, whose blocks started earlier, so progress needs no co-residency
        chain_sweep2_kernel<T, ld_depth<T>(), ld_slots<T>(), BWD><<<W.nb, 128, 0, s>>>(a);
    } else {
        auto kern = chain_sweep_kernel<T, depth<T>(), BWD>;
        const uint32_t grid = persistent_grid(reinterpret_cast<const void*>(kern), W.nb);
        kern<<<grid, 64, 0, s>>>(a);
    }
    record_launch_error(hipGetLastError());
    return W.nb;  // partials: one per bundle
}

template <typename T>
void pack_t(hipStream_t s, const Matrix& A, const T* lu, T* pf, T* pb) {
    const ChainSchedule& C = A.chain;
    const uint16_t* nlow = A.sched.nlow.p;
    chain_pack_kernel<T, false><<<cdiv((uint64_t)C.fwd.nq * 64, 256), 256, 0, s>>>(
        C.fwd.nq, C.fwd.far.p, A.row_ptr.p, nlow, lu, pf);
    chain_pack_kernel<T, true><<<cdiv((uint64_t)C.bwd.nq * 64, 256), 256, 0, s>>>(
        C.bwd.nq, C.bwd.far.p, A.row_ptr.p, nlow, lu, pb);
    record_launch_error(hipGetLastError());
}

}  // namespace

int chain_pack(hipStream_t s, int prec, Matrix& A) {
    ChainSchedule& C = A.chain;
    if (prec == PREC64) {
        const size_t nv = Shape<double>::STEP / sizeof(double);
        if (C.fwd.pk64.alloc((size_t)C.fwd.nq * nv) != hipSuccess ||
            C.bwd.pk64.alloc((size_t)C.bwd.nq * nv) != hipSuccess)
            return FM_EDEVICE;
        pack_t<double>(s, A, A.lu64.p, C.fwd.pk64.p, C.bwd.pk64.p);
    } else {
        const size_t nv = Shape<float>::STEP / sizeof(float);
        if (C.fwd.pk32.alloc((size_t)C.fwd.nq * nv) != hipSuccess ||
            C.bwd.pk32.alloc((size_t)C.bwd.nq * nv) != hipSuccess)
            return FM_EDEVICE;
        pack_t<float>(s, A, A.lu32.p, C.fwd.pk32.p, C.bwd.pk32.p);
    }
    return FM_OK;
}

void launch_chain_fwd(hipStream_t s, int prec, const Matrix& A, const void* rhs, void* z, PcgState* st,
                      bool check_done, int* err) {
    const ChainSchedule& C = A.chain;
    auto run = [&](auto zero) {
        using T = decltype(zero);
        SweepArgs<T> a{};
        a.in0 = (const T*)rhs;
        a.out = (T*)z;
        a.st = st;
        a.check_done = check_done ? 1 : 0;
        a.err = err;
        const T* pk = sizeof(T) == 8 ? (const T*)C.fwd.pk64.p : (const T*)C.fwd.pk32.p;
        launch_sweep<T, false>(s, A, C.fwd, pk, a);
    };
    if (prec == PREC64)
        run(0.0);
    else
        run(0.0f);
}

uint32_t launch_chain_bwd(hipStream_t s, int prec, const Matrix& A, const void* z, const void* r, void* y,
                          double* partials, PcgState* st, bool check_done, int* err) {
    const ChainSchedule& C = A.chain;
    auto run = [&](auto zero) -> uint32_t {
        using T = decltype(zero);
        SweepArgs<T> a{};
        a.in0 = (const T*)z;
        a.in1 = partials ? (const T*)r : nullptr;
        a.out = (T*)y;
        a.partials = partials;
        a.st = st;
        a.check_done = check_done ? 1 : 0;
        a.err = err;
        const T* pk = sizeof(T) == 8 ? (const T*)C.bwd.pk64.p : (const T*)C.bwd.pk32.p;
        return launch_sweep<T, true>(s, A, C.bwd, pk, a);
    };
    return prec == PREC64 ? run(0.0) : run(0.0f);
}

}  // namespace fm
