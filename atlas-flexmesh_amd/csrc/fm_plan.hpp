// fm_plan.hpp -- host-side plan of the chain-bundle triangular sweeps (no HIP
// types, so host-only programs can build and inspect a plan).
#pragma once

#include <cstdint>
#include <string>
#include <vector>

namespace fm {

constexpr int kChainMaxBlk = 8;               // block slots per node (two per lane of a quad)
constexpr int kChainDepBlk = kChainMaxBlk - 1;  // dependency blocks per node and direction: the last
                                                // slot carries the node's diagonal block
constexpr int kChainRing = 8;                 // LDS ring depth in steps
// largest published vector (bytes) the sweep kernels address: their buffer
// offsets are 32-bit, and offsets from 2^31 up mark lanes with no access
constexpr uint32_t kChainMaxBytes = 0x80000000u;
constexpr uint32_t kDepNone = 0xFFFFFFFFu;    // no dependency in this lane / idle lane group
constexpr uint32_t kDepLocal = 0x80000000u;   // | LDS slot: dependency inside the bundle

// Device form of one wave-step (lpn 4), kRecWords 32-bit words:
//   word 2 * lane + j  dependency block j of the lane: the byte offset of an
//                      external dependency's node triple in an FP32 vector
//                      (12 * node, below 2^31), or kDepLocal | LDS ring slot
//                      (kZeroSlot, a slot that holds zeros, when absent)
//   word 128 + g       byte offset of quad g's node triple (12 * node),
//                      kOffNone for an idle quad
// FP64 kernels double the offsets; kOffNone stays past every buffer either way.
constexpr uint32_t kOffNone = 0xFFFFFFF0u;
constexpr uint32_t kZeroSlot = kChainRing * 16;  // the slot after the ring (kChainRing steps x 16 quads)
constexpr uint32_t kRecWords = 128 + 16;

// One sweep direction.  Wave-step q of bundle b is bstart[b] + s for step s <
// bsteps[b]; bundles are listed in processing order.
//   node[q * (64/lpn) + g]            node of lane group g at wave-step q
//   dep[(q * bpl + j) * 64 + lane]    dependency word of block j of the lane:
//       kDepNone, kDepLocal | slot (values in LDS slot (step % ring) * 64/lpn + group),
//       or the node id whose published values are awaited
// Lane gl of a group holds the node's dependency blocks gl, gl + lpn, ...
// (lower blocks in column order for the forward sweep, upper blocks for the
// backward sweep).
struct ChainPlanDir {
    uint32_t nb = 0, nq = 0;
    std::vector<uint32_t> bstart, bsteps, node, dep;
};

struct ChainPlan {
    int lpn = 4;
    uint32_t chains = 0, stride = 0, gj = 0, gk = 0;
    ChainPlanDir fwd, bwd;
};

// Plan for a 3-row node-blocked CSR with sorted columns (only rows 3v are
// read, so a compact view holding the node columns in row 3v suffices).
// Returns false with the reason in `why` when the matrix does not fit.
// Device form of one direction (lpn 4): kRecWords words per wave-step, as
// described above.
void encode_chain_records(const ChainPlanDir& H, int lpn, std::vector<uint32_t>& far);

bool plan_chain_schedule(const std::vector<uint32_t>& rp, const std::vector<uint32_t>& ci, int lpn, ChainPlan& P,
                         std::string& why);

}  // namespace fm
