// The host-side planner of tile passes (qa_fused.hip): turns a run of queued
// gates into a pass program; the device side that runs it is qa_tiledev.hpp.
#pragma once

#include <cstdint>
#include <vector>

#include "qa_common.hpp"
#include "qa_fused.hpp"
#include "qa_span.hpp"
#include "qa_tiledev.hpp"

namespace qa {
namespace fuse {

// ---- host-side planner -------------------------------------------------------

// resident workgroups of a kernel on the whole current device, for
// persistent grids (cached per kernel, block size and LDS bytes)
unsigned residentBlocks(const void* fn, unsigned block, size_t lds);

// true when a pass kernel has no static LDS, so that its dynamic LDS tile
// starts at LDS address 0 (the register-layout exchanges rely on it)
bool ldsTileAtZero(const void* fn);

// qubits whose amplitude pairs an op mixes (all must be local in its pass;
// diagonal ops may involve any qubit)
uint64_t denseMask(const GOp& g);

// every qubit an op involves: dense and diagonal targets, controls, parity
// and Pauli masks
uint64_t supportMask(const GOp& g);

// the longest prefix of ops[0..num) whose dense targets, with the `low`
// lowest qubits, fit kMax local qubits; localMask is set to that local set
// padded to exactly kMax qubits with the lowest free ones
size_t takeRun(const GOp* ops, size_t num, int numQubits, int kMax, int low, uint64_t& localMask);

// the ops of the next pass when ops may be reordered: scanning ops[0..num)
// in order, an op joins the pass when its dense targets fit the local set
// (grown up to kMax qubits, the `low` lowest always in it) and it commutes
// with every op passed over before it.  Two ops are taken to commute when
// neither involves a dense target of the other: on their shared qubits both
// are then block-diagonal, so they commute exactly (their roundings differ
// from the queue order).  Appends the positions taken, in queue order, to
// `picked` and sets localMask as takeRun does.
void pickPass(const GOp* ops, size_t num, int numQubits, int kMax, int low, uint64_t& localMask,
              std::vector<uint32_t>& picked);

// ops[0..num) of one pass (dense targets inside localMask) with every run
// of gates confined to the same one or two qubits (controls included)
// multiplied, in FP64 on the host, into one dense matrix applied as one op:
// each gate joins the open run that shares a qubit with it, runs that a gate
// bridges are joined, and a run is emitted when a gate involving its qubits
// cannot join it (so everything emitted before it commutes with it).  A run
// on one local qubit whose gates also involve one or two non-local qubits
// (as controls or diagonal factors, which makes the run block-diagonal in
// them) becomes one DENSE1V op: a 2x2 matrix per value of those qubits,
// picked per tile.  Runs of diagonal gates only, and runs of one gate, are
// emitted unchanged.
void mergePass(const GOp* ops, size_t num, uint64_t localMask, std::vector<GOp>& out);

// the ops of one pass reordered so that ops on the same register group are
// applied together: rounds that each take, in order, every op that commutes
// with the ops passed over in the round and whose dense targets fit, with
// those taken so far, in RB qubits (a Pauli op applied in LDS ends a round's
// dense ops).  Each round then needs one register layout.  The first
// round leaves the `avoid` qubits (the always-local low qubits) to later
// rounds when it can still take a dense op without them, so that the pass
// loads HBM in its first layout instead of through LDS.
void clusterPass(const GOp* ops, size_t num, uint64_t avoid, std::vector<GOp>& out);

// the program of one pass applying ops[0..num) (dense targets all inside
// localMask) to a state of numQubits index bits: fills every field of `a`
// but the LDS offsets and appends the ops, with their GROUP ops, to `prog`;
// with `phases`, runs of diagonal gates become PHASES ops (their factors are
// then multiplied before they reach the amplitudes, which changes rounding)
void planPass(const GOp* ops, size_t num, int numQubits, uint64_t localMask, int lowQubits, TileArgs& a,
              std::vector<TOp>& prog, bool phases = false);

// the same for a complex64 state, where runs of at least blockMinGates ops
// whose qubits (targets, controls, masks) fit 5 local qubits become one
// BLOCK op; their packed 32x32 matrices are appended to `blocks` and each
// BLOCK op holds its offset there (a.blocks is left for the caller)
void planPassBlocks(const GOp* ops, size_t num, int numQubits, uint64_t localMask, int lowQubits, int blockMinGates,
                    TileArgs& a, std::vector<TOp>& prog, std::vector<float>& blocks);

// the LDS byte offsets of the register layouts of a planned pass (prog[0..num),
// whose first op is a GROUP), for amplitudes of precision prec: the fx words
// of its GROUP, PAULI_LDS and BLOCK ops, and a.rowOff / a.lastOff
void setLdsOffsets(TOp* prog, size_t num, TileArgs& a, int prec);

// ---- span passes (qa_span.hpp) -----------------------------------------------

constexpr int MAX_SPAN_OPS = 128;

// true for the ops a span pass takes: Pauli gadgets and tensors, and phase
// gadgets, without controls
bool spanEligible(const GOp& g);

// the longest run at the front of ops[0..num) (at most MAX_SPAN_OPS) of
// span-eligible ops whose X/Y masks, with the `low` lowest qubits, span at
// most k dimensions, planned as one span pass over cosets of 2^k
// amplitudes: fills every field of `a` but repStep and the ops in `prog`.
// Returns the run's length (0 when n <= k or k is outside [10, 13]).
size_t spanRun(const GOp* ops, size_t num, int numQubits, int k, int low, SpanArgs& a, std::vector<SpanOp>& prog);

// the span-pass kernel and its launch (qa_span.hip)
const void* spanKernel(int prec, bool nt);
void launchSpanAny(int prec, bool nt, unsigned grid, unsigned block, size_t lds, void* amps, const SpanArgs& a,
                   const SpanOp* ops);

}  // namespace fuse
}  // namespace qa
