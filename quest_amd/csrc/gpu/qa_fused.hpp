// Interface between the gate entry points and the fusion queue (qa_fused.hip).
#pragma once

#include <cstdint>

#include "qa_kinds.hpp"

namespace qa {
namespace fuse {


// a queued gate in global index space
struct GOp {
    int kind = 0;
    int t1 = -1, t2 = -1;            // targets (DEPOL1/DEPH1: ket, bra)
    uint64_t cMask = 0, cVal = 0;    // control condition: (i & cMask) == cVal
    uint64_t m1 = 0, m2 = 0;         // PARITY: parity mask; PAULI: maskXY, maskYZ; DENSE1V: variant qubits
    double f[32] = {0};              // matrix / factors (re, im interleaved); DENSE1V: one 2x2
                                     // matrix per value w of the (at most 2) variant qubits, at
                                     // f[8w], bit b of w = the b-th lowest variant qubit
    // DENSEN: targets (matrix index bit b <-> targs[b]), and the row-major
    // 2^n x 2^n matrix {re, im} in device memory staged by stageMatrix
    // (conj: its imaginary parts negated where read)
    int targs[4] = {-1, -1, -1, -1};
    int numTargs = 0;
    int conj = 0;
    const double* mat = nullptr;
};

// copies a DENSEN matrix (count doubles, device or host memory) into the
// library's matrix ring, in stream order; the copy lives until the queue
// that holds the gate is applied (a ring wrap applies every queue first).
// nullptr on failure (error set).
const double* stageMatrix(const void* mat, size_t count);

struct Config {
    bool enabled = false;
    int tileQubits = 12;   // tile = 2^12 amplitudes (64 KiB of LDS in FP64)
    int lowQubits = 4;     // always-local low qubits: tile rows are 2^4 contiguous amplitudes
    int maxQueue = 512;    // flush when this many gates are queued
    int minQubits = 1;
    int prefetch = 2;      // tile passes prefetch each workgroup's next tile into registers: 1 at
                           // the loop head, 2 spread over the program (generated kernels), 3 as 2
                           // with the stores deferred a tile and spread too, 0 not
    bool exact = false;    // passes apply gates in queue order, each with its single-gate arithmetic
    int spanQubits = 12;   // runs of Pauli and phase gadgets may go to span passes over cosets of
                           // 2^spanQubits amplitudes (0 = off)
    int blockMinGates = 0; // complex64 states: runs of at least this many gates on 5 qubits are
                           // multiplied into one 32x32 block applied with FP32 MFMA (0 = off)
};

bool enabledFor(void* amps, int prec, int numQubits);
// appends the gate; `accepted` is false when the gate cannot be fused (the
// caller then flushes this state's queue and runs the gate directly)
int enqueue(void* amps, int prec, int numQubits, const GOp& op, bool& accepted);
int flush(void* amps);
int flushAll();
// true when gates are queued for this state (a flush would launch passes)
bool pending(const void* amps);
void drop(void* amps);
void shutdown();  // discards every queue and frees the program ring

// control condition as masks over the index: (i & m) == v
inline void ctrlMasks(const int* ctrls, const int* states, int nc, uint64_t& m, uint64_t& v) {
    m = v = 0;
    for (int i = 0; i < nc; i++) {
        m |= 1ull << ctrls[i];
        if (!states || states[i]) v |= 1ull << ctrls[i];
    }
}

// queue the gate when fusion is on; `done` tells the caller whether it was
// queued.  A gate the queue refuses first flushes the state's pending gates,
// then the caller runs it with its own kernel.
inline int tryFuse(void* amps, int prec, int nq, const GOp& op, bool& done) {
    done = false;
    if (!enabledFor(amps, prec, nq)) return 0;
    bool accepted = false;
    int rc = enqueue(amps, prec, nq, op, accepted);
    if (rc != 0) return rc;
    if (accepted) {
        done = true;
        return 0;
    }
    return flush(amps);
}

}  // namespace fuse
}  // namespace qa

// gate entry points: return early when the gate was queued
#define QA_TRY_FUSE(amps, prec, nq, op)                                    \
    do {                                                                   \
        bool done_ = false;                                                \
        int rct_ = qa::fuse::tryFuse((amps), (prec), (nq), (op), done_);  \
        if (rct_ != QA_OK) return rct_;                                    \
        if (done_) return QA_OK;                                           \
    } while (0)

// every entry point that reads or writes amplitudes outside the queue
// starts with this: queued gates are applied first
#define QA_FLUSH_ALL()                      \
    do {                                    \
        int rcf_ = qa::fuse::flushAll();    \
        if (rcf_ != QA_OK) return rcf_;     \
    } while (0)
