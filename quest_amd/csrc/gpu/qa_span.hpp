// Span passes: a run of uncontrolled Pauli gadgets (and Pauli tensors) and
// phase gadgets applied in one pass over the state, whatever qubits they
// touch.
//
// A Pauli gadget mixes the amplitude pairs (x, x ^ m) of its X/Y mask m and
// a phase gadget is diagonal.  Let V be a GF(2) subspace of dimension k that
// holds the X/Y masks of a run of gadgets, with basis b_0 .. b_{k-1} in
// reduced echelon form (each b_j has a pivot bit that no other basis vector
// has).  The state splits into the 2^(n-k) cosets rep ^ V, rep ranging over
// the indices with every pivot bit clear, and each gadget of the run maps
// every coset into itself.  A workgroup loads one coset, 2^k amplitudes,
// into LDS at its coordinates s (the amplitude rep ^ sum_j s_j b_j), applies
// the run in queue order and writes it back.  In coordinates a gadget with
// X/Y mask m and Y/Z mask z is the pair update over (s, s ^ c), c = the
// coordinates of m, with the sign of an amplitude the parity of
// (s & d) ^ (rep & z), d_j = parity(b_j & z); a phase gadget is the same
// parity rule on its Z mask.
//
// In LDS the gadgets go in batches whose X/Y coordinates span at most 4
// dimensions (a register basis B_0..B_3 in reduced echelon form over the
// coordinates): each thread takes one coset of span(B) -- its 16
// coordinates s0 ^ S4[v], s0 its thread id deposited over the coordinates
// that are not pivots of B -- into registers, applies the batch there with
// the tile passes' register ops (regPauli, regParity: the pair (v, v ^ c4),
// c4 = the gadget's X/Y coordinates over B, the sign parity(v & r4) ^
// parity(s0 & d) ^ parity(rep & z)) and puts it back: one LDS sweep and
// barrier per batch instead of per gadget.  Each pair gets the arithmetic
// of the single-gate kernels (cmat2T, cmulT), so a span pass is
// bit-identical to the gates one by one.
//
// The first basis vectors are the unit vectors of the lowest qubits, so the
// lanes of a wave read rows of 2^low contiguous amplitudes.  A tile pass
// needs every X/Y qubit of a gadget among its 13 local qubits, so random
// gadgets of 4-12 qubits fit one or two to a tile pass; a span pass takes
// as many as add at most k - low dimensions.
#pragma once

#include "qa_tiledev.hpp"

namespace qa {
namespace fuse {

// the records of a span pass: per batch, a header (SPAN_BATCH), a table
// (the 16 coordinates of the batch's register span, as 16 uint32 in the
// record's 64 bytes) and the batch's gadgets
constexpr int SPAN_BATCH = 200;

struct SpanOp {
    int kind;       // F_PAULI, F_PARITY, or SPAN_BATCH
    uint32_t c;     // PAULI: the X/Y mask over the batch's register basis (1..15);
                    // SPAN_BATCH: the number of gadget records after the table
    uint32_t r;     // the sign (PAULI: Y/Z, PARITY: Z) mask's parity over the batch's
                    // register basis; SPAN_BATCH: the number of non-pivot coordinates
    uint32_t d;     // the sign mask's parity over all k coordinates
    uint64_t z;     // the same mask over the global qubits (the coset's own parity)
    uint64_t dep;   // SPAN_BATCH: the non-pivot coordinates, 4 bits each, ascending
    double f[4];    // PAULI: ampFac, pairFac; PARITY: the factors for even and odd parity
};
static_assert(sizeof(SpanOp) == 64, "span ops are 64 bytes");

struct SpanArgs {
    uint64_t numCosets;     // 2^(n - k)
    uint64_t repMask;       // the non-pivot qubits
    uint64_t repStep;       // the grid size deposited over them
    int k;                  // coordinates; the workgroup has 2^(k - RB) threads
    int numRep;             // number of non-pivot qubits
    int numOps;
    int pad;
    uint64_t basis[MAXQ];   // b_j over the global qubits
    uint64_t top[RA];       // top[i]: the sum of the last RB basis vectors selected by the bits of i
    int repPos[64];         // the non-pivot qubits, ascending
};

// Thread t of the workgroup holds coordinates t + i * 2^(k - RB), i < RA:
// its address part `own` is the sum of b_j over the bits j of t, and
// register i adds top[i]
template <typename T, int NT>
__device__ __forceinline__ void spanLoad(ct<T>* r, const T* __restrict__ amps, uint64_t at, const SpanArgs& a) {
#pragma unroll
    for (int i = 0; i < RA; i++) r[i] = ldx<NT>(amps, at ^ a.top[i]);
}

// the thread's address part, made opaque at each use so that the sixteen
// register addresses are formed there, not hoisted out of the coset loop
// into 32 more live registers
__device__ __forceinline__ uint64_t opaqueOwn(uint64_t own) {
    asm volatile("" : "+v"(own));
    return own;
}

template <typename T, int NT>
__device__ __forceinline__ void spanPassLoop(T* __restrict__ amps, const SpanArgs& a,
                                             const SpanOp* __restrict__ ops) {
    extern __shared__ __align__(16) unsigned char smem[];
    ct<T>* tile = reinterpret_cast<ct<T>*>(smem);
    const uint32_t tid = threadIdx.x, nth = blockDim.x;
    const int lowK = a.k - RB;
    uint64_t own = 0;
    for (int j = 0; j < lowK; j++)
        if ((tid >> j) & 1u) own ^= a.basis[j];
    uint64_t w = blockIdx.x;
    if (w >= a.numCosets) return;
    uint64_t rep = deposit(w, a.repPos, a.numRep);
    ct<T> rn[RA];
    spanLoad<T, NT>(rn, amps, rep ^ own, a);
    waitLoadsBeforeStores();
    for (; w < a.numCosets; w += gridDim.x) {
#pragma unroll
        for (int i = 0; i < RA; i++) tile[tid + i * nth] = rn[i];
        const uint64_t repN = ((rep | ~a.repMask) + a.repStep) & a.repMask;
        if (w + gridDim.x < a.numCosets) spanLoad<T, NT>(rn, amps, repN ^ opaqueOwn(own), a);
        __syncthreads();
        for (int o = 0; o < a.numOps;) {
            // a batch: thread t takes the 16 coordinates s0 ^ S4[v] of the
            // coset of the batch's register span whose representative s0 is
            // t deposited over the non-pivot coordinates, applies the
            // batch's gadgets in registers, and puts them back
            const SpanOp& hb = ops[o];
            const uint32_t* __restrict__ S4 = reinterpret_cast<const uint32_t*>(&ops[o + 1]);
            const int nops = (int)hb.c, ndep = (int)hb.r;
            uint32_t s0 = 0;
            for (int j = 0; j < ndep; j++) s0 |= ((tid >> j) & 1u) << ((hb.dep >> (4 * j)) & 15u);
            ct<T> x[RA];
#pragma unroll
            for (int v = 0; v < RA; v++) x[v] = tile[s0 ^ S4[v]];
            for (int g = o + 2; g < o + 2 + nops; g++) {
                const SpanOp& op = ops[g];
                TOpHead h;
                h.rm1 = op.r;
                h.rm2 = op.r;
                const int par = parity64(rep & op.z) ^ (__popc(s0 & op.d) & 1);
                if (op.kind == F_PARITY) {
                    regParity<T, false>(x, h, cf<T>(op.f, 0), cf<T>(op.f, 1), par);
                } else {
                    switch (op.c) {
#define QA_SPAN_CASE(M) case M: regPauli<T, M, false, false>(x, h, op.f, par, true); break;
                        QA_SPAN_CASE(1) QA_SPAN_CASE(2) QA_SPAN_CASE(3) QA_SPAN_CASE(4) QA_SPAN_CASE(5)
                        QA_SPAN_CASE(6) QA_SPAN_CASE(7) QA_SPAN_CASE(8) QA_SPAN_CASE(9) QA_SPAN_CASE(10)
                        QA_SPAN_CASE(11) QA_SPAN_CASE(12) QA_SPAN_CASE(13) QA_SPAN_CASE(14) QA_SPAN_CASE(15)
#undef QA_SPAN_CASE
                    }
                }
            }
#pragma unroll
            for (int v = 0; v < RA; v++) tile[s0 ^ S4[v]] = x[v];
            __syncthreads();
            o += 2 + nops;
        }
        const uint64_t at = rep ^ opaqueOwn(own);
        // the next coset's loads retire before the stores, which then drain
        // while the next coset is worked on
        waitLoadsBeforeStores();
#pragma unroll
        for (int i = 0; i < RA; i++) stx<NT>(amps, at ^ a.top[i], tile[tid + i * nth]);
        rep = repN;
        // the next coset's LDS writes must not overtake this coset's reads
        __syncthreads();
    }
}

template <typename T, int NT>
__global__ void __launch_bounds__(512) k_spanPass(T* __restrict__ amps, SpanArgs a, const SpanOp* __restrict__ ops) {
    spanPassLoop<T, NT>(amps, a, ops);
}

}  // namespace fuse
}  // namespace qa
