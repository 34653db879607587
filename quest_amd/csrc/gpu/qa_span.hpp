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
// parity rule on its Z mask.  Each pair gets the arithmetic of the
// single-gate kernels (cmat2T, cmulT), so a span pass is bit-identical to
// the gates one by one.
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

// one gadget in coset coordinates
struct SpanOp {
    int kind;       // F_PAULI or F_PARITY
    uint32_t c;     // PAULI: the X/Y mask's coordinates
    uint32_t top;   // PAULI: the highest bit of c (pairs are enumerated with it clear)
    uint32_t d;     // the Y/Z (PARITY: Z) mask's parity over the basis vectors
    uint64_t z;     // the same mask over the global qubits (the coset's own parity)
    uint64_t pad;
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
    const uint32_t half = 1u << (a.k - 1);
    uint64_t w = blockIdx.x;
    if (w >= a.numCosets) return;
    uint64_t rep = deposit(w, a.repPos, a.numRep);
    ct<T> rn[RA];
    spanLoad<T, NT>(rn, amps, rep ^ own, a);
    for (; w < a.numCosets; w += gridDim.x) {
#pragma unroll
        for (int i = 0; i < RA; i++) tile[tid + i * nth] = rn[i];
        const uint64_t repN = ((rep | ~a.repMask) + a.repStep) & a.repMask;
        if (w + gridDim.x < a.numCosets) spanLoad<T, NT>(rn, amps, repN ^ own, a);
        __syncthreads();
        for (int o = 0; o < a.numOps; o++) {
            const SpanOp& op = ops[o];
            const int rp = parity64(rep & op.z);
            if (op.kind == F_PARITY) {
                const ct<T> f0 = cf<T>(op.f, 0), f1 = cf<T>(op.f, 1);
#pragma unroll
                for (int i = 0; i < RA; i++) {
                    const uint32_t s = tid + i * nth;
                    const ct<T> f = ((__popc(s & op.d) & 1) ^ rp) ? f1 : f0;
                    tile[s] = cmulT(f, tile[s]);
                }
                // a run of phase gadgets touches each amplitude from the same thread
                if (o + 1 < a.numOps && ops[o + 1].kind == F_PARITY) continue;
            } else {
                const ct<T> af = cf<T>(op.f, 0), pf = cf<T>(op.f, 1);
                const uint32_t lowMask = (1u << op.top) - 1u;
                for (uint32_t j = tid; j < half; j += nth) {
                    const uint32_t s = ((j & ~lowMask) << 1) | (j & lowMask);
                    const uint32_t v = s ^ op.c;
                    const T sA = ((__popc(s & op.d) & 1) ^ rp) ? (T)-1 : (T)1;
                    const T sB = ((__popc(v & op.d) & 1) ^ rp) ? (T)-1 : (T)1;
                    ct<T> x0 = tile[s], x1 = tile[v];
                    cmat2T(af, ct<T>{pf.re * sB, pf.im * sB}, ct<T>{pf.re * sA, pf.im * sA}, af, x0, x1);
                    tile[s] = x0;
                    tile[v] = x1;
                }
            }
            __syncthreads();
        }
        const uint64_t at = rep ^ own;
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
