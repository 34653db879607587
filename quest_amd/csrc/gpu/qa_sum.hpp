// Compensated FP64 accumulation shared by the reductions (qa_reduce.hip)
// and the serving engine's batched expectations (qa_serve.hip).
#pragma once

#include "qa_common.hpp"

namespace qa {

struct KB {  // compensated accumulator
    double s, c;
};

__device__ __forceinline__ void kadd(KB& a, double v) {
    double t = a.s + v;
    a.c += (fabs(a.s) >= fabs(v)) ? ((a.s - t) + v) : ((v - t) + a.s);
    a.s = t;
}
__device__ __forceinline__ void kmerge(KB& a, KB b) {
    double t = a.s + b.s;
    double e = (fabs(a.s) >= fabs(b.s)) ? ((a.s - t) + b.s) : ((b.s - t) + a.s);
    a.s = t;
    a.c += b.c + e;
}

// reduce NV compensated accumulators across the workgroup; lane 0 of wave 0
// ends holding the result
template <int NV>
__device__ void blockMerge(KB (&acc)[NV]) {
    __shared__ double sh[BLOCK / 64][NV][2];
#pragma unroll
    for (int off = 32; off >= 1; off >>= 1) {
#pragma unroll
        for (int v = 0; v < NV; v++) {
            KB o;
            o.s = __shfl_xor(acc[v].s, off, 64);
            o.c = __shfl_xor(acc[v].c, off, 64);
            kmerge(acc[v], o);
        }
    }
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    if (lane == 0)
        for (int v = 0; v < NV; v++) {
            sh[wave][v][0] = acc[v].s;
            sh[wave][v][1] = acc[v].c;
        }
    __syncthreads();
    if (threadIdx.x == 0)
        for (int w = 1; w < BLOCK / 64; w++)
            for (int v = 0; v < NV; v++) kmerge(acc[v], KB{sh[w][v][0], sh[w][v][1]});
}

}  // namespace qa
