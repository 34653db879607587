// FP64 reductions (probabilities, purity, Pauli expectations, inner
// products, fidelities) replacing the Thrust transform_reduce calls of
// quest/src/gpu/gpu_thrust.cuh:745-1005 and gpu_subroutines.cpp:1440-1723.
//
// Two kernels per reduction, both deterministic for a given problem size:
//  1. a fixed-size grid (at most RED_BLOCKS workgroups) sweeps the terms in a
//     grid-stride loop, 4 independent loads in flight per lane; each lane
//     keeps a compensated (Neumaier) FP64 sum, waves combine lanes with
//     cross-lane shuffles and a two-sum merge, workgroups combine their 4
//     waves through LDS and write one (sum, compensation) pair;
//  2. one workgroup merges the partials in a fixed order and writes the final
//     value, which is copied to pinned host memory.
// Amplitudes are widened to FP64 before squaring, so FP32 states reduce in
// FP64 too.
#include <algorithm>
#include <vector>

#include "qa_common.hpp"
#include "qa_fused.hpp"
#include "qa_sum.hpp"

namespace qa {
namespace {

constexpr int RED_BLOCKS = 1024;
constexpr int MAXV = 2;  // values reduced at once (complex results)
constexpr size_t RESULT_OFF = (size_t)RED_BLOCKS * MAXV * 2;

template <int NV, typename F>
__global__ void __launch_bounds__(BLOCK) k_reduce(F f, uint64_t numWork, double* __restrict__ partials) {
    KB acc[NV];
#pragma unroll
    for (int v = 0; v < NV; v++) acc[v] = KB{0, 0};
    const uint64_t stride = (uint64_t)gridDim.x * BLOCK;
    uint64_t n = (uint64_t)blockIdx.x * BLOCK + threadIdx.x;
    for (; n + 3 * stride < numWork; n += 4 * stride) {
        double t[4][NV];
#pragma unroll
        for (int u = 0; u < 4; u++) f(n + (uint64_t)u * stride, t[u]);
#pragma unroll
        for (int u = 0; u < 4; u++)
#pragma unroll
            for (int v = 0; v < NV; v++) kadd(acc[v], t[u][v]);
    }
    for (; n < numWork; n += stride) {
        double t[NV];
        f(n, t);
#pragma unroll
        for (int v = 0; v < NV; v++) kadd(acc[v], t[v]);
    }
    blockMerge<NV>(acc);
    if (threadIdx.x == 0)
        for (int v = 0; v < NV; v++) {
            partials[(blockIdx.x * NV + v) * 2] = acc[v].s;
            partials[(blockIdx.x * NV + v) * 2 + 1] = acc[v].c;
        }
}

template <int NV>
__global__ void __launch_bounds__(BLOCK) k_finalize(const double* __restrict__ partials, int numParts,
                                                   double* __restrict__ out) {
    KB acc[NV];
#pragma unroll
    for (int v = 0; v < NV; v++) acc[v] = KB{0, 0};
    for (int p = threadIdx.x; p < numParts; p += BLOCK)
        for (int v = 0; v < NV; v++) kmerge(acc[v], KB{partials[(p * NV + v) * 2], partials[(p * NV + v) * 2 + 1]});
    blockMerge<NV>(acc);
    if (threadIdx.x == 0)
        for (int v = 0; v < NV; v++) out[v] = acc[v].s + acc[v].c;
}

template <int NV, typename F>
int runReduce(F f, uint64_t numWork, double* out) {
    double* scratch = devScratch(RESULT_OFF + 8);
    double* host = hostScratch(8);
    if (!scratch || !host) {
        setError("reduction scratch unavailable");
        return QA_ERR_OOM;
    }
    uint64_t want = (numWork + (uint64_t)BLOCK * 4 - 1) / ((uint64_t)BLOCK * 4);
    int blocks = (int)(want < 1 ? 1 : (want > RED_BLOCKS ? RED_BLOCKS : want));
    k_reduce<NV, F><<<blocks, BLOCK, 0, stream()>>>(f, numWork, scratch);
    QA_LAUNCH_CHECK();
    k_finalize<NV><<<1, BLOCK, 0, stream()>>>(scratch, blocks, scratch + RESULT_OFF);
    QA_LAUNCH_CHECK();
    QA_HIP(hipMemcpyAsync(host, scratch + RESULT_OFF, NV * sizeof(double), hipMemcpyDeviceToHost, stream()));
    QA_HIP(hipStreamSynchronize(stream()));
    for (int v = 0; v < NV; v++) out[v] = host[v];
    return QA_OK;
}

// runReduce for NV values (more than MAXV): partials and results in a
// larger slice of the same scratch
template <int NV, typename F>
int runReduceN(F f, uint64_t numWork, double* out) {
    const size_t resOff = (size_t)RED_BLOCKS * NV * 2;
    double* scratch = devScratch(resOff + NV);
    double* host = hostScratch(NV);
    if (!scratch || !host) {
        setError("reduction scratch unavailable");
        return QA_ERR_OOM;
    }
    uint64_t want = (numWork + (uint64_t)BLOCK * 4 - 1) / ((uint64_t)BLOCK * 4);
    int blocks = (int)(want < 1 ? 1 : (want > RED_BLOCKS ? RED_BLOCKS : want));
    k_reduce<NV, F><<<blocks, BLOCK, 0, stream()>>>(f, numWork, scratch);
    QA_LAUNCH_CHECK();
    k_finalize<NV><<<1, BLOCK, 0, stream()>>>(scratch, blocks, scratch + resOff);
    QA_LAUNCH_CHECK();
    QA_HIP(hipMemcpyAsync(host, scratch + resOff, NV * sizeof(double), hipMemcpyDeviceToHost, stream()));
    QA_HIP(hipStreamSynchronize(stream()));
    for (int v = 0; v < NV; v++) out[v] = host[v];
    return QA_OK;
}

template <typename T>
__device__ __forceinline__ cd ldd(const T* __restrict__ a, uint64_t i) {
    ct<T> v = ld(a, i);
    return cd{(double)v.re, (double)v.im};
}
__device__ __forceinline__ double norm2(cd a) { return a.re * a.re + a.im * a.im; }

// ---- term functors --------------------------------------------------------

// both outcome probabilities of PQ qubits q0..q0+PQ-1 in one read of the
// state: |a_n|^2 goes to o[b] when qubit q0 + b is 0 in n, to o[PQ + b]
// when it is 1, so each outcome is the compensated sum of its own terms
// (never a difference of two sums, which would lose a small probability)
constexpr int PQ = 10;
template <typename T>
struct FProbsOfQubits {
    const T* a;
    int q0;
    __device__ void operator()(uint64_t n, double* o) const {
        const double p = norm2(ldd(a, n));
#pragma unroll
        for (int b = 0; b < PQ; b++) {
            const bool one = (n >> (q0 + b)) & 1ull;
            o[b] = one ? 0.0 : p;
            o[PQ + b] = one ? p : 0.0;
        }
    }
};

// Both outcome probabilities of every qubit of a statevector of 10..PQALL
// qubits in one read.  A wave reads 1024-amplitude chunks, lane l the
// amplitudes chunk*1024 + 64 i + l (i < 16), so in each amplitude's index
// bits 0-5 are the lane's, bits 6-9 are i's and the higher bits the
// chunk's.  Each lane keeps: its total (bits 0-5 are constant for the lane;
// the workgroup merge splits the totals by lane bit), the zero- and
// one-sums of bits 6-9 (sums of 8 of its 16 terms per chunk), and for each
// higher bit the zero- and one-sums of its chunk totals.  Every outcome is
// a sum of its own terms, never a difference.  The merges are compensated;
// a lane's chunk totals for the high bits are plain FP64 sums (at most a
// few hundred per lane).
constexpr int PQALL = 32;
constexpr int PQALL_VALUES = 2 * PQALL;  // (zero, one) per qubit, merged 8 at a time
template <typename T>
__global__ void __launch_bounds__(BLOCK) k_probsOfEveryQubit(const T* __restrict__ a, int nq, uint64_t numChunks,
                                                             double* __restrict__ partials) {
    const int lane = threadIdx.x & 63;
    const uint64_t waves = (uint64_t)gridDim.x * (BLOCK / 64);
    KB tot = {0, 0};
    KB mid[8];
#pragma unroll
    for (int j = 0; j < 8; j++) mid[j] = KB{0, 0};
    double hi[2 * (PQALL - 10)];
#pragma unroll
    for (int j = 0; j < 2 * (PQALL - 10); j++) hi[j] = 0;
    for (uint64_t c = (uint64_t)blockIdx.x * (BLOCK / 64) + (threadIdx.x >> 6); c < numChunks; c += waves) {
        const uint64_t base = (c << 10) | (uint64_t)lane;
        double p[16];
#pragma unroll
        for (int i = 0; i < 16; i++) p[i] = norm2(ldd(a, base + 64 * i));
        double g = 0;
#pragma unroll
        for (int i = 0; i < 16; i++) g += p[i];
        kadd(tot, g);
#pragma unroll
        for (int b = 0; b < 4; b++) {
            double z = 0, o = 0;
#pragma unroll
            for (int i = 0; i < 16; i++) {
                if ((i >> b) & 1) o += p[i];
                else z += p[i];
            }
            kadd(mid[2 * b], z);
            kadd(mid[2 * b + 1], o);
        }
#pragma unroll
        for (int b = 0; b < PQALL - 10; b++) {
            const bool one = (c >> b) & 1ull;
            hi[2 * b] += one ? 0.0 : g;
            hi[2 * b + 1] += one ? g : 0.0;
        }
    }
    // value 2q + o: qubit q's outcome o; merged across the workgroup in
    // groups of 8, written group-major: partials[((grp * blocks) + block) * 8 + v][2]
#pragma unroll
    for (int grp = 0; grp < PQALL_VALUES / 8; grp++) {
        KB t[8];
#pragma unroll
        for (int v = 0; v < 8; v++) {
            const int val = grp * 8 + v, q = val >> 1, o = val & 1;
            if (q < 6) t[v] = (((lane >> q) & 1) == o) ? tot : KB{0, 0};
            else if (q < 10) t[v] = mid[2 * (q - 6) + o];
            else t[v] = KB{hi[2 * (q - 10) + o], 0};
        }
        blockMerge<8>(t);
        if (threadIdx.x == 0)
#pragma unroll
            for (int v = 0; v < 8; v++) {
                const size_t at = (((size_t)grp * gridDim.x + blockIdx.x) * 8 + v) * 2;
                partials[at] = t[v].s;
                partials[at + 1] = t[v].c;
            }
        // the next group's merge reuses the workgroup's LDS
        __syncthreads();
    }
    (void)nq;
}

template <typename T>
struct FNormAll {  // |a_n|^2
    const T* a;
    __device__ void operator()(uint64_t n, double* o) const { o[0] = norm2(ldd(a, n)); }
};
template <typename T>
struct FNormOutcome {  // |a_i|^2, i = n with outcome bits inserted
    const T* a;
    InsertSpec ins;
    __device__ void operator()(uint64_t n, double* o) const { o[0] = norm2(ldd(a, insertBits(n, ins))); }
};
template <typename T>
struct FDiagReal {  // Re rho_rr, r = n with outcome bits inserted
    const T* a;
    InsertSpec ins;
    uint64_t diagStride;  // 2^N + 1
    __device__ void operator()(uint64_t n, double* o) const {
        o[0] = ldd(a, insertBits(n, ins) * diagStride).re;
    }
};
template <typename T>
struct FExpecZ {
    const T* a;
    uint64_t mask;
    __device__ void operator()(uint64_t n, double* o) const {
        double v = norm2(ldd(a, n));
        o[0] = parity64(n & mask) ? -v : v;
    }
};
template <typename T>
struct FDensExpecZ {
    const T* a;
    uint64_t mask, diagStride;
    __device__ void operator()(uint64_t n, double* o) const {
        cd v = ldd(a, n * diagStride);
        double s = parity64(n & mask) ? -1.0 : 1.0;
        o[0] = s * v.re;
        o[1] = s * v.im;
    }
};
template <typename T>
struct FExpecPauli {  // sign(j & maskYZ) conj(a_n) a_j, j = n ^ maskXY
    const T* a;
    uint64_t maskXY, maskYZ;
    __device__ void operator()(uint64_t n, double* o) const {
        uint64_t j = n ^ maskXY;
        cd x = ldd(a, n), y = ldd(a, j);
        cd t = cmul(cconj(x), y);
        double s = parity64(j & maskYZ) ? -1.0 : 1.0;
        o[0] = s * t.re;
        o[1] = s * t.im;
    }
};
// the same sum visited one pair at a time: n with the top X/Y bit clear and
// j = n ^ maskXY contribute sign(j) conj(a_n) a_j + sign(n) conj(a_j) a_n,
// so every amplitude is read once (the plain form reads each twice)
template <typename T>
struct FExpecPauliPair {
    const T* a;
    InsertSpec topBit;
    uint64_t maskXY, maskYZ;
    __device__ void operator()(uint64_t w, double* o) const {
        uint64_t n = insertBits(w, topBit);
        uint64_t j = n ^ maskXY;
        cd x = ldd(a, n), y = ldd(a, j);
        cd t1 = cmul(cconj(x), y), t2 = cmul(cconj(y), x);
        double s1 = parity64(j & maskYZ) ? -1.0 : 1.0;
        double s2 = parity64(n & maskYZ) ? -1.0 : 1.0;
        o[0] = s1 * t1.re + s2 * t2.re;
        o[1] = s1 * t1.im + s2 * t2.im;
    }
};
// NT Pauli strings that share their X/Y mask (maskXY), reduced together
// with one read of the state: per term, the same two values (real,
// imaginary) FExpecPauliPair (maskXY != 0) or FExpecPauli (maskXY == 0)
// produce for that term alone, so each term's sum is bitwise the
// single-string result
template <typename T, int NT>
struct FExpecPauliGroupPair {
    const T* a;
    InsertSpec topBit;
    uint64_t maskXY;
    uint64_t maskYZ[NT];
    __device__ void operator()(uint64_t w, double* o) const {
        uint64_t n = insertBits(w, topBit);
        uint64_t j = n ^ maskXY;
        cd x = ldd(a, n), y = ldd(a, j);
        cd t1 = cmul(cconj(x), y), t2 = cmul(cconj(y), x);
#pragma unroll
        for (int t = 0; t < NT; t++) {
            double s1 = parity64(j & maskYZ[t]) ? -1.0 : 1.0;
            double s2 = parity64(n & maskYZ[t]) ? -1.0 : 1.0;
            o[2 * t] = s1 * t1.re + s2 * t2.re;
            o[2 * t + 1] = s1 * t1.im + s2 * t2.im;
        }
    }
};
template <typename T, int NT>
struct FExpecPauliGroupZ {
    const T* a;
    uint64_t maskYZ[NT];
    __device__ void operator()(uint64_t n, double* o) const {
        cd x = ldd(a, n), y = ldd(a, n);
        cd tt = cmul(cconj(x), y);
#pragma unroll
        for (int t = 0; t < NT; t++) {
            double s = parity64(n & maskYZ[t]) ? -1.0 : 1.0;
            o[2 * t] = s * tt.re;
            o[2 * t + 1] = s * tt.im;
        }
    }
};
template <typename T>
struct FDensExpecPauli {  // column n, row i = n ^ maskXY: sign(i & maskYZ) rho_{i,n}
    const T* a;
    uint64_t maskXY, maskYZ;
    int N;
    __device__ void operator()(uint64_t n, double* o) const {
        uint64_t i = n ^ maskXY;
        cd v = ldd(a, i + (n << N));
        double s = parity64(i & maskYZ) ? -1.0 : 1.0;
        o[0] = s * v.re;
        o[1] = s * v.im;
    }
};
template <typename T>
struct FInner {
    const T* a;
    const T* b;
    __device__ void operator()(uint64_t n, double* o) const {
        cd t = cmul(cconj(ldd(a, n)), ldd(b, n));
        o[0] = t.re;
        o[1] = t.im;
    }
};
template <typename T>
struct FHSDist {
    const T* a;
    const T* b;
    __device__ void operator()(uint64_t n, double* o) const {
        cd x = ldd(a, n), y = ldd(b, n);
        o[0] = norm2(cd{x.re - y.re, x.im - y.im});
    }
};
template <typename T>
struct FFidelity {  // cpu_densmatr_calcFidelityWithPureState_sub
    const T* rho;
    const T* psi;
    int N, conj;
    __device__ void operator()(uint64_t n, double* o) const {
        uint64_t r = n & ((1ull << N) - 1), c = n >> N;
        cd rv = ldd(rho, n), pr = ldd(psi, r), pc = ldd(psi, c);
        if (conj) {
            rv = cconj(rv);
            pc = cconj(pc);
        } else {
            pr = cconj(pr);
        }
        cd t = cmul(cmul(rv, pr), pc);
        o[0] = t.re;
        o[1] = t.im;
    }
};
__device__ __forceinline__ cd diagElem(const double* __restrict__ e, uint64_t k, int powMode, double xr,
                                       double xi) {
    double er = e[2 * k], ei = e[2 * k + 1];
    if (powMode == 1) {
        double lr = log(hypot(er, ei)), li = atan2(ei, er);
        double pr = xr * lr - xi * li, pi = xr * li + xi * lr;
        double mag = exp(pr);
        return cd{mag * cos(pi), mag * sin(pi)};
    }
    if (powMode == 2) return cd{pow(er, xr), 0};
    return cd{er, ei};
}
template <typename T>
struct FExpecDiag {
    const T* a;
    const double* e;
    int powMode;
    double xr, xi;
    __device__ void operator()(uint64_t n, double* o) const {
        cd t = cscale(diagElem(e, n, powMode, xr, xi), norm2(ldd(a, n)));
        o[0] = t.re;
        o[1] = t.im;
    }
};
template <typename T>
struct FDensExpecDiag {
    const T* a;
    const double* e;
    int powMode;
    double xr, xi;
    uint64_t diagStride;
    __device__ void operator()(uint64_t n, double* o) const {
        cd t = cmul(diagElem(e, n, powMode, xr, xi), ldd(a, n * diagStride));
        o[0] = t.re;
        o[1] = t.im;
    }
};

// all outcomes of k qubits: register bins for k <= 4 (deterministic),
// global atomics beyond
template <typename T, int K>
__global__ void __launch_bounds__(BLOCK) k_allOutcomesReg(const T* __restrict__ a, uint64_t numTerms,
                                                         TargSpec qs, uint64_t diagStride,
                                                         double* __restrict__ partials) {
    constexpr int B = 1 << K;
    KB acc[B];
#pragma unroll
    for (int b = 0; b < B; b++) acc[b] = KB{0, 0};
    const uint64_t stride = (uint64_t)gridDim.x * BLOCK;
    for (uint64_t n = (uint64_t)blockIdx.x * BLOCK + threadIdx.x; n < numTerms; n += stride) {
        double p = diagStride ? ldd(a, n * diagStride).re : norm2(ldd(a, n));
        uint64_t v = valueOfBits(n, qs);
#pragma unroll
        for (int b = 0; b < B; b++) kadd(acc[b], (v == (uint64_t)b) ? p : 0.0);
    }
    blockMerge<B>(acc);
    if (threadIdx.x == 0)
        for (int b = 0; b < B; b++) {
            partials[(blockIdx.x * B + b) * 2] = acc[b].s;
            partials[(blockIdx.x * B + b) * 2 + 1] = acc[b].c;
        }
}

template <int B>
__global__ void __launch_bounds__(BLOCK) k_finalizeBins(const double* __restrict__ partials, int numParts,
                                                       double* __restrict__ out) {
    // one thread per bin, fixed order over workgroups
    int b = threadIdx.x;
    if (b >= B) return;
    KB acc = {0, 0};
    for (int p = 0; p < numParts; p++) kmerge(acc, KB{partials[(p * B + b) * 2], partials[(p * B + b) * 2 + 1]});
    out[b] = acc.s + acc.c;
}

// outcome tables of more than 4 qubits: thread (b, c) sums, compensated and
// in index order, the terms of outcome b whose other bits r satisfy
// r = c (mod R), so consecutive threads read consecutive amplitudes; the R
// partial sums of each outcome are then merged in order of c.  The order of
// every sum depends on the sizes only, so tables are deterministic.  With
// R = 1 the thread writes its outcome's probability directly.
template <typename T>
__global__ void __launch_bounds__(BLOCK) k_allOutcomesSplit(const T* __restrict__ a, uint64_t restTerms,
                                                            InsertSpec rest, TargSpec qs, uint64_t diagStride,
                                                            int logR, uint64_t numThreads,
                                                            double* __restrict__ dst) {
    const uint64_t t = blockId() * BLOCK + threadIdx.x;
    if (t >= numThreads) return;
    const uint64_t R = 1ull << logR;
    const uint64_t b = t >> logR, c = t & (R - 1);
    uint64_t bits = 0;
    for (int i = 0; i < qs.num; i++) bits |= ((b >> i) & 1ull) << qs.targ[i];
    KB acc = {0, 0};
    for (uint64_t r = c; r < restTerms; r += R) {
        const uint64_t n = insertBits(r, rest) | bits;
        kadd(acc, diagStride ? ldd(a, n * diagStride).re : norm2(ldd(a, n)));
    }
    if (logR == 0) {
        dst[t] = acc.s + acc.c;
    } else {
        dst[2 * t] = acc.s;
        dst[2 * t + 1] = acc.c;
    }
}

__global__ void __launch_bounds__(BLOCK) k_mergeSplitBins(const double* __restrict__ partials, int logR, uint64_t nb,
                                                          double* __restrict__ out) {
    const uint64_t b = (uint64_t)blockIdx.x * BLOCK + threadIdx.x;
    if (b >= nb) return;
    const uint64_t R = 1ull << logR;
    KB acc = {0, 0};
    for (uint64_t c = 0; c < R; c++) kmerge(acc, KB{partials[2 * (b * R + c)], partials[2 * (b * R + c) + 1]});
    out[b] = acc.s + acc.c;
}

// The launch shape of an outcome table of `num` qubits over numTerms terms:
// register bins in `blocks` workgroups for num <= 4, else 2^num << logR
// threads each summing every R-th term of one outcome; `partials` is the
// scratch (doubles) the launches need besides the 2^num results
struct OutcomeShape {
    int blocks = 0;
    int logR = 0;
    int logRest = 0;
    uint64_t threads = 0;
    size_t partials = 0;
};

OutcomeShape outcomeShape(uint64_t numTerms, int num) {
    OutcomeShape s;
    const uint64_t want = (numTerms + (uint64_t)BLOCK * 4 - 1) / ((uint64_t)BLOCK * 4);
    s.blocks = (int)(want < 1 ? 1 : (want > 256 ? 256 : want));
    if (num <= 4) {
        s.partials = (size_t)s.blocks * 16 * 2;
        return s;
    }
    // about 2^20 threads: R partial sums per outcome
    int logTerms = 0;
    while ((1ull << logTerms) < numTerms) logTerms++;
    s.logRest = logTerms - num;
    s.logR = std::max(0, std::min(s.logRest, 20 - num));
    s.threads = (1ull << num) << s.logR;
    s.partials = s.logR ? 2 * (size_t)s.threads : 0;
    return s;
}

// the launches of an outcome table on the library stream: partial sums in
// `scratch` (at least shape.partials doubles, every one written before it
// is read), the 2^num probabilities into the device array `res`
template <typename T>
int launchAllOutcomes(double* res, double* scratch, const OutcomeShape& s, const T* a, uint64_t numTerms,
                      const int* qubits, int num, uint64_t diagStride) {
    TargSpec ts;
    ts.num = num;
    for (int i = 0; i < num; i++) ts.targ[i] = qubits[i];
    const uint64_t nb = 1ull << num;
    if (num <= 4) {
        switch (num) {
#define QA_CASE(K)                                                                                     \
    case K:                                                                                            \
        k_allOutcomesReg<T, K><<<s.blocks, BLOCK, 0, stream()>>>(a, numTerms, ts, diagStride, scratch); \
        k_finalizeBins<1 << K><<<1, BLOCK, 0, stream()>>>(scratch, s.blocks, res);                     \
        break;
            QA_CASE(1)
            QA_CASE(2)
            QA_CASE(3)
            QA_CASE(4)
#undef QA_CASE
        }
        QA_LAUNCH_CHECK();
        return QA_OK;
    }
    // the other bits of a term's index: zeros inserted at the table's qubits
    InsertSpec rest;
    rest.num = num;
    rest.valueMask = 0;
    for (int i = 0; i < num; i++) rest.pos[i] = qubits[i];
    std::sort(rest.pos, rest.pos + num);
    k_allOutcomesSplit<T><<<gridFor(s.threads, 1), BLOCK, 0, stream()>>>(a, 1ull << s.logRest, rest, ts, diagStride,
                                                                        s.logR, s.threads, s.logR ? scratch : res);
    QA_LAUNCH_CHECK();
    if (s.logR) {
        k_mergeSplitBins<<<(unsigned)((nb + BLOCK - 1) / BLOCK), BLOCK, 0, stream()>>>(scratch, s.logR, nb, res);
        QA_LAUNCH_CHECK();
    }
    return QA_OK;
}

template <typename T>
int allOutcomes(double* out, const T* a, uint64_t numTerms, const int* qubits, int num, uint64_t diagStride) {
    const uint64_t nb = 1ull << num;
    const OutcomeShape s = outcomeShape(numTerms, num);
    double* scratch = devScratch(s.partials + (num <= 4 ? 64 : (size_t)nb));
    if (!scratch) return QA_ERR_OOM;
    double* res = scratch + s.partials;
    int rc = launchAllOutcomes<T>(res, scratch, s, a, numTerms, qubits, num, diagStride);
    if (rc != QA_OK) return rc;
    if (num <= 4) {
        double* host = hostScratch(16);
        if (!host) return QA_ERR_OOM;
        QA_HIP(hipMemcpyAsync(host, res, nb * sizeof(double), hipMemcpyDeviceToHost, stream()));
        QA_HIP(hipStreamSynchronize(stream()));
        for (uint64_t b = 0; b < nb; b++) out[b] = host[b];
        return QA_OK;
    }
    QA_HIP(hipMemcpyAsync(out, res, nb * sizeof(double), hipMemcpyDeviceToHost, stream()));
    QA_HIP(hipStreamSynchronize(stream()));
    trimDevScratch();
    return QA_OK;
}

// ---- outcome tables without allocation -----------------------------------
//
// Scratch of qa_statevec_calcProbsOfAllMultiQubitOutcomesNoAlloc: a
// file-scope device array, one copy per device for the life of the
// process.  Its largest use is 2 * 2^20 doubles (num > 4: 2^20 threads,
// one compensated pair each), so every table the allocating entry point
// accepts (1..24 qubits) fits, with the same launch shape and therefore
// the same sums.  Host code reaches it only through the address that
// hipGetSymbolAddress gives, taken once per device by qa_init.
constexpr size_t NOALLOC_SCRATCH_DOUBLES = (size_t)1 << 21;  // 16 MiB
constexpr int NOALLOC_MAX_DEVICES = 64;
__device__ __align__(256) double scratch[NOALLOC_SCRATCH_DOUBLES];
double* g_scratchPtr[NOALLOC_MAX_DEVICES] = {};

}  // namespace

// called by qa_init with `device` current
int initNoAllocScratch(int device) {
    if (device < 0 || device >= NOALLOC_MAX_DEVICES) {
        setError("qa_init: device %d beyond the %d devices of the no-allocation scratch table", device,
                 NOALLOC_MAX_DEVICES);
        return QA_ERR_ARG;
    }
    if (g_scratchPtr[device]) return QA_OK;
    void* p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !p) {
        setError("qa_init: hipGetSymbolAddress of the no-allocation scratch failed: %s", hipGetErrorString(e));
        return QA_ERR_HIP;
    }
    g_scratchPtr[device] = static_cast<double*>(p);
    return QA_OK;
}

namespace {

bool checkArgs(const void* a, int prec, int nq, const char* caller) {
    if (!a || (prec != QA_FP32 && prec != QA_FP64) || nq < 1 || nq > 62) {
        setError("%s: bad qureg arguments", caller);
        return false;
    }
    return true;
}

bool outcomeSpec(InsertSpec& s, int nq, const int* qubits, const int* outcomes, int num, const char* caller) {
    if (num < 0 || num > nq || num > MAXQ) {
        setError("%s: bad number of qubits", caller);
        return false;
    }
    int idx[MAXQ];
    for (int i = 0; i < num; i++) idx[i] = i;
    // insertion sort by qubit
    for (int i = 1; i < num; i++)
        for (int j = i; j > 0 && qubits[idx[j]] < qubits[idx[j - 1]]; j--) {
            int t = idx[j];
            idx[j] = idx[j - 1];
            idx[j - 1] = t;
        }
    s.num = num;
    s.valueMask = 0;
    for (int i = 0; i < num; i++) {
        int q = qubits[idx[i]];
        if (q < 0 || q >= nq || (i && q == s.pos[i - 1])) {
            setError("%s: bad qubit %d", caller, q);
            return false;
        }
        s.pos[i] = q;
        if (outcomes && outcomes[idx[i]]) s.valueMask |= 1ull << q;
    }
    return true;
}

#define QA_DISPATCH(prec, ...)   \
    do {                         \
        if ((prec) == QA_FP64) { \
            typedef double T;    \
            __VA_ARGS__;         \
        } else {                 \
            typedef float T;     \
            __VA_ARGS__;         \
        }                        \
    } while (0)

}  // namespace
}  // namespace qa

using namespace qa;

extern "C" {

int qa_statevec_calcTotalProb(const void* amps, int prec, int nq, double* out) {
    QA_REQUIRE_INIT_RO();
    QA_FLUSH_ALL();
    if (!checkArgs(amps, prec, nq, __func__)) return QA_ERR_ARG;
    int rc;
    QA_DISPATCH(prec, rc = runReduce<1>(FNormAll<T>{(const T*)amps}, 1ull << nq, out));
    return rc;
}

int qa_statevec_calcProbsOfEveryQubit(const void* amps, int prec, int nq, double* out) {
    QA_REQUIRE_INIT_RO();
    QA_FLUSH_ALL();
    if (!checkArgs(amps, prec, nq, __func__)) return QA_ERR_ARG;
    QA_ARG_CHECK(out, "%s: NULL output", __func__);
    if (nq >= 10 && nq <= PQALL) {
        // every qubit in one read (k_probsOfEveryQubit)
        const uint64_t chunks = 1ull << (nq - 10);
        const int blocks = (int)std::min<uint64_t>(RED_BLOCKS, (chunks + BLOCK / 64 - 1) / (BLOCK / 64));
        const int groups = PQALL_VALUES / 8;
        const size_t resOff = (size_t)groups * blocks * 8 * 2;
        double* scratch = devScratch(resOff + PQALL_VALUES);
        double* host = hostScratch(PQALL_VALUES);
        if (!scratch || !host) {
            setError("reduction scratch unavailable");
            return QA_ERR_OOM;
        }
        if (prec == QA_FP64)
            k_probsOfEveryQubit<double><<<blocks, BLOCK, 0, stream()>>>((const double*)amps, nq, chunks, scratch);
        else
            k_probsOfEveryQubit<float><<<blocks, BLOCK, 0, stream()>>>((const float*)amps, nq, chunks, scratch);
        QA_LAUNCH_CHECK();
        for (int g = 0; g < groups; g++)
            k_finalize<8><<<1, BLOCK, 0, stream()>>>(scratch + (size_t)g * blocks * 8 * 2, blocks,
                                                     scratch + resOff + 8 * g);
        QA_LAUNCH_CHECK();
        QA_HIP(hipMemcpyAsync(host, scratch + resOff, PQALL_VALUES * sizeof(double), hipMemcpyDeviceToHost,
                              stream()));
        QA_HIP(hipStreamSynchronize(stream()));
        for (int q = 0; q < nq; q++) {
            out[2 * q] = host[2 * q];
            out[2 * q + 1] = host[2 * q + 1];
        }
        return QA_OK;
    }
    for (int q0 = 0; q0 < nq; q0 += PQ) {
        double v[2 * PQ];
        int rc;
        QA_DISPATCH(prec, rc = runReduceN<2 * PQ>(FProbsOfQubits<T>{(const T*)amps, q0}, 1ull << nq, v));
        if (rc != QA_OK) return rc;
        for (int b = 0; b < PQ && q0 + b < nq; b++) {
            out[2 * (q0 + b)] = v[b];
            out[2 * (q0 + b) + 1] = v[PQ + b];
        }
    }
    return QA_OK;
}

int qa_densmatr_calcTotalProb(const void* amps, int prec, int N, double* out) {
    QA_REQUIRE_INIT_RO();
    QA_FLUSH_ALL();
    if (!checkArgs(amps, prec, 2 * N, __func__)) return QA_ERR_ARG;
    InsertSpec none;
    none.num = 0;
    none.valueMask = 0;
    int rc;
    QA_DISPATCH(prec, rc = runReduce<1>(FDiagReal<T>{(const T*)amps, none, (1ull << N) + 1}, 1ull << N, out));
    return rc;
}

int qa_statevec_calcProbOfMultiQubitOutcome(const void* amps, int prec, int nq, const int* qubits,
                                            const int* outcomes, int num, double* out) {
    QA_REQUIRE_INIT_RO();
    QA_FLUSH_ALL();
    if (!checkArgs(amps, prec, nq, __func__)) return QA_ERR_ARG;
    InsertSpec ins;
    if (!outcomeSpec(ins, nq, qubits, outcomes, num, __func__)) return QA_ERR_ARG;
    int rc;
    QA_DISPATCH(prec, rc = runReduce<1>(FNormOutcome<T>{(const T*)amps, ins}, 1ull << (nq - num), out));
    return rc;
}

int qa_densmatr_calcProbOfMultiQubitOutcome(const void* amps, int prec, int N, const int* qubits,
                                            const int* outcomes, int num, double* out) {
    QA_REQUIRE_INIT_RO();
    QA_FLUSH_ALL();
    if (!checkArgs(amps, prec, 2 * N, __func__)) return QA_ERR_ARG;
    InsertSpec ins;
    if (!outcomeSpec(ins, N, qubits, outcomes, num, __func__)) return QA_ERR_ARG;
    int rc;
    QA_DISPATCH(prec, rc = runReduce<1>(FDiagReal<T>{(const T*)amps, ins, (1ull << N) + 1}, 1ull << (N - num), out));
    return rc;
}

int qa_statevec_calcProbsOfAllMultiQubitOutcomes(double* out, const void* amps, int prec, int nq,
                                                 const int* qubits, int num) {
    QA_REQUIRE_INIT_RO();
    QA_FLUSH_ALL();
    if (!checkArgs(amps, prec, nq, __func__)) return QA_ERR_ARG;
    InsertSpec chk;
    if (!outcomeSpec(chk, nq, qubits, nullptr, num, __func__)) return QA_ERR_ARG;
    QA_ARG_CHECK(num >= 1 && num <= 24, "%s: %d qubits unsupported", __func__, num);
    int rc;
    QA_DISPATCH(prec, rc = allOutcomes<T>(out, (const T*)amps, 1ull << nq, qubits, num, 0));
    return rc;
}

int qa_statevec_calcProbsOfAllMultiQubitOutcomesNoAlloc(double* devOutProbs, const void* amps, int prec, int nq,
                                                        const int* qubits, int num) {
    QA_REQUIRE_INIT_RO();
    if (!checkArgs(amps, prec, nq, __func__)) return QA_ERR_ARG;
    QA_ARG_CHECK(devOutProbs, "%s: NULL output", __func__);
    InsertSpec chk;
    if (!outcomeSpec(chk, nq, qubits, nullptr, num, __func__)) return QA_ERR_ARG;
    QA_ARG_CHECK(num >= 1 && num <= 24, "%s: %d qubits unsupported", __func__, num);
    // a flush could allocate the fusion program ring: the caller flushes first
    QA_ARG_CHECK(!fuse::pending(amps), "%s: gates are queued for this state; call qa_flush(amps) first", __func__);
    const int dev = device();
    double* sp = (dev >= 0 && dev < NOALLOC_MAX_DEVICES) ? g_scratchPtr[dev] : nullptr;
    if (!sp) {
        setError("%s: no scratch address for device %d", __func__, dev);
        return QA_ERR_NOT_INIT;
    }
    const uint64_t numTerms = 1ull << nq;
    const OutcomeShape s = outcomeShape(numTerms, num);
    QA_ARG_CHECK(s.partials <= NOALLOC_SCRATCH_DOUBLES,
                 "%s: a %d-qubit table of a %d-qubit state needs %zu doubles of scratch, more than the %zu of the "
                 "no-allocation scratch", __func__, num, nq, s.partials, NOALLOC_SCRATCH_DOUBLES);
    int rc;
    QA_DISPATCH(prec, rc = launchAllOutcomes<T>(devOutProbs, sp, s, (const T*)amps, numTerms, qubits, num, 0));
    return rc;
}

int qa_densmatr_calcProbsOfAllMultiQubitOutcomes(double* out, const void* amps, int prec, int N,
                                                 const int* qubits, int num) {
    QA_REQUIRE_INIT_RO();
    QA_FLUSH_ALL();
    if (!checkArgs(amps, prec, 2 * N, __func__)) return QA_ERR_ARG;
    InsertSpec chk;
    if (!outcomeSpec(chk, N, qubits, nullptr, num, __func__)) return QA_ERR_ARG;
    QA_ARG_CHECK(num >= 1 && num <= 24, "%s: %d qubits unsupported", __func__, num);
    int rc;
    QA_DISPATCH(prec, rc = allOutcomes<T>(out, (const T*)amps, 1ull << N, qubits, num, (1ull << N) + 1));
    return rc;
}

int qa_statevec_calcInnerProduct(const void* a, const void* b, int prec, int nq, double* out) {
    QA_REQUIRE_INIT_RO();
    QA_FLUSH_ALL();
    if (!checkArgs(a, prec, nq, __func__) || !b) return QA_ERR_ARG;
    int rc;
    QA_DISPATCH(prec, rc = runReduce<2>(FInner<T>{(const T*)a, (const T*)b}, 1ull << nq, out));
    return rc;
}

int qa_densmatr_calcHilbertSchmidtDistance(const void* a, const void* b, int prec, int N, double* out) {
    QA_REQUIRE_INIT_RO();
    QA_FLUSH_ALL();
    if (!checkArgs(a, prec, 2 * N, __func__) || !b) return QA_ERR_ARG;
    int rc;
    QA_DISPATCH(prec, rc = runReduce<1>(FHSDist<T>{(const T*)a, (const T*)b}, 1ull << (2 * N), out));
    return rc;
}

int qa_densmatr_calcFidelityWithPureState(const void* rho, const void* psi, int prec, int N, int conj,
                                          double* out) {
    QA_REQUIRE_INIT_RO();
    QA_FLUSH_ALL();
    if (!checkArgs(rho, prec, 2 * N, __func__) || !psi) return QA_ERR_ARG;
    int rc;
    QA_DISPATCH(prec, rc = runReduce<2>(FFidelity<T>{(const T*)rho, (const T*)psi, N, conj}, 1ull << (2 * N), out));
    return rc;
}

int qa_statevec_calcExpecAnyTargZ(const void* amps, int prec, int nq, uint64_t mask, double* out) {
    QA_REQUIRE_INIT_RO();
    QA_FLUSH_ALL();
    if (!checkArgs(amps, prec, nq, __func__)) return QA_ERR_ARG;
    int rc;
    QA_DISPATCH(prec, rc = runReduce<1>(FExpecZ<T>{(const T*)amps, mask}, 1ull << nq, out));
    return rc;
}

int qa_densmatr_calcExpecAnyTargZ(const void* amps, int prec, int N, uint64_t mask, double* out) {
    QA_REQUIRE_INIT_RO();
    QA_FLUSH_ALL();
    if (!checkArgs(amps, prec, 2 * N, __func__)) return QA_ERR_ARG;
    int rc;
    QA_DISPATCH(prec, rc = runReduce<2>(FDensExpecZ<T>{(const T*)amps, mask, (1ull << N) + 1}, 1ull << N, out));
    return rc;
}

static void timesPowerOfI(double* v, int k) {
    for (int i = 0; i < (k & 3); i++) {
        double re = v[0], im = v[1];
        v[0] = -im;
        v[1] = re;
    }
}

int qa_statevec_calcExpecPauliStr(const void* amps, int prec, int nq, uint64_t maskX, uint64_t maskY,
                                  uint64_t maskZ, double* out) {
    QA_REQUIRE_INIT_RO();
    QA_FLUSH_ALL();
    if (!checkArgs(amps, prec, nq, __func__)) return QA_ERR_ARG;
    int rc;
    uint64_t mxy = maskX | maskY, myz = maskY | maskZ;
    if (mxy) {
        InsertSpec top;
        top.num = 1;
        top.pos[0] = 63 - __builtin_clzll(mxy);
        top.valueMask = 0;
        QA_DISPATCH(prec, rc = runReduce<2>(FExpecPauliPair<T>{(const T*)amps, top, mxy, myz}, 1ull << (nq - 1), out));
    } else {
        QA_DISPATCH(prec, rc = runReduce<2>(FExpecPauli<T>{(const T*)amps, mxy, myz}, 1ull << nq, out));
    }
    if (rc == QA_OK) timesPowerOfI(out, __builtin_popcountll(maskY));
    return rc;
}

}  // extern "C"

namespace qa {
namespace {

// one launch for up to 8 strings sharing maskXY: out[2t], out[2t+1] = the
// sum of string t before its i^#Y factor
template <typename T, int NT>
int expecGroup(const T* a, int nq, uint64_t mxy, const uint64_t* myz, double* out) {
    if (mxy) {
        FExpecPauliGroupPair<T, NT> f;
        f.a = a;
        f.topBit.num = 1;
        f.topBit.pos[0] = 63 - __builtin_clzll(mxy);
        f.topBit.valueMask = 0;
        f.maskXY = mxy;
        for (int t = 0; t < NT; t++) f.maskYZ[t] = myz[t];
        return runReduceN<2 * NT>(f, 1ull << (nq - 1), out);
    }
    FExpecPauliGroupZ<T, NT> f;
    f.a = a;
    for (int t = 0; t < NT; t++) f.maskYZ[t] = myz[t];
    return runReduceN<2 * NT>(f, 1ull << nq, out);
}

template <typename T>
int expecGroupAny(const T* a, int nq, uint64_t mxy, const uint64_t* myz, int nt, double* out) {
    switch (nt) {
        case 1: return expecGroup<T, 1>(a, nq, mxy, myz, out);
        case 2: return expecGroup<T, 2>(a, nq, mxy, myz, out);
        case 3: return expecGroup<T, 3>(a, nq, mxy, myz, out);
        case 4: return expecGroup<T, 4>(a, nq, mxy, myz, out);
        case 5: return expecGroup<T, 5>(a, nq, mxy, myz, out);
        case 6: return expecGroup<T, 6>(a, nq, mxy, myz, out);
        case 7: return expecGroup<T, 7>(a, nq, mxy, myz, out);
        default: return expecGroup<T, 8>(a, nq, mxy, myz, out);
    }
}

}  // namespace
}  // namespace qa

extern "C" {

int qa_statevec_calcExpecPauliStrSum(const void* amps, int prec, int nq, int64_t numTerms, const uint64_t* maskX,
                                     const uint64_t* maskY, const uint64_t* maskZ, const double* coeffs,
                                     double* termValues, double* out) {
    QA_REQUIRE_INIT_RO();
    QA_FLUSH_ALL();
    if (!checkArgs(amps, prec, nq, __func__)) return QA_ERR_ARG;
    QA_ARG_CHECK(numTerms >= 0 && (numTerms == 0 || (maskX && maskY && maskZ && coeffs)),
                 "%s: bad term arrays", __func__);
    const uint64_t all = nq >= 64 ? ~0ull : ((1ull << nq) - 1ull);
    for (int64_t t = 0; t < numTerms; t++)
        QA_ARG_CHECK(((maskX[t] | maskY[t] | maskZ[t]) & ~all) == 0 && !(maskX[t] & maskY[t]) &&
                         !(maskX[t] & maskZ[t]) && !(maskY[t] & maskZ[t]),
                     "%s: term %lld is not a Pauli string on the qureg", __func__, (long long)t);
    // terms grouped by X/Y mask (groups in ascending mask order, each in term
    // order); one launch per 8 terms of a group
    std::vector<int64_t> order(numTerms);
    for (int64_t t = 0; t < numTerms; t++) order[t] = t;
    std::stable_sort(order.begin(), order.end(), [&](int64_t x, int64_t y) {
        return (maskX[x] | maskY[x]) < (maskX[y] | maskY[y]);
    });
    std::vector<double> vals(2 * (size_t)numTerms);
    int rc = QA_OK;
    for (int64_t g = 0; g < numTerms && rc == QA_OK;) {
        const uint64_t mxy = maskX[order[g]] | maskY[order[g]];
        int64_t e = g;
        while (e < numTerms && (maskX[order[e]] | maskY[order[e]]) == mxy && e - g < 8) e++;
        const int nt = (int)(e - g);
        uint64_t myz[8];
        double res[16];
        for (int t = 0; t < nt; t++) myz[t] = maskY[order[g + t]] | maskZ[order[g + t]];
        QA_DISPATCH(prec, rc = expecGroupAny<T>((const T*)amps, nq, mxy, myz, nt, res));
        for (int t = 0; t < nt && rc == QA_OK; t++) {
            const int64_t term = order[g + t];
            double v[2] = {res[2 * t], res[2 * t + 1]};
            timesPowerOfI(v, __builtin_popcountll(maskY[term]));
            vals[2 * term] = v[0];
            vals[2 * term + 1] = v[1];
        }
        g = e;
    }
    if (rc != QA_OK) return rc;
    // sum of coefficient x value, in term order
    double re = 0, im = 0;
    for (int64_t t = 0; t < numTerms; t++) {
        const double cr = coeffs[2 * t], ci = coeffs[2 * t + 1], vr = vals[2 * t], vi = vals[2 * t + 1];
        re += cr * vr - ci * vi;
        im += cr * vi + ci * vr;
    }
    if (termValues)
        for (int64_t t = 0; t < 2 * numTerms; t++) termValues[t] = vals[t];
    out[0] = re;
    out[1] = im;
    return QA_OK;
}

int qa_densmatr_calcExpecPauliStr(const void* amps, int prec, int N, uint64_t maskX, uint64_t maskY,
                                  uint64_t maskZ, double* out) {
    QA_REQUIRE_INIT_RO();
    QA_FLUSH_ALL();
    if (!checkArgs(amps, prec, 2 * N, __func__)) return QA_ERR_ARG;
    int rc;
    QA_DISPATCH(prec, rc = runReduce<2>(FDensExpecPauli<T>{(const T*)amps, maskX | maskY, maskY | maskZ, N},
                                        1ull << N, out));
    if (rc == QA_OK) timesPowerOfI(out, __builtin_popcountll(maskY));
    return rc;
}

int qa_statevec_calcExpecFullStateDiagMatr(const void* amps, int prec, int nq, const void* e, int powMode,
                                           double xr, double xi, double* out) {
    QA_REQUIRE_INIT_RO();
    QA_FLUSH_ALL();
    if (!checkArgs(amps, prec, nq, __func__) || !e) return QA_ERR_ARG;
    int rc;
    QA_DISPATCH(prec, rc = runReduce<2>(FExpecDiag<T>{(const T*)amps, (const double*)e, powMode, xr, xi},
                                        1ull << nq, out));
    return rc;
}

int qa_densmatr_calcExpecFullStateDiagMatr(const void* amps, int prec, int N, const void* e, int powMode,
                                           double xr, double xi, double* out) {
    QA_REQUIRE_INIT_RO();
    QA_FLUSH_ALL();
    if (!checkArgs(amps, prec, 2 * N, __func__) || !e) return QA_ERR_ARG;
    int rc;
    QA_DISPATCH(prec, rc = runReduce<2>(FDensExpecDiag<T>{(const T*)amps, (const double*)e, powMode, xr, xi,
                                                          (1ull << N) + 1},
                                        1ull << N, out));
    return rc;
}

}  // extern "C"
