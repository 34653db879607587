// State initialisation and projector kernels, replacing
// quest/src/gpu/gpu_thrust.cuh:1015-1084 (projectors, uniform/debug/random
// amplitudes) with plain streaming kernels.
#include <cmath>

#include <vector>

#include "qa_common.hpp"
#include "qa_fused.hpp"

namespace qa {
namespace {

template <typename T>
__global__ void __launch_bounds__(BLOCK) k_fill(T* __restrict__ amps, uint64_t num, T re, T im) {
    const uint64_t base = blockId() * (BLOCK * ITEMS) + threadIdx.x;
#pragma unroll
    for (int j = 0; j < ITEMS; j++) {
        uint64_t i = base + (uint64_t)j * BLOCK;
        if (i < num) stnt(amps, i, ct<T>{re, im});
    }
}

// a_i = (2i/10, (2i+1)/10)  (cpu_subroutines.cpp: cpu_statevec_initDebugState_sub)
template <typename T>
__global__ void __launch_bounds__(BLOCK) k_debug(T* __restrict__ amps, uint64_t num) {
    const uint64_t i = blockId() * BLOCK + threadIdx.x;
    if (i >= num) return;
    double re = (double)(2 * i) / 10.;
    double im = (double)(2 * i + 1) / 10.;
    st(amps, i, ct<T>{(T)re, (T)im});
}

// SplitMix64 finaliser (Steele, Lea & Flood 2014) as a counter-based hash
__device__ __forceinline__ uint64_t mix64(uint64_t z) {
    z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ull;
    z = (z ^ (z >> 27)) * 0x94d049bb133111ebull;
    return z ^ (z >> 31);
}
__device__ __forceinline__ double unit(uint64_t x) {  // (0, 1]
    return ((double)(x >> 11) + 1.0) * (1.0 / 9007199254740992.0);
}

// amplitude = sqrt(n1^2 + n2^2) e^{i phi}, n1,n2 ~ N(0,1) (Box-Muller), phi ~ U(0,2pi):
// the distribution of randomiser.cpp rand_getThreadPrivateRandomAmp
template <typename T>
__global__ void __launch_bounds__(BLOCK) k_random(T* __restrict__ amps, uint64_t num, uint64_t seed) {
    const uint64_t i = blockId() * BLOCK + threadIdx.x;
    if (i >= num) return;
    uint64_t k = seed * 0x9e3779b97f4a7c15ull + i * 0xd1b54a32d192ed03ull;
    double u1 = unit(mix64(k ^ 0x1ull)), u3 = unit(mix64(k ^ 0x3ull));
    // n1^2 + n2^2 = -2 ln u1 for a Box-Muller pair, whatever its angle
    double mag = sqrt(-2.0 * log(u1));
    double phi = 2.0 * M_PI * u3;
    st(amps, i, ct<T>{(T)(mag * cos(phi)), (T)(mag * sin(phi))});
}

// keep amplitudes whose qubits equal the outcome bits, scaled by renorm
template <typename T>
__global__ void __launch_bounds__(BLOCK) k_projSV(T* __restrict__ amps, uint64_t num, uint64_t qubitMask,
                                                 uint64_t outcomeMask, T renorm) {
    const uint64_t i = blockId() * BLOCK + threadIdx.x;
    if (i >= num) return;
    T f = ((i & qubitMask) == outcomeMask) ? renorm : (T)0;
    ct<T> a = ld(amps, i);
    st(amps, i, ct<T>{a.re * f, a.im * f});
}

// rho_rc *= renorm [bits(r) == bits(c) == outcome]  (cpu_densmatr_multiQubitProjector_sub)
template <typename T>
__global__ void __launch_bounds__(BLOCK) k_projDM(T* __restrict__ amps, uint64_t num, int N, uint64_t qubitMask,
                                                 uint64_t outcomeMask, T renorm) {
    const uint64_t i = blockId() * BLOCK + threadIdx.x;
    if (i >= num) return;
    uint64_t r = i & ((1ull << N) - 1), c = i >> N;
    bool keep = ((r & qubitMask) == outcomeMask) && ((c & qubitMask) == outcomeMask);
    T f = keep ? renorm : (T)0;
    ct<T> a = ld(amps, i);
    st(amps, i, ct<T>{a.re * f, a.im * f});
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

bool ok(const void* a, int prec, int nq, const char* caller) {
    if (!a || (prec != QA_FP32 && prec != QA_FP64) || nq < 1 || nq > 62) {
        setError("%s: bad qureg arguments", caller);
        return false;
    }
    return true;
}

bool masks(int nq, const int* qubits, const int* outcomes, int num, uint64_t& qm, uint64_t& om,
           const char* caller) {
    qm = om = 0;
    for (int i = 0; i < num; i++) {
        int q = qubits[i];
        if (q < 0 || q >= nq || ((qm >> q) & 1ull) || (outcomes[i] != 0 && outcomes[i] != 1)) {
            setError("%s: bad qubit/outcome", caller);
            return false;
        }
        qm |= 1ull << q;
        if (outcomes[i]) om |= 1ull << q;
    }
    return true;
}

// One term of a PauliStrSum as the kernels read it: its X/Y mask, Y mask,
// Z mask and complex coefficient.
struct PsTerm {
    uint64_t mxy, my, mz;
    double re, im;
};

// <r| sum_j c_j sigma_j |c> (fastmath.hpp:124-196, cpu_subroutines.cpp:63-92):
// sigma_j's element is 0 unless r ^ c is its X/Y mask, else the product of
// its one-qubit elements -- i^(#Y), times -1 per Y with row bit 0 and per Z
// with row bit 1 -- exactly +-1 or +-i; each term adds c_j times it, in term
// order, with the complex product written out as the reference's is
__device__ __forceinline__ cd pauliSumElem(const PsTerm* __restrict__ terms, int64_t numTerms, uint64_t r,
                                           uint64_t c) {
    cd e = cd{0, 0};
    for (int64_t j = 0; j < numTerms; j++) {
        const PsTerm t = terms[j];
        if ((r ^ c) != t.mxy) continue;
        const int ny = __popcll(t.my);
        const int neg = (__popcll(t.my & ~r) + __popcll(t.mz & r) + (ny & 2 ? 1 : 0)) & 1;
        double x = (ny & 1) ? 0.0 : 1.0, y = (ny & 1) ? 1.0 : 0.0;  // i^(#Y) up to its sign
        if (neg) {
            x = -x;
            y = -y;
        }
        e.re = e.re + (t.re * x - t.im * y);
        e.im = e.im + (t.re * y + t.im * x);
    }
    return e;
}

// rho (column-major, index = row + col 2^N) <- the matrix of the sum
template <typename T>
__global__ void __launch_bounds__(BLOCK) k_densPauliSum(T* __restrict__ amps, int N, const PsTerm* __restrict__ terms,
                                                       int64_t numTerms) {
    const uint64_t i = blockId() * BLOCK + threadIdx.x;
    if (i >= (1ull << (2 * N))) return;
    const uint64_t r = i & ((1ull << N) - 1ull), c = i >> N;
    const cd e = pauliSumElem(terms, numTerms, r, c);
    st(amps, i, ct<T>{(T)e.re, (T)e.im});
}

// diagonal elements (complex128) of the sum
__global__ void __launch_bounds__(BLOCK) k_diagPauliSum(double* __restrict__ elems, int n,
                                                       const PsTerm* __restrict__ terms, int64_t numTerms) {
    const uint64_t i = blockId() * BLOCK + threadIdx.x;
    if (i >= (1ull << n)) return;
    const cd e = pauliSumElem(terms, numTerms, i, i);
    elems[2 * i] = e.re;
    elems[2 * i + 1] = e.im;
}

// the terms in device scratch (PsTerm records), or nullptr
const PsTerm* uploadTerms(int64_t numTerms, const uint64_t* maskX, const uint64_t* maskY, const uint64_t* maskZ,
                          const double* coeffs) {
    static_assert(sizeof(PsTerm) == 5 * sizeof(double), "terms pack as five 8-byte words");
    double* d = devScratch((size_t)numTerms * 5);
    if (!d) return nullptr;
    std::vector<PsTerm> h((size_t)numTerms);
    for (int64_t j = 0; j < numTerms; j++)
        h[j] = PsTerm{maskX[j] | maskY[j], maskY[j], maskZ[j], coeffs[2 * j], coeffs[2 * j + 1]};
    if (hipMemcpyAsync(d, h.data(), h.size() * sizeof(PsTerm), hipMemcpyHostToDevice, stream()) != hipSuccess)
        return nullptr;
    // the host array must outlive the copy
    if (hipStreamSynchronize(stream()) != hipSuccess) return nullptr;
    return reinterpret_cast<const PsTerm*>(d);
}

}  // namespace
}  // namespace qa

using namespace qa;

extern "C" {

int qa_statevec_initUniformState(void* amps, int prec, int nq, double re, double im) {
    QA_REQUIRE_INIT();
    QA_FLUSH_ALL();
    if (!ok(amps, prec, nq, __func__)) return QA_ERR_ARG;
    uint64_t n = 1ull << nq;
    QA_DISPATCH(prec, (k_fill<T><<<gridFor(n), BLOCK, 0, stream()>>>((T*)amps, n, (T)re, (T)im)));
    QA_LAUNCH_CHECK();
    return QA_OK;
}

int qa_statevec_initDebugState(void* amps, int prec, int nq) {
    QA_REQUIRE_INIT();
    QA_FLUSH_ALL();
    if (!ok(amps, prec, nq, __func__)) return QA_ERR_ARG;
    uint64_t n = 1ull << nq;
    QA_DISPATCH(prec, (k_debug<T><<<gridFor(n, 1), BLOCK, 0, stream()>>>((T*)amps, n)));
    QA_LAUNCH_CHECK();
    return QA_OK;
}

int qa_statevec_initUnnormalisedUniformlyRandomPureStateAmps(void* amps, int prec, int nq, uint64_t seed) {
    QA_REQUIRE_INIT();
    QA_FLUSH_ALL();
    if (!ok(amps, prec, nq, __func__)) return QA_ERR_ARG;
    uint64_t n = 1ull << nq;
    QA_DISPATCH(prec, (k_random<T><<<gridFor(n, 1), BLOCK, 0, stream()>>>((T*)amps, n, seed)));
    QA_LAUNCH_CHECK();
    return QA_OK;
}

int qa_statevec_multiQubitProjector(void* amps, int prec, int nq, const int* qubits, const int* outcomes, int num,
                                    double prob) {
    QA_REQUIRE_INIT();
    QA_FLUSH_ALL();
    if (!ok(amps, prec, nq, __func__)) return QA_ERR_ARG;
    uint64_t qm, om;
    if (!masks(nq, qubits, outcomes, num, qm, om, __func__)) return QA_ERR_ARG;
    uint64_t n = 1ull << nq;
    double renorm = 1 / std::sqrt(prob);
    QA_DISPATCH(prec, (k_projSV<T><<<gridFor(n, 1), BLOCK, 0, stream()>>>((T*)amps, n, qm, om, (T)renorm)));
    QA_LAUNCH_CHECK();
    return QA_OK;
}

int qa_densmatr_multiQubitProjector(void* amps, int prec, int N, const int* qubits, const int* outcomes, int num,
                                    double prob) {
    QA_REQUIRE_INIT();
    QA_FLUSH_ALL();
    if (!ok(amps, prec, 2 * N, __func__)) return QA_ERR_ARG;
    uint64_t qm, om;
    if (!masks(N, qubits, outcomes, num, qm, om, __func__)) return QA_ERR_ARG;
    uint64_t n = 1ull << (2 * N);
    double renorm = 1 / prob;
    QA_DISPATCH(prec, (k_projDM<T><<<gridFor(n, 1), BLOCK, 0, stream()>>>((T*)amps, n, N, qm, om, (T)renorm)));
    QA_LAUNCH_CHECK();
    return QA_OK;
}

int qa_densmatr_setAmpsToPauliStrSum(void* amps, int prec, int N, int64_t numTerms, const uint64_t* maskX,
                                     const uint64_t* maskY, const uint64_t* maskZ, const double* coeffs) {
    QA_REQUIRE_INIT();
    QA_FLUSH_ALL();
    if (!ok(amps, prec, 2 * N, __func__)) return QA_ERR_ARG;
    QA_ARG_CHECK(numTerms >= 1 && maskX && maskY && maskZ && coeffs, "%s: no terms", __func__);
    for (int64_t j = 0; j < numTerms; j++)
        QA_ARG_CHECK(((maskX[j] | maskY[j] | maskZ[j]) >> N) == 0 && !(maskX[j] & maskY[j]) &&
                         !(maskX[j] & maskZ[j]) && !(maskY[j] & maskZ[j]),
                     "%s: term %lld has bad masks", __func__, (long long)j);
    const PsTerm* t = uploadTerms(numTerms, maskX, maskY, maskZ, coeffs);
    if (!t) {
        setError("%s: term upload failed", __func__);
        return QA_ERR_HIP;
    }
    const uint64_t n = 1ull << (2 * N);
    QA_DISPATCH(prec, (k_densPauliSum<T><<<gridFor(n, 1), BLOCK, 0, stream()>>>((T*)amps, N, t, numTerms)));
    QA_LAUNCH_CHECK();
    return QA_OK;
}

int qa_fullstatediagmatr_setElemsToPauliStrSum(void* devElems, int numQubits, int64_t numTerms, const uint64_t* maskZ,
                                               const double* coeffs) {
    QA_REQUIRE_INIT();
    QA_ARG_CHECK(devElems && numQubits >= 1 && numQubits <= 50, "%s: bad matrix", __func__);
    QA_ARG_CHECK(numTerms >= 1 && maskZ && coeffs, "%s: no terms", __func__);
    for (int64_t j = 0; j < numTerms; j++)
        QA_ARG_CHECK((maskZ[j] >> numQubits) == 0, "%s: term %lld targets a qubit beyond the matrix", __func__,
                     (long long)j);
    std::vector<uint64_t> zero((size_t)numTerms, 0);
    const PsTerm* t = uploadTerms(numTerms, zero.data(), zero.data(), maskZ, coeffs);
    if (!t) {
        setError("%s: term upload failed", __func__);
        return QA_ERR_HIP;
    }
    const uint64_t n = 1ull << numQubits;
    k_diagPauliSum<<<gridFor(n, 1), BLOCK, 0, stream()>>>((double*)devElems, numQubits, t, numTerms);
    QA_LAUNCH_CHECK();
    return QA_OK;
}

}  // extern "C"
