// Statevector gate kernels: any-control dense 1/2/k-target matrices,
// diagonal 1/2/k-target matrices, swaps, Pauli tensors/gadgets and Z/phase
// gadgets.  Each is one streaming pass over the control-satisfied amplitudes
// (HBM-bound), replacing quest/src/gpu/gpu_kernels.cuh:134-691 and their
// launchers in quest/src/gpu/gpu_subroutines.cpp:197-887.
//
// Launch shape: 256-thread workgroups (4 wave64s); every thread owns ITEMS
// work items strided by the workgroup size, issues all of their loads before
// any arithmetic, and writes back with non-temporal stores, so each wave
// instruction moves whole 64-amplitude runs while several independent
// requests per lane are in flight.  Density matrices use these kernels on
// their 2N-bit index (ket qubit q = bit q, bra qubit q = bit q+N).
#include <algorithm>
#include <cmath>
#include <vector>

#include "qa_common.hpp"
#include "qa_fused.hpp"
#include "qa_jit.hpp"

namespace qa {
namespace {

// ---- host-side argument preparation -----------------------------------

// Build the ascending insertion list for (ctrls with states) plus extra
// qubits with fixed values.  Returns false (and sets the error) on a
// repeated or out-of-range qubit.
bool buildInsert(InsertSpec& s, int numQubits, const int* ctrls, const int* ctrlStates, int numCtrls,
                 const int* extra, const int* extraVals, int numExtra, const char* caller) {
    std::vector<std::pair<int, int>> v;
    v.reserve((size_t)(numCtrls + numExtra));
    for (int i = 0; i < numCtrls; i++) v.push_back({ctrls[i], ctrlStates ? (ctrlStates[i] != 0) : 1});
    for (int i = 0; i < numExtra; i++) v.push_back({extra[i], extraVals ? extraVals[i] : 0});
    if ((int)v.size() > MAXQ || (int)v.size() > numQubits) {
        setError("%s: %d qubits exceed the %d-qubit register", caller, (int)v.size(), numQubits);
        return false;
    }
    std::sort(v.begin(), v.end());
    s.num = (int)v.size();
    s.valueMask = 0;
    for (int i = 0; i < s.num; i++) {
        int q = v[(size_t)i].first;
        if (q < 0 || q >= numQubits) {
            setError("%s: qubit %d outside [0,%d)", caller, q, numQubits);
            return false;
        }
        if (i && v[(size_t)i - 1].first == q) {
            setError("%s: qubit %d appears twice", caller, q);
            return false;
        }
        s.pos[i] = q;
        if (v[(size_t)i].second) s.valueMask |= 1ull << q;
    }
    return true;
}

bool checkQureg(const void* amps, int prec, int numQubits, const char* caller) {
    if (!amps) {
        setError("%s: null amplitude pointer", caller);
        return false;
    }
    if (prec != QA_FP32 && prec != QA_FP64) {
        setError("%s: precision %d is neither QA_FP32 nor QA_FP64", caller, prec);
        return false;
    }
    if (numQubits < 1 || numQubits > 62) {
        setError("%s: %d index qubits unsupported", caller, numQubits);
        return false;
    }
    return true;
}

template <typename T>
struct M1 {
    ct<T> m00, m01, m10, m11;
};
template <typename T>
M1<T> toM1(const double* m) {
    return M1<T>{{(T)m[0], (T)m[1]}, {(T)m[2], (T)m[3]}, {(T)m[4], (T)m[5]}, {(T)m[6], (T)m[7]}};
}
template <typename T>
struct M2 {
    ct<T> e[16];
};

// ---- kernels ------------------------------------------------------------

// [a0; a1] <- M [a0; a1] on pairs (i0, i0 | 1<<targ) with ctrls satisfied
// IT items per thread; NT = streaming (non-temporal) accesses.  The host
// picks <2, 1> when every touched index bit is >= 3 and <1, 0> otherwise
// (see ldx/stx).
template <typename T, int IT, int NT>
__global__ void __launch_bounds__(BLOCK) k_oneTargDense(T* __restrict__ amps, uint64_t numWork,
                                                       InsertSpec ins, int targ, M1<T> m) {
    const uint64_t base = blockId() * (BLOCK * IT) + threadIdx.x;
    const uint64_t tbit = 1ull << targ;
    uint64_t i0[IT];
    ct<T> a0[IT], a1[IT];
#pragma unroll
    for (int j = 0; j < IT; j++) {
        uint64_t n = base + (uint64_t)j * BLOCK;
        if (n < numWork) {
            i0[j] = insertBits(n, ins);
            a0[j] = ldx<NT>(amps, i0[j]);
            a1[j] = ldx<NT>(amps, i0[j] | tbit);
        }
    }
#pragma unroll
    for (int j = 0; j < IT; j++) {
        uint64_t n = base + (uint64_t)j * BLOCK;
        if (n < numWork) {
            cmat2T(m.m00, m.m01, m.m10, m.m11, a0[j], a1[j]);
            stx<NT>(amps, i0[j], a0[j]);
            stx<NT>(amps, i0[j] | tbit, a1[j]);
        }
    }
}

// 4-amplitude dense 4x4; matrix index bit0 <-> targ1, bit1 <-> targ2
template <typename T>
__global__ void __launch_bounds__(BLOCK) k_twoTargDense(T* __restrict__ amps, uint64_t numWork,
                                                       InsertSpec ins, int targ1, int targ2, M2<T> m) {
    const uint64_t base = blockId() * (BLOCK * 2) + threadIdx.x;
    const uint64_t b1 = 1ull << targ1, b2 = 1ull << targ2;
    const uint64_t off[4] = {0, b1, b2, b1 | b2};
#pragma unroll
    for (int j = 0; j < 2; j++) {
        uint64_t n = base + (uint64_t)j * BLOCK;
        if (n >= numWork) continue;
        uint64_t i0 = insertBits(n, ins);
        ct<T> a[4];
#pragma unroll
        for (int k = 0; k < 4; k++) a[k] = ldnt(amps, i0 | off[k]);
        cmat4T<T>([&](int e) { return m.e[e]; }, a[0], a[1], a[2], a[3]);
#pragma unroll
        for (int r = 0; r < 4; r++) stnt(amps, i0 | off[r], a[r]);
    }
}

// 2^K-amplitude dense, K = 3..5, matrix (FP64, row-major) read with uniform
// (scalar) loads; offsets[v] = index bits of value v over the unsorted targs
template <typename T, int K>
__global__ void __launch_bounds__(BLOCK) k_fewTargDense(T* __restrict__ amps, uint64_t numWork,
                                                       InsertSpec ins, TargSpec ts,
                                                       const double* __restrict__ mat, int conj) {
    constexpr int D = 1 << K;
    const uint64_t n = blockId() * BLOCK + threadIdx.x;
    if (n >= numWork) return;
    const uint64_t i0 = insertBits(n, ins);
    uint64_t off[D];
#pragma unroll
    for (int v = 0; v < D; v++) {
        uint64_t o = 0;
#pragma unroll
        for (int b = 0; b < K; b++)
            if ((v >> b) & 1) o |= 1ull << ts.targ[b];
        off[v] = o;
    }
    ct<T> a[D];
#pragma unroll
    for (int v = 0; v < D; v++) a[v] = ldnt(amps, i0 | off[v]);
    const double sgn = conj ? -1.0 : 1.0;
    for (int r = 0; r < D; r++) {
        ct<T> acc = {0, 0};
#pragma unroll
        for (int c = 0; c < D; c++) {
            ct<T> e = {(T)mat[2 * (r * D + c)], (T)(sgn * mat[2 * (r * D + c) + 1])};
            acc = cfmaT(e, a[c], acc);
        }
        stnt(amps, i0 | off[r], acc);
    }
}

// 2^K-amplitude dense for K >= 6: one workgroup per group of amplitudes,
// the group staged through LDS, rows split across the workgroup's threads
template <typename T>
__global__ void __launch_bounds__(BLOCK) k_manyTargDense(T* __restrict__ amps, uint64_t numWork,
                                                        InsertSpec ins, TargSpec ts,
                                                        const double* __restrict__ mat, int conj) {
    extern __shared__ unsigned char smem_raw[];
    ct<T>* cache = reinterpret_cast<ct<T>*>(smem_raw);
    const uint64_t D = 1ull << ts.num;
    const double sgn = conj ? -1.0 : 1.0;
    for (uint64_t n = blockIdx.x; n < numWork; n += gridDim.x) {
        const uint64_t i0 = insertBits(n, ins);
        for (uint64_t v = threadIdx.x; v < D; v += BLOCK) cache[v] = ld(amps, setBitsFrom(i0, ts, v));
        __syncthreads();
        for (uint64_t r = threadIdx.x; r < D; r += BLOCK) {
            ct<T> acc = {0, 0};
            const double* row = mat + 2 * r * D;
            for (uint64_t c = 0; c < D; c++) {
                ct<T> e = {(T)row[2 * c], (T)(sgn * row[2 * c + 1])};
                acc = cfmaT(e, cache[c], acc);
            }
            st(amps, setBitsFrom(i0, ts, r), acc);
        }
        __syncthreads();
    }
}

// a <- a * d[bit targ] over control-satisfied amplitudes (targ < 0: the
// target bit was folded into the insertion spec and every amplitude takes d1)
template <typename T, int IT, int NT>
__global__ void __launch_bounds__(BLOCK) k_oneTargDiag(T* __restrict__ amps, uint64_t numWork,
                                                      InsertSpec ins, int targ, ct<T> d0, ct<T> d1) {
    const uint64_t base = blockId() * (BLOCK * IT) + threadIdx.x;
    uint64_t idx[IT];
    ct<T> a[IT];
#pragma unroll
    for (int j = 0; j < IT; j++) {
        uint64_t n = base + (uint64_t)j * BLOCK;
        if (n < numWork) {
            idx[j] = insertBits(n, ins);
            a[j] = ldx<NT>(amps, idx[j]);
        }
    }
#pragma unroll
    for (int j = 0; j < IT; j++) {
        uint64_t n = base + (uint64_t)j * BLOCK;
        if (n < numWork) {
            ct<T> d = (targ < 0 || ((idx[j] >> targ) & 1ull)) ? d1 : d0;
            stx<NT>(amps, idx[j], cmulT(d, a[j]));
        }
    }
}

template <typename T>
__global__ void __launch_bounds__(BLOCK) k_twoTargDiag(T* __restrict__ amps, uint64_t numWork,
                                                      InsertSpec ins, int targ1, int targ2, ct<T> d0,
                                                      ct<T> d1, ct<T> d2, ct<T> d3) {
    const uint64_t base = blockId() * (BLOCK * ITEMS) + threadIdx.x;
#pragma unroll
    for (int j = 0; j < ITEMS; j++) {
        uint64_t n = base + (uint64_t)j * BLOCK;
        if (n >= numWork) continue;
        uint64_t i = insertBits(n, ins);
        int k = (int)(((i >> targ2) & 1ull) << 1 | ((i >> targ1) & 1ull));
        ct<T> d = (k == 0) ? d0 : (k == 1) ? d1 : (k == 2) ? d2 : d3;
        stnt(amps, i, cmulT(d, ldnt(amps, i)));
    }
}

// general diagonal: element index from the (unsorted) target bits, optional
// complex power (principal branch, as std::pow) and conjugation
template <typename T>
__global__ void __launch_bounds__(BLOCK) k_anyTargDiag(T* __restrict__ amps, uint64_t numWork,
                                                      InsertSpec ins, TargSpec ts,
                                                      const double* __restrict__ elems, int conj,
                                                      int hasPower, double expRe, double expIm) {
    const uint64_t n = blockId() * BLOCK + threadIdx.x;
    if (n >= numWork) return;
    uint64_t i = insertBits(n, ins);
    uint64_t t = valueOfBits(i, ts);
    double er = elems[2 * t], ei = elems[2 * t + 1];
    if (hasPower) {
        // z^w = exp(w log z)
        double lr = log(hypot(er, ei)), li = atan2(ei, er);
        double pr = expRe * lr - expIm * li, pi = expRe * li + expIm * lr;
        double mag = exp(pr);
        er = mag * cos(pi);
        ei = mag * sin(pi);
    }
    if (conj) ei = -ei;
    ct<T> d = {(T)er, (T)ei};
    st(amps, i, cmulT(d, ld(amps, i)));
}

template <typename T>
__global__ void __launch_bounds__(BLOCK) k_allTargDiag(T* __restrict__ amps, uint64_t numAmps,
                                                      const double* __restrict__ elems, int hasPower,
                                                      double expRe, double expIm, int densN,
                                                      int multiplyOnly) {
    const uint64_t i = blockId() * BLOCK + threadIdx.x;
    if (i >= numAmps) return;
    auto elem = [&](uint64_t k) {
        double er = elems[2 * k], ei = elems[2 * k + 1];
        if (hasPower) {
            double lr = log(hypot(er, ei)), li = atan2(ei, er);
            double pr = expRe * lr - expIm * li, pi = expRe * li + expIm * lr;
            double mag = exp(pr);
            er = mag * cos(pi);
            ei = mag * sin(pi);
        }
        return cd{er, ei};
    };
    cd f;
    if (densN < 0) {
        f = elem(i);
    } else {
        // rho_rc <- d_r conj(d_c) rho_rc (cpu_subroutines.cpp: densmatr_allTargDiagMatr)
        uint64_t dim = 1ull << densN;
        f = elem(i & (dim - 1));
        if (!multiplyOnly) f = cmul(f, cconj(elem(i >> densN)));
    }
    ct<T> a = ld(amps, i);
    st(amps, i, cmulT(toT<T>(f), a));
}

// swap amplitudes where (targ1, targ2) = (1, 0) and (0, 1)
template <typename T, int IT, int NT>
__global__ void __launch_bounds__(BLOCK) k_swap(T* __restrict__ amps, uint64_t numWork,
                                               InsertSpec ins, uint64_t flip) {
    const uint64_t base = blockId() * (BLOCK * IT) + threadIdx.x;
    uint64_t ia[IT];
    ct<T> a[IT], b[IT];
#pragma unroll
    for (int j = 0; j < IT; j++) {
        uint64_t n = base + (uint64_t)j * BLOCK;
        if (n < numWork) {
            ia[j] = insertBits(n, ins);
            a[j] = ldx<NT>(amps, ia[j]);
            b[j] = ldx<NT>(amps, ia[j] ^ flip);
        }
    }
#pragma unroll
    for (int j = 0; j < IT; j++) {
        uint64_t n = base + (uint64_t)j * BLOCK;
        if (n < numWork) {
            stx<NT>(amps, ia[j], b[j]);
            stx<NT>(amps, ia[j] ^ flip, a[j]);
        }
    }
}

// Pauli tensor / gadget on pairs (iA, iB = iA ^ maskXY): iA enumerates the
// control-satisfied indices whose highest X/Y target bit is 0
// (cpu_subroutines.cpp applyPauliUponAmpPair)
template <typename T, int IT, int NT>
__global__ void __launch_bounds__(BLOCK) k_pauli(T* __restrict__ amps, uint64_t numWork,
                                                InsertSpec ins, uint64_t maskXY, uint64_t maskYZ,
                                                ct<T> ampFac, ct<T> pairFac) {
    const uint64_t base = blockId() * (BLOCK * IT) + threadIdx.x;
    uint64_t ia[IT];
    ct<T> a[IT], b[IT];
#pragma unroll
    for (int j = 0; j < IT; j++) {
        uint64_t n = base + (uint64_t)j * BLOCK;
        if (n < numWork) {
            ia[j] = insertBits(n, ins);
            a[j] = ldx<NT>(amps, ia[j]);
            b[j] = ldx<NT>(amps, ia[j] ^ maskXY);
        }
    }
#pragma unroll
    for (int j = 0; j < IT; j++) {
        uint64_t n = base + (uint64_t)j * BLOCK;
        if (n < numWork) {
            uint64_t ib = ia[j] ^ maskXY;
            T sA = parity64(ia[j] & maskYZ) ? (T)-1 : (T)1;
            T sB = parity64(ib & maskYZ) ? (T)-1 : (T)1;
            ct<T> pB = {pairFac.re * sB, pairFac.im * sB};
            ct<T> pA = {pairFac.re * sA, pairFac.im * sA};
            cmat2T(ampFac, pB, pA, ampFac, a[j], b[j]);
            stx<NT>(amps, ia[j], a[j]);
            stx<NT>(amps, ib, b[j]);
        }
    }
}

// a_i <- a_i * fac[parity(i & targMask)]
template <typename T, int IT, int NT>
__global__ void __launch_bounds__(BLOCK) k_phaseGadget(T* __restrict__ amps, uint64_t numWork,
                                                      InsertSpec ins, uint64_t targMask, ct<T> f0,
                                                      ct<T> f1) {
    const uint64_t base = blockId() * (BLOCK * IT) + threadIdx.x;
    uint64_t idx[IT];
    ct<T> a[IT];
#pragma unroll
    for (int j = 0; j < IT; j++) {
        uint64_t n = base + (uint64_t)j * BLOCK;
        if (n < numWork) {
            idx[j] = insertBits(n, ins);
            a[j] = ldx<NT>(amps, idx[j]);
        }
    }
#pragma unroll
    for (int j = 0; j < IT; j++) {
        uint64_t n = base + (uint64_t)j * BLOCK;
        if (n < numWork) stx<NT>(amps, idx[j], cmulT(parity64(idx[j] & targMask) ? f1 : f0, a[j]));
    }
}

// launch a streaming kernel in its best (items, hint) form given the lowest
// index bit any of its accesses varies or is fixed at
#define QA_LAUNCH_STREAMING(kernel, T, lowBit, work, ...)                                                 \
    do {                                                                                                  \
        if ((lowBit) >= STREAM_LOW_BIT)                                                                   \
            kernel<T, 2, 1><<<gridFor((work), 2), BLOCK, 0, stream()>>>(__VA_ARGS__);                     \
        else                                                                                              \
            kernel<T, 1, 0><<<gridFor((work), 1), BLOCK, 0, stream()>>>(__VA_ARGS__);                     \
    } while (0)

int lowestBit(const InsertSpec& ins, uint64_t extraMask) {
    int low = 64;
    if (ins.num > 0) low = ins.pos[0];
    if (extraMask) {
        int e = __builtin_ctzll(extraMask);
        if (e < low) low = e;
    }
    return low;
}

template <typename T>
__global__ void __launch_bounds__(BLOCK) k_superpose(T* __restrict__ out, const T* __restrict__ in1,
                                                    const T* __restrict__ in2, uint64_t numAmps,
                                                    ct<T> fo, ct<T> f1, ct<T> f2) {
    const uint64_t i = blockId() * BLOCK + threadIdx.x;
    if (i >= numAmps) return;
    ct<T> r = cmulT(fo, ld(out, i));
    r = cfmaT(f1, ld(in1, i), r);
    r = cfmaT(f2, ld(in2, i), r);
    st(out, i, r);
}

// ---- typed launchers ------------------------------------------------------

template <typename T>
void launchOneTargDense(void* amps, uint64_t work, const InsertSpec& ins, int targ, const double* m) {
    // ins already holds the target, so its lowest position covers every access
    QA_LAUNCH_STREAMING(k_oneTargDense, T, lowestBit(ins, 0), work, (T*)amps, work, ins, targ, toM1<T>(m));
}

template <typename T>
void launchTwoTargDense(void* amps, uint64_t work, const InsertSpec& ins, int t1, int t2,
                        const double* m) {
    M2<T> mm;
    for (int k = 0; k < 16; k++) mm.e[k] = ct<T>{(T)m[2 * k], (T)m[2 * k + 1]};
    k_twoTargDense<T><<<gridFor(work, 2), BLOCK, 0, stream()>>>((T*)amps, work, ins, t1, t2, mm);
}

template <typename T>
int launchAnyTargDense(void* amps, uint64_t work, const InsertSpec& ins, const TargSpec& ts,
                       const double* mat, int conj) {
    dim3 g = gridFor(work, 1);
    switch (ts.num) {
        case 3: k_fewTargDense<T, 3><<<g, BLOCK, 0, stream()>>>((T*)amps, work, ins, ts, mat, conj); break;
        case 4: k_fewTargDense<T, 4><<<g, BLOCK, 0, stream()>>>((T*)amps, work, ins, ts, mat, conj); break;
        case 5: k_fewTargDense<T, 5><<<g, BLOCK, 0, stream()>>>((T*)amps, work, ins, ts, mat, conj); break;
        default: {
            size_t lds = sizeof(ct<T>) << ts.num;
            if (lds > 64 * 1024) {
                setError("qa_statevec_anyCtrlAnyTargDenseMatr: %d targets exceed the LDS stage", ts.num);
                return QA_ERR_ARG;
            }
            uint64_t gg = work < 65536 ? work : 65536;
            k_manyTargDense<T><<<(unsigned)gg, BLOCK, lds, stream()>>>((T*)amps, work, ins, ts, mat, conj);
        }
    }
    return QA_OK;
}

#define QA_DISPATCH(prec, ...)                     \
    do {                                           \
        if ((prec) == QA_FP64) {                   \
            typedef double T;                      \
            __VA_ARGS__;                           \
        } else {                                   \
            typedef float T;                       \
            __VA_ARGS__;                           \
        }                                          \
    } while (0)

}  // namespace
}  // namespace qa

using namespace qa;

extern "C" {

int qa_statevec_anyCtrlSwap(void* amps, int prec, int numQubits, const int* ctrls, const int* ctrlStates,
                            int numCtrls, int targ1, int targ2) {
    QA_REQUIRE_INIT();
    if (!checkQureg(amps, prec, numQubits, __func__)) return QA_ERR_ARG;
    int ex[2] = {targ1, targ2}, vals[2] = {1, 0};
    InsertSpec ins;
    if (!buildInsert(ins, numQubits, ctrls, ctrlStates, numCtrls, ex, vals, 2, __func__)) return QA_ERR_ARG;
    {
        fuse::GOp op;
        op.kind = fuse::F_SWAP;
        op.t1 = targ1;
        op.t2 = targ2;
        fuse::ctrlMasks(ctrls, ctrlStates, numCtrls, op.cMask, op.cVal);
        QA_TRY_FUSE(amps, prec, numQubits, op);
    }
    uint64_t work = 1ull << (numQubits - numCtrls - 2);
    uint64_t flip = (1ull << targ1) | (1ull << targ2);
    QA_DISPATCH(prec, QA_LAUNCH_STREAMING(k_swap, T, lowestBit(ins, 0), work, (T*)amps, work, ins, flip));
    QA_LAUNCH_CHECK();
    return QA_OK;
}

int qa_statevec_anyCtrlOneTargDenseMatr(void* amps, int prec, int numQubits, const int* ctrls,
                                        const int* ctrlStates, int numCtrls, int targ, const double* matr) {
    QA_REQUIRE_INIT();
    if (!checkQureg(amps, prec, numQubits, __func__)) return QA_ERR_ARG;
    InsertSpec ins;
    if (!buildInsert(ins, numQubits, ctrls, ctrlStates, numCtrls, &targ, nullptr, 1, __func__))
        return QA_ERR_ARG;
    {
        fuse::GOp op;
        op.kind = fuse::F_DENSE1;
        op.t1 = targ;
        fuse::ctrlMasks(ctrls, ctrlStates, numCtrls, op.cMask, op.cVal);
        for (int i = 0; i < 8; i++) op.f[i] = matr[i];
        QA_TRY_FUSE(amps, prec, numQubits, op);
    }
    uint64_t work = 1ull << (numQubits - numCtrls - 1);
    QA_DISPATCH(prec, launchOneTargDense<T>(amps, work, ins, targ, matr));
    QA_LAUNCH_CHECK();
    return QA_OK;
}

int qa_statevec_anyCtrlTwoTargDenseMatr(void* amps, int prec, int numQubits, const int* ctrls,
                                        const int* ctrlStates, int numCtrls, int targ1, int targ2,
                                        const double* matr) {
    QA_REQUIRE_INIT();
    if (!checkQureg(amps, prec, numQubits, __func__)) return QA_ERR_ARG;
    int ex[2] = {targ1, targ2};
    InsertSpec ins;
    if (!buildInsert(ins, numQubits, ctrls, ctrlStates, numCtrls, ex, nullptr, 2, __func__)) return QA_ERR_ARG;
    {
        fuse::GOp op;
        op.kind = fuse::F_DENSE2;
        op.t1 = targ1;
        op.t2 = targ2;
        fuse::ctrlMasks(ctrls, ctrlStates, numCtrls, op.cMask, op.cVal);
        for (int i = 0; i < 32; i++) op.f[i] = matr[i];
        QA_TRY_FUSE(amps, prec, numQubits, op);
    }
    uint64_t work = 1ull << (numQubits - numCtrls - 2);
    QA_DISPATCH(prec, launchTwoTargDense<T>(amps, work, ins, targ1, targ2, matr));
    QA_LAUNCH_CHECK();
    return QA_OK;
}

int qa_statevec_anyCtrlAnyTargDenseMatr(void* amps, int prec, int numQubits, const int* ctrls,
                                        const int* ctrlStates, int numCtrls, const int* targs, int numTargs,
                                        const void* devMatr, int applyConj) {
    QA_REQUIRE_INIT();
    if (!checkQureg(amps, prec, numQubits, __func__)) return QA_ERR_ARG;
    QA_ARG_CHECK(numTargs >= 1 && numTargs <= 12, "%s: %d targets unsupported", __func__, numTargs);
    QA_ARG_CHECK(devMatr != nullptr, "%s: null matrix", __func__);
    if (numTargs <= 2) {
        // 1 and 2 targets go through the by-value kernels
        double m[32];
        int d = 1 << numTargs;
        QA_HIP(hipMemcpyAsync(m, devMatr, (size_t)(d * d) * 16, hipMemcpyDeviceToHost, stream()));
        QA_HIP(hipStreamSynchronize(stream()));
        if (applyConj)
            for (int k = 0; k < d * d; k++) m[2 * k + 1] = -m[2 * k + 1];
        return numTargs == 1
                   ? qa_statevec_anyCtrlOneTargDenseMatr(amps, prec, numQubits, ctrls, ctrlStates, numCtrls,
                                                         targs[0], m)
                   : qa_statevec_anyCtrlTwoTargDenseMatr(amps, prec, numQubits, ctrls, ctrlStates, numCtrls,
                                                         targs[0], targs[1], m);
    }
    InsertSpec ins;
    if (!buildInsert(ins, numQubits, ctrls, ctrlStates, numCtrls, targs, nullptr, numTargs, __func__))
        return QA_ERR_ARG;
    if ((numTargs == 3 || numTargs == 4) && numCtrls == 0 && jit::settings().mode != 0 && jit::available() &&
        fuse::enabledFor(amps, prec, numQubits)) {
        // queued with a copy of its matrix (the caller's may change or go
        // before the queue is applied); the tile pass applies it in LDS, in
        // the pass's generated kernel (compiled at first sight when the pass
        // runs), so with generated kernels off the gate runs unfused
        const double* staged = fuse::stageMatrix(devMatr, (size_t)2 << (2 * numTargs));
        if (!staged) return QA_ERR_HIP;
        fuse::GOp op;
        op.kind = fuse::F_DENSEN;
        op.numTargs = numTargs;
        for (int b = 0; b < numTargs; b++) op.targs[b] = targs[b];
        op.conj = applyConj ? 1 : 0;
        op.mat = staged;
        QA_TRY_FUSE(amps, prec, numQubits, op);
    }
    QA_FLUSH_ALL();
    TargSpec ts;
    ts.num = numTargs;
    for (int i = 0; i < numTargs; i++) ts.targ[i] = targs[i];
    uint64_t work = 1ull << (numQubits - numCtrls - numTargs);
    int rc = QA_OK;
    QA_DISPATCH(prec, rc = launchAnyTargDense<T>(amps, work, ins, ts, (const double*)devMatr, applyConj));
    if (rc != QA_OK) return rc;
    QA_LAUNCH_CHECK();
    return QA_OK;
}

int qa_statevec_anyCtrlOneTargDiagMatr(void* amps, int prec, int numQubits, const int* ctrls,
                                       const int* ctrlStates, int numCtrls, int targ, const double* e) {
    QA_REQUIRE_INIT();
    if (!checkQureg(amps, prec, numQubits, __func__)) return QA_ERR_ARG;
    QA_ARG_CHECK(targ >= 0 && targ < numQubits, "%s: target %d out of range", __func__, targ);
    for (int i = 0; i < numCtrls; i++)
        QA_ARG_CHECK(ctrls[i] != targ, "%s: control %d equals the target", __func__, targ);
    // a phase-only diagonal diag(1, d1) (phase shifts, CZ, S, T) leaves every
    // amplitude with a 0 target bit exactly unchanged: fold the target into the
    // insertion spec as one more "control" on 1 and touch only the others
    bool phaseOnly = (e[0] == 1.0 && e[1] == 0.0);
    InsertSpec ins;
    int one = 1;
    if (!buildInsert(ins, numQubits, ctrls, ctrlStates, numCtrls, phaseOnly ? &targ : nullptr,
                     phaseOnly ? &one : nullptr, phaseOnly ? 1 : 0, __func__))
        return QA_ERR_ARG;
    {
        fuse::GOp op;
        op.kind = fuse::F_DIAG1;
        op.t1 = targ;
        fuse::ctrlMasks(ctrls, ctrlStates, numCtrls, op.cMask, op.cVal);
        if (phaseOnly) {
            op.cMask |= 1ull << targ;
            op.cVal |= 1ull << targ;
        }
        for (int i = 0; i < 4; i++) op.f[i] = e[i];
        QA_TRY_FUSE(amps, prec, numQubits, op);
    }
    uint64_t work = 1ull << (numQubits - numCtrls - (phaseOnly ? 1 : 0));
    int kt = phaseOnly ? -1 : targ;
    int low = lowestBit(ins, 0);  // runs of consecutive amplitudes break only at inserted bits
    QA_DISPATCH(prec, QA_LAUNCH_STREAMING(k_oneTargDiag, T, low, work, (T*)amps, work, ins, kt,
                                          (ct<T>{(T)e[0], (T)e[1]}), (ct<T>{(T)e[2], (T)e[3]})));
    QA_LAUNCH_CHECK();
    return QA_OK;
}

int qa_statevec_anyCtrlTwoTargDiagMatr(void* amps, int prec, int numQubits, const int* ctrls,
                                       const int* ctrlStates, int numCtrls, int targ1, int targ2,
                                       const double* e) {
    QA_REQUIRE_INIT();
    if (!checkQureg(amps, prec, numQubits, __func__)) return QA_ERR_ARG;
    QA_ARG_CHECK(targ1 >= 0 && targ1 < numQubits && targ2 >= 0 && targ2 < numQubits && targ1 != targ2,
                 "%s: bad targets", __func__);
    InsertSpec ins;
    if (!buildInsert(ins, numQubits, ctrls, ctrlStates, numCtrls, nullptr, nullptr, 0, __func__))
        return QA_ERR_ARG;
    QA_FLUSH_ALL();
    uint64_t work = 1ull << (numQubits - numCtrls);
    QA_DISPATCH(prec, (k_twoTargDiag<T><<<gridFor(work), BLOCK, 0, stream()>>>(
                          (T*)amps, work, ins, targ1, targ2, ct<T>{(T)e[0], (T)e[1]}, ct<T>{(T)e[2], (T)e[3]},
                          ct<T>{(T)e[4], (T)e[5]}, ct<T>{(T)e[6], (T)e[7]})));
    QA_LAUNCH_CHECK();
    return QA_OK;
}

int qa_statevec_anyCtrlAnyTargDiagMatr(void* amps, int prec, int numQubits, const int* ctrls,
                                       const int* ctrlStates, int numCtrls, const int* targs, int numTargs,
                                       const void* devElems, int applyConj, int hasPower, double expRe,
                                       double expIm) {
    QA_REQUIRE_INIT();
    if (!checkQureg(amps, prec, numQubits, __func__)) return QA_ERR_ARG;
    QA_ARG_CHECK(numTargs >= 1 && numTargs <= MAXQ, "%s: bad target count", __func__);
    QA_ARG_CHECK(devElems != nullptr, "%s: null elements", __func__);
    for (int i = 0; i < numTargs; i++) {
        QA_ARG_CHECK(targs[i] >= 0 && targs[i] < numQubits, "%s: target out of range", __func__);
        for (int c = 0; c < numCtrls; c++)
            QA_ARG_CHECK(ctrls[c] != targs[i], "%s: control equals a target", __func__);
    }
    InsertSpec ins;
    if (!buildInsert(ins, numQubits, ctrls, ctrlStates, numCtrls, nullptr, nullptr, 0, __func__))
        return QA_ERR_ARG;
    QA_FLUSH_ALL();
    TargSpec ts;
    ts.num = numTargs;
    for (int i = 0; i < numTargs; i++) ts.targ[i] = targs[i];
    uint64_t work = 1ull << (numQubits - numCtrls);
    QA_DISPATCH(prec, (k_anyTargDiag<T><<<gridFor(work, 1), BLOCK, 0, stream()>>>(
                          (T*)amps, work, ins, ts, (const double*)devElems, applyConj, hasPower, expRe, expIm)));
    QA_LAUNCH_CHECK();
    return QA_OK;
}

int qa_statevec_allTargDiagMatr(void* amps, int prec, int numQubits, const void* devElems, int hasPower,
                                double expRe, double expIm) {
    QA_REQUIRE_INIT();
    if (!checkQureg(amps, prec, numQubits, __func__)) return QA_ERR_ARG;
    QA_FLUSH_ALL();
    uint64_t n = 1ull << numQubits;
    QA_DISPATCH(prec, (k_allTargDiag<T><<<gridFor(n, 1), BLOCK, 0, stream()>>>(
                          (T*)amps, n, (const double*)devElems, hasPower, expRe, expIm, -1, 0)));
    QA_LAUNCH_CHECK();
    return QA_OK;
}

int qa_densmatr_allTargDiagMatr(void* amps, int prec, int numQubitsN, const void* devElems, int hasPower,
                                double expRe, double expIm, int multiplyOnly) {
    QA_REQUIRE_INIT();
    if (!checkQureg(amps, prec, 2 * numQubitsN, __func__)) return QA_ERR_ARG;
    QA_FLUSH_ALL();
    uint64_t n = 1ull << (2 * numQubitsN);
    QA_DISPATCH(prec, (k_allTargDiag<T><<<gridFor(n, 1), BLOCK, 0, stream()>>>(
                          (T*)amps, n, (const double*)devElems, hasPower, expRe, expIm, numQubitsN,
                          multiplyOnly)));
    QA_LAUNCH_CHECK();
    return QA_OK;
}

int qa_statevec_anyCtrlPauliTensorOrGadget(void* amps, int prec, int numQubits, const int* ctrls,
                                           const int* ctrlStates, int numCtrls, uint64_t maskX, uint64_t maskY,
                                           uint64_t maskZ, double ampFacRe, double ampFacIm, double pairFacRe,
                                           double pairFacIm) {
    QA_REQUIRE_INIT();
    if (!checkQureg(amps, prec, numQubits, __func__)) return QA_ERR_ARG;
    uint64_t maskXY = maskX | maskY;
    QA_ARG_CHECK(maskXY != 0, "%s: the Pauli string has no X or Y", __func__);
    QA_ARG_CHECK(((maskX & maskY) | (maskX & maskZ) | (maskY & maskZ)) == 0, "%s: overlapping Pauli masks",
                 __func__);
    QA_ARG_CHECK(numQubits >= 64 || ((maskX | maskY | maskZ) >> numQubits) == 0, "%s: Pauli beyond register",
                 __func__);
    for (int c = 0; c < numCtrls; c++)
        QA_ARG_CHECK(!(((maskX | maskY | maskZ) >> ctrls[c]) & 1ull), "%s: control %d is also a target",
                     __func__, ctrls[c]);
    int top = 63 - __builtin_clzll(maskXY);
    InsertSpec ins;
    if (!buildInsert(ins, numQubits, ctrls, ctrlStates, numCtrls, &top, nullptr, 1, __func__))
        return QA_ERR_ARG;
    // pair factor gains i^#Y (cpu_subroutines.cpp: util_getPowerOfI(y.size()))
    cd pf = {pairFacRe, pairFacIm};
    int ny = __builtin_popcountll(maskY) & 3;
    for (int k = 0; k < ny; k++) pf = cmul(pf, cd{0, 1});
    {
        fuse::GOp op;
        op.kind = fuse::F_PAULI;
        op.m1 = maskXY;
        op.m2 = maskY | maskZ;
        fuse::ctrlMasks(ctrls, ctrlStates, numCtrls, op.cMask, op.cVal);
        op.f[0] = ampFacRe;
        op.f[1] = ampFacIm;
        op.f[2] = pf.re;
        op.f[3] = pf.im;
        QA_TRY_FUSE(amps, prec, numQubits, op);
    }
    uint64_t work = 1ull << (numQubits - numCtrls - 1);
    QA_DISPATCH(prec, QA_LAUNCH_STREAMING(k_pauli, T, lowestBit(ins, maskXY), work, (T*)amps, work, ins, maskXY,
                                          maskY | maskZ, (ct<T>{(T)ampFacRe, (T)ampFacIm}), toT<T>(pf)));
    QA_LAUNCH_CHECK();
    return QA_OK;
}

int qa_statevec_anyCtrlAnyTargZOrPhaseGadget(void* amps, int prec, int numQubits, const int* ctrls,
                                             const int* ctrlStates, int numCtrls, uint64_t targMask,
                                             double fac0Re, double fac0Im, double fac1Re, double fac1Im) {
    QA_REQUIRE_INIT();
    if (!checkQureg(amps, prec, numQubits, __func__)) return QA_ERR_ARG;
    InsertSpec ins;
    if (!buildInsert(ins, numQubits, ctrls, ctrlStates, numCtrls, nullptr, nullptr, 0, __func__))
        return QA_ERR_ARG;
    {
        fuse::GOp op;
        op.kind = fuse::F_PARITY;
        op.m1 = targMask;
        fuse::ctrlMasks(ctrls, ctrlStates, numCtrls, op.cMask, op.cVal);
        op.f[0] = fac0Re;
        op.f[1] = fac0Im;
        op.f[2] = fac1Re;
        op.f[3] = fac1Im;
        QA_TRY_FUSE(amps, prec, numQubits, op);
    }
    uint64_t work = 1ull << (numQubits - numCtrls);
    QA_DISPATCH(prec, QA_LAUNCH_STREAMING(k_phaseGadget, T, lowestBit(ins, 0), work, (T*)amps, work, ins, targMask,
                                          (ct<T>{(T)fac0Re, (T)fac0Im}), (ct<T>{(T)fac1Re, (T)fac1Im})));
    QA_LAUNCH_CHECK();
    return QA_OK;
}

int qa_statevec_setQuregToSuperposition(int prec, int numQubits, double foRe, double foIm, void* out,
                                        double f1Re, double f1Im, const void* in1, double f2Re, double f2Im,
                                        const void* in2) {
    QA_REQUIRE_INIT();
    if (!checkQureg(out, prec, numQubits, __func__)) return QA_ERR_ARG;
    QA_ARG_CHECK(in1 && in2, "%s: null input", __func__);
    QA_FLUSH_ALL();
    uint64_t n = 1ull << numQubits;
    QA_DISPATCH(prec, (k_superpose<T><<<gridFor(n, 1), BLOCK, 0, stream()>>>(
                          (T*)out, (const T*)in1, (const T*)in2, n, ct<T>{(T)foRe, (T)foIm}, ct<T>{(T)f1Re, (T)f1Im},
                          ct<T>{(T)f2Re, (T)f2Im})));
    QA_LAUNCH_CHECK();
    return QA_OK;
}

}  // extern "C"
