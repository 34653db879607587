// Density-matrix superoperator kernels (dephasing, depolarising, Pauli
// channel, damping), mixing and partial trace.  The matrix is flattened
// column-major over 2N index bits: ket qubit q is bit q, bra qubit q is bit
// q+N.  Replaces quest/src/gpu/gpu_kernels.cuh:705-1136 (the
// non-distributed *_subA variants) and gpu_subroutines.cpp:910-1420.
//
// Every channel is one streaming pass: each thread owns the 4 amplitudes
// {ket q, bra q} in {00,01,10,11} (or the 16 of a 2-qubit channel), which
// are at distance 2^q and 2^(q+N) -- the ket stride keeps a wave's loads in
// contiguous runs, the bra stride only selects which runs.
#include <algorithm>
#include <vector>

#include "qa_common.hpp"
#include "qa_fused.hpp"

namespace qa {
namespace {

bool checkDM(const void* amps, int prec, int N, const char* caller) {
    if (!amps || (prec != QA_FP32 && prec != QA_FP64) || N < 1 || 2 * N > 62) {
        setError("%s: bad density-matrix arguments", caller);
        return false;
    }
    return true;
}

InsertSpec zerosAt(std::vector<int> qs) {
    std::sort(qs.begin(), qs.end());
    InsertSpec s;
    s.num = (int)qs.size();
    s.valueMask = 0;
    for (int i = 0; i < s.num; i++) s.pos[i] = qs[(size_t)i];
    return s;
}

// mixDephasing: off-diagonal-in-q entries scale by fac = 1 - 2p
// (cpu_subroutines.cpp: cpu_densmatr_oneQubitDephasing_subA)
template <typename T>
__global__ void __launch_bounds__(BLOCK) k_dephase1(T* __restrict__ amps, uint64_t numWork, InsertSpec ins,
                                                   uint64_t ketBit, uint64_t braBit, T fac) {
    const uint64_t base = blockId() * (BLOCK * ITEMS) + threadIdx.x;
#pragma unroll
    for (int j = 0; j < ITEMS; j++) {
        uint64_t n = base + (uint64_t)j * BLOCK;
        if (n >= numWork) continue;
        uint64_t i00 = insertBits(n, ins);
        uint64_t i01 = i00 | ketBit, i10 = i00 | braBit;
        ct<T> a = ldnt(amps, i01), b = ldnt(amps, i10);
        stnt(amps, i01, ct<T>{a.re * fac, a.im * fac});
        stnt(amps, i10, ct<T>{b.re * fac, b.im * fac});
    }
}

// mixTwoQubitDephasing: amp *= 1 + term * (ketA!=braA | ketB!=braB)
template <typename T>
__global__ void __launch_bounds__(BLOCK) k_dephase2(T* __restrict__ amps, uint64_t numAmps, int kA, int bA,
                                                   int kB, int bB, T term) {
    const uint64_t base = blockId() * (BLOCK * ITEMS) + threadIdx.x;
#pragma unroll
    for (int j = 0; j < ITEMS; j++) {
        uint64_t i = base + (uint64_t)j * BLOCK;
        if (i >= numAmps) continue;
        int flag = (int)((((i >> kA) ^ (i >> bA)) | ((i >> kB) ^ (i >> bB))) & 1ull);
        if (flag) {
            ct<T> a = ld(amps, i);
            T f = (T)1 + term;
            st(amps, i, ct<T>{a.re * f, a.im * f});
        }
    }
}

// mixDepolarising (cpu_densmatr_oneQubitDepolarising_subA):
//   a00 <- c1 a00 + c2 a11, a11 <- c1 a11 + c2 a00, a01,a10 <- c3 a
template <typename T>
__global__ void __launch_bounds__(BLOCK) k_depol1(T* __restrict__ amps, uint64_t numWork, InsertSpec ins,
                                                 uint64_t ketBit, uint64_t braBit, T c1, T c2, T c3) {
    const uint64_t base = blockId() * (BLOCK * 2) + threadIdx.x;
    uint64_t i00[2];
    ct<T> a00[2], a01[2], a10[2], a11[2];
#pragma unroll
    for (int j = 0; j < 2; j++) {
        uint64_t n = base + (uint64_t)j * BLOCK;
        if (n < numWork) {
            i00[j] = insertBits(n, ins);
            a00[j] = ldnt(amps, i00[j]);
            a01[j] = ldnt(amps, i00[j] | ketBit);
            a10[j] = ldnt(amps, i00[j] | braBit);
            a11[j] = ldnt(amps, i00[j] | ketBit | braBit);
        }
    }
#pragma unroll
    for (int j = 0; j < 2; j++) {
        uint64_t n = base + (uint64_t)j * BLOCK;
        if (n < numWork) {
            ct<T> n00 = {fmaT(c1, a00[j].re, c2 * a11[j].re), fmaT(c1, a00[j].im, c2 * a11[j].im)};
            ct<T> n11 = {fmaT(c1, a11[j].re, c2 * a00[j].re), fmaT(c1, a11[j].im, c2 * a00[j].im)};
            stnt(amps, i00[j], n00);
            stnt(amps, i00[j] | ketBit, ct<T>{c3 * a01[j].re, c3 * a01[j].im});
            stnt(amps, i00[j] | braBit, ct<T>{c3 * a10[j].re, c3 * a10[j].im});
            stnt(amps, i00[j] | ketBit | braBit, n11);
        }
    }
}

// mixPaulis (cpu_densmatr_oneQubitPauliChannel_subA)
template <typename T>
__global__ void __launch_bounds__(BLOCK) k_pauliChannel(T* __restrict__ amps, uint64_t numWork,
                                                       InsertSpec ins, uint64_t ketBit, uint64_t braBit,
                                                       T fAA, T fBB, T fAB, T fBA) {
    const uint64_t n = blockId() * BLOCK + threadIdx.x;
    if (n >= numWork) return;
    uint64_t i00 = insertBits(n, ins);
    uint64_t i01 = i00 | ketBit, i10 = i00 | braBit, i11 = i01 | braBit;
    ct<T> a00 = ld(amps, i00), a01 = ld(amps, i01), a10 = ld(amps, i10), a11 = ld(amps, i11);
    st(amps, i00, ct<T>{fAA * a00.re + fBB * a11.re, fAA * a00.im + fBB * a11.im});
    st(amps, i01, ct<T>{fAB * a01.re + fBA * a10.re, fAB * a01.im + fBA * a10.im});
    st(amps, i10, ct<T>{fAB * a10.re + fBA * a01.re, fAB * a10.im + fBA * a01.im});
    st(amps, i11, ct<T>{fAA * a11.re + fBB * a00.re, fAA * a11.im + fBB * a00.im});
}

// mixDamping (cpu_densmatr_oneQubitDamping_subA)
template <typename T>
__global__ void __launch_bounds__(BLOCK) k_damp(T* __restrict__ amps, uint64_t numWork, InsertSpec ins,
                                               uint64_t ketBit, uint64_t braBit, T prob, T c1, T c2) {
    const uint64_t n = blockId() * BLOCK + threadIdx.x;
    if (n >= numWork) return;
    uint64_t i00 = insertBits(n, ins);
    uint64_t i01 = i00 | ketBit, i10 = i00 | braBit, i11 = i01 | braBit;
    ct<T> a00 = ld(amps, i00), a01 = ld(amps, i01), a10 = ld(amps, i10), a11 = ld(amps, i11);
    st(amps, i00, ct<T>{a00.re + prob * a11.re, a00.im + prob * a11.im});
    st(amps, i01, ct<T>{c1 * a01.re, c1 * a01.im});
    st(amps, i10, ct<T>{c1 * a10.re, c1 * a10.im});
    st(amps, i11, ct<T>{c2 * a11.re, c2 * a11.im});
}

// mixTwoQubitDepolarising, fused form of cpu_densmatr_twoQubitDepolarising
// _subA then _subB: every amplitude first scales by (1 + c3*mod), mod = not
// (ket1==bra1 and ket2==bra2); then on the 4 amplitudes with ket=bra on both
// qubits, a <- (c1-c2) a + c2 * sum.  One thread owns the 16 amplitudes of
// one (ket1,ket2,bra1,bra2) group so both steps happen in one pass.
template <typename T>
__global__ void __launch_bounds__(BLOCK) k_depol2(T* __restrict__ amps, uint64_t numWork, InsertSpec ins,
                                                 uint64_t k1, uint64_t k2, uint64_t b1, uint64_t b2, T c1mc2,
                                                 T c2, T c3) {
    const uint64_t n = blockId() * BLOCK + threadIdx.x;
    if (n >= numWork) return;
    const uint64_t i0 = insertBits(n, ins);
    ct<T> diag[4];
    uint64_t diagIdx[4];
    int nd = 0;
#pragma unroll
    for (int v = 0; v < 16; v++) {
        uint64_t i = i0 | ((v & 1) ? k1 : 0) | ((v & 2) ? k2 : 0) | ((v & 4) ? b1 : 0) | ((v & 8) ? b2 : 0);
        bool same = (((v & 1) != 0) == ((v & 4) != 0)) && (((v & 2) != 0) == ((v & 8) != 0));
        ct<T> a = ldnt(amps, i);
        if (same) {
            diag[nd] = a;
            diagIdx[nd] = i;
            nd++;
        } else {
            T f = (T)1 + c3;
            stnt(amps, i, ct<T>{a.re * f, a.im * f});
        }
    }
    ct<T> sum = {diag[0].re + diag[1].re + diag[2].re + diag[3].re,
                 diag[0].im + diag[1].im + diag[2].im + diag[3].im};
#pragma unroll
    for (int k = 0; k < 4; k++)
        stnt(amps, diagIdx[k], ct<T>{c1mc2 * diag[k].re + c2 * sum.re, c1mc2 * diag[k].im + c2 * sum.im});
}

template <typename T>
__global__ void __launch_bounds__(BLOCK) k_mixDM(T* __restrict__ out, const T* __restrict__ in, uint64_t num,
                                                T po, T pi) {
    const uint64_t i = blockId() * BLOCK + threadIdx.x;
    if (i >= num) return;
    ct<T> a = ld(out, i), b = ld(in, i);
    st(out, i, ct<T>{po * a.re + pi * b.re, po * a.im + pi * b.im});
}

// out_rc <- po out_rc + pi psi_r conj(psi_c)
template <typename T>
__global__ void __launch_bounds__(BLOCK) k_mixSV(T* __restrict__ out, const T* __restrict__ psi, uint64_t num,
                                                int N, T po, T pi) {
    const uint64_t i = blockId() * BLOCK + threadIdx.x;
    if (i >= num) return;
    uint64_t r = i & ((1ull << N) - 1), c = i >> N;
    ct<T> a = ld(out, i), pr = ld(psi, r), pc = ld(psi, c);
    ct<T> t = cmulT(pr, ct<T>{pc.re, -pc.im});
    st(out, i, ct<T>{po * a.re + pi * t.re, po * a.im + pi * t.im});
}

// rho <- |psi><psi|
template <typename T>
__global__ void __launch_bounds__(BLOCK) k_pureDM(T* __restrict__ rho, const T* __restrict__ psi, uint64_t num,
                                                 int N) {
    const uint64_t i = blockId() * BLOCK + threadIdx.x;
    if (i >= num) return;
    uint64_t r = i & ((1ull << N) - 1), c = i >> N;
    ct<T> pr = ld(psi, r), pc = ld(psi, c);
    st(rho, i, cmulT(pr, ct<T>{pc.re, -pc.im}));
}

// partial trace (cpu_densmatr_partialTrace_sub): out[n] = sum_j in[k with
// targs and pairTargs both set to j], k = n with zeros inserted at all
// (sorted) traced ket and bra positions
template <typename T>
__global__ void __launch_bounds__(BLOCK) k_partialTrace(const T* __restrict__ in, T* __restrict__ out,
                                                       uint64_t numOut, InsertSpec ins, TargSpec targs,
                                                       TargSpec pairs) {
    const uint64_t n = blockId() * BLOCK + threadIdx.x;
    if (n >= numOut) return;
    uint64_t k = insertBits(n, ins);
    uint64_t numInner = 1ull << targs.num;
    double re = 0, im = 0;
    for (uint64_t j = 0; j < numInner; j++) {
        uint64_t i = setBitsFrom(setBitsFrom(k, targs, j), pairs, j);
        ct<T> a = ld(in, i);
        re += (double)a.re;
        im += (double)a.im;
    }
    st(out, n, ct<T>{(T)re, (T)im});
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

int qa_densmatr_oneQubitDephasing(void* amps, int prec, int N, int q, double prob) {
    QA_REQUIRE_INIT();
    if (!checkDM(amps, prec, N, __func__)) return QA_ERR_ARG;
    QA_ARG_CHECK(q >= 0 && q < N, "%s: qubit out of range", __func__);
    InsertSpec ins = zerosAt({q, q + N});
    uint64_t work = 1ull << (2 * N - 2);
    double fac = 1 - 2 * prob;  // util_getOneQubitDephasingFactor
    {
        fuse::GOp op;
        op.kind = fuse::F_DEPH1;
        op.t1 = q;
        op.t2 = q + N;
        op.f[0] = fac;
        QA_TRY_FUSE(amps, prec, 2 * N, op);
    }
    QA_DISPATCH(prec, (k_dephase1<T><<<gridFor(work), BLOCK, 0, stream()>>>(
                          (T*)amps, work, ins, 1ull << q, 1ull << (q + N), (T)fac)));
    QA_LAUNCH_CHECK();
    return QA_OK;
}

int qa_densmatr_twoQubitDephasing(void* amps, int prec, int N, int q1, int q2, double prob) {
    QA_REQUIRE_INIT();
    QA_FLUSH_ALL();
    if (!checkDM(amps, prec, N, __func__)) return QA_ERR_ARG;
    QA_ARG_CHECK(q1 >= 0 && q1 < N && q2 >= 0 && q2 < N && q1 != q2, "%s: bad qubits", __func__);
    uint64_t num = 1ull << (2 * N);
    double term = -4 * prob / 3;  // util_getTwoQubitDephasingTerm
    QA_DISPATCH(prec, (k_dephase2<T><<<gridFor(num), BLOCK, 0, stream()>>>((T*)amps, num, q1, q1 + N, q2,
                                                                             q2 + N, (T)term)));
    QA_LAUNCH_CHECK();
    return QA_OK;
}

int qa_densmatr_oneQubitDepolarising(void* amps, int prec, int N, int q, double prob) {
    QA_REQUIRE_INIT();
    if (!checkDM(amps, prec, N, __func__)) return QA_ERR_ARG;
    QA_ARG_CHECK(q >= 0 && q < N, "%s: qubit out of range", __func__);
    InsertSpec ins = zerosAt({q, q + N});
    uint64_t work = 1ull << (2 * N - 2);
    // util_getOneQubitDepolarisingFactors (utilities.cpp:932-947)
    double c1 = 1 - (2 * prob / 3), c2 = 2 * prob / 3, c3 = 1 - (4 * prob / 3);
    {
        fuse::GOp op;
        op.kind = fuse::F_DEPOL1;
        op.t1 = q;
        op.t2 = q + N;
        op.f[0] = c1;
        op.f[1] = c2;
        op.f[2] = c3;
        QA_TRY_FUSE(amps, prec, 2 * N, op);
    }
    QA_DISPATCH(prec, (k_depol1<T><<<gridFor(work, 2), BLOCK, 0, stream()>>>(
                          (T*)amps, work, ins, 1ull << q, 1ull << (q + N), (T)c1, (T)c2, (T)c3)));
    QA_LAUNCH_CHECK();
    return QA_OK;
}

int qa_densmatr_twoQubitDepolarising(void* amps, int prec, int N, int q1, int q2, double prob) {
    QA_REQUIRE_INIT();
    QA_FLUSH_ALL();
    if (!checkDM(amps, prec, N, __func__)) return QA_ERR_ARG;
    QA_ARG_CHECK(q1 >= 0 && q1 < N && q2 >= 0 && q2 < N && q1 != q2, "%s: bad qubits", __func__);
    InsertSpec ins = zerosAt({q1, q2, q1 + N, q2 + N});
    uint64_t work = 1ull << (2 * N - 4);
    // util_getTwoQubitDepolarisingFactors (utilities.cpp)
    double c1 = 1 - (4 * prob / 5), c2 = 4 * prob / 15, c3 = -(16 * prob / 15);
    // _subB applies to amplitudes already scaled by (1 + c3*0) = 1 (they are
    // exactly the ones with ket==bra on both qubits), so fusing is exact
    QA_DISPATCH(prec, (k_depol2<T><<<gridFor(work, 1), BLOCK, 0, stream()>>>(
                          (T*)amps, work, ins, 1ull << q1, 1ull << q2, 1ull << (q1 + N), 1ull << (q2 + N),
                          (T)(c1 - c2), (T)c2, (T)c3)));
    QA_LAUNCH_CHECK();
    return QA_OK;
}

int qa_densmatr_oneQubitPauliChannel(void* amps, int prec, int N, int q, double pI, double pX, double pY,
                                     double pZ) {
    QA_REQUIRE_INIT();
    QA_FLUSH_ALL();
    if (!checkDM(amps, prec, N, __func__)) return QA_ERR_ARG;
    QA_ARG_CHECK(q >= 0 && q < N, "%s: qubit out of range", __func__);
    InsertSpec ins = zerosAt({q, q + N});
    uint64_t work = 1ull << (2 * N - 2);
    // util_getOneQubitPauliChannelFactors
    double fAA = pI + pZ, fBB = pX + pY, fAB = pI - pZ, fBA = pX - pY;
    QA_DISPATCH(prec, (k_pauliChannel<T><<<gridFor(work, 1), BLOCK, 0, stream()>>>(
                          (T*)amps, work, ins, 1ull << q, 1ull << (q + N), (T)fAA, (T)fBB, (T)fAB, (T)fBA)));
    QA_LAUNCH_CHECK();
    return QA_OK;
}

int qa_densmatr_oneQubitDamping(void* amps, int prec, int N, int q, double prob) {
    QA_REQUIRE_INIT();
    QA_FLUSH_ALL();
    if (!checkDM(amps, prec, N, __func__)) return QA_ERR_ARG;
    QA_ARG_CHECK(q >= 0 && q < N, "%s: qubit out of range", __func__);
    InsertSpec ins = zerosAt({q, q + N});
    uint64_t work = 1ull << (2 * N - 2);
    double c1 = std::sqrt(1 - prob), c2 = 1 - prob;  // util_getOneQubitDampingFactors
    QA_DISPATCH(prec, (k_damp<T><<<gridFor(work, 1), BLOCK, 0, stream()>>>(
                          (T*)amps, work, ins, 1ull << q, 1ull << (q + N), (T)prob, (T)c1, (T)c2)));
    QA_LAUNCH_CHECK();
    return QA_OK;
}

int qa_densmatr_mixQureg(double outProb, void* out, double inProb, const void* in, int prec, int N) {
    QA_REQUIRE_INIT();
    QA_FLUSH_ALL();
    if (!checkDM(out, prec, N, __func__) || !in) return QA_ERR_ARG;
    uint64_t num = 1ull << (2 * N);
    QA_DISPATCH(prec, (k_mixDM<T><<<gridFor(num, 1), BLOCK, 0, stream()>>>((T*)out, (const T*)in, num,
                                                                             (T)outProb, (T)inProb)));
    QA_LAUNCH_CHECK();
    return QA_OK;
}

int qa_densmatr_mixQuregWithStatevec(double outProb, void* out, double inProb, const void* psi, int prec, int N) {
    QA_REQUIRE_INIT();
    QA_FLUSH_ALL();
    if (!checkDM(out, prec, N, __func__) || !psi) return QA_ERR_ARG;
    uint64_t num = 1ull << (2 * N);
    QA_DISPATCH(prec, (k_mixSV<T><<<gridFor(num, 1), BLOCK, 0, stream()>>>((T*)out, (const T*)psi, num, N,
                                                                             (T)outProb, (T)inProb)));
    QA_LAUNCH_CHECK();
    return QA_OK;
}

int qa_densmatr_initPureState(void* rho, const void* psi, int prec, int N) {
    QA_REQUIRE_INIT();
    QA_FLUSH_ALL();
    if (!checkDM(rho, prec, N, __func__) || !psi) return QA_ERR_ARG;
    uint64_t num = 1ull << (2 * N);
    QA_DISPATCH(prec, (k_pureDM<T><<<gridFor(num, 1), BLOCK, 0, stream()>>>((T*)rho, (const T*)psi, num, N)));
    QA_LAUNCH_CHECK();
    return QA_OK;
}

int qa_densmatr_partialTrace(const void* in, int prec, int inN, void* out, const int* targs, int numTargs) {
    QA_REQUIRE_INIT();
    QA_FLUSH_ALL();
    if (!checkDM(in, prec, inN, __func__) || !out) return QA_ERR_ARG;
    QA_ARG_CHECK(numTargs >= 1 && numTargs < inN, "%s: bad number of traced qubits", __func__);
    std::vector<int> all;
    TargSpec ts, ps;
    ts.num = ps.num = numTargs;
    for (int i = 0; i < numTargs; i++) {
        QA_ARG_CHECK(targs[i] >= 0 && targs[i] < inN, "%s: qubit out of range", __func__);
        ts.targ[i] = targs[i];
        ps.targ[i] = targs[i] + inN;
        all.push_back(targs[i]);
        all.push_back(targs[i] + inN);
    }
    InsertSpec ins = zerosAt(all);
    for (int i = 1; i < ins.num; i++)
        QA_ARG_CHECK(ins.pos[i] != ins.pos[i - 1], "%s: repeated qubit", __func__);
    uint64_t numOut = 1ull << (2 * (inN - numTargs));
    QA_DISPATCH(prec, (k_partialTrace<T><<<gridFor(numOut, 1), BLOCK, 0, stream()>>>(
                          (const T*)in, (T*)out, numOut, ins, ts, ps)));
    QA_LAUNCH_CHECK();
    return QA_OK;
}

}  // extern "C"
