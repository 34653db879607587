// Device-side basics of the gfx950 kernel layer: amplitude vector types,
// complex arithmetic in the storage precision, index bit algebra.  Written
// to compile both as part of the library (hipcc) and inside kernels the
// library generates at run time (hiprtc, which provides the HIP device
// runtime itself and no host headers).
//
// Index algebra follows quest/src/core/bitwise.hpp:34-254 (insertBits over
// ascending qubit positions, setBits over an unsorted target list, parity of
// masked bits); it is restated here for 64-bit device indices.
#pragma once

#ifndef __HIPCC_RTC__
#include <hip/hip_runtime.h>
#endif
#include <cstdint>

namespace qa {

// ---- amplitude storage -------------------------------------------------
// Interleaved {re, im}; native clang vectors so that a whole amplitude moves
// with one global_load/store (dwordx4 for FP64, dwordx2 for FP32).
typedef double v2d __attribute__((ext_vector_type(2)));
typedef float v2f __attribute__((ext_vector_type(2)));

template <typename T> struct Amp;
template <> struct Amp<double> { typedef v2d V; };
template <> struct Amp<float> { typedef v2f V; };

// complex value held in FP64 registers regardless of storage precision
struct cd {
    double re, im;
};

__host__ __device__ __forceinline__ cd mk(double r, double i) { return cd{r, i}; }
__host__ __device__ __forceinline__ cd cmul(cd a, cd b) {
    return cd{a.re * b.re - a.im * b.im, a.re * b.im + a.im * b.re};
}
__host__ __device__ __forceinline__ cd cadd(cd a, cd b) { return cd{a.re + b.re, a.im + b.im}; }
__host__ __device__ __forceinline__ cd cconj(cd a) { return cd{a.re, -a.im}; }
__host__ __device__ __forceinline__ cd cscale(cd a, double s) { return cd{a.re * s, a.im * s}; }

// complex arithmetic in the storage precision T (FP32 gates stay FP32)
template <typename T>
struct ct {
    T re, im;
};
// Fused multiply-adds are written out so that every kernel that applies a
// gate rounds it the same way (the compiler's own contraction choices may
// differ between kernels, the fused-tile and single-gate paths must not).
__device__ __forceinline__ double fmaT(double a, double b, double c) { return __builtin_fma(a, b, c); }
__device__ __forceinline__ float fmaT(float a, float b, float c) { return __builtin_fmaf(a, b, c); }
template <typename T>
__device__ __forceinline__ ct<T> cmulT(ct<T> a, ct<T> b) {
    T ii = a.im * b.im, ir = a.im * b.re;
    return ct<T>{fmaT(a.re, b.re, -ii), fmaT(a.re, b.im, ir)};
}
template <typename T>
__device__ __forceinline__ ct<T> caddT(ct<T> a, ct<T> b) {
    return ct<T>{a.re + b.re, a.im + b.im};
}
template <typename T>
__device__ __forceinline__ ct<T> cfmaT(ct<T> m, ct<T> a, ct<T> acc) {
    // acc + m*a, accumulated term by term
    T re = fmaT(m.re, a.re, acc.re), im = fmaT(m.re, a.im, acc.im);
    return ct<T>{fmaT(-m.im, a.im, re), fmaT(m.im, a.re, im)};
}

// Matrix-vector products over a pair or quad of amplitudes, y = M x, with
// y[k] finished by the two multiply-adds m_kk.re * x[k] that read x[k] last:
// every partial sum is formed before any output, so each output can take
// its input's registers (a kernel that keeps amplitudes in registers across
// many gates then needs no copies).  Row k sums, in order: the cross term of
// Im m_kk, then m_kj x[j] for j != k ascending, then Re m_kk x[k].
template <typename T>
__device__ __forceinline__ ct<T> crowStartT(ct<T> mkk, ct<T> xk) {
    return ct<T>{-(mkk.im * xk.im), mkk.im * xk.re};
}
template <typename T>
__device__ __forceinline__ void crowAddT(ct<T>& acc, ct<T> m, ct<T> x) {
    acc.re = fmaT(m.re, x.re, acc.re);
    acc.re = fmaT(-m.im, x.im, acc.re);
    acc.im = fmaT(m.re, x.im, acc.im);
    acc.im = fmaT(m.im, x.re, acc.im);
}
template <typename T>
__device__ __forceinline__ ct<T> crowEndT(ct<T> mkk, ct<T> xk, ct<T> acc) {
    return ct<T>{fmaT(mkk.re, xk.re, acc.re), fmaT(mkk.re, xk.im, acc.im)};
}
// [x0; x1] <- [[m00, m01], [m10, m11]] [x0; x1]
template <typename T>
__device__ __forceinline__ void cmat2T(ct<T> m00, ct<T> m01, ct<T> m10, ct<T> m11, ct<T>& x0, ct<T>& x1) {
    ct<T> p = crowStartT(m00, x0), q = crowStartT(m11, x1);
    crowAddT(p, m01, x1);
    crowAddT(q, m10, x0);
    x0 = crowEndT(m00, x0, p);
    x1 = crowEndT(m11, x1, q);
}
// x <- M x for 4 amplitudes, M(i) = element i of the row-major 4x4
template <typename T, typename E>
__device__ __forceinline__ void cmat4T(const E& M, ct<T>& x0, ct<T>& x1, ct<T>& x2, ct<T>& x3) {
    ct<T>* x[4] = {&x0, &x1, &x2, &x3};
    ct<T> acc[4];
#pragma unroll
    for (int k = 0; k < 4; k++) {
        acc[k] = crowStartT(M(5 * k), *x[k]);
#pragma unroll
        for (int j = 0; j < 4; j++)
            if (j != k) crowAddT(acc[k], M(4 * k + j), *x[j]);
    }
#pragma unroll
    for (int k = 0; k < 4; k++) *x[k] = crowEndT(M(5 * k), *x[k], acc[k]);
}

template <typename T>
__device__ __forceinline__ ct<T> ld(const T* __restrict__ amps, uint64_t i) {
    typedef typename Amp<T>::V V;
    V v = reinterpret_cast<const V*>(amps)[i];
    return ct<T>{v.x, v.y};
}
template <typename T>
__device__ __forceinline__ void st(T* __restrict__ amps, uint64_t i, ct<T> a) {
    typedef typename Amp<T>::V V;
    V v;
    v.x = a.re;
    v.y = a.im;
    reinterpret_cast<V*>(amps)[i] = v;
}
// streaming (non-temporal) variants for single-pass kernels with no reuse
template <typename T>
__device__ __forceinline__ ct<T> ldnt(const T* __restrict__ amps, uint64_t i) {
    typedef typename Amp<T>::V V;
    V v = __builtin_nontemporal_load(reinterpret_cast<const V*>(amps) + i);
    return ct<T>{v.x, v.y};
}
template <typename T>
__device__ __forceinline__ void stnt(T* __restrict__ amps, uint64_t i, ct<T> a) {
    typedef typename Amp<T>::V V;
    V v;
    v.x = a.re;
    v.y = a.im;
    __builtin_nontemporal_store(v, reinterpret_cast<V*>(amps) + i);
}

// load/store with the streaming hint chosen at compile time.  Non-temporal
// accesses pay off when every 128-byte line a wave touches is fully used by
// that same instruction; when the lowest touched index bit is below 3 (FP64:
// pairs inside one line) the hint costs bandwidth and plain accesses win
// (measured on MI355X at 30 qubits: 5.5 vs 4.0 TB/s for target 0).
template <int NT, typename T>
__device__ __forceinline__ ct<T> ldx(const T* __restrict__ amps, uint64_t i) {
    if constexpr (NT) return ldnt(amps, i);
    else return ld(amps, i);
}
template <int NT, typename T>
__device__ __forceinline__ void stx(T* __restrict__ amps, uint64_t i, ct<T> a) {
    if constexpr (NT) stnt(amps, i, a);
    else st(amps, i, a);
}
constexpr int STREAM_LOW_BIT = 3;

template <typename T>
__host__ __device__ __forceinline__ ct<T> toT(cd a) {
    return ct<T>{(T)a.re, (T)a.im};
}

// ---- index bit algebra -------------------------------------------------
constexpr int MAXQ = 48;

// Ascending qubit positions whose bits get inserted into a work-item index,
// plus the values those bits take (control states, zero target bits).
struct InsertSpec {
    int num;
    int pos[MAXQ];
    uint64_t valueMask;
};

// insert a zero at every pos[j] (ascending), then OR in valueMask
// (bitwise.hpp:insertBitsWithMaskedValues)
__device__ __forceinline__ uint64_t insertBits(uint64_t n, const InsertSpec& s) {
    for (int j = 0; j < s.num; j++) {
        int q = s.pos[j];
        uint64_t lo = n & ((1ull << q) - 1ull);
        n = ((n >> q) << (q + 1)) | lo;
    }
    return n | s.valueMask;
}

// unsorted target list; bit i of a value <-> targ[i]
struct TargSpec {
    int num;
    int targ[MAXQ];
};

// bitwise.hpp:setBits -- set bits targ[i] of n to bit i of v
__device__ __forceinline__ uint64_t setBitsFrom(uint64_t n, const TargSpec& t, uint64_t v) {
    for (int i = 0; i < t.num; i++) {
        uint64_t b = 1ull << t.targ[i];
        n = ((v >> i) & 1ull) ? (n | b) : (n & ~b);
    }
    return n;
}

// bitwise.hpp:getValueOfBits
__device__ __forceinline__ uint64_t valueOfBits(uint64_t n, const TargSpec& t) {
    uint64_t v = 0;
    for (int i = 0; i < t.num; i++) v |= ((n >> t.targ[i]) & 1ull) << i;
    return v;
}

__device__ __forceinline__ int parity64(uint64_t x) { return __popcll(x) & 1; }

}  // namespace qa
