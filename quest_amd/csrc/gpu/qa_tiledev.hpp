// Tile-pass programs: the device side that applies one pass to one tile,
// shared by the gate-fusion queue (qa_fused.hip), the batched serving engine
// (qa_serve.hip) and the pass kernels the library generates at run time
// (qa_jit.hip, compiled by hiprtc: this header includes nothing but the
// other device headers).
#pragma once

#include "qa_dev.hpp"
#include "qa_kinds.hpp"

namespace qa {
namespace fuse {

// ---- tile-space program -----------------------------------------------------
//
// One pass = one launch of k_tilePass over all 2^(n-k) tiles of the state.
// A tile is the 2^k amplitudes that share the values of the n-k non-local
// qubits; its k local qubits are numbered by "slots" 0..k-1 in ascending
// qubit order, and slots 0..low-1 are the contiguous low qubits.
//
// Inside a tile each thread holds RA = 2^RB amplitudes in registers: those
// whose tile indices differ only in the RB "register slots" of the current
// group (the "layout").  A run of gates whose dense targets all lie in those
// slots is applied in registers; a GROUP op changes the layout by writing
// the registers to an LDS copy of the tile and gathering the new layout.
// When the first (last) group's register slots avoid the low slots, its
// layout is coalesced enough to load (store) HBM directly; otherwise the
// tile goes through LDS in the "row" layout whose register slots are the
// top RB slots.  Workgroups are persistent and prefetch their next tile into
// registers before working on the current one, so HBM traffic overlaps the
// register and LDS work.

constexpr int RB = 4;        // register slots per group
constexpr int RA = 1 << RB;  // amplitudes per thread
enum { F_GROUP = 100, F_BLOCK = 101, F_DENSE1N = 102, F_PAULI_LDS = 103, F_PHASES = 104, F_DENSE_LDS = 105 };
constexpr int BQ = 5;  // qubits of an MFMA block (32 x 32 complex)

// the first 64 bytes of an op, fetched with one scalar load
struct TOpHead {
    int kind;
    int rb1, rb2;            // register bits of the dense targets
    uint32_t rcMask, rcVal;  // control condition, register part (over v)
    uint32_t rm1, rm2;       // register parts of m1 / m2
    uint32_t pad;            // DIAG1 with its target in the registers: 1 = phase-only
    uint64_t gcMask, gcVal;  // control condition, non-local part (tested against the tile base)
    uint64_t jcMask, jcVal;  // control condition, tile part outside the register slots
};
static_assert(sizeof(TOpHead) == 64, "op header is one 64-byte scalar load");

// One op: a 64-byte header, a 64-byte body, both fetched with one 64-byte
// scalar load each, and (for two-qubit dense matrices only) the rest of
// the matrix.  Body by kind:
//   DENSE1, DENSE2: b = matrix elements 0..3 (re, im interleaved)
//   DENSE1V: b = the 2x2 matrix for variant 0, fx = variants 0..3; the
//     variant is the value of the non-local qubits rb2 (bit 0) and pad - 1
//     (bit 1, when pad > 0) at the tile
//   DIAG1, PARITY, PAULI: b[0..3] = two complex factors (d0, d1 / f0, f1 /
//     ampFac, pairFac), b[4] = the tile part of the target bit (DIAG1), the
//     parity mask (PARITY) or the maskYZ bits (PAULI) outside the
//     registers, b[5] = its non-local part
//   DEPOL1: b[0..2] = c1, c2, c3;  DEPH1: b[0] = factor
//   GROUP: b[0..1] = the RB tile slots of the register bits (4 ints);
//     fx holds, as 32-bit words, the LDS byte offsets of its layout's
//     registers (words 0..15, loaded after the exchange) and of the layout
//     before it (words 16..31, stored first)
//   PAULI_LDS, BLOCK: fx words 16..31 = the LDS byte offsets of the
//     registers of the layout before them
struct alignas(64) TOp {
    TOpHead h;
    uint64_t b[8];
    double fx[40];  // DENSE2: matrix elements 4..15
};
static_assert(sizeof(TOp) % 64 == 0, "ops stay 64-byte aligned in arrays");

// PHASES: a run of diagonal gates as terms stored from b onwards; term k
// multiplies every amplitude whose index i has (i & mask) == value by f,
// with the condition split into its register (rm, rv), tile (jm, jv) and
// non-local (gm, gv) parts.  The register part is empty or one register bit
// that must be set; terms come sorted by it (empty first, then bits 0..RB-1)
// with the number of each in a nibble of the head's rm1, so each thread
// multiplies the factors of one slot together first.
struct PhaseTerm {
    uint32_t rm, rv, jm, jv;
    uint64_t gm, gv;
    double fre, fim;
};
static_assert(sizeof(PhaseTerm) == 48, "phase terms pack 8 to an op");
constexpr int MAX_PHASE_TERMS = (int)((sizeof(TOp) - sizeof(TOpHead)) / sizeof(PhaseTerm));

typedef uint32_t u32x16 __attribute__((ext_vector_type(16)));

template <int I>
__device__ __forceinline__ uint64_t u64At(const u32x16& v) {
    return ((uint64_t)v[2 * I + 1] << 32) | (uint64_t)v[2 * I];
}
template <int I>
__device__ __forceinline__ double dAt(const u32x16& v) {
    return __longlong_as_double((long long)u64At<I>(v));
}

// an op as fetched: header fields and body in scalar registers
struct OpRegs {
    TOpHead h;
    double f[8];
    uint64_t jm, gm;
};

__device__ __forceinline__ void fetchOp(const TOp* __restrict__ op, u32x16& h, u32x16& b) {
    h = reinterpret_cast<const u32x16*>(op)[0];
    b = reinterpret_cast<const u32x16*>(op)[1];
}

__device__ __forceinline__ void unpackOp(const u32x16& hv, const u32x16& bv, OpRegs& o) {
    o.h.kind = (int)hv[0];
    o.h.rb1 = (int)hv[1];
    o.h.rb2 = (int)hv[2];
    o.h.rcMask = hv[3];
    o.h.rcVal = hv[4];
    o.h.rm1 = hv[5];
    o.h.rm2 = hv[6];
    o.h.pad = hv[7];
    o.h.gcMask = u64At<4>(hv);
    o.h.gcVal = u64At<5>(hv);
    o.h.jcMask = u64At<6>(hv);
    o.h.jcVal = u64At<7>(hv);
    o.f[0] = dAt<0>(bv);
    o.f[1] = dAt<1>(bv);
    o.f[2] = dAt<2>(bv);
    o.f[3] = dAt<3>(bv);
    o.f[4] = dAt<4>(bv);
    o.f[5] = dAt<5>(bv);
    o.f[6] = dAt<6>(bv);
    o.f[7] = dAt<7>(bv);
    o.jm = u64At<4>(bv);
    o.gm = u64At<5>(bv);
}

struct TileArgs {
    int k;            // tile bits
    int numNonLocal;
    int numOps;
    int firstDirect;  // the first group loads its layout from HBM
    int lastDirect;   // the last group stores its layout to HBM
    int usesLds;      // some layout change goes through LDS
    uint64_t numTiles;
    uint64_t nonLocalMask;  // the non-local qubits as a mask
    uint64_t tileStep;      // the grid size deposited over them (tile-pass kernels: the
                            // base of tile t + gridDim.x is a masked add of it)
    const float* blocks;    // packed block matrices of the pass's BLOCK ops
    int localPos[MAXQ];     // global positions of the tile slots (ascending)
    int nonLocalPos[MAXQ];  // global positions of the other qubits (ascending)
    uint32_t rowOff[RA];    // LDS byte offsets of the registers: row layout
    uint32_t lastOff[RA];   // and the layout of the pass's last group
};

__device__ __forceinline__ uint64_t deposit(uint64_t v, const int* pos, int num) {
    uint64_t r = 0;
    for (int i = 0; i < num; i++) r |= ((v >> i) & 1ull) << pos[i];
    return r;
}

// LDS position of tile amplitude j: the low index bits are XORed with the
// next ones, so threads whose amplitudes are 16 apart (register slots 0..3)
// still spread over all banks
// (complex128: 16 bytes, 16 amplitudes span the 64 four-byte banks; complex64:
// 8 bytes, so 32 amplitudes do, and 5 bits are XORed with the next 5)
template <typename T>
__host__ __device__ __forceinline__ uint32_t swz(uint32_t j) {
    if constexpr (sizeof(T) == 8) return j ^ ((j >> 4) & 15u);
    else return j ^ ((j >> 5) & 31u);
}

// a register layout: the tile slots of the register bits (8 bits each, so
// a layout costs one scalar register), and this thread's tile index (jb)
// and global offset (tg) with the register bits zero
struct RegLayout {
    uint32_t slots;
    uint32_t jb;
    uint64_t tg;
};

__device__ __forceinline__ uint32_t slotAt(const RegLayout& L, int b) { return (L.slots >> (8 * b)) & 0xFFu; }

__device__ __forceinline__ void makeLayout(RegLayout& L, const int* slots, const int* localPos, int k, uint32_t tid) {
    uint32_t gm = 0;
    L.slots = 0;
#pragma unroll
    for (int b = 0; b < RB; b++) {
        L.slots |= (uint32_t)slots[b] << (8 * b);
        gm |= 1u << slots[b];
    }
    // the thread id deposited over the other slots, ascending
    L.jb = 0;
    L.tg = 0;
    for (int s = 0, i = 0; s < k; s++) {
        if ((gm >> s) & 1u) continue;
        const uint32_t bit = (tid >> i) & 1u;
        L.jb |= bit << s;
        L.tg |= (uint64_t)bit << localPos[s];
        i++;
    }
}

template <int V>
__device__ __forceinline__ uint32_t offJ(const RegLayout& L) {
    uint32_t o = 0;
#pragma unroll
    for (int b = 0; b < RB; b++)
        if (V & (1 << b)) o |= 1u << slotAt(L, b);
    return o;
}
// pos: the pass's tile slot -> global qubit map
template <int V>
__device__ __forceinline__ uint64_t offQ(const RegLayout& L, const int* pos) {
    uint64_t o = 0;
#pragma unroll
    for (int b = 0; b < RB; b++)
        if (V & (1 << b)) o |= 1ull << pos[slotAt(L, b)];
    return o;
}

// LDS tile accesses move a whole amplitude per instruction (16 bytes in
// FP64, ds_read_b128 / ds_write_b128): split 8-byte halves would make lanes
// 16 apart meet in the same banks under the row swizzle
template <typename T>
__device__ __forceinline__ ct<T> ldsGet(const ct<T>* tile, uint32_t i) {
    const typename Amp<T>::V x = reinterpret_cast<const typename Amp<T>::V*>(tile)[i];
    return ct<T>{x.x, x.y};
}
template <typename T>
__device__ __forceinline__ void ldsPut(ct<T>* tile, uint32_t i, ct<T> a) {
    typename Amp<T>::V x;
    x.x = a.re;
    x.y = a.im;
    reinterpret_cast<typename Amp<T>::V*>(tile)[i] = x;
}

// byte offset of tile amplitude j in LDS: the swizzle is linear over XOR,
// and a register layout's amplitude indices are jb | offJ<V> with disjoint
// bits, so each thread computes swz(jb) once and every register's offset is
// one XOR with a uniform constant (computed on the host per layout)
template <typename T>
__host__ __device__ __forceinline__ uint32_t ldsByte(uint32_t j) {
    return swz<T>(j) * (uint32_t)sizeof(typename Amp<T>::V);
}
// Register-layout exchanges address LDS by byte offset alone: the tile is
// the pass kernels' only LDS (dynamic, so at LDS address 0 when the kernel
// has no static LDS, which the host checks before launching), and an
// access then needs no add of the array's base
typedef __attribute__((address_space(3))) unsigned char LdsByte;
template <typename T>
__device__ __forceinline__ ct<T> ldsGetB(const ct<T>*, uint32_t byteOff) {
    typedef typename Amp<T>::V V;
    const V x = *reinterpret_cast<const __attribute__((address_space(3))) V*>((LdsByte*)nullptr + byteOff);
    return ct<T>{x.x, x.y};
}
template <typename T>
__device__ __forceinline__ void ldsPutB(ct<T>*, uint32_t byteOff, ct<T> a) {
    typedef typename Amp<T>::V V;
    V x;
    x.x = a.re;
    x.y = a.im;
    *reinterpret_cast<__attribute__((address_space(3))) V*>((LdsByte*)nullptr + byteOff) = x;
}

// the thread's part of a layout's LDS offsets, made opaque at each use so
// that the per-register offsets are formed there, not hoisted out of the
// tile loop into sixteen more live registers
__device__ __forceinline__ uint32_t opaqueV(uint32_t sj) {
    asm volatile("" : "+v"(sj));
    return sj;
}

template <typename T, int NT, int V>
struct RegIO {
    // off: the layout's per-register byte offsets (uniform); sj: this
    // thread's part
    __device__ __forceinline__ static void ldsLoad(ct<T>* r, const ct<T>* tile, uint32_t sj, const uint32_t* off) {
        r[V] = ldsGetB(tile, sj ^ off[V]);
        RegIO<T, NT, V + 1>::ldsLoad(r, tile, sj, off);
    }
    __device__ __forceinline__ static void ldsStore(const ct<T>* r, ct<T>* tile, uint32_t sj, const uint32_t* off) {
        ldsPutB(tile, sj ^ off[V], r[V]);
        RegIO<T, NT, V + 1>::ldsStore(r, tile, sj, off);
    }
    __device__ __forceinline__ static void hbmLoad(ct<T>* r, const T* amps, uint64_t at, const RegLayout& L,
                                                   const int* pos) {
        r[V] = ldx<NT>(amps, at | offQ<V>(L, pos));
        RegIO<T, NT, V + 1>::hbmLoad(r, amps, at, L, pos);
    }
    __device__ __forceinline__ static void hbmStore(const ct<T>* r, T* amps, uint64_t at, const RegLayout& L,
                                                    const int* pos) {
        stx<NT>(amps, at | offQ<V>(L, pos), r[V]);
        RegIO<T, NT, V + 1>::hbmStore(r, amps, at, L, pos);
    }
};
template <typename T, int NT>
struct RegIO<T, NT, RA> {
    __device__ __forceinline__ static void ldsLoad(ct<T>*, const ct<T>*, uint32_t, const uint32_t*) {}
    __device__ __forceinline__ static void ldsStore(const ct<T>*, ct<T>*, uint32_t, const uint32_t*) {}
    __device__ __forceinline__ static void hbmLoad(ct<T>*, const T*, uint64_t, const RegLayout&, const int*) {}
    __device__ __forceinline__ static void hbmStore(const ct<T>*, T*, uint64_t, const RegLayout&, const int*) {}
};

// ---- ops on the registers ------------------------------------------------------
//
// Every op is straight-line arithmetic on all RA registers, with no branch
// around any amplitude.  A control condition removes the op's effect where
// it fails by substituting the identity for the op's matrix (1 for its
// factors, "no exchange" for a swap): per register group when the condition
// is on register bits (uniform over the workgroup, scalar selects), and per
// lane when it is on the tile bits outside the registers (J: vector selects
// on a lane predicate).  The identity leaves every finite amplitude value
// unchanged (x * 1 + 0 * y = x; at most the sign of a zero changes), and
// with no branch each amplitude keeps its registers across ops: a branch
// around an update would make the compiler keep a second copy of all RA
// registers for the path that skips it.

template <typename T>
__device__ __forceinline__ ct<T> cf(const double* f, int i) {
    return ct<T>{(T)f[2 * i], (T)f[2 * i + 1]};
}
template <typename T>
__device__ __forceinline__ ct<T> pickC(bool p, ct<T> x, T re, T im) {
    return ct<T>{p ? x.re : re, p ? x.im : im};
}

// a 2x2 matrix [[a, b], [c, d]]
template <typename T>
struct M2 {
    ct<T> a, b, c, d;
};
template <typename T>
__device__ __forceinline__ M2<T> m2At(const double* f) {
    return M2<T>{cf<T>(f, 0), cf<T>(f, 1), cf<T>(f, 2), cf<T>(f, 3)};
}
// m where p holds, else the identity
template <typename T>
__device__ __forceinline__ M2<T> m2Pick(bool p, const M2<T>& m) {
    return M2<T>{pickC(p, m.a, (T)1, (T)0), pickC(p, m.b, (T)0, (T)0), pickC(p, m.c, (T)0, (T)0),
                 pickC(p, m.d, (T)1, (T)0)};
}

// a 4x4 matrix, row-major; element i is a constant once loops are unrolled
template <typename T>
struct M4 {
    ct<T> e[16];
    __device__ __forceinline__ ct<T> operator()(int i) const { return e[i]; }
};
// row-major elements 0..3 in f (the op body), 4..15 in fx
template <typename T>
__device__ __forceinline__ M4<T> m4At(const double* f, const double* __restrict__ fx) {
    M4<T> m;
#pragma unroll
    for (int i = 0; i < 16; i++) m.e[i] = i < 4 ? cf<T>(f, i) : cf<T>(fx, i - 4);
    return m;
}
// register part of an op's control condition for register group v; C =
// false when the op has no control among the register bits
template <bool C>
__device__ __forceinline__ bool regCond(int v, const TOpHead& h) {
    return !C || (((uint32_t)v & h.rcMask) == h.rcVal);
}

// the arithmetic of each gate below is the expression of its single-gate
// kernel (qa_gates.hip / qa_channels.hip), term for term

template <typename T, int B, bool C>
__device__ __forceinline__ void regDense1(ct<T>* r, const TOpHead& h, const M2<T>& m) {
#pragma unroll
    for (int v = 0; v < RA; v++) {
        if (v & (1 << B)) continue;
        const M2<T> mv = C ? m2Pick(regCond<C>(v, h), m) : m;
        cmat2T(mv.a, mv.b, mv.c, mv.d, r[v], r[v | (1 << B)]);
    }
}

// where a condition fails (register group v for C, this lane for J) the
// quad keeps its values: a 4x4 matrix is too large to substitute
template <typename T, int B1, int B2, bool C, bool J>
__device__ __forceinline__ void regDense2(ct<T>* r, const TOpHead& h, const M4<T>& m, bool act) {
    constexpr int off[4] = {0, 1 << B1, 1 << B2, (1 << B1) | (1 << B2)};
#pragma unroll
    for (int v = 0; v < RA; v++) {
        if (v & ((1 << B1) | (1 << B2))) continue;
        if constexpr (C || J) {
            ct<T> y[4] = {r[v | off[0]], r[v | off[1]], r[v | off[2]], r[v | off[3]]};
            cmat4T<T>(m, y[0], y[1], y[2], y[3]);
            const bool p = regCond<C>(v, h) && (!J || act);
#pragma unroll
            for (int i = 0; i < 4; i++) r[v | off[i]] = pickC(p, y[i], r[v | off[i]].re, r[v | off[i]].im);
        } else {
            cmat4T<T>(m, r[v | off[0]], r[v | off[1]], r[v | off[2]], r[v | off[3]]);
        }
    }
}

// act: the lane predicate (true for every lane unless J)
template <typename T, int B1, int B2, bool C, bool J>
__device__ __forceinline__ void regSwap(ct<T>* r, const TOpHead& h, bool act) {
#pragma unroll
    for (int v = 0; v < RA; v++) {
        if (!(v & (1 << B1)) || (v & (1 << B2))) continue;
        constexpr int flip = (1 << B1) | (1 << B2);
        const ct<T> x = r[v], y = r[v ^ flip];
        if constexpr (C || J) {
            const bool p = regCond<C>(v, h) && (!J || act);
            r[v] = ct<T>{p ? y.re : x.re, p ? y.im : x.im};
            r[v ^ flip] = ct<T>{p ? x.re : y.re, p ? x.im : y.im};
        } else {
            r[v] = y;
            r[v ^ flip] = x;
        }
    }
}

// one-qubit depolarising / dephasing on (ket, bra): both formulas are
// symmetric in the two bits, so B1 < B2 covers either order (never
// controlled)
template <typename T, int B1, int B2>
__device__ __forceinline__ void regDepol1(ct<T>* r, const double* f) {
    const T c1 = (T)f[0], c2 = (T)f[1], c3 = (T)f[2];
#pragma unroll
    for (int v = 0; v < RA; v++) {
        if (v & ((1 << B1) | (1 << B2))) continue;
        constexpr int k1 = 1 << B1, k2 = 1 << B2;
        const ct<T> a00 = r[v], a01 = r[v | k1], a10 = r[v | k2], a11 = r[v | k1 | k2];
        r[v] = ct<T>{fmaT(c1, a00.re, c2 * a11.re), fmaT(c1, a00.im, c2 * a11.im)};
        r[v | k1] = ct<T>{c3 * a01.re, c3 * a01.im};
        r[v | k2] = ct<T>{c3 * a10.re, c3 * a10.im};
        r[v | k1 | k2] = ct<T>{fmaT(c1, a11.re, c2 * a00.re), fmaT(c1, a11.im, c2 * a00.im)};
    }
}

template <typename T, int B1, int B2>
__device__ __forceinline__ void regDeph1(ct<T>* r, const double* f) {
    const T fac = (T)f[0];
#pragma unroll
    for (int v = 0; v < RA; v++) {
        if (v & ((1 << B1) | (1 << B2))) continue;
        constexpr int k1 = 1 << B1, k2 = 1 << B2;
        const ct<T> x = r[v | k1], y = r[v | k2];
        r[v | k1] = ct<T>{x.re * fac, x.im * fac};
        r[v | k2] = ct<T>{y.re * fac, y.im * fac};
    }
}

// Pauli tensor / gadget with X/Y bits MXY (all in registers); thrPar is the
// parity of the maskYZ bits outside the registers for this thread.  Where
// the condition fails the pair gets ampFac = 1, pairFac = 0.
template <typename T, int MXY, bool C, bool J>
__device__ __forceinline__ void regPauli(ct<T>* r, const TOpHead& h, const double* f, int thrPar, bool act) {
    constexpr int top = 31 - __builtin_clz((unsigned)MXY);
    const ct<T> ampFac = cf<T>(f, 0), pairFac = cf<T>(f, 1);
#pragma unroll
    for (int v = 0; v < RA; v++) {
        if (v & (1 << top)) continue;
        const int w = v ^ MXY;
        const bool p = regCond<C>(v, h) && (!J || act);
        const ct<T> af = (C || J) ? pickC(p, ampFac, (T)1, (T)0) : ampFac;
        const ct<T> pf = (C || J) ? pickC(p, pairFac, (T)0, (T)0) : pairFac;
        const T sA = ((__popc((uint32_t)v & h.rm2) & 1) ^ thrPar) ? (T)-1 : (T)1;
        const T sB = ((__popc((uint32_t)w & h.rm2) & 1) ^ thrPar) ? (T)-1 : (T)1;
        const ct<T> pB = {pf.re * sB, pf.im * sB};
        const ct<T> pA = {pf.re * sA, pf.im * sA};
        cmat2T(af, pB, pA, af, r[v], r[w]);
    }
}

// one-target diagonal whose target is register bit B; PH: phase-only (the
// amplitudes with bit B clear are left as they are, as in the single-gate
// kernel, which never touches them)
template <typename T, int B, bool C, bool PH>
__device__ __forceinline__ void regDiag1R(ct<T>* r, const TOpHead& h, ct<T> d0, ct<T> d1) {
#pragma unroll
    for (int v = 0; v < RA; v++) {
        if (PH && !(v & (1 << B))) continue;
        const ct<T> d = (v & (1 << B)) ? d1 : d0;
        r[v] = cmulT(C ? pickC(regCond<C>(v, h), d, (T)1, (T)0) : d, r[v]);
    }
}

// a diagonal factor per lane for all of this thread's amplitudes (target
// outside the registers, or a parity over them)
template <typename T, bool C>
__device__ __forceinline__ void regScale(ct<T>* r, const TOpHead& h, ct<T> d) {
#pragma unroll
    for (int v = 0; v < RA; v++) r[v] = cmulT(C ? pickC(regCond<C>(v, h), d, (T)1, (T)0) : d, r[v]);
}

template <typename T, bool C>
__device__ __forceinline__ void regParity(ct<T>* r, const TOpHead& h, ct<T> f0, ct<T> f1, int thrPar) {
#pragma unroll
    for (int v = 0; v < RA; v++) {
        const ct<T> f = ((__popc((uint32_t)v & h.rm1) & 1) ^ thrPar) ? f1 : f0;
        r[v] = cmulT(C ? pickC(regCond<C>(v, h), f, (T)1, (T)0) : f, r[v]);
    }
}

#define QA_SWAP_CASES                                                 \
    case 0 * RB + 1: regSwap<T, 0, 1, C, J>(r, h, act); break;        \
    case 0 * RB + 2: regSwap<T, 0, 2, C, J>(r, h, act); break;        \
    case 0 * RB + 3: regSwap<T, 0, 3, C, J>(r, h, act); break;        \
    case 1 * RB + 2: regSwap<T, 1, 2, C, J>(r, h, act); break;        \
    case 1 * RB + 3: regSwap<T, 1, 3, C, J>(r, h, act); break;        \
    case 2 * RB + 3: regSwap<T, 2, 3, C, J>(r, h, act); break;

#define QA_CHAN_CASES(FN)                                 \
    case 0 * RB + 1: FN<T, 0, 1>(r, f); break;           \
    case 0 * RB + 2: FN<T, 0, 2>(r, f); break;           \
    case 0 * RB + 3: FN<T, 0, 3>(r, f); break;           \
    case 1 * RB + 2: FN<T, 1, 2>(r, f); break;           \
    case 1 * RB + 3: FN<T, 1, 3>(r, f); break;           \
    case 2 * RB + 3: FN<T, 2, 3>(r, f); break;

template <typename T, bool C>
__device__ __forceinline__ void dense1On(ct<T>* r, const TOpHead& h, int b, const M2<T>& m) {
    switch (b) {
        case 0: regDense1<T, 0, C>(r, h, m); break;
        case 1: regDense1<T, 1, C>(r, h, m); break;
        case 2: regDense1<T, 2, C>(r, h, m); break;
        case 3: regDense1<T, 3, C>(r, h, m); break;
    }
}

// applies one op that may carry controls to the registers of one thread;
// C = the op has controls among the register bits, J = among the tile bits
// outside them (act is then this lane's part of the condition)
template <typename T, bool C, bool J>
__device__ __forceinline__ void applyOpC(ct<T>* r, const TOp* __restrict__ op, const OpRegs& o, uint32_t jb,
                                         uint64_t base, bool act) {
    const TOpHead& h = o.h;
    const double* f = o.f;
    switch (h.kind) {
        case F_DENSE1: {
            const M2<T> m = m2At<T>(f);
            dense1On<T, C>(r, h, h.rb1, J ? m2Pick(act, m) : m);
            break;
        }
        case F_DENSE1V: {
            // a 2x2 matrix per value of one or two non-local qubits, all
            // four in fx: the tile's value picks it (uniform over the
            // workgroup)
            uint32_t w = (uint32_t)((base >> h.rb2) & 1ull);
            if (h.pad) w |= (uint32_t)((base >> (h.pad - 1)) & 1ull) << 1;
            const M2<T> m = m2At<T>(op->fx + 8 * w);
            dense1On<T, C>(r, h, h.rb1, J ? m2Pick(act, m) : m);
            break;
        }
        case F_DENSE2: {
            const M4<T> m = m4At<T>(f, op->fx);
            switch (h.rb1 * RB + h.rb2) {
#define QA_D2(B1, B2) case B1 * RB + B2: regDense2<T, B1, B2, C, J>(r, h, m, act); break;
                QA_D2(0, 1) QA_D2(0, 2) QA_D2(0, 3) QA_D2(1, 2) QA_D2(1, 3) QA_D2(2, 3)
                QA_D2(1, 0) QA_D2(2, 0) QA_D2(3, 0) QA_D2(2, 1) QA_D2(3, 1) QA_D2(3, 2)
#undef QA_D2
            }
            break;
        }
        case F_SWAP:
            switch (h.rb1 * RB + h.rb2) { QA_SWAP_CASES }
            break;
        case F_DIAG1: {
            ct<T> d0 = cf<T>(f, 0), d1 = cf<T>(f, 1);
            if (J) {
                d0 = pickC(act, d0, (T)1, (T)0);
                d1 = pickC(act, d1, (T)1, (T)0);
            }
            if (h.rb1 < 0) {
                regScale<T, C>(r, h, ((jb & o.jm) | (base & o.gm)) != 0 ? d1 : d0);
            } else if (h.pad) {
                switch (h.rb1) {
                    case 0: regDiag1R<T, 0, C, true>(r, h, d0, d1); break;
                    case 1: regDiag1R<T, 1, C, true>(r, h, d0, d1); break;
                    case 2: regDiag1R<T, 2, C, true>(r, h, d0, d1); break;
                    case 3: regDiag1R<T, 3, C, true>(r, h, d0, d1); break;
                }
            } else {
                switch (h.rb1) {
                    case 0: regDiag1R<T, 0, C, false>(r, h, d0, d1); break;
                    case 1: regDiag1R<T, 1, C, false>(r, h, d0, d1); break;
                    case 2: regDiag1R<T, 2, C, false>(r, h, d0, d1); break;
                    case 3: regDiag1R<T, 3, C, false>(r, h, d0, d1); break;
                }
            }
            break;
        }
        case F_PARITY: {
            ct<T> f0 = cf<T>(f, 0), f1 = cf<T>(f, 1);
            if (J) {
                f0 = pickC(act, f0, (T)1, (T)0);
                f1 = pickC(act, f1, (T)1, (T)0);
            }
            regParity<T, C>(r, h, f0, f1, parity64((jb & o.jm) ^ (base & o.gm)));
            break;
        }
        case F_PAULI: {
            const int thrPar = parity64((jb & o.jm) ^ (base & o.gm));
            switch (h.rm1) {
#define QA_PAULI_CASE(M) case M: regPauli<T, M, C, J>(r, h, f, thrPar, act); break;
                QA_PAULI_CASE(1) QA_PAULI_CASE(2) QA_PAULI_CASE(3) QA_PAULI_CASE(4) QA_PAULI_CASE(5)
                QA_PAULI_CASE(6) QA_PAULI_CASE(7) QA_PAULI_CASE(8) QA_PAULI_CASE(9) QA_PAULI_CASE(10)
                QA_PAULI_CASE(11) QA_PAULI_CASE(12) QA_PAULI_CASE(13) QA_PAULI_CASE(14) QA_PAULI_CASE(15)
#undef QA_PAULI_CASE
            }
            break;
        }
    }
}

// a run of diagonal gates (PHASES): per-thread factors A for every
// amplitude and B[b] for those with register bit b set (the terms come
// sorted by that slot, with the count of each slot in a nibble of rm1),
// multiplied into one factor per register (subset products) before they
// reach the amplitudes

// one term: acc <- (its condition holds ? its factor : 1) * acc
template <typename T>
__device__ __forceinline__ void phaseTerm(ct<T>& acc, bool on, double fre, double fim) {
    acc = cmulT(ct<T>{on ? (T)fre : (T)1, on ? (T)fim : (T)0}, acc);
}

// acc[0] for every register, acc[1 + b] for registers with bit b set
template <typename T>
__device__ __forceinline__ void phasesFinish(ct<T>* r, const ct<T>* acc) {
    ct<T> fac[RA];
    fac[0] = acc[0];
#pragma unroll
    for (int b = 0; b < RB; b++)
#pragma unroll
        for (int v = 0; v < (1 << b); v++) fac[v | (1 << b)] = cmulT(acc[1 + b], fac[v]);
#pragma unroll
    for (int v = 0; v < RA; v++) r[v] = cmulT(fac[v], r[v]);
}

template <typename T>
__device__ __forceinline__ void applyPhases(ct<T>* r, const TOp* __restrict__ op, const TOpHead& h, uint32_t jb,
                                            uint64_t base) {
    const PhaseTerm* pt = reinterpret_cast<const PhaseTerm*>(op->b);
    ct<T> acc[1 + RB];
    int k = 0;
#pragma unroll
    for (int sl = 0; sl <= RB; sl++) {
        acc[sl] = ct<T>{(T)1, (T)0};
        const int ns = (int)((h.rm1 >> (4 * sl)) & 15u);
        for (int e = 0; e < ns; e++, k++) {
            const PhaseTerm t = pt[k];
            phaseTerm(acc[sl], (jb & t.jm) == t.jv && (base & t.gm) == t.gv, t.fre, t.fim);
        }
    }
    phasesFinish<T>(r, acc);
}

// ops that never carry controls (PHASES, DENSE1N, the channels) are applied
// here; the others by applyOpC
template <typename T>
__device__ __forceinline__ void applyOp(ct<T>* r, const TOp* __restrict__ op, const OpRegs& o, uint32_t jb,
                                        uint64_t base) {
    const TOpHead& h = o.h;
    const double* f = o.f;
    switch (h.kind) {
        case F_PHASES:
            applyPhases<T>(r, op, h, jb, base);
            break;
        case F_DENSE1N: {
            // N = rb2 uncontrolled one-qubit matrices on register bits 0..N-1
            // in that order: straight-line code, one dispatch for N gates
            const double* fx = op->fx;
            switch (h.rb2) {
                case 2:
                    regDense1<T, 0, false>(r, h, m2At<T>(f));
                    regDense1<T, 1, false>(r, h, m2At<T>(fx));
                    break;
                case 3:
                    regDense1<T, 0, false>(r, h, m2At<T>(f));
                    regDense1<T, 1, false>(r, h, m2At<T>(fx));
                    regDense1<T, 2, false>(r, h, m2At<T>(fx + 8));
                    break;
                case 4:
                    regDense1<T, 0, false>(r, h, m2At<T>(f));
                    regDense1<T, 1, false>(r, h, m2At<T>(fx));
                    regDense1<T, 2, false>(r, h, m2At<T>(fx + 8));
                    regDense1<T, 3, false>(r, h, m2At<T>(fx + 16));
                    break;
            }
            break;
        }
        case F_DEPOL1:
            switch (h.rb1 * RB + h.rb2) { QA_CHAN_CASES(regDepol1) }
            break;
        case F_DEPH1:
            switch (h.rb1 * RB + h.rb2) { QA_CHAN_CASES(regDeph1) }
            break;
        default:
            if (h.rcMask) applyOpC<T, true, false>(r, op, o, jb, base, true);
            else applyOpC<T, false, false>(r, op, o, jb, base, true);
    }
}

// an op whose condition involves the tile bits outside the registers: act
// is this lane's part of it
template <typename T>
__device__ __forceinline__ void applyOpLanes(ct<T>* r, const TOp* __restrict__ op, const OpRegs& o, uint32_t jb,
                                             uint64_t base, bool act) {
    if (o.h.rcMask) applyOpC<T, true, true>(r, op, o, jb, base, act);
    else applyOpC<T, false, true>(r, op, o, jb, base, act);
}

// One op with its kind, register bits (RB1, RB2), pad and register mask
// (RM1: Pauli X/Y bits, PHASES slot counts) as constants: the arm of
// applyOp / applyOpC for that op, without the dispatch (for generated
// kernels, which instantiate only the arms their program uses)
template <typename T, int KIND, int RB1, int RB2, unsigned PAD, unsigned RM1, bool C, bool J>
__device__ __forceinline__ void applyFixed(ct<T>* r, const TOp* __restrict__ op, const OpRegs& o, uint32_t jb,
                                           uint64_t base, bool act) {
    const TOpHead& h = o.h;
    const double* f = o.f;
    if constexpr (KIND == F_DENSE1) {
        const M2<T> m = m2At<T>(f);
        regDense1<T, RB1, C>(r, h, J ? m2Pick(act, m) : m);
    } else if constexpr (KIND == F_DENSE1V) {
        uint32_t w = (uint32_t)((base >> RB2) & 1ull);
        if constexpr (PAD != 0) w |= (uint32_t)((base >> (PAD - 1)) & 1ull) << 1;
        const M2<T> m = m2At<T>(op->fx + 8 * w);
        regDense1<T, RB1, C>(r, h, J ? m2Pick(act, m) : m);
    } else if constexpr (KIND == F_DENSE2) {
        regDense2<T, RB1, RB2, C, J>(r, h, m4At<T>(f, op->fx), act);
    } else if constexpr (KIND == F_SWAP) {
        regSwap<T, RB1, RB2, C, J>(r, h, act);
    } else if constexpr (KIND == F_DIAG1) {
        ct<T> d0 = cf<T>(f, 0), d1 = cf<T>(f, 1);
        if (J) {
            d0 = pickC(act, d0, (T)1, (T)0);
            d1 = pickC(act, d1, (T)1, (T)0);
        }
        if constexpr (RB1 < 0) regScale<T, C>(r, h, ((jb & o.jm) | (base & o.gm)) != 0 ? d1 : d0);
        else regDiag1R<T, RB1, C, PAD != 0>(r, h, d0, d1);
    } else if constexpr (KIND == F_PARITY) {
        ct<T> f0 = cf<T>(f, 0), f1 = cf<T>(f, 1);
        if (J) {
            f0 = pickC(act, f0, (T)1, (T)0);
            f1 = pickC(act, f1, (T)1, (T)0);
        }
        regParity<T, C>(r, h, f0, f1, parity64((jb & o.jm) ^ (base & o.gm)));
    } else if constexpr (KIND == F_PAULI) {
        regPauli<T, (int)RM1, C, J>(r, h, f, parity64((jb & o.jm) ^ (base & o.gm)), act);
    } else if constexpr (KIND == F_PHASES) {
        applyPhases<T>(r, op, h, jb, base);
    } else if constexpr (KIND == F_DENSE1N) {
        const double* fx = op->fx;
        regDense1<T, 0, false>(r, h, m2At<T>(f));
        regDense1<T, 1, false>(r, h, m2At<T>(fx));
        if constexpr (RB2 >= 3) regDense1<T, 2, false>(r, h, m2At<T>(fx + 8));
        if constexpr (RB2 >= 4) regDense1<T, 3, false>(r, h, m2At<T>(fx + 16));
    } else if constexpr (KIND == F_DEPOL1) {
        regDepol1<T, RB1, RB2>(r, f);
    } else if constexpr (KIND == F_DEPH1) {
        regDeph1<T, RB1, RB2>(r, f);
    }
}

__host__ __device__ __forceinline__ const int* groupSlots(const TOp& op) { return reinterpret_cast<const int*>(op.b); }
__host__ __device__ __forceinline__ const uint32_t* loadOffsets(const TOp* op) {
    return reinterpret_cast<const uint32_t*>(op->fx);
}
__host__ __device__ __forceinline__ const uint32_t* storeOffsets(const TOp* op) {
    return reinterpret_cast<const uint32_t*>(op->fx) + RA;
}

// A Pauli tensor / gadget with more X/Y qubits than register bits, applied
// to the tile in LDS: every thread takes pairs (jA, jA ^ xy) with the top
// X/Y bit of jA clear; same arithmetic as the single-gate kernel
template <typename T>
__device__ __forceinline__ void ldsPauli(ct<T>* tile, int k, const OpRegs& o, uint64_t xy, uint64_t base,
                                         uint32_t tid, uint32_t nt) {
    const int top = 63 - __clzll((long long)xy);
    const ct<T> ampFac = cf<T>(o.f, 0), pairFac = cf<T>(o.f, 1);
    const int basePar = parity64(base & o.gm);
    const uint32_t half = 1u << (k - 1);
    for (uint32_t q = tid; q < half; q += nt) {
        const uint32_t jA = ((q >> top) << (top + 1)) | (q & ((1u << top) - 1u));
        if ((jA & o.h.jcMask) != o.h.jcVal) continue;
        const uint32_t jB = jA ^ (uint32_t)xy;
        ct<T> va = ldsGet(tile, swz<T>(jA)), vb = ldsGet(tile, swz<T>(jB));
        const T sA = (parity64(jA & o.jm) ^ basePar) ? (T)-1 : (T)1;
        const T sB = (parity64(jB & o.jm) ^ basePar) ? (T)-1 : (T)1;
        const ct<T> pB = {pairFac.re * sB, pairFac.im * sB};
        const ct<T> pA = {pairFac.re * sA, pairFac.im * sA};
        cmat2T(ampFac, pB, pA, ampFac, va, vb);
        ldsPut(tile, swz<T>(jA), va);
        ldsPut(tile, swz<T>(jB), vb);
    }
}

// ---- FP32 MFMA blocks ---------------------------------------------------------
//
// A BLOCK op applies a 32x32 complex matrix U to the 5 tile slots in its
// body (b[0..2] as ints; b[3] = offset of its packed matrix in
// TileArgs::blocks) for every column of the tile (the 32 amplitudes that
// share the other tile bits).  In real form Y = [[Re U, -Im U], [Im U, Re U]]
// [Re X; Im X], 64 x 64 times 64 x 32 columns per wave, as 32 K-steps of
// v_mfma_f32_32x32x2f32 for each of the two 32-row halves.  Lane l holds
// column l%32 of the wave's 32 columns; for K-step s it supplies
// A[l%32 (+32)][2s + l/32] (packed host-side at [s][half][lane]) and
// X[2s + l/32][l%32]; result register i of lane l is row 8(i/4) + 4(l/32) +
// i%4.  The tile is in LDS; waves = columns / 32 exactly (tiles of 2^k,
// k >= 10, with 2^(k-4) threads).
typedef float f32x16 __attribute__((ext_vector_type(16)));

__device__ __forceinline__ void runBlock32(ct<float>* tile, int k, const int* slots, const float* __restrict__ A,
                                           uint32_t tid) {
    const uint32_t lane = tid & 63u, h = lane >> 5;
    const uint32_t col = (tid >> 6) * 32u + (lane & 31u);
    uint32_t sb[BQ], sm = 0;
#pragma unroll
    for (int b = 0; b < BQ; b++) {
        sb[b] = 1u << slots[b];
        sm |= sb[b];
    }
    uint32_t base = 0;
    for (int s = 0, i = 0; s < k; s++) {
        if ((sm >> s) & 1u) continue;
        base |= ((col >> i) & 1u) << s;
        i++;
    }
    const uint32_t hb = h ? sb[0] : 0u;
    float xr[16], xi[16];
#pragma unroll
    for (int t = 0; t < 16; t++) {
        // amplitude 2t + h of the column
        uint32_t j = base | hb;
#pragma unroll
        for (int b = 1; b < BQ; b++)
            if ((t >> (b - 1)) & 1) j |= sb[b];
        const ct<float> v = ldsGet(tile, swz<float>(j));
        xr[t] = v.re;
        xi[t] = v.im;
    }
    f32x16 dTop = {0}, dBot = {0};
#pragma unroll
    for (int s = 0; s < 32; s++) {
        const float aTop = A[(2 * s) * 64 + lane], aBot = A[(2 * s + 1) * 64 + lane];
        const float x = s < 16 ? xr[s] : xi[s - 16];
        dTop = __builtin_amdgcn_mfma_f32_32x32x2f32(aTop, x, dTop, 0, 0, 0);
        dBot = __builtin_amdgcn_mfma_f32_32x32x2f32(aBot, x, dBot, 0, 0, 0);
    }
#pragma unroll
    for (int i = 0; i < 16; i++) {
        const int rowConst = 8 * (i / 4) + (i % 4);  // plus 4h
        uint32_t j = base;
#pragma unroll
        for (int b = 0; b < BQ; b++)
            if ((rowConst >> b) & 1) j |= sb[b];
        if (h) j |= sb[2];  // the 4h term is bit 2 of the row
        ldsPut(tile, swz<float>(j), ct<float>{dTop[i], dBot[i]});
    }
}

// the tile's load layout: the first group's when it loads HBM directly,
// else the row layout whose register slots are the top RB slots
__device__ __forceinline__ void passLayouts(RegLayout& row, RegLayout& first, const TileArgs& a,
                                            const TOp* __restrict__ ops, uint32_t tid) {
    int rowSlots[RB];
#pragma unroll
    for (int b = 0; b < RB; b++) rowSlots[b] = a.k - RB + b;
    makeLayout(row, rowSlots, a.localPos, a.k, tid);
    makeLayout(first, groupSlots(ops[0]), a.localPos, a.k, tid);
}

// ---- one tile of a pass ------------------------------------------------------
//
// A pass program runs on a tile as a sequence of steps.  The interpreter
// (tileBody) takes each op's kind from the program at run time; the pass
// kernels generated at run time (qa_jit.hip) call the same steps in program
// order with each op's kind, register bits and register masks written in as
// constants, so both do the same arithmetic in the same order.

struct TileState {
    RegLayout cur;  // the layout the registers hold (or held last)
    uint32_t sj;    // this thread's part of cur's LDS offsets
    bool inRegs;    // the registers hold `cur` (else the tile is in LDS)
};

// r holds the tile in its load layout: the first group's when it loads
// HBM directly, else the row layout, which goes to LDS first
template <typename T, int NT>
__device__ __forceinline__ void tileBegin(TileState& st, const TileArgs& a, ct<T>* r, ct<T>* tile,
                                          const RegLayout& row, const RegLayout& first) {
    st.cur = first;
    st.sj = ldsByte<T>(first.jb);
    st.inRegs = a.firstDirect;
    if (!a.firstDirect) RegIO<T, NT, 0>::ldsStore(r, tile, opaqueV(ldsByte<T>(row.jb)), a.rowOff);
}

// a GROUP op: the registers go to LDS in their layout and come back in the
// layout whose register bits are the tile slots `slots`
template <typename T, int NT>
__device__ __forceinline__ void stepGroup(TileState& st, const TileArgs& a, const TOp* __restrict__ op,
                                          const int* slots, ct<T>* r, ct<T>* tile, uint32_t tid) {
    if (st.inRegs) RegIO<T, NT, 0>::ldsStore(r, tile, opaqueV(st.sj), storeOffsets(op));
    __syncthreads();
    makeLayout(st.cur, slots, a.localPos, a.k, tid);
    st.sj = ldsByte<T>(st.cur.jb);
    RegIO<T, NT, 0>::ldsLoad(r, tile, opaqueV(st.sj), loadOffsets(op));
    st.inRegs = true;
}

// a PAULI_LDS op (xy: its X/Y bits in tile indices)
template <typename T, int NT>
__device__ __forceinline__ void stepPauliLds(TileState& st, const TileArgs& a, const TOp* __restrict__ op,
                                             const OpRegs& o, uint64_t xy, ct<T>* r, ct<T>* tile, uint64_t base,
                                             uint32_t tid) {
    if (st.inRegs) RegIO<T, NT, 0>::ldsStore(r, tile, opaqueV(st.sj), storeOffsets(op));
    __syncthreads();
    if ((base & o.h.gcMask) == o.h.gcVal) ldsPauli<T>(tile, a.k, o, xy, base, tid, blockDim.x);
    __syncthreads();
    st.inRegs = false;
}

// a DENSE_LDS op: a dense 2^N x 2^N matrix (N = 3 or 4, row-major {re, im}
// in device memory, imaginary parts negated when conj) on the tile slots
// slots[0..N) (matrix bit b <-> slots[b]), applied to the tile in LDS.
// Generated kernels only: in the interpreter its registers would push the
// op loop into scratch memory, so the host runs every pass holding one
// through its generated kernel.
// Each thread takes the groups of 2^N amplitudes that share the other tile
// bits (one group for N = 4, two for N = 3) into its registers and writes
// each row's sum back as it is formed, with the arithmetic of the
// single-gate kernel (k_fewTargDense: terms accumulated from zero in
// column order), so results are bit-identical to it.
template <typename T, int NT, int N>
__device__ __forceinline__ void stepDenseLds(TileState& st, const TOp* __restrict__ op, const uint32_t* slots,
                                             const double* __restrict__ mat, int conj, int k, ct<T>* r,
                                             ct<T>* tile, uint32_t tid) {
    constexpr int D = 1 << N;
    if (st.inRegs) RegIO<T, NT, 0>::ldsStore(r, tile, opaqueV(st.sj), storeOffsets(op));
    __syncthreads();
    uint32_t tmask = 0;
#pragma unroll
    for (int b = 0; b < N; b++) tmask |= 1u << slots[b];
    const double sgn = conj ? -1.0 : 1.0;
    // the matrix read by vector loads (a per-lane offset the compiler cannot
    // see is zero): as scalar loads its 2^(2N) elements would crowd the
    // scalar registers of the whole op loop
    const double2* __restrict__ mrow = reinterpret_cast<const double2*>(mat) + opaqueV(0u);
    const uint32_t groups = 1u << (k - N);
    for (uint32_t g = tid; g < groups; g += blockDim.x) {
        // g with zeros inserted at the target slots
        uint32_t j0 = 0;
        for (int bit = 0, src = 0; bit < k; bit++) {
            if ((tmask >> bit) & 1u) continue;
            j0 |= ((g >> src) & 1u) << bit;
            src++;
        }
        uint32_t sb[N];
#pragma unroll
        for (int b = 0; b < N; b++) sb[b] = 1u << slots[b];
#pragma unroll
        for (int v = 0; v < D; v++) {
            uint32_t o = j0;
#pragma unroll
            for (int b = 0; b < N; b++)
                if ((v >> b) & 1) o |= sb[b];
            r[v] = ldsGet(tile, swz<T>(o));
        }
#pragma unroll 1
        for (int row = 0; row < D; row++) {
            ct<T> acc = {0, 0};
#pragma unroll
            for (int c = 0; c < D; c++) {
                const double2 m = mrow[row * D + c];
                const ct<T> e = {(T)m.x, (T)(sgn * m.y)};
                acc = cfmaT(e, r[c], acc);
            }
            uint32_t o = j0;
#pragma unroll
            for (int b = 0; b < N; b++)
                if ((row >> b) & 1) o |= sb[b];
            ldsPut(tile, swz<T>(o), acc);
        }
    }
    __syncthreads();
    st.inRegs = false;
}

// a BLOCK op (complex64 only): bslots = its 5 tile slots, A = its packed matrix
template <typename T, int NT>
__device__ __forceinline__ void stepBlock(TileState& st, const TOp* __restrict__ op, const int* bslots,
                                          const float* __restrict__ A, int k, ct<T>* r, ct<T>* tile, uint32_t tid) {
    if constexpr (sizeof(T) == 4) {
        if (st.inRegs) RegIO<T, NT, 0>::ldsStore(r, tile, opaqueV(st.sj), storeOffsets(op));
        __syncthreads();
        runBlock32(reinterpret_cast<ct<float>*>(tile), k, bslots, A, tid);
        __syncthreads();
        st.inRegs = false;
    }
}

// a register op: the non-local part of its condition is uniform over the
// workgroup (the whole tile skips it), the part on tile bits outside the
// registers is the same for all of a thread's registers (a lane predicate)
template <typename T>
__device__ __forceinline__ void stepApply(const TileState& st, const TOp* __restrict__ op, const OpRegs& o, ct<T>* r,
                                          uint64_t base) {
    if ((base & o.h.gcMask) != o.h.gcVal) return;
    if (!o.h.jcMask) {
        applyOp<T>(r, op, o, st.cur.jb, base);
        return;
    }
    const bool act = (st.cur.jb & o.h.jcMask) == o.h.jcVal;
    if (__builtin_amdgcn_ballot_w64(act)) applyOpLanes<T>(r, op, o, st.cur.jb, base, act);
}

// stepApply for an op of known shape (generated kernels): GC / J = its
// condition has a non-local part / a part on tile bits outside the
// registers, C = a part on register bits
template <typename T, int KIND, int RB1, int RB2, unsigned PAD, unsigned RM1, bool GC, bool J, bool C>
__device__ __forceinline__ void stepFixed(const TileState& st, const TOp* __restrict__ op, const OpRegs& o, ct<T>* r,
                                          uint64_t base) {
    if constexpr (GC) {
        if ((base & o.h.gcMask) != o.h.gcVal) return;
    }
    if constexpr (J) {
        const bool act = (st.cur.jb & o.h.jcMask) == o.h.jcVal;
        if (__builtin_amdgcn_ballot_w64(act))
            applyFixed<T, KIND, RB1, RB2, PAD, RM1, C, true>(r, op, o, st.cur.jb, base, act);
    } else {
        applyFixed<T, KIND, RB1, RB2, PAD, RM1, C, false>(r, op, o, st.cur.jb, base, true);
    }
}

// Waits for every vector memory load in flight -- in a prefetching pass
// loop, the next tile's -- before this tile's stores are issued.  Loads and
// stores share one counter that retires in issue order, so a wait for the
// prefetched registers placed after the stores (where the compiler would put
// it, at the next iteration's first use) waits for the stores as well, and
// the tile's store traffic would never overlap the next tile's work.  With
// the loads retired here, the stores drain while the next tile computes.
__device__ __forceinline__ void waitLoadsBeforeStores() {
    __builtin_amdgcn_s_waitcnt(0x0F70);  // vmcnt(0); expcnt, lgkmcnt unchanged
}

// the tile back to HBM: directly in the last group's layout, or through
// LDS in the row layout
template <typename T, int NT>
__device__ __forceinline__ void tileEnd(const TileState& st, const TileArgs& a, T* __restrict__ amps, ct<T>* r,
                                        ct<T>* tile, const RegLayout& row, uint64_t base) {
    waitLoadsBeforeStores();
    if (a.lastDirect) {
        RegIO<T, NT, 0>::hbmStore(r, amps, base | st.cur.tg, st.cur, a.localPos);
    } else {
        if (st.inRegs) RegIO<T, NT, 0>::ldsStore(r, tile, opaqueV(st.sj), a.lastOff);
        __syncthreads();
        RegIO<T, NT, 0>::ldsLoad(r, tile, opaqueV(ldsByte<T>(row.jb)), a.rowOff);
        RegIO<T, NT, 0>::hbmStore(r, amps, base | row.tg, row, a.localPos);
    }
}

// the tile's registers put in the layout they are stored from (the last
// group's, or the row layout through LDS), which is returned; nothing stored
template <typename T, int NT>
__device__ __forceinline__ RegLayout tileFinish(const TileState& st, const TileArgs& a, ct<T>* r, ct<T>* tile,
                                                const RegLayout& row) {
    if (a.lastDirect) return st.cur;
    if (st.inRegs) RegIO<T, NT, 0>::ldsStore(r, tile, opaqueV(st.sj), a.lastOff);
    __syncthreads();
    RegIO<T, NT, 0>::ldsLoad(r, tile, opaqueV(ldsByte<T>(row.jb)), a.rowOff);
    return row;
}

// the interpreter: one tile of the pass whose program is ops[0, a.numOps)
// (ops[0] a GROUP op); r holds the tile in its load layout
template <typename T, int NT>
__device__ __forceinline__ void tileBody(T* __restrict__ amps, const TileArgs& a, const TOp* __restrict__ ops,
                                         uint64_t base, ct<T>* r, ct<T>* tile, const RegLayout& row,
                                         const RegLayout& first, uint32_t tid) {
    TileState st;
    tileBegin<T, NT>(st, a, r, tile, row, first);
    for (int o = a.firstDirect ? 1 : 0; o < a.numOps; o++) {
        u32x16 ch, cb;
        fetchOp(ops + o, ch, cb);
        OpRegs op;
        unpackOp(ch, cb, op);
        if (op.h.kind == F_PAULI_LDS) {
            stepPauliLds<T, NT>(st, a, ops + o, op, ((uint64_t)cb[13] << 32) | cb[12], r, tile, base, tid);
        } else if (op.h.kind == F_BLOCK) {
            const int bslots[BQ] = {(int)cb[0], (int)cb[1], (int)cb[2], (int)cb[3], (int)cb[4]};
            stepBlock<T, NT>(st, ops + o, bslots, a.blocks + (cb[6] | ((uint64_t)cb[7] << 32)), a.k, r, tile, tid);

        } else if (op.h.kind == F_GROUP) {
            int slots[RB];
#pragma unroll
            for (int b = 0; b < RB; b++) slots[b] = (int)cb[b];
            stepGroup<T, NT>(st, a, ops + o, slots, r, tile, tid);
        } else {
            stepApply<T>(st, ops + o, op, r, base);
        }
    }
    tileEnd<T, NT>(st, a, amps, r, tile, row, base);
}

// The next tile's loads as a program issues them: part<B, E>() loads
// registers B..E-1.  A generated program calls it between its ops, so the
// loads leave in small groups while the tile is worked on; issued at once
// (sixteen per wave, every wave together), they back up the memory pipeline
// and in-order issue stalls each wave at the group's last load until the
// group has been accepted.
template <typename T, int NT, int V, int E>
struct PrefetchRange {
    __device__ __forceinline__ static void load(ct<T>* rn, const T* amps, uint64_t at, const RegLayout& L,
                                                const int* pos) {
        rn[V] = ldx<NT>(amps, at | offQ<V>(L, pos));
        PrefetchRange<T, NT, V + 1, E>::load(rn, amps, at, L, pos);
    }
};
template <typename T, int NT, int E>
struct PrefetchRange<T, NT, E, E> {
    __device__ __forceinline__ static void load(ct<T>*, const T*, uint64_t, const RegLayout&, const int*) {}
};
template <typename T, int NT, int V, int E>
struct StoreRange {
    __device__ __forceinline__ static void store(const ct<T>* x, T* amps, uint64_t at, const RegLayout& L,
                                                 const int* pos) {
        stx<NT>(amps, at | offQ<V>(L, pos), x[V]);
        StoreRange<T, NT, V + 1, E>::store(x, amps, at, L, pos);
    }
};
template <typename T, int NT, int E>
struct StoreRange<T, NT, E, E> {
    __device__ __forceinline__ static void store(const ct<T>*, T*, uint64_t, const RegLayout&, const int*) {}
};

// a generated program ends with pf.finish: the tile stored (tileEnd)
template <typename T, int NT>
struct TilePrefetch {
    ct<T>* rn;
    const T* amps;
    uint64_t at;
    const RegLayout& L;
    const int* pos;
    template <int B, int E>
    __device__ __forceinline__ void part() {
        PrefetchRange<T, NT, B, E>::load(rn, amps, at, L, pos);
    }
    template <typename U, int NU>
    __device__ __forceinline__ void finish(const TileState& st, const TileArgs& a, U* __restrict__ amps_, ct<U>* r,
                                           ct<U>* tile, const RegLayout& row, uint64_t base) {
        tileEnd<U, NU>(st, a, amps_, r, tile, row, base);
    }
};
// the loop's stand-in when it loads the next tile itself (or not at all)
struct NoPrefetch {
    template <int B, int E>
    __device__ __forceinline__ void part() {}
    template <typename U, int NU>
    __device__ __forceinline__ void finish(const TileState& st, const TileArgs& a, U* __restrict__ amps, ct<U>* r,
                                           ct<U>* tile, const RegLayout& row, uint64_t base) {
        tileEnd<U, NU>(st, a, amps, r, tile, row, base);
    }
};

// PF = 3: two register sets take turns.  While a tile is worked on in one,
// the other holds the previous tile's results; each part<B, E>() stores
// their registers B..E-1 and loads the next tile's into the same registers,
// so stores and loads both leave in small groups through the work.  finish
// leaves the tile's results in its registers (put in their store layout)
// and records where they go.
template <typename T, int NT>
struct TileStream {
    ct<T>* x;           // the other register set
    T* amps;
    bool pending;       // x holds results to store first (not before a workgroup's first tile)
    uint64_t storeAt;
    RegLayout storeL;
    uint64_t loadAt;
    const RegLayout& loadL;
    const int* pos;
    uint64_t doneAt;    // set by finish: where this tile's results go
    RegLayout doneL;
    template <int B, int E>
    __device__ __forceinline__ void part() {
        if (pending) StoreRange<T, NT, B, E>::store(x, amps, storeAt, storeL, pos);
        PrefetchRange<T, NT, B, E>::load(x, amps, loadAt, loadL, pos);
    }
    template <typename U, int NU>
    __device__ __forceinline__ void finish(const TileState& st, const TileArgs& a, U* __restrict__, ct<U>* r,
                                           ct<U>* tile, const RegLayout& row, uint64_t base) {
        doneL = tileFinish<U, NU>(st, a, r, tile, row);
        doneAt = base | doneL.tg;
    }
};

// The persistent pass loop around one tile's program (Body::run), shared by
// the interpreter kernels (qa_tilepass.hpp) and the generated ones.  NT:
// non-temporal HBM loads/stores, used when the contiguous low run of a tile
// row is at least 2^STREAM_LOW_BIT amplitudes.  PF = 1: each workgroup
// prefetches its next tile into registers before working on the current
// one, so HBM traffic overlaps the register and LDS work (16 more registers
// of amplitudes per thread); PF = 2: the same, with the loads issued by the
// program between its ops (TilePrefetch; generated programs only); PF = 3:
// the stores too, a tile later (TileStream; generated programs only);
// PF = 0: load, work, store.
template <typename T, int NT, int PF, typename Body>
__device__ __forceinline__ void tilePassLoop(T* __restrict__ amps, const TileArgs& a, const TOp* __restrict__ ops) {
    extern __shared__ __align__(16) unsigned char smem[];
    ct<T>* tile = reinterpret_cast<ct<T>*>(smem);
    const uint32_t tid = threadIdx.x;
    RegLayout row, first;
    passLayouts(row, first, a, ops, tid);
    const RegLayout& loadL = a.firstDirect ? first : row;
    // tile t's base is t deposited over the non-local positions; the next
    // tile of this workgroup (t + gridDim.x) is one masked add away
    const uint64_t nlMask = a.nonLocalMask, step = a.tileStep;
    uint64_t t = blockIdx.x;
    uint64_t base = deposit(t, a.nonLocalPos, a.numNonLocal);
    if constexpr (PF == 3) {
        ct<T> x[RA], y[RA];
        if (t < a.numTiles) RegIO<T, NT, 0>::hbmLoad(x, amps, base | loadL.tg, loadL, a.localPos);
        bool pendY = false;
        uint64_t atY = 0;
        RegLayout LY = row;
        // two tiles per trip, one in each register set; a workgroup's last
        // tile loads its own rows again into the other set (dropped)
        while (t < a.numTiles) {
            uint64_t baseN = ((base | ~nlMask) + step) & nlMask;
            TileStream<T, NT> sx{y,     amps, pendY, atY, LY, (t + gridDim.x < a.numTiles ? baseN : base) | loadL.tg,
                                 loadL, a.localPos, 0, row};
            Body::template run<T, NT>(amps, a, ops, base, x, tile, row, first, tid, sx);
            base = baseN;
            t += gridDim.x;
            if (a.usesLds) __syncthreads();
            if (t >= a.numTiles) {
                RegIO<T, NT, 0>::hbmStore(x, amps, sx.doneAt, sx.doneL, a.localPos);
                break;
            }
            baseN = ((base | ~nlMask) + step) & nlMask;
            TileStream<T, NT> sy{x,     amps, true, sx.doneAt, sx.doneL,
                                 (t + gridDim.x < a.numTiles ? baseN : base) | loadL.tg, loadL, a.localPos, 0, row};
            Body::template run<T, NT>(amps, a, ops, base, y, tile, row, first, tid, sy);
            base = baseN;
            t += gridDim.x;
            if (a.usesLds) __syncthreads();
            if (t >= a.numTiles) {
                RegIO<T, NT, 0>::hbmStore(y, amps, sy.doneAt, sy.doneL, a.localPos);
                break;
            }
            pendY = true;
            atY = sy.doneAt;
            LY = sy.doneL;
        }
    } else if constexpr (PF) {
        ct<T> rn[RA];
        if (t < a.numTiles) RegIO<T, NT, 0>::hbmLoad(rn, amps, base | loadL.tg, loadL, a.localPos);
        // the first tile's loads retired before the loop, so that the loop
        // head needs no wait (which would also wait for the last tile's stores)
        waitLoadsBeforeStores();
        for (; t < a.numTiles; t += gridDim.x) {
            ct<T> r[RA];
#pragma unroll
            for (int v = 0; v < RA; v++) r[v] = rn[v];
            const uint64_t baseN = ((base | ~nlMask) + step) & nlMask;
            if constexpr (PF == 2) {
                // the program's loads are unconditional: a workgroup's last
                // tile loads its own rows again (read before its stores,
                // then dropped)
                const uint64_t nextAt = (t + gridDim.x < a.numTiles ? baseN : base) | loadL.tg;
                TilePrefetch<T, NT> pf{rn, amps, nextAt, loadL, a.localPos};
                Body::template run<T, NT>(amps, a, ops, base, r, tile, row, first, tid, pf);
            } else {
                if (t + gridDim.x < a.numTiles)
                    RegIO<T, NT, 0>::hbmLoad(rn, amps, baseN | loadL.tg, loadL, a.localPos);
                NoPrefetch pf;
                Body::template run<T, NT>(amps, a, ops, base, r, tile, row, first, tid, pf);
            }
            base = baseN;
            // the next tile's first LDS writes must not overtake this tile's reads
            if (a.usesLds) __syncthreads();
        }
    } else {
        for (; t < a.numTiles; t += gridDim.x) {
            ct<T> r[RA];
            RegIO<T, NT, 0>::hbmLoad(r, amps, base | loadL.tg, loadL, a.localPos);
            NoPrefetch pf;
            Body::template run<T, NT>(amps, a, ops, base, r, tile, row, first, tid, pf);
            base = ((base | ~nlMask) + step) & nlMask;
            if (a.usesLds) __syncthreads();
        }
    }
}

struct Interpreter {
    template <typename T, int NT, typename Pf>
    __device__ __forceinline__ static void run(T* __restrict__ amps, const TileArgs& a, const TOp* __restrict__ ops,
                                               uint64_t base, ct<T>* r, ct<T>* tile, const RegLayout& row,
                                               const RegLayout& first, uint32_t tid, Pf&) {
        tileBody<T, NT>(amps, a, ops, base, r, tile, row, first, tid);
    }
};

}  // namespace fuse
}  // namespace qa
