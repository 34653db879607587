// Gate fusion below the kernel seam.
//
// QuEST applies every gate as its own pass over the state (one kernel per
// gate: quest/src/gpu/gpu_subroutines.cpp launches one per operation).  Here the
// gate entry points of include/quest_amd_gpu.h may instead append the gate to
// a per-state queue; the queue is flushed (applied) before anything else
// touches that state -- every reduction, copy, getter, setter, free, sync --
// or when it grows long.  A flush plans "passes": maximal runs of queued gates
// whose non-diagonal targets all lie in one set L of k "local" qubits (the low
// `lowQubits` qubits are always in L so every tile row is a contiguous run of
// amplitudes).  One pass = one kernel: each workgroup loads one tile (the 2^k
// amplitudes that share the values of the n-k non-local bits) into LDS,
// applies the run of gates in program order on the tile, and writes it back:
// one HBM read and one write for the whole run instead of one per gate.
// Controls, diagonal factors and Pauli signs may involve any qubit: their
// non-local part is a per-tile constant, their local part is read from the
// tile index.
//
// Each gate's arithmetic on an amplitude is the same expression as in the
// single-gate kernels (qa_gates.hip, qa_channels.hip), applied in the same
// order, so fused and unfused runs give identical amplitudes.
#include <algorithm>
#include <cstring>
#include <mutex>
#include <unordered_map>
#include <vector>

#include "qa_common.hpp"
#include "qa_fused.hpp"

namespace qa {
namespace fuse {

// ---- tile-space program -----------------------------------------------------
//
// Inside a pass each thread holds RA = 2^RB amplitudes of the tile in
// registers: those whose tile indices differ only in the RB "register
// slots" of the current group.  A run of gates whose dense targets all lie
// in those slots is applied in registers; a GROUP op changes the register
// slots by writing the registers back to the LDS copy of the tile and
// gathering the new layout.  LDS is therefore read and written once per
// group instead of once per gate.

constexpr int RB = 4;        // register slots per group
constexpr int RA = 1 << RB;  // amplitudes per thread
enum { F_GROUP = 100 };

struct TOp {
    int kind;
    int rb1, rb2;            // register bits of the dense targets
    int slot[RB];            // GROUP: tile slots of register bits 0..RB-1 (ascending)
    uint32_t rcMask, rcVal;  // control condition, register part (over v)
    uint32_t rm1, rm2;       // register parts of the masks below
    uint64_t jcMask, jcVal;  // control condition, tile part outside the register slots
    uint64_t gcMask, gcVal;  // control condition, non-local part (tested against the tile base)
    uint64_t jm1, jm2;       // DIAG1 target / PARITY mask / PAULI maskXY (m1), maskYZ (m2): tile part
    uint64_t gm1, gm2;       // ... non-local part
    double f[32];
};

struct TileArgs {
    int k;          // tile bits
    int low;        // contiguous low bits (always local)
    int numNonLocal;
    int numOps;
    int localPos[MAXQ];     // sorted global positions of the tile bits
    int nonLocalPos[MAXQ];  // sorted global positions of the other bits
};

__device__ __forceinline__ uint64_t deposit(uint64_t v, const int* pos, int num) {
    uint64_t r = 0;
    for (int i = 0; i < num; i++) r |= ((v >> i) & 1ull) << pos[i];
    return r;
}

// LDS position of tile amplitude j: the low 4 index bits are XORed with the
// next 4, so threads whose amplitudes are 16 apart (register slots 0..3)
// still spread over all banks
__device__ __forceinline__ uint32_t swz(uint32_t j) { return j ^ ((j >> 4) & 15u); }

template <int V>
__device__ __forceinline__ uint32_t offOf(const uint32_t* sb) {
    uint32_t o = 0;
#pragma unroll
    for (int b = 0; b < RB; b++)
        if (V & (1 << b)) o |= sb[b];
    return o;
}

template <typename T, int V>
struct RegIO {
    __device__ __forceinline__ static void load(ct<T>* r, const ct<T>* tile, uint32_t jb, const uint32_t* sb) {
        r[V] = tile[swz(jb | offOf<V>(sb))];
        RegIO<T, V + 1>::load(r, tile, jb, sb);
    }
    __device__ __forceinline__ static void store(const ct<T>* r, ct<T>* tile, uint32_t jb, const uint32_t* sb) {
        tile[swz(jb | offOf<V>(sb))] = r[V];
        RegIO<T, V + 1>::store(r, tile, jb, sb);
    }
};
template <typename T>
struct RegIO<T, RA> {
    __device__ __forceinline__ static void load(ct<T>*, const ct<T>*, uint32_t, const uint32_t*) {}
    __device__ __forceinline__ static void store(const ct<T>*, ct<T>*, uint32_t, const uint32_t*) {}
};

__device__ __forceinline__ bool regCond(int v, const TOp& op) { return ((uint32_t)v & op.rcMask) == op.rcVal; }

template <typename T>
__device__ __forceinline__ ct<T> cf(const TOp& op, int i) {
    return ct<T>{(T)op.f[2 * i], (T)op.f[2 * i + 1]};
}

// the arithmetic of each gate below is the expression of its single-gate
// kernel (qa_gates.hip / qa_channels.hip), term for term

template <typename T, int B>
__device__ __forceinline__ void regDense1(ct<T>* r, const TOp& op) {
    const ct<T> m00 = cf<T>(op, 0), m01 = cf<T>(op, 1), m10 = cf<T>(op, 2), m11 = cf<T>(op, 3);
#pragma unroll
    for (int v = 0; v < RA; v++) {
        if (v & (1 << B)) continue;
        if (!regCond(v, op)) continue;
        const ct<T> a0 = r[v], a1 = r[v | (1 << B)];
        ct<T> b0 = cmulT(m00, a0);
        b0 = cfmaT(m01, a1, b0);
        ct<T> b1 = cmulT(m10, a0);
        b1 = cfmaT(m11, a1, b1);
        r[v] = b0;
        r[v | (1 << B)] = b1;
    }
}

template <typename T, int B1, int B2>
__device__ __forceinline__ void regDense2(ct<T>* r, const TOp& op) {
    constexpr int off[4] = {0, 1 << B1, 1 << B2, (1 << B1) | (1 << B2)};
#pragma unroll
    for (int v = 0; v < RA; v++) {
        if (v & ((1 << B1) | (1 << B2))) continue;
        if (!regCond(v, op)) continue;
        ct<T> x[4];
#pragma unroll
        for (int c = 0; c < 4; c++) x[c] = r[v | off[c]];
#pragma unroll
        for (int row = 0; row < 4; row++) {
            ct<T> acc = cmulT(cf<T>(op, 4 * row), x[0]);
#pragma unroll
            for (int c = 1; c < 4; c++) acc = cfmaT(cf<T>(op, 4 * row + c), x[c], acc);
            r[v | off[row]] = acc;
        }
    }
}

template <typename T, int B1, int B2>
__device__ __forceinline__ void regSwap(ct<T>* r, const TOp& op) {
#pragma unroll
    for (int v = 0; v < RA; v++) {
        if (!(v & (1 << B1)) || (v & (1 << B2))) continue;
        if (!regCond(v, op)) continue;
        constexpr int flip = (1 << B1) | (1 << B2);
        const ct<T> x = r[v];
        r[v] = r[v ^ flip];
        r[v ^ flip] = x;
    }
}

// one-qubit depolarising / dephasing on (ket, bra): both formulas are
// symmetric in the two bits, so B1 < B2 covers either order
template <typename T, int B1, int B2>
__device__ __forceinline__ void regDepol1(ct<T>* r, const TOp& op) {
    const T c1 = (T)op.f[0], c2 = (T)op.f[1], c3 = (T)op.f[2];
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
__device__ __forceinline__ void regDeph1(ct<T>* r, const TOp& op) {
    const T fac = (T)op.f[0];
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
// parity of the maskYZ bits outside the registers for this thread
template <typename T, int MXY>
__device__ __forceinline__ void regPauli(ct<T>* r, const TOp& op, int thrPar) {
    constexpr int top = 31 - __builtin_clz((unsigned)MXY);
    const ct<T> ampFac = cf<T>(op, 0), pairFac = cf<T>(op, 1);
#pragma unroll
    for (int v = 0; v < RA; v++) {
        if (v & (1 << top)) continue;
        if (!regCond(v, op)) continue;
        const int w = v ^ MXY;
        const ct<T> va = r[v], vb = r[w];
        const T sA = ((__popc((uint32_t)v & op.rm2) & 1) ^ thrPar) ? (T)-1 : (T)1;
        const T sB = ((__popc((uint32_t)w & op.rm2) & 1) ^ thrPar) ? (T)-1 : (T)1;
        const ct<T> pB = {pairFac.re * sB, pairFac.im * sB};
        const ct<T> pA = {pairFac.re * sA, pairFac.im * sA};
        r[v] = cfmaT(pB, vb, cmulT(ampFac, va));
        r[w] = cfmaT(pA, va, cmulT(ampFac, vb));
    }
}

template <typename T>
__device__ __forceinline__ void regDiag1(ct<T>* r, const TOp& op, bool thrBit) {
    const ct<T> d0 = cf<T>(op, 0), d1 = cf<T>(op, 1);
#pragma unroll
    for (int v = 0; v < RA; v++) {
        if (!regCond(v, op)) continue;
        const bool bit = op.rm1 ? (((uint32_t)v & op.rm1) != 0) : thrBit;
        r[v] = cmulT(r[v], bit ? d1 : d0);
    }
}

template <typename T>
__device__ __forceinline__ void regParity(ct<T>* r, const TOp& op, int thrPar) {
    const ct<T> f0 = cf<T>(op, 0), f1 = cf<T>(op, 1);
#pragma unroll
    for (int v = 0; v < RA; v++) {
        if (!regCond(v, op)) continue;
        r[v] = cmulT(r[v], ((__popc((uint32_t)v & op.rm1) & 1) ^ thrPar) ? f1 : f0);
    }
}

#define QA_PAIR_CASES(FN)                                             \
    case 0 * RB + 1: FN<T, 0, 1>(r, op); break;                       \
    case 0 * RB + 2: FN<T, 0, 2>(r, op); break;                       \
    case 0 * RB + 3: FN<T, 0, 3>(r, op); break;                       \
    case 1 * RB + 2: FN<T, 1, 2>(r, op); break;                       \
    case 1 * RB + 3: FN<T, 1, 3>(r, op); break;                       \
    case 2 * RB + 3: FN<T, 2, 3>(r, op); break;
#define QA_PAIR_CASES_REV(FN)                                         \
    case 1 * RB + 0: FN<T, 1, 0>(r, op); break;                       \
    case 2 * RB + 0: FN<T, 2, 0>(r, op); break;                       \
    case 3 * RB + 0: FN<T, 3, 0>(r, op); break;                       \
    case 2 * RB + 1: FN<T, 2, 1>(r, op); break;                       \
    case 3 * RB + 1: FN<T, 3, 1>(r, op); break;                       \
    case 3 * RB + 2: FN<T, 3, 2>(r, op); break;

// NT: non-temporal HBM loads/stores, used when the contiguous low run of a
// tile row is at least 2^STREAM_LOW_BIT amplitudes
template <typename T, int NT>
__global__ void __launch_bounds__(512) k_tilePass(T* __restrict__ amps, TileArgs a, const TOp* __restrict__ ops) {
    extern __shared__ unsigned char smem[];
    const int k = a.k;
    ct<T>* tile = reinterpret_cast<ct<T>*>(smem);
    uint64_t* table = reinterpret_cast<uint64_t*>(smem + (sizeof(ct<T>) << k));
    const uint32_t nt = blockDim.x;
    const uint32_t tid = threadIdx.x;
    const uint64_t base = deposit((uint64_t)blockIdx.x, a.nonLocalPos, a.numNonLocal);
    // global offset of tile row h (the k-low tile bits above the contiguous
    // low run) = tlo[low half of h] | thi[high half of h]
    const int hb = k - a.low, b0 = hb >> 1, b1 = hb - b0;
    uint64_t* tlo = table;
    uint64_t* thi = table + (1ull << b0);
    const uint32_t lowMask = (1u << a.low) - 1u, b0Mask = (1u << b0) - 1u;
    const int s0 = a.low + b0;
    for (uint32_t h = tid; h < (1u << b0); h += nt) tlo[h] = deposit(h, a.localPos + a.low, b0);
    for (uint32_t h = tid; h < (1u << b1); h += nt) thi[h] = deposit(h, a.localPos + a.low + b0, b1);
    __syncthreads();
    // every thread moves exactly RA = tileSize / nt amplitudes: all RA
    // loads are issued before the first LDS write
    ct<T> r[RA];
#pragma unroll
    for (int i = 0; i < RA; i++) {
        const uint32_t j = tid + (uint32_t)i * nt;
        r[i] = ldx<NT>(amps, base | (j & lowMask) | tlo[(j >> a.low) & b0Mask] | thi[j >> s0]);
    }
#pragma unroll
    for (int i = 0; i < RA; i++) tile[swz(tid + (uint32_t)i * nt)] = r[i];

    uint32_t sb[RB] = {0, 0, 0, 0};
    uint32_t jb = 0;
    bool have = false;
    for (int o = 0; o < a.numOps; o++) {
        const TOp& op = ops[o];
        if (op.kind == F_GROUP) {
            if (have) RegIO<T, 0>::store(r, tile, jb, sb);
            __syncthreads();
            uint32_t gm = 0;
#pragma unroll
            for (int b = 0; b < RB; b++) {
                sb[b] = 1u << op.slot[b];
                gm |= sb[b];
            }
            // this thread's tile index with the register slots zero: its id
            // deposited over the other slots, ascending
            jb = 0;
            for (int s = 0, i = 0; s < k; s++) {
                if ((gm >> s) & 1u) continue;
                jb |= ((tid >> i) & 1u) << s;
                i++;
            }
            RegIO<T, 0>::load(r, tile, jb, sb);
            have = true;
            continue;
        }
        if ((base & op.gcMask) != op.gcVal) continue;      // uniform over the workgroup
        if ((jb & op.jcMask) != op.jcVal) continue;        // uniform over the thread's registers
        switch (op.kind) {
            case F_DENSE1:
                switch (op.rb1) {
                    case 0: regDense1<T, 0>(r, op); break;
                    case 1: regDense1<T, 1>(r, op); break;
                    case 2: regDense1<T, 2>(r, op); break;
                    case 3: regDense1<T, 3>(r, op); break;
                }
                break;
            case F_DENSE2:
                switch (op.rb1 * RB + op.rb2) {
                    QA_PAIR_CASES(regDense2)
                    QA_PAIR_CASES_REV(regDense2)
                }
                break;
            case F_SWAP:
                switch (op.rb1 * RB + op.rb2) { QA_PAIR_CASES(regSwap) }
                break;
            case F_DEPOL1:
                switch (op.rb1 * RB + op.rb2) { QA_PAIR_CASES(regDepol1) }
                break;
            case F_DEPH1:
                switch (op.rb1 * RB + op.rb2) { QA_PAIR_CASES(regDeph1) }
                break;
            case F_DIAG1:
                regDiag1<T>(r, op, ((jb & op.jm1) | (base & op.gm1)) != 0);
                break;
            case F_PARITY:
                regParity<T>(r, op, parity64((jb & op.jm1) ^ (base & op.gm1)));
                break;
            case F_PAULI: {
                const int thrPar = parity64((jb & op.jm2) ^ (base & op.gm2));
                switch (op.rm1) {
#define QA_PAULI_CASE(M) case M: regPauli<T, M>(r, op, thrPar); break;
                    QA_PAULI_CASE(1) QA_PAULI_CASE(2) QA_PAULI_CASE(3) QA_PAULI_CASE(4) QA_PAULI_CASE(5)
                    QA_PAULI_CASE(6) QA_PAULI_CASE(7) QA_PAULI_CASE(8) QA_PAULI_CASE(9) QA_PAULI_CASE(10)
                    QA_PAULI_CASE(11) QA_PAULI_CASE(12) QA_PAULI_CASE(13) QA_PAULI_CASE(14) QA_PAULI_CASE(15)
#undef QA_PAULI_CASE
                }
                break;
            }
        }
    }
    if (have) RegIO<T, 0>::store(r, tile, jb, sb);
    __syncthreads();
#pragma unroll
    for (int i = 0; i < RA; i++) r[i] = tile[swz(tid + (uint32_t)i * nt)];
#pragma unroll
    for (int i = 0; i < RA; i++) {
        const uint32_t j = tid + (uint32_t)i * nt;
        stx<NT>(amps, base | (j & lowMask) | tlo[(j >> a.low) & b0Mask] | thi[j >> s0], r[i]);
    }
}

// ---- host side --------------------------------------------------------------

struct Queue {
    int prec = QA_FP64;
    int numQubits = 0;
    std::vector<GOp> ops;
};

std::mutex g_mu;
std::unordered_map<void*, Queue> g_queues;
Config g_cfg;  // disabled by default

// counters, and optional per-pass HIP events recorded on the library stream
// around each tile-pass launch (for measuring the pass kernel alone)
struct Stats {
    int64_t passes = 0, gates = 0;
    bool timing = false;
    std::vector<std::pair<hipEvent_t, hipEvent_t>> pool;
    size_t used = 0;
};
Stats g_stats;

// program ring: pinned host staging + device copy, reused in order; a wrap
// waits for the stream so no pending copy still reads the reused region
constexpr size_t RING_OPS = 16384;
TOp* g_ringHost = nullptr;
TOp* g_ringDev = nullptr;
size_t g_ringPos = 0;

bool ensureRing() {
    if (g_ringDev) return true;
    if (hipHostMalloc(&g_ringHost, RING_OPS * sizeof(TOp), hipHostMallocDefault) != hipSuccess) return false;
    if (hipMalloc(&g_ringDev, RING_OPS * sizeof(TOp)) != hipSuccess) return false;
    g_ringPos = 0;
    return true;
}

uint64_t denseMask(const GOp& g) {
    switch (g.kind) {
        case F_DENSE1: return 1ull << g.t1;
        case F_DENSE2: case F_SWAP: case F_DEPOL1: case F_DEPH1: return (1ull << g.t1) | (1ull << g.t2);
        case F_PAULI: return g.m1;
        default: return 0;  // diagonal: any qubit
    }
}

// dynamic LDS of one tile pass: the tile plus the two row-offset tables
size_t ldsBytes(int prec, int k, int low) {
    size_t amp = prec == QA_FP64 ? 16 : 8;
    int hb = k - low, b0 = hb >> 1, b1 = hb - b0;
    return (amp << k) + sizeof(uint64_t) * ((1ull << b0) + (1ull << b1));
}

std::vector<TOp> g_prog;  // program of the pass being launched (under g_mu)

// where each qubit of the state sits during one group of a pass
struct Layout {
    int slotOf[64];  // tile slot of a global qubit, -1 if non-local
    int regOf[MAXQ]; // register bit of a tile slot, -1 if outside the registers
};

// split a global mask into register bits, tile bits outside the registers,
// and non-local bits (kept at their global positions)
void splitMask(uint64_t gm, const Layout& L, uint32_t& reg, uint64_t& tile, uint64_t& nonlocal) {
    reg = 0;
    tile = nonlocal = 0;
    while (gm) {
        int q = __builtin_ctzll(gm);
        gm &= gm - 1;
        int s = L.slotOf[q];
        if (s < 0) nonlocal |= 1ull << q;
        else if (L.regOf[s] >= 0) reg |= 1u << L.regOf[s];
        else tile |= 1ull << s;
    }
}

TOp toTile(const GOp& g, const Layout& L) {
    TOp t;
    std::memset(&t, 0, sizeof(t));
    t.kind = g.kind;
    t.rb1 = g.t1 >= 0 ? L.regOf[L.slotOf[g.t1]] : -1;
    t.rb2 = g.t2 >= 0 ? L.regOf[L.slotOf[g.t2]] : -1;
    if ((g.kind == F_SWAP || g.kind == F_DEPOL1 || g.kind == F_DEPH1) && t.rb1 > t.rb2) std::swap(t.rb1, t.rb2);
    uint32_t rv;
    uint64_t tv, gv;
    splitMask(g.cMask, L, t.rcMask, t.jcMask, t.gcMask);
    splitMask(g.cVal, L, rv, tv, gv);
    t.rcVal = rv;
    t.jcVal = tv;
    t.gcVal = gv;
    if (g.kind == F_DIAG1) splitMask(1ull << g.t1, L, t.rm1, t.jm1, t.gm1);
    else splitMask(g.m1, L, t.rm1, t.jm1, t.gm1);
    splitMask(g.m2, L, t.rm2, t.jm2, t.gm2);
    std::memcpy(t.f, g.f, sizeof(t.f));
    return t;
}

// tile slots an op needs in registers
uint32_t regNeed(const GOp& g, const Layout& L) {
    uint64_t m = denseMask(g);
    uint32_t r = 0;
    while (m) {
        int q = __builtin_ctzll(m);
        m &= m - 1;
        r |= 1u << L.slotOf[q];
    }
    return r;
}

// ops [begin, end) of the queue, with every dense target local, as one
// pass: groups of consecutive ops whose dense targets fit RB register slots
int launchPass(void* amps, const Queue& qu, const std::vector<GOp>& ops, size_t begin, size_t end, uint64_t localMask) {
    const int n = qu.numQubits;
    TileArgs a;
    std::memset(&a, 0, sizeof(a));
    Layout L;
    for (int q = 0; q < 64; q++) L.slotOf[q] = -1;
    int k = 0, nl = 0;
    for (int q = 0; q < n; q++) {
        if ((localMask >> q) & 1ull) {
            L.slotOf[q] = k;
            a.localPos[k++] = q;
        } else {
            a.nonLocalPos[nl++] = q;
        }
    }
    a.k = k;
    a.low = std::min(g_cfg.lowQubits, k);
    a.numNonLocal = nl;

    // plan the groups
    std::vector<TOp>& prog = g_prog;
    prog.clear();
    size_t i = begin;
    while (i < end) {
        uint32_t G = 0;
        size_t j = i;
        for (; j < end; j++) {
            uint32_t u = G | regNeed(ops[j], L);
            if (__builtin_popcount(u) > RB) break;
            G = u;
        }
        // pad the register slots with the lowest free ones
        for (int sl = 0; sl < k && __builtin_popcount(G) < RB; sl++) G |= 1u << sl;
        TOp grp;
        std::memset(&grp, 0, sizeof(grp));
        grp.kind = F_GROUP;
        for (int sl = 0; sl < k; sl++) L.regOf[sl] = -1;
        for (int sl = 0, b = 0; sl < k; sl++)
            if ((G >> sl) & 1u) {
                grp.slot[b] = sl;
                L.regOf[sl] = b++;
            }
        prog.push_back(grp);
        for (; i < j; i++) prog.push_back(toTile(ops[i], L));
    }

    size_t num = prog.size();
    if (g_ringPos + num > RING_OPS) {
        QA_HIP(hipStreamSynchronize(stream()));
        g_ringPos = 0;
    }
    TOp* h = g_ringHost + g_ringPos;
    std::memcpy(h, prog.data(), num * sizeof(TOp));
    TOp* d = g_ringDev + g_ringPos;
    g_ringPos += num;
    QA_HIP(hipMemcpyAsync(d, h, num * sizeof(TOp), hipMemcpyHostToDevice, stream()));
    a.numOps = (int)num;
    size_t lds = ldsBytes(qu.prec, k, a.low);
    unsigned grid = (unsigned)(1ull << nl);
    unsigned block = 1u << (k - RB);
    std::pair<hipEvent_t, hipEvent_t>* ev = nullptr;
    if (g_stats.timing) {
        if (g_stats.used == g_stats.pool.size()) {
            std::pair<hipEvent_t, hipEvent_t> e;
            QA_HIP(hipEventCreate(&e.first));
            QA_HIP(hipEventCreate(&e.second));
            g_stats.pool.push_back(e);
        }
        ev = &g_stats.pool[g_stats.used++];
        QA_HIP(hipEventRecord(ev->first, stream()));
    }
    const bool nt = a.low >= STREAM_LOW_BIT;
    if (qu.prec == QA_FP64) {
        if (nt) k_tilePass<double, 1><<<grid, block, lds, stream()>>>((double*)amps, a, d);
        else k_tilePass<double, 0><<<grid, block, lds, stream()>>>((double*)amps, a, d);
    } else {
        if (nt) k_tilePass<float, 1><<<grid, block, lds, stream()>>>((float*)amps, a, d);
        else k_tilePass<float, 0><<<grid, block, lds, stream()>>>((float*)amps, a, d);
    }
    QA_LAUNCH_CHECK();
    if (ev) QA_HIP(hipEventRecord(ev->second, stream()));
    g_stats.passes++;
    g_stats.gates += (int64_t)(end - begin);
    return QA_OK;
}

// greedy planner: extend the local set while the next gate's non-diagonal
// targets still fit in k qubits, then launch the run as one pass
int flushQueue(void* amps, Queue& qu) {
    if (qu.ops.empty()) return QA_OK;
    if (!ensureRing()) {
        setError("fusion: program ring allocation failed");
        return QA_ERR_OOM;
    }
    const int n = qu.numQubits;
    const int kMax = std::min(n, g_cfg.tileQubits);
    const int low = std::min(g_cfg.lowQubits, kMax);
    const uint64_t lowMask = (low >= 64) ? ~0ull : ((1ull << low) - 1ull);
    std::vector<GOp>& ops = qu.ops;
    size_t i = 0;
    int rc = QA_OK;
    while (i < ops.size() && rc == QA_OK) {
        uint64_t L = lowMask;
        size_t start = i;
        while (i < ops.size()) {
            uint64_t need = denseMask(ops[i]);
            uint64_t u = L | need;
            if (__builtin_popcountll(u) > kMax) break;
            L = u;
            i++;
        }
        if (i == start) {
            setError("fusion: a gate needs more local qubits than the tile holds");
            rc = QA_ERR_ARG;
            break;
        }
        // pad the local set to exactly kMax qubits with the lowest free qubits
        for (int q = 0; q < n && __builtin_popcountll(L) < kMax; q++) L |= 1ull << q;
        rc = launchPass(amps, qu, ops, start, i, L);
    }
    ops.clear();
    return rc;
}

}  // namespace fuse

// ---- interface used by the entry points --------------------------------------

namespace fuse {

bool enabledFor(void* amps, int prec, int numQubits) {
    if (!g_cfg.enabled) return false;
    (void)amps;
    (void)prec;
    return numQubits >= std::max(g_cfg.minQubits, RB);
}

int enqueue(void* amps, int prec, int numQubits, const GOp& op, bool& accepted) {
    accepted = false;
    std::lock_guard<std::mutex> lk(g_mu);
    Queue& qu = g_queues[amps];
    if (!qu.ops.empty() && (qu.prec != prec || qu.numQubits != numQubits)) {
        int rc = flushQueue(amps, qu);
        if (rc != QA_OK) return rc;
    }
    qu.prec = prec;
    qu.numQubits = numQubits;
    // a gate whose dense part cannot fit any tile runs unfused
    int kMax = std::min(numQubits, g_cfg.tileQubits);
    int lowCount = std::min(g_cfg.lowQubits, kMax);
    uint64_t lowMask = (1ull << lowCount) - 1ull;
    if (__builtin_popcountll(denseMask(op) | lowMask) > kMax || __builtin_popcountll(denseMask(op)) > RB)
        return QA_OK;
    qu.ops.push_back(op);
    accepted = true;
    if ((int)qu.ops.size() >= g_cfg.maxQueue) return flushQueue(amps, qu);
    return QA_OK;
}

int flush(void* amps) {
    std::lock_guard<std::mutex> lk(g_mu);
    auto it = g_queues.find(amps);
    if (it == g_queues.end()) return QA_OK;
    int rc = flushQueue(amps, it->second);
    if (it->second.ops.empty()) g_queues.erase(it);
    return rc;
}

int flushAll() {
    std::lock_guard<std::mutex> lk(g_mu);
    int rc = QA_OK;
    for (auto& kv : g_queues) {
        int r = flushQueue(kv.first, kv.second);
        if (r != QA_OK) rc = r;
    }
    g_queues.clear();
    return rc;
}

void drop(void* amps) {
    std::lock_guard<std::mutex> lk(g_mu);
    g_queues.erase(amps);
}

void shutdown() {
    std::lock_guard<std::mutex> lk(g_mu);
    g_queues.clear();
    for (auto& e : g_stats.pool) {
        (void)hipEventDestroy(e.first);
        (void)hipEventDestroy(e.second);
    }
    g_stats = Stats();
    if (g_ringHost) (void)hipHostFree(g_ringHost);
    if (g_ringDev) (void)hipFree(g_ringDev);
    g_ringHost = nullptr;
    g_ringDev = nullptr;
    g_ringPos = 0;
}

const Config& config() { return g_cfg; }

}  // namespace fuse
}  // namespace qa

extern "C" {

int qa_setFusion(int enabled, int tileQubits, int lowQubits, int maxQueue) {
    QA_REQUIRE_INIT();
    int rc = qa::fuse::flushAll();
    if (rc != QA_OK) return rc;
    QA_ARG_CHECK(tileQubits >= qa::fuse::RB && tileQubits <= 13, "qa_setFusion: tileQubits must be in [4, 13]");
    QA_ARG_CHECK(lowQubits >= 0 && lowQubits <= tileQubits - 2, "qa_setFusion: lowQubits must be in [0, tileQubits-2]");
    QA_ARG_CHECK(maxQueue >= 1 && maxQueue <= 4096, "qa_setFusion: maxQueue must be in [1, 4096]");
    {
        int dev = 0, maxLds = 0;
        QA_HIP(hipGetDevice(&dev));
        QA_HIP(hipDeviceGetAttribute(&maxLds, hipDeviceAttributeMaxSharedMemoryPerBlock, dev));
        // FP32 tiles hold one more qubit and the same bytes as FP64 tiles
        size_t need = qa::fuse::ldsBytes(QA_FP64, tileQubits, lowQubits);
        QA_ARG_CHECK(need <= (size_t)maxLds, "qa_setFusion: a %d-qubit tile needs %zu bytes of LDS, the device allows %d",
                     tileQubits, need, maxLds);
        if (need > 65536) {
            const void* fns[] = {(const void*)qa::fuse::k_tilePass<double, 0>,
                                 (const void*)qa::fuse::k_tilePass<double, 1>,
                                 (const void*)qa::fuse::k_tilePass<float, 0>,
                                 (const void*)qa::fuse::k_tilePass<float, 1>};
            for (const void* fn : fns)
                QA_HIP(hipFuncSetAttribute(fn, hipFuncAttributeMaxDynamicSharedMemorySize, (int)need));
        }
    }
    std::lock_guard<std::mutex> lk(qa::fuse::g_mu);
    qa::fuse::g_cfg.enabled = enabled != 0;
    qa::fuse::g_cfg.tileQubits = tileQubits;
    qa::fuse::g_cfg.lowQubits = lowQubits;
    qa::fuse::g_cfg.maxQueue = maxQueue;
    return QA_OK;
}

int qa_getFusion(int* enabled, int* tileQubits, int* lowQubits, int* maxQueue) {
    const qa::fuse::Config& c = qa::fuse::config();
    if (enabled) *enabled = c.enabled;
    if (tileQubits) *tileQubits = c.tileQubits;
    if (lowQubits) *lowQubits = c.lowQubits;
    if (maxQueue) *maxQueue = c.maxQueue;
    return QA_OK;
}

int qa_setFusionTiming(int enabled) {
    QA_REQUIRE_INIT();
    QA_FLUSH_ALL();
    std::lock_guard<std::mutex> lk(qa::fuse::g_mu);
    qa::fuse::g_stats.timing = enabled != 0;
    qa::fuse::g_stats.used = 0;
    qa::fuse::g_stats.passes = qa::fuse::g_stats.gates = 0;
    return QA_OK;
}

int qa_getFusionStats(int64_t* passes, int64_t* gates, double* passMs) {
    QA_REQUIRE_INIT();
    QA_FLUSH_ALL();
    std::lock_guard<std::mutex> lk(qa::fuse::g_mu);
    qa::fuse::Stats& st = qa::fuse::g_stats;
    double total = 0;
    for (size_t i = 0; i < st.used; i++) {
        float f = 0;
        QA_HIP(hipEventSynchronize(st.pool[i].second));
        QA_HIP(hipEventElapsedTime(&f, st.pool[i].first, st.pool[i].second));
        total += f;
    }
    if (passes) *passes = st.passes;
    if (gates) *gates = st.gates;
    if (passMs) *passMs = total;
    st.used = 0;
    st.passes = st.gates = 0;
    return QA_OK;
}

int qa_flush(void* amps) {
    QA_REQUIRE_INIT();
    return amps ? qa::fuse::flush(amps) : qa::fuse::flushAll();
}

}  // extern "C"
