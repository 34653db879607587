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
// single-gate kernels (qa_gates.hip, qa_channels.hip).  In exact mode the
// gates are applied in queue order, so fused and unfused runs give identical
// amplitudes; by default a pass also takes later gates that commute with
// those it passes over, and multiplies runs of gates on the same one or two
// qubits into one matrix on the host, which changes rounding only.
#include <algorithm>
#include <cmath>
#include <complex>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <map>
#include <mutex>
#include <tuple>
#include <unordered_map>
#include <vector>

#include "qa_common.hpp"
#include "qa_fused.hpp"
#include "qa_tile.hpp"
#include "qa_tilepass.hpp"
#include "qa_jit.hpp"

namespace qa {
namespace fuse {

const void* tileKernel(int prec, bool nt, bool pf) {
    return prec == QA_FP64 ? tileKernelF64(nt, pf) : tileKernelT<float>(nt, pf);
}

void launchTileAny(int prec, bool nt, bool pf, unsigned grid, unsigned block, size_t lds, void* amps,
                   const TileArgs& a, const TOp* ops) {
    if (prec == QA_FP64) launchTileF64(nt, pf, grid, block, lds, amps, a, ops);
    else launchTileT<float>(nt, pf, grid, block, lds, amps, a, ops);
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
    int64_t spanPasses = 0, spanGates = 0;  // among them, span passes and their gates
    int64_t progOps = 0, layouts = 0;  // program entries of the passes, and their GROUP ops
    double flops = 0;  // algorithmic real floating-point operations of the passes
    bool timing = false;
    std::vector<std::pair<hipEvent_t, hipEvent_t>> pool;
    std::vector<char> poolSpan;  // the pass of pool[i] was a span pass
    size_t used = 0;
    double passMs = 0, spanMs = 0;  // elapsed times collected from the pool
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
        case F_DENSEN: {
            uint64_t m = 0;
            for (int b = 0; b < g.numTargs; b++) m |= 1ull << g.targs[b];
            return m;
        }
        case F_DENSE1: case F_DENSE1V: return 1ull << g.t1;
        case F_DENSE2: case F_SWAP: case F_DEPOL1: case F_DEPH1: return (1ull << g.t1) | (1ull << g.t2);
        case F_PAULI: return g.m1;
        default: return 0;  // diagonal: any qubit
    }
}

// dynamic LDS of one tile pass: the tile plus the two row-offset tables
size_t ldsBytes(int prec, int k) {
    size_t amp = prec == QA_FP64 ? 16 : 8;
    return amp << k;
}

unsigned residentBlocks(const void* fn, unsigned block, size_t lds) {
    static std::mutex mu;
    static std::map<std::tuple<const void*, unsigned, size_t>, unsigned> cache;
    std::lock_guard<std::mutex> lk(mu);
    const auto key = std::make_tuple(fn, block, lds);
    auto it = cache.find(key);
    if (it != cache.end()) return it->second;
    int perCU = 0, dev = 0, cus = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&perCU, fn, (int)block, lds) != hipSuccess || perCU < 1)
        perCU = 1;
    (void)hipGetDevice(&dev);
    if (hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess || cus < 1) cus = 1;
    unsigned r = (unsigned)(perCU * cus);
    cache[key] = r;
    return r;
}

bool ldsTileAtZero(const void* fn) {
    static std::mutex mu;
    static std::map<const void*, bool> cache;
    std::lock_guard<std::mutex> lk(mu);
    auto it = cache.find(fn);
    if (it != cache.end()) return it->second;
    hipFuncAttributes attr;
    const bool ok = hipFuncGetAttributes(&attr, fn) == hipSuccess && attr.sharedSizeBytes == 0;
    cache[fn] = ok;
    return ok;
}

std::vector<TOp> g_prog;  // program of the pass being launched (under g_mu)
std::vector<SpanOp> g_spanProg;  // ops of the span pass being planned (under g_mu)

// packed MFMA block matrices: staged per pass, uploaded through a ring like
// the op ring (a wrap waits for the stream)
constexpr size_t BLOCK_RING_FLOATS = 4u << 20;  // 1024 blocks of 16 KiB
std::vector<float> g_blockData;
float* g_blockHost = nullptr;
float* g_blockDev = nullptr;
size_t g_blockPos = 0;

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
    TOpHead& h = t.h;
    h.kind = g.kind;
    auto regBit = [&](int q) { return (q >= 0 && L.slotOf[q] >= 0) ? L.regOf[L.slotOf[q]] : -1; };
    h.rb1 = regBit(g.t1);
    h.rb2 = regBit(g.t2);
    if ((g.kind == F_SWAP || g.kind == F_DEPOL1 || g.kind == F_DEPH1) && h.rb1 > h.rb2) std::swap(h.rb1, h.rb2);
    uint32_t rv;
    uint64_t tv, gv;
    splitMask(g.cMask, L, h.rcMask, h.jcMask, h.gcMask);
    splitMask(g.cVal, L, rv, tv, gv);
    h.rcVal = rv;
    h.jcVal = tv;
    h.gcVal = gv;
    uint64_t jm1, gm1, jm2, gm2;
    if (g.kind == F_DIAG1) splitMask(1ull << g.t1, L, h.rm1, jm1, gm1);
    else splitMask(g.m1, L, h.rm1, jm1, gm1);
    splitMask(g.m2, L, h.rm2, jm2, gm2);
    std::memcpy(t.b, g.f, 8 * sizeof(double));
    if (g.kind == F_DIAG1 || g.kind == F_PARITY) {
        t.b[4] = jm1;
        t.b[5] = gm1;
    } else if (g.kind == F_PAULI) {
        t.b[4] = jm2;
        t.b[5] = gm2;
    }
    // DENSE2: matrix elements 4..15 (the first 4 are in b)
    if (g.kind == F_DENSE2) std::memcpy(t.fx, g.f + 8, 24 * sizeof(double));
    if (g.kind == F_DENSE1V) {
        // variant qubits (non-local): the lowest in rb2, the other (+1) in pad
        int nv = 0;
        h.pad = 0;
        for (int x = 0; x < 64; x++)
            if ((g.m1 >> x) & 1ull) {
                if (nv == 0) h.rb2 = x;
                else h.pad = (uint32_t)(x + 1);
                nv++;
            }
        // all four variant matrices in fx, so a tile reads its variant's
        // matrix at one computed offset (variant 0 also stays in b)
        std::memcpy(t.fx, g.f, 32 * sizeof(double));
    }
    if (g.kind == F_DIAG1 && h.rb1 >= 0 && ((h.rcMask >> h.rb1) & 1u)) {
        // phase-only diagonal (its target is one of its controls, set to 1):
        // the register kernel skips the amplitudes with the bit clear itself
        h.rcMask &= ~(1u << h.rb1);
        h.rcVal &= ~(1u << h.rb1);
        h.pad = 1;
    }
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

}  // namespace fuse (planner helpers)

size_t fuse::takeRun(const GOp* ops, size_t num, int numQubits, int kMax, int low, uint64_t& localMask) {
    uint64_t L = (low >= 64) ? ~0ull : ((1ull << low) - 1ull);
    size_t i = 0;
    for (; i < num; i++) {
        uint64_t u = L | denseMask(ops[i]);
        if (__builtin_popcountll(u) > kMax) break;
        L = u;
    }
    for (int q = 0; q < numQubits && __builtin_popcountll(L) < kMax; q++) L |= 1ull << q;
    localMask = L;
    return i;
}

void fuse::pickPass(const GOp* ops, size_t num, int numQubits, int kMax, int low, uint64_t& localMask,
                    std::vector<uint32_t>& picked) {
    uint64_t L = (low >= 64) ? ~0ull : ((1ull << low) - 1ull);
    const uint64_t all = (numQubits >= 64) ? ~0ull : ((1ull << numQubits) - 1ull);
    uint64_t skipDense = 0, skipTouch = 0;  // qubits of the ops passed over
    for (size_t i = 0; i < num; i++) {
        const uint64_t d = denseMask(ops[i]);
        const uint64_t u = L | d;
        if ((d & skipTouch) == 0 && (supportMask(ops[i]) & skipDense) == 0 && __builtin_popcountll(u) <= kMax) {
            L = u;
            picked.push_back((uint32_t)i);
            continue;
        }
        skipDense |= d;
        skipTouch |= supportMask(ops[i]);
        // nothing involving a qubit can join once every qubit is a dense
        // target of a skipped op
        if (skipDense == all) break;
    }
    for (int q = 0; q < numQubits && __builtin_popcountll(L) < kMax; q++) L |= 1ull << q;
    localMask = L;
}

namespace fuse {

// the tile slots of localMask, and the pass's fixed fields of `a`
void passShape(int numQubits, uint64_t localMask, TileArgs& a, Layout& L) {
    std::memset(&a, 0, sizeof(a));
    for (int q = 0; q < 64; q++) L.slotOf[q] = -1;
    for (int s = 0; s < MAXQ; s++) L.regOf[s] = -1;
    int k = 0, nl = 0;
    for (int q = 0; q < numQubits; q++) {
        if ((localMask >> q) & 1ull) {
            L.slotOf[q] = k;
            a.localPos[k++] = q;
        } else {
            a.nonLocalPos[nl++] = q;
            a.nonLocalMask |= 1ull << q;
        }
    }
    a.k = k;
    a.numNonLocal = nl;
    a.numTiles = 1ull << nl;
}

// Emits ops[from, to) as register groups (a GROUP op, then the ops whose
// register needs fit RB slots with it, and so on).  Padding takes the
// highest free slots outside the low run first, so that the first and last
// groups can load and store HBM directly.
struct GroupState {
    int numGroups = 0;
    uint32_t firstG = 0, lastG = 0;
    bool anyItem = false;    // a group or an LDS op was emitted
    bool firstIsLds = false, lastIsLds = false;  // the pass starts / ends with an LDS op
};

// a GROUP op whose layout is never used: keeps ops[0] a layout when a pass
// starts with an op that works on LDS
TOp placeholderGroup(int k) {
    TOp grp;
    std::memset(&grp, 0, sizeof(grp));
    grp.h.kind = F_GROUP;
    int* slots = reinterpret_cast<int*>(grp.b);
    for (int b = 0; b < RB; b++) slots[b] = k - RB + b;
    return grp;
}

// a Pauli tensor/gadget with more X/Y qubits than register bits: applied
// to the tile in LDS
bool isLdsOp(const GOp& g) { return (g.kind == F_PAULI && __builtin_popcountll(g.m1) > RB) || g.kind == F_DENSEN; }

// a 3- or 4-qubit dense matrix: applied to the tile in LDS, each thread
// taking the groups of 2^n amplitudes that differ in its target slots
// (body words: the target slots in matrix-bit order, n, conj, the matrix)
TOp ldsDense(const GOp& g, const Layout& L) {
    TOp t;
    std::memset(&t, 0, sizeof(t));
    t.h.kind = F_DENSE_LDS;
    uint32_t* w = reinterpret_cast<uint32_t*>(t.b);
    for (int b = 0; b < 4; b++) w[b] = b < g.numTargs ? (uint32_t)L.slotOf[g.targs[b]] : 0u;
    w[4] = (uint32_t)g.numTargs;
    w[5] = (uint32_t)g.conj;
    const uint64_t p = (uint64_t)(uintptr_t)g.mat;
    w[6] = (uint32_t)p;
    w[7] = (uint32_t)(p >> 32);
    return t;
}

TOp ldsOp(const GOp& g, const Layout& L);

TOp ldsPauli(const GOp& g, const Layout& L) {
    Layout T = L;  // every local bit counted in the tile part
    for (int s = 0; s < MAXQ; s++) T.regOf[s] = -1;
    TOp t = toTile(g, T);
    t.h.kind = F_PAULI_LDS;
    uint32_t r;
    uint64_t tile, nonlocal;
    splitMask(g.m1, T, r, tile, nonlocal);
    t.b[6] = tile;  // X/Y bits (all local)
    return t;
}

TOp ldsOp(const GOp& g, const Layout& L) { return g.kind == F_DENSEN ? ldsDense(g, L) : ldsPauli(g, L); }

// appends the PHASES terms of a diagonal gate (DIAG1, or a PARITY on one
// qubit) under layout L; false, appending nothing, for any other gate or
// when the terms would not fit one op
bool phaseTerms(const GOp& g, const Layout& L, std::vector<PhaseTerm>& terms) {
    uint64_t t;
    std::complex<double> f0, f1;
    if (g.kind == F_DIAG1) {
        t = 1ull << g.t1;
        f0 = {g.f[0], g.f[1]};
        f1 = {g.f[2], g.f[3]};
    } else if (g.kind == F_PARITY && __builtin_popcountll(g.m1) == 1) {
        t = g.m1;
        f0 = {g.f[0], g.f[1]};
        f1 = {g.f[2], g.f[3]};
    } else {
        return false;
    }
    // the target bit set (factor f1) and clear (f0) under the controls;
    // a phase-only gate has its target among its controls
    PhaseTerm add[2];
    int na = 0;
    const uint64_t m = g.cMask | t;
    for (int bit = 1; bit >= 0; bit--) {
        if ((g.cMask & t) && ((g.cVal & t) != 0) != (bit == 1)) continue;
        const std::complex<double> f = bit ? f1 : f0;
        if (f == std::complex<double>(1.0, 0.0)) continue;
        const uint64_t v = (g.cVal & ~t) | (bit ? t : 0);
        PhaseTerm& pt = add[na++];
        uint64_t jm, jv;
        splitMask(m, L, pt.rm, jm, pt.gm);
        splitMask(v, L, pt.rv, jv, pt.gv);
        pt.jm = (uint32_t)jm;
        pt.jv = (uint32_t)jv;
        pt.fre = f.real();
        pt.fim = f.imag();
    }
    // register part: empty, or one bit that must be set
    for (int i = 0; i < na; i++)
        if (__builtin_popcount(add[i].rm) > 1 || add[i].rv != add[i].rm) return false;
    if ((int)terms.size() + na > MAX_PHASE_TERMS) return false;
    for (int i = 0; i < na; i++) terms.push_back(add[i]);
    return true;
}

void emitGroups(const GOp* ops, size_t from, size_t to, Layout& L, int k, int low, std::vector<TOp>& prog,
                GroupState& st, bool phases) {
    size_t i = from;
    while (i < to) {
        if (isLdsOp(ops[i])) {
            prog.push_back(ldsOp(ops[i], L));
            if (st.numGroups == 0 && !st.anyItem) st.firstIsLds = true;
            st.lastIsLds = true;
            st.anyItem = true;
            i++;
            continue;
        }
        uint32_t G = 0;
        size_t j = i;
        for (; j < to; j++) {
            if (isLdsOp(ops[j])) break;
            uint32_t u = G | regNeed(ops[j], L);
            if (__builtin_popcount(u) > RB) break;
            G = u;
        }
        for (int sl = k - 1; sl >= low && __builtin_popcount(G) < RB; sl--) G |= 1u << sl;
        for (int sl = 0; sl < low && __builtin_popcount(G) < RB; sl++) G |= 1u << sl;
        TOp grp;
        std::memset(&grp, 0, sizeof(grp));
        grp.h.kind = F_GROUP;
        for (int sl = 0; sl < k; sl++) L.regOf[sl] = -1;
        int* slots = reinterpret_cast<int*>(grp.b);
        // register bits in the order uncontrolled one-qubit gates first use
        // their slots, so runs of them land on bits 0, 1, 2, ... in order
        int nb = 0;
        for (size_t p = i; p < j && nb < RB; p++) {
            const GOp& g = ops[p];
            if (g.kind != F_DENSE1 || g.cMask) continue;
            const int sl = L.slotOf[g.t1];
            if (L.regOf[sl] >= 0) continue;
            slots[nb] = sl;
            L.regOf[sl] = nb++;
        }
        for (int sl = 0; sl < k; sl++)
            if (((G >> sl) & 1u) && L.regOf[sl] < 0) {
                slots[nb] = sl;
                L.regOf[sl] = nb++;
            }
        prog.push_back(grp);
        if (st.numGroups == 0) st.firstG = G;
        st.lastG = G;
        st.numGroups++;
        st.lastIsLds = false;
        st.anyItem = true;
        while (i < j) {
            if (phases) {
                // a run of diagonal gates becomes one PHASES op
                std::vector<PhaseTerm> terms;
                size_t n = 0;
                while (i + n < j && phaseTerms(ops[i + n], L, terms)) n++;
                if (n >= 2) {
                    std::stable_sort(terms.begin(), terms.end(),
                                     [](const PhaseTerm& x, const PhaseTerm& y) { return x.rm < y.rm; });
                    TOp t;
                    std::memset(&t, 0, sizeof(t));
                    t.h.kind = F_PHASES;
                    t.h.rb1 = (int)terms.size();
                    for (const PhaseTerm& pt : terms) {
                        const int sl = pt.rm ? 1 + __builtin_ctz(pt.rm) : 0;
                        t.h.rm1 += 1u << (4 * sl);
                    }
                    std::memcpy(t.b, terms.data(), terms.size() * sizeof(PhaseTerm));
                    prog.push_back(t);
                    i += n;
                    continue;
                }
            }
            // a run of 2-4 uncontrolled one-qubit gates on register bits
            // 0, 1, ... in order becomes one DENSE1N op (same arithmetic)
            size_t n = 0;
            while (n < (size_t)RB && i + n < j && ops[i + n].kind == F_DENSE1 && !ops[i + n].cMask &&
                   L.regOf[L.slotOf[ops[i + n].t1]] == (int)n)
                n++;
            if (n >= 2) {
                TOp t = toTile(ops[i], L);
                t.h.kind = F_DENSE1N;
                t.h.rb2 = (int)n;
                for (size_t m = 1; m < n; m++) std::memcpy(t.fx + 8 * (m - 1), ops[i + m].f, 8 * sizeof(double));
                prog.push_back(t);
                i += n;
            } else {
                prog.push_back(toTile(ops[i], L));
                i++;
            }
        }
    }
}

uint64_t supportMask(const GOp& g) {
    uint64_t m = g.cMask | denseMask(g);
    switch (g.kind) {
        case F_DENSE1V: m |= g.m1; break;
        case F_DIAG1: m |= 1ull << g.t1; break;
        case F_PARITY: m |= g.m1; break;
        case F_PAULI: m |= g.m1 | g.m2; break;
        default: break;
    }
    return m;
}

// g applied to a vector of 2^nq amplitudes whose bit b is global qubit q[b]
// (every qubit of g is one of them), in FP64, with the same semantics as
// the single-gate kernels
void applySmall(const GOp& g, const int* q, int nq, std::complex<double>* v) {
    typedef std::complex<double> cx;
    auto loc = [&](uint64_t m) {
        uint32_t r = 0;
        for (int b = 0; b < nq; b++)
            if ((m >> q[b]) & 1ull) r |= 1u << b;
        return r;
    };
    auto bitOf = [&](int qubit) {
        for (int b = 0; b < nq; b++)
            if (q[b] == qubit) return b;
        return -1;
    };
    const uint32_t cm = loc(g.cMask), cv = loc(g.cVal);
    const int D = 1 << nq;
    auto fc = [&](int i) { return cx(g.f[2 * i], g.f[2 * i + 1]); };
    switch (g.kind) {
        case F_DENSE1: {
            const uint32_t t = 1u << bitOf(g.t1);
            for (int j = 0; j < D; j++) {
                if ((j & t) || (j & cm) != cv) continue;
                const cx a0 = v[j], a1 = v[j | t];
                v[j] = fc(0) * a0 + fc(1) * a1;
                v[j | t] = fc(2) * a0 + fc(3) * a1;
            }
            break;
        }
        case F_DENSE1V: {
            const uint32_t t = 1u << bitOf(g.t1);
            int vb[2], nv = 0;
            for (int x = 0; x < 64 && nv < 2; x++)
                if ((g.m1 >> x) & 1ull) vb[nv++] = bitOf(x);
            for (int j = 0; j < D; j++) {
                if (j & t) continue;
                int w = 0;
                for (int b = 0; b < nv; b++) w |= ((j >> vb[b]) & 1) << b;
                const double* m = g.f + 8 * w;
                const cx a0 = v[j], a1 = v[j | t];
                v[j] = cx(m[0], m[1]) * a0 + cx(m[2], m[3]) * a1;
                v[j | t] = cx(m[4], m[5]) * a0 + cx(m[6], m[7]) * a1;
            }
            break;
        }
        case F_DENSE2: {
            const uint32_t t1 = 1u << bitOf(g.t1), t2 = 1u << bitOf(g.t2);
            const uint32_t off[4] = {0, t1, t2, t1 | t2};
            for (int j = 0; j < D; j++) {
                if ((j & (t1 | t2)) || (j & cm) != cv) continue;
                cx x[4], y[4];
                for (int c = 0; c < 4; c++) x[c] = v[j | off[c]];
                for (int r = 0; r < 4; r++) {
                    y[r] = 0;
                    for (int c = 0; c < 4; c++) y[r] += fc(4 * r + c) * x[c];
                }
                for (int r = 0; r < 4; r++) v[j | off[r]] = y[r];
            }
            break;
        }
        case F_DIAG1: {
            const uint32_t t = 1u << bitOf(g.t1);
            for (int j = 0; j < D; j++)
                if ((j & cm) == cv) v[j] *= (j & t) ? fc(1) : fc(0);
            break;
        }
        case F_PARITY: {
            const uint32_t m = loc(g.m1);
            for (int j = 0; j < D; j++)
                if ((j & cm) == cv) v[j] *= (__builtin_popcount(j & m) & 1) ? fc(1) : fc(0);
            break;
        }
        case F_PAULI: {
            const uint32_t mxy = loc(g.m1), myz = loc(g.m2);
            const uint32_t top = 1u << (31 - __builtin_clz(mxy));
            for (int j = 0; j < D; j++) {
                if ((j & top) || (j & cm) != cv) continue;
                const int kk = j ^ mxy;
                const double sA = (__builtin_popcount(j & myz) & 1) ? -1 : 1;
                const double sB = (__builtin_popcount(kk & myz) & 1) ? -1 : 1;
                const cx va = v[j], vb = v[kk];
                v[j] = fc(0) * va + fc(1) * sB * vb;
                v[kk] = fc(0) * vb + fc(1) * sA * va;
            }
            break;
        }
        case F_SWAP: {
            const uint32_t t1 = 1u << bitOf(g.t1), t2 = 1u << bitOf(g.t2);
            for (int j = 0; j < D; j++) {
                if (!(j & t1) || (j & t2) || (j & cm) != cv) continue;
                std::swap(v[j], v[j ^ (t1 | t2)]);
            }
            break;
        }
        case F_DEPOL1: {
            const uint32_t k1 = 1u << bitOf(g.t1), k2 = 1u << bitOf(g.t2);
            const double c1 = g.f[0], c2 = g.f[1], c3 = g.f[2];
            for (int j = 0; j < D; j++) {
                if (j & (k1 | k2)) continue;
                const cx a00 = v[j], a11 = v[j | k1 | k2];
                v[j] = c1 * a00 + c2 * a11;
                v[j | k1] *= c3;
                v[j | k2] *= c3;
                v[j | k1 | k2] = c1 * a11 + c2 * a00;
            }
            break;
        }
        case F_DEPH1: {
            const uint32_t k1 = 1u << bitOf(g.t1), k2 = 1u << bitOf(g.t2);
            for (int j = 0; j < D; j++)
                if (((j & k1) != 0) != ((j & k2) != 0)) v[j] *= g.f[0];
            break;
        }
        default: break;
    }
}

// the BLOCK op of ops[from, to), whose qubits all lie in the BQ local
// qubits of S; its matrix is appended to `blocks` as the MFMA operands
TOp makeBlock(const GOp* ops, size_t from, size_t to, uint64_t S, const Layout& L, std::vector<float>& blocks) {
    typedef std::complex<double> cx;
    const int D = 1 << BQ;
    int q[BQ];
    for (int b = 0, qq = 0; qq < 64 && b < BQ; qq++)
        if ((S >> qq) & 1ull) q[b++] = qq;
    // U column by column: the run applied to each basis vector
    std::vector<cx> U(D * D);
    for (int c = 0; c < D; c++) {
        cx v[D];
        for (int j = 0; j < D; j++) v[j] = (j == c) ? 1.0 : 0.0;
        for (size_t i = from; i < to; i++) applySmall(ops[i], q, BQ, v);
        for (int r = 0; r < D; r++) U[r * D + c] = v[r];
    }
    // A[r][kk] of the real form [[Re U, -Im U], [Im U, Re U]], packed as
    // [K-step s][half][lane] = A[half*32 + lane%32][2s + lane/32]
    const size_t off = blocks.size();
    blocks.resize(off + 32 * 2 * 64);
    float* P = blocks.data() + off;
    for (int s = 0; s < 32; s++)
        for (int half = 0; half < 2; half++)
            for (int lane = 0; lane < 64; lane++) {
                const int r = lane % 32, kk = 2 * s + lane / 32;
                const cx u = U[r * D + (kk % 32)];
                double val;
                if (half == 0) val = kk < 32 ? u.real() : -u.imag();
                else val = kk < 32 ? u.imag() : u.real();
                P[(2 * s + half) * 64 + lane] = (float)val;
            }
    TOp t;
    std::memset(&t, 0, sizeof(t));
    t.h.kind = F_BLOCK;
    int* slots = reinterpret_cast<int*>(t.b);
    for (int b = 0; b < BQ; b++) slots[b] = L.slotOf[q[b]];
    t.b[3] = (uint64_t)off;
    return t;
}

// ops whose action is confined to one or two qubits (controls included),
// and those among them that are diagonal
bool mergeable(const GOp& g) {
    switch (g.kind) {
        case F_DENSE1: case F_DENSE2: case F_DIAG1: case F_PARITY: case F_PAULI: case F_SWAP: case F_DEPOL1:
        case F_DEPH1: break;
        default: return false;
    }
    const int n = __builtin_popcountll(supportMask(g));
    return n >= 1 && n <= 2;
}
bool diagonalOp(const GOp& g) { return g.kind == F_DIAG1 || g.kind == F_PARITY || g.kind == F_DEPH1; }

}  // namespace fuse

void fuse::mergePass(const GOp* ops, size_t num, uint64_t localMask, std::vector<GOp>& out) {
    struct Block {
        uint64_t qubits;                // local qubits
        uint64_t variants;              // non-local qubits (members are diagonal in them)
        std::vector<uint32_t> members;  // positions in ops, ascending
        bool dense;                     // some member is not diagonal
    };
    std::vector<Block> open;  // pairwise disjoint qubit sets
    auto emit = [&](const Block& b) {
        if (b.members.size() == 1 || !b.dense) {
            for (uint32_t m : b.members) out.push_back(ops[m]);
            return;
        }
        // the block's local qubits, then its variant qubits, ascending
        int q[4], nq = 0;
        for (int x = 0; x < 64; x++)
            if ((b.qubits >> x) & 1ull) q[nq++] = x;
        const int nl = nq;
        for (int x = 0; x < 64; x++)
            if ((b.variants >> x) & 1ull) q[nq++] = x;
        const int D = 1 << nq, DL = 1 << nl;
        std::complex<double> M[64];
        for (int c = 0; c < D; c++) {
            std::complex<double> v[8] = {};
            v[c] = 1.0;
            for (uint32_t m : b.members) applySmall(ops[m], q, nq, v);
            for (int r = 0; r < D; r++) M[r * D + c] = v[r];
        }
        GOp g;
        g.t1 = q[0];
        if (nq == nl) {
            g.kind = nl == 1 ? F_DENSE1 : F_DENSE2;
            g.t2 = nl == 2 ? q[1] : -1;
            for (int i = 0; i < D * D; i++) {
                g.f[2 * i] = M[i].real();
                g.f[2 * i + 1] = M[i].imag();
            }
        } else {
            // one 2x2 matrix per value w of the variant qubits: the diagonal
            // block of M at local index r | w << 1
            g.kind = F_DENSE1V;
            g.m1 = b.variants;
            for (int w = 0; w < (D >> 1); w++)
                for (int r = 0; r < DL; r++)
                    for (int c = 0; c < DL; c++) {
                        const std::complex<double> e = M[(r | w << 1) * D + (c | w << 1)];
                        g.f[8 * w + 2 * (2 * r + c)] = e.real();
                        g.f[8 * w + 2 * (2 * r + c) + 1] = e.imag();
                    }
        }
        out.push_back(g);
    };
    for (size_t i = 0; i < num; i++) {
        const GOp& g = ops[i];
        const uint64_t T = supportMask(g);
        uint64_t UL = T & localMask, UN = T & ~localMask;
        bool dense = !diagonalOp(g);
        for (const Block& b : open)
            if ((b.qubits | b.variants) & T) {
                UL |= b.qubits;
                UN |= b.variants;
                dense |= b.dense;
            }
        // a block that ends up dense becomes one matrix on its (at most two)
        // local qubits, or, with non-local qubits, one 2x2 matrix per value of
        // at most two of them, chosen per tile
        const int nl = __builtin_popcountll(UL), nn = __builtin_popcountll(UN);
        bool fits = mergeable(g);
        if (fits) {
            if (!dense) fits = nl + nn <= 2;
            else if (nn == 0) fits = nl <= 2;
            else fits = nl <= 1 && nn <= 2;
        }
        Block nb{fits ? UL : (T & localMask), fits ? UN : (T & ~localMask), {}, !diagonalOp(g)};
        for (size_t b = 0; b < open.size();) {
            if (!((open[b].qubits | open[b].variants) & T)) {
                b++;
                continue;
            }
            if (fits) {
                nb.members.insert(nb.members.end(), open[b].members.begin(), open[b].members.end());
                nb.dense |= open[b].dense;
            } else {
                emit(open[b]);
            }
            open.erase(open.begin() + b);
        }
        if (!mergeable(g)) {
            out.push_back(g);
            continue;
        }
        std::sort(nb.members.begin(), nb.members.end());
        nb.members.push_back((uint32_t)i);
        open.push_back(std::move(nb));
    }
    for (const Block& b : open) emit(b);
}

void fuse::clusterPass(const GOp* ops, size_t num, uint64_t avoid, std::vector<GOp>& out) {
    std::vector<uint32_t> rem(num), next, taken;
    for (size_t i = 0; i < num; i++) rem[i] = (uint32_t)i;
    // one round over rem: the ops taken (in order) and the rest; `keepOff`
    // qubits may not be dense targets of the round's ops
    auto round = [&](uint64_t keepOff) {
        uint64_t R = 0, skipDense = 0, skipTouch = 0;
        bool closed = false;  // an LDS op ended this round's register group
        next.clear();
        taken.clear();
        for (uint32_t i : rem) {
            const GOp& g = ops[i];
            const uint64_t d = denseMask(g), t = supportMask(g);
            bool take = (d & skipTouch) == 0 && (t & skipDense) == 0 && (d & keepOff) == 0;
            if (take && d) {
                if (isLdsOp(g)) {
                    take = !closed && R == 0;
                    closed = take;
                } else {
                    take = !closed && __builtin_popcountll(R | d) <= RB;
                }
            }
            if (take) {
                R |= d;
                taken.push_back(i);
            } else {
                skipDense |= d;
                skipTouch |= t;
                next.push_back(i);
            }
        }
        return R;
    };
    bool first = true;
    while (!rem.empty()) {
        // the first round keeps off `avoid` when it still applies a dense op,
        // so the pass can load HBM in its first register layout
        if (!(first && avoid && round(avoid) != 0)) round(0);
        first = false;
        for (uint32_t i : taken) out.push_back(ops[i]);
        rem.swap(next);
    }
}

void fuse::planPass(const GOp* ops, size_t num, int numQubits, uint64_t localMask, int lowQubits, TileArgs& a,
                    std::vector<TOp>& prog, bool phases) {
    Layout L;
    passShape(numQubits, localMask, a, L);
    const int k = a.k;
    const int low = std::min(lowQubits, k);
    const uint32_t lowSlots = (1u << low) - 1u;
    const size_t start = prog.size();
    GroupState st;
    emitGroups(ops, 0, num, L, k, low, prog, st, phases);
    if (st.firstIsLds) prog.insert(prog.begin() + start, placeholderGroup(k));
    a.firstDirect = !st.firstIsLds && (st.firstG & lowSlots) == 0;
    a.lastDirect = !st.lastIsLds && (st.lastG & lowSlots) == 0;
    a.usesLds = !a.firstDirect || !a.lastDirect || st.numGroups > 1 || st.firstIsLds || st.lastIsLds;
    a.numOps = (int)(prog.size() - start);
}

void fuse::planPassBlocks(const GOp* ops, size_t num, int numQubits, uint64_t localMask, int lowQubits,
                          int blockMinGates, TileArgs& a, std::vector<TOp>& prog, std::vector<float>& blocks) {
    Layout L;
    passShape(numQubits, localMask, a, L);
    const int k = a.k;
    const int low = std::min(lowQubits, k);
    const uint32_t lowSlots = (1u << low) - 1u;
    const size_t start = prog.size();
    GroupState st;
    bool firstItemGroup = true, lastItemGroup = true, any = false;
    size_t pending = 0;  // first op not yet emitted
    size_t i = 0;
    while (i < num) {
        // the longest run from i whose qubits fit BQ local qubits
        uint64_t S = 0;
        size_t j = i;
        for (; j < num; j++) {
            const uint64_t sup = supportMask(ops[j]);
            if (sup & ~localMask) break;
            const uint64_t u = S | sup;
            if (__builtin_popcountll(u) > BQ) break;
            S = u;
        }
        if (k >= 10 && (int)(j - i) >= blockMinGates && blockMinGates > 0) {
            if (pending < i) {
                if (!any) firstItemGroup = true;
                emitGroups(ops, pending, i, L, k, low, prog, st, false);
                any = true;
                lastItemGroup = true;
            }
            // pad S with other local qubits
            for (int qq = 0; qq < numQubits && __builtin_popcountll(S) < BQ; qq++)
                if ((localMask >> qq) & 1ull) S |= 1ull << qq;
            if (!any) firstItemGroup = false;
            prog.push_back(makeBlock(ops, i, j, S, L, blocks));
            any = true;
            lastItemGroup = false;
            i = pending = j;
        } else {
            i++;
        }
    }
    if (pending < num) {
        if (!any) firstItemGroup = true;
        emitGroups(ops, pending, num, L, k, low, prog, st, false);
        lastItemGroup = true;
    }
    if (st.firstIsLds && firstItemGroup) firstItemGroup = false;
    if (st.lastIsLds) lastItemGroup = false;
    // a pass that starts or ends with a BLOCK moves the tile through LDS in
    // the row layout; one that starts with a BLOCK still begins with a GROUP
    // op (its layout is never used), so ops[0] is always a layout
    if (!firstItemGroup) prog.insert(prog.begin() + start, placeholderGroup(k));
    a.firstDirect = firstItemGroup && (st.firstG & lowSlots) == 0;
    a.lastDirect = lastItemGroup && (st.lastG & lowSlots) == 0;
    a.usesLds = true;
    a.numOps = (int)(prog.size() - start);
}

namespace fuse {

// algorithmic real flops of one op on a state of n index bits: a complex
// multiply is 6, a complex multiply-add 8; amplitudes failing the controls
// are not counted
double opFlops(const GOp& g, int n) {
    double perAmp = 0;
    switch (g.kind) {
        case F_DENSE1: case F_DENSE1V: perAmp = 16; break;   // 4 complex MACs per pair
        case F_DENSE2: perAmp = 32; break;   // 16 complex MACs per 4 amplitudes
        case F_DIAG1: perAmp = 6; break;     // one complex multiply
        case F_PARITY: perAmp = 6; break;
        case F_PAULI: perAmp = 14; break;    // multiply + multiply-add per amplitude
        case F_DEPOL1: perAmp = 4; break;    // 16 per 4 amplitudes
        case F_DEPH1: perAmp = 1; break;     // 4 per 4 amplitudes
        case F_DENSEN: perAmp = 8.0 * (1 << g.numTargs); break;  // 2^n complex MACs per amplitude
        default: perAmp = 0; break;
    }
    return perAmp * std::ldexp(1.0, n - __builtin_popcountll(g.cMask));
}

void setLdsOffsets(TOp* prog, size_t num, TileArgs& a, int prec) {
    auto offsets = [&](const int* slots, uint32_t* out) {
        for (int v = 0; v < RA; v++) {
            uint32_t j = 0;
            for (int b = 0; b < RB; b++)
                if ((v >> b) & 1) j |= 1u << slots[b];
            out[v] = prec == QA_FP64 ? ldsByte<double>(j) : ldsByte<float>(j);
        }
    };
    int row[RB];
    for (int b = 0; b < RB; b++) row[b] = a.k - RB + b;
    offsets(row, a.rowOff);
    if (num == 0) return;
    int cur[RB];
    std::memcpy(cur, groupSlots(prog[0]), sizeof(cur));
    for (size_t i = 0; i < num; i++) {
        TOp& t = prog[i];
        uint32_t* w = reinterpret_cast<uint32_t*>(t.fx);
        if (t.h.kind == F_GROUP) {
            offsets(cur, w + RA);
            std::memcpy(cur, groupSlots(t), sizeof(cur));
            offsets(cur, w);
        } else if (t.h.kind == F_PAULI_LDS || t.h.kind == F_BLOCK || t.h.kind == F_DENSE_LDS) {
            offsets(cur, w + RA);
        }
    }
    offsets(cur, a.lastOff);
}

bool spanEligible(const GOp& g) { return (g.kind == F_PAULI || g.kind == F_PARITY) && g.cMask == 0; }

bool gateFitsTile(const GOp& g, int n, int kMax, int low) {
    const uint64_t d = denseMask(g);
    const uint64_t lowMask = (1ull << std::min(low, std::min(n, kMax))) - 1ull;
    return __builtin_popcountll(d | lowMask) <= std::min(n, kMax) && (__builtin_popcountll(d) <= RB || g.kind == F_PAULI);
}

// Groups a span pass's gadgets (in coset coordinates) into batches whose
// X/Y coordinates span at most RB dimensions, each run with one LDS sweep:
// a header record (SPAN_BATCH), a table record (S4) and the batch's
// gadgets with their X/Y and sign masks over the batch's own register
// basis (c, r); d stays over all k coordinates.
void batchSpanOps(std::vector<SpanOp>& prog, int k) {
    std::vector<SpanOp> in;
    in.swap(prog);
    size_t i = 0;
    while (i < in.size()) {
        // the batch basis in reduced echelon form (pivot: the highest bit)
        uint32_t bv[RB];
        int bp[RB];
        int nb = 0;
        auto reduce = [&](uint32_t c) {
            for (int q = 0; q < nb; q++)
                if ((c >> bp[q]) & 1u) c ^= bv[q];
            return c;
        };
        auto add = [&](uint32_t c) {
            const int p = 31 - __builtin_clz(c);
            for (int q = 0; q < nb; q++)
                if ((bv[q] >> p) & 1u) bv[q] ^= c;
            bv[nb] = c;
            bp[nb++] = p;
        };
        size_t j = i;
        for (; j < in.size(); j++) {
            if (in[j].kind != F_PAULI) continue;
            const uint32_t c = reduce(in[j].c);
            if (!c) continue;
            if (nb == RB) break;
            add(c);
        }
        // padded with unit vectors of the highest free coordinates (the
        // lanes of a wave then take the lowest ones)
        for (int x = k - 1; x >= 0 && nb < RB; x--) {
            const uint32_t c = reduce(1u << x);
            if (c) add(c);
        }
        uint32_t pivMask = 0;
        for (int q = 0; q < nb; q++) pivMask |= 1u << bp[q];
        SpanOp hdr, tab;
        std::memset(&hdr, 0, sizeof(hdr));
        std::memset(&tab, 0, sizeof(tab));
        hdr.kind = SPAN_BATCH;
        hdr.c = (uint32_t)(j - i);
        int nd = 0;
        for (int x = 0; x < k; x++)
            if (!((pivMask >> x) & 1u)) hdr.dep |= (uint64_t)x << (4 * nd++);
        hdr.r = (uint32_t)nd;
        uint32_t* S4 = reinterpret_cast<uint32_t*>(&tab);
        for (int v = 0; v < RA; v++) {
            uint32_t t = 0;
            for (int q = 0; q < RB; q++)
                if ((v >> q) & 1) t ^= bv[q];
            S4[v] = t;
        }
        prog.push_back(hdr);
        prog.push_back(tab);
        for (size_t m = i; m < j; m++) {
            SpanOp o = in[m];
            uint32_t c4 = 0, r4 = 0;
            for (int q = 0; q < RB; q++) {
                if (o.kind == F_PAULI) c4 |= ((o.c >> bp[q]) & 1u) << q;
                r4 |= (uint32_t)(__builtin_popcount(bv[q] & o.d) & 1) << q;
            }
            o.c = c4;
            o.r = r4;
            prog.push_back(o);
        }
        i = j;
    }
}

size_t spanRun(const GOp* ops, size_t num, int n, int k, int low, SpanArgs& a, std::vector<SpanOp>& prog) {
    prog.clear();
    if (k < 10 || k > 13 || n <= k || low < 0 || low > k - RB) return 0;
    // basis vectors in reduced echelon form: vec[j] has bit piv[j], which no
    // other basis vector has
    uint64_t vec[MAXQ];
    int piv[MAXQ];
    int dim = 0;
    for (int q = 0; q < low; q++) {
        vec[dim] = 1ull << q;
        piv[dim++] = q;
    }
    auto add = [&](uint64_t m) {  // m reduced (no pivot bits), nonzero
        const int p = 63 - __builtin_clzll(m);
        for (int j = 0; j < dim; j++)
            if ((vec[j] >> p) & 1ull) vec[j] ^= m;
        vec[dim] = m;
        piv[dim++] = p;
    };
    size_t run = 0;
    for (; run < num && run < (size_t)MAX_SPAN_OPS; run++) {
        const GOp& g = ops[run];
        if (!spanEligible(g)) break;
        if (g.kind != F_PAULI) continue;
        uint64_t m = g.m1;
        for (int j = 0; j < dim; j++)
            if ((m >> piv[j]) & 1ull) m ^= vec[j];
        if (!m) continue;
        if (dim == k) break;
        add(m);
    }
    if (run == 0) return 0;
    // the rest of the k dimensions: unit vectors of the lowest free qubits
    uint64_t pivMask = 0;
    for (int j = 0; j < dim; j++) pivMask |= 1ull << piv[j];
    for (int q = 0; dim < k && q < n; q++)
        if (!((pivMask >> q) & 1ull)) {
            add(1ull << q);
            pivMask |= 1ull << q;
        }
    // coordinates in pivot order: the lanes of a wave hold the lowest
    // coordinates, so the unit vectors of the low qubits come first
    for (int i = 1; i < dim; i++)
        for (int j = i; j > 0 && piv[j - 1] > piv[j]; j--) {
            std::swap(piv[j - 1], piv[j]);
            std::swap(vec[j - 1], vec[j]);
        }
    std::memset(&a, 0, sizeof(a));
    a.k = k;
    for (int j = 0; j < k; j++) a.basis[j] = vec[j];
    for (int i = 0; i < RA; i++) {
        uint64_t t = 0;
        for (int b = 0; b < RB; b++)
            if ((i >> b) & 1) t ^= vec[k - RB + b];
        a.top[i] = t;
    }
    for (int q = 0; q < n; q++)
        if (!((pivMask >> q) & 1ull)) a.repPos[a.numRep++] = q;
    a.repMask = ((n == 64) ? ~0ull : ((1ull << n) - 1ull)) & ~pivMask;
    a.numCosets = 1ull << (n - k);
    // a mask in V has coordinate j equal to its bit piv[j]; a sign mask's
    // coordinate j is its parity over b_j
    auto coords = [&](uint64_t m) {
        uint32_t c = 0;
        for (int j = 0; j < k; j++) c |= (uint32_t)((m >> piv[j]) & 1ull) << j;
        return c;
    };
    auto parities = [&](uint64_t z) {
        uint32_t d = 0;
        for (int j = 0; j < k; j++) d |= (uint32_t)(__builtin_popcountll(vec[j] & z) & 1) << j;
        return d;
    };
    for (size_t i = 0; i < run; i++) {
        const GOp& g = ops[i];
        SpanOp o;
        std::memset(&o, 0, sizeof(o));
        o.kind = g.kind;
        if (g.kind == F_PAULI) {
            o.c = coords(g.m1);
            o.d = parities(g.m2);
            o.z = g.m2;
        } else {
            o.d = parities(g.m1);
            o.z = g.m1;
        }
        std::memcpy(o.f, g.f, 4 * sizeof(double));
        prog.push_back(o);
    }
    batchSpanOps(prog, k);
    a.numOps = (int)prog.size();
    return run;
}

// launches one pass applying ops [begin, end) of `ops`, which stand for
// numGates queued gates
int launchPass(void* amps, const Queue& qu, const std::vector<GOp>& ops, size_t begin, size_t end, uint64_t localMask,
               size_t numGates) {
    TileArgs a;
    std::vector<TOp>& prog = g_prog;
    prog.clear();
    if (qu.prec == QA_FP32 && g_cfg.blockMinGates > 0) {
        g_blockData.clear();
        planPassBlocks(ops.data() + begin, end - begin, qu.numQubits, localMask, g_cfg.lowQubits, g_cfg.blockMinGates,
                       a, prog, g_blockData);
        if (!g_blockData.empty()) {
            const size_t nf = g_blockData.size();
            if (nf > BLOCK_RING_FLOATS) {
                setError("fusion: a pass holds more MFMA blocks than the block ring");
                return QA_ERR_ARG;
            }
            if (!g_blockDev) {
                if (hipHostMalloc(&g_blockHost, BLOCK_RING_FLOATS * sizeof(float), hipHostMallocDefault) != hipSuccess ||
                    hipMalloc(&g_blockDev, BLOCK_RING_FLOATS * sizeof(float)) != hipSuccess) {
                    setError("fusion: block ring allocation failed");
                    return QA_ERR_OOM;
                }
                g_blockPos = 0;
            }
            if (g_blockPos + nf > BLOCK_RING_FLOATS) {
                QA_HIP(hipStreamSynchronize(stream()));
                g_blockPos = 0;
            }
            std::memcpy(g_blockHost + g_blockPos, g_blockData.data(), nf * sizeof(float));
            QA_HIP(hipMemcpyAsync(g_blockDev + g_blockPos, g_blockHost + g_blockPos, nf * sizeof(float),
                                  hipMemcpyHostToDevice, stream()));
            a.blocks = g_blockDev + g_blockPos;
            g_blockPos += nf;
        }
    } else {
        planPass(ops.data() + begin, end - begin, qu.numQubits, localMask, g_cfg.lowQubits, a, prog, !g_cfg.exact);
    }
    const int k = a.k;

    size_t num = prog.size();
    setLdsOffsets(prog.data(), num, a, qu.prec);
    if (g_ringPos + num > RING_OPS) {
        QA_HIP(hipStreamSynchronize(stream()));
        g_ringPos = 0;
    }
    TOp* h = g_ringHost + g_ringPos;
    std::memcpy(h, prog.data(), num * sizeof(TOp));
    TOp* d = g_ringDev + g_ringPos;
    g_ringPos += num;
    QA_HIP(hipMemcpyAsync(d, h, num * sizeof(TOp), hipMemcpyHostToDevice, stream()));
    const size_t lds = a.usesLds ? ldsBytes(qu.prec, k) : 0;
    const unsigned block = 1u << (k - RB);
    const bool nt = std::min(g_cfg.lowQubits, k) >= STREAM_LOW_BIT;
    const int pf = g_cfg.prefetch;
    const void* fn = tileKernel(qu.prec, nt, pf != 0);
    if (!ldsTileAtZero(fn)) {
        setError("fusion: the tile-pass kernel has static LDS; its LDS tile would not start at address 0");
        return QA_ERR_HIP;
    }
    // the pass's generated kernel when its shape has one (checked for static
    // LDS and scratch when it was loaded), else the interpreter; a pass with
    // a 3- or 4-qubit matrix (DENSE_LDS) has only a generated kernel
    bool needsJit = false;
    for (size_t i = 0; i < num; i++) needsJit = needsJit || prog[i].h.kind == F_DENSE_LDS;
    const void* jfn = jit::kernelFor(qu.prec, nt, pf, a, prog.data(), num, needsJit);
    if (needsJit && !jfn) {
        setError("fusion: no generated kernel for a pass holding a 3- or 4-qubit matrix: %s", jit::lastError());
        return QA_ERR_HIP;
    }
    const unsigned grid = (unsigned)std::min<uint64_t>(
        a.numTiles, jfn ? jit::residentBlocks(jfn, block, lds) : residentBlocks(fn, block, lds));
    a.tileStep = 0;
    for (int i = 0; i < a.numNonLocal && i < 32; i++) a.tileStep |= (uint64_t)((grid >> i) & 1u) << a.nonLocalPos[i];
    std::pair<hipEvent_t, hipEvent_t>* ev = nullptr;
    if (g_stats.timing) {
        if (g_stats.used == g_stats.pool.size()) {
            std::pair<hipEvent_t, hipEvent_t> e;
            QA_HIP(hipEventCreate(&e.first));
            QA_HIP(hipEventCreate(&e.second));
            g_stats.pool.push_back(e);
        }
        g_stats.poolSpan.resize(g_stats.pool.size());
        g_stats.poolSpan[g_stats.used] = 0;
        ev = &g_stats.pool[g_stats.used++];
        QA_HIP(hipEventRecord(ev->first, stream()));
    }
    if (jfn) {
        const int e = jit::launch(jfn, grid, block, lds, amps, a, d);
        if (e != (int)hipSuccess) {
            setError("fusion: launch of a generated pass kernel failed: %s", hipGetErrorString((hipError_t)e));
            return QA_ERR_HIP;
        }
        jit::countLaunch();
    } else {
        launchTileAny(qu.prec, nt, pf, grid, block, lds, amps, a, d);
    }
    QA_LAUNCH_CHECK();
    if (ev) QA_HIP(hipEventRecord(ev->second, stream()));
    g_stats.passes++;
    g_stats.gates += (int64_t)numGates;
    g_stats.progOps += (int64_t)num;
    for (size_t i = 0; i < num; i++) g_stats.layouts += prog[i].h.kind == F_GROUP;
    for (size_t i = begin; i < end; i++) g_stats.flops += opFlops(ops[i], qu.numQubits);
    return QA_OK;
}

// adds the elapsed times of the timed passes so far to passMs (and, for
// span passes, spanMs) and frees their events for reuse; under g_mu
int collectTimes() {
    Stats& st = g_stats;
    for (size_t i = 0; i < st.used; i++) {
        float f = 0;
        QA_HIP(hipEventSynchronize(st.pool[i].second));
        QA_HIP(hipEventElapsedTime(&f, st.pool[i].first, st.pool[i].second));
        st.passMs += f;
        if (i < st.poolSpan.size() && st.poolSpan[i]) st.spanMs += f;
    }
    st.used = 0;
    return QA_OK;
}

// launches one span pass applying ops [begin, begin + prog.size()) of `ops`
// as planned by spanRun
int launchSpan(void* amps, const Queue& qu, const std::vector<GOp>& ops, size_t begin, size_t numGates,
               const SpanArgs& plan, const std::vector<SpanOp>& prog) {
    SpanArgs a = plan;
    const size_t bytes = prog.size() * sizeof(SpanOp);
    const size_t slots = (bytes + sizeof(TOp) - 1) / sizeof(TOp);
    if (g_ringPos + slots > RING_OPS) {
        QA_HIP(hipStreamSynchronize(stream()));
        g_ringPos = 0;
    }
    std::memcpy(g_ringHost + g_ringPos, prog.data(), bytes);
    QA_HIP(hipMemcpyAsync(g_ringDev + g_ringPos, g_ringHost + g_ringPos, bytes, hipMemcpyHostToDevice, stream()));
    const SpanOp* d = reinterpret_cast<const SpanOp*>(g_ringDev + g_ringPos);
    g_ringPos += slots;
    const unsigned block = 1u << (a.k - RB);
    const size_t lds = ldsBytes(qu.prec, a.k);
    const bool nt = std::min(g_cfg.lowQubits, a.k) >= STREAM_LOW_BIT;
    const void* fn = spanKernel(qu.prec, nt);
    if (!ldsTileAtZero(fn)) {
        setError("fusion: the span-pass kernel has static LDS");
        return QA_ERR_HIP;
    }
    const unsigned grid = (unsigned)std::min<uint64_t>(a.numCosets, residentBlocks(fn, block, lds));
    a.repStep = 0;
    for (int i = 0; i < a.numRep && i < 32; i++) a.repStep |= (uint64_t)((grid >> i) & 1u) << a.repPos[i];
    std::pair<hipEvent_t, hipEvent_t>* ev = nullptr;
    if (g_stats.timing) {
        if (g_stats.used == g_stats.pool.size()) {
            std::pair<hipEvent_t, hipEvent_t> e;
            QA_HIP(hipEventCreate(&e.first));
            QA_HIP(hipEventCreate(&e.second));
            g_stats.pool.push_back(e);
        }
        g_stats.poolSpan.resize(g_stats.pool.size());
        g_stats.poolSpan[g_stats.used] = 1;
        ev = &g_stats.pool[g_stats.used++];
        QA_HIP(hipEventRecord(ev->first, stream()));
    }
    launchSpanAny(qu.prec, nt, grid, block, lds, amps, a, d);
    QA_LAUNCH_CHECK();
    if (ev) QA_HIP(hipEventRecord(ev->second, stream()));
    g_stats.passes++;
    g_stats.gates += (int64_t)numGates;
    g_stats.spanPasses++;
    g_stats.spanGates += (int64_t)numGates;
    g_stats.progOps += (int64_t)numGates;
    for (size_t i = begin; i < begin + numGates; i++) g_stats.flops += opFlops(ops[i], qu.numQubits);
    return QA_OK;
}

// greedy planner: in queue order, extend the local set while the next gate's
// dense targets still fit the tile, then launch the run as one pass; unless
// exact, gates that do not fit are passed over, later gates that commute
// with them join the pass, and the pass's runs of gates on the same one or
// two qubits are multiplied into one matrix
int flushQueue(void* amps, Queue& qu) {
    if (qu.ops.empty()) return QA_OK;
    if (!ensureRing()) {
        setError("fusion: program ring allocation failed");
        return QA_ERR_OOM;
    }
    const int n = qu.numQubits;
    const int kMax = std::min(n, g_cfg.tileQubits);
    const int low = std::min(g_cfg.lowQubits, kMax);
    std::vector<GOp>& ops = qu.ops;
    int rc = QA_OK;
    SpanArgs span;
    std::vector<SpanOp>& spanProg = g_spanProg;
    // a span pass instead of the next tile pass when it takes more gates
    // (force: the front gadget fits no tile, so a span pass takes it alone
    // if need be)
    auto spanFirst = [&](size_t from, size_t tileGates, bool force) {
        const int ks = std::min(g_cfg.spanQubits, n - 1);
        if (ks < 10) return (size_t)0;
        const size_t r = spanRun(ops.data() + from, ops.size() - from, n, ks, std::min(low, ks - RB), span, spanProg);
        return ((force && r >= 1) || (r >= 2 && r > tileGates)) ? r : (size_t)0;
    };
    if (g_cfg.exact) {
        size_t i = 0;
        while (i < ops.size() && rc == QA_OK) {
            uint64_t L = 0;
            size_t run = takeRun(ops.data() + i, ops.size() - i, n, kMax, low, L);
            if (run == 0) {
                if (const size_t r = spanFirst(i, 0, true)) {
                    rc = launchSpan(amps, qu, ops, i, r, span, spanProg);
                    i += r;
                    continue;
                }
                setError("fusion: a gate needs more local qubits than the tile holds");
                rc = QA_ERR_ARG;
                break;
            }
            if (const size_t r = spanFirst(i, run, false)) {
                rc = launchSpan(amps, qu, ops, i, r, span, spanProg);
                i += r;
                continue;
            }
            rc = launchPass(amps, qu, ops, i, i + run, L, run);
            i += run;
        }
        ops.clear();
        return rc;
    }
    std::vector<uint32_t> picked;
    std::vector<GOp> pass, rest, merged;
    while (!ops.empty() && rc == QA_OK) {
        uint64_t L = 0;
        if (!gateFitsTile(ops[0], n, kMax, low)) {
            if (const size_t r = spanFirst(0, 0, true)) {
                rc = launchSpan(amps, qu, ops, 0, r, span, spanProg);
                ops.erase(ops.begin(), ops.begin() + (ptrdiff_t)r);
                continue;
            }
        }
        picked.clear();
        pickPass(ops.data(), ops.size(), n, kMax, low, L, picked);
        if (picked.empty()) {
            setError("fusion: a gate needs more local qubits than the tile holds");
            rc = QA_ERR_ARG;
            break;
        }
        if (const size_t r = spanFirst(0, picked.size(), false)) {
            rc = launchSpan(amps, qu, ops, 0, r, span, spanProg);
            ops.erase(ops.begin(), ops.begin() + (ptrdiff_t)r);
            continue;
        }
        pass.clear();
        rest.clear();
        size_t p = 0;
        for (size_t i = 0; i < ops.size(); i++) {
            if (p < picked.size() && picked[p] == i) {
                pass.push_back(ops[i]);
                p++;
            } else {
                rest.push_back(ops[i]);
            }
        }
        const size_t numGates = pass.size();
        merged.clear();
        mergePass(pass.data(), pass.size(), L, merged);
        pass.clear();
        clusterPass(merged.data(), merged.size(), (1ull << low) - 1ull, pass);
        rc = launchPass(amps, qu, pass, 0, pass.size(), L, numGates);
        ops.swap(rest);
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
    // a gate whose dense part cannot fit any tile runs unfused (a Pauli
    // tensor or gadget with more than RB X/Y qubits is applied to the LDS
    // copy of the tile instead of in registers), unless it is a gadget a
    // span pass takes whatever its width
    const int kMax = std::min(numQubits, g_cfg.tileQubits);
    const int lowCount = std::min(g_cfg.lowQubits, kMax);
    if (!gateFitsTile(op, numQubits, kMax, lowCount) &&
        !(spanEligible(op) && std::min(g_cfg.spanQubits, numQubits - 1) >= 10))
        return QA_OK;
    qu.ops.push_back(op);
    accepted = true;
    if ((int)qu.ops.size() >= g_cfg.maxQueue) return flushQueue(amps, qu);
    return QA_OK;
}

// the matrices of queued DENSEN gates: a device ring, filled in order; a
// wrap first applies every queue (their gates point into the ring) and
// waits for the stream
constexpr size_t MAT_RING_DOUBLES = 1u << 18;  // 2 MiB: 512 matrices of 16x16
double* g_matRing = nullptr;
size_t g_matPos = 0;

const double* stageMatrix(const void* mat, size_t count) {
    std::lock_guard<std::mutex> lk(g_mu);
    if (count > MAT_RING_DOUBLES) {
        setError("fusion: a matrix of %zu doubles exceeds the matrix ring", count);
        return nullptr;
    }
    if (!g_matRing) {
        if (hipMalloc(&g_matRing, MAT_RING_DOUBLES * sizeof(double)) != hipSuccess) {
            (void)hipGetLastError();
            g_matRing = nullptr;
            setError("fusion: matrix ring allocation failed");
            return nullptr;
        }
        g_matPos = 0;
    }
    if (g_matPos + count > MAT_RING_DOUBLES) {
        for (auto& kv : g_queues)
            if (flushQueue(kv.first, kv.second) != QA_OK) return nullptr;
        if (hipStreamSynchronize(stream()) != hipSuccess) {
            setError("fusion: stream synchronisation failed");
            return nullptr;
        }
        g_matPos = 0;
    }
    double* dst = g_matRing + g_matPos;
    if (hipMemcpyAsync(dst, mat, count * sizeof(double), hipMemcpyDefault, stream()) != hipSuccess) {
        (void)hipGetLastError();
        setError("fusion: matrix staging copy failed");
        return nullptr;
    }
    // 64-byte aligned entries
    g_matPos += (count + 7) & ~size_t(7);
    return dst;
}

int flush(void* amps) {
    std::lock_guard<std::mutex> lk(g_mu);
    auto it = g_queues.find(amps);
    if (it == g_queues.end()) return QA_OK;
    int rc = flushQueue(amps, it->second);
    if (it->second.ops.empty()) g_queues.erase(it);
    return rc;
}

bool pending(const void* amps) {
    std::lock_guard<std::mutex> lk(g_mu);
    auto it = g_queues.find(const_cast<void*>(amps));
    return it != g_queues.end() && !it->second.ops.empty();
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
    jit::shutdown();
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
    if (g_blockHost) (void)hipHostFree(g_blockHost);
    if (g_blockDev) (void)hipFree(g_blockDev);
    g_blockHost = nullptr;
    g_blockDev = nullptr;
    g_blockPos = 0;
    if (g_matRing) (void)hipFree(g_matRing);
    g_matRing = nullptr;
    g_matPos = 0;
}

// lets the tile-pass and span-pass kernels use the LDS of a 2^tileQubits
// tile (complex128 tiles are the larger ones)
int allowTileLds(int tileQubits) {
    int dev = 0, maxLds = 0;
    QA_HIP(hipGetDevice(&dev));
    QA_HIP(hipDeviceGetAttribute(&maxLds, hipDeviceAttributeMaxSharedMemoryPerBlock, dev));
    const size_t need = ldsBytes(QA_FP64, tileQubits);
    QA_ARG_CHECK(need <= (size_t)maxLds, "fusion: a %d-qubit tile needs %zu bytes of LDS, the device allows %d",
                 tileQubits, need, maxLds);
    if (need > 65536) {
        for (int v = 0; v < 8; v++) {
            const void* fn = tileKernel((v & 4) ? QA_FP64 : QA_FP32, v & 2, v & 1);
            QA_HIP(hipFuncSetAttribute(fn, hipFuncAttributeMaxDynamicSharedMemorySize, (int)need));
        }
        for (int v = 0; v < 4; v++)
            QA_HIP(hipFuncSetAttribute(spanKernel((v & 2) ? QA_FP64 : QA_FP32, v & 1),
                                       hipFuncAttributeMaxDynamicSharedMemorySize, (int)need));
    }
    return QA_OK;
}

// ---- autotuner ----------------------------------------------------------------

struct TuneChoice {
    int tile, low, span, blocks;
    double ms;
};
std::map<std::string, TuneChoice> g_tuned;  // (device, precision, qubits, gate mix) -> choice

// the gate mix of a batch: tenths of its gates per kind group
std::string tuneKey(int dev, int prec, int n, const std::vector<GOp>& ops) {
    int cnt[6] = {0, 0, 0, 0, 0, 0};
    for (const GOp& g : ops) {
        const int k = g.kind;
        const int grp = k == F_DENSE1 || k == F_DENSE1V ? 0
                        : k == F_DENSE2                  ? 1
                        : k == F_DIAG1                   ? 2
                        : k == F_SWAP                    ? 3
                        : k == F_PAULI || k == F_PARITY  ? 4
                                                         : 5;
        cnt[grp]++;
    }
    char buf[160];
    const double tot = ops.empty() ? 1.0 : (double)ops.size();
    snprintf(buf, sizeof buf, "dev%d p%d n%d mix%d.%d.%d.%d.%d.%d", dev, prec, n, (int)std::lround(10 * cnt[0] / tot),
             (int)std::lround(10 * cnt[1] / tot), (int)std::lround(10 * cnt[2] / tot),
             (int)std::lround(10 * cnt[3] / tot), (int)std::lround(10 * cnt[4] / tot),
             (int)std::lround(10 * cnt[5] / tot));
    return buf;
}

// whether every gate of the batch is taken by a queue of this shape (the
// queue runs a gate whose dense part does not fit a tile unfused)
bool batchFits(const std::vector<GOp>& ops, int n, int tile, int low, int spanQubits) {
    const bool spans = std::min(spanQubits, n - 1) >= 10;
    for (const GOp& g : ops)
        if (!gateFitsTile(g, n, tile, low) && !(spans && spanEligible(g))) return false;
    return true;
}

}  // namespace fuse
}  // namespace qa

extern "C" int qa_tuneFusion(void* amps, int flags, qa_fusionTrial* trials, int maxTrials, int* numTrials,
                             int* best) {
    using namespace qa;
    using namespace qa::fuse;
    QA_REQUIRE_INIT();
    QA_ARG_CHECK(amps && numTrials && best && (trials || maxTrials == 0) && maxTrials >= 0,
                 "qa_tuneFusion: null argument");
    *numTrials = 0;
    *best = -1;
    std::lock_guard<std::mutex> lk(g_mu);
    auto it = g_queues.find(amps);
    QA_ARG_CHECK(it != g_queues.end() && !it->second.ops.empty(),
                 "qa_tuneFusion: no gates are queued for this state (queue a representative batch first)");
    Queue& qu = it->second;
    const int n = qu.numQubits, prec = qu.prec;
    const std::vector<GOp> batch = qu.ops;
    // other states' queued gates were taken under the current shape: apply
    // them before it changes
    for (auto& kv : g_queues)
        if (kv.first != amps && !kv.second.ops.empty()) {
            const int r = flushQueue(kv.first, kv.second);
            if (r != QA_OK) return r;
        }
    int dev = 0;
    QA_HIP(hipGetDevice(&dev));
    const std::string key = tuneKey(dev, prec, n, batch);
    const Config start = g_cfg;
    auto report = [&](const TuneChoice& c) {
        if (*numTrials < maxTrials) {
            qa_fusionTrial& t = trials[*numTrials];
            t.tileQubits = c.tile;
            t.lowQubits = c.low;
            t.spanQubits = c.span;
            t.blockMinGates = c.blocks;
            t.ms = c.ms;
        }
        return (*numTrials)++;
    };
    auto apply = [&](const TuneChoice& c) {
        g_cfg.tileQubits = c.tile;
        g_cfg.lowQubits = c.low;
        g_cfg.spanQubits = c.span;
        g_cfg.blockMinGates = c.blocks;
    };
    auto cached = g_tuned.find(key);
    if (!(flags & 1) && cached != g_tuned.end()) {
        int rc = allowTileLds(cached->second.tile);
        if (rc != QA_OK) return rc;
        apply(cached->second);
        *best = report(cached->second);
        return flushQueue(amps, qu);
    }
    // a scratch copy of the state: every candidate applies the batch to it
    const size_t bytes = (size_t)(prec == QA_FP64 ? 16 : 8) << n;
    void* scratch = nullptr;
    if (hipMalloc(&scratch, bytes) != hipSuccess || !scratch) {
        (void)hipGetLastError();
        setError("qa_tuneFusion: no room for a %zu-byte scratch copy of the state", bytes);
        return QA_ERR_OOM;
    }
    const jit::Settings js = jit::settings();
    if (js.mode == 1) jit::setSettings(jit::Settings{1, 1});
    hipEvent_t e0 = nullptr, e1 = nullptr;
    int rc = QA_OK;
    std::map<std::tuple<int, int, int, int>, double> seen;
    // the batch's milliseconds under one shape: once to warm (the generated
    // kernels of its passes are compiled meanwhile and waited for), once timed
    auto time = [&](TuneChoice c) -> double {
        const auto k = std::make_tuple(c.tile, c.low, c.span, c.blocks);
        auto f = seen.find(k);
        if (f != seen.end()) return f->second;
        double ms = -1;
        if (batchFits(batch, n, c.tile, c.low, c.span) && allowTileLds(c.tile) == QA_OK) {
            apply(c);
            for (int run = 0; run < 2 && rc == QA_OK; run++) {
                Queue q;
                q.prec = prec;
                q.numQubits = n;
                q.ops = batch;
                if (run == 1) rc = hipEventRecord(e0, stream()) == hipSuccess ? QA_OK : QA_ERR_HIP;
                if (rc == QA_OK) rc = flushQueue(scratch, q);
                if (run == 0) {
                    jit::waitIdle();
                    continue;
                }
                if (rc == QA_OK) rc = hipEventRecord(e1, stream()) == hipSuccess ? QA_OK : QA_ERR_HIP;
                float f32 = 0;
                if (rc == QA_OK && (hipEventSynchronize(e1) != hipSuccess || hipEventElapsedTime(&f32, e0, e1) != hipSuccess))
                    rc = QA_ERR_HIP;
                ms = f32;
            }
        }
        c.ms = ms;
        report(c);
        seen[k] = ms;
        // QA_TUNE_VERBOSE: a line per shape timed, for long tuning runs
        if (std::getenv("QA_TUNE_VERBOSE")) {
            std::fprintf(stderr, "qa_tuneFusion: tile %d low %d span %d blocks %d: %.2f ms\n", c.tile, c.low, c.span,
                         c.blocks, ms);
            std::fflush(stderr);
        }
        return ms;
    };
    TuneChoice cur{start.tileQubits, start.lowQubits, start.spanQubits, start.blockMinGates, -1};
    if (hipEventCreate(&e0) != hipSuccess || hipEventCreate(&e1) != hipSuccess ||
        hipMemcpyAsync(scratch, amps, bytes, hipMemcpyDeviceToDevice, stream()) != hipSuccess)
        rc = QA_ERR_HIP;
    double bestMs = rc == QA_OK ? time(cur) : -1;
    // coordinate descent: tile size, then the always-local low qubits, then
    // the span-pass coset size, then (complex64) the MFMA block threshold
    const std::vector<std::vector<int>> values = {{13, 12, 11}, {2, 3, 4}, {13, 12, 0}, {0, 4}};
    for (int dim = 0; dim < 4 && rc == QA_OK; dim++) {
        if (dim == 3 && prec != QA_FP32) break;
        for (int v : values[dim]) {
            TuneChoice c = cur;
            (dim == 0 ? c.tile : dim == 1 ? c.low : dim == 2 ? c.span : c.blocks) = v;
            if (c.low > c.tile - 2 || c.tile > n) continue;
            const double ms = time(c);
            if (rc != QA_OK) break;
            if (ms >= 0 && (bestMs < 0 || ms < bestMs)) {
                bestMs = ms;
                cur = c;
            }
        }
    }
    if (e0) (void)hipEventDestroy(e0);
    if (e1) (void)hipEventDestroy(e1);
    (void)hipStreamSynchronize(stream());
    (void)hipFree(scratch);
    jit::setSettings(js);
    if (rc == QA_OK && bestMs < 0) {
        setError("qa_tuneFusion: no candidate shape takes every queued gate");
        rc = QA_ERR_ARG;
    }
    if (rc != QA_OK) {
        g_cfg = start;
        (void)allowTileLds(start.tileQubits);
        return rc;
    }
    cur.ms = bestMs;
    apply(cur);
    rc = allowTileLds(cur.tile);
    if (rc != QA_OK) return rc;
    g_tuned[key] = cur;
    for (int i = 0; i < std::min(*numTrials, maxTrials); i++) {
        const qa_fusionTrial& t = trials[i];
        if (t.tileQubits == cur.tile && t.lowQubits == cur.low && t.spanQubits == cur.span &&
            t.blockMinGates == cur.blocks)
            *best = i;
    }
    return flushQueue(amps, qu);
}

namespace qa {
namespace fuse {

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
    rc = qa::fuse::allowTileLds(tileQubits);
    if (rc != QA_OK) return rc;
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
    qa::fuse::g_stats.progOps = qa::fuse::g_stats.layouts = 0;
    qa::fuse::g_stats.spanPasses = qa::fuse::g_stats.spanGates = 0;
    qa::fuse::g_stats.passMs = qa::fuse::g_stats.spanMs = 0;
    qa::fuse::g_stats.flops = 0;
    return QA_OK;
}

int qa_getFusionStats(int64_t* passes, int64_t* gates, double* passMs, double* flops) {
    QA_REQUIRE_INIT();
    QA_FLUSH_ALL();
    std::lock_guard<std::mutex> lk(qa::fuse::g_mu);
    qa::fuse::Stats& st = qa::fuse::g_stats;
    const int rc = qa::fuse::collectTimes();
    if (rc != QA_OK) return rc;
    if (passes) *passes = st.passes;
    if (gates) *gates = st.gates;
    if (passMs) *passMs = st.passMs;
    if (flops) *flops = st.flops;
    st.passes = st.gates = 0;
    st.passMs = 0;
    st.flops = 0;
    return QA_OK;
}

int qa_getFusionProgramStats(int64_t* ops, int64_t* layouts) {
    QA_REQUIRE_INIT();
    QA_FLUSH_ALL();
    std::lock_guard<std::mutex> lk(qa::fuse::g_mu);
    qa::fuse::Stats& st = qa::fuse::g_stats;
    if (ops) *ops = st.progOps;
    if (layouts) *layouts = st.layouts;
    st.progOps = st.layouts = 0;
    return QA_OK;
}

int qa_setFusionPrefetch(int enabled) {
    QA_REQUIRE_INIT();
    QA_FLUSH_ALL();
    std::lock_guard<std::mutex> lk(qa::fuse::g_mu);
    if (enabled < 0 || enabled > 3) {
        qa::setError("qa_setFusionPrefetch: mode %d is not 0 (off), 1 (at the loop head), 2 (spread loads) or 3 "
                     "(spread loads and stores)", enabled);
        return QA_ERR_ARG;
    }
    qa::fuse::g_cfg.prefetch = enabled;
    return QA_OK;
}

int qa_setFusionSpan(int spanQubits) {
    QA_REQUIRE_INIT();
    QA_ARG_CHECK(spanQubits == 0 || (spanQubits >= 10 && spanQubits <= 13),
                 "qa_setFusionSpan: spanQubits must be 0 (off) or in [10, 13]");
    QA_FLUSH_ALL();
    std::lock_guard<std::mutex> lk(qa::fuse::g_mu);
    qa::fuse::g_cfg.spanQubits = spanQubits;
    return QA_OK;
}

int qa_getFusionSpanStats(int64_t* passes, int64_t* gates, double* passMs) {
    QA_REQUIRE_INIT();
    QA_FLUSH_ALL();
    std::lock_guard<std::mutex> lk(qa::fuse::g_mu);
    qa::fuse::Stats& st = qa::fuse::g_stats;
    const int rc = qa::fuse::collectTimes();
    if (rc != QA_OK) return rc;
    if (passes) *passes = st.spanPasses;
    if (gates) *gates = st.spanGates;
    if (passMs) *passMs = st.spanMs;
    st.spanPasses = st.spanGates = 0;
    st.spanMs = 0;
    return QA_OK;
}

int qa_setFusionExact(int exact) {
    QA_REQUIRE_INIT();
    QA_FLUSH_ALL();
    std::lock_guard<std::mutex> lk(qa::fuse::g_mu);
    qa::fuse::g_cfg.exact = exact != 0;
    return QA_OK;
}

int qa_setFusionBlocks(int minGates) {
    QA_REQUIRE_INIT();
    QA_FLUSH_ALL();
    QA_ARG_CHECK(minGates >= 0 && minGates <= 4096, "qa_setFusionBlocks: minGates must be in [0, 4096]");
    std::lock_guard<std::mutex> lk(qa::fuse::g_mu);
    qa::fuse::g_cfg.blockMinGates = minGates;
    return QA_OK;
}

int qa_flush(void* amps) {
    QA_REQUIRE_INIT();
    return amps ? qa::fuse::flush(amps) : qa::fuse::flushAll();
}

}  // extern "C"
