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

struct TOp {
    int kind;
    int p1, p2;             // local slots (tile-index bits)
    int top;                // PAULI: slot of the highest local X/Y bit
    uint64_t jcMask, jcVal;  // local part of the control condition
    uint64_t gcMask, gcVal;  // non-local part (tested against the tile base)
    uint64_t jm1, gm1;       // DIAG1 target bit / PARITY mask / PAULI maskXY (local only)
    uint64_t jm2, gm2;       // PAULI maskYZ
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
__device__ __forceinline__ uint64_t insertZero(uint64_t q, int p) {
    return ((q >> p) << (p + 1)) | (q & ((1ull << p) - 1ull));
}

template <typename T>
__device__ __forceinline__ ct<T> ldsLd(const ct<T>* tile, uint64_t j) { return tile[j]; }

template <typename T>
__global__ void __launch_bounds__(BLOCK) k_tilePass(T* __restrict__ amps, TileArgs a, const TOp* __restrict__ ops) {
    extern __shared__ unsigned char smem[];
    const int k = a.k;
    const uint64_t tileSize = 1ull << k;
    ct<T>* tile = reinterpret_cast<ct<T>*>(smem);
    uint64_t* table = reinterpret_cast<uint64_t*>(smem + (sizeof(ct<T>) << k));
    const int nt = (int)blockDim.x;
    const int tid = (int)threadIdx.x;
    const uint64_t base = deposit((uint64_t)blockIdx.x, a.nonLocalPos, a.numNonLocal);
    // global offset of tile row h (the k-low tile bits above the contiguous
    // low run) = tlo[low half of h] | thi[high half of h]
    const int hb = k - a.low, b0 = hb >> 1, b1 = hb - b0;
    uint64_t* tlo = table;
    uint64_t* thi = table + (1ull << b0);
    const uint64_t lowMask = (1ull << a.low) - 1ull, b0Mask = (1ull << b0) - 1ull;
    const int s0 = a.low + b0;
    for (uint64_t h = (uint64_t)tid; h < (1ull << b0); h += (uint64_t)nt)
        tlo[h] = deposit(h, a.localPos + a.low, b0);
    for (uint64_t h = (uint64_t)tid; h < (1ull << b1); h += (uint64_t)nt)
        thi[h] = deposit(h, a.localPos + a.low + b0, b1);
    __syncthreads();
    for (uint64_t j = (uint64_t)tid; j < tileSize; j += (uint64_t)nt)
        tile[j] = ld(amps, base | (j & lowMask) | tlo[(j >> a.low) & b0Mask] | thi[j >> s0]);
    __syncthreads();

    for (int o = 0; o < a.numOps; o++) {
        const TOp& op = ops[o];
        if ((base & op.gcMask) != op.gcVal) continue;  // uniform across the workgroup
        switch (op.kind) {
            case F_DENSE1: {
                const uint64_t tb = 1ull << op.p1;
                ct<T> m00 = {(T)op.f[0], (T)op.f[1]}, m01 = {(T)op.f[2], (T)op.f[3]};
                ct<T> m10 = {(T)op.f[4], (T)op.f[5]}, m11 = {(T)op.f[6], (T)op.f[7]};
                for (uint64_t q = (uint64_t)tid; q < (tileSize >> 1); q += (uint64_t)nt) {
                    uint64_t j0 = insertZero(q, op.p1);
                    if ((j0 & op.jcMask) != op.jcVal) continue;
                    ct<T> a0 = tile[j0], a1 = tile[j0 | tb];
                    ct<T> b0 = cmulT(m00, a0);
                    b0 = cfmaT(m01, a1, b0);
                    ct<T> b1 = cmulT(m10, a0);
                    b1 = cfmaT(m11, a1, b1);
                    tile[j0] = b0;
                    tile[j0 | tb] = b1;
                }
                break;
            }
            case F_DENSE2: {
                const int lo = op.p1 < op.p2 ? op.p1 : op.p2, hi = op.p1 < op.p2 ? op.p2 : op.p1;
                const uint64_t b1 = 1ull << op.p1, b2 = 1ull << op.p2;
                const uint64_t off[4] = {0, b1, b2, b1 | b2};
                for (uint64_t q = (uint64_t)tid; q < (tileSize >> 2); q += (uint64_t)nt) {
                    uint64_t j0 = insertZero(insertZero(q, lo), hi);
                    if ((j0 & op.jcMask) != op.jcVal) continue;
                    ct<T> v[4];
#pragma unroll
                    for (int c = 0; c < 4; c++) v[c] = tile[j0 | off[c]];
#pragma unroll
                    for (int r = 0; r < 4; r++) {
                        ct<T> acc = cmulT(ct<T>{(T)op.f[8 * r], (T)op.f[8 * r + 1]}, v[0]);
#pragma unroll
                        for (int c = 1; c < 4; c++)
                            acc = cfmaT(ct<T>{(T)op.f[8 * r + 2 * c], (T)op.f[8 * r + 2 * c + 1]}, v[c], acc);
                        tile[j0 | off[r]] = acc;
                    }
                }
                break;
            }
            case F_DIAG1: {
                ct<T> d0 = {(T)op.f[0], (T)op.f[1]}, d1 = {(T)op.f[2], (T)op.f[3]};
                const bool baseBit = (base & op.gm1) != 0;
                for (uint64_t j = (uint64_t)tid; j < tileSize; j += (uint64_t)nt) {
                    if ((j & op.jcMask) != op.jcVal) continue;
                    bool bit = op.jm1 ? ((j & op.jm1) != 0) : baseBit;
                    tile[j] = cmulT(tile[j], bit ? d1 : d0);
                }
                break;
            }
            case F_PARITY: {
                ct<T> f0 = {(T)op.f[0], (T)op.f[1]}, f1 = {(T)op.f[2], (T)op.f[3]};
                const int basePar = parity64(base & op.gm1);
                for (uint64_t j = (uint64_t)tid; j < tileSize; j += (uint64_t)nt) {
                    if ((j & op.jcMask) != op.jcVal) continue;
                    tile[j] = cmulT(tile[j], (parity64(j & op.jm1) ^ basePar) ? f1 : f0);
                }
                break;
            }
            case F_PAULI: {
                ct<T> ampFac = {(T)op.f[0], (T)op.f[1]}, pairFac = {(T)op.f[2], (T)op.f[3]};
                const int basePar = parity64(base & op.gm2);
                for (uint64_t q = (uint64_t)tid; q < (tileSize >> 1); q += (uint64_t)nt) {
                    uint64_t jA = insertZero(q, op.top);
                    if ((jA & op.jcMask) != op.jcVal) continue;
                    uint64_t jB = jA ^ op.jm1;
                    ct<T> va = tile[jA], vb = tile[jB];
                    T sA = (parity64(jA & op.jm2) ^ basePar) ? (T)-1 : (T)1;
                    T sB = (parity64(jB & op.jm2) ^ basePar) ? (T)-1 : (T)1;
                    ct<T> pB = {pairFac.re * sB, pairFac.im * sB};
                    ct<T> pA = {pairFac.re * sA, pairFac.im * sA};
                    tile[jA] = cfmaT(pB, vb, cmulT(ampFac, va));
                    tile[jB] = cfmaT(pA, va, cmulT(ampFac, vb));
                }
                break;
            }
            case F_SWAP: {
                // (p1 = 1, p2 = 0) <-> (p1 = 0, p2 = 1)
                const int lo = op.p1 < op.p2 ? op.p1 : op.p2, hi = op.p1 < op.p2 ? op.p2 : op.p1;
                const uint64_t flip = (1ull << op.p1) | (1ull << op.p2);
                for (uint64_t q = (uint64_t)tid; q < (tileSize >> 2); q += (uint64_t)nt) {
                    uint64_t j = insertZero(insertZero(q, lo), hi) | (1ull << op.p1);
                    if ((j & op.jcMask) != op.jcVal) continue;
                    ct<T> x = tile[j], y = tile[j ^ flip];
                    tile[j] = y;
                    tile[j ^ flip] = x;
                }
                break;
            }
            case F_DEPOL1: {
                const T c1 = (T)op.f[0], c2 = (T)op.f[1], c3 = (T)op.f[2];
                const int lo = op.p1 < op.p2 ? op.p1 : op.p2, hi = op.p1 < op.p2 ? op.p2 : op.p1;
                const uint64_t kb = 1ull << op.p1, bb = 1ull << op.p2;
                for (uint64_t q = (uint64_t)tid; q < (tileSize >> 2); q += (uint64_t)nt) {
                    uint64_t j00 = insertZero(insertZero(q, lo), hi);
                    ct<T> a00 = tile[j00], a01 = tile[j00 | kb], a10 = tile[j00 | bb], a11 = tile[j00 | kb | bb];
                    tile[j00] = ct<T>{fmaT(c1, a00.re, c2 * a11.re), fmaT(c1, a00.im, c2 * a11.im)};
                    tile[j00 | kb] = ct<T>{c3 * a01.re, c3 * a01.im};
                    tile[j00 | bb] = ct<T>{c3 * a10.re, c3 * a10.im};
                    tile[j00 | kb | bb] = ct<T>{fmaT(c1, a11.re, c2 * a00.re), fmaT(c1, a11.im, c2 * a00.im)};
                }
                break;
            }
            case F_DEPH1: {
                const T fac = (T)op.f[0];
                const int lo = op.p1 < op.p2 ? op.p1 : op.p2, hi = op.p1 < op.p2 ? op.p2 : op.p1;
                const uint64_t kb = 1ull << op.p1, bb = 1ull << op.p2;
                for (uint64_t q = (uint64_t)tid; q < (tileSize >> 2); q += (uint64_t)nt) {
                    uint64_t j00 = insertZero(insertZero(q, lo), hi);
                    ct<T> x = tile[j00 | kb], y = tile[j00 | bb];
                    tile[j00 | kb] = ct<T>{x.re * fac, x.im * fac};
                    tile[j00 | bb] = ct<T>{y.re * fac, y.im * fac};
                }
                break;
            }
        }
        __syncthreads();
    }

    for (uint64_t j = (uint64_t)tid; j < tileSize; j += (uint64_t)nt)
        st(amps, base | (j & lowMask) | tlo[(j >> a.low) & b0Mask] | thi[j >> s0], tile[j]);
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

// split a global mask into the tile-index part (bits remapped to slots) and
// the non-local part (kept in global positions)
void splitMask(uint64_t gm, const int* slotOf, uint64_t& local, uint64_t& nonlocal) {
    local = nonlocal = 0;
    while (gm) {
        int q = __builtin_ctzll(gm);
        gm &= gm - 1;
        if (slotOf[q] >= 0) local |= 1ull << slotOf[q];
        else nonlocal |= 1ull << q;
    }
}

TOp toTile(const GOp& g, const int* slotOf) {
    TOp t;
    std::memset(&t, 0, sizeof(t));
    t.kind = g.kind;
    t.p1 = g.t1 >= 0 ? slotOf[g.t1] : -1;
    t.p2 = g.t2 >= 0 ? slotOf[g.t2] : -1;
    uint64_t lm, gm, lv, gv;
    splitMask(g.cMask, slotOf, lm, gm);
    splitMask(g.cVal, slotOf, lv, gv);
    t.jcMask = lm;
    t.jcVal = lv;
    t.gcMask = gm;
    t.gcVal = gv;
    splitMask(g.m1, slotOf, t.jm1, t.gm1);
    splitMask(g.m2, slotOf, t.jm2, t.gm2);
    if (g.kind == F_DIAG1) {
        // target bit: local slot mask in jm1, else its global mask in gm1
        t.jm1 = t.gm1 = 0;
        if (slotOf[g.t1] >= 0) t.jm1 = 1ull << slotOf[g.t1];
        else t.gm1 = 1ull << g.t1;
    }
    if (g.kind == F_PAULI) t.top = 63 - __builtin_clzll(t.jm1);
    std::memcpy(t.f, g.f, sizeof(t.f));
    return t;
}

// dynamic LDS of one tile pass: the tile plus the two row-offset tables
size_t ldsBytes(int prec, int k, int low) {
    size_t amp = prec == QA_FP64 ? 16 : 8;
    int hb = k - low, b0 = hb >> 1, b1 = hb - b0;
    return (amp << k) + sizeof(uint64_t) * ((1ull << b0) + (1ull << b1));
}

int launchPass(void* amps, const Queue& qu, const std::vector<GOp>& ops, size_t begin, size_t end, uint64_t localMask) {
    const int n = qu.numQubits;
    TileArgs a;
    std::memset(&a, 0, sizeof(a));
    int slotOf[64];
    for (int q = 0; q < 64; q++) slotOf[q] = -1;
    int k = 0, nl = 0;
    for (int q = 0; q < n; q++) {
        if ((localMask >> q) & 1ull) {
            slotOf[q] = k;
            a.localPos[k++] = q;
        } else {
            a.nonLocalPos[nl++] = q;
        }
    }
    a.k = k;
    a.low = std::min(g_cfg.lowQubits, k);
    a.numNonLocal = nl;
    size_t num = end - begin;
    if (g_ringPos + num > RING_OPS) {
        QA_HIP(hipStreamSynchronize(stream()));
        g_ringPos = 0;
    }
    TOp* h = g_ringHost + g_ringPos;
    for (size_t i = 0; i < num; i++) h[i] = toTile(ops[begin + i], slotOf);
    TOp* d = g_ringDev + g_ringPos;
    g_ringPos += num;
    QA_HIP(hipMemcpyAsync(d, h, num * sizeof(TOp), hipMemcpyHostToDevice, stream()));
    a.numOps = (int)num;
    size_t lds = ldsBytes(qu.prec, k, a.low);
    unsigned grid = (unsigned)(1ull << nl);
    unsigned block = (unsigned)std::min<uint64_t>(BLOCK, 1ull << k);
    if (qu.prec == QA_FP64)
        k_tilePass<double><<<grid, block, lds, stream()>>>((double*)amps, a, d);
    else
        k_tilePass<float><<<grid, block, lds, stream()>>>((float*)amps, a, d);
    QA_LAUNCH_CHECK();
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
    const int kMax = std::min(n, qu.prec == QA_FP64 ? g_cfg.tileQubits : g_cfg.tileQubits + 1);
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
    return numQubits >= g_cfg.minQubits;
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
    int kMax = std::min(numQubits, prec == QA_FP64 ? g_cfg.tileQubits : g_cfg.tileQubits + 1);
    int lowCount = std::min(g_cfg.lowQubits, kMax);
    uint64_t lowMask = (1ull << lowCount) - 1ull;
    if (__builtin_popcountll(denseMask(op) | lowMask) > kMax) return QA_OK;
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
    QA_ARG_CHECK(tileQubits >= 2 && tileQubits <= 13, "qa_setFusion: tileQubits must be in [2, 13]");
    QA_ARG_CHECK(lowQubits >= 0 && lowQubits <= tileQubits - 2, "qa_setFusion: lowQubits must be in [0, tileQubits-2]");
    QA_ARG_CHECK(maxQueue >= 1 && maxQueue <= 4096, "qa_setFusion: maxQueue must be in [1, 4096]");
    {
        int dev = 0, maxLds = 0;
        QA_HIP(hipGetDevice(&dev));
        QA_HIP(hipDeviceGetAttribute(&maxLds, hipDeviceAttributeMaxSharedMemoryPerBlock, dev));
        // FP32 tiles hold one more qubit and the same bytes as FP64 tiles
        size_t need = std::max(qa::fuse::ldsBytes(QA_FP64, tileQubits, lowQubits),
                               qa::fuse::ldsBytes(QA_FP32, tileQubits + 1, lowQubits));
        QA_ARG_CHECK(need <= (size_t)maxLds, "qa_setFusion: a %d-qubit tile needs %zu bytes of LDS, the device allows %d",
                     tileQubits, need, maxLds);
        if (need > 65536) {
            QA_HIP(hipFuncSetAttribute((const void*)qa::fuse::k_tilePass<double>,
                                       hipFuncAttributeMaxDynamicSharedMemorySize, (int)need));
            QA_HIP(hipFuncSetAttribute((const void*)qa::fuse::k_tilePass<float>,
                                       hipFuncAttributeMaxDynamicSharedMemorySize, (int)need));
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

int qa_flush(void* amps) {
    QA_REQUIRE_INIT();
    return amps ? qa::fuse::flush(amps) : qa::fuse::flushAll();
}

}  // extern "C"
