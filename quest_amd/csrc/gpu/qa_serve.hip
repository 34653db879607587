// Batched circuit serving (include/quest_amd_serve.h).  No counterpart in
// QuEST, which simulates one Qureg at a time with one kernel per gate.
//
// Device side: k_batchPass applies one tile pass to each of many states in
// one launch (workgroup g finds its state's pass by binary search over the
// passes' first tile indices, then runs the shared tile body of
// qa_tile.hpp); k_batchExpec / k_batchExpecFinal compute the Pauli
// expectations of the states whose gates are done, in FP64 with compensated
// sums, chunk by chunk in an order fixed by each state alone.
//
// Host side: one engine thread per engine.  Each step admits queued
// requests into a buddy-allocated HBM arena, plans one pass per running
// request with the fusion planner (all gates at once for states of at most
// tileQubits qubits), uploads the step's program with one copy, launches the
// passes (one launch per tile size), the reductions and one copy of the
// results back, and records an event.  Two steps are in flight at a time:
// while the GPU runs one, the engine plans the next; results are delivered
// and memory is freed when a step's event completes.
#include <algorithm>
#include <array>
#include <atomic>
#include <chrono>
#include <cmath>
#include <condition_variable>
#include <cstring>
#include <deque>
#include <map>
#include <mutex>
#include <set>
#include <string>
#include <thread>
#include <unordered_map>
#include <vector>

#include "qa_common.hpp"
#include "qa_sum.hpp"
#include "qa_tile.hpp"
#include "quest_amd_serve.h"

namespace qa {
namespace serve {

using namespace fuse;
using Clock = std::chrono::steady_clock;

// ---- device ---------------------------------------------------------------

struct PassDesc {
    void* amps;
    uint64_t tileBegin;  // first workgroup-tile index of this pass in the launch
    uint32_t opsOffset;  // into the step's op array
    int initZero;        // the state starts here as |0...0> (nothing is read)
    TileArgs a;
};

struct ExpJob {
    const void* amps;
    uint64_t maskXY, maskYZ;
    uint64_t numWork;    // pairs (X/Y present) or amplitudes (Z only)
    uint64_t chunkBegin;  // first chunk of this job in the launch
    uint32_t numChunks;
    int top;             // highest X/Y bit, -1 for Z-only strings
    int numY;
    int out;             // result index
};

constexpr int EXP_CHUNK = BLOCK * 16;  // work items per chunk

template <typename T, int V>
struct RegZero {
    __device__ __forceinline__ static void run(ct<T>* r, uint64_t at, const RegLayout& L, const int* pos) {
        r[V] = ((at | offQ<V>(L, pos)) == 0) ? ct<T>{(T)1, (T)0} : ct<T>{(T)0, (T)0};
        RegZero<T, V + 1>::run(r, at, L, pos);
    }
};
template <typename T>
struct RegZero<T, RA> {
    __device__ __forceinline__ static void run(ct<T>*, uint64_t, const RegLayout&, const int*) {}
};

template <typename T>
__global__ void __launch_bounds__(512) k_batchPass(const PassDesc* __restrict__ passes, int numPasses,
                                                   const TOp* __restrict__ ops, uint64_t totalTiles) {
    extern __shared__ __align__(16) unsigned char smem[];
    ct<T>* tile = reinterpret_cast<ct<T>*>(smem);
    const uint32_t tid = threadIdx.x;
    for (uint64_t g = blockIdx.x; g < totalTiles; g += gridDim.x) {
        int lo = 0, hi = numPasses - 1;
        while (lo < hi) {
            const int mid = (lo + hi + 1) >> 1;
            if (passes[mid].tileBegin <= g) lo = mid;
            else hi = mid - 1;
        }
        const PassDesc& P = passes[lo];
        const TileArgs& a = P.a;
        const TOp* po = ops + P.opsOffset;
        T* amps = (T*)P.amps;
        const uint64_t base = deposit(g - P.tileBegin, a.nonLocalPos, a.numNonLocal);
        RegLayout row, first;
        passLayouts(row, first, a, po, tid);
        const RegLayout& L = a.firstDirect ? first : row;
        ct<T> r[RA];
        if (P.initZero) RegZero<T, 0>::run(r, base | L.tg, L, a.localPos);
        else RegIO<T, 0, 0>::hbmLoad(r, amps, base | L.tg, L, a.localPos);
        tileBody<T, 0>(amps, a, po, base, r, tile, row, first, tid);
        // the next tile's first LDS writes must not overtake this tile's reads
        if (a.usesLds) __syncthreads();
    }
}

template <typename T>
__device__ __forceinline__ cd ldc(const T* a, uint64_t i) {
    return cd{(double)a[2 * i], (double)a[2 * i + 1]};
}

// partial sums (re, im) of each chunk of each job
template <typename T>
__global__ void __launch_bounds__(BLOCK) k_batchExpec(const ExpJob* __restrict__ jobs, int numJobs,
                                                     uint64_t totalChunks, double* __restrict__ partials) {
    for (uint64_t c = blockIdx.x; c < totalChunks; c += gridDim.x) {
        int lo = 0, hi = numJobs - 1;
        while (lo < hi) {
            const int mid = (lo + hi + 1) >> 1;
            if (jobs[mid].chunkBegin <= c) lo = mid;
            else hi = mid - 1;
        }
        const ExpJob& J = jobs[lo];
        const T* a = (const T*)J.amps;
        const uint64_t w0 = (c - J.chunkBegin) * (uint64_t)EXP_CHUNK;
        const uint64_t w1 = std::min<uint64_t>(w0 + EXP_CHUNK, J.numWork);
        KB acc[2] = {{0, 0}, {0, 0}};
        for (uint64_t w = w0 + threadIdx.x; w < w1; w += BLOCK) {
            if (J.top < 0) {
                // Z-only string: sign(n & maskZ) |a_n|^2
                const cd x = ldc(a, w);
                const double s = parity64(w & J.maskYZ) ? -1.0 : 1.0;
                kadd(acc[0], s * (x.re * x.re + x.im * x.im));
            } else {
                // one pair (n, j = n ^ maskXY) with the top X/Y bit of n clear:
                // sign(j) conj(a_n) a_j + sign(n) conj(a_j) a_n
                const uint64_t n = ((w >> J.top) << (J.top + 1)) | (w & ((1ull << J.top) - 1ull));
                const uint64_t j = n ^ J.maskXY;
                const cd x = ldc(a, n), y = ldc(a, j);
                const cd t1 = cmul(cconj(x), y), t2 = cmul(cconj(y), x);
                const double s1 = parity64(j & J.maskYZ) ? -1.0 : 1.0;
                const double s2 = parity64(n & J.maskYZ) ? -1.0 : 1.0;
                kadd(acc[0], s1 * t1.re + s2 * t2.re);
                kadd(acc[1], s1 * t1.im + s2 * t2.im);
            }
        }
        blockMerge<2>(acc);
        if (threadIdx.x == 0) {
            partials[4 * c + 0] = acc[0].s;
            partials[4 * c + 1] = acc[0].c;
            partials[4 * c + 2] = acc[1].s;
            partials[4 * c + 3] = acc[1].c;
        }
        __syncthreads();  // blockMerge's LDS is reused by the next chunk
    }
}

// one workgroup per job: its chunks merged in order, times i^numY, real part
__global__ void __launch_bounds__(BLOCK) k_batchExpecFinal(const ExpJob* __restrict__ jobs,
                                                          const double* __restrict__ partials,
                                                          double* __restrict__ out) {
    const ExpJob& J = jobs[blockIdx.x];
    KB acc[2] = {{0, 0}, {0, 0}};
    for (uint32_t p = threadIdx.x; p < J.numChunks; p += BLOCK) {
        const double* q = partials + 4 * (J.chunkBegin + p);
        kmerge(acc[0], KB{q[0], q[1]});
        kmerge(acc[1], KB{q[2], q[3]});
    }
    blockMerge<2>(acc);
    if (threadIdx.x == 0) {
        const double re = acc[0].s + acc[0].c, im = acc[1].s + acc[1].c;
        // Re(i^numY (re + i im))
        double v;
        switch (J.numY & 3) {
            case 0: v = re; break;
            case 1: v = -im; break;
            case 2: v = -re; break;
            default: v = im; break;
        }
        out[J.out] = v;
    }
}

// ---- host -------------------------------------------------------------------

// power-of-two blocks over [0, size): top-level blocks are the binary
// digits of size, largest first, so every block is aligned to its size
class Buddy {
public:
    void init(uint64_t size, int minOrder) {
        minOrder_ = minOrder;
        free_.assign(64, {});
        tops_.clear();
        uint64_t off = 0;
        for (int o = 63; o >= minOrder; o--)
            if ((size >> o) & 1ull) {
                free_[o].insert(off);
                tops_.push_back({off, o});
                off += 1ull << o;
            }
    }
    int orderFor(uint64_t bytes) const {
        int o = minOrder_;
        while ((1ull << o) < bytes) o++;
        return o;
    }
    int largest() const {
        for (int o = 63; o >= minOrder_; o--)
            if (!free_[o].empty()) return o;
        return -1;
    }
    int maxOrder() const { return tops_.empty() ? -1 : tops_.front().second; }
    bool alloc(int order, uint64_t& off) {
        int o = order;
        while (o < 64 && free_[o].empty()) o++;
        if (o >= 64) return false;
        off = *free_[o].begin();
        free_[o].erase(free_[o].begin());
        while (o > order) {
            o--;
            free_[o].insert(off + (1ull << o));
        }
        used_[off] = order;
        return true;
    }
    void release(uint64_t off) {
        auto it = used_.find(off);
        if (it == used_.end()) return;
        int o = it->second;
        used_.erase(it);
        const int top = topOrder(off);
        while (o < top) {
            const uint64_t buddy = off ^ (1ull << o);
            auto f = free_[o].find(buddy);
            if (f == free_[o].end()) break;
            free_[o].erase(f);
            off = std::min(off, buddy);
            o++;
        }
        free_[o].insert(off);
    }

private:
    int topOrder(uint64_t off) const {
        for (auto& t : tops_)
            if (off >= t.first && off < t.first + (1ull << t.second)) return t.second;
        return minOrder_;
    }
    int minOrder_ = 0;
    std::vector<std::set<uint64_t>> free_;
    std::vector<std::pair<uint64_t, int>> tops_;
    std::unordered_map<uint64_t, int> used_;
};

enum Phase { QUEUED, RUNNING, REDUCING, DONE, FAILED };

struct Request {
    int64_t id = 0;
    int n = 0;
    std::vector<GOp> ops;  // not yet applied
    bool firstPass = true;
    std::vector<std::array<uint64_t, 3>> paulis;
    uint64_t off = 0;
    void* amps = nullptr;
    Phase phase = QUEUED;
    int outBase = 0;
    std::vector<double> result;
    Clock::time_point submitted, finished;
};

struct StepBuf {
    unsigned char* host = nullptr;  // pinned: [passes][ops][jobs]
    unsigned char* dev = nullptr;
    size_t cap = 0;
    double* partials = nullptr;
    size_t partialsCap = 0;  // chunks
    double* outDev = nullptr;
    double* outHost = nullptr;
    size_t outCap = 0;
    hipEvent_t done = nullptr;
    // this slot's executable graphs (graphs on), one per launch signature:
    // (upload, number of pass launches, reduction) -> graph
    std::map<int, hipGraphExec_t> execs;
    std::vector<Request*> finishing;
};

constexpr int NSTEP = 3;

struct Engine {
    int prec = QA_FP32, tileQ = 12, lowQ = 4, maxActive = 256;
    size_t ampBytes = 8;
    int device = 0;
    hipStream_t stream = nullptr;
    unsigned char* arena = nullptr;
    size_t arenaBytes = 0;
    Buddy buddy;

    std::mutex mu;
    std::condition_variable cvWork, cvDone;
    std::deque<Request*> submitted;                 // under mu
    std::unordered_map<int64_t, Request*> byId;     // under mu
    int64_t nextId = 1;                             // under mu
    bool stop = false;                              // under mu
    std::string error;                              // under mu

    std::deque<Request*> queued;  // engine thread
    std::vector<Request*> running;
    StepBuf steps[NSTEP];
    int stepHead = 0;
    std::deque<int> inflight;
    std::thread th;

    std::atomic<bool> graphs{false};  // launch each step as a captured HIP graph
    std::atomic<int64_t> nSteps{0}, nPasses{0}, nAdmitted{0}, nCompleted{0}, maxRunning{0};
    std::atomic<int64_t> nGraphLaunches{0}, nGraphInstances{0}, nGraphUpdates{0};
    std::atomic<int64_t> planNs{0}, waitNs{0};  // engine thread: planning + launching, waiting for steps
};

bool hipOk(hipError_t e, std::string& err, const char* what) {
    if (e == hipSuccess) return true;
    err = std::string(what) + ": " + hipGetErrorString(e);
    (void)hipGetLastError();
    return false;
}

// grows a step buffer's host/device program area to at least `bytes`
bool reserve(StepBuf& b, size_t bytes, size_t chunks, size_t outs, std::string& err) {
    if (bytes > b.cap) {
        if (b.host) (void)hipHostFree(b.host);
        if (b.dev) (void)hipFree(b.dev);
        b.host = b.dev = nullptr;
        b.cap = 0;
        size_t cap = std::max<size_t>(bytes * 2, 1 << 20);
        if (!hipOk(hipHostMalloc(&b.host, cap, hipHostMallocDefault), err, "serve: pinned program buffer")) return false;
        if (!hipOk(hipMalloc(&b.dev, cap), err, "serve: device program buffer")) return false;
        b.cap = cap;
    }
    if (chunks > b.partialsCap) {
        if (b.partials) (void)hipFree(b.partials);
        b.partials = nullptr;
        size_t cap = std::max<size_t>(chunks * 2, 4096);
        if (!hipOk(hipMalloc(&b.partials, cap * 4 * sizeof(double)), err, "serve: reduction partials")) return false;
        b.partialsCap = cap;
    }
    if (outs > b.outCap) {
        if (b.outDev) (void)hipFree(b.outDev);
        if (b.outHost) (void)hipHostFree(b.outHost);
        b.outDev = b.outHost = nullptr;
        size_t cap = std::max<size_t>(outs * 2, 1024);
        if (!hipOk(hipMalloc(&b.outDev, cap * sizeof(double)), err, "serve: results")) return false;
        if (!hipOk(hipHostMalloc(&b.outHost, cap * sizeof(double), hipHostMallocDefault), err, "serve: pinned results"))
            return false;
        b.outCap = cap;
    }
    return true;
}

template <typename T>
const void* passKernel() {
    return (const void*)k_batchPass<T>;
}

// Captures the step's launches (`issue`) into a graph and launches it as
// one of the slot's executable graphs: the one of the step's launch
// signature, updated in place (hipGraphExecUpdate: new grid sizes,
// arguments and copy sizes), else a new instantiation.  The slot is never
// in flight here (the engine retires a slot's step before reusing it).  The
// graph holds copies and kernels only.
template <typename Issue>
bool launchGraph(Engine& E, StepBuf& sb, int signature, Issue& issue, std::string& err) {
    if (!hipOk(hipStreamBeginCapture(E.stream, hipStreamCaptureModeThreadLocal), err, "serve: begin capture"))
        return false;
    const bool ok = issue();
    hipGraph_t g = nullptr;
    const hipError_t ec = hipStreamEndCapture(E.stream, &g);
    if (!ok || ec != hipSuccess) {
        if (g) (void)hipGraphDestroy(g);
        if (ok) hipOk(ec, err, "serve: end capture");
        return false;
    }
    bool ready = false;
    hipGraphExec_t& exec = sb.execs[signature];
    if (exec) {
        hipGraphNode_t errNode = nullptr;
        hipGraphExecUpdateResult res = hipGraphExecUpdateError;
        if (hipGraphExecUpdate(exec, g, &errNode, &res) == hipSuccess && res == hipGraphExecUpdateSuccess) {
            ready = true;
            E.nGraphUpdates++;
        } else {
            (void)hipGetLastError();
            (void)hipGraphExecDestroy(exec);
            exec = nullptr;
        }
    }
    if (!ready) {
        if (!hipOk(hipGraphInstantiate(&exec, g, nullptr, nullptr, 0), err, "serve: graph instantiation")) {
            exec = nullptr;
            (void)hipGraphDestroy(g);
            return false;
        }
        E.nGraphInstances++;
    }
    (void)hipGraphDestroy(g);
    if (!hipOk(hipGraphLaunch(exec, E.stream), err, "serve: graph launch")) return false;
    E.nGraphLaunches++;
    return true;
}

// plans and launches one step; false with `err` set on failure
bool launchStep(Engine& E, StepBuf& sb, std::string& err) {
    sb.finishing.clear();
    std::map<int, std::vector<PassDesc>> byK;
    std::vector<TOp> prog;
    std::vector<ExpJob> jobs;
    std::vector<uint32_t> picked;
    std::vector<GOp> pass, rest, merged;
    uint64_t chunks = 0;
    int outs = 0;
    for (Request* r : E.running) {
        if (r->phase != RUNNING) continue;
        if (!r->ops.empty()) {
            const int kMax = std::min(r->n, E.tileQ);
            const int low = std::min(E.lowQ, kMax);
            uint64_t L = 0;
            picked.clear();
            pickPass(r->ops.data(), r->ops.size(), r->n, kMax, low, L, picked);
            if (picked.empty()) {
                err = "serve: a gate needs more local qubits than the tile holds";
                return false;
            }
            pass.clear();
            rest.clear();
            size_t p = 0;
            for (size_t i = 0; i < r->ops.size(); i++) {
                if (p < picked.size() && picked[p] == i) {
                    pass.push_back(r->ops[i]);
                    p++;
                } else {
                    rest.push_back(r->ops[i]);
                }
            }
            PassDesc pd;
            std::memset(&pd, 0, sizeof(pd));
            pd.amps = r->amps;
            pd.initZero = r->firstPass ? 1 : 0;
            pd.opsOffset = (uint32_t)prog.size();
            merged.clear();
            mergePass(pass.data(), pass.size(), L, merged);
            pass.clear();
            clusterPass(merged.data(), merged.size(), (1ull << low) - 1ull, pass);
            planPass(pass.data(), pass.size(), r->n, L, E.lowQ, pd.a, prog, true);
            setLdsOffsets(prog.data() + pd.opsOffset, prog.size() - pd.opsOffset, pd.a, E.prec);
            r->firstPass = false;
            r->ops.swap(rest);
            byK[pd.a.k].push_back(pd);
        }
        if (r->ops.empty()) {
            r->phase = REDUCING;
            r->outBase = outs;
            for (auto& p : r->paulis) {
                ExpJob j;
                std::memset(&j, 0, sizeof(j));
                j.amps = r->amps;
                const uint64_t xy = p[0] | p[1];
                j.maskXY = xy;
                j.maskYZ = p[1] | p[2];
                j.top = xy ? 63 - __builtin_clzll(xy) : -1;
                j.numY = __builtin_popcountll(p[1]);
                j.numWork = xy ? (1ull << (r->n - 1)) : (1ull << r->n);
                j.numChunks = (uint32_t)((j.numWork + EXP_CHUNK - 1) / EXP_CHUNK);
                j.chunkBegin = chunks;
                chunks += j.numChunks;
                j.out = outs++;
                jobs.push_back(j);
            }
            sb.finishing.push_back(r);
        }
    }
    size_t numPasses = 0;
    for (auto& kv : byK) numPasses += kv.second.size();
    const size_t passBytes = numPasses * sizeof(PassDesc);
    const size_t opBytes = prog.size() * sizeof(TOp);
    const size_t jobBytes = jobs.size() * sizeof(ExpJob);
    if (!reserve(sb, passBytes + opBytes + jobBytes + 256, chunks, (size_t)outs, err)) return false;
    // layout: passes (grouped by k), ops, jobs
    size_t at = 0;
    std::vector<std::pair<int, size_t>> groupAt;  // (k, index of first pass)
    for (auto& kv : byK) {
        groupAt.push_back({kv.first, at / sizeof(PassDesc)});
        uint64_t tiles = 0;
        for (auto& pd : kv.second) {
            pd.tileBegin = tiles;
            tiles += pd.a.numTiles;
        }
        std::memcpy(sb.host + at, kv.second.data(), kv.second.size() * sizeof(PassDesc));
        at += kv.second.size() * sizeof(PassDesc);
    }
    const size_t opsAt = (at + 63) & ~size_t(63);  // ops are 64-byte aligned
    if (opBytes) std::memcpy(sb.host + opsAt, prog.data(), opBytes);
    const size_t jobsAt = (opsAt + opBytes + 63) & ~size_t(63);
    if (jobBytes) std::memcpy(sb.host + jobsAt, jobs.data(), jobBytes);
    const size_t total = jobsAt + jobBytes;

    // launch shapes, worked out before any launch (a captured step may not
    // query the runtime for them)
    struct PassLaunch {
        const PassDesc* dp;
        int num;
        uint64_t tiles;
        unsigned grid, block;
        size_t lds;
    };
    std::vector<PassLaunch> launches;
    const void* fn = E.prec == QA_FP64 ? passKernel<double>() : passKernel<float>();
    if (!groupAt.empty() && !ldsTileAtZero(fn)) {
        err = "serve: the pass kernel has static LDS; its LDS tile would not start at address 0";
        return false;
    }
    for (auto& g : groupAt) {
        const int k = g.first;
        const std::vector<PassDesc>& v = byK[k];
        uint64_t tiles = 0;
        bool lds = false;
        for (auto& pd : v) {
            tiles += pd.a.numTiles;
            lds = lds || pd.a.usesLds;
        }
        PassLaunch l;
        l.block = 1u << (k - RB);
        l.lds = lds ? (E.ampBytes << k) : 0;
        l.grid = (unsigned)std::min<uint64_t>(tiles, residentBlocks(fn, l.block, l.lds));
        l.dp = reinterpret_cast<const PassDesc*>(sb.dev) + g.second;
        l.num = (int)v.size();
        l.tiles = tiles;
        launches.push_back(l);
    }
    const TOp* devOps = reinterpret_cast<const TOp*>(sb.dev + opsAt);
    const ExpJob* dj = reinterpret_cast<const ExpJob*>(sb.dev + jobsAt);
    const unsigned expGrid = (unsigned)std::min<uint64_t>(chunks, 8192);

    // the step's copies and launches on the engine stream, in order
    auto issue = [&]() -> bool {
        if (total && !hipOk(hipMemcpyAsync(sb.dev, sb.host, total, hipMemcpyHostToDevice, E.stream), err,
                            "serve: upload"))
            return false;
        for (const PassLaunch& l : launches) {
            if (E.prec == QA_FP64)
                k_batchPass<double><<<l.grid, l.block, l.lds, E.stream>>>(l.dp, l.num, devOps, l.tiles);
            else
                k_batchPass<float><<<l.grid, l.block, l.lds, E.stream>>>(l.dp, l.num, devOps, l.tiles);
            if (!hipOk(hipGetLastError(), err, "serve: pass launch")) return false;
        }
        if (!jobs.empty()) {
            if (E.prec == QA_FP64)
                k_batchExpec<double><<<expGrid, BLOCK, 0, E.stream>>>(dj, (int)jobs.size(), chunks, sb.partials);
            else
                k_batchExpec<float><<<expGrid, BLOCK, 0, E.stream>>>(dj, (int)jobs.size(), chunks, sb.partials);
            if (!hipOk(hipGetLastError(), err, "serve: expectation launch")) return false;
            k_batchExpecFinal<<<(unsigned)jobs.size(), BLOCK, 0, E.stream>>>(dj, sb.partials, sb.outDev);
            if (!hipOk(hipGetLastError(), err, "serve: expectation merge launch")) return false;
            if (!hipOk(hipMemcpyAsync(sb.outHost, sb.outDev, (size_t)outs * sizeof(double), hipMemcpyDeviceToHost,
                                      E.stream),
                       err, "serve: result copy"))
                return false;
        }
        return true;
    };
    const bool any = total || !launches.empty() || !jobs.empty();
    if (any && E.graphs.load()) {
        const int signature = (total ? 1 : 0) | (jobs.empty() ? 0 : 2) | ((int)launches.size() << 2);
        if (!launchGraph(E, sb, signature, issue, err)) return false;
    } else if (any && !issue()) {
        return false;
    }
    for (const PassLaunch& l : launches) E.nPasses += (int64_t)l.num;
    if (!hipOk(hipEventRecord(sb.done, E.stream), err, "serve: step event")) return false;
    E.nSteps++;
    return true;
}

// waits for the oldest step in flight, delivers its results, frees memory
bool retireStep(Engine& E, std::string& err) {
    const int idx = E.inflight.front();
    E.inflight.pop_front();
    StepBuf& sb = E.steps[idx];
    const Clock::time_point t0 = Clock::now();
    if (!hipOk(hipEventSynchronize(sb.done), err, "serve: step wait")) return false;
    const Clock::time_point now = Clock::now();
    E.waitNs += std::chrono::duration_cast<std::chrono::nanoseconds>(now - t0).count();
    {
        std::lock_guard<std::mutex> lk(E.mu);
        for (Request* r : sb.finishing) {
            r->result.assign(sb.outHost + r->outBase, sb.outHost + r->outBase + r->paulis.size());
            r->finished = now;
            r->phase = DONE;
            E.buddy.release(r->off);
            r->amps = nullptr;
        }
    }
    E.nCompleted += (int64_t)sb.finishing.size();
    E.running.erase(std::remove_if(E.running.begin(), E.running.end(), [](Request* r) { return r->phase == DONE; }),
                    E.running.end());
    sb.finishing.clear();
    E.cvDone.notify_all();
    return true;
}

void failAll(Engine& E, const std::string& err) {
    std::lock_guard<std::mutex> lk(E.mu);
    E.error = err;
    for (auto& kv : E.byId)
        if (kv.second->phase != DONE) kv.second->phase = FAILED;
    E.stop = true;
    E.cvDone.notify_all();
}

void engineLoop(Engine* e) {
    Engine& E = *e;
    std::string err;
    if (hipSetDevice(E.device) != hipSuccess) {
        failAll(E, "serve: cannot bind the engine thread to its device");
        return;
    }
    for (;;) {
        {
            std::unique_lock<std::mutex> lk(E.mu);
            while (!E.submitted.empty()) {
                E.queued.push_back(E.submitted.front());
                E.submitted.pop_front();
            }
            if (E.queued.empty() && E.running.empty() && E.inflight.empty()) {
                if (E.stop) break;
                E.cvWork.wait(lk, [&] { return E.stop || !E.submitted.empty(); });
                continue;
            }
        }
        // admission, first come first served, by slots and arena memory
        while (!E.queued.empty() && (int)E.running.size() < E.maxActive) {
            Request* r = E.queued.front();
            uint64_t off = 0;
            bool ok;
            {
                std::lock_guard<std::mutex> lk(E.mu);
                ok = E.buddy.alloc(E.buddy.orderFor(E.ampBytes << r->n), off);
            }
            if (!ok) break;
            r->off = off;
            r->amps = E.arena + off;
            r->phase = RUNNING;
            E.running.push_back(r);
            E.queued.pop_front();
            E.nAdmitted++;
        }
        E.maxRunning = std::max<int64_t>(E.maxRunning.load(), (int64_t)E.running.size());
        const bool work = std::any_of(E.running.begin(), E.running.end(), [](Request* r) { return r->phase == RUNNING; });
        if (work) {
            if ((int)E.inflight.size() == NSTEP && !retireStep(E, err)) break;
            StepBuf& sb = E.steps[E.stepHead];
            const Clock::time_point t0 = Clock::now();
            if (!launchStep(E, sb, err)) break;
            E.planNs += std::chrono::duration_cast<std::chrono::nanoseconds>(Clock::now() - t0).count();
            E.inflight.push_back(E.stepHead);
            E.stepHead = (E.stepHead + 1) % NSTEP;
        }
        // keep at most two steps in flight; with nothing new to plan, drain
        if (!E.inflight.empty() && (!work || E.inflight.size() >= 2))
            if (!retireStep(E, err)) break;
    }
    if (!err.empty()) failAll(E, err);
}

void toMatrix(const qa_serveGate& g, double* m) {
    const double h = 1.0 / std::sqrt(2.0);
    const double c = std::cos(g.angle / 2), s = std::sin(g.angle / 2);
    std::memset(m, 0, 8 * sizeof(double));
    switch (g.kind) {
        case QA_SG_MATR1: std::memcpy(m, g.m, 8 * sizeof(double)); break;
        case QA_SG_H: m[0] = h; m[2] = h; m[4] = h; m[6] = -h; break;
        case QA_SG_RX: m[0] = c; m[3] = -s; m[5] = -s; m[6] = c; break;
        case QA_SG_RY: m[0] = c; m[2] = -s; m[4] = s; m[6] = c; break;
        case QA_SG_RZ: m[0] = c; m[1] = -s; m[6] = c; m[7] = s; break;
        case QA_SG_PHASE: m[0] = 1; m[6] = std::cos(g.angle); m[7] = std::sin(g.angle); break;
        default: break;
    }
}

// b <- a * b for 2x2 complex row-major {re,im}
void mul2(const double* a, double* b) {
    double r[8];
    for (int i = 0; i < 2; i++)
        for (int j = 0; j < 2; j++) {
            double re = 0, im = 0;
            for (int k = 0; k < 2; k++) {
                const double ar = a[4 * i + 2 * k], ai = a[4 * i + 2 * k + 1];
                const double br = b[4 * k + 2 * j], bi = b[4 * k + 2 * j + 1];
                re += ar * br - ai * bi;
                im += ar * bi + ai * br;
            }
            r[4 * i + 2 * j] = re;
            r[4 * i + 2 * j + 1] = im;
        }
    std::memcpy(b, r, sizeof(r));
}

// the circuit as fusable queue ops; consecutive one-qubit gates on the same
// qubit are multiplied into one matrix (in FP64) first
bool translate(const qa_serveCircuit& c, std::vector<GOp>& ops, std::string& err) {
    const int n = c.numQubits;
    for (int i = 0; i < c.numGates; i++) {
        const qa_serveGate& g = c.gates[i];
        auto qok = [&](int q) { return q >= 0 && q < n; };
        GOp op;
        switch (g.kind) {
            case QA_SG_MATR1: case QA_SG_H: case QA_SG_RX: case QA_SG_RY: case QA_SG_RZ: case QA_SG_PHASE: {
                if (!qok(g.t1)) {
                    err = "serve: gate qubit out of range";
                    return false;
                }
                double m[8];
                toMatrix(g, m);
                if (!ops.empty() && ops.back().kind == F_DENSE1 && ops.back().t1 == g.t1 && ops.back().cMask == 0) {
                    mul2(m, ops.back().f);
                    continue;
                }
                op.kind = F_DENSE1;
                op.t1 = g.t1;
                std::memcpy(op.f, m, sizeof(m));
                break;
            }
            case QA_SG_CZ:
                if (!qok(g.t1) || !qok(g.t2) || g.t1 == g.t2) {
                    err = "serve: CZ qubits invalid";
                    return false;
                }
                // phase-only diagonal on t2 over amplitudes with both bits set
                op.kind = F_DIAG1;
                op.t1 = g.t2;
                op.cMask = op.cVal = (1ull << g.t1) | (1ull << g.t2);
                op.f[0] = 1;
                op.f[2] = -1;
                break;
            case QA_SG_CX:
                if (!qok(g.t1) || !qok(g.t2) || g.t1 == g.t2) {
                    err = "serve: CX qubits invalid";
                    return false;
                }
                op.kind = F_DENSE1;
                op.t1 = g.t2;
                op.cMask = op.cVal = 1ull << g.t1;
                op.f[2] = 1;
                op.f[4] = 1;
                break;
            default:
                err = "serve: unknown gate kind";
                return false;
        }
        ops.push_back(op);
    }
    if (ops.empty()) {
        // an identity diagonal, so the state is still created by one pass
        GOp op;
        op.kind = F_DIAG1;
        op.t1 = 0;
        op.f[0] = 1;
        op.f[2] = 1;
        ops.push_back(op);
    }
    return true;
}

}  // namespace serve
}  // namespace qa

using namespace qa;
using namespace qa::serve;

extern "C" {

int qa_serveCreate(void** engine, size_t arenaBytes, int maxActive, int tileQubits, int lowQubits, int precision) {
    QA_REQUIRE_INIT();
    QA_ARG_CHECK(engine != nullptr, "qa_serveCreate: null engine pointer");
    *engine = nullptr;
    QA_ARG_CHECK(precision == QA_FP32 || precision == QA_FP64, "qa_serveCreate: precision must be 1 or 2");
    QA_ARG_CHECK(tileQubits >= RB && tileQubits <= 13, "qa_serveCreate: tileQubits must be in [4, 13]");
    QA_ARG_CHECK(lowQubits >= 0 && lowQubits <= tileQubits - 2, "qa_serveCreate: lowQubits must be in [0, tileQubits-2]");
    QA_ARG_CHECK(maxActive >= 1 && maxActive <= 65536, "qa_serveCreate: maxActive must be in [1, 65536]");
    const size_t ampBytes = precision == QA_FP64 ? 16 : 8;
    const size_t minBlock = ampBytes << RB;
    QA_ARG_CHECK(arenaBytes >= minBlock, "qa_serveCreate: arena smaller than one state");
    Engine* E = new Engine();
    E->prec = precision;
    E->ampBytes = ampBytes;
    E->tileQ = tileQubits;
    E->lowQ = lowQubits;
    E->maxActive = maxActive;
    const size_t arena = arenaBytes & ~(minBlock - 1);
    E->arenaBytes = arena;
    auto fail = [&](int rc) {
        if (E->arena) (void)hipFree(E->arena);
        if (E->stream) (void)hipStreamDestroy(E->stream);
        for (auto& s : E->steps)
            if (s.done) (void)hipEventDestroy(s.done);
        delete E;
        return rc;
    };
    if (hipGetDevice(&E->device) != hipSuccess) return fail((setError("qa_serveCreate: no device"), QA_ERR_HIP));
    hipError_t e = hipMalloc(&E->arena, arena);
    if (e != hipSuccess) {
        (void)hipGetLastError();
        setError("qa_serveCreate: the %zu-byte arena does not fit in device memory", arena);
        return fail(QA_ERR_OOM);
    }
    if (hipStreamCreateWithFlags(&E->stream, hipStreamNonBlocking) != hipSuccess)
        return fail((setError("qa_serveCreate: stream creation failed"), QA_ERR_HIP));
    for (auto& s : E->steps)
        if (hipEventCreateWithFlags(&s.done, hipEventDisableTiming) != hipSuccess)
            return fail((setError("qa_serveCreate: event creation failed"), QA_ERR_HIP));
    const size_t lds = ampBytes << tileQubits;
    if (lds > 65536) {
        if (hipFuncSetAttribute((const void*)k_batchPass<float>, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds) !=
                hipSuccess ||
            hipFuncSetAttribute((const void*)k_batchPass<double>, hipFuncAttributeMaxDynamicSharedMemorySize,
                                (int)lds) != hipSuccess)
            return fail((setError("qa_serveCreate: %zu bytes of LDS per workgroup are not available", lds), QA_ERR_ARG));
    }
    E->buddy.init(arena, __builtin_ctzll(minBlock));
    E->th = std::thread(engineLoop, E);
    *engine = E;
    return QA_OK;
}

int qa_serveDestroy(void* engine) {
    QA_ARG_CHECK(engine != nullptr, "qa_serveDestroy: null engine");
    Engine* E = (Engine*)engine;
    {
        std::lock_guard<std::mutex> lk(E->mu);
        E->stop = true;
    }
    E->cvWork.notify_all();
    if (E->th.joinable()) E->th.join();
    (void)hipStreamSynchronize(E->stream);
    for (auto& kv : E->byId) delete kv.second;
    for (auto& s : E->steps) {
        if (s.host) (void)hipHostFree(s.host);
        if (s.dev) (void)hipFree(s.dev);
        if (s.partials) (void)hipFree(s.partials);
        if (s.outDev) (void)hipFree(s.outDev);
        if (s.outHost) (void)hipHostFree(s.outHost);
        if (s.done) (void)hipEventDestroy(s.done);
        for (auto& kv : s.execs)
            if (kv.second) (void)hipGraphExecDestroy(kv.second);
    }
    (void)hipFree(E->arena);
    (void)hipStreamDestroy(E->stream);
    delete E;
    return QA_OK;
}

int qa_serveSubmit(void* engine, const qa_serveCircuit* c, int64_t* id) {
    QA_ARG_CHECK(engine && c && id, "qa_serveSubmit: null argument");
    Engine* E = (Engine*)engine;
    QA_ARG_CHECK(c->numQubits >= RB && c->numQubits <= 30, "qa_serveSubmit: numQubits must be in [4, 30]");
    QA_ARG_CHECK(c->numGates >= 0 && (c->numGates == 0 || c->gates), "qa_serveSubmit: bad gate list");
    QA_ARG_CHECK(c->numPaulis >= 0 && (c->numPaulis == 0 || (c->pauliX && c->pauliY && c->pauliZ)),
                 "qa_serveSubmit: bad Pauli list");
    QA_ARG_CHECK((E->ampBytes << c->numQubits) <= E->arenaBytes &&
                     E->buddy.orderFor(E->ampBytes << c->numQubits) <= E->buddy.maxOrder(),
                 "qa_serveSubmit: a %d-qubit state does not fit the engine's arena", c->numQubits);
    Request* r = new Request();
    r->n = c->numQubits;
    std::string err;
    if (!translate(*c, r->ops, err)) {
        delete r;
        setError("%s", err.c_str());
        return QA_ERR_ARG;
    }
    const uint64_t all = (c->numQubits >= 64) ? ~0ull : ((1ull << c->numQubits) - 1ull);
    for (int i = 0; i < c->numPaulis; i++) {
        const uint64_t x = c->pauliX[i], y = c->pauliY[i], z = c->pauliZ[i];
        if (((x | y | z) & ~all) || (x & y) || (x & z) || (y & z)) {
            delete r;
            setError("qa_serveSubmit: Pauli string %d has overlapping or out-of-range masks", i);
            return QA_ERR_ARG;
        }
        r->paulis.push_back({x, y, z});
    }
    {
        std::lock_guard<std::mutex> lk(E->mu);
        if (!E->error.empty()) {
            delete r;
            setError("%s", E->error.c_str());
            return QA_ERR_HIP;
        }
        r->id = E->nextId++;
        r->submitted = Clock::now();
        E->byId[r->id] = r;
        E->submitted.push_back(r);
        *id = r->id;
    }
    E->cvWork.notify_one();
    return QA_OK;
}

int qa_servePoll(void* engine, int64_t id, int* done) {
    QA_ARG_CHECK(engine && done, "qa_servePoll: null argument");
    Engine* E = (Engine*)engine;
    std::lock_guard<std::mutex> lk(E->mu);
    auto it = E->byId.find(id);
    QA_ARG_CHECK(it != E->byId.end(), "qa_servePoll: unknown request %lld", (long long)id);
    if (it->second->phase == FAILED) {
        setError("%s", E->error.c_str());
        return QA_ERR_HIP;
    }
    *done = it->second->phase == DONE;
    return QA_OK;
}

int qa_serveWait(void* engine, int64_t id, double* expectations, double* latencyMs) {
    QA_ARG_CHECK(engine, "qa_serveWait: null engine");
    Engine* E = (Engine*)engine;
    std::unique_lock<std::mutex> lk(E->mu);
    auto it = E->byId.find(id);
    QA_ARG_CHECK(it != E->byId.end(), "qa_serveWait: unknown request %lld", (long long)id);
    Request* r = it->second;
    E->cvDone.wait(lk, [&] { return r->phase == DONE || r->phase == FAILED; });
    if (r->phase == FAILED) {
        setError("%s", E->error.c_str());
        return QA_ERR_HIP;
    }
    if (expectations) std::copy(r->result.begin(), r->result.end(), expectations);
    if (latencyMs) *latencyMs = std::chrono::duration<double, std::milli>(r->finished - r->submitted).count();
    E->byId.erase(it);
    delete r;
    return QA_OK;
}

int qa_serveSetGraphs(void* engine, int enabled) {
    QA_ARG_CHECK(engine, "qa_serveSetGraphs: null engine");
    ((Engine*)engine)->graphs = enabled != 0;
    return QA_OK;
}

int qa_serveGraphStats(void* engine, int64_t* launches, int64_t* instantiations, int64_t* updates) {
    QA_ARG_CHECK(engine, "qa_serveGraphStats: null engine");
    Engine* E = (Engine*)engine;
    if (launches) *launches = E->nGraphLaunches;
    if (instantiations) *instantiations = E->nGraphInstances;
    if (updates) *updates = E->nGraphUpdates;
    return QA_OK;
}

int qa_serveTimes(void* engine, double* planMs, double* waitMs) {
    QA_ARG_CHECK(engine, "qa_serveTimes: null engine");
    Engine* E = (Engine*)engine;
    if (planMs) *planMs = E->planNs.load() / 1e6;
    if (waitMs) *waitMs = E->waitNs.load() / 1e6;
    return QA_OK;
}

int qa_serveStats(void* engine, int64_t* steps, int64_t* passes, int64_t* admitted, int64_t* completed,
                  int64_t* maxRunning) {
    QA_ARG_CHECK(engine, "qa_serveStats: null engine");
    Engine* E = (Engine*)engine;
    if (steps) *steps = E->nSteps;
    if (passes) *passes = E->nPasses;
    if (admitted) *admitted = E->nAdmitted;
    if (completed) *completed = E->nCompleted;
    if (maxRunning) *maxRunning = E->maxRunning;
    return QA_OK;
}

}  // extern "C"
