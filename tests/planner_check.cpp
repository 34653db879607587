// CPU check of the fusion planner's host-side transformations (test
// infrastructure; built and run by tests/test_planner.py, never shipped).
//
// Random queues of the fusable gate kinds (with controls and masks on local
// and non-local qubits) on a small state are planned the way a flush plans
// them -- pickPass, then mergePass and clusterPass (default mode) or
// takeRun (exact mode) -- and the planned op sequence is applied to a copy
// of a random state.  Both the queue and the planned sequence are applied
// with this file's own full-state restatement of each op's semantics (the
// single-gate kernels' formulas), independent of the planner's helpers; the
// two results must agree to rounding.  Prints one line per case and exits
// non-zero on the first mismatch.
#include <cmath>
#include <complex>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <random>
#include <vector>

#include "qa_tile.hpp"
#include "qa_jit.hpp"

#include <hip/hiprtc.h>
#include <map>
#include <string>

// the device headers the library embeds for its generated pass kernels
extern "C" const int qa_jit_num_headers;
extern "C" const char* const qa_jit_header_names[];
extern "C" const char* const qa_jit_header_sources[];

using namespace qa::fuse;
typedef std::complex<double> cx;

static int parity(uint64_t x) { return __builtin_popcountll(x) & 1; }

// one op on the full state, as the single-gate kernels define it
static void apply(const GOp& g, std::vector<cx>& a) {
    const uint64_t N = a.size();
    auto ok = [&](uint64_t i) { return (i & g.cMask) == g.cVal; };
    auto f = [&](int k) { return cx(g.f[2 * k], g.f[2 * k + 1]); };
    switch (g.kind) {
        case F_DENSE1: {
            const uint64_t t = 1ull << g.t1;
            for (uint64_t i = 0; i < N; i++) {
                if ((i & t) || !ok(i)) continue;
                const cx x0 = a[i], x1 = a[i | t];
                a[i] = f(0) * x0 + f(1) * x1;
                a[i | t] = f(2) * x0 + f(3) * x1;
            }
            break;
        }
        case F_DENSE1V: {
            const uint64_t t = 1ull << g.t1;
            for (uint64_t i = 0; i < N; i++) {
                if (i & t) continue;
                int w = 0, b = 0;
                for (int q = 0; q < 64; q++)
                    if ((g.m1 >> q) & 1ull) w |= (int)((i >> q) & 1ull) << b++;
                const double* m = g.f + 8 * w;
                const cx x0 = a[i], x1 = a[i | t];
                a[i] = cx(m[0], m[1]) * x0 + cx(m[2], m[3]) * x1;
                a[i | t] = cx(m[4], m[5]) * x0 + cx(m[6], m[7]) * x1;
            }
            break;
        }
        case F_DENSEN: {
            const int D = 1 << g.numTargs;
            uint64_t tm = 0;
            for (int b = 0; b < g.numTargs; b++) tm |= 1ull << g.targs[b];
            std::vector<cx> x(D), y(D);
            for (uint64_t i = 0; i < N; i++) {
                if (i & tm) continue;
                auto at = [&](int v) {
                    uint64_t j = i;
                    for (int b = 0; b < g.numTargs; b++)
                        if ((v >> b) & 1) j |= 1ull << g.targs[b];
                    return j;
                };
                for (int c = 0; c < D; c++) x[c] = a[at(c)];
                for (int rr = 0; rr < D; rr++) {
                    y[rr] = 0;
                    for (int c = 0; c < D; c++)
                        y[rr] += cx(g.mat[2 * (rr * D + c)], (g.conj ? -1 : 1) * g.mat[2 * (rr * D + c) + 1]) * x[c];
                }
                for (int rr = 0; rr < D; rr++) a[at(rr)] = y[rr];
            }
            break;
        }
        case F_DENSE2: {
            const uint64_t t1 = 1ull << g.t1, t2 = 1ull << g.t2;
            const uint64_t off[4] = {0, t1, t2, t1 | t2};
            for (uint64_t i = 0; i < N; i++) {
                if ((i & (t1 | t2)) || !ok(i)) continue;
                cx x[4], y[4];
                for (int c = 0; c < 4; c++) x[c] = a[i | off[c]];
                for (int r = 0; r < 4; r++) {
                    y[r] = 0;
                    for (int c = 0; c < 4; c++) y[r] += f(4 * r + c) * x[c];
                }
                for (int r = 0; r < 4; r++) a[i | off[r]] = y[r];
            }
            break;
        }
        case F_DIAG1:
            for (uint64_t i = 0; i < N; i++)
                if (ok(i)) a[i] *= ((i >> g.t1) & 1ull) ? f(1) : f(0);
            break;
        case F_PARITY:
            for (uint64_t i = 0; i < N; i++)
                if (ok(i)) a[i] *= parity(i & g.m1) ? f(1) : f(0);
            break;
        case F_PAULI: {
            const uint64_t top = 1ull << (63 - __builtin_clzll(g.m1));
            for (uint64_t i = 0; i < N; i++) {
                if ((i & top) || !ok(i)) continue;
                const uint64_t j = i ^ g.m1;
                const double sA = parity(i & g.m2) ? -1 : 1, sB = parity(j & g.m2) ? -1 : 1;
                const cx xa = a[i], xb = a[j];
                a[i] = f(0) * xa + f(1) * sB * xb;
                a[j] = f(0) * xb + f(1) * sA * xa;
            }
            break;
        }
        case F_SWAP: {
            const uint64_t t1 = 1ull << g.t1, t2 = 1ull << g.t2;
            for (uint64_t i = 0; i < N; i++)
                if ((i & t1) && !(i & t2) && ok(i)) std::swap(a[i], a[i ^ (t1 | t2)]);
            break;
        }
        case F_DEPOL1: {
            const uint64_t k1 = 1ull << g.t1, k2 = 1ull << g.t2;
            const double c1 = g.f[0], c2 = g.f[1], c3 = g.f[2];
            for (uint64_t i = 0; i < N; i++) {
                if (i & (k1 | k2)) continue;
                const cx a00 = a[i], a11 = a[i | k1 | k2];
                a[i] = c1 * a00 + c2 * a11;
                a[i | k1] *= c3;
                a[i | k2] *= c3;
                a[i | k1 | k2] = c1 * a11 + c2 * a00;
            }
            break;
        }
        case F_DEPH1: {
            const uint64_t k1 = 1ull << g.t1, k2 = 1ull << g.t2;
            for (uint64_t i = 0; i < N; i++)
                if (((i & k1) != 0) != ((i & k2) != 0)) a[i] *= g.f[0];
            break;
        }
        default:
            std::fprintf(stderr, "unknown op kind %d\n", g.kind);
            std::exit(3);
    }
}

struct Rng {
    std::mt19937_64 g;
    explicit Rng(uint64_t s) : g(s) {}
    int below(int n) { return (int)(g() % (uint64_t)n); }
    double unit() { return std::uniform_real_distribution<double>(-1, 1)(g); }
};

static void randomUnitary2(Rng& r, double* f) {
    // e^{i a} [[cos t e^{i b}, -sin t e^{-i c}], [sin t e^{i c}, cos t e^{-i b}]]
    const double t = r.unit() * 3, b = r.unit() * 3, c = r.unit() * 3, ph = r.unit() * 3;
    const cx e(std::cos(ph), std::sin(ph));
    const cx m[4] = {e * std::cos(t) * cx(std::cos(b), std::sin(b)), -e * std::sin(t) * cx(std::cos(c), -std::sin(c)),
                     e * std::sin(t) * cx(std::cos(c), std::sin(c)), e * std::cos(t) * cx(std::cos(b), -std::sin(b))};
    for (int i = 0; i < 4; i++) {
        f[2 * i] = m[i].real();
        f[2 * i + 1] = m[i].imag();
    }
}

// a random fusable gate; `dens` adds the one-qubit channels on (q, q + n/2)
static GOp randomGate(Rng& r, int n, bool dens, int hot) {
    GOp g;
    auto qubit = [&]() { return r.below(4) == 0 ? r.below(n) : n - 1 - r.below(hot); };
    auto other = [&](int a) {
        int b;
        do b = qubit(); while (b == a);
        return b;
    };
    auto ctrl = [&](uint64_t busy) {
        if (r.below(3)) return;
        int c;
        do c = r.below(n); while ((busy >> c) & 1ull);
        g.cMask |= 1ull << c;
        if (r.below(4)) g.cVal |= 1ull << c;
    };
    if (n >= 5 && r.below(12) == 0) {
        // an uncontrolled 3- or 4-qubit dense matrix (its matrix kept in a
        // pool that outlives the queue)
        static std::vector<std::vector<double>> pool;
        g.kind = F_DENSEN;
        g.numTargs = 3 + r.below(2);
        for (int b = 0; b < g.numTargs; b++) {
            int q;
            bool dup;
            do {
                q = qubit();
                dup = false;
                for (int c = 0; c < b; c++) dup = dup || g.targs[c] == q;
            } while (dup);
            g.targs[b] = q;
        }
        const int D = 1 << g.numTargs;
        pool.emplace_back(2 * D * D);
        for (double& x : pool.back()) x = r.unit() / D;
        g.mat = pool.back().data();
        g.conj = r.below(2);
        return g;
    }
    const int kind = r.below(dens ? 10 : 8);
    switch (kind) {
        case 0: case 1:
            g.kind = F_DENSE1;
            g.t1 = qubit();
            randomUnitary2(r, g.f);
            ctrl(1ull << g.t1);
            break;
        case 2: {
            g.kind = F_DENSE2;
            g.t1 = qubit();
            g.t2 = other(g.t1);
            for (int i = 0; i < 32; i++) g.f[i] = r.unit();
            ctrl((1ull << g.t1) | (1ull << g.t2));
            break;
        }
        case 3: {
            g.kind = F_DIAG1;
            g.t1 = qubit();
            const double a0 = r.unit() * 3, a1 = r.unit() * 3;
            const bool phaseOnly = r.below(2);
            g.f[0] = phaseOnly ? 1 : std::cos(a0);
            g.f[1] = phaseOnly ? 0 : std::sin(a0);
            g.f[2] = std::cos(a1);
            g.f[3] = std::sin(a1);
            ctrl(1ull << g.t1);
            if (phaseOnly) {
                g.cMask |= 1ull << g.t1;
                g.cVal |= 1ull << g.t1;
            }
            break;
        }
        case 4: {
            g.kind = F_PARITY;
            const int a = qubit();
            g.m1 = 1ull << a;
            if (r.below(2)) g.m1 |= 1ull << other(a);
            const double th = r.unit() * 3;
            g.f[0] = std::cos(th);
            g.f[1] = -std::sin(th);
            g.f[2] = std::cos(th);
            g.f[3] = std::sin(th);
            ctrl(g.m1);
            break;
        }
        case 5: {
            g.kind = F_PAULI;
            const int a = qubit(), b = other(a);
            g.m1 = 1ull << a;
            if (r.below(2)) g.m1 |= 1ull << b;
            g.m2 = (r.below(2) ? (1ull << a) : 0) | (r.below(2) ? (1ull << r.below(n)) : 0);
            const double th = r.unit() * 3;
            g.f[0] = std::cos(th);
            g.f[1] = 0;
            g.f[2] = 0;
            g.f[3] = std::sin(th);
            ctrl(g.m1 | g.m2);
            break;
        }
        case 6:
            g.kind = F_SWAP;
            g.t1 = qubit();
            g.t2 = other(g.t1);
            ctrl((1ull << g.t1) | (1ull << g.t2));
            break;
        case 7: {
            // a controlled phase, the workload's most common diagonal
            g.kind = F_DIAG1;
            g.t1 = qubit();
            const int c = other(g.t1);
            g.cMask = g.cVal = (1ull << g.t1) | (1ull << c);
            const double th = r.unit() * 3;
            g.f[0] = 1;
            g.f[2] = std::cos(th);
            g.f[3] = std::sin(th);
            break;
        }
        default: {
            g.kind = r.below(2) ? F_DEPOL1 : F_DEPH1;
            const int q = r.below(n / 2);
            g.t1 = q;
            g.t2 = q + n / 2;
            const double p = 0.05 * (r.unit() + 1);
            g.f[0] = g.kind == F_DEPOL1 ? 1 - 2 * p / 3 : 1 - 2 * p;
            g.f[1] = 2 * p / 3;
            g.f[2] = 1 - 4 * p / 3;
            break;
        }
    }
    return g;
}

// With "jit": every planned pass's program (complex128, and complex64 with
// MFMA blocks) also gets its generated kernel source; programs of equal
// shape key must have equal source (what makes the per-shape kernel cache
// sound), and up to maxCompiles distinct sources are compiled for gfx950
// with hiprtc, as the library does at run time (loading needs a GPU and is
// left to the GPU tests).
struct JitCheck {
    bool on = false;
    int maxCompiles = 0, compiled = 0;
    std::map<std::string, std::string> byKey;
    int check(const GOp* ops, size_t num, int n, uint64_t L, int low, bool exact, int c) {
        for (int prec : {QA_FP64, QA_FP32}) {
            TileArgs a;
            std::vector<TOp> prog;
            std::vector<float> blocks;
            if (prec == QA_FP32) planPassBlocks(ops, num, n, L, low, 2, a, prog, blocks);
            else planPass(ops, num, n, L, low, a, prog, !exact);
            setLdsOffsets(prog.data(), prog.size(), a, prec);
            const bool nt = low >= 3;
            const std::string key = qa::jit::shapeKey(prec, nt, 2, a, prog.data(), prog.size());
            const std::string src = qa::jit::generate(prec, nt, 2, a, prog.data(), prog.size());
            auto it = byKey.find(key);
            if (it != byKey.end()) {
                if (it->second != src) {
                    std::printf("case %d: two programs of one shape key generated different sources\n", c);
                    return 1;
                }
                continue;
            }
            byKey[key] = src;
            if (const char* dir = std::getenv("QA_JIT_DUMP_DIR")) {
                // the source, for compiling outside (test infrastructure)
                char path[512];
                std::snprintf(path, sizeof path, "%s/pass_%04zu.hip", dir, byKey.size());
                if (FILE* fp = std::fopen(path, "w")) {
                    std::fputs(src.c_str(), fp);
                    std::fclose(fp);
                }
            }
            bool dense = false;
            for (const TOp& t : prog) dense = dense || t.h.kind == F_DENSE_LDS;
            // every pass with a 3-4 qubit matrix compiles (only generated
            // kernels run them), a few others
            if (compiled >= maxCompiles && !(dense && denseCompiled < 3)) continue;
            hiprtcProgram p;
            hiprtcCreateProgram(&p, src.c_str(), "qa_jit_pass.hip", qa_jit_num_headers, qa_jit_header_sources,
                                qa_jit_header_names);
            const char* opts[] = {"--offload-arch=gfx950", "-O3", "-std=c++17", "-Xclang", "-target-feature",
                                  "-Xclang", "-packed-fp32-ops", "-mllvm", "-simplifycfg-sink-common=false"};
            const hiprtcResult rc = hiprtcCompileProgram(p, 9, opts);
            if (rc != HIPRTC_SUCCESS) {
                size_t ln = 0;
                hiprtcGetProgramLogSize(p, &ln);
                std::string log(ln + 1, '\0');
                hiprtcGetProgramLog(p, &log[0]);
                std::printf("case %d: generated kernel does not compile (%s):\n%s\n", c, hiprtcGetErrorString(rc),
                            log.substr(0, 2000).c_str());
                return 1;
            }
            size_t cn = 0;
            hiprtcGetCodeSize(p, &cn);
            std::vector<char> code(cn);
            hiprtcGetCode(p, code.data());
            hiprtcDestroyProgram(&p);
            if (dense) {
                const long scratch = privateSegmentSize(code);
                if (scratch != 0) {
                    std::printf("case %d: generated kernel with a 3-4 qubit matrix uses %ld bytes of scratch\n", c,
                                scratch);
                    return 1;
                }
                denseCompiled++;
            } else {
                compiled++;
            }
        }
        return 0;
    }
    int denseCompiled = 0;

    // private segment size of the kernel descriptor "qa_jit_pass.kd" in an
    // amdgcn code object (ELF64), -1 if not found
    static long privateSegmentSize(const std::vector<char>& elf) {
        if (elf.size() < 64 || std::memcmp(elf.data(), "\x7f" "ELF", 4) != 0) return -1;
        auto u16 = [&](size_t o) { uint16_t v; std::memcpy(&v, elf.data() + o, 2); return v; };
        auto u32 = [&](size_t o) { uint32_t v; std::memcpy(&v, elf.data() + o, 4); return v; };
        auto u64 = [&](size_t o) { uint64_t v; std::memcpy(&v, elf.data() + o, 8); return v; };
        const uint64_t shoff = u64(0x28);
        const int shentsize = u16(0x3a), shnum = u16(0x3c);
        for (int i = 0; i < shnum; i++) {
            const size_t sh = shoff + (size_t)i * shentsize;
            if (u32(sh + 4) != 2) continue;  // SHT_SYMTAB
            const uint64_t off = u64(sh + 0x18), size = u64(sh + 0x20), entsize = u64(sh + 0x38);
            const uint32_t strIdx = u32(sh + 0x28);
            const uint64_t strOff = u64(shoff + (size_t)strIdx * shentsize + 0x18);
            for (uint64_t q = 0; q + entsize <= size; q += entsize) {
                const char* name = elf.data() + strOff + u32(off + q);
                if (std::strcmp(name, "qa_jit_pass.kd") != 0) continue;
                const uint16_t secIdx = u16(off + q + 6);
                const uint64_t value = u64(off + q + 8);
                const size_t sec = shoff + (size_t)secIdx * shentsize;
                const uint64_t secAddr = u64(sec + 0x10), secOff = u64(sec + 0x18);
                return (long)u32(secOff + (value - secAddr) + 4);
            }
        }
        return -1;
    }
};

// ---- span passes ----------------------------------------------------------

static int rankGf2(std::vector<uint64_t> v) {
    int rank = 0;
    for (int bit = 63; bit >= 0; bit--) {
        size_t p = v.size();
        for (size_t i = 0; i < v.size(); i++)
            if ((v[i] >> bit) & 1ull) {
                p = i;
                break;
            }
        if (p == v.size()) continue;
        const uint64_t pv = v[p];
        v.erase(v.begin() + (ptrdiff_t)p);
        for (auto& x : v)
            if ((x >> bit) & 1ull) x ^= pv;
        rank++;
    }
    return rank;
}

// a gadget over random qubits: PAULI (X/Y somewhere) or PARITY
static GOp randomGadget(Rng& r, int n, int maxQubits) {
    GOp g;
    const int k = 1 + r.below(maxQubits);
    uint64_t x = 0, y = 0, z = 0;
    for (int j = 0; j < k; j++) {
        const uint64_t b = 1ull << r.below(n);
        if ((x | y | z) & b) continue;
        const int c = r.below(3);
        (c == 0 ? x : c == 1 ? y : z) |= b;
    }
    const double th = r.unit() * 3;
    if (x | y) {
        g.kind = F_PAULI;
        g.m1 = x | y;
        g.m2 = y | z;
        g.f[0] = std::cos(th);
        g.f[3] = std::sin(th);
    } else {
        g.kind = F_PARITY;
        g.m1 = z ? z : 1ull;
        g.f[0] = std::cos(th);
        g.f[1] = -std::sin(th);
        g.f[2] = std::cos(th);
        g.f[3] = std::sin(th);
    }
    return g;
}

static uint64_t depositBits(uint64_t v, const int* pos, int num) {
    uint64_t out = 0;
    for (int i = 0; i < num; i++) out |= ((v >> i) & 1ull) << pos[i];
    return out;
}

// one span pass on the full state, following the kernel's indexing: a grid
// of `grid` workgroups, each walking its cosets by masked adds; thread t,
// register i holds coordinate t + i * 2^(k-4) at rep ^ own(t) ^ top[i].
// Every amplitude must be visited exactly once.
static bool emulateSpan(const SpanArgs& a, const std::vector<SpanOp>& prog, uint64_t grid, std::vector<cx>& amps,
                        int c) {
    const uint32_t nth = 1u << (a.k - RB), size = 1u << a.k, half = size / 2;
    std::vector<uint8_t> seen(amps.size(), 0);
    std::vector<cx> tile(size);
    const uint64_t step = depositBits(grid, a.repPos, a.numRep);
    for (uint64_t b = 0; b < grid && b < a.numCosets; b++) {
        uint64_t rep = depositBits(b, a.repPos, a.numRep);
        for (uint64_t w = b; w < a.numCosets; w += grid) {
            std::vector<uint64_t> addr(size);
            for (uint32_t t = 0; t < nth; t++) {
                uint64_t own = 0;
                for (int j = 0; j < a.k - RB; j++)
                    if ((t >> j) & 1u) own ^= a.basis[j];
                for (int i = 0; i < RA; i++) {
                    const uint32_t s = t + (uint32_t)i * nth;
                    addr[s] = rep ^ own ^ a.top[i];
                    if (addr[s] >= amps.size() || seen[addr[s]]++) {
                        std::printf("case %d: coset %llu visits amplitude %llu twice or out of range\n", c,
                                    (unsigned long long)w, (unsigned long long)addr[s]);
                        return false;
                    }
                    tile[s] = amps[addr[s]];
                }
            }
            for (size_t o = 0; o < prog.size();) {
                // a batch: header, table, gadgets (the kernel's register batches)
                const SpanOp& hb = prog[o];
                if (hb.kind != SPAN_BATCH || o + 2 + hb.c > prog.size()) {
                    std::printf("case %d: malformed span program at record %zu\n", c, o);
                    return false;
                }
                const uint32_t* S4 = reinterpret_cast<const uint32_t*>(&prog[o + 1]);
                std::vector<uint8_t> cover(size, 0);
                for (uint32_t t = 0; t < nth; t++) {
                    uint32_t s0 = 0;
                    for (uint32_t j = 0; j < hb.r; j++) s0 |= ((t >> j) & 1u) << ((hb.dep >> (4 * j)) & 15u);
                    cx x[RA];
                    for (int v = 0; v < RA; v++) {
                        const uint32_t sc = s0 ^ S4[v];
                        if (sc >= size || cover[sc]++) {
                            std::printf("case %d: batch coordinate %u out of range or taken twice\n", c, sc);
                            return false;
                        }
                        x[v] = tile[sc];
                    }
                    for (uint32_t g = 0; g < hb.c; g++) {
                        const SpanOp& op = prog[o + 2 + g];
                        const int par = parity(rep & op.z) ^ parity(s0 & op.d);
                        const cx f0(op.f[0], op.f[1]), f1(op.f[2], op.f[3]);
                        if (op.kind == F_PARITY) {
                            for (int v = 0; v < RA; v++) x[v] *= (parity((uint64_t)v & op.r) ^ par) ? f1 : f0;
                            continue;
                        }
                        const int top = 31 - __builtin_clz(op.c);
                        for (int v = 0; v < RA; v++) {
                            if ((v >> top) & 1) continue;
                            const int w = v ^ (int)op.c;
                            const double sA = (parity((uint64_t)v & op.r) ^ par) ? -1 : 1;
                            const double sB = (parity((uint64_t)w & op.r) ^ par) ? -1 : 1;
                            const cx xa = x[v], xb = x[w];
                            x[v] = f0 * xa + f1 * sB * xb;
                            x[w] = f0 * xb + f1 * sA * xa;
                        }
                    }
                    for (int v = 0; v < RA; v++) tile[s0 ^ S4[v]] = x[v];
                }
                o += 2 + hb.c;
            }
            for (uint32_t s = 0; s < size; s++) amps[addr[s]] = tile[s];
            rep = ((rep | ~a.repMask) + step) & a.repMask;
        }
    }
    for (size_t i = 0; i < seen.size(); i++)
        if (seen[i] != 1) {
            std::printf("case %d: amplitude %zu visited %d times\n", c, i, (int)seen[i]);
            return false;
        }
    return true;
}

// random runs of gadgets (with an occasional op a span pass does not take)
// planned by spanRun as a flush does, each span pass emulated on the full
// state; must agree with the gates one by one, and each run must be the
// longest whose X/Y masks with the low qubits span at most k dimensions
static int spanMain(int cases, uint64_t seed0) {
    int checked = 0, spanPasses = 0;
    for (int c = 0; c < cases; c++) {
        Rng r(seed0 * 7919ull + (uint64_t)c);
        const int n = 11 + r.below(6);
        const int k = 10 + r.below(std::min(4, n - 10));
        const int low = r.below(4);
        const int num = 10 + r.below(60);
        std::vector<GOp> queue;
        for (int i = 0; i < num; i++) {
            if (r.below(12) == 0) {
                GOp g = randomGadget(r, n, 6);
                g.cMask = g.cVal = 0;
                if (r.below(2)) {
                    g.kind = F_DENSE1;
                    g.t1 = r.below(n);
                    g.m1 = g.m2 = 0;
                    randomUnitary2(r, g.f);
                } else {
                    // a controlled gadget: not taken
                    uint64_t free = ((1ull << n) - 1ull) & ~(g.m1 | g.m2);
                    if (!free) continue;
                    int q = 0;
                    while (!((free >> q) & 1ull)) q++;
                    g.cMask = g.cVal = 1ull << q;
                }
                queue.push_back(g);
                continue;
            }
            queue.push_back(randomGadget(r, n, 1 + r.below(std::min(n, 12))));
        }
        std::vector<cx> want(1ull << n);
        for (auto& x : want) x = cx(r.unit(), r.unit());
        std::vector<cx> got = want;
        for (const GOp& g : queue) ::apply(g, want);
        std::vector<GOp> ops = queue;
        SpanArgs a;
        std::vector<SpanOp> prog;
        while (!ops.empty()) {
            const size_t run = spanRun(ops.data(), ops.size(), n, k, low, a, prog);
            std::vector<uint64_t> masks;
            for (int q = 0; q < low; q++) masks.push_back(1ull << q);
            for (size_t i = 0; i < run; i++)
                if (ops[i].kind == F_PAULI) masks.push_back(ops[i].m1);
            if (run && rankGf2(masks) > k) {
                std::printf("case %d: a run of %zu spans %d > %d dimensions\n", c, run, rankGf2(masks), k);
                return 1;
            }
            if (run < ops.size() && run < (size_t)MAX_SPAN_OPS && spanEligible(ops[run])) {
                masks.push_back(ops[run].m1);
                if (ops[run].kind != F_PAULI || rankGf2(masks) <= k) {
                    std::printf("case %d: the run of %zu stops before an op that fits\n", c, run);
                    return 1;
                }
            }
            if (run == 0) {
                if (spanEligible(ops[0])) {
                    std::printf("case %d: an eligible op starts no run\n", c);
                    return 1;
                }
                ::apply(ops[0], got);
                ops.erase(ops.begin());
                continue;
            }
            size_t gadgets = 0;
            for (const SpanOp& o : prog) gadgets += o.kind == F_PAULI || o.kind == F_PARITY;
            if (gadgets != run || a.numOps != (int)prog.size() || a.k != k) {
                std::printf("case %d: plan of %zu ops has %zu gadgets in %zu records\n", c, run, gadgets,
                            prog.size());
                return 1;
            }
            const uint64_t grid = 1 + (uint64_t)r.below((int)std::min<uint64_t>(a.numCosets, 40));
            if (!emulateSpan(a, prog, grid, got, c)) return 1;
            spanPasses++;
            ops.erase(ops.begin(), ops.begin() + (ptrdiff_t)run);
        }
        double err = 0, norm = 0;
        for (size_t i = 0; i < got.size(); i++) {
            err = std::max(err, std::abs(got[i] - want[i]));
            norm = std::max(norm, std::abs(want[i]));
        }
        if (!(err <= 1e-12 * std::max(1.0, norm))) {
            std::printf("case %d (n=%d k=%d low=%d gates=%zu): max error %.3e\n", c, n, k, low, queue.size(), err);
            return 1;
        }
        checked++;
    }
    std::printf("span check: %d cases agree (%d span passes)\n", checked, spanPasses);
    return 0;
}

int main(int argc, char** argv) {
    const int cases = argc > 1 ? std::atoi(argv[1]) : 200;
    const uint64_t seed0 = argc > 2 ? std::strtoull(argv[2], nullptr, 10) : 1;
    if (argc > 3 && std::string(argv[3]) == "span") return spanMain(cases, seed0);
    JitCheck jit;
    if (argc > 3 && std::string(argv[3]) == "jit") {
        jit.on = true;
        jit.maxCompiles = argc > 4 ? std::atoi(argv[4]) : 4;
    }
    int checked = 0;
    for (int c = 0; c < cases; c++) {
        Rng r(seed0 * 1000003ull + (uint64_t)c);
        const bool dens = c % 4 == 3;
        // QA_PLANNER_WIDE: wider states and tiles (up to 14 qubits, 13-qubit tiles)
        const bool wide = std::getenv("QA_PLANNER_WIDE") != nullptr;
        const int n = wide ? (dens ? 2 * (5 + r.below(3)) : 11 + r.below(4)) : (dens ? 2 * (3 + r.below(3)) : 6 + r.below(5));
        const int kMax = wide ? std::min(n, 9 + r.below(5)) : std::min(n, 4 + r.below(5));
        const int low = r.below(std::max(1, kMax - 1));
        const bool exact = c % 5 == 4;
        const int num = 20 + r.below(120);
        std::vector<GOp> queue;
        const uint64_t lowMask = (1ull << low) - 1ull;
        for (int i = 0; i < num; i++) {
            GOp g = randomGate(r, n, dens, std::max(2, kMax / 2));
            // the queue takes only gates whose dense targets fit a tile
            if (__builtin_popcountll(denseMask(g) | lowMask) > kMax) continue;
            queue.push_back(g);
        }
        std::vector<cx> a(1ull << n);
        for (auto& x : a) x = cx(r.unit(), r.unit());
        std::vector<cx> want = a, got = a;
        for (const GOp& g : queue) ::apply(g, want);

        std::vector<GOp> ops = queue, pass, rest, merged, planned;
        std::vector<uint32_t> picked;
        int passes = 0;
        while (!ops.empty()) {
            uint64_t L = 0;
            pass.clear();
            rest.clear();
            if (exact) {
                const size_t run = takeRun(ops.data(), ops.size(), n, kMax, low, L);
                if (run == 0) {
                    std::printf("case %d: empty run\n", c);
                    return 1;
                }
                pass.assign(ops.begin(), ops.begin() + run);
                rest.assign(ops.begin() + run, ops.end());
                planned = pass;
            } else {
                picked.clear();
                pickPass(ops.data(), ops.size(), n, kMax, low, L, picked);
                if (picked.empty()) {
                    std::printf("case %d: empty pass\n", c);
                    return 1;
                }
                size_t p = 0;
                for (size_t i = 0; i < ops.size(); i++) {
                    if (p < picked.size() && picked[p] == i) {
                        pass.push_back(ops[i]);
                        p++;
                    } else {
                        rest.push_back(ops[i]);
                    }
                }
                merged.clear();
                mergePass(pass.data(), pass.size(), L, merged);
                planned.clear();
                clusterPass(merged.data(), merged.size(), lowMask, planned);
                if (planned.size() > pass.size()) {
                    std::printf("case %d: planning added ops (%zu > %zu)\n", c, planned.size(), pass.size());
                    return 1;
                }
            }
            if (jit.on && jit.check(planned.data(), planned.size(), n, L, low, exact, c)) return 1;
            // every planned op's dense targets are local to the pass
            for (const GOp& g : planned) {
                if (denseMask(g) & ~L) {
                    std::printf("case %d: op kind %d has dense targets outside the local set\n", c, g.kind);
                    return 1;
                }
                apply(g, got);
            }
            if (__builtin_popcountll(L) != kMax || (L & lowMask) != lowMask) {
                std::printf("case %d: local set %llx is not %d qubits with the low ones\n", c, (unsigned long long)L,
                            kMax);
                return 1;
            }
            ops.swap(rest);
            passes++;
        }
        double err = 0, norm = 0;
        for (size_t i = 0; i < a.size(); i++) {
            err = std::max(err, std::abs(got[i] - want[i]));
            norm = std::max(norm, std::abs(want[i]));
        }
        if (!(err <= 1e-12 * std::max(1.0, norm))) {
            std::printf("case %d (n=%d k=%d low=%d exact=%d gates=%zu): max error %.3e\n", c, n, kMax, low, exact,
                        queue.size(), err);
            return 1;
        }
        checked++;
    }
    if (jit.on)
        std::printf("jit check: %zu shapes, %d generated kernels compiled (%d with 3-4 qubit matrices)\n",
                    jit.byKey.size(), jit.compiled, jit.denseCompiled);
    std::printf("planner check: %d cases agree\n", checked);
    return 0;
}
