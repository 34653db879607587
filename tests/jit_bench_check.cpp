// CPU check of the pass kernels generated for the benchmark's circuit family
// (test infrastructure; built and run by tests/test_planner.py, never
// shipped).
//
// Builds the gate queue a 30-qubit layered circuit puts on the fusion queue
// (layers of a Haar-random one-qubit matrix on every qubit, two-qubit phase
// shifts on random disjoint pairs and one 3-control one-qubit matrix, then
// the full quantum Fourier transform and random Pauli gadgets of 4-12
// qubits), in the forms the library's entry points queue them, plans its
// passes as the default planner does (pickPass, mergePass, clusterPass,
// planPass at 2^13-amplitude tiles with 3 always-local qubits, flushing
// every 2048 gates), and generates every pass's kernel source.  Each
// distinct source is compiled for gfx950 with hiprtc, and its kernel
// descriptor must show no scratch memory (a generated kernel that would use
// scratch is refused by the library, leaving its passes in the
// interpreter).  Prints one line per kernel; exits non-zero on a failure.
//
//   jit_bench_check <layers> <gadgets> <maxCompiles> [dumpDir]
#include <hip/hiprtc.h>

#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <random>
#include <set>
#include <string>
#include <vector>

#include "qa_jit.hpp"
#include "qa_tile.hpp"

extern "C" const int qa_jit_num_headers;
extern "C" const char* const qa_jit_header_names[];
extern "C" const char* const qa_jit_header_sources[];

using namespace qa::fuse;

namespace {

constexpr int N = 30, TILE = 13, LOW = 3, MAXQ_QUEUE = 2048;
// the library's default prefetch mode (Config::prefetch)
#ifndef PF_MODE
#define PF_MODE 2
#endif

std::mt19937_64 rng(2024);
double unit() { return std::uniform_real_distribution<double>(0, 1)(rng); }
int below(int n) { return (int)(rng() % (uint64_t)n); }

GOp dense1(int t, uint64_t ctrls) {
    GOp g;
    g.kind = F_DENSE1;
    g.t1 = t;
    g.cMask = g.cVal = ctrls;
    for (int i = 0; i < 8; i++) g.f[i] = unit() - 0.5;
    return g;
}
// applyTwoQubitPhaseShift(a, b): diag(1, e^{i theta}) on a controlled by b,
// queued phase-only (the target folded into the controls)
GOp phase2(int a, int b, double th) {
    GOp g;
    g.kind = F_DIAG1;
    g.t1 = a;
    g.cMask = g.cVal = (1ull << a) | (1ull << b);
    g.f[0] = 1;
    g.f[2] = std::cos(th);
    g.f[3] = std::sin(th);
    return g;
}

std::vector<GOp> benchQueue(int layers, int gadgets) {
    std::vector<GOp> q;
    for (int l = 0; l < layers; l++) {
        for (int t = 0; t < N; t++) q.push_back(dense1(t, 0));
        std::vector<int> p(N);
        for (int i = 0; i < N; i++) p[i] = i;
        std::shuffle(p.begin(), p.end(), rng);
        for (int k = 0; k < N / 2; k++) q.push_back(phase2(p[2 * k], p[2 * k + 1], 6.28 * unit()));
        std::shuffle(p.begin(), p.end(), rng);
        q.push_back(dense1(p[3], (1ull << p[0]) | (1ull << p[1]) | (1ull << p[2])));
    }
    // full QFT, in the host layer's order
    for (int n = N - 1; n >= 0; n--) {
        q.push_back(dense1(n, 0));
        for (int m = 0; m < n; m++) q.push_back(phase2(n, n - m - 1, M_PI / (double)(1ll << (m + 1))));
    }
    for (int n = 0; n < N / 2; n++) {
        GOp g;
        g.kind = F_SWAP;
        g.t1 = n;
        g.t2 = N - 1 - n;
        q.push_back(g);
    }
    for (int i = 0; i < gadgets; i++) {
        const int k = 4 + below(9);
        std::vector<int> p(N);
        for (int j = 0; j < N; j++) p[j] = j;
        std::shuffle(p.begin(), p.end(), rng);
        uint64_t x = 0, y = 0, z = 0;
        for (int j = 0; j < k; j++) {
            const int c = below(3);
            (c == 0 ? x : c == 1 ? y : z) |= 1ull << p[j];
        }
        GOp g;
        const double th = 6.28 * unit();
        if (x | y) {
            g.kind = F_PAULI;
            g.m1 = x | y;
            g.m2 = y | z;
        } else {
            g.kind = F_PARITY;
            g.m1 = z;
        }
        g.f[0] = std::cos(th);
        g.f[3] = std::sin(th);
        q.push_back(g);
    }
    return q;
}

// private segment size of the kernel descriptor "qa_jit_pass.kd" in an
// amdgcn code object (ELF64), -1 if not found
long privateSegmentSize(const std::vector<char>& elf) {
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
        for (uint64_t s = 0; s + entsize <= size; s += entsize) {
            const char* name = elf.data() + strOff + u32(off + s);
            if (std::strcmp(name, "qa_jit_pass.kd") != 0) continue;
            const uint16_t secIdx = u16(off + s + 6);
            const uint64_t value = u64(off + s + 8);
            const size_t sec = shoff + (size_t)secIdx * shentsize;
            const uint64_t secAddr = u64(sec + 0x10), secOff = u64(sec + 0x18);
            return (long)u32(secOff + (value - secAddr) + 4);  // private_segment_fixed_size
        }
    }
    return -1;
}

// compiles a generated kernel for gfx950: its scratch bytes, or -1 if it
// does not compile
long compileScratch(const std::string& src) {
    hiprtcProgram p;
    hiprtcCreateProgram(&p, src.c_str(), "qa_jit_pass.hip", qa_jit_num_headers, qa_jit_header_sources,
                        qa_jit_header_names);
    const char* opts[] = {"--offload-arch=gfx950", "-O3", "-std=c++17", "-Xclang", "-target-feature",
                          "-Xclang", "-packed-fp32-ops", "-mllvm", "-simplifycfg-sink-common=false"};
    if (hiprtcCompileProgram(p, 9, opts) != HIPRTC_SUCCESS) {
        hiprtcDestroyProgram(&p);
        return -1;
    }
    size_t n = 0;
    hiprtcGetCodeSize(p, &n);
    std::vector<char> code(n);
    hiprtcGetCode(p, code.data());
    hiprtcDestroyProgram(&p);
    return privateSegmentSize(code);
}

}  // namespace

int main(int argc, char** argv) {
    const int layers = argc > 1 ? std::atoi(argv[1]) : 40;
    const int gadgets = argc > 2 ? std::atoi(argv[2]) : 64;
    const int maxCompiles = argc > 3 ? std::atoi(argv[3]) : 1000;
    const char* dumpDir = argc > 4 ? argv[4] : nullptr;
    std::vector<GOp> all = benchQueue(layers, gadgets);
    std::set<std::string> seen;
    int passes = 0, compiled = 0, failures = 0;
    for (size_t from = 0; from < all.size(); from += MAXQ_QUEUE) {
        std::vector<GOp> ops(all.begin() + from, all.begin() + std::min(all.size(), from + MAXQ_QUEUE));
        std::vector<GOp> pass, rest, merged, planned;
        std::vector<uint32_t> picked;
        while (!ops.empty()) {
            uint64_t L = 0;
            picked.clear();
            pickPass(ops.data(), ops.size(), N, TILE, LOW, L, picked);
            if (picked.empty()) {
                std::printf("empty pass\n");
                return 1;
            }
            pass.clear();
            rest.clear();
            for (size_t i = 0, p = 0; i < ops.size(); i++) {
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
            clusterPass(merged.data(), merged.size(), (1ull << LOW) - 1ull, planned);
            TileArgs a;
            std::vector<TOp> prog;
            planPass(planned.data(), planned.size(), N, L, LOW, a, prog, true);
            setLdsOffsets(prog.data(), prog.size(), a, QA_FP64);
            ops.swap(rest);
            passes++;
            const std::string key = qa::jit::shapeKey(QA_FP64, true, PF_MODE, a, prog.data(), prog.size());
            if (!seen.insert(key).second) continue;
            const std::string src = qa::jit::generate(QA_FP64, true, PF_MODE, a, prog.data(), prog.size());
            if (dumpDir) {
                char path[512];
                std::snprintf(path, sizeof path, "%s/bench_pass_%03d.hip", dumpDir, passes);
                if (FILE* fp = std::fopen(path, "w")) {
                    std::fputs(src.c_str(), fp);
                    std::fclose(fp);
                }
            }
            if (compiled >= maxCompiles) continue;
            // as the library does: the kernel of the default prefetch mode
            // (PF_MODE), and when it would use scratch the same program with
            // less prefetch
            const long scratch = compileScratch(src);
            long lean = -2;
            if (scratch != 0) lean = compileScratch(qa::jit::generate(QA_FP64, true, PF_MODE == 3 ? 2 : 0, a, prog.data(), prog.size()));
            compiled++;
            std::printf("pass %d: %zu ops, scratch %ld bytes%s\n", passes, prog.size(), scratch,
                        lean == -2 ? "" : lean == 0 ? "; with less prefetch none" : "; with less prefetch too");
            if (scratch != 0 && lean != 0) failures++;
        }
    }
    std::printf("bench check: %d passes, %zu shapes, %d compiled, %d with scratch or failed\n", passes, seen.size(),
                compiled, failures);
    return failures ? 1 : 0;
}
