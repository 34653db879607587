// CPU check that a process ends cleanly while pass kernels are still being
// compiled (test infrastructure; built and run by tests/test_planner.py).
//
// Queues compiles of many distinct pass programs through the library's own
// generated-kernel path (qa::jit::kernelFor, background compile threads,
// compiles in qa_jitc child processes) and returns from main after
// `delayMs`, without qa_finalize, while compiles are in flight.  Without a
// GPU the module loads fail after each compile; the compiles, the threads
// and the exit sequence are the library's.  The process must exit with
// status 0, at once (the compiler processes in flight are killed).
//
//   jit_exit_check <shapes> <delayMs>
#include <chrono>
#include <cstdio>
#include <cstdlib>
#include <random>
#include <thread>
#include <vector>

#include "qa_jit.hpp"
#include "qa_tile.hpp"

using namespace qa::fuse;

int main(int argc, char** argv) {
    const int shapes = argc > 1 ? std::atoi(argv[1]) : 20;
    const int delayMs = argc > 2 ? std::atoi(argv[2]) : 500;
    constexpr int N = 30, TILE = 13, LOW = 3;
    std::mt19937_64 rng(7);
    qa::jit::setSettings(qa::jit::Settings{1, 1});
    for (int s = 0; s < shapes; s++) {
        // one-qubit gates on a varying set of qubits: a distinct program shape each
        std::vector<GOp> ops;
        for (int i = 0; i < 6 + s % 7; i++) {
            GOp g;
            g.kind = F_DENSE1;
            g.t1 = (i * 5 + s) % N;
            for (int j = 0; j < 8; j++) g.f[j] = std::uniform_real_distribution<double>(-1, 1)(rng);
            ops.push_back(g);
        }
        std::vector<uint32_t> picked;
        uint64_t L = 0;
        pickPass(ops.data(), ops.size(), N, TILE, LOW, L, picked);
        std::vector<GOp> pass;
        for (uint32_t p : picked) pass.push_back(ops[p]);
        TileArgs a;
        std::vector<TOp> prog;
        planPass(pass.data(), pass.size(), N, L, LOW, a, prog, true);
        setLdsOffsets(prog.data(), prog.size(), a, QA_FP64);
        qa::jit::kernelFor(QA_FP64, true, 2, a, prog.data(), prog.size());
    }
    std::this_thread::sleep_for(std::chrono::milliseconds(delayMs));
    std::printf("returning with compiles queued\n");
    return 0;
}
