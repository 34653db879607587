// Pass kernels generated at run time (qa_jit.hip).
//
// The interpreter kernel (k_tilePass) reads each op's kind from the pass
// program and dispatches to its code; that dispatch, and the register
// copies the compiler places where all op kinds join, cost about as much as
// a one-qubit gate's arithmetic per op.  A pass program's "shape" -- its op
// kinds, register bits, register masks, group layouts and which conditions
// are present -- is written into HIP source as constants, around the same
// device steps the interpreter calls (qa_tiledev.hpp), and compiled with
// hiprtc into a straight-line kernel for that shape.  Matrices, factors and
// the remaining masks stay in the program the kernel reads, so one kernel
// serves every pass of the same shape, with the interpreter's arithmetic
// bit for bit.  Kernels are cached per shape for the life of the library;
// compiles run in child processes (qa_jitc, hiprtc) driven by background
// threads, and a pass whose shape has no kernel
// yet runs in the interpreter.
#pragma once

#include <cstdint>
#include <string>

#include "qa_tiledev.hpp"

namespace qa {
namespace jit {

struct Settings {
    int mode = 1;     // 0 off, 1 background compiles, 2 compile at first use and wait
    int minSeen = 2;  // background mode: compile a shape once planned this many times
};

// the settings (qa_setFusionJit)
Settings settings();
// whether the compiler program (qa_jitc) is there to run
bool available();
void setSettings(const Settings& s);

// The kernel for this pass, or nullptr when the interpreter runs it (JIT
// off, the shape not compiled yet, or its compile failed).  Counts the
// sighting, queues a compile per the settings, and in mode 2 waits for it;
// force: compile now and wait whatever the settings (a pass the interpreter
// cannot run).
const void* kernelFor(int prec, bool nt, int pf, const fuse::TileArgs& a, const fuse::TOp* prog, size_t num,
                      bool force = false);

// launches a generated kernel (returns a HIP error code)
int launch(const void* kernel, unsigned grid, unsigned block, size_t lds, void* amps, const fuse::TileArgs& a,
           const fuse::TOp* ops);

// resident workgroups of a generated kernel on the whole device
unsigned residentBlocks(const void* kernel, unsigned block, size_t lds);

// blocks until every queued compile has finished
void waitIdle();

struct Stats {
    int64_t kernels = 0;    // compiled and loaded
    int64_t failures = 0;   // compiles or loads that failed (those shapes stay interpreted)
    int64_t launches = 0;   // passes run by generated kernels
    double compileSeconds = 0;
};
Stats stats();
void countLaunch();
const char* lastError();

// stops the compile threads and unloads every module (library finalize)
void shutdown();

// the source of a pass's kernel and its cache key (exposed for tests)
std::string shapeKey(int prec, bool nt, int pf, const fuse::TileArgs& a, const fuse::TOp* prog, size_t num);
// pf: 0 no prefetch, 1 the next tile loaded at the loop head, 2 its loads
// spread over the program in PF_CHUNKS groups, 3 as 2 with each group
// after the stores of the previous tile's results (two register sets)
constexpr int PF_CHUNKS = 4;
std::string generate(int prec, bool nt, int pf, const fuse::TileArgs& a, const fuse::TOp* prog, size_t num);

}  // namespace jit
}  // namespace qa
