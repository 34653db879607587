// qa_jitc: the compiler process of the generated pass kernels (qa_jit.hip).
//
// The library does not compile in its own process: a compile thread there
// would still be inside the compiler while the process exits, when the
// compiler's static state is torn down under it (with crashes and hangs at
// exit as the result).  Each compile runs this program instead: it reads
// one kernel source from stdin (an 8-byte little-endian length, then the
// text), compiles it for gfx950 with hiprtc against the device headers
// embedded in it, and writes to stdout a 4-byte status (0: compiled), an
// 8-byte length and then the code object, or the compiler's log.  It never
// touches the GPU.
#include <hip/hiprtc.h>

#include <cstdint>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

extern "C" const int qa_jit_num_headers;
extern "C" const char* const qa_jit_header_names[];
extern "C" const char* const qa_jit_header_sources[];

namespace {

bool readAll(void* p, size_t n) {
    return std::fread(p, 1, n, stdin) == n;
}

void reply(uint32_t status, const char* data, uint64_t n) {
    std::fwrite(&status, 4, 1, stdout);
    std::fwrite(&n, 8, 1, stdout);
    if (n) std::fwrite(data, 1, n, stdout);
    std::fflush(stdout);
}

}  // namespace

int main() {
    uint64_t n = 0;
    if (!readAll(&n, 8) || n > (64u << 20)) {
        reply(2, "qa_jitc: bad request", 20);
        return 2;
    }
    std::string src(n, '\0');
    if (n && !readAll(&src[0], n)) {
        reply(2, "qa_jitc: short source", 21);
        return 2;
    }
    hiprtcProgram p;
    if (hiprtcCreateProgram(&p, src.c_str(), "qa_jit_pass.hip", qa_jit_num_headers, qa_jit_header_sources,
                            qa_jit_header_names) != HIPRTC_SUCCESS) {
        reply(1, "hiprtcCreateProgram failed", 26);
        return 1;
    }
    // the options the library's own tile-pass kernels are built with, but the
    // default scheduler: straight-line programs scheduled for ILP can need
    // more than the 256 registers of two waves per SIMD and spill
    const char* opts[] = {"--offload-arch=gfx950", "-O3", "-std=c++17", "-Xclang", "-target-feature",
                          "-Xclang", "-packed-fp32-ops", "-mllvm", "-simplifycfg-sink-common=false"};
    const hiprtcResult rc = hiprtcCompileProgram(p, (int)(sizeof opts / sizeof opts[0]), opts);
    if (rc != HIPRTC_SUCCESS) {
        size_t ln = 0;
        hiprtcGetProgramLogSize(p, &ln);
        std::string log = std::string("hiprtc: ") + hiprtcGetErrorString(rc) + ": ";
        std::string body(ln + 1, '\0');
        if (ln) hiprtcGetProgramLog(p, &body[0]);
        log += body.substr(0, 400);
        hiprtcDestroyProgram(&p);
        reply(1, log.data(), log.size());
        return 1;
    }
    size_t cn = 0;
    hiprtcGetCodeSize(p, &cn);
    std::vector<char> code(cn);
    if (cn) hiprtcGetCode(p, code.data());
    hiprtcDestroyProgram(&p);
    reply(0, code.data(), code.size());
    return 0;
}
