// Pass kernels generated at run time: source generation from a pass
// program's shape, compiles in qa_jitc child processes driven by background
// threads, a per-shape cache of loaded kernels, and their launches.  See
// qa_jit.hpp.
#include <dlfcn.h>
#include <signal.h>
#include <spawn.h>
#include <sys/socket.h>
#include <sys/wait.h>
#include <unistd.h>

#include <algorithm>
#include <chrono>
#include <condition_variable>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <deque>
#include <mutex>
#include <string>
#include <thread>
#include <unordered_map>
#include <vector>

#include "qa_common.hpp"
#include "qa_jit.hpp"

extern char** environ;

// the device headers of the generated kernels, embedded at build time
extern "C" const int qa_jit_num_headers;
extern "C" const char* const qa_jit_header_names[];
extern "C" const char* const qa_jit_header_sources[];

namespace qa {
namespace jit {
namespace {

using fuse::TOp;
using fuse::TileArgs;

enum State { NONE, QUEUED, READY, FAILED };

struct Entry {
    int seen = 0;
    State state = NONE;
    hipModule_t module = nullptr;
    hipFunction_t fn = nullptr;
    unsigned residentBlock = 0, residentValue = 0;  // residentBlocks cache (block size -> blocks)
    size_t residentLds = 0;
};

struct Job {
    std::string key;
    std::string src;      // the kernel as asked for
    std::string srcLean;  // the same program with less prefetch (or empty)
    int prec;
};

std::mutex g_mu;  // everything below
std::condition_variable g_work, g_done;
std::unordered_map<std::string, Entry> g_cache;
std::unordered_map<const void*, Entry*> g_byFn;
std::deque<Job> g_jobs;
int g_inFlight = 0;
std::vector<std::thread> g_workers;
bool g_stop = false;
bool g_exiting = false;  // set at exit: no compile threads start again
bool g_exitRegistered = false;
Settings g_settings;
Stats g_stats;
std::string g_lastError;

// Stops and joins the compile threads; at exit (exiting) also kills the
// compiler processes in flight, whose results would not be used.  Compiles
// run in child processes (qa_jitc), so nothing of the compiler lives in
// this process; a compile thread only waits on its child, and a thread
// whose child was killed ends at once.  Registered with atexit when the
// threads start (the HIP runtime the threads load modules with is
// initialised before, so it is torn down after this runs) and run by
// qa_finalize (stopWorkers(false)).
std::vector<pid_t> g_children;  // compiler processes in flight (under g_mu)

void stopWorkers(bool exiting) {
    std::vector<std::thread> workers;
    {
        std::lock_guard<std::mutex> lk(g_mu);
        g_stop = true;
        g_exiting = g_exiting || exiting;
        g_jobs.clear();
        workers.swap(g_workers);
        if (exiting)
            for (pid_t c : g_children) (void)kill(c, SIGKILL);
    }
    g_work.notify_all();
    for (auto& w : workers)
        if (w.get_id() != std::this_thread::get_id()) w.join();
        else w.detach();
}

// the same for a library unloaded without qa_finalize and without exit
// (a joinable std::thread may not be destroyed); declared after the state
// it uses, so it is destroyed first
struct WorkerJoin {
    ~WorkerJoin() { stopWorkers(true); }
} g_workerJoin;

void stopAtExit() { stopWorkers(true); }

// the compiler program: QA_JITC, else qa_jitc beside this library
std::string compilerPath() {
    if (const char* p = std::getenv("QA_JITC")) return p;
    Dl_info info;
    if (dladdr((const void*)&compilerPath, &info) && info.dli_fname) {
        std::string lib = info.dli_fname;
        const size_t slash = lib.rfind('/');
        return (slash == std::string::npos ? std::string(".") : lib.substr(0, slash)) + "/qa_jitc";
    }
    return "qa_jitc";
}

}  // namespace

bool available() {
    static const bool ok = access(compilerPath().c_str(), X_OK) == 0;
    return ok;
}

namespace {

bool sendAll(int fd, const void* p, size_t n) {
    const char* c = (const char*)p;
    while (n) {
        const ssize_t w = send(fd, c, n, MSG_NOSIGNAL);
        if (w <= 0) return false;
        c += w;
        n -= (size_t)w;
    }
    return true;
}
bool recvAll(int fd, void* p, size_t n) {
    char* c = (char*)p;
    while (n) {
        const ssize_t r = recv(fd, c, n, 0);
        if (r <= 0) return false;
        c += r;
        n -= (size_t)r;
    }
    return true;
}

// compiles one source in a qa_jitc child process: the code object, or
// false with `err` set
bool runCompiler(const std::string& src, std::vector<char>& code, std::string& err) {
    static const std::string path = compilerPath();
    int sv[2];
    if (socketpair(AF_UNIX, SOCK_STREAM | SOCK_CLOEXEC, 0, sv) != 0) {
        err = "socketpair failed";
        return false;
    }
    posix_spawn_file_actions_t fa;
    posix_spawn_file_actions_init(&fa);
    posix_spawn_file_actions_adddup2(&fa, sv[1], 0);
    posix_spawn_file_actions_adddup2(&fa, sv[1], 1);
    // the child gets no other descriptor of this process (device files included)
    posix_spawn_file_actions_addclosefrom_np(&fa, 3);
    char* argv[] = {const_cast<char*>(path.c_str()), nullptr};
    pid_t pid = -1;
    int rc;
    {
        std::lock_guard<std::mutex> lk(g_mu);
        rc = g_exiting ? -1 : posix_spawn(&pid, path.c_str(), &fa, nullptr, argv, environ);
        if (rc == 0) g_children.push_back(pid);
    }
    posix_spawn_file_actions_destroy(&fa);
    close(sv[1]);
    if (rc != 0) {
        close(sv[0]);
        err = "cannot start the kernel compiler " + path;
        return false;
    }
    const uint64_t n = src.size();
    uint32_t status = 99;
    uint64_t len = 0;
    bool ok = sendAll(sv[0], &n, 8) && sendAll(sv[0], src.data(), src.size());
    shutdown(sv[0], SHUT_WR);
    ok = ok && recvAll(sv[0], &status, 4) && recvAll(sv[0], &len, 8) && len <= (256u << 20);
    std::vector<char> payload;
    if (ok) {
        payload.resize(len);
        ok = recvAll(sv[0], payload.data(), len);
    }
    close(sv[0]);
    int ws = 0;
    (void)waitpid(pid, &ws, 0);
    {
        std::lock_guard<std::mutex> lk(g_mu);
        g_children.erase(std::remove(g_children.begin(), g_children.end(), pid), g_children.end());
    }
    if (!ok) {
        err = "the kernel compiler " + path + " ended without a result";
        return false;
    }
    if (status != 0) {
        err.assign(payload.begin(), payload.end());
        return false;
    }
    code.swap(payload);
    return true;
}

// ---- shapes ---------------------------------------------------------------

bool hasGc(const TOp& t) { return t.h.gcMask || t.h.gcVal; }
bool hasJc(const TOp& t) { return t.h.jcMask || t.h.jcVal; }

// the number of terms of a PHASES op (a nibble per slot in rm1)
int phaseCount(const TOp& t) {
    int n = 0;
    for (int sl = 0; sl <= fuse::RB; sl++) n += (int)((t.h.rm1 >> (4 * sl)) & 15u);
    return n;
}

// the shape of one op: what its generated code writes in as constants
std::string opShape(const TOp& t) {
    char buf[192];
    const int* w = reinterpret_cast<const int*>(t.b);
    switch (t.h.kind) {
        case fuse::F_GROUP:
            snprintf(buf, sizeof buf, "G%d,%d,%d,%d;", w[0], w[1], w[2], w[3]);
            break;
        case fuse::F_PAULI_LDS:
            snprintf(buf, sizeof buf, "L;");
            break;
        case fuse::F_DENSE_LDS:
            snprintf(buf, sizeof buf, "D%d,%d,%d,%d,%d;", w[4], w[0], w[1], w[2], w[3]);
            break;
        case fuse::F_BLOCK:
            snprintf(buf, sizeof buf, "B%d,%d,%d,%d,%d;", w[0], w[1], w[2], w[3], w[4]);
            break;
        case fuse::F_PHASES: {
            // every term's slot and condition masks are written in; only
            // the factors are read from the program
            // (and the op's own non-local condition, empty as planned)
            snprintf(buf, sizeof buf, "P%llx,%llx,%d|", (unsigned long long)t.h.gcMask,
                     (unsigned long long)t.h.gcVal, hasJc(t) ? 1 : 0);
            std::string k = buf;
            const fuse::PhaseTerm* pt = reinterpret_cast<const fuse::PhaseTerm*>(t.b);
            for (int i = 0; i < phaseCount(t); i++) {
                snprintf(buf, sizeof buf, "%x,%x,%x,%llx,%llx,", pt[i].rm, pt[i].jm, pt[i].jv,
                         (unsigned long long)pt[i].gm, (unsigned long long)pt[i].gv);
                k += buf;
            }
            return k + ";";
        }
        default:
            snprintf(buf, sizeof buf, "O%d,%d,%d,%u,%u,%u,%u,%u,%d,%d;", t.h.kind, t.h.rb1, t.h.rb2, t.h.pad, t.h.rm1,
                     t.h.rm2, t.h.rcMask, t.h.rcVal, hasGc(t) ? 1 : 0, hasJc(t) ? 1 : 0);
    }
    return buf;
}

const char* typeName(int prec) { return prec == QA_FP64 ? "double" : "float"; }

// ---- compiles -------------------------------------------------------------

uint64_t fnv1a(uint64_t h, const char* p, size_t n) {
    for (size_t i = 0; i < n; i++) h = (h ^ (unsigned char)p[i]) * 1099511628211ull;
    return h;
}

// the FNV-1a hash of the device headers compiled into every generated
// kernel (their names and text, as embedded in this library)
uint64_t headersHash() {
    static const uint64_t h = [] {
        uint64_t v = 1469598103934665603ull;
        for (int i = 0; i < qa_jit_num_headers; i++) {
            v = fnv1a(v, qa_jit_header_names[i], std::strlen(qa_jit_header_names[i]) + 1);
            v = fnv1a(v, qa_jit_header_sources[i], std::strlen(qa_jit_header_sources[i]) + 1);
        }
        return v;
    }();
    return h;
}

// QA_JIT_CACHE_DIR: code objects kept across processes, one file per
// kernel (named by a 64-bit FNV-1a hash of the source, which holds the
// whole shape, continued from the hash of the device headers it includes,
// so that a library built from other headers finds none of the old
// objects), written to a temporary name and renamed into place
std::string cachePath(const std::string& src) {
    const char* dir = std::getenv("QA_JIT_CACHE_DIR");
    if (!dir || !*dir) return std::string();
    const uint64_t h = fnv1a(headersHash(), src.data(), src.size());
    char name[64];
    snprintf(name, sizeof name, "/qa_jit_%016llx_%zu.co", (unsigned long long)h, src.size());
    return std::string(dir) + name;
}

bool cacheLoad(const std::string& path, std::vector<char>& code) {
    if (path.empty()) return false;
    FILE* f = std::fopen(path.c_str(), "rb");
    if (!f) return false;
    std::fseek(f, 0, SEEK_END);
    const long n = std::ftell(f);
    std::fseek(f, 0, SEEK_SET);
    bool ok = n > 4;
    if (ok) {
        code.resize((size_t)n);
        ok = std::fread(code.data(), 1, (size_t)n, f) == (size_t)n && std::memcmp(code.data(), "\x7f" "ELF", 4) == 0;
    }
    std::fclose(f);
    return ok;
}

void cacheStore(const std::string& path, const std::vector<char>& code) {
    if (path.empty()) return;
    char tmp[64];
    snprintf(tmp, sizeof tmp, ".tmp%d_%p", (int)getpid(), (const void*)&code);
    const std::string t = path + tmp;
    FILE* f = std::fopen(t.c_str(), "wb");
    if (!f) return;
    const bool ok = std::fwrite(code.data(), 1, code.size(), f) == code.size();
    std::fclose(f);
    if (!ok || std::rename(t.c_str(), path.c_str()) != 0) (void)std::remove(t.c_str());
}

bool compileSource(const std::string& src, hipModule_t& mod, hipFunction_t& fn, std::string& err) {
    std::vector<char> code;
    const std::string cached = cachePath(src);
    if (!cacheLoad(cached, code)) {
        if (!runCompiler(src, code, err)) return false;
        cacheStore(cached, code);
    }
    {
        std::lock_guard<std::mutex> lk(g_mu);
        if (g_exiting) {
            err = "the process is exiting";
            return false;
        }
    }
    if (hipModuleLoadData(&mod, code.data()) != hipSuccess) {
        err = "hipModuleLoadData failed";
        return false;
    }
    if (hipModuleGetFunction(&fn, mod, "qa_jit_pass") != hipSuccess) {
        err = "hipModuleGetFunction failed";
        (void)hipModuleUnload(mod);
        return false;
    }
    // the tile must sit at LDS address 0 (no static LDS) and the register
    // array must not have gone to scratch memory
    int shared = -1, local = -1;
    (void)hipFuncGetAttribute(&shared, HIP_FUNC_ATTRIBUTE_SHARED_SIZE_BYTES, fn);
    (void)hipFuncGetAttribute(&local, HIP_FUNC_ATTRIBUTE_LOCAL_SIZE_BYTES, fn);
    if (shared != 0 || local != 0) {
        char buf[128];
        snprintf(buf, sizeof buf, "generated kernel has %d bytes of static LDS and %d bytes of scratch", shared, local);
        err = buf;
        (void)hipModuleUnload(mod);
        return false;
    }
    return true;
}

// A program too large for the registers of the prefetching kernel (it
// would spill to scratch) is compiled again with less: a kernel with
// deferred stores (pf 3) as one with spread loads only (pf 2), which keeps
// fewer addresses live; otherwise without the prefetch, which frees the 64
// registers of the next tile.
bool compileOne(const Job& job, hipModule_t& mod, hipFunction_t& fn, std::string& err) {
    if (compileSource(job.src, mod, fn, err)) return true;
    if (job.srcLean.empty()) return false;
    std::string err2;
    if (compileSource(job.srcLean, mod, fn, err2)) return true;
    err += "; with less prefetch: " + err2;
    return false;
}

void worker(int device) {
    (void)hipSetDevice(device);
    for (;;) {
        Job job;
        {
            std::unique_lock<std::mutex> lk(g_mu);
            g_work.wait(lk, [] { return g_stop || !g_jobs.empty(); });
            if (g_stop) return;
            job = std::move(g_jobs.front());
            g_jobs.pop_front();
            g_inFlight++;
        }
        const auto t0 = std::chrono::steady_clock::now();
        hipModule_t mod = nullptr;
        hipFunction_t fn = nullptr;
        std::string err;
        const bool ok = compileOne(job, mod, fn, err);
        const double dt = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
        {
            std::lock_guard<std::mutex> lk(g_mu);
            Entry& e = g_cache[job.key];
            if (ok) {
                e.state = READY;
                e.module = mod;
                e.fn = fn;
                g_byFn[(const void*)fn] = &e;
                g_stats.kernels++;
            } else {
                e.state = FAILED;
                g_stats.failures++;
                g_lastError = err;
                // QA_JIT_DUMP_DIR: keep the source of each kernel that failed, for diagnosis
                if (const char* dir = std::getenv("QA_JIT_DUMP_DIR")) {
                    char path[512];
                    snprintf(path, sizeof path, "%s/failed_%lld.hip", dir, (long long)g_stats.failures);
                    if (FILE* fp = std::fopen(path, "w")) {
                        std::fprintf(fp, "// %s\n", err.c_str());
                        std::fputs(job.src.c_str(), fp);
                        std::fclose(fp);
                    }
                }
            }
            g_stats.compileSeconds += dt;
            g_inFlight--;
        }
        g_done.notify_all();
    }
}

// under g_mu
void ensureWorkers() {
    if (!g_workers.empty() || g_exiting) return;
    int dev = 0;
    (void)hipGetDevice(&dev);
    unsigned hw = std::thread::hardware_concurrency();
    int n = (int)(hw > 2 ? hw - 1 : 1);
    if (n > 16) n = 16;
    if (!g_exitRegistered) {
        (void)std::atexit(stopAtExit);
        g_exitRegistered = true;
    }
    g_stop = false;
    for (int i = 0; i < n; i++) g_workers.emplace_back(worker, dev);
}

}  // namespace

// ---- source generation ----------------------------------------------------

std::string shapeKey(int prec, bool nt, int pf, const TileArgs& a, const TOp* prog, size_t num) {
    char buf[128];
    snprintf(buf, sizeof buf, "p%d n%d f%d k%d d%d%d%d o%d|", prec, nt ? 1 : 0, pf, a.k, a.firstDirect,
             a.lastDirect, a.usesLds, a.numOps);
    std::string k = buf;
    for (size_t i = 0; i < num; i++) k += opShape(prog[i]);
    return k;
}

// ops that exchange the tile through LDS between workgroup barriers
static bool exchanges(const TOp& t) {
    return t.h.kind == fuse::F_GROUP || t.h.kind == fuse::F_PAULI_LDS || t.h.kind == fuse::F_DENSE_LDS;
}

std::string generate(int prec, bool nt, int pf, const TileArgs& a, const TOp* prog, size_t num) {
    std::string s;
    s.reserve(4096 + num * 512);
    s += "#include \"qa_tiledev.hpp\"\n"
         "namespace qa {\nnamespace fuse {\n"
         "struct JitProgram {\n"
         "    template <typename T, int NT, typename Pf>\n"
         "    __device__ __forceinline__ static void run(T* __restrict__ amps, const TileArgs& args,\n"
         "                                               const TOp* __restrict__ ops, uint64_t base, ct<T>* r,\n"
         "                                               ct<T>* tile, const RegLayout& row, const RegLayout& first,\n"
         "                                               uint32_t tid, Pf& pf) {\n"
         // the program and the pass arguments addressed through a value that
         // changes per tile (base >> 62 is 0 for every state): loop-invariant
         // addresses would be hoisted out of the tile loop and, one per op,
         // held in scalar registers for the whole loop
         "        ops += (size_t)(base >> 62);\n"
         "        const TileArgs& a = *(&args + (size_t)(base >> 62));\n"
         "        TileState st;\n"
         "        tileBegin<T, NT>(st, a, r, tile, row, first);\n";
    char buf[512];
    // pf >= 2: the next tile's loads in PF_CHUNKS groups (pf 3: each group
    // after the stores of the previous tile's results held in the same
    // registers).  The ops that exchange the tile through LDS (with a
    // workgroup barrier, which loads are not moved across) cut the program
    // into S = exchanges + 1 segments; group c starts segment
    // c * (S - 1) / PF_CHUNKS (c * S / PF_CHUNKS for S < 3), so the loads
    // leave through the tile's work and none in its last segment, where
    // they could still be in flight at the stores (measured 0.3% faster
    // per pass than spreading over all S segments).
    const size_t first = a.firstDirect ? 1 : 0;
    int S = 1;
    for (size_t i = first; i < num; i++)
        if (exchanges(prog[i])) S++;
    const int spread = S >= 3 ? S - 1 : S;
    int chunk = 0, seg = 0;
    auto loadsUpTo = [&](int upToSeg) {
        for (; pf >= 2 && chunk < PF_CHUNKS && chunk * spread / PF_CHUNKS <= upToSeg; chunk++) {
            snprintf(buf, sizeof buf, "        pf.template part<%d, %d>();\n", chunk * (fuse::RA / PF_CHUNKS),
                     (chunk + 1) * (fuse::RA / PF_CHUNKS));
            s += buf;
        }
    };
    for (size_t i = first; i < num; i++) {
        loadsUpTo(seg);
        const TOp& t = prog[i];
        if (exchanges(t)) seg++;
        const int* w = reinterpret_cast<const int*>(t.b);
        switch (t.h.kind) {
            case fuse::F_GROUP:
                snprintf(buf, sizeof buf,
                         "        { const int slots[RB] = {%d, %d, %d, %d};\n"
                         "          stepGroup<T, NT>(st, a, ops + %zu, slots, r, tile, tid); }\n",
                         w[0], w[1], w[2], w[3], i);
                s += buf;
                break;
            case fuse::F_PAULI_LDS:
                snprintf(buf, sizeof buf,
                         "        { const TOp* op_ = ops + %zu; u32x16 ch, cb; fetchOp(op_, ch, cb); OpRegs o;\n"
                         "          unpackOp(ch, cb, o);\n"
                         "          stepPauliLds<T, NT>(st, a, op_, o, ((uint64_t)cb[13] << 32) | cb[12], r, tile,\n"
                         "                              base, tid); }\n",
                         i);
                s += buf;
                break;
            case fuse::F_DENSE_LDS:
                snprintf(buf, sizeof buf,
                         "        { const TOp* op_ = ops + %zu; u32x16 ch, cb; fetchOp(op_, ch, cb);\n"
                         "          const uint32_t dslots[4] = {%du, %du, %du, %du};\n"
                         "          stepDenseLds<T, NT, %d>(st, op_, dslots,\n"
                         "              reinterpret_cast<const double*>(cb[6] | ((uint64_t)cb[7] << 32)), (int)cb[5], a.k,\n"
                         "              r, tile, tid); }\n",
                         i, w[0], w[1], w[2], w[3], w[4]);
                s += buf;
                break;
            case fuse::F_BLOCK:
                snprintf(buf, sizeof buf,
                         "        { const TOp* op_ = ops + %zu; u32x16 ch, cb; fetchOp(op_, ch, cb);\n"
                         "          const int bslots[BQ] = {%d, %d, %d, %d, %d};\n"
                         "          stepBlock<T, NT>(st, op_, bslots, a.blocks + (cb[6] | ((uint64_t)cb[7] << 32)),\n"
                         "                           a.k, r, tile, tid); }\n",
                         i, w[0], w[1], w[2], w[3], w[4]);
                s += buf;
                break;
            case fuse::F_PHASES: {
                if (hasJc(t)) {
                    // planned PHASES ops carry no op-level lane condition
                    // (their terms hold every condition); keep the interpreter
                    s.clear();
                    return s;
                }
                // each term as straight-line code with its condition masks
                // written in (the interpreter's applyPhases, unrolled)
                const fuse::PhaseTerm* pt = reinterpret_cast<const fuse::PhaseTerm*>(t.b);
                snprintf(buf, sizeof buf,
                         "        { const PhaseTerm* pt = reinterpret_cast<const PhaseTerm*>((ops + %zu)->b);\n"
                         "          const uint32_t jb = st.cur.jb;\n"
                         "          if ((base & 0x%llxull) == 0x%llxull) {\n"
                         "          ct<T> acc[1 + RB];\n"
                         "          for (int s_ = 0; s_ <= RB; s_++) acc[s_] = ct<T>{(T)1, (T)0};\n",
                         i, (unsigned long long)t.h.gcMask, (unsigned long long)t.h.gcVal);
                s += buf;
                int k = 0;
                for (int sl = 0; sl <= fuse::RB; sl++) {
                    const int ns = (int)((t.h.rm1 >> (4 * sl)) & 15u);
                    for (int e = 0; e < ns; e++, k++) {
                        snprintf(buf, sizeof buf,
                                 "          phaseTerm(acc[%d], (jb & 0x%xu) == 0x%xu && (base & 0x%llxull) == 0x%llxull,\n"
                                 "                    pt[%d].fre, pt[%d].fim);\n",
                                 sl, pt[k].jm, pt[k].jv, (unsigned long long)pt[k].gm, (unsigned long long)pt[k].gv,
                                 k, k);
                        s += buf;
                    }
                }
                s += "          phasesFinish<T>(r, acc); } }\n";
                break;
            }
            default:
                // the op's arm alone, its shape fields as constants
                snprintf(buf, sizeof buf,
                         "        { const TOp* op_ = ops + %zu; u32x16 ch, cb; fetchOp(op_, ch, cb); OpRegs o;\n"
                         "          unpackOp(ch, cb, o);\n"
                         "          o.h.kind = %d; o.h.rb1 = %d; o.h.rb2 = %d; o.h.pad = %uu; o.h.rm1 = %uu;\n"
                         "          o.h.rm2 = %uu; o.h.rcMask = %uu; o.h.rcVal = %uu;\n"
                         "          stepFixed<T, %d, %d, %d, %uu, %uu, %s, %s, %s>(st, op_, o, r, base); }\n",
                         i, t.h.kind, t.h.rb1, t.h.rb2, t.h.pad, t.h.rm1, t.h.rm2, t.h.rcMask, t.h.rcVal, t.h.kind,
                         t.h.rb1, t.h.rb2, t.h.pad, t.h.rm1, hasGc(t) ? "true" : "false", hasJc(t) ? "true" : "false",
                         t.h.rcMask ? "true" : "false");
                s += buf;
        }
    }
    loadsUpTo(S);
    s += "        pf.template finish<T, NT>(st, a, amps, r, tile, row, base);\n"
         "    }\n"
         "};\n"
         "}  // namespace fuse\n}  // namespace qa\n";
    snprintf(buf, sizeof buf,
             "extern \"C\" __global__ void __launch_bounds__(512)\n"
             "qa_jit_pass(%s* __restrict__ amps, qa::fuse::TileArgs a, const qa::fuse::TOp* __restrict__ ops) {\n"
             "    qa::fuse::tilePassLoop<%s, %d, %d, qa::fuse::JitProgram>(amps, a, ops);\n"
             "}\n",
             typeName(prec), typeName(prec), nt ? 1 : 0, pf);
    s += buf;
    return s;
}

// ---- the cache ------------------------------------------------------------

Settings settings() {
    std::lock_guard<std::mutex> lk(g_mu);
    return g_settings;
}

void setSettings(const Settings& s) {
    std::lock_guard<std::mutex> lk(g_mu);
    g_settings = s;
}

const void* kernelFor(int prec, bool nt, int pf, const TileArgs& a, const TOp* prog, size_t num, bool force) {
    Settings s = settings();
    if (force) s.mode = 2;
    if (s.mode == 0) return nullptr;
    const std::string key = shapeKey(prec, nt, pf, a, prog, num);
    std::unique_lock<std::mutex> lk(g_mu);
    Entry& e = g_cache[key];  // element references survive rehashing
    e.seen++;
    if (e.state == READY) return (const void*)e.fn;
    if (e.state == NONE && !g_exiting && (s.mode == 2 || e.seen >= s.minSeen)) {
        e.state = QUEUED;
        lk.unlock();
        std::string src = generate(prec, nt, pf, a, prog, num);
        std::string lean = pf ? generate(prec, nt, pf == 3 ? 2 : 0, a, prog, num) : std::string();
        lk.lock();
        if (src.empty()) {
            e.state = FAILED;
            return nullptr;
        }
        g_jobs.push_back(Job{key, std::move(src), std::move(lean), prec});
        ensureWorkers();
        g_work.notify_one();
    }
    if (s.mode == 2 && e.state == QUEUED) g_done.wait(lk, [&] { return e.state == READY || e.state == FAILED; });
    return e.state == READY ? (const void*)e.fn : nullptr;
}

int launch(const void* kernel, unsigned grid, unsigned block, size_t lds, void* amps, const TileArgs& a,
           const TOp* ops) {
    void* am = amps;
    TileArgs aa = a;
    const TOp* op = ops;
    void* args[] = {&am, &aa, &op};
    return (int)hipModuleLaunchKernel((hipFunction_t)kernel, grid, 1, 1, block, 1, 1, (unsigned)lds, stream(), args,
                                      nullptr);
}

unsigned residentBlocks(const void* kernel, unsigned block, size_t lds) {
    std::lock_guard<std::mutex> lk(g_mu);
    auto it = g_byFn.find(kernel);
    Entry* e = it == g_byFn.end() ? nullptr : it->second;
    if (e && e->residentBlock == block && e->residentLds == lds && e->residentValue) return e->residentValue;
    int perCU = 0, dev = 0, cus = 0;
    if (hipModuleOccupancyMaxActiveBlocksPerMultiprocessor(&perCU, (hipFunction_t)kernel, (int)block, lds) !=
            hipSuccess ||
        perCU < 1)
        perCU = 1;
    (void)hipGetDevice(&dev);
    if (hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess || cus < 1) cus = 1;
    const unsigned r = (unsigned)(perCU * cus);
    if (e) {
        e->residentBlock = block;
        e->residentLds = lds;
        e->residentValue = r;
    }
    return r;
}

void waitIdle() {
    std::unique_lock<std::mutex> lk(g_mu);
    g_done.wait(lk, [] { return g_jobs.empty() && g_inFlight == 0; });
}

Stats stats() {
    std::lock_guard<std::mutex> lk(g_mu);
    return g_stats;
}

void countLaunch() {
    std::lock_guard<std::mutex> lk(g_mu);
    g_stats.launches++;
}

const char* lastError() {
    static thread_local std::string copy;
    std::lock_guard<std::mutex> lk(g_mu);
    copy = g_lastError;
    return copy.c_str();
}

void shutdown() {
    stopWorkers(false);
    std::lock_guard<std::mutex> lk(g_mu);
    for (auto& kv : g_cache)
        if (kv.second.module) (void)hipModuleUnload(kv.second.module);
    g_cache.clear();
    g_byFn.clear();
    g_inFlight = 0;
    g_stop = false;
    g_stats = Stats();
    g_lastError.clear();
}

}  // namespace jit
}  // namespace qa

using namespace qa;

extern "C" {

int qa_setFusionJit(int mode, int minSeen) {
    QA_ARG_CHECK(mode >= 0 && mode <= 2, "qa_setFusionJit: mode must be 0 (off), 1 (background) or 2 (at first use)");
    QA_ARG_CHECK(minSeen >= 1, "qa_setFusionJit: minSeen must be at least 1");
    jit::Settings s;
    s.mode = mode;
    s.minSeen = minSeen;
    jit::setSettings(s);
    return QA_OK;
}

int qa_getFusionJit(int* mode, int* minSeen) {
    const jit::Settings s = jit::settings();
    if (mode) *mode = s.mode;
    if (minSeen) *minSeen = s.minSeen;
    return QA_OK;
}

int qa_fusionJitWait(void) {
    jit::waitIdle();
    return QA_OK;
}

int qa_getFusionJitStats(int64_t* kernels, int64_t* failures, int64_t* launches, double* compileSeconds) {
    const jit::Stats s = jit::stats();
    if (kernels) *kernels = s.kernels;
    if (failures) *failures = s.failures;
    if (launches) *launches = s.launches;
    if (compileSeconds) *compileSeconds = s.compileSeconds;
    return QA_OK;
}

const char* qa_fusionJitLastError(void) { return jit::lastError(); }

}  // extern "C"
