// Environment, error handling, validation helpers, qureg registry and host
// RNG of the host layer.
//
// Behaviour follows quest/src/api/environment.cpp (init/sync/finalize),
// quest/src/api/debug.cpp (seeds, validation toggles, error handler),
// quest/src/core/validation.cpp:1073-1100 (default handler prints and
// exits), quest/src/core/randomiser.cpp:48-137 (one mt19937_64 seeded via
// std::seed_seq; outcome = U[0,1) > p0) and quest/src/api/qureg.cpp:37-83
// (Qureg field population for a single node).
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <mutex>
#include <random>
#include <unordered_map>

#include "host_core.hpp"

namespace qah {

namespace {

bool g_envInit = false;
bool g_envFinalized = false;
bool g_validate = true;
const double DEFAULT_EPS = 1e-12;  // precision.h:133-142 for FP64
double g_eps = DEFAULT_EPS;
void (*g_handler)(const char*, const char*) = nullptr;

std::vector<unsigned> g_seeds;
std::mt19937_64 g_gen;

std::mutex g_regMutex;
std::unordered_map<const void*, int> g_precision;  // gpuAmps -> QA_FP32/QA_FP64

void defaultHandler(const char* func, const char* msg) {
    std::printf("\nQuEST encountered a validation error during function '%s':\n%s\nExiting...\n", func, msg);
    std::fflush(stdout);
    std::exit(EXIT_FAILURE);
}

}  // namespace

void fail(const std::string& msg) { throw ApiError{msg}; }

void handleError(const char* func, const std::string& msg) {
    (g_handler ? g_handler : defaultHandler)(func, msg.c_str());
}

void check(int rc) {
    if (rc != QA_OK) fail(std::string("GPU backend: ") + qa_lastError());
}

bool envInit() { return g_envInit; }
void requireEnv() {
    if (!g_envInit) fail("The QuEST environment has not been initialised; call initQuESTEnv() first.");
}
bool validating() { return g_validate; }
double epsilon() { return g_eps; }
double epsilonFor(const Qureg& q) {
    const double FP32_EPS = 1e-5;
    return (precOf(q) == QA_FP32 && g_eps < FP32_EPS) ? FP32_EPS : g_eps;
}

int precOf(const Qureg& q) {
    std::lock_guard<std::mutex> lk(g_regMutex);
    auto it = g_precision.find(q.gpuAmps);
    return it == g_precision.end() ? QA_FP64 : it->second;
}
void registerQureg(const Qureg& q, int prec) {
    std::lock_guard<std::mutex> lk(g_regMutex);
    g_precision[q.gpuAmps] = prec;
}
void unregisterQureg(const Qureg& q) {
    std::lock_guard<std::mutex> lk(g_regMutex);
    g_precision.erase(q.gpuAmps);
}
bool isKnownQureg(const Qureg& q) {
    std::lock_guard<std::mutex> lk(g_regMutex);
    return g_precision.count(q.gpuAmps) != 0;
}

Qureg makeQureg(int numQubits, int isDensMatr, int prec) {
    requireEnv();
    if (numQubits < 1) fail("Invalid number of qubits (" + std::to_string(numQubits) + "). Must create >= 1 qubit.");
    if (prec != QA_FP32 && prec != QA_FP64) fail("Invalid precision: must be 1 (complex64) or 2 (complex128).");
    int bits = isDensMatr ? 2 * numQubits : numQubits;
    if (bits > 40)
        fail("Cannot create a Qureg of " + std::to_string(numQubits) + " qubits: its " +
             std::to_string(bits) + " index bits exceed the single-GPU limit.");
    Qureg q;
    std::memset(&q, 0, sizeof(q));
    q.isMultithreaded = 0;
    q.isGpuAccelerated = 1;
    q.isDistributed = 0;
    q.rank = 0;
    q.numNodes = 1;
    q.logNumNodes = 0;
    q.isDensityMatrix = isDensMatr ? 1 : 0;
    q.numQubits = numQubits;
    q.logNumAmps = bits;
    q.numAmps = 1LL << bits;
    q.logNumAmpsPerNode = bits;
    q.numAmpsPerNode = q.numAmps;
    q.logNumColsPerNode = isDensMatr ? numQubits : 0;
    size_t bytes = (size_t)q.numAmps * (prec == QA_FP64 ? 16 : 8);
    size_t freeB = 0, totalB = 0;
    check(qa_getMemInfo(&freeB, &totalB));
    if (bytes > freeB)
        fail("Creating a Qureg of " + std::to_string(numQubits) + " qubits requires " + std::to_string(bytes) +
             " bytes of GPU memory, but only " + std::to_string(freeB) + " are free.");
    void* d = nullptr;
    int rc = qa_malloc(&d, bytes);
    if (rc == QA_ERR_OOM) fail(std::string("Allocation of the GPU amplitudes failed: ") + qa_lastError());
    check(rc);
    // host mirror (qureg.cpp:156-164 always allocates one); calloc'd pages are
    // only committed when first touched, so a mirror that is never synced
    // costs no host memory
    q.cpuAmps = (qcomp*)std::calloc((size_t)q.numAmps, sizeof(qcomp));
    q.gpuAmps = (qcomp*)d;
    q.cpuCommBuffer = nullptr;
    q.gpuCommBuffer = nullptr;
    registerQureg(q, prec);
    // initZeroState
    if (isDensMatr) {
        check(qa_statevec_initUniformState(d, prec, bits, 0, 0));
        double one[2] = {1, 0};
        check(qa_setAmps(d, prec, 0, one, 1));
    } else {
        check(qa_statevec_initUniformState(d, prec, bits, 0, 0));
        double one[2] = {1, 0};
        check(qa_setAmps(d, prec, 0, one, 1));
    }
    return q;
}

// ---- validation -------------------------------------------------------------

void validateQureg(const Qureg& q) {
    requireEnv();
    if (!q.gpuAmps || !isKnownQureg(q))
        fail("Invalid Qureg; its GPU memory is NULL or it was not created by createQureg (or was destroyed).");
}
void validateTarget(const Qureg& q, int t) {
    validateQureg(q);
    if (t < 0 || t >= q.numQubits)
        fail("Invalid target qubit (" + std::to_string(t) + "). Must be >= 0 and < " + std::to_string(q.numQubits) + ".");
}
void validateTargets(const Qureg& q, const int* t, int n) {
    validateQureg(q);
    if (n < 1 || n > q.numQubits) fail("Invalid number of target qubits (" + std::to_string(n) + ").");
    if (!t) fail("Received a NULL list of target qubits.");
    for (int i = 0; i < n; i++) {
        validateTarget(q, t[i]);
        for (int j = 0; j < i; j++)
            if (t[i] == t[j]) fail("The target qubits contained duplicates (" + std::to_string(t[i]) + ").");
    }
}
void validateCtrlsAndTargs(const Qureg& q, const int* c, int nc, const int* t, int nt) {
    validateTargets(q, t, nt);
    if (nc < 0 || nc + nt > q.numQubits) fail("Invalid number of control qubits (" + std::to_string(nc) + ").");
    if (nc > 0 && !c) fail("Received a NULL list of control qubits.");
    for (int i = 0; i < nc; i++) {
        if (c[i] < 0 || c[i] >= q.numQubits)
            fail("Invalid control qubit (" + std::to_string(c[i]) + "). Must be >= 0 and < " +
                 std::to_string(q.numQubits) + ".");
        for (int j = 0; j < i; j++)
            if (c[i] == c[j]) fail("The control qubits contained duplicates.");
        for (int j = 0; j < nt; j++)
            if (c[i] == t[j]) fail("A qubit (" + std::to_string(c[i]) + ") was both a control and a target.");
    }
}
void validateCtrlStates(const int* s, int n) {
    if (!s) return;
    for (int i = 0; i < n; i++)
        if (s[i] != 0 && s[i] != 1) fail("Invalid control-qubit state (" + std::to_string(s[i]) + "). Must be 0 or 1.");
}
void validateDensMatr(const Qureg& q) {
    validateQureg(q);
    if (!q.isDensityMatrix) fail("Expected a density matrix Qureg but received a statevector.");
}
void validateStateVec(const Qureg& q) {
    validateQureg(q);
    if (q.isDensityMatrix) fail("Expected a statevector Qureg but received a density matrix.");
}
void validateProb(double p, double maxProb, const char* what) {
    if (!(p >= 0)) fail(std::string("The ") + what + " probability is negative or NaN.");
    if (p > maxProb)
        fail(std::string("The ") + what + " probability (" + std::to_string(p) + ") exceeds the maximum of " +
             std::to_string(maxProb) + ".");
}
void validateSameDims(const Qureg& a, const Qureg& b) {
    if (a.numQubits != b.numQubits) fail("The Quregs have different numbers of qubits.");
}
void validateOutcome(int o) {
    if (o != 0 && o != 1) fail("Invalid measurement outcome (" + std::to_string(o) + "). Must be 0 or 1.");
}

static bool approx(std::complex<double> a, std::complex<double> b, double eps) { return std::norm(a - b) <= eps; }

bool isUnitary(const qcomp* m, qindex dim) {
    for (qindex r = 0; r < dim; r++)
        for (qindex c = 0; c < dim; c++) {
            qcomp e = 0;
            for (qindex i = 0; i < dim; i++) e += m[r * dim + i] * std::conj(m[c * dim + i]);
            if (!approx(e, qcomp(r == c, 0), g_eps)) return false;
        }
    return true;
}
bool isUnitaryDiag(const qcomp* d, qindex dim) {
    for (qindex i = 0; i < dim; i++)
        if (std::abs(std::abs(d[i]) - 1) > g_eps) return false;
    return true;
}

// ---- Pauli strings ---------------------------------------------------------

int pauliAt(PauliStr s, int q) {
    return q < 32 ? (int)((s.lowPaulis >> (2 * q)) & 3ull) : (int)((s.highPaulis >> (2 * (q - 32))) & 3ull);
}
void pauliMasks(PauliStr s, uint64_t& mx, uint64_t& my, uint64_t& mz) {
    mx = my = mz = 0;
    for (int q = 0; q < 64; q++) {
        int p = pauliAt(s, q);
        if (p == 1) mx |= 1ull << q;
        if (p == 2) my |= 1ull << q;
        if (p == 3) mz |= 1ull << q;
    }
}
bool pauliHasOddY(PauliStr s) {
    uint64_t x, y, z;
    pauliMasks(s, x, y, z);
    return __builtin_popcountll(y) & 1;
}
// paulis.cpp:239-262: move every Pauli up by `shift` qubits
PauliStr pauliShifted(PauliStr s, int shift) {
    PauliStr o{0, 0};
    for (int q = 0; q < 64; q++) {
        int p = pauliAt(s, q);
        if (!p) continue;
        int t = q + shift;
        if (t >= 64) fail("A Pauli string was shifted beyond 64 qubits.");
        if (t < 32) o.lowPaulis |= (PAULI_MASK_TYPE)p << (2 * t);
        else o.highPaulis |= (PAULI_MASK_TYPE)p << (2 * (t - 32));
    }
    return o;
}
int pauliMaxIndex(PauliStr s) {
    for (int q = 63; q >= 0; q--)
        if (pauliAt(s, q)) return q;
    return -1;
}

// ---- host RNG ----------------------------------------------------------------

int randomQubitOutcome(double probOfZero) {
    std::uniform_real_distribution<double> distrib(0, 1);
    double sample = distrib(g_gen);
    return sample > probOfZero;
}

qindex randomMultiQubitOutcome(const vector<double>& probs) {
    std::uniform_real_distribution<double> distrib(0, 1);
    double sample = distrib(g_gen);
    double cum = 0;
    for (size_t i = 0; i < probs.size(); i++) {
        cum += probs[i];
        if (sample < cum) return (qindex)i;
    }
    if (1 - cum <= g_eps) return (qindex)probs.size() - 1;
    fail("The outcome probabilities were not normalised.");
}

unsigned randomSharedSeed() {
    std::uniform_int_distribution<unsigned> distrib(0, std::numeric_limits<unsigned>::max());
    return distrib(g_gen);
}

static void setSeedsInternal(const std::vector<unsigned>& seeds) {
    g_seeds = seeds;
    std::seed_seq seq(seeds.begin(), seeds.end());
    g_gen.seed(seq);
}

}  // namespace qah

using namespace qah;

extern "C" {

void initCustomQuESTEnv(int useDistrib, int useGpuAccel, int useMultithread) {
    QAH_API({
        if (g_envInit) fail("The QuEST environment has already been initialised.");
        if (g_envFinalized) fail("The QuEST environment cannot be re-initialised after finalizeQuESTEnv().");
        if (useDistrib == 1) fail("This build has no distribution: useDistrib must be 0 or -1.");
        if (useGpuAccel == 0) fail("This build runs every Qureg on the GPU: useGpuAccel must be 1 or -1.");
        (void)useMultithread;
        const char* dev = std::getenv("QUEST_AMD_DEVICE");
        int rc = qa_init(dev ? std::atoi(dev) : 0);
        if (rc != QA_OK) fail(std::string("Could not bind a GPU: ") + qa_lastError());
        // gate fusion is on unless QUEST_AMD_FUSION=0; "tile,low,queue" sets its shape
        const char* fz = std::getenv("QUEST_AMD_FUSION");
        int on = 1, tile = 12, low = 4, queue = 512;
        if (fz && *fz) {
            int t = 0, l = 0, q = 0;
            if (std::sscanf(fz, "%d,%d,%d", &t, &l, &q) == 3) {
                tile = t;
                low = l;
                queue = q;
            } else {
                on = std::atoi(fz) != 0;
            }
        }
        check(qa_setFusion(on, tile, low, queue));
        std::random_device rd;
        setSeedsInternal({rd(), rd(), rd(), rd()});
        g_envInit = true;
    });
}

void initQuESTEnv(void) { initCustomQuESTEnv(0, -1, -1); }

void finalizeQuESTEnv(void) {
    QAH_API({
        requireEnv();
        releaseInitScratch();
        check(qa_finalize());
        g_envInit = false;
        g_envFinalized = true;
    });
}

void syncQuESTEnv(void) {
    QAH_API({
        requireEnv();
        check(qa_sync());
    });
}

int isQuESTEnvInit(void) { return g_envInit ? 1 : 0; }

QuESTEnv getQuESTEnv(void) {
    QAH_API({
        requireEnv();
        QuESTEnv e;
        e.isMultithreaded = 0;
        e.isGpuAccelerated = 1;
        e.isDistributed = 0;
        e.rank = 0;
        e.numNodes = 1;
        return e;
    });
}

void getEnvironmentString(char str[200]) {
    QAH_API({
        requireEnv();
        int cus = 0;
        char name[96] = "";
        check(qa_getDeviceInfo(&cus, name, (int)sizeof(name)));
        std::snprintf(str, 200, "CUDA=0 HIP=1 OMP=0 MPI=0 threads=1 ranks=1 GPU=%s CUs=%d precision=2", name, cus);
    });
}

void reportQuESTEnv(void) {
    QAH_API({
        requireEnv();
        char s[200];
        getEnvironmentString(s);
        size_t f = 0, t = 0;
        check(qa_getMemInfo(&f, &t));
        std::printf("QuESTEnv:\n  backend: MI355X HIP (single GPU)\n  %s\n  GPU memory: %zu free of %zu bytes\n", s, f, t);
    });
}

void setSeeds(unsigned* seeds, int numSeeds) {
    QAH_API({
        requireEnv();
        if (numSeeds <= 0) fail("Invalid number of random seeds; must be >= 1.");
        setSeedsInternal(std::vector<unsigned>(seeds, seeds + numSeeds));
    });
}

void setSeedsToDefault(void) {
    QAH_API({
        requireEnv();
        std::random_device rd;
        setSeedsInternal({rd(), rd(), rd(), rd()});
    });
}

int getNumSeeds(void) {
    QAH_API({
        requireEnv();
        return (int)g_seeds.size();
    });
}

void getSeeds(unsigned* seeds) {
    QAH_API({
        requireEnv();
        for (size_t i = 0; i < g_seeds.size(); i++) seeds[i] = g_seeds[i];
    });
}

void setInputErrorHandler(void (*callback)(const char* func, const char* msg)) { g_handler = callback; }
void setValidationOn(void) { g_validate = true; }
void setValidationOff(void) { g_validate = false; }
void setValidationEpsilonToDefault(void) { g_eps = DEFAULT_EPS; }
void setValidationEpsilon(qreal eps) {
    QAH_API({
        if (eps < 0) fail("Invalid validation epsilon; must be >= 0.");
        g_eps = eps;
    });
}
qreal getValidationEpsilon(void) { return g_eps; }
qindex getGpuCacheSize(void) { return 0; }
void clearGpuCache(void) {}

void setQuESTFusion(int enabled, int tileQubits, int lowQubits, int maxQueue) {
    QAH_API({
        requireEnv();
        check(qa_setFusion(enabled, tileQubits, lowQubits, maxQueue));
    });
}

void getQuESTFusion(int* enabled, int* tileQubits, int* lowQubits, int* maxQueue) {
    QAH_API({
        requireEnv();
        check(qa_getFusion(enabled, tileQubits, lowQubits, maxQueue));
    });
}

void* getQuESTStream(void) { return qa_getStream(); }

}  // extern "C"

void setSeeds(std::vector<unsigned> seeds) { setSeeds(seeds.data(), (int)seeds.size()); }
std::vector<unsigned> getSeeds() { return qah::g_seeds; }
