// Internal interface of the host layer: error handling in QuEST's style,
// validation helpers, the qureg registry, and the thin "accelerator" that
// forwards to the C ABI of the kernel layer (include/quest_amd_gpu.h).
//
// The host layer mirrors QuEST's own split (quest/src/api -> core/localiser
// -> core/accelerator -> gpu_*): API functions validate their inputs
// (core/validation.cpp), route statevector and density-matrix cases
// (core/localiser.cpp, single-node branch only) and call qa_* entry points,
// which play the role of gpu_subroutines.cpp.
#pragma once

#include <complex>
#include <cstdint>
#include <string>
#include <vector>

#include "quest.h"
#include "quest_amd_ext.h"
#include "quest_amd_gpu.h"

namespace qah {

using std::vector;

// ---- errors ---------------------------------------------------------------
// A failed validation throws ApiError; every exported function catches it at
// its boundary and calls the input-error handler (default: report and exit,
// as validation.cpp:1073-1100).  If a user handler returns, the function
// returns without effect.
struct ApiError {
    std::string msg;
};
[[noreturn]] void fail(const std::string& msg);
void handleError(const char* func, const std::string& msg);

template <class F>
auto guard(const char* func, F&& f) -> decltype(f()) {
    using R = decltype(f());
    try {
        return f();
    } catch (ApiError& e) {
        handleError(func, e.msg);
        if constexpr (!std::is_void_v<R>) return R{};
    }
}

// kernel-layer return code -> error
void check(int rc);

// ---- environment / validation settings -------------------------------------
bool envInit();
void requireEnv();
bool validating();
double epsilon();
// numerical tolerance for results of one qureg: the reference's default for
// its amplitude precision (precision.h:131-141: 1e-5 for FP32) unless the
// user set a looser one
double epsilonFor(const Qureg& q);

// ---- qureg bookkeeping ------------------------------------------------------
int precOf(const Qureg& q);  // QA_FP64 / QA_FP32
void registerQureg(const Qureg& q, int prec);
void unregisterQureg(const Qureg& q);
bool isKnownQureg(const Qureg& q);
inline void* dev(const Qureg& q) { return (void*)q.gpuAmps; }
inline int idxBits(const Qureg& q) { return q.isDensityMatrix ? 2 * q.numQubits : q.numQubits; }
Qureg makeQureg(int numQubits, int isDensMatr, int prec);

// ---- validation (subset of core/validation.cpp) -----------------------------
void validateQureg(const Qureg& q);
void validateTarget(const Qureg& q, int t);
void validateTargets(const Qureg& q, const int* t, int n);
void validateCtrlsAndTargs(const Qureg& q, const int* c, int nc, const int* t, int nt);
void validateCtrlStates(const int* s, int n);
void validateDensMatr(const Qureg& q);
void validateStateVec(const Qureg& q);
void validateProb(double p, double maxProb, const char* what);
void validateSameDims(const Qureg& a, const Qureg& b);
void validateOutcome(int o);

// unitarity (utilities.cpp:418-466): |(M M^dagger)_rc - delta_rc|^2 <= eps
bool isUnitary(const qcomp* rowMajor, qindex dim);
bool isUnitaryDiag(const qcomp* diag, qindex dim);

// ---- misc -------------------------------------------------------------------
inline vector<int> vec(const int* p, int n) { return (p && n > 0) ? vector<int>(p, p + n) : vector<int>(); }
inline vector<int> braOf(const vector<int>& v, const Qureg& q) {
    vector<int> o(v);
    for (auto& x : o) x += q.numQubits;
    return o;
}
inline const double* dbl(const qcomp* p) { return reinterpret_cast<const double*>(p); }

// PauliStr helpers (paulis.cpp:98-331)
int pauliAt(PauliStr s, int q);
void pauliMasks(PauliStr s, uint64_t& mx, uint64_t& my, uint64_t& mz);
bool pauliHasOddY(PauliStr s);
PauliStr pauliShifted(PauliStr s, int shift);
int pauliMaxIndex(PauliStr s);

// host RNG (randomiser.cpp): global mt19937_64 seeded through seed_seq
int randomQubitOutcome(double probOfZero);
qindex randomMultiQubitOutcome(const vector<double>& probs);
unsigned randomSharedSeed();
// frees the density-matrix initialisers' statevector scratch (qureg.cpp)
void releaseInitScratch();

}  // namespace qah

// wrap an exported function's body: QAH_API({ ... });
#define QAH_API(...) return qah::guard(__func__, [&]() __VA_ARGS__)
