// QuEST's accelerator seam, routed to the MI355X kernel layer.
//
// This translation unit is what a QuEST v4.1.0 maintainer compiles in place
// of quest/src/gpu/gpu_config.cpp and quest/src/gpu/gpu_subroutines.cpp: it
// defines every function that quest/src/gpu/gpu_config.hpp:41-120 and
// quest/src/gpu/gpu_subroutines.hpp:24-196 declare, with the template
// instantiations that core/accelerator.cpp binds to (the INSTANTIATE_*
// macros of core/accelerator.hpp:30-135), each forwarding to one extern "C"
// entry point of include/quest_amd_gpu.h.  It is built against QuEST's own
// headers (the include path is the root of a QuEST checkout) and linked with
// QuEST's other sources and libquest_amd_gpu.so; nothing of this file is in
// libquest_amd_gpu.so itself.
//
// The seam covers one GPU: QuEST's communication routines (comm buffer
// packing, the _subB.. variants that consult a distributed Qureg's buffer)
// are called only when a Qureg is distributed, which an environment bound
// to one GPU never creates; they report an internal error here.
//
// Where QuEST splits one operation into two passes that its localiser
// always issues back to back on a single node (two-qubit depolarising,
// localiser.cpp:1551-1552: _subA scales, _subB mixes), the kernel layer
// applies both in one pass, so _subA does the whole operation and _subB
// has nothing left to do.
//
// Amplitudes are complex128 (QuEST built with FLOAT_PRECISION=2); a
// density matrix's flat amplitudes are statevector amplitudes over
// logNumAmpsPerNode index bits, which is how the statevector routines
// below are applied to both kinds of Qureg.

#include "quest/include/types.h"
#include "quest/include/qureg.h"
#include "quest/include/matrices.h"
#include "quest/include/channels.h"
#include "quest/include/paulis.h"

#include "quest/src/core/accelerator.hpp"
#include "quest/src/core/errors.hpp"
#include "quest/src/core/randomiser.hpp"
#include "quest/src/gpu/gpu_config.hpp"
#include "quest/src/gpu/gpu_subroutines.hpp"

#include "quest_amd_gpu.h"

#include <cstdint>
#include <cstdlib>
#include <cstdio>
#include <vector>

using std::vector;

static_assert(sizeof(qreal) == 8, "the seam forwards complex128 amplitudes (FLOAT_PRECISION=2)");
static_assert(sizeof(qcomp) == 16, "qcomp is {re, im} of two doubles");

namespace {

constexpr int PREC = QA_FP64;

// a failing kernel-layer call ends the program the way a failing CUDA call
// does in QuEST (gpu_config.hpp:28-31: CUDA_CHECK -> error_cudaCallFailed)
void check(int rc, const char* call, const char* caller, int line) {
    if (rc != QA_OK) error_cudaCallFailed(qa_lastError(), call, caller, __FILE__, line);
}
#define QA_CALL(expr) check((expr), #expr, __func__, __LINE__)

[[noreturn]] void notOnOneGpu(const char* caller) {
    error_cudaCallFailed("a distributed routine was called, but this GPU backend serves one GPU", "(none)", caller,
                         __FILE__, __LINE__);
    std::abort();
}

void* amps(const Qureg& q) { return (void*)q.gpuAmps; }
int bits(const Qureg& q) { return (int)q.logNumAmpsPerNode; }  // index bits of the flat amplitudes
const double* dbl(const qcomp* p) { return reinterpret_cast<const double*>(p); }

uint64_t maskOf(const vector<int>& qs) {
    uint64_t m = 0;
    for (int q : qs) m |= 1ull << q;
    return m;
}

// the X, Y and Z bit masks of a Pauli string (2 bits per qubit, I=0 X=1
// Y=2 Z=3, qubits below 32 in lowPaulis: paulis.cpp:306-331)
void pauliMasks(PauliStr s, uint64_t& x, uint64_t& y, uint64_t& z) {
    x = y = z = 0;
    for (int q = 0; q < 64; q++) {
        const PAULI_MASK_TYPE word = q < 32 ? s.lowPaulis : s.highPaulis;
        const int code = (int)((word >> (2 * (q % 32))) & 3u);
        if (code == 1) x |= 1ull << q;
        if (code == 2) y |= 1ull << q;
        if (code == 3) z |= 1ull << q;
    }
}

void sumMasks(const PauliStrSum& sum, vector<uint64_t>& x, vector<uint64_t>& y, vector<uint64_t>& z) {
    x.resize(sum.numTerms);
    y.resize(sum.numTerms);
    z.resize(sum.numTerms);
    for (qindex t = 0; t < sum.numTerms; t++) pauliMasks(sum.strings[t], x[t], y[t], z[t]);
}

qcomp toComp(const double* v) { return qcomp(v[0], v[1]); }

}  // namespace


/*
 * HARDWARE AVAILABILITY (gpu_config.hpp:41-65)
 */

bool gpu_isGpuCompiled() { return true; }
bool gpu_isCuQuantumCompiled() { return false; }
bool gpu_isGpuAvailable() { return qa_isGpuAvailable() != 0; }
bool gpu_isDirectGpuCommPossible() { return false; }

int gpu_getNumberOfLocalGpus() { return qa_isGpuAvailable() ? 1 : 0; }

// QuEST reports the CUDA compute capability; an AMD device has none, and
// the value is only printed (environment.cpp reportQuESTEnv)
int gpu_getComputeCapability() { return 0; }

size_t gpu_getCurrentAvailableMemoryInBytes() {
    size_t freeB = 0, totalB = 0;
    QA_CALL(qa_getMemInfo(&freeB, &totalB));
    return freeB;
}

size_t gpu_getTotalMemoryInBytes() {
    size_t freeB = 0, totalB = 0;
    QA_CALL(qa_getMemInfo(&freeB, &totalB));
    return totalB;
}

bool gpu_doesGpuSupportMemPools() { return false; }

// 256 compute units x 2048 resident threads each (32 wave64s)
qindex gpu_getMaxNumConcurrentThreads() {
    int cus = 0;
    char name[64];
    QA_CALL(qa_getDeviceInfo(&cus, name, (int)sizeof name));
    return (qindex)cus * 2048;
}


/*
 * ENVIRONMENT MANAGEMENT (gpu_config.hpp:72-80)
 */

void gpu_bindLocalGPUsToNodes() {
    const char* dev = std::getenv("QUEST_AMD_DEVICE");
    QA_CALL(qa_init(dev ? std::atoi(dev) : 0));
    // gate fusion, as the quest.h layer of this backend enables it; the
    // variable QUEST_AMD_FUSION=0 turns it off, "tile,low,queue" shapes it
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
    QA_CALL(qa_setFusion(on, tile, low, queue));
}

bool gpu_areAnyNodesBoundToSameGpu() { return false; }

void gpu_sync() { QA_CALL(qa_sync()); }

void gpu_initCuQuantum() {}
void gpu_finalizeCuQuantum() {}


/*
 * MEMORY MANAGEMENT (gpu_config.hpp:87-112)
 */

qcomp* gpu_allocArray(qindex numLocalAmps) {
    void* p = nullptr;
    // an allocation that does not fit returns NULL, which QuEST validates
    if (qa_malloc(&p, (size_t)numLocalAmps * sizeof(qcomp)) != QA_OK) return nullptr;
    return (qcomp*)p;
}

void gpu_deallocArray(qcomp* amps) { QA_CALL(qa_free(amps)); }

void gpu_copyArray(qcomp* dest, qcomp* src, qindex dim) {
    QA_CALL(qa_memcpyDtoD(dest, src, (size_t)dim * sizeof(qcomp)));
    QA_CALL(qa_sync());
}

void gpu_copyCpuToGpu(qcomp* cpuArr, qcomp* gpuArr, qindex numElems) {
    QA_CALL(qa_memcpyHtoD(gpuArr, cpuArr, (size_t)numElems * sizeof(qcomp)));
}
void gpu_copyGpuToCpu(qcomp* gpuArr, qcomp* cpuArr, qindex numElems) {
    QA_CALL(qa_memcpyDtoH(cpuArr, gpuArr, (size_t)numElems * sizeof(qcomp)));
}

void gpu_copyGpuToCpu(Qureg, qcomp* gpuArr, qcomp* cpuArr, qindex numElems) {
    gpu_copyGpuToCpu(gpuArr, cpuArr, numElems);
}
void gpu_copyGpuToCpu(Qureg qureg) { gpu_copyGpuToCpu(qureg.gpuAmps, qureg.cpuAmps, qureg.numAmpsPerNode); }

void gpu_copyCpuToGpu(Qureg, qcomp* cpuArr, qcomp* gpuArr, qindex numElems) {
    gpu_copyCpuToGpu(cpuArr, gpuArr, numElems);
}
void gpu_copyCpuToGpu(Qureg qureg) { gpu_copyCpuToGpu(qureg.cpuAmps, qureg.gpuAmps, qureg.numAmpsPerNode); }

void gpu_copyCpuToGpu(CompMatr m) { gpu_copyCpuToGpu(m.cpuElemsFlat, m.gpuElemsFlat, m.numRows * m.numRows); }
void gpu_copyGpuToCpu(CompMatr m) { gpu_copyGpuToCpu(m.gpuElemsFlat, m.cpuElemsFlat, m.numRows * m.numRows); }

void gpu_copyCpuToGpu(DiagMatr m) { gpu_copyCpuToGpu(m.cpuElems, m.gpuElems, m.numElems); }
void gpu_copyGpuToCpu(DiagMatr m) { gpu_copyGpuToCpu(m.gpuElems, m.cpuElems, m.numElems); }

void gpu_copyCpuToGpu(SuperOp op) { gpu_copyCpuToGpu(op.cpuElemsFlat, op.gpuElemsFlat, op.numRows * op.numRows); }
void gpu_copyGpuToCpu(SuperOp op) { gpu_copyGpuToCpu(op.gpuElemsFlat, op.cpuElemsFlat, op.numRows * op.numRows); }

void gpu_copyCpuToGpu(FullStateDiagMatr m) { gpu_copyCpuToGpu(m.cpuElems, m.gpuElems, m.numElemsPerNode); }


/*
 * CACHE MANAGEMENT (gpu_config.hpp:115-120)
 *
 * QuEST keeps a device cache for its any-target dense-matrix kernel; the
 * kernel layer stages such matrices through LDS and owns its scratch, so
 * there is no cache to hand out or clear.
 */

qcomp* gpu_getCacheOfSize(qindex, qindex) { return nullptr; }
void gpu_clearCache() {}
size_t gpu_getCacheMemoryInBytes() { return 0; }


/*
 * GETTERS AND SETTERS (gpu_subroutines.hpp:24-36)
 */

qcomp gpu_statevec_getAmp_sub(Qureg qureg, qindex ind) {
    double v[2];
    QA_CALL(qa_getAmps(v, amps(qureg), PREC, ind, 1));
    return toComp(v);
}

void gpu_densmatr_setAmpsToPauliStrSum_sub(Qureg qureg, PauliStrSum sum) {
    vector<uint64_t> x, y, z;
    sumMasks(sum, x, y, z);
    QA_CALL(qa_densmatr_setAmpsToPauliStrSum(amps(qureg), PREC, qureg.numQubits, sum.numTerms, x.data(), y.data(),
                                             z.data(), dbl(sum.coeffs)));
}

void gpu_fullstatediagmatr_setElemsToPauliStrSum(FullStateDiagMatr out, PauliStrSum in) {
    vector<uint64_t> x, y, z;
    sumMasks(in, x, y, z);
    QA_CALL(qa_fullstatediagmatr_setElemsToPauliStrSum(out.gpuElems, out.numQubits, in.numTerms, z.data(),
                                                       dbl(in.coeffs)));
}


/*
 * COMMUNICATION BUFFER PACKING (gpu_subroutines.hpp:39-44): distributed only
 */

template <int NumQubits>
qindex gpu_statevec_packAmpsIntoBuffer(Qureg, vector<int>, vector<int>) {
    notOnOneGpu(__func__);
}

qindex gpu_statevec_packPairSummedAmpsIntoBuffer(Qureg, int, int, int, int) { notOnOneGpu(__func__); }

INSTANTIATE_FUNC_OPTIMISED_FOR_NUM_TARGS(qindex, gpu_statevec_packAmpsIntoBuffer, (Qureg, vector<int>, vector<int>))


/*
 * SWAPS (gpu_subroutines.hpp:47-53)
 */

template <int NumCtrls>
void gpu_statevec_anyCtrlSwap_subA(Qureg qureg, vector<int> ctrls, vector<int> ctrlStates, int targ1, int targ2) {
    QA_CALL(qa_statevec_anyCtrlSwap(amps(qureg), PREC, bits(qureg), ctrls.data(), ctrlStates.data(), (int)ctrls.size(),
                                    targ1, targ2));
}
template <int NumCtrls>
void gpu_statevec_anyCtrlSwap_subB(Qureg, vector<int>, vector<int>) {
    notOnOneGpu(__func__);
}
template <int NumCtrls>
void gpu_statevec_anyCtrlSwap_subC(Qureg, vector<int>, vector<int>, int, int) {
    notOnOneGpu(__func__);
}

INSTANTIATE_FUNC_OPTIMISED_FOR_NUM_CTRLS(void, gpu_statevec_anyCtrlSwap_subA,
                                         (Qureg qureg, vector<int> ctrls, vector<int> ctrlStates, int targ1, int targ2))
INSTANTIATE_FUNC_OPTIMISED_FOR_NUM_CTRLS(void, gpu_statevec_anyCtrlSwap_subB,
                                         (Qureg qureg, vector<int> ctrls, vector<int> ctrlStates))
INSTANTIATE_FUNC_OPTIMISED_FOR_NUM_CTRLS(void, gpu_statevec_anyCtrlSwap_subC,
                                         (Qureg qureg, vector<int> ctrls, vector<int> ctrlStates, int targ, int targState))


/*
 * DENSE MATRIX (gpu_subroutines.hpp:56-66)
 */

template <int NumCtrls>
void gpu_statevec_anyCtrlOneTargDenseMatr_subA(Qureg qureg, vector<int> ctrls, vector<int> ctrlStates, int targ,
                                               CompMatr1 matr) {
    QA_CALL(qa_statevec_anyCtrlOneTargDenseMatr(amps(qureg), PREC, bits(qureg), ctrls.data(), ctrlStates.data(),
                                                (int)ctrls.size(), targ, dbl(&matr.elems[0][0])));
}
template <int NumCtrls>
void gpu_statevec_anyCtrlOneTargDenseMatr_subB(Qureg, vector<int>, vector<int>, qcomp, qcomp) {
    notOnOneGpu(__func__);
}

template <int NumCtrls>
void gpu_statevec_anyCtrlTwoTargDenseMatr_sub(Qureg qureg, vector<int> ctrls, vector<int> ctrlStates, int targ1,
                                              int targ2, CompMatr2 matr) {
    QA_CALL(qa_statevec_anyCtrlTwoTargDenseMatr(amps(qureg), PREC, bits(qureg), ctrls.data(), ctrlStates.data(),
                                                (int)ctrls.size(), targ1, targ2, dbl(&matr.elems[0][0])));
}

template <int NumCtrls, int NumTargs, bool ApplyConj>
void gpu_statevec_anyCtrlAnyTargDenseMatr_sub(Qureg qureg, vector<int> ctrls, vector<int> ctrlStates,
                                              vector<int> targs, CompMatr matr) {
    QA_CALL(qa_statevec_anyCtrlAnyTargDenseMatr(amps(qureg), PREC, bits(qureg), ctrls.data(), ctrlStates.data(),
                                                (int)ctrls.size(), targs.data(), (int)targs.size(), matr.gpuElemsFlat,
                                                ApplyConj ? 1 : 0));
}

INSTANTIATE_FUNC_OPTIMISED_FOR_NUM_CTRLS(void, gpu_statevec_anyCtrlOneTargDenseMatr_subA,
                                         (Qureg, vector<int>, vector<int>, int, CompMatr1))
INSTANTIATE_FUNC_OPTIMISED_FOR_NUM_CTRLS(void, gpu_statevec_anyCtrlOneTargDenseMatr_subB,
                                         (Qureg, vector<int>, vector<int>, qcomp, qcomp))
INSTANTIATE_FUNC_OPTIMISED_FOR_NUM_CTRLS(void, gpu_statevec_anyCtrlTwoTargDenseMatr_sub,
                                         (Qureg, vector<int>, vector<int>, int, int, CompMatr2))
INSTANTIATE_CONJUGABLE_FUNC_OPTIMISED_FOR_NUM_CTRLS_AND_TARGS(void, gpu_statevec_anyCtrlAnyTargDenseMatr_sub,
                                                              (Qureg, vector<int>, vector<int>, vector<int>, CompMatr))


/*
 * DIAGONAL MATRIX (gpu_subroutines.hpp:69-81)
 */

template <int NumCtrls>
void gpu_statevec_anyCtrlOneTargDiagMatr_sub(Qureg qureg, vector<int> ctrls, vector<int> ctrlStates, int targ,
                                             DiagMatr1 matr) {
    QA_CALL(qa_statevec_anyCtrlOneTargDiagMatr(amps(qureg), PREC, bits(qureg), ctrls.data(), ctrlStates.data(),
                                               (int)ctrls.size(), targ, dbl(matr.elems)));
}

template <int NumCtrls>
void gpu_statevec_anyCtrlTwoTargDiagMatr_sub(Qureg qureg, vector<int> ctrls, vector<int> ctrlStates, int targ1,
                                             int targ2, DiagMatr2 matr) {
    QA_CALL(qa_statevec_anyCtrlTwoTargDiagMatr(amps(qureg), PREC, bits(qureg), ctrls.data(), ctrlStates.data(),
                                               (int)ctrls.size(), targ1, targ2, dbl(matr.elems)));
}

template <int NumCtrls, int NumTargs, bool ApplyConj, bool HasPower>
void gpu_statevec_anyCtrlAnyTargDiagMatr_sub(Qureg qureg, vector<int> ctrls, vector<int> ctrlStates,
                                             vector<int> targs, DiagMatr matr, qcomp exponent) {
    QA_CALL(qa_statevec_anyCtrlAnyTargDiagMatr(amps(qureg), PREC, bits(qureg), ctrls.data(), ctrlStates.data(),
                                               (int)ctrls.size(), targs.data(), (int)targs.size(), matr.gpuElems,
                                               ApplyConj ? 1 : 0, HasPower ? 1 : 0, std::real(exponent),
                                               std::imag(exponent)));
}

template <bool HasPower>
void gpu_statevec_allTargDiagMatr_sub(Qureg qureg, FullStateDiagMatr matr, qcomp exponent) {
    QA_CALL(qa_statevec_allTargDiagMatr(amps(qureg), PREC, bits(qureg), matr.gpuElems, HasPower ? 1 : 0,
                                        std::real(exponent), std::imag(exponent)));
}

template <bool HasPower, bool MultiplyOnly>
void gpu_densmatr_allTargDiagMatr_sub(Qureg qureg, FullStateDiagMatr matr, qcomp exponent) {
    QA_CALL(qa_densmatr_allTargDiagMatr(amps(qureg), PREC, qureg.numQubits, matr.gpuElems, HasPower ? 1 : 0,
                                        std::real(exponent), std::imag(exponent), MultiplyOnly ? 1 : 0));
}

INSTANTIATE_FUNC_OPTIMISED_FOR_NUM_CTRLS(void, gpu_statevec_anyCtrlOneTargDiagMatr_sub,
                                         (Qureg, vector<int>, vector<int>, int, DiagMatr1))
INSTANTIATE_FUNC_OPTIMISED_FOR_NUM_CTRLS(void, gpu_statevec_anyCtrlTwoTargDiagMatr_sub,
                                         (Qureg, vector<int>, vector<int>, int, int, DiagMatr2))
INSTANTIATE_EXPONENTIABLE_CONJUGABLE_FUNC_OPTIMISED_FOR_NUM_CTRLS_AND_TARGS(
    void, gpu_statevec_anyCtrlAnyTargDiagMatr_sub, (Qureg, vector<int>, vector<int>, vector<int>, DiagMatr, qcomp))

template void gpu_statevec_allTargDiagMatr_sub<true>(Qureg, FullStateDiagMatr, qcomp);
template void gpu_statevec_allTargDiagMatr_sub<false>(Qureg, FullStateDiagMatr, qcomp);
template void gpu_densmatr_allTargDiagMatr_sub<true, true>(Qureg, FullStateDiagMatr, qcomp);
template void gpu_densmatr_allTargDiagMatr_sub<true, false>(Qureg, FullStateDiagMatr, qcomp);
template void gpu_densmatr_allTargDiagMatr_sub<false, true>(Qureg, FullStateDiagMatr, qcomp);
template void gpu_densmatr_allTargDiagMatr_sub<false, false>(Qureg, FullStateDiagMatr, qcomp);


/*
 * PAULI TENSOR AND GADGET (gpu_subroutines.hpp:84-91)
 */

template <int NumCtrls, int NumTargs>
void gpu_statevector_anyCtrlPauliTensorOrGadget_subA(Qureg qureg, vector<int> ctrls, vector<int> ctrlStates,
                                                      vector<int> x, vector<int> y, vector<int> z, qcomp ampFac,
                                                      qcomp pairAmpFac) {
    QA_CALL(qa_statevec_anyCtrlPauliTensorOrGadget(amps(qureg), PREC, bits(qureg), ctrls.data(), ctrlStates.data(),
                                                   (int)ctrls.size(), maskOf(x), maskOf(y), maskOf(z),
                                                   std::real(ampFac), std::imag(ampFac), std::real(pairAmpFac),
                                                   std::imag(pairAmpFac)));
}

template <int NumCtrls>
void gpu_statevector_anyCtrlPauliTensorOrGadget_subB(Qureg, vector<int>, vector<int>, vector<int>, vector<int>,
                                                      vector<int>, qcomp, qcomp, qindex) {
    notOnOneGpu(__func__);
}

template <int NumCtrls>
void gpu_statevector_anyCtrlAnyTargZOrPhaseGadget_sub(Qureg qureg, vector<int> ctrls, vector<int> ctrlStates,
                                                       vector<int> targs, qcomp fac0, qcomp fac1) {
    QA_CALL(qa_statevec_anyCtrlAnyTargZOrPhaseGadget(amps(qureg), PREC, bits(qureg), ctrls.data(), ctrlStates.data(),
                                                     (int)ctrls.size(), maskOf(targs), std::real(fac0), std::imag(fac0),
                                                     std::real(fac1), std::imag(fac1)));
}

INSTANTIATE_FUNC_OPTIMISED_FOR_NUM_CTRLS_AND_TARGS(void, gpu_statevector_anyCtrlPauliTensorOrGadget_subA,
                                                   (Qureg, vector<int>, vector<int>, vector<int>, vector<int>,
                                                    vector<int>, qcomp, qcomp))
INSTANTIATE_FUNC_OPTIMISED_FOR_NUM_CTRLS(void, gpu_statevector_anyCtrlPauliTensorOrGadget_subB,
                                         (Qureg, vector<int>, vector<int>, vector<int>, vector<int>, vector<int>,
                                          qcomp, qcomp, qindex))
INSTANTIATE_FUNC_OPTIMISED_FOR_NUM_CTRLS(void, gpu_statevector_anyCtrlAnyTargZOrPhaseGadget_sub,
                                         (Qureg, vector<int>, vector<int>, vector<int>, qcomp, qcomp))


/*
 * QUREG COMBINATION (gpu_subroutines.hpp:94-100)
 */

void gpu_statevec_setQuregToSuperposition_sub(qcomp facOut, Qureg outQureg, qcomp fac1, Qureg inQureg1, qcomp fac2,
                                              Qureg inQureg2) {
    QA_CALL(qa_statevec_setQuregToSuperposition(PREC, bits(outQureg), std::real(facOut), std::imag(facOut),
                                                amps(outQureg), std::real(fac1), std::imag(fac1), amps(inQureg1),
                                                std::real(fac2), std::imag(fac2), amps(inQureg2)));
}

void gpu_densmatr_mixQureg_subA(qreal outProb, Qureg outQureg, qreal inProb, Qureg inDensMatr) {
    QA_CALL(qa_densmatr_mixQureg(outProb, amps(outQureg), inProb, amps(inDensMatr), PREC, outQureg.numQubits));
}
void gpu_densmatr_mixQureg_subB(qreal outProb, Qureg outQureg, qreal inProb, Qureg inStateVec) {
    QA_CALL(qa_densmatr_mixQuregWithStatevec(outProb, amps(outQureg), inProb, amps(inStateVec), PREC,
                                             outQureg.numQubits));
}
void gpu_densmatr_mixQureg_subC(qreal, Qureg, qreal) { notOnOneGpu(__func__); }


/*
 * DECOHERENCE (gpu_subroutines.hpp:103-128); on one GPU every bra qubit is
 * local, so only the _subA forms are reached
 */

void gpu_densmatr_oneQubitDephasing_subA(Qureg qureg, int qubit, qreal prob) {
    QA_CALL(qa_densmatr_oneQubitDephasing(amps(qureg), PREC, qureg.numQubits, qubit, prob));
}
void gpu_densmatr_oneQubitDephasing_subB(Qureg, int, qreal) { notOnOneGpu(__func__); }

void gpu_densmatr_twoQubitDephasing_subA(Qureg qureg, int qubitA, int qubitB, qreal prob) {
    QA_CALL(qa_densmatr_twoQubitDephasing(amps(qureg), PREC, qureg.numQubits, qubitA, qubitB, prob));
}
void gpu_densmatr_twoQubitDephasing_subB(Qureg, int, int, qreal) { notOnOneGpu(__func__); }

void gpu_densmatr_oneQubitDepolarising_subA(Qureg qureg, int qubit, qreal prob) {
    QA_CALL(qa_densmatr_oneQubitDepolarising(amps(qureg), PREC, qureg.numQubits, qubit, prob));
}
void gpu_densmatr_oneQubitDepolarising_subB(Qureg, int, qreal) { notOnOneGpu(__func__); }

// _subA scales and _subB mixes (cpu_subroutines.cpp:1264-1326); the kernel
// layer does both in one pass, in that order, when _subA is called
void gpu_densmatr_twoQubitDepolarising_subA(Qureg qureg, int qubit1, int qubit2, qreal prob) {
    QA_CALL(qa_densmatr_twoQubitDepolarising(amps(qureg), PREC, qureg.numQubits, qubit1, qubit2, prob));
}
void gpu_densmatr_twoQubitDepolarising_subB(Qureg, int, int, qreal) {}
void gpu_densmatr_twoQubitDepolarising_subC(Qureg, int, int, qreal) { notOnOneGpu(__func__); }
void gpu_densmatr_twoQubitDepolarising_subD(Qureg, int, int, qreal) { notOnOneGpu(__func__); }
void gpu_densmatr_twoQubitDepolarising_subE(Qureg, int, int, qreal) { notOnOneGpu(__func__); }
void gpu_densmatr_twoQubitDepolarising_subF(Qureg, int, int, qreal) { notOnOneGpu(__func__); }

void gpu_densmatr_oneQubitPauliChannel_subA(Qureg qureg, int ketQubit, qreal pI, qreal pX, qreal pY, qreal pZ) {
    QA_CALL(qa_densmatr_oneQubitPauliChannel(amps(qureg), PREC, qureg.numQubits, ketQubit, pI, pX, pY, pZ));
}
void gpu_densmatr_oneQubitPauliChannel_subB(Qureg, int, qreal, qreal, qreal, qreal) { notOnOneGpu(__func__); }

void gpu_densmatr_oneQubitDamping_subA(Qureg qureg, int qubit, qreal prob) {
    QA_CALL(qa_densmatr_oneQubitDamping(amps(qureg), PREC, qureg.numQubits, qubit, prob));
}
void gpu_densmatr_oneQubitDamping_subB(Qureg, int, qreal) { notOnOneGpu(__func__); }
void gpu_densmatr_oneQubitDamping_subC(Qureg, int, qreal) { notOnOneGpu(__func__); }
void gpu_densmatr_oneQubitDamping_subD(Qureg, int, qreal) { notOnOneGpu(__func__); }


/*
 * PARTIAL TRACE (gpu_subroutines.hpp:131-133); pairTargs are the bra
 * qubits of targs, which the kernel layer derives itself
 */

template <int NumTargs>
void gpu_densmatr_partialTrace_sub(Qureg inQureg, Qureg outQureg, vector<int> targs, vector<int>) {
    QA_CALL(qa_densmatr_partialTrace(amps(inQureg), PREC, inQureg.numQubits, amps(outQureg), targs.data(),
                                     (int)targs.size()));
}

INSTANTIATE_FUNC_OPTIMISED_FOR_NUM_TARGS(void, gpu_densmatr_partialTrace_sub, (Qureg, Qureg, vector<int>, vector<int>))


/*
 * PROBABILITIES (gpu_subroutines.hpp:136-146)
 */

qreal gpu_statevec_calcTotalProb_sub(Qureg qureg) {
    double v = 0;
    QA_CALL(qa_statevec_calcTotalProb(amps(qureg), PREC, bits(qureg), &v));
    return v;
}
qreal gpu_densmatr_calcTotalProb_sub(Qureg qureg) {
    double v = 0;
    QA_CALL(qa_densmatr_calcTotalProb(amps(qureg), PREC, qureg.numQubits, &v));
    return v;
}

template <int NumQubits>
qreal gpu_statevec_calcProbOfMultiQubitOutcome_sub(Qureg qureg, vector<int> qubits, vector<int> outcomes) {
    double v = 0;
    QA_CALL(qa_statevec_calcProbOfMultiQubitOutcome(amps(qureg), PREC, bits(qureg), qubits.data(), outcomes.data(),
                                                    (int)qubits.size(), &v));
    return v;
}
template <int NumQubits>
qreal gpu_densmatr_calcProbOfMultiQubitOutcome_sub(Qureg qureg, vector<int> qubits, vector<int> outcomes) {
    double v = 0;
    QA_CALL(qa_densmatr_calcProbOfMultiQubitOutcome(amps(qureg), PREC, qureg.numQubits, qubits.data(),
                                                    outcomes.data(), (int)qubits.size(), &v));
    return v;
}

template <int NumQubits>
void gpu_statevec_calcProbsOfAllMultiQubitOutcomes_sub(qreal* outProbs, Qureg qureg, vector<int> qubits) {
    QA_CALL(qa_statevec_calcProbsOfAllMultiQubitOutcomes(outProbs, amps(qureg), PREC, bits(qureg), qubits.data(),
                                                         (int)qubits.size()));
}
template <int NumQubits>
void gpu_densmatr_calcProbsOfAllMultiQubitOutcomes_sub(qreal* outProbs, Qureg qureg, vector<int> qubits) {
    QA_CALL(qa_densmatr_calcProbsOfAllMultiQubitOutcomes(outProbs, amps(qureg), PREC, qureg.numQubits,
                                                         qubits.data(), (int)qubits.size()));
}

INSTANTIATE_FUNC_OPTIMISED_FOR_NUM_TARGS(qreal, gpu_statevec_calcProbOfMultiQubitOutcome_sub,
                                         (Qureg, vector<int>, vector<int>))
INSTANTIATE_FUNC_OPTIMISED_FOR_NUM_TARGS(qreal, gpu_densmatr_calcProbOfMultiQubitOutcome_sub,
                                         (Qureg, vector<int>, vector<int>))
INSTANTIATE_FUNC_OPTIMISED_FOR_NUM_TARGS(void, gpu_statevec_calcProbsOfAllMultiQubitOutcomes_sub,
                                         (qreal * outProbs, Qureg, vector<int>))
INSTANTIATE_FUNC_OPTIMISED_FOR_NUM_TARGS(void, gpu_densmatr_calcProbsOfAllMultiQubitOutcomes_sub,
                                         (qreal * outProbs, Qureg, vector<int>))


/*
 * INNER PRODUCTS (gpu_subroutines.hpp:149-156)
 */

qcomp gpu_statevec_calcInnerProduct_sub(Qureg quregA, Qureg quregB) {
    double v[2];
    QA_CALL(qa_statevec_calcInnerProduct(amps(quregA), amps(quregB), PREC, bits(quregA), v));
    return toComp(v);
}

qreal gpu_densmatr_calcHilbertSchmidtDistance_sub(Qureg quregA, Qureg quregB) {
    double v[2] = {0, 0};
    QA_CALL(qa_densmatr_calcHilbertSchmidtDistance(amps(quregA), amps(quregB), PREC, quregA.numQubits, v));
    return v[0];  // the sum of |A - B|^2, not yet square-rooted (as QuEST's)
}

template <bool Conj>
qcomp gpu_densmatr_calcFidelityWithPureState_sub(Qureg rho, Qureg psi) {
    double v[2];
    QA_CALL(qa_densmatr_calcFidelityWithPureState(amps(rho), amps(psi), PREC, rho.numQubits, Conj ? 1 : 0, v));
    return toComp(v);
}

template qcomp gpu_densmatr_calcFidelityWithPureState_sub<true>(Qureg, Qureg);
template qcomp gpu_densmatr_calcFidelityWithPureState_sub<false>(Qureg, Qureg);


/*
 * EXPECTATION VALUES (gpu_subroutines.hpp:159-173); the Pauli-string
 * reductions include the i^#Y factor, as QuEST's do
 */

qreal gpu_statevec_calcExpecAnyTargZ_sub(Qureg qureg, vector<int> targs) {
    double v = 0;
    QA_CALL(qa_statevec_calcExpecAnyTargZ(amps(qureg), PREC, bits(qureg), maskOf(targs), &v));
    return v;
}
qcomp gpu_densmatr_calcExpecAnyTargZ_sub(Qureg qureg, vector<int> targs) {
    double v[2];
    QA_CALL(qa_densmatr_calcExpecAnyTargZ(amps(qureg), PREC, qureg.numQubits, maskOf(targs), v));
    return toComp(v);
}

qcomp gpu_statevec_calcExpecPauliStr_subA(Qureg qureg, vector<int> x, vector<int> y, vector<int> z) {
    double v[2];
    QA_CALL(qa_statevec_calcExpecPauliStr(amps(qureg), PREC, bits(qureg), maskOf(x), maskOf(y), maskOf(z), v));
    return toComp(v);
}
qcomp gpu_statevec_calcExpecPauliStr_subB(Qureg, vector<int>, vector<int>, vector<int>) { notOnOneGpu(__func__); }
qcomp gpu_densmatr_calcExpecPauliStr_sub(Qureg qureg, vector<int> x, vector<int> y, vector<int> z) {
    double v[2];
    QA_CALL(qa_densmatr_calcExpecPauliStr(amps(qureg), PREC, qureg.numQubits, maskOf(x), maskOf(y), maskOf(z), v));
    return toComp(v);
}

// powMode of the kernel layer: 0 no power, 1 complex power, 2 real power of
// the elements' real parts
template <bool HasPower, bool UseRealPow>
qcomp gpu_statevec_calcExpecFullStateDiagMatr_sub(Qureg qureg, FullStateDiagMatr matr, qcomp exponent) {
    double v[2];
    QA_CALL(qa_statevec_calcExpecFullStateDiagMatr(amps(qureg), PREC, bits(qureg), matr.gpuElems,
                                                   HasPower ? (UseRealPow ? 2 : 1) : 0, std::real(exponent),
                                                   std::imag(exponent), v));
    return toComp(v);
}
template <bool HasPower, bool UseRealPow>
qcomp gpu_densmatr_calcExpecFullStateDiagMatr_sub(Qureg qureg, FullStateDiagMatr matr, qcomp exponent) {
    double v[2];
    QA_CALL(qa_densmatr_calcExpecFullStateDiagMatr(amps(qureg), PREC, qureg.numQubits, matr.gpuElems,
                                                   HasPower ? (UseRealPow ? 2 : 1) : 0, std::real(exponent),
                                                   std::imag(exponent), v));
    return toComp(v);
}

template qcomp gpu_statevec_calcExpecFullStateDiagMatr_sub<true, true>(Qureg, FullStateDiagMatr, qcomp);
template qcomp gpu_statevec_calcExpecFullStateDiagMatr_sub<true, false>(Qureg, FullStateDiagMatr, qcomp);
template qcomp gpu_statevec_calcExpecFullStateDiagMatr_sub<false, false>(Qureg, FullStateDiagMatr, qcomp);
template qcomp gpu_statevec_calcExpecFullStateDiagMatr_sub<false, true>(Qureg, FullStateDiagMatr, qcomp);
template qcomp gpu_densmatr_calcExpecFullStateDiagMatr_sub<true, true>(Qureg, FullStateDiagMatr, qcomp);
template qcomp gpu_densmatr_calcExpecFullStateDiagMatr_sub<true, false>(Qureg, FullStateDiagMatr, qcomp);
template qcomp gpu_densmatr_calcExpecFullStateDiagMatr_sub<false, false>(Qureg, FullStateDiagMatr, qcomp);
template qcomp gpu_densmatr_calcExpecFullStateDiagMatr_sub<false, true>(Qureg, FullStateDiagMatr, qcomp);


/*
 * PROJECTORS (gpu_subroutines.hpp:180-183)
 */

template <int NumQubits>
void gpu_statevec_multiQubitProjector_sub(Qureg qureg, vector<int> qubits, vector<int> outcomes, qreal prob) {
    QA_CALL(qa_statevec_multiQubitProjector(amps(qureg), PREC, bits(qureg), qubits.data(), outcomes.data(),
                                            (int)qubits.size(), prob));
}
template <int NumQubits>
void gpu_densmatr_multiQubitProjector_sub(Qureg qureg, vector<int> qubits, vector<int> outcomes, qreal prob) {
    QA_CALL(qa_densmatr_multiQubitProjector(amps(qureg), PREC, qureg.numQubits, qubits.data(), outcomes.data(),
                                            (int)qubits.size(), prob));
}

INSTANTIATE_FUNC_OPTIMISED_FOR_NUM_TARGS(void, gpu_statevec_multiQubitProjector_sub,
                                         (Qureg qureg, vector<int> qubits, vector<int> outcomes, qreal prob))
INSTANTIATE_FUNC_OPTIMISED_FOR_NUM_TARGS(void, gpu_densmatr_multiQubitProjector_sub,
                                         (Qureg qureg, vector<int> qubits, vector<int> outcomes, qreal prob))


/*
 * STATE INITIALISATION (gpu_subroutines.hpp:186-192)
 */

void gpu_statevec_initUniformState_sub(Qureg qureg, qcomp amp) {
    QA_CALL(qa_statevec_initUniformState(amps(qureg), PREC, bits(qureg), std::real(amp), std::imag(amp)));
}

void gpu_statevec_initDebugState_sub(Qureg qureg) {
    QA_CALL(qa_statevec_initDebugState(amps(qureg), PREC, bits(qureg)));
}

// one seed from QuEST's measurement RNG, as its GPU path draws
// (gpu_thrust.cuh:1073-1084, randomiser.cpp:175-193)
void gpu_statevec_initUnnormalisedUniformlyRandomPureStateAmps_sub(Qureg qureg) {
    const unsigned seed = rand_getThreadSharedRandomSeed(qureg.isDistributed);
    QA_CALL(qa_statevec_initUnnormalisedUniformlyRandomPureStateAmps(amps(qureg), PREC, bits(qureg), seed));
}
