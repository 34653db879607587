// Prints the size of every by-value struct of the quest.h API and the offset
// of each of its fields, compiled either against this backend's
// include/quest.h or (with QA_LAYOUT_REFERENCE) against QuEST's own headers
// (quest/include/*.h of a v4.1.0 checkout, FLOAT_PRECISION=2); the two
// outputs must be identical for the header to be a drop-in.
#include <cstddef>
#include <cstdio>

#ifdef QA_LAYOUT_REFERENCE
#include "quest/include/types.h"
#include "quest/include/environment.h"
#include "quest/include/qureg.h"
#include "quest/include/matrices.h"
#include "quest/include/paulis.h"
#include "quest/include/channels.h"
#else
#include "quest.h"
#endif

#define SIZE(T) std::printf("sizeof %s %zu\n", #T, sizeof(T))
#define OFF(T, F) std::printf("offsetof %s.%s %zu\n", #T, #F, offsetof(T, F))

int main() {
    SIZE(qreal); SIZE(qindex); SIZE(qcomp);
    SIZE(QuESTEnv);
    OFF(QuESTEnv, isMultithreaded); OFF(QuESTEnv, isGpuAccelerated); OFF(QuESTEnv, isDistributed);
    OFF(QuESTEnv, rank); OFF(QuESTEnv, numNodes);
    SIZE(Qureg);
    OFF(Qureg, isMultithreaded); OFF(Qureg, isGpuAccelerated); OFF(Qureg, isDistributed); OFF(Qureg, rank);
    OFF(Qureg, numNodes); OFF(Qureg, logNumNodes); OFF(Qureg, isDensityMatrix); OFF(Qureg, numQubits);
    OFF(Qureg, numAmps); OFF(Qureg, logNumAmps); OFF(Qureg, numAmpsPerNode); OFF(Qureg, logNumAmpsPerNode);
    OFF(Qureg, logNumColsPerNode); OFF(Qureg, cpuAmps); OFF(Qureg, gpuAmps); OFF(Qureg, cpuCommBuffer);
    OFF(Qureg, gpuCommBuffer);
    SIZE(CompMatr1); OFF(CompMatr1, numQubits); OFF(CompMatr1, numRows); OFF(CompMatr1, elems);
    SIZE(CompMatr2); OFF(CompMatr2, numQubits); OFF(CompMatr2, numRows); OFF(CompMatr2, elems);
    SIZE(CompMatr);
    OFF(CompMatr, numQubits); OFF(CompMatr, numRows); OFF(CompMatr, isApproxUnitary);
    OFF(CompMatr, isApproxHermitian); OFF(CompMatr, wasGpuSynced); OFF(CompMatr, cpuElems);
    OFF(CompMatr, cpuElemsFlat); OFF(CompMatr, gpuElemsFlat);
    SIZE(DiagMatr1); OFF(DiagMatr1, numQubits); OFF(DiagMatr1, numElems); OFF(DiagMatr1, elems);
    SIZE(DiagMatr2); OFF(DiagMatr2, numQubits); OFF(DiagMatr2, numElems); OFF(DiagMatr2, elems);
    SIZE(DiagMatr);
    OFF(DiagMatr, numQubits); OFF(DiagMatr, numElems); OFF(DiagMatr, isApproxUnitary);
    OFF(DiagMatr, isApproxHermitian); OFF(DiagMatr, isApproxNonZero); OFF(DiagMatr, isStrictlyNonNegative);
    OFF(DiagMatr, wasGpuSynced); OFF(DiagMatr, cpuElems); OFF(DiagMatr, gpuElems);
    SIZE(FullStateDiagMatr);
    OFF(FullStateDiagMatr, numQubits); OFF(FullStateDiagMatr, numElems); OFF(FullStateDiagMatr, isGpuAccelerated);
    OFF(FullStateDiagMatr, isMultithreaded); OFF(FullStateDiagMatr, isDistributed);
    OFF(FullStateDiagMatr, numElemsPerNode); OFF(FullStateDiagMatr, isApproxUnitary);
    OFF(FullStateDiagMatr, isApproxHermitian); OFF(FullStateDiagMatr, isApproxNonZero);
    OFF(FullStateDiagMatr, isStrictlyNonNegative); OFF(FullStateDiagMatr, wasGpuSynced);
    OFF(FullStateDiagMatr, cpuElems); OFF(FullStateDiagMatr, gpuElems);
    SIZE(PauliStr); OFF(PauliStr, lowPaulis); OFF(PauliStr, highPaulis);
    SIZE(PauliStrSum);
    OFF(PauliStrSum, numTerms); OFF(PauliStrSum, strings); OFF(PauliStrSum, coeffs);
    OFF(PauliStrSum, isApproxHermitian);
    SIZE(SuperOp);
    OFF(SuperOp, numQubits); OFF(SuperOp, numRows); OFF(SuperOp, cpuElems); OFF(SuperOp, cpuElemsFlat);
    OFF(SuperOp, gpuElemsFlat); OFF(SuperOp, wasGpuSynced);
    SIZE(KrausMap);
    OFF(KrausMap, numQubits); OFF(KrausMap, numMatrices); OFF(KrausMap, numRows); OFF(KrausMap, matrices);
    OFF(KrausMap, superop); OFF(KrausMap, isApproxCPTP);
    return 0;
}
