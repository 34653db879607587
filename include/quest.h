/*
 * quest.h -- QuEST v4 API subset served by the MI355X backend.
 *
 * A program written against QuEST's quest.h (reference: quest/include/
 * quest.h.in:74-110 and the headers it includes) compiles unchanged against
 * this header for every function declared here, and links against
 * libquest_amd.so instead of libQuEST.so.  The struct layouts (Qureg,
 * CompMatr1/2, CompMatr, DiagMatr1/2, DiagMatr, FullStateDiagMatr, PauliStr,
 * PauliStrSum, SuperOp, KrausMap, QuESTEnv) and the signatures are those of
 * the reference, because a drop-in must be ABI compatible with code that
 * fills these structs or passes them by value.  Precision is fixed at
 * qreal = double (FLOAT_PRECISION 2); per-qureg FP32 amplitudes are an
 * extension (quest_amd_ext.h) that keeps these signatures.
 *
 * Every qureg lives in GPU memory.  There is no CPU execution path: with no
 * usable GPU, initQuESTEnv() reports a validation error.
 */
#ifndef QUEST_H
#define QUEST_H

#define QUEST_AMD 1
#define FLOAT_PRECISION 2
#define COMPILE_MPI 0
#define COMPILE_OPENMP 0
#define COMPILE_CUDA 0
#define COMPILE_CUQUANTUM 0

#ifdef __cplusplus
#include <complex>
#include <string>
#include <vector>
#else
#include <complex.h>
#endif

/* ---------------------------------------------------------------- types */
/* quest/include/types.h:31-73, precision.h:36-96 */
typedef double qreal;
typedef long long int qindex;
#define PAULI_MASK_TYPE long long unsigned int
#ifdef __cplusplus
typedef std::complex<double> qcomp;
#else
typedef double _Complex qcomp;
#endif

#ifdef __cplusplus
static inline qcomp getQcomp(qreal re, qreal im) { return qcomp(re, im); }
static inline qcomp operator""_i(long double y) { return qcomp(0, static_cast<qreal>(y)); }
static inline qcomp operator""_i(unsigned long long int y) { return qcomp(0, static_cast<qreal>(y)); }
#else
static inline qcomp getQcomp(qreal re, qreal im) { return re + I * im; }
#endif

#ifdef __cplusplus
extern "C" {
#endif

/* quest/include/environment.h */
typedef struct {
    int isMultithreaded;
    int isGpuAccelerated;
    int isDistributed;
    int rank;
    int numNodes;
} QuESTEnv;

/* quest/include/qureg.h:49-80 */
typedef struct {
    int isMultithreaded;
    int isGpuAccelerated;
    int isDistributed;
    int rank;
    int numNodes;
    int logNumNodes;
    int isDensityMatrix;
    int numQubits;
    qindex numAmps;
    qindex logNumAmps;
    qindex numAmpsPerNode;
    qindex logNumAmpsPerNode;
    qindex logNumColsPerNode;
    qcomp* cpuAmps;
    qcomp* gpuAmps;
    qcomp* cpuCommBuffer;
    qcomp* gpuCommBuffer;
} Qureg;

/* quest/include/matrices.h:68-186 */
typedef struct {
    int numQubits;
    qindex numRows;
    qcomp elems[2][2];
} CompMatr1;

typedef struct {
    int numQubits;
    qindex numRows;
    qcomp elems[4][4];
} CompMatr2;

typedef struct {
    int numQubits;
    qindex numRows;
    int* isApproxUnitary;
    int* isApproxHermitian;
    int* wasGpuSynced;
    qcomp** cpuElems;
    qcomp* cpuElemsFlat;
    qcomp* gpuElemsFlat;
} CompMatr;

typedef struct {
    int numQubits;
    qindex numElems;
    qcomp elems[2];
} DiagMatr1;

typedef struct {
    int numQubits;
    qindex numElems;
    qcomp elems[4];
} DiagMatr2;

typedef struct {
    int numQubits;
    qindex numElems;
    int* isApproxUnitary;
    int* isApproxHermitian;
    int* isApproxNonZero;
    int* isStrictlyNonNegative;
    int* wasGpuSynced;
    qcomp* cpuElems;
    qcomp* gpuElems;
} DiagMatr;

typedef struct {
    int numQubits;
    qindex numElems;
    int isGpuAccelerated;
    int isMultithreaded;
    int isDistributed;
    qindex numElemsPerNode;
    int* isApproxUnitary;
    int* isApproxHermitian;
    int* isApproxNonZero;
    int* isStrictlyNonNegative;
    int* wasGpuSynced;
    qcomp* cpuElems;
    qcomp* gpuElems;
} FullStateDiagMatr;

/* quest/include/paulis.h:53-80: 2 bits per qubit (I=0 X=1 Y=2 Z=3),
   qubits 0..31 in lowPaulis, 32..63 in highPaulis */
typedef struct {
    PAULI_MASK_TYPE lowPaulis;
    PAULI_MASK_TYPE highPaulis;
} PauliStr;

typedef struct {
    qindex numTerms;
    PauliStr* strings;
    qcomp* coeffs;
    int* isApproxHermitian;
} PauliStrSum;

/* quest/include/channels.h:74-118 */
typedef struct {
    int numQubits;
    qindex numRows;
    qcomp** cpuElems;
    qcomp* cpuElemsFlat;
    qcomp* gpuElemsFlat;
    int* wasGpuSynced;
} SuperOp;

typedef struct {
    int numQubits;
    int numMatrices;
    qindex numRows;
    qcomp*** matrices;
    SuperOp superop;
    int* isApproxCPTP;
} KrausMap;

/* ---------------------------------------------------------- environment */
void initQuESTEnv(void);
void initCustomQuESTEnv(int useDistrib, int useGpuAccel, int useMultithread);
void finalizeQuESTEnv(void);
void syncQuESTEnv(void);
void reportQuESTEnv(void);
int isQuESTEnvInit(void);
QuESTEnv getQuESTEnv(void);

/* ---------------------------------------------------------------- debug */
void setSeeds(unsigned* seeds, int numSeeds);
void setSeedsToDefault(void);
void getSeeds(unsigned* seeds);
int getNumSeeds(void);
void setInputErrorHandler(void (*callback)(const char* func, const char* msg));
void setValidationOn(void);
void setValidationOff(void);
void setValidationEpsilonToDefault(void);
void setValidationEpsilon(qreal eps);
qreal getValidationEpsilon(void);
qindex getGpuCacheSize(void);
void clearGpuCache(void);
void getEnvironmentString(char str[200]);

/* ---------------------------------------------------------------- qureg */
Qureg createQureg(int numQubits);
Qureg createDensityQureg(int numQubits);
Qureg createForcedQureg(int numQubits);
Qureg createForcedDensityQureg(int numQubits);
Qureg createCustomQureg(int numQubits, int isDensMatr, int useDistrib, int useGpuAccel, int useMultithread);
Qureg createCloneQureg(Qureg qureg);
void destroyQureg(Qureg qureg);
void reportQuregParams(Qureg qureg);
void reportQureg(Qureg qureg);
void syncQuregToGpu(Qureg qureg);
void syncQuregFromGpu(Qureg qureg);
void syncSubQuregToGpu(Qureg qureg, qindex localStartInd, qindex numLocalAmps);
void syncSubQuregFromGpu(Qureg qureg, qindex localStartInd, qindex numLocalAmps);
void getQuregAmps(qcomp* outAmps, Qureg qureg, qindex startInd, qindex numAmps);
void getDensityQuregAmps(qcomp** outAmps, Qureg qureg, qindex startRow, qindex startCol, qindex numRows,
                         qindex numCols);
void _wrap_getQuregAmp(qcomp* out, Qureg qureg, qindex index);
void _wrap_getDensityQuregAmp(qcomp* out, Qureg qureg, qindex row, qindex column);

/* -------------------------------------------------------- initialisation */
void initBlankState(Qureg qureg);
void initZeroState(Qureg qureg);
void initPlusState(Qureg qureg);
void initPureState(Qureg qureg, Qureg pure);
void initClassicalState(Qureg qureg, qindex stateInd);
void initDebugState(Qureg qureg);
void initArbitraryPureState(Qureg qureg, qcomp* amps);
void initRandomPureState(Qureg qureg);
void initRandomMixedState(Qureg qureg, qindex numPureStates);
void setQuregAmps(Qureg qureg, qindex startInd, qcomp* amps, qindex numAmps);
void setDensityQuregAmps(Qureg qureg, qindex startRow, qindex startCol, qcomp** amps, qindex numRows,
                         qindex numCols);
void setDensityQuregFlatAmps(Qureg qureg, qindex startInd, qcomp* amps, qindex numAmps);
void setQuregToClone(Qureg targetQureg, Qureg copyQureg);
void setQuregToSuperposition(qcomp facOut, Qureg out, qcomp fac1, Qureg qureg1, qcomp fac2, Qureg qureg2);
qreal setQuregToRenormalized(Qureg qureg);
void setQuregToPauliStrSum(Qureg qureg, PauliStrSum sum); /* initialisations.cpp:216-227 */
void setQuregToPartialTrace(Qureg out, Qureg in, int* traceOutQubits, int numTraceQubits);
void setQuregToReducedDensityMatrix(Qureg out, Qureg in, int* retainQubits, int numRetainQubits);

/* ------------------------------------------------------------- matrices */
void _validateNewNestedElemsPtrNotNull(qcomp** ptrs, int numQubits, const char* caller);
void _validateNewElemsPtrNotNull(qcomp* ptr, const char* caller);
CompMatr createCompMatr(int numQubits);
DiagMatr createDiagMatr(int numQubits);
FullStateDiagMatr createFullStateDiagMatr(int numQubits);
FullStateDiagMatr createCustomFullStateDiagMatr(int numQubits, int useDistrib, int useGpuAccel, int useMultithread);
void destroyCompMatr(CompMatr matrix);
void destroyDiagMatr(DiagMatr matrix);
void destroyFullStateDiagMatr(FullStateDiagMatr matrix);
void syncCompMatr(CompMatr matr);
void syncDiagMatr(DiagMatr matr);
void syncFullStateDiagMatr(FullStateDiagMatr matr);
void setCompMatr(CompMatr matr, qcomp** vals);
void setDiagMatr(DiagMatr out, qcomp* in);
void setFullStateDiagMatr(FullStateDiagMatr out, qindex startInd, qcomp* in, qindex numElems);
/* matrices.cpp:668-730: diagonals as functions of integer variables (each
   variable the next numQubitsPerVar[v] index bits, lowest first) */
void setDiagMatrFromMultiVarFunc(DiagMatr out, qcomp (*func)(qindex*), int* numQubitsPerVar, int numVars,
                                 int areSigned);
void setDiagMatrFromMultiDimLists(DiagMatr out, void* lists, int* numQubitsPerDim, int numDims);
void setFullStateDiagMatrFromMultiVarFunc(FullStateDiagMatr out, qcomp (*func)(qindex*), int* numQubitsPerVar,
                                          int numVars, int areSigned);
void setFullStateDiagMatrFromMultiDimLists(FullStateDiagMatr out, void* lists, int* numQubitsPerDim, int numDims);
/* matrices.cpp:635-665: the diagonal of a sum of I/Z Pauli strings */
void setFullStateDiagMatrFromPauliStrSum(FullStateDiagMatr out, PauliStrSum in);
FullStateDiagMatr createFullStateDiagMatrFromPauliStrSum(PauliStrSum in);

/* --------------------------------------------------------------- paulis */
PauliStr getPauliStr(const char* paulis, int* indices, int numPaulis);
PauliStrSum createPauliStrSum(PauliStr* strings, qcomp* coeffs, qindex numTerms);
PauliStrSum createInlinePauliStrSum(const char* str);
void destroyPauliStrSum(PauliStrSum sum);
void reportPauliStr(PauliStr str);
void reportPauliStrSum(PauliStrSum str);

/* ------------------------------------------------------------- channels */
KrausMap createKrausMap(int numQubits, int numOperators);
void syncKrausMap(KrausMap map);
void destroyKrausMap(KrausMap map);
void reportKrausMap(KrausMap map);
SuperOp createSuperOp(int numQubits);
void syncSuperOp(SuperOp op);
void destroySuperOp(SuperOp op);
void reportSuperOp(SuperOp op);
void setKrausMap(KrausMap map, qcomp*** matrices);
void setSuperOp(SuperOp op, qcomp** matrix);

/* ----------------------------------------------------------- operations */
void multiplyCompMatr1(Qureg qureg, int target, CompMatr1 matrix);
void applyCompMatr1(Qureg qureg, int target, CompMatr1 matrix);
void applyControlledCompMatr1(Qureg qureg, int control, int target, CompMatr1 matrix);
void applyMultiControlledCompMatr1(Qureg qureg, int* controls, int numControls, int target, CompMatr1 matrix);
void applyMultiStateControlledCompMatr1(Qureg qureg, int* controls, int* states, int numControls, int target,
                                        CompMatr1 matrix);
void multiplyCompMatr2(Qureg qureg, int target1, int target2, CompMatr2 matr);
void applyCompMatr2(Qureg qureg, int target1, int target2, CompMatr2 matrix);
void applyControlledCompMatr2(Qureg qureg, int control, int target1, int target2, CompMatr2 matr);
void applyMultiControlledCompMatr2(Qureg qureg, int* controls, int numControls, int target1, int target2,
                                   CompMatr2 matr);
void applyMultiStateControlledCompMatr2(Qureg qureg, int* controls, int* states, int numControls, int target1,
                                        int target2, CompMatr2 matr);
void multiplyCompMatr(Qureg qureg, int* targets, int numTargets, CompMatr matrix);
void applyCompMatr(Qureg qureg, int* targets, int numTargets, CompMatr matr);
void applyControlledCompMatr(Qureg qureg, int control, int* targets, int numTargets, CompMatr matr);
void applyMultiControlledCompMatr(Qureg qureg, int* controls, int numControls, int* targets, int numTargets,
                                  CompMatr matr);
void applyMultiStateControlledCompMatr(Qureg qureg, int* controls, int* states, int numControls, int* targets,
                                       int numTargets, CompMatr matr);
void multiplyDiagMatr1(Qureg qureg, int target, DiagMatr1 matr);
void applyDiagMatr1(Qureg qureg, int target, DiagMatr1 matr);
void applyControlledDiagMatr1(Qureg qureg, int control, int target, DiagMatr1 matr);
void applyMultiControlledDiagMatr1(Qureg qureg, int* controls, int numControls, int target, DiagMatr1 matr);
void applyMultiStateControlledDiagMatr1(Qureg qureg, int* controls, int* states, int numControls, int target,
                                        DiagMatr1 matr);
void multiplyDiagMatr2(Qureg qureg, int target1, int target2, DiagMatr2 matr);
void applyDiagMatr2(Qureg qureg, int target1, int target2, DiagMatr2 matr);
void applyControlledDiagMatr2(Qureg qureg, int control, int target1, int target2, DiagMatr2 matr);
void applyMultiControlledDiagMatr2(Qureg qureg, int* controls, int numControls, int target1, int target2,
                                   DiagMatr2 matr);
void applyMultiStateControlledDiagMatr2(Qureg qureg, int* controls, int* states, int numControls, int target1,
                                        int target2, DiagMatr2 matr);
void multiplyDiagMatr(Qureg qureg, int* targets, int numTargets, DiagMatr matrix);
void applyDiagMatr(Qureg qureg, int* targets, int numTargets, DiagMatr matrix);
void applyControlledDiagMatr(Qureg qureg, int control, int* targets, int numTargets, DiagMatr matrix);
void applyMultiControlledDiagMatr(Qureg qureg, int* controls, int numControls, int* targets, int numTargets,
                                  DiagMatr matrix);
void applyMultiStateControlledDiagMatr(Qureg qureg, int* controls, int* states, int numControls, int* targets,
                                       int numTargets, DiagMatr matrix);
/* DiagMatr powers (operations.cpp:405-500): elements raised to a complex
   exponent (pow(qcomp, qcomp)) before use; the apply forms require a
   unitary matrix and an (approximately) real exponent */
void multiplyDiagMatrPower(Qureg qureg, int* targets, int numTargets, DiagMatr matrix, qcomp exponent);
void applyDiagMatrPower(Qureg qureg, int* targets, int numTargets, DiagMatr matrix, qcomp exponent);
void applyControlledDiagMatrPower(Qureg qureg, int control, int* targets, int numTargets, DiagMatr matrix,
                                  qcomp exponent);
void applyMultiControlledDiagMatrPower(Qureg qureg, int* controls, int numControls, int* targets, int numTargets,
                                       DiagMatr matrix, qcomp exponent);
void applyMultiStateControlledDiagMatrPower(Qureg qureg, int* controls, int* states, int numControls, int* targets,
                                            int numTargets, DiagMatr matrix, qcomp exponent);
void multiplyFullStateDiagMatr(Qureg qureg, FullStateDiagMatr matrix);
void applyFullStateDiagMatr(Qureg qureg, FullStateDiagMatr matrix);
/* operations.cpp:545-576 */
void multiplyFullStateDiagMatrPower(Qureg qureg, FullStateDiagMatr matrix, qcomp exponent);
void applyFullStateDiagMatrPower(Qureg qureg, FullStateDiagMatr matrix, qcomp exponent);
void applyS(Qureg qureg, int target);
void applyControlledS(Qureg qureg, int control, int target);
void applyMultiControlledS(Qureg qureg, int* controls, int numControls, int target);
void applyMultiStateControlledS(Qureg qureg, int* controls, int* states, int numControls, int target);
void applyT(Qureg qureg, int target);
void applyControlledT(Qureg qureg, int control, int target);
void applyMultiControlledT(Qureg qureg, int* controls, int numControls, int target);
void applyMultiStateControlledT(Qureg qureg, int* controls, int* states, int numControls, int target);
void applyHadamard(Qureg qureg, int target);
void applyControlledHadamard(Qureg qureg, int control, int target);
void applyMultiControlledHadamard(Qureg qureg, int* controls, int numControls, int target);
void applyMultiStateControlledHadamard(Qureg qureg, int* controls, int* states, int numControls, int target);
void multiplySwap(Qureg qureg, int qubit1, int qubit2);
void applySwap(Qureg qureg, int qubit1, int qubit2);
void applyControlledSwap(Qureg qureg, int control, int qubit1, int qubit2);
void applyMultiControlledSwap(Qureg qureg, int* controls, int numControls, int qubit1, int qubit2);
void applyMultiStateControlledSwap(Qureg qureg, int* controls, int* states, int numControls, int qubit1,
                                   int qubit2);
void applySqrtSwap(Qureg qureg, int qubit1, int qubit2);
/* operations.cpp:825-860 */
void applyControlledSqrtSwap(Qureg qureg, int control, int qubit1, int qubit2);
void applyMultiControlledSqrtSwap(Qureg qureg, int* controls, int numControls, int qubit1, int qubit2);
void applyMultiStateControlledSqrtSwap(Qureg qureg, int* controls, int* states, int numControls, int qubit1,
                                       int qubit2);
void applyPauliX(Qureg qureg, int target);
void applyPauliY(Qureg qureg, int target);
void applyPauliZ(Qureg qureg, int target);
void applyControlledPauliX(Qureg qureg, int control, int target);
void applyControlledPauliY(Qureg qureg, int control, int target);
void applyControlledPauliZ(Qureg qureg, int control, int target);
void applyMultiControlledPauliX(Qureg qureg, int* controls, int numControls, int target);
void applyMultiControlledPauliY(Qureg qureg, int* controls, int numControls, int target);
void applyMultiControlledPauliZ(Qureg qureg, int* controls, int numControls, int target);
void applyMultiStateControlledPauliX(Qureg qureg, int* controls, int* states, int numControls, int target);
void applyMultiStateControlledPauliY(Qureg qureg, int* controls, int* states, int numControls, int target);
void applyMultiStateControlledPauliZ(Qureg qureg, int* controls, int* states, int numControls, int target);
void multiplyPauliStr(Qureg qureg, PauliStr str);
void applyPauliStr(Qureg qureg, PauliStr str);
void applyControlledPauliStr(Qureg qureg, int control, PauliStr str);
void applyMultiControlledPauliStr(Qureg qureg, int* controls, int numControls, PauliStr str);
void applyMultiStateControlledPauliStr(Qureg qureg, int* controls, int* states, int numControls, PauliStr str);
void applyRotateX(Qureg qureg, int target, qreal angle);
void applyRotateY(Qureg qureg, int target, qreal angle);
void applyRotateZ(Qureg qureg, int target, qreal angle);
void applyControlledRotateX(Qureg qureg, int control, int target, qreal angle);
void applyControlledRotateY(Qureg qureg, int control, int target, qreal angle);
void applyControlledRotateZ(Qureg qureg, int control, int target, qreal angle);
void applyMultiControlledRotateX(Qureg qureg, int* controls, int numControls, int target, qreal angle);
void applyMultiControlledRotateY(Qureg qureg, int* controls, int numControls, int target, qreal angle);
void applyMultiControlledRotateZ(Qureg qureg, int* controls, int numControls, int target, qreal angle);
void applyMultiStateControlledRotateX(Qureg qureg, int* controls, int* states, int numControls, int target,
                                      qreal angle);
void applyMultiStateControlledRotateY(Qureg qureg, int* controls, int* states, int numControls, int target,
                                      qreal angle);
void applyMultiStateControlledRotateZ(Qureg qureg, int* controls, int* states, int numControls, int target,
                                      qreal angle);
void applyRotateAroundAxis(Qureg qureg, int targ, qreal angle, qreal axisX, qreal axisY, qreal axisZ);
void applyControlledRotateAroundAxis(Qureg qureg, int ctrl, int targ, qreal angle, qreal axisX, qreal axisY,
                                     qreal axisZ);
void applyMultiControlledRotateAroundAxis(Qureg qureg, int* ctrls, int numCtrls, int targ, qreal angle,
                                          qreal axisX, qreal axisY, qreal axisZ);
void applyMultiStateControlledRotateAroundAxis(Qureg qureg, int* ctrls, int* states, int numCtrls, int targ,
                                               qreal angle, qreal axisX, qreal axisY, qreal axisZ);
void multiplyPauliGadget(Qureg qureg, PauliStr str, qreal angle);
void applyPauliGadget(Qureg qureg, PauliStr str, qreal angle);
void applyControlledPauliGadget(Qureg qureg, int control, PauliStr str, qreal angle);
void applyMultiControlledPauliGadget(Qureg qureg, int* controls, int numControls, PauliStr str, qreal angle);
void applyMultiStateControlledPauliGadget(Qureg qureg, int* controls, int* states, int numControls, PauliStr str,
                                          qreal angle);
void multiplyPhaseGadget(Qureg qureg, int* targets, int numTargets, qreal angle);
void applyPhaseGadget(Qureg qureg, int* targets, int numTargets, qreal angle);
void applyControlledPhaseGadget(Qureg qureg, int control, int* targets, int numTargets, qreal angle);
void applyMultiControlledPhaseGadget(Qureg qureg, int* controls, int numControls, int* targets, int numTargets,
                                     qreal angle);
void applyMultiStateControlledPhaseGadget(Qureg qureg, int* controls, int* states, int numControls,
                                          int* targets, int numTargets, qreal angle);
/* Pauli string sums (operations.cpp:1112-1197): qureg <- sum * qureg through
   a workspace of the same shape (on a density matrix, left multiplication),
   and exp(i angle sum) by a Trotter-Suzuki product of Pauli gadgets */
void multiplyPauliStrSum(Qureg qureg, PauliStrSum sum, Qureg workspace);
void applyTrotterizedPauliStrSumGadget(Qureg qureg, PauliStrSum sum, qreal angle, int order, int reps);
void applyPhaseFlip(Qureg qureg, int target);
void applyPhaseShift(Qureg qureg, int target, qreal angle);
void applyTwoQubitPhaseFlip(Qureg qureg, int target1, int target2);
void applyTwoQubitPhaseShift(Qureg qureg, int target1, int target2, qreal angle);
void applyMultiQubitPhaseFlip(Qureg qureg, int* targets, int numTargets);
void applyMultiQubitPhaseShift(Qureg qureg, int* targets, int numTargets, qreal angle);
void multiplyMultiQubitNot(Qureg qureg, int* targets, int numTargets);
void applyMultiQubitNot(Qureg qureg, int* targets, int numTargets);
void applyControlledMultiQubitNot(Qureg qureg, int control, int* targets, int numTargets);
void applyMultiControlledMultiQubitNot(Qureg qureg, int* controls, int numControls, int* targets, int numTargets);
void applyMultiStateControlledMultiQubitNot(Qureg qureg, int* controls, int* states, int numControls,
                                            int* targets, int numTargets);
void applyQubitProjector(Qureg qureg, int target, int outcome);
void applyMultiQubitProjector(Qureg qureg, int* qubits, int* outcomes, int numQubits);
int applyQubitMeasurement(Qureg qureg, int target);
int applyQubitMeasurementAndGetProb(Qureg qureg, int target, qreal* probability);
qreal applyForcedQubitMeasurement(Qureg qureg, int target, int outcome);
qindex applyMultiQubitMeasurement(Qureg qureg, int* qubits, int numQubits);
qindex applyMultiQubitMeasurementAndGetProb(Qureg qureg, int* qubits, int numQubits, qreal* probability);
qreal applyForcedMultiQubitMeasurement(Qureg qureg, int* qubits, int* outcomes, int numQubits);
void applyQuantumFourierTransform(Qureg qureg, int* targets, int numTargets);
void applyFullQuantumFourierTransform(Qureg qureg);

/* ---------------------------------------------------------- decoherence */
void mixDephasing(Qureg qureg, int qubit, qreal prob);
void mixTwoQubitDephasing(Qureg qureg, int qubit1, int qubit2, qreal prob);
void mixDepolarising(Qureg qureg, int qubit, qreal prob);
void mixTwoQubitDepolarising(Qureg qureg, int qubit1, int qubit2, qreal prob);
void mixDamping(Qureg qureg, int qubit, qreal prob);
void mixPaulis(Qureg qureg, int qubit, qreal probX, qreal probY, qreal probZ);
void mixQureg(Qureg qureg, Qureg other, qreal prob);
void mixKrausMap(Qureg qureg, int* qubits, int numQubits, KrausMap map);
void mixSuperOp(Qureg qureg, int* targets, int numTargets, SuperOp superop);

/* --------------------------------------------------------- calculations */
qreal calcExpecPauliStr(Qureg qureg, PauliStr str);
qreal calcExpecPauliStrSum(Qureg qureg, PauliStrSum sum);
qreal calcExpecFullStateDiagMatr(Qureg qureg, FullStateDiagMatr matr);
qreal calcExpecFullStateDiagMatrPower(Qureg qureg, FullStateDiagMatr matr, qreal exponent);
qreal calcProbOfBasisState(Qureg qureg, qindex index);
qreal calcProbOfQubitOutcome(Qureg qureg, int qubit, int outcome);
qreal calcProbOfMultiQubitOutcome(Qureg qureg, int* qubits, int* outcomes, int numQubits);
void calcProbsOfAllMultiQubitOutcomes(qreal* outcomeProbs, Qureg qureg, int* qubits, int numQubits);
qreal calcTotalProb(Qureg qureg);
qreal calcPurity(Qureg qureg);
qreal calcFidelity(Qureg qureg, Qureg other);
qreal calcDistance(Qureg qureg1, Qureg qureg2);
Qureg calcPartialTrace(Qureg qureg, int* traceOutQubits, int numTraceQubits);
Qureg calcReducedDensityMatrix(Qureg qureg, int* retainQubits, int numRetainQubits);
void _wrap_calcInnerProduct(Qureg bra, Qureg ket, qcomp* out);
void _wrap_calcExpecNonHermitianPauliStrSum(qcomp* out, Qureg qureg, PauliStrSum sum);
void _wrap_calcExpecNonHermitianFullStateDiagMatr(qcomp* out, Qureg qureg, FullStateDiagMatr matr);
/* calculations.cpp:98-115 */
void _wrap_calcExpecNonHermitianFullStateDiagMatrPower(qcomp* out, Qureg qureg, FullStateDiagMatr matr,
                                                       qcomp exponent);

#ifdef __cplusplus
}  /* extern "C" */
#endif

/* ---------------------------------- complex-returning functions (C++) */
#ifdef __cplusplus
qcomp getQuregAmp(Qureg qureg, qindex index);
qcomp getDensityQuregAmp(Qureg qureg, qindex row, qindex column);
qcomp calcInnerProduct(Qureg qureg1, Qureg qureg2);
qcomp calcExpecNonHermitianPauliStrSum(Qureg qureg, PauliStrSum sum);
qcomp calcExpecNonHermitianFullStateDiagMatr(Qureg qureg, FullStateDiagMatr matr);
qcomp calcExpecNonHermitianFullStateDiagMatrPower(Qureg qureg, FullStateDiagMatr matr, qcomp exponent);
#else
static inline qcomp getQuregAmp(Qureg qureg, qindex index) {
    qcomp out;
    _wrap_getQuregAmp(&out, qureg, index);
    return out;
}
static inline qcomp getDensityQuregAmp(Qureg qureg, qindex row, qindex column) {
    qcomp out;
    _wrap_getDensityQuregAmp(&out, qureg, row, column);
    return out;
}
static inline qcomp calcInnerProduct(Qureg bra, Qureg ket) {
    qcomp out;
    _wrap_calcInnerProduct(bra, ket, &out);
    return out;
}
static inline qcomp calcExpecNonHermitianPauliStrSum(Qureg qureg, PauliStrSum sum) {
    qcomp out;
    _wrap_calcExpecNonHermitianPauliStrSum(&out, qureg, sum);
    return out;
}
static inline qcomp calcExpecNonHermitianFullStateDiagMatr(Qureg qureg, FullStateDiagMatr matr) {
    qcomp out;
    _wrap_calcExpecNonHermitianFullStateDiagMatr(&out, qureg, matr);
    return out;
}
static inline qcomp calcExpecNonHermitianFullStateDiagMatrPower(Qureg qureg, FullStateDiagMatr matr,
                                                               qcomp exponent) {
    qcomp out;
    _wrap_calcExpecNonHermitianFullStateDiagMatrPower(&out, qureg, matr, exponent);
    return out;
}
#endif

/* -------------------------------- inline matrix constructors (header-only,
   quest/include/matrices.h:190-330) */
static inline CompMatr1 getCompMatr1(qcomp** in) {
    _validateNewNestedElemsPtrNotNull(in, 1, "getCompMatr1");
    CompMatr1 out;
    out.numQubits = 1;
    out.numRows = 2;
    for (int r = 0; r < 2; r++)
        for (int c = 0; c < 2; c++) out.elems[r][c] = in[r][c];
    return out;
}
static inline CompMatr2 getCompMatr2(qcomp** in) {
    _validateNewNestedElemsPtrNotNull(in, 2, "getCompMatr2");
    CompMatr2 out;
    out.numQubits = 2;
    out.numRows = 4;
    for (int r = 0; r < 4; r++)
        for (int c = 0; c < 4; c++) out.elems[r][c] = in[r][c];
    return out;
}
static inline DiagMatr1 getDiagMatr1(qcomp* in) {
    _validateNewElemsPtrNotNull(in, "getDiagMatr1");
    DiagMatr1 out;
    out.numQubits = 1;
    out.numElems = 2;
    for (int i = 0; i < 2; i++) out.elems[i] = in[i];
    return out;
}
static inline DiagMatr2 getDiagMatr2(qcomp* in) {
    _validateNewElemsPtrNotNull(in, "getDiagMatr2");
    DiagMatr2 out;
    out.numQubits = 2;
    out.numElems = 4;
    for (int i = 0; i < 4; i++) out.elems[i] = in[i];
    return out;
}

#ifdef __cplusplus
/* C++ overloads taking std containers (the C++ sections of quest/include) */
CompMatr1 getCompMatr1(std::vector<std::vector<qcomp>> in);
CompMatr2 getCompMatr2(std::vector<std::vector<qcomp>> in);
DiagMatr1 getDiagMatr1(std::vector<qcomp> in);
DiagMatr2 getDiagMatr2(std::vector<qcomp> in);
void setCompMatr(CompMatr out, std::vector<std::vector<qcomp>> in);
void setDiagMatr(DiagMatr out, std::vector<qcomp> in);
void setKrausMap(KrausMap map, std::vector<std::vector<std::vector<qcomp>>> matrices);
void setSuperOp(SuperOp op, std::vector<std::vector<qcomp>> matrix);
CompMatr createInlineCompMatr(int numQb, std::vector<std::vector<qcomp>> elems);
DiagMatr createInlineDiagMatr(int numQb, std::vector<qcomp> elems);
KrausMap createInlineKrausMap(int numQb, int numOps, std::vector<std::vector<std::vector<qcomp>>> matrices);
SuperOp createInlineSuperOp(int numQb, std::vector<std::vector<qcomp>> matrix);
void setInlineCompMatr(CompMatr matr, int numQb, std::vector<std::vector<qcomp>> in);
void setInlineDiagMatr(DiagMatr matr, int numQb, std::vector<qcomp> in);
void setInlineFullStateDiagMatr(FullStateDiagMatr matr, qindex startInd, qindex numElems, std::vector<qcomp> in);
void setInlineKrausMap(KrausMap map, int numQb, int numOps, std::vector<std::vector<std::vector<qcomp>>> matrices);
void setInlineSuperOp(SuperOp op, int numQb, std::vector<std::vector<qcomp>> matrix);
PauliStr getPauliStr(std::string paulis, int* indices, int numPaulis);
PauliStr getPauliStr(std::string paulis, std::vector<int> indices);
PauliStr getPauliStr(std::string paulis);
void setSeeds(std::vector<unsigned> seeds);
std::vector<unsigned> getSeeds();
void setQuregAmps(Qureg qureg, qindex startInd, std::vector<qcomp> amps);
std::vector<qcomp> getQuregAmps(Qureg qureg, qindex startInd, qindex numAmps);
void applyMultiControlledCompMatr1(Qureg qureg, std::vector<int> controls, int target, CompMatr1 matrix);
void applyMultiControlledCompMatr2(Qureg qureg, std::vector<int> controls, int target1, int target2,
                                   CompMatr2 matr);
void applyCompMatr(Qureg qureg, std::vector<int> targets, CompMatr matr);
void applyMultiQubitPhaseShift(Qureg qureg, std::vector<int> targets, qreal angle);
void applyPhaseGadget(Qureg qureg, std::vector<int> targets, qreal angle);
void applyQuantumFourierTransform(Qureg qureg, std::vector<int> targets);
void mixKrausMap(Qureg qureg, std::vector<int> targets, KrausMap map);
qreal calcProbOfMultiQubitOutcome(Qureg qureg, std::vector<int> qubits, std::vector<int> outcomes);
std::vector<qreal> calcProbsOfAllMultiQubitOutcomes(Qureg qureg, std::vector<int> qubits);
#define getInlineCompMatr1(...) getCompMatr1(__VA_ARGS__)
#define getInlineCompMatr2(...) getCompMatr2(__VA_ARGS__)
#define getInlineDiagMatr1(...) getDiagMatr1(__VA_ARGS__)
#define getInlineDiagMatr2(...) getDiagMatr2(__VA_ARGS__)
#define getInlinePauliStr(str, ...) getPauliStr(str, __VA_ARGS__)
#endif

#endif /* QUEST_H */
