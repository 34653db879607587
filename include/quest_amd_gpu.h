/*
 * quest_amd_gpu.h -- the C ABI of the MI355X (gfx950) kernel layer.
 *
 * This is the drop-in boundary for QuEST's GPU backend.  QuEST's host code
 * reaches its GPU backend only through the gpu_* functions declared in
 * quest/src/gpu/gpu_subroutines.hpp:24-196 (state updates and reductions)
 * and quest/src/gpu/gpu_config.hpp:41-120 (device, memory and copies),
 * called from quest/src/core/accelerator.cpp (e.g. :313-336, :345-365,
 * :513-532, :729-778, :857-1007, :1104-1143) and quest/src/api/qureg.cpp.
 * Each entry point below replaces one of those, with C linkage and plain
 * arguments only:
 *
 *   - a qureg is (amps, prec, numQubits): `amps` is a device pointer to
 *     2^numQubits interleaved {re,im} amplitudes, `prec` is QA_FP64
 *     (complex128) or QA_FP32 (complex64), and numQubits is the number of
 *     index bits (n for a statevector, 2N for an N-qubit density matrix,
 *     whose amplitude (row r, column c) sits at r + c*2^N);
 *   - qubit lists are (const int* list, int count) host arrays, controls
 *     come with an optional list of control states (NULL = all ones);
 *   - 1- and 2-qubit matrices travel by value as host arrays of doubles
 *     (row-major, re/im interleaved), larger matrices as device pointers to
 *     complex128 row-major arrays (the gpuElemsFlat of QuEST's CompMatr,
 *     DiagMatr and SuperOp);
 *   - every function returns QA_OK (0) or an error code; qa_lastError()
 *     gives the message of the last failure on the calling thread;
 *   - updates are enqueued on the library's stream and return at once;
 *     reductions write FP64 results (also for FP32 amplitudes) into host
 *     memory and therefore synchronise that stream.
 *
 * Nothing here allocates device memory per call: scratch for reductions
 * and staging is owned by the library and created by qa_init().
 */
#ifndef QUEST_AMD_GPU_H
#define QUEST_AMD_GPU_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define QA_FP32 1
#define QA_FP64 2

#define QA_OK            0
#define QA_ERR_ARG       1
#define QA_ERR_HIP       2
#define QA_ERR_OOM       3
#define QA_ERR_NOT_INIT  4
#define QA_ERR_NO_DEVICE 5

/* ------------------------------------------------------------------ */
/* device, memory, copies        (gpu_config.hpp:41-120)              */
/* ------------------------------------------------------------------ */

const char* qa_lastError(void);

/* gpu_isGpuCompiled / gpu_isGpuAvailable (gpu_config.hpp:41-45) */
int qa_isGpuAvailable(void);
/* gpu_bindLocalGPUsToNodes (gpu_config.cpp:332-353): selects the device,
   creates the library stream and the reduction scratch. Idempotent. */
int qa_init(int device);
int qa_finalize(void);
/* a counter that changes whenever an entry point that may write amplitudes
   runs (no counterpart in QuEST): results computed at equal values still
   hold; a host layer keys caches of reductions on it */
uint64_t qa_stateEpoch(void);
/* gpu_getCurrentAvailableMemoryInBytes / gpu_getTotalMemoryInBytes */
int qa_getMemInfo(size_t* freeBytes, size_t* totalBytes);
/* number of compute units and device name, for reporting */
int qa_getDeviceInfo(int* numCUs, char* name, int nameLen);
/* gpu_allocArray / gpu_deallocArray (gpu_config.cpp:399-433): returns
   QA_ERR_OOM (and *ptr = NULL) when the allocation does not fit */
int qa_malloc(void** ptr, size_t numBytes);
int qa_free(void* ptr);
/* gpu_copyCpuToGpu / gpu_copyGpuToCpu / gpu_copyArray: synchronous */
int qa_memcpyHtoD(void* dst, const void* src, size_t numBytes);
int qa_memcpyDtoH(void* dst, const void* src, size_t numBytes);
int qa_memcpyDtoD(void* dst, const void* src, size_t numBytes);
/* gpu_sync (gpu_config.cpp:382-390) */
int qa_sync(void);
/* the hipStream_t every entry point enqueues on */
void* qa_getStream(void);

/* amplitude transfer with conversion to/from complex128 host arrays
   (accel_statevec_getAmps_sub / setAmps_sub, accelerator.cpp:203-230) */
int qa_getAmps(double* outComplex128, const void* amps, int prec, int64_t startInd, int64_t numAmps);
int qa_setAmps(void* amps, int prec, int64_t startInd, const double* inComplex128, int64_t numAmps);
/* strided 2D block of a density matrix: rows [r0, r0+nr) x cols [c0, c0+nc),
   written row-major into out (numRows x numCols) */
int qa_getDensityAmps(double* outComplex128, const void* amps, int prec, int numQubits,
                      int64_t startRow, int64_t startCol, int64_t numRows, int64_t numCols);
/* copy between quregs of possibly different precision */
int qa_copyAmps(void* dst, int dstPrec, const void* src, int srcPrec, int64_t numAmps);

/* event timing on the library stream (for benchmarks) */
int qa_timerStart(void);
int qa_timerStop(double* elapsedMs);
int qa_eventCreate(void** event);
int qa_eventRecord(void* event);
/* waits for `stop`, then gives the milliseconds between the two records */
int qa_eventElapsedMs(void* start, void* stop, double* elapsedMs);
int qa_eventDestroy(void* event);

/* ------------------------------------------------------------------ */
/* gate fusion (no counterpart in QuEST, which runs one kernel per    */
/* gate: gpu_subroutines.cpp:197-887)                                 */
/* ------------------------------------------------------------------ */
/* When enabled, the one/two-qubit dense, one-qubit diagonal, swap, Pauli
   tensor/gadget, phase gadget, one-qubit dephasing and one-qubit
   depolarising entry points queue their gate instead of launching it.
   Queued gates are applied by tiled passes over the state: each workgroup holds a
   tile of 2^tileQubits amplitudes (whose low `lowQubits` qubits are always
   in the tile) and applies runs of gates whose dense targets lie in 4 of
   the tile's qubits in registers.  A pass takes the queued gates in order
   while their dense targets fit one tile.  Unless fusion is exact
   (qa_setFusionExact), it also takes later gates that commute with the
   ones passed over, and multiplies (FP64, on the host) each run of its
   gates confined to the same one or two qubits into one matrix, so results
   differ from the unfused kernels by rounding only; exact fusion gives
   bit-identical results.  A gate is queued when its
   dense targets number at most 4, or it is a Pauli tensor/gadget whose X/Y
   qubits, with the low qubits, fit one tile; every other entry point
   (reductions, copies, getters, setters, sync, event records) first applies
   what is queued.  Queues are per state, states of fewer than 4 qubits are
   never queued, and a queue is flushed after `maxQueue` gates.  Changing
   the configuration flushes every queue.  tileQubits in [4,13], lowQubits
   in [0, tileQubits-2], maxQueue in [1,4096]; fusion is off after qa_init. */
int qa_setFusion(int enabled, int tileQubits, int lowQubits, int maxQueue);
int qa_getFusion(int* enabled, int* tileQubits, int* lowQubits, int* maxQueue);
/* applies the gates queued on `amps` (NULL: on every state) */
int qa_flush(void* amps);
/* complex64 states only: within a pass, every run of at least `minGates`
   queued gates whose qubits (targets, controls, masks) fit 5 of the tile's
   qubits is multiplied on the host (FP64) into one 32x32 complex block and
   applied with FP32 MFMA instead of gate by gate (tiles of at least 2^10
   amplitudes).  Results then differ from the unfused kernels by rounding.
   0 (the default) turns blocks off. */
int qa_setFusionBlocks(int minGates);
/* how tile passes prefetch each workgroup's next tile into registers while
   it works on the current one: 2 (the default) with the loads spread over
   the pass's generated program in four groups, 3 as 2 with each tile's
   stores deferred into the next tile's work (two register sets taking
   turns), 1 all at the start of the tile, 0 not at all (load, work,
   store).  Passes that run in the interpreter treat 2 and 3 as 1.  Other
   values: QA_ERR_ARG. */
int qa_setFusionPrefetch(int mode);
/* exact = 1: passes take queued gates in queue order and apply each with
   the arithmetic of its single-gate kernel.  exact = 0 (the default):
   passes may take gates out of order -- two gates are reordered only when
   neither involves a dense target of the other (on shared qubits both are
   then diagonal or controls), which makes them commute exactly -- and runs
   of gates on the same one or two qubits become one matrix. */
int qa_setFusionExact(int exact);
/* counters of tile passes launched, the gates they applied and the real
   flops of the ops they executed (gates after merging; a complex multiply
   counts 6, a complex multiply-add 8) since the last qa_getFusionStats
   or qa_setFusionTiming (which reset them); with
   timing on, each pass launch is bracketed by HIP events on the library
   stream and passMs sums their elapsed times (qa_getFusionStats waits for
   them) */
/* Pass kernels generated at run time: a pass program's shape (op kinds,
   register bits and masks, group layouts, which conditions are present) is
   compiled with hiprtc (in a qa_jitc child process, beside the library or
   at $QA_JITC; with $QA_JIT_CACHE_DIR set, code objects are also kept there
   for later processes) into a straight-line kernel, cached per shape, with
   the interpreter kernel's arithmetic bit for bit.  mode 0: off (every pass
   in the interpreter); 1 (default): compiles on background threads once a
   shape has been planned minSeen times (default 2), passes of a shape
   without a kernel yet run in the interpreter; 2: compile at first use and
   wait.  A shape whose compile fails stays interpreted
   (qa_fusionJitLastError). */
int qa_setFusionJit(int mode, int minSeen);
int qa_getFusionJit(int* mode, int* minSeen);
/* waits until every queued compile has finished */
int qa_fusionJitWait(void);
/* kernels compiled, compiles failed, passes run by generated kernels, and the
   compile threads' total seconds */
int qa_getFusionJitStats(int64_t* kernels, int64_t* failures, int64_t* launches, double* compileSeconds);
const char* qa_fusionJitLastError(void);
int qa_setFusionTiming(int enabled);
int qa_getFusionStats(int64_t* passes, int64_t* gates, double* passMs, double* flops);
/* program size of the tile passes since the last call (which resets them):
   ops the passes executed (gates after merging, plus layout changes) and,
   among them, register-layout changes through LDS */
int qa_getFusionProgramStats(int64_t* ops, int64_t* layouts);
/* span passes over cosets of 2^spanQubits amplitudes (10..13; 12, the
   default: two workgroups per CU overlap one's LDS work with the other's
   HBM traffic) or none (0).  A run of uncontrolled Pauli gadgets, Pauli
   tensors and phase gadgets at the front of the queue whose X/Y masks, with
   the always-local low qubits, span at most spanQubits dimensions over
   GF(2) may be applied as one span pass: each workgroup loads one coset of
   that subspace into LDS and applies the run in queue order with the
   single-gate arithmetic, so the result is bit-identical to the gates one
   by one.  The flush takes it instead of the next tile pass when it holds
   more gates.  (No counterpart in the reference, which applies each gadget
   as its own kernel: quest/src/gpu/gpu_subroutines.cpp:780 and :865.) */
int qa_setFusionSpan(int spanQubits);
/* Fusion autotuner (no counterpart in the reference, which has no fusion).
   Queue a representative batch of gates for the state `amps` (one
   iteration of the circuit, with maxQueue large enough to hold it), then
   call: the batch is applied to a scratch copy of the state under
   candidate fusion shapes -- by coordinate descent from the current
   configuration over tileQubits (13, 12, 11), lowQubits (2, 3, 4),
   spanQubits (13, 12, 0) and, for complex64 states, the MFMA block
   threshold (0, 4) -- once to warm (generated kernels are compiled and
   waited for) and once timed with HIP events; the fastest becomes the
   fusion configuration (maxQueue is kept) and the queued batch is then
   applied to `amps` with it.  Choices are cached per device, precision,
   qubit count and gate mix (tenths of the batch per gate kind); a cached
   choice is applied without timing unless flags & 1.  trials[0..maxTrials)
   receives the shapes tried, in order, with the batch's milliseconds (-1
   where a queued gate does not fit the shape), *numTrials their count and
   *best the index of the one chosen.  Needs device memory for one more
   copy of the state; holds the fusion lock (other threads' gates wait)
   while it runs. */
typedef struct {
    int tileQubits, lowQubits, spanQubits, blockMinGates;
    double ms;
} qa_fusionTrial;
int qa_tuneFusion(void* amps, int flags, qa_fusionTrial* trials, int maxTrials, int* numTrials, int* best);
/* span passes among the passes, the gates they applied and (with timing
   on) their summed elapsed milliseconds, since the last call (which resets
   them; qa_getFusionStats counts them too) */
int qa_getFusionSpanStats(int64_t* passes, int64_t* gates, double* passMs);

/* ------------------------------------------------------------------ */
/* statevector gates             (gpu_subroutines.hpp:33-70)          */
/* ------------------------------------------------------------------ */

/* gpu_statevec_anyCtrlSwap_subA */
int qa_statevec_anyCtrlSwap(void* amps, int prec, int numQubits,
                            const int* ctrls, const int* ctrlStates, int numCtrls,
                            int targ1, int targ2);

/* gpu_statevec_anyCtrlOneTargDenseMatr_subA; matr = 8 doubles */
int qa_statevec_anyCtrlOneTargDenseMatr(void* amps, int prec, int numQubits,
                                        const int* ctrls, const int* ctrlStates, int numCtrls,
                                        int targ, const double* matr);

/* gpu_statevec_anyCtrlTwoTargDenseMatr_sub; matr = 32 doubles; matrix
   index bit 0 <-> targ1, bit 1 <-> targ2 */
int qa_statevec_anyCtrlTwoTargDenseMatr(void* amps, int prec, int numQubits,
                                        const int* ctrls, const int* ctrlStates, int numCtrls,
                                        int targ1, int targ2, const double* matr);

/* gpu_statevec_anyCtrlAnyTargDenseMatr_sub; devMatr = device complex128
   2^k x 2^k row-major; matrix index bit i <-> targs[i] (unsorted) */
int qa_statevec_anyCtrlAnyTargDenseMatr(void* amps, int prec, int numQubits,
                                        const int* ctrls, const int* ctrlStates, int numCtrls,
                                        const int* targs, int numTargs,
                                        const void* devMatr, int applyConj);

/* gpu_statevec_anyCtrlOneTargDiagMatr_sub; elems = 4 doubles */
int qa_statevec_anyCtrlOneTargDiagMatr(void* amps, int prec, int numQubits,
                                       const int* ctrls, const int* ctrlStates, int numCtrls,
                                       int targ, const double* elems);

/* gpu_statevec_anyCtrlTwoTargDiagMatr_sub; elems = 8 doubles */
int qa_statevec_anyCtrlTwoTargDiagMatr(void* amps, int prec, int numQubits,
                                       const int* ctrls, const int* ctrlStates, int numCtrls,
                                       int targ1, int targ2, const double* elems);

/* gpu_statevec_anyCtrlAnyTargDiagMatr_sub; devElems = device complex128[2^k];
   element index bit i <-> targs[i]; optional complex power and conjugation */
int qa_statevec_anyCtrlAnyTargDiagMatr(void* amps, int prec, int numQubits,
                                       const int* ctrls, const int* ctrlStates, int numCtrls,
                                       const int* targs, int numTargs,
                                       const void* devElems, int applyConj,
                                       int hasPower, double expRe, double expIm);

/* gpu_statevec_allTargDiagMatr_sub / gpu_densmatr_allTargDiagMatr_sub */
int qa_statevec_allTargDiagMatr(void* amps, int prec, int numQubits, const void* devElems,
                                int hasPower, double expRe, double expIm);
int qa_densmatr_allTargDiagMatr(void* amps, int prec, int numQubitsN, const void* devElems,
                                int hasPower, double expRe, double expIm, int multiplyOnly);

/* gpu_statevector_anyCtrlPauliTensorOrGadget_subA: for every pair
   (iA, iB = iA ^ maskXY) a_A <- ampFac a_A + pairFac i^#Y (+-) a_B with the
   sign of (pair index & maskYZ); requires maskX|maskY != 0 */
int qa_statevec_anyCtrlPauliTensorOrGadget(void* amps, int prec, int numQubits,
                                           const int* ctrls, const int* ctrlStates, int numCtrls,
                                           uint64_t maskX, uint64_t maskY, uint64_t maskZ,
                                           double ampFacRe, double ampFacIm,
                                           double pairFacRe, double pairFacIm);

/* gpu_statevector_anyCtrlAnyTargZOrPhaseGadget_sub:
   a_i <- a_i * (parity(i & targMask) ? fac1 : fac0) */
int qa_statevec_anyCtrlAnyTargZOrPhaseGadget(void* amps, int prec, int numQubits,
                                             const int* ctrls, const int* ctrlStates, int numCtrls,
                                             uint64_t targMask,
                                             double fac0Re, double fac0Im,
                                             double fac1Re, double fac1Im);

/* gpu_statevec_setQuregToSuperposition_sub: out <- fOut out + f1 in1 + f2 in2 */
int qa_statevec_setQuregToSuperposition(int prec, int numQubits,
                                        double facOutRe, double facOutIm, void* out,
                                        double fac1Re, double fac1Im, const void* in1,
                                        double fac2Re, double fac2Im, const void* in2);

/* ------------------------------------------------------------------ */
/* density-matrix channels       (gpu_subroutines.hpp:75-112)         */
/* numQubitsN is N (the matrix is 2^N x 2^N, 2N index bits)           */
/* ------------------------------------------------------------------ */

int qa_densmatr_mixQureg(double outProb, void* outDM, double inProb, const void* inDM,
                         int prec, int numQubitsN);
int qa_densmatr_mixQuregWithStatevec(double outProb, void* outDM, double inProb, const void* inSV,
                                     int prec, int numQubitsN);
int qa_densmatr_oneQubitDephasing(void* amps, int prec, int numQubitsN, int qubit, double prob);
int qa_densmatr_twoQubitDephasing(void* amps, int prec, int numQubitsN, int qubit1, int qubit2, double prob);
int qa_densmatr_oneQubitDepolarising(void* amps, int prec, int numQubitsN, int qubit, double prob);
int qa_densmatr_twoQubitDepolarising(void* amps, int prec, int numQubitsN, int qubit1, int qubit2, double prob);
int qa_densmatr_oneQubitPauliChannel(void* amps, int prec, int numQubitsN, int qubit,
                                     double pI, double pX, double pY, double pZ);
int qa_densmatr_oneQubitDamping(void* amps, int prec, int numQubitsN, int qubit, double prob);
/* gpu_densmatr_partialTrace_sub: out (numQubitsN - numTargs qubits) */
int qa_densmatr_partialTrace(const void* inAmps, int prec, int inNumQubitsN, void* outAmps,
                             const int* targs, int numTargs);

/* ------------------------------------------------------------------ */
/* reductions (FP64 results)     (gpu_subroutines.hpp:136-183)        */
/* ------------------------------------------------------------------ */

int qa_statevec_calcTotalProb(const void* amps, int prec, int numQubits, double* out);
int qa_densmatr_calcTotalProb(const void* amps, int prec, int numQubitsN, double* out);
/* every qubit's outcome probabilities in one read of the state per 10 qubits
   (gpu_statevec_calcProbOfMultiQubitOutcome_sub for each single qubit, as
   calcProbOfQubitOutcome asks it): out[2q] = P(q = 0), out[2q+1] = P(q = 1),
   each the compensated sum of its own amplitudes' |a|^2 */
int qa_statevec_calcProbsOfEveryQubit(const void* amps, int prec, int numQubits, double* out);
int qa_statevec_calcProbOfMultiQubitOutcome(const void* amps, int prec, int numQubits,
                                            const int* qubits, const int* outcomes, int num,
                                            double* out);
int qa_densmatr_calcProbOfMultiQubitOutcome(const void* amps, int prec, int numQubitsN,
                                            const int* qubits, const int* outcomes, int num,
                                            double* out);
int qa_statevec_calcProbsOfAllMultiQubitOutcomes(double* outProbs, const void* amps, int prec,
                                                 int numQubits, const int* qubits, int num);
int qa_densmatr_calcProbsOfAllMultiQubitOutcomes(double* outProbs, const void* amps, int prec,
                                                 int numQubitsN, const int* qubits, int num);
/* No counterpart in the reference: for callers that cannot allocate (code
   under HIP graph capture, tight serving loops, harnesses that forbid
   hipMalloc).  The statevector outcome table of
   qa_statevec_calcProbsOfAllMultiQubitOutcomes, with the same kernels and
   launch shape (so the same sums), written into the caller's DEVICE array
   devOutProbs (2^num doubles) on the library stream; returns without
   synchronising (qa_sync, or an event on qa_getStream(), orders reads).
   It allocates nothing: its partial sums live in a 16 MiB device array that
   each device holds for the life of the process; tables of 1..24 qubits
   fit, anything larger is rejected with QA_ERR_ARG.  Every call on a device
   shares that array, and every partial is written before it is read, so
   nothing needs clearing between calls; calls whose kernels overlap on one
   device (the same scratch in flight on two streams) race.  Gates still
   queued for this state are an error (a flush may allocate): call
   qa_flush(amps) first. */
int qa_statevec_calcProbsOfAllMultiQubitOutcomesNoAlloc(double* devOutProbs, const void* amps, int prec,
                                                        int numQubits, const int* qubits, int num);
/* out[2] = sum conj(a_i) b_i */
int qa_statevec_calcInnerProduct(const void* ampsA, const void* ampsB, int prec, int numQubits,
                                 double* out);
int qa_densmatr_calcHilbertSchmidtDistance(const void* ampsA, const void* ampsB, int prec,
                                           int numQubitsN, double* out);
/* out[2] = sum_rc rho_rc conj(psi_r) psi_c (conj=0) or its conjugate form (conj=1) */
int qa_densmatr_calcFidelityWithPureState(const void* rho, const void* psi, int prec,
                                          int numQubitsN, int conj, double* out);
int qa_statevec_calcExpecAnyTargZ(const void* amps, int prec, int numQubits, uint64_t targMask,
                                  double* out);
/* out[2] */
int qa_densmatr_calcExpecAnyTargZ(const void* amps, int prec, int numQubitsN, uint64_t targMask,
                                  double* out);
/* out[2], includes the i^#Y factor */
int qa_statevec_calcExpecPauliStr(const void* amps, int prec, int numQubits,
                                  uint64_t maskX, uint64_t maskY, uint64_t maskZ, double* out);
int qa_densmatr_calcExpecPauliStr(const void* amps, int prec, int numQubitsN,
                                  uint64_t maskX, uint64_t maskY, uint64_t maskZ, double* out);
/* localiser_statevec_calcExpecPauliStrSum (localiser.cpp:2052-2126):
   out[2] = sum_t coeffs[t] <P_t> over numTerms Pauli strings given as X, Y, Z
   bit masks, coeffs complex (re, im interleaved).  Strings that share their
   X/Y mask mix the same amplitude pairs; up to 8 of them are reduced in one
   read of the state, each term's value bitwise equal to
   qa_statevec_calcExpecPauliStr's.  The weighted sum is taken in term
   order.  termValues (2 doubles per term, may be NULL) receives every <P_t>. */
int qa_statevec_calcExpecPauliStrSum(const void* amps, int prec, int numQubits, int64_t numTerms,
                                     const uint64_t* maskX, const uint64_t* maskY, const uint64_t* maskZ,
                                     const double* coeffs, double* termValues, double* out);
/* out[2]; powMode 0 = none, 1 = complex power, 2 = real power of the real part */
int qa_statevec_calcExpecFullStateDiagMatr(const void* amps, int prec, int numQubits,
                                           const void* devElems, int powMode,
                                           double expRe, double expIm, double* out);
int qa_densmatr_calcExpecFullStateDiagMatr(const void* amps, int prec, int numQubitsN,
                                           const void* devElems, int powMode,
                                           double expRe, double expIm, double* out);

/* ------------------------------------------------------------------ */
/* projectors and initialisation (gpu_subroutines.hpp:185-196)        */
/* ------------------------------------------------------------------ */

int qa_statevec_multiQubitProjector(void* amps, int prec, int numQubits,
                                    const int* qubits, const int* outcomes, int num, double prob);
int qa_densmatr_multiQubitProjector(void* amps, int prec, int numQubitsN,
                                    const int* qubits, const int* outcomes, int num, double prob);
int qa_statevec_initUniformState(void* amps, int prec, int numQubits, double re, double im);
/* gpu_densmatr_setAmpsToPauliStrSum_sub (cpu_subroutines.cpp:63-92,
   fastmath.hpp:124-196): rho_rc <- sum_j c_j <r|sigma_j|c>, terms as X, Y,
   Z masks (host arrays) and complex coefficients {re, im} (host) */
int qa_densmatr_setAmpsToPauliStrSum(void* amps, int prec, int numQubitsN, int64_t numTerms,
                                     const uint64_t* maskX, const uint64_t* maskY, const uint64_t* maskZ,
                                     const double* coeffs);
/* gpu_fullstatediagmatr_setElemsToPauliStrSum (cpu_subroutines.cpp:95-116):
   devElems = device complex128[2^numQubits] <- the diagonal of a sum of
   Z strings (masks and coefficients on the host) */
int qa_fullstatediagmatr_setElemsToPauliStrSum(void* devElems, int numQubits, int64_t numTerms,
                                               const uint64_t* maskZ, const double* coeffs);
int qa_statevec_initDebugState(void* amps, int prec, int numQubits);
/* unnormalised Haar-distributed amplitudes from a counter-based generator
   keyed by (seed, amplitude index): reproducible for a given seed,
   independent of launch shape */
int qa_statevec_initUnnormalisedUniformlyRandomPureStateAmps(void* amps, int prec, int numQubits,
                                                             uint64_t seed);
/* rho <- |psi><psi| */
int qa_densmatr_initPureState(void* rho, const void* psi, int prec, int numQubitsN);

#ifdef __cplusplus
}
#endif

#endif /* QUEST_AMD_GPU_H */
