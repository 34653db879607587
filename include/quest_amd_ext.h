/*
 * quest_amd_ext.h -- extensions of the QuEST API that the MI355X backend
 * adds on top of quest.h.  None of them has a counterpart in the reference,
 * whose amplitude precision is one compile-time switch for the whole
 * library (quest/include/precision.h:88-96).
 */
#ifndef QUEST_AMD_EXT_H
#define QUEST_AMD_EXT_H

#include "quest.h"

#ifdef __cplusplus
extern "C" {
#endif

/* Amplitude precision of one qureg: 2 = complex128 (default), 1 = complex64.
   Gates on an FP32 qureg run in FP32; every reduction (probabilities,
   purity, expectations, norms) accumulates in FP64 with compensated
   summation.  The host mirror (cpuAmps) and every getter/setter keep using
   complex128 (qcomp), converting on the copy. */
Qureg createCustomPrecisionQureg(int numQubits, int isDensMatr, int precision);
int getQuregPrecision(Qureg qureg);

/* Bytes of device memory a qureg of this shape would take. */
qindex getQuregDeviceBytes(int numQubits, int isDensMatr, int precision);

/* Device-side handle of the library stream (a hipStream_t).  Gates may be
   queued for fusion: syncQuESTEnv() (or any amplitude read) applies them. */
void* getQuESTStream(void);

/* Gate fusion (no counterpart in the reference, which launches one kernel
   per gate).  On by default with 2^12-amplitude tiles whose low 4 qubits are
   always local and at most 512 queued gates; the environment variable
   QUEST_AMD_FUSION=0 turns it off at initQuESTEnv, "tile,low,queue" sets
   the shape.  Fused and unfused runs give identical amplitudes. */
void setQuESTFusion(int enabled, int tileQubits, int lowQubits, int maxQueue);
void getQuESTFusion(int* enabled, int* tileQubits, int* lowQubits, int* maxQueue);

#ifdef __cplusplus
}
#endif

#endif /* QUEST_AMD_EXT_H */
