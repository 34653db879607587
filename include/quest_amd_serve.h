/*
 * quest_amd_serve.h -- batched circuit serving on one MI355X.
 *
 * No counterpart in QuEST: the reference runs one Qureg at a time, one
 * kernel per gate (quest/src/gpu/gpu_subroutines.cpp:197-887), and has no
 * request loop.  This engine keeps many small statevectors resident in one
 * HBM arena and advances all of them together:
 *
 *   - a request is a circuit (one- and two-qubit gates on n qubits, 4 <= n
 *     <= 30) plus Pauli strings whose expectation values it returns;
 *   - requests are admitted first come, first served while their state fits
 *     the arena (power-of-two blocks of a buddy allocator) and fewer than
 *     `maxActive` are running; a request that arrives while others run is
 *     admitted as soon as a slot and memory free up (continuous batching);
 *   - every engine step launches one batched tile pass that advances every
 *     running request by one run of gates (all gates of a state of at most
 *     tileQubits qubits; otherwise the gates whose targets fit one tile,
 *     later gates joining when they commute with those passed over, and
 *     runs of gates on the same one or two qubits multiplied into one
 *     matrix),
 *     then one batched reduction for the requests whose gates are done;
 *   - expectations accumulate in FP64 with compensated summation, in an
 *     order fixed by the request alone, so a request's results do not depend
 *     on what else is running;
 *   - amplitudes are complex64 (precision 1) or complex128 (precision 2).
 *
 * All functions return QA_OK (0) or an error code (quest_amd_gpu.h), and
 * qa_lastError() describes the last failure of the calling thread.  The
 * engine runs on its own host thread and HIP stream; submit, poll and wait
 * may be called from any thread.
 */
#ifndef QUEST_AMD_SERVE_H
#define QUEST_AMD_SERVE_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* gate kinds of a served circuit */
#define QA_SG_MATR1 1 /* dense 2x2 on t1: m[0..7] row-major {re,im} */
#define QA_SG_H     2 /* Hadamard on t1 */
#define QA_SG_RX    3 /* exp(-i angle/2 X) on t1 */
#define QA_SG_RY    4 /* exp(-i angle/2 Y) on t1 */
#define QA_SG_RZ    5 /* exp(-i angle/2 Z) on t1 */
#define QA_SG_CZ    6 /* controlled Z on (t1, t2) */
#define QA_SG_CX    7 /* controlled X, control t1, target t2 */
#define QA_SG_PHASE 8 /* diag(1, e^{i angle}) on t1 */

typedef struct {
    int kind;
    int t1, t2;
    double angle;
    double m[8];
} qa_serveGate;

typedef struct {
    int numQubits;
    int numGates;
    const qa_serveGate* gates;
    /* Pauli strings as bit masks over the qubits, one (X, Y, Z) triple each */
    int numPaulis;
    const uint64_t* pauliX;
    const uint64_t* pauliY;
    const uint64_t* pauliZ;
} qa_serveCircuit;

/* Creates an engine on the current device: an arena of `arenaBytes` of HBM,
   at most `maxActive` requests running at once, tiles of 2^tileQubits
   amplitudes (4..13) whose `lowQubits` lowest qubits are always in the
   tile, amplitudes of `precision` (1 = complex64, 2 = complex128). */
int qa_serveCreate(void** engine, size_t arenaBytes, int maxActive, int tileQubits, int lowQubits, int precision);
/* Finishes every submitted request, stops the engine and frees its memory. */
int qa_serveDestroy(void* engine);
/* Queues a request (the circuit is copied); its id is written to *id. */
int qa_serveSubmit(void* engine, const qa_serveCircuit* circuit, int64_t* id);
/* *done = 1 once the request's results are available. */
int qa_servePoll(void* engine, int64_t id, int* done);
/* Waits for a request and writes its numPaulis expectation values and its
   latency (submission to results, milliseconds); the request is then
   forgotten by the engine. */
int qa_serveWait(void* engine, int64_t id, double* expectations, double* latencyMs);
/* Engine counters since creation: steps, tile passes, states admitted,
   requests completed, and the largest number running at once. */
int qa_serveStats(void* engine, int64_t* steps, int64_t* passes, int64_t* admitted, int64_t* completed,
                  int64_t* maxRunning);
/* Milliseconds the engine thread has spent planning and launching steps,
   and waiting for steps to finish on the GPU. */
int qa_serveTimes(void* engine, double* planMs, double* waitMs);
/* enabled = 1: each step's copies and launches (program upload, one batched
   tile pass per tile size, the expectation reduction and its merge, the
   result copy) are captured into a HIP graph and launched as one graph;
   each of the engine's step slots keeps one executable graph per launch
   signature (upload or not, number of pass launches, reduction or not) and
   updates it in place (hipGraphExecUpdate) with the new step's sizes and
   arguments, instantiating anew when the update is refused.  0 (the
   default): launched one by one.
   Takes effect from the next step; results are the same either way. */
int qa_serveSetGraphs(void* engine, int enabled);
/* Graph launches, instantiations and in-place updates since creation. */
int qa_serveGraphStats(void* engine, int64_t* launches, int64_t* instantiations, int64_t* updates);

#ifdef __cplusplus
}
#endif

#endif /* QUEST_AMD_SERVE_H */
