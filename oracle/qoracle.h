/*
 * qoracle.h -- TEST INFRASTRUCTURE ONLY.  A plain-C (C11 + OpenMP) CPU
 * restatement of QuEST v4.1.0's statevector / density-matrix routines
 * (quest/src/cpu/cpu_subroutines.cpp), used as the checker of the HIP path by
 * tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg.  Nothing
 * in the product links or calls it.
 *
 * Amplitudes are interleaved complex128 ("double _Complex"); a density matrix
 * of N qubits is passed as 2N index qubits (column-major flattening: ket
 * qubit q = bit q, bra qubit q = bit q+N).  Qubit lists are int arrays;
 * `states` may be NULL (all controls on 1).
 */
#ifndef QORACLE_H
#define QORACLE_H

#include <complex.h>
#include <stdint.h>

typedef double _Complex qo_c;

/* statevector updates */
void qo_dense1(qo_c* a, int nq, const int* ctrls, const int* states, int nc, int targ, const qo_c* m4);
void qo_dense2(qo_c* a, int nq, const int* ctrls, const int* states, int nc, int t1, int t2, const qo_c* m16);
void qo_denseN(qo_c* a, int nq, const int* ctrls, const int* states, int nc, const int* targs, int nt,
               const qo_c* m, int conj);
void qo_diag1(qo_c* a, int nq, const int* ctrls, const int* states, int nc, int targ, const qo_c* d2);
void qo_diag2(qo_c* a, int nq, const int* ctrls, const int* states, int nc, int t1, int t2, const qo_c* d4);
void qo_diagN(qo_c* a, int nq, const int* ctrls, const int* states, int nc, const int* targs, int nt,
              const qo_c* d, int conj, int hasPower, double expRe, double expIm);
void qo_swap(qo_c* a, int nq, const int* ctrls, const int* states, int nc, int t1, int t2);
void qo_pauli(qo_c* a, int nq, const int* ctrls, const int* states, int nc, uint64_t mx, uint64_t my,
              uint64_t mz, double ampFacRe, double ampFacIm, double pairFacRe, double pairFacIm);
void qo_phaseGadget(qo_c* a, int nq, const int* ctrls, const int* states, int nc, uint64_t targMask,
                    double f0Re, double f0Im, double f1Re, double f1Im);
void qo_projector(qo_c* a, int nq, const int* qubits, const int* outcomes, int num, double prob);

/* density-matrix channels (N = number of qubits, 2N index bits) */
void qo_dm_dephase1(qo_c* a, int N, int q, double p);
void qo_dm_dephase2(qo_c* a, int N, int q1, int q2, double p);
void qo_dm_depol1(qo_c* a, int N, int q, double p);
void qo_dm_depol2(qo_c* a, int N, int q1, int q2, double p);
void qo_dm_pauliChannel(qo_c* a, int N, int q, double pI, double pX, double pY, double pZ);
void qo_dm_damp(qo_c* a, int N, int q, double p);
void qo_dm_projector(qo_c* a, int N, const int* qubits, const int* outcomes, int num, double prob);

/* reductions */
double qo_totalProb(const qo_c* a, int nq);
double qo_dm_totalProb(const qo_c* a, int N);
double qo_probOutcome(const qo_c* a, int nq, const int* qubits, const int* outcomes, int num);
double qo_dm_probOutcome(const qo_c* a, int N, const int* qubits, const int* outcomes, int num);
void qo_probsAllOutcomes(double* out, const qo_c* a, int nq, const int* qubits, int num);
double qo_expecZ(const qo_c* a, int nq, uint64_t mask);
void qo_expecPauli(double* out2, const qo_c* a, int nq, uint64_t mx, uint64_t my, uint64_t mz);
void qo_dm_expecPauli(double* out2, const qo_c* a, int N, uint64_t mx, uint64_t my, uint64_t mz);
void qo_innerProduct(double* out2, const qo_c* a, const qo_c* b, int nq);
void qo_dm_probsAllOutcomes(double* out, const qo_c* a, int nq, const int* qubits, int num);
double qo_dm_hsDistanceSq(const qo_c* a, const qo_c* b, int nq);
void qo_dm_fidelityPure(double* out2, const qo_c* rho, const qo_c* psi, int nq, int conjugate);

/* initialisation */
void qo_initDebug(qo_c* a, int nq);

#endif
