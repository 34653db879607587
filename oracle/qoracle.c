/*
 * qoracle.c -- TEST INFRASTRUCTURE ONLY (see qoracle.h).  CPU restatement of
 * QuEST v4.1.0's OpenMP routines; each function names the reference lines it
 * follows.  It is the checker of the HIP path: only tests/, smoke() and
 * bench.py's cpu_baseline leg load it.
 *
 * Parity pins: tests/test_oracle.py checks every function here against (a)
 * fixtures produced by the reference's own CPU build (oracle/_ref, driven by
 * oracle/circuit_driver.cpp), (b) dense-matrix brute force at small sizes and
 * (c) closed forms (debug state, DFT, channel limits).
 */
#include "qoracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>

/* ---------------------------------------------------------------- bits */
/* quest/src/core/bitwise.hpp:34-254 */

typedef struct {
    int num;
    int pos[64];
    uint64_t val;
} ins_t;

static void ins_add(ins_t* s, int q, int v) {
    /* keep ascending order */
    int i = s->num++;
    while (i > 0 && s->pos[i - 1] > q) {
        s->pos[i] = s->pos[i - 1];
        i--;
    }
    s->pos[i] = q;
    if (v) s->val |= 1ull << q;
}

static ins_t ins_make(const int* ctrls, const int* states, int nc) {
    ins_t s;
    s.num = 0;
    s.val = 0;
    for (int i = 0; i < nc; i++) ins_add(&s, ctrls[i], states ? states[i] : 1);
    return s;
}

/* insertBitsWithMaskedValues (bitwise.hpp) */
static inline uint64_t ins_apply(uint64_t n, const ins_t* s) {
    for (int j = 0; j < s->num; j++) {
        int q = s->pos[j];
        uint64_t lo = n & ((1ull << q) - 1ull);
        n = ((n >> q) << (q + 1)) | lo;
    }
    return n | s->val;
}

static inline int parity(uint64_t x) { return __builtin_parityll(x); }

/* setBits over an unsorted list: bit i of v -> position t[i] */
static inline uint64_t set_bits(uint64_t n, const int* t, int nt, uint64_t v) {
    for (int i = 0; i < nt; i++) {
        uint64_t b = 1ull << t[i];
        n = ((v >> i) & 1ull) ? (n | b) : (n & ~b);
    }
    return n;
}

static inline uint64_t value_of_bits(uint64_t n, const int* t, int nt) {
    uint64_t v = 0;
    for (int i = 0; i < nt; i++) v |= ((n >> t[i]) & 1ull) << i;
    return v;
}

/* ------------------------------------------------------- statevector gates */

/* cpu_subroutines.cpp: cpu_statevec_anyCtrlOneTargDenseMatr_subA (:362-391) */
void qo_dense1(qo_c* a, int nq, const int* ctrls, const int* states, int nc, int targ, const qo_c* m) {
    ins_t s = ins_make(ctrls, states, nc);
    ins_add(&s, targ, 0);
    long long numIts = 1ll << (nq - nc - 1);
    uint64_t tb = 1ull << targ;
#pragma omp parallel for
    for (long long n = 0; n < numIts; n++) {
        uint64_t i0 = ins_apply((uint64_t)n, &s), i1 = i0 | tb;
        qo_c a0 = a[i0], a1 = a[i1];
        a[i0] = m[0] * a0 + m[1] * a1;
        a[i1] = m[2] * a0 + m[3] * a1;
    }
}

/* cpu_statevec_anyCtrlTwoTargDenseMatr_sub (:435-480): index bit0 <-> t1 */
void qo_dense2(qo_c* a, int nq, const int* ctrls, const int* states, int nc, int t1, int t2, const qo_c* m) {
    ins_t s = ins_make(ctrls, states, nc);
    ins_add(&s, t1, 0);
    ins_add(&s, t2, 0);
    long long numIts = 1ll << (nq - nc - 2);
    uint64_t b1 = 1ull << t1, b2 = 1ull << t2;
#pragma omp parallel for
    for (long long n = 0; n < numIts; n++) {
        uint64_t i00 = ins_apply((uint64_t)n, &s);
        uint64_t idx[4] = {i00, i00 | b1, i00 | b2, i00 | b1 | b2};
        qo_c v[4];
        for (int k = 0; k < 4; k++) v[k] = a[idx[k]];
        for (int r = 0; r < 4; r++)
            a[idx[r]] = m[4 * r] * v[0] + m[4 * r + 1] * v[1] + m[4 * r + 2] * v[2] + m[4 * r + 3] * v[3];
    }
}

/* cpu_statevec_anyCtrlAnyTargDenseMatr_sub (:483-578): matrix index bit i <-> targs[i] */
void qo_denseN(qo_c* a, int nq, const int* ctrls, const int* states, int nc, const int* targs, int nt,
               const qo_c* m, int applyConj) {
    ins_t s = ins_make(ctrls, states, nc);
    for (int i = 0; i < nt; i++) ins_add(&s, targs[i], 0);
    long long numIts = 1ll << (nq - nc - nt);
    long long dim = 1ll << nt;
#pragma omp parallel
    {
        qo_c* cache = (qo_c*)malloc(sizeof(qo_c) * (size_t)dim);
#pragma omp for
        for (long long n = 0; n < numIts; n++) {
            uint64_t i0 = ins_apply((uint64_t)n, &s);
            for (long long j = 0; j < dim; j++) cache[j] = a[set_bits(i0, targs, nt, (uint64_t)j)];
            for (long long k = 0; k < dim; k++) {
                qo_c acc = 0;
                for (long long j = 0; j < dim; j++) {
                    qo_c e = m[k * dim + j];
                    if (applyConj) e = conj(e);
                    acc += e * cache[j];
                }
                a[set_bits(i0, targs, nt, (uint64_t)k)] = acc;
            }
        }
        free(cache);
    }
}

/* cpu_statevec_anyCtrlOneTargDiagMatr_sub (:582-617) */
void qo_diag1(qo_c* a, int nq, const int* ctrls, const int* states, int nc, int targ, const qo_c* d) {
    ins_t s = ins_make(ctrls, states, nc);
    long long numIts = 1ll << (nq - nc);
#pragma omp parallel for
    for (long long n = 0; n < numIts; n++) {
        uint64_t j = ins_apply((uint64_t)n, &s);
        a[j] *= d[(j >> targ) & 1ull];
    }
}

/* cpu_statevec_anyCtrlTwoTargDiagMatr_sub (:620-655) */
void qo_diag2(qo_c* a, int nq, const int* ctrls, const int* states, int nc, int t1, int t2, const qo_c* d) {
    ins_t s = ins_make(ctrls, states, nc);
    long long numIts = 1ll << (nq - nc);
#pragma omp parallel for
    for (long long n = 0; n < numIts; n++) {
        uint64_t j = ins_apply((uint64_t)n, &s);
        int k = (int)((((j >> t2) & 1ull) << 1) | ((j >> t1) & 1ull));
        a[j] *= d[k];
    }
}

/* cpu_statevec_anyCtrlAnyTargDiagMatr_sub (:658-712): optional power then conj */
void qo_diagN(qo_c* a, int nq, const int* ctrls, const int* states, int nc, const int* targs, int nt,
              const qo_c* d, int applyConj, int hasPower, double expRe, double expIm) {
    ins_t s = ins_make(ctrls, states, nc);
    long long numIts = 1ll << (nq - nc);
    qo_c ex = expRe + I * expIm;
#pragma omp parallel for
    for (long long n = 0; n < numIts; n++) {
        uint64_t j = ins_apply((uint64_t)n, &s);
        qo_c e = d[value_of_bits(j, targs, nt)];
        if (hasPower) e = cpow(e, ex);
        if (applyConj) e = conj(e);
        a[j] *= e;
    }
}

/* cpu_statevec_anyCtrlSwap_subA (:259-284) */
void qo_swap(qo_c* a, int nq, const int* ctrls, const int* states, int nc, int t1, int t2) {
    ins_t s = ins_make(ctrls, states, nc);
    ins_add(&s, t2, 0);
    ins_add(&s, t1, 1);
    long long numIts = 1ll << (nq - nc - 2);
    uint64_t flip = (1ull << t1) | (1ull << t2);
#pragma omp parallel for
    for (long long n = 0; n < numIts; n++) {
        uint64_t i01 = ins_apply((uint64_t)n, &s), i10 = i01 ^ flip;
        qo_c t = a[i01];
        a[i01] = a[i10];
        a[i10] = t;
    }
}

/* cpu_statevector_anyCtrlPauliTensorOrGadget_subA (:803-907) and
   applyPauliUponAmpPair: pairFac gains i^#Y */
void qo_pauli(qo_c* a, int nq, const int* ctrls, const int* states, int nc, uint64_t mx, uint64_t my,
              uint64_t mz, double afr, double afi, double pfr, double pfi) {
    uint64_t mxy = mx | my, myz = my | mz;
    int targs[64], nt = 0;
    for (int q = 0; q < 64; q++)
        if ((mxy >> q) & 1ull) targs[nt++] = q;
    ins_t s = ins_make(ctrls, states, nc);
    for (int i = 0; i < nt; i++) ins_add(&s, targs[i], 0);
    qo_c ampFac = afr + I * afi, pairFac = pfr + I * pfi;
    int ny = __builtin_popcountll(my) & 3;
    for (int k = 0; k < ny; k++) pairFac *= I;
    long long numOuter = 1ll << (nq - nc - nt);
    long long numInner = (1ll << nt) / 2;
#pragma omp parallel for
    for (long long n = 0; n < numOuter; n++) {
        uint64_t i0 = ins_apply((uint64_t)n, &s);
        for (long long v = 0; v < numInner; v++) {
            uint64_t iA = set_bits(i0, targs, nt, (uint64_t)v);
            uint64_t iB = iA ^ mxy;
            int sA = 1 - 2 * parity(iA & myz), sB = 1 - 2 * parity(iB & myz);
            qo_c aA = a[iA], aB = a[iB];
            a[iA] = ampFac * aA + pairFac * sB * aB;
            a[iB] = ampFac * aB + pairFac * sA * aA;
        }
    }
}

/* cpu_statevector_anyCtrlAnyTargZOrPhaseGadget_sub (:963-1000) */
void qo_phaseGadget(qo_c* a, int nq, const int* ctrls, const int* states, int nc, uint64_t mask, double f0r,
                    double f0i, double f1r, double f1i) {
    ins_t s = ins_make(ctrls, states, nc);
    qo_c f[2] = {f0r + I * f0i, f1r + I * f1i};
    long long numIts = 1ll << (nq - nc);
#pragma omp parallel for
    for (long long n = 0; n < numIts; n++) {
        uint64_t i = ins_apply((uint64_t)n, &s);
        a[i] *= f[parity(i & mask)];
    }
}

/* cpu_statevec_multiQubitProjector_sub (:2282-2309) */
void qo_projector(qo_c* a, int nq, const int* qubits, const int* outcomes, int num, double prob) {
    double renorm = 1 / sqrt(prob);
    uint64_t retain = 0;
    for (int i = 0; i < num; i++) retain |= (uint64_t)outcomes[i] << i;
    long long numIts = 1ll << nq;
#pragma omp parallel for
    for (long long n = 0; n < numIts; n++) {
        uint64_t v = value_of_bits((uint64_t)n, qubits, num);
        a[n] *= renorm * (double)(v == retain);
    }
}

/* ---------------------------------------------------------- channels */

/* cpu_densmatr_oneQubitDephasing_subA (:1077-1107), fac = 1 - 2p */
void qo_dm_dephase1(qo_c* a, int N, int q, double p) {
    double fac = 1 - 2 * p;
    int bra = q + N;
    ins_t s = {0, {0}, 0};
    ins_add(&s, q, 0);
    ins_add(&s, bra, 0);
    long long numIts = 1ll << (2 * N - 2);
#pragma omp parallel for
    for (long long n = 0; n < numIts; n++) {
        uint64_t i00 = ins_apply((uint64_t)n, &s);
        a[i00 | (1ull << q)] *= fac;
        a[i00 | (1ull << bra)] *= fac;
    }
}

/* cpu_densmatr_twoQubitDephasing_subB (:1148-1181), term = -4p/3 */
void qo_dm_dephase2(qo_c* a, int N, int qa, int qb, double p) {
    double term = -4 * p / 3;
    long long numIts = 1ll << (2 * N);
#pragma omp parallel for
    for (long long n = 0; n < numIts; n++) {
        uint64_t i = (uint64_t)n;
        int bitA = (int)(((i >> qa) ^ (i >> (qa + N))) & 1ull);
        int bitB = (int)(((i >> qb) ^ (i >> (qb + N))) & 1ull);
        int flag = bitA | bitB;
        a[n] *= 1 + (flag * term);
    }
}

/* cpu_densmatr_oneQubitDepolarising_subA (:1183-1216); factors utilities.cpp:932-947 */
void qo_dm_depol1(qo_c* a, int N, int q, double p) {
    double c1 = 1 - (2 * p / 3), c2 = 2 * p / 3, c3 = 1 - (4 * p / 3);
    int bra = q + N;
    ins_t s = {0, {0}, 0};
    ins_add(&s, q, 0);
    ins_add(&s, bra, 0);
    long long numIts = 1ll << (2 * N - 2);
#pragma omp parallel for
    for (long long n = 0; n < numIts; n++) {
        uint64_t i00 = ins_apply((uint64_t)n, &s);
        uint64_t i01 = i00 | (1ull << q), i10 = i00 | (1ull << bra), i11 = i01 | (1ull << bra);
        qo_c a00 = a[i00], a11 = a[i11];
        a[i00] = (c1 * a00) + (c2 * a11);
        a[i01] *= c3;
        a[i10] *= c3;
        a[i11] = (c1 * a11) + (c2 * a00);
    }
}

/* cpu_densmatr_twoQubitDepolarising_subA then _subB (:1264-1327);
   factors utilities.cpp (c1 = 1-4p/5, c2 = 4p/15, c3 = -16p/15) */
void qo_dm_depol2(qo_c* a, int N, int q1, int q2, double p) {
    double c1 = 1 - (4 * p / 5), c2 = 4 * p / 15, c3 = -(16 * p / 15);
    int b1 = q1 + N, b2 = q2 + N;
    long long num = 1ll << (2 * N);
#pragma omp parallel for
    for (long long n = 0; n < num; n++) {
        int f1 = ((n >> q1) & 1) == ((n >> b1) & 1);
        int f2 = ((n >> q2) & 1) == ((n >> b2) & 1);
        int mod = !(f1 & f2);
        a[n] *= 1 + c3 * mod;
    }
    c1 -= c2;
    ins_t s = {0, {0}, 0};
    ins_add(&s, q1, 0);
    ins_add(&s, q2, 0);
    ins_add(&s, b1, 0);
    ins_add(&s, b2, 0);
    long long numIts = 1ll << (2 * N - 4);
#pragma omp parallel for
    for (long long n = 0; n < numIts; n++) {
        uint64_t i0000 = ins_apply((uint64_t)n, &s);
        uint64_t i0101 = i0000 | (1ull << b1) | (1ull << q1);
        uint64_t i1010 = i0000 | (1ull << b2) | (1ull << q2);
        uint64_t i1111 = i0101 | (1ull << b2) | (1ull << q2);
        qo_c term = a[i0000] + a[i0101] + a[i1010] + a[i1111];
        a[i0000] = c1 * a[i0000] + c2 * term;
        a[i0101] = c1 * a[i0101] + c2 * term;
        a[i1010] = c1 * a[i1010] + c2 * term;
        a[i1111] = c1 * a[i1111] + c2 * term;
    }
}

/* cpu_densmatr_oneQubitPauliChannel_subA (:1455-1498) */
void qo_dm_pauliChannel(qo_c* a, int N, int q, double pI, double pX, double pY, double pZ) {
    double fAA = pI + pZ, fBB = pX + pY, fAB = pI - pZ, fBA = pX - pY;
    int bra = q + N;
    ins_t s = {0, {0}, 0};
    ins_add(&s, q, 0);
    ins_add(&s, bra, 0);
    long long numIts = 1ll << (2 * N - 2);
#pragma omp parallel for
    for (long long n = 0; n < numIts; n++) {
        uint64_t i00 = ins_apply((uint64_t)n, &s);
        uint64_t i01 = i00 | (1ull << q), i10 = i00 | (1ull << bra), i11 = i01 | (1ull << bra);
        qo_c a00 = a[i00], a01 = a[i01], a10 = a[i10], a11 = a[i11];
        a[i00] = fAA * a00 + fBB * a11;
        a[i01] = fAB * a01 + fBA * a10;
        a[i10] = fAB * a10 + fBA * a01;
        a[i11] = fAA * a11 + fBB * a00;
    }
}

/* cpu_densmatr_oneQubitDamping_subA (:1550-1579) */
void qo_dm_damp(qo_c* a, int N, int q, double p) {
    double c1 = sqrt(1 - p), c2 = 1 - p;
    int bra = q + N;
    ins_t s = {0, {0}, 0};
    ins_add(&s, q, 0);
    ins_add(&s, bra, 0);
    long long numIts = 1ll << (2 * N - 2);
#pragma omp parallel for
    for (long long n = 0; n < numIts; n++) {
        uint64_t i00 = ins_apply((uint64_t)n, &s);
        uint64_t i01 = i00 | (1ull << q), i10 = i00 | (1ull << bra), i11 = i01 | (1ull << bra);
        a[i00] += p * a[i11];
        a[i01] *= c1;
        a[i10] *= c1;
        a[i11] *= c2;
    }
}

/* cpu_densmatr_multiQubitProjector_sub (:2311-2354) */
void qo_dm_projector(qo_c* a, int N, const int* qubits, const int* outcomes, int num, double prob) {
    double renorm = 1 / prob;
    uint64_t retain = 0;
    for (int i = 0; i < num; i++) retain |= (uint64_t)outcomes[i] << i;
    long long numIts = 1ll << (2 * N);
    uint64_t rowMask = (1ull << N) - 1;
#pragma omp parallel for
    for (long long n = 0; n < numIts; n++) {
        uint64_t r = (uint64_t)n & rowMask, c = (uint64_t)n >> N;
        uint64_t v1 = value_of_bits(r, qubits, num), v2 = value_of_bits(c, qubits, num);
        a[n] *= renorm * (double)(v1 == v2) * (double)(retain == v1);
    }
}

/* -------------------------------------------------------- reductions */
/* cpu_subroutines.cpp:1704-1824, 2006-2166 (OpenMP reduction(+)) */

double qo_totalProb(const qo_c* a, int nq) {
    double p = 0;
    long long num = 1ll << nq;
#pragma omp parallel for reduction(+ : p)
    for (long long n = 0; n < num; n++) p += creal(a[n]) * creal(a[n]) + cimag(a[n]) * cimag(a[n]);
    return p;
}

double qo_dm_totalProb(const qo_c* a, int N) {
    double p = 0;
    long long dim = 1ll << N;
#pragma omp parallel for reduction(+ : p)
    for (long long n = 0; n < dim; n++) p += creal(a[n * (dim + 1)]);
    return p;
}

double qo_probOutcome(const qo_c* a, int nq, const int* qubits, const int* outcomes, int num) {
    ins_t s = {0, {0}, 0};
    for (int i = 0; i < num; i++) ins_add(&s, qubits[i], outcomes[i]);
    double p = 0;
    long long numIts = 1ll << (nq - num);
#pragma omp parallel for reduction(+ : p)
    for (long long n = 0; n < numIts; n++) {
        qo_c v = a[ins_apply((uint64_t)n, &s)];
        p += creal(v) * creal(v) + cimag(v) * cimag(v);
    }
    return p;
}

double qo_dm_probOutcome(const qo_c* a, int N, const int* qubits, const int* outcomes, int num) {
    ins_t s = {0, {0}, 0};
    for (int i = 0; i < num; i++) ins_add(&s, qubits[i], outcomes[i]);
    double p = 0;
    long long dim = 1ll << N;
    long long numIts = 1ll << (N - num);
#pragma omp parallel for reduction(+ : p)
    for (long long n = 0; n < numIts; n++) p += creal(a[ins_apply((uint64_t)n, &s) * (uint64_t)(dim + 1)]);
    return p;
}

void qo_probsAllOutcomes(double* out, const qo_c* a, int nq, const int* qubits, int num) {
    long long nb = 1ll << num;
    for (long long b = 0; b < nb; b++) out[b] = 0;
    long long numIts = 1ll << nq;
    for (long long n = 0; n < numIts; n++) {
        double p = creal(a[n]) * creal(a[n]) + cimag(a[n]) * cimag(a[n]);
        out[value_of_bits((uint64_t)n, qubits, num)] += p;
    }
}

double qo_expecZ(const qo_c* a, int nq, uint64_t mask) {
    double v = 0;
    long long num = 1ll << nq;
#pragma omp parallel for reduction(+ : v)
    for (long long n = 0; n < num; n++) {
        double p = creal(a[n]) * creal(a[n]) + cimag(a[n]) * cimag(a[n]);
        v += (1 - 2 * parity((uint64_t)n & mask)) * p;
    }
    return v;
}

/* cpu_statevec_calcExpecPauliStr_subA (:2059-2088) */
void qo_expecPauli(double* out, const qo_c* a, int nq, uint64_t mx, uint64_t my, uint64_t mz) {
    uint64_t mxy = mx | my, myz = my | mz;
    double re = 0, im = 0;
    long long num = 1ll << nq;
#pragma omp parallel for reduction(+ : re, im)
    for (long long n = 0; n < num; n++) {
        uint64_t j = (uint64_t)n ^ mxy;
        int sign = 1 - 2 * parity(j & myz);
        qo_c t = sign * conj(a[n]) * a[j];
        re += creal(t);
        im += cimag(t);
    }
    qo_c v = re + I * im;
    int ny = __builtin_popcountll(my) & 3;
    for (int k = 0; k < ny; k++) v *= I;
    out[0] = creal(v);
    out[1] = cimag(v);
}

/* cpu_densmatr_calcExpecPauliStr_sub (:2129-2174) */
void qo_dm_expecPauli(double* out, const qo_c* a, int N, uint64_t mx, uint64_t my, uint64_t mz) {
    uint64_t mxy = mx | my, myz = my | mz;
    double re = 0, im = 0;
    long long dim = 1ll << N;
#pragma omp parallel for reduction(+ : re, im)
    for (long long n = 0; n < dim; n++) {
        uint64_t i = (uint64_t)n ^ mxy;
        int sign = 1 - 2 * parity(i & myz);
        qo_c t = sign * a[i + (uint64_t)n * (uint64_t)dim];
        re += creal(t);
        im += cimag(t);
    }
    qo_c v = re + I * im;
    int ny = __builtin_popcountll(my) & 3;
    for (int k = 0; k < ny; k++) v *= I;
    out[0] = creal(v);
    out[1] = cimag(v);
}

/* cpu_densmatr_calcProbsOfAllMultiQubitOutcomes_sub (:1866-1903): the
   diagonal element of every column, binned by the outcome of its row */
void qo_dm_probsAllOutcomes(double* out, const qo_c* a, int nq, const int* qubits, int num) {
    long long nb = 1ll << num;
    for (long long b = 0; b < nb; b++) out[b] = 0;
    long long dim = 1ll << nq;
    for (long long n = 0; n < dim; n++) out[value_of_bits((uint64_t)n, qubits, num)] += creal(a[n * (dim + 1)]);
}

/* cpu_densmatr_calcHilbertSchmidtDistance_sub (:1940-1953): sum |A-B|^2 */
double qo_dm_hsDistanceSq(const qo_c* a, const qo_c* b, int nq) {
    double d = 0;
    long long num = 1ll << (2 * nq);
#pragma omp parallel for reduction(+ : d)
    for (long long n = 0; n < num; n++) {
        qo_c t = a[n] - b[n];
        d += creal(t) * creal(t) + cimag(t) * cimag(t);
    }
    return d;
}

/* cpu_densmatr_calcFidelityWithPureState_sub (:1956-1994): <psi|rho|psi>,
   or <psi|rho^dagger|psi> when conjugated */
void qo_dm_fidelityPure(double* out, const qo_c* rho, const qo_c* psi, int nq, int conjugate) {
    double re = 0, im = 0;
    long long num = 1ll << (2 * nq);
    long long dim = 1ll << nq;
#pragma omp parallel for reduction(+ : re, im)
    for (long long i = 0; i < num; i++) {
        long long r = i & (dim - 1), c = i >> nq;
        qo_c rhoAmp = rho[i], rowAmp = psi[r], colAmp = psi[c];
        if (conjugate) {
            rhoAmp = conj(rhoAmp);
            colAmp = conj(colAmp);
        } else {
            rowAmp = conj(rowAmp);
        }
        qo_c t = rhoAmp * rowAmp * colAmp;
        re += creal(t);
        im += cimag(t);
    }
    out[0] = re;
    out[1] = im;
}

/* cpu_statevec_calcInnerProduct_sub (:1919-1937): sum conj(a) b; density
   matrices use it on their flat amplitudes (Tr(A^dagger B)) */
void qo_innerProduct(double* out, const qo_c* a, const qo_c* b, int nq) {
    double re = 0, im = 0;
    long long num = 1ll << nq;
#pragma omp parallel for reduction(+ : re, im)
    for (long long n = 0; n < num; n++) {
        qo_c t = conj(a[n]) * b[n];
        re += creal(t);
        im += cimag(t);
    }
    out[0] = re;
    out[1] = im;
}

/* cpu_statevec_initDebugState_sub (:2364-2377) */
void qo_initDebug(qo_c* a, int nq) {
    long long num = 1ll << nq;
    for (long long i = 0; i < num; i++) a[i] = (2 * i / 10.) + I * ((2 * i + 1) / 10.);
}
