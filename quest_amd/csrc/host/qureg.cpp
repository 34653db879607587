// Qureg lifecycle, amplitude access and state initialisation
// (quest/src/api/qureg.cpp:143-457, quest/src/api/initialisations.cpp).
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>

#include "host_core.hpp"

using namespace qah;

namespace {

// statevector scratch of the density-matrix initialisers (a random or
// given pure state before rho = |psi><psi|): allocated on first need, grown
// only when a larger density matrix asks, freed by finalizeQuESTEnv, so
// repeated initialisations allocate nothing
void* g_initScratch = nullptr;
size_t g_initScratchBytes = 0;

void* initScratch(size_t bytes) {
    if (bytes > g_initScratchBytes) {
        if (g_initScratch) check(qa_free(g_initScratch));
        g_initScratch = nullptr;
        g_initScratchBytes = 0;
        check(qa_malloc(&g_initScratch, bytes));
        g_initScratchBytes = bytes;
    }
    return g_initScratch;
}

void fillZero(const Qureg& q) { check(qa_statevec_initUniformState(dev(q), precOf(q), idxBits(q), 0, 0)); }

void setOne(const Qureg& q, qindex ind, double re, double im) {
    double v[2] = {re, im};
    check(qa_setAmps(dev(q), precOf(q), ind, v, 1));
}

void validateAmpRange(const Qureg& q, qindex start, qindex num) {
    if (start < 0 || num < 0 || start + num > q.numAmps)
        fail("Invalid amplitude range [" + std::to_string(start) + ", " + std::to_string(start + num) +
             ") for a Qureg of " + std::to_string(q.numAmps) + " amplitudes.");
}

}  // namespace

namespace qah {
void releaseInitScratch() {
    if (g_initScratch) qa_free(g_initScratch);
    g_initScratch = nullptr;
    g_initScratchBytes = 0;
}
}  // namespace qah

extern "C" {

Qureg createCustomQureg(int numQubits, int isDensMatr, int useDistrib, int useGpuAccel, int useMultithread) {
    QAH_API({
        if (useDistrib == 1) fail("This build has no distribution: useDistrib must be 0 or -1.");
        if (useGpuAccel == 0) fail("This build runs every Qureg on the GPU: useGpuAccel must be 1 or -1.");
        (void)useMultithread;
        return makeQureg(numQubits, isDensMatr, QA_FP64);
    });
}

Qureg createQureg(int numQubits) { return createCustomQureg(numQubits, 0, -1, -1, -1); }
Qureg createDensityQureg(int numQubits) { return createCustomQureg(numQubits, 1, -1, -1, -1); }
Qureg createForcedQureg(int numQubits) { return createCustomQureg(numQubits, 0, 0, 1, 0); }
Qureg createForcedDensityQureg(int numQubits) { return createCustomQureg(numQubits, 1, 0, 1, 0); }

Qureg createCustomPrecisionQureg(int numQubits, int isDensMatr, int precision) {
    QAH_API({ return makeQureg(numQubits, isDensMatr, precision); });
}

int getQuregPrecision(Qureg qureg) {
    QAH_API({
        validateQureg(qureg);
        return precOf(qureg) == QA_FP64 ? 2 : 1;
    });
}

qindex getQuregDeviceBytes(int numQubits, int isDensMatr, int precision) {
    int bits = isDensMatr ? 2 * numQubits : numQubits;
    return (qindex)(1LL << bits) * (precision == 1 ? 8 : 16);
}

Qureg createCloneQureg(Qureg qureg) {
    QAH_API({
        validateQureg(qureg);
        Qureg c = makeQureg(qureg.numQubits, qureg.isDensityMatrix, precOf(qureg));
        check(qa_copyAmps(dev(c), precOf(c), dev(qureg), precOf(qureg), qureg.numAmps));
        return c;
    });
}

void destroyQureg(Qureg qureg) {
    QAH_API({
        validateQureg(qureg);
        unregisterQureg(qureg);
        check(qa_free(dev(qureg)));
        std::free(qureg.cpuAmps);
    });
}

void reportQuregParams(Qureg qureg) {
    QAH_API({
        validateQureg(qureg);
        int p = precOf(qureg);
        std::printf("Qureg (%s, %d qubit%s, %lld amps, %s, GPU):\n  %lld bytes of GPU memory\n",
                    qureg.isDensityMatrix ? "density matrix" : "statevector", qureg.numQubits,
                    qureg.numQubits == 1 ? "" : "s", qureg.numAmps, p == QA_FP64 ? "complex128" : "complex64",
                    (long long)(qureg.numAmps * (p == QA_FP64 ? 16 : 8)));
    });
}

void reportQureg(Qureg qureg) {
    QAH_API({
        validateQureg(qureg);
        reportQuregParams(qureg);
        qindex n = qureg.numAmps < 16 ? qureg.numAmps : 16;
        std::vector<double> a((size_t)(2 * n));
        check(qa_getAmps(a.data(), dev(qureg), precOf(qureg), 0, n));
        for (qindex i = 0; i < n; i++) std::printf("  %lld: %.10g%+.10gi\n", i, a[(size_t)(2 * i)], a[(size_t)(2 * i + 1)]);
        if (n < qureg.numAmps) std::printf("  ...\n");
    });
}

void syncSubQuregToGpu(Qureg qureg, qindex start, qindex num) {
    QAH_API({
        validateQureg(qureg);
        validateAmpRange(qureg, start, num);
        if (!qureg.cpuAmps) fail("This Qureg has no host mirror.");
        check(qa_setAmps(dev(qureg), precOf(qureg), start, dbl(qureg.cpuAmps + start), num));
    });
}
void syncSubQuregFromGpu(Qureg qureg, qindex start, qindex num) {
    QAH_API({
        validateQureg(qureg);
        validateAmpRange(qureg, start, num);
        if (!qureg.cpuAmps) fail("This Qureg has no host mirror.");
        check(qa_getAmps((double*)(qureg.cpuAmps + start), dev(qureg), precOf(qureg), start, num));
    });
}
void syncQuregToGpu(Qureg qureg) { syncSubQuregToGpu(qureg, 0, qureg.numAmps); }
void syncQuregFromGpu(Qureg qureg) { syncSubQuregFromGpu(qureg, 0, qureg.numAmps); }

void getQuregAmps(qcomp* outAmps, Qureg qureg, qindex startInd, qindex numAmps) {
    QAH_API({
        validateStateVec(qureg);
        validateAmpRange(qureg, startInd, numAmps);
        check(qa_getAmps((double*)outAmps, dev(qureg), precOf(qureg), startInd, numAmps));
    });
}

void getDensityQuregAmps(qcomp** outAmps, Qureg qureg, qindex startRow, qindex startCol, qindex numRows,
                         qindex numCols) {
    QAH_API({
        validateDensMatr(qureg);
        std::vector<double> buf((size_t)(2 * numRows * numCols));
        check(qa_getDensityAmps(buf.data(), dev(qureg), precOf(qureg), qureg.numQubits, startRow, startCol, numRows,
                                numCols));
        for (qindex r = 0; r < numRows; r++)
            for (qindex c = 0; c < numCols; c++)
                outAmps[r][c] = qcomp(buf[(size_t)(2 * (r * numCols + c))], buf[(size_t)(2 * (r * numCols + c) + 1)]);
    });
}

void _wrap_getQuregAmp(qcomp* out, Qureg qureg, qindex index) {
    QAH_API({
        validateStateVec(qureg);
        validateAmpRange(qureg, index, 1);
        check(qa_getAmps((double*)out, dev(qureg), precOf(qureg), index, 1));
    });
}

void _wrap_getDensityQuregAmp(qcomp* out, Qureg qureg, qindex row, qindex column) {
    QAH_API({
        validateDensMatr(qureg);
        qindex dim = 1LL << qureg.numQubits;
        if (row < 0 || row >= dim || column < 0 || column >= dim) fail("Invalid density-matrix row or column.");
        check(qa_getAmps((double*)out, dev(qureg), precOf(qureg), row + column * dim, 1));
    });
}

// ---- initialisation (initialisations.cpp) ----------------------------------

void initBlankState(Qureg qureg) {
    QAH_API({
        validateQureg(qureg);
        fillZero(qureg);
    });
}

void initZeroState(Qureg qureg) { initClassicalState(qureg, 0); }

void initPlusState(Qureg qureg) {
    QAH_API({
        validateQureg(qureg);
        // statevector: 1/sqrt(2^n) everywhere; density matrix: 1/2^N everywhere
        double v = qureg.isDensityMatrix ? 1.0 / (double)(1LL << qureg.numQubits)
                                         : 1.0 / std::sqrt((double)qureg.numAmps);
        check(qa_statevec_initUniformState(dev(qureg), precOf(qureg), idxBits(qureg), v, 0));
    });
}

void initClassicalState(Qureg qureg, qindex stateInd) {
    QAH_API({
        validateQureg(qureg);
        qindex dim = 1LL << qureg.numQubits;
        if (stateInd < 0 || stateInd >= dim) fail("Invalid classical state index.");
        fillZero(qureg);
        qindex ind = qureg.isDensityMatrix ? stateInd * (dim + 1) : stateInd;
        setOne(qureg, ind, 1, 0);
    });
}

void initDebugState(Qureg qureg) {
    QAH_API({
        validateQureg(qureg);
        check(qa_statevec_initDebugState(dev(qureg), precOf(qureg), idxBits(qureg)));
    });
}

void initArbitraryPureState(Qureg qureg, qcomp* amps) {
    QAH_API({
        validateQureg(qureg);
        qindex dim = 1LL << qureg.numQubits;
        if (!qureg.isDensityMatrix) {
            check(qa_setAmps(dev(qureg), precOf(qureg), 0, dbl(amps), dim));
            return;
        }
        // rho = |psi><psi| through the library's statevector scratch
        int p = precOf(qureg);
        void* psi = initScratch((size_t)dim * (p == QA_FP64 ? 16 : 8));
        check(qa_setAmps(psi, p, 0, dbl(amps), dim));
        check(qa_densmatr_initPureState(dev(qureg), psi, p, qureg.numQubits));
    });
}

void initPureState(Qureg qureg, Qureg pure) {
    QAH_API({
        validateQureg(qureg);
        validateStateVec(pure);
        validateSameDims(qureg, pure);
        if (precOf(qureg) != precOf(pure)) fail("The Quregs have different precisions.");
        if (qureg.isDensityMatrix)
            check(qa_densmatr_initPureState(dev(qureg), dev(pure), precOf(qureg), qureg.numQubits));
        else
            check(qa_copyAmps(dev(qureg), precOf(qureg), dev(pure), precOf(pure), qureg.numAmps));
    });
}

qreal setQuregToRenormalized(Qureg qureg);

// one normalised random statevector of n qubits at psi, from one seed of
// the measurement RNG (initialisations.cpp:121-133: the unnormalised
// random amplitudes, then setQuregToRenormalized's 1/sqrt(total))
static void randomPureInto(void* psi, int p, int n) {
    unsigned seed = randomSharedSeed();
    check(qa_statevec_initUnnormalisedUniformlyRandomPureStateAmps(psi, p, n, seed));
    double norm = 0;
    check(qa_statevec_calcTotalProb(psi, p, n, &norm));
    if (norm <= 0) fail("The random state has zero norm.");
    check(qa_statevec_setQuregToSuperposition(p, n, 1 / std::sqrt(norm), 0, psi, 0, 0, psi, 0, 0, psi));
}

void initRandomPureState(Qureg qureg) {
    QAH_API({
        validateQureg(qureg);
        int p = precOf(qureg);
        if (!qureg.isDensityMatrix) {
            randomPureInto(dev(qureg), p, qureg.numQubits);
            return;
        }
        // localiser.cpp:774-798: a random statevector, then rho = |psi><psi|
        qindex dim = 1LL << qureg.numQubits;
        void* psi = initScratch((size_t)dim * (p == QA_FP64 ? 16 : 8));
        randomPureInto(psi, p, qureg.numQubits);
        check(qa_densmatr_initPureState(dev(qureg), psi, p, qureg.numQubits));
    });
}

void initRandomMixedState(Qureg qureg, qindex numPureStates) {
    QAH_API({
        validateDensMatr(qureg);
        if (numPureStates < 1) fail("Invalid number of pure states; must be >= 1.");
        // localiser.cpp:801-827: from a blank matrix, mix in numPureStates
        // random pure states with weight 1/numPureStates each (one seed of
        // the measurement RNG per state), then renormalise by the trace
        qindex dim = 1LL << qureg.numQubits;
        int p = precOf(qureg);
        void* psi = initScratch((size_t)dim * (p == QA_FP64 ? 16 : 8));
        check(qa_statevec_initUniformState(dev(qureg), p, 2 * qureg.numQubits, 0, 0));
        for (qindex k = 0; k < numPureStates; k++) {
            randomPureInto(psi, p, qureg.numQubits);
            check(qa_densmatr_mixQuregWithStatevec(1.0, dev(qureg), 1.0 / (double)numPureStates, psi, p,
                                                   qureg.numQubits));
        }
        double tr = 0;
        check(qa_densmatr_calcTotalProb(dev(qureg), p, qureg.numQubits, &tr));
        if (tr <= 0) fail("The random mixed state has zero trace.");
        check(qa_statevec_setQuregToSuperposition(p, 2 * qureg.numQubits, 1 / tr, 0, dev(qureg), 0, 0, dev(qureg), 0,
                                                  0, dev(qureg)));
    });
}

void setQuregAmps(Qureg qureg, qindex startInd, qcomp* amps, qindex numAmps) {
    QAH_API({
        validateStateVec(qureg);
        validateAmpRange(qureg, startInd, numAmps);
        check(qa_setAmps(dev(qureg), precOf(qureg), startInd, dbl(amps), numAmps));
    });
}

void setDensityQuregFlatAmps(Qureg qureg, qindex startInd, qcomp* amps, qindex numAmps) {
    QAH_API({
        validateDensMatr(qureg);
        validateAmpRange(qureg, startInd, numAmps);
        check(qa_setAmps(dev(qureg), precOf(qureg), startInd, dbl(amps), numAmps));
    });
}

void setDensityQuregAmps(Qureg qureg, qindex startRow, qindex startCol, qcomp** amps, qindex numRows,
                         qindex numCols) {
    QAH_API({
        validateDensMatr(qureg);
        qindex dim = 1LL << qureg.numQubits;
        if (startRow < 0 || startCol < 0 || numRows < 0 || numCols < 0 || startRow + numRows > dim ||
            startCol + numCols > dim)
            fail("Invalid density-matrix block.");
        // each column segment is contiguous in the column-major layout
        std::vector<qcomp> col((size_t)numRows);
        for (qindex c = 0; c < numCols; c++) {
            for (qindex r = 0; r < numRows; r++) col[(size_t)r] = amps[r][c];
            check(qa_setAmps(dev(qureg), precOf(qureg), startRow + (startCol + c) * dim, dbl(col.data()), numRows));
        }
    });
}

void setQuregToClone(Qureg target, Qureg copy) {
    QAH_API({
        validateQureg(target);
        validateQureg(copy);
        if (target.isDensityMatrix != copy.isDensityMatrix) fail("The Quregs must both be statevectors or density matrices.");
        validateSameDims(target, copy);
        check(qa_copyAmps(dev(target), precOf(target), dev(copy), precOf(copy), target.numAmps));
    });
}

void setQuregToSuperposition(qcomp facOut, Qureg out, qcomp fac1, Qureg q1, qcomp fac2, Qureg q2) {
    QAH_API({
        validateStateVec(out);
        validateStateVec(q1);
        validateStateVec(q2);
        validateSameDims(out, q1);
        validateSameDims(out, q2);
        if (precOf(out) != precOf(q1) || precOf(out) != precOf(q2)) fail("The Quregs have different precisions.");
        check(qa_statevec_setQuregToSuperposition(precOf(out), out.numQubits, facOut.real(), facOut.imag(), dev(out),
                                                  fac1.real(), fac1.imag(), dev(q1), fac2.real(), fac2.imag(),
                                                  dev(q2)));
    });
}

// initialisations.cpp:216-227: rho <- the matrix of the sum (which may be
// non-Hermitian)
void setQuregToPauliStrSum(Qureg qureg, PauliStrSum sum) {
    QAH_API({
        validateQureg(qureg);
        if (!qureg.isDensityMatrix) fail("The Qureg must be a density matrix.");
        if (sum.numTerms < 1 || !sum.strings || !sum.coeffs) fail("Invalid PauliStrSum.");
        std::vector<uint64_t> mx((size_t)sum.numTerms), my((size_t)sum.numTerms), mz((size_t)sum.numTerms);
        std::vector<double> c(2 * (size_t)sum.numTerms);
        for (qindex t = 0; t < sum.numTerms; t++) {
            if (pauliMaxIndex(sum.strings[t]) >= qureg.numQubits)
                fail("The PauliStrSum targets qubits beyond the " + std::to_string(qureg.numQubits) + "-qubit Qureg.");
            pauliMasks(sum.strings[t], mx[t], my[t], mz[t]);
            c[2 * t] = sum.coeffs[t].real();
            c[2 * t + 1] = sum.coeffs[t].imag();
        }
        check(qa_densmatr_setAmpsToPauliStrSum(dev(qureg), precOf(qureg), qureg.numQubits, sum.numTerms, mx.data(),
                                               my.data(), mz.data(), c.data()));
    });
}

qreal setQuregToRenormalized(Qureg qureg) {
    QAH_API({
        validateQureg(qureg);
        int p = precOf(qureg);
        double prob = 0;
        if (qureg.isDensityMatrix) check(qa_densmatr_calcTotalProb(dev(qureg), p, qureg.numQubits, &prob));
        else check(qa_statevec_calcTotalProb(dev(qureg), p, qureg.numQubits, &prob));
        if (std::abs(prob) <= epsilonFor(qureg)) fail("The Qureg has (near) zero probability and cannot be renormalised.");
        double f = qureg.isDensityMatrix ? 1 / prob : 1 / std::sqrt(prob);
        check(qa_statevec_setQuregToSuperposition(p, idxBits(qureg), f, 0, dev(qureg), 0, 0, dev(qureg), 0, 0,
                                                  dev(qureg)));
        return 1 / prob;
    });
}

void setQuregToPartialTrace(Qureg out, Qureg in, int* traceOutQubits, int numTraceQubits) {
    QAH_API({
        validateDensMatr(in);
        validateDensMatr(out);
        validateTargets(in, traceOutQubits, numTraceQubits);
        if (out.numQubits != in.numQubits - numTraceQubits) fail("The output Qureg has the wrong number of qubits.");
        if (precOf(out) != precOf(in)) fail("The Quregs have different precisions.");
        check(qa_densmatr_partialTrace(dev(in), precOf(in), in.numQubits, dev(out), traceOutQubits, numTraceQubits));
    });
}

void setQuregToReducedDensityMatrix(Qureg out, Qureg in, int* retainQubits, int numRetainQubits) {
    QAH_API({
        validateDensMatr(in);
        validateTargets(in, retainQubits, numRetainQubits);
        std::vector<int> trace;
        for (int q = 0; q < in.numQubits; q++) {
            bool keep = false;
            for (int i = 0; i < numRetainQubits; i++) keep |= (retainQubits[i] == q);
            if (!keep) trace.push_back(q);
        }
        setQuregToPartialTrace(out, in, trace.data(), (int)trace.size());
    });
}

}  // extern "C"

qcomp getQuregAmp(Qureg qureg, qindex index) {
    qcomp out = 0;
    _wrap_getQuregAmp(&out, qureg, index);
    return out;
}
qcomp getDensityQuregAmp(Qureg qureg, qindex row, qindex column) {
    qcomp out = 0;
    _wrap_getDensityQuregAmp(&out, qureg, row, column);
    return out;
}
void setQuregAmps(Qureg qureg, qindex startInd, std::vector<qcomp> amps) {
    setQuregAmps(qureg, startInd, amps.data(), (qindex)amps.size());
}
std::vector<qcomp> getQuregAmps(Qureg qureg, qindex startInd, qindex numAmps) {
    std::vector<qcomp> out((size_t)(numAmps > 0 ? numAmps : 0));
    getQuregAmps(out.data(), qureg, startInd, numAmps);
    return out;
}
