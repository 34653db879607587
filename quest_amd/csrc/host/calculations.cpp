// Decoherence channels (quest/src/api/decoherence.cpp, localiser.cpp:639-767)
// and calculations (quest/src/api/calculations.cpp, localiser.cpp:827-1045).
#include <cmath>

#include "host_core.hpp"

using namespace qah;

namespace {

const double REDUCTION_EPS_FACTOR = 100;  // validation.cpp:1137

bool isCPTP(KrausMap map) {
    if (*map.isApproxCPTP != -1) return *map.isApproxCPTP;
    *map.isApproxCPTP = 1;
    for (qindex r = 0; r < map.numRows; r++)
        for (qindex c = 0; c < map.numRows; c++) {
            qcomp e = 0;
            for (int n = 0; n < map.numMatrices; n++)
                for (qindex k = 0; k < map.numRows; k++) e += std::conj(map.matrices[n][k][r]) * map.matrices[n][k][c];
            if (std::norm(e - qcomp(r == c, 0)) > epsilon()) {
                *map.isApproxCPTP = 0;
                return false;
            }
        }
    return true;
}

// superoperator on targets ket ++ bra (localiser.cpp:759-768)
void applySuperOp(const Qureg& q, const SuperOp& op, const int* targs, int nt) {
    vector<int> all = vec(targs, nt);
    for (int i = 0; i < nt; i++) all.push_back(targs[i] + q.numQubits);
    if (all.size() == 2) {
        double d[32];
        for (int r = 0; r < 4; r++)
            for (int c = 0; c < 4; c++) {
                d[2 * (4 * r + c)] = op.cpuElems[r][c].real();
                d[2 * (4 * r + c) + 1] = op.cpuElems[r][c].imag();
            }
        check(qa_statevec_anyCtrlTwoTargDenseMatr(dev(q), precOf(q), idxBits(q), nullptr, nullptr, 0, all[0], all[1], d));
        return;
    }
    check(qa_statevec_anyCtrlAnyTargDenseMatr(dev(q), precOf(q), idxBits(q), nullptr, nullptr, 0, all.data(),
                                              (int)all.size(), op.gpuElemsFlat, 0));
}

qcomp expecPauli(const Qureg& q, PauliStr str) {
    uint64_t mx, my, mz;
    pauliMasks(str, mx, my, mz);
    double out[2] = {0, 0};
    int p = precOf(q);
    if (q.isDensityMatrix) {
        // localiser.cpp:891-898: total prob, Z-only diagonal sum, or general
        if (!(mx | my | mz)) check(qa_densmatr_calcTotalProb(dev(q), p, q.numQubits, out));
        else if (!(mx | my)) check(qa_densmatr_calcExpecAnyTargZ(dev(q), p, q.numQubits, mz, out));
        else check(qa_densmatr_calcExpecPauliStr(dev(q), p, q.numQubits, mx, my, mz, out));
    } else {
        // localiser.cpp:883-890
        if (!(mx | my | mz)) check(qa_statevec_calcTotalProb(dev(q), p, q.numQubits, out));
        else if (!(mx | my)) check(qa_statevec_calcExpecAnyTargZ(dev(q), p, q.numQubits, mz, out));
        else check(qa_statevec_calcExpecPauliStr(dev(q), p, q.numQubits, mx, my, mz, out));
    }
    return qcomp(out[0], out[1]);
}

void validatePauliStrOnQureg(const Qureg& q, PauliStr str) {
    validateQureg(q);
    if (pauliMaxIndex(str) >= q.numQubits) fail("The Pauli string targets qubits beyond the Qureg.");
}

}  // namespace

extern "C" {

// ---- decoherence ------------------------------------------------------------------

void mixDephasing(Qureg q, int t, qreal prob) {
    QAH_API({
        validateDensMatr(q);
        validateTarget(q, t);
        validateProb(prob, 0.5, "dephasing");
        if (prob == 0) return;
        check(qa_densmatr_oneQubitDephasing(dev(q), precOf(q), q.numQubits, t, prob));
    });
}

void mixTwoQubitDephasing(Qureg q, int t1, int t2, qreal prob) {
    QAH_API({
        validateDensMatr(q);
        int t[2] = {t1, t2};
        validateTargets(q, t, 2);
        validateProb(prob, 0.75, "two-qubit dephasing");
        if (prob == 0) return;
        check(qa_densmatr_twoQubitDephasing(dev(q), precOf(q), q.numQubits, t1, t2, prob));
    });
}

void mixDepolarising(Qureg q, int t, qreal prob) {
    QAH_API({
        validateDensMatr(q);
        validateTarget(q, t);
        validateProb(prob, 0.75, "depolarising");
        if (prob == 0) return;
        check(qa_densmatr_oneQubitDepolarising(dev(q), precOf(q), q.numQubits, t, prob));
    });
}

void mixTwoQubitDepolarising(Qureg q, int t1, int t2, qreal prob) {
    QAH_API({
        validateDensMatr(q);
        int t[2] = {t1, t2};
        validateTargets(q, t, 2);
        validateProb(prob, 15. / 16, "two-qubit depolarising");
        if (prob == 0) return;
        check(qa_densmatr_twoQubitDepolarising(dev(q), precOf(q), q.numQubits, t1, t2, prob));
    });
}

void mixDamping(Qureg q, int t, qreal prob) {
    QAH_API({
        validateDensMatr(q);
        validateTarget(q, t);
        validateProb(prob, 1.0, "damping");
        if (prob == 0) return;
        check(qa_densmatr_oneQubitDamping(dev(q), precOf(q), q.numQubits, t, prob));
    });
}

void mixPaulis(Qureg q, int t, qreal px, qreal py, qreal pz) {
    QAH_API({
        validateDensMatr(q);
        validateTarget(q, t);
        validateProb(px, 1, "X");
        validateProb(py, 1, "Y");
        validateProb(pz, 1, "Z");
        if (px + py + pz > 1) fail("The Pauli channel probabilities exceed one in total.");
        double pI = 1 - px - py - pz;
        if (std::max(px, std::max(py, pz)) > pI) fail("The Pauli channel probabilities exceed maximal mixing.");
        if (px == 0 && py == 0 && pz == 0) return;
        check(qa_densmatr_oneQubitPauliChannel(dev(q), precOf(q), q.numQubits, t, pI, px, py, pz));
    });
}

void mixQureg(Qureg out, Qureg in, qreal prob) {
    QAH_API({
        validateDensMatr(out);
        validateQureg(in);
        validateSameDims(out, in);
        validateProb(prob, 1, "mixing");
        if (precOf(out) != precOf(in)) fail("The Quregs have different precisions.");
        if (in.isDensityMatrix)
            check(qa_densmatr_mixQureg(1 - prob, dev(out), prob, dev(in), precOf(out), out.numQubits));
        else
            check(qa_densmatr_mixQuregWithStatevec(1 - prob, dev(out), prob, dev(in), precOf(out), out.numQubits));
    });
}

void mixKrausMap(Qureg q, int* targs, int nt, KrausMap map) {
    QAH_API({
        validateDensMatr(q);
        validateTargets(q, targs, nt);
        if (map.numQubits != nt) fail("The KrausMap size does not match the number of target qubits.");
        if (!*map.superop.wasGpuSynced) fail("The KrausMap has not been set or synced.");
        if (validating() && !isCPTP(map)) fail("The KrausMap was not (approximately) completely positive and trace preserving.");
        applySuperOp(q, map.superop, targs, nt);
    });
}

void mixSuperOp(Qureg q, int* targs, int nt, SuperOp op) {
    QAH_API({
        validateDensMatr(q);
        validateTargets(q, targs, nt);
        if (op.numQubits != nt) fail("The SuperOp size does not match the number of target qubits.");
        if (!*op.wasGpuSynced) fail("The SuperOp has not been set or synced.");
        applySuperOp(q, op, targs, nt);
    });
}

// ---- calculations -------------------------------------------------------------------

qreal calcExpecPauliStr(Qureg q, PauliStr str) {
    QAH_API({
        validatePauliStrOnQureg(q, str);
        qcomp v = expecPauli(q, str);
        if (validating() && std::abs(v.imag()) > REDUCTION_EPS_FACTOR * epsilonFor(q))
            fail(q.isDensityMatrix ? "The calculated expectation value was not approximately real; the density matrix "
                                     "may be unnormalised or non-Hermitian."
                                   : "The calculated expectation value was not approximately real.");
        return v.real();
    });
}

void _wrap_calcExpecNonHermitianPauliStrSum(qcomp* out, Qureg q, PauliStrSum sum) {
    QAH_API({
        validateQureg(q);
        for (qindex t = 0; t < sum.numTerms; t++)
            if (pauliMaxIndex(sum.strings[t]) >= q.numQubits) fail("A Pauli string targets qubits beyond the Qureg.");
        if (!q.isDensityMatrix) {
            // localiser.cpp:2052-2126: strings sharing their X/Y mask are
            // reduced together, one read of the state per group of 8
            vector<uint64_t> mx(sum.numTerms), my(sum.numTerms), mz(sum.numTerms);
            vector<double> c(2 * sum.numTerms);
            for (qindex t = 0; t < sum.numTerms; t++) {
                pauliMasks(sum.strings[t], mx[t], my[t], mz[t]);
                c[2 * t] = sum.coeffs[t].real();
                c[2 * t + 1] = sum.coeffs[t].imag();
            }
            double v[2] = {0, 0};
            check(qa_statevec_calcExpecPauliStrSum(dev(q), precOf(q), q.numQubits, sum.numTerms, mx.data(), my.data(),
                                                   mz.data(), c.data(), nullptr, v));
            *out = qcomp(v[0], v[1]);
            return;
        }
        // localiser.cpp:2128-2151: one diagonal-sized sum per term
        qcomp tot = 0;
        for (qindex t = 0; t < sum.numTerms; t++) tot += sum.coeffs[t] * expecPauli(q, sum.strings[t]);
        *out = tot;
    });
}

qreal calcExpecPauliStrSum(Qureg q, PauliStrSum sum) {
    QAH_API({
        validateQureg(q);
        if (validating())
            for (qindex t = 0; t < sum.numTerms; t++)
                if (std::abs(sum.coeffs[t].imag()) > epsilon()) fail("The PauliStrSum is not Hermitian.");
        qcomp v = 0;
        _wrap_calcExpecNonHermitianPauliStrSum(&v, q, sum);
        if (validating() && std::abs(v.imag()) > REDUCTION_EPS_FACTOR * epsilonFor(q))
            fail("The calculated expectation value was not approximately real.");
        return v.real();
    });
}

void _wrap_calcExpecNonHermitianFullStateDiagMatr(qcomp* out, Qureg q, FullStateDiagMatr m) {
    QAH_API({
        validateQureg(q);
        if (m.numQubits != q.numQubits) fail("The FullStateDiagMatr and Qureg have different numbers of qubits.");
        double v[2];
        if (q.isDensityMatrix)
            check(qa_densmatr_calcExpecFullStateDiagMatr(dev(q), precOf(q), q.numQubits, m.gpuElems, 0, 1, 0, v));
        else
            check(qa_statevec_calcExpecFullStateDiagMatr(dev(q), precOf(q), q.numQubits, m.gpuElems, 0, 1, 0, v));
        *out = qcomp(v[0], v[1]);
    });
}

// calculations.cpp:98-115: sum of d_i^expo |a_i|^2 (rho_ii on a density
// matrix) with the complex power
void _wrap_calcExpecNonHermitianFullStateDiagMatrPower(qcomp* out, Qureg q, FullStateDiagMatr m, qcomp expo) {
    QAH_API({
        validateQureg(q);
        if (m.numQubits != q.numQubits) fail("The FullStateDiagMatr and Qureg have different numbers of qubits.");
        double v[2];
        if (q.isDensityMatrix)
            check(qa_densmatr_calcExpecFullStateDiagMatr(dev(q), precOf(q), q.numQubits, m.gpuElems, 1, expo.real(),
                                                         expo.imag(), v));
        else
            check(qa_statevec_calcExpecFullStateDiagMatr(dev(q), precOf(q), q.numQubits, m.gpuElems, 1, expo.real(),
                                                         expo.imag(), v));
        *out = qcomp(v[0], v[1]);
    });
}

qreal calcExpecFullStateDiagMatr(Qureg q, FullStateDiagMatr m) {
    qcomp v = 0;
    _wrap_calcExpecNonHermitianFullStateDiagMatr(&v, q, m);
    return v.real();
}

qreal calcExpecFullStateDiagMatrPower(Qureg q, FullStateDiagMatr m, qreal expo) {
    QAH_API({
        validateQureg(q);
        if (m.numQubits != q.numQubits) fail("The FullStateDiagMatr and Qureg have different numbers of qubits.");
        double v[2];
        if (q.isDensityMatrix)
            check(qa_densmatr_calcExpecFullStateDiagMatr(dev(q), precOf(q), q.numQubits, m.gpuElems, 2, expo, 0, v));
        else
            check(qa_statevec_calcExpecFullStateDiagMatr(dev(q), precOf(q), q.numQubits, m.gpuElems, 2, expo, 0, v));
        return v[0];
    });
}

qreal calcProbOfBasisState(Qureg q, qindex index) {
    QAH_API({
        validateQureg(q);
        qindex dim = 1LL << q.numQubits;
        if (index < 0 || index >= dim) fail("Invalid basis state index.");
        double a[2];
        qindex i = q.isDensityMatrix ? index * (1 + dim) : index;
        check(qa_getAmps(a, dev(q), precOf(q), i, 1));
        return q.isDensityMatrix ? a[0] : a[0] * a[0] + a[1] * a[1];
    });
}

qreal calcProbOfMultiQubitOutcome(Qureg q, int* qubits, int* outcomes, int n) {
    QAH_API({
        validateTargets(q, qubits, n);
        for (int i = 0; i < n; i++) validateOutcome(outcomes[i]);
        double p = 0;
        if (q.isDensityMatrix)
            check(qa_densmatr_calcProbOfMultiQubitOutcome(dev(q), precOf(q), q.numQubits, qubits, outcomes, n, &p));
        else
            check(qa_statevec_calcProbOfMultiQubitOutcome(dev(q), precOf(q), q.numQubits, qubits, outcomes, n, &p));
        return p;
    });
}

// calculations.cpp:241-262.  Callers ask this for qubit after qubit, so on
// statevectors of at least 16 qubits the first call computes every qubit's
// probabilities with one read of the state per 10 qubits and later calls
// take them from a cache, valid while the kernel layer's state epoch is
// unchanged (any entry point that may write amplitudes moves it)
qreal calcProbOfQubitOutcome(Qureg q, int qubit, int outcome) {
    if (q.isDensityMatrix || q.numQubits < 16) return calcProbOfMultiQubitOutcome(q, &qubit, &outcome, 1);
    QAH_API({
        validateTarget(q, qubit);
        validateOutcome(outcome);
        static struct {
            const void* amps = nullptr;
            uint64_t epoch = 0;
            int numQubits = 0, prec = 0;
            std::vector<double> p;
        } cache;
        const uint64_t epoch = qa_stateEpoch();
        if (cache.amps != dev(q) || cache.epoch != epoch || cache.numQubits != q.numQubits ||
            cache.prec != precOf(q)) {
            cache.amps = nullptr;
            cache.p.assign(2 * (size_t)q.numQubits, 0.0);
            check(qa_statevec_calcProbsOfEveryQubit(dev(q), precOf(q), q.numQubits, cache.p.data()));
            cache.amps = dev(q);
            cache.epoch = epoch;
            cache.numQubits = q.numQubits;
            cache.prec = precOf(q);
        }
        return cache.p[2 * (size_t)qubit + (size_t)outcome];
    });
}

void calcProbsOfAllMultiQubitOutcomes(qreal* out, Qureg q, int* qubits, int n) {
    QAH_API({
        validateTargets(q, qubits, n);
        if (q.isDensityMatrix)
            check(qa_densmatr_calcProbsOfAllMultiQubitOutcomes(out, dev(q), precOf(q), q.numQubits, qubits, n));
        else
            check(qa_statevec_calcProbsOfAllMultiQubitOutcomes(out, dev(q), precOf(q), q.numQubits, qubits, n));
    });
}

qreal calcTotalProb(Qureg q) {
    QAH_API({
        validateQureg(q);
        double p = 0;
        if (q.isDensityMatrix) check(qa_densmatr_calcTotalProb(dev(q), precOf(q), q.numQubits, &p));
        else check(qa_statevec_calcTotalProb(dev(q), precOf(q), q.numQubits, &p));
        return p;
    });
}

// calculations.cpp:293-300: sum |a|^2 over all amplitudes; squared for a statevector
qreal calcPurity(Qureg q) {
    QAH_API({
        validateQureg(q);
        double p = 0;
        check(qa_statevec_calcTotalProb(dev(q), precOf(q), idxBits(q), &p));
        return q.isDensityMatrix ? p : p * p;
    });
}

void _wrap_calcInnerProduct(Qureg a, Qureg b, qcomp* out) {
    QAH_API({
        validateQureg(a);
        validateQureg(b);
        validateSameDims(a, b);
        if (precOf(a) != precOf(b)) fail("The Quregs have different precisions.");
        double v[2];
        if (a.isDensityMatrix == b.isDensityMatrix) {
            check(qa_statevec_calcInnerProduct(dev(a), dev(b), precOf(a), idxBits(a), v));
        } else if (a.isDensityMatrix) {
            check(qa_densmatr_calcFidelityWithPureState(dev(a), dev(b), precOf(a), a.numQubits, 1, v));
        } else {
            check(qa_densmatr_calcFidelityWithPureState(dev(b), dev(a), precOf(a), a.numQubits, 0, v));
        }
        *out = qcomp(v[0], v[1]);
    });
}

qreal calcFidelity(Qureg a, Qureg b) {
    QAH_API({
        validateQureg(a);
        validateQureg(b);
        validateSameDims(a, b);
        if (a.isDensityMatrix && b.isDensityMatrix) fail("calcFidelity needs at least one statevector.");
        if (precOf(a) != precOf(b)) fail("The Quregs have different precisions.");
        double v[2];
        if (!a.isDensityMatrix && !b.isDensityMatrix) {
            check(qa_statevec_calcInnerProduct(dev(a), dev(b), precOf(a), a.numQubits, v));
            return v[0] * v[0] + v[1] * v[1];
        }
        const Qureg& rho = a.isDensityMatrix ? a : b;
        const Qureg& psi = a.isDensityMatrix ? b : a;
        check(qa_densmatr_calcFidelityWithPureState(dev(rho), dev(psi), precOf(a), rho.numQubits, 0, v));
        return v[0];
    });
}

qreal calcDistance(Qureg a, Qureg b) {
    QAH_API({
        validateQureg(a);
        validateQureg(b);
        validateSameDims(a, b);
        if (precOf(a) != precOf(b)) fail("The Quregs have different precisions.");
        double v[2];
        if (a.isDensityMatrix && b.isDensityMatrix) {
            check(qa_densmatr_calcHilbertSchmidtDistance(dev(a), dev(b), precOf(a), a.numQubits, v));
            return std::sqrt(v[0]);
        }
        if (!a.isDensityMatrix && !b.isDensityMatrix) {
            check(qa_statevec_calcInnerProduct(dev(a), dev(b), precOf(a), a.numQubits, v));
            double mag = std::sqrt(v[0] * v[0] + v[1] * v[1]);
            mag = mag > 1 ? 1 : mag;
            return std::sqrt(2 - 2 * mag);
        }
        const Qureg& rho = a.isDensityMatrix ? a : b;
        const Qureg& psi = a.isDensityMatrix ? b : a;
        check(qa_densmatr_calcFidelityWithPureState(dev(rho), dev(psi), precOf(a), rho.numQubits, 0, v));
        double re = v[0] > 1 ? 1 : v[0];
        return std::sqrt(1 - re);
    });
}

Qureg createCustomPrecisionQureg(int numQubits, int isDensMatr, int precision);
void setQuregToPartialTrace(Qureg out, Qureg in, int* traceOutQubits, int numTraceQubits);

Qureg calcPartialTrace(Qureg q, int* trace, int n) {
    QAH_API({
        validateDensMatr(q);
        validateTargets(q, trace, n);
        if (n >= q.numQubits) fail("Cannot trace out every qubit.");
        Qureg out = makeQureg(q.numQubits - n, 1, precOf(q));
        check(qa_densmatr_partialTrace(dev(q), precOf(q), q.numQubits, dev(out), trace, n));
        return out;
    });
}

Qureg calcReducedDensityMatrix(Qureg q, int* retain, int n) {
    QAH_API({
        validateDensMatr(q);
        validateTargets(q, retain, n);
        vector<int> trace;
        for (int k = 0; k < q.numQubits; k++) {
            bool keep = false;
            for (int i = 0; i < n; i++) keep |= retain[i] == k;
            if (!keep) trace.push_back(k);
        }
        return calcPartialTrace(q, trace.data(), (int)trace.size());
    });
}

}  // extern "C"

qcomp calcInnerProduct(Qureg a, Qureg b) {
    qcomp out = 0;
    _wrap_calcInnerProduct(a, b, &out);
    return out;
}
qcomp calcExpecNonHermitianPauliStrSum(Qureg q, PauliStrSum sum) {
    qcomp out = 0;
    _wrap_calcExpecNonHermitianPauliStrSum(&out, q, sum);
    return out;
}
qcomp calcExpecNonHermitianFullStateDiagMatrPower(Qureg q, FullStateDiagMatr m, qcomp expo) {
    qcomp out = 0;
    _wrap_calcExpecNonHermitianFullStateDiagMatrPower(&out, q, m, expo);
    return out;
}
qcomp calcExpecNonHermitianFullStateDiagMatr(Qureg q, FullStateDiagMatr m) {
    qcomp out = 0;
    _wrap_calcExpecNonHermitianFullStateDiagMatr(&out, q, m);
    return out;
}
void mixKrausMap(Qureg q, std::vector<int> t, KrausMap map) { mixKrausMap(q, t.data(), (int)t.size(), map); }
qreal calcProbOfMultiQubitOutcome(Qureg q, std::vector<int> qs, std::vector<int> os) {
    if (qs.size() != os.size()) {
        qah::handleError(__func__, "The number of outcomes does not match the number of qubits.");
        return 0;
    }
    return calcProbOfMultiQubitOutcome(q, qs.data(), os.data(), (int)qs.size());
}
std::vector<qreal> calcProbsOfAllMultiQubitOutcomes(Qureg q, std::vector<int> qs) {
    std::vector<qreal> out((size_t)1 << qs.size());
    calcProbsOfAllMultiQubitOutcomes(out.data(), q, qs.data(), (int)qs.size());
    return out;
}
