// Unitary and non-unitary operations (quest/src/api/operations.cpp) with the
// single-node routing of quest/src/core/localiser.cpp:395-620.
//
// A density matrix rho (2N index bits) receives U rho U^dagger as two passes
// of the statevector kernels: U on the ket qubits, then conj(U) on the bra
// qubits q+N (operations.cpp:39-62).  Gadgets flip their phase on the bra
// pass unless the string has an odd number of Y (operations.cpp:1428-1477);
// an uncontrolled Pauli tensor acts on ket and bra together in one pass with
// factor (-1)^#Y.
#include <cmath>
#include <cstring>

#include "host_core.hpp"

using namespace qah;

namespace {

const double PI = 3.14159265358979323846;

const int* statesPtr(const vector<int>& s) { return s.empty() ? nullptr : s.data(); }

// ---- accelerator-level calls on index bits --------------------------------

void denseOne(const Qureg& q, const vector<int>& c, const vector<int>& s, int t, const qcomp (&m)[2][2], bool conj) {
    double d[8];
    std::memcpy(d, m, sizeof(d));
    if (conj)
        for (int k = 1; k < 8; k += 2) d[k] = -d[k];
    check(qa_statevec_anyCtrlOneTargDenseMatr(dev(q), precOf(q), idxBits(q), c.data(), statesPtr(s), (int)c.size(), t, d));
}

void denseTwo(const Qureg& q, const vector<int>& c, const vector<int>& s, int t1, int t2, const qcomp (&m)[4][4],
              bool conj) {
    double d[32];
    std::memcpy(d, m, sizeof(d));
    if (conj)
        for (int k = 1; k < 32; k += 2) d[k] = -d[k];
    check(qa_statevec_anyCtrlTwoTargDenseMatr(dev(q), precOf(q), idxBits(q), c.data(), statesPtr(s), (int)c.size(), t1,
                                              t2, d));
}

void applyMatr(const Qureg& q, const vector<int>& c, const vector<int>& s, const vector<int>& t, const CompMatr1& m,
               bool conj) {
    denseOne(q, c, s, t[0], m.elems, conj);
}
void applyMatr(const Qureg& q, const vector<int>& c, const vector<int>& s, const vector<int>& t, const CompMatr2& m,
               bool conj) {
    denseTwo(q, c, s, t[0], t[1], m.elems, conj);
}
void applyMatr(const Qureg& q, const vector<int>& c, const vector<int>& s, const vector<int>& t, const CompMatr& m,
               bool conj) {
    // localiser.cpp:473-481: one and two targets use the by-value kernels
    if (t.size() == 1) {
        qcomp e[2][2] = {{m.cpuElems[0][0], m.cpuElems[0][1]}, {m.cpuElems[1][0], m.cpuElems[1][1]}};
        denseOne(q, c, s, t[0], e, conj);
    } else if (t.size() == 2) {
        qcomp e[4][4];
        for (int r = 0; r < 4; r++)
            for (int k = 0; k < 4; k++) e[r][k] = m.cpuElems[r][k];
        denseTwo(q, c, s, t[0], t[1], e, conj);
    } else {
        check(qa_statevec_anyCtrlAnyTargDenseMatr(dev(q), precOf(q), idxBits(q), c.data(), statesPtr(s), (int)c.size(),
                                                  t.data(), (int)t.size(), m.gpuElemsFlat, conj ? 1 : 0));
    }
}
void applyMatr(const Qureg& q, const vector<int>& c, const vector<int>& s, const vector<int>& t, const DiagMatr1& m,
               bool conj) {
    double d[4] = {m.elems[0].real(), m.elems[0].imag(), m.elems[1].real(), m.elems[1].imag()};
    if (conj) d[1] = -d[1], d[3] = -d[3];
    check(qa_statevec_anyCtrlOneTargDiagMatr(dev(q), precOf(q), idxBits(q), c.data(), statesPtr(s), (int)c.size(), t[0], d));
}
void applyMatr(const Qureg& q, const vector<int>& c, const vector<int>& s, const vector<int>& t, const DiagMatr2& m,
               bool conj) {
    double d[8];
    std::memcpy(d, m.elems, sizeof(d));
    if (conj)
        for (int k = 1; k < 8; k += 2) d[k] = -d[k];
    check(qa_statevec_anyCtrlTwoTargDiagMatr(dev(q), precOf(q), idxBits(q), c.data(), statesPtr(s), (int)c.size(), t[0],
                                             t[1], d));
}
void applyDiagPow(const Qureg& q, const vector<int>& c, const vector<int>& s, const vector<int>& t, const DiagMatr& m,
                  qcomp expo, bool hasPower, bool conj) {
    check(qa_statevec_anyCtrlAnyTargDiagMatr(dev(q), precOf(q), idxBits(q), c.data(), statesPtr(s), (int)c.size(),
                                             t.data(), (int)t.size(), m.gpuElems, conj ? 1 : 0, hasPower ? 1 : 0,
                                             expo.real(), expo.imag()));
}
void applyMatr(const Qureg& q, const vector<int>& c, const vector<int>& s, const vector<int>& t, const DiagMatr& m,
               bool conj) {
    applyDiagPow(q, c, s, t, m, 1, false, conj);
}

// ---- unitarity checks --------------------------------------------------------
void requireUnitary(const CompMatr1& m) {
    if (validating() && !isUnitary(&m.elems[0][0], 2)) fail("The given matrix was not (approximately) unitary.");
}
void requireUnitary(const CompMatr2& m) {
    if (validating() && !isUnitary(&m.elems[0][0], 4)) fail("The given matrix was not (approximately) unitary.");
}
void requireUnitary(const CompMatr& m) {
    if (!validating()) return;
    if (!*m.wasGpuSynced) fail("The CompMatr has not been synced to the GPU (call setCompMatr or syncCompMatr).");
    if (*m.isApproxUnitary == -1) *m.isApproxUnitary = isUnitary(m.cpuElemsFlat, m.numRows) ? 1 : 0;
    if (!*m.isApproxUnitary) fail("The given matrix was not (approximately) unitary.");
}
void requireUnitary(const DiagMatr1& m) {
    if (validating() && !isUnitaryDiag(m.elems, 2)) fail("The given matrix was not (approximately) unitary.");
}
void requireUnitary(const DiagMatr2& m) {
    if (validating() && !isUnitaryDiag(m.elems, 4)) fail("The given matrix was not (approximately) unitary.");
}
void requireUnitary(const DiagMatr& m) {
    if (!validating()) return;
    if (!*m.wasGpuSynced) fail("The DiagMatr has not been synced to the GPU.");
    if (*m.isApproxUnitary == -1) *m.isApproxUnitary = isUnitaryDiag(m.cpuElems, m.numElems) ? 1 : 0;
    if (!*m.isApproxUnitary) fail("The given matrix was not (approximately) unitary.");
}

int numTargsOf(const CompMatr1&) { return 1; }
int numTargsOf(const CompMatr2&) { return 2; }
int numTargsOf(const CompMatr& m) { return m.numQubits; }
int numTargsOf(const DiagMatr1&) { return 1; }
int numTargsOf(const DiagMatr2&) { return 2; }
int numTargsOf(const DiagMatr& m) { return m.numQubits; }

// operations.cpp:18-37 validateAndApplyAnyCtrlAnyTargUnitaryMatrix
template <class M>
void applyUnitary(const Qureg& q, const int* ctrls, const int* states, int nc, const int* targs, int nt, const M& m) {
    validateCtrlsAndTargs(q, ctrls, nc, targs, nt);
    validateCtrlStates(states, nc);
    if (numTargsOf(m) != nt) fail("The matrix size does not match the number of target qubits.");
    requireUnitary(m);
    auto c = vec(ctrls, nc), s = states ? vec(states, nc) : vector<int>(), t = vec(targs, nt);
    applyMatr(q, c, s, t, m, false);
    if (!q.isDensityMatrix) return;
    applyMatr(q, braOf(c, q), s, braOf(t, q), m, true);
}

// non-unitary left-multiplication (multiply*): ket only, no validation of unitarity
template <class M>
void multiplyOnly(const Qureg& q, const int* targs, int nt, const M& m) {
    validateTargets(q, targs, nt);
    if (numTargsOf(m) != nt) fail("The matrix size does not match the number of target qubits.");
    applyMatr(q, {}, {}, vec(targs, nt), m, false);
}

// ---- Pauli tensors and gadgets (localiser.cpp:559-620) ---------------------------

void ctrlZTensorOrGadget(const Qureg& q, const vector<int>& c, const vector<int>& s, uint64_t targMask, bool isGadget,
                         double phase) {
    qcomp f0 = isGadget ? std::exp(qcomp(0, +phase)) : qcomp(+1, 0);
    qcomp f1 = isGadget ? std::exp(qcomp(0, -phase)) : qcomp(-1, 0);
    check(qa_statevec_anyCtrlAnyTargZOrPhaseGadget(dev(q), precOf(q), idxBits(q), c.data(), statesPtr(s), (int)c.size(),
                                                   targMask, f0.real(), f0.imag(), f1.real(), f1.imag()));
}

void ctrlPauliTensorOrGadget(const Qureg& q, const vector<int>& c, const vector<int>& s, PauliStr str, qcomp ampFac,
                             qcomp pairFac) {
    uint64_t mx, my, mz;
    pauliMasks(str, mx, my, mz);
    check(qa_statevec_anyCtrlPauliTensorOrGadget(dev(q), precOf(q), idxBits(q), c.data(), statesPtr(s), (int)c.size(), mx,
                                                 my, mz, ampFac.real(), ampFac.imag(), pairFac.real(), pairFac.imag()));
}

void ctrlPauliTensor(const Qureg& q, const vector<int>& c, const vector<int>& s, PauliStr str, qcomp factor) {
    uint64_t mx, my, mz;
    pauliMasks(str, mx, my, mz);
    if (mx | my) {
        ctrlPauliTensorOrGadget(q, c, s, str, qcomp(0, 0) * factor, qcomp(1, 0) * factor);
    } else {
        if (factor != qcomp(1, 0)) fail("Internal error: a Z tensor received a non-unity factor.");
        ctrlZTensorOrGadget(q, c, s, mz, false, 0);
    }
}

void ctrlPauliGadget(const Qureg& q, const vector<int>& c, const vector<int>& s, PauliStr str, double phase) {
    uint64_t mx, my, mz;
    pauliMasks(str, mx, my, mz);
    if (!(mx | my)) {
        ctrlZTensorOrGadget(q, c, s, mz, true, phase);
        return;
    }
    ctrlPauliTensorOrGadget(q, c, s, str, qcomp(std::cos(phase), 0), qcomp(0, std::sin(phase)));
}

void validatePauliTargets(const Qureg& q, PauliStr str, const int* ctrls, int nc) {
    validateQureg(q);
    int top = pauliMaxIndex(str);
    if (top >= q.numQubits)
        fail("The Pauli string targets qubit " + std::to_string(top) + " beyond the " + std::to_string(q.numQubits) +
             "-qubit Qureg.");
    if (nc < 0 || (nc > 0 && !ctrls)) fail("Invalid control qubits.");
    for (int i = 0; i < nc; i++) {
        if (ctrls[i] < 0 || ctrls[i] >= q.numQubits) fail("Invalid control qubit (" + std::to_string(ctrls[i]) + ").");
        if (pauliAt(str, ctrls[i])) fail("A control qubit overlaps a non-identity Pauli.");
        for (int j = 0; j < i; j++)
            if (ctrls[i] == ctrls[j]) fail("The control qubits contained duplicates.");
    }
}

double phaseFromAngle(double angle) { return -angle / 2; }  // utilities.cpp:911-914

void pauliGadgetImpl(const Qureg& q, const int* ctrls, const int* states, int nc, PauliStr str, double angle) {
    validatePauliTargets(q, str, ctrls, nc);
    validateCtrlStates(states, nc);
    if (str.lowPaulis == 0 && str.highPaulis == 0 && nc == 0 && q.isDensityMatrix) return;
    double phase = phaseFromAngle(angle);
    auto c = vec(ctrls, nc), s = states ? vec(states, nc) : vector<int>();
    ctrlPauliGadget(q, c, s, str, phase);
    if (!q.isDensityMatrix) return;
    phase *= pauliHasOddY(str) ? 1 : -1;
    ctrlPauliGadget(q, braOf(c, q), s, pauliShifted(str, q.numQubits), phase);
}

void pauliStrImpl(const Qureg& q, const int* ctrls, const int* states, int nc, PauliStr str) {
    validatePauliTargets(q, str, ctrls, nc);
    validateCtrlStates(states, nc);
    qcomp factor = 1;
    auto c = vec(ctrls, nc), s = states ? vec(states, nc) : vector<int>();
    if (q.isDensityMatrix && nc == 0) {
        factor = pauliHasOddY(str) ? -1 : 1;
        PauliStr sh = pauliShifted(str, q.numQubits);
        str.lowPaulis |= sh.lowPaulis;
        str.highPaulis |= sh.highPaulis;
    }
    ctrlPauliTensor(q, c, s, str, factor);
    if (q.isDensityMatrix && nc > 0) {
        factor = pauliHasOddY(str) ? -1 : 1;
        ctrlPauliTensor(q, braOf(c, q), s, pauliShifted(str, q.numQubits), factor);
    }
}

void phaseGadgetImpl(const Qureg& q, const int* ctrls, const int* states, int nc, const int* targs, int nt,
                     double angle) {
    validateCtrlsAndTargs(q, ctrls, nc, targs, nt);
    validateCtrlStates(states, nc);
    double phase = phaseFromAngle(angle);
    auto c = vec(ctrls, nc), s = states ? vec(states, nc) : vector<int>();
    uint64_t mask = 0;
    for (int i = 0; i < nt; i++) mask |= 1ull << targs[i];
    ctrlZTensorOrGadget(q, c, s, mask, true, phase);
    if (!q.isDensityMatrix) return;
    ctrlZTensorOrGadget(q, braOf(c, q), s, mask << q.numQubits, true, -phase);
}

void swapImpl(const Qureg& q, const int* ctrls, const int* states, int nc, int a, int b) {
    int t[2] = {a, b};
    validateCtrlsAndTargs(q, ctrls, nc, t, 2);
    validateCtrlStates(states, nc);
    auto c = vec(ctrls, nc), s = states ? vec(states, nc) : vector<int>();
    if (a > b) std::swap(a, b);
    check(qa_statevec_anyCtrlSwap(dev(q), precOf(q), idxBits(q), c.data(), statesPtr(s), nc, a, b));
    if (!q.isDensityMatrix) return;
    auto cb = braOf(c, q);
    check(qa_statevec_anyCtrlSwap(dev(q), precOf(q), idxBits(q), cb.data(), statesPtr(s), nc, a + q.numQubits,
                                  b + q.numQubits));
}

CompMatr1 m1(qcomp a, qcomp b, qcomp c, qcomp d) {
    CompMatr1 m;
    m.numQubits = 1;
    m.numRows = 2;
    m.elems[0][0] = a;
    m.elems[0][1] = b;
    m.elems[1][0] = c;
    m.elems[1][1] = d;
    return m;
}
DiagMatr1 d1(qcomp a, qcomp b) {
    DiagMatr1 m;
    m.numQubits = 1;
    m.numElems = 2;
    m.elems[0] = a;
    m.elems[1] = b;
    return m;
}

PauliStr singlePauli(int code, int target) {
    PauliStr s{0, 0};
    if (target < 32) s.lowPaulis = (PAULI_MASK_TYPE)code << (2 * target);
    else s.highPaulis = (PAULI_MASK_TYPE)code << (2 * (target - 32));
    return s;
}

// projector / measurement helpers (localiser.cpp:1031-1045)
void projector(const Qureg& q, const vector<int>& qubits, const vector<int>& outcomes, double prob) {
    if (q.isDensityMatrix)
        check(qa_densmatr_multiQubitProjector(dev(q), precOf(q), q.numQubits, qubits.data(), outcomes.data(),
                                              (int)qubits.size(), prob));
    else
        check(qa_statevec_multiQubitProjector(dev(q), precOf(q), q.numQubits, qubits.data(), outcomes.data(),
                                              (int)qubits.size(), prob));
}

}  // namespace

extern "C" {

// ---- dense matrices -----------------------------------------------------------

void multiplyCompMatr1(Qureg q, int t, CompMatr1 m) { QAH_API({ multiplyOnly(q, &t, 1, m); }); }
void applyCompMatr1(Qureg q, int t, CompMatr1 m) { QAH_API({ applyUnitary(q, nullptr, nullptr, 0, &t, 1, m); }); }
void applyControlledCompMatr1(Qureg q, int c, int t, CompMatr1 m) { QAH_API({ applyUnitary(q, &c, nullptr, 1, &t, 1, m); }); }
void applyMultiControlledCompMatr1(Qureg q, int* c, int nc, int t, CompMatr1 m) {
    QAH_API({ applyUnitary(q, c, nullptr, nc, &t, 1, m); });
}
void applyMultiStateControlledCompMatr1(Qureg q, int* c, int* s, int nc, int t, CompMatr1 m) {
    QAH_API({ applyUnitary(q, c, s, nc, &t, 1, m); });
}

void multiplyCompMatr2(Qureg q, int t1, int t2, CompMatr2 m) {
    QAH_API({
        int t[2] = {t1, t2};
        multiplyOnly(q, t, 2, m);
    });
}
void applyCompMatr2(Qureg q, int t1, int t2, CompMatr2 m) {
    QAH_API({
        int t[2] = {t1, t2};
        applyUnitary(q, nullptr, nullptr, 0, t, 2, m);
    });
}
void applyControlledCompMatr2(Qureg q, int c, int t1, int t2, CompMatr2 m) {
    QAH_API({
        int t[2] = {t1, t2};
        applyUnitary(q, &c, nullptr, 1, t, 2, m);
    });
}
void applyMultiControlledCompMatr2(Qureg q, int* c, int nc, int t1, int t2, CompMatr2 m) {
    QAH_API({
        int t[2] = {t1, t2};
        applyUnitary(q, c, nullptr, nc, t, 2, m);
    });
}
void applyMultiStateControlledCompMatr2(Qureg q, int* c, int* s, int nc, int t1, int t2, CompMatr2 m) {
    QAH_API({
        int t[2] = {t1, t2};
        applyUnitary(q, c, s, nc, t, 2, m);
    });
}

void multiplyCompMatr(Qureg q, int* t, int nt, CompMatr m) {
    QAH_API({
        if (!*m.wasGpuSynced) fail("The CompMatr has not been synced to the GPU.");
        multiplyOnly(q, t, nt, m);
    });
}
void applyCompMatr(Qureg q, int* t, int nt, CompMatr m) { QAH_API({ applyUnitary(q, nullptr, nullptr, 0, t, nt, m); }); }
void applyControlledCompMatr(Qureg q, int c, int* t, int nt, CompMatr m) {
    QAH_API({ applyUnitary(q, &c, nullptr, 1, t, nt, m); });
}
void applyMultiControlledCompMatr(Qureg q, int* c, int nc, int* t, int nt, CompMatr m) {
    QAH_API({ applyUnitary(q, c, nullptr, nc, t, nt, m); });
}
void applyMultiStateControlledCompMatr(Qureg q, int* c, int* s, int nc, int* t, int nt, CompMatr m) {
    QAH_API({ applyUnitary(q, c, s, nc, t, nt, m); });
}

// ---- diagonal matrices ----------------------------------------------------------

void multiplyDiagMatr1(Qureg q, int t, DiagMatr1 m) { QAH_API({ multiplyOnly(q, &t, 1, m); }); }
void applyDiagMatr1(Qureg q, int t, DiagMatr1 m) { QAH_API({ applyUnitary(q, nullptr, nullptr, 0, &t, 1, m); }); }
void applyControlledDiagMatr1(Qureg q, int c, int t, DiagMatr1 m) { QAH_API({ applyUnitary(q, &c, nullptr, 1, &t, 1, m); }); }
void applyMultiControlledDiagMatr1(Qureg q, int* c, int nc, int t, DiagMatr1 m) {
    QAH_API({ applyUnitary(q, c, nullptr, nc, &t, 1, m); });
}
void applyMultiStateControlledDiagMatr1(Qureg q, int* c, int* s, int nc, int t, DiagMatr1 m) {
    QAH_API({ applyUnitary(q, c, s, nc, &t, 1, m); });
}
void multiplyDiagMatr2(Qureg q, int t1, int t2, DiagMatr2 m) {
    QAH_API({
        int t[2] = {t1, t2};
        multiplyOnly(q, t, 2, m);
    });
}
void applyDiagMatr2(Qureg q, int t1, int t2, DiagMatr2 m) {
    QAH_API({
        int t[2] = {t1, t2};
        applyUnitary(q, nullptr, nullptr, 0, t, 2, m);
    });
}
void applyControlledDiagMatr2(Qureg q, int c, int t1, int t2, DiagMatr2 m) {
    QAH_API({
        int t[2] = {t1, t2};
        applyUnitary(q, &c, nullptr, 1, t, 2, m);
    });
}
void applyMultiControlledDiagMatr2(Qureg q, int* c, int nc, int t1, int t2, DiagMatr2 m) {
    QAH_API({
        int t[2] = {t1, t2};
        applyUnitary(q, c, nullptr, nc, t, 2, m);
    });
}
void applyMultiStateControlledDiagMatr2(Qureg q, int* c, int* s, int nc, int t1, int t2, DiagMatr2 m) {
    QAH_API({
        int t[2] = {t1, t2};
        applyUnitary(q, c, s, nc, t, 2, m);
    });
}
void multiplyDiagMatr(Qureg q, int* t, int nt, DiagMatr m) { QAH_API({ multiplyOnly(q, t, nt, m); }); }
void applyDiagMatr(Qureg q, int* t, int nt, DiagMatr m) { QAH_API({ applyUnitary(q, nullptr, nullptr, 0, t, nt, m); }); }
void applyControlledDiagMatr(Qureg q, int c, int* t, int nt, DiagMatr m) {
    QAH_API({ applyUnitary(q, &c, nullptr, 1, t, nt, m); });
}
void applyMultiControlledDiagMatr(Qureg q, int* c, int nc, int* t, int nt, DiagMatr m) {
    QAH_API({ applyUnitary(q, c, nullptr, nc, t, nt, m); });
}
void applyMultiStateControlledDiagMatr(Qureg q, int* c, int* s, int nc, int* t, int nt, DiagMatr m) {
    QAH_API({ applyUnitary(q, c, s, nc, t, nt, m); });
}

// ---- DiagMatr powers (operations.cpp:405-500, 545-576) --------------------------------

namespace {

// validation.cpp:2408-2431: a strictly real, negative exponent needs every
// element of (approximately) non-zero magnitude
void requireNonDiverging(const qcomp* elems, qindex num, qcomp expo) {
    if (!validating() || expo.imag() != 0 || expo.real() >= 0) return;
    for (qindex i = 0; i < num; i++)
        if (std::abs(elems[i]) <= epsilon())
            fail("The diagonal matrix has an element of (approximately) zero magnitude, which diverges under the "
                 "real, negative exponent.");
}

// validation.cpp:2367-2379
void requireRealExponent(qcomp expo) {
    if (validating() && std::abs(expo.imag()) > epsilon())
        fail("The exponent of a unitary diagonal matrix must be (approximately) real.");
}

void requireSynced(const DiagMatr& m) {
    if (!*m.wasGpuSynced) fail("The DiagMatr has not been synced to the GPU.");
}

void requireFullState(const Qureg& q, const FullStateDiagMatr& m) {
    validateQureg(q);
    if (m.numQubits != q.numQubits) fail("The FullStateDiagMatr and Qureg have different numbers of qubits.");
    if (!*m.wasGpuSynced) fail("The FullStateDiagMatr has not been synced to the GPU.");
}

void requireUnitaryFullState(FullStateDiagMatr& m) {
    if (!validating()) return;
    if (*m.isApproxUnitary == -1) *m.isApproxUnitary = isUnitaryDiag(m.cpuElems, m.numElems) ? 1 : 0;
    if (!*m.isApproxUnitary) fail("The given matrix was not (approximately) unitary.");
}

}  // namespace

void multiplyDiagMatrPower(Qureg q, int* t, int nt, DiagMatr m, qcomp expo) {
    QAH_API({
        validateTargets(q, t, nt);
        if (m.numQubits != nt) fail("The matrix size does not match the number of target qubits.");
        requireSynced(m);
        requireNonDiverging(m.cpuElems, m.numElems, expo);
        applyDiagPow(q, {}, {}, vec(t, nt), m, expo, true, false);
    });
}
// d^expo on the kets, conj(d^expo) on the bras of a density matrix
void applyMultiStateControlledDiagMatrPower(Qureg q, int* c, int* s, int nc, int* t, int nt, DiagMatr m,
                                            qcomp expo) {
    QAH_API({
        validateCtrlsAndTargs(q, c, nc, t, nt);
        validateCtrlStates(s, nc);
        if (m.numQubits != nt) fail("The matrix size does not match the number of target qubits.");
        requireSynced(m);
        requireUnitary(m);
        requireNonDiverging(m.cpuElems, m.numElems, expo);
        requireRealExponent(expo);
        auto cv = vec(c, nc), sv = s ? vec(s, nc) : vector<int>(), tv = vec(t, nt);
        applyDiagPow(q, cv, sv, tv, m, expo, true, false);
        if (!q.isDensityMatrix) return;
        applyDiagPow(q, braOf(cv, q), sv, braOf(tv, q), m, expo, true, true);
    });
}
void applyDiagMatrPower(Qureg q, int* t, int nt, DiagMatr m, qcomp expo) {
    applyMultiStateControlledDiagMatrPower(q, nullptr, nullptr, 0, t, nt, m, expo);
}
void applyControlledDiagMatrPower(Qureg q, int c, int* t, int nt, DiagMatr m, qcomp expo) {
    applyMultiStateControlledDiagMatrPower(q, &c, nullptr, 1, t, nt, m, expo);
}
void applyMultiControlledDiagMatrPower(Qureg q, int* c, int nc, int* t, int nt, DiagMatr m, qcomp expo) {
    applyMultiStateControlledDiagMatrPower(q, c, nullptr, nc, t, nt, m, expo);
}

void multiplyFullStateDiagMatrPower(Qureg q, FullStateDiagMatr m, qcomp expo) {
    QAH_API({
        requireFullState(q, m);
        requireNonDiverging(m.cpuElems, m.numElems, expo);
        if (q.isDensityMatrix)
            check(qa_densmatr_allTargDiagMatr(dev(q), precOf(q), q.numQubits, m.gpuElems, 1, expo.real(), expo.imag(),
                                              1));
        else
            check(qa_statevec_allTargDiagMatr(dev(q), precOf(q), q.numQubits, m.gpuElems, 1, expo.real(),
                                              expo.imag()));
    });
}
void applyFullStateDiagMatrPower(Qureg q, FullStateDiagMatr m, qcomp expo) {
    QAH_API({
        requireFullState(q, m);
        requireUnitaryFullState(m);
        requireRealExponent(expo);
        requireNonDiverging(m.cpuElems, m.numElems, expo);
        if (q.isDensityMatrix)
            check(qa_densmatr_allTargDiagMatr(dev(q), precOf(q), q.numQubits, m.gpuElems, 1, expo.real(), expo.imag(),
                                              0));
        else
            check(qa_statevec_allTargDiagMatr(dev(q), precOf(q), q.numQubits, m.gpuElems, 1, expo.real(),
                                              expo.imag()));
    });
}

void multiplyFullStateDiagMatr(Qureg q, FullStateDiagMatr m) {
    QAH_API({
        validateQureg(q);
        if (m.numQubits != q.numQubits) fail("The FullStateDiagMatr and Qureg have different numbers of qubits.");
        if (q.isDensityMatrix)
            check(qa_densmatr_allTargDiagMatr(dev(q), precOf(q), q.numQubits, m.gpuElems, 0, 1, 0, 1));
        else
            check(qa_statevec_allTargDiagMatr(dev(q), precOf(q), q.numQubits, m.gpuElems, 0, 1, 0));
    });
}
void applyFullStateDiagMatr(Qureg q, FullStateDiagMatr m) {
    QAH_API({
        validateQureg(q);
        if (m.numQubits != q.numQubits) fail("The FullStateDiagMatr and Qureg have different numbers of qubits.");
        if (validating()) {
            if (*m.isApproxUnitary == -1) *m.isApproxUnitary = isUnitaryDiag(m.cpuElems, m.numElems) ? 1 : 0;
            if (!*m.isApproxUnitary) fail("The given matrix was not (approximately) unitary.");
        }
        if (q.isDensityMatrix)
            check(qa_densmatr_allTargDiagMatr(dev(q), precOf(q), q.numQubits, m.gpuElems, 0, 1, 0, 0));
        else
            check(qa_statevec_allTargDiagMatr(dev(q), precOf(q), q.numQubits, m.gpuElems, 0, 1, 0));
    });
}

// ---- named single-qubit gates (operations.cpp:346-430) ------------------------

void applyMultiStateControlledS(Qureg q, int* c, int* s, int nc, int t) {
    QAH_API({ applyUnitary(q, c, s, nc, &t, 1, d1(1, qcomp(0, 1))); });
}
void applyS(Qureg q, int t) { applyMultiStateControlledS(q, nullptr, nullptr, 0, t); }
void applyControlledS(Qureg q, int c, int t) { applyMultiStateControlledS(q, &c, nullptr, 1, t); }
void applyMultiControlledS(Qureg q, int* c, int nc, int t) { applyMultiStateControlledS(q, c, nullptr, nc, t); }

void applyMultiStateControlledT(Qureg q, int* c, int* s, int nc, int t) {
    QAH_API({ applyUnitary(q, c, s, nc, &t, 1, d1(1, qcomp(1 / std::sqrt(2), 0) + qcomp(0, 1) / std::sqrt(2))); });
}
void applyT(Qureg q, int t) { applyMultiStateControlledT(q, nullptr, nullptr, 0, t); }
void applyControlledT(Qureg q, int c, int t) { applyMultiStateControlledT(q, &c, nullptr, 1, t); }
void applyMultiControlledT(Qureg q, int* c, int nc, int t) { applyMultiStateControlledT(q, c, nullptr, nc, t); }

void applyMultiStateControlledHadamard(Qureg q, int* c, int* s, int nc, int t) {
    QAH_API({
        qcomp a = 1 / std::sqrt(2);
        applyUnitary(q, c, s, nc, &t, 1, m1(a, a, a, -a));
    });
}
void applyHadamard(Qureg q, int t) { applyMultiStateControlledHadamard(q, nullptr, nullptr, 0, t); }
void applyControlledHadamard(Qureg q, int c, int t) { applyMultiStateControlledHadamard(q, &c, nullptr, 1, t); }
void applyMultiControlledHadamard(Qureg q, int* c, int nc, int t) { applyMultiStateControlledHadamard(q, c, nullptr, nc, t); }

// ---- swaps ------------------------------------------------------------------------

void multiplySwap(Qureg q, int a, int b) {
    QAH_API({
        int t[2] = {a, b};
        validateTargets(q, t, 2);
        if (a > b) std::swap(a, b);
        check(qa_statevec_anyCtrlSwap(dev(q), precOf(q), idxBits(q), nullptr, nullptr, 0, a, b));
    });
}
void applyMultiStateControlledSwap(Qureg q, int* c, int* s, int nc, int a, int b) {
    QAH_API({ swapImpl(q, c, s, nc, a, b); });
}
void applySwap(Qureg q, int a, int b) { applyMultiStateControlledSwap(q, nullptr, nullptr, 0, a, b); }
void applyControlledSwap(Qureg q, int c, int a, int b) { applyMultiStateControlledSwap(q, &c, nullptr, 1, a, b); }
void applyMultiControlledSwap(Qureg q, int* c, int nc, int a, int b) { applyMultiStateControlledSwap(q, c, nullptr, nc, a, b); }

// operations.cpp:841-860: the dense two-qubit root of SWAP
void applyMultiStateControlledSqrtSwap(Qureg q, int* c, int* s, int nc, int a, int b) {
    QAH_API({
        CompMatr2 m;
        m.numQubits = 2;
        m.numRows = 4;
        for (int r = 0; r < 4; r++)
            for (int k = 0; k < 4; k++) m.elems[r][k] = 0;
        m.elems[0][0] = m.elems[3][3] = 1;
        m.elems[1][1] = m.elems[2][2] = qcomp(.5, .5);
        m.elems[1][2] = m.elems[2][1] = qcomp(.5, -.5);
        int t[2] = {a, b};
        applyUnitary(q, c, s, nc, t, 2, m);
    });
}
void applyControlledSqrtSwap(Qureg q, int c, int a, int b) { applyMultiStateControlledSqrtSwap(q, &c, nullptr, 1, a, b); }
void applyMultiControlledSqrtSwap(Qureg q, int* c, int nc, int a, int b) {
    applyMultiStateControlledSqrtSwap(q, c, nullptr, nc, a, b);
}
void applySqrtSwap(Qureg q, int a, int b) {
    QAH_API({ applyMultiStateControlledSqrtSwap(q, nullptr, nullptr, 0, a, b); });
}

// ---- Paulis -------------------------------------------------------------------------

void applyMultiStateControlledPauliX(Qureg q, int* c, int* s, int nc, int t) {
    QAH_API({
        validateCtrlsAndTargs(q, c, nc, &t, 1);
        pauliStrImpl(q, c, s, nc, singlePauli(1, t));
    });
}
void applyMultiStateControlledPauliY(Qureg q, int* c, int* s, int nc, int t) {
    QAH_API({
        validateCtrlsAndTargs(q, c, nc, &t, 1);
        pauliStrImpl(q, c, s, nc, singlePauli(2, t));
    });
}
void applyMultiStateControlledPauliZ(Qureg q, int* c, int* s, int nc, int t) {
    QAH_API({
        validateCtrlsAndTargs(q, c, nc, &t, 1);
        // operations.cpp:569-575: a diagonal matrix diag(1,-1)
        applyUnitary(q, c, s, nc, &t, 1, d1(1, -1));
    });
}
void applyPauliX(Qureg q, int t) { applyMultiStateControlledPauliX(q, nullptr, nullptr, 0, t); }
void applyPauliY(Qureg q, int t) { applyMultiStateControlledPauliY(q, nullptr, nullptr, 0, t); }
void applyPauliZ(Qureg q, int t) { applyMultiStateControlledPauliZ(q, nullptr, nullptr, 0, t); }
void applyControlledPauliX(Qureg q, int c, int t) { applyMultiStateControlledPauliX(q, &c, nullptr, 1, t); }
void applyControlledPauliY(Qureg q, int c, int t) { applyMultiStateControlledPauliY(q, &c, nullptr, 1, t); }
void applyControlledPauliZ(Qureg q, int c, int t) { applyMultiStateControlledPauliZ(q, &c, nullptr, 1, t); }
void applyMultiControlledPauliX(Qureg q, int* c, int nc, int t) { applyMultiStateControlledPauliX(q, c, nullptr, nc, t); }
void applyMultiControlledPauliY(Qureg q, int* c, int nc, int t) { applyMultiStateControlledPauliY(q, c, nullptr, nc, t); }
void applyMultiControlledPauliZ(Qureg q, int* c, int nc, int t) { applyMultiStateControlledPauliZ(q, c, nullptr, nc, t); }

void multiplyPauliStr(Qureg q, PauliStr str) {
    QAH_API({
        validatePauliTargets(q, str, nullptr, 0);
        ctrlPauliTensor(q, {}, {}, str, 1);
    });
}
void applyMultiStateControlledPauliStr(Qureg q, int* c, int* s, int nc, PauliStr str) {
    QAH_API({ pauliStrImpl(q, c, s, nc, str); });
}
void applyPauliStr(Qureg q, PauliStr str) { applyMultiStateControlledPauliStr(q, nullptr, nullptr, 0, str); }
void applyControlledPauliStr(Qureg q, int c, PauliStr str) { applyMultiStateControlledPauliStr(q, &c, nullptr, 1, str); }
void applyMultiControlledPauliStr(Qureg q, int* c, int nc, PauliStr str) {
    applyMultiStateControlledPauliStr(q, c, nullptr, nc, str);
}

// ---- rotations (gadgets on single Paulis, operations.cpp:750-767) --------------

void applyMultiStateControlledRotateX(Qureg q, int* c, int* s, int nc, int t, qreal a) {
    QAH_API({
        validateCtrlsAndTargs(q, c, nc, &t, 1);
        pauliGadgetImpl(q, c, s, nc, singlePauli(1, t), a);
    });
}
void applyMultiStateControlledRotateY(Qureg q, int* c, int* s, int nc, int t, qreal a) {
    QAH_API({
        validateCtrlsAndTargs(q, c, nc, &t, 1);
        pauliGadgetImpl(q, c, s, nc, singlePauli(2, t), a);
    });
}
void applyMultiStateControlledRotateZ(Qureg q, int* c, int* s, int nc, int t, qreal a) {
    QAH_API({
        validateCtrlsAndTargs(q, c, nc, &t, 1);
        pauliGadgetImpl(q, c, s, nc, singlePauli(3, t), a);
    });
}
void applyRotateX(Qureg q, int t, qreal a) { applyMultiStateControlledRotateX(q, nullptr, nullptr, 0, t, a); }
void applyRotateY(Qureg q, int t, qreal a) { applyMultiStateControlledRotateY(q, nullptr, nullptr, 0, t, a); }
void applyRotateZ(Qureg q, int t, qreal a) { applyMultiStateControlledRotateZ(q, nullptr, nullptr, 0, t, a); }
void applyControlledRotateX(Qureg q, int c, int t, qreal a) { applyMultiStateControlledRotateX(q, &c, nullptr, 1, t, a); }
void applyControlledRotateY(Qureg q, int c, int t, qreal a) { applyMultiStateControlledRotateY(q, &c, nullptr, 1, t, a); }
void applyControlledRotateZ(Qureg q, int c, int t, qreal a) { applyMultiStateControlledRotateZ(q, &c, nullptr, 1, t, a); }
void applyMultiControlledRotateX(Qureg q, int* c, int nc, int t, qreal a) {
    applyMultiStateControlledRotateX(q, c, nullptr, nc, t, a);
}
void applyMultiControlledRotateY(Qureg q, int* c, int nc, int t, qreal a) {
    applyMultiStateControlledRotateY(q, c, nullptr, nc, t, a);
}
void applyMultiControlledRotateZ(Qureg q, int* c, int nc, int t, qreal a) {
    applyMultiStateControlledRotateZ(q, c, nullptr, nc, t, a);
}

// operations.cpp:809-823
void applyMultiStateControlledRotateAroundAxis(Qureg q, int* c, int* s, int nc, int t, qreal angle, qreal x, qreal y,
                                               qreal z) {
    QAH_API({
        if (x == 0 && y == 0 && z == 0) fail("The rotation axis must not be the zero vector.");
        double norm = std::sqrt(x * x + y * y + z * z);
        double co = std::cos(angle / 2), si = std::sin(angle / 2);
        qcomp I(0, 1);
        qcomp u11 = co - (si * z * I) / norm;
        qcomp u12 = -(si * (y + x * I)) / norm;
        qcomp u21 = (si * (y - x * I)) / norm;
        qcomp u22 = co + (si * z * I) / norm;
        applyUnitary(q, c, s, nc, &t, 1, m1(u11, u12, u21, u22));
    });
}
void applyRotateAroundAxis(Qureg q, int t, qreal a, qreal x, qreal y, qreal z) {
    applyMultiStateControlledRotateAroundAxis(q, nullptr, nullptr, 0, t, a, x, y, z);
}
void applyControlledRotateAroundAxis(Qureg q, int c, int t, qreal a, qreal x, qreal y, qreal z) {
    applyMultiStateControlledRotateAroundAxis(q, &c, nullptr, 1, t, a, x, y, z);
}
void applyMultiControlledRotateAroundAxis(Qureg q, int* c, int nc, int t, qreal a, qreal x, qreal y, qreal z) {
    applyMultiStateControlledRotateAroundAxis(q, c, nullptr, nc, t, a, x, y, z);
}

// ---- gadgets ---------------------------------------------------------------------------

void multiplyPauliGadget(Qureg q, PauliStr str, qreal angle) {
    QAH_API({
        validatePauliTargets(q, str, nullptr, 0);
        ctrlPauliGadget(q, {}, {}, str, phaseFromAngle(angle));
    });
}
void applyMultiStateControlledPauliGadget(Qureg q, int* c, int* s, int nc, PauliStr str, qreal angle) {
    QAH_API({ pauliGadgetImpl(q, c, s, nc, str, angle); });
}
void applyPauliGadget(Qureg q, PauliStr str, qreal angle) {
    applyMultiStateControlledPauliGadget(q, nullptr, nullptr, 0, str, angle);
}
void applyControlledPauliGadget(Qureg q, int c, PauliStr str, qreal angle) {
    applyMultiStateControlledPauliGadget(q, &c, nullptr, 1, str, angle);
}
void applyMultiControlledPauliGadget(Qureg q, int* c, int nc, PauliStr str, qreal angle) {
    applyMultiStateControlledPauliGadget(q, c, nullptr, nc, str, angle);
}

// ---- Pauli string sums (operations.cpp:1112-1197) ---------------------------------------

namespace {

void validatePauliSum(const Qureg& q, const PauliStrSum& sum) {
    validateQureg(q);
    if (sum.numTerms < 1 || !sum.strings || !sum.coeffs) fail("Invalid PauliStrSum.");
    for (qindex t = 0; t < sum.numTerms; t++)
        if (pauliMaxIndex(sum.strings[t]) >= q.numQubits)
            fail("The PauliStrSum targets qubits beyond the " + std::to_string(q.numQubits) + "-qubit Qureg.");
}

// operations.cpp:1133-1142: the terms in order (or reversed), each as the
// gadget exp(i angle c_j sigma_j) (gadget angle 2 angle Re c_j)
void firstOrderTrotter(const Qureg& q, const PauliStrSum& sum, double angle, bool reverse) {
    for (qindex i = 0; i < sum.numTerms; i++) {
        const qindex j = reverse ? sum.numTerms - i - 1 : i;
        pauliGadgetImpl(q, nullptr, nullptr, 0, sum.strings[j], 2 * angle * sum.coeffs[j].real());
    }
}

// operations.cpp:1144-1167.  Orders above 2 apply, as the reference does,
// five first-order sweeps in reversed term order with angles p, p, 1-4p,
// p, p (times angle), p = 1 / (4 - 4^(1/(order-1)))
void higherOrderTrotter(const Qureg& q, const PauliStrSum& sum, double angle, int order) {
    if (order == 1) {
        firstOrderTrotter(q, sum, angle, false);
    } else if (order == 2) {
        firstOrderTrotter(q, sum, angle / 2, false);
        firstOrderTrotter(q, sum, angle / 2, true);
    } else {
        const double p = 1. / (4 - std::pow(4, 1. / (order - 1)));
        const double a = p * angle, b = (1 - 4 * p) * angle;
        const bool reverse = (order - 2) != 0;
        firstOrderTrotter(q, sum, a, reverse);
        firstOrderTrotter(q, sum, a, reverse);
        firstOrderTrotter(q, sum, b, reverse);
        firstOrderTrotter(q, sum, a, reverse);
        firstOrderTrotter(q, sum, a, reverse);
    }
}

}  // namespace

// operations.cpp:1112-1131: the workspace takes a copy of qureg, qureg is
// zeroed, and each term is applied to the workspace, added into qureg with
// its coefficient and undone (Pauli strings are their own inverses); the
// strings act on the state's index bits, so on a density matrix this is
// left multiplication
void multiplyPauliStrSum(Qureg q, PauliStrSum sum, Qureg ws) {
    QAH_API({
        validatePauliSum(q, sum);
        validateQureg(ws);
        if (ws.numQubits != q.numQubits || ws.isDensityMatrix != q.isDensityMatrix || precOf(ws) != precOf(q))
            fail("The workspace Qureg is incompatible with the Qureg (its qubits, type or precision differ).");
        if (dev(ws) == dev(q)) fail("The workspace must be a different Qureg.");
        const int p = precOf(q), nb = idxBits(q);
        check(qa_statevec_setQuregToSuperposition(p, nb, 0, 0, dev(ws), 1, 0, dev(q), 0, 0, dev(q)));
        check(qa_statevec_initUniformState(dev(q), p, nb, 0, 0));
        for (qindex t = 0; t < sum.numTerms; t++) {
            ctrlPauliTensor(ws, {}, {}, sum.strings[t], 1);
            check(qa_statevec_setQuregToSuperposition(p, nb, 1, 0, dev(q), sum.coeffs[t].real(), sum.coeffs[t].imag(),
                                                      dev(ws), 0, 0, dev(ws)));
            ctrlPauliTensor(ws, {}, {}, sum.strings[t], 1);
        }
    });
}

// operations.cpp:1169-1197
void applyTrotterizedPauliStrSumGadget(Qureg q, PauliStrSum sum, qreal angle, int order, int reps) {
    QAH_API({
        validatePauliSum(q, sum);
        if (validating())
            for (qindex t = 0; t < sum.numTerms; t++)
                if (std::abs(sum.coeffs[t].imag()) > epsilon()) fail("The PauliStrSum is not Hermitian.");
        if (order <= 0 || (order % 2 != 0 && order != 1))
            fail("The Trotter order (" + std::to_string(order) + ") must be 1, or positive and even.");
        if (reps <= 0) fail("The number of Trotter repetitions (" + std::to_string(reps) + ") must be positive.");
        if (angle == 0) return;
        for (int r = 0; r < reps; r++) higherOrderTrotter(q, sum, angle / reps, order);
    });
}

void multiplyPhaseGadget(Qureg q, int* t, int nt, qreal angle) {
    QAH_API({
        validateTargets(q, t, nt);
        uint64_t mask = 0;
        for (int i = 0; i < nt; i++) mask |= 1ull << t[i];
        ctrlZTensorOrGadget(q, {}, {}, mask, true, phaseFromAngle(angle));
    });
}
void applyMultiStateControlledPhaseGadget(Qureg q, int* c, int* s, int nc, int* t, int nt, qreal angle) {
    QAH_API({ phaseGadgetImpl(q, c, s, nc, t, nt, angle); });
}
void applyPhaseGadget(Qureg q, int* t, int nt, qreal angle) {
    applyMultiStateControlledPhaseGadget(q, nullptr, nullptr, 0, t, nt, angle);
}
void applyControlledPhaseGadget(Qureg q, int c, int* t, int nt, qreal angle) {
    applyMultiStateControlledPhaseGadget(q, &c, nullptr, 1, t, nt, angle);
}
void applyMultiControlledPhaseGadget(Qureg q, int* c, int nc, int* t, int nt, qreal angle) {
    applyMultiStateControlledPhaseGadget(q, c, nullptr, nc, t, nt, angle);
}

// ---- phase shifts and flips (operations.cpp:935-975) ------------------------------------

void applyMultiQubitPhaseShift(Qureg q, int* t, int nt, qreal angle) {
    QAH_API({
        validateTargets(q, t, nt);
        applyUnitary(q, t + 1, nullptr, nt - 1, &t[0], 1, d1(1, std::exp(qcomp(0, angle))));
    });
}
void applyPhaseShift(Qureg q, int t, qreal angle) { applyMultiQubitPhaseShift(q, &t, 1, angle); }
void applyTwoQubitPhaseShift(Qureg q, int t1, int t2, qreal angle) {
    int t[2] = {t1, t2};
    applyMultiQubitPhaseShift(q, t, 2, angle);
}
void applyMultiQubitPhaseFlip(Qureg q, int* t, int nt) {
    QAH_API({
        validateTargets(q, t, nt);
        applyUnitary(q, t + 1, nullptr, nt - 1, &t[0], 1, d1(1, -1));
    });
}
void applyPhaseFlip(Qureg q, int t) { applyMultiQubitPhaseFlip(q, &t, 1); }
void applyTwoQubitPhaseFlip(Qureg q, int t1, int t2) {
    int t[2] = {t1, t2};
    applyMultiQubitPhaseFlip(q, t, 2);
}

// ---- multi-qubit NOT (a Pauli X string) ------------------------------------------------

static PauliStr xString(const Qureg& q, const int* t, int nt) {
    validateTargets(q, t, nt);
    PauliStr s{0, 0};
    for (int i = 0; i < nt; i++) {
        PauliStr one = singlePauli(1, t[i]);
        s.lowPaulis |= one.lowPaulis;
        s.highPaulis |= one.highPaulis;
    }
    return s;
}
void multiplyMultiQubitNot(Qureg q, int* t, int nt) {
    QAH_API({ ctrlPauliTensor(q, {}, {}, xString(q, t, nt), 1); });
}
void applyMultiStateControlledMultiQubitNot(Qureg q, int* c, int* s, int nc, int* t, int nt) {
    QAH_API({
        validateCtrlsAndTargs(q, c, nc, t, nt);
        pauliStrImpl(q, c, s, nc, xString(q, t, nt));
    });
}
void applyMultiQubitNot(Qureg q, int* t, int nt) { applyMultiStateControlledMultiQubitNot(q, nullptr, nullptr, 0, t, nt); }
void applyControlledMultiQubitNot(Qureg q, int c, int* t, int nt) {
    applyMultiStateControlledMultiQubitNot(q, &c, nullptr, 1, t, nt);
}
void applyMultiControlledMultiQubitNot(Qureg q, int* c, int nc, int* t, int nt) {
    applyMultiStateControlledMultiQubitNot(q, c, nullptr, nc, t, nt);
}

// ---- projectors and measurement (operations.cpp:1024-1127) --------------------------

void applyQubitProjector(Qureg q, int t, int o) {
    QAH_API({
        validateTarget(q, t);
        validateOutcome(o);
        projector(q, {t}, {o}, 1);
    });
}
void applyMultiQubitProjector(Qureg q, int* qs, int* os, int n) {
    QAH_API({
        validateTargets(q, qs, n);
        for (int i = 0; i < n; i++) validateOutcome(os[i]);
        projector(q, vec(qs, n), vec(os, n), 1);
    });
}

qreal calcProbOfQubitOutcome(Qureg qureg, int qubit, int outcome);
qreal calcProbOfMultiQubitOutcome(Qureg qureg, int* qubits, int* outcomes, int numQubits);

int applyQubitMeasurementAndGetProb(Qureg q, int t, qreal* probability) {
    QAH_API({
        validateTarget(q, t);
        // each probability from its own reduction (the reference's
        // arithmetic), not from calcProbOfQubitOutcome's per-state cache
        int zero = 0, one = 1;
        double p0 = calcProbOfMultiQubitOutcome(q, &t, &zero, 1);
        double p1 = calcProbOfMultiQubitOutcome(q, &t, &one, 1);
        if (validating() && std::abs(p0 + p1 - 1) > epsilonFor(q))
            fail("The state was unnormalised such that the probabilities of possible outcomes did not sum to one, and "
                 "ergo cannot be meaningfully sampled.");
        int outcome = randomQubitOutcome(p0);
        double p = outcome ? p1 : p0;
        *probability = p;
        projector(q, {t}, {outcome}, p);
        return outcome;
    });
}
int applyQubitMeasurement(Qureg q, int t) {
    qreal p = 0;
    return applyQubitMeasurementAndGetProb(q, t, &p);
}
qreal applyForcedQubitMeasurement(Qureg q, int t, int o) {
    QAH_API({
        validateTarget(q, t);
        validateOutcome(o);
        double p = calcProbOfMultiQubitOutcome(q, &t, &o, 1);
        if (validating() && p <= epsilonFor(q)) fail("The forced measurement outcome has (near) zero probability.");
        projector(q, {t}, {o}, p);
        return p;
    });
}

void calcProbsOfAllMultiQubitOutcomes(qreal* outcomeProbs, Qureg qureg, int* qubits, int numQubits);

qindex applyMultiQubitMeasurementAndGetProb(Qureg q, int* qs, int n, qreal* probability) {
    QAH_API({
        validateTargets(q, qs, n);
        vector<double> probs((size_t)(1LL << n));
        calcProbsOfAllMultiQubitOutcomes(probs.data(), q, qs, n);
        double tot = 0;
        for (double p : probs) tot += p;
        if (validating() && std::abs(tot - 1) > epsilonFor(q))
            fail("The state was unnormalised such that the probabilities of possible outcomes did not sum to one, and "
                 "ergo cannot be meaningfully sampled.");
        qindex outcome = randomMultiQubitOutcome(probs);
        *probability = probs[(size_t)outcome];
        vector<int> os((size_t)n);
        for (int i = 0; i < n; i++) os[(size_t)i] = (int)((outcome >> i) & 1);
        projector(q, vec(qs, n), os, *probability);
        return outcome;
    });
}
qindex applyMultiQubitMeasurement(Qureg q, int* qs, int n) {
    qreal p = 0;
    return applyMultiQubitMeasurementAndGetProb(q, qs, n, &p);
}
qreal applyForcedMultiQubitMeasurement(Qureg q, int* qs, int* os, int n) {
    QAH_API({
        validateTargets(q, qs, n);
        for (int i = 0; i < n; i++) validateOutcome(os[i]);
        double p = calcProbOfMultiQubitOutcome(q, qs, os, n);
        if (validating() && p <= epsilonFor(q)) fail("The forced measurement outcomes have (near) zero probability.");
        projector(q, vec(qs, n), vec(os, n), p);
        return p;
    });
}

// ---- QFT (operations.cpp:1934-1964): n H, n(n-1)/2 phase shifts, n/2 swaps -----------------

void applyQuantumFourierTransform(Qureg q, int* t, int nt) {
    QAH_API({
        validateTargets(q, t, nt);
        for (int n = nt - 1; n >= 0; n--) {
            applyHadamard(q, t[n]);
            for (int m = 0; m < n; m++) {
                double arg = PI / (double)(1LL << (m + 1));
                applyTwoQubitPhaseShift(q, t[n], t[n - m - 1], arg);
            }
        }
        int mid = nt / 2;
        for (int n = 0; n < mid; n++) applySwap(q, t[n], t[nt - 1 - n]);
    });
}
void applyFullQuantumFourierTransform(Qureg q) {
    QAH_API({
        validateQureg(q);
        vector<int> t((size_t)q.numQubits);
        for (int i = 0; i < q.numQubits; i++) t[(size_t)i] = i;
        applyQuantumFourierTransform(q, t.data(), q.numQubits);
    });
}

}  // extern "C"

// ---- C++ overloads --------------------------------------------------------------------------
void applyMultiControlledCompMatr1(Qureg q, std::vector<int> c, int t, CompMatr1 m) {
    applyMultiControlledCompMatr1(q, c.data(), (int)c.size(), t, m);
}
void applyMultiControlledCompMatr2(Qureg q, std::vector<int> c, int t1, int t2, CompMatr2 m) {
    applyMultiControlledCompMatr2(q, c.data(), (int)c.size(), t1, t2, m);
}
void applyCompMatr(Qureg q, std::vector<int> t, CompMatr m) { applyCompMatr(q, t.data(), (int)t.size(), m); }
void applyMultiQubitPhaseShift(Qureg q, std::vector<int> t, qreal a) {
    applyMultiQubitPhaseShift(q, t.data(), (int)t.size(), a);
}
void applyPhaseGadget(Qureg q, std::vector<int> t, qreal a) { applyPhaseGadget(q, t.data(), (int)t.size(), a); }
void applyQuantumFourierTransform(Qureg q, std::vector<int> t) { applyQuantumFourierTransform(q, t.data(), (int)t.size()); }
