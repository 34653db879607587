// Heap matrices, Pauli strings and channels: creation, setting, syncing to
// the GPU and destruction (quest/src/api/matrices.cpp, paulis.cpp:98-385,
// channels.cpp:176-313, utilities.cpp:771-801 for the superoperator).
#include <cstdio>
#include <cstdlib>
#include <algorithm>
#include <cstring>
#include <string>

#include "host_core.hpp"

using namespace qah;

namespace {

const int UNKNOWN = -1;  // validate_STRUCT_PROPERTY_UNKNOWN_FLAG

int* newFlag(int v) {
    int* p = (int*)std::malloc(sizeof(int));
    if (!p) fail("Host allocation failed.");
    *p = v;
    return p;
}

qcomp** newNested(qindex dim, qcomp*& flat) {
    flat = (qcomp*)std::calloc((size_t)(dim * dim), sizeof(qcomp));
    qcomp** rows = (qcomp**)std::malloc(sizeof(qcomp*) * (size_t)dim);
    if (!flat || !rows) fail("Host allocation failed.");
    for (qindex r = 0; r < dim; r++) rows[r] = flat + r * dim;
    return rows;
}

void* devAlloc(size_t bytes) {
    void* d = nullptr;
    int rc = qa_malloc(&d, bytes);
    if (rc == QA_ERR_OOM) fail("Insufficient GPU memory for the matrix.");
    check(rc);
    return d;
}

int pauliCode(char c) {
    switch (c) {
        case 'I': case 'i': case '0': return 0;
        case 'X': case 'x': case '1': return 1;
        case 'Y': case 'y': case '2': return 2;
        case 'Z': case 'z': case '3': return 3;
    }
    fail(std::string("Invalid Pauli character '") + c + "'; must be one of IXYZ (or 0123).");
}

// superop += conj(K) (x) K for every Kraus operator (utilities.cpp:771-801):
// S[i*d + k][j*d + l] += conj(K_ij) K_kl
void buildSuperOp(SuperOp op, qcomp*** ks, int numOps, qindex d) {
    qindex D = d * d;
    for (qindex r = 0; r < D; r++)
        for (qindex c = 0; c < D; c++) op.cpuElems[r][c] = 0;
    for (int n = 0; n < numOps; n++)
        for (qindex i = 0; i < d; i++)
            for (qindex j = 0; j < d; j++)
                for (qindex k = 0; k < d; k++)
                    for (qindex l = 0; l < d; l++) op.cpuElems[i * d + k][j * d + l] += std::conj(ks[n][i][j]) * ks[n][k][l];
}

}  // namespace

namespace {

// validation.cpp:2170-2188: at least one variable, each of at least one
// qubit, together the matrix's qubits
void validateVars(int matrQubits, const int* numQubitsPerVar, int numVars) {
    if (numVars <= 0 || !numQubitsPerVar) fail("The number of variables must be positive.");
    int total = 0;
    for (int v = 0; v < numVars; v++) {
        if (numQubitsPerVar[v] <= 0)
            fail("Variable " + std::to_string(v) + " must have a positive number of qubits.");
        total += numQubitsPerVar[v];
    }
    if (total != matrQubits)
        fail("The variables' qubits (" + std::to_string(total) + ") differ from the matrix's (" +
             std::to_string(matrQubits) + ").");
}

// fastmath.hpp:100-115: variable n takes the next numQubitsPerVar[n] index
// bits, lowest first; with areSigned, a variable whose top bit is set has
// 2^(bits-1) subtracted, as the reference computes it
void varValues(qindex index, const int* numQubitsPerVar, int numVars, bool areSigned, qindex* out) {
    qindex rest = index;
    for (int n = 0; n < numVars; n++) {
        out[n] = rest & ((1LL << numQubitsPerVar[n]) - 1);
        rest >>= numQubitsPerVar[n];
    }
    if (areSigned)
        for (int n = 0; n < numVars; n++)
            if ((out[n] >> (numQubitsPerVar[n] - 1)) & 1) out[n] -= 1LL << (numQubitsPerVar[n] - 1);
}

// utilities.cpp:307-315: lists[i0][i1]...[i_{n-1}]
qcomp nestedElem(void* lists, const qindex* inds, int n) {
    if (n == 1) return static_cast<qcomp*>(lists)[inds[0]];
    return nestedElem(static_cast<void**>(lists)[inds[0]], inds + 1, n - 1);
}

template <class F>
void fillDiag(qcomp* elems, qindex numElems, const int* numQubitsPerVar, int numVars, bool areSigned, F&& elemOf) {
    std::vector<qindex> vals((size_t)numVars);
    for (qindex i = 0; i < numElems; i++) {
        varValues(i, numQubitsPerVar, numVars, areSigned, vals.data());
        elems[i] = elemOf(vals.data());
    }
}

}  // namespace

extern "C" {

void _validateNewNestedElemsPtrNotNull(qcomp** ptrs, int numQubits, const char* caller) {
    QAH_API({
        if (!ptrs) fail(std::string(caller) + ": received a NULL matrix pointer.");
        for (int r = 0; r < (1 << numQubits); r++)
            if (!ptrs[r]) fail(std::string(caller) + ": a row pointer of the matrix was NULL.");
    });
}
void _validateNewElemsPtrNotNull(qcomp* ptr, const char* caller) {
    QAH_API({
        if (!ptr) fail(std::string(caller) + ": received a NULL pointer.");
    });
}

// ---- CompMatr -----------------------------------------------------------------

CompMatr createCompMatr(int numQubits) {
    QAH_API({
        requireEnv();
        if (numQubits < 1 || numQubits > 12) fail("Invalid number of qubits for a CompMatr (1..12 supported).");
        CompMatr m;
        m.numQubits = numQubits;
        m.numRows = 1LL << numQubits;
        m.isApproxUnitary = newFlag(UNKNOWN);
        m.isApproxHermitian = newFlag(UNKNOWN);
        m.wasGpuSynced = newFlag(0);
        m.cpuElems = newNested(m.numRows, m.cpuElemsFlat);
        m.gpuElemsFlat = (qcomp*)devAlloc((size_t)(m.numRows * m.numRows) * 16);
        return m;
    });
}

void destroyCompMatr(CompMatr m) {
    QAH_API({
        check(qa_free(m.gpuElemsFlat));
        std::free(m.cpuElemsFlat);
        std::free(m.cpuElems);
        std::free(m.isApproxUnitary);
        std::free(m.isApproxHermitian);
        std::free(m.wasGpuSynced);
    });
}

void syncCompMatr(CompMatr m) {
    QAH_API({
        check(qa_memcpyHtoD(m.gpuElemsFlat, m.cpuElemsFlat, (size_t)(m.numRows * m.numRows) * 16));
        *m.wasGpuSynced = 1;
        *m.isApproxUnitary = UNKNOWN;
        *m.isApproxHermitian = UNKNOWN;
    });
}

void setCompMatr(CompMatr m, qcomp** vals) {
    QAH_API({
        if (!vals) fail("setCompMatr: NULL elements.");
        for (qindex r = 0; r < m.numRows; r++)
            for (qindex c = 0; c < m.numRows; c++) m.cpuElems[r][c] = vals[r][c];
        syncCompMatr(m);
    });
}

// ---- DiagMatr -------------------------------------------------------------------

DiagMatr createDiagMatr(int numQubits) {
    QAH_API({
        requireEnv();
        if (numQubits < 1 || numQubits > 30) fail("Invalid number of qubits for a DiagMatr.");
        DiagMatr m;
        m.numQubits = numQubits;
        m.numElems = 1LL << numQubits;
        m.isApproxUnitary = newFlag(UNKNOWN);
        m.isApproxHermitian = newFlag(UNKNOWN);
        m.isApproxNonZero = newFlag(UNKNOWN);
        m.isStrictlyNonNegative = newFlag(UNKNOWN);
        m.wasGpuSynced = newFlag(0);
        m.cpuElems = (qcomp*)std::calloc((size_t)m.numElems, sizeof(qcomp));
        if (!m.cpuElems) fail("Host allocation failed.");
        m.gpuElems = (qcomp*)devAlloc((size_t)m.numElems * 16);
        return m;
    });
}

void destroyDiagMatr(DiagMatr m) {
    QAH_API({
        check(qa_free(m.gpuElems));
        std::free(m.cpuElems);
        std::free(m.isApproxUnitary);
        std::free(m.isApproxHermitian);
        std::free(m.isApproxNonZero);
        std::free(m.isStrictlyNonNegative);
        std::free(m.wasGpuSynced);
    });
}

void syncDiagMatr(DiagMatr m) {
    QAH_API({
        check(qa_memcpyHtoD(m.gpuElems, m.cpuElems, (size_t)m.numElems * 16));
        *m.wasGpuSynced = 1;
        *m.isApproxUnitary = UNKNOWN;
        *m.isApproxHermitian = UNKNOWN;
        *m.isApproxNonZero = UNKNOWN;
        *m.isStrictlyNonNegative = UNKNOWN;
    });
}

void setDiagMatr(DiagMatr m, qcomp* in) {
    QAH_API({
        if (!in) fail("setDiagMatr: NULL elements.");
        std::memcpy(m.cpuElems, in, (size_t)m.numElems * sizeof(qcomp));
        syncDiagMatr(m);
    });
}

// ---- FullStateDiagMatr ---------------------------------------------------------

FullStateDiagMatr createCustomFullStateDiagMatr(int numQubits, int useDistrib, int useGpuAccel, int useMultithread) {
    QAH_API({
        requireEnv();
        (void)useMultithread;
        if (useDistrib == 1) fail("This build has no distribution.");
        if (useGpuAccel == 0) fail("This build keeps every FullStateDiagMatr on the GPU.");
        if (numQubits < 1 || numQubits > 36) fail("Invalid number of qubits for a FullStateDiagMatr.");
        FullStateDiagMatr m;
        m.numQubits = numQubits;
        m.numElems = 1LL << numQubits;
        m.isGpuAccelerated = 1;
        m.isMultithreaded = 0;
        m.isDistributed = 0;
        m.numElemsPerNode = m.numElems;
        m.isApproxUnitary = newFlag(UNKNOWN);
        m.isApproxHermitian = newFlag(UNKNOWN);
        m.isApproxNonZero = newFlag(UNKNOWN);
        m.isStrictlyNonNegative = newFlag(UNKNOWN);
        m.wasGpuSynced = newFlag(0);
        m.cpuElems = (qcomp*)std::calloc((size_t)m.numElems, sizeof(qcomp));
        if (!m.cpuElems) fail("Host allocation failed.");
        m.gpuElems = (qcomp*)devAlloc((size_t)m.numElems * 16);
        return m;
    });
}

FullStateDiagMatr createFullStateDiagMatr(int numQubits) { return createCustomFullStateDiagMatr(numQubits, 0, 1, 0); }

void destroyFullStateDiagMatr(FullStateDiagMatr m) {
    QAH_API({
        check(qa_free(m.gpuElems));
        std::free(m.cpuElems);
        std::free(m.isApproxUnitary);
        std::free(m.isApproxHermitian);
        std::free(m.isApproxNonZero);
        std::free(m.isStrictlyNonNegative);
        std::free(m.wasGpuSynced);
    });
}

void syncFullStateDiagMatr(FullStateDiagMatr m) {
    QAH_API({
        check(qa_memcpyHtoD(m.gpuElems, m.cpuElems, (size_t)m.numElems * 16));
        *m.wasGpuSynced = 1;
        *m.isApproxUnitary = UNKNOWN;
        *m.isApproxHermitian = UNKNOWN;
    });
}

void setFullStateDiagMatr(FullStateDiagMatr m, qindex start, qcomp* in, qindex num) {
    QAH_API({
        if (start < 0 || num < 0 || start + num > m.numElems) fail("Invalid FullStateDiagMatr range.");
        std::memcpy(m.cpuElems + start, in, (size_t)num * sizeof(qcomp));
        check(qa_memcpyHtoD(m.gpuElems + start, in, (size_t)num * 16));
        *m.wasGpuSynced = 1;
        *m.isApproxUnitary = UNKNOWN;
        *m.isApproxHermitian = UNKNOWN;
    });
}

// ---- diagonal matrices from functions of integer variables (matrices.cpp:668-730)

// the callback is called once per element, in index order, on the host
void setDiagMatrFromMultiVarFunc(DiagMatr out, qcomp (*f)(qindex*), int* numQubitsPerVar, int numVars,
                                 int areSigned) {
    QAH_API({
        validateVars(out.numQubits, numQubitsPerVar, numVars);
        if (areSigned != 0 && areSigned != 1) fail("areSigned must be 0 or 1.");
        if (!f) fail("NULL callback function.");
        fillDiag(out.cpuElems, out.numElems, numQubitsPerVar, numVars, areSigned, [&](qindex* v) { return f(v); });
        syncDiagMatr(out);
    });
}
void setFullStateDiagMatrFromMultiVarFunc(FullStateDiagMatr out, qcomp (*f)(qindex*), int* numQubitsPerVar,
                                          int numVars, int areSigned) {
    QAH_API({
        validateVars(out.numQubits, numQubitsPerVar, numVars);
        if (areSigned != 0 && areSigned != 1) fail("areSigned must be 0 or 1.");
        if (!f) fail("NULL callback function.");
        fillDiag(out.cpuElems, out.numElems, numQubitsPerVar, numVars, areSigned, [&](qindex* v) { return f(v); });
        syncFullStateDiagMatr(out);
    });
}
// lists nested numDims deep, indexed by the variables' (unsigned) values
void setDiagMatrFromMultiDimLists(DiagMatr out, void* lists, int* numQubitsPerDim, int numDims) {
    QAH_API({
        validateVars(out.numQubits, numQubitsPerDim, numDims);
        if (!lists) fail("NULL lists.");
        fillDiag(out.cpuElems, out.numElems, numQubitsPerDim, numDims, false,
                 [&](qindex* v) { return nestedElem(lists, v, numDims); });
        syncDiagMatr(out);
    });
}
void setFullStateDiagMatrFromMultiDimLists(FullStateDiagMatr out, void* lists, int* numQubitsPerDim, int numDims) {
    QAH_API({
        validateVars(out.numQubits, numQubitsPerDim, numDims);
        if (!lists) fail("NULL lists.");
        fillDiag(out.cpuElems, out.numElems, numQubitsPerDim, numDims, false,
                 [&](qindex* v) { return nestedElem(lists, v, numDims); });
        syncFullStateDiagMatr(out);
    });
}

// matrices.cpp:635-665: the diagonal of a sum of I/Z strings, computed on
// the GPU and mirrored to the host elements
void setFullStateDiagMatrFromPauliStrSum(FullStateDiagMatr m, PauliStrSum sum) {
    QAH_API({
        if (sum.numTerms < 1 || !sum.strings || !sum.coeffs) fail("Invalid PauliStrSum.");
        std::vector<uint64_t> mz((size_t)sum.numTerms);
        std::vector<double> c(2 * (size_t)sum.numTerms);
        for (qindex t = 0; t < sum.numTerms; t++) {
            uint64_t x, y, z;
            pauliMasks(sum.strings[t], x, y, z);
            if (x | y) fail("The PauliStrSum contains X or Y Paulis, so it cannot initialise a diagonal matrix.");
            if (pauliMaxIndex(sum.strings[t]) >= m.numQubits)
                fail("The PauliStrSum targets qubits beyond the " + std::to_string(m.numQubits) +
                     "-qubit FullStateDiagMatr.");
            mz[t] = z;
            c[2 * t] = sum.coeffs[t].real();
            c[2 * t + 1] = sum.coeffs[t].imag();
        }
        check(qa_fullstatediagmatr_setElemsToPauliStrSum(m.gpuElems, m.numQubits, sum.numTerms, mz.data(), c.data()));
        check(qa_memcpyDtoH(m.cpuElems, m.gpuElems, (size_t)m.numElems * 16));
        *m.wasGpuSynced = 1;
        *m.isApproxUnitary = UNKNOWN;
        *m.isApproxHermitian = UNKNOWN;
    });
}

FullStateDiagMatr createFullStateDiagMatrFromPauliStrSum(PauliStrSum sum) {
    QAH_API({
        if (sum.numTerms < 1 || !sum.strings || !sum.coeffs) fail("Invalid PauliStrSum.");
        int top = -1;
        for (qindex t = 0; t < sum.numTerms; t++) top = std::max(top, pauliMaxIndex(sum.strings[t]));
        FullStateDiagMatr m = createFullStateDiagMatr(top + 1);
        setFullStateDiagMatrFromPauliStrSum(m, sum);
        return m;
    });
}

// ---- Pauli strings (paulis.cpp:306-331: string char i acts on indices[i]) ----

PauliStr getPauliStr(const char* paulis, int* indices, int numPaulis) {
    QAH_API({
        if (numPaulis < 1 || numPaulis > 64) fail("Invalid number of Paulis (must be 1..64).");
        if (!paulis || !indices) fail("getPauliStr: NULL arguments.");
        PauliStr s{0, 0};
        for (int i = 0; i < numPaulis; i++) {
            if (!paulis[i]) fail("getPauliStr: the Pauli string is shorter than numPaulis.");
            int t = indices[i];
            if (t < 0 || t >= 64) fail("Invalid Pauli index (must be 0..63).");
            for (int j = 0; j < i; j++)
                if (indices[j] == t) fail("The Pauli indices contained a duplicate.");
            PAULI_MASK_TYPE p = (PAULI_MASK_TYPE)pauliCode(paulis[i]);
            if (t < 32) s.lowPaulis |= p << (2 * t);
            else s.highPaulis |= p << (2 * (t - 32));
        }
        return s;
    });
}

PauliStrSum createPauliStrSum(PauliStr* strings, qcomp* coeffs, qindex numTerms) {
    QAH_API({
        if (numTerms < 1) fail("A PauliStrSum needs at least one term.");
        PauliStrSum s;
        s.numTerms = numTerms;
        s.strings = (PauliStr*)std::malloc(sizeof(PauliStr) * (size_t)numTerms);
        s.coeffs = (qcomp*)std::malloc(sizeof(qcomp) * (size_t)numTerms);
        s.isApproxHermitian = newFlag(UNKNOWN);
        if (!s.strings || !s.coeffs) fail("Host allocation failed.");
        std::memcpy(s.strings, strings, sizeof(PauliStr) * (size_t)numTerms);
        std::memcpy(s.coeffs, coeffs, sizeof(qcomp) * (size_t)numTerms);
        return s;
    });
}

// "coeff PAULIS" per line; the rightmost Pauli acts on qubit 0
// (paulis.cpp:374-385)
PauliStrSum createInlinePauliStrSum(const char* str) {
    QAH_API({
        if (!str) fail("createInlinePauliStrSum: NULL string.");
        std::vector<PauliStr> strs;
        std::vector<qcomp> coeffs;
        const char* p = str;
        while (*p) {
            while (*p == ' ' || *p == '\t' || *p == '\n' || *p == '\r') p++;
            if (!*p) break;
            char* end = nullptr;
            double re = std::strtod(p, &end);
            if (end == p) fail("Could not parse a PauliStrSum coefficient.");
            p = end;
            double im = 0;
            if (*p == '+' || *p == '-') {
                const char* q = p;
                double v = std::strtod(q, &end);
                if (end != q && (*end == 'i' || *end == 'j')) {
                    im = v;
                    p = end + 1;
                }
            }
            while (*p == ' ' || *p == '\t') p++;
            std::string letters;
            while (*p && *p != '\n' && *p != '\r') {
                if (*p != ' ' && *p != '\t') letters.push_back(*p);
                p++;
            }
            if (letters.empty() || letters.size() > 64) fail("Invalid Pauli string in PauliStrSum.");
            PauliStr s{0, 0};
            int n = (int)letters.size();
            for (int i = 0; i < n; i++) {
                int q = n - 1 - i;
                PAULI_MASK_TYPE code = (PAULI_MASK_TYPE)pauliCode(letters[(size_t)i]);
                if (q < 32) s.lowPaulis |= code << (2 * q);
                else s.highPaulis |= code << (2 * (q - 32));
            }
            strs.push_back(s);
            coeffs.push_back(qcomp(re, im));
        }
        return createPauliStrSum(strs.data(), coeffs.data(), (qindex)strs.size());
    });
}

void destroyPauliStrSum(PauliStrSum sum) {
    std::free(sum.strings);
    std::free(sum.coeffs);
    std::free(sum.isApproxHermitian);
}

void reportPauliStr(PauliStr str) {
    int top = pauliMaxIndex(str);
    std::string s;
    for (int q = (top < 0 ? 0 : top); q >= 0; q--) s.push_back("IXYZ"[pauliAt(str, q)]);
    std::printf("%s\n", s.c_str());
}

void reportPauliStrSum(PauliStrSum sum) {
    for (qindex t = 0; t < sum.numTerms; t++) {
        std::printf("%+.6g%+.6gi ", sum.coeffs[t].real(), sum.coeffs[t].imag());
        reportPauliStr(sum.strings[t]);
    }
}

// ---- SuperOp / KrausMap ---------------------------------------------------------

SuperOp createSuperOp(int numQubits) {
    QAH_API({
        requireEnv();
        if (numQubits < 1 || numQubits > 6) fail("Invalid number of qubits for a SuperOp (1..6 supported).");
        SuperOp op;
        op.numQubits = numQubits;
        op.numRows = 1LL << (2 * numQubits);
        op.cpuElems = newNested(op.numRows, op.cpuElemsFlat);
        op.gpuElemsFlat = (qcomp*)devAlloc((size_t)(op.numRows * op.numRows) * 16);
        op.wasGpuSynced = newFlag(0);
        return op;
    });
}

void syncSuperOp(SuperOp op) {
    QAH_API({
        check(qa_memcpyHtoD(op.gpuElemsFlat, op.cpuElemsFlat, (size_t)(op.numRows * op.numRows) * 16));
        *op.wasGpuSynced = 1;
    });
}

void destroySuperOp(SuperOp op) {
    QAH_API({
        check(qa_free(op.gpuElemsFlat));
        std::free(op.cpuElemsFlat);
        std::free(op.cpuElems);
        std::free(op.wasGpuSynced);
    });
}

void setSuperOp(SuperOp op, qcomp** matrix) {
    QAH_API({
        for (qindex r = 0; r < op.numRows; r++)
            for (qindex c = 0; c < op.numRows; c++) op.cpuElems[r][c] = matrix[r][c];
        syncSuperOp(op);
    });
}

void reportSuperOp(SuperOp op) { std::printf("SuperOp (%d qubits, %lld x %lld)\n", op.numQubits, op.numRows, op.numRows); }

KrausMap createKrausMap(int numQubits, int numOperators) {
    QAH_API({
        requireEnv();
        if (numQubits < 1 || numQubits > 6) fail("Invalid number of qubits for a KrausMap (1..6 supported).");
        if (numOperators < 1) fail("A KrausMap needs at least one operator.");
        KrausMap m;
        m.numQubits = numQubits;
        m.numMatrices = numOperators;
        m.numRows = 1LL << numQubits;
        m.matrices = (qcomp***)std::malloc(sizeof(qcomp**) * (size_t)numOperators);
        if (!m.matrices) fail("Host allocation failed.");
        for (int n = 0; n < numOperators; n++) {
            qcomp* flat = nullptr;
            m.matrices[n] = newNested(m.numRows, flat);
        }
        m.superop = createSuperOp(numQubits);
        m.isApproxCPTP = newFlag(UNKNOWN);
        return m;
    });
}

void syncKrausMap(KrausMap map) {
    QAH_API({
        buildSuperOp(map.superop, map.matrices, map.numMatrices, map.numRows);
        syncSuperOp(map.superop);
        *map.isApproxCPTP = UNKNOWN;
    });
}

void setKrausMap(KrausMap map, qcomp*** matrices) {
    QAH_API({
        for (int n = 0; n < map.numMatrices; n++)
            for (qindex r = 0; r < map.numRows; r++)
                for (qindex c = 0; c < map.numRows; c++) map.matrices[n][r][c] = matrices[n][r][c];
        syncKrausMap(map);
    });
}

void destroyKrausMap(KrausMap map) {
    QAH_API({
        for (int n = 0; n < map.numMatrices; n++) {
            std::free(map.matrices[n][0]);
            std::free(map.matrices[n]);
        }
        std::free(map.matrices);
        destroySuperOp(map.superop);
        std::free(map.isApproxCPTP);
    });
}

void reportKrausMap(KrausMap map) {
    std::printf("KrausMap (%d qubits, %d operators)\n", map.numQubits, map.numMatrices);
}

}  // extern "C"

// ---- C++ overloads ----------------------------------------------------------------

CompMatr1 getCompMatr1(std::vector<std::vector<qcomp>> in) {
    CompMatr1 m;
    m.numQubits = 1;
    m.numRows = 2;
    for (int r = 0; r < 2; r++)
        for (int c = 0; c < 2; c++) m.elems[r][c] = in.at((size_t)r).at((size_t)c);
    return m;
}
CompMatr2 getCompMatr2(std::vector<std::vector<qcomp>> in) {
    CompMatr2 m;
    m.numQubits = 2;
    m.numRows = 4;
    for (int r = 0; r < 4; r++)
        for (int c = 0; c < 4; c++) m.elems[r][c] = in.at((size_t)r).at((size_t)c);
    return m;
}
DiagMatr1 getDiagMatr1(std::vector<qcomp> in) {
    DiagMatr1 m;
    m.numQubits = 1;
    m.numElems = 2;
    for (int i = 0; i < 2; i++) m.elems[i] = in.at((size_t)i);
    return m;
}
DiagMatr2 getDiagMatr2(std::vector<qcomp> in) {
    DiagMatr2 m;
    m.numQubits = 2;
    m.numElems = 4;
    for (int i = 0; i < 4; i++) m.elems[i] = in.at((size_t)i);
    return m;
}
// the C++ container forms check the containers' sizes against the matrix
// (matrices.cpp / channels.cpp validate_matrixNumNewElems and friends)
namespace {
void requireSquare(const std::vector<std::vector<qcomp>>& in, qindex dim, const char* what) {
    if ((qindex)in.size() != dim) fail(std::string("The ") + what + " has the wrong number of rows.");
    for (const auto& r : in)
        if ((qindex)r.size() != dim) fail(std::string("A row of the ") + what + " has the wrong number of elements.");
}
}  // namespace

void setCompMatr(CompMatr out, std::vector<std::vector<qcomp>> in) {
    QAH_API({
        requireSquare(in, out.numRows, "CompMatr");
        std::vector<qcomp*> rows;
        for (auto& r : in) rows.push_back(r.data());
        setCompMatr(out, rows.data());
    });
}
void setDiagMatr(DiagMatr out, std::vector<qcomp> in) {
    QAH_API({
        if ((qindex)in.size() != out.numElems) fail("The DiagMatr needs exactly " + std::to_string(out.numElems) +
                                                      " elements.");
        setDiagMatr(out, in.data());
    });
}
void setKrausMap(KrausMap map, std::vector<std::vector<std::vector<qcomp>>> ms) {
    QAH_API({
        if ((int)ms.size() != map.numMatrices) fail("The KrausMap needs exactly " + std::to_string(map.numMatrices) +
                                                     " matrices.");
        for (auto& m : ms) requireSquare(m, map.numRows, "Kraus operator");
        std::vector<std::vector<qcomp*>> rows(ms.size());
        std::vector<qcomp**> ptrs;
        for (size_t n = 0; n < ms.size(); n++) {
            for (auto& r : ms[n]) rows[n].push_back(r.data());
            ptrs.push_back(rows[n].data());
        }
        setKrausMap(map, ptrs.data());
    });
}
void setSuperOp(SuperOp op, std::vector<std::vector<qcomp>> matrix) {
    QAH_API({
        requireSquare(matrix, op.numRows, "SuperOp");
        std::vector<qcomp*> rows;
        for (auto& r : matrix) rows.push_back(r.data());
        setSuperOp(op, rows.data());
    });
}

// inline constructors and setters (matrices.h:1027-1301, channels.h:483-562)
CompMatr createInlineCompMatr(int numQb, std::vector<std::vector<qcomp>> elems) {
    QAH_API({
        CompMatr m = createCompMatr(numQb);
        setCompMatr(m, elems);
        return m;
    });
}
DiagMatr createInlineDiagMatr(int numQb, std::vector<qcomp> elems) {
    QAH_API({
        DiagMatr m = createDiagMatr(numQb);
        setDiagMatr(m, elems);
        return m;
    });
}
KrausMap createInlineKrausMap(int numQb, int numOps, std::vector<std::vector<std::vector<qcomp>>> matrices) {
    QAH_API({
        KrausMap m = createKrausMap(numQb, numOps);
        setKrausMap(m, matrices);
        return m;
    });
}
SuperOp createInlineSuperOp(int numQb, std::vector<std::vector<qcomp>> matrix) {
    QAH_API({
        SuperOp m = createSuperOp(numQb);
        setSuperOp(m, matrix);
        return m;
    });
}
void setInlineCompMatr(CompMatr m, int numQb, std::vector<std::vector<qcomp>> in) {
    QAH_API({
        if (numQb != m.numQubits) fail("The number of qubits does not match the CompMatr's.");
        setCompMatr(m, in);
    });
}
void setInlineDiagMatr(DiagMatr m, int numQb, std::vector<qcomp> in) {
    QAH_API({
        if (numQb != m.numQubits) fail("The number of qubits does not match the DiagMatr's.");
        setDiagMatr(m, in);
    });
}
void setInlineFullStateDiagMatr(FullStateDiagMatr m, qindex startInd, qindex numElems, std::vector<qcomp> in) {
    QAH_API({
        if ((qindex)in.size() != numElems) fail("The list length does not match numElems.");
        setFullStateDiagMatr(m, startInd, in.data(), numElems);
    });
}
void setInlineKrausMap(KrausMap m, int numQb, int numOps, std::vector<std::vector<std::vector<qcomp>>> in) {
    QAH_API({
        if (numQb != m.numQubits || numOps != m.numMatrices)
            fail("The number of qubits or operators does not match the KrausMap's.");
        setKrausMap(m, in);
    });
}
void setInlineSuperOp(SuperOp m, int numQb, std::vector<std::vector<qcomp>> in) {
    QAH_API({
        if (numQb != m.numQubits) fail("The number of qubits does not match the SuperOp's.");
        setSuperOp(m, in);
    });
}
PauliStr getPauliStr(std::string paulis, int* indices, int numPaulis) {
    return getPauliStr(paulis.c_str(), indices, numPaulis);
}
PauliStr getPauliStr(std::string paulis, std::vector<int> indices) {
    return getPauliStr(paulis.c_str(), indices.data(), (int)indices.size());
}
PauliStr getPauliStr(std::string paulis) {
    std::vector<int> inds(paulis.size());
    for (size_t i = 0; i < paulis.size(); i++) inds[i] = (int)(paulis.size() - 1 - i);
    return getPauliStr(paulis.c_str(), inds.data(), (int)inds.size());
}
