// Test infrastructure only (built by oracle/Makefile into oracle/_ref/).
//
// Replays a circuit text file through QuEST's own public C/C++ API
// (quest/include/*.h of the reference, linked against the reference's FP64
// OpenMP CPU build) and prints every requested result with 17 significant
// digits, so that fixtures generated from it pin the repository's oracle and
// HIP path to the reference's behaviour.
//
// Circuit format: one operation per line, whitespace separated, '#' comments.
// The op names and argument orders are defined in quest_amd/circuits.py
// (CircuitWriter), which is the single writer of these files.

#include "quest/include/modes.h"
#include "quest/include/precision.h"
#include "quest/include/types.h"
#include "quest/include/environment.h"
#include "quest/include/qureg.h"
#include "quest/include/matrices.h"
#include "quest/include/paulis.h"
#include "quest/include/channels.h"
#include "quest/include/operations.h"
#include "quest/include/decoherence.h"
#include "quest/include/calculations.h"
#include "quest/include/initialisations.h"
#include "quest/include/debug.h"

#include <chrono>
#include <cstdio>
#include <cstdlib>
#include <fstream>
#include <sstream>
#include <string>
#include <vector>

using std::string;
using std::vector;

namespace {

struct Tok {
    std::istringstream in;
    explicit Tok(const string& s) : in(s) {}
    int i() { long long v; in >> v; return (int) v; }
    long long ll() { long long v; in >> v; return v; }
    double d() { double v; in >> v; return v; }
    string s() { string v; in >> v; return v; }
    vector<int> ints(int n) { vector<int> v(n); for (auto& x : v) x = i(); return v; }
    vector<qcomp> comps(long long n) {
        vector<qcomp> v(n);
        for (auto& x : v) { double re = d(); double im = d(); x = qcomp(re, im); }
        return v;
    }
};

CompMatr1 readM1(Tok& t) {
    auto e = t.comps(4);
    return getCompMatr1(vector<vector<qcomp>>{{e[0], e[1]}, {e[2], e[3]}});
}

CompMatr2 readM2(Tok& t) {
    auto e = t.comps(16);
    vector<vector<qcomp>> m(4, vector<qcomp>(4));
    for (int r = 0; r < 4; r++) for (int c = 0; c < 4; c++) m[r][c] = e[4*r + c];
    return getCompMatr2(m);
}

void printReal(const string& label, double v) { std::printf("R %s %.17g\n", label.c_str(), v); }

// QD_DEPLOY=gpu: every Qureg on the GPU (createCustomQureg with
// useGpuAccel = 1), whatever its size, so that a build whose accelerator
// seam is another GPU backend runs every operation there; otherwise QuEST's
// own automatic deployment
bool forceGpu() {
    const char* d = std::getenv("QD_DEPLOY");
    return d && std::string(d) == "gpu";
}

Qureg makeQureg(int n, int dm) {
    if (!forceGpu()) return dm ? createDensityQureg(n) : createQureg(n);
    Qureg q = createCustomQureg(n, dm, 0, 1, 0);
    if (!q.isGpuAccelerated) {
        std::fprintf(stderr, "QD_DEPLOY=gpu but the Qureg is not GPU-accelerated\n");
        std::exit(5);
    }
    return q;
}

}  // namespace

int main(int argc, char** argv) {
    if (argc < 2) { std::fprintf(stderr, "usage: %s circuit.txt\n", argv[0]); return 2; }
    std::ifstream file(argv[1]);
    if (!file) { std::fprintf(stderr, "cannot open %s\n", argv[1]); return 2; }

    initQuESTEnv();
    Qureg q{};
    Qureg aux{};
    bool have = false, haveAux = false;
    std::chrono::steady_clock::time_point t0 = std::chrono::steady_clock::now();
    long lineNo = 0;
    string line;
    while (std::getline(file, line)) {
        lineNo++;
        auto hash = line.find('#');
        if (hash != string::npos) line = line.substr(0, hash);
        Tok t(line);
        string op = t.s();
        if (op.empty()) continue;

        if (op == "QUREG") {
            int n = t.i(); int dm = t.i();
            q = makeQureg(n, dm);
            have = true;
            if (forceGpu()) std::fprintf(stderr, "qureg of %d qubits on the GPU\n", n);
        } else if (op == "INITAMPS") {
            string path = t.s();
            std::ifstream bin(path, std::ios::binary);
            vector<qcomp> amps(q.numAmps);
            bin.read(reinterpret_cast<char*>(amps.data()), (std::streamsize)(amps.size() * sizeof(qcomp)));
            if (!bin) { std::fprintf(stderr, "short read %s\n", path.c_str()); return 3; }
            if (q.isDensityMatrix) setDensityQuregFlatAmps(q, 0, amps.data(), q.numAmps);
            else setQuregAmps(q, 0, amps.data(), q.numAmps);
        } else if (op == "INITZERO") { initZeroState(q);
        } else if (op == "INITPLUS") { initPlusState(q);
        } else if (op == "INITDEBUG") { initDebugState(q);
        } else if (op == "INITCLASSICAL") { initClassicalState(q, t.ll());
        } else if (op == "U1") { int tg = t.i(); applyCompMatr1(q, tg, readM1(t));
        } else if (op == "CU1") {
            int nc = t.i(); auto c = t.ints(nc); int tg = t.i();
            applyMultiControlledCompMatr1(q, c.data(), nc, tg, readM1(t));
        } else if (op == "SCU1") {
            int nc = t.i(); auto c = t.ints(nc); auto s = t.ints(nc); int tg = t.i();
            applyMultiStateControlledCompMatr1(q, c.data(), s.data(), nc, tg, readM1(t));
        } else if (op == "U2") { int a = t.i(); int b = t.i(); applyCompMatr2(q, a, b, readM2(t));
        } else if (op == "CU2") {
            int nc = t.i(); auto c = t.ints(nc); int a = t.i(); int b = t.i();
            applyMultiControlledCompMatr2(q, c.data(), nc, a, b, readM2(t));
        } else if (op == "UN" || op == "CUN") {
            int nc = 0; vector<int> c;
            if (op == "CUN") { nc = t.i(); c = t.ints(nc); }
            int nt = t.i(); auto tg = t.ints(nt);
            long long dim = 1LL << nt;
            auto e = t.comps(dim * dim);
            CompMatr m = createCompMatr(nt);
            vector<vector<qcomp>> rows(dim, vector<qcomp>(dim));
            for (long long r = 0; r < dim; r++) for (long long cc = 0; cc < dim; cc++) rows[r][cc] = e[r*dim + cc];
            setCompMatr(m, rows);
            if (op == "UN") applyCompMatr(q, tg.data(), nt, m);
            else applyMultiControlledCompMatr(q, c.data(), nc, tg.data(), nt, m);
            destroyCompMatr(m);
        } else if (op == "D1") { int tg = t.i(); auto e = t.comps(2); applyDiagMatr1(q, tg, getDiagMatr1(e));
        } else if (op == "CD1") {
            int nc = t.i(); auto c = t.ints(nc); int tg = t.i(); auto e = t.comps(2);
            applyMultiControlledDiagMatr1(q, c.data(), nc, tg, getDiagMatr1(e));
        } else if (op == "D2") { int a = t.i(); int b = t.i(); auto e = t.comps(4); applyDiagMatr2(q, a, b, getDiagMatr2(e));
        } else if (op == "DN") {
            int nt = t.i(); auto tg = t.ints(nt); auto e = t.comps(1LL << nt);
            DiagMatr m = createDiagMatr(nt); setDiagMatr(m, e);
            applyDiagMatr(q, tg.data(), nt, m); destroyDiagMatr(m);
        } else if (op == "H") { applyHadamard(q, t.i());
        } else if (op == "X") { applyPauliX(q, t.i());
        } else if (op == "Y") { applyPauliY(q, t.i());
        } else if (op == "Z") { applyPauliZ(q, t.i());
        } else if (op == "S") { applyS(q, t.i());
        } else if (op == "T") { applyT(q, t.i());
        } else if (op == "CX") { int c = t.i(); applyControlledPauliX(q, c, t.i());
        } else if (op == "CZ") { int c = t.i(); applyControlledPauliZ(q, c, t.i());
        } else if (op == "PS") { int tg = t.i(); applyPhaseShift(q, tg, t.d());
        } else if (op == "PS2") { int a = t.i(); int b = t.i(); applyTwoQubitPhaseShift(q, a, b, t.d());
        } else if (op == "PSN") { int nt = t.i(); auto tg = t.ints(nt); applyMultiQubitPhaseShift(q, tg.data(), nt, t.d());
        } else if (op == "RX") { int tg = t.i(); applyRotateX(q, tg, t.d());
        } else if (op == "RY") { int tg = t.i(); applyRotateY(q, tg, t.d());
        } else if (op == "RZ") { int tg = t.i(); applyRotateZ(q, tg, t.d());
        } else if (op == "SWAP") { int a = t.i(); applySwap(q, a, t.i());
        } else if (op == "CSWAP") {
            int nc = t.i(); auto c = t.ints(nc); int a = t.i(); int b = t.i();
            applyMultiControlledSwap(q, c.data(), nc, a, b);
        } else if (op == "PSTR" || op == "CPSTR" || op == "PG" || op == "CPG") {
            int nc = 0; vector<int> c;
            if (op[0] == 'C') { nc = t.i(); c = t.ints(nc); }
            string chars = t.s(); int nt = t.i(); auto tg = t.ints(nt);
            PauliStr str = getPauliStr(chars.c_str(), tg.data(), nt);
            if (op == "PSTR") applyPauliStr(q, str);
            else if (op == "CPSTR") applyMultiControlledPauliStr(q, c.data(), nc, str);
            else if (op == "PG") applyPauliGadget(q, str, t.d());
            else applyMultiControlledPauliGadget(q, c.data(), nc, str, t.d());
        } else if (op == "PHG") { int nt = t.i(); auto tg = t.ints(nt); applyPhaseGadget(q, tg.data(), nt, t.d());
        } else if (op == "QFT") { int nt = t.i(); auto tg = t.ints(nt); applyQuantumFourierTransform(q, tg.data(), nt);
        } else if (op == "FQFT") { applyFullQuantumFourierTransform(q);
        } else if (op == "DEPH") { int a = t.i(); mixDephasing(q, a, t.d());
        } else if (op == "DEPH2") { int a = t.i(); int b = t.i(); mixTwoQubitDephasing(q, a, b, t.d());
        } else if (op == "DEPOL") { int a = t.i(); mixDepolarising(q, a, t.d());
        } else if (op == "DEPOL2") { int a = t.i(); int b = t.i(); mixTwoQubitDepolarising(q, a, b, t.d());
        } else if (op == "DAMP") { int a = t.i(); mixDamping(q, a, t.d());
        } else if (op == "PAULIS") { int a = t.i(); double x = t.d(); double y = t.d(); mixPaulis(q, a, x, y, t.d());
        } else if (op == "KRAUS") {
            int nq = t.i(); auto tg = t.ints(nq); int nops = t.i();
            long long dim = 1LL << nq;
            KrausMap map = createKrausMap(nq, nops);
            vector<vector<vector<qcomp>>> ms(nops, vector<vector<qcomp>>(dim, vector<qcomp>(dim)));
            for (int k = 0; k < nops; k++) {
                auto e = t.comps(dim * dim);
                for (long long r = 0; r < dim; r++) for (long long cc = 0; cc < dim; cc++) ms[k][r][cc] = e[r*dim + cc];
            }
            setKrausMap(map, ms);
            mixKrausMap(q, tg.data(), nq, map);
            destroyKrausMap(map);
        } else if (op == "PROJ") { int a = t.i(); applyQubitProjector(q, a, t.i());
        } else if (op == "SEEDS") {
            int ns = t.i(); vector<unsigned> s(ns); for (auto& x : s) x = (unsigned) t.ll();
            setSeeds(s.data(), ns);
        } else if (op == "MEAS") {
            int a = t.i(); int o = applyQubitMeasurement(q, a);
            std::printf("R meas %d\n", o);
        } else if (op == "MEASN") {
            int nq = t.i(); auto qs = t.ints(nq);
            std::printf("R measn %lld\n", (long long) applyMultiQubitMeasurement(q, qs.data(), nq));
        } else if (op == "RANDPURE" || op == "RANDMIXED") {
            // a scratch qureg: only the RNG stream it advances is observed
            int n = t.i(); long long k = t.ll();
            Qureg s = makeQureg(n, (op == "RANDMIXED" || k) ? 1 : 0);
            if (op == "RANDPURE") initRandomPureState(s);
            else initRandomMixedState(s, k);
            destroyQureg(s);
        } else if (op == "AUXFILE" || op == "AUXCOPY") {
            // the auxiliary qureg of the two-state calculations: amplitudes
            // from a file, or a clone of the main qureg
            if (haveAux) destroyQureg(aux);
            if (op == "AUXFILE") {
                int n = t.i(); int dm = t.i(); string path = t.s();
                aux = makeQureg(n, dm);
                std::ifstream bin(path, std::ios::binary);
                vector<qcomp> amps(aux.numAmps);
                bin.read(reinterpret_cast<char*>(amps.data()), (std::streamsize)(amps.size() * sizeof(qcomp)));
                if (!bin) { std::fprintf(stderr, "short read %s\n", path.c_str()); return 3; }
                if (dm) setDensityQuregFlatAmps(aux, 0, amps.data(), aux.numAmps);
                else setQuregAmps(aux, 0, amps.data(), aux.numAmps);
            } else {
                aux = createCloneQureg(q);
            }
            haveAux = true;
        } else if (op == "OUT_INNER") {
            qcomp z = calcInnerProduct(q, aux);
            std::printf("R inner %.17g %.17g\n", std::real(z), std::imag(z));
        } else if (op == "OUT_FIDELITY") { printReal("fidelity", calcFidelity(q, aux));
        } else if (op == "OUT_DISTANCE") { printReal("distance", calcDistance(q, aux));
        } else if (op == "OUT_PROB") {
            int a = t.i(); int o = t.i(); printReal("prob", calcProbOfQubitOutcome(q, a, o));
        } else if (op == "OUT_EXPEC") {
            string chars = t.s(); int nt = t.i(); auto tg = t.ints(nt);
            printReal("expec", calcExpecPauliStr(q, getPauliStr(chars.c_str(), tg.data(), nt)));
        } else if (op == "OUT_PURITY") { printReal("purity", calcPurity(q));
        } else if (op == "OUT_TOTALPROB") { printReal("totalprob", calcTotalProb(q));
        } else if (op == "OUT_PROBS") {
            int nq = t.i(); auto qs = t.ints(nq);
            vector<qreal> p(1LL << nq);
            calcProbsOfAllMultiQubitOutcomes(p.data(), q, qs.data(), nq);
            std::printf("R probs");
            for (auto v : p) std::printf(" %.17g", v);
            std::printf("\n");
        } else if (op == "OUT_AMPS") {
            string path = t.s();
            vector<qcomp> amps(q.numAmps);
            if (q.isDensityMatrix) {
                // flat column-major order, identical to the in-memory layout
                long long dim = 1LL << q.numQubits;
                for (long long c = 0; c < dim; c++)
                    for (long long r = 0; r < dim; r++) amps[r + c*dim] = getDensityQuregAmp(q, r, c);
            } else {
                getQuregAmps(amps.data(), q, 0, q.numAmps);
            }
            std::ofstream out(path, std::ios::binary);
            out.write(reinterpret_cast<const char*>(amps.data()), (std::streamsize)(amps.size() * sizeof(qcomp)));
        } else if (op == "TIMER_START") {
            syncQuESTEnv(); t0 = std::chrono::steady_clock::now();
        } else if (op == "TIMER_STOP") {
            syncQuESTEnv();
            double s = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
            std::printf("T %s %.9g\n", t.s().c_str(), s);
        } else {
            std::fprintf(stderr, "line %ld: unknown op '%s'\n", lineNo, op.c_str());
            return 4;
        }
        std::fflush(stdout);
    }
    if (haveAux) destroyQureg(aux);
    if (have) destroyQureg(q);
    finalizeQuESTEnv();
    return 0;
}
