// A C++ program written against include/quest.h as QuEST v4's C++ users
// write it: inline constructors from initializer lists (matrices.h:1027-1301,
// channels.h:483-562), applied on the GPU and checked against hand-computed
// values; a mismatched container size must reach the input-error handler.
// Prints "cpp api ok" on success.
#include <cmath>
#include <cstdio>
#include <stdexcept>
#include <string>

#include "quest.h"

static void handler(const char* func, const char* msg) { throw std::runtime_error(std::string(func) + ": " + msg); }

static int fails = 0;
static void expect(bool ok, const char* what) {
    if (!ok) {
        std::printf("FAILED: %s\n", what);
        fails++;
    }
}
static bool near(qcomp a, qcomp b) { return std::abs(a - b) < 1e-12; }

int main() {
    initQuESTEnv();
    setInputErrorHandler(handler);
    const double h = 1 / std::sqrt(2.0);

    // Hadamard as an inline CompMatr on qubit 1 of |000>
    Qureg q = createQureg(3);
    initZeroState(q);
    CompMatr H = createInlineCompMatr(1, {{h, h}, {h, -h}});
    int t1[] = {1};
    applyCompMatr(q, t1, 1, H);
    expect(near(getQuregAmp(q, 0), h) && near(getQuregAmp(q, 2), h), "inline CompMatr");

    // diag(1, i) via setInlineDiagMatr on qubit 1
    DiagMatr D = createDiagMatr(1);
    setInlineDiagMatr(D, 1, {1, qcomp(0, 1)});
    applyDiagMatr(q, t1, 1, D);
    expect(near(getQuregAmp(q, 2), qcomp(0, h)), "inline DiagMatr");

    // a FullStateDiagMatr set piecewise from lists
    FullStateDiagMatr F = createFullStateDiagMatr(3);
    setInlineFullStateDiagMatr(F, 0, 4, {1, 1, 2, 1});
    setInlineFullStateDiagMatr(F, 4, 4, {1, 1, 1, 1});
    multiplyFullStateDiagMatr(q, F);
    expect(near(getQuregAmp(q, 2), qcomp(0, 2 * h)), "inline FullStateDiagMatr");

    // full dephasing of qubit 0 as an inline KrausMap and as an inline SuperOp
    Qureg rho = createDensityQureg(1);
    initPlusState(rho);
    KrausMap K = createInlineKrausMap(1, 2, {{{h, 0}, {0, h}}, {{h, 0}, {0, -h}}});
    int t0[] = {0};
    mixKrausMap(rho, t0, 1, K);
    expect(std::abs(calcPurity(rho) - 0.5) < 1e-12, "inline KrausMap");
    initPlusState(rho);
    SuperOp S = createInlineSuperOp(1, {{1, 0, 0, 0}, {0, 0, 0, 0}, {0, 0, 0, 0}, {0, 0, 0, 1}});
    mixSuperOp(rho, t0, 1, S);
    expect(std::abs(calcPurity(rho) - 0.5) < 1e-12, "inline SuperOp");

    // a wrong-sized list is an input error, not an out-of-bounds read
    bool threw = false;
    try {
        setInlineDiagMatr(D, 1, {1, 2, 3});
    } catch (const std::runtime_error&) {
        threw = true;
    }
    expect(threw, "size mismatch reported");
    threw = false;
    try {
        setInlineCompMatr(H, 1, {{1, 0}});
    } catch (const std::runtime_error&) {
        threw = true;
    }
    expect(threw, "row-count mismatch reported");

    destroyCompMatr(H);
    destroyDiagMatr(D);
    destroyFullStateDiagMatr(F);
    destroyKrausMap(K);
    destroySuperOp(S);
    destroyQureg(q);
    destroyQureg(rho);
    finalizeQuESTEnv();
    if (fails) return 1;
    std::printf("cpp api ok\n");
    return 0;
}
