"""calcExpecPauliStrSum on statevectors: terms that share their X/Y mask are
reduced together (up to 8 per read of the state).  Checked against the CPU
oracle (the sum of coefficient x expectation over the terms, 1e-10 for
complex128 amplitudes, 1e-5 for complex64) and, term by term, bitwise
against the single-string reduction qa_statevec_calcExpecPauliStr."""

import ctypes as C
import sys

import numpy as np
import pytest

from quest_amd import circuits as qc

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def grouped_terms(n: int, seed: int):
    """Strings in groups sharing X/Y masks (each group's strings differ in
    their X<->Y and Z choices), Z-only strings and the identity."""
    rng = qc.SplitMix64(seed)
    terms = []
    for g in range(7):
        k = 1 + rng.randint(min(5, n))
        xy = rng.sample(n, k)
        rest = [q for q in range(n) if q not in xy]
        for _ in range(3 + rng.randint(9)):
            chars, targs = [], []
            for q in xy:
                chars.append("XY"[rng.randint(2)])
                targs.append(q)
            for q in rest:
                if rng.randint(3) == 0:
                    chars.append("Z")
                    targs.append(q)
            terms.append(("".join(chars), targs, 2 * rng.uniform() - 1))
    for _ in range(5):
        targs = rng.sample(n, 1 + rng.randint(n))
        terms.append(("Z" * len(targs), targs, 2 * rng.uniform() - 1))
    terms.append(("I", [0], 0.25))
    return terms


def make_sum(api, terms, complex_coeffs=None):
    strs = (api.PauliStr * len(terms))(*[api.pauli_str(c, t) for c, t, _ in terms])
    cs = complex_coeffs if complex_coeffs is not None else [complex(w) for _, _, w in terms]
    coeffs = (api.qcomp * len(terms))(*[api.qcomp.of(z) for z in cs])
    return api.createPauliStrSum(strs, coeffs, len(terms))


@pytest.mark.parametrize("n,prec,tol", [(13, 2, 1e-10), (17, 2, 1e-10), (12, 1, 1e-5)])
def test_grouped_sum_matches_oracle(api, n, prec, tol):
    from oracle import qoracle as O
    say(f"[gpu] pauli sum n={n} prec={prec}")
    terms = grouped_terms(n, 50 + n)
    amps = qc.random_state(n, 7 + n)
    st = O.State(n)
    st.a[:] = amps
    want = sum(w * st.expec(c, t) for c, t, w in terms)
    q = api.createCustomPrecisionQureg(n, 0, prec)
    s = make_sum(api, terms)
    try:
        api.set_amps(q, amps)
        got = api.calcExpecPauliStrSum(q, s)
    finally:
        api.destroyPauliStrSum(s)
        api.destroyQureg(q)
    assert abs(got - want) <= tol, (got, want)


@pytest.mark.parametrize("prec", [2, 1])
def test_grouped_terms_equal_single_string_reductions(api, prec):
    """Each term's value from the grouped reduction is bitwise the value of
    qa_statevec_calcExpecPauliStr on the same string, and the weighted sum
    with complex coefficients is taken in term order."""
    say(f"[gpu] pauli sum per-term prec={prec}")
    n = 15
    terms = grouped_terms(n, 91)
    lib = api.gpu()
    q = api.createCustomPrecisionQureg(n, 0, prec)
    try:
        api.set_amps(q, qc.random_state(n, 3))
        from oracle import qoracle as O
        masks = [O.masks(c, t) for c, t, _ in terms]
        nt = len(terms)
        mx = (C.c_uint64 * nt)(*[m[0] for m in masks])
        my = (C.c_uint64 * nt)(*[m[1] for m in masks])
        mz = (C.c_uint64 * nt)(*[m[2] for m in masks])
        rng = np.random.default_rng(5)
        cz = rng.normal(size=nt) + 1j * rng.normal(size=nt)
        coeffs = (C.c_double * (2 * nt))(*[v for z in cz for v in (z.real, z.imag)])
        vals = (C.c_double * (2 * nt))()
        out = (C.c_double * 2)()
        rc = lib.qa_statevec_calcExpecPauliStrSum(C.c_void_p(q.gpuAmps), prec, n, C.c_int64(nt), mx, my, mz,
                                                  coeffs, vals, out)
        assert rc == 0
        single = (C.c_double * 2)()
        total = 0j
        for t in range(nt):
            assert lib.qa_statevec_calcExpecPauliStr(C.c_void_p(q.gpuAmps), prec, n, C.c_uint64(mx[t]),
                                                     C.c_uint64(my[t]), C.c_uint64(mz[t]), single) == 0
            assert (vals[2 * t], vals[2 * t + 1]) == (single[0], single[1]), t
            total += cz[t] * complex(vals[2 * t], vals[2 * t + 1])
        assert abs(complex(out[0], out[1]) - total) <= 1e-12 * max(1.0, abs(total))
    finally:
        api.destroyQureg(q)


def sum_terms(n: int, seed: int, complex_coeffs: bool):
    """Strings of 1..n random Paulis with random coefficients."""
    rng = np.random.default_rng(seed)
    terms = []
    for _ in range(9):
        k = int(rng.integers(1, n + 1))
        targs = [int(t) for t in rng.choice(n, size=k, replace=False)]
        chars = "".join(rng.choice(list("XYZ"), size=k))
        c = rng.normal() + (1j * rng.normal() if complex_coeffs else 0)
        terms.append((chars, targs, complex(c)))
    return terms


def rel_l2(a, b):
    return float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300))


@pytest.mark.parametrize("dens,prec,tol", [(False, 2, 1e-12), (True, 2, 1e-12), (False, 1, 1e-5)])
@pytest.mark.parametrize("order,reps", [(1, 2), (2, 1), (4, 1), (6, 1)])
def test_trotterized_gadget_matches_oracle(api, dens, prec, tol, order, reps):
    """applyTrotterizedPauliStrSumGadget (operations.cpp:1169-1197) against
    the oracle's restatement, which is pinned to products of matrix
    exponentials on the CPU; the gadgets run through the fused tile passes."""
    from oracle import qoracle as O
    n = 5 if dens else 11
    say(f"[gpu] trotter dens={dens} prec={prec} order={order} reps={reps}")
    terms = sum_terms(n, 40 + order, False)
    init = qc.random_state(n, 9, dens)
    st = O.State(n, dens)
    st.a[:] = init
    st.trotter(terms, 0.41, order, reps)
    q = api.createCustomPrecisionQureg(n, int(dens), prec)
    s = make_sum(api, terms, [c for _, _, c in terms])
    try:
        api.set_amps(q, init)
        api.applyTrotterizedPauliStrSumGadget(q, s, 0.41, order, reps)
        got = api.get_amps(q)
    finally:
        api.destroyPauliStrSum(s)
        api.destroyQureg(q)
    assert rel_l2(got, st.a) <= tol


@pytest.mark.parametrize("dens,prec,tol", [(False, 2, 1e-12), (True, 2, 1e-12), (False, 1, 1e-5)])
def test_multiply_pauli_sum_matches_oracle(api, dens, prec, tol):
    """multiplyPauliStrSum (operations.cpp:1112-1131): qureg <- H qureg with
    complex coefficients, left multiplication on a density matrix."""
    from oracle import qoracle as O
    n = 5 if dens else 12
    say(f"[gpu] multiply pauli sum dens={dens} prec={prec}")
    terms = sum_terms(n, 77, True)
    init = qc.random_state(n, 5, dens)
    st = O.State(n, dens)
    st.a[:] = init
    st.multiply_pauli_sum(terms)
    q = api.createCustomPrecisionQureg(n, int(dens), prec)
    ws = api.createCustomPrecisionQureg(n, int(dens), prec)
    s = make_sum(api, terms, [c for _, _, c in terms])
    try:
        api.set_amps(q, init)
        api.multiplyPauliStrSum(q, s, ws)
        got = api.get_amps(q)
    finally:
        api.destroyPauliStrSum(s)
        api.destroyQureg(ws)
        api.destroyQureg(q)
    assert rel_l2(got, st.a) <= tol


def test_pauli_sum_operation_errors(api):
    """validation.cpp:3706-3711 (Trotter order 1 or even, reps > 0),
    3272-3280 (Hermitian sum), 3800-3808 (compatible workspace)."""
    say("[gpu] pauli sum validation")
    q = api.createQureg(4)
    w = api.createQureg(5)
    herm = make_sum(api, [("XZ", [0, 2], 0.5), ("Y", [1], -0.25)])
    nonh = make_sum(api, [("XZ", [0, 2], 0.5)], [0.5 + 0.5j])
    wide = make_sum(api, [("X", [6], 1.0)])
    try:
        with pytest.raises(api.QuESTError, match="order"):
            api.applyTrotterizedPauliStrSumGadget(q, herm, 0.3, 3, 1)
        with pytest.raises(api.QuESTError, match="order"):
            api.applyTrotterizedPauliStrSumGadget(q, herm, 0.3, 0, 1)
        with pytest.raises(api.QuESTError, match="repetitions"):
            api.applyTrotterizedPauliStrSumGadget(q, herm, 0.3, 2, 0)
        with pytest.raises(api.QuESTError, match="Hermitian"):
            api.applyTrotterizedPauliStrSumGadget(q, nonh, 0.3, 2, 1)
        with pytest.raises(api.QuESTError, match="beyond"):
            api.applyTrotterizedPauliStrSumGadget(q, wide, 0.3, 2, 1)
        with pytest.raises(api.QuESTError, match="workspace"):
            api.multiplyPauliStrSum(q, herm, w)
    finally:
        for s in (herm, nonh, wide):
            api.destroyPauliStrSum(s)
        api.destroyQureg(w)
        api.destroyQureg(q)


PAULI = {"I": np.eye(2), "X": np.array([[0, 1], [1, 0]]), "Y": np.array([[0, -1j], [1j, 0]]),
         "Z": np.diag([1.0, -1.0])}


def pauli_op(n, chars, targs):
    per = {t: c for c, t in zip(chars, targs)}
    out = np.array([[1.0 + 0j]])
    for q in range(n - 1, -1, -1):
        out = np.kron(out, PAULI[per.get(q, "I")])
    return out


@pytest.mark.parametrize("prec,tol", [(2, 1e-13), (1, 1e-5)])
def test_set_qureg_to_pauli_str_sum(api, prec, tol):
    """setQuregToPauliStrSum (initialisations.cpp:216-227): the density
    matrix becomes sum_j c_j sigma_j, against Kronecker products in numpy
    (column-major flattening, complex coefficients)."""
    n = 4
    say(f"[gpu] setQuregToPauliStrSum prec={prec}")
    terms = sum_terms(n, 31, True)
    want = sum(c * pauli_op(n, ch, t) for ch, t, c in terms).reshape(-1, order="F")
    q = api.createCustomPrecisionQureg(n, 1, prec)
    s = make_sum(api, terms, [c for _, _, c in terms])
    try:
        api.setQuregToPauliStrSum(q, s)
        got = api.get_amps(q)
    finally:
        api.destroyPauliStrSum(s)
        api.destroyQureg(q)
    assert np.max(np.abs(got - want)) <= tol * max(1.0, np.max(np.abs(want)))


def test_full_state_diag_from_z_sum(api):
    """createFullStateDiagMatrFromPauliStrSum / setFullStateDiagMatrFromPauliStrSum
    (matrices.cpp:635-665): the diagonal of a sum of I/Z strings, checked
    against numpy and through applying it (multiplyFullStateDiagMatr)."""
    n = 9
    say("[gpu] full-state diagonal from a Z sum")
    rng = np.random.default_rng(4)
    terms = []
    for _ in range(6):
        k = int(rng.integers(1, n + 1))
        targs = [int(t) for t in rng.choice(n, size=k, replace=False)]
        terms.append(("Z" * k, targs, complex(rng.normal(), rng.normal())))
    terms.append(("Z", [n - 1], 0.5 + 0j))
    want = np.diag(sum(c * pauli_op(n, ch, t) for ch, t, c in terms))
    s = make_sum(api, terms, [c for _, _, c in terms])
    m = api.createFullStateDiagMatrFromPauliStrSum(s)
    q = api.createQureg(n)
    xs = make_sum(api, [("XZ", [0, 1], 1.0)])
    try:
        assert m.numQubits == n
        elems = np.ctypeslib.as_array(C.cast(m.cpuElems, C.POINTER(C.c_double)), shape=(2 << n,))
        got = elems[0::2] + 1j * elems[1::2]
        assert np.max(np.abs(got - want)) <= 1e-13
        init = qc.random_state(n, 3)
        api.set_amps(q, init)
        api.multiplyFullStateDiagMatr(q, m)
        assert rel_l2(api.get_amps(q), init * want) <= 1e-13
        with pytest.raises(api.QuESTError, match="X or Y"):
            api.setFullStateDiagMatrFromPauliStrSum(m, xs)
        sv = api.createQureg(3)
        try:
            with pytest.raises(api.QuESTError, match="density matrix"):
                api.setQuregToPauliStrSum(sv, xs)
        finally:
            api.destroyQureg(sv)
    finally:
        api.destroyPauliStrSum(xs)
        api.destroyPauliStrSum(s)
        api.destroyFullStateDiagMatr(m)
        api.destroyQureg(q)
