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
