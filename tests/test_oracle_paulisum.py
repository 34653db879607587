"""Pins the oracle's Pauli-string-sum operations (qoracle.State.trotter and
multiply_pauli_sum, restating operations.cpp:1112-1197) against brute force
that shares none of their code: full 2^n x 2^n Pauli operators built from
numpy Kronecker products and matrix exponentials from scipy.

  - multiplyPauliStrSum: H psi, and H rho (left multiplication) for a
    density matrix, with H = sum_j c_j sigma_j (complex c_j allowed);
  - applyTrotterizedPauliStrSumGadget: the product, in the reference's sweep
    order, of expm(-i a c_j sigma_j) -- applyPauliGadget(P, t) is
    exp(-i t/2 P) (operations.cpp:1428-1477, utilities.cpp:911-914) and each
    term is a gadget of angle 2 a c_j (operations.cpp:1133-1142) -- for
    orders 1, 2, 4, 6 (orders above 2 as the reference's code applies them:
    five reversed first-order sweeps with angles p, p, 1-4p, p, p);
  - and the mathematics: a sum of commuting Z strings is exponentiated
    exactly by one first-order sweep, and order 2 with many repetitions
    converges to expm(-i angle H) for non-commuting terms.
"""

import numpy as np
import pytest
from scipy.linalg import expm

from quest_amd import circuits as qc
from oracle import qoracle as O

PAULI = {"I": np.eye(2), "X": np.array([[0, 1], [1, 0]]), "Y": np.array([[0, -1j], [1j, 0]]),
         "Z": np.diag([1.0, -1.0])}


def pauli_op(n, chars, targs):
    """kron over qubits n-1..0 (qubit q is index bit q)."""
    per = {t: c for c, t in zip(chars, targs)}
    out = np.array([[1.0 + 0j]])
    for q in range(n - 1, -1, -1):
        out = np.kron(out, PAULI[per.get(q, "I")])
    return out


def random_terms(n, num, seed, complex_coeffs=False, letters="XYZ"):
    rng = np.random.default_rng(seed)
    terms = []
    for _ in range(num):
        k = int(rng.integers(1, n + 1))
        targs = [int(t) for t in rng.choice(n, size=k, replace=False)]
        chars = "".join(rng.choice(list(letters), size=k))
        c = rng.normal() + (1j * rng.normal() if complex_coeffs else 0)
        terms.append((chars, targs, complex(c)))
    return terms


def hamiltonian(n, terms):
    return sum(c * pauli_op(n, ch, t) for ch, t, c in terms)


def as_matrix(a, n):
    return a.reshape(1 << n, 1 << n, order="F")


def sweeps(angle, order, reps):
    """(angle, reversed) of every first-order sweep, in application order."""
    out = []
    step = angle / reps
    for _ in range(reps):
        if order == 1:
            out.append((step, False))
        elif order == 2:
            out += [(step / 2, False), (step / 2, True)]
        else:
            p = 1.0 / (4 - 4 ** (1.0 / (order - 1)))
            out += [(p * step, True), (p * step, True), ((1 - 4 * p) * step, True), (p * step, True),
                    (p * step, True)]
    return out


def brute_trotter(n, terms, angle, order, reps):
    U = np.eye(1 << n, dtype=complex)
    for a, rev in sweeps(angle, order, reps):
        seq = list(reversed(terms)) if rev else terms
        for ch, t, c in seq:
            U = expm(-1j * a * c.real * pauli_op(n, ch, t)) @ U
    return U


@pytest.mark.parametrize("dens", [False, True])
@pytest.mark.parametrize("seed", [1, 2])
def test_multiply_pauli_sum_is_left_multiplication(dens, seed):
    n = 4 if dens else 6
    terms = random_terms(n, 7, seed, complex_coeffs=True)
    st = O.State(n, dens)
    st.a[:] = qc.random_state(n, seed + 10, dens)
    H = hamiltonian(n, terms)
    want = H @ (as_matrix(st.a, n) if dens else st.a)
    st.multiply_pauli_sum(terms)
    got = as_matrix(st.a, n) if dens else st.a
    assert np.max(np.abs(got - want)) <= 1e-12 * max(1.0, np.max(np.abs(want)))


@pytest.mark.parametrize("dens", [False, True])
@pytest.mark.parametrize("order,reps", [(1, 1), (1, 3), (2, 2), (4, 1), (6, 2)])
def test_trotter_matches_product_of_exponentials(dens, order, reps):
    n = 4 if dens else 5
    terms = random_terms(n, 5, 100 + order + reps)
    angle = 0.37
    st = O.State(n, dens)
    st.a[:] = qc.random_state(n, 7, dens)
    U = brute_trotter(n, terms, angle, order, reps)
    want = U @ st.a if not dens else (U @ as_matrix(st.a, n) @ U.conj().T).reshape(-1, order="F")
    st.trotter(terms, angle, order, reps)
    assert np.max(np.abs(st.a - want)) <= 1e-12


def test_commuting_z_sum_is_exact_in_one_sweep():
    n = 6
    terms = random_terms(n, 6, 5, letters="Z")
    st = O.State(n)
    st.a[:] = qc.random_state(n, 3)
    want = expm(-1j * 0.9 * hamiltonian(n, terms)) @ st.a
    st.trotter(terms, 0.9, 1, 1)
    assert np.max(np.abs(st.a - want)) <= 1e-12


def test_second_order_converges_to_the_exponential():
    n = 4
    terms = random_terms(n, 6, 9)
    st = O.State(n)
    st.a[:] = qc.random_state(n, 4)
    want = expm(-1j * 0.3 * hamiltonian(n, terms)) @ st.a
    errs = []
    for reps in (10, 40):
        s = O.State(n)
        s.a[:] = st.a
        s.trotter(terms, 0.3, 2, reps)
        errs.append(np.linalg.norm(s.a - want))
    # second order: error ~ 1/reps^2
    assert errs[1] < errs[0] / 8 and errs[1] < 1e-3


def test_zero_angle_is_identity():
    st = O.State(3)
    st.a[:] = qc.random_state(3, 1)
    before = st.a.copy()
    st.trotter(random_terms(3, 3, 2), 0.0, 2, 3)
    assert np.array_equal(st.a, before)
