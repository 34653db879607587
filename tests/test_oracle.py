"""Pins the CPU oracle (oracle/qoracle.c + oracle/qoracle.py) before it is
trusted as the checker of the HIP path.

  1. against the golden fixtures written by QuEST's own CPU build
     (tests/golden/make_golden.py, oracle/_ref);
  2. against brute force: full 2^n x 2^n operators built independently with
     numpy Kronecker products (the method of the reference's tests,
     tests/utils/evolve.cpp:117-196), density matrices as U rho U^dagger and
     channels as sum_k K rho K^dagger (evolve.cpp:252-264);
  3. against closed forms: the debug state (cpu_subroutines.cpp:2364-2377),
     the QFT of a basis state (a DFT column), channel limits
     (tests/integration/densitymatrix.cpp:221-227).
A planted mistake in the oracle (a transposed matrix) must fail check 1.
"""

import math

import numpy as np
import pytest

from conftest import GOLDEN_CASES, SAMPLED_CASES, compare_results, load_golden, rel_l2
from quest_amd import circuits as qc
from oracle import qoracle as O

PAULI = {"I": np.eye(2), "X": np.array([[0, 1], [1, 0]]), "Y": np.array([[0, -1j], [1j, 0]]),
         "Z": np.diag([1, -1])}


# ------------------------------------------------------------- 1. golden

@pytest.mark.parametrize("name", [c for c in GOLDEN_CASES if c not in SAMPLED_CASES])
def test_oracle_matches_reference_fixture(name):
    circ, init, amps, sample, results = load_golden(name)
    got, res = O.run(circ, init_amps=init) if init is not None else O.run(circ)
    if sample is not None:
        got = got[sample]
    assert rel_l2(got, amps) <= 1e-12
    compare_results(res, results, tol=1e-10)


def test_planted_transpose_is_caught():
    """A plausible mistake (applying M^T instead of M) must fail the fixture check."""
    circ, init, amps, _, _ = load_golden("layered_n10")
    bad = qc.Circuit(circ.num_qubits, circ.dens, [], circ.init, circ.name)
    for op in circ.ops:
        if op[0] == "U1":
            m = np.array(op[2])
            op = ("U1", op[1], m.T.tolist())
        bad.ops.append(op)
    got, _ = O.run(bad, init_amps=init)
    assert rel_l2(got, amps) > 1e-3


# ------------------------------------------------------------- 2. brute force

def op_on(n, mats: dict, ctrls=(), states=None):
    """Full operator: mats maps qubit -> 2x2 (or one 2^k matrix on a qubit
    tuple), applied when every control is in its state."""
    dim = 1 << n
    out = np.zeros((dim, dim), dtype=complex)
    states = list(states) if states is not None else [1] * len(ctrls)
    for col in range(dim):
        if any(((col >> c) & 1) != s for c, s in zip(ctrls, states)):
            out[col, col] = 1
            continue
        vec = np.zeros(dim, dtype=complex)
        vec[col] = 1
        for targs, m in mats.items():
            k = len(targs)
            nv = np.zeros(dim, dtype=complex)
            for i in np.nonzero(vec)[0]:
                sub = sum(((i >> t) & 1) << b for b, t in enumerate(targs))
                base = i
                for t in targs:
                    base &= ~(1 << t)
                for r in range(1 << k):
                    j = base
                    for b, t in enumerate(targs):
                        if (r >> b) & 1:
                            j |= 1 << t
                    nv[j] += m[r, sub] * vec[i]
            vec = nv
        out[:, col] = vec
    return out


def brute_apply(state, dens, n, U):
    if not dens:
        return U @ state
    rho = state.reshape(1 << n, 1 << n, order="F")
    return (U @ rho @ U.conj().T).reshape(-1, order="F")


def random_state_vec(n, seed, dens):
    return qc.random_state(n, seed, dens)


@pytest.mark.parametrize("dens", [False, True])
@pytest.mark.parametrize("seed", [1, 2, 3])
def test_oracle_ops_match_brute_force(dens, seed):
    n = 3 if dens else 5
    rng = qc.SplitMix64(seed)
    init = random_state_vec(n, seed, dens)
    c = qc.Circuit(n, dens, [], ("random", seed), "bf")
    c.ops = [("U1", 1, qc.haar_unitary(rng, 2)), ("CU1", [0], 2, qc.haar_unitary(rng, 2)),
             ("U2", 2, 0, qc.haar_unitary(rng, 4)), ("UN", [1, 0, 2], qc.haar_unitary(rng, 8)),
             ("PS2", 0, 2, rng.angle()), ("SWAP", 0, 2), ("PG", "XYZ", [2, 0, 1], rng.angle()),
             ("PSTR", "YZ", [0, 1]), ("PHG", [0, 2], rng.angle()), ("RY", 1, rng.angle()), ("QFT", [0, 1, 2])]
    got, _ = O.run(c, init_amps=init)
    state = init.copy()
    for op in c.ops:
        k = op[0]
        if k == "U1":
            U = op_on(n, {(op[1],): np.array(op[2])})
        elif k == "CU1":
            U = op_on(n, {(op[2],): np.array(op[3])}, ctrls=op[1])
        elif k in ("U2",):
            U = op_on(n, {(op[1], op[2]): np.array(op[3])})
        elif k == "UN":
            U = op_on(n, {tuple(op[1]): np.array(op[2])})
        elif k == "PS2":
            U = op_on(n, {(op[1],): np.diag([1, np.exp(1j * op[3])])}, ctrls=[op[2]])
        elif k == "SWAP":
            U = op_on(n, {(op[1], op[2]): np.array([[1, 0, 0, 0], [0, 0, 1, 0], [0, 1, 0, 0], [0, 0, 0, 1]])})
        elif k in ("PG", "PSTR", "PHG", "RY"):
            chars, targs = (("Z" * len(op[1]), op[1]) if k == "PHG" else ("Y", [op[1]]) if k == "RY"
                            else (op[1], op[2]))
            P = op_on(n, {(t,): PAULI[ch] for ch, t in zip(chars, targs)})
            if k == "PSTR":
                U = P
            else:
                theta = op[-1]
                U = math.cos(theta / 2) * np.eye(1 << n) - 1j * math.sin(theta / 2) * P
        elif k == "QFT":
            targs = op[1]
            m = len(targs)
            F = np.array([[np.exp(2j * np.pi * r * s / (1 << m)) for s in range(1 << m)]
                          for r in range(1 << m)]) / math.sqrt(1 << m)
            U = op_on(n, {tuple(targs): F})
        state = brute_apply(state, dens, n, U)
    assert rel_l2(got, state) < 1e-12


@pytest.mark.parametrize("kind", ["DEPH", "DEPH2", "DEPOL", "DEPOL2", "DAMP", "PAULIS", "KRAUS1", "KRAUS2"])
def test_oracle_channels_match_kraus_sum(kind):
    n = 3
    rho0 = qc.random_state(n, 11, True)
    rho = rho0.reshape(8, 8, order="F")
    rng = qc.SplitMix64(5)
    I, X, Y, Z = (PAULI[c] for c in "IXYZ")
    p = 0.2
    if kind == "DEPH":
        op, ks = ("DEPH", 1, p), [(math.sqrt(1 - p) * I, (1,)), (math.sqrt(p) * Z, (1,))]
    elif kind == "DEPH2":
        op = ("DEPH2", 0, 2, p)
        ks = [(math.sqrt(1 - p) * np.kron(I, I), (0, 2))] + [
            (math.sqrt(p / 3) * np.kron(b, a), (0, 2)) for a, b in ((Z, I), (I, Z), (Z, Z))]
    elif kind == "DEPOL":
        op, ks = ("DEPOL", 2, p), [(math.sqrt(1 - p) * I, (2,))] + [(math.sqrt(p / 3) * P, (2,)) for P in (X, Y, Z)]
    elif kind == "DEPOL2":
        paulis = [np.kron(b, a) for a in (I, X, Y, Z) for b in (I, X, Y, Z)]
        op = ("DEPOL2", 1, 0, p)
        ks = [(math.sqrt(1 - p) * paulis[0], (1, 0))] + [(math.sqrt(p / 15) * P, (1, 0)) for P in paulis[1:]]
    elif kind == "DAMP":
        op = ("DAMP", 0, p)
        ks = [(np.array([[1, 0], [0, math.sqrt(1 - p)]]), (0,)), (np.array([[0, math.sqrt(p)], [0, 0]]), (0,))]
    elif kind == "PAULIS":
        px, py, pz = 0.05, 0.1, 0.15
        op = ("PAULIS", 1, px, py, pz)
        ks = [(math.sqrt(1 - px - py - pz) * I, (1,)), (math.sqrt(px) * X, (1,)), (math.sqrt(py) * Y, (1,)),
              (math.sqrt(pz) * Z, (1,))]
    elif kind == "KRAUS1":
        mats = qc.kraus_from_isometry(qc.random_isometry(rng, 6, 2), 3, 2)
        op, ks = ("KRAUS", [2], mats), [(np.array(m), (2,)) for m in mats]
    else:
        mats = qc.kraus_from_isometry(qc.random_isometry(rng, 16, 4), 4, 4)
        op, ks = ("KRAUS", [2, 0], mats), [(np.array(m), (2, 0)) for m in mats]
    c = qc.Circuit(n, True, [op], ("random", 0), "ch")
    got, _ = O.run(c, init_amps=rho0)
    want = sum(op_on(n, {t: K}) @ rho @ op_on(n, {t: K}).conj().T for K, t in ks)
    assert rel_l2(got, want.reshape(-1, order="F")) < 1e-12


# ------------------------------------------------------------- 3. closed forms

def test_debug_state_closed_form():
    c = qc.Circuit(4, False, [], ("debug",), "dbg")
    got, _ = O.run(c)
    i = np.arange(16)
    want = np.empty(16, dtype=complex)
    want.real = (2 * i) / 10.0
    want.imag = (2 * i + 1) / 10.0
    np.testing.assert_array_equal(got, want)


def test_qft_of_basis_state_is_dft_column():
    n, k = 6, 13
    c = qc.Circuit(n, False, [("FQFT",)], ("zero",), "qft")
    st = O.State(n)
    st.a[:] = 0
    st.a[k] = 1
    for op in c.ops:
        O.apply_op(st, op, [])
    want = np.exp(2j * np.pi * k * np.arange(1 << n) / (1 << n)) / math.sqrt(1 << n)
    assert rel_l2(st.a, want) < 1e-13


def test_channel_limits():
    # full depolarising (p=3/4) maximally mixes the qubit; full damping restores |0>
    st = O.State(2, True)
    st.a[:] = qc.random_state(2, 3, True)
    O.apply_op(st, ("DEPOL", 0, 0.75), [])
    rho = st.a.reshape(4, 4, order="F")
    red = rho.reshape(2, 2, 2, 2, order="F")  # indices (r0, r1, c0, c1)
    r0 = np.einsum("ajbj->ab", red)
    np.testing.assert_allclose(r0, np.eye(2) / 2, atol=1e-14)
    O.apply_op(st, ("DAMP", 1, 1.0), [])
    res = []
    O.apply_op(st, ("OUT_PROB", 1, 0), res)
    assert abs(res[0][1] - 1) < 1e-14


def test_pure_state_purity_and_normalisation():
    st = O.State(4, True)
    st.a[:] = qc.random_state(4, 9, True)
    assert abs(st.purity() - 1) < 1e-13
    assert abs(st.total_prob() - 1) < 1e-13


# ------------------------------------------------------------- 4. two-state calculations

def test_overlaps_reduce_to_closed_forms():
    """calculations.cpp:49-61, 303-364 on states whose answers are known:
    rho = |psi><psi| makes the mixed forms equal the statevector ones
    (<phi|rho|phi> = |<phi|psi>|^2, Tr(rho sigma) = |<psi|phi>|^2), a state
    has fidelity 1 and distance 0 with itself, and the density-matrix
    outcome table of a pure state is the statevector's."""
    n = 4
    psi = qc.random_state(n, 11)
    phi = qc.random_state(n, 12)
    rho = qc.random_state(n, 11, dens=True)
    sigma = qc.random_state(n, 12, dens=True)
    sv = O.State(n, False)
    sv.a[:] = psi
    dm = O.State(n, True)
    dm.a[:] = rho
    ov = np.vdot(psi, phi)
    assert abs(sv.inner(phi, False) - ov) < 1e-14
    assert abs(sv.fidelity(phi, False) - abs(ov) ** 2) < 1e-14
    assert abs(sv.distance(phi, False) - math.sqrt(2 - 2 * abs(ov))) < 1e-14
    assert abs(sv.fidelity(psi, False) - 1) < 1e-14 and sv.distance(psi, False) < 1e-7
    # density matrices: Tr(rho^dagger sigma) of pure states, HS distance
    assert abs(dm.inner(sigma, True) - abs(ov) ** 2) < 1e-14
    hs = math.sqrt(2 - 2 * abs(ov) ** 2)  # ||rho - sigma||_F of two pure states
    assert abs(dm.distance(sigma, True) - hs) < 1e-13
    # mixed forms: <phi|rho|phi>, purified distance, both argument orders
    assert abs(dm.fidelity(phi, False) - abs(ov) ** 2) < 1e-14
    assert abs(dm.inner(phi, False) - abs(ov) ** 2) < 1e-14
    assert abs(dm.distance(phi, False) - math.sqrt(1 - abs(ov) ** 2)) < 1e-14
    sv2 = O.State(n, False)
    sv2.a[:] = phi
    assert abs(sv2.fidelity(rho, True) - abs(ov) ** 2) < 1e-14
    # outcome tables: pure density matrix = statevector, unsorted qubits
    for qs in ([2], [3, 0], [1, 3, 0, 2]):
        _, r1 = O.run(qc.Circuit(n, False, [("OUT_PROBS", qs)], ("random", 11)))
        _, r2 = O.run(qc.Circuit(n, True, [("OUT_PROBS", qs)], ("random", 11)))
        np.testing.assert_allclose(r1[0][1], r2[0][1], atol=1e-15)
        assert abs(sum(r2[0][1]) - 1) < 1e-14
        # against an explicit sum over the basis states
        want = np.zeros(1 << len(qs))
        for i, p in enumerate(np.abs(psi) ** 2):
            want[sum(((i >> q) & 1) << b for b, q in enumerate(qs))] += p
        np.testing.assert_allclose(r1[0][1], want, atol=1e-15)
