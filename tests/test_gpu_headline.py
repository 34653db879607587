"""The benchmark's own configuration checked directly: the 30-qubit
complex128 circuit step runs with fusion tiles of 2^13 amplitudes, 3
always-local low qubits and a 2048-gate queue (13, 3, 2048), planned by the
default (reordering, merging) planner.  Here that shape, in both the default
and the exact planner, is held against the reference's golden fixtures and
the oracle, and the full 30-qubit step (the same circuit family and size the
benchmark times) is checked through size-independent properties: the norm,
the single-qubit outcome probabilities, and the step followed by its inverse
returning the initial state.

Tolerances: amplitudes 1e-12 relative L2, results 1e-10 (as the other parity
tests); at 30 qubits norm and fidelity within 1e-10.
"""

import sys

import numpy as np
import pytest

from conftest import GOLDEN_CASES, compare_results, load_golden, rel_l2
from quest_amd import circuits as qc

pytestmark = pytest.mark.gpu

AMP_TOL = 1e-12
SCALAR_TOL = 1e-10
BENCH_SHAPE = (13, 3, 2048)


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


@pytest.fixture
def bench_fusion(api):
    def use(exact):
        api.set_fusion(True, *BENCH_SHAPE)
        api.set_fusion_exact(exact)
    yield use
    api.set_fusion(True)
    api.set_fusion_exact(False)


@pytest.mark.parametrize("exact", [False, True])
@pytest.mark.parametrize("name", GOLDEN_CASES)
def test_golden_fixture_at_bench_shape(api, bench_fusion, name, exact):
    say(f"[gpu] golden {name} at {BENCH_SHAPE} exact={exact}")
    bench_fusion(exact)
    circ, init, amps, sample, results = load_golden(name)
    q = api.createDensityQureg(circ.num_qubits) if circ.dens else api.createQureg(circ.num_qubits)
    try:
        if init is not None:
            api.set_amps(q, init)
            circ.init = ("given",)
        q, res = qc.run(circ, q)
        got = api.get_amps(q)
        if sample is not None:
            got = got[sample]
        assert rel_l2(got, amps) <= AMP_TOL
        compare_results(res, results, tol=SCALAR_TOL)
    finally:
        api.destroyQureg(q)


@pytest.mark.parametrize("exact", [False, True])
@pytest.mark.parametrize("n,seed", [(14, 1), (17, 2), (20, 3)])
def test_layered_family_at_bench_shape(api, bench_fusion, n, seed, exact):
    from oracle import qoracle as O
    say(f"[gpu] layered n={n} seed={seed} at {BENCH_SHAPE} exact={exact}")
    bench_fusion(exact)
    c = qc.layered(n, 6, seed, ctrl3=True, qft=True, gadgets=8, gadget_range=(4, min(12, n)), num_expec=4)
    want, wres = O.run(c)
    q, res = qc.run(c)
    try:
        assert rel_l2(api.get_amps(q), want) <= AMP_TOL
        compare_results(res, wres, tol=SCALAR_TOL)
    finally:
        api.destroyQureg(q)


def test_benchmark_step_at_30_qubits(api, bench_fusion):
    """The benchmark's step (its circuit family: 40 layers of 30 Haar
    one-qubit gates, 15 two-qubit phase shifts and a 3-control gate, a full
    QFT, 64 Pauli gadgets of 4-12 qubits) on a 30-qubit complex128 state at
    the benchmark's fusion shape: the norm stays 1, every qubit's outcome
    probabilities sum to the total, and the step followed by its inverse
    returns the initial state (fidelity 1)."""
    n = 30
    say(f"[gpu] benchmark step n={n} at {BENCH_SHAPE}")
    bench_fusion(False)
    c = qc.layered(n, 40, 2024, ctrl3=True, qft=True, gadgets=64, gadget_range=(4, 12), num_expec=0, probs=False)
    q = api.createQureg(n)
    start = None
    try:
        api.initRandomPureState(q)
        start = api.createCloneQureg(q)
        qc.run(qc.Circuit(n, False, c.ops, ("given",), "c2_step"), q)
        norm = api.calcTotalProb(q)
        say(f"[gpu] norm after the step: {norm!r}")
        assert abs(norm - 1) <= SCALAR_TOL
        p0 = np.array([api.calcProbOfQubitOutcome(q, t, 0) for t in range(n)])
        p1 = np.array([api.calcProbOfQubitOutcome(q, t, 1) for t in range(n)])
        assert np.all(p0 >= 0) and np.all(p1 >= 0)
        assert np.max(np.abs(p0 + p1 - norm)) <= SCALAR_TOL
        # the Haar layers leave no qubit near a basis state
        assert np.all((p0 > 1e-3) & (p0 < 1 - 1e-3))
        qc.run(qc.Circuit(n, False, qc.inverse_ops(c), ("given",), "c2_inverse"), q)
        fid = api.calcFidelity(q, start)
        say(f"[gpu] fidelity of step + inverse with the start: {fid!r}")
        assert abs(fid - 1) <= SCALAR_TOL
    finally:
        if start is not None:
            api.destroyQureg(start)
        api.destroyQureg(q)
