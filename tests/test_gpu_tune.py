"""The fusion autotuner (qa_tuneFusion): it times a queued batch of gates
under candidate fusion shapes on a scratch copy of the state, keeps the
fastest and applies the batch with it.

In exact fusion mode every shape applies each gate with its single-gate
kernel's arithmetic in queue order, so the state after tuning must equal
the unfused run bitwise whatever shape was chosen; the choice must be the
fastest shape timed; a second batch with the same gate mix reuses the
cached choice without timing; an empty queue is an input error.
"""

import sys

import numpy as np
import pytest

from quest_amd import circuits as qc

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


@pytest.fixture(autouse=True)
def restore(api):
    yield
    api.set_fusion(True)
    api.set_fusion_exact(False)
    api.set_fusion_span(True)
    api.set_fusion_jit(1, 2)


def gates_only(c):
    return qc.Circuit(c.num_qubits, c.dens, [op for op in c.ops if not op[0].startswith("OUT_")], c.init, c.name)


def test_tuned_exact_run_matches_unfused(api):
    c = gates_only(qc.layered(18, 4, 31, ctrl3=True, qft=True, gadgets=12, gadget_range=(4, 12), num_expec=0,
                              probs=False))
    say(f"tune case: {c.num_qubits} qubits, {len(c.ops)} gates")
    api.set_fusion(False)
    q = api.createQureg(c.num_qubits)
    q, _ = qc.run(c, q)
    want = api.get_amps(q)
    api.destroyQureg(q)

    api.set_fusion(True, 13, 3, 4096)
    api.set_fusion_exact(True)
    api.set_fusion_jit(0)
    q = api.createQureg(c.num_qubits)
    try:
        q, _ = qc.run(c, q)  # queued, not applied: no reduction flushes it
        res = api.tune_fusion(q, retune=True)
        got = api.get_amps(q)
        say(f"  trials: {[(t['tile_qubits'], t['low_qubits'], t['span_qubits'], round(t['ms'], 2)) for t in res['trials']]}")
        say(f"  chose {res['best']}")
        timed = [t for t in res["trials"] if t["ms"] >= 0]
        assert len(res["trials"]) >= 5 and res["best"] is not None
        assert res["best"]["ms"] == min(t["ms"] for t in timed)
        cfg = api.get_fusion()
        assert (cfg[1], cfg[2]) == (res["best"]["tile_qubits"], res["best"]["low_qubits"])
        assert np.array_equal(got, want), f"max |diff| {np.max(np.abs(got - want))}"

        # the same gate mix again: the cached choice, no timing
        q2 = api.createQureg(c.num_qubits)
        q2, _ = qc.run(c, q2)
        again = api.tune_fusion(q2)
        assert len(again["trials"]) == 1 and again["best"]["tile_qubits"] == res["best"]["tile_qubits"]
        api.destroyQureg(q2)
    finally:
        api.destroyQureg(q)


def test_tune_at_28_qubits(api):
    """The tuner's scratch copy of a 28-qubit complex128 state is 4 GiB (a
    size that overflows 32-bit arithmetic): tuned in exact mode, the state
    must be the unfused one -- every qubit's probabilities equal."""
    n = 28
    ops = []
    rng = qc.SplitMix64(5)
    for i in range(24):
        t = rng.randint(n)
        ops.append(("U1", t, qc.haar_unitary(rng, 2)))
        if i % 3 == 0:
            k = 4 + rng.randint(6)
            ops.append(("PG", qc.random_pauli_chars(rng, k, allow_identity=False), rng.sample(n, k), rng.angle()))
    c = qc.Circuit(n, False, ops, ("plus",), "tune28")
    say(f"tune case: {n} qubits, {len(ops)} gates")
    probs = []
    for tuned in (False, True):
        api.set_fusion(tuned, 13, 3, 4096)
        api.set_fusion_exact(True)
        api.set_fusion_jit(0)
        q = api.createQureg(n)
        try:
            q, _ = qc.run(c, q)
            if tuned:
                res = api.tune_fusion(q, retune=True)
                say(f"  chose {res['best']}")
            probs.append([api.calcProbOfQubitOutcome(q, t, 0) for t in range(n)])
        finally:
            api.destroyQureg(q)
    assert probs[0] == probs[1]


def test_tune_needs_queued_gates(api):
    api.set_fusion(True)
    q = api.createQureg(12)
    try:
        api.get_amps(q)  # nothing queued
        with pytest.raises(api.QuESTError):
            api.tune_fusion(q)
    finally:
        api.destroyQureg(q)
