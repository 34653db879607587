"""Span passes against the single-gate kernels.

A run of uncontrolled Pauli gadgets, Pauli tensors and phase gadgets whose
X/Y masks (with the always-local low qubits) span at most spanQubits
dimensions is applied as one span pass: one coset of that subspace per
workgroup, the gadgets in queue order with the single-gate kernels'
arithmetic.  With exact fusion the amplitudes must therefore be bitwise
equal to the unfused run (each gadget its own kernel, as the reference's
quest/src/gpu/gpu_subroutines.cpp:780 and :865 launch them); in the default
mode, where tile passes between the runs may reorder and merge gates, equal
to rounding.  Every case must also have used span passes.
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


def gadget_circuit(n, num, seed, dens=False, others=True, max_k=12):
    """Random Pauli gadgets (most), phase gadgets and Pauli tensors on up to
    max_k qubits, with an occasional one-qubit unitary or controlled gadget
    (which end a span run)."""
    rng = qc.SplitMix64(seed)
    c = qc.Circuit(n, dens, [], ("zero",) if dens else ("random", seed), f"gadgets_n{n}_s{seed}")
    if dens:
        for q in range(n):
            c.ops.append(("H", q))
    for _ in range(num):
        r = rng.randint(20)
        k = 1 + rng.randint(min(max_k, n))
        targs = rng.sample(n, k)
        if r < 12:
            c.ops.append(("PG", qc.random_pauli_chars(rng, k, allow_identity=False), targs, rng.angle()))
        elif r < 15:
            c.ops.append(("PHG", targs, rng.angle()))
        elif r < 17:
            c.ops.append(("PSTR", qc.random_pauli_chars(rng, k, allow_identity=False), targs))
        elif others and r == 17:
            c.ops.append(("U1", targs[0], qc.haar_unitary(rng, 2)))
        elif others and r == 18 and k >= 2 and k < n:
            rest = [q for q in range(n) if q not in targs]
            c.ops.append(("CPG", [rest[0]], qc.random_pauli_chars(rng, k, allow_identity=False), targs,
                          rng.angle()))
        else:
            c.ops.append(("PG", qc.random_pauli_chars(rng, k, allow_identity=False), targs, rng.angle()))
    return c


def run(api, c, fused, exact=True, span=True, precision=2, shape=(13, 3, 2048)):
    api.set_fusion(fused, *shape)
    api.set_fusion_exact(exact)
    api.set_fusion_span(span)
    q = api.createCustomPrecisionQureg(c.num_qubits, int(c.dens), precision)
    try:
        api.fusion_span_stats()
        q, _ = qc.run(c, q)
        amps = api.get_amps(q)
        return amps, api.fusion_span_stats()
    finally:
        api.destroyQureg(q)


EXACT_CASES = [
    # (name, circuit factory, precision, tile shape, span coset qubits)
    ("sv14", lambda: gadget_circuit(14, 120, 1), 2, (13, 3, 2048), 13),
    ("sv18_wide_gadgets", lambda: gadget_circuit(18, 160, 2, max_k=18), 2, (13, 3, 2048), 13),
    ("sv16_tile10_low0", lambda: gadget_circuit(16, 120, 3), 2, (10, 0, 2048), 13),
    ("sv16_tile11_low2", lambda: gadget_circuit(16, 120, 4), 2, (11, 2, 2048), 13),
    ("sv16_span12", lambda: gadget_circuit(16, 120, 13), 2, (13, 3, 2048), 12),
    ("sv15_span10_low4", lambda: gadget_circuit(15, 120, 14), 2, (12, 4, 2048), 10),
    ("sv16_fp32", lambda: gadget_circuit(16, 120, 5), 1, (13, 3, 2048), 13),
    ("sv16_fp32_span11", lambda: gadget_circuit(16, 120, 15), 1, (13, 3, 2048), 11),
    ("dm8", lambda: gadget_circuit(8, 80, 6, dens=True), 2, (13, 3, 2048), 13),
    ("sv20_only_gadgets", lambda: gadget_circuit(20, 200, 7, others=False), 2, (13, 3, 2048), 13),
    # gadgets wider than the tile: queued all the same, taken by span passes
    ("sv16_tile8_wide_gadgets", lambda: gadget_circuit(16, 120, 16, max_k=16), 2, (8, 2, 2048), 12),
    ("sv14_tile6_wide_gadgets_fp32", lambda: gadget_circuit(14, 100, 17, max_k=14), 1, (6, 2, 2048), 11),
]


@pytest.mark.parametrize("name,make,precision,shape,span", EXACT_CASES, ids=[c[0] for c in EXACT_CASES])
def test_span_passes_bitwise_equal_to_single_gates(api, name, make, precision, shape, span):
    c = make()
    say(f"span case {name}: {c.num_qubits} qubits, {len(c.ops)} ops, precision {precision}, tile {shape}, "
        f"span {span}")
    want, _ = run(api, c, fused=False, precision=precision, shape=shape)
    got, st = run(api, c, fused=True, exact=True, span=span, precision=precision, shape=shape)
    say(f"  {st['passes']} span passes took {st['gates']} gates")
    assert st["passes"] > 0
    assert np.array_equal(got, want), f"max |diff| {np.max(np.abs(got - want))}"


def test_span_passes_in_default_mode(api):
    """Default (reordering) fusion with and without span passes, and the
    30-qubit benchmark circuit family's shape at 20 qubits (layers, QFT and
    gadgets of 4-12 qubits)."""
    for c in [gadget_circuit(18, 160, 11),
              qc.layered(20, 3, 5, ctrl3=True, qft=True, gadgets=24, gadget_range=(4, 12), num_expec=0,
                         probs=False)]:
        say(f"span default-mode case {c.name}: {c.num_qubits} qubits, {len(c.ops)} ops")
        want, _ = run(api, c, fused=False)
        off, st0 = run(api, c, fused=True, exact=False, span=False)
        on, st1 = run(api, c, fused=True, exact=False, span=True)
        say(f"  span passes {st1['passes']} ({st1['gates']} gates); without: {st0['passes']}")
        assert st0["passes"] == 0 and st1["passes"] > 0
        assert np.allclose(on, want, rtol=0, atol=1e-11), np.max(np.abs(on - want))
        assert np.allclose(off, want, rtol=0, atol=1e-11), np.max(np.abs(off - want))


def test_span_setting_round_trip(api):
    api.set_fusion_span(False)
    c = gadget_circuit(14, 40, 12)
    _, st = run(api, c, fused=True, exact=True, span=False)
    assert st["passes"] == 0
    g = api.gpu()
    for bad in (1, 9, 14, -1):
        assert g.qa_setFusionSpan(bad) != 0
