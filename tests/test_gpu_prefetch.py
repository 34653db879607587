"""Tile-pass prefetch modes against the single-gate kernels.

A persistent workgroup of a tile pass works through tiles t, t + grid, ...
and can load its next tile while working on the current one: all loads at
the tile's start (mode 1), spread over the generated program between its
LDS exchanges (mode 2, the default), or with the stores deferred into the
next tile's work as well, two register sets taking turns (mode 3).  None of
this changes any arithmetic, so in exact fusion mode every mode must give
the unfused amplitudes bitwise.  The sizes give a workgroup one tile (the
turn-taking loop leaves after its first set), two and four tiles (it leaves
after the second), on a device whose grid is a power of two.
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
    api.set_fusion_prefetch(2)
    api.set_fusion_jit(1, 2)


def amps_after(api, c, fused, prefetch=2):
    api.set_fusion(fused, 13, 3, 2048)
    api.set_fusion_exact(True)
    api.set_fusion_prefetch(prefetch)
    api.set_fusion_jit(2)
    q = api.createQureg(c.num_qubits)
    try:
        q, _ = qc.run(c, q)
        return api.get_amps(q)
    finally:
        api.destroyQureg(q)


@pytest.mark.parametrize("n", [21, 22, 23])
def test_prefetch_modes_bitwise_equal_to_single_gates(api, n):
    c = qc.layered(n, 2, 40 + n, ctrl3=True, qft=False, gadgets=6, gadget_range=(4, 10), num_expec=0,
                   probs=False)
    c = qc.Circuit(c.num_qubits, c.dens, [op for op in c.ops if not op[0].startswith("OUT_")], c.init, c.name)
    say(f"prefetch case: {n} qubits, {len(c.ops)} gates")
    want = amps_after(api, c, fused=False)
    for mode in (0, 1, 2, 3):
        got = amps_after(api, c, fused=True, prefetch=mode)
        say(f"  mode {mode}: max |diff| {np.max(np.abs(got - want))}")
        assert np.array_equal(got, want), f"mode {mode}"


def test_prefetch_mode_range(api):
    g = api.gpu()
    for bad in (-1, 4):
        assert g.qa_setFusionPrefetch(bad) != 0
