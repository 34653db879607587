"""Dense matrices on 3 and 4 qubits inside fused passes (DENSE_LDS ops).

An uncontrolled applyCompMatr on 3 or 4 targets -- and so every 2-qubit
Kraus map of a density matrix, a 16x16 superoperator on (q1, q2, q1+N,
q2+N) -- is queued with a copy of its matrix and applied by the tile pass
to the tile in LDS, each thread taking the groups of 2^n amplitudes that
share the other tile bits, with the single-gate kernel's arithmetic
(k_fewTargDense: terms accumulated from zero in column order).  Only
generated pass kernels run such ops, so their passes compile at first
sight.  Exact fusion must therefore give bitwise the unfused amplitudes;
the default planner equal to rounding; and the passes must have held the
ops (no unfused 3-4 target kernel launched).
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
    api.set_fusion_jit(1, 2)


def dense_circuit(n, num, seed, dens=False):
    rng = qc.SplitMix64(seed)
    c = qc.Circuit(n, dens, [], ("zero",) if dens else ("random", seed), f"dense_n{n}_s{seed}")
    if dens:
        for q in range(n):
            c.ops.append(("H", q))
    for i in range(num):
        r = rng.randint(6)
        if r < 2:
            k = 3 + rng.randint(2)
            c.ops.append(("UN", rng.sample(n, k), qc.haar_unitary(rng, 1 << k)))
        elif r == 2 and dens:
            qs = rng.sample(n, 2)
            c.ops.append(("KRAUS", qs, qc.kraus_from_isometry(qc.random_isometry(rng, 16, 4), 4, 4)))
        elif r == 3:
            c.ops.append(("U2", *rng.sample(n, 2), qc.haar_unitary(rng, 4)))
        else:
            c.ops.append(("U1", rng.randint(n), qc.haar_unitary(rng, 2)))
    return c


def run(api, c, fused, exact=True, precision=2, shape=(13, 3, 2048)):
    api.set_fusion(fused, *shape)
    api.set_fusion_exact(exact)
    api.set_fusion_jit(1, 1)
    q = api.createCustomPrecisionQureg(c.num_qubits, int(c.dens), precision)
    try:
        api.fusion_stats()
        q, _ = qc.run(c, q)
        amps = api.get_amps(q)
        return amps, api.fusion_stats()
    finally:
        api.destroyQureg(q)


CASES = [
    # (name, circuit factory, precision, tile shape)
    ("sv14", lambda: dense_circuit(14, 60, 1), 2, (13, 3, 2048)),
    ("sv12_tile8", lambda: dense_circuit(12, 60, 2), 2, (8, 2, 2048)),
    ("sv15_fp32", lambda: dense_circuit(15, 60, 3), 1, (12, 4, 2048)),
    ("dm6_kraus", lambda: dense_circuit(6, 40, 4, dens=True), 2, (12, 3, 2048)),
    ("dm7_kraus_fp32", lambda: dense_circuit(7, 40, 5, dens=True), 1, (13, 3, 2048)),
]


@pytest.mark.parametrize("name,make,precision,shape", CASES, ids=[c[0] for c in CASES])
def test_dense_lds_ops_bitwise_equal_to_single_gates(api, name, make, precision, shape):
    c = make()
    say(f"dense-lds case {name}: {c.num_qubits} qubits, {len(c.ops)} ops, precision {precision}, tile {shape}")
    want, _ = run(api, c, fused=False, precision=precision, shape=shape)
    got, st = run(api, c, fused=True, exact=True, precision=precision, shape=shape)
    say(f"  {st['passes']} passes for {st['gates']} gates")
    if not c.dens:
        assert st["gates"] == len(c.ops), st  # every gate went through the passes
    assert np.array_equal(got, want), f"max |diff| {np.max(np.abs(got - want))}"


def test_dense_lds_ops_in_default_mode(api):
    for c in (dense_circuit(16, 80, 11), dense_circuit(7, 50, 12, dens=True)):
        say(f"dense-lds default-mode case {c.name}")
        want, _ = run(api, c, fused=False)
        got, st = run(api, c, fused=True, exact=False)
        assert st["passes"] > 0
        assert np.allclose(got, want, rtol=0, atol=1e-11), np.max(np.abs(got - want))


def test_noisy_density_family_kraus_fused(api):
    """The density-matrix benchmark family (Kraus map every 2 layers): fused
    default mode against the unfused run."""
    c = qc.noisy_density(7, 6, 21, kraus_every=2, num_expec=0)
    c.ops = [op for op in c.ops if not op[0].startswith("OUT_")]
    want, _ = run(api, c, fused=False)
    got, st = run(api, c, fused=True, exact=False)
    assert np.allclose(got, want, rtol=0, atol=1e-11), np.max(np.abs(got - want))
