"""Gate fusion: gates queued per state and applied by LDS-tiled passes must
give amplitudes bit-identical to the single-gate kernels (fusion off), for
every tile shape, queue length, precision and state kind, and must agree
with the CPU oracle within the usual 1e-12 relative L2.

The circuits mix every fusable gate kind (one/two-qubit dense, one-qubit
diagonal and phase-only diagonal, swap, Pauli tensor and gadget, phase
gadget, one-qubit dephasing/depolarising) with controls on local and
non-local qubits, and interleave gates that force a flush (multi-target
dense and diagonal matrices, two-qubit channels, Kraus maps, reductions).
"""

import sys

import numpy as np
import pytest

from conftest import compare_results, rel_l2
from quest_amd import circuits as qc

pytestmark = pytest.mark.gpu

AMP_TOL = 1e-12
SCALAR_TOL = 1e-10

# (tileQubits, lowQubits): default, smallest, no always-local rows, widest, and
# the benchmark's (13, 3)
TILE_SHAPES = [(12, 4), (4, 0), (6, 2), (8, 5), (13, 8), (11, 3), (13, 3)]


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def fusion_mix(n: int, num: int, seed: int, dens: bool = False, flushers: bool = True) -> qc.Circuit:
    """Random circuit over the fusable gate kinds (oracle-supported ops only)."""
    rng = qc.SplitMix64(seed)
    c = qc.Circuit(n, dens, [], ("zero",) if dens else ("random", seed), f"fusion_n{n}_s{seed}")
    if dens:
        for q in range(n):
            c.ops.append(("H", q))
    for i in range(num):
        r = rng.randint(16 if dens else 14)
        t = rng.randint(n)
        others = [q for q in range(n) if q != t]
        u = others[rng.randint(len(others))] if others else t
        if r == 0:
            c.ops.append(("U1", t, qc.haar_unitary(rng, 2)))
        elif r == 1 and n >= 2:
            nc = 1 + rng.randint(min(3, n - 1))
            ctrls = [others[j] for j in rng.sample(len(others), nc)]
            c.ops.append(("CU1", ctrls, t, qc.haar_unitary(rng, 2)))
        elif r == 2 and n >= 2:
            c.ops.append(("U2", t, u, qc.haar_unitary(rng, 4)))
        elif r == 3:
            a = rng.angle()
            c.ops.append(("D1", t, [complex(np.cos(a), np.sin(a)), complex(np.cos(2 * a), -np.sin(2 * a))]))
        elif r == 4:
            c.ops.append(("PS", t, rng.angle()))
        elif r == 5:
            c.ops.append((["H", "X", "Y", "Z", "S", "T"][rng.randint(6)], t))
        elif r == 6 and n >= 2:
            c.ops.append((["CZ", "CX", "SWAP"][rng.randint(3)], t, u))
        elif r == 7 and n >= 3:
            nc = 1 + rng.randint(min(2, n - 2))
            rest = [q for q in others if q != u]
            ctrls = [rest[j] for j in rng.sample(len(rest), nc)]
            c.ops.append(("CSWAP", ctrls, t, u))
        elif r == 8:
            k = 1 + rng.randint(min(4, n))
            targs = rng.sample(n, k)
            c.ops.append(("PSTR", qc.random_pauli_chars(rng, k, allow_identity=False), targs))
        elif r == 9:
            k = 1 + rng.randint(min(5, n))
            targs = rng.sample(n, k)
            c.ops.append(("PG", qc.random_pauli_chars(rng, k, allow_identity=False), targs, rng.angle()))
        elif r == 10 and n >= 3:
            k = 1 + rng.randint(min(3, n - 1))
            targs = rng.sample(n, k)
            rest = [q for q in range(n) if q not in targs]
            ctrls = [rest[j] for j in rng.sample(len(rest), 1 + rng.randint(min(2, len(rest))))]
            c.ops.append(("CPG", ctrls, qc.random_pauli_chars(rng, k, allow_identity=False), targs, rng.angle()))
        elif r == 11:
            k = 1 + rng.randint(min(6, n))
            c.ops.append(("PHG", rng.sample(n, k), rng.angle()))
        elif r == 12:
            c.ops.append((["RX", "RY", "RZ"][rng.randint(3)], t, rng.angle()))
        elif r == 13 and flushers and n >= 3 and rng.randint(4) == 0:
            # gates the queue does not take: they flush it first
            k = rng.randint(3)
            if k == 0:
                c.ops.append(("UN", rng.sample(n, 3), qc.haar_unitary(rng, 8)))
            elif k == 1:
                c.ops.append(("PS2", t, u, rng.angle()))
            else:
                c.ops.append(("OUT_TOTALPROB",))
        elif r == 14:
            c.ops.append(("DEPOL", t, 0.05 * rng.uniform()))
        elif r == 15:
            c.ops.append(("DEPH", t, 0.05 * rng.uniform()))
    if dens and flushers and n >= 2:
        c.ops.append(("DEPOL2", 0, n - 1, 0.03))
        c.ops.append(("DEPOL", n // 2, 0.02))
    c.ops.append(("OUT_TOTALPROB",))
    c.ops.append(("OUT_EXPEC", qc.random_pauli_chars(rng, n), list(range(n))))
    return c


def run_amps(api, c, fusion, precision=2, exact=True):
    """Amplitudes (complex128 host array) and results of circuit c; fused
    runs are exact (queue order, single-gate arithmetic) unless `exact` is
    false (commuting gates reordered, runs on one or two qubits merged)."""
    if fusion is None:
        api.set_fusion(False)
    else:
        api.set_fusion(True, *fusion)
    api.set_fusion_exact(exact)
    q = api.createCustomPrecisionQureg(c.num_qubits, int(c.dens), precision)
    try:
        q, res = qc.run(c, q)
        return api.get_amps(q), res
    finally:
        api.destroyQureg(q)
        api.set_fusion(True)
        api.set_fusion_exact(False)


@pytest.fixture(autouse=True)
def default_fusion(api):
    yield
    api.set_fusion(True)
    api.set_fusion_exact(False)


def check_fused(got, gres, ref, rres, exact):
    """Exact fused amplitudes are bit-identical to the unfused kernels;
    reordered and merged, they differ by rounding only."""
    if not exact:
        assert rel_l2(got, ref) <= AMP_TOL
        compare_results(gres, rres, tol=SCALAR_TOL)
    else:
        assert np.array_equal(got, ref), f"max |diff| {np.max(np.abs(got - ref))}"
        compare_results(gres, rres, tol=0.0)


@pytest.mark.parametrize("exact", [True, False])
@pytest.mark.parametrize("shape", TILE_SHAPES)
@pytest.mark.parametrize("n,seed", [(10, 11), (16, 12), (21, 13)])
def test_statevector_fused_matches_unfused(api, shape, n, seed, exact):
    say(f"[gpu] fusion sv n={n} seed={seed} shape={shape} exact={exact}")
    c = fusion_mix(n, 150, seed)
    ref, rres = run_amps(api, c, None)
    got, gres = run_amps(api, c, shape, exact=exact)
    check_fused(got, gres, ref, rres, exact)


@pytest.mark.parametrize("exact", [True, False])
@pytest.mark.parametrize("shape", TILE_SHAPES)
@pytest.mark.parametrize("n,seed", [(3, 21), (6, 22), (9, 23)])
def test_density_matrix_fused_matches_unfused(api, shape, n, seed, exact):
    say(f"[gpu] fusion dm n={n} seed={seed} shape={shape} exact={exact}")
    c = fusion_mix(n, 120, seed, dens=True)
    ref, rres = run_amps(api, c, None)
    got, gres = run_amps(api, c, shape, exact=exact)
    check_fused(got, gres, ref, rres, exact)


@pytest.mark.parametrize("n,seed,dens", [(12, 31, False), (17, 32, False), (7, 33, True)])
def test_fused_matches_oracle(api, n, seed, dens):
    from oracle import qoracle as O
    say(f"[gpu] fusion vs oracle n={n} seed={seed} dens={dens}")
    c = fusion_mix(n, 200, seed, dens=dens)
    want, wres = O.run(c)
    got, gres = run_amps(api, c, (12, 4), exact=False)
    assert rel_l2(got, want) <= AMP_TOL
    compare_results(gres, wres, tol=SCALAR_TOL)


def reorder_traps(n: int, num: int, seed: int) -> qc.Circuit:
    """Gates crowded onto a few high qubits, mixing dense gates with gates
    that are diagonal on (or controlled by) the same qubits: a pass that
    passes over a dense gate on qubit q may take later gates that avoid q or
    that are diagonal on q only when q is no dense target of theirs and they
    involve no dense target of the skipped gate."""
    rng = qc.SplitMix64(seed)
    c = qc.Circuit(n, False, [], ("random", seed), f"traps_n{n}_s{seed}")
    hot = list(range(n - 6, n)) + [0, 1]
    for _ in range(num):
        t, u = (hot[j] for j in rng.sample(len(hot), 2))
        r = rng.randint(9)
        if r == 0:
            c.ops.append(("U1", t, qc.haar_unitary(rng, 2)))
        elif r == 1:
            c.ops.append(("CU1", [u], t, qc.haar_unitary(rng, 2)))
        elif r == 2:
            c.ops.append(("PS", t, rng.angle()))
        elif r == 3:
            c.ops.append(("CZ", t, u))
        elif r == 4:
            c.ops.append(("SWAP", t, u))
        elif r == 5:
            c.ops.append(("U2", t, u, qc.haar_unitary(rng, 4)))
        elif r == 6:
            c.ops.append(("PG", qc.random_pauli_chars(rng, 2, allow_identity=False), [t, u], rng.angle()))
        elif r == 7:
            c.ops.append(("PHG", [t, u], rng.angle()))
        else:
            c.ops.append(("CX", u, t))
    c.ops.append(("OUT_EXPEC", qc.random_pauli_chars(rng, n), list(range(n))))
    return c


@pytest.mark.parametrize("shape", [(4, 0), (5, 2), (6, 2), (8, 4)])
@pytest.mark.parametrize("seed", [81, 82])
def test_reordered_passes_respect_gate_order(api, shape, seed):
    """Small tiles over a circuit crowded onto few qubits force many passes
    in which gates are passed over; reordered passes match the oracle and
    the unfused kernels, and take fewer passes than in-order ones."""
    from oracle import qoracle as O
    say(f"[gpu] fusion reorder traps shape={shape} seed={seed}")
    c = reorder_traps(12, 300, seed)
    want, wres = O.run(c)
    ref, _ = run_amps(api, c, None)
    passes = {}
    for exact in (True, False):
        api.fusion_stats()
        got, gres = run_amps(api, c, shape, exact=exact)
        passes[exact] = api.fusion_stats()["passes"]
        assert rel_l2(got, want) <= AMP_TOL
        assert rel_l2(got, ref) <= AMP_TOL
        compare_results(gres, wres, tol=SCALAR_TOL)
    assert passes[False] <= passes[True], passes


@pytest.mark.parametrize("shape", [(12, 4), (5, 1)])
@pytest.mark.parametrize("dens", [False, True])
def test_fp32_fused_is_bit_identical(api, shape, dens):
    n = 7 if dens else 15
    say(f"[gpu] fusion fp32 n={n} dens={dens} shape={shape}")
    c = fusion_mix(n, 150, 41 + dens, dens=dens)
    ref, _ = run_amps(api, c, None, precision=1)
    got, _ = run_amps(api, c, shape, precision=1)
    assert np.array_equal(got, ref)


@pytest.mark.parametrize("max_queue", [1, 2, 7, 4096])
def test_queue_length_does_not_change_results(api, max_queue):
    say(f"[gpu] fusion maxQueue={max_queue}")
    c = fusion_mix(14, 120, 51)
    ref, _ = run_amps(api, c, None)
    got, _ = run_amps(api, c, (10, 3, max_queue))
    assert np.array_equal(got, ref)
    got, _ = run_amps(api, c, (10, 3, max_queue), exact=False)
    assert rel_l2(got, ref) <= AMP_TOL


def test_interleaved_states_keep_separate_queues(api):
    """Gates on two quregs interleaved: each queue holds only its own state's
    gates, and a reduction on one state flushes both."""
    say("[gpu] fusion interleaved quregs")
    ca, cb = fusion_mix(13, 80, 61, flushers=False), fusion_mix(11, 80, 62, flushers=False)
    ra, _ = run_amps(api, ca, None)
    rb, _ = run_amps(api, cb, None)
    api.set_fusion(True, 9, 3, 512)
    api.set_fusion_exact(True)
    qa = api.createQureg(13)
    qb = api.createQureg(11)
    try:
        api.set_amps(qa, qc.initial_amps(ca))
        api.set_amps(qb, qc.initial_amps(cb))
        sink = []
        gates_a = [op for op in ca.ops if not op[0].startswith("OUT_")]
        gates_b = [op for op in cb.ops if not op[0].startswith("OUT_")]
        for i in range(max(len(gates_a), len(gates_b))):
            if i < len(gates_a):
                qc.apply_op(api, qa, gates_a[i], sink)
            if i < len(gates_b):
                qc.apply_op(api, qb, gates_b[i], sink)
            if i == 40:
                api.calcTotalProb(qa)
        assert np.array_equal(api.get_amps(qa), ra)
        assert np.array_equal(api.get_amps(qb), rb)
    finally:
        api.destroyQureg(qa)
        api.destroyQureg(qb)


def test_destroying_a_state_with_queued_gates(api):
    """Queued gates of a destroyed qureg are discarded, not applied to
    memory that a later qureg reuses."""
    say("[gpu] fusion destroy with pending gates")
    api.set_fusion(True, 12, 4, 4096)
    q = api.createQureg(16)
    api.initPlusState(q)
    for t in range(16):
        api.applyHadamard(q, t)
    api.destroyQureg(q)
    q2 = api.createQureg(16)
    try:
        api.initZeroState(q2)
        amps = api.get_amps(q2)
        assert amps[0] == 1 and np.count_nonzero(amps) == 1
    finally:
        api.destroyQureg(q2)


def test_fusion_at_26_qubits(api):
    """Full-size property: the layered benchmark family (one-qubit Haar
    gates and two-qubit phases) at 26 qubits, fused (exact: bit-identical;
    reordered and merged: within rounding) against unfused."""
    say("[gpu] fusion 26q layered")
    c = qc.layered(26, 3, 71, ctrl3=True, probs=False, num_expec=1, init_seed=9)
    c.init = ("plus",)
    ref, rres = run_amps(api, c, None)
    for exact in (True, False):
        got, gres = run_amps(api, c, (12, 4), exact=exact)
        check_fused(got, gres, ref, rres, exact)
        assert abs(np.vdot(got, got).real - 1) < 1e-10


def test_benchmark_family_at_28_qubits(api):
    """Full-size property for the default planner: the headline workload's
    family (layers of Haar gates, two-qubit phases and 3-control gates, the
    full QFT, 64 Pauli gadgets on 4-12 qubits, probabilities and Pauli
    expectations) at 28 qubits with the benchmark's tile shape, default
    (reordered and merged) against exact fusion (bit-identical to the
    single-gate kernels): amplitudes within 1e-12 relative L2, results within
    1e-10, the norm kept."""
    say("[gpu] fusion 28q benchmark family")
    c = qc.layered(28, 6, 2024, ctrl3=True, qft=True, gadgets=16, gadget_range=(4, 12), num_expec=4,
                   probs=True, init_seed=5)
    c.init = ("plus",)
    ref, rres = run_amps(api, c, (13, 3, 2048), exact=True)
    got, gres = run_amps(api, c, (13, 3, 2048), exact=False)
    check_fused(got, gres, ref, rres, exact=False)
    assert abs(np.vdot(got, got).real - 1) < 1e-10


def test_fusion_config_validation(api):
    from quest_amd.api import QuESTError
    say("[gpu] fusion config validation")
    for bad in [(1, 0, 8), (14, 4, 8), (12, 11, 8), (12, 4, 0), (12, 4, 5000)]:
        with pytest.raises(QuESTError):
            api.set_fusion(True, *bad)
    api.set_fusion(True, 10, 2, 64)
    assert api.get_fusion() == (1, 10, 2, 64)


@pytest.mark.parametrize("min_gates", [2, 4])
@pytest.mark.parametrize("dens", [False, True])
def test_fp32_mfma_blocks_track_oracle(api, min_gates, dens):
    """complex64 states with runs of gates on 5 qubits multiplied into
    32x32 blocks applied with FP32 MFMA: within 1e-5 (relative L2 of the
    amplitudes, absolute on results) of the FP64 oracle, for tiles of 2^10
    to 2^12 amplitudes."""
    from oracle import qoracle as O
    n = 7 if dens else 15
    say(f"[gpu] fusion fp32 mfma blocks n={n} dens={dens} min_gates={min_gates}")
    c = fusion_mix(n, 160, 81 + dens, dens=dens)
    want, wres = O.run(c)
    lib = api.gpu()
    for tile in (10, 12):
        api.set_fusion(True, tile, 4, 512)
        assert lib.qa_setFusionBlocks(min_gates) == 0
        try:
            q = api.createCustomPrecisionQureg(n, int(dens), 1)
            q, res = qc.run(c, q)
            got = api.get_amps(q)
            api.destroyQureg(q)
        finally:
            lib.qa_setFusionBlocks(0)
        assert rel_l2(got, want) <= 1e-5, tile
        compare_results(res, wres, tol=1e-5)


def test_fp32_mfma_blocks_layered_family(api):
    """The plain-method family (one-qubit Haar gates and two-qubit phases)
    at 16 qubits: the blocks path and the gate-by-gate complex64 path agree
    to 1e-5 and both track the complex128 oracle."""
    from oracle import qoracle as O
    say("[gpu] fusion fp32 mfma layered")
    c = qc.layered(16, 6, 91, probs=True, num_expec=2)
    want, wres = O.run(c)
    lib = api.gpu()
    api.set_fusion(True, 12, 4, 512)
    outs = []
    for mg in (0, 3):
        assert lib.qa_setFusionBlocks(mg) == 0
        try:
            q = api.createCustomPrecisionQureg(16, 0, 1)
            q, res = qc.run(c, q)
            outs.append(api.get_amps(q))
            api.destroyQureg(q)
        finally:
            lib.qa_setFusionBlocks(0)
        assert rel_l2(outs[-1], want) <= 1e-5
        compare_results(res, wres, tol=1e-5)
    assert rel_l2(outs[0], outs[1]) <= 1e-5


@pytest.mark.parametrize("shape", [(12, 4), (8, 2), (13, 5)])
@pytest.mark.parametrize("dens", [False, True])
def test_wide_pauli_gadgets_fused_in_lds(api, shape, dens):
    """Pauli tensors and gadgets on more X/Y qubits than a thread's register
    bits are applied to the tile in LDS: bit-identical to the single-gate
    kernel, also with controls and when they open or close a pass."""
    from oracle import qoracle as O
    n = 6 if dens else 14
    say(f"[gpu] fusion wide pauli n={n} dens={dens} shape={shape}")
    rng = qc.SplitMix64(97 + dens)
    c = qc.Circuit(n, dens, [], ("zero",) if dens else ("random", 5), "wide_pauli")
    c.ops.append(("PG", "XYXYXYX"[: min(7, n)], list(range(min(7, n))), rng.angle()))
    for i in range(40):
        k = 5 + rng.randint(min(6, n - 5))
        targs = rng.sample(n, k)
        chars = "".join("XY"[rng.randint(2)] for _ in range(k))
        if i % 3 == 0:
            c.ops.append(("PG", chars, targs, rng.angle()))
        elif i % 3 == 1:
            rest = [q for q in range(n) if q not in targs]
            ctrls = [rest[j] for j in rng.sample(len(rest), 1)] if rest else []
            c.ops.append(("CPG", ctrls, chars, targs, rng.angle()) if ctrls else ("PSTR", chars, targs))
        else:
            c.ops.append(("PSTR", chars, targs))
        t = rng.randint(n)
        c.ops.append(("U1", t, qc.haar_unitary(rng, 2)))
    c.ops.append(("PG", "YZXY"[: min(4, n)] + "X", list(range(min(4, n))) + [n - 1], rng.angle()))
    c.ops.append(("OUT_TOTALPROB",))
    c.ops.append(("OUT_EXPEC", qc.random_pauli_chars(rng, n), list(range(n))))
    ref, rres = run_amps(api, c, None)
    got, gres = run_amps(api, c, shape)
    assert np.array_equal(got, ref), f"max |diff| {np.max(np.abs(got - ref))}"
    compare_results(gres, rres, tol=0.0)
    want, wres = O.run(c)
    assert rel_l2(got, want) <= AMP_TOL
