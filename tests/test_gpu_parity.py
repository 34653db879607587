"""Parity of the HIP path (through the quest.h C API of libquest_amd.so and
the kernel C ABI of libquest_amd_gpu.so) with QuEST's own CPU build (golden
fixtures) and with the CPU oracle on the same seeded inputs.

Tolerances (north star): amplitudes within 1e-12 relative L2, probabilities
and expectations within 1e-10, measurement outcomes bit-exact.  FP32
amplitude quregs: expectations within 1e-5 of the FP64 oracle.
"""

import sys

import numpy as np
import pytest

from conftest import GOLDEN_CASES, compare_results, load_golden, rel_l2
from quest_amd import circuits as qc

pytestmark = pytest.mark.gpu

AMP_TOL = 1e-12
SCALAR_TOL = 1e-10


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


@pytest.mark.parametrize("name", GOLDEN_CASES)
def test_golden_fixture_on_gpu(api, name):
    say(f"[gpu] golden {name}")
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


@pytest.mark.parametrize("n,seed", [(14, 1), (17, 2), (20, 3)])
def test_layered_family_vs_oracle(api, n, seed):
    from oracle import qoracle as O
    say(f"[gpu] layered n={n} seed={seed}")
    c = qc.layered(n, 6, seed, ctrl3=True, qft=True, gadgets=8, gadget_range=(4, min(12, n)), num_expec=4)
    want, wres = O.run(c)
    q, res = qc.run(c)
    try:
        assert rel_l2(api.get_amps(q), want) <= AMP_TOL
        compare_results(res, wres, tol=SCALAR_TOL)
    finally:
        api.destroyQureg(q)


@pytest.mark.parametrize("n,seed", [(6, 4), (8, 5)])
def test_density_family_vs_oracle(api, n, seed):
    from oracle import qoracle as O
    say(f"[gpu] density n={n} seed={seed}")
    c = qc.noisy_density(n, 10, seed)
    want, wres = O.run(c)
    q, res = qc.run(c)
    try:
        assert rel_l2(api.get_amps(q), want) <= AMP_TOL
        compare_results(res, wres, tol=SCALAR_TOL)
    finally:
        api.destroyQureg(q)


@pytest.mark.parametrize("nc", [0, 1, 2, 3, 5])
@pytest.mark.parametrize("dens", [False, True])
def test_controls_and_targets_everywhere(api, nc, dens):
    """Every target position including 0 and n-1 (low targets exercise the
    intra-wave pair layout, high ones the widest strides), with nc controls
    in random states."""
    from oracle import qoracle as O
    n = 6 if dens else 12
    say(f"[gpu] controls nc={nc} dens={dens}")
    rng = qc.SplitMix64(100 + nc)
    c = qc.Circuit(n, dens, [], ("random", 7 + nc), "ctrl")
    for t in range(n):
        others = [q for q in range(n) if q != t]
        ctrls = [others[i] for i in rng.sample(len(others), nc)]
        c.ops.append(("CU1", ctrls, t, qc.haar_unitary(rng, 2)))
        t2 = others[rng.randint(len(others))]
        if t2 not in ctrls:
            c.ops.append(("U2", t, t2, qc.haar_unitary(rng, 4)))
        c.ops.append(("CPG", ctrls, "X", [t], rng.angle()))
        c.ops.append(("PHG", [t, t2], rng.angle()))
        if len(ctrls) >= 1:
            c.ops.append(("CSWAP", ctrls, t, t2) if t2 not in ctrls else ("SWAP", t, t2))
    for t in range(0, n - 2, 2):
        c.ops.append(("UN", [t + 2, t, t + 1], qc.haar_unitary(rng, 8)))
    c.ops += [("OUT_PROB", q, 1) for q in range(n)] + [("OUT_EXPEC", "XYZ" * (n // 3), list(range(3 * (n // 3))))]
    want, wres = O.run(c)
    q, res = qc.run(c)
    try:
        assert rel_l2(api.get_amps(q), want) <= AMP_TOL
        compare_results(res, wres, tol=SCALAR_TOL)
    finally:
        api.destroyQureg(q)


@pytest.mark.parametrize("k", [3, 4, 5, 6, 7])
def test_many_target_dense(api, k):
    from oracle import qoracle as O
    say(f"[gpu] dense k={k}")
    rng = qc.SplitMix64(k)
    n = 12
    c = qc.Circuit(n, False, [], ("random", k), "dense")
    for rep in range(3):
        targs = rng.sample(n, k)
        c.ops.append(("UN", targs, qc.haar_unitary(rng, 1 << k)))
    want, _ = O.run(c)
    q, _ = qc.run(c)
    try:
        assert rel_l2(api.get_amps(q), want) <= AMP_TOL
    finally:
        api.destroyQureg(q)


def test_fp32_quregs_track_fp64_oracle(api):
    """FP32 amplitudes, FP64 reductions: expectations within 1e-5."""
    from oracle import qoracle as O
    say("[gpu] fp32 layered")
    c = qc.layered(16, 8, 77, ctrl3=True, gadgets=8, gadget_range=(2, 8), num_expec=8)
    want, wres = O.run(c)
    q = api.createCustomPrecisionQureg(16, 0, 1)
    try:
        assert api.getQuregPrecision(q) == 1
        q, res = qc.run(c, q)
        got = api.get_amps(q)
        # ~400 FP32 gates: infidelity grows like (gates x 2^-24)^2-ish; measured ~7e-8
        fid = abs(np.vdot(want, got)) ** 2
        assert 1 - fid < 1e-6
        compare_results(res, wres, tol=1e-5)
    finally:
        api.destroyQureg(q)


def test_fp32_density_matrix(api):
    from oracle import qoracle as O
    say("[gpu] fp32 density")
    c = qc.noisy_density(6, 6, 31)
    want, wres = O.run(c)
    q = api.createCustomPrecisionQureg(6, 1, 1)
    try:
        q, res = qc.run(c, q)
        compare_results(res, wres, tol=1e-5)
    finally:
        api.destroyQureg(q)


def test_size_independent_properties_at_26_qubits(api):
    """At 26 qubits (1 GiB), where the oracle is slow: unitarity keeps the norm,
    a circuit followed by its inverse returns the initial state, and the
    single-qubit probabilities of both outcomes sum to one."""
    say("[gpu] properties n=26")
    n = 26
    rng = qc.SplitMix64(26)
    q = api.createQureg(n)
    try:
        api.initRandomPureState(q)
        ref = api.createCloneQureg(q)
        ops = []
        for layer in range(3):
            for t in range(n):
                ops.append(("U1", t, qc.haar_unitary(rng, 2)))
            perm = rng.permutation(n)
            for k in range(n // 2):
                ops.append(("PS2", perm[2 * k], perm[2 * k + 1], rng.angle()))
            ops.append(("PG", "XYZY", rng.sample(n, 4), rng.angle()))
        for op in ops:
            qc.apply_op(api, q, op, [])
        assert abs(api.calcTotalProb(q) - 1) < 1e-10
        for q_ in (0, 13, 25):
            assert abs(api.calcProbOfQubitOutcome(q, q_, 0) + api.calcProbOfQubitOutcome(q, q_, 1) - 1) < 1e-10
        # inverse: reverse order with adjoint matrices / negated angles
        for op in reversed(ops):
            if op[0] == "U1":
                qc.apply_op(api, q, ("U1", op[1], np.array(op[2]).conj().T.tolist()), [])
            elif op[0] == "PS2":
                qc.apply_op(api, q, ("PS2", op[1], op[2], -op[3]), [])
            else:
                qc.apply_op(api, q, ("PG", op[1], op[2], -op[3]), [])
        fid = abs(api.inner_product(ref, q)) ** 2
        assert abs(fid - 1) < 1e-10
        api.destroyQureg(ref)
    finally:
        api.destroyQureg(q)


def test_measurement_is_seeded(api):
    say("[gpu] measurement determinism")
    outs = []
    for rep in range(2):
        q = api.createQureg(8)
        api.initPlusState(q)
        api.set_seeds([1, 2, 3])
        outs.append([api.applyQubitMeasurement(q, t) for t in range(8)])
        assert abs(api.calcTotalProb(q) - 1) < 1e-12
        api.destroyQureg(q)
    assert outs[0] == outs[1]


def test_qubit_probabilities_cache_follows_the_state(api):
    """calcProbOfQubitOutcome on statevectors of >= 16 qubits computes every
    qubit's probabilities at once and keeps them while no entry point has
    written amplitudes: interleaved with gates, amplitude setters, a
    measurement and a second state, every value must match the single
    reduction (calcProbOfMultiQubitOutcome) and the oracle."""
    from oracle import qoracle as O
    n = 17
    say("[gpu] qubit probability cache")
    q = api.createQureg(n)
    r = api.createQureg(n)
    try:
        init = qc.random_state(n, 41)
        api.set_amps(q, init)
        api.set_amps(r, qc.random_state(n, 42))

        def check_all(state_amps):
            st = O.State(n)
            st.a[:] = state_amps
            for t in range(n):
                for o in (0, 1):
                    got = api.calcProbOfQubitOutcome(q, t, o)
                    single = api.calcProbOfMultiQubitOutcome(q, api.ints([t]), api.ints([o]), 1)
                    assert abs(got - single) <= 1e-14, (t, o)
                    assert abs(got - st.prob(t, o)) <= 1e-12, (t, o)

        check_all(init)
        api.applyHadamard(q, 3)
        api.applyCompMatr1(q, 16, api.compmatr1(qc.haar_unitary(qc.SplitMix64(9), 2)))
        check_all(api.get_amps(q))
        api.calcProbOfQubitOutcome(r, 0, 0)  # another state in between
        check_all(api.get_amps(q))
        new = qc.random_state(n, 43)
        api.set_amps(q, new)
        check_all(new)
        api.applyQubitMeasurement(q, 5)
        check_all(api.get_amps(q))
    finally:
        api.destroyQureg(r)
        api.destroyQureg(q)


@pytest.mark.parametrize("n,precision", [(10, 2), (11, 1), (16, 2), (20, 2), (23, 1), (26, 2)])
def test_every_qubit_probabilities_in_one_read(api, n, precision):
    """qa_statevec_calcProbsOfEveryQubit (one read of the state for up to 32
    qubits) against the marginals of the amplitudes, summed in FP64 with
    numpy: both outcomes of every qubit, within 1e-13 (complex128) and 1e-7
    (complex64, whose amplitudes are rounded)."""
    import ctypes as C
    say(f"[gpu] every-qubit probabilities n={n} precision={precision}")
    q = api.createCustomPrecisionQureg(n, 0, precision)
    try:
        api.set_amps(q, qc.random_state(n, 100 + n))
        amps = api.get_amps(q).astype(np.complex128)
        out = np.zeros(2 * n)
        rc = api.gpu().qa_statevec_calcProbsOfEveryQubit(C.c_void_p(q.gpuAmps), precision, n,
                                                          out.ctypes.data_as(C.c_void_p))
        assert rc == 0
        p = np.abs(amps) ** 2
        idx = np.arange(1 << n)
        tol = 1e-13 if precision == 2 else 1e-7
        for t in range(n):
            one = ((idx >> t) & 1).astype(bool)
            assert abs(out[2 * t] - p[~one].sum()) <= tol, t
            assert abs(out[2 * t + 1] - p[one].sum()) <= tol, t
    finally:
        api.destroyQureg(q)


def test_qureg_size_limits(api):
    """Sizes at and past the limits are input errors, never a crash: no
    qubits, more index bits than one GPU addresses, and a state larger than
    the free HBM (2^36 complex128 amplitudes = 1 TiB > 288 GB); after them a
    1-qubit and a 30-qubit state still work."""
    say("[gpu] qureg size limits")
    with pytest.raises(api.QuESTError, match="qubit"):
        api.createQureg(0)
    with pytest.raises(api.QuESTError, match="limit"):
        api.createQureg(41)
    with pytest.raises(api.QuESTError, match="limit"):
        api.createDensityQureg(21)
    with pytest.raises(api.QuESTError, match="GPU memory"):
        api.createQureg(36)
    q = api.createQureg(1)
    try:
        api.applyHadamard(q, 0)
        assert abs(api.calcProbOfQubitOutcome(q, 0, 1) - 0.5) < 1e-15
    finally:
        api.destroyQureg(q)
    q = api.createQureg(30)
    try:
        api.applyHadamard(q, 29)
        assert abs(api.calcProbOfQubitOutcome(q, 29, 0) - 0.5) < 1e-12
    finally:
        api.destroyQureg(q)


def test_input_errors_raise(api):
    say("[gpu] validation")
    q = api.createQureg(4)
    try:
        with pytest.raises(api.QuESTError, match="target"):
            api.applyHadamard(q, 9)
        with pytest.raises(api.QuESTError, match="unitary"):
            api.applyCompMatr1(q, 0, api.compmatr1([[1, 1], [0, 1]]))
        with pytest.raises(api.QuESTError):
            api.applyMultiControlledCompMatr1(q, api.ints([0]), 1, 0, api.compmatr1(np.eye(2)))
        with pytest.raises(api.QuESTError):
            api.mixDepolarising(q, 0, 0.1)  # statevector
    finally:
        api.destroyQureg(q)
