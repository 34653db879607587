"""The batched serving engine (include/quest_amd_serve.h): results against
the CPU oracle, the same results alone and in a mix of concurrent requests,
admission under a small HBM arena, and input errors.

Tolerances: complex128 engine within 1e-10 of the FP64 oracle (the engine
multiplies consecutive one-qubit gates on a qubit into one matrix, so its
rounding differs from the per-gate oracle); complex64 engine within 1e-5.
"""

import sys
import threading
import time

import numpy as np
import pytest

from quest_amd import circuits as qc
from quest_amd import serve as S

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def oracle_expecs(c):
    from oracle import qoracle as O
    _, res = O.run(c)
    return np.array([v for (lab, v) in res if lab == "expec"])


def run_alone(eng, c):
    rid = eng.submit(c.num_qubits, c.ops, S.paulis_of(c))
    return eng.wait(rid)[0]


@pytest.mark.parametrize("precision,tol", [(2, 1e-10), (1, 1e-5)])
def test_engine_matches_oracle(api, precision, tol):
    say(f"[gpu] serve vs oracle precision={precision}")
    reqs = S.serving_requests(10, 5, qubits=(10, 16), depth=(2, 8))
    with S.Engine(1 << 30, precision=precision) as eng:
        ids = [eng.submit(c.num_qubits, c.ops, S.paulis_of(c)) for c in reqs]
        for c, rid in zip(reqs, ids):
            got, lat = eng.wait(rid)
            want = oracle_expecs(c)
            assert np.max(np.abs(got - want)) <= tol, (c.name, got, want)
            assert lat >= 0


@pytest.mark.parametrize("tile", [6, 9, 12])
def test_multi_pass_states(api, tile):
    """States wider than the tile take several passes per request."""
    say(f"[gpu] serve multi-pass tile={tile}")
    reqs = [qc.variational(n, 3, 100 + n, num_expec=8) for n in (13, 17, 20)]
    with S.Engine(1 << 30, tile_qubits=tile, low_qubits=min(4, tile - 2), precision=2) as eng:
        for c in reqs:
            got = run_alone(eng, c)
            assert np.max(np.abs(got - oracle_expecs(c))) <= 1e-10, c.name


def test_same_alone_and_in_the_mix(api):
    """A request's results do not depend on what else is running: the same
    circuit alone, then submitted from another thread in the middle of a
    stream of others, gives bitwise identical expectations."""
    say("[gpu] serve alone vs mix")
    mix = S.serving_requests(48, 9, qubits=(10, 20), depth=(4, 16))
    probe = S.serving_requests(1, 77, qubits=(18, 18), depth=(12, 12))[0]
    with S.Engine(1 << 31, max_active=32) as eng:
        alone = run_alone(eng, probe)
        ids = {}

        def feeder():
            for i, c in enumerate(mix):
                ids[i] = eng.submit(c.num_qubits, c.ops, S.paulis_of(c))
                time.sleep(0.0005)

        th = threading.Thread(target=feeder)
        th.start()
        time.sleep(0.005)
        pid = eng.submit(probe.num_qubits, probe.ops, S.paulis_of(probe))
        th.join()
        in_mix, _ = eng.wait(pid)
        others = {i: eng.wait(rid)[0] for i, rid in ids.items()}
        st = eng.stats()
    assert np.array_equal(alone, in_mix)
    assert st["max_running"] > 1, st
    for i in (0, 17, 47):
        assert np.max(np.abs(others[i] - oracle_expecs(mix[i]))) <= 1e-5


def test_small_arena_admits_as_memory_frees(api):
    """An arena that holds two 16-qubit complex64 states (2 x 512 KiB) still
    serves ten such requests, never more than two at once."""
    say("[gpu] serve small arena")
    reqs = [qc.variational(16, 4, 300 + i, num_expec=4) for i in range(10)]
    with S.Engine(2 * 8 * (1 << 16), max_active=64) as eng:
        ids = [eng.submit(c.num_qubits, c.ops, S.paulis_of(c)) for c in reqs]
        outs = [eng.wait(r)[0] for r in ids]
        st = eng.stats()
    assert st["max_running"] <= 2 and st["completed"] == 10, st
    for c, o in zip(reqs, outs):
        assert np.max(np.abs(o - oracle_expecs(c))) <= 1e-5


def test_submit_errors(api):
    say("[gpu] serve errors")
    with S.Engine(1 << 24) as eng:
        with pytest.raises(S.ServeError):
            eng.submit(3, [("H", 0)], [])
        with pytest.raises(S.ServeError):
            eng.submit(30, [("H", 0)], [])  # 8 GiB state, 16 MiB arena
        with pytest.raises(S.ServeError):
            eng.submit(10, [("H", 12)], [])
        with pytest.raises(S.ServeError):
            eng.submit(10, [("CZ", 2, 2)], [])
        # still serving after the refusals
        got = run_alone(eng, qc.variational(10, 2, 5, num_expec=2))
        assert got.shape == (2,)


def test_empty_circuit_reads_the_zero_state(api):
    """A request with no gates still gets one pass that creates |0...0>, so
    it never reads what an earlier request left in its arena block: after a
    request has used the block, an empty circuit gives <Z...Z> = 1 and
    <X> = 0 exactly."""
    say("[gpu] serve empty circuit")
    n = 12
    busy = qc.variational(n, 6, 41, num_expec=2)
    paulis = [("Z" * n, list(range(n))), ("X", [3]), ("ZZ", [0, n - 1])]
    with S.Engine(16 << n, max_active=1, precision=2) as eng:  # room for exactly one state
        run_alone(eng, busy)
        rid = eng.submit(n, [], paulis)
        got, _ = eng.wait(rid)
    assert list(got) == [1.0, 0.0, 1.0], got


@pytest.mark.parametrize("precision", [2, 1])
def test_graph_steps_match_eager_steps(api, precision):
    """Steps launched as captured HIP graphs (each step slot's executable
    graph updated in place between steps, or instantiated anew when the
    launches change) give bitwise the results of steps launched one by one.
    The graph engine runs many steps (each a graph replay), with requests of
    1-3 tile sizes per step, and eager steps of the same engine run in
    between (graphs switched off and on again mid-stream)."""
    say(f"[gpu] serve graphs vs eager precision={precision}")
    reqs = S.serving_requests(40, 21, qubits=(8, 20), depth=(2, 12))
    with S.Engine(1 << 30, max_active=16, tile_qubits=12, precision=precision) as eng:
        want = [run_alone(eng, c) for c in reqs]
    with S.Engine(1 << 30, max_active=16, tile_qubits=12, precision=precision) as eng:
        eng.set_graphs(True)
        got = []
        for i, c in enumerate(reqs):
            if i == 15:
                eng.set_graphs(False)
            if i == 20:
                eng.set_graphs(True)
            got.append(run_alone(eng, c))
        # and a concurrent batch: steps with several requests and tile sizes
        ids = [eng.submit(c.num_qubits, c.ops, S.paulis_of(c)) for c in reqs]
        batch = [eng.wait(rid)[0] for rid in ids]
        gs = eng.graph_stats()
        st = eng.stats()
    say(f"  {st['steps']} steps, graphs: {gs}")
    assert gs["launches"] > 0 and gs["updates"] > 0, gs
    assert gs["instantiations"] + gs["updates"] == gs["launches"], gs
    for i, c in enumerate(reqs):
        assert np.array_equal(got[i], want[i]), (i, c.name)
        assert np.array_equal(batch[i], want[i]), (i, c.name)
