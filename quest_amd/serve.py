"""ctypes mirror of include/quest_amd_serve.h (the batched serving engine)
and the variational request family of the serving workload.

The engine runs on its own native thread: `Engine.submit` returns at once,
`Engine.wait` blocks (outside the GIL) until a request's expectation values
are back.
"""

from __future__ import annotations

import ctypes as C

import numpy as np

from . import api as _api
from . import circuits as _qc

SG = {"U1": 1, "H": 2, "RX": 3, "RY": 4, "RZ": 5, "CZ": 6, "CX": 7, "PS": 8}


class ServeGate(C.Structure):
    _fields_ = [("kind", C.c_int), ("t1", C.c_int), ("t2", C.c_int), ("angle", C.c_double), ("m", C.c_double * 8)]


class ServeCircuit(C.Structure):
    _fields_ = [("numQubits", C.c_int), ("numGates", C.c_int), ("gates", C.POINTER(ServeGate)),
                ("numPaulis", C.c_int), ("pauliX", C.POINTER(C.c_uint64)), ("pauliY", C.POINTER(C.c_uint64)),
                ("pauliZ", C.POINTER(C.c_uint64))]


def pauli_masks(chars: str, targets) -> tuple[int, int, int]:
    """(X, Y, Z) bit masks of a Pauli string given as one character per
    target (the layout of quest.h's getPauliStr(chars, targets))."""
    x = y = z = 0
    for ch, t in zip(chars, targets):
        bit = 1 << int(t)
        ch = ch.upper()
        if ch == "X":
            x |= bit
        elif ch == "Y":
            y |= bit
        elif ch == "Z":
            z |= bit
        elif ch != "I":
            raise ValueError(f"not a Pauli: {ch!r}")
    return x, y, z


def pack_gates(ops) -> "C.Array":
    """Circuit ops (the tuples of quest_amd.circuits) as ServeGate records;
    OUT_* ops are skipped."""
    gates = [op for op in ops if not op[0].startswith("OUT_")]
    arr = (ServeGate * max(1, len(gates)))()
    for g, op in zip(arr, gates):
        k = op[0]
        if k not in SG:
            raise ValueError(f"the serving engine does not take {k}")
        g.kind = SG[k]
        if k in ("CZ", "CX"):
            g.t1, g.t2 = int(op[1]), int(op[2])
        else:
            g.t1 = int(op[1])
            if k == "U1":
                m = np.asarray(op[2], dtype=complex).reshape(2, 2)
                for i, v in enumerate(m.ravel()):
                    g.m[2 * i], g.m[2 * i + 1] = v.real, v.imag
            elif k in ("RX", "RY", "RZ", "PS"):
                g.angle = float(op[2])
    return arr, len(gates)


class Engine:
    def __init__(self, arena_bytes: int, max_active: int = 256, tile_qubits: int = 12, low_qubits: int = 4,
                 precision: int = 1):
        _api.ensure_env()
        self._lib = _lib()
        h = C.c_void_p()
        rc = self._lib.qa_serveCreate(C.byref(h), C.c_size_t(int(arena_bytes)), int(max_active), int(tile_qubits),
                                      int(low_qubits), int(precision))
        _check(rc, "qa_serveCreate")
        self._h = h
        self._npaulis = {}

    def submit(self, num_qubits: int, ops, paulis) -> int:
        """paulis: list of (chars, targets)."""
        gates, ng = pack_gates(ops)
        masks = [pauli_masks(c, t) for c, t in paulis]
        npl = max(1, len(masks))
        xs = (C.c_uint64 * npl)(*[m[0] for m in masks])
        ys = (C.c_uint64 * npl)(*[m[1] for m in masks])
        zs = (C.c_uint64 * npl)(*[m[2] for m in masks])
        circ = ServeCircuit(int(num_qubits), ng, gates, len(masks), xs, ys, zs)
        rid = C.c_int64()
        _check(self._lib.qa_serveSubmit(self._h, C.byref(circ), C.byref(rid)), "qa_serveSubmit")
        self._npaulis[rid.value] = len(masks)
        return rid.value

    def poll(self, rid: int) -> bool:
        done = C.c_int()
        _check(self._lib.qa_servePoll(self._h, C.c_int64(rid), C.byref(done)), "qa_servePoll")
        return bool(done.value)

    def wait(self, rid: int) -> tuple[np.ndarray, float]:
        n = self._npaulis.pop(rid)
        out = np.zeros(max(1, n))
        lat = C.c_double()
        _check(self._lib.qa_serveWait(self._h, C.c_int64(rid), out.ctypes.data_as(C.POINTER(C.c_double)),
                                      C.byref(lat)), "qa_serveWait")
        return out[:n], lat.value

    def stats(self) -> dict:
        v = [C.c_int64() for _ in range(5)]
        _check(self._lib.qa_serveStats(self._h, *[C.byref(x) for x in v]), "qa_serveStats")
        out = dict(zip(["steps", "passes", "admitted", "completed", "max_running"], [x.value for x in v]))
        pm, wm = C.c_double(), C.c_double()
        _check(self._lib.qa_serveTimes(self._h, C.byref(pm), C.byref(wm)), "qa_serveTimes")
        out.update(plan_ms=pm.value, wait_ms=wm.value)
        return out

    def set_graphs(self, enabled: bool) -> None:
        """Launch each engine step as one captured HIP graph (True) or launch
        by launch (False, the default)."""
        _check(self._lib.qa_serveSetGraphs(self._h, int(bool(enabled))), "qa_serveSetGraphs")

    def graph_stats(self) -> dict:
        v = [C.c_int64() for _ in range(3)]
        _check(self._lib.qa_serveGraphStats(self._h, *[C.byref(x) for x in v]), "qa_serveGraphStats")
        return dict(zip(["launches", "instantiations", "updates"], [x.value for x in v]))

    def close(self) -> None:
        if self._h:
            _check(self._lib.qa_serveDestroy(self._h), "qa_serveDestroy")
            self._h = None

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()


class ServeError(RuntimeError):
    pass


def _check(rc: int, what: str) -> None:
    if rc != 0:
        raise ServeError(f"{what} failed ({rc}): {_lib().qa_lastError().decode()}")


_LIB = None


def _lib():
    global _LIB
    if _LIB is None:
        lib = _api.gpu()
        lib.qa_serveCreate.argtypes = [C.POINTER(C.c_void_p), C.c_size_t, C.c_int, C.c_int, C.c_int, C.c_int]
        lib.qa_serveDestroy.argtypes = [C.c_void_p]
        lib.qa_serveSubmit.argtypes = [C.c_void_p, C.POINTER(ServeCircuit), C.POINTER(C.c_int64)]
        lib.qa_servePoll.argtypes = [C.c_void_p, C.c_int64, C.POINTER(C.c_int)]
        lib.qa_serveWait.argtypes = [C.c_void_p, C.c_int64, C.POINTER(C.c_double), C.POINTER(C.c_double)]
        lib.qa_serveStats.argtypes = [C.c_void_p] + [C.POINTER(C.c_int64)] * 5
        lib.qa_serveTimes.argtypes = [C.c_void_p, C.POINTER(C.c_double), C.POINTER(C.c_double)]
        lib.qa_serveSetGraphs.argtypes = [C.c_void_p, C.c_int]
        lib.qa_serveGraphStats.argtypes = [C.c_void_p] + [C.POINTER(C.c_int64)] * 3
        lib.qa_lastError.restype = C.c_char_p
        _LIB = lib
    return _LIB


def serving_requests(count: int, seed: int, qubits=(10, 24), depth=(4, 32), num_paulis: int = 8):
    """The serving workload's requests: variational circuits with qubit count
    and depth drawn uniformly from the given inclusive ranges, each layer
    RY and RZ (angles U(0, 2pi)) on every qubit then a CZ ladder, and
    `num_paulis` random Pauli strings over all qubits.  Returns a list of
    quest_amd.circuits.Circuit (the strings are its OUT_EXPEC ops)."""
    rng = _qc.SplitMix64(seed)
    out = []
    for i in range(count):
        n = qubits[0] + rng.randint(qubits[1] - qubits[0] + 1)
        d = depth[0] + rng.randint(depth[1] - depth[0] + 1)
        c = _qc.variational(n, d, (seed << 20) ^ i, num_expec=num_paulis)
        c.name = f"serve_{seed}_{i}_n{n}_d{d}"
        out.append(c)
    return out


def paulis_of(c) -> list:
    return [(op[1], op[2]) for op in c.ops if op[0] == "OUT_EXPEC"]
