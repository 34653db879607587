"""TEST INFRASTRUCTURE ONLY: the CPU oracle.

ctypes front end of oracle/build/liboracle.so (the C restatement of QuEST's
cpu_subroutines.cpp, oracle/qoracle.c) plus the API-level semantics of
QuEST's operations.cpp / decoherence.cpp / calculations.cpp restated on top
of it (the density-matrix second pass, gadget phases, QFT decomposition...).
Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg import
this module; the product never does.

run(circuit) replays a quest_amd.circuits.Circuit on a numpy complex128
array and returns (amps, results) in the same form as circuits.run.
"""

from __future__ import annotations

import ctypes as C
import math
import subprocess
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
LIB = HERE / "build" / "liboracle.so"
REF_DRIVER = HERE / "_ref" / "circuit_driver"
# the reference's front end on this backend's kernels (oracle/Makefile ref-amd)
REF_AMD_DRIVER = HERE / "_ref" / "circuit_driver_amd"


def build() -> None:
    subprocess.run(["make", "-s", "-C", str(HERE), "all"], check=True)


def _load():
    if not LIB.exists():
        build()
    lib = C.CDLL(str(LIB))
    P = C.c_void_p
    I = C.c_int
    IP = C.POINTER(C.c_int)
    D = C.c_double
    U = C.c_uint64
    sigs = {
        "qo_dense1": [P, I, IP, IP, I, I, P],
        "qo_dense2": [P, I, IP, IP, I, I, I, P],
        "qo_denseN": [P, I, IP, IP, I, IP, I, P, I],
        "qo_diag1": [P, I, IP, IP, I, I, P],
        "qo_diag2": [P, I, IP, IP, I, I, I, P],
        "qo_diagN": [P, I, IP, IP, I, IP, I, P, I, I, D, D],
        "qo_swap": [P, I, IP, IP, I, I, I],
        "qo_pauli": [P, I, IP, IP, I, U, U, U, D, D, D, D],
        "qo_phaseGadget": [P, I, IP, IP, I, U, D, D, D, D],
        "qo_projector": [P, I, IP, IP, I, D],
        "qo_dm_dephase1": [P, I, I, D], "qo_dm_dephase2": [P, I, I, I, D],
        "qo_dm_depol1": [P, I, I, D], "qo_dm_depol2": [P, I, I, I, D],
        "qo_dm_pauliChannel": [P, I, I, D, D, D, D], "qo_dm_damp": [P, I, I, D],
        "qo_dm_projector": [P, I, IP, IP, I, D],
        "qo_probsAllOutcomes": [P, P, I, IP, I],
        "qo_expecPauli": [P, P, I, U, U, U], "qo_dm_expecPauli": [P, P, I, U, U, U],
        "qo_innerProduct": [P, P, P, I], "qo_initDebug": [P, I],
        "qo_dm_probsAllOutcomes": [P, P, I, IP, I], "qo_dm_fidelityPure": [P, P, P, I, I],
    }
    for name, args in sigs.items():
        f = getattr(lib, name)
        f.argtypes = args
        f.restype = None
    for name, args in {"qo_totalProb": [P, I], "qo_dm_totalProb": [P, I],
                       "qo_probOutcome": [P, I, IP, IP, I], "qo_dm_probOutcome": [P, I, IP, IP, I],
                       "qo_expecZ": [P, I, U], "qo_dm_hsDistanceSq": [P, P, I]}.items():
        f = getattr(lib, name)
        f.argtypes = args
        f.restype = C.c_double
    return lib


_lib = None


def lib():
    global _lib
    if _lib is None:
        _lib = _load()
    return _lib


def _ip(v):
    v = [int(x) for x in v]
    return (C.c_int * max(1, len(v)))(*v)


def _p(a: np.ndarray):
    return C.c_void_p(a.ctypes.data)


def _cmat(m) -> np.ndarray:
    return np.ascontiguousarray(np.asarray(m, dtype=np.complex128))


PAULI = {"I": 0, "X": 1, "Y": 2, "Z": 3}


def masks(chars: str, targets) -> tuple[int, int, int]:
    mx = my = mz = 0
    for ch, t in zip(chars, targets):
        p = PAULI[ch.upper()]
        if p == 1:
            mx |= 1 << t
        elif p == 2:
            my |= 1 << t
        elif p == 3:
            mz |= 1 << t
    return mx, my, mz


class State:
    """A statevector (dens=False) or density matrix on numpy, mutated through
    the C oracle with the API semantics of QuEST's operations.cpp."""

    def __init__(self, num_qubits: int, dens: bool = False):
        self.n = num_qubits
        self.dens = dens
        self.bits = 2 * num_qubits if dens else num_qubits
        self.a = np.zeros(1 << self.bits, dtype=np.complex128)
        self.a[0] = 1

    # ---- statevector-level passes on index bits
    def _dense1(self, c, s, t, m, conj=False):
        m = _cmat(m).reshape(4)
        if conj:
            m = m.conj().copy()
        lib().qo_dense1(_p(self.a), self.bits, _ip(c), _ip(s) if s else None, len(c), t, _p(m))

    def _dense2(self, c, s, t1, t2, m, conj=False):
        m = _cmat(m).reshape(16)
        if conj:
            m = m.conj().copy()
        lib().qo_dense2(_p(self.a), self.bits, _ip(c), _ip(s) if s else None, len(c), t1, t2, _p(m))

    def _denseN(self, c, s, t, m, conj=False):
        m = _cmat(m).reshape(-1)
        lib().qo_denseN(_p(self.a), self.bits, _ip(c), _ip(s) if s else None, len(c), _ip(t), len(t), _p(m),
                        int(conj))

    def _diag1(self, c, s, t, d, conj=False):
        d = _cmat(d).reshape(2)
        if conj:
            d = d.conj().copy()
        lib().qo_diag1(_p(self.a), self.bits, _ip(c), _ip(s) if s else None, len(c), t, _p(d))

    def _pauli(self, c, s, mx, my, mz, amp_fac, pair_fac):
        lib().qo_pauli(_p(self.a), self.bits, _ip(c), _ip(s) if s else None, len(c), mx, my, mz,
                       amp_fac.real, amp_fac.imag, pair_fac.real, pair_fac.imag)

    def _phase(self, c, s, mask, f0, f1):
        lib().qo_phaseGadget(_p(self.a), self.bits, _ip(c), _ip(s) if s else None, len(c), mask,
                             f0.real, f0.imag, f1.real, f1.imag)

    def _bra(self, qs):
        return [q + self.n for q in qs]

    # ---- API-level operations (operations.cpp)
    def unitary(self, ctrls, targs, m, kind="dense", states=None):
        """operations.cpp:18-37: U on kets, conj(U) on bras for a density matrix."""
        def apply(c, t, conj):
            if kind == "diag":
                self._diag1(c, states, t[0], m, conj)
            elif len(t) == 1:
                self._dense1(c, states, t[0], m, conj)
            elif len(t) == 2:
                self._dense2(c, states, t[0], t[1], m, conj)
            else:
                self._denseN(c, states, t, m, conj)
        apply(list(ctrls), list(targs), False)
        if self.dens:
            apply(self._bra(ctrls), self._bra(targs), True)

    def diag_n(self, ctrls, targs, d, power=None, states=None, multiply=False):
        """applyDiagMatr / applyDiagMatrPower (operations.cpp:468-500):
        cpu_statevec_anyCtrlAnyTargDiagMatr_sub with the optional complex
        power, conj (after the power) on the bra pass; multiply = ket only"""
        d = _cmat(d).reshape(-1)
        hp = 0 if power is None else 1
        e = complex(0 if power is None else power)
        s = list(states) if states is not None else None
        lib().qo_diagN(_p(self.a), self.bits, _ip(ctrls), _ip(s) if s else None, len(ctrls), _ip(targs), len(targs),
                       _p(d), 0, hp, e.real, e.imag)
        if self.dens and not multiply:
            lib().qo_diagN(_p(self.a), self.bits, _ip(self._bra(ctrls)), _ip(s) if s else None, len(ctrls),
                           _ip(self._bra(targs)), len(targs), _p(d), 1, hp, e.real, e.imag)

    def swap(self, ctrls, a, b):
        if a > b:
            a, b = b, a
        lib().qo_swap(_p(self.a), self.bits, _ip(ctrls), None, len(ctrls), a, b)
        if self.dens:
            lib().qo_swap(_p(self.a), self.bits, _ip(self._bra(ctrls)), None, len(ctrls), a + self.n, b + self.n)

    def _z_or_phase(self, c, mask, is_gadget, phase):
        f0 = complex(math.cos(phase), math.sin(phase)) if is_gadget else 1
        f1 = complex(math.cos(phase), -math.sin(phase)) if is_gadget else -1
        self._phase(c, None, mask, complex(f0), complex(f1))

    def _pauli_tensor(self, c, mx, my, mz, factor):
        """localiser.cpp:596-607"""
        if mx | my:
            self._pauli(c, None, mx, my, mz, 0 * factor, 1 * factor)
        else:
            assert factor == 1
            self._z_or_phase(c, mz, False, 0)

    def _pauli_gadget(self, c, mx, my, mz, phase):
        """localiser.cpp:613-620"""
        if not (mx | my):
            self._z_or_phase(c, mz, True, phase)
            return
        self._pauli(c, None, mx, my, mz, complex(math.cos(phase), 0), complex(0, math.sin(phase)))

    def pauli_str(self, ctrls, chars, targs):
        """operations.cpp:619-640"""
        mx, my, mz = masks(chars, targs)
        odd_y = bin(my).count("1") % 2 == 1
        factor = 1
        c = list(ctrls)
        if self.dens and not c:
            factor = -1 if odd_y else 1
            mx, my, mz = mx | (mx << self.n), my | (my << self.n), mz | (mz << self.n)
        self._pauli_tensor(c, mx, my, mz, factor)
        if self.dens and c:
            factor = -1 if odd_y else 1
            self._pauli_tensor(self._bra(c), mx << self.n, my << self.n, mz << self.n, factor)

    def pauli_gadget(self, ctrls, chars, targs, angle):
        """operations.cpp:854-870"""
        mx, my, mz = masks(chars, targs)
        if not (mx | my | mz) and not ctrls and self.dens:
            return
        phase = -angle / 2
        self._pauli_gadget(list(ctrls), mx, my, mz, phase)
        if not self.dens:
            return
        phase *= 1 if bin(my).count("1") % 2 == 1 else -1
        self._pauli_gadget(self._bra(ctrls), mx << self.n, my << self.n, mz << self.n, phase)

    # ---- Pauli string sums; terms are (chars, targets, complex coefficient)
    def multiply_pauli_sum(self, terms):
        """operations.cpp:1112-1131: a workspace copy, each string applied to
        it, added in with its coefficient and undone; strings act on the
        index bits (left multiplication of a density matrix)."""
        ws = State(self.n, self.dens)
        ws.a = self.a.copy()
        self.a = np.zeros_like(self.a)
        for chars, targs, coeff in terms:
            mx, my, mz = masks(chars, targs)
            ws._pauli_tensor([], mx, my, mz, 1)
            self.a = self.a + complex(coeff) * ws.a
            ws._pauli_tensor([], mx, my, mz, 1)

    def _first_order_trotter(self, terms, angle, reverse):
        """operations.cpp:1133-1142"""
        order = range(len(terms) - 1, -1, -1) if reverse else range(len(terms))
        for j in order:
            chars, targs, coeff = terms[j]
            self.pauli_gadget([], chars, targs, 2 * angle * complex(coeff).real)

    def trotter(self, terms, angle, order, reps):
        """operations.cpp:1144-1197.  Orders above 2 follow the reference's
        code (operations.cpp:1156-1165): five first-order sweeps in reversed
        term order with angles p, p, 1-4p, p, p (times angle), not the
        recursive formula its documentation gives."""
        if angle == 0:
            return
        step = angle / reps
        for _ in range(reps):
            if order == 1:
                self._first_order_trotter(terms, step, False)
            elif order == 2:
                self._first_order_trotter(terms, step / 2, False)
                self._first_order_trotter(terms, step / 2, True)
            else:
                p = 1.0 / (4 - 4 ** (1.0 / (order - 1)))
                a, b = p * step, (1 - 4 * p) * step
                for x in (a, a, b, a, a):
                    self._first_order_trotter(terms, x, True)

    def phase_gadget(self, ctrls, targs, angle):
        phase = -angle / 2
        mask = sum(1 << t for t in targs)
        self._z_or_phase(list(ctrls), mask, True, phase)
        if self.dens:
            self._z_or_phase(self._bra(ctrls), mask << self.n, True, -phase)

    def phase_shift(self, targs, angle):
        """operations.cpp:946-952: (k-1)-controlled diag(1, e^{i angle}) on targs[0]"""
        self.unitary(targs[1:], [targs[0]], [1, complex(math.cos(angle), math.sin(angle))], kind="diag")

    def qft(self, targs):
        """operations.cpp:1128-1141"""
        h = 1 / math.sqrt(2)
        nt = len(targs)
        for n in range(nt - 1, -1, -1):
            self.unitary([], [targs[n]], [[h, h], [h, -h]])
            for m in range(n):
                self.phase_shift([targs[n], targs[n - m - 1]], math.pi / (1 << (m + 1)))
        for n in range(nt // 2):
            self.swap([], targs[n], targs[nt - 1 - n])

    def kraus(self, targs, ops):
        """channels.cpp superop (utilities.cpp:771-801) on ket++bra targets"""
        d = 1 << len(targs)
        S = np.zeros((d * d, d * d), dtype=np.complex128)
        for k in ops:
            k = np.asarray(k, dtype=np.complex128)
            S += np.kron(k.conj(), k)
        allt = list(targs) + self._bra(targs)
        if len(allt) == 2:
            self._dense2([], None, allt[0], allt[1], S)
        else:
            self._denseN([], None, allt, S)

    # ---- reductions
    def prob(self, qubit, outcome):
        if self.dens:
            return lib().qo_dm_probOutcome(_p(self.a), self.n, _ip([qubit]), _ip([outcome]), 1)
        return lib().qo_probOutcome(_p(self.a), self.n, _ip([qubit]), _ip([outcome]), 1)

    def expec(self, chars, targs):
        mx, my, mz = masks(chars, targs)
        out = np.zeros(2)
        if self.dens:
            if not (mx | my | mz):
                return lib().qo_dm_totalProb(_p(self.a), self.n)
            lib().qo_dm_expecPauli(_p(out), _p(self.a), self.n, mx, my, mz)
            return out[0]
        if not (mx | my | mz):
            return lib().qo_totalProb(_p(self.a), self.n)
        if not (mx | my):
            return lib().qo_expecZ(_p(self.a), self.n, mz)
        lib().qo_expecPauli(_p(out), _p(self.a), self.n, mx, my, mz)
        return out[0]

    def purity(self):
        p = lib().qo_totalProb(_p(self.a), self.bits)
        return p if self.dens else p * p

    # calculations.cpp:49-61: <A|B>, or Tr(A^dagger B) of two density
    # matrices (their flat inner product); with one density matrix rho and
    # one statevector psi, <psi|rho^dagger|psi> (rho first) or <psi|rho|psi>
    def _pure_fid(self, rho, psi, conjugate):
        out = np.zeros(2)
        lib().qo_dm_fidelityPure(_p(out), _p(rho), _p(psi), self.n, int(conjugate))
        return complex(out[0], out[1])

    def inner(self, other, other_dens):
        if self.dens == other_dens:
            out = np.zeros(2)
            lib().qo_innerProduct(_p(out), _p(self.a), _p(other), self.bits)
            return complex(out[0], out[1])
        if self.dens:
            return self._pure_fid(self.a, other, True)
        return self._pure_fid(other, self.a, False)

    # calculations.cpp:303-328: |<A|B>|^2, or Re<psi|rho|psi>
    def fidelity(self, other, other_dens):
        if self.dens and other_dens:
            raise ValueError("fidelity of two density matrices is not defined by the reference")
        if not self.dens and not other_dens:
            return abs(self.inner(other, other_dens)) ** 2
        rho, psi = (self.a, other) if self.dens else (other, self.a)
        return self._pure_fid(rho, psi, False).real

    # calculations.cpp:331-364: Hilbert-Schmidt, Bures or purified distance
    def distance(self, other, other_dens):
        if self.dens and other_dens:
            return math.sqrt(lib().qo_dm_hsDistanceSq(_p(self.a), _p(other), self.n))
        if not self.dens and not other_dens:
            mag = min(abs(self.inner(other, other_dens)), 1.0)
            return math.sqrt(2 - 2 * mag)
        rho, psi = (self.a, other) if self.dens else (other, self.a)
        re = min(self._pure_fid(rho, psi, False).real, 1.0)
        return math.sqrt(1 - re)

    def total_prob(self):
        if self.dens:
            return lib().qo_dm_totalProb(_p(self.a), self.n)
        return lib().qo_totalProb(_p(self.a), self.n)


def run(circuit, init_amps: np.ndarray | None = None):
    """Replay a quest_amd.circuits.Circuit on the oracle."""
    st = State(circuit.num_qubits, circuit.dens)
    kind = circuit.init[0]
    if kind == "debug":
        lib().qo_initDebug(_p(st.a), st.bits)
    elif kind == "plus":
        st.a[:] = (1.0 / (1 << circuit.num_qubits)) if circuit.dens else 1.0 / math.sqrt(1 << circuit.num_qubits)
    elif kind == "random":
        if init_amps is None:
            from quest_amd.circuits import initial_amps
            init_amps = initial_amps(circuit)
        st.a[:] = init_amps
    results = []
    for op in circuit.ops:
        apply_op(st, op, results)
    return st.a, results


def apply_op(st: State, op, results: list) -> None:
    k = op[0]
    h = 1 / math.sqrt(2)
    if k == "U1":
        st.unitary([], [op[1]], op[2])
    elif k == "CU1":
        st.unitary(op[1], [op[2]], op[3])
    elif k == "U2":
        st.unitary([], [op[1], op[2]], op[3])
    elif k == "UN":
        st.unitary([], op[1], op[2])
    elif k == "D1":
        st.unitary([], [op[1]], op[2], kind="diag")
    elif k == "DN":
        st.diag_n([], op[1], op[2])
    elif k == "H":
        st.unitary([], [op[1]], [[h, h], [h, -h]])
    elif k == "X":
        st.pauli_str([], "X", [op[1]])
    elif k == "Y":
        st.pauli_str([], "Y", [op[1]])
    elif k == "Z":
        st.unitary([], [op[1]], [1, -1], kind="diag")
    elif k == "S":
        st.unitary([], [op[1]], [1, 1j], kind="diag")
    elif k == "T":
        st.unitary([], [op[1]], [1, complex(1 / math.sqrt(2), 1 / math.sqrt(2))], kind="diag")
    elif k == "CZ":
        st.unitary([op[1]], [op[2]], [1, -1], kind="diag")
    elif k == "CX":
        st.pauli_str([op[1]], "X", [op[2]])
    elif k == "PS":
        st.phase_shift([op[1]], op[2])
    elif k == "PS2":
        st.phase_shift([op[1], op[2]], op[3])
    elif k in ("RX", "RY", "RZ"):
        st.pauli_gadget([], k[1], [op[1]], op[2])
    elif k == "SWAP":
        st.swap([], op[1], op[2])
    elif k == "CSWAP":
        st.swap(op[1], op[2], op[3])
    elif k == "PSTR":
        st.pauli_str([], op[1], op[2])
    elif k == "CPSTR":
        st.pauli_str(op[1], op[2], op[3])
    elif k == "PG":
        st.pauli_gadget([], op[1], op[2], op[3])
    elif k == "CPG":
        st.pauli_gadget(op[1], op[2], op[3], op[4])
    elif k == "PHG":
        st.phase_gadget([], op[1], op[2])
    elif k == "QFT":
        st.qft(op[1])
    elif k == "FQFT":
        st.qft(list(range(st.n)))
    elif k == "DEPH":
        if op[2]:
            lib().qo_dm_dephase1(_p(st.a), st.n, op[1], op[2])
    elif k == "DEPH2":
        if op[3]:
            lib().qo_dm_dephase2(_p(st.a), st.n, op[1], op[2], op[3])
    elif k == "DEPOL":
        if op[2]:
            lib().qo_dm_depol1(_p(st.a), st.n, op[1], op[2])
    elif k == "DEPOL2":
        if op[3]:
            lib().qo_dm_depol2(_p(st.a), st.n, op[1], op[2], op[3])
    elif k == "DAMP":
        if op[2]:
            lib().qo_dm_damp(_p(st.a), st.n, op[1], op[2])
    elif k == "PAULIS":
        px, py, pz = op[2], op[3], op[4]
        if px or py or pz:
            lib().qo_dm_pauliChannel(_p(st.a), st.n, op[1], 1 - px - py - pz, px, py, pz)
    elif k == "KRAUS":
        st.kraus(op[1], op[2])
    elif k == "PROJ":
        if st.dens:
            lib().qo_dm_projector(_p(st.a), st.n, _ip([op[1]]), _ip([op[2]]), 1, 1.0)
        else:
            lib().qo_projector(_p(st.a), st.n, _ip([op[1]]), _ip([op[2]]), 1, 1.0)
    elif k == "OUT_PROB":
        results.append(("prob", st.prob(op[1], op[2])))
    elif k == "OUT_EXPEC":
        results.append(("expec", st.expec(op[1], op[2])))
    elif k == "OUT_PURITY":
        results.append(("purity", st.purity()))
    elif k == "OUT_TOTALPROB":
        results.append(("totalprob", st.total_prob()))
    elif k == "OUT_PROBS":
        out = np.zeros(1 << len(op[1]))
        if st.dens:
            lib().qo_dm_probsAllOutcomes(_p(out), _p(st.a), st.n, _ip(op[1]), len(op[1]))
        else:
            lib().qo_probsAllOutcomes(_p(out), _p(st.a), st.n, _ip(op[1]), len(op[1]))
        results.append(("probs", list(out)))
    elif k in ("AUX", "AUXCOPY"):
        if k == "AUX":
            from quest_amd.circuits import random_state
            st.aux = (np.ascontiguousarray(random_state(op[1], op[3], op[2]), dtype=np.complex128), op[2])
        else:
            st.aux = (st.a.copy(), st.dens)
    elif k == "OUT_INNER":
        z = st.inner(*st.aux)
        results.append(("inner", [z.real, z.imag]))
    elif k == "OUT_FIDELITY":
        results.append(("fidelity", st.fidelity(*st.aux)))
    elif k == "OUT_DISTANCE":
        results.append(("distance", st.distance(*st.aux)))
    elif k in ("SEEDS", "MEAS", "MEASN", "RANDPURE", "RANDMIXED"):
        raise NotImplementedError("measurement sampling is checked against reference fixtures directly")
    else:
        raise ValueError(f"unknown op {k}")


def run_reference(circuit, workdir: Path, init_amps: np.ndarray | None = None, want_amps: bool = True,
                  driver: Path | None = None, env: dict | None = None):
    """Run the circuit through the reference's own CPU build (oracle/_ref),
    or through another build of the reference's driver (`driver`, with extra
    environment `env`).  Returns (amps or None, results, stdout)."""
    from quest_amd.circuits import to_text, parse_results, initial_amps
    workdir.mkdir(parents=True, exist_ok=True)
    init_path = None
    if circuit.init[0] == "random":
        amps = init_amps if init_amps is not None else initial_amps(circuit)
        init_path = workdir / f"{circuit.name}_init.bin"
        np.ascontiguousarray(amps, dtype=np.complex128).tofile(init_path)
    out_path = workdir / f"{circuit.name}_out.bin" if want_amps else None
    txt = workdir / f"{circuit.name}.txt"
    text = to_text(circuit, str(init_path) if init_path else None, str(out_path) if out_path else None)
    # AUX n dm seed -> AUXFILE n dm path: the driver reads the auxiliary
    # qureg's amplitudes (the project's seeded random state) from a file
    lines = []
    for i, line in enumerate(text.splitlines()):
        t = line.split()
        if t and t[0] == "AUX":
            from quest_amd.circuits import random_state
            n, dm, seed = int(t[1]), bool(int(t[2])), int(t[3])
            path = workdir / f"{circuit.name}_aux{i}.bin"
            np.ascontiguousarray(random_state(n, seed, dm), dtype=np.complex128).tofile(path)
            line = f"AUXFILE {n} {int(dm)} {path}"
        lines.append(line)
    txt.write_text("\n".join(lines) + "\n")
    import os
    full_env = dict(os.environ, **(env or {}))
    proc = subprocess.run([str(driver or REF_DRIVER), str(txt)], capture_output=True, text=True, env=full_env)
    if proc.returncode != 0:
        raise RuntimeError(f"{driver or REF_DRIVER} failed ({proc.returncode}):\n{proc.stdout[-2000:]}\n{proc.stderr[-2000:]}")
    amps_out = np.fromfile(out_path, dtype=np.complex128) if out_path else None
    return amps_out, parse_results(proc.stdout), proc.stdout + proc.stderr
