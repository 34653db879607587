"""Seeded circuit families, the circuit text format, and replay onto the API.

One project RNG (SplitMix64, Steele, Lea & Flood 2014) drives every random
choice, so a (family, size, seed) triple names one circuit on every machine:

  * layered(...)      -- the statevector family: per layer a Haar-random 2x2
                         unitary (applyCompMatr1) on every qubit, then
                         applyTwoQubitPhaseShift(U(0,2pi)) on n//2 random
                         disjoint pairs; optionally one 3-control
                         applyMultiControlledCompMatr1 per layer, a full QFT
                         and random Pauli gadgets; outputs single-qubit
                         probabilities and Pauli-string expectations.
  * noisy_density(...) -- the density-matrix family: Haar 1-qubit gates and
                         CZ pairs, mixDepolarising(U(0,0.01)) after every gate
                         on each gate qubit, a 2-qubit mixKrausMap (4
                         operators from a random 16x4 isometry) every 5
                         layers; outputs purity, expectations, probabilities.
  * variational(...)  -- RY(U(0,2pi)), RZ(U(0,2pi)) on all qubits then a CZ
                         ladder, per layer; 8 random Pauli expectations.

Haar matrices come from Gram-Schmidt on complex Gaussian matrices (Mezzadri,
"How to generate random matrices from the classical compact groups", 2007:
QR with a positive real R diagonal), computed in plain Python floats so that
they are bit-identical everywhere.

A circuit is a list of ops (tuples).  `to_text` writes the line format read
by the reference-side driver (one op per line, reals with 17 significant
digits); `run` replays the ops on a Qureg through quest_amd.api.
"""

from __future__ import annotations

import cmath
import math
from dataclasses import dataclass, field

import numpy as np

MASK64 = (1 << 64) - 1
GOLDEN = 0x9E3779B97F4A7C15


class SplitMix64:
    def __init__(self, seed: int):
        self.state = seed & MASK64

    def next_u64(self) -> int:
        self.state = (self.state + GOLDEN) & MASK64
        z = self.state
        z = ((z ^ (z >> 30)) * 0xBF58476D1CE4E5B9) & MASK64
        z = ((z ^ (z >> 27)) * 0x94D049BB133111EB) & MASK64
        return z ^ (z >> 31)

    def uniform(self) -> float:
        """[0, 1) with 53 random bits."""
        return (self.next_u64() >> 11) * (1.0 / 9007199254740992.0)

    def angle(self) -> float:
        return 2 * math.pi * self.uniform()

    def randint(self, n: int) -> int:
        return self.next_u64() % n

    def gauss_pair(self) -> tuple[float, float]:
        """Box-Muller."""
        u1 = 1.0 - self.uniform()
        u2 = self.uniform()
        r = math.sqrt(-2.0 * math.log(u1))
        return r * math.cos(2 * math.pi * u2), r * math.sin(2 * math.pi * u2)

    def cgauss(self) -> complex:
        a, b = self.gauss_pair()
        return complex(a, b) / math.sqrt(2)

    def permutation(self, n: int) -> list[int]:
        p = list(range(n))
        for i in range(n - 1, 0, -1):  # Fisher-Yates
            j = self.randint(i + 1)
            p[i], p[j] = p[j], p[i]
        return p

    def sample(self, n: int, k: int) -> list[int]:
        return self.permutation(n)[:k]


def _gram_schmidt(cols: list[list[complex]]) -> list[list[complex]]:
    out: list[list[complex]] = []
    for v in cols:
        w = list(v)
        for u in out:
            dot = sum(u[i].conjugate() * w[i] for i in range(len(w)))
            w = [w[i] - dot * u[i] for i in range(len(w))]
        nrm = math.sqrt(sum(abs(x) ** 2 for x in w))
        out.append([x / nrm for x in w])
    return out


def haar_unitary(rng: SplitMix64, dim: int) -> list[list[complex]]:
    """dim x dim Haar-random unitary, row-major nested lists."""
    cols = [[rng.cgauss() for _ in range(dim)] for _ in range(dim)]
    q = _gram_schmidt(cols)
    return [[q[c][r] for c in range(dim)] for r in range(dim)]


def random_isometry(rng: SplitMix64, rows: int, cols: int) -> list[list[complex]]:
    """rows x cols matrix with orthonormal columns (Haar on the Stiefel manifold)."""
    vecs = [[rng.cgauss() for _ in range(rows)] for _ in range(cols)]
    q = _gram_schmidt(vecs)
    return [[q[c][r] for c in range(cols)] for r in range(rows)]


def kraus_from_isometry(v: list[list[complex]], num_ops: int, dim: int) -> list[list[list[complex]]]:
    """K_i = rows [i*dim, (i+1)*dim) of a (num_ops*dim) x dim isometry: sum K^dag K = I."""
    return [[v[i * dim + r][:] for r in range(dim)] for i in range(num_ops)]


def random_pauli_chars(rng: SplitMix64, n: int, allow_identity: bool = True) -> str:
    alphabet = "IXYZ" if allow_identity else "XYZ"
    return "".join(alphabet[rng.randint(len(alphabet))] for _ in range(n))


# ------------------------------------------------------------ initial states

def random_state(num_qubits: int, seed: int, dens: bool = False) -> np.ndarray:
    """Normalised complex-Gaussian statevector keyed by (seed, index); for a
    density matrix, the pure state |psi><psi| flattened column-major."""
    n = 1 << num_qubits
    i = np.arange(n, dtype=np.uint64)
    with np.errstate(over="ignore"):
        base = (np.uint64(seed & MASK64) * np.uint64(GOLDEN)) + i * np.uint64(0xD1B54A32D192ED03)

        def mix(z):
            z = (z ^ (z >> np.uint64(30))) * np.uint64(0xBF58476D1CE4E5B9)
            z = (z ^ (z >> np.uint64(27))) * np.uint64(0x94D049BB133111EB)
            return z ^ (z >> np.uint64(31))

        x1 = mix(base ^ np.uint64(1))
        x2 = mix(base ^ np.uint64(2))
    u1 = ((x1 >> np.uint64(11)).astype(np.float64) + 1.0) * (1.0 / 9007199254740992.0)
    u2 = (x2 >> np.uint64(11)).astype(np.float64) * (1.0 / 9007199254740992.0)
    r = np.sqrt(-2.0 * np.log(u1))
    psi = r * np.cos(2 * np.pi * u2) + 1j * (r * np.sin(2 * np.pi * u2))
    psi /= np.sqrt(np.sum(np.abs(psi) ** 2))
    if not dens:
        return psi
    return np.outer(psi, psi.conj()).reshape(-1, order="F")


# ------------------------------------------------------------ circuits

@dataclass
class Circuit:
    num_qubits: int
    dens: bool = False
    ops: list = field(default_factory=list)
    init: tuple = ("zero",)        # ("zero",) | ("random", seed) | ("debug",) | ("amps", path)
    name: str = ""

    def gate_count(self) -> int:
        return sum(1 for op in self.ops if not op[0].startswith("OUT_") and op[0] not in ("SEEDS",))


def layered(n: int, depth: int, seed: int, *, ctrl3: bool = False, qft: bool = False, gadgets: int = 0,
            gadget_range=(4, 12), num_expec: int = 1, probs: bool = True, init_seed: int | None = None) -> Circuit:
    rng = SplitMix64(seed)
    c = Circuit(n, False, [], ("random", init_seed if init_seed is not None else seed ^ 0x5EED),
                f"layered_n{n}_d{depth}_s{seed}")
    for _ in range(depth):
        for q in range(n):
            c.ops.append(("U1", q, haar_unitary(rng, 2)))
        perm = rng.permutation(n)
        for k in range(n // 2):
            c.ops.append(("PS2", perm[2 * k], perm[2 * k + 1], rng.angle()))
        if ctrl3 and n >= 4:
            qs = rng.sample(n, 4)
            c.ops.append(("CU1", qs[:3], qs[3], haar_unitary(rng, 2)))
    if qft:
        c.ops.append(("FQFT",))
    lo, hi = gadget_range
    for _ in range(gadgets):
        k = min(n, lo + rng.randint(hi - lo + 1))
        targs = rng.sample(n, k)
        c.ops.append(("PG", random_pauli_chars(rng, k, allow_identity=False), targs, rng.angle()))
    if probs:
        for q in range(n):
            c.ops.append(("OUT_PROB", q, 0))
    for _ in range(num_expec):
        c.ops.append(("OUT_EXPEC", random_pauli_chars(rng, n), list(range(n))))
    return c


def noisy_density(n: int, depth: int, seed: int, *, kraus_every: int = 5, num_expec: int = 8) -> Circuit:
    rng = SplitMix64(seed)
    c = Circuit(n, True, [], ("zero",), f"density_n{n}_d{depth}_s{seed}")
    for layer in range(depth):
        for q in range(n):
            c.ops.append(("U1", q, haar_unitary(rng, 2)))
            c.ops.append(("DEPOL", q, 0.01 * rng.uniform()))
        perm = rng.permutation(n)
        for k in range(n // 2):
            a, b = perm[2 * k], perm[2 * k + 1]
            c.ops.append(("CZ", a, b))
            c.ops.append(("DEPOL", a, 0.01 * rng.uniform()))
            c.ops.append(("DEPOL", b, 0.01 * rng.uniform()))
        if kraus_every and (layer + 1) % kraus_every == 0 and n >= 2:
            qs = rng.sample(n, 2)
            ops = kraus_from_isometry(random_isometry(rng, 16, 4), 4, 4)
            c.ops.append(("KRAUS", qs, ops))
    c.ops.append(("OUT_PURITY",))
    for _ in range(num_expec):
        c.ops.append(("OUT_EXPEC", random_pauli_chars(rng, n), list(range(n))))
    for q in range(n):
        c.ops.append(("OUT_PROB", q, 0))
    return c


def variational(n: int, depth: int, seed: int, *, num_expec: int = 8) -> Circuit:
    rng = SplitMix64(seed)
    c = Circuit(n, False, [], ("zero",), f"variational_n{n}_d{depth}_s{seed}")
    for _ in range(depth):
        for q in range(n):
            c.ops.append(("RY", q, rng.angle()))
            c.ops.append(("RZ", q, rng.angle()))
        for q in range(n - 1):
            c.ops.append(("CZ", q, q + 1))
    for _ in range(num_expec):
        c.ops.append(("OUT_EXPEC", random_pauli_chars(rng, n), list(range(n))))
    return c


def _dagger(m):
    return [[complex(m[c][r]).conjugate() for c in range(len(m))] for r in range(len(m))]


def inverse_ops(c: Circuit) -> list:
    """The gates of c undone, in reverse order: matrices by their adjoints,
    phase shifts and Pauli gadgets by their negated angles, a full QFT by its
    gate sequence (operations.cpp:1934-1953) reversed and conjugated.
    Outputs are dropped; only unitary ops are supported."""
    out = []
    for op in reversed(c.ops):
        k = op[0]
        if k.startswith("OUT_"):
            continue
        if k == "U1":
            out.append(("U1", op[1], _dagger(op[2])))
        elif k == "CU1":
            out.append(("CU1", op[1], op[2], _dagger(op[3])))
        elif k == "PS2":
            out.append(("PS2", op[1], op[2], -op[3]))
        elif k == "PG":
            out.append(("PG", op[1], op[2], -op[3]))
        elif k == "FQFT":
            n = c.num_qubits
            fwd = []
            for t in range(n - 1, -1, -1):
                fwd.append(("H", t))
                for m in range(t):
                    fwd.append(("PS2", t, t - m - 1, math.pi / (1 << (m + 1))))
            fwd += [("SWAP", t, n - 1 - t) for t in range(n // 2)]
            for g in reversed(fwd):
                out.append(("PS2", g[1], g[2], -g[3]) if g[0] == "PS2" else g)
        else:
            raise ValueError(f"inverse_ops: unsupported op {k}")
    return out


# ------------------------------------------------------------ text format

def _fmt(x: float) -> str:
    return repr(float(x))  # shortest repr round-trips exactly


def _cfmt(z: complex) -> str:
    return f"{_fmt(z.real)} {_fmt(z.imag)}"


def _mat(m) -> str:
    return " ".join(_cfmt(complex(v)) for row in m for v in row)


def op_to_line(op) -> str:
    k = op[0]
    if k == "U1":
        return f"U1 {op[1]} {_mat(op[2])}"
    if k == "CU1":
        return f"CU1 {len(op[1])} {' '.join(map(str, op[1]))} {op[2]} {_mat(op[3])}"
    if k == "U2":
        return f"U2 {op[1]} {op[2]} {_mat(op[3])}"
    if k == "UN":
        return f"UN {len(op[1])} {' '.join(map(str, op[1]))} {_mat(op[2])}"
    if k == "D1":
        return f"D1 {op[1]} {' '.join(_cfmt(complex(v)) for v in op[2])}"
    if k == "DN":
        return f"DN {len(op[1])} {' '.join(map(str, op[1]))} {' '.join(_cfmt(complex(v)) for v in op[2])}"
    if k in ("H", "X", "Y", "Z", "S", "T"):
        return f"{k} {op[1]}"
    if k in ("CZ", "CX", "SWAP"):
        return f"{k} {op[1]} {op[2]}"
    if k in ("PS", "RX", "RY", "RZ"):
        return f"{k} {op[1]} {_fmt(op[2])}"
    if k == "PS2":
        return f"PS2 {op[1]} {op[2]} {_fmt(op[3])}"
    if k == "CSWAP":
        return f"CSWAP {len(op[1])} {' '.join(map(str, op[1]))} {op[2]} {op[3]}"
    if k == "PSTR":
        return f"PSTR {op[1]} {len(op[2])} {' '.join(map(str, op[2]))}"
    if k == "CPSTR":
        return f"CPSTR {len(op[1])} {' '.join(map(str, op[1]))} {op[2]} {len(op[3])} {' '.join(map(str, op[3]))}"
    if k == "PG":
        return f"PG {op[1]} {len(op[2])} {' '.join(map(str, op[2]))} {_fmt(op[3])}"
    if k == "CPG":
        return (f"CPG {len(op[1])} {' '.join(map(str, op[1]))} {op[2]} {len(op[3])} "
                f"{' '.join(map(str, op[3]))} {_fmt(op[4])}")
    if k == "PHG":
        return f"PHG {len(op[1])} {' '.join(map(str, op[1]))} {_fmt(op[2])}"
    if k == "QFT":
        return f"QFT {len(op[1])} {' '.join(map(str, op[1]))}"
    if k in ("FQFT", "OUT_PURITY", "OUT_TOTALPROB"):
        return k
    if k in ("DEPH", "DEPOL", "DAMP"):
        return f"{k} {op[1]} {_fmt(op[2])}"
    if k in ("DEPH2", "DEPOL2"):
        return f"{k} {op[1]} {op[2]} {_fmt(op[3])}"
    if k == "PAULIS":
        return f"PAULIS {op[1]} {_fmt(op[2])} {_fmt(op[3])} {_fmt(op[4])}"
    if k == "KRAUS":
        mats = " ".join(_mat(m) for m in op[2])
        return f"KRAUS {len(op[1])} {' '.join(map(str, op[1]))} {len(op[2])} {mats}"
    if k == "PROJ":
        return f"PROJ {op[1]} {op[2]}"
    if k == "SEEDS":
        return f"SEEDS {len(op[1])} {' '.join(map(str, op[1]))}"
    if k == "MEAS":
        return f"MEAS {op[1]}"
    if k == "OUT_PROB":
        return f"OUT_PROB {op[1]} {op[2]}"
    if k == "OUT_EXPEC":
        return f"OUT_EXPEC {op[1]} {len(op[2])} {' '.join(map(str, op[2]))}"
    if k == "OUT_PROBS":
        return f"OUT_PROBS {len(op[1])} {' '.join(map(str, op[1]))}"
    if k == "MEASN":
        return f"MEASN {len(op[1])} {' '.join(map(str, op[1]))}"
    if k == "RANDPURE":
        return f"RANDPURE {op[1]} {int(op[2])}"
    if k == "RANDMIXED":
        return f"RANDMIXED {op[1]} {op[2]}"
    if k == "AUX":
        return f"AUX {op[1]} {int(op[2])} {op[3]}"
    if k in ("AUXCOPY", "OUT_INNER", "OUT_FIDELITY", "OUT_DISTANCE"):
        return k
    raise ValueError(f"unknown op {k}")


def to_text(c: Circuit, init_amps_path: str | None = None, out_amps_path: str | None = None) -> str:
    lines = [f"# {c.name}", f"QUREG {c.num_qubits} {int(c.dens)}"]
    if c.init[0] == "zero":
        lines.append("INITZERO")
    elif c.init[0] == "debug":
        lines.append("INITDEBUG")
    elif c.init[0] == "plus":
        lines.append("INITPLUS")
    else:
        if init_amps_path is None:
            raise ValueError("a random initial state needs an amplitude file")
        lines.append(f"INITAMPS {init_amps_path}")
    lines += [op_to_line(op) for op in c.ops]
    if out_amps_path:
        lines.append(f"OUT_AMPS {out_amps_path}")
    return "\n".join(lines) + "\n"


def from_text(text: str) -> Circuit:
    """Inverse of to_text (the INITAMPS path, if any, is kept in init)."""
    c = None
    name = ""
    for raw in text.splitlines():
        if c is None and raw.startswith("#") and not name:
            name = raw[1:].strip()
        line = raw.split("#", 1)[0].strip()
        if not line:
            continue
        t = line.split()
        k = t[0]
        pos = 1

        def take(n=1):
            nonlocal pos
            v = t[pos:pos + n]
            pos += n
            return v

        def ints(n):
            return [int(x) for x in take(n)]

        def comps(n):
            v = [float(x) for x in take(2 * n)]
            return [complex(v[2 * i], v[2 * i + 1]) for i in range(n)]

        def mat(d):
            e = comps(d * d)
            return [e[r * d:(r + 1) * d] for r in range(d)]

        if k == "QUREG":
            c = Circuit(int(t[1]), bool(int(t[2])), [], ("zero",), name)
            continue
        if k == "INITZERO":
            c.init = ("zero",)
        elif k == "INITDEBUG":
            c.init = ("debug",)
        elif k == "INITPLUS":
            c.init = ("plus",)
        elif k == "INITAMPS":
            c.init = ("amps", t[1])
        elif k == "OUT_AMPS":
            continue
        elif k == "U1":
            tg = ints(1)[0]
            c.ops.append(("U1", tg, mat(2)))
        elif k == "CU1":
            nc = int(take()[0])
            cs = ints(nc)
            tg = ints(1)[0]
            c.ops.append(("CU1", cs, tg, mat(2)))
        elif k == "U2":
            a, b = ints(2)
            c.ops.append(("U2", a, b, mat(4)))
        elif k == "UN":
            nt = int(take()[0])
            ts = ints(nt)
            c.ops.append(("UN", ts, mat(1 << nt)))
        elif k == "D1":
            tg = ints(1)[0]
            c.ops.append(("D1", tg, comps(2)))
        elif k == "DN":
            nt = int(take()[0])
            ts = ints(nt)
            c.ops.append(("DN", ts, comps(1 << nt)))
        elif k in ("H", "X", "Y", "Z", "S", "T", "MEAS"):
            c.ops.append((k, int(t[1])))
        elif k in ("CZ", "CX", "SWAP", "PROJ", "OUT_PROB"):
            c.ops.append((k, int(t[1]), int(t[2])))
        elif k in ("PS", "RX", "RY", "RZ", "DEPH", "DEPOL", "DAMP"):
            c.ops.append((k, int(t[1]), float(t[2])))
        elif k in ("PS2", "DEPH2", "DEPOL2"):
            c.ops.append((k, int(t[1]), int(t[2]), float(t[3])))
        elif k == "PAULIS":
            c.ops.append((k, int(t[1]), float(t[2]), float(t[3]), float(t[4])))
        elif k == "CSWAP":
            nc = int(take()[0])
            cs = ints(nc)
            a, b = ints(2)
            c.ops.append(("CSWAP", cs, a, b))
        elif k in ("PSTR", "PG", "OUT_EXPEC"):
            chars = take()[0]
            nt = int(take()[0])
            ts = ints(nt)
            c.ops.append((k, chars, ts) if k != "PG" else (k, chars, ts, float(take()[0])))
        elif k in ("CPSTR", "CPG"):
            nc = int(take()[0])
            cs = ints(nc)
            chars = take()[0]
            nt = int(take()[0])
            ts = ints(nt)
            c.ops.append((k, cs, chars, ts) if k == "CPSTR" else (k, cs, chars, ts, float(take()[0])))
        elif k == "PHG":
            nt = int(take()[0])
            ts = ints(nt)
            c.ops.append(("PHG", ts, float(take()[0])))
        elif k == "QFT":
            nt = int(take()[0])
            c.ops.append(("QFT", ints(nt)))
        elif k in ("FQFT", "OUT_PURITY", "OUT_TOTALPROB"):
            c.ops.append((k,))
        elif k == "KRAUS":
            nq = int(take()[0])
            ts = ints(nq)
            nops = int(take()[0])
            c.ops.append(("KRAUS", ts, [mat(1 << nq) for _ in range(nops)]))
        elif k == "SEEDS":
            ns = int(take()[0])
            c.ops.append(("SEEDS", ints(ns)))
        elif k in ("OUT_PROBS", "MEASN"):
            nq = int(take()[0])
            c.ops.append((k, ints(nq)))
        elif k == "RANDPURE":
            c.ops.append((k, int(t[1]), bool(int(t[2]))))
        elif k == "RANDMIXED":
            c.ops.append((k, int(t[1]), int(t[2])))
        elif k == "AUX":
            c.ops.append((k, int(t[1]), bool(int(t[2])), int(t[3])))
        elif k in ("AUXCOPY", "OUT_INNER", "OUT_FIDELITY", "OUT_DISTANCE"):
            c.ops.append((k,))
        else:
            raise ValueError(f"unknown op line: {line}")
    return c


def initial_amps(c: Circuit) -> np.ndarray | None:
    if c.init[0] == "random":
        return random_state(c.num_qubits, c.init[1], c.dens)
    return None


# ------------------------------------------------------------ replay

def run(c: Circuit, q=None, record: list | None = None, api=None):
    """Apply the circuit to a Qureg of the MI355X backend (creates one, in
    FP64, unless given).  Returns (qureg, results) where results lists the
    OUT_* / MEAS values in order."""
    if api is None:
        from . import api as api_mod
        api = api_mod
    if q is None:
        q = (api.createDensityQureg(c.num_qubits) if c.dens else api.createQureg(c.num_qubits))
    init = c.init[0]
    if init == "zero":
        api.initZeroState(q)
    elif init == "debug":
        api.initDebugState(q)
    elif init == "plus":
        api.initPlusState(q)
    elif init == "random":
        api.set_amps(q, initial_amps(c))
    results = [] if record is None else record
    ctx = {}
    try:
        for op in c.ops:
            apply_op(api, q, op, results, ctx)
    finally:
        if "aux" in ctx:
            api.destroyQureg(ctx.pop("aux"))
    return q, results


def apply_op(api, q, op, results: list, ctx: dict | None = None) -> None:
    """One op on qureg q; ctx holds the auxiliary qureg of AUX/AUXCOPY."""
    k = op[0]
    A = api
    ctx = {} if ctx is None else ctx
    if k == "U1":
        A.applyCompMatr1(q, op[1], A.compmatr1(op[2]))
    elif k == "CU1":
        A.applyMultiControlledCompMatr1(q, A.ints(op[1]), len(op[1]), op[2], A.compmatr1(op[3]))
    elif k == "U2":
        A.applyCompMatr2(q, op[1], op[2], A.compmatr2(op[3]))
    elif k == "UN":
        m = A.make_compmatr(np.array(op[2], dtype=complex))
        A.applyCompMatr(q, A.ints(op[1]), len(op[1]), m)
        A.destroyCompMatr(m)
    elif k == "D1":
        A.applyDiagMatr1(q, op[1], A.diagmatr1(op[2]))
    elif k == "DN":
        m = A.make_diagmatr(np.array(op[2], dtype=complex))
        A.applyDiagMatr(q, A.ints(op[1]), len(op[1]), m)
        A.destroyDiagMatr(m)
    elif k == "H":
        A.applyHadamard(q, op[1])
    elif k == "X":
        A.applyPauliX(q, op[1])
    elif k == "Y":
        A.applyPauliY(q, op[1])
    elif k == "Z":
        A.applyPauliZ(q, op[1])
    elif k == "S":
        A.applyS(q, op[1])
    elif k == "T":
        A.applyT(q, op[1])
    elif k == "CZ":
        A.applyControlledPauliZ(q, op[1], op[2])
    elif k == "CX":
        A.applyControlledPauliX(q, op[1], op[2])
    elif k == "PS":
        A.applyPhaseShift(q, op[1], op[2])
    elif k == "PS2":
        A.applyTwoQubitPhaseShift(q, op[1], op[2], op[3])
    elif k == "RX":
        A.applyRotateX(q, op[1], op[2])
    elif k == "RY":
        A.applyRotateY(q, op[1], op[2])
    elif k == "RZ":
        A.applyRotateZ(q, op[1], op[2])
    elif k == "SWAP":
        A.applySwap(q, op[1], op[2])
    elif k == "CSWAP":
        A.applyMultiControlledSwap(q, A.ints(op[1]), len(op[1]), op[2], op[3])
    elif k == "PSTR":
        A.applyPauliStr(q, A.pauli_str(op[1], op[2]))
    elif k == "CPSTR":
        A.applyMultiControlledPauliStr(q, A.ints(op[1]), len(op[1]), A.pauli_str(op[2], op[3]))
    elif k == "PG":
        A.applyPauliGadget(q, A.pauli_str(op[1], op[2]), op[3])
    elif k == "CPG":
        A.applyMultiControlledPauliGadget(q, A.ints(op[1]), len(op[1]), A.pauli_str(op[2], op[3]), op[4])
    elif k == "PHG":
        A.applyPhaseGadget(q, A.ints(op[1]), len(op[1]), op[2])
    elif k == "QFT":
        A.applyQuantumFourierTransform(q, A.ints(op[1]), len(op[1]))
    elif k == "FQFT":
        A.applyFullQuantumFourierTransform(q)
    elif k == "DEPH":
        A.mixDephasing(q, op[1], op[2])
    elif k == "DEPH2":
        A.mixTwoQubitDephasing(q, op[1], op[2], op[3])
    elif k == "DEPOL":
        A.mixDepolarising(q, op[1], op[2])
    elif k == "DEPOL2":
        A.mixTwoQubitDepolarising(q, op[1], op[2], op[3])
    elif k == "DAMP":
        A.mixDamping(q, op[1], op[2])
    elif k == "PAULIS":
        A.mixPaulis(q, op[1], op[2], op[3], op[4])
    elif k == "KRAUS":
        km = A.make_krausmap([np.array(m, dtype=complex) for m in op[2]])
        A.mixKrausMap(q, A.ints(op[1]), len(op[1]), km)
        A.destroyKrausMap(km)
    elif k == "PROJ":
        A.applyQubitProjector(q, op[1], op[2])
    elif k == "SEEDS":
        A.set_seeds(op[1])
    elif k == "MEAS":
        results.append(("meas", A.applyQubitMeasurement(q, op[1])))
    elif k == "OUT_PROB":
        results.append(("prob", A.calcProbOfQubitOutcome(q, op[1], op[2])))
    elif k == "OUT_EXPEC":
        results.append(("expec", A.calcExpecPauliStr(q, A.pauli_str(op[1], op[2]))))
    elif k == "OUT_PURITY":
        results.append(("purity", A.calcPurity(q)))
    elif k == "OUT_TOTALPROB":
        results.append(("totalprob", A.calcTotalProb(q)))
    elif k == "OUT_PROBS":
        results.append(("probs", list(A.probs_of_all_outcomes(q, op[1]))))
    elif k == "MEASN":
        results.append(("measn", A.applyMultiQubitMeasurement(q, A.ints(op[1]), len(op[1]))))
    elif k in ("RANDPURE", "RANDMIXED"):
        # a scratch qureg: only the measurement RNG stream it advances matters
        s = A.createDensityQureg(op[1]) if (k == "RANDMIXED" or op[2]) else A.createQureg(op[1])
        try:
            if k == "RANDPURE":
                A.initRandomPureState(s)
            else:
                A.initRandomMixedState(s, op[2])
        finally:
            A.destroyQureg(s)
    elif k in ("AUX", "AUXCOPY"):
        if "aux" in ctx:
            A.destroyQureg(ctx.pop("aux"))
        if k == "AUX":
            ctx["aux"] = A.createDensityQureg(op[1]) if op[2] else A.createQureg(op[1])
            A.set_amps(ctx["aux"], random_state(op[1], op[3], op[2]))
        else:
            ctx["aux"] = A.createCloneQureg(q)
    elif k == "OUT_INNER":
        z = A.inner_product(q, ctx["aux"])
        results.append(("inner", [z.real, z.imag]))
    elif k == "OUT_FIDELITY":
        results.append(("fidelity", A.calcFidelity(q, ctx["aux"])))
    elif k == "OUT_DISTANCE":
        results.append(("distance", A.calcDistance(q, ctx["aux"])))
    else:
        raise ValueError(f"unknown op {k}")


def parse_results(text: str) -> list:
    """Parse 'R label values...' lines printed by the reference-side driver."""
    out = []
    for line in text.splitlines():
        parts = line.split()
        if not parts or parts[0] != "R":
            continue
        label = parts[1]
        if label in ("meas", "measn"):
            out.append((label, int(parts[2])))
        elif label in ("probs", "inner"):
            out.append((label, [float(v) for v in parts[2:]]))
        else:
            out.append((label, float(parts[2])))
    return out


# ------------------------------------------------------------ byte accounting

def gate_bytes(op, num_qubits: int, dens: bool, amp_bytes: int = 16) -> int:
    """Algorithmic HBM bytes of one op as dispatched (read + write of the
    amplitudes it touches), following the per-gate formulas of the
    measurement plan: dense/gadget/diagonal 2*S*2^(n-c); swap 2*S*2^(n-c-1);
    density matrices pay two passes over 2N index bits for unitaries and one
    for channels; reductions read S*2^(n-1) (one qubit outcome) or S*2^n."""
    n = 2 * num_qubits if dens else num_qubits
    passes = 2 if dens else 1
    full = amp_bytes * (1 << n)
    k = op[0]
    if k in ("U1", "U2", "UN", "D1", "DN", "H", "X", "Y", "S", "T", "RX", "RY", "RZ", "PS", "PG", "PSTR",
             "PHG", "Z"):
        return passes * 2 * full
    if k == "CU1":
        return passes * 2 * (full >> len(op[1]))
    if k in ("PS2", "CZ", "CX"):
        return passes * 2 * (full >> 1)
    if k == "SWAP":
        return passes * (full >> 1)
    if k == "FQFT":
        m = num_qubits
        return passes * (m * 2 * full + (m * (m - 1) // 2) * (full) + (m // 2) * (full >> 1))
    if k in ("DEPOL", "DEPOL2", "KRAUS", "PAULIS", "DAMP"):
        return 2 * full
    if k in ("DEPH",):
        return full
    if k == "OUT_PROB":
        return (amp_bytes * (1 << num_qubits)) // 2 if not dens else 128 * (1 << num_qubits) // 2
    if k == "OUT_EXPEC":
        return amp_bytes * (1 << num_qubits) if not dens else 128 * (1 << num_qubits)
    if k in ("OUT_PURITY", "OUT_TOTALPROB"):
        return full
    return 0


def count_gates(c: Circuit) -> int:
    """Gates as the API dispatches them (a full QFT counts its n H, n(n-1)/2
    phase shifts and n/2 swaps; outputs and channels are not gates)."""
    total = 0
    for op in c.ops:
        k = op[0]
        if k.startswith("OUT_") or k in ("SEEDS", "MEAS", "MEASN", "RANDPURE", "RANDMIXED", "AUX", "AUXCOPY", "DEPOL", "DEPOL2", "DEPH", "DEPH2", "DAMP", "PAULIS",
                                         "KRAUS"):
            continue
        if k == "FQFT":
            n = c.num_qubits
            total += n + n * (n - 1) // 2 + n // 2
        else:
            total += 1
    return total
