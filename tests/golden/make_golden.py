"""Write the golden fixtures of tests/golden/ from QuEST's own CPU build.

Runs in the development container only (the reference does not travel):
builds oracle/_ref (QuEST v4.1.0 FP64 + OpenMP, compiled from its sources by
oracle/Makefile), replays each seeded circuit through oracle/circuit_driver
and stores, per case, one .npz holding

  circuit   the circuit text (exact 17-digit reals) that was replayed
  init      the initial amplitudes (when not a closed-form state)
  amps      final amplitudes (small cases) or a fixed sample of them
  sample    indices of the sampled amplitudes (large cases)
  labels / values   every OUT_* / MEAS result, in order

Usage:  python tests/golden/make_golden.py [case names...]   (all by default)
"""

from __future__ import annotations

import subprocess
import sys
import tempfile
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parents[2]
sys.path.insert(0, str(ROOT))

from quest_amd import circuits as qc  # noqa: E402
from oracle import qoracle  # noqa: E402

OUT = Path(__file__).resolve().parent


def cases():
    # statevector layer family (configs: 20-qubit plain CPU case and its
    # 30-qubit GPU form, here at sizes whose final state fits a fixture)
    yield "layered_n10", qc.layered(10, 6, 101, ctrl3=True, qft=True, gadgets=8, gadget_range=(2, 8),
                                     num_expec=4), True
    yield "layered_n12", qc.layered(12, 4, 202, ctrl3=True, qft=True, gadgets=4, gadget_range=(4, 12),
                                     num_expec=4), True
    yield "density_n5", qc.noisy_density(5, 10, 303), True
    yield "variational_n8", qc.variational(8, 6, 404), True
    yield "allops_sv_n7", allops(7, False, 505), True
    yield "allops_dm_n4", allops(4, True, 606), True
    yield "measure_sv_n6", measure_case(6, False, 707), True
    yield "measure_dm_n3", measure_case(3, True, 808), True
    # the 20-qubit plain-CPU configuration at full size: results + a sample
    yield "layered_n20_c1", qc.layered(20, 20, 909, num_expec=1), False
    yield "random_seq_sv_n5", random_seq_case(5, 1001), True
    yield "measn_sv_n6", measn_case(6, False, 1102), True
    yield "measn_dm_n3", measn_case(3, True, 1203), True
    yield "overlap_sv_n6", overlap_case(6, False, 1304), True
    yield "overlap_dm_n3", overlap_case(3, True, 1405), True
    yield "probs_sv_n12", probs_case(12, False, 1506), True
    yield "probs_dm_n5", probs_case(5, True, 1607), True


def allops(n: int, dens: bool, seed: int) -> qc.Circuit:
    """One of every operation the circuit format carries."""
    rng = qc.SplitMix64(seed)
    c = qc.Circuit(n, dens, [], ("random", seed), f"allops_{'dm' if dens else 'sv'}_n{n}")
    ops = c.ops
    H = qc.haar_unitary
    ops += [("U1", 0, H(rng, 2)), ("CU1", [1, 2], 3, H(rng, 2)), ("U2", 2, 0, H(rng, 4)),
            ("UN", [3, 1, 0], H(rng, 8)) if n >= 4 else ("U2", 1, 0, H(rng, 4)),
            ("D1", 1, [complex(np.cos(a), np.sin(a)) for a in (rng.angle(), rng.angle())]),
            ("DN", [2, 0], [complex(np.cos(a), np.sin(a)) for a in (rng.angle() for _ in range(4))]),
            ("H", 3), ("X", 0), ("Y", 1), ("Z", 2), ("S", 3), ("T", 0), ("CX", 1, 3), ("CZ", 0, 2),
            ("PS", 1, rng.angle()), ("PS2", 3, 1, rng.angle()), ("RX", 0, rng.angle()), ("RY", 2, rng.angle()),
            ("RZ", 1, rng.angle()), ("SWAP", 0, 3), ("CSWAP", [2], 1, 3),
            ("PSTR", "XYZ", [2, 0, 3]), ("CPSTR", [1], "YY", [0, 2]), ("PSTR", "ZZ", [0, 3]),
            ("PG", "XZY", [1, 3, 0], rng.angle()), ("PG", "ZZZ", [0, 1, 2], rng.angle()),
            ("CPG", [3], "XY", [0, 1], rng.angle()), ("PHG", [0, 2, 3], rng.angle()), ("QFT", [1, 2, 3])]
    if n > 4:
        ops += [("U1", q, H(rng, 2)) for q in range(4, n)]
        ops += [("PG", "XYZX", [n - 1, 0, 4, 2], rng.angle()), ("CU1", [0, 1, 2], n - 1, H(rng, 2))]
    if dens:
        ops += [("DEPH", 0, 0.1), ("DEPH2", 1, 3, 0.2), ("DEPOL", 2, 0.3), ("DEPOL2", 0, 3, 0.4),
                ("DAMP", 1, 0.25), ("PAULIS", 3, 0.05, 0.1, 0.15),
                ("KRAUS", [2], qc.kraus_from_isometry(qc.random_isometry(rng, 6, 2), 3, 2)),
                ("KRAUS", [3, 0], qc.kraus_from_isometry(qc.random_isometry(rng, 16, 4), 4, 4))]
    ops += [("OUT_TOTALPROB",), ("OUT_PURITY",)]
    ops += [("OUT_PROB", q, o) for q in range(n) for o in (0, 1)]
    ops += [("OUT_EXPEC", qc.random_pauli_chars(rng, n), list(range(n))) for _ in range(6)]
    ops += [("OUT_EXPEC", "Z" * 2, [0, n - 1]), ("OUT_EXPEC", "I" * n, list(range(n)))]
    if not dens:
        ops += [("OUT_PROBS", [0, 2]), ("OUT_PROBS", [3, 1, 0])]
    return c


def measure_case(n: int, dens: bool, seed: int) -> qc.Circuit:
    """Seeded measurements: outcomes must be bit-exact (randomiser.cpp)."""
    rng = qc.SplitMix64(seed)
    c = qc.Circuit(n, dens, [], ("random", seed), f"measure_{'dm' if dens else 'sv'}_n{n}")
    c.ops.append(("SEEDS", [seed, seed + 1, 7]))
    for q in range(n):
        c.ops.append(("MEAS", q))
        c.ops.append(("U1", q, qc.haar_unitary(rng, 2)))
        c.ops.append(("MEAS", (q + 1) % n))
    c.ops.append(("PROJ", 0, 1))
    c.ops.append(("OUT_TOTALPROB",))
    return c


def random_seq_case(n: int, seed: int) -> qc.Circuit:
    """Seeded measurements interleaved with random-state initialisations of
    scratch quregs: each initRandomPureState draws one seed from the
    measurement RNG (randomiser.cpp:175-193, initialisations.cpp:121-133),
    initRandomMixedState one per pure state it mixes (localiser.cpp:801-827),
    so the outcomes pin how far each advances the stream."""
    rng = qc.SplitMix64(seed)
    c = qc.Circuit(n, False, [], ("random", seed), f"random_seq_sv_n{n}")
    c.ops.append(("SEEDS", [seed, 11, 12345]))
    c.ops.append(("MEAS", 0))
    c.ops.append(("RANDPURE", 4, False))
    for q in range(n):
        c.ops.append(("U1", q, qc.haar_unitary(rng, 2)))
        c.ops.append(("MEAS", q))
    c.ops.append(("RANDPURE", 3, True))
    c.ops.append(("MEAS", 1))
    c.ops.append(("RANDMIXED", 3, 3))
    for q in range(n):
        c.ops.append(("U1", q, qc.haar_unitary(rng, 2)))
    c.ops.append(("MEAS", 2))
    c.ops.append(("RANDPURE", 2, False))
    c.ops.append(("MEASN", [4, 0, 3]))
    c.ops.append(("OUT_TOTALPROB",))
    return c


def measn_case(n: int, dens: bool, seed: int) -> qc.Circuit:
    """Seeded multi-qubit measurements (operations.cpp:1848-1925, outcome
    sampled by randomiser.cpp:140-166): outcomes bit-exact, then the
    collapsed state."""
    rng = qc.SplitMix64(seed)
    c = qc.Circuit(n, dens, [], ("random", seed), f"measn_{'dm' if dens else 'sv'}_n{n}")
    c.ops.append(("SEEDS", [seed, 3]))
    groups = [[1, 0], [n - 1, 1, 0], [2]] if n >= 3 else [[1, 0], [0]]
    for g in groups:
        for q in range(n):
            c.ops.append(("U1", q, qc.haar_unitary(rng, 2)))
        c.ops.append(("MEASN", g))
        c.ops.append(("OUT_PROBS", sorted(g)))
    c.ops.append(("MEAS", 0))
    c.ops.append(("OUT_TOTALPROB",))
    return c


def overlap_case(n: int, dens: bool, seed: int) -> qc.Circuit:
    """Inner products, fidelities and distances (calculations.cpp:49-61,
    303-364) between the main qureg and an auxiliary one of every kind the
    reference accepts: statevector pairs (Bures distance), density-matrix
    pairs (Hilbert-Schmidt distance), and mixed pairs (<psi|rho|psi>,
    purified distance)."""
    rng = qc.SplitMix64(seed)
    c = qc.Circuit(n, dens, [], ("random", seed), f"overlap_{'dm' if dens else 'sv'}_n{n}")
    ops = c.ops
    if dens:
        for q in range(n):
            ops.append(("DEPOL", q, 0.05 + 0.1 * rng.uniform()))
    outs = [("OUT_INNER",), ("OUT_DISTANCE",)]
    ops.append(("AUX", n, dens, seed + 1))
    ops += outs + ([] if dens else [("OUT_FIDELITY",)])
    ops.append(("AUX", n, not dens, seed + 2))
    ops += outs + [("OUT_FIDELITY",)]
    ops.append(("AUXCOPY",))
    for q in range(n):
        ops.append(("U1", q, qc.haar_unitary(rng, 2)))
    ops += outs + ([] if dens else [("OUT_FIDELITY",)])
    return c


def probs_case(n: int, dens: bool, seed: int) -> qc.Circuit:
    """All-outcome probability tables of 1 to 8 qubits, unsorted qubit lists
    included (cpu_subroutines.cpp:1828-1903)."""
    rng = qc.SplitMix64(seed)
    c = qc.noisy_density(n, 3, seed) if dens else qc.layered(n, 3, seed, num_expec=0, probs=False)
    c.name = f"probs_{'dm' if dens else 'sv'}_n{n}"
    c.ops = [op for op in c.ops if not op[0].startswith("OUT_")]
    for k in range(1, min(n, 8) + 1):
        c.ops.append(("OUT_PROBS", rng.sample(n, k)))
    return c


def main() -> None:
    subprocess.run(["make", "-s", "-C", str(ROOT / "oracle"), "ref"], check=True)
    with tempfile.TemporaryDirectory() as tmp:
        only = set(sys.argv[1:])
        for name, circ, full in cases():
            if only and name not in only:
                continue
            circ.name = name
            init = qc.initial_amps(circ)
            amps, results, _ = qoracle.run_reference(circ, Path(tmp), init_amps=init, want_amps=True)
            # results flattened: values[offsets[i]:offsets[i+1]] belongs to labels[i]
            flat, offs = [], [0]
            for r in results:
                flat.extend(np.atleast_1d(np.asarray(r[1], dtype=np.float64)).tolist())
                offs.append(len(flat))
            payload = {
                "circuit": np.array(qc.to_text(circ, "init.bin" if init is not None else None)),
                "labels": np.array([r[0] for r in results]),
                "values": np.array(flat, dtype=np.float64),
                "offsets": np.array(offs, dtype=np.int64),
            }
            if init is not None:
                payload["init_seed"] = np.array([circ.init[1]], dtype=np.uint64)
                if full:
                    payload["init"] = init
            if full:
                payload["amps"] = amps
            else:
                idx = np.unique(np.concatenate([np.arange(1024),
                                                np.linspace(0, amps.size - 1, 3072).astype(np.int64)]))
                payload["sample"] = idx
                payload["amps"] = amps[idx]
                payload["norm2"] = np.array([np.sum(np.abs(amps) ** 2)])
            np.savez_compressed(OUT / f"{name}.npz", **payload)
            print(f"wrote {name}.npz ({len(results)} results, {amps.size} amps)")


if __name__ == "__main__":
    main()
