#!/usr/bin/env python3
"""Benchmark of the drop-in GPU path: 30-qubit FP64 statevector gates/s.

Workload (one "step"): on a 30-qubit complex128 statevector (16 GiB, resident
in HBM), 40 layers of [a Haar-random applyCompMatr1 on each of the 30 qubits,
applyTwoQubitPhaseShift(U(0,2pi)) on 15 random disjoint pairs, one 3-control
applyMultiControlledCompMatr1], then applyFullQuantumFourierTransform (30 H,
435 phase shifts, 15 swaps), then 64 applyPauliGadget on 4-12 random qubits,
then calcProbOfQubitOutcome on all 30 qubits and 16 calcExpecPauliStr.
2184 gates and 46 reductions per step, every one a separate call of the
quest.h API, i.e. one pass over the state each (no fusion).

Output: one JSON line (see DESIGN.md for the field definitions):
  value      gates/s over the timed steps (HIP events on the library stream
             bracket them; wall clock agrees after syncQuESTEnv)
  roofline   the dominant kernel (1-target dense, applyCompMatr1): algorithmic
             bytes per launch 2 * 16 B * 2^30 = 32 GiB, divided by its average
             duration measured with HIP events around each layer's block of 30
             launches inside the timed steps
  cpu_baseline  QuEST's own FP64 OpenMP CPU build (oracle/_ref) on a bounded
             sample of the same workload at the same size, on this host
"""

from __future__ import annotations

import argparse
import json
import os
import subprocess
import sys
import tempfile
import time
from pathlib import Path

ROOT = Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT))

METRIC = "30q FP64 gates/s & % HBM roofline; batched circuits/s + p99 latency"
HBM_PEAK = 8.0e12


def log(msg: str) -> None:
    print(msg, file=sys.stderr, flush=True)


def build_step(Q, qc, circ, q):
    """Prebuilt (callable, args) list for one step, plus the indices of the
    calls that open/close each layer's block of applyCompMatr1 launches."""
    calls = []
    u1_blocks = []
    i = 0
    ops = circ.ops
    while i < len(ops):
        op = ops[i]
        if op[0] == "U1":
            start = len(calls)
            while i < len(ops) and ops[i][0] == "U1":
                calls.append((Q.applyCompMatr1, (q, ops[i][1], Q.compmatr1(ops[i][2]))))
                i += 1
            u1_blocks.append((start, len(calls)))
            continue
        k = op[0]
        if k == "PS2":
            calls.append((Q.applyTwoQubitPhaseShift, (q, op[1], op[2], op[3])))
        elif k == "CU1":
            calls.append((Q.applyMultiControlledCompMatr1, (q, Q.ints(op[1]), len(op[1]), op[2], Q.compmatr1(op[3]))))
        elif k == "FQFT":
            calls.append((Q.applyFullQuantumFourierTransform, (q,)))
        elif k == "PG":
            calls.append((Q.applyPauliGadget, (q, Q.pauli_str(op[1], op[2]), op[3])))
        elif k == "OUT_PROB":
            calls.append((Q.calcProbOfQubitOutcome, (q, op[1], op[2])))
        elif k == "OUT_EXPEC":
            calls.append((Q.calcExpecPauliStr, (q, Q.pauli_str(op[1], op[2]))))
        else:
            raise ValueError(k)
        i += 1
    return calls, u1_blocks


def cpu_baseline(circ, n: int) -> dict | None:
    """Bounded sample of the same workload on the host cores: the first 8
    applyCompMatr1 and 4 applyTwoQubitPhaseShift of layer 0 at n qubits,
    through QuEST's own CPU build if present, else through the oracle port."""
    from quest_amd import circuits as qc
    sample_ops = [op for op in circ.ops if op[0] == "U1"][:8] + [op for op in circ.ops if op[0] == "PS2"][:4]
    threads = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
    driver = ROOT / "oracle" / "_ref" / "circuit_driver"
    desc = f"{len(sample_ops)} gates (8 applyCompMatr1 + 4 applyTwoQubitPhaseShift) on a {n}-qubit FP64 statevector"
    if driver.exists():
        lines = [f"QUREG {n} 0", "INITPLUS", "TIMER_START"] + [qc.op_to_line(op) for op in sample_ops]
        lines.append("TIMER_STOP sample")
        with tempfile.NamedTemporaryFile("w", suffix=".txt", delete=False) as f:
            f.write("\n".join(lines) + "\n")
            path = f.name
        env = dict(os.environ, OMP_NUM_THREADS=str(threads), OMP_PROC_BIND="spread", OMP_PLACES="cores",
                   OMP_DYNAMIC="false")
        try:
            out = subprocess.run([str(driver), path], capture_output=True, text=True, env=env, timeout=600)
        finally:
            os.unlink(path)
        if out.returncode == 0:
            secs = [float(l.split()[2]) for l in out.stdout.splitlines() if l.startswith("T sample")]
            if secs:
                return {"value": len(sample_ops) / secs[0], "unit": "gates/s", "cores": threads,
                        "kind": "reference", "sample": desc + f", QuEST v4.1.0 FP64 OpenMP build, {secs[0]:.2f} s"}
        log(f"reference CPU build failed ({out.returncode}); falling back to the oracle port")
    from oracle import qoracle as O
    st = O.State(n)
    st.a[:] = 1 / (2 ** (n / 2))
    t0 = time.perf_counter()
    for op in sample_ops:
        O.apply_op(st, op, [])
    secs = time.perf_counter() - t0
    return {"value": len(sample_ops) / secs, "unit": "gates/s", "cores": threads, "kind": "port",
            "sample": desc + f", C restatement (OpenMP), {secs:.2f} s"}


def kernel_sweep(Q, n: int, reps: int = 5) -> None:
    """Per-target timing of the dominant kernel plus a plain device copy of
    known size (the calibration dispatch for rocprofv3 --pmc FETCH_SIZE /
    WRITE_SIZE runs): one JSON line per measurement on stdout."""
    import ctypes as C
    gpu = Q.gpu()

    def ev():
        e = C.c_void_p()
        gpu.qa_eventCreate(C.byref(e))
        return e

    e0, e1 = ev(), ev()
    ms = C.c_double()
    q = Q.createQureg(n)
    Q.initPlusState(q)
    state = 16 * (1 << n)
    # createCloneQureg copies the amplitudes with one device-to-device copy
    # dispatch: reads `state` bytes, writes `state` bytes
    for _ in range(2):
        gpu.qa_eventRecord(e0)
        c = Q.createCloneQureg(q)
        gpu.qa_eventRecord(e1)
        gpu.qa_eventElapsedMs(e0, e1, C.byref(ms))
        Q.destroyQureg(c)
        print(json.dumps({"kernel": "device copy (createCloneQureg)", "read_bytes": state, "write_bytes": state,
                          "ms": ms.value, "GBps": 2 * state / (ms.value * 1e-3) / 1e9}), flush=True)
    h = 2 ** -0.5
    m = Q.compmatr1([[h, h], [h, -h]])
    for t in range(n):
        gpu.qa_eventRecord(e0)
        for _ in range(reps):
            Q.applyCompMatr1(q, t, m)
        gpu.qa_eventRecord(e1)
        gpu.qa_eventElapsedMs(e0, e1, C.byref(ms))
        avg = ms.value / reps
        print(json.dumps({"kernel": "k_oneTargDense<double>", "target": t, "bytes": 2 * state, "ms": avg,
                          "GBps": 2 * state / (avg * 1e-3) / 1e9}), flush=True)
    Q.destroyQureg(q)


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--mode", choices=["circuit", "sweep"], default="circuit",
                    help="circuit: the headline workload; sweep: per-target dense-kernel timing + copy calibration")
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=2)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--qubits", type=int, default=30)
    ap.add_argument("--depth", type=int, default=40)
    ap.add_argument("--seed", type=int, default=2024)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    args = ap.parse_args()

    if args.gpus != 1:
        print(f"bench.py measures one MI355X: --gpus {args.gpus} is not supported by this single-GPU build "
              "(the statevector path would shard by prefix qubits; see DESIGN.md)", file=sys.stderr)
        return 2

    from quest_amd import api as Q, circuits as qc

    n = args.qubits
    if args.mode == "sweep":
        Q.ensure_env()
        kernel_sweep(Q, n)
        return 0
    circ = qc.layered(n, args.depth, args.seed, ctrl3=True, qft=True, gadgets=64, gadget_range=(4, 12),
                      num_expec=16, probs=True)
    gates = qc.count_gates(circ)
    reductions = sum(1 for op in circ.ops if op[0].startswith("OUT_"))
    alg_bytes = sum(qc.gate_bytes(op, n, False) for op in circ.ops)

    Q.ensure_env()
    env = Q.getQuESTEnv()
    log(f"bench: {n} qubits, {gates} gates + {reductions} reductions per step, GPU={env.isGpuAccelerated}")
    q = Q.createQureg(n)
    Q.initRandomPureState(q)
    calls, blocks = build_step(Q, qc, circ, q)
    gpu = Q.gpu()
    import ctypes as C
    P = C.c_void_p

    def mk_event():
        e = P()
        if gpu.qa_eventCreate(C.byref(e)) != 0:
            raise RuntimeError("event create failed")
        return e

    for w in range(args.warmup):
        t = time.perf_counter()
        for f, a in calls:
            f(*a)
        Q.syncQuESTEnv()
        log(f"warmup step {w}: {time.perf_counter() - t:.2f} s")

    ev_start, ev_stop = mk_event(), mk_event()
    blk_events = [(mk_event(), mk_event()) for _ in range(len(blocks) * args.steps)]
    blk_sizes = []
    Q.syncQuESTEnv()
    t0 = time.perf_counter()
    gpu.qa_eventRecord(ev_start)
    bi = 0
    for s in range(args.steps):
        starts = {b[0]: j for j, b in enumerate(blocks)}
        ends = {b[1] - 1: j for j, b in enumerate(blocks)}
        for ci, (f, a) in enumerate(calls):
            if ci in starts:
                gpu.qa_eventRecord(blk_events[bi][0])
            f(*a)
            if ci in ends:
                gpu.qa_eventRecord(blk_events[bi][1])
                b = blocks[ends[ci]]
                blk_sizes.append(b[1] - b[0])
                bi += 1
        log(f"step {s} enqueued")
    gpu.qa_eventRecord(ev_stop)
    Q.syncQuESTEnv()
    wall = time.perf_counter() - t0
    ms = C.c_double()
    gpu.qa_eventElapsedMs(ev_start, ev_stop, C.byref(ms))
    total_ms = ms.value
    blk_ms = 0.0
    for (a, b) in blk_events:
        gpu.qa_eventElapsedMs(a, b, C.byref(ms))
        blk_ms += ms.value
    launches = sum(blk_sizes)
    dense_avg_s = blk_ms / 1e3 / launches
    dense_bytes = 2 * 16 * (1 << n)
    achieved = dense_bytes / dense_avg_s
    value = args.steps * gates / (total_ms / 1e3)
    log(f"timed: {total_ms:.1f} ms (wall {wall * 1e3:.1f} ms) for {args.steps} steps; "
        f"applyCompMatr1 avg {dense_avg_s * 1e3:.3f} ms over {launches} launches = {achieved / 1e9:.0f} GB/s")

    cpu = None if args.no_cpu_baseline else cpu_baseline(circ, n)
    Q.destroyQureg(q)

    # HBM bytes per launch of the dominant kernel, from the calibrated
    # rocprofv3 FETCH_SIZE / WRITE_SIZE runs committed under profiles/
    traffic = None
    tfile = ROOT / "profiles" / f"k_oneTargDense_{n}q_fp64_traffic.json"
    if tfile.exists():
        traffic = json.loads(tfile.read_text()).get("traffic_bytes_per_dispatch")

    out = {
        "metric": METRIC,
        "value": value,
        "unit": "gates/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": total_ms / args.steps,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic",
        "config": {
            "workload": (f"{n}-qubit complex128 statevector: {args.depth} layers x ({n} Haar applyCompMatr1 + "
                         f"{n // 2} applyTwoQubitPhaseShift + 1 3-control applyMultiControlledCompMatr1), "
                         "applyFullQuantumFourierTransform, 64 applyPauliGadget (4-12 qubits); "
                         f"{n} calcProbOfQubitOutcome + 16 calcExpecPauliStr"),
            "gates_per_step": gates,
            "reductions_per_step": reductions,
            "state_bytes": 16 * (1 << n),
            "parallelism": "single GPU",
        },
        "hbm_fraction_all_ops": (args.steps * alg_bytes / (total_ms / 1e3)) / HBM_PEAK,
        "roofline": {
            "bound": "hbm",
            "achieved": achieved / 1e9,
            "peak": HBM_PEAK / 1e9,
            "unit": "GB/s",
            "frac": achieved / HBM_PEAK,
            "traffic": traffic,
            "kernel": "k_oneTargDense<double> (applyCompMatr1)",
            "bytes_per_launch": dense_bytes,
            "avg_launch_ms": dense_avg_s * 1e3,
        },
        "cpu_baseline": cpu,
    }
    print(json.dumps(out), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
