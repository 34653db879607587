#!/usr/bin/env python3
"""Benchmark of the drop-in GPU path: 30-qubit FP64 statevector gates/s.

Workload (one "step"): on a 30-qubit complex128 statevector (16 GiB, resident
in HBM), 40 layers of [a Haar-random applyCompMatr1 on each of the 30 qubits,
applyTwoQubitPhaseShift(U(0,2pi)) on 15 random disjoint pairs, one 3-control
applyMultiControlledCompMatr1], then applyFullQuantumFourierTransform (30 H,
435 phase shifts, 15 swaps), then 64 applyPauliGadget on 4-12 random qubits,
then calcProbOfQubitOutcome on all 30 qubits and 16 calcExpecPauliStr.
2384 gates and 46 reductions per step, every one a separate call of the
quest.h API.  The library queues the gates and applies each maximal run
whose dense targets fit one 2^12-amplitude tile (low 4 qubits always in the
tile) as one LDS-tiled pass over the state (k_tilePass): one HBM read and
write of the 16 GiB state per pass instead of per gate.

Output: one JSON line:
  value      gates/s over the timed steps (HIP events on the library stream
             bracket them; wall clock agrees after syncQuESTEnv)
  roofline   the dominant kernel (the fused tile pass k_tilePass<double>):
             algorithmic bytes per launch 2 * 16 B * 2^30 = 32 GiB (every pass
             reads and writes the whole state once), divided by its average
             duration, measured with HIP events around every pass launch
             inside the timed steps
  unfused    the same step once more with fusion off (one kernel per gate)
  hbm_fraction_unfused_bytes  the step's algorithmic bytes counted as if
             every gate and reduction were its own pass, per second, over
             8 TB/s (above 1 when fusion saves passes)
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
FP64_PEAK = 78.6e12


def log(msg: str) -> None:
    print(msg, file=sys.stderr, flush=True)


def build_step(Q, qc, circ, q):
    """Prebuilt (callable, args) list for one step: each entry is one quest.h
    call, with its matrices and Pauli strings built beforehand."""
    calls = []
    for op in circ.ops:
        k = op[0]
        if k == "U1":
            calls.append((Q.applyCompMatr1, (q, op[1], Q.compmatr1(op[2]))))
        elif k == "PS2":
            calls.append((Q.applyTwoQubitPhaseShift, (q, op[1], op[2], op[3])))
        elif k == "CU1":
            calls.append((Q.applyMultiControlledCompMatr1, (q, Q.ints(op[1]), len(op[1]), op[2], Q.compmatr1(op[3]))))
        elif k == "FQFT":
            calls.append((Q.applyFullQuantumFourierTransform, (q,)))
        elif k == "PG":
            calls.append((Q.applyPauliGadget, (q, Q.pauli_str(op[1], op[2]), op[3])))
        elif k == "CZ":
            calls.append((Q.applyControlledPauliZ, (q, op[1], op[2])))
        elif k == "DEPOL":
            calls.append((Q.mixDepolarising, (q, op[1], op[2])))
        elif k == "KRAUS":
            import numpy as np
            km = Q.make_krausmap([np.array(m, dtype=complex) for m in op[2]])
            calls.append((Q.mixKrausMap, (q, Q.ints(op[1]), len(op[1]), km)))
        elif k == "OUT_PROB":
            calls.append((Q.calcProbOfQubitOutcome, (q, op[1], op[2])))
        elif k == "OUT_EXPEC":
            calls.append((Q.calcExpecPauliStr, (q, Q.pauli_str(op[1], op[2]))))
        elif k == "OUT_PURITY":
            calls.append((Q.calcPurity, (q,)))
        else:
            raise ValueError(k)
    return calls


def cpu_baseline(circ, n: int) -> dict | None:
    """Bounded, stratified sample of the step on the host cores, through
    QuEST's own CPU build if present, else through the oracle's C
    restatement, each segment timed on its own: the first 2 layers (their
    applyCompMatr1, applyTwoQubitPhaseShift and 3-control
    applyMultiControlledCompMatr1 calls), the QFT's first 2 Hadamards with
    their phase shifts in the order QuEST's applyQuantumFourierTransform
    issues them, 4 of the step's Pauli gadgets, and 3 of its reductions (2
    qubit probabilities, 1 Pauli expectation).  The step's time is
    estimated from the segments' rates weighted by the step's own counts
    (layer gates, QFT gates, gadgets, reductions); value = the step's gates
    over that time, so it reflects the step's mix rather than one gate
    kind."""
    import math
    from quest_amd import circuits as qc
    layer_ops, seen = [], 0
    for op in circ.ops:
        if op[0] not in ("U1", "PS2", "CU1"):
            break
        layer_ops.append(op)
        if op[0] == "CU1":
            seen += 1
            if seen == 2:
                break
    qft_ops = []
    for t in (n - 1, n - 2):
        qft_ops.append(("H", t))
        qft_ops += [("PS2", t, t - m - 1, math.pi / (1 << (m + 1))) for m in range(t)]
    gadget_ops = [op for op in circ.ops if op[0] == "PG"][:4]
    red_ops = [op for op in circ.ops if op[0] == "OUT_PROB"][:2] + [op for op in circ.ops if op[0] == "OUT_EXPEC"][:1]
    segments = [("layers", layer_ops), ("qft", qft_ops), ("gadgets", gadget_ops), ("reductions", red_ops)]
    step = {"layers": sum(1 for op in circ.ops if op[0] in ("U1", "PS2", "CU1")),
            "qft": (n + n * (n - 1) // 2 + n // 2) if any(op[0] == "FQFT" for op in circ.ops) else 0,
            "gadgets": sum(1 for op in circ.ops if op[0] == "PG"),
            "reductions": sum(1 for op in circ.ops if op[0].startswith("OUT_"))}
    step_gates = step["layers"] + step["qft"] + step["gadgets"]
    threads = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
    driver = ROOT / "oracle" / "_ref" / "circuit_driver"
    secs = {}
    kind = "port"
    if driver.exists():
        lines = [f"QUREG {n} 0", "INITPLUS"]
        for name, ops in segments:
            lines += ["TIMER_START"] + [qc.op_to_line(op) for op in ops] + [f"TIMER_STOP {name}"]
        with tempfile.NamedTemporaryFile("w", suffix=".txt", delete=False) as f:
            f.write("\n".join(lines) + "\n")
            path = f.name
        env = dict(os.environ, OMP_NUM_THREADS=str(threads), OMP_PROC_BIND="spread", OMP_PLACES="cores",
                   OMP_DYNAMIC="false")
        try:
            out = subprocess.run([str(driver), path], capture_output=True, text=True, env=env, timeout=900)
        finally:
            os.unlink(path)
        if out.returncode == 0:
            for l in out.stdout.splitlines():
                if l.startswith("T "):
                    secs[l.split()[1]] = float(l.split()[2])
            kind = "reference"
        else:
            log(f"reference CPU build failed ({out.returncode}); falling back to the oracle port")
    if len(secs) != len(segments):
        from oracle import qoracle as O
        st = O.State(n)
        st.a[:] = 1 / (2 ** (n / 2))
        secs, kind = {}, "port"
        for name, ops in segments:
            t0 = time.perf_counter()
            for op in ops:
                O.apply_op(st, op, [])
            secs[name] = time.perf_counter() - t0
    est = sum(step[name] * secs[name] / max(1, len(ops)) for name, ops in segments)
    sample = (f"{n}-qubit FP64 statevector, segments timed apart: 2 layers ({len(layer_ops)} gates, "
              f"{secs['layers']:.2f} s), QFT start ({len(qft_ops)} gates, {secs['qft']:.2f} s), "
              f"{len(gadget_ops)} Pauli gadgets ({secs['gadgets']:.2f} s), {len(red_ops)} reductions "
              f"({secs['reductions']:.2f} s); step time estimated from the segment rates at the step's counts "
              f"({step['layers']} layer gates, {step['qft']} QFT gates, {step['gadgets']} gadgets, "
              f"{step['reductions']} reductions) = {est:.1f} s; "
              + ("QuEST v4.1.0 FP64 OpenMP build" if kind == "reference" else "C restatement (OpenMP)"))
    return {"value": step_gates / est, "unit": "gates/s", "cores": threads, "kind": kind, "sample": sample}


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


def timed_steps(Q, gpu, calls, steps: int) -> float:
    """Milliseconds between HIP events recorded on the library stream before
    the first and after the last call of `steps` steps (the stop record
    applies any queued gates first)."""
    import ctypes as C
    e0, e1 = C.c_void_p(), C.c_void_p()
    gpu.qa_eventCreate(C.byref(e0))
    gpu.qa_eventCreate(C.byref(e1))
    Q.syncQuESTEnv()
    gpu.qa_eventRecord(e0)
    for s in range(steps):
        for f, a in calls:
            f(*a)
        log(f"step {s} enqueued")
    gpu.qa_eventRecord(e1)
    Q.syncQuESTEnv()
    ms = C.c_double()
    gpu.qa_eventElapsedMs(e0, e1, C.byref(ms))
    gpu.qa_eventDestroy(e0)
    gpu.qa_eventDestroy(e1)
    return ms.value


def serve_bench(args) -> int:
    """Serving workload: `args.circuits` variational requests (qubits
    U{10..24}, depth U{4..32}, RY+RZ on every qubit then a CZ ladder per
    layer, 8 Pauli expectations each) arriving as a Poisson stream at
    `args.rate` per second, served by the batched complex64 engine;
    compared with running a sample of the same requests one at a time as
    complex128 quregs through the quest.h API, one kernel per gate."""
    import ctypes as C
    import threading

    import numpy as np

    from quest_amd import api as Q, circuits as qc, serve as S

    Q.ensure_env()
    reqs = S.serving_requests(args.circuits, args.seed)
    packed = []
    for c in reqs:
        gates, ng = S.pack_gates(c.ops)
        masks = [S.pauli_masks(ch, t) for ch, t in S.paulis_of(c)]
        xs = (C.c_uint64 * len(masks))(*[m[0] for m in masks])
        ys = (C.c_uint64 * len(masks))(*[m[1] for m in masks])
        zs = (C.c_uint64 * len(masks))(*[m[2] for m in masks])
        packed.append((S.ServeCircuit(c.num_qubits, ng, gates, len(masks), xs, ys, zs), gates, xs, ys, zs))
    rng = np.random.default_rng(args.seed)
    arrivals = np.cumsum(rng.exponential(1.0 / args.rate, size=len(reqs)))
    total_gates = sum(qc.count_gates(c) for c in reqs)
    log(f"serve: {len(reqs)} requests, {total_gates} gates, Poisson rate {args.rate}/s "
        f"(last arrival at {arrivals[-1]:.2f} s)")
    lib = S._lib()
    eng = S.Engine(args.arena_gib << 30, max_active=args.max_active, tile_qubits=args.tile, low_qubits=4,
                   precision=1)
    eng.set_graphs(bool(args.graphs))
    # warm the engine's kernels on a few requests outside the measurement
    for c in reqs[:4]:
        eng.wait(eng.submit(c.num_qubits, c.ops, S.paulis_of(c)))
    ids = [0] * len(reqs)
    t0 = time.perf_counter()

    def feeder():
        for i, (circ, *_keep) in enumerate(packed):
            delay = t0 + arrivals[i] - time.perf_counter()
            if delay > 0:
                time.sleep(delay)
            rid = C.c_int64()
            rc = lib.qa_serveSubmit(eng._h, C.byref(circ), C.byref(rid))
            if rc != 0:
                raise RuntimeError(lib.qa_lastError().decode())
            ids[i] = rid.value

    th = threading.Thread(target=feeder)
    th.start()
    th.join()
    results, lat = [], []
    out = np.zeros(8)
    L = C.c_double()
    for rid in ids:
        rc = lib.qa_serveWait(eng._h, C.c_int64(rid), out.ctypes.data_as(C.POINTER(C.c_double)), C.byref(L))
        if rc != 0:
            raise RuntimeError(lib.qa_lastError().decode())
        results.append(out.copy())
        lat.append(L.value)
    wall = time.perf_counter() - t0
    st = eng.stats()
    st["graphs"] = eng.graph_stats()
    eng.close()
    lat = np.array(lat)
    value = len(reqs) / wall
    log(f"served {len(reqs)} in {wall:.2f} s = {value:.1f} circuits/s; latency p50 {np.percentile(lat, 50):.2f} ms, "
        f"p99 {np.percentile(lat, 99):.2f} ms; engine {st}")

    # baseline: a stratified sample of the same requests, one at a time,
    # complex128 quregs through quest.h with fusion off (one kernel per gate)
    order = sorted(range(len(reqs)), key=lambda i: reqs[i].num_qubits)
    sample = [order[int(j * (len(order) - 1) / (args.baseline_sample - 1))] for j in range(args.baseline_sample)]
    fusion = Q.get_fusion()
    Q.set_fusion(False)
    base_t, max_err = 0.0, 0.0
    for i in sample:
        c = reqs[i]
        q = Q.createQureg(c.num_qubits)
        Q.syncQuESTEnv()
        t = time.perf_counter()
        _, res = qc.run(c, q)
        Q.syncQuESTEnv()
        base_t += time.perf_counter() - t
        Q.destroyQureg(q)
        want = np.array([v for (lab, v) in res if lab == "expec"])
        max_err = max(max_err, float(np.max(np.abs(results[i] - want))))
    Q.set_fusion(bool(fusion[0]), *fusion[1:])
    # the sample's mean time per request, weighted to the full stream's mix:
    # every request of the stream is represented by its nearest-size sample
    mean_sample = base_t / len(sample)
    baseline = 1.0 / mean_sample
    log(f"baseline (sequential complex128, one kernel per gate): {len(sample)} requests in {base_t:.2f} s "
        f"= {baseline:.2f} circuits/s; max |expectation error| complex64 engine vs complex128: {max_err:.2e}")
    print(json.dumps({
        "metric": "batched circuits/s + p99 latency",
        "value": value,
        "unit": "circuits/s",
        "n_gpus": 1,
        "higher_is_better": True,
        "dtype": "c64 amplitudes, f64 reductions",
        "data": "synthetic",
        "config": {"workload": (f"{len(reqs)} variational requests, qubits U{{10..24}}, depth U{{4..32}}, "
                                "RY+RZ on every qubit then a CZ ladder per layer, 8 Pauli expectations each, "
                                f"Poisson arrivals at {args.rate}/s"),
                   "arena_gib": args.arena_gib, "max_active": args.max_active, "tile_qubits": args.tile,
                   "graphs": bool(args.graphs)},
        "latency_ms": {"p50": float(np.percentile(lat, 50)), "p99": float(np.percentile(lat, 99)),
                       "max": float(lat.max())},
        "gates_per_s": total_gates / wall,
        "engine": st,
        "baseline": {"value": baseline, "unit": "circuits/s",
                     "kind": "sequential complex128 quregs through quest.h, one kernel per gate",
                     "sample": f"{len(sample)} requests spread over the qubit-count range, {base_t:.2f} s"},
        "speedup_vs_baseline": value / baseline,
        "max_abs_expectation_error": max_err,
    }), flush=True)
    return 0


def precision_bench(args) -> int:
    """Precision form: the plain-method circuit family (20 layers of a Haar
    applyCompMatr1 on every qubit then applyTwoQubitPhaseShift on n/2
    disjoint random pairs; probabilities of every qubit and one Pauli
    expectation) on a complex64 statevector at n qubits (default 33, 64 GiB)
    against the same circuit on complex128 at n qubits (128 GiB); then the
    error of complex64 against complex128 at 30 qubits on the same initial
    state: state infidelity and the largest absolute error of the outputs."""
    from quest_amd import api as Q, circuits as qc
    import numpy as np

    Q.ensure_env()
    n = args.qubits if args.qubits != 30 else 33
    depth = args.depth if args.depth != 40 else 20

    def family(m):
        return qc.layered(m, depth, args.seed, ctrl3=False, qft=False, gadgets=0, num_expec=1, probs=True)

    def run_calls(calls):
        return [f(*a) for f, a in calls]

    gpu = Q.gpu()
    import ctypes as C
    circ = family(n)
    gates = qc.count_gates(circ)
    rates = {}
    for label, prec, blocks in (("complex64_mfma_blocks", 1, args.block_min_gates), ("complex64", 1, 0),
                                ("complex128", 2, 0)):
        gpu.qa_setFusionBlocks(int(blocks))
        q = Q.createCustomPrecisionQureg(n, 0, prec)
        Q.initRandomPureState(q)
        calls = build_step(Q, qc, circ, q)
        run_calls(calls)
        Q.syncQuESTEnv()
        jit_ready(Q, args)
        gpu.qa_setFusionTiming(1)
        ms = timed_steps(Q, gpu, calls, 1)
        passes, fg, pms, pfl = C.c_int64(), C.c_int64(), C.c_double(), C.c_double()
        gpu.qa_getFusionStats(C.byref(passes), C.byref(fg), C.byref(pms), C.byref(pfl))
        gpu.qa_setFusionTiming(0)
        gpu.qa_setFusionBlocks(0)
        Q.destroyQureg(q)
        amp = 8 if prec == 1 else 16
        avg = pms.value / max(1, passes.value)
        rates[label] = {"gates_per_s": gates / (ms / 1e3), "ms_per_step": ms, "passes": passes.value,
                        "pass_avg_ms": avg,
                        "pass_GBps": 2 * amp * (1 << n) / (avg * 1e-3) / 1e9 if passes.value else None}
        log(f"{n} qubits {label}: {ms:.1f} ms = {rates[label]['gates_per_s']:.1f} gates/s, "
            f"{passes.value} passes, {avg:.2f} ms each")

    # error at 30 qubits: same initial state, same circuit, both precisions
    m = 30
    c30 = family(m)
    q64 = Q.createCustomPrecisionQureg(m, 0, 2)
    q32 = Q.createCustomPrecisionQureg(m, 0, 1)
    Q.initRandomPureState(q64)
    Q.setQuregToClone(q32, q64)
    out64 = [v for v in run_calls(build_step(Q, qc, c30, q64)) if v is not None]
    gpu.qa_setFusionBlocks(int(args.block_min_gates))
    out32 = [v for v in run_calls(build_step(Q, qc, c30, q32)) if v is not None]
    gpu.qa_setFusionBlocks(0)
    # the complex64 result widened to complex128, then <psi64|psi32> on the GPU
    q32w = Q.createCustomPrecisionQureg(m, 0, 2)
    Q.setQuregToClone(q32w, q32)
    Q.destroyQureg(q32)
    ip = Q.inner_product(q64, q32w)
    # fidelity of the normalised states (the complex64 norm drifts by rounding)
    norm32 = Q.calcTotalProb(q32w)
    norm_err = abs(norm32 - 1.0)
    infidelity = 1.0 - abs(ip) ** 2 / (Q.calcTotalProb(q64) * norm32)
    Q.destroyQureg(q64)
    Q.destroyQureg(q32w)
    err_out = float(np.max(np.abs(np.array(out32) - np.array(out64))))
    log(f"30 qubits: infidelity {infidelity:.2e}, complex64 norm error {norm_err:.2e}, max |output error| "
        f"complex64 vs complex128 = {err_out:.2e}")
    print(json.dumps({
        "metric": f"{n}-qubit complex64 gates/s (fused), against complex128; 5-qubit FP32 MFMA blocks beside",
        "value": max(rates["complex64"]["gates_per_s"], rates["complex64_mfma_blocks"]["gates_per_s"]),
        "unit": "gates/s", "n_gpus": 1, "higher_is_better": True,
        "dtype": "c64 amplitudes, f64 reductions", "data": "synthetic",
        "config": {"workload": (f"{n}-qubit statevector, {depth} layers x ({n} Haar applyCompMatr1 + {n // 2} "
                                f"applyTwoQubitPhaseShift); {n} calcProbOfQubitOutcome + 1 calcExpecPauliStr"),
                   "gates_per_step": gates},
        "block_min_gates": args.block_min_gates,
        **rates,
        "error_at_30_qubits": {"state_infidelity": infidelity, "norm_error": norm_err, "max_abs_output_error": err_out,
                               "reference": "the same circuit on a complex128 qureg from the same initial state"},
    }), flush=True)
    return 0


def apply_fusion_args(Q, args) -> bool:
    """Sets the fusion configuration the options name; True when fusion is on."""
    fused = args.fusion != "off"
    if fused:
        tile, low, queue = (int(x) for x in args.fusion.split(","))
        Q.set_fusion(True, tile, low, queue)
    else:
        Q.set_fusion(False)
    Q.gpu().qa_setFusionPrefetch(int(args.prefetch))
    Q.set_fusion_exact(bool(args.exact))
    Q.set_fusion_span(int(args.span))
    # pass kernels generated at run time: every program shape is compiled in
    # the background at first sight (warmup steps run in the interpreter
    # meanwhile; jit_ready() waits for the compiles before the timed steps)
    Q.set_fusion_jit(int(args.jit), 1)
    return fused


def jit_ready(Q, args) -> dict | None:
    """Waits for the queued pass-kernel compiles (outside the timed region)
    and returns the JIT counters so far."""
    if not args.jit:
        return None
    t = time.perf_counter()
    Q.fusion_jit_wait()
    st = Q.fusion_jit_stats()
    st["wait_seconds"] = time.perf_counter() - t
    log(f"jit: {st['kernels']} pass kernels compiled ({st['compile_seconds']:.1f} s of compile threads, "
        f"waited {st['wait_seconds']:.1f} s after warmup), {st['failures']} failed")
    return st


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--rate", type=float, default=2500.0, help="serve mode: Poisson arrival rate (1/s)")
    ap.add_argument("--circuits", type=int, default=4096, help="serve mode: number of requests")
    ap.add_argument("--arena-gib", type=int, default=48, help="serve mode: HBM arena of the engine")
    ap.add_argument("--max-active", type=int, default=256, help="serve mode: requests running at once")
    ap.add_argument("--tile", type=int, default=13, help="serve mode: tile qubits")
    ap.add_argument("--baseline-sample", type=int, default=24, help="serve mode: requests run sequentially")
    ap.add_argument("--graphs", type=int, default=0,
                    help="serve mode: launch each engine step as one captured HIP graph (1) or launch by launch (0)")
    ap.add_argument("--mode", choices=["circuit", "density", "serve", "precision", "sweep"], default="circuit",
                    help="circuit: the headline 30-qubit statevector workload; density: the 14-qubit noisy "
                         "density-matrix workload; sweep: per-target dense-kernel timing + copy calibration")
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--qubits", type=int, default=30)
    ap.add_argument("--depth", type=int, default=40)
    ap.add_argument("--seed", type=int, default=2024)
    ap.add_argument("--fusion", default=None,
                    help="tileQubits,lowQubits,maxQueue of the gate-fusion queue, or 'off' (default: 13,3,2048 "
                         "for the complex128 modes, 12,4,2048 for the complex64 precision mode, the shapes "
                         "measured fastest for each)")
    ap.add_argument("--no-unfused", action="store_true", help="skip the one-step unfused comparison")
    ap.add_argument("--prefetch", type=int, default=2,
                    help="tile passes prefetch the next tile: 2 loads spread over the program, 3 as 2 with the "
                         "stores deferred a tile and spread too, 1 at the tile's start, 0 not")
    ap.add_argument("--exact", type=int, default=0,
                    help="exact fusion (1): gates in queue order with single-gate arithmetic; 0: passes reorder "
                         "commuting gates and merge runs on the same one or two qubits into one matrix")
    ap.add_argument("--block-min-gates", type=int, default=4,
                    help="precision mode: runs of at least this many gates on 5 qubits become FP32 MFMA blocks")
    ap.add_argument("--jit", type=int, default=1,
                    help="pass kernels generated at run time for each program shape (1) or every pass in the "
                         "interpreter kernel (0)")
    ap.add_argument("--span", type=int, default=12,
                    help="span passes for runs of Pauli/phase gadgets over cosets of 2^SPAN amplitudes "
                         "(10..13), 0 = off")
    ap.add_argument("--tune", type=int, default=0,
                    help="run the fusion autotuner (qa_tuneFusion) on one step's gates before the warmup and use "
                         "its choice of tile, low and span qubits")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--calibration-copy", action="store_true",
                    help="copy the state once (one device copy dispatch of known size) before the timed steps, "
                         "as the calibration point of rocprofv3 FETCH_SIZE / WRITE_SIZE runs")
    args = ap.parse_args()
    if args.fusion is None:
        args.fusion = "12,4,2048" if args.mode == "precision" else "13,3,2048"

    if args.gpus != 1:
        print(f"bench.py measures one MI355X: --gpus {args.gpus} is not supported by this single-GPU build "
              "(a multi-GPU statevector would shard by its highest qubits)", file=sys.stderr)
        return 2

    if args.mode == "serve":
        return serve_bench(args)

    from quest_amd import api as Q, circuits as qc

    if args.mode == "precision":
        Q.ensure_env()
        apply_fusion_args(Q, args)
        return precision_bench(args)

    n = args.qubits
    if args.mode == "sweep":
        Q.ensure_env()
        Q.set_fusion(False)
        kernel_sweep(Q, n)
        return 0
    dens = args.mode == "density"
    if dens:
        # 14-qubit density matrix: depth 20 of [Haar applyCompMatr1 + mixDepolarising on every qubit, CZ +
        # mixDepolarising on both qubits of 7 random disjoint pairs], a 2-qubit mixKrausMap every 5 layers
        n = args.qubits if args.qubits != 30 else 14
        depth = args.depth if args.depth != 40 else 20
        circ = qc.noisy_density(n, depth, args.seed, kraus_every=5, num_expec=8)
    else:
        circ = qc.layered(n, args.depth, args.seed, ctrl3=True, qft=True, gadgets=64, gadget_range=(4, 12),
                          num_expec=16, probs=True)
    gates = qc.count_gates(circ)
    if dens:
        # gates and channels, each one quest.h call
        gates += sum(1 for op in circ.ops if op[0] in ("DEPOL", "KRAUS"))
    reductions = sum(1 for op in circ.ops if op[0].startswith("OUT_"))
    # algorithmic HBM bytes of the step with every gate and reduction as its
    # own pass (the unfused formula per op)
    alg_bytes = sum(qc.gate_bytes(op, n, dens) for op in circ.ops)

    Q.ensure_env()
    fused = apply_fusion_args(Q, args)
    fusion_cfg = Q.get_fusion()
    env = Q.getQuESTEnv()
    log(f"bench: {n} qubits, {gates} gates + {reductions} reductions per step, GPU={env.isGpuAccelerated}, "
        f"fusion={fusion_cfg}")
    q = Q.createDensityQureg(n) if dens else Q.createQureg(n)
    Q.initRandomPureState(q)
    calls = build_step(Q, qc, circ, q)
    gpu = Q.gpu()
    import ctypes as C

    if args.calibration_copy:
        Q.destroyQureg(Q.createCloneQureg(q))
        Q.syncQuESTEnv()

    tuner = None
    if args.tune and fused:
        # the fusion autotuner on one step's gates (queued whole, then timed
        # under candidate shapes on a scratch copy of the state)
        t = time.perf_counter()
        os.environ["QA_TUNE_VERBOSE"] = "1"
        Q.set_fusion(True, fusion_cfg[1], fusion_cfg[2], 4096)
        gate_ops = [op for op in circ.ops if not op[0].startswith("OUT_")]
        for f, a in build_step(Q, qc, qc.Circuit(circ.num_qubits, circ.dens, gate_ops, circ.init, circ.name), q):
            f(*a)
        tuner = Q.tune_fusion(q, retune=True)
        cfg = Q.get_fusion()
        Q.set_fusion(True, cfg[1], cfg[2], fusion_cfg[3])
        fusion_cfg = Q.get_fusion()
        args.span = tuner["best"]["span_qubits"]
        tuner["seconds"] = time.perf_counter() - t
        for tr in tuner["trials"]:
            log(f"tuner: tile {tr['tile_qubits']} low {tr['low_qubits']} span {tr['span_qubits']} "
                f"blocks {tr['block_min_gates']}: {tr['ms']:.1f} ms")
        log(f"tuner: chose {tuner['best']} in {tuner['seconds']:.1f} s")

    for w in range(args.warmup):
        t = time.perf_counter()
        for f, a in calls:
            f(*a)
        Q.syncQuESTEnv()
        log(f"warmup step {w}: {time.perf_counter() - t:.2f} s")

    jit0 = jit_ready(Q, args)

    # timed region: every tile pass is bracketed by its own pair of HIP
    # events on the library stream (qa_setFusionTiming)
    gpu.qa_setFusionTiming(1)
    t0 = time.perf_counter()
    total_ms = timed_steps(Q, gpu, calls, args.steps)
    wall = time.perf_counter() - t0
    passes, fgates, pass_ms, pass_flops = C.c_int64(), C.c_int64(), C.c_double(), C.c_double()
    gpu.qa_getFusionStats(C.byref(passes), C.byref(fgates), C.byref(pass_ms), C.byref(pass_flops))
    prog_ops, layouts = C.c_int64(), C.c_int64()
    gpu.qa_getFusionProgramStats(C.byref(prog_ops), C.byref(layouts))
    span = Q.fusion_span_stats()
    gpu.qa_setFusionTiming(0)
    jit = None
    if jit0 is not None:
        j1 = Q.fusion_jit_stats()
        jit = {"kernels": j1["kernels"], "failures": j1["failures"],
               "compile_thread_seconds": round(j1["compile_seconds"], 2),
               "wait_after_warmup_seconds": round(jit0["wait_seconds"], 2),
               "timed_passes_in_generated_kernels": j1["launches"] - jit0["launches"]}
    value = args.steps * gates / (total_ms / 1e3)
    state = 16 * (1 << (2 * n if dens else n))
    pass_bytes = 2 * state
    log(f"timed: {total_ms:.1f} ms (wall {wall * 1e3:.1f} ms) for {args.steps} steps = {value:.1f} gates/s")

    # tile passes (the dominant kernel) and span passes (runs of gadgets) apart
    tile_passes = passes.value - span["passes"]
    tile_ms = pass_ms.value - span["pass_ms"]
    roofline = None
    if tile_passes:
        avg_s = tile_ms / 1e3 / tile_passes
        achieved = pass_bytes / avg_s
        log(f"tile passes: {tile_passes} ({fgates.value - span['gates']} gates, "
            f"{(fgates.value - span['gates']) / tile_passes:.2f} per pass), avg {avg_s * 1e3:.3f} ms = "
            f"{achieved / 1e9:.0f} GB/s; passes are {100 * pass_ms.value / total_ms:.1f}% of the timed region; "
            f"per pass {(prog_ops.value - span['gates']) / tile_passes:.1f} program ops, "
            f"{layouts.value / tile_passes:.2f} register-layout changes")
        traffic = None
        tfile = ROOT / "profiles" / f"tile_pass_{n}q_fp64_traffic.json"
        if tfile.exists():
            traffic = json.loads(tfile.read_text()).get("traffic_bytes_per_dispatch")
        roofline = {
            "bound": "hbm",
            "achieved": achieved / 1e9,
            "peak": HBM_PEAK / 1e9,
            "unit": "GB/s",
            "frac": achieved / HBM_PEAK,
            "traffic": traffic,
            "kernel": ("tile pass (qa_jit_pass generated per program shape, else k_tilePass<double>)"
                       if args.jit else "k_tilePass<double> (fused gate pass)"),
            "bytes_per_launch": pass_bytes,
            "avg_launch_ms": avg_s * 1e3,
            "launches": tile_passes,
            "program_ops_per_launch": (prog_ops.value - span["gates"]) / tile_passes,
            "layout_changes_per_launch": layouts.value / tile_passes,
            "share_of_step_time": tile_ms / total_ms,
            # the same launches against the FP64 vector peak: the gates a
            # pass applies are FP64 arithmetic on the registers
            "fp64_TFLOPs": pass_flops.value / (pass_ms.value / 1e3) / 1e12,
            "fp64_frac": pass_flops.value / (pass_ms.value / 1e3) / FP64_PEAK,
        }
    span_out = None
    if span["passes"]:
        avg_s = span["pass_ms"] / 1e3 / span["passes"]
        span_out = {"kernel": "k_spanPass<double>", "launches": span["passes"], "gates": span["gates"],
                    "avg_launch_ms": avg_s * 1e3, "achieved_GBps": pass_bytes / avg_s / 1e9,
                    "frac": pass_bytes / avg_s / HBM_PEAK, "share_of_step_time": span["pass_ms"] / total_ms}
        log(f"span passes: {span['passes']} ({span['gates']} gates), avg {avg_s * 1e3:.3f} ms = "
            f"{pass_bytes / avg_s / 1e9:.0f} GB/s")

    unfused = None
    if fused and not args.no_unfused and not dens:
        Q.set_fusion(False)
        ums = timed_steps(Q, gpu, calls, 1)
        unfused = {"value": gates / (ums / 1e3), "ms_per_step": ums}
        log(f"unfused: {ums:.1f} ms per step = {unfused['value']:.1f} gates/s")
        Q.set_fusion(True, *fusion_cfg[1:])

    if fused and not args.no_unfused and dens:
        Q.set_fusion(False)
        ums = timed_steps(Q, gpu, calls, args.steps)
        unfused = {"value": args.steps * gates / (ums / 1e3), "ms_per_step": ums / args.steps}
        log(f"unfused: {ums / args.steps:.1f} ms per step = {unfused['value']:.1f} ops/s")
        Q.set_fusion(True, *fusion_cfg[1:])
    cpu = None if (args.no_cpu_baseline or dens) else cpu_baseline(circ, n)
    Q.destroyQureg(q)

    out = {
        "metric": METRIC if not dens else "14-qubit FP64 density-matrix ops/s (gates and channels)",
        "value": value,
        "unit": "gates/s" if not dens else "ops/s",
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
            "workload": ((f"{n}-qubit complex128 density matrix: {depth} layers x ({n} Haar applyCompMatr1 and "
                          f"{n} mixDepolarising, {n // 2} CZ and {2 * (n // 2)} mixDepolarising), a 2-qubit "
                          "4-operator mixKrausMap every 5 layers; calcPurity + 8 calcExpecPauliStr + "
                          f"{n} calcProbOfQubitOutcome") if dens else
                         (f"{n}-qubit complex128 statevector: {args.depth} layers x ({n} Haar applyCompMatr1 + "
                          f"{n // 2} applyTwoQubitPhaseShift + 1 3-control applyMultiControlledCompMatr1), "
                          "applyFullQuantumFourierTransform, 64 applyPauliGadget (4-12 qubits); "
                          f"{n} calcProbOfQubitOutcome + 16 calcExpecPauliStr")),
            ("ops_per_step" if dens else "gates_per_step"): gates,
            "reductions_per_step": reductions,
            "state_bytes": state,
            "parallelism": "single GPU",
            "fusion": ({"tile_qubits": fusion_cfg[1], "low_qubits": fusion_cfg[2], "max_queue": fusion_cfg[3],
                        "exact": bool(args.exact), "span_qubits": args.span,
                        "passes_per_step": passes.value / args.steps,
                        "gates_per_pass": (fgates.value / passes.value) if passes.value else None}
                       if fused else "off"),
        },
        "jit": jit,
        "tuner": tuner,
        "hbm_fraction_unfused_bytes": (args.steps * alg_bytes / (total_ms / 1e3)) / HBM_PEAK,
        "roofline": roofline,
        "span_passes": span_out,
        "unfused": unfused,
        "cpu_baseline": cpu,
    }
    print(json.dumps(out), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
