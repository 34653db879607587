"""Pass kernels generated at run time must do exactly what the interpreter
kernel does: each circuit runs once with every pass in the interpreter
(queuing a compile for each program shape), then, once the compiles are
done, again with every pass in its generated kernel; amplitudes and
results must be bitwise equal, and every pass of the second run must have
used a generated kernel.

The circuits cover every op kind a pass program holds (dense one- and
two-qubit, variant and multi-gate one-qubit ops, diagonals and PHASES runs,
swaps, Pauli tensors and gadgets in registers and in LDS, phase gadgets,
depolarising and dephasing), controls on register, tile and non-local
bits, exact and reordering planners, complex128 and complex64 (with FP32
MFMA blocks), statevectors and density matrices.
"""

import sys

import numpy as np
import pytest

from quest_amd import circuits as qc
from test_gpu_fusion import fusion_mix

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


@pytest.fixture(autouse=True)
def restore(api):
    yield
    api.set_fusion(True)
    api.set_fusion_exact(False)
    api.gpu().qa_setFusionBlocks(0)
    api.set_fusion_jit(1, 2)


def run(api, c, shape, exact, precision, blocks):
    api.set_fusion(True, *shape)
    api.set_fusion_exact(exact)
    api.gpu().qa_setFusionBlocks(blocks)
    q = api.createCustomPrecisionQureg(c.num_qubits, int(c.dens), precision)
    try:
        q, res = qc.run(c, q)
        return api.get_amps(q), res
    finally:
        api.destroyQureg(q)


CASES = [
    # (name, circuit factory, tile shape, exact, precision, MFMA block minimum)
    ("mix_sv12_exact", lambda: fusion_mix(12, 160, 7), (6, 2, 2048), True, 2, 0),
    ("mix_sv12", lambda: fusion_mix(12, 160, 8), (6, 2, 2048), False, 2, 0),
    ("mix_sv13_tile12", lambda: fusion_mix(13, 200, 9), (12, 4, 2048), False, 2, 0),
    ("mix_dm5", lambda: fusion_mix(5, 120, 10, dens=True), (6, 2, 2048), False, 2, 0),
    ("mix_dm5_exact", lambda: fusion_mix(5, 120, 11, dens=True), (8, 3, 2048), True, 2, 0),
    ("mix_sv12_fp32_blocks", lambda: fusion_mix(12, 160, 12, flushers=False), (10, 3, 2048), False, 1, 2),
    ("layered_sv20_bench_shape",
     lambda: qc.layered(20, 4, 5, ctrl3=True, qft=True, gadgets=8, gadget_range=(4, 12), num_expec=2),
     (13, 3, 2048), False, 2, 0),
    ("noisy_dm7", lambda: qc.noisy_density(7, 4, 6, kraus_every=2, num_expec=2), (12, 4, 2048), False, 2, 0),
]


@pytest.mark.parametrize("name,make,shape,exact,precision,blocks", CASES, ids=[c[0] for c in CASES])
def test_generated_kernels_match_the_interpreter(api, name, make, shape, exact, precision, blocks):
    c = make()
    say(f"jit case {name}: {c.num_qubits} qubits, {len(c.ops)} ops, tile {shape}, exact {exact}, "
        f"precision {precision}, blocks {blocks}")
    api.set_fusion_jit(1, 1)  # every shape compiled in the background at first sight
    api.fusion_stats()
    before = api.fusion_jit_stats()
    want, wres = run(api, c, shape, exact, precision, blocks)
    api.fusion_jit_wait()
    mid = api.fusion_jit_stats()
    assert mid["failures"] == before["failures"], mid["last_error"]
    passes = api.fusion_stats()["passes"]
    api.fusion_span_stats()
    got, gres = run(api, c, shape, exact, precision, blocks)
    after = api.fusion_jit_stats()
    passes2 = api.fusion_stats()["passes"]
    # span passes (runs of gadgets) have a kernel of their own
    spans = api.fusion_span_stats()["passes"]
    say(f"  {passes} passes; {mid['kernels'] - before['kernels']} kernels compiled "
        f"({mid['compile_seconds'] - before['compile_seconds']:.1f} s of compile threads); "
        f"second run {after['launches'] - mid['launches']} of {passes2 - spans} tile passes in generated kernels "
        f"({spans} span passes)")
    assert passes2 == passes and passes > 0
    assert after["launches"] - mid["launches"] == passes2 - spans
    assert np.array_equal(got, want), f"max |diff| {np.max(np.abs(got - want))}"
    assert len(gres) == len(wres)
    for (lg, vg), (lw, vw) in zip(gres, wres):
        assert lg == lw
        assert np.array_equal(np.atleast_1d(vg), np.atleast_1d(vw)), (lg, vg, vw)


def test_jit_settings_validation(api):
    g = api.gpu()
    assert g.qa_setFusionJit(3, 1) != 0
    assert g.qa_setFusionJit(1, 0) != 0
    api.set_fusion_jit(0)
    import ctypes as C
    m, s = C.c_int(), C.c_int()
    g.qa_getFusionJit(C.byref(m), C.byref(s))
    assert (m.value, s.value) == (0, 2)


EXIT_SCRIPT = r"""
import sys
sys.path[:0] = [{root!r}, {tests!r}]
from quest_amd import api, circuits as qc
from test_gpu_fusion import fusion_mix
api.ensure_env()
api.set_fusion(True, 6, 2, 2048)
api.set_fusion_jit(1, 1)
c = fusion_mix(12, 160, 21)
q = api.createCustomPrecisionQureg(c.num_qubits, 0, 2)
qc.run(c, q)
print("exiting with", api.fusion_jit_stats()["kernels"], "kernels compiled", flush=True)
"""


def test_exit_with_compiles_in_flight(tmp_path):
    """A program that ends while pass kernels are still being compiled in
    the background, without qa_finalize (as a QuEST program through the
    accelerator bridge does), must exit cleanly: the compile threads are
    stopped before the HIP runtime and the compiler are torn down."""
    import os
    import subprocess
    root = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
    script = tmp_path / "exit_in_flight.py"
    script.write_text(EXIT_SCRIPT.format(root=root, tests=os.path.join(root, "tests")))
    say("jit exit case: 12-qubit fusion mix, every shape queued at first sight, exit without finalize")
    proc = subprocess.run([sys.executable, str(script)], capture_output=True, text=True, timeout=110)
    say(proc.stdout.strip()[-300:])
    assert proc.returncode == 0, proc.stderr[-2000:]
    assert "exiting with" in proc.stdout
