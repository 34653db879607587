"""CPU check of the gate-fusion planner (no GPU): random queues of every
fusable gate kind, with controls and masks on local and non-local qubits,
planned as a flush plans them (default mode: commuting gates taken out of
order, runs on one or two qubits merged -- including the per-tile variant
matrices of DENSE1V -- and ops clustered by register layout; exact mode:
queue prefixes) must act on a random state exactly as the queue does, to
rounding (1e-12).  The planner runs from the built library; the state is
updated by tests/planner_check.cpp's own restatement of each op."""

import shutil
import subprocess
from pathlib import Path

import pytest

from quest_amd import GPU_LIB

ROOT = Path(__file__).resolve().parents[1]
HIPCC = shutil.which("hipcc") or "/opt/rocm/bin/hipcc"


@pytest.fixture(scope="module")
def planner_check(tmp_path_factory):
    if not Path(HIPCC).exists():
        pytest.skip("hipcc not available")
    exe = tmp_path_factory.mktemp("planner") / "planner_check"
    cmd = [HIPCC, "-std=c++17", "-O2", f"-I{ROOT / 'quest_amd' / 'csrc' / 'gpu'}", f"-I{ROOT / 'include'}",
           str(ROOT / "tests" / "planner_check.cpp"), "-o", str(exe), f"-L{GPU_LIB.parent}", "-lquest_amd_gpu", "-lhiprtc",
           f"-Wl,-rpath,{GPU_LIB.parent}"]
    subprocess.run(cmd, check=True, capture_output=True, text=True, timeout=300)
    return exe


@pytest.mark.parametrize("seed", [1, 2, 3])
def test_planned_passes_act_like_the_queue(planner_check, seed):
    out = subprocess.run([str(planner_check), "300", str(seed)], capture_output=True, text=True, timeout=120)
    assert out.returncode == 0, out.stdout + out.stderr
    assert "300 cases agree" in out.stdout


@pytest.mark.parametrize("seed", [1, 2])
def test_span_passes_act_like_the_gadgets(planner_check, seed):
    """Runs of Pauli and phase gadgets planned as span passes (cosets of the
    span of their X/Y masks, applied in coordinates) and emulated with the
    kernel's indexing: every amplitude visited once, results equal to the
    gadgets one by one to rounding, and every run as long as k dimensions
    allow."""
    out = subprocess.run([str(planner_check), "300", str(seed), "span"], capture_output=True, text=True, timeout=300)
    assert out.returncode == 0, out.stdout + out.stderr
    assert "300 cases agree" in out.stdout


def test_generated_pass_kernels_compile_and_cache_soundly(planner_check):
    """Every planned pass program also gets its generated kernel source:
    programs with equal shape keys must generate equal sources (the per-shape
    kernel cache depends on it), and a few distinct sources are compiled for
    gfx950 with hiprtc, as the library does at run time."""
    out = subprocess.run([str(planner_check), "60", "5", "jit", "4"], capture_output=True, text=True, timeout=600)
    assert out.returncode == 0, out.stdout + out.stderr
    assert "4 generated kernels compiled" in out.stdout


@pytest.fixture(scope="module")
def jit_bench_check(tmp_path_factory):
    if not Path(HIPCC).exists():
        pytest.skip("hipcc not available")
    exe = tmp_path_factory.mktemp("jitbench") / "jit_bench_check"
    cmd = [HIPCC, "-std=c++17", "-O2", f"-I{ROOT / 'quest_amd' / 'csrc' / 'gpu'}", f"-I{ROOT / 'include'}",
           str(ROOT / "tests" / "jit_bench_check.cpp"), "-o", str(exe), f"-L{GPU_LIB.parent}", "-lquest_amd_gpu",
           "-lhiprtc", f"-Wl,-rpath,{GPU_LIB.parent}"]
    subprocess.run(cmd, check=True, capture_output=True, text=True, timeout=300)
    return exe


def test_benchmark_pass_kernels_compile_without_scratch(jit_bench_check):
    """The 30-qubit benchmark family's pass programs (a few layers, the QFT
    and some gadgets), planned as the default planner plans them: every
    distinct shape's generated kernel compiles for gfx950 and uses no
    scratch memory (with its prefetch, or without it where the registers
    run out, as the library falls back)."""
    out = subprocess.run([str(jit_bench_check), "4", "8", "1000"], capture_output=True, text=True, timeout=900)
    assert out.returncode == 0, out.stdout[-3000:] + out.stderr[-2000:]
    assert "0 with scratch or failed" in out.stdout


@pytest.fixture(scope="module")
def jit_exit_check(tmp_path_factory):
    if not Path(HIPCC).exists():
        pytest.skip("hipcc not available")
    exe = tmp_path_factory.mktemp("jitexit") / "jit_exit_check"
    cmd = [HIPCC, "-std=c++17", "-O1", f"-I{ROOT / 'quest_amd' / 'csrc' / 'gpu'}", f"-I{ROOT / 'include'}",
           str(ROOT / "tests" / "jit_exit_check.cpp"), "-o", str(exe), f"-L{GPU_LIB.parent}", "-lquest_amd_gpu",
           f"-Wl,-rpath,{GPU_LIB.parent}"]
    subprocess.run(cmd, check=True, capture_output=True, text=True, timeout=300)
    return exe


@pytest.mark.parametrize("delay_ms", [100, 400, 900, 1500, 2500])
def test_exit_while_pass_kernels_compile(jit_exit_check, delay_ms):
    """A process that ends while generated pass kernels are being compiled
    (no qa_finalize) exits with status 0 and at once: the compiles run in
    qa_jitc child processes, which the exit handler kills.  (With compiles
    inside the process, such exits crashed or hung when the compiler's
    state was torn down under the compile threads.)"""
    import time
    t = time.perf_counter()
    out = subprocess.run([str(jit_exit_check), "20", str(delay_ms)], capture_output=True, text=True, timeout=60)
    wall = time.perf_counter() - t
    assert out.returncode == 0, (out.returncode, out.stdout, out.stderr[-2000:])
    assert "returning with compiles queued" in out.stdout
    assert wall < delay_ms / 1e3 + 10, wall


def test_compiler_process_replies():
    """qa_jitc, the compiler process: a kernel source gets status 0 and a
    gfx950 code object (an ELF image); a source with an error gets status 1
    and the compiler's log."""
    import struct
    jitc = GPU_LIB.parent / "qa_jitc"
    assert jitc.exists(), "quest_amd/lib/qa_jitc is not built"

    def ask(src):
        data = src.encode()
        out = subprocess.run([str(jitc)], input=struct.pack("<Q", len(data)) + data, capture_output=True,
                             timeout=120).stdout
        status, n = struct.unpack("<IQ", out[:12])
        return status, out[12:12 + n]

    status, code = ask('extern "C" __global__ void qa_jit_pass(double* a) { a[threadIdx.x] *= 2.0; }')
    assert status == 0 and code[:4] == b"\x7fELF", (status, code[:80])
    status, log = ask('extern "C" __global__ void qa_jit_pass(double* a) { undeclared_name(a); }')
    assert status == 1 and b"undeclared_name" in log, log[:400]


def test_generated_kernel_cache_dir(jit_exit_check, tmp_path):
    """With QA_JIT_CACHE_DIR set, compiled pass kernels are kept there as code
    objects (ELF images), one per kernel source, for later processes."""
    import os
    env = dict(os.environ, QA_JIT_CACHE_DIR=str(tmp_path))
    out = subprocess.run([str(jit_exit_check), "4", "20000"], capture_output=True, text=True, timeout=120, env=env)
    assert out.returncode == 0, out.stderr[-2000:]
    files = sorted(tmp_path.glob("qa_jit_*.co"))
    assert len(files) >= 4, [f.name for f in tmp_path.iterdir()]
    assert all(f.read_bytes()[:4] == b"\x7fELF" for f in files)
    assert not list(tmp_path.glob("*.tmp*"))
