"""Build the two native libraries in-tree.

  quest_amd/lib/libquest_amd_gpu.so   HIP kernels for gfx950 + the C ABI of
                                      include/quest_amd_gpu.h
  quest_amd/lib/libquest_amd.so       the host C++ layer implementing the
                                      quest.h API (include/quest.h) on top of
                                      that C ABI only

Invoked by __graft_entry__.build() and `python -m quest_amd.build`.  Always
recompiles (objects are cheap; hipcc runs in parallel per source file).
"""

from __future__ import annotations

import concurrent.futures as cf
import os
import shutil
import subprocess
import sys
from pathlib import Path

PKG = Path(__file__).resolve().parent
ROOT = PKG.parent
CSRC = PKG / "csrc"
LIB = PKG / "lib"
OBJ = PKG / "lib" / "obj"
INCLUDE = ROOT / "include"

ROCM = os.environ.get("ROCM_PATH", "/opt/rocm")
HIPCC = shutil.which("hipcc") or f"{ROCM}/bin/hipcc"
CXX = os.environ.get("CXX", "g++")

ARCH = "gfx950"
# packed-FP32 instructions stay off: -packed-fp32-ops disables the feature
HIP_FLAGS = [
    f"--offload-arch={ARCH}",
    "-O3",
    "-std=c++17",
    "-fPIC",
    "-Wall",
    "-Wno-unused-function",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-munsafe-fp-atomics",
    f"-I{INCLUDE}",
    f"-I{CSRC / 'gpu'}",
]
HOST_FLAGS = ["-O2", "-std=c++17", "-fPIC", "-Wall", "-Wextra", "-Wno-unused-parameter", f"-I{INCLUDE}"]

GPU_LIB = LIB / "libquest_amd_gpu.so"
HOST_LIB = LIB / "libquest_amd.so"


def _run(cmd: list[str]) -> None:
    proc = subprocess.run(cmd, capture_output=True, text=True)
    if proc.returncode != 0:
        sys.stderr.write(" ".join(cmd) + "\n" + proc.stdout + proc.stderr)
        raise RuntimeError(f"build step failed: {cmd[0]} ... {cmd[-1]}")
    if proc.stderr.strip():
        # keep compiler warnings visible, minus hipcc's host-pass note about
        # the device-only target feature
        lines = [l for l in proc.stderr.splitlines() if "packed-fp32-ops" not in l and l.strip()]
        if lines:
            sys.stderr.write("\n".join(lines) + "\n")


# The complex128 tile-pass kernels (register-resident tiles, two waves per
# SIMD) are scheduled for instruction-level parallelism rather than
# occupancy, which their LDS tile fixes anyway: measured on MI355X, 30-qubit
# circuit passes 1.3% faster and 14-qubit density-matrix passes 2% faster
# (the complex64 instances, in qa_fused.hip and qa_serve.hip, measured
# slower so, and keep the default).
#
# The tile-pass kernels (qa_fused*.hip, qa_serve.hip) keep a tile's
# amplitudes in an array of registers that every op indexes with constants.
# LLVM's CFG simplification would merge the per-register-bit arms of an op
# switch (identical code up to which registers) into one block addressing
# the array through pointer phis, which forces the array into scratch
# memory; common-code sinking is therefore off for those files.
NO_SINK = ["-mllvm", "-simplifycfg-sink-common=false"]
FILE_FLAGS = {
    "qa_fused_f64.hip": ["-mllvm", "-amdgpu-sched-strategy=max-ilp", *NO_SINK],
    "qa_fused.hip": NO_SINK,
    "qa_serve.hip": NO_SINK,
}


def _compile_hip(src: Path) -> Path:
    out = OBJ / (src.stem + ".o")
    _run([HIPCC, *HIP_FLAGS, *FILE_FLAGS.get(src.name, []), "-c", str(src), "-o", str(out)])
    return out


def _compile_host(src: Path) -> Path:
    out = OBJ / (src.stem + ".host.o")
    _run([CXX, *HOST_FLAGS, "-c", str(src), "-o", str(out)])
    return out


def build(verbose: bool = True) -> None:
    OBJ.mkdir(parents=True, exist_ok=True)
    hip_srcs = sorted((CSRC / "gpu").glob("*.hip"))
    host_srcs = sorted((CSRC / "host").glob("*.cpp"))
    jobs = int(os.environ.get("MAX_JOBS", "8"))
    with cf.ThreadPoolExecutor(max_workers=max(1, min(jobs, 16))) as ex:
        hip_objs = list(ex.map(_compile_hip, hip_srcs))
        host_objs = list(ex.map(_compile_host, host_srcs))
    _run([HIPCC, f"--offload-arch={ARCH}", "-shared", "-fPIC", "-o", str(GPU_LIB), *map(str, hip_objs)])
    if host_objs:
        _run([CXX, "-shared", "-fPIC", "-o", str(HOST_LIB), *map(str, host_objs),
              f"-L{LIB}", "-lquest_amd_gpu", "-Wl,-rpath,$ORIGIN"])
    if verbose:
        print(f"built {GPU_LIB.relative_to(ROOT)} ({len(hip_objs)} HIP objects)"
              + (f", {HOST_LIB.relative_to(ROOT)} ({len(host_objs)} host objects)" if host_objs else ""))


if __name__ == "__main__":
    build()
