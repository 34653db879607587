"""The C++ container forms of quest.h (inline constructors and setters from
initializer lists, as QuEST v4's C++ users write them), compiled with g++
against include/quest.h and libquest_amd.so and run on the GPU:
tests/cpp_api_check.cpp applies each to a small state and checks the
amplitudes and purities by hand, and that a wrong-sized list reaches the
input-error handler."""

import shutil
import subprocess
from pathlib import Path

import pytest

from quest_amd import HOST_LIB

pytestmark = pytest.mark.gpu
ROOT = Path(__file__).resolve().parents[1]


def test_cpp_inline_constructors(tmp_path):
    gxx = shutil.which("g++")
    if not gxx:
        pytest.skip("g++ not available")
    exe = tmp_path / "cpp_api_check"
    subprocess.run([gxx, "-std=c++17", "-O1", str(ROOT / "tests" / "cpp_api_check.cpp"), f"-I{ROOT / 'include'}",
                    f"-L{HOST_LIB.parent}", "-lquest_amd", f"-Wl,-rpath,{HOST_LIB.parent}", "-o", str(exe)],
                   check=True, capture_output=True, text=True, timeout=300)
    print("[gpu] cpp api check", flush=True)
    out = subprocess.run([str(exe)], capture_output=True, text=True, timeout=120)
    assert out.returncode == 0, out.stdout + out.stderr
    assert "cpp api ok" in out.stdout
