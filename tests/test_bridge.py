"""The accelerator bridge (quest_amd/csrc/bridge/gpu_bridge.cpp), checked
without a GPU against QuEST's own sources, which only the development
container holds (these tests skip elsewhere):

  * compiled against QuEST's headers, it defines every routine that
    quest/src/gpu/gpu_subroutines.hpp and gpu_config.hpp declare, and QuEST's
    api/core/cpu/comm objects link with it and libquest_amd_gpu.so into a
    complete library with no undefined gpu_* symbol (oracle/Makefile
    ref-amd);
  * that library runs a fixture circuit through QuEST's front end (with no
    GPU visible QuEST deploys to its CPU backend; on a GPU box
    tests/test_gpu_quest_frontend.py runs every fixture on the MI355X
    kernels through the same library).
"""

import re
import shutil
import subprocess
from pathlib import Path

import pytest

from conftest import compare_results, load_golden, rel_l2

ROOT = Path(__file__).resolve().parents[1]
ORACLE = ROOT / "oracle"
REF = Path("/root/reference")

needs_ref = pytest.mark.skipif(not (REF / "quest" / "src").is_dir(),
                               reason="QuEST's sources are only present in the development container")


@pytest.fixture(scope="module")
def built():
    subprocess.run(["make", "-s", "-j8", "-C", str(ORACLE), "ref"], check=True, capture_output=True, timeout=900)
    subprocess.run(["make", "-s", "-C", str(ORACLE), "ref-amd"], check=True, capture_output=True, timeout=900)
    return ORACLE / "_ref"


def declared_gpu_routines() -> set:
    names = set()
    for h in ("quest/src/gpu/gpu_subroutines.hpp", "quest/src/gpu/gpu_config.hpp"):
        text = (REF / h).read_text()
        text = re.sub(r"//.*", "", text)
        names |= set(re.findall(r"\b(gpu_\w+)\s*\(", text))
    return names


@needs_ref
def test_bridge_defines_every_gpu_routine(built):
    nm = shutil.which("nm")
    assert nm
    out = subprocess.run([nm, "-C", "--defined-only", str(built / "obj_amd" / "gpu_bridge.o")], capture_output=True,
                         text=True, check=True).stdout
    defined = set(re.findall(r"\b(gpu_\w+)(?:<[^>]*>)?\(", out))
    names = declared_gpu_routines()
    assert len(names) >= 80
    assert names <= defined, f"declared by QuEST but not defined by the bridge: {sorted(names - defined)}"


@needs_ref
def test_linked_library_has_no_undefined_gpu_routine(built):
    out = subprocess.run(["nm", "-C", "-D", "--undefined-only", str(built / "libQuEST_amd.so")], capture_output=True,
                         text=True, check=True).stdout
    assert not re.findall(r"\bgpu_\w+", out)
    # and every kernel-layer entry it calls is exported by libquest_amd_gpu.so
    need = set(re.findall(r"\b(qa_\w+)", out))
    have = subprocess.run(["nm", "-D", "--defined-only", str(ROOT / "quest_amd" / "lib" / "libquest_amd_gpu.so")],
                          capture_output=True, text=True, check=True).stdout
    assert need and need <= set(re.findall(r"\b(qa_\w+)", have))


@needs_ref
def test_quest_front_end_with_bridge_runs(built, tmp_path):
    from oracle import qoracle as O
    circ, init, amps, sample, results = load_golden("allops_sv_n7")
    got, res, _ = O.run_reference(circ, tmp_path, init_amps=init, driver=O.REF_AMD_DRIVER)
    assert rel_l2(got, amps) <= 1e-12
    compare_results(res, results, tol=1e-10)
