"""CPU-only checks of the host side: the C-ABI libraries load and export every
symbol their headers declare (no compute calls), the circuit tooling is
deterministic and self-consistent, and the product refuses to run without a
GPU instead of falling back to the CPU."""

import ctypes
import subprocess
import re
from pathlib import Path

import numpy as np
import pytest

from quest_amd import GPU_LIB, HOST_LIB, circuits as qc

ROOT = Path(__file__).resolve().parents[1]
INCLUDE = ROOT / "include"


def declared_functions(header: Path) -> list[str]:
    """Names of the exported C functions a header declares."""
    text = header.read_text()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    text = re.sub(r"//[^\n]*", "", text)
    text = re.sub(r"^\s*#[^\n]*", "", text, flags=re.M)
    names = []
    # declarations end in ');' (definitions open a brace); C++-linkage overloads
    # (std:: containers, qcomp returns) are not C symbols
    for m in re.finditer(r"(?:^|(?<=[;{}]))\s*([A-Za-z_][\w \t\*]*?)\b([A-Za-z_]\w*)\s*\(([^;{}]*?)\)\s*;",
                         text, flags=re.S | re.M):
        ret, name, args = m.group(1).strip(), m.group(2), m.group(3)
        if "std::" in args or "std::" in ret or ret.startswith("qcomp") or "static" in ret or "typedef" in ret:
            continue
        if not ret or ret.split()[-1] in ("return", "else"):
            continue
        if name in ("if", "for", "while", "return", "sizeof", "defined"):
            continue
        names.append(name)
    return sorted(set(names))


@pytest.mark.parametrize("header,lib", [("quest_amd_gpu.h", GPU_LIB), ("quest.h", HOST_LIB),
                                        ("quest_amd_ext.h", HOST_LIB), ("quest_amd_serve.h", GPU_LIB)])
def test_library_exports_every_declared_symbol(header, lib):
    assert lib.exists(), f"{lib} is not built (python -m quest_amd.build)"
    handle = ctypes.CDLL(str(lib))
    names = declared_functions(INCLUDE / header)
    assert len(names) >= 4
    missing = [n for n in names if not hasattr(handle, n)]
    assert not missing, f"{header}: not exported by {lib.name}: {missing}"


def test_gpu_layer_exports_cover_reference_boundary():
    """Every single-node gpu_* routine on the path has a qa_* replacement."""
    names = set(declared_functions(INCLUDE / "quest_amd_gpu.h"))
    for need in ["qa_statevec_anyCtrlOneTargDenseMatr", "qa_statevec_anyCtrlTwoTargDenseMatr",
                 "qa_statevec_anyCtrlAnyTargDenseMatr", "qa_statevec_anyCtrlOneTargDiagMatr",
                 "qa_statevec_anyCtrlPauliTensorOrGadget", "qa_statevec_anyCtrlAnyTargZOrPhaseGadget",
                 "qa_statevec_anyCtrlSwap", "qa_densmatr_oneQubitDepolarising", "qa_densmatr_oneQubitDephasing",
                 "qa_statevec_calcProbOfMultiQubitOutcome", "qa_statevec_calcExpecPauliStr",
                 "qa_densmatr_calcExpecPauliStr", "qa_statevec_calcTotalProb", "qa_malloc", "qa_free"]:
        assert need in names


REF_DEFS = ["-DFLOAT_PRECISION=2", "-DCOMPILE_MPI=0", "-DCOMPILE_OPENMP=1", "-DCOMPILE_CUDA=0",
            "-DCOMPILE_CUQUANTUM=0", "-DINCLUDE_DEPRECATED_FUNCTIONS=0"]


def layout(tmp_path, reference: bool) -> dict:
    """sizeof / offsetof of every by-value quest.h struct, from
    tests/layout_probe.cpp compiled against include/quest.h or QuEST's own
    headers."""
    exe = tmp_path / ("probe_ref" if reference else "probe")
    flags = ["-DQA_LAYOUT_REFERENCE", *REF_DEFS, "-I/root/reference"] if reference else [f"-I{INCLUDE}"]
    subprocess.run(["g++", "-std=c++17", *flags, str(Path(__file__).parent / "layout_probe.cpp"), "-o", str(exe)],
                   check=True, capture_output=True, text=True)
    out = subprocess.run([str(exe)], check=True, capture_output=True, text=True).stdout
    return {" ".join(l.split()[:2]): int(l.split()[2]) for l in out.splitlines()}


@pytest.mark.skipif(not Path("/root/reference/quest/include").is_dir(),
                    reason="QuEST's headers are only present in the development container")
def test_struct_layouts_match_reference(tmp_path):
    """Every struct size and field offset of include/quest.h equals QuEST's
    (quest/include/qureg.h:49-80, matrices.h:68-230, paulis.h:53-80,
    channels.h:74-118, environment.h), compiled from both headers."""
    ours, ref = layout(tmp_path, False), layout(tmp_path, True)
    assert len(ref) >= 90
    assert ours == ref


def test_ctypes_layouts_match_header(tmp_path):
    """The ctypes mirrors in quest_amd/api.py have the header's sizes and
    field offsets."""
    from quest_amd import api as Q
    lay = layout(tmp_path, False)
    for name in ("Qureg", "CompMatr1", "CompMatr2", "CompMatr", "DiagMatr1", "DiagMatr2", "DiagMatr",
                 "FullStateDiagMatr", "PauliStr", "PauliStrSum", "SuperOp", "KrausMap", "QuESTEnv"):
        cls = getattr(Q, name)
        assert ctypes.sizeof(cls) == lay[f"sizeof {name}"], name
        for fname, _ in cls._fields_:
            key = f"offsetof {name}.{fname}"
            if key in lay:
                assert getattr(cls, fname).offset == lay[key], key


def test_splitmix64_known_answer():
    # first outputs of SplitMix64 from state 0 (Steele, Lea & Flood 2014)
    r = qc.SplitMix64(0)
    assert r.next_u64() == 0xE220A8397B1DCDAF
    assert r.next_u64() == 0x6E789E6AA1B965F4


def test_haar_unitaries_and_isometries():
    rng = qc.SplitMix64(3)
    for d in (2, 4, 8):
        u = np.array(qc.haar_unitary(rng, d))
        np.testing.assert_allclose(u @ u.conj().T, np.eye(d), atol=1e-14)
    v = np.array(qc.random_isometry(rng, 16, 4))
    np.testing.assert_allclose(v.conj().T @ v, np.eye(4), atol=1e-14)
    ks = qc.kraus_from_isometry(v.tolist(), 4, 4)
    s = sum(np.array(k).conj().T @ np.array(k) for k in ks)
    np.testing.assert_allclose(s, np.eye(4), atol=1e-14)


def test_circuit_text_round_trip():
    for c in (qc.layered(6, 2, 1, ctrl3=True, qft=True, gadgets=3, gadget_range=(2, 4)),
              qc.noisy_density(4, 5, 2), qc.variational(5, 2, 3)):
        text = qc.to_text(c, "x.bin")
        back = qc.from_text(text)
        assert qc.to_text(back, "x.bin") == text


def test_random_state_is_normalised_and_seeded():
    a = qc.random_state(8, 5)
    b = qc.random_state(8, 5)
    c = qc.random_state(8, 6)
    assert abs(np.vdot(a, a).real - 1) < 1e-14
    np.testing.assert_array_equal(a, b)
    assert not np.allclose(a, c)


def test_gate_counting():
    c = qc.layered(30, 40, 1, ctrl3=True, qft=True, gadgets=64)
    # 40 layers x (30 U1 + 15 PS2 + 1 CU1) + QFT (30 + 435 + 15) + 64 gadgets
    assert qc.count_gates(c) == 40 * 46 + 480 + 64


def test_no_cpu_fallback_without_gpu():
    """With no usable GPU, initQuESTEnv reports an input error (no silent CPU path)."""
    from quest_amd import api as Q
    if Q.gpu().qa_isGpuAvailable():
        pytest.skip("a GPU is visible here")
    with pytest.raises(Q.QuESTError, match="GPU"):
        Q.initQuESTEnv()


def test_serve_request_packing():
    """Host-side packing of serving requests (no GPU): Pauli strings to
    (X, Y, Z) masks, circuit ops to gate records, and the request family's
    shape (qubits and depth within their ranges, 8 strings each)."""
    from quest_amd import serve as S
    assert S.pauli_masks("XYZI", [0, 1, 2, 3]) == (0b0001, 0b0010, 0b0100)
    assert S.pauli_masks("zx", [5, 2]) == (1 << 2, 0, 1 << 5)
    with pytest.raises(ValueError):
        S.pauli_masks("Q", [0])
    arr, n = S.pack_gates([("RY", 3, 0.5), ("CZ", 1, 2), ("U1", 0, [[0, 1], [1, 0]]), ("OUT_PROB", 0, 0)])
    assert n == 3
    assert (arr[0].kind, arr[0].t1, arr[0].angle) == (S.SG["RY"], 3, 0.5)
    assert (arr[1].kind, arr[1].t1, arr[1].t2) == (S.SG["CZ"], 1, 2)
    assert list(arr[2].m) == [0, 0, 1, 0, 1, 0, 0, 0]
    with pytest.raises(ValueError):
        S.pack_gates([("KRAUS", [0, 1], [])])
    reqs = S.serving_requests(50, 3)
    for c in reqs:
        assert 10 <= c.num_qubits <= 24
        layers = sum(1 for op in c.ops if op[0] == "RY") // c.num_qubits
        assert 4 <= layers <= 32
        assert len(S.paulis_of(c)) == 8
    assert len({c.num_qubits for c in reqs}) > 5
