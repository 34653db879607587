"""CPU-only checks of the host side: the C-ABI libraries load and export every
symbol their headers declare (no compute calls), the circuit tooling is
deterministic and self-consistent, and the product refuses to run without a
GPU instead of falling back to the CPU."""

import ctypes
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


def test_struct_layouts_match_reference():
    """Sizes of the by-value structs of quest.h (qureg.h:49-80, matrices.h:68-98)."""
    from quest_amd import api as Q
    assert ctypes.sizeof(Q.Qureg) == 8 * 4 + 5 * 8 + 4 * 8
    assert ctypes.sizeof(Q.CompMatr1) == 16 + 4 * 16
    assert ctypes.sizeof(Q.CompMatr2) == 16 + 16 * 16
    assert ctypes.sizeof(Q.PauliStr) == 16


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
