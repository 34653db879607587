"""The allocation-free outcome-table entry point against the allocating one.

qa_statevec_calcProbsOfAllMultiQubitOutcomesNoAlloc runs the kernels and
launch shape of qa_statevec_calcProbsOfAllMultiQubitOutcomes with its
partial sums in the library's static device scratch and its results in a
caller-owned device array, so both must give bitwise-equal tables.  Two
calls in a row on one device (the second on a smaller state and table) show
that what the first call left in the shared scratch does not leak into the
second.  Everything goes through the C ABI of include/quest_amd_gpu.h.
"""

import ctypes as C

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

QA_FP32, QA_FP64 = 1, 2


@pytest.fixture(scope="module")
def lib():
    from quest_amd import api as Q
    g = Q.gpu()
    assert g.qa_init(0) == 0, g.qa_lastError()
    return g


def _err(g) -> str:
    return g.qa_lastError().decode()


def _state(g, nq: int, prec: int, seed: int):
    rng = np.random.default_rng(seed)
    v = rng.standard_normal(2 ** nq) + 1j * rng.standard_normal(2 ** nq)
    # a few very small amplitudes, so small outcome probabilities are exercised
    v[rng.integers(0, 2 ** nq, 64)] *= 1e-9
    v /= np.linalg.norm(v)
    host = np.ascontiguousarray(v.astype(np.complex128))
    amps = C.c_void_p()
    width = 16 if prec == QA_FP64 else 8
    assert g.qa_malloc(C.byref(amps), C.c_size_t(width * 2 ** nq)) == 0, _err(g)
    assert g.qa_setAmps(amps, prec, C.c_int64(0), host.ctypes.data_as(C.c_void_p), C.c_int64(2 ** nq)) == 0, _err(g)
    return amps


def _tables(g, amps, prec: int, nq: int, qubits):
    num = len(qubits)
    qs = (C.c_int * num)(*qubits)
    want = np.zeros(2 ** num)
    assert g.qa_statevec_calcProbsOfAllMultiQubitOutcomes(want.ctypes.data_as(C.c_void_p), amps, prec, nq, qs,
                                                          num) == 0, _err(g)
    dev = C.c_void_p()
    assert g.qa_malloc(C.byref(dev), C.c_size_t(8 * 2 ** num)) == 0, _err(g)
    try:
        rc = g.qa_statevec_calcProbsOfAllMultiQubitOutcomesNoAlloc(dev, amps, prec, nq, qs, num)
        assert rc == 0, _err(g)
        assert g.qa_sync() == 0, _err(g)
        got = np.zeros(2 ** num)
        assert g.qa_memcpyDtoH(got.ctypes.data_as(C.c_void_p), dev, C.c_size_t(8 * 2 ** num)) == 0, _err(g)
    finally:
        g.qa_free(dev)
    return got, want


@pytest.mark.parametrize("prec", [QA_FP64, QA_FP32])
def test_no_alloc_tables_equal_the_allocating_entry_point(lib, prec):
    g = lib
    # first call: a 7-qubit table of a 20-qubit state (the split-bin path,
    # whose partials fill the whole scratch)
    for nq, qubits, seed in ((20, [3, 17, 5, 11, 0, 19, 8], 11), (12, [9, 2, 4], 12)):
        print(f"no-alloc table: prec {prec}, {nq} qubits, table qubits {qubits}", flush=True)
        amps = _state(g, nq, prec, seed)
        try:
            got, want = _tables(g, amps, prec, nq, qubits)
        finally:
            g.qa_free(amps)
        assert np.array_equal(got, want), np.max(np.abs(got - want))
        assert abs(got.sum() - 1.0) < (1e-10 if prec == QA_FP64 else 1e-5)


def test_no_alloc_rejects_what_it_cannot_do(lib):
    g = lib
    nq = 8
    amps = _state(g, nq, QA_FP64, 3)
    dev = C.c_void_p()
    assert g.qa_malloc(C.byref(dev), C.c_size_t(8 * 2 ** nq)) == 0, _err(g)
    try:
        qs = (C.c_int * 2)(1, 1)  # repeated qubit
        assert g.qa_statevec_calcProbsOfAllMultiQubitOutcomesNoAlloc(dev, amps, QA_FP64, nq, qs, 2) != 0
        qs = (C.c_int * 1)(0)
        assert g.qa_statevec_calcProbsOfAllMultiQubitOutcomesNoAlloc(None, amps, QA_FP64, nq, qs, 1) != 0
        assert "NULL output" in _err(g)
    finally:
        g.qa_free(dev)
        g.qa_free(amps)
