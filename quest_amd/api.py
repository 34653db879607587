"""ctypes mirror of include/quest.h (the QuEST v4 API subset) and of the
kernel-layer C ABI include/quest_amd_gpu.h.

Structs are declared with the reference layouts (quest/include/qureg.h:49-80,
matrices.h:68-186, paulis.h:53-80, channels.h:74-118); functions keep the
reference names and argument order.  A QuEST input error raises
QuESTError(func, msg) instead of exiting, through setInputErrorHandler
(quest/include/debug.h), the hook the reference's own tests use to turn
validation errors into exceptions (tests/main.cpp:63-66).
"""

from __future__ import annotations

import ctypes as C
import threading

import numpy as np

from . import GPU_LIB, HOST_LIB

if not HOST_LIB.exists() or not GPU_LIB.exists():
    raise ImportError(
        f"quest_amd native libraries are not built ({HOST_LIB.name}, {GPU_LIB.name}); "
        "run `python -m quest_amd.build`")

_gpu = C.CDLL(str(GPU_LIB), mode=C.RTLD_GLOBAL)
_lib = C.CDLL(str(HOST_LIB), mode=C.RTLD_GLOBAL)

qreal = C.c_double
qindex = C.c_longlong
cint = C.c_int
cdbl = C.c_double
ptr = C.c_void_p
IntP = C.POINTER(C.c_int)
u64 = C.c_uint64


class qcomp(C.Structure):
    """std::complex<double> / double _Complex: two doubles."""
    _fields_ = [("re", C.c_double), ("im", C.c_double)]

    @classmethod
    def of(cls, z):
        z = complex(z)
        return cls(z.real, z.imag)

    def value(self) -> complex:
        return complex(self.re, self.im)


class QuESTEnv(C.Structure):
    _fields_ = [("isMultithreaded", cint), ("isGpuAccelerated", cint), ("isDistributed", cint),
                ("rank", cint), ("numNodes", cint)]


class Qureg(C.Structure):
    _fields_ = [("isMultithreaded", cint), ("isGpuAccelerated", cint), ("isDistributed", cint),
                ("rank", cint), ("numNodes", cint), ("logNumNodes", cint), ("isDensityMatrix", cint),
                ("numQubits", cint), ("numAmps", qindex), ("logNumAmps", qindex),
                ("numAmpsPerNode", qindex), ("logNumAmpsPerNode", qindex), ("logNumColsPerNode", qindex),
                ("cpuAmps", ptr), ("gpuAmps", ptr), ("cpuCommBuffer", ptr), ("gpuCommBuffer", ptr)]


class CompMatr1(C.Structure):
    _fields_ = [("numQubits", cint), ("numRows", qindex), ("elems", qcomp * 2 * 2)]


class CompMatr2(C.Structure):
    _fields_ = [("numQubits", cint), ("numRows", qindex), ("elems", qcomp * 4 * 4)]


class CompMatr(C.Structure):
    _fields_ = [("numQubits", cint), ("numRows", qindex), ("isApproxUnitary", IntP),
                ("isApproxHermitian", IntP), ("wasGpuSynced", IntP),
                ("cpuElems", C.POINTER(C.POINTER(qcomp))), ("cpuElemsFlat", C.POINTER(qcomp)),
                ("gpuElemsFlat", ptr)]


class DiagMatr1(C.Structure):
    _fields_ = [("numQubits", cint), ("numElems", qindex), ("elems", qcomp * 2)]


class DiagMatr2(C.Structure):
    _fields_ = [("numQubits", cint), ("numElems", qindex), ("elems", qcomp * 4)]


class DiagMatr(C.Structure):
    _fields_ = [("numQubits", cint), ("numElems", qindex), ("isApproxUnitary", IntP),
                ("isApproxHermitian", IntP), ("isApproxNonZero", IntP), ("isStrictlyNonNegative", IntP),
                ("wasGpuSynced", IntP), ("cpuElems", C.POINTER(qcomp)), ("gpuElems", ptr)]


class FullStateDiagMatr(C.Structure):
    _fields_ = [("numQubits", cint), ("numElems", qindex), ("isGpuAccelerated", cint),
                ("isMultithreaded", cint), ("isDistributed", cint), ("numElemsPerNode", qindex),
                ("isApproxUnitary", IntP), ("isApproxHermitian", IntP), ("isApproxNonZero", IntP),
                ("isStrictlyNonNegative", IntP), ("wasGpuSynced", IntP), ("cpuElems", C.POINTER(qcomp)),
                ("gpuElems", ptr)]


class PauliStr(C.Structure):
    _fields_ = [("lowPaulis", C.c_ulonglong), ("highPaulis", C.c_ulonglong)]


class PauliStrSum(C.Structure):
    _fields_ = [("numTerms", qindex), ("strings", C.POINTER(PauliStr)), ("coeffs", C.POINTER(qcomp)),
                ("isApproxHermitian", IntP)]


class SuperOp(C.Structure):
    _fields_ = [("numQubits", cint), ("numRows", qindex), ("cpuElems", C.POINTER(C.POINTER(qcomp))),
                ("cpuElemsFlat", C.POINTER(qcomp)), ("gpuElemsFlat", ptr), ("wasGpuSynced", IntP)]


class KrausMap(C.Structure):
    _fields_ = [("numQubits", cint), ("numMatrices", cint), ("numRows", qindex),
                ("matrices", C.POINTER(C.POINTER(C.POINTER(qcomp)))), ("superop", SuperOp),
                ("isApproxCPTP", IntP)]


# ---------------------------------------------------------------- errors

class QuESTError(RuntimeError):
    def __init__(self, func: str, msg: str):
        super().__init__(f"{func}: {msg}")
        self.func = func
        self.msg = msg


_err = threading.local()
_HANDLER_T = C.CFUNCTYPE(None, C.c_char_p, C.c_char_p)


def _on_error(func, msg):
    _err.last = (func.decode(), msg.decode())


_handler = _HANDLER_T(_on_error)


def _raise_pending():
    e = getattr(_err, "last", None)
    if e is not None:
        _err.last = None
        raise QuESTError(*e)


# ------------------------------------------------------------- signatures

_SIGS = {
    # environment / debug
    "initQuESTEnv": (None, []), "initCustomQuESTEnv": (None, [cint, cint, cint]),
    "finalizeQuESTEnv": (None, []), "syncQuESTEnv": (None, []), "reportQuESTEnv": (None, []),
    "isQuESTEnvInit": (cint, []), "getQuESTEnv": (QuESTEnv, []),
    "getEnvironmentString": (None, [C.c_char_p]),
    "setSeeds": (None, [C.POINTER(C.c_uint), cint]), "setSeedsToDefault": (None, []),
    "getSeeds": (None, [C.POINTER(C.c_uint)]), "getNumSeeds": (cint, []),
    "setInputErrorHandler": (None, [_HANDLER_T]),
    "setValidationOn": (None, []), "setValidationOff": (None, []),
    "setValidationEpsilon": (None, [qreal]), "setValidationEpsilonToDefault": (None, []),
    "getValidationEpsilon": (qreal, []),
    # qureg
    "createQureg": (Qureg, [cint]), "createDensityQureg": (Qureg, [cint]),
    "createForcedQureg": (Qureg, [cint]), "createForcedDensityQureg": (Qureg, [cint]),
    "createCustomQureg": (Qureg, [cint, cint, cint, cint, cint]), "createCloneQureg": (Qureg, [Qureg]),
    "createCustomPrecisionQureg": (Qureg, [cint, cint, cint]), "getQuregPrecision": (cint, [Qureg]),
    "getQuregDeviceBytes": (qindex, [cint, cint, cint]),
    "destroyQureg": (None, [Qureg]), "reportQuregParams": (None, [Qureg]), "reportQureg": (None, [Qureg]),
    "syncQuregToGpu": (None, [Qureg]), "syncQuregFromGpu": (None, [Qureg]),
    "getQuregAmps": (None, [ptr, Qureg, qindex, qindex]),
    "_wrap_getQuregAmp": (None, [C.POINTER(qcomp), Qureg, qindex]),
    "_wrap_getDensityQuregAmp": (None, [C.POINTER(qcomp), Qureg, qindex, qindex]),
    # initialisation
    "initBlankState": (None, [Qureg]), "initZeroState": (None, [Qureg]), "initPlusState": (None, [Qureg]),
    "initPureState": (None, [Qureg, Qureg]), "initClassicalState": (None, [Qureg, qindex]),
    "initDebugState": (None, [Qureg]), "initArbitraryPureState": (None, [Qureg, ptr]),
    "initRandomPureState": (None, [Qureg]), "initRandomMixedState": (None, [Qureg, qindex]),
    "setQuregAmps": (None, [Qureg, qindex, ptr, qindex]),
    "setDensityQuregFlatAmps": (None, [Qureg, qindex, ptr, qindex]),
    "setQuregToClone": (None, [Qureg, Qureg]),
    "setQuregToSuperposition": (None, [qcomp, Qureg, qcomp, Qureg, qcomp, Qureg]),
    "setQuregToRenormalized": (qreal, [Qureg]),
    "setQuregToPartialTrace": (None, [Qureg, Qureg, IntP, cint]),
    # matrices / paulis / channels
    "createCompMatr": (CompMatr, [cint]), "destroyCompMatr": (None, [CompMatr]),
    "setCompMatr": (None, [CompMatr, C.POINTER(C.POINTER(qcomp))]), "syncCompMatr": (None, [CompMatr]),
    "createDiagMatr": (DiagMatr, [cint]), "destroyDiagMatr": (None, [DiagMatr]),
    "setDiagMatr": (None, [DiagMatr, C.POINTER(qcomp)]),
    "createFullStateDiagMatr": (FullStateDiagMatr, [cint]),
    "destroyFullStateDiagMatr": (None, [FullStateDiagMatr]),
    "setFullStateDiagMatr": (None, [FullStateDiagMatr, qindex, C.POINTER(qcomp), qindex]),
    "setFullStateDiagMatrFromPauliStrSum": (None, [FullStateDiagMatr, PauliStrSum]),
    "setDiagMatrFromMultiVarFunc": (None, [DiagMatr, C.c_void_p, IntP, cint, cint]),
    "setDiagMatrFromMultiDimLists": (None, [DiagMatr, C.c_void_p, IntP, cint]),
    "setFullStateDiagMatrFromMultiVarFunc": (None, [FullStateDiagMatr, C.c_void_p, IntP, cint, cint]),
    "setFullStateDiagMatrFromMultiDimLists": (None, [FullStateDiagMatr, C.c_void_p, IntP, cint]),
    "createFullStateDiagMatrFromPauliStrSum": (FullStateDiagMatr, [PauliStrSum]),
    "setQuregToPauliStrSum": (None, [Qureg, PauliStrSum]),
    "getPauliStr": (PauliStr, [C.c_char_p, IntP, cint]),
    "createPauliStrSum": (PauliStrSum, [C.POINTER(PauliStr), C.POINTER(qcomp), qindex]),
    "createInlinePauliStrSum": (PauliStrSum, [C.c_char_p]),
    "destroyPauliStrSum": (None, [PauliStrSum]),
    "createKrausMap": (KrausMap, [cint, cint]), "destroyKrausMap": (None, [KrausMap]),
    "setKrausMap": (None, [KrausMap, C.POINTER(C.POINTER(C.POINTER(qcomp)))]),
    "createSuperOp": (SuperOp, [cint]), "destroySuperOp": (None, [SuperOp]),
    "setSuperOp": (None, [SuperOp, C.POINTER(C.POINTER(qcomp))]),
    # operations
    "multiplyCompMatr1": (None, [Qureg, cint, CompMatr1]),
    "applyCompMatr1": (None, [Qureg, cint, CompMatr1]),
    "applyControlledCompMatr1": (None, [Qureg, cint, cint, CompMatr1]),
    "applyMultiControlledCompMatr1": (None, [Qureg, IntP, cint, cint, CompMatr1]),
    "applyMultiStateControlledCompMatr1": (None, [Qureg, IntP, IntP, cint, cint, CompMatr1]),
    "applyCompMatr2": (None, [Qureg, cint, cint, CompMatr2]),
    "applyControlledCompMatr2": (None, [Qureg, cint, cint, cint, CompMatr2]),
    "applyMultiControlledCompMatr2": (None, [Qureg, IntP, cint, cint, cint, CompMatr2]),
    "applyCompMatr": (None, [Qureg, IntP, cint, CompMatr]),
    "applyMultiControlledCompMatr": (None, [Qureg, IntP, cint, IntP, cint, CompMatr]),
    "applyDiagMatr1": (None, [Qureg, cint, DiagMatr1]),
    "applyMultiControlledDiagMatr1": (None, [Qureg, IntP, cint, cint, DiagMatr1]),
    "applyDiagMatr2": (None, [Qureg, cint, cint, DiagMatr2]),
    "applyDiagMatr": (None, [Qureg, IntP, cint, DiagMatr]),
    "applyFullStateDiagMatr": (None, [Qureg, FullStateDiagMatr]),
    "multiplyFullStateDiagMatr": (None, [Qureg, FullStateDiagMatr]),
    "multiplyDiagMatr": (None, [Qureg, IntP, cint, DiagMatr]),
    "multiplyDiagMatrPower": (None, [Qureg, IntP, cint, DiagMatr, qcomp]),
    "applyDiagMatrPower": (None, [Qureg, IntP, cint, DiagMatr, qcomp]),
    "applyControlledDiagMatrPower": (None, [Qureg, cint, IntP, cint, DiagMatr, qcomp]),
    "applyMultiControlledDiagMatrPower": (None, [Qureg, IntP, cint, IntP, cint, DiagMatr, qcomp]),
    "applyMultiStateControlledDiagMatrPower": (None, [Qureg, IntP, IntP, cint, IntP, cint, DiagMatr, qcomp]),
    "multiplyFullStateDiagMatrPower": (None, [Qureg, FullStateDiagMatr, qcomp]),
    "applyFullStateDiagMatrPower": (None, [Qureg, FullStateDiagMatr, qcomp]),
    "applySqrtSwap": (None, [Qureg, cint, cint]),
    "applyControlledSqrtSwap": (None, [Qureg, cint, cint, cint]),
    "applyMultiControlledSqrtSwap": (None, [Qureg, IntP, cint, cint, cint]),
    "applyMultiStateControlledSqrtSwap": (None, [Qureg, IntP, IntP, cint, cint, cint]),
    "applyHadamard": (None, [Qureg, cint]), "applyPauliX": (None, [Qureg, cint]),
    "applyPauliY": (None, [Qureg, cint]), "applyPauliZ": (None, [Qureg, cint]),
    "applyS": (None, [Qureg, cint]), "applyT": (None, [Qureg, cint]),
    "applyControlledPauliX": (None, [Qureg, cint, cint]), "applyControlledPauliZ": (None, [Qureg, cint, cint]),
    "applyPhaseShift": (None, [Qureg, cint, qreal]),
    "applyTwoQubitPhaseShift": (None, [Qureg, cint, cint, qreal]),
    "applyMultiQubitPhaseShift": (None, [Qureg, IntP, cint, qreal]),
    "applyRotateX": (None, [Qureg, cint, qreal]), "applyRotateY": (None, [Qureg, cint, qreal]),
    "applyRotateZ": (None, [Qureg, cint, qreal]),
    "applySwap": (None, [Qureg, cint, cint]),
    "applyMultiControlledSwap": (None, [Qureg, IntP, cint, cint, cint]),
    "applyPauliStr": (None, [Qureg, PauliStr]),
    "applyMultiControlledPauliStr": (None, [Qureg, IntP, cint, PauliStr]),
    "applyPauliGadget": (None, [Qureg, PauliStr, qreal]),
    "multiplyPauliStrSum": (None, [Qureg, PauliStrSum, Qureg]),
    "applyTrotterizedPauliStrSumGadget": (None, [Qureg, PauliStrSum, qreal, cint, cint]),
    "applyMultiControlledPauliGadget": (None, [Qureg, IntP, cint, PauliStr, qreal]),
    "applyPhaseGadget": (None, [Qureg, IntP, cint, qreal]),
    "applyQuantumFourierTransform": (None, [Qureg, IntP, cint]),
    "applyFullQuantumFourierTransform": (None, [Qureg]),
    "applyQubitProjector": (None, [Qureg, cint, cint]),
    "applyQubitMeasurement": (cint, [Qureg, cint]),
    "applyForcedQubitMeasurement": (qreal, [Qureg, cint, cint]),
    "applyMultiQubitMeasurement": (qindex, [Qureg, IntP, cint]),
    "applyForcedMultiQubitMeasurement": (qreal, [Qureg, IntP, IntP, cint]),
    # decoherence
    "mixDephasing": (None, [Qureg, cint, qreal]), "mixTwoQubitDephasing": (None, [Qureg, cint, cint, qreal]),
    "mixDepolarising": (None, [Qureg, cint, qreal]),
    "mixTwoQubitDepolarising": (None, [Qureg, cint, cint, qreal]),
    "mixDamping": (None, [Qureg, cint, qreal]), "mixPaulis": (None, [Qureg, cint, qreal, qreal, qreal]),
    "mixQureg": (None, [Qureg, Qureg, qreal]), "mixKrausMap": (None, [Qureg, IntP, cint, KrausMap]),
    "mixSuperOp": (None, [Qureg, IntP, cint, SuperOp]),
    # calculations
    "calcExpecPauliStr": (qreal, [Qureg, PauliStr]), "calcExpecPauliStrSum": (qreal, [Qureg, PauliStrSum]),
    "calcProbOfBasisState": (qreal, [Qureg, qindex]),
    "calcProbOfQubitOutcome": (qreal, [Qureg, cint, cint]),
    "calcProbOfMultiQubitOutcome": (qreal, [Qureg, IntP, IntP, cint]),
    "calcProbsOfAllMultiQubitOutcomes": (None, [C.POINTER(qreal), Qureg, IntP, cint]),
    "calcTotalProb": (qreal, [Qureg]), "calcPurity": (qreal, [Qureg]),
    "calcFidelity": (qreal, [Qureg, Qureg]), "calcDistance": (qreal, [Qureg, Qureg]),
    "_wrap_calcInnerProduct": (None, [Qureg, Qureg, C.POINTER(qcomp)]),
    "calcExpecFullStateDiagMatr": (qreal, [Qureg, FullStateDiagMatr]),
    "calcExpecFullStateDiagMatrPower": (qreal, [Qureg, FullStateDiagMatr, qreal]),
    "_wrap_calcExpecNonHermitianFullStateDiagMatr": (None, [C.POINTER(qcomp), Qureg, FullStateDiagMatr]),
    "_wrap_calcExpecNonHermitianFullStateDiagMatrPower": (None, [C.POINTER(qcomp), Qureg, FullStateDiagMatr, qcomp]),
    "calcPartialTrace": (Qureg, [Qureg, IntP, cint]),
    "getQuESTStream": (ptr, []),
    "setQuESTFusion": (None, [cint, cint, cint, cint]),
    "getQuESTFusion": (None, [IntP, IntP, IntP, IntP]),
}

_raw = {}
for _name, (_res, _args) in _SIGS.items():
    _f = getattr(_lib, _name)
    _f.restype = _res
    _f.argtypes = _args
    _raw[_name] = _f


def _wrap(name):
    f = _raw[name]

    def call(*args):
        _err.last = None
        out = f(*args)
        _raise_pending()
        return out

    call.__name__ = name
    return call


for _name in _SIGS:
    if not _name.startswith("_"):
        globals()[_name] = _wrap(_name)

_raw["setInputErrorHandler"](_handler)


# ------------------------------------------------------------- helpers

def ints(values) -> "C.Array":
    values = [int(v) for v in values]
    return (C.c_int * max(1, len(values)))(*values)


def compmatr1(m) -> CompMatr1:
    m = np.asarray(m, dtype=complex).reshape(2, 2)
    out = CompMatr1()
    out.numQubits, out.numRows = 1, 2
    for r in range(2):
        for c in range(2):
            out.elems[r][c] = qcomp.of(m[r, c])
    return out


def compmatr2(m) -> CompMatr2:
    m = np.asarray(m, dtype=complex).reshape(4, 4)
    out = CompMatr2()
    out.numQubits, out.numRows = 2, 4
    for r in range(4):
        for c in range(4):
            out.elems[r][c] = qcomp.of(m[r, c])
    return out


def diagmatr1(d) -> DiagMatr1:
    out = DiagMatr1()
    out.numQubits, out.numElems = 1, 2
    for i, v in enumerate(d):
        out.elems[i] = qcomp.of(v)
    return out


def diagmatr2(d) -> DiagMatr2:
    out = DiagMatr2()
    out.numQubits, out.numElems = 2, 4
    for i, v in enumerate(d):
        out.elems[i] = qcomp.of(v)
    return out


def _rows(m: np.ndarray):
    """qcomp** view of a C-contiguous complex128 matrix (keeps refs alive)."""
    m = np.ascontiguousarray(m, dtype=np.complex128)
    rowptrs = (C.POINTER(qcomp) * m.shape[0])(
        *[C.cast(m[r].ctypes.data, C.POINTER(qcomp)) for r in range(m.shape[0])])
    return m, rowptrs


def make_compmatr(m) -> CompMatr:
    m = np.asarray(m, dtype=complex)
    n = int(round(np.log2(m.shape[0])))
    cm = createCompMatr(n)
    keep, rows = _rows(m)
    setCompMatr(cm, rows)
    return cm


def make_diagmatr(d) -> DiagMatr:
    d = np.ascontiguousarray(np.asarray(d, dtype=np.complex128))
    n = int(round(np.log2(d.shape[0])))
    dm = createDiagMatr(n)
    setDiagMatr(dm, C.cast(d.ctypes.data, C.POINTER(qcomp)))
    return dm


def make_krausmap(ops) -> KrausMap:
    ops = [np.ascontiguousarray(np.asarray(k, dtype=np.complex128)) for k in ops]
    n = int(round(np.log2(ops[0].shape[0])))
    km = createKrausMap(n, len(ops))
    keep = []
    mats = (C.POINTER(C.POINTER(qcomp)) * len(ops))()
    for i, k in enumerate(ops):
        arr, rows = _rows(k)
        keep.append((arr, rows))
        mats[i] = C.cast(rows, C.POINTER(C.POINTER(qcomp)))
    setKrausMap(km, mats)
    return km


def pauli_str(chars: str, targets) -> PauliStr:
    t = ints(targets)
    return getPauliStr(chars.encode(), t, len(chars))


def get_amps(q: Qureg) -> np.ndarray:
    """All amplitudes as complex128 (density matrices flat column-major)."""
    out = np.empty(q.numAmps, dtype=np.complex128)
    rc = _gpu.qa_getAmps(out.ctypes.data_as(C.c_void_p), C.c_void_p(q.gpuAmps),
                         C.c_int(getQuregPrecision(q)), C.c_int64(0), C.c_int64(q.numAmps))
    if rc != 0:
        raise RuntimeError(_gpu_last_error())
    return out


def set_amps(q: Qureg, amps) -> None:
    a = np.ascontiguousarray(np.asarray(amps, dtype=np.complex128).reshape(-1))
    if a.size != q.numAmps:
        raise ValueError(f"expected {q.numAmps} amplitudes, got {a.size}")
    if q.isDensityMatrix:
        setDensityQuregFlatAmps(q, 0, a.ctypes.data_as(C.c_void_p), a.size)
    else:
        setQuregAmps(q, 0, a.ctypes.data_as(C.c_void_p), a.size)


def inner_product(a: Qureg, b: Qureg) -> complex:
    out = qcomp()
    _err.last = None
    _raw["_wrap_calcInnerProduct"](a, b, C.byref(out))
    _raise_pending()
    return out.value()


def make_fullstate_diag(n: int, d) -> FullStateDiagMatr:
    d = np.ascontiguousarray(np.asarray(d, dtype=np.complex128))
    m = createFullStateDiagMatr(n)
    setFullStateDiagMatr(m, 0, C.cast(d.ctypes.data, C.POINTER(qcomp)), d.size)
    return m


def expec_nonhermitian_fullstate_diag(q: Qureg, m: FullStateDiagMatr, exponent=None) -> complex:
    out = qcomp()
    _err.last = None
    if exponent is None:
        _raw["_wrap_calcExpecNonHermitianFullStateDiagMatr"](C.byref(out), q, m)
    else:
        _raw["_wrap_calcExpecNonHermitianFullStateDiagMatrPower"](C.byref(out), q, m, qcomp.of(exponent))
    _raise_pending()
    return out.value()


def probs_of_all_outcomes(q: Qureg, qubits) -> np.ndarray:
    out = np.empty(1 << len(qubits), dtype=np.float64)
    calcProbsOfAllMultiQubitOutcomes(out.ctypes.data_as(C.POINTER(qreal)), q, ints(qubits), len(qubits))
    return out


def set_seeds(seeds) -> None:
    arr = (C.c_uint * len(seeds))(*[int(s) & 0xFFFFFFFF for s in seeds])
    setSeeds(arr, len(seeds))


# ------------------------------------------------- kernel-layer C ABI

_gpu.qa_lastError.restype = C.c_char_p


def _gpu_last_error() -> str:
    return _gpu.qa_lastError().decode()


def gpu():
    """The raw kernel-layer library (include/quest_amd_gpu.h)."""
    return _gpu


def get_fusion() -> tuple:
    """(enabled, tileQubits, lowQubits, maxQueue) of the gate-fusion queue."""
    v = [C.c_int(0) for _ in range(4)]
    getQuESTFusion(*[C.byref(x) for x in v])
    return tuple(x.value for x in v)


def set_fusion(enabled: bool, tile_qubits: int = 12, low_qubits: int = 4, max_queue: int = 512) -> None:
    setQuESTFusion(int(bool(enabled)), int(tile_qubits), int(low_qubits), int(max_queue))


def set_fusion_span(span_qubits=12) -> None:
    """Span passes for runs of Pauli and phase gadgets over cosets of
    2^span_qubits amplitudes (10..13; True = 12, the default), or none
    (False / 0)."""
    q = 12 if span_qubits is True else int(span_qubits)
    rc = gpu().qa_setFusionSpan(q)
    if rc != 0:
        raise QuESTError("qa_setFusionSpan", f"failed ({rc}): {_gpu_last_error()}")


class FusionTrial(C.Structure):
    _fields_ = [("tileQubits", C.c_int), ("lowQubits", C.c_int), ("spanQubits", C.c_int),
                ("blockMinGates", C.c_int), ("ms", C.c_double)]


def tune_fusion(qureg, retune: bool = False) -> dict:
    """Runs the fusion autotuner on the gates queued for `qureg` (see
    qa_tuneFusion) and returns {"trials": [...], "best": {...}}."""
    trials = (FusionTrial * 32)()
    num, best = C.c_int(), C.c_int()
    rc = gpu().qa_tuneFusion(C.c_void_p(qureg.gpuAmps), int(bool(retune)), trials, 32, C.byref(num),
                             C.byref(best))
    if rc != 0:
        raise QuESTError("qa_tuneFusion", f"failed ({rc}): {_gpu_last_error()}")
    out = [{"tile_qubits": t.tileQubits, "low_qubits": t.lowQubits, "span_qubits": t.spanQubits,
            "block_min_gates": t.blockMinGates, "ms": t.ms} for t in trials[:min(num.value, 32)]]
    return {"trials": out, "best": out[best.value] if 0 <= best.value < len(out) else None}


def fusion_span_stats() -> dict:
    """Span passes launched, the gates they applied and (with pass timing
    on) their milliseconds since the last call (which resets them)."""
    passes, gates, ms = C.c_int64(), C.c_int64(), C.c_double()
    rc = gpu().qa_getFusionSpanStats(C.byref(passes), C.byref(gates), C.byref(ms))
    if rc != 0:
        raise QuESTError("qa_getFusionSpanStats", f"failed ({rc}): {_gpu_last_error()}")
    return {"passes": passes.value, "gates": gates.value, "pass_ms": ms.value}


def fusion_stats() -> dict:
    """Tile passes launched, the gates they applied, their milliseconds (with
    pass timing on), the flops and program ops they executed and their
    register-layout changes, since the last call (which resets the
    counters)."""
    passes, gates, ms, flops = C.c_int64(), C.c_int64(), C.c_double(), C.c_double()
    rc = gpu().qa_getFusionStats(C.byref(passes), C.byref(gates), C.byref(ms), C.byref(flops))
    if rc != 0:
        raise QuESTError("qa_getFusionStats", f"failed ({rc}): {_gpu_last_error()}")
    ops, layouts = C.c_int64(), C.c_int64()
    rc = gpu().qa_getFusionProgramStats(C.byref(ops), C.byref(layouts))
    if rc != 0:
        raise QuESTError("qa_getFusionProgramStats", f"failed ({rc}): {_gpu_last_error()}")
    return {"passes": passes.value, "gates": gates.value, "pass_ms": ms.value, "flops": flops.value,
            "ops": ops.value, "layouts": layouts.value}


def set_fusion_exact(exact: bool) -> None:
    """Exact fusion applies queued gates in order with their single-gate
    arithmetic (bit-identical to unfused runs); otherwise (the default)
    passes reorder commuting gates and multiply runs of gates on the same
    one or two qubits into one matrix."""
    rc = gpu().qa_setFusionExact(C.c_int(int(bool(exact))))
    if rc != 0:
        raise QuESTError("qa_setFusionExact", f"failed ({rc}): {_gpu_last_error()}")


def set_fusion_prefetch(mode: int = 2) -> None:
    """How tile passes load each workgroup's next tile while working on the
    current one (qa_setFusionPrefetch): 2 (the default) loads spread over
    the generated program, 3 the stores deferred a tile and spread too, 1
    all loads at the tile's start, 0 no prefetch."""
    rc = gpu().qa_setFusionPrefetch(C.c_int(int(mode)))
    if rc != 0:
        raise QuESTError("qa_setFusionPrefetch", f"failed ({rc}): {_gpu_last_error()}")


def set_fusion_jit(mode: int, min_seen: int = 2) -> None:
    """Pass kernels generated at run time (qa_setFusionJit): mode 0 off, 1
    background compiles once a program shape has been planned min_seen
    times, 2 compile at first use and wait."""
    rc = gpu().qa_setFusionJit(C.c_int(int(mode)), C.c_int(int(min_seen)))
    if rc != 0:
        raise QuESTError("qa_setFusionJit", f"failed ({rc}): {_gpu_last_error()}")


def fusion_jit_wait() -> None:
    """Blocks until every queued pass-kernel compile has finished."""
    gpu().qa_fusionJitWait()


def fusion_jit_stats() -> dict:
    """Generated pass kernels compiled, failed compiles, passes they ran and
    the compile threads' seconds, since the library was initialised."""
    k, f, n, sec = C.c_int64(), C.c_int64(), C.c_int64(), C.c_double()
    gpu().qa_getFusionJitStats(C.byref(k), C.byref(f), C.byref(n), C.byref(sec))
    gpu().qa_fusionJitLastError.restype = C.c_char_p
    return {"kernels": k.value, "failures": f.value, "launches": n.value, "compile_seconds": sec.value,
            "last_error": (gpu().qa_fusionJitLastError() or b"").decode()}


def ensure_env() -> None:
    if not isQuESTEnvInit():
        initQuESTEnv()
