"""Python bindings of the MI355X QuEST backend.

`libquest_amd.so` exports the quest.h C API (include/quest.h) and
`libquest_amd_gpu.so` the kernel-layer C ABI (include/quest_amd_gpu.h).
This module loads them with ctypes, declares the struct layouts and
signatures, and turns QuEST input errors into Python exceptions.  There is
no fallback: if the native libraries are missing, importing `quest_amd.api`
raises.

Names follow the reference API one-to-one (quest/include/*.h), so a test
written against this module reads like QuEST's own tests:

    from quest_amd import api as Q
    Q.initQuESTEnv()
    q = Q.createQureg(10)
    Q.applyHadamard(q, 0)
    p = Q.calcProbOfQubitOutcome(q, 0, 0)
"""

from pathlib import Path

PKG_DIR = Path(__file__).resolve().parent
LIB_DIR = PKG_DIR / "lib"
HOST_LIB = LIB_DIR / "libquest_amd.so"
GPU_LIB = LIB_DIR / "libquest_amd_gpu.so"

__all__ = ["PKG_DIR", "LIB_DIR", "HOST_LIB", "GPU_LIB"]
