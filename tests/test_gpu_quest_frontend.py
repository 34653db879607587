"""QuEST's own front end on this backend's GPU kernels.

oracle/_ref/circuit_driver_amd is QuEST v4.1.0's api/, core/, cpu/ and comm/
sources (compiled here from the reference by oracle/Makefile's ref-amd
target) with its gpu/*.cpp replaced by this backend's accelerator bridge
(quest_amd/csrc/bridge/gpu_bridge.cpp) and linked with libquest_amd_gpu.so:
QuEST's validation, localiser and accelerator dispatch call the MI355X
kernels through the gpu_* seam, exactly as they call QuEST's CUDA
kernels.  Every golden fixture (written by QuEST's own CPU build) replayed
through it with every Qureg forced onto the GPU must reproduce the fixture:
amplitudes within 1e-12 relative L2, results within 1e-10, measurement
outcomes bit for bit (the bridge draws the random-state seed from QuEST's own
generator, so the measurement stream stays aligned).
"""

import sys

import pytest

from conftest import GOLDEN_CASES, compare_results, load_golden, rel_l2

pytestmark = pytest.mark.gpu

AMP_TOL = 1e-12
SCALAR_TOL = 1e-10


@pytest.mark.parametrize("name", GOLDEN_CASES)
def test_fixture_through_quest_front_end(name, tmp_path):
    from oracle import qoracle as O
    print(f"[gpu] QuEST front end on the MI355X kernels: {name}", flush=True)
    sys.stdout.flush()
    assert O.REF_AMD_DRIVER.exists(), "oracle/_ref/circuit_driver_amd is not built (make -C oracle ref-amd)"
    circ, init, amps, sample, results = load_golden(name)
    got, res, out = O.run_reference(circ, tmp_path, init_amps=init, driver=O.REF_AMD_DRIVER,
                                    env={"QD_DEPLOY": "gpu"})
    assert f"qureg of {circ.num_qubits} qubits on the GPU" in out
    if sample is not None:
        got = got[sample]
    assert rel_l2(got, amps) <= AMP_TOL
    compare_results(res, results, tol=SCALAR_TOL)
