import sys
from pathlib import Path

import numpy as np
import pytest

ROOT = Path(__file__).resolve().parents[1]
if str(ROOT) not in sys.path:
    sys.path.insert(0, str(ROOT))

GOLDEN = Path(__file__).resolve().parent / "golden"


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs through the HIP kernel layer)")


@pytest.fixture(scope="session")
def api():
    """The quest_amd API with an initialised environment (GPU tests only)."""
    from quest_amd import api as Q
    Q.ensure_env()
    return Q


def load_golden(name: str):
    """(circuit, init amps or None, expected amps, sample indices or None, results)."""
    from quest_amd import circuits as qc
    d = np.load(GOLDEN / f"{name}.npz", allow_pickle=False)
    circ = qc.from_text(str(d["circuit"]))
    circ.name = name
    init = d["init"] if "init" in d else None
    if init is None and "init_seed" in d:
        # large cases regenerate their initial state from the recorded seed
        init = qc.random_state(circ.num_qubits, int(d["init_seed"][0]), circ.dens)
    if init is not None:
        circ.init = ("random", 0)  # amplitudes supplied explicitly
    elif circ.init[0] == "amps":
        raise AssertionError(f"{name}: fixture lacks its initial amplitudes")
    labels = [str(x) for x in d["labels"]]
    vals, offs = d["values"], d["offsets"]
    results = []
    for i, lab in enumerate(labels):
        v = vals[offs[i]:offs[i + 1]]
        if lab in ("meas", "measn"):
            results.append((lab, int(v[0])))
        elif lab in ("probs", "inner"):
            results.append((lab, list(v)))
        else:
            results.append((lab, float(v[0])))
    sample = d["sample"] if "sample" in d else None
    return circ, init, d["amps"], sample, results


GOLDEN_CASES = ["layered_n10", "layered_n12", "density_n5", "variational_n8", "allops_sv_n7", "allops_dm_n4",
                "measure_sv_n6", "measure_dm_n3", "layered_n20_c1", "random_seq_sv_n5", "measn_sv_n6",
                "measn_dm_n3", "overlap_sv_n6", "overlap_dm_n3", "probs_sv_n12", "probs_dm_n5"]
# fixtures whose outcomes come from the measurement RNG (pinned on the GPU
# path directly; the oracle has no sampling)
SAMPLED_CASES = {"measure_sv_n6", "measure_dm_n3", "random_seq_sv_n5", "measn_sv_n6", "measn_dm_n3"}


def rel_l2(a, b) -> float:
    a = np.asarray(a)
    b = np.asarray(b)
    return float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300))


def compare_results(got, want, tol=1e-10):
    assert len(got) == len(want), (len(got), len(want))
    for (lg, vg), (lw, vw) in zip(got, want):
        assert lg == lw, (lg, lw)
        if lg in ("meas", "measn"):
            assert vg == vw, f"measurement outcome {vg} != reference {vw}"
        else:
            np.testing.assert_allclose(np.atleast_1d(vg), np.atleast_1d(vw), rtol=0, atol=tol, err_msg=lg)
