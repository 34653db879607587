"""Pins the oracle's powered diagonal matrices (qoracle.State.diag_n with
power, restating cpu_statevec_anyCtrlAnyTargDiagMatr_sub's optional
pow(qcomp, qcomp) then conj, operations.cpp:468-500) against full operators
built independently: controlled diag(d**p) from numpy, applied as U psi,
U rho U^dagger, or U rho for the multiply form."""

import numpy as np
import pytest

from quest_amd import circuits as qc
from oracle import qoracle as O


def controlled_diag(n, ctrls, states, targs, d):
    dim = 1 << n
    u = np.ones(dim, dtype=complex)
    for i in range(dim):
        if any(((i >> c) & 1) != s for c, s in zip(ctrls, states)):
            continue
        k = sum(((i >> t) & 1) << b for b, t in enumerate(targs))
        u[i] = d[k]
    return np.diag(u)


@pytest.mark.parametrize("dens", [False, True])
@pytest.mark.parametrize("power", [0.5, -1.25, 2.0, 0.3 + 0.2j])
@pytest.mark.parametrize("multiply", [False, True])
def test_powered_diagonal_matches_full_operator(dens, power, multiply):
    n = 3 if dens else 5
    rng = np.random.default_rng(11)
    targs = [3, 0] if not dens else [2, 0]
    ctrls, states = [1], [0]
    d = rng.normal(size=4) + 1j * rng.normal(size=4)
    st = O.State(n, dens)
    st.a[:] = qc.random_state(n, 4, dens)
    U = controlled_diag(n, ctrls, states, targs, d ** power)
    if not dens:
        want = U @ st.a
    else:
        rho = st.a.reshape(1 << n, 1 << n, order="F")
        want = (U @ rho if multiply else U @ rho @ U.conj().T).reshape(-1, order="F")
    st.diag_n(ctrls, targs, d, power=power, states=states, multiply=multiply)
    assert np.max(np.abs(st.a - want)) <= 1e-12 * max(1.0, np.max(np.abs(want)))
