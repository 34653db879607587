"""Initialisations on the GPU against what fixes them without the
reference's own random streams:

  * initDebugState: the closed form a_i = 2i/10 + i(2i+1)/10 of
    cpu_subroutines.cpp:2364-2376, bit for bit, on statevectors and on the
    flat amplitudes of density matrices (initialisations.cpp:96-100);
  * initRandomPureState: the reference draws one seed from the measurement
    RNG (randomiser.cpp:175-193) and fills complex Gaussian amplitudes
    (magnitude sqrt(n1^2 + n2^2), uniform phase), then renormalises
    (initialisations.cpp:121-133).  The state is not reproducible across
    backends (its values depend on the backend's generator), so its
    distribution is checked: the norm is 1; N|a|^2 is exponentially
    distributed (the Haar measure), so sum |a|^4 = 2/(N+1) up to sampling
    noise of relative size sqrt(5/N), and P(N|a|^2 > 1) = e^-1; phases are
    uniform, so |sum a| stays O(1).  The same seeds give the same state.
    Its RNG consumption is pinned bit-exactly by the random_seq_sv_n5
    fixture (measurement outcomes around random initialisations);
  * initRandomMixedState (localiser.cpp:801-827): trace 1, Hermitian,
    positive semi-definite, purity between 1/2^n and 1.
"""

import sys

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def debug_closed_form(num: int) -> np.ndarray:
    i = np.arange(num, dtype=np.float64)
    return (2 * i) / 10.0 + 1j * ((2 * i + 1) / 10.0)


@pytest.mark.parametrize("n,dens", [(1, False), (7, False), (20, False), (1, True), (5, True)])
def test_debug_state_is_the_closed_form(api, n, dens):
    say(f"[gpu] initDebugState n={n} dens={dens}")
    q = api.createDensityQureg(n) if dens else api.createQureg(n)
    try:
        api.initDebugState(q)
        got = api.get_amps(q)
        want = debug_closed_form(q.numAmps)
        assert np.array_equal(got.real, want.real) and np.array_equal(got.imag, want.imag)
    finally:
        api.destroyQureg(q)


@pytest.mark.parametrize("n", [12, 20])
def test_random_pure_state_distribution(api, n):
    say(f"[gpu] initRandomPureState distribution n={n}")
    api.set_seeds([1234, 5])
    q = api.createQureg(n)
    try:
        api.initRandomPureState(q)
        a = api.get_amps(q)
        N = a.size
        p = np.abs(a) ** 2
        assert abs(np.sum(p) - 1) < 1e-12
        assert abs(api.calcTotalProb(q) - 1) < 1e-12
        # sum |a|^4 of a Haar state: 2/(N+1), relative spread sqrt(5/N)
        rel = np.sum(p * p) * (N + 1) / 2 - 1
        assert abs(rel) < 10 * np.sqrt(5.0 / N), rel
        # N|a|^2 ~ Exp(1): fraction above 1 is 1/e
        frac = np.mean(N * p > 1)
        assert abs(frac - np.exp(-1)) < 10 * np.sqrt(np.exp(-1) * (1 - np.exp(-1)) / N), frac
        # uniform phases: the plain sum is a unit complex Gaussian
        assert abs(np.sum(a)) < 6
        # and not a degenerate state: the largest |a|^2 is O(log N / N)
        assert np.max(p) < 40 * np.log(N) / N
    finally:
        api.destroyQureg(q)


def test_random_pure_state_follows_the_seeds(api):
    say("[gpu] initRandomPureState seeds")
    n = 10
    q = api.createQureg(n)
    try:
        api.set_seeds([77, 1])
        api.initRandomPureState(q)
        a1 = api.get_amps(q)
        api.initRandomPureState(q)
        a2 = api.get_amps(q)
        api.set_seeds([77, 1])
        api.initRandomPureState(q)
        a3 = api.get_amps(q)
        assert np.array_equal(a1, a3)
        assert np.max(np.abs(a1 - a2)) > 1e-3
    finally:
        api.destroyQureg(q)


def test_random_pure_density_matrix(api):
    say("[gpu] initRandomPureState density matrix")
    n = 5
    q = api.createDensityQureg(n)
    try:
        api.set_seeds([9, 9])
        api.initRandomPureState(q)
        rho = api.get_amps(q).reshape(1 << n, 1 << n, order="F")
        assert abs(np.trace(rho) - 1) < 1e-12
        assert np.allclose(rho, rho.conj().T, atol=1e-14)
        assert abs(api.calcPurity(q) - 1) < 1e-12
        w = np.linalg.eigvalsh(rho)
        assert abs(w[-1] - 1) < 1e-12 and np.all(np.abs(w[:-1]) < 1e-12)
    finally:
        api.destroyQureg(q)


@pytest.mark.parametrize("num_pure", [1, 3, 40])
def test_random_mixed_state(api, num_pure):
    say(f"[gpu] initRandomMixedState k={num_pure}")
    n = 4
    q = api.createDensityQureg(n)
    try:
        api.set_seeds([31, num_pure])
        api.initRandomMixedState(q, num_pure)
        rho = api.get_amps(q).reshape(1 << n, 1 << n, order="F")
        assert abs(np.trace(rho) - 1) < 1e-12
        assert np.allclose(rho, rho.conj().T, atol=1e-14)
        w = np.linalg.eigvalsh(rho)
        assert w[0] > -1e-12
        # rank at most num_pure
        assert np.sum(w > 1e-10) == min(num_pure, 1 << n)
        purity = api.calcPurity(q)
        assert 1 / (1 << n) - 1e-12 <= purity <= 1 + 1e-12
        if num_pure == 1:
            assert abs(purity - 1) < 1e-12
    finally:
        api.destroyQureg(q)
