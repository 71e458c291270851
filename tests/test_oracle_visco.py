"""The oracle's visco.SLS restatement against the closed forms the reference
pins it with (internal/visco/visco_test.go:9-147), and the product's host
copy of the same arithmetic (gf_visco_fit, gf_ricker) against the oracle."""

import cmath
import math

import numpy as np
import pytest


def phase_velocity(O, s, v, fref, f):
    c = O.sls_velocity(s, v, fref, f)
    return 1 / (1 / c).real


@pytest.mark.parametrize("q", [5, 25, 100, 1000])
@pytest.mark.parametrize("f0", [1, 30, 250])
def test_fit_hits_the_target_exactly(oracle, q, f0):
    s = oracle.visco_fit(q, f0)
    assert abs(oracle.sls_q(s, f0) - q) <= 1e-12 * q


def test_fitted_frequency_is_the_minimum(oracle):
    s = oracle.visco_fit(25, 30)
    base = oracle.sls_q(s, 30)
    for f in [1, 10, 20, 29, 31, 45, 90, 900]:
        assert oracle.sls_q(s, f) >= base
    assert oracle.sls_q(s, 3) >= 2 * base


def test_modulus_limits(oracle):
    s = oracle.visco_fit(25, 30)
    assert oracle.sls_modulus(s, 0) == 1
    want = 1 + oracle.sls_delta(s)
    assert abs(oracle.sls_modulus(s, 1e12).real - want) <= 1e-6 * want
    assert oracle.sls_delta(s) > 0


@pytest.mark.parametrize("v", [120, 200, 800])
def test_phase_velocity_is_anchored_at_the_reference(oracle, v):
    s = oracle.visco_fit(25, 30)
    assert abs(phase_velocity(oracle, s, v, 30, 30) - v) <= 1e-10 * v


def test_velocity_dispersion_is_monotonic(oracle):
    s = oracle.visco_fit(25, 30)
    prev, f = 0.0, 0.5
    while f < 500:
        v = phase_velocity(oracle, s, 200, 30, f)
        assert v > prev
        prev, f = v, f * 1.2
    rise = phase_velocity(oracle, s, 200, 30, 500) / phase_velocity(oracle, s, 200, 30, 0.5) - 1
    assert 0.01 <= rise <= 0.15


def test_attenuation_has_the_right_sign(oracle):
    s = oracle.visco_fit(25, 30)
    c = oracle.sls_velocity(s, 200, 30, 60)
    assert c.imag > 0
    w = 2 * math.pi * 60
    k = w / c
    lam = 2 * math.pi / k.real
    decay = math.exp(k.imag * lam)
    want = math.exp(-math.pi / oracle.sls_q(s, 60))
    assert abs(decay - want) <= 0.01 * want


def test_elastic_limit(oracle, fdtd):
    s = fdtd.SLS(1e-3, 1e-3)
    assert oracle.sls_modulus(s, 50) == 1
    assert math.isinf(oracle.sls_q(s, 50))
    assert abs(phase_velocity(oracle, s, 200, 30, 500) - 200) <= 1e-12


def test_moduli_agree_with_the_complex_velocity(oracle):
    s = oracle.visco_fit(25, 30)
    rho, v, fref = 1700, 200, 30
    mr, mu = oracle.sls_moduli(s, v, fref, rho)
    want = mr * (1 + oracle.sls_delta(s))
    assert abs(mu - want) <= 1e-9 * want
    for f in [1, 30, 200]:
        c = oracle.sls_velocity(s, v, fref, f)
        got = rho * c * c
        want = mr * oracle.sls_modulus(s, f)
        assert abs(got - want) <= 1e-9 * abs(want)


def test_fit_rejects_nonsense(oracle, fdtd):
    with pytest.raises(oracle.OracleError, match="visco: Q must be positive, got 0"):
        oracle.visco_fit(0, 30)
    with pytest.raises(oracle.OracleError, match="visco: reference frequency must be positive, got 0 Hz"):
        oracle.visco_fit(25, 0)
    # the product's host copy says the same
    with pytest.raises(fdtd.FdtdError, match="visco: Q must be positive, got 0"):
        fdtd.Fit(0, 30)


@pytest.mark.parametrize("q,f0", [(5, 1), (25, 30), (30, 30), (1000, 250)])
def test_product_fit_equals_oracle(oracle, fdtd, q, f0):
    a, b = fdtd.Fit(q, f0), oracle.visco_fit(q, f0)
    assert (a.TauSigma, a.TauEps) == (b.TauSigma, b.TauEps)


def test_ricker_closed_form(oracle, fdtd):
    """fdtd.Ricker (run.go:19-31): peak at 1.5/f0, zero crossings at
    1.5/f0 +- 1/(pi f0 sqrt 2), and the product's copy equals the oracle's."""
    f0, dt, n, peak = 40.0, 1e-4, 2000, 800.0
    w = oracle.ricker(peak, f0, dt, n)
    np.testing.assert_array_equal(w, fdtd.Ricker(peak, f0, dt, n))
    t = np.arange(n) * dt
    tau = t - 1.5 / f0
    a = math.pi ** 2 * f0 ** 2
    np.testing.assert_allclose(w, peak * (1 - 2 * a * tau ** 2) * np.exp(-a * tau ** 2), rtol=1e-13, atol=1e-12)
    assert int(np.argmax(w)) == round(1.5 / f0 / dt)
    # zero net impulse (to truncation)
    assert abs(w.sum() * dt) < 1e-9 * peak


def test_ricker_spectrum_matches_the_dft_of_the_series(oracle, fdtd):
    """RickerSpectrum (run.go:41-46) is the transform of Ricker with its delay:
    a direct DFT of the sampled series must agree to sampling accuracy."""
    f0, dt, n, peak = 25.0, 5e-5, 8000, 1000.0
    w = oracle.ricker(peak, f0, dt, n)
    t = np.arange(n) * dt
    for f in [5.0, 25.0, 60.0]:
        direct = dt * np.sum(w * np.exp(-2j * math.pi * f * t))
        want = oracle.ricker_spectrum(peak, f0, f)
        assert abs(direct - want) <= 1e-6 * abs(want)
        assert fdtd.RickerSpectrum(peak, f0, f) == want
