"""The footstep force of config 2 (a restatement of internal/grf's vertical
Stance profile) against the reference's own tests, grf_test.go:20-301: the
momentum balance, the force-plate shape, continuity at contact, the heel
transient, sample placement and scaling.  The fore-aft shear and the
per-field overrides (hump width, taper, duty factor, transient switch) are
not restated: the FDTD source is the vertical force at the defaults."""

import math

import numpy as np
import pytest

import cosim_amd.workloads as W

G = 9.80665


def walker(mass):
    """grf.Walker (grf.go:134-142)."""
    return W.Stance(Mass=mass)


def peak_force(s):
    """Stance.PeakForce (grf.go:424-431): the largest of 20,000 midpoint samples."""
    tau = (np.arange(20000) + 0.5) / 20000
    return float(np.max(s.profile_at(tau))) * s.body_weight()


@pytest.mark.parametrize("mass", [50, 75, 95, 120])
def test_impulse_balances_body_weight_over_gait_cycle(mass):
    assert abs(walker(mass).impulse_ratio() - 1) <= 1e-3


@pytest.mark.parametrize("kw", [
    dict(Mass=75, Duration=0.50, FirstPeak=1.30, SecondPeak=1.25, MidstanceValley=0.65),
    dict(Mass=60, Duration=0.70, FirstPeak=1.08, SecondPeak=1.06, MidstanceValley=0.85),
    dict(Mass=90, Duration=0.62, FirstPeak=1.20, SecondPeak=1.15, MidstanceValley=0.72),
])
def test_impulse_balance_across_plausible_gaits(kw):
    assert abs(W.Stance(**kw).impulse_ratio() - 1) <= 0.06


def test_profile_shape_matches_force_plate_literature():
    s = walker(75)
    assert abs(float(s.profile_at(0.25)) - s.FirstPeak) <= 0.01
    assert abs(float(s.profile_at(0.75)) - s.SecondPeak) <= 0.01
    mid = float(s.profile_at(0.5))
    assert s.MidstanceValley <= mid <= s.MidstanceValley + 0.02
    assert mid < float(s.profile_at(0.25)) and mid < float(s.profile_at(0.75))
    nominal = s.FirstPeak * s.body_weight()
    assert abs(peak_force(s) - nominal) / nominal <= 0.02


def test_force_is_continuous_at_contact_boundaries():
    s = walker(75)
    for tau in [-0.1, -1e-9, 0, 1, 1 + 1e-9, 1.5]:
        assert float(s.profile_at(tau)) == 0.0
    prev = math.inf
    for tau in [1e-4, 1e-5, 1e-6, 1e-7, 1e-8]:
        got = float(s.profile_at(tau))
        assert 0 < got < prev
        prev = got
    assert float(s.profile_at(1e-8)) <= 1e-6
    for tau in [1 - 1e-6, 1 - 1e-8]:
        assert float(s.profile_at(tau)) <= 1e-4


def test_heel_strike_transient():
    s = walker(75)
    tau = np.arange(1, 20000) / 20000
    tr = s.transient_at(tau)
    k = int(np.argmax(tr))
    assert abs(tr[k] - 0.35) <= 1e-3                      # defaultTransientPeak
    want = 0.012 / s.Duration                              # defaultTransientRise
    assert abs(tau[k] - want) / want <= 0.15
    assert float(s.profile_at(want)) >= 0.30               # dominates the first tens of ms
    assert abs(float(s.transient_at(0.080 / s.Duration))) <= 0.05   # rung down by 80 ms


def test_sample_places_stance_correctly():
    fs, lead, tail = 1000.0, 0.2, 0.3
    s = walker(75)
    got = s.sample(fs, lead, tail)
    assert len(got) == int(round((lead + s.Duration + tail) * fs))
    assert np.all(got[: int(lead * fs)] == 0)
    assert np.all(got[int(math.ceil((lead + s.Duration) * fs)) + 1:] == 0)
    at = int(np.argmax(got))
    assert abs(at - (lead + 0.25 * s.Duration) * fs) <= 5
    assert abs(float(got[at]) - peak_force(s)) / float(got[at]) <= 2e-3


def test_force_scales_with_body_weight():
    assert abs(peak_force(walker(100)) / peak_force(walker(50)) - 2) <= 1e-9
    assert abs(walker(75).body_weight() - 75 * G) <= 1e-9


def test_config2_footstep_is_momentum_balanced():
    """Config 2's stance solves its valley from the balance (BalancedValley,
    grf.go:394-416): the ratio is then 1 to the bisection's precision, and the
    valley lies below both peaks (Validate, grf.go:155)."""
    s = W.c2_stance()
    assert abs(s.impulse_ratio() - 1) <= 1e-12
    assert 0 < s.MidstanceValley < min(s.FirstPeak, s.SecondPeak)
    assert s.Mass == 80 and s.FirstPeak == 1.2 and s.SecondPeak == 1.1 and s.TransientRise == 0.025
