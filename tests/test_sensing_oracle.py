"""The runtime-synthesis oracle (oracle/sensing_oracle.py) pinned against the
reference's source tests (internal/source/source_test.go) and closed forms,
and the engine's host-side logic (sensing.py: the walking schedule, the
footfall forces, the argument checks) against the oracle.  CPU only.

  * stride, cadence and speed agree (source_test.go:18-37); stances overlap
    in double support (:40-54); feet alternate either side of the path with
    stable names (:57-87); windows tile without gaps or repeats (:141-177);
    a walk past closes then opens and RangeSpan brackets it (:179-213); a
    walk can be bounded (:240-262);
  * the range quantum rounds half away from zero, with a floor of one
    quantum (sensing.go:178, 186-190);
  * Reference: a unit force sample returns the impulse response delayed to
    it (a closed form), a contact's output equals the brute-force double sum,
    contacts outside the span are ignored, and moving the receiver equals
    moving every contact the other way;
  * the heading rotates the shear but not the weight, force is zero outside
    contact (source_test.go:90-139), and Validate's errors (:215-237);
  * the analytic propagation: the Poisson solid's Rayleigh velocity in closed
    form, the library's root equal to the oracle's, free-surface tractions
    vanishing (green_test.go:21-61), the Poisson solid's ellipticity 0.68127
    at every frequency and scale (:63-84), the radial-force response the
    ellipticity times a quarter cycle (:503-551), spreading as r^-1/2 with the
    constant-Q loss (:109-194), soil validation texts (soil.go:72-89).
"""

import math

import numpy as np
import pytest

import sensing_oracle as SO
from conftest import load_package

pkg = load_package()
se, wl, fd = pkg.sensing, pkg.workloads, pkg.fdtd


def walker():
    """source_test.go:11-13: grf.Walker(75), 1.3 m/s from (-8, 3), heading 0."""
    return SO.WalkSpec(duration=0.62, speed=1.3, start_x=-8, start_y=3)


def product_walker():
    return se.Walk(Stance=wl.Stance(Mass=75), Speed=1.3, StartX=-8, StartY=3)


def test_speed_cadence_and_stance_are_consistent():
    w = walker()
    cycle = w.duration / SO.STANCE_FRACTION
    assert abs(SO.stride_or_default(w) - w.speed * cycle) <= 1e-12
    assert abs(2 * SO.step_period(w) - cycle) <= 1e-12
    c0, c100 = SO.walk_contact_at(w, 0), SO.walk_contact_at(w, 100)
    v = math.hypot(c100.x - c0.x, c100.y - c0.y) / (c100.start - c0.start)
    assert abs(v - w.speed) / w.speed <= 1e-9


def test_stances_overlap_in_double_support():
    w = walker()
    a, b = SO.walk_contact_at(w, 0), SO.walk_contact_at(w, 1)
    assert a.end > b.start
    frac = (a.end - b.start) / (2 * SO.step_period(w))
    assert 0.05 <= frac <= 0.25


def test_feet_alternate_either_side_of_the_path():
    w = walker()
    for n in range(6):
        c = SO.walk_contact_at(w, n)
        want = SO.DEFAULT_WIDTH / 2 * (-1 if n % 2 else 1)
        assert abs((c.y - w.start_y) - want) <= 1e-12
        assert c.id == ("right_foot" if n % 2 else "left_foot")


def test_contact_windows_tile_without_gaps_or_repeats():
    w = walker()
    whole = SO.walk_contacts(w, 0, 10.0)
    assert len(whole) >= 10
    pieces, t0 = [], 0.0
    while t0 < 10.0:
        pieces += SO.walk_contacts(w, t0, t0 + 0.01)
        t0 += 0.01
    assert [(c.id, c.start) for c in pieces] == [(c.id, c.start) for c in whole]
    assert all(whole[i].start >= whole[i - 1].start for i in range(1, len(whole)))
    assert SO.walk_contacts(w, 5, 5) == [] and SO.walk_contacts(w, 5, 4) == []


def test_walk_past_range_closes_then_opens():
    w = walker()
    w.start_x, w.start_y = -10, 4
    cs = SO.walk_contacts(w, 0, 16)
    ranges = [math.hypot(c.x, c.y) for c in cs]
    assert len(ranges) >= 8
    at = int(np.argmin(ranges))
    assert 0 < at < len(ranges) - 1
    assert abs(ranges[at] - 4) <= 0.5
    lo, hi, ok = SO.range_span(cs, 0, 0)
    assert ok and abs(lo - ranges[at]) <= 1e-9 and hi >= 9
    assert SO.range_span(SO.walk_contacts(w, 100, 100), 0, 0)[2] is False


def test_walk_can_be_bounded():
    w = walker()
    w.until = 5
    cs = SO.walk_contacts(w, 0, 100)
    assert cs and all(c.start < 5 for c in cs)
    assert SO.walk_contacts(w, 5, 100) == [] and SO.walk_contacts(w, 6, 7) == []
    assert len(SO.walk_contacts(walker(), 0, 100)) > len(cs)


def test_range_quantum_rounds_half_away_from_zero():
    # 0.05 / 0.02 and 8.01 / 0.02 are exact halves in float64 (2.5, 400.5):
    # math.Round takes them up, where round-half-even would not.
    assert SO.quantized_range(0.05) == 3 * 0.02
    assert SO.quantized_range(8.01) == 401 * 0.02
    assert SO.quantized_range(8.009) == 400 * 0.02
    assert SO.quantized_range(0.0) == 0.02 and SO.quantized_range(0.009) == 0.02 and SO.quantized_range(0.01) == 0.02
    # One quantum of range is one cache slot.
    assert SO.quantized_range(4.0) != SO.quantized_range(4.02)


def test_impulse_length_rule():
    # far / c_R + 2 s of coda at fs, capped at three quarters of the transform.
    assert SO.impulse_length(5.0, 180.0, 1000.0, 0) == math.ceil((5 / 180 + 2) * 1000)
    assert SO.impulse_length(5.0, 180.0, 1000.0, 1024) == 768


def _impulse(rq):
    """A range-dependent test response: a decaying sinusoid delayed by rq."""
    i = np.arange(300)
    d = int(rq * 10)
    h = np.zeros(300)
    h[d:] = np.exp(-(i[: 300 - d]) / 40) * np.sin(0.3 * i[: 300 - d] + rq)
    return h


def test_reference_unit_force_returns_the_delayed_response():
    fs, n = 1000.0, 800
    c = SO.OContact("a", 3.0, 4.0, 0.1, 0.2)
    s0 = 137
    force = lambda t: (np.round(t * fs) == s0).astype(float)  # noqa: E731
    got = SO.reference([c], [force], n, fs, _impulse)
    h = _impulse(5.0)
    want = np.zeros(n)
    want[s0:s0 + len(h)] = h[: n - s0]
    assert np.array_equal(got, want)


def test_reference_equals_the_brute_force_sum():
    fs, n = 500.0, 400
    rng = np.random.default_rng(4)
    g = rng.standard_normal(n)
    cs = [SO.OContact("a", 1.0, 2.0, 0.0, 0.3), SO.OContact("b", -2.5, 0.7, 0.5, 0.9),
          SO.OContact("late", 1.0, 1.0, n / fs, n / fs + 0.1)]
    forces = [lambda t: g * (t < 0.3), lambda t: np.cos(t * 40) * (t >= 0.5), lambda t: np.ones_like(t)]
    got = SO.reference(cs, forces, n, fs, _impulse)
    t = np.arange(n) / fs
    want = np.zeros(n)
    for c, f in zip(cs[:2], forces[:2]):
        x, h = f(t), _impulse(SO.quantized_range(math.hypot(c.x, c.y)))
        for s in range(n):
            want[s] += sum(x[m] * h[s - m] for m in range(max(0, s - len(h) + 1), s + 1))
    assert np.max(np.abs(got - want)) <= 1e-12 * np.max(np.abs(want))


def test_moving_the_receiver_is_moving_the_contacts():
    fs, n = 500.0, 300
    cs = [SO.OContact("a", 1.3, 2.0, 0.0, 0.3), SO.OContact("b", -2.5, 0.7, 0.1, 0.4)]
    forces = [lambda t: np.sin(30 * t) * (t < 0.3), lambda t: (t >= 0.1) * (t < 0.4) * 1.0]
    moved = [SO.OContact(c.id, c.x - 0.4, c.y + 1.1, c.start, c.end) for c in cs]
    a = SO.reference(cs, forces, n, fs, _impulse, rx=0.4, ry=-1.1)
    b = SO.reference(moved, forces, n, fs, _impulse)
    assert np.array_equal(a, b)


# ---- the product's host logic against the oracle -------------------------


def test_walk_schedule_equals_the_oracle():
    for kw in ({}, {"Heading": 0.7, "Width": 0.15, "StrideLength": 1.4}, {"Until": 4.0}):
        w = product_walker()
        o = walker()
        for k, v in kw.items():
            setattr(w, k, v)
            setattr(o, {"Heading": "heading", "Width": "width", "StrideLength": "stride", "Until": "until"}[k], v)
        got = w.Contacts(0, 12)
        want = SO.walk_contacts(o, 0, 12)
        assert [(c.ID, c.X, c.Y, c.Start, c.End()) for c in got] == [(c.id, c.x, c.y, c.start, c.end) for c in want]
        assert w.StepPeriod() == SO.step_period(o)
        lo, hi, ok = se.RangeSpan(w, 0.5, -0.2, 0, 12)
        assert (lo, hi, ok) == SO.range_span(want, 0.5, -0.2)


def test_heading_orients_the_shear_but_not_the_weight():
    stance = wl.Stance(Mass=75)
    at = 0.25 * stance.Duration
    east = se.Footfall(stance, 0.0).ForceAt(at)
    north = se.Footfall(stance, math.pi / 2).ForceAt(at)
    assert abs(east[2] - north[2]) <= 1e-12
    mag = math.hypot(east[0], east[1])
    assert mag >= 1
    assert abs(math.hypot(north[0], north[1]) - mag) <= 1e-9
    assert abs(east[1]) <= 1e-9 and abs(north[0]) <= 1e-9
    assert east[0] < 0


def test_force_is_zero_outside_contact():
    f = se.Footfall(wl.Stance(Mass=75), 0.7)
    for at in (-1, -1e-9, 0, f.Duration(), f.Duration() + 1):
        assert all(float(v) == 0 for v in f.ForceAt(at)), at
    c = se.Contact("x", 0, 0, 5.0, f)
    assert all(float(v) == 0 for v in c.ForceAt(4.9))
    assert any(float(v) != 0 for v in c.ForceAt(5.2))
    assert c.End() == 5 + f.Duration()


@pytest.mark.parametrize("field,value,message", [
    ("Speed", 0, "source: walking speed must be positive, got 0 m/s"),
    ("Width", -1, "source: stance width must not be negative, got -1 m"),
    ("StrideLength", -1, "source: stride length must not be negative, got -1 m"),
])
def test_validate_catches_impossible_walks(field, value, message):
    w = product_walker()
    setattr(w, field, value)
    with pytest.raises(fd.FdtdError) as e:
        w.Validate()
    assert str(e.value) == message
    product_walker().Validate()


def test_validate_catches_an_impossible_gait():
    w = product_walker()
    w.Stance.Mass = 0
    with pytest.raises(fd.FdtdError) as e:
        w.Validate()
    assert str(e.value) == "grf: mass must be positive, got 0 kg"


def test_engine_arguments_are_checked_before_the_device():
    w = product_walker()
    kw = dict(sample_rate=1000.0, walk_duration=10.0, geometry_range=10.0, rayleigh_velocity=180.0)
    with pytest.raises(fd.FdtdError) as e:
        se.NewEngineWith(w, 0, object(), **kw)
    assert str(e.value) == "sensing: chunk size must be positive, got 0"
    with pytest.raises(fd.FdtdError) as e:
        se.NewEngineWith(None, 16, object(), **kw)
    assert str(e.value) == "sensing: a schedule is required"
    with pytest.raises(fd.FdtdError) as e:
        se.NewEngineWith(w, 16, None, **kw)
    assert str(e.value) == "sensing: a propagation model is required"
    with pytest.raises(fd.FdtdError) as e:
        se.FromBank(None)
    assert str(e.value) == "sensing: bank is nil"
    # The impulse-response length NewEngineWith derives from the schedule.
    _, far, _ = SO.range_span(SO.walk_contacts(walker(), 0, 10.0), 0, 0)
    assert se.ImpulseLength(w, 10.0, 10.0, 180.0, 1000.0) == SO.impulse_length(far, 180.0, 1000.0, 0)
    assert se.ImpulseLength(se.Fixed([]), 10.0, 7.5, 180.0, 1000.0) == SO.impulse_length(7.5, 180.0, 1000.0, 0)


# ---- the analytic propagation (green.HalfSpaceGF over soil.HalfSpace) -----

SOILS = {"Poisson solid": SO.OHalfSpace(math.sqrt(3) * 200, 200, 1900, 1e9),
         "dry sand": SO.OHalfSpace(350, 150, 1600, 20), "loam": SO.OHalfSpace(500, 200, 1700, 25),
         "weathered rock": SO.OHalfSpace(1800, 800, 2300, 80)}


def test_rayleigh_velocity_of_a_poisson_solid():
    """The closed form c_R = Vs sqrt(2 - 2/sqrt 3) (soil.go:62-66's fixture),
    and the library's bisection equal to the oracle's."""
    h = SOILS["Poisson solid"]
    assert abs(SO.rayleigh_velocity(h) / 200 - math.sqrt(2 - 2 / math.sqrt(3))) <= 1e-12
    for name, o in SOILS.items():
        got = se.HalfSpace(o.vp, o.vs, o.rho, o.qs).RayleighVelocity()
        assert got == SO.rayleigh_velocity(o), name


def test_free_surface_tractions_vanish():
    """green_test.go:21-61: the mode shapes satisfy the free-surface conditions."""
    for name, h in SOILS.items():
        mu, lam = h.rho * h.vs ** 2, h.rho * (h.vp ** 2 - 2 * h.vs ** 2)
        for f in (5, 20, 100):
            k, q, s, xq, xs, zq, zs = SO.eigenfunctions(h, 2 * math.pi * f)
            xp, zp = -q * xq - s * xs, -q * zq - s * zs
            v0, h0 = zq + zs, xq + xs
            assert abs(xp + k * v0) / (abs(xp) + abs(k * v0)) <= 1e-12, (name, f)
            normal = lam * (-k * h0 + zp) + 2 * mu * zp
            assert abs(normal) / (abs(lam * k * h0) + abs(2 * mu * zp)) <= 1e-12, (name, f)


def test_ellipticity_of_a_poisson_solid():
    """green_test.go:63-84: 0.68127 at every frequency and medium scale."""
    h = SOILS["Poisson solid"]
    for f in (1, 20, 500):
        e = SO.eigenfunctions(h, 2 * math.pi * f)
        assert abs(abs((e[3] + e[4]) / (e[5] + e[6])) - 0.68127) <= 1e-4
    big = SO.OHalfSpace(math.sqrt(3) * 2000, 2000, 2600, 1e9)
    e1, e2 = SO.eigenfunctions(h, 2 * math.pi * 20), SO.eigenfunctions(big, 2 * math.pi * 20)
    assert abs(abs((e1[3] + e1[4]) / (e1[5] + e1[6])) - abs((e2[3] + e2[4]) / (e2[5] + e2[6]))) <= 1e-12


def test_radial_force_response_is_ellipticity_times_a_quarter_cycle():
    """green_test.go:503-551."""
    import cmath

    for name in ("Poisson solid", "dry sand", "loam"):
        h = SOILS[name]
        for f in (5, 20, 100):
            v, r = SO.velocity_response(h, 10, f), SO.radial_force_response(h, 10, f)
            e = SO.eigenfunctions(h, 2 * math.pi * f)
            assert abs(abs(r) / abs(v) - abs((e[3] + e[4]) / (e[5] + e[6]))) <= 1e-12
            d = (cmath.phase(r) - cmath.phase(v)) * 180 / math.pi
            d = (d + 180) % 360 - 180
            assert abs(abs(d) - 90) <= 1e-9
    h = SOILS["Poisson solid"]
    assert abs(abs(SO.radial_force_response(h, 10, 20)) / abs(SO.velocity_response(h, 10, 20)) - 0.68127) <= 1e-4


def test_geometric_spreading_and_attenuation():
    """green_test.go:109-194: |G| sqrt(r) e^{-Im k r} is range-independent,
    with Im k = -|k| sin(pi g / 2), g = atan(1/Q) / pi (Kjartansson)."""
    h = SOILS["loam"]
    f = 30.0
    v = [SO.velocity_response(h, r, f) for r in (5.0, 10.0, 20.0, 40.0)]
    cr = SO.rayleigh_velocity(h)
    g = math.atan(1 / h.qs) / math.pi
    mag = 2 * math.pi * f / (cr * math.cos(math.pi * g / 2))  # |k| at the reference frequency
    kim = -mag * math.sin(math.pi * g / 2)
    norm = [abs(x) * math.sqrt(r) * math.exp(-kim * r) for x, r in zip(v, (5.0, 10.0, 20.0, 40.0))]
    assert max(norm) / min(norm) - 1 <= 1e-12


def test_soil_validation_texts():
    """soil.go:72-89, through the library."""
    cases = [(se.HalfSpace(500, 0, 1700, 25), "soil: Vs must be positive, got 0 m/s"),
             (se.HalfSpace(0, 200, 1700, 25), "soil: Vp must be positive, got 0 m/s"),
             (se.HalfSpace(500, 200, 0, 25), "soil: density must be positive, got 0 kg/m^3"),
             (se.HalfSpace(250, 200, 1700, 25),
              "soil: Vp=250 m/s must exceed sqrt(2)*Vs=282.842712474619 m/s for a non-negative Poisson ratio"),
             (se.HalfSpace(500, 200, 1700, 0), "soil: Qs must be positive, got 0"),
             (se.HalfSpace(500, 200, 1700, 25, -1), "soil: Qp must not be negative, got -1")]
    for h, message in cases:
        with pytest.raises(fd.FdtdError) as e:
            h.RayleighVelocity()
        assert str(e.value) == message


# ---- the sensor (geophone.Geophone), evaluated at the response bins --------


def isolated():
    g = se.SM24()
    g.Coupling = se.Coupling()
    return g


def test_transducer_analytic_limits():
    """geophone_test.go:27-74: flat above resonance, S / (2 zeta) and +90
    degrees at resonance, 12 dB per octave below, nothing at DC."""
    import cmath

    g = isolated()
    G, zeta = g.Sensitivity, g.Damping()
    h = complex(g.TransducerResponse(500))
    assert abs(abs(h) - G) / G <= 1e-3 and abs(cmath.phase(h) * 180 / math.pi) <= 1
    h0 = complex(g.TransducerResponse(g.NaturalFreq))
    assert abs(abs(h0) - G / (2 * zeta)) / (G / (2 * zeta)) <= 1e-12
    assert abs(cmath.phase(h0) * 180 / math.pi - 90) <= 1e-9
    lo, hi = abs(complex(g.TransducerResponse(0.1))), abs(complex(g.TransducerResponse(0.2)))
    assert abs(hi / lo - 4) <= 0.05
    assert abs(lo - G * (0.1 / g.NaturalFreq) ** 2) / (G * (0.1 / g.NaturalFreq) ** 2) <= 5e-3
    assert abs(complex(g.TransducerResponse(0.0))) == 0


def test_shunt_damping_round_trip_and_limits():
    """geophone_test.go:76-115."""
    g = isolated()
    for want in (0.5, 0.6, 0.7, 0.9):
        loaded = isolated()
        loaded.ShuntResistance = g.ShuntForDamping(want)
        assert abs(loaded.Damping() - want) <= 1e-12
    assert 1000 <= se.SM24().ShuntForDamping(0.7) <= 10000
    for bad in (se.SM24().OpenCircuitDamping, 0.1):
        with pytest.raises(fd.FdtdError):
            se.SM24().ShuntForDamping(bad)


def test_coupling_transparent_in_band_and_poor_coupling_distorts():
    """geophone_test.go:117-146."""
    g = se.SM24()
    for f in (1, 10, 50, 100):
        assert 0.95 <= abs(complex(g.CouplingResponse(f))) <= 1.35
    assert abs(complex(g.CouplingResponse(2000))) <= 0.05
    assert abs(complex(isolated().CouplingResponse(10))) == 1
    g.Coupling = se.Coupling(30, 0.2)
    assert abs(complex(g.CouplingResponse(30))) >= 2 and abs(complex(g.CouplingResponse(100))) <= 0.2


def test_engine_rejects_an_invalid_sensor():
    """sensing_test.go:289-300: NewEngine validates the sensor first."""
    bad = se.SM24()
    bad.CoilResistance = 0
    with pytest.raises(fd.FdtdError) as e:
        se.NewEngineWith(product_walker(), 16, object(), sample_rate=1000.0, walk_duration=10.0, geometry_range=10.0,
                         rayleigh_velocity=180.0, sensor=bad)
    assert str(e.value) == "geophone: coil resistance must be positive, got 0 ohm"
