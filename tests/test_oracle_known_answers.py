"""The FP64 oracle pinned against the reference's own FDTD tests
(internal/fdtd/fdtd_test.go:57-316) and the numbers those tests logged in the
reference's generated validation report (docs/validation-report.md:78-81,
186, 191-192, 197-199).  A faithful restatement of the Go scheme reproduces
the logged values to every printed digit; each test formats its numbers with
the Go test's own format string and compares the text.

What each pins: Boussinesq the source normalisation, axis rule, free surface
and half-cell receiver extrapolation; Rayleigh the propagation speed;
C-PML the absorbing layer; attenuation the memory variables; stability the
time step; zero relaxation the elastic switch.
"""

import cmath
import math

import numpy as np
import pytest

from media import loam, poisson_ratio, rayleigh_velocity


def peak_at(v, dt, t0):
    """peakAt (fdtd_test.go:21-38): parabola-refined time of the largest |v|."""
    a = np.abs(v)
    best = int(np.argmax(a))
    amp = a[best]
    at = t0 + best * dt
    if 0 < best < len(v) - 1:
        l, r = a[best - 1], a[best + 1]
        d = l - 2 * amp + r
        if d != 0:
            at += -0.5 * (r - l) / d * dt
    return amp, at


def dft(v, f, dt, t0):
    """dft (fdtd_test.go:225-232)."""
    t = t0 + np.arange(len(v)) * dt
    return np.sum(v * np.exp(-2j * math.pi * f * t)) * dt


def test_boussinesq_static_limit(oracle, fdtd):
    m = fdtd.Model(Stack=loam(), MaxRange=40, Depth=40, DominantFreq=25)
    m.Spacing = oracle.spacing_for(m, 75, 15)
    dt = oracle.OracleSim(m).Dt()
    steps = int(0.3 / dt)
    sigma = 1 / (2 * math.pi * 25)
    force = [1000 * 0.5 * (1 + math.erf((i * dt - 5 * sigma) / (math.sqrt(2) * sigma))) for i in range(steps)]
    res = oracle.run(m, force, [1, 2, 3, 5], steps)
    mu = 1700 * 200.0 * 200.0
    nu = poisson_ratio(500, 200)
    lines = []
    for k, r in enumerate(res.snapped):
        u = 0.0
        for v in res.traces[k]:
            u += v * dt
        want = 1000 * (1 - nu) / (2 * math.pi * mu * r)
        ratio = u / want
        assert abs(ratio - 1) <= 0.01
        lines.append("r=%4.1f m: static displacement %.4e m, Boussinesq %.4e m, ratio %.4f" % (r, u, want, ratio))
    assert lines == [
        "r= 1.1 m: static displacement 1.3096e-06 m, Boussinesq 1.3061e-06 m, ratio 1.0027",
        "r= 2.0 m: static displacement 7.1257e-07 m, Boussinesq 7.1241e-07 m, ratio 1.0002",
        "r= 3.0 m: static displacement 4.6082e-07 m, Boussinesq 4.6097e-07 m, ratio 0.9997",
        "r= 5.0 m: static displacement 2.7964e-07 m, Boussinesq 2.7988e-07 m, ratio 0.9992",
    ]


def _stability_run(oracle, fdtd, courant):
    m = fdtd.Model(Stack=loam(), MaxRange=6, Depth=6, Spacing=0.15, DominantFreq=30, Courant=courant)
    s = oracle.OracleSim(m)
    steps = int(0.15 / s.Dt())
    s.Drive(oracle.ricker(1000, 30, s.Dt(), steps))
    for _ in range(steps):
        s.Step()
    return s.MaxSpeed()


def test_stability_limit(oracle, fdtd):
    stable = _stability_run(oracle, fdtd, 0.5)
    assert not math.isnan(stable) and stable <= 1
    bad = _stability_run(oracle, fdtd, 1.4)
    assert math.isnan(bad) or bad >= 1e3 * stable


def test_rayleigh_velocity(oracle, fdtd):
    m = fdtd.Model(Stack=loam(), MaxRange=22, Depth=16, DominantFreq=30)
    m.Spacing = oracle.spacing_for(m, 90, 20)
    dt = oracle.OracleSim(m).Dt()
    steps = int(0.35 / dt)
    force = oracle.ricker(1000, 30, dt, steps)
    res = oracle.run(m, force, [8, 10, 12, 14, 16], steps)
    sx = sy = sxx = sxy = 0.0
    n = float(len(res.snapped))
    for k, r in enumerate(res.snapped):
        _, at = peak_at(res.traces[k], res.dt, res.t0)
        sx, sy, sxx, sxy = sx + r, sy + at, sxx + r * r, sxy + r * at
    slope = (n * sxy - sx * sy) / (n * sxx - sx * sx)
    got = 1 / slope
    want = rayleigh_velocity(500, 200)
    assert abs(got / want - 1) <= 0.02
    line = "Rayleigh velocity: grid %.2f m/s, secular equation %.2f m/s (%.2f%%)" % (got, want, 100 * (got / want - 1))
    assert line == "Rayleigh velocity: grid 187.18 m/s, secular equation 188.57 m/s (-0.74%)"


def test_pml_reflection(oracle, fdtd):
    def shoot(rmax):
        m = fdtd.Model(Stack=loam(), MaxRange=rmax, Depth=14, DominantFreq=30, Spacing=0.15)
        dt = oracle.OracleSim(m).Dt()
        steps = int(0.45 / dt)
        return oracle.run(m, oracle.ricker(1000, 30, dt, steps), [4, 8], steps)

    near, far = shoot(12), shoot(36)
    lines = []
    for k in range(2):
        peak = float(np.max(np.abs(far.traces[k])))
        diff = float(np.max(np.abs(near.traces[k] - far.traces[k])))
        db = 20 * math.log10(diff / peak)
        assert db <= -40
        lines.append("r=%4.1f m: boundary returns %.2e of a peak of %.2e (%.1f dB)" % (near.snapped[k], diff, peak, db))
    assert lines == [
        "r= 4.0 m: boundary returns 2.46e-08 of a peak of 9.04e-05 (-71.3 dB)",
        "r= 7.9 m: boundary returns 1.75e-08 of a peak of 5.84e-05 (-70.5 dB)",
    ]


def test_attenuation_matches_the_material_model(oracle, fdtd):
    f0, qt, fref = 30.0, 25.0, 30.0
    sls = oracle.visco_fit(qt, fref)
    ranges = [6, 18]

    def shoot(relax):
        m = fdtd.Model(Stack=loam(), Relax=relax, RefFreq=fref, MaxRange=26, Depth=18, DominantFreq=f0, Spacing=0.12)
        dt = oracle.OracleSim(m).Dt()
        steps = int(0.45 / dt)
        return oracle.run(m, oracle.ricker(1000, f0, dt, steps), ranges, steps)

    elastic, lossy = shoot(fdtd.SLS()), shoot(sls)
    c_r = rayleigh_velocity(500, 200)
    dr = ranges[1] - ranges[0]
    lines = []
    for f in [15, 30, 50]:
        def ratio(res):
            a = abs(dft(res.traces[0], f, res.dt, res.t0))
            b = abs(dft(res.traces[1], f, res.dt, res.t0))
            return b / a

        got = ratio(lossy) / ratio(elastic)
        k = 2 * math.pi * f / oracle.sls_velocity(sls, c_r, fref, f)
        want = math.exp(k.imag * dr)
        assert abs(got / want - 1) <= 0.05
        lines.append("%4.0f Hz: measured decay over %g m %.4f, material model %.4f (Q = %.1f)"
                     % (f, dr, got, want, oracle.sls_q(sls, f)))
    assert lines == [
        "  15 Hz: measured decay over 12 m 0.8959, material model 0.9075 (Q = 31.3)",
        "  30 Hz: measured decay over 12 m 0.7860, material model 0.7868 (Q = 25.0)",
        "  50 Hz: measured decay over 12 m 0.7171, material model 0.7051 (Q = 28.3)",
    ]


def test_zero_relaxation_is_elastic(oracle, fdtd):
    m = fdtd.Model(Stack=loam(), MaxRange=6, Depth=6, Spacing=0.15, DominantFreq=30)
    s = oracle.OracleSim(m)
    assert not s.lossy()
    assert not s.has_memory()
    bad = fdtd.Model(Stack=loam(), MaxRange=6, Depth=6, Spacing=0.15, Relax=fdtd.SLS(1e-2, 1e-3))
    with pytest.raises(oracle.OracleError, match="fdtd: TauEps must be at least TauSigma for a dissipative solid"):
        oracle.OracleSim(bad)
