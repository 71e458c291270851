"""The GPU f-k solver against the FP64 oracle, and the properties the
reference's f-k tests assert, run on the GPU path itself.

Tolerances: the GPU computes in float64 with fused multiply-adds, the
device's libm and the quadrature summed per workgroup; the oracle in float64
without fusion, glibc's libm and a sequential sum.  The layered solve
amplifies rounding where it is ill-conditioned: past k*h = 16
(fk.go:225-247), where P and SV eigenvectors nearly coincide (k far above
the body-wave wavenumbers, fk_test.go:44-54), and where the integral cancels
(low frequencies).  Each comparison is therefore calibrated by the oracle's
own sensitivity to rounding: the oracle built with fused multiply-adds
(oracle/Makefile) and the oracle on the medium moved by about one ulp give
two measures of it, and the GPU must agree with the oracle within 10 times
the larger, or 1e-10 relative (what the two libms alone leave in
well-conditioned cases), whichever is larger.  Measured, the GPU's
difference is within a factor 2 of the fused build's everywhere.  Where the
solve is well conditioned (point responses with k*h <= 16) the bound is a
flat 1e-9.  Integrated results are compared with the oracle summing in the
GPU's order, and that order is checked against the Go order separately.
"""

import math

import numpy as np
import pytest

from conftest import load_package

pkg = load_package()
fd, fk = pkg.fdtd, pkg.fk

pytestmark = pytest.mark.gpu

TWO_PI = 2 * math.pi

LOAM = [fd.Layer(0, 500, 200, 1700, 25)]
LAYERED = [fd.Layer(1.5, 300, 150, 1700, 20), fd.Layer(4, 500, 250, 1850, 40), fd.Layer(0, 800, 400, 2000, 80)]
SLS = fd.Fit(25, 30)

MEDIA = {
    "loam": fk.Medium(Stack=LOAM),
    "loam-lossless": fk.Medium(Stack=[fd.Layer(0, 500, 200, 1700, 1e9)], DefaultQ=1e9),
    "layered": fk.Medium(Stack=LAYERED),
    "layered-sls": fk.Medium(Stack=LAYERED, Relax=SLS, RefFreq=30),
    "rock-over-soil": fk.Medium(Stack=[fd.Layer(3, 400, 160, 1600, 0), fd.Layer(0, 2200, 1100, 2400, 0)],
                                DefaultQ=12),
}


def rel(a, b):
    return np.max(np.abs(np.asarray(a) - np.asarray(b))) / np.max(np.abs(np.asarray(b)))


def nudged(m, e):
    """The medium with every velocity and density scaled by (1 + e)."""
    st = [fd.Layer(l.Thickness, l.Vp * (1 + e), l.Vs * (1 + e), l.Density * (1 + e), l.Qs, l.Qp) for l in m.Stack]
    return fk.Medium(Stack=st, RefFreq=m.RefFreq, DefaultQ=m.DefaultQ, Relax=m.Relax)


NUDGES = (2.3e-16, -3.4e-16)


def bound(call, m, want):
    """10x the oracle's own rounding sensitivity, at least 1e-10.  `call`
    takes (medium, fma) and evaluates the oracle."""
    spread = max([rel(call(m, True), want)] + [rel(call(nudged(m, e), False), want) for e in NUDGES])
    return max(10 * spread, 1e-10)


@pytest.mark.parametrize("name", list(MEDIA))
def test_surface_response_equals_the_oracle(oracle, name):
    m = MEDIA[name]
    print(f"[fk] surface response {name}", flush=True)
    ks = np.array([1e-4, 0.003, 0.05, 0.3, 1.0, 2.5, 7.0, 20.0, 60.0])
    fs = np.array([0.5, 5.0, 20.0, 47.3, 90.0])
    K, F = np.meshgrid(ks, fs)
    got = m.SurfaceResponses(K, F)
    # Past k*h = 16 (the reference's MaxLayerPhase, fk.go:248) the direct
    # propagator amplifies rounding (fk.go:225-247), so there the two FP64
    # evaluations agree less closely.
    hmax = max([l.Thickness for l in m.Stack[:-1]], default=0.0)
    worst = {True: 0.0, False: 0.0}
    for i in range(K.shape[0]):
        for j in range(K.shape[1]):
            k, f = K[i, j], F[i, j]
            want = oracle.fk_surface_response(m, k, f)
            err = abs(got[i, j] - want) / abs(want)
            well = k * hmax <= 16
            worst[well] = max(worst[well], err)
            tol = 1e-9 if well else bound(lambda mm, fma: oracle.fk_surface_response(mm, k, f, fma), m, want)
            assert err <= tol, (k, f, got[i, j], want, tol)
    print(f"[fk] {name}: worst rel diff {worst[True]:.1e} (k*h <= 16), {worst[False]:.1e} beyond", flush=True)


@pytest.mark.parametrize("name", ["loam", "layered", "layered-sls", "rock-over-soil"])
def test_multi_equals_the_oracle(oracle, name):
    m = MEDIA[name]
    ranges = [0.5, 1.0, 2.25, 7.0, 16.0, 33.3, 50.0]
    freqs = [3.0, 12.5, 30.0, 61.0, 80.0]
    print(f"[fk] multi {name}", flush=True)
    got = m.VerticalDisplacementMulti(ranges, freqs)
    assert got.shape == (len(freqs), len(ranges))
    for i, f in enumerate(freqs):
        call = lambda mm, fma: oracle.fk_multi_workgroup_order(mm, ranges, f, None, fma)
        want = call(m, False)
        tol = bound(call, m, want)
        print(f"[fk] {name} {f} Hz: rel diff {rel(got[i], want):.1e}, bound {tol:.1e}", flush=True)
        assert rel(got[i], want) <= tol, (f, rel(got[i], want), tol)
        # The GPU's summation order against the reference's sequential one.
        assert rel(want, oracle.fk_vertical_displacement_multi(m, ranges, f)) < 1e-12
    st = fk.Stats()
    assert st["frequencies"] == len(freqs)
    assert st["samples"] == sum(m.GridFor(ranges, f).Samples for f in freqs)
    assert st["weights"] == st["samples"] * len(ranges)
    assert st["kernel_ms"] > 0


def test_single_range_and_static_coefficient_equal_the_oracle(oracle):
    m = MEDIA["layered"]
    for r, f, opt in [(3.0, 25.0, fk.Integration()), (12.0, 70.0, fk.Integration(Samples=3000, KMaxFactor=20))]:
        got = m.VerticalDisplacement(r, f, opt)
        call = lambda mm, fma: oracle.fk_vd_workgroup_order(mm, r, f, opt, fma)
        want = call(m, False)
        assert abs(got - want) / abs(want) <= bound(call, m, want)
        assert abs(want - oracle.fk_vertical_displacement(m, r, f, opt)) / abs(want) < 1e-12
    got = m.StaticCoefficient(40.0, 300.0)
    call = lambda mm, fma: oracle.fk_static_coefficient(mm, 40.0, 300.0, fma)
    want = call(m, False)
    assert abs(got - want) / abs(want) <= bound(call, m, want)


def test_a_batch_equals_its_frequencies_one_at_a_time():
    m = MEDIA["layered-sls"]
    ranges = [1.0, 4.0, 9.5]
    freqs = [7.0, 19.0, 44.0, 71.0]
    batch = m.VerticalDisplacementMulti(ranges, freqs)
    for i, f in enumerate(freqs):
        one = m.VerticalDisplacementMulti(ranges, f)
        assert np.array_equal(batch[i], one)


@pytest.mark.parametrize("name,vp,vs,rho", [("loam", 500, 200, 1700), ("dry sand", 350, 150, 1600),
                                            ("weathered rock", 1800, 800, 2300)])
def test_boussinesq_limits_on_the_gpu(name, vp, vs, rho):
    m = fk.Medium(Stack=[fd.Layer(0, vp, vs, rho, 1e9)], DefaultQ=1e9)
    nu = (vp * vp - 2 * vs * vs) / (2 * (vp * vp - vs * vs))
    coef = (1 - nu) / (TWO_PI * rho * vs * vs)
    u = m.SurfaceResponses([1.0, 10.0, 200.0], 0.05)
    assert np.all(np.abs(np.array([1.0, 10.0, 200.0]) * np.abs(u) - coef) / coef < 1e-3)
    if name == "loam":
        for r in (2, 5, 10, 20):
            got = abs(m.VerticalDisplacement(r, 0.02, fk.Integration(Samples=8000)))
            assert abs(got - coef / r) / (coef / r) < 0.02


def test_quadrature_is_second_order_on_the_gpu():
    m = MEDIA["loam"]
    r, f = 16.0, 60.0
    n0 = m.GridFor([r], f).Samples
    vd = lambda n: m.VerticalDisplacement(r, f, fk.Integration(Samples=n))
    ref = (4 * vd(32 * n0) - vd(16 * n0)) / 3
    errs = [abs(vd(n) - ref) / abs(ref) for n in (n0, 2 * n0, 4 * n0)]
    assert errs[0] < 1e-3
    assert all(math.log2(a / b) > 1.5 for a, b in zip(errs, errs[1:]))


def test_identical_layers_equal_their_half_space_on_the_gpu():
    half = fk.Medium(Stack=LOAM, DefaultQ=30)
    lay = fd.Layer(2, 500, 200, 1700, 25)
    split = fk.Medium(Stack=[lay, lay, fd.Layer(0, 500, 200, 1700, 25)], DefaultQ=30)
    a = half.VerticalDisplacementMulti([10.0], [10.0, 30.0, 100.0, 200.0])
    b = split.VerticalDisplacementMulti([10.0], [10.0, 30.0, 100.0, 200.0])
    assert np.all(np.abs(a - b) / np.abs(a) < 1e-5)


ERRORS = [
    ("multi", dict(ranges=[3.0, 0.0], freqs=[20.0]), "fk: range must be positive, got 0 m"),
    ("multi", dict(ranges=[], freqs=[20.0]), "fk: no ranges given"),
    ("multi", dict(ranges=[3.0], freqs=[20.0, -5.0]), None),
    ("multi", dict(ranges=[3.0], freqs=[20.0, 600.0]), None),
    ("vd", dict(r=0.0, f=10.0), "fk: range must be positive, got 0 m"),
    ("vd", dict(r=5.0, f=0.0), "fk: frequency must be positive, got 0 Hz"),
    ("surface", dict(k=0.0, f=10.0), "fk: wavenumber and frequency must be positive, got 0 and 10"),
]


@pytest.mark.parametrize("kind,args,message", ERRORS)
def test_errors_carry_the_reference_text(oracle, kind, args, message):
    m = MEDIA["layered"]
    with pytest.raises(fd.FdtdError) as got:
        if kind == "multi":
            m.VerticalDisplacementMulti(args["ranges"], args["freqs"])
        elif kind == "vd":
            m.VerticalDisplacement(args["r"], args["f"])
        else:
            m.SurfaceResponse(args["k"], args["f"])
    if message is None:
        # The first frequency the Go calls would fail on, as the oracle says.
        with pytest.raises(oracle.OracleError) as want:
            for f in args["freqs"]:
                oracle.fk_vertical_displacement_multi(m, args["ranges"], f)
        message = str(want.value)
    assert str(got.value) == message
    bad = fk.Medium(Stack=[fd.Layer(0, 100, 200, 1700)])
    with pytest.raises(fd.FdtdError, match=r"layer 0: Vp=100 must exceed sqrt\(2\)\*Vs=282.842712474619"):
        bad.VerticalDisplacementMulti([1.0], 10.0)


def test_synthesise_equals_the_oracle(oracle):
    # geofdtd -compare's reference on a small case: loam with a fitted SLS,
    # a 40 Hz Ricker, dt 2e-4, 1500 instants (nfft 2048), band 140 Hz.
    dt, n, f0, band, peak = 2e-4, 1500, 40.0, 140.0, 800.0
    ranges = [1.0, 2.5, 6.0]
    sls = fd.Fit(30, 30)
    t0 = dt / 2
    print("[fk] synthesise", flush=True)
    got = fk.Synthesise(LOAM, sls, 30, ranges, peak, f0, band, dt, t0, n, 0)
    med = fk.Medium(Stack=LOAM, Relax=sls, RefFreq=30)
    want = oracle.fk_synthesise(med, ranges, peak, f0, band, dt, t0, n)
    assert got.shape == want.shape == (3, n)
    for i in range(3):
        assert np.max(np.abs(got[i] - want[i])) <= 1e-9 * np.max(np.abs(want[i])), i


def test_fill_bank_equals_the_oracle_bank():
    """cmd/geobank's build loop on the device (gf_fk_bank_fill) against the
    oracle filling the same grid as the reference's smallBank does, to float32
    storage; bin 0 and the bins above the band stay zero; the range-sampling
    check refuses a grid too coarse for the band."""
    from test_bank_lookup_oracle import small_bank

    layers = [fd.Layer(0, 500, 200, 1700, 25)]
    h = fd.BankHeader(Solver="geosim fk (layered wavenumber integration)", Medium=layers, SampleRateHz=2000,
                      Samples=512, MinM=2, MaxM=12, Count=101,
                      Component="vertical surface displacement per unit vertical surface point force")
    print("[fk] fill bank 101 ranges x 25 bins", flush=True)
    got = fk.FillBank(fk.Medium(Stack=layers, DefaultQ=30), h, 100)
    d = small_bank(2, 12, 101, 100)
    want = d[..., 0].astype(np.float64) + 1j * d[..., 1].astype(np.float64)
    assert got.shape == want.shape == (101, 257)
    assert np.all(got[:, 0] == 0) and np.all(got[:, 26:] == 0)
    for k in range(1, 26):
        assert np.max(np.abs(got[:, k] - want[:, k])) <= 1e-6 * np.max(np.abs(want[:, k])), k
    coarse = fd.BankHeader(Solver="t", Medium=layers, SampleRateHz=2000, Samples=512, MinM=1, MaxM=40, Count=40,
                           Component="x")
    with pytest.raises(fd.FdtdError) as e:
        fk.FillBank(fk.Medium(Stack=layers), coarse, 300)
    assert str(e.value) == ("bank: range spacing 1.0000 m is too coarse to unwrap phase up to 300 Hz; "
                            "need under 0.1450 m")
    layered = [fd.Layer(2, 500, 200, 1700, 25), fd.Layer(0, 900, 400, 2000, 60)]
    hl = fd.BankHeader(Solver="t", Medium=layered, SampleRateHz=2000, Samples=512, MinM=2, MaxM=4, Count=41,
                       Component="x")
    with pytest.raises(fd.FdtdError) as e:
        fk.FillBank(fk.Medium(Stack=layered), hl, 400)
    assert str(e.value).startswith("bin 57 (222.656 Hz): fk: layer 0 is 2 m thick, which is k*h = 16.1 at 222.65625 Hz")
