"""The device sensing engine (sensing.Engine over FromBank) against the
oracle's restatement of sensing.Reference, the monolithic sum the
reference's own tests hold the streaming engine to
(sensing_test.go:48-77), and the reference's engine properties: voices are
retired (:174-199), start times follow the sample count (:201-221), a bank
drives synthesis with a believable geophone signal (bank_test.go:66-121),
errors carry the Go text.

Tolerance: the reference's own, 1e-7 of the peak (sensing_test.go:70), which
is the float32 output's rounding; beyond it every sample must equal the
oracle's sum rounded to float32 within one float32 step (the device sums
directly in FP64, the oracle through numpy's convolution).
"""

import math

import numpy as np
import pytest

import sensing_oracle as SO
from conftest import load_package
from test_bank_lookup_oracle import RATE, SAMPLES, small_bank
from test_gpu_bank_lookup import device_bank

pkg = load_package()
fd, bk, fk, se, wl = pkg.fdtd, pkg.bank, pkg.fk, pkg.sensing, pkg.workloads

pytestmark = pytest.mark.gpu

C_R = 186.0  # Rayleigh velocity of the loam test medium (Vs 200, Vp 500), about 0.93 Vs


def geophone(f):
    f = np.asarray(f, dtype=np.float64)
    return 28.8 * (1j * f / 4.5) ** 2 / (1 + 1j * 0.7 * (f / 4.5) - (f / 4.5) ** 2)


def walk_past(range_m=4.0, approach=4.0, speed=1.3):
    """config.Resolved.WalkPast (config.go:368-381) for a 75 kg walker."""
    return se.Walk(Stance=wl.Stance(Mass=75), Speed=speed, StartX=-approach, StartY=range_m, Width=0.12,
                   Until=2 * approach / speed)


def oracle_walk(w):
    return SO.WalkSpec(duration=w.Stance.Duration, speed=w.Speed, start_x=w.StartX, start_y=w.StartY,
                       heading=w.Heading, stride=w.StrideLength, width=w.Width, until=w.Until)


def engine(b, w, chunk, fs=RATE, receivers=((0.0, 0.0),), sensor=geophone):
    return se.NewEngineWith(w, chunk, se.FromBank(b), sample_rate=fs, walk_duration=w.Until,
                            geometry_range=w.StartY, rayleigh_velocity=C_R, sensor=sensor, receivers=receivers)


def run(e, n):
    out = []
    while sum(x.shape[-1] for x in out) < n:
        out.append(e.Next().copy())
    return np.concatenate(out, axis=-1)[..., :n]


def oracle_reference(ob, rate, w, n, fs, ir_len, rx=0.0, ry=0.0, sensor=geophone):
    """sensing.Reference on the oracle: the oracle's schedule, each contact's
    vertical force from the same footfall, the oracle bank's responses."""
    contacts = SO.walk_contacts(oracle_walk(w), 0, n / fs)
    foot = se.Footfall(w.Stance, w.Heading)
    forces = [lambda t, c=c: foot.ForceAt(t - c.start)[2] for c in contacts]
    s = None if sensor is None else sensor(np.arange(ob.bins) * fs / (2 * (ob.bins - 1)))
    import oracle as O

    return SO.reference(contacts, forces, n, fs, lambda rq: O.bank_impulse(ob, rate, [rq], ir_len, fs, s)[0], rx, ry)


def assert_matches(got, want):
    scale = np.max(np.abs(want))
    assert scale > 0
    err = np.max(np.abs(got.astype(np.float64) - want))
    assert err <= 1e-7 * scale, err / scale
    w32 = want.astype(np.float32)
    step = np.spacing(np.abs(w32)).astype(np.float64)
    assert np.all(np.abs(got.astype(np.float64) - w32.astype(np.float64)) <= step + 1e-12 * scale)
    return err / scale


def test_streaming_matches_the_reference_across_voices(oracle):
    data = small_bank(2, 12, 101, 100)
    b = device_bank(data, 2, 12)
    ob = oracle.OracleBank(2, 12, data)
    w = walk_past()
    n = 8000
    want = oracle_reference(ob, RATE, w, n, RATE, 3 * SAMPLES // 4)
    for chunk in (10, 25, 100):
        e = engine(b, w, chunk)
        assert e.ImpulseLength() == 3 * SAMPLES // 4
        got = run(e, n)
        rel = assert_matches(got, want)
        print(f"[sensing] chunk {chunk}: max |engine - reference| = {rel:.2e} of the peak, "
              f"{e.Responses()} responses", flush=True)
        e.close()
    b.close()


def test_many_receivers_in_one_engine(oracle):
    data = small_bank(2, 12, 101, 100)
    b = device_bank(data, 2, 12)
    ob = oracle.OracleBank(2, 12, data)
    w = walk_past()
    rec = [(0.0, 0.0), (1.5, 0.3), (-2.0, -0.5), (3.0, 1.0)]
    n = 5000
    e = engine(b, w, 20, receivers=rec)
    got = run(e, n)
    assert got.shape == (4, n)
    for q, (rx, ry) in enumerate(rec):
        print(f"[sensing] receiver {q} at ({rx}, {ry})", flush=True)
        assert_matches(got[q], oracle_reference(ob, RATE, w, n, RATE, 3 * SAMPLES // 4, rx, ry))
    # Receiver 0 alone gives the same samples bit for bit.
    one = engine(b, w, 20)
    assert np.array_equal(run(one, n), got[0])
    one.close()
    e.close()
    b.close()


def test_voices_are_retired_and_the_coda_ends():
    data = small_bank(2, 12, 101, 100)
    b = device_bank(data, 2, 12)
    w = walk_past()
    e = engine(b, w, 50)
    most, trail = 0, []
    total = int(math.ceil((w.Until + 1.5) * RATE / 50))
    for k in range(total):
        y = e.Next()
        most = max(most, e.ActiveVoices())
        if k * 50 / RATE > w.Until + w.Stance.Duration + e.ImpulseLength() / RATE + 0.05:
            trail.append(y)
    assert 0 < most <= 12
    assert e.ActiveVoices() == 0
    assert trail and all(np.all(t == 0) for t in trail)
    e.close()
    b.close()


def test_start_time_tracks_samples_not_calls():
    data = small_bank(2, 12, 101, 100)
    b = device_bank(data, 2, 12)
    e = engine(b, walk_past(), 32)
    for i in range(20):
        assert abs(e.StartTime() - i * 32 / RATE) <= 1e-12
        e.Next()
    assert abs(e.ChunkDuration() - 32 / RATE) <= 1e-12
    assert e.ChunkSize() == 32 and e.SampleRate() == RATE
    assert e.Propagation().startswith("bank: test, 101 ranges 2-12 m")
    e.close()
    b.close()


def test_errors_carry_the_reference_text():
    data = small_bank(2, 12, 101, 100)
    b = device_bank(data, 2, 12)
    e = engine(b, walk_past(), 20)
    with pytest.raises(fd.FdtdError) as err:
        e.Next(np.zeros(7, dtype=np.float32))
    assert str(err.value) == "sensing: output length 7, want 20"
    import ctypes as C

    lib = fd.load_library()
    buf = np.zeros(7, dtype=np.float32)
    assert lib.gf_engine_next(e._h, buf.ctypes.data_as(C.POINTER(C.c_float)), 7) != 0
    assert lib.gf_last_error().decode() == "sensing: output length 7, want 20"
    e.close()
    # A contact on the receiver (admitted directly: warming it would first ask
    # the bank for 0.02 m, outside this one).
    quiet = se.Walk(Stance=wl.Stance(Mass=75), Speed=1.3, StartX=3, StartY=4, Until=0.1)
    e = se.NewEngineWith(quiet, 20, se.FromBank(b), sample_rate=RATE, walk_duration=0.1, geometry_range=4,
                         rayleigh_velocity=C_R, receivers=[(5.0, 0.0)])
    f = np.ones(10)
    rc = lib.gf_engine_admit(e._h, b"left_foot", 5.0, 0.0, f.ctypes.data_as(C.POINTER(C.c_double)), 10, 0.5)
    assert rc != 0 and lib.gf_last_error().decode() == 'sensing: contact "left_foot" at the receiver\'s own position'
    e.close()
    # Off the bank's frequency grid.
    with pytest.raises(fd.FdtdError) as err:
        engine(b, walk_past(), 20, fs=1.5 * RATE)
    assert str(err.value) == "sensing: 5.859375 Hz does not land on the bank's frequency grid (spacing 3.90625 Hz)"
    # A walk beyond the bank's ranges: the first contact's quantized range.
    far = walk_past(range_m=15.0)
    c0 = far.ContactAt(0)
    rq = SO.quantized_range(math.hypot(c0.X, c0.Y))
    with pytest.raises(fd.FdtdError) as err:
        engine(b, far, 20)
    assert str(err.value) == f"bank: range {fd._go_g(rq)} m is outside the bank's 2 to 12 m"
    b.close()


def test_a_bank_drives_synthesis(oracle):
    """bank_test.go:66-121: a bank of layered wavenumber integrals (here filled
    on the device by fk.FillBank), a walk past at 4 m, a geophone: microvolts
    to millivolts, and the streamed output equal to the oracle's Reference
    on the same bank."""
    medium = fk.Medium(Stack=[fd.Layer(0, 500, 200, 1700, 25)], DefaultQ=30)
    limit = bk.RangeNyquist(0.87 * 200, 80) / 2
    count = int(math.ceil((12 - 1) / limit)) + 1
    h = fd.BankHeader(Solver="test", Medium=[fd.Layer(0, 500, 200, 1700, 25)], SampleRateHz=2000, Samples=1024,
                      MinM=1, MaxM=12, Count=count,
                      Component="vertical surface displacement per unit vertical surface point force")
    bk.CheckRangeSampling(h, 80)
    z = fk.FillBank(medium, h, 80)
    b = bk.Bank(h, z)
    data = np.stack([z.real, z.imag], axis=-1).astype(np.float32)
    ob = oracle.OracleBank(1, 12, data)
    w = walk_past()
    e = engine(b, w, 20, fs=2000.0)
    assert e.ImpulseLength() == 768
    n = int(w.Until * 2000)
    got = run(e, n)
    peak = float(np.max(np.abs(got)))
    print(f"[sensing] f-k bank of {count} ranges, walk past at 4 m: peak {peak:.3e} V, "
          f"{e.Responses()} responses", flush=True)
    assert 1e-6 <= peak <= 1e-1
    assert_matches(got, oracle_reference(ob, 2000.0, w, n, 2000.0, 768))
    e.close()
    b.close()


def fixed_engine(b, contacts, n, chunk=25):
    sch = se.Fixed(contacts)
    return se.NewEngineWith(sch, chunk, se.FromBank(b), sample_rate=RATE, walk_duration=n / RATE,
                            geometry_range=8.0, rayleigh_velocity=C_R, sensor=geophone)


def test_voice_slots_are_bounded_and_refusal_is_clean():
    """The device engine holds at most 64 live voices (a walker keeps about
    four); the 65th admission is refused with a stated error, and the engine
    goes on mixing the 64 it holds."""
    import ctypes as C

    lib = fd.load_library()
    b = device_bank(small_bank(2, 12, 101, 100), 2, 12)
    e = engine(b, walk_past(), 20)
    f = np.ones(40)
    fp = f.ctypes.data_as(C.POINTER(C.c_double))
    for i in range(64):
        assert lib.gf_engine_admit(e._h, b"c%d" % i, 3.0 + 0.1 * i, 4.0, fp, 40, 0.5) == 0, lib.gf_last_error()
    assert lib.gf_engine_admit(e._h, b"c64", 9.5, 4.0, fp, 40, 0.5) != 0
    assert lib.gf_last_error().decode() == "geofdtd: more than 64 voices at once"
    out = np.zeros((20, 20), dtype=np.float32)  # 0.2 s: past the arrivals at 5-10 m
    for row in out:
        assert lib.gf_engine_next(e._h, row.ctypes.data_as(C.POINTER(C.c_float)), 20) == 0, lib.gf_last_error()
    assert np.all(np.isfinite(out)) and np.max(np.abs(out)) > 0
    e.close()
    b.close()


def test_contacts_superpose():
    """sensing_test.go:79-117 on the device engine: two contacts together give
    the sum of each alone, so voices share no state.  Bound: the reference's
    1e-7 of the peak, doubled because each of the three outputs is rounded to
    float32 on its own."""
    b = device_bank(small_bank(2, 12, 101, 100), 2, 12)
    stance = wl.Stance(Mass=75)
    a = se.Contact("a", -3, 8, 0.1, se.Footfall(stance, 0.0))
    c = se.Contact("b", 2, 9, 0.4, se.Footfall(stance, 0.0))
    n = 4000
    outs = []
    for cs in ([a], [c], [a, c]):
        e = fixed_engine(b, cs, n)
        outs.append(run(e, n).astype(np.float64))
        e.close()
    scale = np.max(np.abs(outs[2]))
    assert scale > 0
    assert np.max(np.abs(outs[2] - (outs[0] + outs[1]))) <= 2e-7 * scale
    b.close()


def test_walk_past_envelope_peaks_at_closest_approach():
    """sensing_test.go:119-172: a walk past rises and falls, loudest within
    two footfalls of the closest, and range moves the amplitude.  The
    reference passes at 6 m with its far-field model; this bank's layered
    response changes little between 6 and 8.5 m, where the overlap of
    successive footfalls sets the loudest window instead (the oracle's
    Reference on the same bank puts it four footfalls early too), so the walk
    passes at 3 m, where range dominates."""
    b = device_bank(small_bank(2, 12, 101, 100), 2, 12)
    w = walk_past(range_m=3.0, approach=6.0)
    n = int(w.Until * RATE)
    e = engine(b, w, 50)
    trace = run(e, n)
    e.close()
    pts = []
    for c in w.Contacts(0, n / RATE):
        lo = int(c.Start * RATE)
        hi = min(lo + int(0.5 * RATE), n)
        if lo < hi:
            pts.append((math.hypot(c.X, c.Y), float(np.max(np.abs(trace[lo:hi])))))
    assert len(pts) >= 8
    loudest = max(range(len(pts)), key=lambda i: pts[i][1])
    closest = min(range(len(pts)), key=lambda i: pts[i][0])
    print(f"[sensing] walk past: {len(pts)} footfalls, loudest {loudest}, closest {closest}", flush=True)
    assert abs(loudest - closest) <= 2
    assert pts[0][1] < pts[loudest][1] and pts[-1][1] < pts[loudest][1]
    assert pts[closest][1] > 1.1 * pts[0][1]
    b.close()


def test_range_quantisation_error_is_bounded():
    """sensing_test.go:233-287 on the device engine: one quantum of range
    moves the energy by under 0.5 % and the arrival by at most one sample
    plus the true delay.  The reference takes the pair at 8 m; this test
    bank's transform is 512 samples (0.256 s), so its responses stop at
    0.19 s, and beyond about 6 m where that cut falls moves with range by
    enough to change the energy 0.5-2 % per quantum (the oracle's Reference
    on the same bank shows the same).  The pair is taken at 4 m."""
    b = device_bank(small_bank(2, 12, 101, 100), 2, 12)
    stance = wl.Stance(Mass=75)
    n = 3000

    def at(y):
        e = fixed_engine(b, [se.Contact("f", 0, y, 0.05, se.Footfall(stance, 0.0))], n)
        out = run(e, n).astype(np.float64)
        e.close()
        return math.sqrt(np.mean(out ** 2)), out

    p0, t0 = at(4.0)
    p1, t1 = at(4.0 + se.RangeQuantum)
    assert abs(p1 - p0) / p0 < 0.005
    true_shift = se.RangeQuantum / C_R * RATE
    assert true_shift <= 1
    xc = np.correlate(t1, t0, mode="full")
    lag = int(np.argmax(xc)) - (n - 1)
    print(f"[sensing] one quantum: energy {abs(p1 - p0) / p0:.2e}, lag {lag} samples", flush=True)
    assert abs(lag) <= 1 + true_shift
    b.close()


def test_reference_on_the_device(oracle):
    """sensing.Reference (sensing.go:344-386) built from device pieces (the
    bank's impulse responses, one whole-record convolution per contact)
    against the oracle's restatement, within 1e-12 of the peak (both FP64,
    direct sums against numpy's); and the streaming engine against it within
    the reference's 1e-7."""
    data = small_bank(2, 12, 101, 100)
    b = device_bank(data, 2, 12)
    ob = oracle.OracleBank(2, 12, data)
    w = walk_past()
    n = 6000
    kw = dict(sample_rate=RATE, walk_duration=w.Until, geometry_range=w.StartY, rayleigh_velocity=C_R,
              sensor=geophone)
    got = se.Reference(w, n, se.FromBank(b), **kw)
    want = oracle_reference(ob, RATE, w, n, RATE, 3 * SAMPLES // 4)
    scale = np.max(np.abs(want))
    assert np.max(np.abs(got - want)) <= 1e-12 * scale
    e = engine(b, w, 25)
    assert_matches(run(e, n), got)
    e.close()
    off = se.Reference(w, n, se.FromBank(b), receiver=(1.5, 0.3), **kw)
    assert np.max(np.abs(off - oracle_reference(ob, RATE, w, n, RATE, 3 * SAMPLES // 4, 1.5, 0.3))) <= 1e-12 * scale
    b.close()


def test_mix_kernel_at_every_workgroup_shape(oracle):
    """k_engine_mix gives each 1,024-thread workgroup np outputs of one
    receiver, np = clamp(ceil(chunk x receivers / 256), 1, 64): a line of 64
    receivers at 20 samples (np = 5, the benchmark's shape), 16 receivers at
    300 samples (np = 19, ragged last block) and 64 at 256 (np = 64); sampled
    receivers against the oracle's Reference."""
    data = small_bank(2, 12, 101, 100)
    b = device_bank(data, 2, 12)
    ob = oracle.OracleBank(2, 12, data)
    w = walk_past()
    n = 3000
    for R, chunk in ((64, 20), (16, 300), (64, 256)):
        rec = [(-3.0 + 6.0 * q / (R - 1), -0.5 + 0.01 * q) for q in range(R)]
        print(f"[sensing] {R} receivers x chunk {chunk}", flush=True)
        e = engine(b, w, chunk, receivers=rec)
        got = run(e, n)
        e.close()
        for q in range(0, R, max(1, R // 6)):
            assert_matches(got[q], oracle_reference(ob, RATE, w, n, RATE, 3 * SAMPLES // 4, *rec[q]))
    b.close()


def test_analytic_engine_matches_the_reference():
    """sensing.NewEngine's default propagation (sensing.Analytic over
    green.HalfSpaceGF of loam, propagation.go:38-62) on the device: both
    response components built there at the transform dsp picks
    (NextPow2(4 irLen)), each voice driven by its vertical force and by its
    force along the contact-to-receiver line (sensing.go:264-270), against
    the oracle's Reference with both components, at the reference's 1e-7 of
    the peak and within one float32 step; and the radial part is there (the
    same run without it differs)."""
    soil = se.Loam()
    oh = SO.OHalfSpace(500, 200, 1700, 25)
    fs = 1000.0
    w = walk_past(range_m=4.0, approach=2.0)
    n = 4000
    sm24 = se.SM24()  # the reference's geophone, coupling included
    for rec in ([(0.0, 0.0)], [(0.0, 0.0), (2.0, 1.0)]):
        e = se.NewEngine(w, 25, soil, sample_rate=fs, walk_duration=w.Until, geometry_range=4.0, sensor=sm24,
                         receivers=rec)
        ir_len = e.ImpulseLength()
        nfft = 1
        while nfft < 4 * ir_len:
            nfft <<= 1
        import ctypes as C

        t_n = C.c_int32()
        fd._check(fd.load_library().gf_engine_transform(e._h, C.byref(t_n)))
        assert t_n.value == nfft
        got = run(e, n)
        e.close()
        got = np.atleast_2d(got)
        contacts = SO.walk_contacts(oracle_walk(w), 0, n / fs)
        foot = se.Footfall(w.Stance, w.Heading)
        forces = [lambda t, c=c: foot.ForceAt(t - c.start) for c in contacts]

        def pair(rq):
            v = SO.causal_impulse_response(0, ir_len, fs,
                                           lambda f: SO.velocity_response(oh, rq, f) * complex(sm24.Response(f)))
            r = SO.causal_impulse_response(0, ir_len, fs,
                                           lambda f: SO.radial_force_response(oh, rq, f) * complex(sm24.Response(f)))
            return v, r

        for q, (rx, ry) in enumerate(rec):
            want = SO.reference_xyz(contacts, forces, n, fs, pair, rx, ry)
            rel = assert_matches(got[q], want)
            print(f"[sensing] analytic, receiver {q}: {rel:.2e} of the peak", flush=True)
            vert_only = SO.reference_xyz(contacts, forces, n, fs, lambda rq: (pair(rq)[0], 0 * pair(rq)[1]), rx, ry)
            assert np.max(np.abs(vert_only - want)) > 1e-3 * np.max(np.abs(want))


def analytic_fixed(contacts, n, fs=1000.0, chunk=50):
    sch = se.Fixed(contacts)
    return se.NewEngine(sch, chunk, se.Loam(), sample_rate=fs, walk_duration=n / fs, geometry_range=8.0,
                        sensor=geophone)


def test_analytic_walk_past_envelope_peaks_at_closest_approach():
    """sensing_test.go:119-172 as the reference runs it: the analytic
    propagation, a walk past at 6 m, 1 kHz, chunks of 50."""
    w = walk_past(range_m=6.0, approach=6.0)
    fs = 1000.0
    n = int(w.Until * fs)
    e = se.NewEngine(w, 50, se.Loam(), sample_rate=fs, walk_duration=w.Until, geometry_range=6.0, sensor=geophone)
    trace = run(e, n)
    e.close()
    pts = []
    for c in w.Contacts(0, n / fs):
        lo = int(c.Start * fs)
        hi = min(lo + int(0.5 * fs), n)
        if lo < hi:
            pts.append((math.hypot(c.X, c.Y), float(np.max(np.abs(trace[lo:hi])))))
    assert len(pts) >= 8
    loudest = max(range(len(pts)), key=lambda i: pts[i][1])
    closest = min(range(len(pts)), key=lambda i: pts[i][0])
    assert abs(loudest - closest) <= 2
    assert pts[0][1] < pts[loudest][1] and pts[-1][1] < pts[loudest][1]
    assert pts[closest][1] > 1.1 * pts[0][1]


def test_shear_nulls_abeam_and_contributes_in_line():
    """sensing_test.go:419-463 on the device engine: one footfall heading +x
    at 8 m; abeam the shear leaves the energy unchanged (it projects to zero
    on the line of sight), in line it moves it by 0.1-5 %."""
    n = 4000

    def energy(x, y, ap):
        st = wl.Stance(Mass=75, APPeak=ap)
        e = analytic_fixed([se.Contact("f", x, y, 0.05, se.Footfall(st, 0.0))], n)
        out = run(e, n).astype(np.float64)
        e.close()
        return float(np.sum(out * out))

    abeam_off, abeam_on = energy(0, 8, -1), energy(0, 8, 0.35)
    assert abs(abeam_on - abeam_off) / abeam_off <= 1e-6
    inline_off, inline_on = energy(-8, 0, -1), energy(-8, 0, 0.35)
    rel = abs(inline_on - inline_off) / inline_off
    assert 1e-3 <= rel <= 0.05, rel


def test_surrogate_spectra_drive_synthesis(oracle):
    """The engine as a consumer of the surrogate's output: the network's
    spectra for one soil at the bank's 50 ranges (1-50 m, 500 Hz, 512 bins)
    form a bank in HBM, and a walk past at 4 m through it equals the oracle's
    Reference over the same numbers.  The network is freshly initialised
    (seed 0): the check is that its output, in the bank format, is what the
    engine streams, not what a trained network predicts."""
    import cosim_amd.surrogate as sur
    p = 0.5 * (sur.PARAM_RANGES[:, 0] + sur.PARAM_RANGES[:, 1])
    r = np.linspace(sur.RANGES[0], sur.RANGES[1], int(sur.RANGES[2]))
    P = np.repeat(p[None, :], len(r), axis=0)
    net = sur.Network(max_batch=64, seed=0)
    try:
        scale, delay = sur.prior(P, r)
        G = net.Predict(sur.features(P, r), scale, delay)
    finally:
        net.close()
    assert G.shape == (len(r), sur.BINS) and np.all(np.isfinite(G)) and np.any(G)
    layer = fd.Layer(0, p[8] * p[0], p[0], p[9], p[3])
    h = fd.BankHeader(Solver="surrogate", Medium=[layer], SampleRateHz=sur.SAMPLE_RATE_HZ, Samples=sur.SAMPLES,
                      MinM=float(r[0]), MaxM=float(r[-1]), Count=len(r),
                      Component="vertical surface displacement per unit vertical surface point force")
    b = bk.Bank(h, G)
    data = np.stack([G.real, G.imag], axis=-1).astype(np.float32)
    ob = oracle.OracleBank(float(r[0]), float(r[-1]), data)
    fs = sur.SAMPLE_RATE_HZ
    w = walk_past()
    e = se.NewEngineWith(w, 10, se.FromBank(b), sample_rate=fs, walk_duration=w.Until, geometry_range=w.StartY,
                         rayleigh_velocity=sur.RAYLEIGH_FRACTION * p[0], sensor=geophone)
    ir = e.ImpulseLength()
    assert ir == 3 * sur.SAMPLES // 4
    n = int(w.Until * fs)
    got = run(e, n)
    print(f"[sensing] surrogate bank, walk past at 4 m: {e.Responses()} responses of {ir} taps, "
          f"peak {float(np.max(np.abs(got))):.3e}", flush=True)
    assert_matches(got, oracle_reference(ob, fs, w, n, fs, ir))
    e.close()
    b.close()
