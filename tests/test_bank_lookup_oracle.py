"""The bank interpolation oracle (oracle/bank_oracle.c) pinned against the
reference's own bank_test.go interpolation tests, on banks filled by the f-k
oracle as the reference's smallBank fills them (bank_test.go:17-52).  CPU.

  * exact on grid points to 1e-12 (bank_test.go:318-336);
  * within the recommended range spacing (RangeNyquist(0.87 Vs, 100 Hz) / 2)
    the amplitude error is under 3 % and the phase error under 0.03 rad
    (bank_test.go:129-164);
  * past the limit the phase error saturates (over 1 rad at 3.7x) while it stays
    under 0.05 rad at 0.9x (bank_test.go:174-198);
  * interpolating log-magnitude and unwrapped phase beats averaging complex
    values at least five-fold (bank_test.go:204-244);
  * the response keeps decaying across the claimed band (bank_test.go:402-448);
  * the bin and range errors carry the reference's texts (bank_test.go:338-356).
"""

import functools
import math

import numpy as np
import pytest

from conftest import load_package

pkg = load_package()
fd, fk = pkg.fdtd, pkg.fk

RATE, SAMPLES = 2000.0, 512
BINS = SAMPLES // 2 + 1
LOAM_VS = 200.0


def range_nyquist(slowest, max_freq):
    """RangeNyquist (interp.go:120-125)."""
    return math.inf if max_freq <= 0 else slowest / (2 * max_freq)


@functools.lru_cache(maxsize=None)
def small_bank(r_min, r_max, count, max_freq):
    """smallBank: loam, DefaultQ 30, bins 1..max_freq filled by the f-k
    oracle's VerticalDisplacementMulti, stored as float32 like bank.Set."""
    import oracle as O

    m = fk.Medium(Stack=[fd.Layer(0, 500, 200, 1700, 25)], DefaultQ=30)
    ranges = [r_min + (r_max - r_min) * i / (count - 1) for i in range(count)]
    data = np.zeros((count, BINS, 2), dtype=np.float32)
    top = int(max_freq / (RATE / SAMPLES))
    for k in range(1, min(top, BINS - 1) + 1):
        v = O.fk_vertical_displacement_multi(m, ranges, k * RATE / SAMPLES)
        data[:, k, 0] = v.real
        data[:, k, 1] = v.imag
    return data


def bank(oracle, r_min, r_max, count, max_freq):
    return oracle.OracleBank(r_min, r_max, small_bank(r_min, r_max, count, max_freq))


def decimate(oracle, b, stride):
    count = (b.count - 1) // stride + 1
    return oracle.OracleBank(b.r_min, b.range_at(stride * (count - 1)), b.data[::stride][:count])


def compare(fine, coarse, stride):
    worst_amp = worst_phase = 0.0
    n = 0
    for i in range(fine.count):
        if i % stride == 0:
            continue
        r = fine.range_at(i)
        if r > coarse.r_max:
            continue
        for k in range(1, fine.bins):
            want = fine.At(i, k)
            if want == 0:
                continue
            got = coarse.Response(r, k)
            amp = abs(abs(got) - abs(want)) / abs(want)
            ph = math.atan2(got.imag, got.real) - math.atan2(want.imag, want.real)
            ph -= 2 * math.pi * round(ph / (2 * math.pi))
            worst_amp, worst_phase = max(worst_amp, amp), max(worst_phase, abs(ph))
            n += 1
    return worst_amp, worst_phase, n


def test_response_is_exact_on_grid_points(oracle):
    b = bank(oracle, 2, 8, 9, 60)
    for i in range(b.count):
        for k in range(1, b.bins):
            want = b.At(i, k)
            if want == 0:
                continue
            assert abs(b.Response(b.range_at(i), k) - want) / abs(want) <= 1e-12


@pytest.mark.slow
def test_interpolation_error_within_the_recommended_spacing(oracle):
    fine = bank(oracle, 2, 12, 101, 100)
    limit = range_nyquist(0.87 * LOAM_VS, 100) / 2
    at_limit = 0.0
    for stride in (2, 4, 8, 16):
        coarse = decimate(oracle, fine, stride)
        amp, phase, n = compare(fine, coarse, stride)
        assert n > 0
        spacing = (coarse.r_max - coarse.r_min) / (coarse.count - 1)
        if spacing <= limit:
            at_limit = max(at_limit, phase)
            assert amp < 0.03, (stride, amp)
    assert at_limit < 0.03


@pytest.mark.slow
def test_exceeding_the_sampling_limit_fails_catastrophically(oracle):
    fine = bank(oracle, 2, 12, 101, 100)
    limit = range_nyquist(0.87 * LOAM_VS, 100) / 2
    within, beyond = decimate(oracle, fine, 4), decimate(oracle, fine, 16)
    assert (within.r_max - within.r_min) / (within.count - 1) <= limit
    assert (beyond.r_max - beyond.r_min) / (beyond.count - 1) > limit
    assert compare(fine, within, 4)[1] < 0.05
    assert compare(fine, beyond, 16)[1] > 1.0


@pytest.mark.slow
def test_phase_aware_interpolation_beats_complex_averaging(oracle):
    stride = 4
    fine = bank(oracle, 2, 12, 101, 100)
    coarse = decimate(oracle, fine, stride)
    spacing = (coarse.r_max - coarse.r_min) / (coarse.count - 1)
    ours = naive = 0.0
    n = 0
    for i in range(fine.count):
        if i % stride == 0:
            continue
        r = fine.range_at(i)
        if r > coarse.r_max:
            continue
        x = (r - coarse.r_min) / spacing
        lo = int(math.floor(x))
        lo, frac = (coarse.count - 2, 1.0) if lo >= coarse.count - 1 else (lo, x - lo)
        for k in range(1, fine.bins):
            want = fine.At(i, k)
            if want == 0:
                continue
            got = coarse.Response(r, k)
            a, bb = coarse.At(lo, k), coarse.At(lo + 1, k)
            flat = a + frac * (bb - a)
            ours += abs(got - want) / abs(want)
            naive += abs(flat - want) / abs(want)
            n += 1
    assert n > 0
    assert naive / n >= 5 * (ours / n)


@pytest.mark.slow
def test_response_decays_across_the_claimed_band(oracle):
    b = bank(oracle, 2, 12, 41, 100)
    top = int(100 / (RATE / SAMPLES))
    for r in (4.0, 8.0, 12.0):
        lo, hi = b.Response(r, top // 4), b.Response(r, top)
        assert abs(lo) > 0 and abs(hi) > 0
        assert 1e-6 <= abs(hi) / abs(lo) <= 10


def test_rejects_bad_queries(oracle):
    b = bank(oracle, 2, 8, 5, 40)
    cases = [((1.0, 5), "bank: range 1 m is outside the bank's 2 to 8 m"),
             ((100.0, 5), "bank: range 100 m is outside the bank's 2 to 8 m"),
             ((4.0, -1), "bank: bin -1 out of range [0, 257)"),
             ((4.0, BINS), "bank: bin 257 out of range [0, 257)")]
    for (r, k), message in cases:
        with pytest.raises(oracle.OracleError) as e:
            b.Response(r, k)
        assert str(e.value) == message
    # Above the modelled band the bank is zero and stays zero.
    assert b.Response(5.0, 200) == 0


def test_range_nyquist():
    assert abs(range_nyquist(180, 300) - 0.3) < 1e-12
    assert math.isinf(range_nyquist(180, 0))


def test_host_range_sampling_rule():
    """CheckRangeSampling (interp.go:131-141) through the library: a 1 m grid
    over loam is too coarse at 300 Hz and fine at 20 Hz (bank_test.go:246-270)."""
    bk = pkg.bank
    h = fd.BankHeader(Solver="test", Medium=[fd.Layer(0, 500, 200, 1700, 25)], SampleRateHz=2000, Samples=512,
                      MinM=1, MaxM=40, Count=40, Component="test")
    with pytest.raises(fd.FdtdError) as e:
        bk.CheckRangeSampling(h, 300)
    assert str(e.value) == ("bank: range spacing 1.0000 m is too coarse to unwrap phase up to 300 Hz; "
                            "need under 0.1450 m")
    bk.CheckRangeSampling(h, 20)
    assert abs(bk.RangeNyquist(180, 300) - 0.3) < 1e-12
    assert math.isinf(bk.RangeNyquist(180, 0))


def test_bank_impulse_of_a_unit_velocity_bank_is_the_taper(oracle):
    """A bank holding 1/(i 2 pi f) at every bin has a vertical velocity
    response of 1, so its causal impulse response is the inverse transform of
    dsp.CausalImpulseResponseAt's taper alone (dsp.go:245-258) -- a closed form
    that pins the i w factor, the taper and the scaling of the composition."""
    rate, samples = 2000.0, 512
    bins = samples // 2 + 1
    f = np.arange(bins) * rate / samples
    z = np.zeros((3, bins), dtype=complex)
    z[:, 1:] = 1 / (1j * 2 * np.pi * f[1:])
    data = np.stack([z.real, z.imag], axis=-1).astype(np.float32)
    ob = oracle.OracleBank(1, 3, data)
    got = oracle.bank_impulse(ob, rate, [1.0, 2.5], 300, rate)
    nyq, frm = rate / 2, 0.8 * rate / 2
    taper = np.where(f > frm, 0.5 * (1 + np.cos(np.pi * (f - frm) / (nyq - frm))), 1.0)
    taper[0] = 0.0
    want = np.fft.irfft(taper, samples)[:300]
    for q in range(2):
        assert np.max(np.abs(got[q] - want)) <= 1e-6 * np.max(np.abs(want))
