"""The device bank lookup against the oracle's restatement of bank.Bank's
interpolation, and the reference's bank_test.go properties on the GPU path.

Tolerance: the device's log, atan2, exp and sincos against glibc's; the
phase is a running sum along range, so values agree to 1e-12 relative (exact
zeros stay exact).
"""

import math

import numpy as np
import pytest

from conftest import load_package
from test_bank_lookup_oracle import BINS, LOAM_VS, RATE, SAMPLES, range_nyquist, small_bank

pkg = load_package()
fd, bk = pkg.fdtd, pkg.bank

pytestmark = pytest.mark.gpu


def header(r_min, r_max, count, samples=SAMPLES):
    return fd.BankHeader(Solver="test", Medium=[fd.Layer(0, 500, 200, 1700, 25)], SampleRateHz=RATE,
                         Samples=samples, MinM=r_min, MaxM=r_max, Count=count, Component="test")


def device_bank(data, r_min, r_max):
    z = data[..., 0].astype(np.float64) + 1j * data[..., 1].astype(np.float64)
    return bk.Bank(header(r_min, r_max, data.shape[0], 2 * (data.shape[1] - 1)), z)


def test_points_and_spectra_equal_the_oracle(oracle):
    data = small_bank(2, 12, 101, 100)
    print("[bank] 101 ranges x 257 bins", flush=True)
    dev = device_bank(data, 2, 12)
    ref = oracle.OracleBank(2, 12, data)
    rng = np.random.default_rng(7)
    r = rng.uniform(2, 12, 3000)
    r[:5] = [2.0, 12.0, 2.1, 11.999999, 7.0]
    k = rng.integers(0, BINS, 3000)
    got = dev.Responses(r, k)
    for i in range(len(r)):
        want = ref.Response(r[i], int(k[i]))
        if want == 0:
            assert got[i] == 0
        else:
            assert abs(got[i] - want) <= 1e-12 * abs(want), (r[i], k[i], got[i], want)
    st = dev.Stats()
    assert st["kernel_ms"] > 0 and st["bytes"] > 0
    spec = dev.Spectra(r[:64])
    assert spec.shape == (64, BINS)
    for q in range(64):
        assert np.array_equal(spec[q], dev.Responses(np.full(BINS, r[q]), np.arange(BINS)))
    dev.close()


def test_exact_on_grid_points_and_within_the_interpolation_budget(oracle):
    data = small_bank(2, 12, 101, 100)
    fine = oracle.OracleBank(2, 12, data)
    limit = range_nyquist(0.87 * LOAM_VS, 100) / 2
    for stride in (2, 4, 16):
        count = (fine.count - 1) // stride + 1
        r_max = fine.range_at(stride * (count - 1))
        coarse = device_bank(data[::stride][:count], 2, r_max)
        grid = [2 + (r_max - 2) * i / (count - 1) for i in range(count)]
        # Exact on its own grid points.
        g = coarse.Spectra(grid)
        for i in range(count):
            want = data[stride * i, :, 0].astype(np.float64) + 1j * data[stride * i, :, 1].astype(np.float64)
            nz = want != 0
            assert np.all(np.abs(g[i, nz] - want[nz]) <= 1e-12 * np.abs(want[nz]))
        # Between them, the reference's budget inside the sampling limit and
        # saturation beyond it.
        mid = [fine.range_at(i) for i in range(fine.count) if i % stride and fine.range_at(i) <= r_max]
        idx = [i for i in range(fine.count) if i % stride and fine.range_at(i) <= r_max]
        got = coarse.Spectra(mid)
        want = data[idx, :, 0].astype(np.float64) + 1j * data[idx, :, 1].astype(np.float64)
        nz = want != 0
        amp = np.max(np.abs(np.abs(got[nz]) - np.abs(want[nz])) / np.abs(want[nz]))
        ph = np.angle(got[nz]) - np.angle(want[nz])
        ph = np.max(np.abs(ph - 2 * np.pi * np.round(ph / (2 * np.pi))))
        spacing = (r_max - 2) / (count - 1)
        print(f"[bank] stride {stride}: spacing {spacing:.3f} m (limit {limit:.3f}), amp {amp:.2e}, phase {ph:.2e}",
              flush=True)
        if spacing <= limit:
            assert amp < 0.03 and ph < 0.03
        if stride == 16:
            assert ph > 1.0
        coarse.close()


def test_errors_carry_the_reference_text(oracle):
    data = small_bank(2, 8, 5, 40)
    dev = device_bank(data, 2, 8)
    cases = [([4.0, 1.0], [5, 5], "bank: range 1 m is outside the bank's 2 to 8 m"),
             ([4.0, 100.0, 4.0], [5, 5, -1], "bank: range 100 m is outside the bank's 2 to 8 m"),
             ([4.0], [-1], "bank: bin -1 out of range [0, 257)"),
             ([4.0], [BINS], "bank: bin 257 out of range [0, 257)")]
    for r, k, message in cases:
        with pytest.raises(fd.FdtdError) as e:
            dev.Responses(r, k)
        assert str(e.value) == message
    with pytest.raises(fd.FdtdError) as e:
        dev.Spectra([3.0, 9.5])
    assert str(e.value) == "bank: range 9.5 m is outside the bank's 2 to 8 m"
    dev.close()
    with pytest.raises(fd.FdtdError) as e:
        bk.Bank(fd.BankHeader(Solver="t", Medium=[], SampleRateHz=2000, Samples=511, MinM=1, MaxM=2, Count=2,
                              Component="x"), np.zeros((2, 256)))
    assert "bank: samples must be positive and even, got 511" in str(e.value)


def test_a_large_bank_equals_the_oracle(oracle):
    """The reference's benchmark bank (bank_test.go:358-382): 270 ranges 1-40 m,
    4096 samples, value 1/(i+1) + i k 1e-4; 4096 query ranges, every bin."""
    count, samples = 270, 4096
    bins = samples // 2 + 1
    i = np.arange(count)[:, None]
    k = np.arange(bins)[None, :]
    z = (1.0 / (i + 1)) + 1j * (k * 1e-4)
    data = np.stack([z.real, z.imag], axis=-1).astype(np.float32)
    dev = device_bank(data, 1, 40)
    ref = oracle.OracleBank(1, 40, data)
    r = np.random.default_rng(3).uniform(1, 40, 4096)
    got = dev.Spectra(r)
    for q in range(0, 4096, 97):
        for kk in (0, 1, 500, 1333, bins - 1):
            want = ref.Response(r[q], kk)
            assert abs(got[q, kk] - want) <= 1e-12 * abs(want)
    st = dev.Stats()
    print(f"[bank] 4096 ranges x {bins} bins: {st['kernel_ms']:.3f} ms, {st['bytes'] / st['kernel_ms'] / 1e6:.0f} GB/s",
          flush=True)
    dev.close()


def test_impulse_responses_equal_the_oracle(oracle):
    """FromBank + CausalImpulseResponseAt on the device against the oracle's
    composition: the f-k bank of smallBank, a geophone-like sensor, ranges
    between grid points; then the chain into the streaming convolver."""
    data = small_bank(2, 12, 101, 100)
    dev = device_bank(data, 2, 12)
    ob = oracle.OracleBank(2, 12, data)
    f = np.arange(BINS) * RATE / SAMPLES
    sensor = 28.8 * (1j * f / 4.5) ** 2 / (1 + 1j * 0.7 * (f / 4.5) - (f / 4.5) ** 2)  # 4.5 Hz, damping 0.35
    ranges = [2.0, 3.37, 7.0, 11.9]
    for sen in (None, sensor):
        got = dev.ImpulseResponses(ranges, 400, RATE, sen)
        want = oracle.bank_impulse(ob, RATE, ranges, 400, RATE, sen)
        for q in range(len(ranges)):
            assert np.max(np.abs(got[q] - want[q])) <= 1e-12 * np.max(np.abs(want[q])), (q, sen is None)
    # A bank grid twice as fine as the transform's (fs = 2 x rate): every
    # frequency still lands on a bin, the upper half answers 0.
    got = dev.ImpulseResponses([5.0], 256, 2 * RATE)
    want = oracle.bank_impulse(ob, RATE, [5.0], 256, 2 * RATE)
    assert np.max(np.abs(got - want)) <= 1e-12 * np.max(np.abs(want))
    with pytest.raises(fd.FdtdError) as e:
        dev.ImpulseResponses([5.0], 256, 1.5 * RATE)
    assert str(e.value) == "sensing: 5.859375 Hz does not land on the bank's frequency grid (spacing 3.90625 Hz)"
    with pytest.raises(fd.FdtdError) as e:
        dev.ImpulseResponses([5.0, 13.0], 256, RATE)
    assert str(e.value) == "bank: range 13 m is outside the bank's 2 to 12 m"
    with pytest.raises(fd.FdtdError) as e:
        dev.ImpulseResponses([5.0], 600, RATE)
    assert str(e.value) == "dsp: transform size 512 is shorter than the 600 samples to keep"
    # The chain: bank -> impulse responses -> convolvers, all on the device.
    st = load_package().stream
    irs = dev.ImpulseResponses(ranges, 400, RATE, sensor)
    conv = st.NewConvolver(irs, 20)
    x = np.random.default_rng(2).standard_normal((len(ranges), 1000))
    y = np.concatenate([conv.Process(x[:, k * 20:(k + 1) * 20]) for k in range(50)], axis=1)
    for q in range(len(ranges)):
        want_y = np.convolve(x[q], irs[q])[:1000]
        assert np.max(np.abs(y[q] - want_y)) <= 1e-12 * np.max(np.abs(want_y))
    conv.close()
    dev.close()


def test_long_transform_impulse_responses_equal_the_oracle(oracle):
    """A bank of 8,192 samples: its impulse responses take the rocFFT path
    (transforms above 4,096 points) and must equal the oracle's direct
    composition as the shorter ones do."""
    count, samples = 12, 8192
    bins = samples // 2 + 1
    i = np.arange(count)[:, None]
    k = np.arange(bins)[None, :]
    # a smooth, range-dependent spectrum with a travel-time phase
    z = np.exp(-k / 900.0) / (1 + i) * np.exp(-1j * (k * (0.002 + 0.0004 * i)))
    z[:, 0] = 0
    data = np.stack([z.real, z.imag], axis=-1).astype(np.float32)
    h = fd.BankHeader(Solver="test", Medium=[fd.Layer(0, 500, 200, 1700, 25)], SampleRateHz=RATE, Samples=samples,
                      MinM=2, MaxM=13, Count=count, Component="test")
    dev = bk.Bank(h, data[..., 0].astype(np.float64) + 1j * data[..., 1].astype(np.float64))
    ob = oracle.OracleBank(2, 13, data)
    ranges = [2.0, 4.3, 9.9]
    f = np.arange(bins) * RATE / samples
    sensor = 1.0 / (1 + 1j * f / 40.0)
    for sen in (None, sensor):
        got = dev.ImpulseResponses(ranges, 5000, RATE, sen)
        want = oracle.bank_impulse(ob, RATE, ranges, 5000, RATE, sen)
        for q in range(len(ranges)):
            assert np.max(np.abs(got[q] - want[q])) <= 1e-12 * np.max(np.abs(want[q])), (q, sen is None)
    dev.close()
