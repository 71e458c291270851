"""Bank generation on the GPU: the trace-to-spectrum reduction against an
analytic pair and an FP64 restatement, the whole pipeline against the oracle,
and config-3 soils against the oracle."""

import math
import os

import numpy as np
import pytest

from media import rel_l2

pytestmark = pytest.mark.gpu


def ricker_and_derivative(peak, f0, t):
    """fdtd.Ricker's wavelet (run.go:19-31) and its exact time derivative."""
    tau = t - 1.5 / f0
    a = math.pi ** 2 * f0 ** 2
    e = np.exp(-a * tau ** 2)
    return peak * (1 - 2 * a * tau ** 2) * e, peak * e * (-6 * a * tau + 4 * a * a * tau ** 3)


def numpy_g(traces, dt, t0, force, freqs):
    """FP64 restatement of the reduction: G = V / (i w W)."""
    n = traces.shape[1]
    t = t0 + np.arange(n) * dt
    tf = np.arange(len(force)) * dt
    out = []
    for f in freqs:
        w = 2 * math.pi * f
        v = dt * (traces.astype(np.float64) @ np.exp(-1j * w * t))
        wf = dt * np.sum(force * np.exp(-1j * w * tf))
        out.append(v / (1j * w * wf))
    return np.array(out).T


def test_spectrum_recovers_an_analytic_delay(fdtd):
    """v(t) = g0 F'(t - tau) has G(f) = g0 exp(-i w tau) exactly: the
    reduction must return that delay and gain (conventions, sign, T0)."""
    dt, n, t0 = 1e-4, 20000, 0.5e-4
    g0, tau = 3.0e-9, 0.15
    force, _ = ricker_and_derivative(800.0, 40.0, np.arange(n) * dt)
    _, dv = ricker_and_derivative(800.0, 40.0, t0 + np.arange(n) * dt - tau)
    traces = np.stack([g0 * dv, 2 * g0 * dv]).astype(np.float32)
    freqs = np.arange(5.0, 101.0, 5.0)
    got = fdtd.Spectrum(traces, dt, t0, force, freqs)
    want = g0 * np.exp(-2j * math.pi * freqs * tau)
    err = np.abs(got[0] - want) / np.abs(want)
    print(f"analytic delay: max rel err {err.max():.2e}")
    assert err.max() <= 1e-5
    np.testing.assert_allclose(got[1], 2 * got[0], rtol=1e-12)
    ref = numpy_g(traces, dt, t0, force, freqs)
    assert np.max(np.abs(got - ref) / np.abs(ref)) <= 1e-10


def small_models(fdtd):
    two = [fdtd.Layer(1.2, 420, 190, 1700, 20, 40), fdtd.Layer(0, 700, 320, 2000, 60, 120)]
    return [
        fdtd.Model(Stack=two, GslsMechanisms=3, RefFreq=30, MaxRange=12.0, Depth=8.0, Spacing=0.1, DominantFreq=40),
        fdtd.Model(Stack=[fdtd.Layer(0, 500, 200, 1700, 25)], Relax=fdtd.Fit(25, 30), MaxRange=9.0, Depth=6.0,
                   Spacing=0.08, DominantFreq=40),
    ]


def test_bank_generation_matches_the_oracle_pipeline(fdtd, oracle):
    spec = fdtd.BankSpec(SampleRateHz=500.0, Samples=256, MaxFreqHz=100.0, MinM=1.0, MaxM=5.0, Count=5)
    shots = []
    for m in small_models(fdtd):
        dt = m.Grid()[0]
        steps = int(0.35 / dt)
        shots.append(fdtd.Shot(m, fdtd.Ricker(1000, 40, dt, steps), [], steps))
    spectra, snapped, results = fdtd.GenerateBanks(shots, spec)
    bins = spec.Samples // 2 + 1
    top = int(spec.MaxFreqHz / (spec.SampleRateHz / spec.Samples))
    freqs = np.arange(1, top + 1) * spec.SampleRateHz / spec.Samples
    ranges = np.linspace(spec.MinM, spec.MaxM, spec.Count)
    for i, sh in enumerate(shots):
        assert isinstance(results[i], fdtd.Result)
        want_tr = oracle.run(sh.Model, sh.Force, ranges, sh.Steps)
        np.testing.assert_array_equal(snapped[i], want_tr.snapped)
        want = numpy_g(want_tr.traces, want_tr.dt, want_tr.t0, np.asarray(sh.Force), freqs)
        got = spectra[i]
        assert got.shape == (spec.Count, bins)
        assert not np.any(got[:, 0]) and not np.any(got[:, top + 1:])
        errs = [rel_l2(got[r, 1:top + 1], want[r]) for r in range(spec.Count)]
        print(f"model {i}: bank rel-L2 over bins 1..{top}: max {max(errs):.2e}")
        assert max(errs) <= 1e-4


def test_generated_bank_writes_a_readable_file(fdtd, tmp_path):
    from test_bank_format import read_bank

    import cosim_amd.workloads as W

    spec = fdtd.BankSpec(SampleRateHz=500.0, Samples=256, MaxFreqHz=100.0, MinM=1.0, MaxM=5.0, Count=5)
    m = small_models(fdtd)[0]
    dt = m.Grid()[0]
    steps = int(0.2 / dt)
    spectra, snapped, res = fdtd.GenerateBanks([fdtd.Shot(m, fdtd.Ricker(1000, 40, dt, steps), [], steps)], spec)
    h = fdtd.BankHeader(Solver="geofdtd (MI355X FDTD)", Medium=m.Stack, SampleRateHz=spec.SampleRateHz,
                        Samples=spec.Samples, MinM=spec.MinM, MaxM=spec.MaxM, Count=spec.Count,
                        Notes="3 GSLS mechanisms over 5-150 Hz; bins above 100 Hz zero")
    path = str(tmp_path / "fdtd.bank")
    fdtd.WriteBank(path, h, spectra[0])
    header, data, _ = read_bank(path)
    np.testing.assert_array_equal(data, spectra[0].astype(np.complex64))
    assert header["units"] == "m/N" and header["ranges"] == {"min_m": 1, "max_m": 5, "count": 5}


def test_config3_soils_match_the_oracle(fdtd, oracle):
    """Two seeded config-3 soils, batched together on the GPU, against the
    FP64 oracle over the first 500 steps at five bank ranges."""
    import cosim_amd.workloads as W

    ws = W.c3(n_models=2, first=3)
    ranges = [1.0, 2.0, 5.0, 10.0, 20.0]
    shots = [fdtd.Shot(w.model, w.force, ranges, 500) for w in ws]
    got = fdtd.RunBatch(shots)
    for w, sh, g in zip(ws, shots, got):
        want = oracle.run(w.model, w.force, ranges, 500)
        from test_gpu_parity import compare

        compare(w.name, [t.Vertical for t in g.Traces], want.traces)


def test_batch_handle_in_slices_equals_one_call(fdtd):
    """gf_batch_* in slices (a third, two thirds, the rest), then reset and
    run again whole: every pass returns gf_run_batch's traces bit for bit, and
    results are refused before the batch has run to its end."""
    import cosim_amd.workloads as W

    ws = W.c3(n_models=2, first=3)
    ranges = [1.0, 2.0, 5.0, 10.0, 20.0]
    shots = [fdtd.Shot(w.model, w.force, ranges, 400 + 77 * i) for i, w in enumerate(ws)]
    want = fdtd.RunBatch(shots)
    b = fdtd.Batch(shots)
    try:
        b.Advance(1, 3)
        with pytest.raises(fdtd.FdtdError):
            b.Results()
        b.Advance(2, 3)
        b.Advance(1, 1)
        st = b.Stats()
        assert st["models"] == 2 and st["step_launches"] > 0 and st["algorithmic_bytes"] > 0
        first = b.Results()
        b.Reset()
        b.Advance(1, 1)
        again = b.Results()
    finally:
        b.close()
    for w, g, a, r in zip(ws, first, again, want):
        for tg, ta, tr in zip(g.Traces, a.Traces, r.Traces):
            np.testing.assert_array_equal(tg.Vertical, tr.Vertical)
            np.testing.assert_array_equal(ta.Vertical, tr.Vertical)
        assert g.Dt == r.Dt and g.Peak == r.Peak
    print(f"batch handle: {len(ws)} soils, slices and a rerun equal one gf_run_batch call")


def test_batch_handle_spectra_equal_bank_generation(fdtd):
    spec = fdtd.BankSpec(SampleRateHz=500.0, Samples=256, MaxFreqHz=100.0, MinM=1.0, MaxM=5.0, Count=5)
    shots = []
    for m in small_models(fdtd):
        dt = m.Grid()[0]
        steps = int(0.2 / dt)
        shots.append(fdtd.Shot(m, fdtd.Ricker(1000, 40, dt, steps), [], steps))
    want, want_snapped, _ = fdtd.GenerateBanks(shots, spec)
    b = fdtd.Batch(shots, spec)
    try:
        b.Advance(1, 2)
        b.Advance(1, 1)
        got, snapped, results = b.Spectra()
    finally:
        b.close()
    assert all(isinstance(r, fdtd.Result) for r in results)
    np.testing.assert_array_equal(snapped, want_snapped)
    np.testing.assert_array_equal(got, want)


def test_large_batch_runs_all_together_with_the_two_step_kernel(fdtd, monkeypatch):
    """A batch whose state passes 1 GB is planned as one group advanced by the
    two-step kernel (no environment settings); its traces equal a forced
    cache-sized-chunk plan on the one-step kernel and single runs, bit for
    bit.  The first config-3 soils until 1.1 GB of state (72 B per cell), a
    few hundred steps each, odd counts so that launches end mid-pair."""
    import cosim_amd.workloads as W

    ws, cells = [], 0
    for w in W.c3(n_models=40):
        ws.append(w)
        _, nr, nz = w.model.Grid()
        cells += nr * nz
        if cells * 72 > 1.1e9:
            break
    assert cells * 72 > 1.1e9
    ranges = [1.0, 5.0, 20.0]
    shots = [fdtd.Shot(w.model, w.force, ranges, 201 + 38 * (i % 5)) for i, w in enumerate(ws)]
    for var in ("GEOFDTD_CACHE_MB", "GEOFDTD_STEPS_PER_LAUNCH", "GEOFDTD_TILE_ROWS", "GEOFDTD_TUNE"):
        monkeypatch.delenv(var, raising=False)
    b = fdtd.Batch(shots)
    try:
        b.Advance(1, 1)
        st = b.Stats()
        together = b.Results()
    finally:
        b.close()
    print(f"{len(ws)} soils, {cells / 1e6:.1f} M cells: steps per launch {st['steps_per_launch']}, "
          f"tile rows {st['tile_rows']}", flush=True)
    assert st["steps_per_launch"] == 2
    monkeypatch.setenv("GEOFDTD_CACHE_MB", "160")
    monkeypatch.setenv("GEOFDTD_STEPS_PER_LAUNCH", "1")
    chunked = fdtd.RunBatch(shots)
    monkeypatch.delenv("GEOFDTD_CACHE_MB")
    for i in (0, len(ws) - 1):
        single = fdtd.Run(shots[i])
        for a, c, s in zip(together[i].Traces, chunked[i].Traces, single.Traces):
            np.testing.assert_array_equal(a.Vertical, c.Vertical)
            np.testing.assert_array_equal(a.Vertical, s.Vertical)
    for a, c in zip(together, chunked):
        for ta, tc in zip(a.Traces, c.Traces):
            np.testing.assert_array_equal(ta.Vertical, tc.Vertical)
