"""The surrogate on the CPU: its FP64 oracle pinned by finite differences and
closed forms, the layered soil family, the target normalization, and the
error gate's fallback with injected predictions.  No GPU."""

import math

import numpy as np
import pytest

import cosim_amd.fdtd as fd
import cosim_amd.surrogate as S
import surrogate_oracle as O

# The network's layer table as the library defines it (checked against a live
# handle by the GPU tests): (k_in, n_out, gelu, rows per query, bias channels).
ARCH = [(16, 512, 1, 1, 512)] + [(512, 512, 1, 1, 512)] * 5 + [
    (512, 2048, 1, 1, 2048), (32, 64, 1, 64, 32), (32, 8, 0, 128, 2)]


def layers(arch=ARCH):
    out, off = [], 0
    for k, n, g, rows, bm in arch:
        w = off
        off += n * k
        b = off
        off += (bm + 7) // 8 * 8
        out.append(O.Layer(k, n, bool(g), rows, bm, w, b))
    return out, off


def small_arch():
    """The same structure at small width, so finite differences are cheap."""
    return [(16, 24, 1, 1, 24), (24, 32, 1, 1, 32), (8, 16, 1, 4, 8), (8, 8, 1, 8, 4), (4, 16, 0, 16, 2)]


def random_params(L, P, rng, scale=0.5):
    p = np.zeros(P)
    for l in L:
        fan = 15 if l is L[0] else l.k
        w = rng.normal(0, scale / math.sqrt(fan), (l.n, l.k))
        if l is L[0]:
            w[:, 15:] = 0
        p[l.w_off: l.w_off + l.n * l.k] = w.ravel()
        p[l.b_off: l.b_off + l.bias_mod] = rng.normal(0, 0.1, l.bias_mod)
    return p


def test_full_architecture_produces_512_complex_bins():
    L, P = layers()
    assert P == 2374952
    rng = np.random.default_rng(0)
    p = random_params(L, P, rng)
    o, _ = O.forward(p, L, rng.uniform(-1, 1, (3, 15)))
    assert o.shape == (3, 1024)
    # the decoder's row chaining: each layer's output rows x columns equal
    # the next layer's input rows x k
    for a, b in zip(L, L[1:]):
        assert a.rows * a.n == b.rows * b.k


def test_gradient_matches_central_differences():
    """Exact FP64 backward pass against (f(p+h) - f(p-h)) / 2h, every layer."""
    arch = small_arch()
    L, P = layers(arch)
    # the small decoder ends at 16*16/2 = 128 complex bins
    rng = np.random.default_rng(1)
    p = random_params(L, P, rng, scale=1.0)
    x = rng.uniform(-1, 1, (3, 15))
    out_len = L[-1].rows * L[-1].n
    t = rng.normal(0, 1, (3, out_len))
    f0, g = O.loss_grad(p, L, x, t)
    h = 1e-6
    for l in L:
        idx = list(rng.choice(l.n * l.k, size=6, replace=False) + l.w_off) + [l.b_off, l.b_off + l.bias_mod - 1]
        for i in idx:
            if l is L[0] and (i - l.w_off) % l.k >= 15 and i < l.b_off:
                continue  # padded inputs: zero gradient by construction
            pp, pm = p.copy(), p.copy()
            pp[i] += h
            pm[i] -= h
            fd_ = (O.loss_grad(pp, L, x, t)[0] - O.loss_grad(pm, L, x, t)[0]) / (2 * h)
            assert abs(fd_ - g[i]) <= 1e-6 * max(1.0, abs(g[i])) + 1e-8, (i, fd_, g[i])


def test_loss_is_the_relative_l2_squared():
    t = np.array([[3.0, 4.0], [1.0, 0.0]])
    o = np.array([[3.0, 0.0], [0.0, 0.0]])
    np.testing.assert_allclose(O.loss(o, t), [16 / 25, 1.0])


def test_bf16_rounding_is_nearest_even():
    one = 1.0
    ulp = 2.0 ** -7
    vals = np.array([one, one + ulp / 2, one + 3 * ulp / 2, one + ulp / 4, -2.5, 3.0e38, 1e-39])
    got = O.bf16(vals)
    np.testing.assert_array_equal(got[:4], [one, one, one + 2 * ulp, one])
    assert got[4] == -2.5
    assert np.isfinite(got[5])
    assert got[6] != 0 and O.bf16(np.array([2.0 ** -130]))[0] == 2.0 ** -130  # FP32 subnormals keep their bits


def test_normalization_keeps_the_relative_l2():
    """A per-query scale and a per-bin unit phase factor leave |a-b|/|b|
    unchanged: the loss in the network's domain is the physical one."""
    rng = np.random.default_rng(2)
    n = 5
    params = np.column_stack([rng.uniform(lo, hi, n) for lo, hi in S.PARAM_RANGES])
    r = rng.uniform(1, 50, n)
    scale, delay = S.prior(params, r)
    G = (rng.normal(size=(n, S.BINS)) + 1j * rng.normal(size=(n, S.BINS))) * 1e-9
    H = G + (rng.normal(size=(n, S.BINS)) + 1j * rng.normal(size=(n, S.BINS))) * 1e-10
    a, b = S.normalize(H, scale, delay), S.normalize(G, scale, delay)
    np.testing.assert_allclose(S.rel_l2(a, b), S.rel_l2(S.band_limit(H), S.band_limit(G)), rtol=1e-5)
    # the target is zero outside the band the FDTD bank determines
    m = S.band_mask()
    assert not np.any(a.reshape(n, S.BINS, 2)[:, ~m]) and np.all(a.reshape(n, S.BINS, 2)[:, m] != 0)
    # the prior's amplitude is the static Boussinesq factor 1/(2 pi mu1 r)
    mu1 = params[:, 9] * params[:, 0] ** 2
    np.testing.assert_allclose(scale, 1 / (2 * math.pi * mu1 * r))
    np.testing.assert_allclose(delay, r / (0.92 * params[:, 0]))


def test_features_map_the_sampling_box_to_unit_interval():
    lo, hi = S.PARAM_RANGES[:, 0], S.PARAM_RANGES[:, 1]
    x = S.features(np.vstack([lo, hi]), np.array([1.0, 50.0]))
    keep = [0, 1, 2, 3, 4, 5, 8, 9, 10]
    np.testing.assert_allclose(x[:, keep], [[-1] * 9, [1] * 9], atol=1e-6)
    # thicknesses as the grid realizes them: within a cell of the drawn value
    h = S.grid_spacing(np.vstack([lo, hi]))
    for row, p in enumerate([lo, hi]):
        for k in (6, 7):
            span = hi[k] - lo[k]
            got = lo[k] + (x[row, k] + 1) / 2 * span
            assert abs(got - p[k]) <= h[row] + 1e-9
    np.testing.assert_allclose(x[:, 11], [2 / 50 - 1, 1], atol=1e-6)  # r itself
    # travel times r / Vs: the longest (50 m at the slowest Vs1) maps to 1
    np.testing.assert_allclose(x[1, 12], 2 * (50 / 200) / (50 / 120) - 1, atol=1e-6)
    assert np.all(np.abs(x[:, [0, 1, 2, 3, 4, 5, 8, 9, 10, 11, 12, 13, 14]]) <= 1 + 1e-6)


def test_grid_thicknesses_follow_the_node_rule():
    """Node j (depth j h) is in layer 1 iff j h < h1 (an interface node goes
    to the layer below); checked by counting nodes directly, including
    thicknesses that are exact multiples of h."""
    rng = np.random.default_rng(11)
    P = np.column_stack([rng.uniform(lo, hi, 200) for lo, hi in S.PARAM_RANGES])
    h = S.grid_spacing(P)
    P[:5, 6] = h[:5] * np.array([4, 7, 10, 13, 16])   # exact multiples
    t = S.grid_thicknesses(P)
    for i in range(len(P)):
        j = np.arange(0, 200)
        n1 = int(np.sum(j * h[i] < P[i, 6]))
        n12 = int(np.sum(j * h[i] < P[i, 6] + P[i, 7]))
        assert t[i, 0] == n1 * h[i] and t[i, 1] == (n12 - n1) * h[i]


def test_layered_family_is_deterministic_and_inside_its_box():
    for i in range(40):
        s1, p1 = fd.SampleLayeredSoil(7, i)
        s2, p2 = fd.SampleLayeredSoil(7, i)
        assert p1 == p2 and s1 == s2
        p = np.asarray(p1)
        assert np.all(p >= S.PARAM_RANGES[:, 0]) and np.all(p <= S.PARAM_RANGES[:, 1])
        fd.ValidateStack(s1.Stack)
        assert [l.Vs for l in s1.Stack] == list(p[:3])
        assert [l.Qp for l in s1.Stack] == [2 * q for q in p[3:6]]
        assert s1.Stack[0].Thickness == p[6] and s1.Stack[1].Thickness == p[7]
        assert [l.Density for l in s1.Stack] == [p[9], p[9] + 150, p[9] + 300]
        assert s1.Spacing == pytest.approx(0.9 * p[0] / 1000)
        assert abs(s1.MaxRange - 80.0) <= s1.Spacing and abs(s1.Depth - 30.0) <= s1.Spacing and s1.PMLCells == 20
    # a different stream from gf_soil_sample's
    assert fd.SampleSoil(7, 0).Stack[0].Vs != fd.SampleLayeredSoil(7, 0)[0].Stack[0].Vs


def _gate_case():
    S_, R = 50, 4
    rng = np.random.default_rng(3)
    truth = (rng.normal(size=(S_, R, S.BINS)) + 1j * rng.normal(size=(S_, R, S.BINS))).astype(np.complex64)
    truth = S.band_limit(truth)
    pred = truth.copy()
    bad = set(range(0, S_, 3))
    for s in bad:
        pred[s, 1] *= 1.10   # 10 % off on one range
    for s in range(1, S_, 3):
        pred[s] *= 1.005     # 0.5 %: within the bound
    calls = []

    def solver(soils):
        calls.append(list(soils))
        return truth[np.asarray(soils)], None, 0.0, {}

    return S_, R, truth, pred, bad, calls, solver


def test_gate_falls_back_for_injected_bad_predictions():
    """The error gate with the network's output replaced and escalation off:
    sampled soils whose predictions are off by more than the bound get their
    FDTD solution, the others keep the prediction.  The solver is a stand-in
    returning known truth, so this runs without a GPU."""
    S_, R, truth, pred, bad, calls, solver = _gate_case()
    soils = list(range(S_))
    params = np.zeros((S_, 10))
    ranges = np.ones((S_, R))
    out, rep = S.serve(None, params, ranges, soils, gate_fraction=0.5, bound=0.02, seed=4, solver=solver,
                       predictions=pred.reshape(S_ * R, S.BINS), escalate=False)
    assert len(rep.sampled_soils) == 25 and len(calls) == 1 and calls[0] == rep.sampled_soils
    assert rep.failed_soils == sorted(s for s in rep.sampled_soils if s in bad)
    assert not rep.escalated and rep.fdtd_soils == rep.failed_soils
    for s in range(S_):
        if s in rep.failed_soils:
            np.testing.assert_array_equal(out[s], truth[s])
        else:
            np.testing.assert_array_equal(out[s], pred[s])
    assert rep.sampled_errors.shape == (25 * R,)
    # a gate fraction of 1% samples ceil(0.01 S) soils
    _, rep1 = S.serve(None, params, ranges, soils, gate_fraction=0.01, solver=solver,
                      predictions=pred.reshape(S_ * R, S.BINS), escalate=False)
    assert len(rep1.sampled_soils) == 1


def test_gate_escalates_a_request_the_sample_shows_out_of_bound():
    """With escalation (the default), one failing sampled soil sends the whole
    request to the FDTD: every soil is answered by its solution, in one more
    solver call over the soils not yet solved."""
    S_, R, truth, pred, bad, calls, solver = _gate_case()
    soils = list(range(S_))
    out, rep = S.serve(None, np.zeros((S_, 10)), np.ones((S_, R)), soils, gate_fraction=0.5, bound=0.02, seed=4,
                       solver=solver, predictions=pred.reshape(S_ * R, S.BINS))
    assert rep.failed_soils and rep.escalated and rep.fdtd_soils == soils
    assert len(calls) == 2 and sorted(calls[0] + calls[1]) == soils and not set(calls[0]) & set(calls[1])
    np.testing.assert_array_equal(out, truth)


def test_gate_serves_predictions_when_the_sample_passes():
    """Predictions within the bound on every sampled soil: no escalation, the
    network's answers are served, and the report states what the sample
    vouches for (1 - 0.05^(1/k) of the soils at 95 % confidence)."""
    S_, R, truth, pred, bad, calls, solver = _gate_case()
    for s in bad:
        pred[s] = truth[s] * 1.01
    soils = list(range(S_))
    out, rep = S.serve(None, np.zeros((S_, 10)), np.ones((S_, R)), soils, gate_fraction=0.2, bound=0.02, seed=5,
                       solver=solver, predictions=pred.reshape(S_ * R, S.BINS))
    assert not rep.failed_soils and not rep.escalated and rep.fdtd_soils == []
    np.testing.assert_array_equal(out, pred)
    assert len(calls) == 1 and rep.sampled_errors.max() < 0.02
    assert rep.vouch_bound() == pytest.approx(1 - 0.05 ** (1 / 10))


def test_polar_targets_reconstruct_the_spectrum_and_the_loss():
    """exp(a + i phi) gives back the normalized target in band, and the
    weighted squares of a small perturbation of (a, phi) equal the relative L2
    squared of the complex spectrum to first order."""
    rng = np.random.default_rng(9)
    n = 4
    params = np.column_stack([rng.uniform(lo, hi, n) for lo, hi in S.PARAM_RANGES])
    r = rng.uniform(1, 50, n)
    scale, delay = S.prior(params, r)
    f = np.arange(S.BINS) * S.BIN_HZ
    G = scale[:, None] * (1 + 0.3 * np.cos(f / 7.0))[None, :] * np.exp(-2j * np.pi * f[None, :] * (delay[:, None] + 0.01))
    t = S.normalize(G, scale, delay)
    pt, w = S.polar_targets(t)
    m = S.band_mask()
    T = t[:, 0::2] + 1j * t[:, 1::2]
    rec = np.exp(pt[:, 0::2] + 1j * pt[:, 1::2])
    np.testing.assert_allclose(rec[:, m], T[:, m], rtol=1e-5, atol=1e-6)
    np.testing.assert_allclose(w.sum(axis=1), 2.0, rtol=1e-6)   # both entries of each bin
    assert np.all(w[:, 0::2][:, ~m] == 0)
    # first-order equivalence of the weighted squares and the relative L2
    eps = 1e-3 * rng.normal(size=pt.shape)
    P = np.exp((pt + eps)[:, 0::2] + 1j * (pt + eps)[:, 1::2])
    rel2 = np.sum(np.abs(P[:, m] - T[:, m]) ** 2, axis=1) / np.sum(np.abs(T[:, m]) ** 2, axis=1)
    wsq = np.sum(w * eps ** 2, axis=1)
    np.testing.assert_allclose(wsq, rel2, rtol=1e-2)


def test_weighted_gradient_matches_central_differences():
    L, P = layers(small_arch())
    rng = np.random.default_rng(10)
    p = random_params(L, P, rng, scale=1.0)
    x = rng.uniform(-1, 1, (3, 15))
    out_len = L[-1].rows * L[-1].n
    t = rng.normal(0, 1, (3, out_len))
    w = rng.uniform(0, 1, (3, out_len)) / out_len
    _, g = O.loss_grad(p, L, x, t, w=w)
    h = 1e-6
    for l in L:
        for i in list(rng.choice(l.n * l.k, size=4, replace=False) + l.w_off) + [l.b_off]:
            if l is L[0] and (i - l.w_off) % l.k >= 15 and i < l.b_off:
                continue
            pp, pm = p.copy(), p.copy()
            pp[i] += h
            pm[i] -= h
            fd_ = (O.loss_grad(pp, L, x, t, w=w)[0] - O.loss_grad(pm, L, x, t, w=w)[0]) / (2 * h)
            assert abs(fd_ - g[i]) <= 1e-6 * max(1.0, abs(g[i])) + 1e-9, (i, fd_, g[i])


def test_magnitude_guard_sends_blown_up_soils_to_the_fdtd():
    """A soil with a prediction whose normalized norm exceeds the limit is
    re-solved and answered by the FDTD even when the 1 % sample misses it."""
    S_, R = 40, 3
    rng = np.random.default_rng(14)
    params = np.column_stack([rng.uniform(lo, hi, S_) for lo, hi in S.PARAM_RANGES])
    ranges = np.tile([2.0, 10.0, 40.0], (S_, 1))
    scale, _ = S.prior(np.repeat(params, R, axis=0), ranges.reshape(-1))
    truth = (rng.normal(size=(S_, R, S.BINS)) + 1j * rng.normal(size=(S_, R, S.BINS))).astype(np.complex64)
    truth *= scale.reshape(S_, R, 1).astype(np.float32)        # normalized norm ~ sqrt(2 * 512)
    truth = S.band_limit(truth)
    pred = truth.copy()
    blown = [5, 17, 33]
    for s in blown:
        pred[s, 2] *= 100.0
    calls = []

    def solver(soils):
        calls.append(list(soils))
        return truth[np.asarray(soils)], None, 0.0, {}

    limit = 4.0 * np.sqrt(2 * S.BINS) * 1.2
    out, rep = S.serve(None, params, ranges, list(range(S_)), gate_fraction=0.01, seed=1, solver=solver,
                       predictions=pred.reshape(S_ * R, S.BINS), norm_limit=limit)
    assert set(blown) <= set(rep.guarded_soils) | set(rep.sampled_soils)
    assert len(calls) == 1 and sorted(calls[0]) == sorted(rep.sampled_soils + rep.guarded_soils)
    for s in blown:
        np.testing.assert_array_equal(out[s], truth[s])
    others = [s for s in range(S_) if s not in rep.guarded_soils and s not in rep.failed_soils]
    for s in others:
        np.testing.assert_array_equal(out[s], pred[s])
    # without the limit, unsampled blown-up soils keep their predictions
    out2, rep2 = S.serve(None, params, ranges, list(range(S_)), gate_fraction=0.01, seed=1, solver=solver,
                         predictions=pred.reshape(S_ * R, S.BINS))
    assert rep2.guarded_soils == []
    missed = [s for s in blown if s not in rep2.sampled_soils]
    assert missed and all(np.array_equal(out2[s], pred[s]) for s in missed)
