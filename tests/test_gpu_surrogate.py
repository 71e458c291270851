"""The surrogate on the GPU: the BF16 MFMA GEMM against FP64, the network's
forward pass and gradients against the CPU oracle, training, and what the
trained network gains over solving every query with the FDTD."""

import math
import time

import numpy as np
import pytest

import surrogate_oracle as O
from test_surrogate_oracle import ARCH, layers, random_params

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def S():
    import cosim_amd.surrogate as S

    return S


def bf16_bits(x):
    """(raw uint16 bits, the values they hold as float64)."""
    f = np.asarray(x, dtype=np.float32)
    u = f.view(np.uint32).astype(np.uint64)
    r = ((u + 0x7FFF + ((u >> 16) & 1)) >> 16).astype(np.uint16)
    return r, (r.astype(np.uint32) << 16).view(np.float32).astype(np.float64)


@pytest.mark.parametrize("M,N,K,at,bnk,splits", [
    (128, 128, 32, 0, 1, 1),
    (200, 512, 512, 0, 1, 1),     # M not a tile multiple, full encoder width
    (96, 64, 40, 0, 0, 1),        # K not a multiple of 32, B stored [K][N]
    (64, 8, 16, 0, 1, 1),         # the output layer's shape
    (512, 512, 1000, 1, 0, 4),    # a weight gradient: A stored [K][M], split K
    (32, 16, 4096, 1, 0, 7),
    (8, 32, 8000, 1, 0, 5),       # the 8-output convolution's weight gradient (K split over wavefronts)
    (64, 32, 6000, 1, 0, 3),      # the 64-output convolution's weight gradient
    (48, 24, 520, 1, 0, 1),
    (136, 40, 24, 1, 1, 2),
])
def test_gemm_matches_fp64_within_fp32_accumulation(fdtd, M, N, K, at, bnk, splits):
    """Operands are exact BF16 values, so every product is exact in FP32 and
    the only error is the FP32 accumulation: |C - C64| <= K 2^-24 sum|a b|
    (a first-order bound, with a factor 2 for the split sums)."""
    import ctypes as C

    rng = np.random.default_rng(M * 7 + N * 3 + K)
    a_bits, a = bf16_bits(rng.normal(size=(K, M) if at else (M, K)))
    b_bits, b = bf16_bits(rng.normal(size=(N, K) if bnk else (K, N)))
    A = a.T if at else a          # [M][K]
    Bm = b.T if bnk else b        # [K][N]
    want = A @ Bm
    bound = 2 * K * 2.0 ** -24 * (np.abs(A) @ np.abs(Bm)) + 1e-30
    out = np.zeros((M, N), dtype=np.float32)
    lda = M if at else K
    ldb = K if bnk else N
    lib = fdtd.load_library()
    pa = np.ascontiguousarray(a_bits).ctypes.data_as(C.POINTER(C.c_uint16))
    pb = np.ascontiguousarray(b_bits).ctypes.data_as(C.POINTER(C.c_uint16))
    print(f"gemm M={M} N={N} K={K} at={at} b_nk={bnk} splits={splits}", flush=True)
    fdtd._check(lib.gf_gemm_bf16(M, N, K, pa, lda, at, pb, ldb, bnk, splits,
                                 out.ctypes.data_as(C.POINTER(C.c_float))))
    err = np.abs(out - want)
    assert np.all(err <= bound), f"worst excess {np.max(err / bound):.3f} of the bound"


def test_gemm_rejects_shapes_it_cannot_vectorise(fdtd):
    import ctypes as C

    lib = fdtd.load_library()
    buf = np.zeros(4096, dtype=np.uint16).ctypes.data_as(C.POINTER(C.c_uint16))
    out = np.zeros(4096, dtype=np.float32).ctypes.data_as(C.POINTER(C.c_float))
    assert lib.gf_gemm_bf16(16, 16, 12, buf, 16, 0, buf, 16, 1, 1, out) == fdtd.GF_ERR_ARGUMENT
    assert lib.gf_gemm_bf16(12, 16, 16, buf, 16, 1, buf, 16, 1, 1, out) == fdtd.GF_ERR_ARGUMENT


@pytest.fixture(scope="module")
def net(S):
    n = S.Network(max_batch=64, seed=3)
    yield n
    n.close()


def test_layer_table_is_the_documented_architecture(net):
    L, P = layers(ARCH)
    got = [(l.k_in, l.n_out, l.gelu, l.rows_per_query, l.bias_channels, l.w_offset, l.b_offset)
           for l in net.Layers()]
    assert got == [(l.k, l.n, int(l.gelu), l.rows, l.bias_mod, l.w_off, l.b_off) for l in L]
    assert net.ParamCount() == P


def test_forward_matches_the_oracle(net):
    L = O.layers_from(None, net)
    rng = np.random.default_rng(5)
    p = random_params(L, net.ParamCount(), rng, scale=1.0).astype(np.float32)
    net.SetParams(p)
    np.testing.assert_array_equal(net.Params(), p)
    x = rng.uniform(-1, 1, (40, 15)).astype(np.float32)
    got = net.Forward(x)
    want_bf, _ = O.forward(p.astype(np.float64), L, x.astype(np.float64), bf=True)
    want, _ = O.forward(p.astype(np.float64), L, x.astype(np.float64), bf=False)
    e_bf = np.linalg.norm(got - want_bf, axis=1) / np.linalg.norm(want_bf, axis=1)
    e = np.linalg.norm(got - want, axis=1) / np.linalg.norm(want, axis=1)
    print(f"forward rel-L2 vs BF16-emulating oracle: max {e_bf.max():.2e}; vs FP64: max {e.max():.2e}", flush=True)
    assert e_bf.max() <= 3e-3
    assert e.max() <= 3e-2


def test_gradients_match_the_oracle(net):
    """Each layer's weight and bias gradient against the oracle's, as norm-wise
    relative error: at most 1e-2 against exact FP64 and 5e-3 against the
    oracle with the GPU's BF16 operand rounding."""
    L = O.layers_from(None, net)
    rng = np.random.default_rng(6)
    p = random_params(L, net.ParamCount(), rng, scale=1.0).astype(np.float32)
    net.SetParams(p)
    B = 32
    x = rng.uniform(-1, 1, (B, 15)).astype(np.float32)
    t = rng.normal(0, 1, (B, 1024)).astype(np.float32)
    loss, g = net.LossGrad(x, t)
    l_bf, g_bf = O.loss_grad(p.astype(np.float64), L, x.astype(np.float64), t.astype(np.float64), bf=True)
    l64, g64 = O.loss_grad(p.astype(np.float64), L, x.astype(np.float64), t.astype(np.float64), bf=False)
    assert loss == pytest.approx(l64, rel=2e-2) and loss == pytest.approx(l_bf, rel=2e-3)
    for i, l in enumerate(L):
        for name, a, b in (("W", l.w_off, l.w_off + l.n * l.k), ("b", l.b_off, l.b_off + l.bias_mod)):
            ref, ref_bf, got = g64[a:b], g_bf[a:b], g[a:b].astype(np.float64)
            r64 = np.linalg.norm(got - ref) / np.linalg.norm(ref)
            rbf = np.linalg.norm(got - ref_bf) / np.linalg.norm(ref_bf)
            print(f"layer {i} d{name}: rel err {r64:.2e} (FP64), {rbf:.2e} (BF16 oracle)", flush=True)
            assert r64 <= 1e-2 and rbf <= 5e-3, (i, name)


def test_weighted_gradients_match_the_oracle(net):
    L = O.layers_from(None, net)
    rng = np.random.default_rng(12)
    p = random_params(L, net.ParamCount(), rng, scale=1.0).astype(np.float32)
    net.SetParams(p)
    B = 16
    x = rng.uniform(-1, 1, (B, 15)).astype(np.float32)
    t = rng.normal(0, 1, (B, 1024)).astype(np.float32)
    w = (rng.uniform(0, 1, (B, 1024)) / 512).astype(np.float32)
    loss, g = net.LossGrad(x, t, w)
    l_bf, g_bf = O.loss_grad(p.astype(np.float64), L, x.astype(np.float64), t.astype(np.float64), bf=True,
                             w=w.astype(np.float64))
    assert loss == pytest.approx(l_bf, rel=2e-3)
    for i, l in enumerate(L):
        for name, a, b in (("W", l.w_off, l.w_off + l.n * l.k), ("b", l.b_off, l.b_off + l.bias_mod)):
            r = np.linalg.norm(g[a:b] - g_bf[a:b]) / np.linalg.norm(g_bf[a:b])
            assert r <= 5e-3, (i, name, r)


def test_polar_predict_exponentiates(S):
    net = S.Network(max_batch=64, seed=3, polar=True)
    rng = np.random.default_rng(13)
    x = rng.uniform(-1, 1, (16, 15)).astype(np.float32)
    scale = rng.uniform(1e-9, 1e-8, 16)
    delay = rng.uniform(0.0, 0.4, 16)
    o = net.Forward(x).astype(np.float64)
    f = np.arange(S.BINS) * S.BIN_HZ
    want = scale[:, None] * np.exp(o[:, 0::2] + 1j * (o[:, 1::2] - 2 * np.pi * f[None, :] * delay[:, None]))
    got = net.Predict(x, scale, delay)
    m = S.band_mask()
    assert np.all(got[:, ~m] == 0)
    np.testing.assert_allclose(got[:, m], want[:, m], rtol=2e-5, atol=1e-6 * np.abs(want).max())
    net.close()


def test_predict_applies_the_prior(S):
    net = S.Network(max_batch=64, seed=3, band=None)
    rng = np.random.default_rng(7)
    x = rng.uniform(-1, 1, (24, 15)).astype(np.float32)
    scale = rng.uniform(1e-9, 1e-8, 24)
    delay = rng.uniform(0.0, 0.4, 24)
    o = net.Forward(x).astype(np.float64).view(np.complex128)
    f = np.arange(S.BINS) * S.BIN_HZ
    want = scale[:, None] * o * np.exp(-2j * np.pi * f[None, :] * delay[:, None])
    got = net.Predict(x, scale, delay)
    np.testing.assert_allclose(got, want, rtol=1e-5, atol=1e-6 * np.abs(want).max())
    net.close()
    banded = S.Network(max_batch=64, seed=3)
    got = banded.Predict(x, scale, delay)
    m = S.band_mask()
    assert np.all(got[:, ~m] == 0) and m.sum() == 194
    np.testing.assert_allclose(got[:, m], want[:, m], rtol=1e-5, atol=1e-6 * np.abs(want).max())
    banded.close()


def test_training_reduces_the_loss(S):
    """Adam on a smooth synthetic map from 11 inputs to spectra: the loss falls
    by an order of magnitude in a few hundred steps."""
    n = S.Network(max_batch=256, seed=11)
    try:
        rng = np.random.default_rng(8)
        x = rng.uniform(-1, 1, (2048, 15)).astype(np.float32)
        k = np.arange(512)
        amp = np.exp(-k / (80 + 40 * x[:, :1]))
        ph = 2 * np.pi * k * (0.01 + 0.005 * x[:, 1:2])
        t = np.empty((2048, 1024), dtype=np.float32)
        t[:, 0::2] = amp * np.cos(ph)
        t[:, 1::2] = amp * np.sin(ph)
        losses = n.Train(x, t, steps=400, batch=256, lr=1e-3, lr_min=1e-4, seed=1)
        first, last = losses[:10].mean(), losses[-10:].mean()
        st = n.Stats()
        print(f"training: loss {first:.3f} -> {last:.4f}; {st['train_ms'] / st['train_steps']:.3f} ms per step",
              flush=True)
        assert np.all(np.isfinite(losses))
        assert last < 0.1 * first
    finally:
        n.close()


def test_training_is_deterministic(S):
    """Training runs on two streams (each layer's weight gradient beside the
    input gradient, three gradient sets, the weight transposes beside the
    next forward pass) with the steps enqueued back to back.  Every
    reduction has a fixed order, so two runs from the same seeds give the
    same losses and parameters bit for bit; a missing wait between the
    streams would show here as a difference.  Several epochs of a small set,
    so the double-buffered batch order is reused."""
    rng = np.random.default_rng(31)
    x = rng.uniform(-1, 1, (5000, S.FEATURES)).astype(np.float32)
    w = rng.normal(0, 1, (S.FEATURES, 2 * S.BINS)).astype(np.float32)
    t = np.tanh(x @ w).astype(np.float32)
    runs = []
    for _ in range(2):
        n = S.Network(max_batch=4096, seed=5)
        try:
            losses = n.Train(x, t, steps=300, batch=1024, lr=2e-3, lr_min=4e-5, seed=6)
            runs.append((losses, n.Params()))
        finally:
            n.close()
    (l0, p0), (l1, p1) = runs
    print(f"deterministic training: loss {l0[:10].mean():.4f} -> {l0[-10:].mean():.5f}", flush=True)
    assert np.all(np.isfinite(l0)) and l0[-10:].mean() < 0.5 * l0[:10].mean()
    np.testing.assert_array_equal(l0, l1)
    np.testing.assert_array_equal(p0, p1)


def test_trained_network_gains_over_fdtd_only(S):
    """On FDTD-generated spectra of the layered family, against the method
    without the network -- the analytic prior alone (the Boussinesq amplitude
    1/(2 pi mu1 r) with the top layer's Rayleigh delay) -- and against the
    batched FDTD's cost: trained on 384 soils, the network's median held-out
    error is at most 0.6 of the prior's and below the untrained network's,
    and answering a query costs under 1 % of solving it."""
    train = S.solve_set(S.soil_set(0, 0, 384), seconds=1.0)
    test = S.solve_set(S.soil_set(0, 10_000, 24), seconds=1.0)
    x, t, _, _, _ = S.dataset(train)
    xt, tt, sc, dl, Gt = S.dataset(test)
    f = np.arange(S.BINS) * S.BIN_HZ
    prior_only = S.band_limit(sc[:, None] * np.exp(-2j * np.pi * f[None, :] * dl[:, None]))
    n = S.Network(max_batch=4096, seed=5)
    try:
        untrained = S.rel_l2(n.Predict(xt, sc, dl), Gt)
        losses = n.Train(x, t, steps=3000, batch=2048, lr=1e-3, lr_min=2e-5, seed=2)
        after = S.rel_l2(n.Predict(xt, sc, dl), Gt)
        n.Predict(np.repeat(xt, 8, axis=0), np.repeat(sc, 8), np.repeat(dl, 8))
        st = n.Stats()
    finally:
        n.close()
    base = S.rel_l2(prior_only, Gt)
    fdtd_per_trace = train.fdtd_seconds / t.shape[0]
    sur_per_trace = st["forward_ms"] / 1e3 / st["forward_queries"]
    print(f"held-out rel-L2 median: untrained {np.median(untrained):.3f}, prior only {np.median(base):.3f}, "
          f"trained {np.median(after):.3f} (p90 {np.quantile(after, 0.9):.3f}); train loss "
          f"{losses[:20].mean():.3f} -> {losses[-20:].mean():.4f}; FDTD {fdtd_per_trace * 1e3:.2f} ms/trace, "
          f"surrogate {sur_per_trace * 1e6:.3f} us/trace ({fdtd_per_trace / sur_per_trace:.0f}x)", flush=True)
    assert np.median(after) <= 0.6 * np.median(base)
    assert np.median(after) < np.median(untrained)
    assert fdtd_per_trace / sur_per_trace >= 100


@pytest.fixture(scope="module")
def held_out(S):
    """165 held-out soils of the layered family (8,250 queries: one inference
    batch of 8,192 and a ragged one) with their FDTD banks."""
    return S.solve_set(S.soil_set(0, 2_000_000, 165), seconds=1.0)


def test_receiver_ranges_equal_the_batch_snapping(S, held_out):
    """The host's snapped receiver ranges (what a query names before any
    solve) equal the ranges the batched FDTD recorded at."""
    np.testing.assert_array_equal(S.receiver_ranges(held_out.soils), held_out.ranges)


def test_served_request_with_degraded_weights_falls_back_to_the_fdtd(S, held_out):
    """Config 5's service on the device with the real batched-FDTD re-solve:
    an untrained network (degraded weights) fails the 1 % sample, so the
    request escalates and every soil is answered by its FDTD solution; what
    is served then equals the soils' FDTD banks (batch = single, bit for
    bit), so its error against the FDTD is zero at p50 and p99."""
    h = held_out
    net = S.Network(max_batch=8192, seed=3)
    try:
        served, rep = S.serve(net, h.params, h.ranges, h.soils, gate_fraction=0.01, bound=0.02, seed=1)
        st = net.Stats()
    finally:
        net.close()
    truth = S.band_limit(h.spectra)
    err = S.rel_l2(served.reshape(-1, S.BINS), truth.reshape(-1, S.BINS))
    print(f"degraded weights: {len(rep.sampled_soils)} sampled, {len(rep.failed_soils)} failed (sampled rel-L2 "
          f"max {rep.sampled_errors.max():.2f}), escalated {rep.escalated}, {len(rep.fdtd_soils)} of {rep.soils} "
          f"soils from the FDTD; served rel-L2 p50 {np.median(err):.1e} p99 {np.quantile(err, 0.99):.1e}; "
          f"inference {st['forward_queries']} queries", flush=True)
    assert len(rep.sampled_soils) == 2 and rep.failed_soils and rep.escalated
    assert rep.fdtd_soils == list(range(165))
    assert st["forward_queries"] >= 8250
    np.testing.assert_array_equal(served, truth)


def test_served_request_within_the_bound_keeps_the_predictions(S, held_out):
    """The same request with predictions within the bound (the FDTD banks
    perturbed by 1 % in amplitude and 0.005 rad in phase, standing in for a
    network inside its bound): the sampled soils are re-solved by the
    batched FDTD and pass, nothing escalates, and the predictions are served."""
    h = held_out
    rng = np.random.default_rng(4)
    truth = S.band_limit(h.spectra)
    pert = (1 + 0.01 * rng.uniform(-1, 1, truth.shape)) * np.exp(0.005j * rng.normal(size=truth.shape))
    pred = (truth * pert).astype(np.complex64)
    served, rep = S.serve(None, h.params, h.ranges, h.soils, gate_fraction=0.01, bound=0.02, seed=1,
                          predictions=pred.reshape(-1, S.BINS))
    print(f"within bound: sampled {rep.sampled_soils}, rel-L2 max {rep.sampled_errors.max():.4f}, "
          f"vouches for < {rep.vouch_bound():.2f} failing soils at 95 %", flush=True)
    assert not rep.failed_soils and not rep.escalated and rep.fdtd_soils == []
    assert 0.004 < rep.sampled_errors.max() < 0.02
    np.testing.assert_array_equal(served, pred)


def test_fused_inference_equals_the_per_layer_gemms(S, monkeypatch):
    """The fused encoder kernel (seven layers in one launch, activations in
    LDS, fragment-packed weights, the FP32 features rounded to BF16 inside
    it) gives the per-layer GEMMs' bits, whose features a separate launch
    rounds: the same MFMA k-step order, the same bias + GELU epilogue and
    BF16 rounding.  8,200 queries: a full batch of 8,192 and a ragged one of
    8."""
    rng = np.random.default_rng(21)
    net = S.Network(max_batch=8192, seed=7)
    try:
        x = rng.uniform(-1, 1, (8200, S.FEATURES)).astype(np.float32)
        fused = net.Forward(x)
        monkeypatch.setenv("GEOFDTD_SUR_UNFUSED", "1")
        per_layer = net.Forward(x)
    finally:
        net.close()
    print(f"fused vs per-layer: max |diff| {np.max(np.abs(fused - per_layer)):.3e}", flush=True)
    np.testing.assert_array_equal(fused, per_layer)


@pytest.mark.parametrize("polar", [False, True])
def test_fused_predict_equals_the_per_layer_path(S, monkeypatch, polar):
    """Predict with the de-normalization fused into the network's launch gives
    the bits of the per-layer GEMMs followed by the separate de-normalization
    kernel (both call one device function for a bin, products not
    contracted).  8,203 queries: a full batch and a ragged one of 11."""
    rng = np.random.default_rng(22)
    net = S.Network(max_batch=8192, seed=9, polar=polar)
    try:
        x = rng.uniform(-1, 1, (8203, S.FEATURES)).astype(np.float32)
        scale = rng.uniform(1e-9, 1e-8, 8203)
        delay = rng.uniform(0.0, 0.4, 8203)
        fused = net.Predict(x, scale, delay)
        monkeypatch.setenv("GEOFDTD_SUR_UNFUSED", "1")
        per_layer = net.Predict(x, scale, delay)
    finally:
        net.close()
    m = S.band_mask()
    assert np.all(fused[:, ~m] == 0) and np.any(fused[:, m] != 0)
    print(f"fused vs per-layer predict (polar={polar}): max |diff| {np.max(np.abs(fused - per_layer)):.3e}",
          flush=True)
    np.testing.assert_array_equal(fused, per_layer)
