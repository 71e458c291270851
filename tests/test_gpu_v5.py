"""V5 on the GPU: the FDTD solver against the frequency-wavenumber solver, two
methods that share no code (internal/fdtd/v5_test.go).

Restated from the reference's test: each grid run is compared receiver by
receiver with the f-k synthesis on the same instants (peak ratio, lag by
cross-correlation with parabolic refinement, rms residual after removing the
lag); two spacings are extrapolated to zero spacing at first order
(richardson, v5_test.go:187) and the extrapolated peak ratio must be 1 and the
residual 0 within 2 % (homogeneous) or 3 % (layered); the first-order
exponent itself is measured over three spacings.  Both paths here are the
GPU's: k_step in float32 and the f-k kernels in float64.
"""

import math

import numpy as np
import pytest

from conftest import load_package

pkg = load_package()
fd, fk = pkg.fdtd, pkg.fk

pytestmark = pytest.mark.gpu

V5_FREQ, V5_BAND, V5_REF_FREQ, V5_Q = 25.0, 90.0, 30.0, 25.0
LOAM = [fd.Layer(0, 500, 200, 1700, 25)]


def compare(got, want, dt):
    """agreement of one receiver (v5_test.go:95-145): peak ratio, lag (s),
    residual rms after the lag."""
    got = np.asarray(got, dtype=np.float64)
    want = np.asarray(want, dtype=np.float64)
    n = len(want)
    peak = np.max(np.abs(got)) / np.max(np.abs(want))
    max_lag = n // 8

    def corr(k):
        # sum over i of got[i] * want[i - k]
        if k >= 0:
            return float(np.dot(got[k:], want[: n - k]))
        return float(np.dot(got[: n + k], want[-k:]))

    best, best_at = -math.inf, 0
    for k in range(-max_lag, max_lag + 1):
        c = corr(k)
        if c > best:
            best, best_at = c, k
    l, r = corr(best_at - 1), corr(best_at + 1)
    lag = (best_at - 0.5 * (r - l) / (l - 2 * best + r)) * dt if l - 2 * best + r != 0 else best_at * dt
    v = np.zeros(n)
    if best_at >= 0:
        v[best_at:] = want[: n - best_at]
    else:
        v[: n + best_at] = want[-best_at:]
    residual = math.sqrt(np.sum((got - v) ** 2) / np.sum(v * v))
    return peak, lag, residual


def shoot(stack, sls, spacing, r_max, depth, ranges, seconds):
    """One grid run and its comparison with the f-k path (v5_test.go:148-179)."""
    m = fd.Model(Stack=stack, Relax=sls, RefFreq=V5_REF_FREQ, MaxRange=r_max, Depth=depth, Spacing=spacing,
                 DominantFreq=V5_FREQ)
    dt = m.Grid()[0]
    steps = int(seconds / dt)
    force = fd.Ricker(1000, V5_FREQ, dt, steps)
    print(f"[v5] grid h={spacing} m, {steps} steps", flush=True)
    res = fd.Run(fd.Shot(m, force, ranges, steps))
    snapped = [tr.Range for tr in res.Traces]
    ref = fk.Synthesise(stack, sls, V5_REF_FREQ, snapped, 1000, V5_FREQ, V5_BAND, dt, res.T0, steps, 0)
    return [compare(tr.Vertical, ref[i], dt) for i, tr in enumerate(res.Traces)], snapped


def richardson(coarse, fine):
    return 2 * fine - coarse


def check(label, coarse, fine, ranges, tol):
    for i, r in enumerate(ranges):
        p = richardson(coarse[i][0], fine[i][0])
        lag = richardson(coarse[i][1], fine[i][1])
        res = richardson(coarse[i][2], fine[i][2])
        print(f"[v5] {label} r={r:5.2f} m: peak ratio {coarse[i][0]:.4f} -> {fine[i][0]:.4f} (extrapolated {p:.4f}), "
              f"lag {1e3 * coarse[i][1]:+.3f} -> {1e3 * fine[i][1]:+.3f} ms (extrapolated {1e3 * lag:+.3f} ms), "
              f"residual {100 * coarse[i][2]:.2f}% -> {100 * fine[i][2]:.2f}% (extrapolated {100 * res:.2f}%)",
              flush=True)
        assert abs(p - 1) <= tol, (label, r, p)
        assert abs(res) <= tol, (label, r, res)


def test_v5_homogeneous():
    sls = fd.Fit(V5_Q, V5_REF_FREQ)
    ranges = [4.0, 8.0, 16.0]
    coarse, snapped = shoot(LOAM, sls, 0.111, 26, 18, ranges, 0.3)
    fine, _ = shoot(LOAM, sls, 0.0555, 26, 18, ranges, 0.3)
    check("homogeneous", coarse, fine, snapped, 0.02)


def test_v5_layered():
    sls = fd.Fit(V5_Q, V5_REF_FREQ)
    stack = [fd.Layer(2, 350, 140, 1600), fd.Layer(0, 800, 320, 2000)]
    coarse, snapped = shoot(stack, sls, 0.1, 16, 10, [3.0, 6.0, 10.0], 0.3)
    fine, _ = shoot(stack, sls, 0.05, 16, 10, [3.0, 6.0, 10.0], 0.3)
    check("layered", coarse, fine, snapped, 0.03)


def test_convergence_is_first_order():
    sls = fd.Fit(V5_Q, V5_REF_FREQ)
    errs = []
    for h in (0.222, 0.111, 0.0555):
        a, _ = shoot(LOAM, sls, h, 20, 14, [8.0], 0.3)
        errs.append(abs(a[0][0] - 1))
        print(f"[v5] h={h:.4f} m: peak ratio {a[0][0]:.4f}, residual {100 * a[0][2]:.2f}%", flush=True)
    for e0, e1 in zip(errs, errs[1:]):
        order = math.log2(e0 / e1)
        print(f"[v5] halving the spacing reduced the amplitude error {e0 / e1:.2f}x (order {order:.2f})", flush=True)
        assert 0.7 <= order <= 1.6
