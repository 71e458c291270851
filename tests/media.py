"""Media and models the tests share, named after the reference's fixtures."""

import math

from conftest import load_package

fd = load_package().fdtd


def loam():
    """soil.Loam (internal/soil/soil.go:50-52): Vp 500, Vs 200, rho 1700, Qs 25."""
    return [fd.Layer(0, 500, 200, 1700, 25)]


def poisson_ratio(vp, vs):
    """soil.HalfSpace.PoissonRatio (soil.go:103-106)."""
    vp2, vs2 = vp * vp, vs * vs
    return (vp2 - 2 * vs2) / (2 * (vp2 - vs2))


def rayleigh_velocity(vp, vs):
    """soil.HalfSpace.RayleighVelocity (soil.go:136-164): bisection on the
    irrational secular equation (2-x)^2 = 4 sqrt(1-x) sqrt(1-x k^2)."""
    k2 = (vs / vp) ** 2

    def f(x):
        return (2 - x) ** 2 - 4 * math.sqrt(1 - x) * math.sqrt(1 - x * k2)

    lo, hi = 1e-9, 1.0
    assert f(lo) < 0 < f(hi)
    for _ in range(200):
        mid = 0.5 * (lo + hi)
        if mid == lo or mid == hi:
            break
        if f(mid) < 0:
            lo = mid
        else:
            hi = mid
    return math.sqrt(0.5 * (lo + hi)) * vs


def rel_l2(a, b):
    import numpy as np

    # complex spectra keep their imaginary parts
    a, b = np.asarray(a), np.asarray(b)
    t = np.complex128 if np.iscomplexobj(a) or np.iscomplexobj(b) else np.float64
    a, b = a.astype(t), b.astype(t)
    return float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300))
