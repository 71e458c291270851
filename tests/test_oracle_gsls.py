"""The attenuation extension (per-layer Q, separate P and S relaxation, N
mechanisms) in the FP64 oracle.  The reference cannot express it, so it is
pinned by what the mathematics fixes:

  * the fitted relaxation spectrum's Q(f) against the constant-Q target over
    the fit band (the generalized Maxwell body's exact Q relation);
  * the static limit: a held load settles to Boussinesq's displacement with
    the RELAXED moduli, mu_R = mu_U (1 - sum Y), which pins the unrelaxed
    anchoring and the long-time sum of every memory variable at once;
  * the reference's own attenuation test (fdtd_test.go:243-293) run on the
    extension with Qp = Qs, where a scalar relaxation carries the Rayleigh
    velocity along unchanged, against the extension's complex modulus;
  * elastic layers (Q = 0) stay exactly elastic.
"""

import cmath
import math

import numpy as np
import pytest

from media import loam, poisson_ratio, rayleigh_velocity
from test_oracle_known_answers import dft


def gsls_model(fdtd, q, qp=0.0, nmech=3, **kw):
    return fdtd.Model(Stack=[fdtd.Layer(0, 500, 200, 1700, q, qp)], GslsMechanisms=nmech, GslsFmin=5, GslsFmax=150,
                      RefFreq=30, **kw)


def modulus_r(oracle, model, q, f):
    """R(f) = 1 - sum Y_l w_l / (w_l + i w), normalised to the unrelaxed modulus."""
    w, y = oracle.gsls_weights(model, q)
    om = 2 * math.pi * f
    return 1 - sum(y[l] * w[l] / (w[l] + 1j * om) for l in range(len(w)))


@pytest.mark.parametrize("q", [10, 20, 30, 60, 80])
@pytest.mark.parametrize("nmech", [3, 4])
def test_fitted_q_is_constant_over_the_band(oracle, fdtd, q, nmech):
    m = gsls_model(fdtd, q, nmech=nmech)
    fs = np.geomspace(5, 150, 41)
    got = np.array([oracle.gsls_q(m, q, f) for f in fs])
    # Q = Re/Im of R; compare with the spectrum evaluated independently here
    r = np.array([modulus_r(oracle, m, q, f) for f in fs])
    np.testing.assert_allclose(got, r.real / r.imag, rtol=1e-12)
    tol = 0.04  # least-squares fit at 2N-1 frequencies: about 3-4 % between them
    assert np.max(np.abs(got / q - 1)) <= tol


def test_single_mechanism_is_a_standard_linear_solid(oracle, fdtd):
    """N = 1: R(w) = 1 - Y w1/(w1 + i w) is an SLS with TauSigma = 1/w1 and
    Delta = Y/(1-Y); its Q must equal visco.SLS.Q of those times."""
    m = gsls_model(fdtd, 30, nmech=1)
    w, y = oracle.gsls_weights(m, 30)
    ts = 1 / w[0]
    te = ts * (1 + y[0] / (1 - y[0]))
    sls = fdtd.SLS(ts, te)
    for f in [5, 20, 27.386, 60, 150]:
        assert oracle.gsls_q(m, 30, f) == pytest.approx(oracle.sls_q(sls, f), rel=1e-12)


def test_static_limit_uses_relaxed_moduli(oracle, fdtd):
    """TestBoussinesqStaticLimit (fdtd_test.go:57-102) on a 3-mechanism
    medium with Qp != Qs: the settled displacement is Boussinesq's with the
    relaxed shear modulus and relaxed Poisson ratio."""
    q, qp = 25.0, 50.0
    m = fdtd.Model(Stack=[fdtd.Layer(0, 500, 200, 1700, q, qp)], GslsMechanisms=3, GslsFmin=5, GslsFmax=150,
                   RefFreq=30, MaxRange=40, Depth=40, DominantFreq=25)
    m.Spacing = oracle.spacing_for(m, 75, 15)
    dt = oracle.OracleSim(m).Dt()
    steps = int(0.3 / dt)
    sigma = 1 / (2 * math.pi * 25)
    force = [1000 * 0.5 * (1 + math.erf((i * dt - 5 * sigma) / (math.sqrt(2) * sigma))) for i in range(steps)]
    res = oracle.run(m, force, [1, 2, 3, 5], steps)
    # relaxed moduli from the spectrum: M_R = M_U (1 - sum Y), M_U from the
    # phase-velocity anchoring at 30 Hz
    def relaxed(v, qq):
        r_ref = modulus_r(oracle, m, qq, 30.0)
        mu_u = 1700 * (v * (1 / cmath.sqrt(r_ref)).real) ** 2
        return mu_u * modulus_r(oracle, m, qq, 0.0).real
    mu_r, m_r = relaxed(200.0, q), relaxed(500.0, qp)
    lam_r = m_r - 2 * mu_r
    nu_r = lam_r / (2 * (lam_r + mu_r))
    for k, r in enumerate(res.snapped):
        u = float(np.sum(res.traces[k]) * dt)
        want = 1000 * (1 - nu_r) / (2 * math.pi * mu_r * r)
        # the elastic run of the same test sits within 0.3 % (1.0027 .. 0.9992)
        assert abs(u / want - 1) <= 0.01, (r, u / want)


def test_attenuation_matches_the_extension_material_model(oracle, fdtd):
    """The reference's attenuation test, on the extension with Qp = Qs = 25."""
    f0, fref = 30.0, 30.0
    ranges = [6, 18]

    def shoot(nmech):
        kw = dict(MaxRange=26, Depth=18, DominantFreq=f0, Spacing=0.12)
        if nmech:
            m = gsls_model(fdtd, 25, nmech=nmech, **kw)
        else:
            m = fdtd.Model(Stack=loam(), RefFreq=fref, **kw)
        dt = oracle.OracleSim(m).Dt()
        steps = int(0.45 / dt)
        return oracle.run(m, oracle.ricker(1000, f0, dt, steps), ranges, steps), m

    (elastic, _), (lossy, m) = shoot(0), shoot(3)
    c_r = rayleigh_velocity(500, 200)
    s_ref = (1 / cmath.sqrt(modulus_r(oracle, m, 25, fref))).real
    dr = ranges[1] - ranges[0]
    for f in [15, 30, 50]:
        def ratio(res):
            return abs(dft(res.traces[1], f, res.dt, res.t0)) / abs(dft(res.traces[0], f, res.dt, res.t0))

        got = ratio(lossy) / ratio(elastic)
        c = c_r * s_ref * cmath.sqrt(modulus_r(oracle, m, 25, f))
        want = math.exp((2 * math.pi * f / c).imag * dr)
        assert abs(got / want - 1) <= 0.05, (f, got, want)


def test_lossless_layers_stay_elastic(oracle, fdtd):
    m = gsls_model(fdtd, 0.0, MaxRange=6, Depth=6, Spacing=0.15, DominantFreq=30)
    s = oracle.OracleSim(m)
    assert not s.lossy() and not s.has_memory()
    e = fdtd.Model(Stack=loam(), MaxRange=6, Depth=6, Spacing=0.15, DominantFreq=30)
    dt = s.Dt()
    assert dt == oracle.OracleSim(e).Dt()
    force = oracle.ricker(1000, 30, dt, 300)
    a = oracle.run(m, force, [1, 2], 300).traces
    b = oracle.run(e, force, [1, 2], 300).traces
    np.testing.assert_array_equal(a, b)
