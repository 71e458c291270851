"""GPU parity: the FP32 HIP path against the FP64 oracle on the same inputs.

Tolerance: relative L2 of each receiver trace <= 1e-4 (the target stated for
FP32 receivers against the FP64 reference scheme).
"""

import math

import numpy as np
import pytest

from media import loam, rel_l2

pytestmark = pytest.mark.gpu

TOL = 1e-4


def _report(name, errs):
    errs = np.asarray(errs)
    print(f"{name}: rel-L2 p50 {np.median(errs):.2e} max {errs.max():.2e}", flush=True)


def test_c1_reference_physics_full(fdtd, oracle):
    """Config 1-R, all 10,000 steps, 49 receivers: every trace within 1e-4."""
    import cosim_amd.workloads as W

    w = W.c1("R")
    got = fdtd.Run(w.shot())
    want = oracle.run(w.model, w.force, w.ranges, w.steps)
    assert got.Dt == want.dt and got.Cells == want.cells == (520, 270)
    errs = [rel_l2(got.Traces[k].Vertical, want.traces[k]) for k in range(len(w.ranges))]
    _report("c1-R", errs)
    assert max(errs) <= TOL
    np.testing.assert_array_equal([t.Range for t in got.Traces], want.snapped)


def test_boussinesq_setup_matches_oracle(fdtd, oracle):
    """The TestBoussinesqStaticLimit shot (fdtd_test.go:57-102) on both paths."""
    m = fdtd.Model(Stack=loam(), MaxRange=40, Depth=40, DominantFreq=25)
    m.Spacing = m.SpacingFor(75, 15)
    dt = fdtd.New(m).Dt()
    steps = int(0.3 / dt)
    sigma = 1 / (2 * math.pi * 25)
    force = [1000 * 0.5 * (1 + math.erf((i * dt - 5 * sigma) / (math.sqrt(2) * sigma))) for i in range(steps)]
    got = fdtd.Run(fdtd.Shot(m, force, [1, 2, 3, 5], steps))
    want = oracle.run(m, force, [1, 2, 3, 5], steps)
    errs = [rel_l2(got.Traces[k].Vertical, want.traces[k]) for k in range(4)]
    _report("boussinesq", errs)
    assert max(errs) <= TOL
