"""Whole shots on the GPU against the FP64 oracle, as fdtd.Run drives them
(internal/fdtd/run.go:88-127): config 2 on its full 2048 x 1024 grid for all
80,000 steps (2 s), both physics variants, and the two config-3 soils with
the fewest cell-steps over their whole second, batched together.

The oracle's traces come from tests/golden/full_traces (written once by
tests/golden/make_full_traces.py; a config-2 shot is an hour of CPU), every
`decimate`-th sample in float64.  The tolerance is the path's: relative L2
<= 1e-4 per trace.  Each trace's late window (after the direct arrivals have
passed every receiver) is also compared, as the peak there over the trace's
peak: it is where a growing C-PML mode or FP32 drift would show first.
"""

import json
import os

import numpy as np
import pytest

from conftest import ROOT
from media import rel_l2

pytestmark = pytest.mark.gpu

GOLD = os.path.join(ROOT, "tests", "golden", "full_traces")
TOL = 1e-4


def golden(name):
    z = np.load(os.path.join(GOLD, f"{name}.npz"))
    with open(os.path.join(GOLD, f"{name}.json")) as f:
        info = json.load(f)
    return z, info


def check(name, got, z, info):
    """got: [receivers][steps] float64 from the GPU."""
    d = int(z["decimate"])
    want = z["traces"]
    got = np.asarray(got)
    assert got.shape[1] == int(z["steps"])
    g = got[:, ::d]
    errs = np.array([rel_l2(g[k], want[k]) for k in range(len(want))])
    n0 = int(round(info["late_window_start_s"] / float(z["dt"])))
    late_got = np.max(np.abs(got[:, n0:]), axis=1) / np.max(np.abs(got), axis=1)
    late_want = np.asarray(info["late_over_peak"])
    print(f"{name}: {len(errs)} traces x {got.shape[1]} steps, rel-L2 p50 {np.median(errs):.2e} "
          f"max {errs.max():.2e}; late-window peak / peak: GPU {late_got.max():.2e}, oracle {late_want.max():.2e}",
          flush=True)
    assert errs.max() <= TOL, f"{name}: worst rel-L2 {errs.max():.3e}"
    # peaks at full rate agree, and the late window is as quiet as the oracle's
    np.testing.assert_allclose(np.max(np.abs(got), axis=1), z["peak"], rtol=2e-4)
    assert np.all(np.abs(late_got - late_want) <= 1e-3 * 1.0 + 1e-2 * late_want)


@pytest.mark.parametrize("variant", ["R", "X"])
def test_c2_whole_shot(fdtd, variant):
    """Config 2, all 80,000 steps, ten receivers from 0.5 to 50 m."""
    import cosim_amd.workloads as W

    name = f"c2-{variant}"
    if not os.path.exists(os.path.join(GOLD, f"{name}.npz")):
        pytest.skip(f"{name} golden traces not written yet")
    z, info = golden(name)
    w = W.c2(variant)
    assert w.steps == int(z["steps"]) == 80_000
    ranges = [float(r) for r in z["ranges"]]
    got = fdtd.Run(fdtd.Shot(w.model, w.force, ranges, w.steps))
    assert got.Dt == float(z["dt"]) and got.Cells == (2048, 1024)
    np.testing.assert_array_equal([t.Range for t in got.Traces], z["snapped"])
    check(name, [t.Vertical for t in got.Traces], z, info)


def test_c3_smallest_soils_whole_second(fdtd):
    """Config-3 soils 204 and 153 (seed 0), batched together, over their
    whole second (6,934 and 7,025 steps), ten bank ranges each."""
    import cosim_amd.workloads as W

    names = ["c3-X-204", "c3-X-153"]
    ws = [W.c3(n_models=1, first=204)[0], W.c3(n_models=1, first=153)[0]]
    golds = [golden(n) for n in names]
    shots = [fdtd.Shot(w.model, w.force, [float(r) for r in z["ranges"]], w.steps) for w, (z, _) in zip(ws, golds)]
    for w, (z, _) in zip(ws, golds):
        assert w.steps == int(z["steps"])
    got = fdtd.RunBatch(shots)
    for name, g, (z, info) in zip(names, got, golds):
        assert g.Dt == float(z["dt"])
        check(name, [t.Vertical for t in g.Traces], z, info)
