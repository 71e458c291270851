"""The step kernels' arithmetic and indexing, checked on the CPU.

tests/emu runs the per-lane phases of csrc/stencil.h and csrc/march2.h and
the coefficient builders of csrc/coef.h exactly as the gfx950 kernels string
them together, 64 lanes at a time, with every device array a separate
allocation of the size the device arena gives it.  Its traces must match the
FP64 oracle to the FP32 bound, for every tile height; the two-step
(temporally blocked) kernel must give the one-step kernel's bits; the
AddressSanitizer build must run clean.
"""

import os
import subprocess
import sys

import numpy as np
import pytest

from conftest import ROOT
from media import loam, rel_l2

sys.path.insert(0, os.path.join(ROOT, "tests", "emu"))


def cases(fdtd):
    two = [fdtd.Layer(1.3, 420, 190, 1700, 20, 35), fdtd.Layer(0, 700, 320, 2000, 60, 100)]
    return {
        "elastic": fdtd.Model(Stack=loam(), MaxRange=6, Depth=6, Spacing=0.15, DominantFreq=30),
        "sls": fdtd.Model(Stack=loam(), Relax=fdtd.Fit(25, 30), MaxRange=6, Depth=6, Spacing=0.15, DominantFreq=30),
        "gsls-layered": fdtd.Model(Stack=two, GslsMechanisms=3, MaxRange=7.3, Depth=5.1, Spacing=0.1, PMLCells=12),
        "sls-layered": fdtd.Model(Stack=two, Relax=fdtd.Fit(40, 30), MaxRange=9.9, Depth=4.3, Spacing=0.1),
        # 221 x 121 cells: tiles clear of the axis, the surface and the
        # C-PML strips take the kernels' case-free interior path
        "sls-interior": fdtd.Model(Stack=loam(), Relax=fdtd.Fit(25, 30), MaxRange=24, Depth=12, Spacing=0.12,
                                   DominantFreq=30),
    }


@pytest.mark.parametrize("name", ["elastic", "sls", "gsls-layered", "sls-layered"])
@pytest.mark.parametrize("th", [1, 3, 8])
def test_emulated_kernel_matches_oracle(fdtd, oracle, name, th):
    import emu

    m = cases(fdtd)[name]
    o = oracle.OracleSim(m)
    dt = o.Dt()
    steps = 500
    force = oracle.ricker(1000, 35, dt, steps)
    ranges = [0.5, 1.0, 2.5, 4.0]
    want = oracle.run(m, force, ranges, steps)
    cols = [o.Column(r)[0] for r in ranges]
    got, probes = emu.run(m, force, cols, steps, th)
    for k in range(len(ranges)):
        assert rel_l2(got[k], want.traces[k]) <= 1e-5, (name, th, k)
    # the divergence probe saw the oracle's MaxSpeed at step 0 and 256
    assert probes[0] > 0 and probes[1] > 0


@pytest.mark.parametrize("name", ["elastic", "sls", "gsls-layered", "sls-layered", "sls-interior"])
def test_two_step_kernel_gives_the_one_step_bits(fdtd, oracle, name):
    """Two steps per launch (march2.h) with every tile height, over an odd
    number of steps (the last launch takes one), equal one step per launch
    bit for bit -- traces and divergence probes -- since both run the same
    per-cell arithmetic; and the traces match the oracle."""
    import emu

    m = cases(fdtd)[name]
    o = oracle.OracleSim(m)
    dt = o.Dt()
    steps = 301
    force = oracle.ricker(1000, 35, dt, steps)
    ranges = [0.5, 1.0, 2.5, 4.0]
    if name == "sls-interior":
        ranges, steps = [0.5, 1.0, 2.0, 3.0], 151
        force = force[:steps]
    cols = [o.Column(r)[0] for r in ranges]
    want = oracle.run(m, force, ranges, steps)
    one, p1 = emu.run(m, force, cols, steps, 3, span=1)
    for th in (1, 2, 5, 16):
        two, p2 = emu.run(m, force, cols, steps, th, span=2)
        np.testing.assert_array_equal(two, one, err_msg=f"{name} th={th}")
        np.testing.assert_array_equal(p2, p1)
    for k in range(len(ranges)):
        assert rel_l2(one[k], want.traces[k]) <= 1e-5, (name, k)


def test_emulated_kernel_runs_clean_under_asan(tmp_path):
    """The same emulation built with -fsanitize=address, in a child process
    with the sanitizer runtime preloaded."""
    libasan = subprocess.run(["gcc", "-print-file-name=libasan.so"], capture_output=True, text=True).stdout.strip()
    if not os.path.isabs(libasan) or not os.path.exists(libasan):
        pytest.skip("libasan is not available")
    script = tmp_path / "asan_case.py"
    script.write_text(
        "import sys\n"
        f"sys.path[:0] = [{os.path.join(ROOT, 'tests')!r}, {os.path.join(ROOT, 'tests', 'emu')!r}, "
        f"{os.path.join(ROOT, 'oracle')!r}]\n"
        "import conftest, emu, oracle as O\n"
        "fd = conftest.load_package().fdtd\n"
        "two = [fd.Layer(1.3, 420, 190, 1700, 20, 35), fd.Layer(0, 700, 320, 2000, 60, 100)]\n"
        "for m in [fd.Model(Stack=two, GslsMechanisms=3, MaxRange=7.3, Depth=5.1, Spacing=0.1, PMLCells=12),\n"
        "          fd.Model(Stack=two, Relax=fd.Fit(40, 30), MaxRange=3.9, Depth=2.3, Spacing=0.1, PMLCells=5),\n"
        "          fd.Model(Stack=two, GslsMechanisms=3, MaxRange=16, Depth=9, Spacing=0.1, PMLCells=12)]:\n"
        "    dt = O.OracleSim(m).Dt()\n"
        "    for span, ths in ((1, (1, 2, 5, 64)), (2, (1, 2, 3, 7, 64))):\n"
        "        for th in ths:\n"
        "            emu.run(m, O.ricker(1000, 35, dt, 41), [0, 3, 17], 41, th, asan=True, span=span)\n"
        "print('asan ok')\n"
    )
    env = dict(os.environ, LD_PRELOAD=libasan, ASAN_OPTIONS="detect_leaks=0:abort_on_error=1")
    r = subprocess.run([sys.executable, str(script)], capture_output=True, text=True, env=env, timeout=600)
    assert r.returncode == 0 and "asan ok" in r.stdout, r.stderr[-3000:]
