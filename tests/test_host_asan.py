"""The library's host code on the CPU under AddressSanitizer.

tests/hipstub builds csrc/run.cpp, api.cpp, tune.cpp, model.cpp, sampler.cpp
and bank.cpp against a recording stand-in for the HIP runtime and for the
kernel launchers (g++ -fsanitize=address, no device code).  Device memory is
host heap memory, so every copy, fill and free the host code issues is
checked by ASan against the allocation it names; the stub also keeps each
stream's unfinished operations and fails on a synchronous copy or another
stream's operation that touches memory one of them still uses, on anything
but kernels recorded into a captured graph, and on a step launch (direct or
replayed) whose descriptors, tile-to-model table, field planes, C-PML strips,
memory variables, tables, trace slots or probe slots are not live
allocations of the sizes the device layout gives them, whose tiling is not
the launched kernel's, or which reads another buffer than the one the
model's previous launch wrote.

The scenarios drive the batched driver as the config-3 bank generator runs
it: cache-sized chunks with the tile-height tuner re-planning the group
between trial launches (the path a profiled run once ended in a host fault
on), the chunk budget tuned over trial chunks re-planned between trial
launches, all models advanced together, mixed kernel variants with failing shots
and a bank generation, the step API, and the batch as a handle advanced in
pieces, reset half-way and after its end, and run again; each with one and
with two time steps per launch.
"""

import os
import subprocess

import pytest

from conftest import ROOT

HERE = os.path.join(ROOT, "tests", "hipstub")


@pytest.fixture(scope="module")
def driver():
    subprocess.run(["make", "-s", "-C", HERE], check=True, capture_output=True)
    return os.path.join(HERE, "lib", "batch_driver")


@pytest.mark.parametrize("span,split", [(1, False), (2, False), (2, True)], ids=["span1", "span2", "span2-split"])
@pytest.mark.parametrize("scenario", ["chunked", "together", "mixed", "sim", "handle", "tuned"])
def test_host_path_is_clean_under_asan(driver, scenario, span, split):
    """split: two-step launches always split into interior and other tiles
    (the stand-in checks both lists against the tiles and the interior test)."""
    env = dict(os.environ, ASAN_OPTIONS="detect_leaks=1:abort_on_error=0", GEOFDTD_STEPS_PER_LAUNCH=str(span))
    env.pop("GEOFDTD_TILE_ROWS", None)
    env.pop("GEOFDTD_SPLIT_MIN_TILES", None)
    if split:
        env["GEOFDTD_SPLIT_MIN_TILES"] = "0"
    p = subprocess.run([driver, scenario], capture_output=True, text=True, timeout=300, env=env)
    print(p.stdout + p.stderr[-3000:])
    assert p.returncode == 0, p.stderr[-3000:]
    assert f"batch_driver {scenario}: clean" in p.stdout
    assert ", 0 live," in p.stdout  # every device allocation was freed
