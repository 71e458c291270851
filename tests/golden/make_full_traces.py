"""Writes the full-duration FP64 oracle traces the GPU parity tests compare
against: config 2 (2048 x 1024 cells, 80,000 steps, both physics variants)
at ten receivers from 0.5 to 50 m, and the two config-3 soils with the
fewest cell-steps over their whole second.

The oracle needs one to two hours per config-2 shot on one core (a quarter
of that with its row-parallel build on OMP_NUM_THREADS threads, which gives
the same bits), so the traces are computed once and committed; the GPU box
never runs the oracle for them.  Each file holds every `decimate`-th sample (the signals are
band-limited far below the decimated Nyquist rate, so the relative L2 over
the kept samples measures the same error) in float64, the oracle's peak per
trace at full rate, and the late-window diagnostics.

    python tests/golden/make_full_traces.py c2-R c2-X c3

Config-2 shots run step by step through the oracle's Sim (the same calls, in
the same order, as its Run) and checkpoint their state every few thousand
steps into GOLDEN_CHECKPOINTS (default /tmp/golden_checkpoints), so an
interrupted run continues where it stopped.
"""

import json
import os
import sys
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
OUT = os.path.join(HERE, "full_traces")
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.join(ROOT, "oracle"))

C2_RANGES = [0.5, 1.0, 2.0, 3.5, 5.0, 8.0, 12.5, 20.0, 32.0, 50.0]
C3_RANGES = [1.0, 2.0, 3.0, 5.0, 8.0, 12.0, 18.0, 25.0, 35.0, 50.0]
# The config-3 soils (seed 0) with the fewest cell-steps: 683 x 349 cells for
# 6,934 steps and 697 x 356 for 7,025 steps.
C3_SOILS = [204, 153]


def late_window(traces, dt, start_s):
    """Peak |v| after `start_s` over the trace's overall peak, per trace."""
    n0 = int(round(start_s / dt))
    full = np.max(np.abs(traces), axis=1)
    late = np.max(np.abs(traces[:, n0:]), axis=1) if n0 < traces.shape[1] else np.zeros(len(traces))
    return (late / full).tolist()


class _Run:
    def __init__(self, traces, snapped, peak, dt, cells):
        self.traces, self.snapped, self.peak, self.dt, self.cells = traces, snapped, peak, dt, cells


def run_resumable(name, w, ranges, block=2000):
    """fdtd.Run's loop (run.go:104-126) over the oracle's Sim: step, record
    the receivers, probe MaxSpeed every 256 steps; state checkpointed every
    `block` steps."""
    import oracle as O

    ck = os.environ.get("GOLDEN_CHECKPOINTS", "/tmp/golden_checkpoints")
    os.makedirs(ck, exist_ok=True)
    state, meta = os.path.join(ck, f"{name}.state"), os.path.join(ck, f"{name}.npz")
    o = O.OracleSim(w.model)
    o.Drive(w.force)
    cols, snapped = zip(*[o.Column(r) for r in ranges])
    steps = w.steps
    traces = np.zeros((len(ranges), steps))
    peak, start = 0.0, 0
    if os.path.exists(state) and os.path.exists(meta):
        z = np.load(meta)
        start = int(z["done"])
        traces[:, :start] = z["traces"]
        peak = float(z["peak"])
        o.checkpoint(state, save=False)
        print(f"{name}: resuming at step {start}", flush=True)
    t = time.time()
    for n in range(start, steps):
        o.Step()
        for i, c in enumerate(cols):
            traces[i, n] = o.SurfaceVelocity(c)
        if n % 256 == 0:
            p = o.MaxSpeed()
            peak = p if (p > peak or p != p) else peak
            if not np.isfinite(p):
                raise SystemExit(f"{name}: diverged at step {n} of {steps}")
        if (n + 1) % block == 0 and n + 1 < steps:
            o.checkpoint(state + ".tmp", save=True)
            np.savez(meta + ".tmp.npz", traces=traces[:, :n + 1], peak=peak, done=n + 1)
            os.replace(state + ".tmp", state)
            os.replace(meta + ".tmp.npz", meta)
            print(f"{name}: step {n + 1} of {steps} ({time.time() - t:.0f} s this sitting)", flush=True)
    return _Run(traces, np.asarray(snapped), peak, o.Dt(), o.Cells())


def write(name, w, ranges, decimate, late_start, resumable=False):
    import oracle as O

    O.use_row_threads()
    t = time.time()
    r = run_resumable(name, w, ranges) if resumable else O.run(w.model, w.force, ranges, w.steps)
    el = time.time() - t
    tr = r.traces
    np.savez_compressed(
        os.path.join(OUT, f"{name}.npz"),
        traces=tr[:, ::decimate].copy(),
        decimate=np.int64(decimate),
        ranges=np.asarray(ranges, dtype=np.float64),
        snapped=r.snapped,
        dt=np.float64(r.dt),
        steps=np.int64(w.steps),
        peak=np.max(np.abs(tr), axis=1),
    )
    info = {
        "workload": name,
        "cells": list(r.cells),
        "steps": int(w.steps),
        "dt": r.dt,
        "ranges_m": list(ranges),
        "decimate": decimate,
        "oracle_seconds": round(el, 1),
        "late_window_start_s": late_start,
        "late_over_peak": late_window(tr, r.dt, late_start),
        "max_speed_peak": r.peak,
    }
    with open(os.path.join(OUT, f"{name}.json"), "w") as f:
        json.dump(info, f, indent=1)
    print(f"{name}: {r.cells} x {w.steps} steps in {el:.0f} s; late/peak "
          f"{min(info['late_over_peak']):.2e}..{max(info['late_over_peak']):.2e}", flush=True)


def main(which):
    import __graft_entry__ as g

    g._package()
    from cosim_amd import workloads as W

    os.makedirs(OUT, exist_ok=True)
    for name in which:
        if name in ("c2-R", "c2-X"):
            write(name, W.c2(name[-1]), C2_RANGES, 4, 1.5, resumable=True)
        elif name == "c3":
            for i in C3_SOILS:
                w = W.c3(n_models=1, first=i)[0]
                write(f"c3-X-{i}", w, C3_RANGES, 2, 0.75)
        else:
            raise SystemExit(f"unknown workload {name}")


if __name__ == "__main__":
    main(sys.argv[1:] or ["c2-R", "c2-X", "c3"])
