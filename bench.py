#!/usr/bin/env python3
"""Benchmark: Green's-function traces per second of the MI355X FDTD solver.

The default workload is config 3, batched bank generation: 256 seeded soil
models (600-2048 x 300-1024 cells, GSLS attenuation, 1 s each, 50 receivers)
advanced together and reduced on the device to bank spectra.  The batch is
set up before the timed region (every model's state, tables, source series
and receiver columns resident in HBM, launches planned and tuned); one
benchmark "step" advances every model by 1/steps of its run, so the timed
steps generate the whole batch once, the last step ending with the spectra.

Printed (one JSON line on stdout):
  value       traces/s over the timed steps (wall clock; every step ends with
              a device synchronization)
  roofline    the step kernel's algorithmic HBM bytes per launch over its
              average launch time (HIP events around each chunk's step loop
              on the solver's stream), against MI355X's 8 TB/s; `traffic` is
              the rocprofv3-measured HBM bytes per launch when a committed
              profile summary provides it
  cpu_baseline the FP64 CPU restatement of the reference scheme on the same
              workload, a bounded sample of time steps on every worker thread,
              scaled by cell-steps
  accuracy    whole-shot traces of two of the batch's soils against the FP64
              oracle's committed traces
  also        config 2 (2048 x 1024, 80,000 steps) as single shots; fk,
              bank_lookup, stream, engine: the adjacent components
"""

from __future__ import annotations

import argparse
import importlib.util
import json
import math
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
PKG_DIR = os.path.join(ROOT, "co-simulation-of-seismic-sensors-and-ros-for-robotic-human-detection_amd")
METRIC = "GF traces/s: batched FDTD vs surrogate at stated L2 err; stencil HBM GB/s"
HBM_PEAK_GBS = 8000.0
SUR_WEIGHTS = os.path.join(PKG_DIR, "weights", "surrogate_layered.npz")


def log(msg: str) -> None:
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def load_package():
    name = "cosim_amd"
    if name in sys.modules:
        return sys.modules[name]
    spec = importlib.util.spec_from_file_location(
        name, os.path.join(PKG_DIR, "__init__.py"), submodule_search_locations=[PKG_DIR]
    )
    mod = importlib.util.module_from_spec(spec)
    sys.modules[name] = mod
    spec.loader.exec_module(mod)
    return mod


def workload(name: str):
    import cosim_amd.workloads as W

    kind, variant = name.split("-")
    if kind == "c1":
        return W.c1(variant)
    if kind == "c2":
        return W.c2(variant)
    if kind == "c2w":
        # config 2's medium, source and receivers on a grid four times as wide
        # and twice as deep (8093 x 2015 cells, 1.2 GB of state for GSLS):
        # a working set far beyond the 256 MB last-level cache
        w = W.c2(variant)
        w.model.MaxRange *= 4
        w.model.Depth *= 2
        w.name = f"c2w-{variant}"
        return w
    raise SystemExit(f"unknown workload {name}")


def describe(w, nr, nz, dt) -> dict:
    return {
        "workload": f"{w.name}: {nr}x{nz} cells, h={w.model.Spacing:g} m, {w.steps} steps of dt={dt:.3g} s, "
                    f"{len(w.ranges)} receivers",
        "grid": [nr, nz],
        "time_steps_per_shot": w.steps,
        "receivers": len(w.ranges),
        "physics": "one scalar SLS (reference form)" if w.model.GslsMechanisms == 0
        else f"GSLS, {w.model.GslsMechanisms} mechanisms, per-layer Qs/Qp",
        "source": "GRF footstep" if w.name.startswith("c2") else "Ricker",
    }


def kernel_name(span: int) -> str:
    """The step kernel of `span` time steps per launch."""
    return "k_step2" if span == 2 else "k_step"


def traffic_from_profiles(workload_name: str, span: int = 1):
    """HBM bytes per step launch measured by rocprofv3 --pmc, if committed."""
    path = os.path.join(ROOT, "profiles", f"{kernel_name(span)}_{workload_name}_traffic.json")
    if not os.path.exists(path):
        return None
    with open(path) as fh:
        return json.load(fh).get("hbm_bytes_per_launch")


def traffic_ratio_from_profiles(workload_name: str, span: int = 1):
    """Measured HBM bytes over algorithmic bytes for a workload whose launches
    differ in size (config 3: the set of running models shrinks), if committed."""
    path = os.path.join(ROOT, "profiles", f"{kernel_name(span)}_{workload_name}_traffic.json")
    if not os.path.exists(path):
        return None
    with open(path) as fh:
        d = json.load(fh)
    return d["hbm_bytes_per_launch"] / d["algorithmic_bytes_per_launch"]


def cpu_baseline(w, gpu_traces, sample_steps: int, threads: int) -> tuple:
    """Time the FP64 oracle on `threads` concurrent copies of the shot for
    `sample_steps` time steps; scale to whole shots.  Also compares the GPU's
    traces over that window with the oracle's."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import numpy as np
    import oracle as O

    O.lib()
    steps = [sample_steps] * threads
    t0 = time.perf_counter()
    outs = O.run_many([w.model] * threads, [w.force] * threads, w.ranges, steps, threads)
    elapsed = time.perf_counter() - t0
    shot_seconds = elapsed * w.steps / sample_steps  # each thread ran one shot's prefix
    traces_per_s = threads * len(w.ranges) / shot_seconds
    nr, nz = w.model.Grid()[1:]
    cell_steps_per_s_core = nr * nz * sample_steps / elapsed
    ref = outs[0]
    # Only receivers the wavefield has reached inside the window: the others
    # hold nothing but the stencil's exponentially small precursor.
    peaks = np.max(np.abs(ref), axis=1)
    errs = []
    for k in range(len(w.ranges)):
        b = ref[k]
        if peaks[k] > 1e-3 * peaks.max():
            errs.append(float(np.linalg.norm(gpu_traces[k, :sample_steps] - b) / np.linalg.norm(b)))
    base = {
        "value": traces_per_s,
        "unit": "traces/s",
        "cores": threads,
        "kind": "port",
        "sample": f"{threads} concurrent shots x {sample_steps} of {w.steps} time steps "
                  f"({elapsed:.1f} s), scaled by time steps; one shot per thread",
        "cell_steps_per_s_per_core": cell_steps_per_s_core,
    }
    acc = {
        "rel_l2_max": max(errs) if errs else None,
        "rel_l2_median": float(np.median(errs)) if errs else None,
        "traces_compared": len(errs),
        "window_steps": sample_steps,
        "against": "FP64 CPU restatement of the reference scheme, same shot",
    }
    return base, acc


C3_GOLDEN = ("c3-X-204", "c3-X-153")


def c3_accuracy(fd) -> dict:
    """The two config-3 soils with the fewest cell-steps, batched on the GPU
    over their whole second, against the FP64 oracle's committed traces
    (tests/golden/full_traces, ten receivers each)."""
    import numpy as np

    import cosim_amd.workloads as W

    gold = [np.load(os.path.join(ROOT, "tests", "golden", "full_traces", f"{n}.npz")) for n in C3_GOLDEN]
    ws = [W.c3(n_models=1, first=int(n.rsplit("-", 1)[1]))[0] for n in C3_GOLDEN]
    shots = [fd.Shot(w.model, w.force, [float(r) for r in z["ranges"]], w.steps) for w, z in zip(ws, gold)]
    errs = []
    for g, z in zip(fd.RunBatch(shots), gold):
        d = int(z["decimate"])
        for k, t in enumerate(g.Traces):
            a, b = t.Vertical[::d], z["traces"][k]
            errs.append(float(np.linalg.norm(a - b) / np.linalg.norm(b)))
    return {"rel_l2_max": max(errs), "rel_l2_median": float(np.median(errs)), "traces_compared": len(errs),
            "window_steps": "whole shots (6,934 and 7,025 steps)",
            "against": "FP64 CPU restatement of the reference scheme, soils 204 and 153 of the batch"}


def bench_c3(args, fd) -> dict:
    """Config 3: batched bank generation over seeded soils.  The batch is set
    up once (every model's arrays and tables in HBM, launches planned and
    tuned) outside the timed region; a bench step advances every chunk of
    models by 1/steps of its longest run, so `steps` steps generate the whole
    batch of banks, the last one ending with the device reduction of every
    trace to its bank spectrum.  Warm-up steps run the first fractions of the
    batch, which is then reset to step 0."""
    import numpy as np

    import cosim_amd.workloads as W

    ws = W.c3(n_models=args.c3_models, seconds=args.c3_seconds)
    spec = W.c3_bank_spec()
    shots = [w.shot() for w in ws]
    grids = [w.model.Grid() for w in ws]
    cells = [g[1] * g[2] for g in grids]
    cell_steps = float(sum(c * w.steps for c, w in zip(cells, ws)))
    log(f"c3: {len(ws)} models, {sum(cells) / 1e6:.1f} M cells, {cell_steps:.3g} cell-steps")
    t0 = time.perf_counter()
    batch = fd.Batch(shots, spec)
    log(f"set-up and launch planning {time.perf_counter() - t0:.1f} s")
    K, Wm = max(args.steps, 1), max(args.warmup, 0)
    for i in range(min(Wm, K)):
        batch.Advance(i + 1, K)
        log(f"warmup step {i + 1}/{Wm}")
    batch.Reset()
    t0 = time.perf_counter()
    for i in range(K):
        batch.Advance(i + 1, K)
        if i + 1 == K:
            spectra, snapped, res = batch.Spectra()
        log(f"step {i + 1}/{K}: {time.perf_counter() - t0:.1f} s")
    wall = time.perf_counter() - t0
    st = fd.BatchStats()
    batch.close()
    bad = [r for r in res if isinstance(r, fd.FdtdError)]
    if bad:
        raise SystemExit(f"{len(bad)} models failed: {bad[0]}")
    if not np.all(np.isfinite(spectra)):
        raise SystemExit("non-finite spectra")
    traces = len(ws) * spec.Count
    span = st["steps_per_launch"]
    launch_s = st["kernel_ms"] / 1e3 / max(st["step_launches"], 1)
    # algorithmic bytes of one launch: 144 B per cell-step (each state array
    # read and written once per step) for every cell-step the launch takes;
    # a two-step launch moves the state only once for its two steps
    # (pass_bytes), which is what lets it beat this figure
    per_launch = st["algorithmic_bytes"] / max(st["step_launches"], 1)
    pass_per_launch = st["pass_bytes"] / max(st["step_launches"], 1)
    achieved = per_launch / launch_s / 1e9
    lockstep = st["budget_mb"] <= 0  # the library's plan: every model advanced together
    ratio = traffic_ratio_from_profiles("c3" if lockstep else "c3_chunked", span)
    kern = kernel_name(span)
    what = "one step per launch" if span == 1 else "two steps per launch, the state read and written once per launch"
    out = {
        "metric": METRIC,
        "value": traces / wall,
        "unit": "traces/s",
        "n_gpus": 1,
        "steps": K,
        "warmup": Wm,
        "ms_per_step": wall * 1e3 / K,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {
            "workload": f"c3: batched bank generation, {len(ws)} seeded soils (Philox seed 0), 600-2048 x 300-1024 "
                        f"cells, {args.c3_seconds:g} s each, {spec.Count} receivers, banks of "
                        f"{spec.Samples // 2 + 1} bins; a step advances every model by 1/{K} of its run",
            "models": len(ws), "cells": sum(cells), "cell_steps": cell_steps,
            "physics": "GSLS, 3 mechanisms, per-layer Qs, Qp = 2 Qs",
        },
        "roofline": {
            "bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
            "frac": achieved / HBM_PEAK_GBS, "traffic": None if ratio is None else ratio * per_launch,
            "kernel": f"{kern} over every model of a chunk still running ({what}, graph blocks of 100 steps)"
            if not lockstep else f"{kern} over every model still running ({what})",
            "steps_per_launch": span, "tile_rows": st["tile_rows"], "tiles_per_workgroup": st["waves"],
            "chunk_budget_mb": st["budget_mb"], "algorithmic_bytes": st["algorithmic_bytes"],
            "step_loop_ms": st["kernel_ms"], "step_launches": st["step_launches"],
            "avg_launch_us": launch_s * 1e6, "algorithmic_bytes_per_launch": per_launch,
            # the state read and written once per launch, over the same time
            "state_bytes_per_launch": pass_per_launch, "state_pass_gbs": pass_per_launch / launch_s / 1e9,
        },
        "cells_per_s": cell_steps / wall,
    }
    if not args.no_cpu:
        sys.path.insert(0, os.path.join(ROOT, "oracle"))
        import oracle as O

        threads = args.cpu_threads or min(16, os.cpu_count() or 1)
        sample = ws[:threads]
        steps = [args.cpu_sample_steps] * len(sample)
        t1 = time.perf_counter()
        O.run_many([w.model for w in sample], [w.force for w in sample], [1.0, 25.0, 50.0], steps, threads)
        el = time.perf_counter() - t1
        rate = sum(c * args.cpu_sample_steps for c in cells[:threads]) / el  # cell-steps/s, all threads
        out["cpu_baseline"] = {
            "value": traces / (cell_steps / rate), "unit": "traces/s", "cores": threads, "kind": "port",
            "sample": f"the first {len(sample)} soils, {args.cpu_sample_steps} steps each, one per thread ({el:.1f} s); "
                      f"scaled to the batch by cell-steps",
            "cell_steps_per_s": rate,
        }
        out["speedup_vs_cpu"] = out["value"] / out["cpu_baseline"]["value"]
    else:
        out["cpu_baseline"] = None
    if os.path.exists(os.path.join(ROOT, "tests", "golden", "full_traces", f"{C3_GOLDEN[0]}.npz")):
        out["accuracy"] = c3_accuracy(fd)
    return out


FP64_PEAK_TFLOPS = 78.6


def fk_workload(fd):
    """The frequency-wavenumber bank of the config-2 layered medium: geobank's
    frequency grid (2000 Hz, 4096 samples: 0.488 Hz bins, cmd/geobank/main.go)
    up to 80 Hz (the layer-phase guard allows 83 Hz for the 4 m layer), 50
    ranges 1-50 m, Kjartansson constant Q from each layer's Qs."""
    fd_ = load_package().fdtd
    vs, rho, qs, th = [150.0, 250.0, 400.0], [1700.0, 1850.0, 2000.0], [20.0, 40.0, 80.0], [1.5, 4.0, 0.0]
    stack = [fd_.Layer(th[i], 2 * vs[i], vs[i], rho[i], qs[i]) for i in range(3)]
    medium = load_package().fk.Medium(Stack=stack)
    bin_hz = 2000.0 / 4096
    freqs = [b * bin_hz for b in range(1, int(80.0 / bin_hz) + 1)]
    ranges = [float(r) for r in range(1, 51)]
    return medium, freqs, ranges


def bench_fk(args, fd, standalone=True):
    """f-k bank generation (the reference's bank filler, cmd/geobank over
    fk.VerticalDisplacementMulti): every frequency of one bank in one set of
    launches; a step is one bank."""
    import numpy as np

    fk = load_package().fk
    medium, freqs, ranges = fk_workload(fd)
    for i in range(args.warmup):
        medium.VerticalDisplacementMulti(ranges, freqs)
    t0 = time.perf_counter()
    kernel_ms = 0.0
    for i in range(args.steps):
        g = medium.VerticalDisplacementMulti(ranges, freqs)
        kernel_ms += fk.Stats()["kernel_ms"]
    wall = time.perf_counter() - t0
    st = fk.Stats()
    if not np.all(np.isfinite(g)):
        raise SystemExit("non-finite f-k bank")
    # FP64 VALU work per bank, counted by rocprofv3 (executed, not algorithmic:
    # the f-k solve has no closed-form operation count to quote).
    prof = os.path.join(ROOT, "profiles", "fk_bank_c2_fp64.json")
    flops = None
    if os.path.exists(prof):
        with open(prof) as fh:
            flops = json.load(fh)
    log(f"fk: {len(freqs)} bins x {len(ranges)} ranges, {st['samples']} samples per bank, "
        f"{args.steps} banks in {wall:.3f} s ({kernel_ms:.1f} ms of kernels)")
    res = {
        "value": len(ranges) * args.steps / wall,
        "unit": "traces/s",
        "ms_per_step": wall * 1e3 / args.steps,
        "config": {
            "workload": "fk: one Green's-function bank of the config-2 layered medium by wavenumber integration; "
                        f"{len(freqs)} bins of 0.488 Hz to 80 Hz, {len(ranges)} ranges 1-50 m",
            "bins": len(freqs), "ranges": len(ranges), "samples_per_bank": st["samples"],
            "hankel_terms_per_bank": st["weights"],
        },
        "kernel_ms_per_bank": kernel_ms / args.steps,
        "solve_ms_per_bank": st["solve_ms"],
        "hankel_ms_per_bank": st["hankel_ms"],
        "samples_per_s": st["samples"] * args.steps / wall,
        "dtype": "f64",
        "roofline": None,
    }
    if flops:
        t = (st["solve_ms"] + st["hankel_ms"]) / 1e3
        achieved = (flops["solve_fp64_flop_per_bank"] + flops["hankel_fp64_flop_per_bank"]) / t / 1e12
        res["roofline"] = {
            "bound": "fp64-valu", "achieved": achieved, "peak": FP64_PEAK_TFLOPS, "unit": "TFLOP/s",
            "frac": achieved / FP64_PEAK_TFLOPS, "traffic": None,
            "kernel": "k_fk_solve + k_fk_hankel (last bank's HIP-event times)",
            "flop_count": "executed FP64 VALU FLOP from rocprofv3 counters (profiles/fk_bank_c2_fp64.json)",
        }
    if not args.no_cpu:
        sys.path.insert(0, os.path.join(ROOT, "oracle"))
        import oracle as O
        from concurrent.futures import ThreadPoolExecutor

        threads = args.cpu_threads or min(16, os.cpu_count() or 1)
        sample = freqs[::10]
        n_all = sum(medium.GridFor(ranges, f).Samples for f in freqs)
        n_sample = sum(medium.GridFor(ranges, f).Samples for f in sample)
        t1 = time.perf_counter()
        with ThreadPoolExecutor(threads) as ex:
            list(ex.map(lambda f: O.fk_vertical_displacement_multi(medium, ranges, f), sample))
        el = time.perf_counter() - t1
        bank_s = el * n_all / n_sample
        res["cpu_baseline"] = {
            "value": len(ranges) / bank_s, "unit": "traces/s", "cores": threads, "kind": "port",
            "sample": f"every 10th bin ({len(sample)} bins, {n_sample} samples) over {threads} threads in "
                      f"{el:.1f} s, scaled to the bank by wavenumber samples",
        }
        res["speedup_vs_cpu"] = res["value"] / res["cpu_baseline"]["value"]
    if standalone:
        out = {"metric": METRIC, "value": res["value"], "unit": "traces/s", "n_gpus": 1, "steps": args.steps,
               "warmup": args.warmup, "ms_per_step": res["ms_per_step"], "higher_is_better": True,
               "scaling": "weak", "vs_baseline": None, "dtype": "f64", "data": "synthetic",
               "config": res["config"], **{k: v for k, v in res.items() if k not in (
                   "value", "unit", "ms_per_step", "config", "dtype")}}
        print(json.dumps(out), flush=True)
        return 0
    return res


def bench_bank_lookup(args, fd):
    """Batched bank lookup (bank.Bank.Response at every bin for many ranges):
    the reference's benchmark bank (bank_test.go:358-382: 270 ranges 1-40 m,
    4096 samples, so 2049 bins), 8192 query ranges per call; a step is one
    call.  `value` is responses per second of kernel time with the result in
    HBM; the PCIe-inclusive rate (FP64 results copied to the host) beside it."""
    import numpy as np

    bkm = load_package().bank
    count, samples = 270, 4096
    bins = samples // 2 + 1
    i = np.arange(count)[:, None]
    k = np.arange(bins)[None, :]
    z = (1.0 / (i + 1)) + 1j * (k * 1e-4)
    h = fd.BankHeader(Solver="bench", Medium=[fd.Layer(0, 500, 200, 1700, 25)], SampleRateHz=2000, Samples=samples,
                      MinM=1, MaxM=40, Count=count, Component="x", Units="m/N")
    dev = bkm.Bank(h, z)
    nq = 8192
    r = np.random.default_rng(0).uniform(1, 40, nq)
    for _ in range(args.warmup):
        dev.Spectra(r)
    kernel_ms, t0 = 0.0, time.perf_counter()
    for _ in range(args.steps):
        dev.Spectra(r)
        kernel_ms += dev.Stats()["kernel_ms"]
    wall = time.perf_counter() - t0
    st = dev.Stats()
    dev.close()
    per_call = nq * bins
    # Unique bytes: both FP64 tables read once and the FP64 complex output
    # written once (st["bytes"] counts 48 B per value, with no row reuse).
    unique = count * bins * 16 + per_call * 16
    achieved = unique / (st["kernel_ms"] / 1e3) / 1e9
    traffic = None
    prof = os.path.join(ROOT, "profiles", "k_bank_spectra_traffic.json")
    if os.path.exists(prof):
        with open(prof) as fh:
            traffic = json.load(fh)["hbm_bytes_per_launch"]
    res = {
        "value": per_call * args.steps / (kernel_ms / 1e3), "unit": "responses/s",
        "pcie_inclusive_responses_per_s": per_call * args.steps / wall,
        "config": {"workload": f"bank lookup: {count} ranges x {bins} bins bank, {nq} query ranges x every bin "
                               "per call", "responses_per_call": per_call},
        "kernel_ms_per_call": kernel_ms / args.steps,
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS, "traffic": traffic, "kernel": "k_bank_spectra",
                     "algorithmic_bytes_per_launch": unique,
                     "bytes_rule": "log|G| and phase tables (FP64) read once, FP64 complex output written once",
                     "note": "measured traffic (tables re-read from L2 misses, output writes) runs at about a "
                             "plain copy's rate for this size; 89 VALU instructions per value"},
        "dtype": "f64",
    }
    if not args.no_cpu:
        sys.path.insert(0, os.path.join(ROOT, "oracle"))
        import oracle as O
        from concurrent.futures import ThreadPoolExecutor

        threads = args.cpu_threads or min(16, os.cpu_count() or 1)
        data = np.stack([z.real, z.imag], axis=-1).astype(np.float32)
        ob = O.OracleBank(1, 40, data)
        sample = np.array_split(r, threads)
        reps = 5
        t1 = time.perf_counter()
        with ThreadPoolExecutor(threads) as ex:
            for _ in range(reps):
                list(ex.map(ob.Spectra, sample))
        el = time.perf_counter() - t1
        res["cpu_baseline"] = {"value": reps * nq * bins / el, "unit": "responses/s", "cores": threads,
                               "kind": "port", "sample": f"the {nq} query ranges x every bin, {reps} times, over "
                                                         f"{threads} threads ({el:.2f} s)"}
        res["speedup_vs_cpu"] = res["value"] / res["cpu_baseline"]["value"]
    return res


def bench_stream(args, fd):
    """Streaming convolution (stream.Convolver) for a 64-receiver array, each
    with its own 1 s response at 2 kHz (2000 taps): the latency of one
    20-sample chunk (10 ms of signal, the reference's default) for every
    receiver, and the offline rate of ProcessAll over 60 s of input."""
    import numpy as np

    stm = load_package().stream
    R, L, chunk, fs = 64, 2000, 20, 2000.0
    i = np.arange(L)
    h = np.stack([np.exp(-i / (200 + 10 * r)) * np.sin(i * (0.05 + 0.003 * r) + r) for r in range(R)])
    x = np.random.default_rng(1).standard_normal((R, int(60 * fs)))
    c = stm.NewConvolver(h, chunk)
    for k in range(20):
        c.Process(x[:, k * chunk:(k + 1) * chunk])
    n_chunks = 500
    t0 = time.perf_counter()
    kern = 0.0
    for k in range(n_chunks):
        c.Process(x[:, k * chunk:(k + 1) * chunk])
        kern += c.KernelMs()
    wall = time.perf_counter() - t0
    c.Reset()
    t1 = time.perf_counter()
    c.ProcessAll(x)
    wall_all = time.perf_counter() - t1
    k_all = c.KernelMs()
    c.close()
    macs = float(R) * x.shape[1] * L
    achieved = 2 * macs / (k_all / 1e3) / 1e12
    res = {
        "value": R * x.shape[1] / (k_all / 1e3), "unit": "output samples/s (ProcessAll, kernel time)",
        "chunk_latency_us": wall / n_chunks * 1e6, "chunk_kernel_us": kern / n_chunks * 1e3,
        "real_time_factor": (chunk / fs) / (wall / n_chunks),
        "process_all_wall_s": wall_all,
        "config": {"workload": f"stream: {R} receivers, {L}-tap responses, chunks of {chunk} samples at 2 kHz; "
                               "ProcessAll over 60 s"},
        "roofline": {"bound": "fp64-valu", "achieved": achieved, "peak": FP64_PEAK_TFLOPS, "unit": "TFLOP/s",
                     "frac": achieved / FP64_PEAK_TFLOPS, "traffic": None, "kernel": "k_fir (ProcessAll)",
                     "flop_rule": "2 x taps per output sample"},
        "dtype": "f64",
    }
    if not args.no_cpu:
        sys.path.insert(0, os.path.join(ROOT, "oracle"))
        import stream_oracle as SO
        from concurrent.futures import ThreadPoolExecutor

        threads = args.cpu_threads or min(16, os.cpu_count() or 1)
        n = 4000
        t2 = time.perf_counter()
        with ThreadPoolExecutor(threads) as ex:
            list(ex.map(lambda r: SO.Convolver(h[r], chunk).ProcessAll(x[r, :n]), range(threads)))
        el = time.perf_counter() - t2
        res["cpu_baseline"] = {"value": threads * n / el, "unit": "output samples/s", "cores": threads,
                               "kind": "port", "sample": f"{threads} receivers x {n} samples in {chunk}-sample "
                                                         f"chunks, one receiver per thread ({el:.2f} s; numpy "
                                                         "FFTs, Python per chunk)"}
        res["speedup_vs_cpu"] = res["value"] / res["cpu_baseline"]["value"]
    return res


def bench_engine(args, fd):
    """Runtime synthesis (sensing.Engine over a bank) heard by a 64-geophone
    line: a 75 kg walker at 1.3 m/s passing 4 m from the line's centre over
    16 m of path, chunks of 20 samples at 2 kHz, impulse responses from an
    f-k bank of the loam medium (1-25 m, 4096-sample transform, so 3072-tap
    responses).  Reports the latency of one chunk for every receiver, the
    mix kernel's share, and the offline rate over the whole walk."""
    import numpy as np

    pkg = load_package()
    fk, bk, se, wl = pkg.fk, pkg.bank, pkg.sensing, pkg.workloads
    fs, chunk, R = 2000.0, 20, 64
    layer = fd.Layer(0, 500, 200, 1700, 25)
    limit = bk.RangeNyquist(0.87 * 200, 80) / 2
    count = int(math.ceil((25 - 1) / limit)) + 1
    h = fd.BankHeader(Solver="fk", Medium=[layer], SampleRateHz=fs, Samples=4096, MinM=1, MaxM=25, Count=count,
                      Component="vertical surface displacement per unit vertical surface point force")
    z = fk.FillBank(fk.Medium(Stack=[layer], DefaultQ=30), h, 80)
    b = bk.Bank(h, z)
    walk = se.Walk(Stance=wl.Stance(Mass=75), Speed=1.3, StartX=-8, StartY=4, Width=0.12, Until=16 / 1.3)
    rec = [(-15.75 + 0.5 * q, 0.0) for q in range(R)]
    geophone = lambda f: 28.8 * (1j * f / 4.5) ** 2 / (1 + 1j * 0.7 * (f / 4.5) - (f / 4.5) ** 2)  # noqa: E731
    total = int(math.ceil((walk.Until + 1.6) * fs / chunk))

    def one_walk():
        e = se.NewEngineWith(walk, chunk, se.FromBank(b), sample_rate=fs, walk_duration=walk.Until,
                             geometry_range=4, rayleigh_velocity=186.0, sensor=geophone, receivers=rec)
        lat = []
        for _ in range(total):
            t0 = time.perf_counter()
            e.Next()
            lat.append(time.perf_counter() - t0)
        st = e.Stats()
        info = (e.ImpulseLength(), e.Responses())
        e.close()
        return np.array(lat), st, info

    one_walk()
    t0 = time.perf_counter()
    lat, st, (ir_len, responses) = one_walk()
    wall = time.perf_counter() - t0
    samples = R * total * chunk
    achieved = 2 * st["macs"] / (st["mix_ms_total"] / 1e3) / 1e12
    busy = lat
    res = {
        "value": samples / float(np.sum(lat)), "unit": "output samples/s (Next over the whole walk, wall clock)",
        "chunk_latency_us": {"p50": float(np.percentile(busy, 50) * 1e6), "p99": float(np.percentile(busy, 99) * 1e6),
                             "max": float(np.max(busy) * 1e6)},
        "mix_kernel_us_per_chunk": st["mix_ms_total"] / total * 1e3,
        "real_time_factor": (chunk / fs) / float(np.percentile(busy, 99)),
        "walk_wall_s": wall, "responses_built": responses, "build_ms": st["build_ms"],
        "config": {"workload": f"engine: a walk past 4 m from a line of {R} geophones (0.5 m apart), 16 m of path "
                               f"at 1.3 m/s, chunks of {chunk} samples at 2 kHz; impulse responses of {ir_len} taps "
                               f"from an f-k bank of {count} ranges 1-25 m"},
        "roofline": {"bound": "fp64-valu", "achieved": achieved, "peak": FP64_PEAK_TFLOPS, "unit": "TFLOP/s",
                     "frac": achieved / FP64_PEAK_TFLOPS, "traffic": None, "kernel": "k_engine_mix",
                     "flop_rule": "2 x the multiply-adds where force and response are both defined"},
        "dtype": "f64",
    }
    if not args.no_cpu:
        sys.path.insert(0, os.path.join(ROOT, "oracle"))
        import oracle as O
        import sensing_oracle as SO

        # sensing.Reference on the oracle for a bounded sample of receivers:
        # responses built first (untimed), then the per-contact whole-record
        # convolutions and the sum, one receiver at a time on one core.
        data = np.stack([z.real, z.imag], axis=-1).astype(np.float32)
        ob = O.OracleBank(1, 25, data)
        n = total * chunk
        ow = SO.WalkSpec(duration=walk.Stance.Duration, speed=walk.Speed, start_x=walk.StartX,
                         start_y=walk.StartY, width=walk.Width, until=walk.Until)
        contacts = SO.walk_contacts(ow, 0, n / fs)
        foot = se.Footfall(walk.Stance, 0.0)
        forces = [lambda t, c=c: foot.ForceAt(t - c.start)[2] for c in contacts]
        sens = geophone(np.arange(ob.bins) * fs / 4096)
        sample = [0, R // 2]
        cache = {}
        for q in sample:
            for c in contacts:
                rq = SO.quantized_range(math.hypot(rec[q][0] - c.x, rec[q][1] - c.y))
                if rq not in cache:
                    cache[rq] = O.bank_impulse(ob, fs, [rq], ir_len, fs, sens)[0]
        t2 = time.perf_counter()
        for q in sample:
            SO.reference(contacts, forces, n, fs, lambda rq: cache[rq], rec[q][0], rec[q][1])
        el = time.perf_counter() - t2
        res["cpu_baseline"] = {"value": len(sample) * n / el, "unit": "output samples/s", "cores": 1,
                               "kind": "port", "sample": f"{len(sample)} receivers x {n} samples, the whole walk "
                                                         f"({el:.2f} s): sensing.Reference restated (per contact a "
                                                         "whole-record numpy convolution), responses prebuilt"}
        res["speedup_vs_cpu"] = res["value"] / res["cpu_baseline"]["value"]
    b.close()
    # The same walk and line through sensing.NewEngine's analytic half-space
    # (loam): both response components built on the device (32,768-point
    # transforms at this irLen), each voice convolving its vertical force and
    # its force along the line to every receiver.
    t0 = time.perf_counter()
    e = se.NewEngine(walk, chunk, se.Loam(), sample_rate=fs, walk_duration=walk.Until, geometry_range=4,
                     sensor=geophone, receivers=rec)
    create_s = time.perf_counter() - t0
    alat = []
    for _ in range(total):
        t1 = time.perf_counter()
        e.Next()
        alat.append(time.perf_counter() - t1)
    ast = e.Stats()
    a_ir = e.ImpulseLength()
    e.close()
    alat = np.array(alat)
    res["analytic"] = {
        "workload": f"the same walk and line, sensing.NewEngine over the loam half-space: {a_ir}-tap vertical and "
                    "radial-force responses",
        "value": R * total * chunk / float(np.sum(alat)), "unit": "output samples/s",
        "create_s": create_s, "build_ms": ast["build_ms"], "responses_built": ast["built"],
        "chunk_latency_us": {"p50": float(np.percentile(alat, 50) * 1e6),
                             "p99": float(np.percentile(alat, 99) * 1e6), "max": float(np.max(alat) * 1e6)},
        "mix_kernel_us_per_chunk": ast["mix_ms_total"] / total * 1e3,
        "mix_tflops": 2 * ast["macs"] / (ast["mix_ms_total"] / 1e3) / 1e12,
    }
    return res


MFMA_BF16_PEAK_TFLOPS = 2500.0


def bench_surrogate(args, fd) -> int:
    """Configs 4 and 5: train the surrogate on banks the batched FDTD generates
    for `--sur-soils` layered soils, then serve `--sur-queries` held-out
    (soil, range) queries behind the 1 % FDTD error gate."""
    import numpy as np

    import cosim_amd.surrogate as S

    if args.sur_speed_only:
        return surrogate_speed(args, S)
    R = S.RANGES[2]
    n_held = -(-args.sur_queries // R)
    log(f"surrogate: FDTD banks for {args.sur_soils} training soils")
    prog = lambda d, n, s: log(f"  FDTD banks {d}/{n} soils, {s:.0f} s")
    train = S.solve_set(S.soil_set(0, 0, args.sur_soils), progress=prog)
    fdtd_rate = args.sur_soils * R / train.fdtd_seconds
    qc = S.soil_qc(train)
    x, t, _, _, _ = S.dataset(S.subset(train, qc))
    limit = S.norm_bound(t)
    log(f"  {train.fdtd_seconds:.1f} s ({fdtd_rate:.0f} traces/s); {qc.sum()} of {len(qc)} soils pass the bank QC; "
        f"training {args.sur_steps} steps")
    net = S.Network(max_batch=8192, seed=1)
    losses = net.Train(x, t, steps=args.sur_steps, batch=4096, lr=args.sur_lr, lr_min=args.sur_lr / 50, seed=2)
    tr = net.Stats()
    log(f"  loss {losses[:50].mean():.4f} -> {losses[-50:].mean():.4f} in {tr['train_ms'] / 1e3:.1f} s of device time")
    log(f"held-out: FDTD truth for {n_held} soils (evaluation only)")
    held = S.solve_set(S.soil_set(0, 1_000_000, n_held), progress=prog)
    xh, _, sch, dlh, Gh = S.dataset(held)
    xh, sch, dlh, Gh = xh[:args.sur_queries], sch[:args.sur_queries], dlh[:args.sur_queries], Gh[:args.sur_queries]
    # inference rate: device time of the predict kernels with queries in HBM
    for _ in range(2):
        pred = net.Predict(xh, sch, dlh)
    inf = net.Stats()
    inf_rate = inf["forward_queries"] / (inf["forward_ms"] / 1e3)
    err = S.rel_l2(pred, Gh)
    # the served job: every query answered, 1 % of soils re-solved by the FDTD
    t0 = time.perf_counter()
    served, rep = S.serve(net, held.params, held.ranges, held.soils, gate_fraction=0.01, bound=0.02, seed=3,
                          norm_limit=limit)
    wall = time.perf_counter() - t0
    served = served.reshape(-1, S.BINS)[:args.sur_queries]
    err_served = S.rel_l2(served, Gh)
    held_qc = S.soil_qc(held)
    qmask = np.repeat(held_qc, R)[:args.sur_queries]
    err_qc = err[qmask]
    fb = set(rep.failed_soils)
    achieved = tr["flops_per_query"] * inf_rate / 1e12
    out = {
        "metric": METRIC,
        "value": args.sur_queries / wall,
        "unit": "traces/s",
        "n_gpus": 1,
        "steps": 1,
        "warmup": 1,
        "ms_per_step": wall * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "bf16",
        "data": "synthetic",
        "config": {
            "workload": f"surrogate: trained on FDTD banks of {args.sur_soils} layered soils x {R} ranges "
                        f"({args.sur_soils * R} spectra, 512 bins 0-250 Hz), serving {args.sur_queries} held-out "
                        f"queries, BF16 inference batch 8192, 1% of soils re-solved by the batched FP32 FDTD, "
                        f"fallback above 2% relative L2",
            "train_soils": args.sur_soils, "queries": args.sur_queries, "train_steps": args.sur_steps,
            "train_batch": 4096,
        },
        "roofline": {
            "bound": "mfma", "achieved": achieved, "peak": MFMA_BF16_PEAK_TFLOPS, "unit": "TFLOP/s",
            "frac": achieved / MFMA_BF16_PEAK_TFLOPS, "traffic": None,
            "kernel": "k_encoder_fused: the whole network and the de-normalization in one launch per batch (BF16 MFMA, FP32 accumulation)",
            "flops_per_query": tr["flops_per_query"],
        },
        "surrogate": {
            "inference_traces_per_s": inf_rate,
            "fdtd_traces_per_s": fdtd_rate,
            "inference_vs_fdtd_per_trace": inf_rate / fdtd_rate,
            "served_vs_fdtd_per_trace": (args.sur_queries / wall) / fdtd_rate,
            "rel_l2_p50": float(np.median(err)), "rel_l2_p99": float(np.quantile(err, 0.99)),
            "fraction_within_2pct": float(np.mean(err <= 0.02)),
            "served_rel_l2_p50": float(np.median(err_served)), "served_rel_l2_p99": float(np.quantile(err_served, 0.99)),
            "gate_sampled_soils": len(rep.sampled_soils), "gate_fallback_soils": len(fb),
            "magnitude_guard_soils": len(rep.guarded_soils), "train_soils_passing_qc": int(qc.sum()),
            "held_out_soils_passing_qc": int(held_qc.sum()),
            "rel_l2_p50_qc_soils": float(np.median(err_qc)), "rel_l2_p99_qc_soils": float(np.quantile(err_qc, 0.99)),
            "gate_seconds": rep.gate_seconds, "surrogate_seconds": rep.surrogate_seconds,
            "train_loss_first": float(losses[:50].mean()), "train_loss_last": float(losses[-50:].mean()),
            "train_ms_per_step": tr["train_ms"] / max(tr["train_steps"], 1),
            "train_tflops": 3 * tr["flops_per_query"] * tr["train_queries"] / (tr["train_ms"] / 1e3) / 1e12,
        },
        "cpu_baseline": None,
    }
    if args.sur_save:
        S.save_weights(args.sur_save, net, limit, {
            "trained_by": "bench.py --workload surrogate (this project's BF16 MFMA forward/backward kernels and Adam)",
            "command": " ".join(sys.argv),
            "train_soils": args.sur_soils, "train_soils_passing_qc": int(qc.sum()), "train_steps": args.sur_steps,
            "train_batch": 4096, "lr": args.sur_lr, "held_out_first": 1_000_000, "held_out_soils": n_held,
            "held_out_queries": args.sur_queries,
            "held_out_rel_l2_p50": float(np.median(err)), "held_out_rel_l2_p99": float(np.quantile(err, 0.99)),
            "held_out_fraction_within_2pct": float(np.mean(err <= 0.02)),
        })
        log(f"weights saved to {args.sur_save}")
    if not args.no_cpu:
        # the FP64 CPU oracle's forward pass on the trained weights, on a
        # sample of the same queries (numpy matmuls on the host's cores)
        sys.path.insert(0, os.path.join(ROOT, "oracle"))
        import surrogate_oracle as SO

        layers = SO.layers_from(None, net)
        p = net.Params().astype(np.float64)
        sample = xh[:2048].astype(np.float64)
        t1 = time.perf_counter()
        SO.forward(p, layers, sample)
        el = time.perf_counter() - t1
        out["cpu_baseline"] = {
            "value": len(sample) / el, "unit": "traces/s", "cores": os.cpu_count() or 1, "kind": "port",
            "sample": f"FP64 forward of the trained network on {len(sample)} of the queries ({el:.2f} s, numpy BLAS "
                      f"on every host core)",
        }
        out["speedup_vs_cpu"] = inf_rate / out["cpu_baseline"]["value"]
    net.close()
    print(json.dumps(out), flush=True)
    return 0


def bench_served(args, fd) -> dict:
    """Config 5: --sur-queries held-out (soil, range) queries of the layered
    family served by the network with the committed weights, at inference
    batch 8,192, behind the 1 % FDTD error gate (escalating the request to the
    batched FDTD when a sampled soil exceeds 2 %).  Timed from the queries to
    the served spectra, FDTD re-solves included."""
    import numpy as np

    import cosim_amd.surrogate as S

    R = S.RANGES[2]
    n_soils = -(-args.sur_queries // R)
    net, limit, meta = S.load_weights(args.sur_weights)
    held = S.soil_set(0, meta.get("held_out_first", 1_000_000), n_soils)
    ranges = S.receiver_ranges(held.soils)
    x = S.features(np.repeat(held.params, R, axis=0), ranges.reshape(-1))
    sc, dl = S.prior(np.repeat(held.params, R, axis=0), ranges.reshape(-1))
    # the inference rate over the stated query count (whole batches of 8,192);
    # the request below serves every range of every soil
    q_inf = min(args.sur_queries, len(x))
    for _ in range(2):
        net.Predict(x[:q_inf], sc[:q_inf], dl[:q_inf])
    inf = net.Stats()
    inf_rate = inf["forward_queries"] / (inf["forward_ms"] / 1e3)
    t0 = time.perf_counter()
    served, rep = S.serve(net, held.params, ranges, held.soils, gate_fraction=0.01, bound=0.02, seed=3,
                          norm_limit=limit)
    wall = time.perf_counter() - t0
    net.close()
    queries = n_soils * R
    achieved = inf["flops_per_query"] * inf_rate / 1e12
    known = sorted(set(rep.sampled_soils) | set(rep.fdtd_soils))
    out = {
        "workload": f"config 5: {queries} held-out queries ({n_soils} soils x {R} ranges), BF16 inference batch "
                    f"8192, 1% of soils re-solved by the batched FP32 FDTD, the request escalated to the FDTD when "
                    f"a sampled soil exceeds 2% relative L2",
        "value": queries / wall, "unit": "traces/s", "seconds": wall,
        "inference_traces_per_s": inf_rate, "inference_queries": q_inf,
        "roofline": {"bound": "mfma", "achieved": achieved, "peak": MFMA_BF16_PEAK_TFLOPS, "unit": "TFLOP/s",
                     "frac": achieved / MFMA_BF16_PEAK_TFLOPS, "traffic": None,
                     "kernel": "k_encoder_fused: the whole network and the de-normalization in one launch per batch (BF16 MFMA, FP32 accumulation)",
                     "flops_per_query": inf["flops_per_query"]},
        "gate": {"sampled_soils": len(rep.sampled_soils), "failed_soils": len(rep.failed_soils),
                 "escalated": rep.escalated, "fdtd_soils": len(rep.fdtd_soils),
                 "magnitude_guard_soils": len(rep.guarded_soils),
                 "sampled_rel_l2_p50": float(np.median(rep.sampled_errors)) if len(rep.sampled_errors) else None,
                 "sampled_rel_l2_max": float(np.max(rep.sampled_errors)) if len(rep.sampled_errors) else None,
                 "vouch_bound_95": rep.vouch_bound() if not rep.escalated else None,
                 "surrogate_seconds": rep.surrogate_seconds, "fdtd_seconds": rep.gate_seconds},
        "served_from_network": queries - R * len(rep.fdtd_soils),
        "weights": {k: meta[k] for k in ("train_soils", "train_steps", "held_out_rel_l2_p50",
                                        "held_out_rel_l2_p99", "held_out_fraction_within_2pct") if k in meta},
    }
    if rep.fdtd_soils:
        # the batched FDTD's rate on these soils, from the gate's own solves
        out["fdtd_traces_per_s_same_soils"] = R * len(rep.fdtd_soils) / max(rep.gate_seconds, 1e-9)
        out["served_vs_fdtd_same_soils"] = out["value"] / out["fdtd_traces_per_s_same_soils"]
    if known:
        # served answers against the FDTD where the run solved the soil
        truth_soils = S.solve([held.soils[i] for i in known[:min(len(known), 16)]])[0]
        e = S.rel_l2(served[known[:len(truth_soils)]].reshape(-1, S.BINS),
                     S.band_limit(truth_soils).reshape(-1, S.BINS))
        out["served_rel_l2_vs_fdtd"] = {"p50": float(np.median(e)), "p99": float(np.quantile(e, 0.99)),
                                        "queries": int(len(e)),
                                        "over": "the first 16 soils the gate solved (sampled or escalated)"}
    return out


def surrogate_speed(args, S) -> int:
    """The surrogate's kernels alone, on synthetic inputs and targets (their
    speed does not depend on the weights): inference over --sur-queries at
    batch 8192, then 50 training steps at batch 4096."""
    import numpy as np

    rng = np.random.default_rng(0)
    net = S.Network(max_batch=8192, seed=1)
    Q = args.sur_queries
    x = rng.uniform(-1, 1, (Q, S.FEATURES)).astype(np.float32)
    sc, dl = rng.uniform(1e-9, 1e-8, Q), rng.uniform(0, 0.4, Q)
    for _ in range(3):
        net.Predict(x, sc, dl)
    inf = net.Stats()
    rate = inf["forward_queries"] / (inf["forward_ms"] / 1e3)
    t = rng.normal(0, 1, (16384, 2 * S.BINS)).astype(np.float32)
    net.Train(x[:16384], t, steps=50, batch=4096, lr=1e-4, lr_min=1e-5, seed=1)
    tr = net.Stats()
    net.close()
    achieved = inf["flops_per_query"] * rate / 1e12
    print(json.dumps({
        "probe": "surrogate kernels", "queries": Q, "inference_traces_per_s": rate,
        "roofline": {"bound": "mfma", "achieved": achieved, "peak": MFMA_BF16_PEAK_TFLOPS, "unit": "TFLOP/s",
                     "frac": achieved / MFMA_BF16_PEAK_TFLOPS, "flops_per_query": inf["flops_per_query"]},
        "train_ms_per_step": tr["train_ms"] / tr["train_steps"],
        "train_tflops": 3 * inf["flops_per_query"] * tr["train_queries"] / (tr["train_ms"] / 1e3) / 1e12,
    }), flush=True)
    return 0


def ceiling(args) -> int:
    """Bandwidth ceilings: device copies over sizes below and above the 256 MB
    last-level cache, and the step kernel's rate as its state outgrows it."""
    import torch

    for mib in (64, 128, 256, 512, 1024, 4096):
        n = mib * (1 << 20) // 4
        x = torch.ones(n, dtype=torch.float32, device="cuda")
        y = torch.empty_like(x)
        for _ in range(3):
            y.copy_(x)
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        reps = 10
        e0.record()
        for _ in range(reps):
            y.copy_(x)
        e1.record()
        torch.cuda.synchronize()
        us = e0.elapsed_time(e1) * 1e3 / reps
        print(json.dumps({"probe": "copy", "mib": mib, "us": us, "gbs": 2 * 4 * n / us / 1e3}), flush=True)
        del x, y
        torch.cuda.empty_cache()
    pkg = load_package()
    fd = pkg.fdtd
    import cosim_amd.workloads as W

    steps = 1000
    for fr, fz in ((0.5, 0.5), (1, 1), (2, 1), (2, 2), (4, 2)):
        w = W.c2("X", steps)
        w.model.MaxRange *= fr
        w.model.Depth *= fz
        dt, nr, nz = w.model.Grid()
        sim = fd.New(w.model)
        sim.Drive(w.force[:steps])
        sim.SetReceivers([sim.Column(1.0)[0]], steps)
        rows = {}
        for th in (1, 2, 4, 8):
            sim.SetTileRows(th)
            sim.Reset()
            sim.Advance(200, True)
            sim.SyncTiming()
            sim.Reset()
            sim.Advance(steps, True)
            ms, n = sim.SyncTiming()
            rows[th] = sim.StepBytes() / (ms / 1e3 / n) / 1e9
        print(json.dumps({"probe": "k_step", "grid": [nr, nz], "state_mb": sim.StepBytes() / 2e6,
                          "gbs_by_tile_rows": rows}), flush=True)
        sim.close()
    return 0


def measure_shot(args, fd, name):
    """One single-shot workload (c1/c2): --warmup untimed and --steps timed whole shots."""
    import numpy as np

    w = workload(name)
    if args.time_steps:
        w.steps = args.time_steps
    dt, nr, nz = w.model.Grid()
    log(f"workload {w.name}: {nr}x{nz}, {w.steps} steps, {len(w.ranges)} receivers")
    sim = fd.New(w.model)
    sim.Drive(w.force)
    cols = [sim.Column(r)[0] for r in w.ranges]
    sim.SetReceivers(cols, w.steps)
    use_graph = not args.no_graph
    step_bytes = sim.StepBytes()
    env_rows = os.environ.get("GEOFDTD_TILE_ROWS")
    tile_rows = int(env_rows) if env_rows else sim.Autotune(10)
    if env_rows:
        sim.SetTileRows(tile_rows)
    shape = sim.LaunchShape()
    span = shape["steps_per_launch"]
    log(f"launch shape {shape}")
    for i in range(args.warmup):
        sim.Reset()
        sim.Advance(w.steps, use_graph)
        sim.SyncTiming()
        log(f"warmup shot {i + 1}/{args.warmup} done")
    sim.SyncTiming()
    t0 = time.perf_counter()
    for i in range(args.steps):
        sim.Reset()
        sim.Advance(w.steps, use_graph)
    kernel_ms, launches = sim.SyncTiming()
    wall = time.perf_counter() - t0
    log(f"timed {args.steps} shots in {wall:.3f} s; {launches} step launches in {kernel_ms:.1f} ms of device time")
    traces = sim.ReadTraces(len(cols))
    sim.close()
    if not np.all(np.isfinite(traces)):
        raise SystemExit("non-finite traces")
    avg_launch_s = kernel_ms / 1e3 / max(launches, 1)
    # algorithmic bytes of one launch: every step it takes reads and writes
    # the state once (step_bytes); a two-step launch moves the state once
    # for both (state_pass_gbs)
    achieved = step_bytes * span / avg_launch_s / 1e9
    roof = {
        "bound": "hbm",
        "achieved": achieved,
        "peak": HBM_PEAK_GBS,
        "unit": "GB/s",
        "frac": achieved / HBM_PEAK_GBS,
        "traffic": traffic_from_profiles(w.name, span),
        "kernel": f"{kernel_name(span)} (fused velocity+stress+memory-variable step"
                  + (")" if span == 1 else "s, two per launch)"),
        "steps_per_launch": span,
        "tile_rows": shape["tile_rows"],
        "tiles_per_workgroup": shape["waves"],
        "algorithmic_bytes_per_launch": step_bytes * span,
        "avg_launch_us": avg_launch_s * 1e6,
        "state_bytes_per_launch": step_bytes,
        "state_pass_gbs": step_bytes / avg_launch_s / 1e9,
    }
    res = {
        "value": len(w.ranges) * args.steps / wall,
        "ms_per_step": wall * 1e3 / args.steps,
        "config": describe(w, nr, nz, dt),
        "roofline": roof,
        "cells_per_s": nr * nz * w.steps * args.steps / wall,
    }
    return w, traces, res



def main() -> int:
    ap = argparse.ArgumentParser(description=__doc__, formatter_class=argparse.RawDescriptionHelpFormatter)
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10,
                    help="timed steps: c3, each advances the whole batch by 1/steps of its run; c1/c2, whole shots")
    ap.add_argument("--warmup", type=int, default=1, help="untimed steps")
    ap.add_argument("--workload", default="c3",
                    help="c3 (default: batched bank generation, 256 seeded soils), c2-X (config 2 as stated, "
                         "per-layer Q), c2-R (the reference's one-SLS physics), c2w-X / c2w-R (config 2 on a grid "
                         "4x wider and 2x deeper, 1.2 GB of state), c1-R, c1-X, surrogate (train on FDTD banks, serve held-out "
                         "queries behind the FDTD error gate), fk (one bank by wavenumber integration), bank-lookup "
                         "(batched interpolation in a resident bank), stream (chunked convolution), engine (runtime "
                         "synthesis of a walk past a geophone line)")
    ap.add_argument("--also", default="c2-X", help="single-shot workload measured over --also-shots shots and "
                    "reported under 'also' (empty to skip)")
    ap.add_argument("--also-shots", type=int, default=3)
    ap.add_argument("--no-graph", action="store_true", help="plain launches instead of HIP graphs")
    ap.add_argument("--no-fk", dest="fk", action="store_false",
                    help="skip the f-k bank line reported under 'fk' (the wavenumber-integration bank filler)")
    ap.add_argument("--cpu-sample-steps", type=int, default=400)
    ap.add_argument("--cpu-threads", type=int, default=0, help="0: min(16, cpu_count)")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--time-steps", type=int, default=0,
                    help="shorten every shot to this many time steps (profiling runs; recorded in config)")
    ap.add_argument("--c3-models", type=int, default=256, help="config 3: soil models in the batch")
    ap.add_argument("--sur-soils", type=int, default=4096, help="surrogate: training soils (FDTD-solved)")
    ap.add_argument("--sur-queries", type=int, default=65536, help="surrogate: held-out queries served")
    ap.add_argument("--sur-steps", type=int, default=60000, help="surrogate: Adam steps (batch 4096)")
    ap.add_argument("--sur-lr", type=float, default=2e-3, help="surrogate: peak learning rate")
    ap.add_argument("--sur-save", default="", help="surrogate: write the trained weights to this .npz")
    ap.add_argument("--sur-weights", default=SUR_WEIGHTS,
                    help="weights the served line of the default run loads (written by --workload surrogate "
                         "--sur-save)")
    ap.add_argument("--sur-speed-only", action="store_true",
                    help="surrogate: time the kernels on synthetic inputs only (no FDTD data, no accuracy)")
    ap.add_argument("--c3-seconds", type=float, default=1.0, help="config 3: simulated seconds per model")
    ap.add_argument("--calibrate-copy", type=int, default=0, metavar="MIB",
                    help="instead of the benchmark, run 5 plain device copies of MIB MiB (the HBM counter "
                         "calibration for rocprofv3 --pmc runs) and exit")
    ap.add_argument("--ceiling", action="store_true",
                    help="instead of the benchmark, measure what bounds the step kernel: device copies of 64 MiB "
                         "to 4 GiB, and k_step on the config-2 (GSLS) medium at five grid extents (81 MB to 2.4 GB "
                         "of state per step), one JSON line each")
    args = ap.parse_args()

    if args.ceiling:
        return ceiling(args)

    if args.calibrate_copy:
        import torch

        n = args.calibrate_copy * (1 << 20) // 4
        x = torch.ones(n, dtype=torch.float32, device="cuda")
        y = torch.empty_like(x)
        for _ in range(5):
            y.copy_(x)
        torch.cuda.synchronize()
        print(json.dumps({"copy_bytes_read": 4 * n, "copy_bytes_written": 4 * n, "copies": 5}))
        return 0

    if args.gpus != 1:
        print(f"bench.py runs on one GPU only (asked for {args.gpus}); every shot is independent, so "
              "N GPUs would each run their own shots with no exchange", file=sys.stderr)
        return 2

    pkg = load_package()
    fd = pkg.fdtd
    import numpy as np

    if args.workload == "c3":
        out = bench_c3(args, fd)
        if args.also:
            shots_args = argparse.Namespace(**vars(args))
            shots_args.steps, shots_args.warmup = args.also_shots, 1
            _, _, also = measure_shot(shots_args, fd, args.also)
            out["also"] = {"workload": args.also, **also}
        if args.sur_weights and os.path.exists(args.sur_weights):
            out["surrogate"] = bench_served(args, fd)
            # per trace against the config-3 batch (larger grids than this
            # family's; the same-soil ratio is served_vs_fdtd_same_soils)
            out["surrogate"]["inference_vs_config3_fdtd_per_trace"] = (
                out["surrogate"]["inference_traces_per_s"] / out["value"])
        if args.fk:
            out["fk"] = bench_fk(args, fd, standalone=False)
            out["bank_lookup"] = bench_bank_lookup(args, fd)
            out["stream"] = bench_stream(args, fd)
            out["engine"] = bench_engine(args, fd)
        print(json.dumps(out), flush=True)
        return 0
    if args.workload == "surrogate":
        return bench_surrogate(args, fd)
    if args.workload == "fk":
        return bench_fk(args, fd)
    if args.workload == "stream":
        res = bench_stream(args, fd)
        print(json.dumps({"metric": METRIC, "n_gpus": 1, "higher_is_better": True, "scaling": "weak",
                          "vs_baseline": None, "data": "synthetic", **res}), flush=True)
        return 0
    if args.workload == "engine":
        res = bench_engine(args, fd)
        print(json.dumps({"metric": METRIC, "n_gpus": 1, "higher_is_better": True, "scaling": "weak",
                          "vs_baseline": None, "data": "synthetic", **res}), flush=True)
        return 0
    if args.workload == "bank-lookup":
        res = bench_bank_lookup(args, fd)
        print(json.dumps({"metric": METRIC, "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
                          "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "data": "synthetic",
                          **res}), flush=True)
        return 0

    w, traces, main_res = measure_shot(args, fd, args.workload)
    out = {
        "metric": METRIC,
        "value": main_res["value"],
        "unit": "traces/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": main_res["ms_per_step"],
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": main_res["config"],
        "roofline": main_res["roofline"],
        "cells_per_s": main_res["cells_per_s"],
    }
    value = out["value"]
    if not args.no_cpu:
        threads = args.cpu_threads or min(16, os.cpu_count() or 1)
        log(f"cpu baseline: {threads} threads x {args.cpu_sample_steps} steps")
        base, acc = cpu_baseline(w, traces, args.cpu_sample_steps, threads)
        out["cpu_baseline"] = base
        out["accuracy"] = acc
        out["speedup_vs_cpu"] = value / base["value"]
    else:
        out["cpu_baseline"] = None
    if args.also:
        _, _, also = measure_shot(args, fd, args.also)
        out["also"] = {"workload": args.also, **also}
    if args.fk:
        out["fk"] = bench_fk(args, fd, standalone=False)
        out["bank_lookup"] = bench_bank_lookup(args, fd)
        out["stream"] = bench_stream(args, fd)
        out["engine"] = bench_engine(args, fd)
    print(json.dumps(out), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
