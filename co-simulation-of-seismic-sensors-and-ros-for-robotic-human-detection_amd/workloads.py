"""The named workloads: models, sources and receivers of the benchmark
configurations, built through the solver's own API.

    c1(variant)   homogeneous half-space, 520 x 270 cells at 0.1 m, 10,000 steps
    c2(variant)   two layers over a half-space, 2048 x 1024 cells at 0.05 m
    c3(...)       seeded soil models for batched bank generation

variant "R" is the physics the reference can express (one scalar standard
linear solid, Qp = Qs, one Q for the whole model); "X" is the extension
(per-layer Q, Qp != Qs, three mechanisms fitted over 5-150 Hz).
"""

from __future__ import annotations

import math
from dataclasses import dataclass
from typing import List

import numpy as np

from . import fdtd as fd


@dataclass
class Workload:
    name: str
    model: fd.Model
    force: np.ndarray
    ranges: List[float]
    steps: int

    def shot(self) -> fd.Shot:
        return fd.Shot(self.model, self.force, self.ranges, self.steps)


def _courant_for(model: fd.Model, dt: float) -> float:
    """The Courant factor that makes New's time step equal dt."""
    probe = fd.New(model)
    try:
        base = probe.Dt() / (model.Courant if model.Courant > 0 else 0.5)
    finally:
        probe.close()
    return dt / base


def c1(variant: str = "R", steps: int = 10_000) -> Workload:
    """Config 1: Vp 400, Vs 200, rho 1800; 520 x 270 cells at h = 0.1 m with a
    20-cell C-PML (MaxRange 49.9 m, Depth 24.9 m); vertical Ricker force of
    800 N at 40 Hz on the axis; dt = 1e-4 s; v_z receivers every metre from 1
    to 49 m (the reference rejects 50 m: it is not inside MaxRange)."""
    if variant == "R":
        stack = [fd.Layer(0, 400, 200, 1800, 30)]
        model = fd.Model(Stack=stack, Relax=fd.Fit(30, 30), RefFreq=30)
    elif variant == "X":
        stack = [fd.Layer(0, 400, 200, 1800, 30, 60)]
        model = fd.Model(Stack=stack, RefFreq=30, GslsMechanisms=3, GslsFmin=5, GslsFmax=150)
    else:
        raise ValueError(f"unknown variant {variant!r}")
    model.MaxRange, model.Depth, model.Spacing = 49.9, 24.9, 0.1
    model.PMLCells, model.DominantFreq = 20, 40
    model.Courant = _courant_for(model, 1e-4)
    dt = fd.New(model).Dt()
    force = fd.Ricker(800, 40, dt, steps)
    ranges = [float(r) for r in range(1, 50)]
    return Workload(f"c1-{variant}", model, force, ranges, steps)
