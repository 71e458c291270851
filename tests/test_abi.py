"""The C ABI: libgeofdtd.so loads and exports every function include/geofdtd.h
declares; the ctypes structs have the C layout; the host-side arithmetic
(model validation, grid, time step, spacing) matches the oracle and the
reference's error texts.  None of this launches a kernel."""

import ctypes as C
import os
import re
import subprocess

import numpy as np
import pytest

from conftest import ROOT
from media import loam

HEADER = os.path.join(ROOT, "include", "geofdtd.h")


def declared_functions():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(gf_[a-z0-9_]+)\s*\(", text)))


def test_library_exports_every_declared_symbol(fdtd):
    lib = fdtd.load_library()
    names = declared_functions()
    assert len(names) >= 25
    for name in names:
        assert hasattr(lib, name), name
    assert set(names) == set(fdtd.SIGNATURES), set(names) ^ set(fdtd.SIGNATURES)
    assert lib.gf_abi_version() == 1


def test_struct_layout_matches_c(fdtd, tmp_path):
    src = tmp_path / "layout.c"
    src.write_text(
        '#include <stdio.h>\n#include <stddef.h>\n#include "geofdtd.h"\n'
        "int main(void){printf(\"%zu %zu %zu %zu %zu %zu %zu %zu %zu\\n\", sizeof(gf_layer), sizeof(gf_model),"
        " sizeof(gf_shot), sizeof(gf_result), offsetof(gf_model, courant), offsetof(gf_model, gsls_fmax_hz),"
        " offsetof(gf_result, status), offsetof(gf_result, error), offsetof(gf_shot, steps)); return 0;}\n"
    )
    exe = tmp_path / "layout"
    subprocess.run(["gcc", "-I", os.path.join(ROOT, "include"), str(src), "-o", str(exe)], check=True)
    got = [int(x) for x in subprocess.run([str(exe)], capture_output=True, text=True, check=True).stdout.split()]
    want = [C.sizeof(fdtd.GfLayer), C.sizeof(fdtd.GfModel), C.sizeof(fdtd.GfShot), C.sizeof(fdtd.GfResult),
            fdtd.GfModel.courant.offset, fdtd.GfModel.gsls_fmax_hz.offset, fdtd.GfResult.status.offset,
            fdtd.GfResult.error.offset, fdtd.GfShot.steps.offset]
    assert got == want


@pytest.mark.parametrize("kw", [
    dict(MaxRange=40, Depth=40, Spacing=0.1778, DominantFreq=25),
    dict(MaxRange=49.9, Depth=24.9, Spacing=0.1, Courant=0.5751),
    dict(MaxRange=6, Depth=6, Spacing=0.15, Courant=1.4, PMLCells=7, PMLReflection=1e-3),
])
@pytest.mark.parametrize("atten", ["elastic", "sls", "gsls"])
def test_grid_and_time_step_equal_the_oracle(fdtd, oracle, kw, atten):
    m = fdtd.Model(Stack=[fdtd.Layer(1.2, 420, 190, 1700, 25, 40), fdtd.Layer(0, 700, 320, 2000, 60)], **kw)
    if atten == "sls":
        m.Relax = fdtd.Fit(25, 30)
    if atten == "gsls":
        m.GslsMechanisms = 3
    dt, nr, nz = m.Grid()
    o = oracle.OracleSim(m)
    assert (dt, nr, nz) == (o.Dt(), *o.Cells())


def test_spacing_for_equals_the_oracle(fdtd, oracle):
    for relax in [fdtd.SLS(), fdtd.Fit(25, 30)]:
        m = fdtd.Model(Stack=loam(), Relax=relax)
        assert m.SpacingFor(75, 15) == oracle.spacing_for(m, 75, 15)
    with pytest.raises(fdtd.FdtdError, match="fdtd: bandwidth and cells per wavelength must be positive"):
        fdtd.Model(Stack=loam()).SpacingFor(0, 15)


# Each case is a model the reference's fdtd.New (fdtd.go:45-57) or
# layer.Stack.Validate (layer.go:53-73) rejects, with its exact message.
REJECTED = [
    (dict(Stack=[]), "layer: a stack needs at least a half-space"),
    (dict(Stack=[(0, 500, 0, 1700)]), "layer 0: Vs must be positive, got 0 m/s"),
    (dict(Stack=[(0, -1, 200, 1700)]), "layer 0: Vp must be positive, got -1 m/s"),
    (dict(Stack=[(0, 500, 200, 0)]), "layer 0: density must be positive, got 0 kg/m^3"),
    (dict(Stack=[(0, 280, 200, 1700)]),
     "layer 0: Vp=280 must exceed sqrt(2)*Vs=282.842712474619 for a non-negative Poisson ratio"),
    (dict(Stack=[(0, 500, 200, 1700), (0, 500, 200, 1700)]), "layer 0: thickness must be positive, got 0 m"),
    (dict(Spacing=0), "fdtd: spacing must be positive, got 0 m"),
    (dict(Spacing=-0.25), "fdtd: spacing must be positive, got -0.25 m"),
    (dict(MaxRange=0), "fdtd: domain must be positive, got 0 m by 6 m"),
    (dict(Depth=-2.5), "fdtd: domain must be positive, got 6 m by -2.5 m"),
    (dict(Relax=(1e-2, 1e-3)), "fdtd: TauEps must be at least TauSigma for a dissipative solid"),
]


@pytest.mark.parametrize("change,message", REJECTED)
def test_invalid_models_are_rejected_with_the_reference_message(fdtd, oracle, change, message):
    kw = dict(Stack=loam(), MaxRange=6, Depth=6, Spacing=0.15)
    for k, v in change.items():
        if k == "Stack":
            v = [fdtd.Layer(*l) for l in v]
        if k == "Relax":
            v = fdtd.SLS(*v)
        kw[k] = v
    m = fdtd.Model(**kw)
    with pytest.raises(fdtd.FdtdError) as got:
        fdtd.New(m)
    assert got.value.status == fdtd.GF_ERR_INVALID
    assert str(got.value) == message
    with pytest.raises(fdtd.FdtdError) as run_err:
        fdtd.Run(fdtd.Shot(m, [0.0], [1.0], 10))
    assert str(run_err.value) == message
    with pytest.raises(oracle.OracleError) as want:
        oracle.OracleSim(m)
    assert str(want.value) == message


def test_ricker_rejects_nonsense(fdtd):
    with pytest.raises(fdtd.FdtdError, match="fdtd: Ricker needs a positive frequency, step and length"):
        fdtd.Ricker(1000, 0, 1e-4, 10)
