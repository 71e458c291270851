"""The bank file writer against the reference's format (internal/bank/
format.go:32-46, 115-240; py/bankfmt/bankfmt.py) and its cross-language
fixtures, copied from the reference's testdata/bank/ into tests/golden/bank/.

The Go fixture pins the writer byte for byte: it is the conformance pattern
(conformance_test.go:10-12) under the Go fixture's header, written by Go's
encoding/json and binary.Write.  The Python fixture is read back value by
value, as py/bankfmt/conformance.py does."""

import json
import math
import os
import struct

import numpy as np
import pytest

from conftest import ROOT

GOLDEN = os.path.join(ROOT, "tests", "golden", "bank")


def pattern(i, k):
    """conformPattern (internal/bank/conformance_test.go:10-12)."""
    return complex((i + 1) * 10.0 ** (k % 5 - 2), -(k + 1) * 0.25 + i)


def read_bank(path):
    """A reader restating bankfmt.read (py/bankfmt/bankfmt.py:55-78)."""
    raw = open(path, "rb").read()
    assert raw[:8] == b"GEOBANK\x00"
    version, hlen = struct.unpack_from("<II", raw, 8)
    assert version == 1
    header = json.loads(raw[16:16 + hlen])
    off = 16 + hlen
    if off % 4096:
        off += 4096 - off % 4096
    count, bins = header["ranges"]["count"], header["samples"] // 2 + 1
    flat = np.frombuffer(raw, dtype="<f4", count=2 * count * bins, offset=off).reshape(count, bins, 2)
    return header, flat[..., 0] + 1j * flat[..., 1], raw


def go_fixture_header(fdtd):
    return fdtd.BankHeader(Solver="go conformance fixture", Medium=[fdtd.Layer(0, 500, 200, 1700)],
                           SampleRateHz=2000, Samples=32, MinM=1, MaxM=4, Count=7,
                           Component="conformance fixture, not physical", Units="m/N")


def pattern_data(count, bins):
    return np.array([[pattern(i, k) for k in range(bins)] for i in range(count)], dtype=np.complex64)


def test_writer_reproduces_the_go_fixture_byte_for_byte(fdtd, tmp_path):
    h = go_fixture_header(fdtd)
    out = tmp_path / "ours.bank"
    fdtd.WriteBank(str(out), h, pattern_data(7, 17))
    assert out.read_bytes() == open(os.path.join(GOLDEN, "go_written.bank"), "rb").read()


@pytest.mark.parametrize("name", ["go_written.bank", "python_written.bank"])
def test_fixtures_hold_the_conformance_pattern(name):
    header, data, _ = read_bank(os.path.join(GOLDEN, name))
    assert header["ranges"]["count"] == 7 and header["samples"] == 32
    worst = 0.0
    for i in range(7):
        for k in range(17):
            want = np.complex64(pattern(i, k))
            worst = max(worst, abs(data[i, k] - want) / max(abs(want), 1e-30))
    assert worst <= 1e-6


def test_written_bank_round_trips_through_a_bankfmt_style_reader(fdtd, tmp_path):
    rng = np.random.default_rng(7)
    h = fdtd.BankHeader(Solver="geofdtd test", Medium=[fdtd.Layer(1.5, 300, 150, 1700, 20), fdtd.Layer(0, 800, 400, 2000, 80)],
                        SampleRateHz=500, Samples=1024, MinM=1, MaxM=50, Count=50, Notes="seed 7 <&>")
    data = (rng.standard_normal((50, 513)) + 1j * rng.standard_normal((50, 513))).astype(np.complex64) * 1e-9
    fdtd.WriteBank(str(tmp_path / "b.bank"), h, data)
    header, back, raw = read_bank(str(tmp_path / "b.bank"))
    np.testing.assert_array_equal(back, data)
    assert header["provenance"] == {"solver": "geofdtd test", "notes": "seed 7 <&>"}
    assert header["medium"][0] == {"thickness_m": 1.5, "vp_mps": 300, "vs_mps": 150, "density_kgm3": 1700, "qs": 20}
    assert header["component"] == "vertical surface displacement per unit vertical surface point force"
    assert header["units"] == "m/N"
    # Go's encoding/json escapes <, > and & in strings
    assert b'"notes":"seed 7 \\u003c\\u0026\\u003e"' in raw
    assert (16 + struct.unpack_from("<I", raw, 12)[0]) <= 4096 and len(raw) == 4096 + 50 * 513 * 8


@pytest.mark.parametrize("x,want", [
    (2000.0, "2000"), (1.0, "1"), (0.0, "0"), (0.1, "0.1"), (1.5, "1.5"), (-2.25, "-2.25"),
    (1e-6, "0.000001"), (1e-7, "1e-7"), (1.5e-10, "1.5e-10"), (123456789012.0, "123456789012"),
    (1e20, "100000000000000000000"), (1e21, "1e+21"), (3.5e22, "3.5e+22"), (0.30000000000000004, "0.30000000000000004"),
])
def test_header_numbers_use_go_json_formatting(fdtd, x, want):
    h = fdtd.BankHeader(Solver="s", Medium=[], SampleRateHz=x if x > 0 else 1, Samples=2, MinM=1, MaxM=2, Count=2)
    text = fdtd.BankHeaderJSON(h)
    if x > 0:
        assert f'"sample_rate_hz":{want},' in text
    assert '"medium":null' in text


REJECT = [
    (dict(Samples=33), "bank: samples must be positive and even, got 33"),
    (dict(Samples=0), "bank: samples must be positive and even, got 0"),
    (dict(SampleRateHz=0), "bank: sample rate must be positive, got 0"),
    (dict(Count=0), "bank: range grid must have at least one entry, got 0"),
    (dict(MinM=0), "bank: minimum range must be positive, got 0 m"),
    (dict(MaxM=1), "bank: maximum range 1 m must exceed minimum 1 m"),
    (dict(Component=""), "bank: component must be named"),
    (dict(Medium=[(0, 250, 200, 1700)]),
     "bank: layer 0: Vp=250 must exceed sqrt(2)*Vs=282.842712474619 for a non-negative Poisson ratio"),
]


@pytest.mark.parametrize("change,message", REJECT)
def test_invalid_headers_are_rejected_like_header_validate(fdtd, tmp_path, change, message):
    kw = dict(Solver="s", Medium=[fdtd.Layer(0, 500, 200, 1700)], SampleRateHz=2000, Samples=32, MinM=1, MaxM=4,
              Count=7, Component="c")
    for k, v in change.items():
        kw[k] = [fdtd.Layer(*l) for l in v] if k == "Medium" else v
    h = fdtd.BankHeader(**kw)
    data = np.zeros((max(h.Count, 0), h.Samples // 2 + 1), dtype=np.complex64)
    with pytest.raises(fdtd.FdtdError) as e:
        fdtd.WriteBank(str(tmp_path / "x.bank"), h, data)
    assert str(e.value) == message
