"""The streaming-convolution oracle (oracle/stream_oracle.py) pinned against
the reference's own stream tests (internal/stream/stream_test.go), and the
device convolver's host-side argument checks.  CPU only.

  * chunked output equals the whole-record convolution within 1e-9 of the peak
    for chunk sizes 1 to 1024 (stream_test.go:22-55);
  * the output is independent of the chunk size (stream_test.go:59-91);
  * no discontinuity at chunk boundaries (stream_test.go:95-128);
  * Reset returns the filter to its initial state (stream_test.go:222-239);
  * the partition count and the argument errors (stream_test.go:165-168,
    241-258).
"""

import math

import numpy as np
import pytest

import stream_oracle as S
from conftest import load_package


def decaying(n, tau, w):
    i = np.arange(n)
    return np.exp(-i / tau) * np.sin(i * w)


def test_chunked_equals_monolithic():
    impulse = decaying(700, 120, 0.21)
    x = np.random.default_rng(12).standard_normal(4096)
    want = S.Monolithic(x, impulse)
    scale = np.max(np.abs(want))
    for chunk in (1, 2, 7, 16, 20, 64, 128, 512, 1024):
        got = S.Convolver(impulse, chunk).ProcessAll(x)
        assert np.max(np.abs(got - want)) <= 1e-9 * scale, chunk


def test_output_is_independent_of_chunk_size():
    impulse = decaying(300, 60, 0.4)
    x = np.random.default_rng(34).standard_normal(3000)
    want = S.Convolver(impulse, 64).ProcessAll(x)
    scale = np.max(np.abs(want))
    for chunk in (3, 17, 50, 200, 333):
        assert np.max(np.abs(S.Convolver(impulse, chunk).ProcessAll(x) - want)) <= 1e-9 * scale


def test_no_discontinuity_at_chunk_boundaries():
    chunk = 32
    out = S.Convolver(decaying(400, 90, 0.3), chunk).ProcessAll(np.sin(2 * math.pi * 5 * np.arange(2048) / 2000))
    d = np.abs(np.diff(out))
    at = d[np.arange(1, len(out)) % chunk == 0].max()
    elsewhere = d[np.arange(1, len(out)) % chunk != 0].max()
    assert at <= 1.5 * elsewhere


def test_reset_clears_state():
    c = S.Convolver([1, 0.5, 0.25, 0.125], 4)
    first = c.ProcessAll([1, 2, 3, 4, 5, 6, 7, 8])
    c.Reset()
    assert np.max(np.abs(c.ProcessAll([1, 2, 3, 4, 5, 6, 7, 8]) - first)) <= 1e-12


def test_partitions_and_argument_errors():
    assert S.Convolver(np.ones(700), 64).Partitions() == 11
    with pytest.raises(ValueError, match="stream: chunk size must be positive, got 0"):
        S.Convolver([1.0], 0)
    with pytest.raises(ValueError, match="stream: impulse response must not be empty"):
        S.Convolver([], 8)


def test_device_convolver_argument_errors():
    """The checks that run before any device work, with the reference's texts."""
    st = load_package().stream
    fd = load_package().fdtd
    with pytest.raises(fd.FdtdError) as e:
        st.NewConvolver([1.0], 0)
    assert str(e.value) == "stream: chunk size must be positive, got 0"
    with pytest.raises(fd.FdtdError) as e:
        st.NewConvolver([], 8)
    assert str(e.value) == "stream: impulse response must not be empty"
