"""The device streaming convolver against the oracle (the reference's
partitioned overlap-add restated with numpy's FFT) and against the
whole-record convolution, as the reference's stream tests require.

Tolerance: the device sums taps directly in FP64, the oracle goes through
FFTs; both equal the whole-record convolution to rounding, and the reference's
own bound is 1e-9 of the peak (stream_test.go:48).  Measured differences are
near 1e-15 of the peak; the tests hold them to 1e-12.
"""

import numpy as np
import pytest

import stream_oracle as S
from conftest import load_package

st = load_package().stream
fd = load_package().fdtd

pytestmark = pytest.mark.gpu


def decaying(n, tau, w, phase=0.0):
    i = np.arange(n)
    return np.exp(-i / tau) * np.sin(i * w + phase)


def close(got, want, tol=1e-12):
    scale = np.max(np.abs(want))
    return np.max(np.abs(np.asarray(got) - want)) <= tol * scale


def test_process_all_equals_the_oracle_and_the_whole_record():
    impulse = decaying(700, 120, 0.21)
    x = np.random.default_rng(12).standard_normal(4096)
    mono = S.Monolithic(x, impulse)
    for chunk in (1, 2, 7, 16, 20, 64, 128, 512, 1024):
        print(f"[stream] chunk {chunk}", flush=True)
        c = st.NewConvolver(impulse, chunk)
        assert c.ChunkSize() == chunk and c.Partitions() == (700 + chunk - 1) // chunk
        got = c.ProcessAll(x)
        assert close(got, S.Convolver(impulse, chunk).ProcessAll(x)), chunk
        assert close(got, mono), chunk
        c.close()


def test_streaming_many_receivers_chunk_by_chunk():
    """64 receivers, each its own response, fed 20-sample chunks (10 ms at
    2 kHz), then a ProcessAll continuing from that state, then Reset."""
    rng = np.random.default_rng(5)
    R, L, chunk = 64, 2000, 20
    h = np.stack([decaying(L, 200 + 10 * r, 0.05 + 0.003 * r, r) for r in range(R)])
    x = rng.standard_normal((R, 3000))
    c = st.NewConvolver(h, chunk)
    oracles = [S.Convolver(h[r], chunk) for r in range(R)]
    out = np.concatenate([c.Process(x[:, k * chunk:(k + 1) * chunk]) for k in range(100)], axis=1)
    want = np.stack([np.concatenate([o.Process(x[r, k * chunk:(k + 1) * chunk]) for k in range(100)])
                     for r, o in enumerate(oracles)])
    for r in range(R):
        assert close(out[r], want[r]), r
    rest = c.ProcessAll(x[:, 2000:])
    want_rest = np.stack([o.ProcessAll(x[r, 2000:]) for r, o in enumerate(oracles)])
    for r in range(R):
        assert close(rest[r], want_rest[r]), r
        assert close(np.concatenate([out[r], rest[r]]), S.Monolithic(x[r], h[r]))
    c.Reset()
    again = np.concatenate([c.Process(x[:, k * chunk:(k + 1) * chunk]) for k in range(10)], axis=1)
    assert np.array_equal(again, out[:, :200])
    print(f"[stream] 64 receivers x 2000 taps, one 20-sample chunk: {c.KernelMs() * 1e3:.1f} us of kernels",
          flush=True)
    c.close()


def test_chunk_length_is_checked():
    c = st.NewConvolver([1.0, 2.0], 4)
    with pytest.raises(fd.FdtdError) as e:
        c.Process(np.zeros(3))
    assert str(e.value) == "stream: chunk length mismatch: in 3, out 4, want 4"
    c.close()


@pytest.mark.parametrize("chunk", [1, 64, 65, 300])
def test_process_chunk_by_chunk_at_both_kernel_forms(chunk):
    """Chunks of at most 64 samples take the latency kernel, longer ones the
    tiled kernel over the ring; both must follow the oracle chunk by chunk."""
    impulse = decaying(900, 150, 0.13)
    x = np.random.default_rng(chunk).standard_normal(6 * chunk + 5)
    c = st.NewConvolver(np.stack([impulse, -0.5 * impulse]), chunk)
    o = S.Convolver(impulse, chunk)
    for k in range(6):
        seg = x[k * chunk:(k + 1) * chunk]
        got = c.Process(np.stack([seg, seg]))
        want = o.Process(seg)
        assert close(got[0], want), (chunk, k)
        assert close(got[1], -0.5 * want), (chunk, k)
    c.close()


def test_chunk_kernel_at_every_workgroup_shape():
    """The latency kernel gives each 1,024-thread workgroup np outputs of one
    receiver, np = clamp(ceil(chunk x receivers / 256), 1, chunk): 300
    receivers x 64 samples take np = 64 (16 slices per output), 40 x 48 take
    np = 8, 3 x 7 take np = 1 (1,024 slices); all must follow the oracle."""
    rng = np.random.default_rng(9)
    for R, chunk, L in ((300, 64, 700), (40, 48, 1500), (3, 7, 400)):
        print(f"[stream] {R} receivers x chunk {chunk}", flush=True)
        h = np.stack([decaying(L, 50 + r % 97, 0.02 + 0.001 * (r % 31), r) for r in range(R)])
        x = rng.standard_normal((R, 4 * chunk))
        c = st.NewConvolver(h, chunk)
        got = np.concatenate([c.Process(x[:, k * chunk:(k + 1) * chunk]) for k in range(4)], axis=1)
        for r in range(0, R, max(1, R // 12)):
            assert close(got[r], S.Monolithic(x[r], h[r])), (R, chunk, r)
        c.close()
