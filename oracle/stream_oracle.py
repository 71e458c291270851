"""CPU restatement of the reference's stream.Convolver (test infrastructure
only): uniformly partitioned overlap-add, as internal/stream/stream.go:72-196
computes it, with numpy's real FFT in place of gonum's (the same transform:
gonum's Coefficients is numpy's rfft, its Sequence is n times numpy's irfft).

Nothing in the product imports this module.  Pinned by
tests/test_stream_oracle.py against the reference's own stream tests.
"""

from __future__ import annotations

import numpy as np


class Convolver:
    """stream.Convolver (stream.go:47-70) for one receiver."""

    def __init__(self, impulse, chunk: int):
        if chunk <= 0:
            raise ValueError(f"stream: chunk size must be positive, got {chunk}")
        h = np.asarray(impulse, dtype=np.float64)
        if h.size == 0:
            raise ValueError("stream: impulse response must not be empty")
        n = 2 * chunk
        parts = (h.size + chunk - 1) // chunk
        self.chunk, self.n = chunk, n
        self.blocks = []
        for p in range(parts):
            block = np.zeros(n)
            seg = h[p * chunk: min((p + 1) * chunk, h.size)]
            block[: seg.size] = seg
            self.blocks.append(np.fft.rfft(block))
        self.fdl = [np.zeros(n // 2 + 1, dtype=complex) for _ in range(parts)]
        self.pos = 0
        self.tail = np.zeros(chunk)

    def Partitions(self) -> int:
        return len(self.blocks)

    def Process(self, x) -> np.ndarray:
        """stream.go:120-162."""
        x = np.asarray(x, dtype=np.float64)
        if x.size != self.chunk:
            raise ValueError(f"stream: chunk length mismatch: in {x.size}, out {self.chunk}, want {self.chunk}")
        buf = np.zeros(self.n)
        buf[: self.chunk] = x
        self.fdl[self.pos] = np.fft.rfft(buf)
        acc = np.zeros(self.n // 2 + 1, dtype=complex)
        P = len(self.fdl)
        for p in range(P):
            acc += self.fdl[(self.pos - p + 2 * P) % P] * self.blocks[p]
        self.pos = (self.pos + 1) % P
        seq = np.fft.irfft(acc, self.n)  # gonum's Sequence is n times this, then scaled by 1/n
        out = seq[: self.chunk] + self.tail
        self.tail = seq[self.chunk:].copy()
        return out

    def Reset(self):
        for s in self.fdl:
            s[:] = 0
        self.pos = 0
        self.tail[:] = 0

    def ProcessAll(self, x) -> np.ndarray:
        """stream.go:175-186."""
        x = np.asarray(x, dtype=np.float64)
        chunks = (x.size + self.chunk - 1) // self.chunk
        src = np.zeros(chunks * self.chunk)
        src[: x.size] = x
        out = np.concatenate([self.Process(src[k * self.chunk:(k + 1) * self.chunk]) for k in range(chunks)])
        return out[: x.size]


def Monolithic(x, impulse) -> np.ndarray:
    """stream.Monolithic (stream.go:190-196): the whole convolution truncated
    to the input length, here by direct summation."""
    full = np.convolve(np.asarray(x, dtype=np.float64), np.asarray(impulse, dtype=np.float64))
    return full[: len(x)]
