"""CPU oracle of the Green's-function surrogate (test infrastructure only).

FP64 numpy forward and backward passes of the network the library trains on
the GPU, from the same flat parameter vector and layer table
(gf_sur_layer): an encoder of GELU layers and a decoder of transposed
convolutions whose kernel equals their stride (each a GEMM over (query,
position) rows whose output columns are ordered (tap, channel)).  The loss is
the relative spectral L2, mean over queries of |o - t|^2 / |t|^2.

`bf16=True` rounds every GEMM operand to BF16 (round to nearest even) where
the GPU kernels do -- layer inputs and weights going forward; the output
gradient, layer inputs and weights going backward -- so that the GPU can be
compared at FP32-accumulation precision; `bf16=False` is exact FP64, which
the finite-difference tests pin.  The surrogate has no counterpart in the
reference; its checks are mathematical (finite differences, closed forms).

Nothing in the product imports this module.
"""

from __future__ import annotations

from dataclasses import dataclass
from typing import List, Sequence

import numpy as np

GELU_C = np.sqrt(2.0 / np.pi)
GELU_A = 0.044715
IN_PAD = 16
FEATURES = 15


@dataclass
class Layer:
    k: int
    n: int
    gelu: bool
    rows: int
    bias_mod: int
    w_off: int
    b_off: int


def layers_from(fd, net) -> List[Layer]:
    """The layer table of a live surrogate handle."""
    return [Layer(l.k_in, l.n_out, bool(l.gelu), l.rows_per_query, l.bias_channels, l.w_offset, l.b_offset)
            for l in net.Layers()]


def bf16(x: np.ndarray) -> np.ndarray:
    """Round to BF16 (nearest even), returned as float64."""
    f = np.asarray(x, dtype=np.float32)
    u = f.view(np.uint32).astype(np.uint64)
    r = ((u + 0x7FFF + ((u >> 16) & 1)) >> 16) << 16
    out = r.astype(np.uint32).view(np.float32).astype(np.float64)
    nan = np.isnan(f)
    if nan.any():
        out[nan] = np.nan
    return out


def gelu(z):
    return 0.5 * z * (1.0 + np.tanh(GELU_C * (z + GELU_A * z ** 3)))


def gelu_grad(z):
    t = np.tanh(GELU_C * (z + GELU_A * z ** 3))
    return 0.5 * (1.0 + t) + 0.5 * z * (1.0 - t * t) * GELU_C * (1.0 + 3.0 * GELU_A * z * z)


def _w(params, L: Layer):
    return params[L.w_off: L.w_off + L.n * L.k].reshape(L.n, L.k)


def _b(params, L: Layer):
    return params[L.b_off: L.b_off + L.bias_mod][np.arange(L.n) % L.bias_mod]


def forward(params: np.ndarray, layers: Sequence[Layer], x: np.ndarray, bf: bool = False):
    """(output [B][1024], cache) for features x [B][15]."""
    p = np.asarray(params, dtype=np.float64)
    B = x.shape[0]
    a = np.zeros((B, IN_PAD))
    a[:, : x.shape[1]] = x
    q = bf16 if bf else (lambda v: v)
    acts, zs = [], []
    for L in layers:
        A = q(a.reshape(B * L.rows, L.k))
        z = A @ q(_w(p, L)).T + _b(p, L)
        acts.append(A)
        zs.append(z)
        a = gelu(z) if L.gelu else z
    return a.reshape(B, -1), (acts, zs)


def loss(o: np.ndarray, t: np.ndarray) -> np.ndarray:
    """Per-query relative spectral L2 squared."""
    return np.sum((o - t) ** 2, axis=1) / np.sum(t ** 2, axis=1)


def loss_grad(params, layers, x, t, bf: bool = False, w=None):
    """(mean loss, flat gradient) of the batch x [B][15] against t [B][1024]:
    the relative L2 |o - t|^2 / |t|^2 per query, or with weights w the
    weighted squares sum_i w_i (o_i - t_i)^2."""
    p = np.asarray(params, dtype=np.float64)
    o, (acts, zs) = forward(p, layers, x, bf)
    B = x.shape[0]
    if w is None:
        den = np.sum(t ** 2, axis=1)
        L_b = np.sum((o - t) ** 2, axis=1) / den
        d = 2.0 * (o - t) / den[:, None] / B
    else:
        L_b = np.sum(w * (o - t) ** 2, axis=1)
        d = 2.0 * w * (o - t) / B
    q = bf16 if bf else (lambda v: v)
    g = np.zeros_like(p)
    for i in range(len(layers) - 1, -1, -1):
        L = layers[i]
        dz = d.reshape(B * L.rows, L.n)
        g[L.w_off: L.w_off + L.n * L.k] = (q(dz).T @ acts[i]).ravel()
        db = dz.sum(axis=0)
        g[L.b_off: L.b_off + L.bias_mod] = np.bincount(np.arange(L.n) % L.bias_mod, weights=db,
                                                       minlength=L.bias_mod)
        if i == 0:
            break
        da = q(dz) @ q(_w(p, L))
        prev = layers[i - 1]
        d = da.reshape(B * prev.rows, prev.n) * gelu_grad(zs[i - 1])
    return float(np.mean(L_b)), g
