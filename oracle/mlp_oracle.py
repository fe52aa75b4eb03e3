"""FP64 restatement of the learned-bandwidth MLP (forward, loss, gradient).

TEST INFRASTRUCTURE ONLY: the checker of the HIP MLP kernels
(entropical_amd/csrc/mlp.hip).  Parameter layout matches the library: per
layer W[n_in][n_out] (row-major) followed by b[n_out]; tanh hidden units,
identity output, loss = mean squared error.  With ``bf16=True`` the weights
(not the biases) are rounded to bfloat16 (nearest, ties to even) before use,
as the device does; the gradient is then the exact FP64 gradient with
respect to those rounded weights.  Pinned by central finite differences
(tests/test_oracle.py).
"""
from __future__ import annotations

import numpy as np


def bf16_round(w: np.ndarray) -> np.ndarray:
    u = np.ascontiguousarray(w, dtype=np.float32).view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) & 0xFFFF0000
    return u.astype(np.uint32).view(np.float32).astype(np.float64)


def unpack(params: np.ndarray, dims):
    p = np.asarray(params, dtype=np.float64)
    out, off = [], 0
    for a, b in zip(dims[:-1], dims[1:]):
        W = p[off: off + a * b].reshape(a, b)
        off += a * b
        bias = p[off: off + b]
        off += b
        out.append((W, bias))
    assert off == p.size
    return out


def forward(params, dims, X, bf16=True):
    layers = unpack(params, dims)
    acts = [np.asarray(X, dtype=np.float64)]
    for li, (W, b) in enumerate(layers):
        Wu = bf16_round(W) if bf16 else W
        z = acts[-1] @ Wu + b
        acts.append(np.tanh(z) if li + 1 < len(layers) else z)
    return acts


def loss_and_grad(params, dims, X, t, bf16=True):
    layers = unpack(params, dims)
    acts = forward(params, dims, X, bf16)
    y = acts[-1][:, 0]
    t = np.asarray(t, dtype=np.float64)
    B = X.shape[0]
    loss = float(np.mean((y - t) ** 2))
    dz = (2.0 * (y - t) / B)[:, None]
    grads = []
    for li in range(len(layers) - 1, -1, -1):
        W, _ = layers[li]
        Wu = bf16_round(W) if bf16 else W
        gW = acts[li].T @ dz
        gb = dz.sum(axis=0)
        grads.append((gW, gb))
        if li > 0:
            dz = (dz @ Wu.T) * (1.0 - acts[li] ** 2)
    grads.reverse()
    flat = np.concatenate([np.concatenate([gW.reshape(-1), gb]) for gW, gb in grads])
    return loss, flat
