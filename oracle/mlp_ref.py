"""CPU restatement of the planner's estimator network for checking the HIP kernels.

TEST INFRASTRUCTURE ONLY: imported by tests/ as the checker; nothing in the engine calls it.

The network (lapis-silo_amd/csrc/mlp.h, mlp.hip): inputs -> ReLU hidden layers -> one linear
output; parameters flat, per layer W (out x in, row-major) then b.  Loss: mean squared error.
Everything here runs in float64 on the parameter values the kernels use, so the kernels' FP32
results differ from it only by rounding.  The analytic gradients below are pinned against central
finite differences by tests/test_mlp_oracle.py.
"""
from __future__ import annotations

import numpy as np


def layer_shapes(dims):
    return [(dims[l + 1], dims[l]) for l in range(len(dims) - 1)]


def unpack(params, dims):
    out = []
    o = 0
    for (n_out, n_in) in layer_shapes(dims):
        w = params[o:o + n_out * n_in].reshape(n_out, n_in)
        o += n_out * n_in
        b = params[o:o + n_out]
        o += n_out
        out.append((w, b))
    assert o == len(params)
    return out


def forward(params, dims, x):
    """Returns the output and every layer's activations (input first)."""
    acts = [np.asarray(x, np.float64)]
    layers = unpack(np.asarray(params, np.float64), dims)
    for i, (w, b) in enumerate(layers):
        z = acts[-1] @ w.T + b
        if i + 1 < len(layers):
            z = np.maximum(z, 0.0)
        acts.append(z)
    return acts[-1][:, 0], acts


def loss_and_grads(params, dims, x, y):
    out, acts = forward(params, dims, x)
    n = len(y)
    err = out - np.asarray(y, np.float64)
    loss = float(np.mean(err ** 2))
    layers = unpack(np.asarray(params, np.float64), dims)
    grads = []
    dz = (2.0 * err / n)[:, None]
    for i in range(len(layers) - 1, -1, -1):
        w, _ = layers[i]
        a_in = acts[i]
        grads.append((dz.T @ a_in, dz.sum(axis=0)))
        if i == 0:
            break
        da = dz @ w
        dz = da * (acts[i] > 0)
    grads.reverse()
    flat = np.concatenate([np.concatenate([gw.ravel(), gb]) for gw, gb in grads])
    return loss, flat


def bf16_round(x):
    """Round-to-nearest-even to bfloat16, returned as float32."""
    a = np.ascontiguousarray(x, dtype=np.float32).view(np.uint32).astype(np.uint64)
    a = (a + 0x7FFF + ((a >> 16) & 1)) & 0xFFFF0000
    return a.astype(np.uint32).view(np.float32)
