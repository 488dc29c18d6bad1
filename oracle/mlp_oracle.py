"""FP64 CPU restatement of the surrogate network's arithmetic.

TEST INFRASTRUCTURE ONLY: ``tests/`` use it to check the GPU training and
inference kernels (``openrcwa_amd/csrc/mlp.hip``).  The reference has no
learned surrogate, so there is no reference file to follow; this module states
the network the C ABI of ``include/openrcwa_surrogate.h`` documents:

  x -> [Linear(width) -> GELU] x depth -> Linear(n_out),
  GELU(z) = z (1 + erf(z / sqrt 2)) / 2,
  loss = mean over rows and outputs of (y - t)^2,

with one member's parameters flat in the order W_0 [n_in][width], b_0, ...,
W_depth [width][n_out], b_depth (row-major, W[i][j] maps input i to output j).

Pinning: ``tests/test_surrogate.py`` checks the analytic backward pass against
central finite differences of the loss (float64), GELU against its integral
definition z Phi(z) by quadrature, and a one-layer network (depth 1, width w)
against its closed form.
"""
from __future__ import annotations

import math

import numpy as np
from scipy.special import erf


def gelu(z):
    return 0.5 * z * (1.0 + erf(z / math.sqrt(2.0)))


def gelu_grad(z):
    return 0.5 * (1.0 + erf(z / math.sqrt(2.0))) + z * np.exp(-0.5 * z * z) / math.sqrt(2.0 * math.pi)


def shapes(n_in, width, depth, n_out):
    """(K, N) of every linear layer."""
    dims = [n_in] + [width] * depth + [n_out]
    return [(dims[i], dims[i + 1]) for i in range(depth + 1)]


def param_count(n_in, width, depth, n_out):
    return sum(k * n + n for k, n in shapes(n_in, width, depth, n_out))


def unpack(p, n_in, width, depth, n_out):
    """Flat parameters of one member -> [(W, b), ...]."""
    out, o = [], 0
    for k, n in shapes(n_in, width, depth, n_out):
        W = np.asarray(p[o:o + k * n], dtype=np.float64).reshape(k, n)
        o += k * n
        b = np.asarray(p[o:o + n], dtype=np.float64)
        o += n
        out.append((W, b))
    return out


def forward(p, x, n_in, width, depth, n_out):
    """Output y [rows][n_out] and the cache of pre-activations / activations."""
    layers = unpack(p, n_in, width, depth, n_out)
    a = np.asarray(x, dtype=np.float64)
    acts, pre = [a], []
    for l, (W, b) in enumerate(layers):
        z = a @ W + b
        if l < depth:
            pre.append(z)
            a = gelu(z)
            acts.append(a)
        else:
            a = z
    return a, (acts, pre)


def loss(p, x, t, n_in, width, depth, n_out):
    y, _ = forward(p, x, n_in, width, depth, n_out)
    d = y - np.asarray(t, dtype=np.float64)
    return float(np.mean(d * d))


def gradient(p, x, t, n_in, width, depth, n_out):
    """Loss and its gradient (flat, the parameter order) by backpropagation."""
    layers = unpack(p, n_in, width, depth, n_out)
    y, (acts, pre) = forward(p, x, n_in, width, depth, n_out)
    t = np.asarray(t, dtype=np.float64)
    rows = y.shape[0]
    d = y - t
    L = float(np.mean(d * d))
    G = 2.0 * d / (rows * n_out)
    grads = [None] * (depth + 1)
    for l in range(depth, -1, -1):
        W, _ = layers[l]
        grads[l] = (acts[l].T @ G, G.sum(axis=0))
        if l > 0:
            G = (G @ W.T) * gelu_grad(pre[l - 1])
    flat = np.concatenate([np.concatenate([gw.ravel(), gb]) for gw, gb in grads])
    return L, flat


def he_init(n_in, width, depth, n_out, rng):
    """N(0, 2 / fan_in) weights, zero biases (one member)."""
    parts = []
    for k, n in shapes(n_in, width, depth, n_out):
        parts.append(rng.normal(0.0, math.sqrt(2.0 / k), size=k * n))
        parts.append(np.zeros(n))
    return np.concatenate(parts)
