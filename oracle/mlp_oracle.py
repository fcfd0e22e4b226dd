"""CPU restatement of the learned correction's network (float64 numpy).

TEST INFRASTRUCTURE ONLY: imported by tests/ as the checker of the HIP
forward, backward, Adam and BF16 inference kernels (``csrc/mlp.hip``). The
product path never imports it.

The network has no counterpart in the reference. Its arithmetic is fixed by
the definitions below (the same ones ``include/vpbs.h`` states):

    x' = (x − in_shift)·in_scale
    h_0 = x',  h_l = silu(W_{l-1} h_{l-1} + b_{l-1})  (l = 1..n_layers),  silu(z) = z/(1 + e^{−z})
    y' = W_L h_L + b_L,  y = y'·out_scale
    loss = (1/B) Σ_b Σ_c (y'_bc − t_bc/out_scale_c)²

with the parameters packed as [W_0, b_0, W_1, b_1, …, W_L, b_L], W_l row-major
[out][in]. Adam follows Kingma & Ba (2015), Algorithm 1, with the bias
corrections applied to the step. The backward pass here is pinned by central
finite differences (``tests/test_mlp.py``).
"""

from __future__ import annotations

from typing import List, Tuple

import numpy as np


def layer_dims(n_in: int, H: int, nl: int, n_out: int) -> List[Tuple[int, int]]:
    """(out, in) of each of the nl + 1 linear layers."""
    return [(H if l < nl else n_out, n_in if l == 0 else H) for l in range(nl + 1)]


def num_params(n_in, H, nl, n_out) -> int:
    return sum(o * i + o for o, i in layer_dims(n_in, H, nl, n_out))


def unpack(params, dims):
    out, off = [], 0
    for o, i in dims:
        W = params[off : off + o * i].reshape(o, i)
        off += o * i
        b = params[off : off + o]
        off += o
        out.append((W, b))
    return out


def init_params(n_in, H, nl, n_out, seed=0, n_used_in=None) -> np.ndarray:
    """He-uniform weights scaled for SiLU, zero biases; inputs past n_used_in get zero weights."""
    rng = np.random.default_rng(seed)
    dims = layer_dims(n_in, H, nl, n_out)
    parts = []
    for l, (o, i) in enumerate(dims):
        fan = n_used_in if (l == 0 and n_used_in) else i
        lim = np.sqrt(6.0 / fan) if l < nl else np.sqrt(3.0 / fan)
        W = rng.uniform(-lim, lim, size=(o, i))
        if l == 0 and n_used_in:
            W[:, n_used_in:] = 0.0
        parts += [W.ravel(), np.zeros(o)]
    return np.concatenate(parts)


def silu(z):
    return z / (1.0 + np.exp(-z))


def dsilu(z):
    s = 1.0 / (1.0 + np.exp(-z))
    return s * (1.0 + z * (1.0 - s))


def forward(params, dims, X, in_shift, in_scale, out_scale):
    """Returns y (physical units) and the normalised output y'."""
    h = (np.asarray(X, np.float64) - in_shift) * in_scale
    layers = unpack(np.asarray(params, np.float64), dims)
    for W, b in layers[:-1]:
        h = silu(h @ W.T + b)
    W, b = layers[-1]
    yn = h @ W.T + b
    return yn * out_scale, yn


def loss_and_grad(params, dims, X, T, in_shift, in_scale, out_scale, dsilu_fn=dsilu):
    """Mean loss over the B samples and its exact gradient (backpropagation)."""
    params = np.asarray(params, np.float64)
    layers = unpack(params, dims)
    h = (np.asarray(X, np.float64) - in_shift) * in_scale
    tn = np.asarray(T, np.float64) / out_scale
    B = h.shape[0]
    hs, zs = [h], []
    for W, b in layers[:-1]:
        z = h @ W.T + b
        zs.append(z)
        h = silu(z)
        hs.append(h)
    W, b = layers[-1]
    yn = h @ W.T + b
    d = yn - tn
    loss = float(np.sum(d * d) / B)
    dz = 2.0 * d / B
    grads = []
    for l in range(len(layers) - 1, -1, -1):
        W, _ = layers[l]
        grads.append((dz.T @ hs[l], dz.sum(axis=0)))
        if l == 0:
            break
        dz = (dz @ W) * dsilu_fn(zs[l - 1])
    grads.reverse()
    return loss, np.concatenate([np.concatenate([gW.ravel(), gb]) for gW, gb in grads])


def adam_step(params, grad, m, v, t, lr, b1=0.9, b2=0.999, eps=1e-8):
    """One Adam step (t counts from 1); returns the new (params, m, v)."""
    m = b1 * m + (1 - b1) * grad
    v = b2 * v + (1 - b2) * grad * grad
    mhat = m / (1 - b1 ** t)
    vhat = v / (1 - b2 ** t)
    return params - lr * mhat / (np.sqrt(vhat) + eps), m, v


def bf16(x):
    """Round to the nearest bfloat16 (ties to even), returned as float64."""
    a = np.asarray(x, np.float32)
    u = a.view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) & 0xFFFF0000
    return u.astype(np.uint32).view(np.float32).astype(np.float64)


def infer_bf16(params, dims, X, in_shift, in_scale, out_scale):
    """The inference arithmetic of the BF16 kernel: normalised inputs, weights and hidden
    activations rounded to bfloat16, products accumulated exactly (the kernel accumulates in
    FP32), biases and the output scale in FP32."""
    h = bf16((np.asarray(X, np.float32) - np.float32(in_shift)) * np.float32(in_scale))
    layers = unpack(np.asarray(params, np.float64), dims)
    for W, b in layers[:-1]:
        h = bf16(silu(h @ bf16(W).T + np.float32(b)))
    W, b = layers[-1]
    return (h @ bf16(W).T + np.float32(b)) * np.float32(out_scale)
