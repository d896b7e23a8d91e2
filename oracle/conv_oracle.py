"""FP64 CPU reference of the learned psi preconditioner network.

TEST INFRASTRUCTURE ONLY (the checker for the BF16-MFMA conv kernels; never
called by the product path).  There is no reference counterpart: the network is
an addition of this build.  Forward and backward are written from the
definitions below; the backward is pinned by finite differences in
tests/test_oracle_kat.py.

Network (periodic 3x3 convolutions, channels 2 -> 16 -> 16 -> 16 -> 9):
    a0 = x                                  (B, 2, H, W)
    z_l = conv3(W_l, a_{l-1}) + b_l
    a_l = gelu(z_l)            l = 1, 2, 3
    y   = z_4                               (B, 9, H, W)   (3x3 stencil weights per pixel)
conv3(W, a)[b, o, i, j] = sum_{c, di, dj} W[o, c, di, dj] a[b, c, i+di-1, j+dj-1]  (indices mod H, W)
gelu(z) = z/2 (1 + erf(z / sqrt 2)).
"""
from __future__ import annotations

import math

import numpy as np
from scipy.special import erf

CHANNELS = (2, 16, 16, 16, 9)


def gelu(z):
    return 0.5 * z * (1.0 + erf(z / math.sqrt(2.0)))


def gelu_grad(z):
    return 0.5 * (1.0 + erf(z / math.sqrt(2.0))) + z * np.exp(-0.5 * z * z) / math.sqrt(2.0 * math.pi)


def conv3(W, a):
    """Periodic 3x3 convolution (cross-correlation), W: (O, C, 3, 3), a: (B, C, H, W)."""
    B, C, H, Wd = a.shape
    out = np.zeros((B, W.shape[0], H, Wd))
    for di in range(3):
        for dj in range(3):
            sh = np.roll(a, shift=(1 - di, 1 - dj), axis=(2, 3))   # sh[..., i, j] = a[..., i+di-1, j+dj-1]
            out += np.einsum("oc,bchw->bohw", W[:, :, di, dj], sh)
    return out


def conv3_grad_input(W, g):
    """d/da of sum(g * conv3(W, a)): correlation with the flipped, transposed kernel."""
    B, O, H, Wd = g.shape
    out = np.zeros((B, W.shape[1], H, Wd))
    for di in range(3):
        for dj in range(3):
            sh = np.roll(g, shift=(di - 1, dj - 1), axis=(2, 3))   # sh[..., i, j] = g[..., i-di+1, j-dj+1]
            out += np.einsum("oc,bohw->bchw", W[:, :, di, dj], sh)
    return out


def conv3_grad_weight(a, g):
    """d/dW of sum(g * conv3(W, a))."""
    O = g.shape[1]
    C = a.shape[1]
    dW = np.zeros((O, C, 3, 3))
    for di in range(3):
        for dj in range(3):
            sh = np.roll(a, shift=(1 - di, 1 - dj), axis=(2, 3))
            dW[:, :, di, dj] = np.einsum("bohw,bchw->oc", g, sh)
    return dW


def init_params(seed=0):
    rng = np.random.default_rng(seed)
    params = []
    for l in range(4):
        cin, cout = CHANNELS[l], CHANNELS[l + 1]
        std = math.sqrt(2.0 / (9 * cin))
        params.append((rng.standard_normal((cout, cin, 3, 3)) * std, np.zeros(cout)))
    return params


def forward(params, x):
    acts = [x]
    pre = []
    a = x
    for l, (W, b) in enumerate(params):
        z = conv3(W, a) + b[None, :, None, None]
        pre.append(z)
        a = gelu(z) if l < 3 else z
        acts.append(a)
    return a, (acts, pre)


def backward(params, cache, gy):
    """Gradients of sum(gy * y) w.r.t. every weight and bias, and w.r.t. the input."""
    acts, pre = cache
    grads = [None] * 4
    g = gy
    for l in range(3, -1, -1):
        if l < 3:
            g = g * gelu_grad(pre[l])
        W, _ = params[l]
        grads[l] = (conv3_grad_weight(acts[l], g), g.sum(axis=(0, 2, 3)))
        g = conv3_grad_input(W, g)
    return grads, g


def mse_loss(params, x, target):
    y, cache = forward(params, x)
    d = y - target
    n = d.size
    loss = float(np.sum(d * d) / n)
    grads, _ = backward(params, cache, 2.0 * d / n)
    return loss, grads
