"""CPU oracle of the learned Davidson preconditioner (test infrastructure).

This module is only imported by tests: it restates, in float64 numpy, the
computation that the HIP kernels of dvr3d_amd/csrc/k_mlp.hip perform in FP32
(training) and BF16/FP32 (inference), so that forward values and gradients can
be checked.  The reference program has no counterpart: its final solve is the
dense dsyev call (source/dvr3drjz_segmented.f90:2699 -> 2638-2659); the
learned part belongs to the block-Davidson replacement of that call.

Network (per basis function j of the contracted 3D basis, per Ritz pair s):
  f = features(j, s)                            (NF = 6)
  h1 = tanh(W1^T f + b1)                        (128)
  h2 = tanh(W2^T h1 + b2)                       (128)
  w  = exp(clip(W3^T h2 + b3, -ZMAX, ZMAX))      preconditioner weight
The correction of Ritz pair s is t = (H - sigma I)^-1 (w o r).

Training loss per pair: 1 - cos^2 between the correction, projected out of the
Ritz vector x, and the true error e (x* - x, projected out of x).
"""
from __future__ import annotations

import numpy as np

NF = 6
WIDTH = 128
ZMAX = 4.0
DELTA = 1.0e-3      # energy scale of the features (Eh)


def features(diag, e2, ir1, n2d, nr1, nmax, theta, sigma):
    """Feature matrix (N, NF) for one Ritz pair.

    diag: H3 diagonal; e2: parent 2D eigenvalue of each contracted function;
    ir1: r1 grid index of its block; n2d: its index among the block's chosen
    2D functions; nr1: number of r1 points; nmax: largest block size."""
    a = np.arcsinh
    return np.stack([
        a((diag - theta) / DELTA),
        a((diag - sigma) / DELTA),
        a((e2 - theta) / DELTA),
        ir1 / max(nr1 - 1, 1),
        n2d / max(nmax - 1, 1),
        np.full_like(diag, a((sigma - theta) / DELTA)),
    ], axis=1)


def init_params(seed=0):
    rng = np.random.default_rng(seed)
    return dict(
        W1=rng.standard_normal((NF, WIDTH)) / np.sqrt(NF),
        b1=np.zeros(WIDTH),
        W2=rng.standard_normal((WIDTH, WIDTH)) / np.sqrt(WIDTH),
        b2=np.zeros(WIDTH),
        W3=np.zeros(WIDTH),
        b3=np.zeros(1),
    )


def forward(p, f):
    a1 = f @ p["W1"] + p["b1"]
    h1 = np.tanh(a1)
    a2 = h1 @ p["W2"] + p["b2"]
    h2 = np.tanh(a2)
    z = h2 @ p["W3"] + p["b3"][0]
    zc = np.clip(z, -ZMAX, ZMAX)
    w = np.exp(zc)
    return w, dict(f=f, h1=h1, h2=h2, z=z, w=w)


def backward(p, cache, gw):
    """Gradients of sum(gw * w) with respect to every parameter."""
    f, h1, h2, z, w = (cache[k] for k in ("f", "h1", "h2", "z", "w"))
    gz = gw * w * ((z > -ZMAX) & (z < ZMAX))
    g = dict(W3=h2.T @ gz, b3=np.array([gz.sum()]))
    ga2 = np.outer(gz, p["W3"]) * (1.0 - h2 * h2)
    g["W2"] = h1.T @ ga2
    g["b2"] = ga2.sum(axis=0)
    ga1 = (ga2 @ p["W2"].T) * (1.0 - h1 * h1)
    g["W1"] = f.T @ ga1
    g["b1"] = ga1.sum(axis=0)
    return g


def loss_and_grad_t(t, x, e):
    """Per-column 1 - cos^2(P t, e), P = I - x x^T (e is orthogonal to x);
    returns (loss per column, dL/dt)."""
    tp = t - x * (x * t).sum(axis=0)
    c = (tp * e).sum(axis=0)
    tt = (tp * tp).sum(axis=0)
    ee = (e * e).sum(axis=0)
    loss = 1.0 - c * c / (tt * ee)
    gtp = -2.0 * c / (tt * ee) * e + 2.0 * c * c / (tt * tt * ee) * tp
    gt = gtp - x * (x * gtp).sum(axis=0)
    return loss, gt


def adam(p, g, m, v, step, lr, b1=0.9, b2=0.999, eps=1e-8):
    for k in p:
        m[k] = b1 * m[k] + (1 - b1) * g[k]
        v[k] = b2 * v[k] + (1 - b2) * g[k] * g[k]
        mh = m[k] / (1 - b1 ** step)
        vh = v[k] / (1 - b2 ** step)
        p[k] = p[k] - lr * mh / (np.sqrt(vh) + eps)
