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
  w  = W3^T h2 + b3                             preconditioner weight
The correction of Ritz pair (theta, x) with residual r is the learned
diagonal preconditioner t_j = w_j r_j / (D_j - theta) (D = diag H, the
denominator clamped away from zero); with W3 = 0, b3 = 1 it is the classical
Davidson preconditioner.

Training loss per pair: 1 - cos^2 between the correction, projected out of the
Ritz vector x, and the true error e (x* - x, projected out of x).
"""
from __future__ import annotations

import numpy as np

NF = 6
WIDTH = 128
DELTA = 1.0e-3      # energy scale of the features (Eh)
DMIN = 1.0e-4       # smallest |D_j - theta| of the classical denominator (Eh)


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
        b3=np.ones(1),
    )


def forward(p, f):
    a1 = f @ p["W1"] + p["b1"]
    h1 = np.tanh(a1)
    a2 = h1 @ p["W2"] + p["b2"]
    h2 = np.tanh(a2)
    w = h2 @ p["W3"] + p["b3"][0]
    return w, dict(f=f, h1=h1, h2=h2, w=w)


def backward(p, cache, gw):
    """Gradients of sum(gw * w) with respect to every parameter."""
    f, h1, h2 = (cache[k] for k in ("f", "h1", "h2"))
    gz = gw
    g = dict(W3=h2.T @ gz, b3=np.array([gz.sum()]))
    ga2 = np.outer(gz, p["W3"]) * (1.0 - h2 * h2)
    g["W2"] = h1.T @ ga2
    g["b2"] = ga2.sum(axis=0)
    ga1 = (ga2 @ p["W2"].T) * (1.0 - h1 * h1)
    g["W1"] = f.T @ ga1
    g["b1"] = ga1.sum(axis=0)
    return g


def classical(diag, theta):
    """1 / (D - theta) with the denominator kept at least DMIN from zero."""
    den = diag - theta
    den = np.where(np.abs(den) < DMIN, np.copysign(DMIN, den), den)
    return 1.0 / den


def flatten(p):
    """Parameter layout of the device kernels (float32): W1 (NF x 128,
    row-major), b1, W2 (128 x 128, row-major: [k, o]), b2, W3, b3."""
    return np.concatenate([p["W1"].ravel(), p["b1"], p["W2"].ravel(), p["b2"], p["W3"],
                           p["b3"]]).astype(np.float32)


def unflatten(v):
    v = np.asarray(v, dtype=np.float64)
    o = 0
    out = {}
    for k, shp in (("W1", (NF, WIDTH)), ("b1", (WIDTH,)), ("W2", (WIDTH, WIDTH)),
                   ("b2", (WIDTH,)), ("W3", (WIDTH,)), ("b3", (1,))):
        n = int(np.prod(shp))
        out[k] = v[o:o + n].reshape(shp)
        o += n
    return out


NPARAM = NF * WIDTH + WIDTH + WIDTH * WIDTH + WIDTH + WIDTH + 1


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
