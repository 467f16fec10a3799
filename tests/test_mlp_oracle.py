"""Pins the CPU oracle of the learned preconditioner (oracle/mlp_oracle.py)
before it is used to check the GPU kernels: analytic gradients against
central finite differences, the loss gradient likewise, and the identity of
the untrained network with the classical Davidson preconditioner."""
import numpy as np

from oracle import mlp_oracle as M


def _random_params(seed):
    rng = np.random.default_rng(seed)
    p = M.init_params(seed)
    p["W3"] = rng.standard_normal(M.WIDTH) * 0.3
    p["b1"] = rng.standard_normal(M.WIDTH) * 0.1
    p["b2"] = rng.standard_normal(M.WIDTH) * 0.1
    return p


def test_backward_matches_finite_differences():
    rng = np.random.default_rng(1)
    p = _random_params(1)
    f = rng.standard_normal((7, M.NF))
    gw = rng.standard_normal(7)
    _, cache = M.forward(p, f)
    g = M.backward(p, cache, gw)
    h = 1e-6
    for name, idx in [("W1", (2, 5)), ("b1", (9,)), ("W2", (3, 77)), ("b2", (100,)),
                      ("W3", (11,)), ("b3", (0,))]:
        q = {k: v.copy() for k, v in p.items()}
        q[name][idx] += h
        wp, _ = M.forward(q, f)
        q[name][idx] -= 2 * h
        wm, _ = M.forward(q, f)
        fd = ((wp - wm) / (2 * h)) @ gw
        assert abs(fd - g[name][idx]) <= 1e-6 * max(1.0, abs(fd)), name


def test_loss_gradient_matches_finite_differences():
    rng = np.random.default_rng(2)
    n, s = 30, 3
    x = rng.standard_normal((n, s))
    x /= np.linalg.norm(x, axis=0)
    e = rng.standard_normal((n, s))
    e -= x * (x * e).sum(axis=0)
    t = rng.standard_normal((n, s))
    loss, gt = M.loss_and_grad_t(t, x, e)
    h = 1e-6
    for (i, c) in [(0, 0), (7, 1), (29, 2)]:
        tp = t.copy(); tp[i, c] += h
        tm = t.copy(); tm[i, c] -= h
        fd = (M.loss_and_grad_t(tp, x, e)[0][c] - M.loss_and_grad_t(tm, x, e)[0][c]) / (2 * h)
        assert abs(fd - gt[i, c]) <= 1e-6
    # a correction parallel to the error has zero loss
    l0, _ = M.loss_and_grad_t(e + 0.3 * x, x, e)
    assert np.abs(l0).max() < 1e-12


def test_untrained_network_is_classical_preconditioner():
    p = M.init_params(0)
    f = np.random.default_rng(3).standard_normal((50, M.NF))
    w, _ = M.forward(p, f)
    assert np.all(w == 1.0)
    d = np.array([0.5, 0.30005, 0.2])
    assert np.allclose(M.classical(d, 0.3), [1 / 0.2, 1 / M.DMIN, -1 / 0.1])


def test_flatten_roundtrip():
    p = _random_params(4)
    v = M.flatten(p)
    assert v.size == M.NPARAM == 17537
    q = M.unflatten(v)
    for k in p:
        assert np.allclose(q[k], p[k], atol=1e-6)


def test_device_layout_initialiser_matches_oracle_layout():
    """_lib.mlp_init lays the parameters out as the oracle's flatten() does
    (same shapes, W3 = 0, b3 = 1, zero biases) and with the oracle's fan-in
    scaling."""
    from dvr3d_amd import _lib
    p = _lib.mlp_init(3)
    assert p.dtype == np.float32 and p.size == M.NPARAM
    u = M.unflatten(p)
    assert np.all(u["b1"] == 0) and np.all(u["b2"] == 0) and np.all(u["W3"] == 0)
    assert u["b3"][0] == 1.0
    assert abs(u["W1"].std() * np.sqrt(M.NF) - 1.0) < 0.15
    assert abs(u["W2"].std() * np.sqrt(M.WIDTH) - 1.0) < 0.05
