"""Two-stage tridiagonalisation of the plain final solve (band reduction to
bandwidth 64 on MFMA, then bulge chasing), through dvr_syev_batch with
eigenvalues only and through the pipeline: eigenvalues against LAPACK
(numpy eigvalsh) and closed forms.  The routine it replaces is the dsyev of
diag3d (source/dvr3drjz_segmented.f90:2699 -> diag 2638-2659), whose DSYTRD
reduces in one stage (source/dsyev.f:6209)."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _sym(rng, n, spread=1.0):
    x = rng.standard_normal((n, n))
    return 0.5 * (x + x.T) + np.diag(np.linspace(0.0, spread * n, n))


@pytest.mark.parametrize("n", [256, 300, 513, 1000, 2049])
def test_twostage_random_matches_lapack(n):
    from dvr3d_amd import _lib
    rng = np.random.default_rng(n)
    a = _sym(rng, n)
    k = min(n, 200)
    w = _lib.syev_batch([a], [k], vectors=False)[0]
    ref = np.linalg.eigvalsh(a)[:k]
    scale = np.abs(np.linalg.eigvalsh(a)).max()
    err = np.abs(w - ref).max()
    print(f"n = {n}: max |dw| = {err:.2e} ({err / scale:.2e} of ||A||)")
    assert err <= 2e-13 * n ** 0.5 * scale


def test_twostage_clustered_spectrum():
    """Q diag(w) Q^T with triple-degenerate and near-degenerate eigenvalues."""
    from dvr3d_amd import _lib
    rng = np.random.default_rng(3)
    n = 700
    w0 = np.sort(np.concatenate([np.full(3, 0.5), 0.5 + 1e-9 * np.arange(4),
                                 rng.uniform(-2.0, 3.0, n - 7)]))
    q, _ = np.linalg.qr(rng.standard_normal((n, n)))
    a = (q * w0) @ q.T
    a = 0.5 * (a + a.T)
    w = _lib.syev_batch([a], [n], vectors=False)[0]
    assert np.abs(w - w0).max() <= 1e-12 * np.abs(w0).max() * 10


def test_twostage_tridiagonal_closed_form():
    """(2, -1) tridiagonal: every panel below the band is zero (skipped);
    eigenvalues 2 - 2 cos(k pi / (n + 1))."""
    from dvr3d_amd import _lib
    n = 400
    a = 2.0 * np.eye(n) - np.eye(n, k=1) - np.eye(n, k=-1)
    w = _lib.syev_batch([a], [n], vectors=False)[0]
    k = np.arange(1, n + 1)
    assert np.abs(w - (2.0 - 2.0 * np.cos(k * np.pi / (n + 1)))).max() <= 1e-13


def test_twostage_rank_deficient_panel_falls_back():
    """Columns 10..19 of the first panel zero below the band: Cholesky QR of
    that panel breaks down and the problem goes to the one-stage reduction;
    the eigenvalues are still right."""
    from dvr3d_amd import _lib
    rng = np.random.default_rng(5)
    n = 320
    a = _sym(rng, n)
    a[64:, 10:20] = 0.0
    a[10:20, 64:] = 0.0
    w = _lib.syev_batch([a], [50], vectors=False)[0]
    assert np.abs(w - np.linalg.eigvalsh(a)[:50]).max() <= 1e-12 * n


def test_twostage_batch_of_sizes():
    from dvr3d_amd import _lib
    rng = np.random.default_rng(11)
    sizes = [256, 333, 640]
    mats = [_sym(rng, n, 0.1) for n in sizes]
    ws = _lib.syev_batch(mats, [20, 333, 7], vectors=False)
    for a, w, k in zip(mats, ws, [20, 333, 7]):
        assert np.abs(w - np.linalg.eigvalsh(a)[:k]).max() <= 1e-12 * len(a)
