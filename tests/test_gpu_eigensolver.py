"""GPU parity of the batched dense eigensolver (dvr_syev_batch / dvr_diag).

The replaced routine is dsyev('V','L') behind SUBROUTINE diag
(source/dvr3drjz_segmented.f90:2638-2659).  LAPACK (numpy.linalg.eigh) is the
check: eigenvalues to a few ulps of ||A||, eigenvectors by residual and
orthogonality (their signs are arbitrary in both).
"""
import os

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


@pytest.fixture(params=["multi", "wg"], autouse=True)
def solver_kind(request, monkeypatch):
    """Run every case through both the multi-launch and the single-work-group
    solver."""
    monkeypatch.setenv("DVR_SYEV", request.param)
    return request.param


def _sym(n, rng, spread=1.0):
    a = rng.standard_normal((n, n)) * spread
    return 0.5 * (a + a.T)


def _check(a, w, z, nv, tol=1e-12):
    ref = np.linalg.eigvalsh(a)[:nv]
    nrm = max(np.abs(ref).max(), 1.0)
    assert np.abs(w - ref).max() <= tol * nrm * max(1, a.shape[0]) ** 0.5
    if z is not None and nv:
        r = a @ z - z * w[None, :]
        assert np.abs(r).max() <= 1e-11 * nrm * a.shape[0] ** 0.5
        o = z.T @ z - np.eye(nv)
        assert np.abs(o).max() <= 1e-10


def test_dvr_diag_closed_form():
    from dvr3d_amd import _lib
    # tridiagonal (2,-1): eigenvalues 2 - 2 cos(k pi/(n+1))
    n = 3
    a = np.diag([2.0] * n) + np.diag([-1.0] * (n - 1), -1)
    e, v = _lib.diag(a)
    assert np.allclose(e, [2 - np.sqrt(2), 2, 2 + np.sqrt(2)], atol=1e-14)
    n = 50
    a = np.diag([2.0] * n) + np.diag([-1.0] * (n - 1), -1)
    e, v = _lib.diag(a)
    k = np.arange(1, n + 1)
    assert np.abs(e - (2 - 2 * np.cos(k * np.pi / (n + 1)))).max() < 1e-13
    full = a + np.tril(a, -1).T
    assert np.abs(full @ v - v * e).max() < 1e-12
    # already tridiagonal past the LDS-resident tail (every reflector is the
    # identity, tau = 0, in both the column launches and the tail)
    n = 300
    a = np.diag([2.0] * n) + np.diag([-1.0] * (n - 1), -1)
    e, v = _lib.diag(a)
    k = np.arange(1, n + 1)
    assert np.abs(e - (2 - 2 * np.cos(k * np.pi / (n + 1)))).max() < 1e-13
    full = a + np.tril(a, -1).T
    assert np.abs(full @ v - v * e).max() < 1e-12


@pytest.mark.parametrize("sizes", [[1], [2], [3, 5], [17, 64, 1, 130], [257, 31],
                                   [128, 129, 127, 2], [700, 128, 129, 300]])
def test_syev_batch_random(sizes):
    from dvr3d_amd import _lib
    rng = np.random.default_rng(sum(sizes))
    mats = [_sym(n, rng) for n in sizes]
    ws, zs = _lib.syev_batch(mats)
    for a, w, z in zip(mats, ws, zs):
        _check(a, w, z, a.shape[0])


@pytest.mark.parametrize("tri", ["fused", "twolaunch"])
def test_syev_small_problems_column_launches(monkeypatch, tri):
    """Problems of at most 512 rows run one launch per column (every
    work-group repeats the column phase); DVR_TRI=twolaunch keeps the
    column and trailing launches apart.  Both agree with LAPACK, in a batch
    that mixes sizes below the LDS tail (100), across it (129, 300) and at
    the one-launch limit (512)."""
    from dvr3d_amd import _lib
    monkeypatch.setenv("DVR_TRI", tri)
    rng = np.random.default_rng(448)
    mats = [_sym(nn, rng) for nn in (448, 100, 129, 300, 512)]
    ws, zs = _lib.syev_batch(mats)
    for a, w, z in zip(mats, ws, zs):
        _check(a, w, z, a.shape[0])


@pytest.mark.parametrize("sizes", [[700], [1500, 600, 513, 100], [2203, 640, 129, 1]])
def test_syev_blocked_reduction(monkeypatch, sizes, solver_kind):
    """The blocked (panel) reduction of problems over 512 rows
    (DVR_TRI=blocked; the default for batches of at least 8 problems over
    1024 rows with one over 2048: the trailing square
    updated once per 32-column panel by an MFMA GEMM, read once per column):
    eigenpairs against LAPACK in batches that mix sizes above and below 512,
    whose LDS tails start inside different panels, and against the
    column-by-column reduction (DVR_TRI=column).  Multi-
    launch solver only: the single-work-group one has no column launches."""
    if solver_kind != "multi":
        pytest.skip("the blocked reduction belongs to the multi-launch solver")
    from dvr3d_amd import _lib
    rng = np.random.default_rng(sum(sizes) + 5)
    mats = [_sym(nn, rng) for nn in sizes]
    monkeypatch.setenv("DVR_TWOSTAGE_VEC", "0")    # eigenvectors from the one-stage reduction
    monkeypatch.setenv("DVR_TRI", "blocked")
    ws, zs = _lib.syev_batch(mats)
    for a, w, z in zip(mats, ws, zs):
        _check(a, w, z, a.shape[0])
    monkeypatch.setenv("DVR_TRI", "column")
    ws0, _ = _lib.syev_batch(mats)
    for w, w0, a in zip(ws, ws0, mats):
        assert np.abs(w - w0).max() <= 1e-12 * max(1.0, np.abs(w0).max()) * a.shape[0] ** 0.5


def test_syev_batch_partial_and_clusters():
    from dvr3d_amd import _lib
    rng = np.random.default_rng(7)
    n = 300
    q, _ = np.linalg.qr(rng.standard_normal((n, n)))
    lam = np.sort(rng.uniform(0.0, 1.0, n))
    lam[10] = lam[9] + 1e-13          # near-degenerate pair
    lam[20:23] = lam[19]              # exact triple degeneracy
    lam = np.sort(lam)
    a = (q * lam) @ q.T
    a = 0.5 * (a + a.T)
    ws, zs = _lib.syev_batch([a, a[:120, :120]], nvec=[40, 7])
    _check(a, ws[0], zs[0], 40, tol=2e-12)
    _check(a[:120, :120], ws[1], zs[1], 7, tol=2e-12)


@pytest.mark.parametrize("sizes,nvec", [([256], [256]), ([300, 280], [300, 50]), ([517], [200]),
                                        ([1027], [333]), ([2203], [100])])
def test_syev_twostage_vectors(monkeypatch, sizes, nvec, solver_kind):
    """Eigenvectors through the two-stage reduction (every problem of order
    >= 256): inverse iteration on the tridiagonal matrix, then back through
    the bulge chase (blocks of 64 sweeps on wavefronts) and the band
    reduction (three GEMMs per panel).  Against LAPACK (values, residuals,
    orthogonality), with all vectors of a 256 matrix and partial sets at
    sizes whose last panel is short (300, 517, 1027, 2203), and against the
    one-stage reduction's vectors (DVR_TWOSTAGE_VEC=0) up to sign."""
    if solver_kind != "multi":
        pytest.skip("the two-stage path does not depend on the one-stage solver kind")
    from dvr3d_amd import _lib
    rng = np.random.default_rng(sum(sizes) + 11)
    mats = [_sym(nn, rng) for nn in sizes]
    ws, zs = _lib.syev_batch(mats, nvec=nvec)
    for a, w, z, nv in zip(mats, ws, zs, nvec):
        _check(a, w, z, nv)
    monkeypatch.setenv("DVR_TWOSTAGE_VEC", "0")
    ws0, zs0 = _lib.syev_batch(mats, nvec=nvec)
    for a, w, z, w0, z0 in zip(mats, ws, zs, ws0, zs0):
        nrm = max(1.0, np.abs(w0).max())
        assert np.abs(w - w0).max() <= 1e-12 * nrm * a.shape[0] ** 0.5
        # well separated levels: the same vector up to sign
        gap = np.minimum(np.diff(np.r_[-np.inf, w0]), np.diff(np.r_[w0, np.inf]))
        sep = gap > 1e-4 * nrm
        assert sep.sum() >= len(w0) // 2
        ov = np.abs(np.sum(z * z0, axis=0))
        assert np.abs(ov[sep] - 1.0).max() <= 1e-9


def test_syev_twostage_vectors_clusters(solver_kind):
    """Two-stage eigenvectors of a 400 matrix with a near-degenerate pair, an
    exact triple and an already tridiagonal 300 matrix (every reflector the
    identity) among the wanted levels."""
    if solver_kind != "multi":
        pytest.skip("the two-stage path does not depend on the one-stage solver kind")
    from dvr3d_amd import _lib
    rng = np.random.default_rng(17)
    n = 400
    q, _ = np.linalg.qr(rng.standard_normal((n, n)))
    lam = np.sort(rng.uniform(0.0, 1.0, n))
    lam[10] = lam[9] + 1e-13
    lam[20:23] = lam[19]
    lam = np.sort(lam)
    a = (q * lam) @ q.T
    a = 0.5 * (a + a.T)
    t = np.diag([2.0] * 300) + np.diag([-1.0] * 299, -1) + np.diag([-1.0] * 299, 1)
    ws, zs = _lib.syev_batch([a, t], nvec=[40, 30])
    _check(a, ws[0], zs[0], 40, tol=2e-12)
    _check(t, ws[1], zs[1], 30, tol=2e-12)


def test_stein_wave_matches_thread(monkeypatch, solver_kind):
    """Inverse iteration with one wavefront per cluster (recurrences staged
    through LDS, DVR_STEIN=wave) against one thread per cluster
    (DVR_STEIN=thread) on a batch of problems over 853 rows reduced in one
    stage, one of them with a near-degenerate pair and an exact triple among
    the wanted levels (Gram-Schmidt inside the cluster): LAPACK checks for
    both, and the same vectors up to sign for separated levels."""
    if solver_kind != "multi":
        pytest.skip("inverse iteration of the multi-launch solver")
    from dvr3d_amd import _lib
    rng = np.random.default_rng(29)
    n = 1000
    q, _ = np.linalg.qr(rng.standard_normal((n, n)))
    lam = np.sort(rng.uniform(0.0, 1.0, n))
    lam[30] = lam[29] + 1e-13
    lam[50:53] = lam[49]
    lam = np.sort(lam)
    c = (q * lam) @ q.T
    mats = [_sym(1500, rng), 0.5 * (c + c.T), _sym(1027, rng)]
    nvec = [300, 80, 333]
    monkeypatch.setenv("DVR_TWOSTAGE_VEC", "0")
    res = {}
    for kind in ("wave", "thread"):
        monkeypatch.setenv("DVR_STEIN", kind)
        ws, zs = _lib.syev_batch(mats, nvec=nvec)
        for a, w, z, nv in zip(mats, ws, zs, nvec):
            _check(a, w, z, nv, tol=2e-12)
        res[kind] = (ws, zs)
    for w, z, w0, z0 in zip(*res["wave"], *res["thread"]):
        assert np.array_equal(w, w0)
        nrm = max(1.0, np.abs(w0).max())
        gap = np.minimum(np.diff(np.r_[-np.inf, w0]), np.diff(np.r_[w0, np.inf]))
        sep = gap > 1e-4 * nrm
        ov = np.abs(np.sum(z * z0, axis=0))
        assert np.abs(ov[sep] - 1.0).max() <= 1e-9
