"""dvr_syev_batch_noalloc (no device allocation; static 32 MiB workspace per
GPU) against dvr_syev_batch on the same problems: two calls in a row, the
second smaller, so that it runs on what the first left in the workspace."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _problems(rng, sizes):
    mats = []
    for n in sizes:
        x = rng.standard_normal((n, n))
        a = 0.5 * (x + x.T) + np.diag(np.linspace(0.0, 4.0 * n, n))
        mats.append(a)
    return mats


def _compare(mats, nvec):
    from dvr3d_amd import _lib
    w0, z0 = _lib.syev_batch(mats, nvec)
    w1, z1 = _lib.syev_batch(mats, nvec, noalloc=True)
    for p, a in enumerate(mats):
        scale = np.abs(np.linalg.eigvalsh(a)).max()
        assert len(w1[p]) == nvec[p]
        assert np.abs(w1[p] - w0[p]).max(initial=0.0) <= 1e-12 * scale, p
        for k in range(nvec[p]):
            s = np.sign(z0[p][:, k] @ z1[p][:, k])
            assert np.abs(z1[p][:, k] - s * z0[p][:, k]).max() <= 1e-9, (p, k)
        if nvec[p]:
            r = a @ z1[p] - z1[p] * w1[p]
            assert np.abs(r).max() <= 1e-11 * scale


def test_noalloc_matches_syev_batch_two_calls():
    rng = np.random.default_rng(7)
    sizes = [128] * 5 + [80] * 9 + [40, 17, 1, 128, 128, 3]
    mats = _problems(rng, sizes)
    nvec = [int(rng.integers(0, n + 1)) for n in sizes]
    nvec[0], nvec[5] = 128, 80
    _compare(mats, nvec)
    # second call, smaller: the workspace keeps the first call's contents
    sizes2 = [50, 50, 50, 7]
    mats2 = _problems(rng, sizes2)
    _compare(mats2, [50, 10, 0, 7])


def test_noalloc_rejects_oversized_requests():
    from dvr3d_amd import _lib
    rng = np.random.default_rng(1)
    # 40 problems of order 128 with all vectors need 40 * 131200 doubles > 4 Mi
    with pytest.raises(_lib.DvrError) as e:
        _lib.syev_batch(_problems(rng, [128] * 40), [128] * 40, noalloc=True)
    assert e.value.code == _lib.DVR_ERR_ARG
    with pytest.raises(_lib.DvrError) as e:
        _lib.syev_batch(_problems(rng, [129]), [1], noalloc=True)
    assert e.value.code == _lib.DVR_ERR_ARG
