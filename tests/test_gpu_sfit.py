"""The in-process SFIT fit (dvr3d_amd/sfit.py; SFIT_DVR3D/H2S/dvr3d_sfit.f90)
on the GPU: its Jacobian, Hellmann-Feynman expectation values of the
parameter derivatives of the surface computed by dvr_xpect3 on the DVR
grid, against central finite differences of the term values computed by the
full pipeline; and the fit itself, which must recover the parameters that
produced a set of term values from a displaced start."""
import os

import numpy as np
import pytest

from tests import helpers as H

pytestmark = pytest.mark.gpu

TRUE = np.array([0.20, 1.10, 1.81, 0.080, np.cos(np.radians(104.5)), 0.010])


def _fit(params, ivar, obs=None):
    from dvr3d_amd import sfit
    return sfit.VibFit(os.path.join(H.JOBS, "c0.job"), sfit.AB2Surface, params, ivar,
                       obs or {n: 0.0 for n in range(1, 16)})


def test_hellmann_feynman_jacobian_matches_finite_differences():
    from dvr3d_amd import sfit
    f = _fit(TRUE, [True] * 6)
    ev, D = f.levels_and_derivatives(TRUE)
    n = np.arange(1, 11)
    for j in (0, 1, 3, 5):                   # De, a, kb, c
        h = 1e-4 * abs(TRUE[j])
        pp, pm = TRUE.copy(), TRUE.copy()
        pp[j] += h
        pm[j] -= h
        ep, _ = f.levels_and_derivatives(pp)
        em, _ = f.levels_and_derivatives(pm)
        fd = ((ep[n] - ep[0]) - (em[n] - em[0])) / (2 * h)
        hf = D[n, j] - D[0, j]
        err = np.abs(fd - hf).max() / np.abs(fd).max()
        print(f"{sfit.AB2Surface.names[j]}: |HF - FD| / max |FD| = {err:.1e}")
        assert err < 1e-2


def test_fit_recovers_the_parameters_from_a_displaced_start():
    f0 = _fit(TRUE, [True] * 6)
    ev, _ = f0.levels_and_derivatives(TRUE)
    from dvr3d_amd import sfit
    obs = {int(n): float((ev[n] - ev[0]) * sfit.AUTOCM) for n in range(1, 16)}
    start = TRUE.copy()
    start[0] *= 1.004
    start[3] *= 0.995
    start[5] *= 1.01
    ivar = [True, False, False, True, False, True]
    f = _fit(start, ivar, obs)
    hist = f.fit(itmax=8)
    for i, s in enumerate(hist):
        print(f"iteration {i}: rms {s.rms:.3e} cm-1, stadev {s.stadev:.3e}")
    rel = np.abs(f.params - TRUE) / np.abs(TRUE)
    print("relative parameter errors:", rel)
    assert hist[0].rms > 1.0
    assert hist[-1].rms < 1e-4 and rel.max() < 1e-6
