"""Host side of the in-process SFIT fit (dvr3d_amd/sfit.py): the parameter
derivatives of the AB2 surface, which are the properties whose expectation
values form the fit's Jacobian, against central finite differences of the
surface itself."""
import numpy as np

from dvr3d_amd.sfit import AB2Surface


def test_ab2_parameter_derivatives_match_finite_differences():
    p = np.array([0.20, 1.10, 1.81, 0.080, np.cos(np.radians(104.5)), 0.010])
    rng = np.random.default_rng(7)
    r1 = rng.uniform(1.3, 2.8, 50)
    r2 = rng.uniform(1.3, 2.8, 50)
    x = rng.uniform(-0.99, 0.99, 50)
    g = AB2Surface.grad(p, r1, r2, x)
    for j in range(len(p)):
        h = 1e-6 * max(1.0, abs(p[j]))
        pp, pm = p.copy(), p.copy()
        pp[j] += h
        pm[j] -= h
        fd = (AB2Surface.value(pp, r1, r2, x) - AB2Surface.value(pm, r1, r2, x)) / (2 * h)
        assert np.abs(fd - g[:, j]).max() <= 1e-7 * max(1.0, np.abs(fd).max()), AB2Surface.names[j]
