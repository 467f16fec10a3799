"""The DIPOLE3 restatement (oracle/dipole3_oracle.py) pinned against the
reference DIPOLE3 run on the wavefunction files of reference DVR3DRJZ and
ROTLEV3 runs (tests/golden/dip_ref.json: the reference's own H3+ J=0 -> J=1
test and three ABC pairings with a synthetic dipole), and against what the
mathematics fixes: Gauss-Legendre points, 3-j symbol identities, the
orthonormality of the associated Legendre functions under the quadrature and
the overlap limit of a constant dipole."""
import json
import os
import re
import tempfile

import numpy as np
import pytest

from oracle import dipole3_oracle as P
from tests import helpers as H
from tests.golden import make_dip_golden as G

GOLD = json.load(open(os.path.join(H.GOLD, "dip_ref.json")))
INPUT = {"h3p": P.DipInput(80, 8, 8, 0, 0, ezero=4361.60387517),
         "k01": P.DipInput(60), "k12": P.DipInput(60, 10, 0, 3, 0), "k22": P.DipInput(61)}


def synth_dipole(r1, r2, x, nu):
    """oracle/dipd_synth.f90 (the synthetic surface of the ABC cases)."""
    u1, u2 = r1 - 2.18, r2 - 3.20
    if nu == 0:
        return 0.35 + 0.12 * u1 - 0.08 * u2 + 0.21 * x + 0.05 * u1 * x
    return np.sqrt(np.maximum(0.0, 1.0 - x * x)) * (0.18 + 0.07 * u1 + 0.04 * u2 - 0.09 * x)


def table_dipole():
    """The reference's H3+ surface as tabulated on the h3p grid."""
    t = np.load(os.path.join(H.GOLD, "h3p_dipgrid.npz"))

    def f(r1, r2, x, nu):
        R2, R1, X = np.meshgrid(t["r2"], t["r1"], t["x"], indexing="ij")
        assert r1.shape == R1.shape and np.array_equal(r1, R1) and np.array_equal(r2, R2)
        assert np.array_equal(x, X)
        return t["v"][nu]
    return f


@pytest.fixture(scope="module")
def waves():
    if not G.built():
        pytest.skip("reference executables not built (oracle/Makefile)")
    root = tempfile.mkdtemp(prefix="dvr3d_dipt_")
    return root, {n: G.run_wave(n, os.path.join(root, n)) for n in G.RUNS}


def _files(waves, name):
    root, w = waves
    _, _, (kr, kf), (br, bf) = G.CASES[name]
    return P.read_wave(os.path.join(w[kr], kf)), P.read_wave(os.path.join(w[br], bf))


@pytest.mark.parametrize("name", ["h3p", "k01", "k12", "k22"])
def test_oracle_matches_reference_dipole3(waves, name):
    """Line strengths (the SPECTRA file's sint, full precision), energies and
    the block sequence against the reference DIPOLE3 on the same files."""
    ket, bra = _files(waves, name)
    g = GOLD[name]
    res = P.run(ket, bra, INPUT[name], table_dipole() if name == "h3p" else synth_dipole)
    sp = P.spect(res, INPUT[name].ezero)
    S = np.array(g["sint"])
    assert sp["sint"].shape == S.shape
    err = np.abs(sp["sint"] - S).max() / np.abs(S).max()
    print(f"{name}: {S.shape} line strengths, max |dS|/max S = {err:.1e}")
    assert err < 1e-12
    assert np.array_equal(res["e1"], g["e1"]) and np.array_equal(res["e2"], g["e2"])
    assert list(res["ibase"]) == g["ibase"] and res["zembed"] == g["zembed"]
    ref = [tuple(int(v) for v in re.findall(r"k[12] =\s*(\d+)", l)) for l in g["blocks"]]
    mine = [(p.b1 + 1 - ket.kmin, p.b2 + 1 - bra.kmin) for p in res["pairs"]]
    assert mine == ref


def test_gauss_legendre_points():
    """The Stroud-Secrest search gives the Gauss-Legendre rule: points to
    rounding, weights to the accuracy of the pre-final Newton derivative."""
    for n in (8, 40, 61 + 1, 80):
        x, w, csa, tsa = P.gauss_legendre(n)
        xr, wr = np.polynomial.legendre.leggauss(n)
        o = np.argsort(x)
        assert np.abs(x[o] - xr).max() < 5e-15
        assert np.abs(w[o] - wr).max() < 1e-10
        assert abs(csa - 1.0) < 1e-10 and tsa == 1.0


def test_threej_identities():
    """Closed form (j j 0; m -m 0) = (-1)^(j-m)/sqrt(2j+1), the parity zero
    of (j1 j2 j3; 0 0 0) and the orthogonality
    sum_{m1,m2} (2 j3 + 1) (j1 j2 j3; m1 m2 m3)(j1 j2 j3'; m1 m2 m3) = delta."""
    b = P.binomials(20)
    for j in range(5):
        for m in range(-j, j + 1):
            assert abs(P.threej(j, j, 0, m, -m, 0, b) - (-1) ** (j - m) / np.sqrt(2 * j + 1)) < 1e-14
    assert P.threej(1, 1, 1, 0, 0, 0, b) == 0.0 and P.threej(2, 1, 2, 0, 0, 0, b) == 0.0
    for j1 in range(4):
        for j2 in range(4):
            for j3 in range(abs(j1 - j2), j1 + j2 + 1):
                for j3p in range(abs(j1 - j2), j1 + j2 + 1):
                    for m3 in range(-min(j3, j3p), min(j3, j3p) + 1):
                        s = sum((2 * j3 + 1) * P.threej(j1, j2, j3, m1, -m1 - m3, m3, b)
                                * P.threej(j1, j2, j3p, m1, -m1 - m3, m3, b)
                                for m1 in range(-j1, j1 + 1))
                        assert abs(s - (j3 == j3p)) < 1e-13, (j1, j2, j3, j3p, m3)


def test_associated_legendre_orthonormal_under_quadrature():
    """sum_p w_p P_l^m(x_p) P_l'^m(x_p) = delta_ll' (60 points: exact up to
    degree 119), to the 1e-11 accuracy of the reference's weights (formed
    from the derivative before the last Newton step)."""
    x, w, _, _ = P.gauss_legendre(60)
    for m in (0, 1, 2, 3):
        pl = P.asleg(40, x, m)
        G_ = pl.T @ (w[:, None] * pl)
        assert np.abs(G_ - np.eye(41)).max() < 1e-10, m


def test_constant_dipole_gives_the_coefficient_overlap(waves):
    """With mu_z = c and mu_x = 0 the quadrature integrates the orthonormal
    angular functions exactly: tz of each nu = 0 block is c * factor times
    the overlap of the two files' FBR coefficients, tx = 0."""
    ket, bra = _files(waves, "k12")
    c = 0.731
    inp = P.DipInput(60, 10, 0, 3, 0)
    res = P.run(ket, bra, inp, lambda r1, r2, x, nu: np.full(r1.shape, c if nu == 0 else 0.0))
    tz = np.zeros_like(res["tz"])
    for p in res["pairs"]:
        if p.comp != 0:
            continue
        assert ket.nbass[p.b1] == bra.nbass[p.b2]
        c1 = P.fbr_block(ket, p.b1, 10, 3)
        c2 = P.fbr_block(bra, p.b2, bra.neval, 1)
        tz += c * p.factor * c1 @ c2.T
    assert np.abs(res["tx"]).max() == 0.0
    assert np.abs(res["tz"] - tz).max() < 1e-10 * np.abs(tz).max()


def test_selection_rules_and_unsupported_inputs_are_refused(waves):
    ket, bra = _files(waves, "k12")
    with pytest.raises(ValueError, match="selection rules"):
        P.run(ket, ket, P.DipInput(60), synth_dipole)          # J=1 e -> J=1 e
    j0, _ = _files(waves, "k01")
    with pytest.raises(ValueError, match="j = 0 -> 0"):
        P.run(j0, j0, P.DipInput(60), synth_dipole)
