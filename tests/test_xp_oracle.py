"""The XPECT3 restatement (oracle/xpect3_oracle.py) pinned against the
reference XPECT3 run on the wavefunction files of reference DVR3DRJZ and
ROTLEV3 runs (tests/golden/xp_ref.json: the reference's own water J=0 test
with its property routine, a J=0 jhmain run, and ROTLEV3 output of J=1
(IDIA=2) and J=2 (both parity sets)), and against what the mathematics
fixes: the quadrature rule, the norm of every wavefunction (property 1 = 1)
and the exchange symmetry of the homonuclear case."""
import json
import os
import tempfile

import numpy as np
import pytest

from oracle import xpect3_oracle as X
from tests import helpers as H
from tests.golden import make_xp_golden as G

GOLD = json.load(open(os.path.join(H.GOLD, "xp_ref.json")))
INPUT = {"water": X.XpInput(38, 4, 29, 0, (1, 2, 3, 4), zfit=True),
         "c0": X.XpInput(0, 5, 5, 0, (1, 2, 3, 4, 5)),
         "k2s": X.XpInput(0, 5, 5, 0, (1, 2, 3, 4, 5)),
         "k2f": X.XpInput(0, 3, 5, 12, (1, 4, 5)),
         "h3p": X.XpInput(40, 5, 5, 0, (1, 2, 3, 4, 5))}


def synth_prop(iprop):
    """oracle/prop_synth.f90."""
    def f(r1, r2, x):
        u1, u2 = r1 - 2.18, r2 - 3.20
        ft = [np.ones_like(x), u1, u2, x, u1 * u2 + 0.3 * x * x]
        return np.stack([ft[i - 1] for i in iprop], axis=-1)
    return f


def water_prop():
    """The reference's water property routine tabulated on the deck's grid."""
    t = np.load(os.path.join(H.GOLD, "water_propgrid.npz"))
    r, xa, v = t["r"], t["x"], t["v"]
    ro = np.argsort(r)
    xo = np.argsort(xa)

    def f(r1, r2, x):
        i1 = ro[np.searchsorted(r[ro], r1)]
        i2 = ro[np.searchsorted(r[ro], r2)]
        ig = xo[np.searchsorted(xa[xo], x)]
        assert np.array_equal(r[i1], r1) and np.array_equal(r[i2], r2) and np.array_equal(xa[ig], x)
        return v[i1, i2, ig]
    return f


def prop_of(name):
    return water_prop() if name == "water" else synth_prop(INPUT[name].iprop)


@pytest.fixture(scope="module")
def waves():
    if not G.built():
        pytest.skip("reference executables not built (oracle/Makefile)")
    root = tempfile.mkdtemp(prefix="dvr3d_xpt_")
    return {n: G.run_wave(n, os.path.join(root, n)) for n in G.RUNS}


def _ket(waves, name):
    _, _, (wr, wf) = G.CASES[name]
    return X.read_ket(os.path.join(waves[wr], wf))


@pytest.mark.parametrize("name", ["water", "c0", "k2s", "k2f", "h3p"])
def test_oracle_matches_reference_xpect3(waves, name):
    """Every expectation value of the stdout table (16 digits) against the
    reference XPECT3 on the same file; for the water fit run, the ground
    state values (ITRA0) and the differences from them."""
    k = _ket(waves, name)
    res = X.run(k, INPUT[name], prop_of(name))
    g = GOLD[name]
    ref = np.array(g["table"])[:, 1:]
    P = res["props"]
    if name == "water":
        assert np.abs(P[0] - np.array(g["itra0"])).max() <= 1e-14 * np.abs(P[0]).max()
        P = P[1:] - P[0]
    assert P.shape == ref.shape
    err = np.abs(P - ref).max() / np.abs(ref).max()
    print(f"{name}: {P.shape} expectation values, max rel dev {err:.1e}")
    assert err < 1e-13


def test_quadrature_is_gauss_legendre():
    for n in (22, 42, 62):
        x, w, csa, tsa = X.quadrature(n)
        xr, wr = np.polynomial.legendre.leggauss(n)
        assert np.abs(x - xr).max() < 5e-15 and np.abs(w - wr).max() < 1e-10
        assert abs(csa - 2.0) < 1e-10 and tsa == 2.0


@pytest.mark.parametrize("name", ["c0", "k2s", "h3p"])
def test_norm_and_exchange_symmetry(waves, name):
    """Property 1 (= 1) gives the norm of every level (1 to the quadrature's
    accuracy); for the homonuclear H3+ case the odd property x averages to
    exactly 0 under the exchange symmetrisation."""
    k = _ket(waves, name)
    res = X.run(k, INPUT[name], prop_of(name))
    assert np.abs(res["props"][:, 0] - 1.0).max() < 1e-11
    if name == "h3p":
        assert np.abs(res["props"][:, 3]).max() == 0.0


def test_jacobi_j0_files_are_refused(waves):
    """A J = 0 IWAVE file of IDIA > 0: the reference reads its energy-count
    record as the angular grid and aborts; the restatement refuses it."""
    d = tempfile.mkdtemp(prefix="dvr3d_xpk_")
    G.RUNS["k0s"] = ("dvr3drjz_ref", "k0s.job", "abc_k2s.par", None)
    try:
        G.run_wave("k0s", d)
    finally:
        del G.RUNS["k0s"]
    with pytest.raises(ValueError, match="IDIA > 0"):
        X.read_ket(os.path.join(d, "fort.26"))
