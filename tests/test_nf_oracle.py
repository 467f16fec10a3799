"""The symmetrised-Radau restatement (oracle/nfmain_oracle.py, nfmain
D:3006-3646 of source/dvr3drjz_segmented.f90) against the reference
executable's runs (tests/golden/nf_ref.json, nf_iwave_ref.npz) and against
the direct-product DVR it contracts (no truncation)."""
import copy
import json
import os

import numpy as np
import pytest

from oracle import dvr3d_oracle as O
from oracle import nfmain_oracle as NF
from tests import helpers as H


def _gold():
    with open(os.path.join(H.GOLD, "nf_ref.json")) as f:
        return json.load(f)


def _water():
    job = O.load_job(os.path.join(H.JOBS, "water_J0.job"))
    v = np.load(os.path.join(H.GOLD, "water_vgrid.npz"))["v"]
    return job, v


def test_water_deck_matches_reference():
    """The reference's own IDIA = -2 water deck (J = 0, even), with the
    reference's water POTV tabulated on the grid: angles' 2D selections,
    nham3 and the 25 printed levels."""
    job, v = _water()
    g = _gold()["water_J0"]
    r = NF.run_nf(job, vgrid=v, keep_vectors=True)
    assert r["iv2"].tolist() == g["iv2"] and r["nham3"] == g["nham3"] == 500
    assert np.abs(r["eval"] - np.array(g["eval_au"])).max() * O.AUTOCM < 1e-6
    # transr's first wavefunction record (index: radial function outer,
    # angle inner), sign-aligned
    iw = np.load(os.path.join(H.GOLD, "nf_iwave_ref.npz"))
    assert iw["water_J0_iang_nbass"].tolist() == [r["iang"], r["iang"] * r["basis"].sizes.npair]
    assert iw["water_J0_iv2"].tolist() == r["iv2"].tolist()
    for k, key in enumerate(("water_J0_psi1", "water_J0_psi2")):
        a, b = r["wvfunc"][k], iw[key]
        assert np.abs(a - np.sign(a @ b) * b).max() < 1e-8


def test_synthetic_radau_deck_matches_reference():
    """c0r: the c0 grid and surface in symmetrised Radau coordinates."""
    job = O.load_job(os.path.join(H.JOBS, "c0r.job"))
    pes = O.PES.from_par(open(os.path.join(H.JOBS, "ab2_c0.par")).read())
    g = _gold()["c0r"]
    r = NF.run_nf(job, pes)
    assert r["iv2"].tolist() == g["iv2"] and r["nham3"] == g["nham3"] == 1000
    assert np.abs(r["eval"] - np.array(g["eval_au"])).max() * O.AUTOCM < 1e-6


def test_no_truncation_even_and_odd_make_the_direct_product():
    """Without truncation (zall) the even (ipar = 0) and odd (ipar = 1)
    symmetrised bases together span the direct-product DVR of both bonds and
    the angle: their levels are exactly its spectrum.  Pins the
    symmetrisation of blc2d1 (signs, 1/sqrt(2) of the r1 = r2 functions)
    and the angular coupling of bloc3d."""
    job = O.load_job(os.path.join(H.JOBS, "c0r.job"))
    job["npnt2"] = job["npnt1"] = 6
    job["nalf"] = 8
    job["nml"]["zall"] = True
    pes = O.PES.from_par(open(os.path.join(H.JOBS, "ab2_c0.par")).read())
    lev = []
    for ipar in (0, 1):
        j = copy.deepcopy(job)
        j["ipar"] = ipar
        j["neval"] = 10000
        lev.append(NF.run_nf(j, pes)["eval"])
    got = np.sort(np.concatenate(lev))
    b = NF.nf_basis(job)
    n, na = b.sizes.npnt, b.sizes.nalf
    mu = b.consts.ur1
    I, Ia = np.eye(n), np.eye(na)
    w = 0.5 / (b.r * b.r * mu)
    # |a, b, g> with g fastest: kinetic energy of both bonds, the angular
    # term (w_a + w_b) L and the potential on the diagonal
    Hd = (np.kron(np.kron(b.hr, I), Ia) + np.kron(np.kron(I, b.hr), Ia)
          + np.kron(np.kron(np.diag(w), I) + np.kron(I, np.diag(w)), b.htheta))
    A, B, G = np.meshgrid(np.arange(n), np.arange(n), np.arange(na), indexing="ij")
    Hd += np.diag(pes(b.r[A.ravel()], b.r[B.ravel()], b.xalf[G.ravel()]))
    ref = np.linalg.eigvalsh(Hd)
    assert len(got) == len(ref)
    assert np.abs(got - ref).max() < 1e-10
