"""The ROTLEV3 restatement (oracle/rotlev3_oracle.py) pinned against the
reference ROTLEV3 run on the IWAVE files of reference DVR3DRJZ runs
(tests/golden/rot_ref.json, rot_vec_ref.npz: the reference's own H3+ J=1
rot.job and a J=2 ABC deck, both parities), and against properties the
mathematics fixes: the Coriolis-free limit, the orthogonality of the
DVR -> FBR transformation, energy ordering of the selection."""
import json
import os
import shutil
import subprocess

import numpy as np
import pytest

from oracle import rotlev3_oracle as R
from tests import helpers as H

REFD = os.path.join(H.REPO, "oracle", "_ref")
DECKS = {"h3p": ("dvr3drjz_h3p", "h3p_J1.job", None),
         "k2s": ("dvr3drjz_ref", "k2s.job", "abc_k2s.par")}
ROT_INPUT = {"h3p": R.RotInput(nvib=400, neval=40, kmin=2, ibass=700, ezero=4362.208278,
                               ztran=True),
             "k2s": R.RotInput(nvib=30, neval=20, kmin=2, ibass=70, neval2=15, ztran=True)}


def _iwave(tmp_path, name):
    exe, job, par = DECKS[name]
    if not os.path.exists(os.path.join(REFD, exe)):
        pytest.skip("reference executables not built (oracle/Makefile)")
    shutil.copy(os.path.join(H.JOBS, job), tmp_path / "dvr.job")
    if par:
        shutil.copy(os.path.join(H.JOBS, par), tmp_path / "pes.par")
    with open(tmp_path / "dvr.job") as fi:
        subprocess.run([os.path.join(REFD, exe)], stdin=fi, cwd=tmp_path,
                       stdout=subprocess.DEVNULL, check=True, timeout=900)
    return R.read_iwave(tmp_path / "fort.26")


@pytest.fixture(scope="module")
def k2s(tmp_path_factory):
    return _iwave(tmp_path_factory.mktemp("k2s"), "k2s")


@pytest.mark.parametrize("name", ["k2s", "h3p"])
def test_oracle_matches_reference_rotlev3(tmp_path, name):
    """Selection, levels of both parity sets and the DIPOLE3 vectors (KVEC)
    against the reference ROTLEV3 on the same IWAVE file."""
    w = _iwave(tmp_path, name)
    g = json.load(open(os.path.join(H.GOLD, "rot_ref.json")))[name]
    z = np.load(os.path.join(H.GOLD, "rot_vec_ref.npz"))
    sel, rh, sets = R.run(w, ROT_INPUT[name])
    assert sel.mvib.tolist() == g["mvib"]
    assert len(sets) == len(g["sets"]) == 2
    for s, (mine, ref) in enumerate(zip(sets, g["sets"])):
        ev = np.array(ref["eval_au"])
        assert len(mine["eval"]) == len(ev)
        # printed as D24.12: 12 digits after the point
        assert np.abs(mine["eval"] - ev).max() <= 6e-13, s
        assert np.abs(mine["eval"] - z[f"{name}_s{s}_keval"]).max() < 1e-14
        if mine["kind"] != "solve":
            continue
        for b, d in enumerate(mine["d"]):
            ref_d = z[f"{name}_s{s}_d{b}"]
            for i in range(2):
                a = d[:, i]
                assert np.abs(a - np.sign(a @ ref_d[i]) * ref_d[i]).max() < 1e-11, (s, b, i)
        v = mine["vec"]
        row = 0
        for b in range(int(z[f"{name}_s{s}_njblk"][0])):
            ref_c = z[f"{name}_s{s}_jc{b}"]
            m = ref_c.shape[1]
            for i in range(2):
                a = v[row:row + m, i]
                assert np.abs(a - np.sign(a @ ref_c[i]) * ref_c[i]).max() < 1e-10
            row += m


def test_coriolis_free_limit_gives_the_vibrational_levels(k2s):
    """With the r^-2 array zeroed there is no coupling between k blocks: the
    levels are the selected vibrational levels, sorted."""
    w = k2s
    w0 = R.Iwave(**{**w.__dict__, "rm2": np.zeros_like(w.rm2), "blocks": w.blocks})
    inp = R.RotInput(nvib=30, neval=40, kmin=2, ibass=70, ztran=True)
    sel, rh, sets = R.run(w0, inp)
    assert np.array_equal(sets[0]["eval"], np.sort(np.concatenate(sel.evib))[:40])
    _, rh1, _ = R.run(w, inp)
    # the coupled matrix differs only off the diagonal blocks
    m = np.concatenate([[0], np.cumsum(sel.mvib)])
    for b in range(len(sel.mvib)):
        s = slice(m[b], m[b + 1])
        assert np.array_equal(rh1.H[s, s], rh.H[s, s])
    assert np.abs(rh1.H - rh.H).max() > 1e-4


def test_dvr_to_fbr_transformation_is_orthogonal(k2s):
    """Every block of the J=2 run has nang = idvr: the DVR -> associated
    Legendre transformation is then orthogonal, so each vibrational vector
    keeps unit norm and the vectors stay orthonormal (a wrong row of the
    stored pleg*w array or a wrong jstart breaks this)."""
    w = k2s
    for k, b in enumerate(w.blocks):
        coef, jstart, nang = R.jtran(w, b, k, 10)
        assert nang == b.idvr and jstart == k
        G = coef.T @ coef
        assert np.abs(G - np.eye(10)).max() < 1e-12


def test_selection_is_by_energy(k2s):
    """IBASS = 50 of 90: every kept level lies at or below every dropped one,
    and on equal energies the lower k block wins."""
    inp = R.RotInput(nvib=30, neval=10, kmin=2, ibass=50)
    sel = R.select(k2s, inp)
    assert sel.mvib.sum() == 50 and sel.nbass == 50
    kept = np.concatenate([b.eval[:m] for b, m in zip(sel.blocks, sel.mvib)])
    dropped = np.concatenate([b.eval[m:30] for b, m in zip(sel.blocks, sel.mvib)])
    assert kept.max() <= dropped.min()
    w = R.Iwave(**{**k2s.__dict__})
    tie = [R.KBlock(**{**b.__dict__, "eval": np.arange(5, dtype=float)}) for b in k2s.blocks]
    w.blocks = tie
    sel = R.select(w, R.RotInput(nvib=5, neval=5, kmin=2, ibass=4))
    assert sel.mvib.tolist() == [2, 1, 1]
