"""The rotational step on the GPU (dvr_rotlev3, source/rotlev3.f90) against
the oracle on the same IWAVE input and against the reference ROTLEV3's runs
(tests/golden/rot_ref.json, rot_vec_ref.npz).

The IWAVE inputs are written on the box by the drop-in executable from the
reference's own H3+ J=1 deck (the reference's H3+ POTV given as a table,
kind 3 of the example POTV) and from the J=2 ABC deck k2s; their
vibrational levels and vectors match the reference DVR3DRJZ's to 1e-7
(tests/test_gpu_dropin.py), so the reference ROTLEV3 levels are matched to
1e-4 cm^-1 and the DIPOLE3 vectors to 1e-6."""
import json
import os
import shutil
import subprocess

import numpy as np
import pytest

from oracle import rotlev3_oracle as R
from tests import helpers as H

pytestmark = pytest.mark.gpu

EXE = os.path.join(H.REPO, "dvr3d_amd", "host", "dvr3drjz_amd")
INPUT = {"h3p": R.RotInput(nvib=400, neval=40, kmin=2, ibass=700, ezero=4362.208278,
                           ztran=True),
         "k2s": R.RotInput(nvib=30, neval=20, kmin=2, ibass=70, neval2=15, ztran=True)}


def _iwave(tmp, name):
    job = {"h3p": "h3p_J1", "k2s": "k2s"}[name]
    if name == "h3p":
        H.write_h3p_table(tmp, job)
    else:
        shutil.copy(os.path.join(H.JOBS, "abc_k2s.par"), tmp / "pes.par")
    shutil.copy(os.path.join(H.JOBS, job + ".job"), tmp / "in.job")
    with open(tmp / "in.job") as fi, open(tmp / "out.txt", "w") as fo:
        r = subprocess.run([EXE], stdin=fi, stdout=fo, stderr=subprocess.STDOUT, cwd=tmp,
                           timeout=300)
    assert r.returncode == 0, open(tmp / "out.txt").read()[-2000:]
    return R.read_iwave(tmp / "fort.26")


@pytest.fixture(scope="module", params=["k2s", "h3p"])
def case(request, tmp_path_factory):
    name = request.param
    w = _iwave(tmp_path_factory.mktemp(name), name)
    inp = INPUT[name]
    sel, rh, sets = R.run(w, inp)
    return name, w, inp, sel, rh, sets


def _device(w, inp, sel, nvec=1):
    two = inp.kmin > 1 and w.jrot > 1
    ib2 = sel.nbass - int(sel.mvib[0])
    dims = dict(nblk=sel.nblk, kfirst=0 if inp.kmin > 0 else 1, jrot=w.jrot, idia=w.idia,
                ipar=w.ipar, npnt1=w.npnt1, npnt2=w.npnt2, zembed=int(w.zembed),
                zquad2=int(w.zquad2), nlim=w.nlim, nsets=2 if two else 1,
                neval=(sel.neval, min(sel.neval2, ib2)), nvec=nvec)
    blocks = [dict(maxleg=b.maxleg, idvr=b.idvr, lincr=b.lincr, mvib=int(m))
              for b, m in zip(sel.blocks, sel.mvib)]
    vecs = [np.concatenate([np.frombuffer(b.vecs[i], "<f8") for i in range(int(m))])
            if m else np.zeros(0) for b, m in zip(sel.blocks, sel.mvib)]
    from dvr3d_amd import _lib
    return _lib.rotlev3(dims, blocks, [b.pleg for b in sel.blocks], vecs, sel.evib, w.rm2)


def _aligned(a, b):
    return np.abs(a - np.sign(np.sum(a * b, axis=0)) * b).max()


def test_device_matches_oracle_on_the_same_input(case):
    """Levels, eigenvectors and DSTORE vectors of both parity sets against
    the oracle run on the same IWAVE file."""
    name, w, inp, sel, rh, sets = case
    out, jn = _device(w, inp, sel)
    assert jn[:, 1].tolist() == rh.nang and jn[:, 0].tolist() == rh.jstart
    solved = [s for s in sets if s["kind"] == "solve"]
    assert len(out) == len(solved)
    for dev, ora in zip(out, solved):
        print(f"{name}: {len(dev['eval'])} levels, max |dE| = "
              f"{np.abs(dev['eval'] - ora['eval']).max() * R.AUTOCM:.2e} cm-1")
        assert np.abs(dev["eval"] - ora["eval"]).max() < 1e-11
        assert _aligned(dev["vec"], ora["vec"]) < 1e-8
        for a, b in zip(dev["d"], ora["d"]):
            assert a.shape == b.shape
            assert _aligned(a, b) < 1e-8


def test_device_matches_reference_rotlev3(case):
    """Against the reference ROTLEV3 run on the reference DVR3DRJZ's IWAVE:
    the selection, all levels of the solved sets within 1e-4 cm^-1 and the
    first two DIPOLE3 vectors of every k block within 1e-6."""
    name, w, inp, sel, rh, sets = case
    g = json.load(open(os.path.join(H.GOLD, "rot_ref.json")))[name]
    z = np.load(os.path.join(H.GOLD, "rot_vec_ref.npz"))
    assert sel.mvib.tolist() == g["mvib"]
    out, _ = _device(w, inp, sel)
    for s, dev in enumerate(out):
        ev = np.array(g["sets"][s]["eval_au"])
        assert np.abs(dev["eval"] - ev).max() * R.AUTOCM < 1e-4
        for b, d in enumerate(dev["d"]):
            ref = z[f"{name}_s{s}_d{b}"]
            for i in range(2):
                a = d[:, i]
                assert np.abs(a - np.sign(a @ ref[i]) * ref[i]).max() < 1e-6, (s, b, i)


def test_values_only_matches_the_vector_run(case):
    """nvec = 0 (levels only; the two-stage solver where the set is large
    enough) gives the levels of the run with vectors."""
    name, w, inp, sel, rh, sets = case
    a, _ = _device(w, inp, sel, nvec=1)
    b, _ = _device(w, inp, sel, nvec=0)
    for x, y in zip(a, b):
        assert np.abs(x["eval"] - y["eval"]).max() < 1e-11
