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
import re
import shutil
import subprocess

import numpy as np
import pytest

from oracle import rotlev3_oracle as R
from tests import helpers as H

pytestmark = pytest.mark.gpu

EXE = os.path.join(H.REPO, "dvr3d_amd", "host", "dvr3drjz_amd")
ROT = os.path.join(H.REPO, "dvr3d_amd", "host", "rotlev3_amd")
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
    tmp = tmp_path_factory.mktemp(name)
    w = _iwave(tmp, name)
    inp = INPUT[name]
    sel, rh, sets = R.run(w, inp)
    DIRS[name] = tmp
    return name, w, inp, sel, rh, sets


DIRS = {}   # the IWAVE directory of each case


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


def _report_levels(text):
    out = []
    for part in text.split("EIGENVALUES IN HARTREES")[1:]:
        vals = []
        for line in part.split("\n")[1:]:
            if "wavenumbers" in line:
                break
            vals += [float(x.replace("D", "E")) for x in line.split()]
        out.append(np.array(vals))
    return out


def test_rotlev3_executable_against_reference(case):
    """The drop-in rotlev3_amd on the reference's own H3+ rot.job and the
    J=2 deck: stdout levels of both parity sets against the reference run,
    the KVEC files (header records byte for byte, levels, the first two
    DIPOLE3 vectors of every k block sign-aligned), the JVEC headers and
    the magnitudes of the JVEC coefficients (the sign of each vibrational
    function is arbitrary in both programs), the J=1f copy of DSTORE1, and
    the ILEV records appended after DVR3DRJZ's."""
    from tests import fortran_io as F
    name, w, inp, sel, rh, sets = case
    d = DIRS[name]
    g = json.load(open(os.path.join(H.GOLD, "rot_ref.json")))[name]
    z = np.load(os.path.join(H.GOLD, "rot_vec_ref.npz"))
    deck = {"h3p": "h3p_rot.job", "k2s": "k2s_rot.job"}[name]
    text = open(os.path.join(H.JOBS, deck)).read().replace("zdcore=.true.,",
                                                           "zdcore=.true., zpfun=.true.,")
    (d / "rot.job").write_text(text)
    for f in ("fort.14", "fort.8", "fort.9", "fort.3", "fort.2"):
        if (d / f).exists():
            (d / f).unlink()
    (d / "fort.14").write_text("dvr3drjz levels\n")
    with open(d / "rot.job") as fi, open(d / "rot.out", "w") as fo:
        r = subprocess.run([ROT], stdin=fi, stdout=fo, stderr=subprocess.STDOUT, cwd=d,
                           timeout=300)
    out = open(d / "rot.out").read()
    assert r.returncode == 0, out[-3000:]
    lev = _report_levels(out)
    assert len(lev) == 2
    for s_, (mine, ref) in enumerate(zip(lev, g["sets"])):
        ref = np.array(ref["eval_au"])
        assert len(mine) == len(ref)
        assert np.abs(mine - ref).max() * R.AUTOCM < 1e-4, s_
    assert [int(b) - int(a) + 1 for a, b in
            re.findall(r"I RUNS FROM\s+(\d+) TO\s+(\d+)", out)] == g["mvib"]
    for s_, kf in enumerate(("fort.8", "fort.9")):
        key = f"{name}_s{s_}"
        recs = F.read_records(d / kf)
        nrec, nh = (int(x) for x in z[f"{key}_nrec"])
        assert len(recs) == nrec
        j1f = nh == 11
        for i in range(nh - 1):
            a, b = np.frombuffer(recs[i], np.uint8), z[f"{key}_k{i + 1}"]
            assert len(a) == len(b), (kf, i)
            if j1f and i in (6, 8):    # the k = 1 block's angular points and pleg*w
                x, y = np.frombuffer(a.tobytes(), "<f8"), np.frombuffer(b.tobytes(), "<f8")
                assert np.abs(x - y).max() < 1e-13, (kf, i)
            else:
                assert np.array_equal(a, b), (kf, i)
        ev = F.doubles(recs[nh - 1])
        assert np.abs(ev - z[f"{key}_keval"]).max() * R.AUTOCM < 1e-4
        nval = len(ev)
        for b in range(int(z[f"{key}_ndblk"][0])):
            ref = z[f"{key}_d{b}"]
            if j1f:
                mine = np.stack([F.doubles(recs[nh]), F.doubles(recs[nh + 1])])
            else:
                mine = F.doubles(recs[nh + b]).reshape(nval, -1)[:2]
            for i in range(2):
                x = mine[i]
                assert np.abs(x - np.sign(x @ ref[i]) * ref[i]).max() < 1e-6, (kf, b, i)
        if j1f:
            continue
        jr = F.read_records(d / ("fort.3" if s_ == 0 else "fort.2"))
        assert np.array_equal(F.ints(jr[0]), z[f"{key}_j1"])
        assert np.array_equal(F.ints(jr[1]), z[f"{key}_j2"])
        assert np.abs(F.doubles(jr[2]) - z[f"{key}_jeval"]).max() * R.AUTOCM < 1e-4
        for b in range(int(z[f"{key}_njblk"][0])):
            c = F.doubles(jr[3 + b]).reshape(nval, -1)[:2]
            assert np.abs(np.abs(c) - np.abs(z[f"{key}_jc{b}"])).max() < 1e-6
    ilev = open(d / "fort.14").read().splitlines()
    assert ilev[0] == "dvr3drjz levels"
    heads = [l for l in ilev[1:] if len(l.split()) == 7]
    assert len(heads) == 2
    h0 = [int(x) for x in heads[0].split()]
    assert h0 == [w.jrot, 0, w.idia, w.ipar, 0, len(g["sets"][0]["eval_au"]), sel.nbass]
