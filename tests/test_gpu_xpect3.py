"""The expectation-value step on the GPU (dvr_xpect3 and the drop-in
xpect3_amd, source/xpect3.f90) against the reference XPECT3's runs
(tests/golden/xp_ref.json) and against the oracle on the same input.

The wavefunction files are written on the box by the drop-in chain: the
drop-in DVR3DRJZ from the reference's own water J=0 deck (IDIA = -2 nfmain;
the reference's water POTV as a table), from c0 (IDIA = -1 jhmain), from
k2s and from the reference's H3+ J=1 deck, and the drop-in ROTLEV3 from the
J > 0 ones.  Their levels and vectors match the reference programs' to 1e-7
(tests/test_gpu_dropin.py, tests/test_gpu_nfmain.py,
tests/test_gpu_rotlev3.py); the expectation values are quadratic in them, so
the reference's are matched to 1e-6 of each column's largest value."""
import json
import os
import re
import shutil
import subprocess

import numpy as np
import pytest

from oracle import xpect3_oracle as X
from tests import helpers as H

pytestmark = pytest.mark.gpu

EXE = os.path.join(H.REPO, "dvr3d_amd", "host", "dvr3drjz_amd")
ROT = os.path.join(H.REPO, "dvr3d_amd", "host", "rotlev3_amd")
XP = os.path.join(H.REPO, "dvr3d_amd", "host", "xpect3_amd")
GOLD = json.load(open(os.path.join(H.GOLD, "xp_ref.json")))
RUNS = {"water": ("water_J0", None), "c0": ("c0", None), "k2s": ("k2s", "k2s_rot"),
        "h3pJ1": ("h3p_J1", "h3p_rot")}
CASES = {"water": ("water_xp0", ("water", "fort.26")), "c0": ("xp_synth", ("c0", "fort.26")),
         "k2s": ("xp_synth", ("k2s", "fort.8")), "k2f": ("xp_synth3", ("k2s", "fort.9")),
         "h3p": ("xp_h3p", ("h3pJ1", "fort.8"))}
INPUT = {"water": X.XpInput(38, 4, 29, 0, (1, 2, 3, 4), zfit=True),
         "c0": X.XpInput(0, 5, 5, 0, (1, 2, 3, 4, 5)),
         "k2s": X.XpInput(0, 5, 5, 0, (1, 2, 3, 4, 5)),
         "k2f": X.XpInput(0, 3, 5, 12, (1, 4, 5)),
         "h3p": X.XpInput(40, 5, 5, 0, (1, 2, 3, 4, 5))}


def _run(cmd, stdin, cwd):
    with open(stdin) as fi, open(os.path.join(cwd, "out.txt"), "w") as fo:
        r = subprocess.run([cmd], stdin=fi, stdout=fo, stderr=subprocess.STDOUT, cwd=cwd,
                           timeout=300)
    text = open(os.path.join(cwd, "out.txt")).read()
    assert r.returncode == 0, text[-3000:]
    return text


@pytest.fixture(scope="module")
def waves(tmp_path_factory):
    out = {}
    for name, (job, rot) in RUNS.items():
        d = tmp_path_factory.mktemp(name)
        if name == "water":
            H.write_water_potv_table(d)
        elif name.startswith("h3p"):
            H.write_h3p_table(d, job)
        else:
            par = "ab2_c0.par" if name == "c0" else "abc_k2s.par"
            shutil.copy(os.path.join(H.JOBS, par), d / "pes.par")
        print(f"drop-in chain: {job}", flush=True)
        _run(EXE, os.path.join(H.JOBS, job + ".job"), d)
        if rot:
            _run(ROT, os.path.join(H.JOBS, rot + ".job"), d)
        out[name] = d
    return out


def _table(text):
    rows = [l.split() for l in text.split("energy   property")[-1].splitlines()
            if re.match(r"^\s*\d+ [ab][12]\s", l)]
    return np.array([[float(v.replace("D", "E")) for v in r[3:]] for r in rows])


def _numbers(path):
    out = []
    for tok in open(path).read().replace("D", "E").split():
        try:
            out.append(float(tok))
        except ValueError:
            pass
    return np.array(out)


def _prop(name):
    if name == "water":
        from tests.test_xp_oracle import water_prop
        return water_prop()
    from tests.test_xp_oracle import synth_prop
    return synth_prop(INPUT[name].iprop)


@pytest.mark.parametrize("name", ["water", "c0", "k2s", "k2f", "h3p"])
def test_xpect3_executable_against_reference(tmp_path, waves, name):
    """The drop-in XPECT3 on the drop-in chain's files: the stdout table
    (energies from ILEV's zero, expectation values) and the ITRA file (and
    ITRA0 for the water fit run) against the reference XPECT3 on the
    reference chain's files."""
    job, (wr, wf) = CASES[name]
    shutil.copy(waves[wr] / wf, tmp_path / "fort.11")
    if wf == "fort.26":
        shutil.copy(waves[wr] / "fort.14", tmp_path / "fort.14")
    if name == "water":
        H.write_water_prop_table(tmp_path)
    else:
        (tmp_path / "prop.par").write_text("1\n")
    print(f"xpect3_amd: {name}", flush=True)
    text = _run(XP, os.path.join(H.JOBS, job + ".job"), tmp_path)
    g = GOLD[name]
    mine, ref = _table(text), np.array(g["table"])
    assert mine.shape == ref.shape
    assert np.abs(mine[:, 0] - ref[:, 0]).max() <= 2e-3           # f10.3 energies
    scale = np.abs(ref[:, 1:]).max(axis=0)
    scale[scale == 0.0] = 1.0                # columns that vanish by symmetry
    err = (np.abs(mine[:, 1:] - ref[:, 1:]) / scale).max()
    print(f"{name}: {ref.shape[0]} levels, max dev / column max = {err:.1e}")
    assert err < 1e-6
    a, b = _numbers(tmp_path / "fort.12"), np.array(g["itra"])
    assert a.shape == b.shape
    assert np.abs(a - b).max() <= 1e-6 * np.abs(b).max()
    if name == "water":
        c, d = _numbers(tmp_path / "fort.28"), np.array(g["itra0"])
        assert np.abs(c - d).max() <= 1e-6 * np.abs(d).max()


@pytest.mark.parametrize("name", ["water", "c0", "k2s", "k2f", "h3p"])
def test_device_matches_oracle_on_the_same_input(waves, name):
    """dvr_xpect3 through the C-ABI against the oracle's DSRD/REST/TRANS on
    the same file: every expectation value to 1e-12 of its column's largest."""
    from dvr3d_amd import _lib
    job, (wr, wf) = CASES[name]
    k = X.read_ket(waves[wr] / wf)
    prop = _prop(name)
    ref = X.run(k, INPUT[name], prop)["props"]
    npts, npot, ne, blocks, P = X.device_inputs(k, INPUT[name], prop)
    out = _lib.xpect3(npts, npot, ne, blocks, P)
    scale = np.abs(ref).max(axis=0)
    scale[scale == 0.0] = 1.0
    err = (np.abs(out - ref) / scale).max()
    print(f"{name}: {ref.shape}, max dev / column max = {err:.1e}")
    assert err < 1e-12
