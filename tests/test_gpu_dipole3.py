"""The line-strength step on the GPU (dvr_dipole3 and the drop-in
dipole3_amd, source/dipole3_segmented.f90) against the reference DIPOLE3's
runs (tests/golden/dip_ref.json) and against the oracle on the same input.

The wavefunction files are written on the box by the drop-in chain: the
drop-in DVR3DRJZ from the reference's own H3+ J=0 and J=1 decks (the
reference's H3+ POTV given as a table) and from the ABC decks k0s, k1s and
k2s, and the drop-in ROTLEV3 from the J > 0 ones.  Their levels and vectors
match the reference programs' to 1e-7 (tests/test_gpu_dropin.py,
tests/test_gpu_rotlev3.py), so the reference's line strengths are matched to
1e-6 of the largest one; on the same files the device agrees with the
oracle to 1e-10."""
import json
import os
import re
import shutil
import subprocess

import numpy as np
import pytest

from oracle import dipole3_oracle as P
from tests import helpers as H

pytestmark = pytest.mark.gpu

EXE = os.path.join(H.REPO, "dvr3d_amd", "host", "dvr3drjz_amd")
ROT = os.path.join(H.REPO, "dvr3d_amd", "host", "rotlev3_amd")
DIP = os.path.join(H.REPO, "dvr3d_amd", "host", "dipole3_amd")
GOLD = json.load(open(os.path.join(H.GOLD, "dip_ref.json")))
# wavefunction runs: deck, ROTLEV3 deck
RUNS = {"h3pJ0": ("h3p_J0", None), "h3pJ1": ("h3p_J1", "h3p_rot"), "k0s": ("k0s", None),
        "k1s": ("k1s", "k1s_rot"), "k2s": ("k2s", "k2s_rot")}
CASES = {"h3p": ("h3p_dip", ("h3pJ0", "fort.26"), ("h3pJ1", "fort.8")),
         "k01": ("dip_k01", ("k0s", "fort.26"), ("k1s", "fort.8")),
         "k12": ("dip_k12", ("k1s", "fort.8"), ("k2s", "fort.8")),
         "k22": ("dip_k22", ("k2s", "fort.8"), ("k2s", "fort.9"))}
INPUT = {"h3p": P.DipInput(80, 8, 8, 0, 0, ezero=4361.60387517),
         "k01": P.DipInput(60), "k12": P.DipInput(60, 10, 0, 3, 0), "k22": P.DipInput(61)}


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
        if name.startswith("h3p"):
            H.write_h3p_table(d, job)
        else:
            shutil.copy(os.path.join(H.JOBS, "abc_k2s.par"), d / "pes.par")
        print(f"drop-in chain: {job}", flush=True)
        _run(EXE, os.path.join(H.JOBS, job + ".job"), d)
        if rot:
            _run(ROT, os.path.join(H.JOBS, rot + ".job"), d)
        out[name] = d
    return out


def _dipole_table(d):
    """dip.par (kind 3 of the example DIPD) and its table: the reference's
    H3+ surface on the h3p integration grid (tests/golden/h3p_dipgrid.npz)."""
    t = np.load(os.path.join(H.GOLD, "h3p_dipgrid.npz"))
    R2, R1, X = np.meshgrid(t["r2"], t["r1"], t["x"], indexing="ij")
    rows = [np.stack([R1.ravel(), R2.ravel(), X.ravel(), np.full(R1.size, float(nu)),
                      t["v"][nu].ravel()], axis=1) for nu in (0, 1)]
    tab = np.concatenate(rows)
    with open(os.path.join(d, "dtable.bin"), "wb") as f:
        np.array([len(tab)], dtype="<f8").tofile(f)
        tab.astype("<f8").tofile(f)
    with open(os.path.join(d, "dip.par"), "w") as f:
        f.write("3\n" + os.path.join(str(d), "dtable.bin") + "\n")


def _case_dir(tmp_path, waves, name):
    job, (kr, kf), (br, bf) = CASES[name]
    shutil.copy(waves[kr] / kf, tmp_path / "fort.11")
    shutil.copy(waves[br] / bf, tmp_path / "fort.12")
    if name == "h3p":
        _dipole_table(tmp_path)
    else:
        (tmp_path / "dip.par").write_text("1\n")
    return os.path.join(H.JOBS, job + ".job")


def _spectra(path):
    recs = P.read_records(path)
    ints = np.frombuffer(recs[0][:36], "<i4").tolist()
    tail = np.frombuffer(recs[0][44:56], "<i4").tolist()
    S = np.stack([P._d(r) for r in recs[3:]], axis=1)
    return ints, bool(tail[0]), tail[1:], P._d(recs[1]), P._d(recs[2]), S


@pytest.mark.parametrize("name", ["h3p", "k01", "k12", "k22"])
def test_dipole3_executable_against_reference(tmp_path, waves, name):
    """The drop-in DIPOLE3 on the drop-in chain's files: the SPECTRA record
    header, energies, block lines and line strengths against the reference
    DIPOLE3 run on the reference chain's files."""
    job = _case_dir(tmp_path, waves, name)
    print(f"dipole3_amd: {name}", flush=True)
    text = _run(DIP, job, tmp_path)
    g = GOLD[name]
    ints, zemb, ib, e1, e2, S = _spectra(tmp_path / "fort.13")
    assert ints == g["header"] and zemb == g["zembed"] and ib == g["ibase"]
    assert np.abs(e1 - g["e1"]).max() * P.AUTOCM < 1e-4
    assert np.abs(e2 - g["e2"]).max() * P.AUTOCM < 1e-4
    Sr = np.array(g["sint"])
    err = np.abs(S - Sr).max() / np.abs(Sr).max()
    print(f"{name}: {Sr.shape} line strengths, max |dS|/max S = {err:.1e}")
    assert err < 1e-6
    assert re.findall(r"Block\s+\d+ k1 =\s*\d+ to k2 =\s*\d+ (?:completed|skipped)", text) == \
        g["blocks"]
    # the printed table: one row per pair, S(f-i) in debye^2 consistent with sint
    rows = [l.split() for l in text.split("a-coefficient")[-1].splitlines()
            if re.match(r"^\s*\d+\s+\d+\s", l)]
    assert len(rows) == S.size


@pytest.mark.parametrize("name", ["h3p", "k01", "k12", "k22"])
def test_device_matches_oracle_on_the_same_input(tmp_path, waves, name):
    """dvr_dipole3 through the C-ABI against the oracle's DSRD/TRANS on the
    same files: z and x moments to 1e-10 of the largest."""
    from dvr3d_amd import _lib
    job, (kr, kf), (br, bf) = CASES[name]
    ket = P.read_wave(waves[kr] / kf)
    bra = P.read_wave(waves[br] / bf)
    inp = INPUT[name]
    if name == "h3p":
        t = np.load(os.path.join(H.GOLD, "h3p_dipgrid.npz"))
        dipd = lambda r1, r2, x, nu: t["v"][nu]          # noqa: E731
    else:
        from tests.test_dip_oracle import synth_dipole as dipd
    res = P.run(ket, bra, inp, dipd)
    ib1, ib2 = res["ibase"]
    ne1, ne2 = res["tz"].shape
    xd, wtd, ipot = res["xd"], res["wtd"], res["ipot"]
    kb = P.device_blocks(ket, ne1, ib1, ket.ipar, ket.ipar, xd, ipot)
    bb = P.device_blocks(bra, ne2, ib2, bra.ipar, ket.ipar, xd, ipot)
    nrad = ket.npnt1 * ket.npnt2
    dip = np.stack([P.dipole_grid(dipd, bra, xd, wtd, nu, ipot) for nu in (0, 1)])
    pairs = [(p.b1, p.b2, p.comp, p.factor) for p in res["pairs"]]
    tz, tx = _lib.dipole3(nrad, ipot, kb, bb, dip, pairs)
    scale = max(np.abs(res["tz"]).max(), np.abs(res["tx"]).max())
    ez = np.abs(tz - res["tz"]).max() / scale
    ex = np.abs(tx - res["tx"]).max() / scale
    print(f"{name}: tz {ez:.1e}, tx {ex:.1e}")
    assert ez < 1e-10 and ex < 1e-10


def test_restart_continues_the_sums(tmp_path, waves):
    """nblock = 2 then a restart (ires = 1) from the scratch file gives the
    line strengths of one full run (RDSCR/WRSCR, P:2022-2041)."""
    job = _case_dir(tmp_path, waves, "k12")
    deck = open(job).read()
    first = deck.replace("zstart=.true.,", "zstart=.true., nblock=2,")
    (tmp_path / "a.job").write_text(first)
    text = _run(DIP, tmp_path / "a.job", tmp_path)
    assert "2 blocks calculated" in text
    (tmp_path / "b.job").write_text(deck.replace("zstart=.true.,", "zstart=.true., ires=1,"))
    text = _run(DIP, tmp_path / "b.job", tmp_path)
    assert "Block   3" in text and "Block   5" in text and "Block   1 " not in text
    _, _, _, _, _, S = _spectra(tmp_path / "fort.13")
    Sr = np.array(GOLD["k12"]["sint"])
    assert np.abs(S - Sr).max() / np.abs(Sr).max() < 1e-6
