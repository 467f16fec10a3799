"""The drop-in executable dvr3drjz_amd against the reference executable's
outputs on the same .job deck: the stdout eigenvalue report, the ILEV file
and the IWAVE records that ROTLEV3/DIPOLE3/XPECT3 read."""
import json
import os
import re
import shutil
import subprocess

import numpy as np
import pytest

from oracle import dvr3d_oracle as O
from tests import fortran_io as F
from tests import helpers as H

pytestmark = pytest.mark.gpu

EXE = os.path.join(H.REPO, "dvr3d_amd", "host", "dvr3drjz_amd")


@pytest.fixture(scope="module")
def c0_run(tmp_path_factory):
    d = tmp_path_factory.mktemp("c0")
    shutil.copy(os.path.join(H.JOBS, "c0.job"), d / "c0.job")
    shutil.copy(os.path.join(H.JOBS, "ab2_c0.par"), d / "pes.par")
    with open(d / "c0.job") as fi, open(d / "c0.out", "w") as fo:
        r = subprocess.run([EXE], stdin=fi, stdout=fo, stderr=subprocess.STDOUT, cwd=d, timeout=120)
    out = open(d / "c0.out").read()
    assert r.returncode == 0, out[-2000:]
    return d, out


def _block(text, head, stop):
    i = text.index(head)
    vals = []
    for line in text[i:].split("\n")[1:]:
        if stop in line:
            break
        vals += [float(x.replace("D", "E")) for x in line.split()]
    return np.array(vals)


def test_stdout_levels_match_reference(c0_run):
    d, out = c0_run
    g = H.golden("c0")
    cm = _block(out, "eigenvalues in wavenumbers", "Band origins")
    au = _block(out, "eigenvalues in hartrees", "wavenumbers")
    assert len(cm) == len(au) == 100
    assert np.abs(cm - np.array(g["eval_cm"])).max() < 1e-4
    assert np.abs(au - np.array(g["eval_au"])).max() < 1e-9


def test_ilev_file(c0_run):
    d, out = c0_run
    lines = open(d / "fort.14").read().split("\n")
    assert lines[0].startswith(" c0: AB2")                      # title (9a8)
    assert lines[1] == "   0   0   0   0   1 100"                # 6i4: J ip jdia jpar isym meval
    vals = [float(x.replace("D", "E")) for l in lines[2:] for x in re.findall(r"\S+", l)]
    g = H.golden("c0")
    assert len(vals) == 100
    assert np.abs(np.array(vals) - np.array(g["eval_au"])).max() < 1e-9


def test_iwave_records_match_reference(c0_run):
    d, out = c0_run
    ref = np.load(os.path.join(H.GOLD, "c0_iwave_ref.npz"))
    recs = F.read_records(d / "fort.26")
    assert len(recs) == int(ref["nrec"][0])
    assert [len(r) for r in recs[:12]] == ref["reclens"].tolist()
    assert np.array_equal(F.ints(recs[0]), ref["rec1"])
    r2 = recs[1]
    assert np.array_equal(np.frombuffer(r2[0:12], "<i4") != 0, ref["rec2_logicals"] != 0)
    assert np.array_equal(np.frombuffer(r2[12:52], "<f8"), ref["rec2_reals"])
    assert np.array_equal(np.frombuffer(r2[52:60], "<i4") != 0, ref["rec2_logicals_tail"] != 0)
    assert np.array_equal(F.doubles(recs[2]), ref["rec3"])
    assert np.array_equal(F.ints(recs[3]), ref["rec4"])
    assert np.array_equal(F.doubles(recs[4]), ref["rec5_r1"])
    assert np.array_equal(F.doubles(recs[5]), ref["rec6_r2"])
    assert np.array_equal(F.doubles(recs[6]), ref["rec7_xalf"])
    assert np.array_equal(F.ints(recs[7]), ref["rec8"])
    assert np.abs(F.doubles(recs[8]) - ref["rec9_plegw"]).max() < 1e-14
    assert np.array_equal(F.ints(recs[9]), ref["rec10_meval"])
    assert np.abs(F.doubles(recs[10]) - ref["rec11_eval"]).max() < 1e-9
    for k, name in ((11, "psi_level1"), (12, "psi_level2")):
        a, b = F.doubles(recs[k]), ref[name]
        s = np.sign(a @ b)                                   # eigenvector sign is arbitrary
        assert np.abs(a - s * b).max() < 1e-7


def test_c1_pipeline_matches_reference():
    """Full c1 workload on the device: 500 levels within 1e-4 cm-1 of the
    reference executable, identical truncation."""
    job, pes = H.job_and_pes("c1")
    inp = H.device_inputs(job, pes)
    ctx = H.run_device(inp)
    g = H.golden("c1")
    iv1, iv2, n3 = ctx.sizes()
    assert iv1.sum(axis=1).tolist() == g["nham2"]
    assert iv2.tolist() == g["iv2"] and n3 == g["nham3"]
    ev = ctx.eigenvalues() * O.AUTOCM
    err = np.abs(ev - np.array(g["eval_cm"]))
    print("c1 max |dE| (cm-1):", err.max(), "stage ms:", ctx.stage_times())
    assert len(ev) == 500 and err.max() < 1e-4


def _run_exe(tmp, cfg, par, edit=None):
    shutil.copy(os.path.join(H.JOBS, par), tmp / "pes.par")
    text = open(os.path.join(H.JOBS, cfg + ".job")).read()
    if edit:
        text = edit(text)
    (tmp / (cfg + ".job")).write_text(text)
    with open(tmp / (cfg + ".job")) as fi, open(tmp / (cfg + ".out"), "w") as fo:
        r = subprocess.run([EXE], stdin=fi, stdout=fo, stderr=subprocess.STDOUT, cwd=tmp,
                           timeout=300)
    out = open(tmp / (cfg + ".out")).read()
    assert r.returncode == 0, out[-2000:]
    return out


def _levels_per_k(out):
    blocks = []
    for part in out.split("eigenvalues in hartrees")[1:]:
        vals = []
        for line in part.split("\n")[1:]:
            if "wavenumbers" in line:
                break
            vals += [float(x.replace("D", "E")) for x in line.split()]
        blocks.append(np.array(vals))
    return blocks


def test_k2s_jrot_levels_and_iwave_match_reference(tmp_path):
    """J = 2 with zrot (k = 0, 1, 2 blocks run concurrently): per-k levels and
    the IWAVE layout of the reference (r**-2 record, angular records per k)."""
    out = _run_exe(tmp_path, "k2s", "abc_k2s.par")
    g = json.load(open(os.path.join(H.GOLD, "k2s_ref.json")))["eval_au"]
    blocks = _levels_per_k(out)
    assert len(blocks) == 3
    # the reference's per-k heading of a zrot run (D:2682-2684, format 1040)
    heads = [l for l in out.splitlines() if "Solutions with J =" in l]
    assert heads == [f"     Solutions with J =  2 k = {k:2d}" for k in range(3)]
    for k, b in enumerate(blocks):
        assert len(b) == 30
        assert np.abs(b - np.array(g[str(k)])).max() * O.AUTOCM < 1e-4
    ref = np.load(os.path.join(H.GOLD, "k2s_iwave_ref.npz"))
    recs = F.read_records(tmp_path / "fort.26")
    assert len(recs) == int(ref["nrec"][0])
    assert np.array_equal(F.ints(recs[0]), ref["rec1"])
    r2 = np.frombuffer(recs[1], np.uint8)
    assert len(r2) == len(ref["rec2"])
    assert np.array_equal(np.frombuffer(recs[1][12:52], "<f8"),
                          np.frombuffer(ref["rec2"][12:52].tobytes(), "<f8"))
    assert np.array_equal(F.doubles(recs[2]), ref["rec3"])
    assert np.abs(F.doubles(recs[3]) - ref["rec4_r2m2"]).max() < 1e-15
    assert np.array_equal(F.doubles(recs[4]), ref["rec5_r1"])
    assert np.array_equal(F.doubles(recs[5]), ref["rec6_r2"])
    pos = 6
    for k in range(3):
        assert np.abs(F.doubles(recs[pos]) - ref[f"k{k}_xalf"]).max() < 1e-14
        assert np.array_equal(F.ints(recs[pos + 1]), ref[f"k{k}_kzrec"])
        assert np.abs(F.doubles(recs[pos + 2]) - ref[f"k{k}_plegw"]).max() < 1e-13
        assert np.array_equal(F.ints(recs[pos + 3]), ref[f"k{k}_meval"])
        assert np.abs(F.doubles(recs[pos + 4]) - ref[f"k{k}_eval"]).max() < 1e-9
        a, b = F.doubles(recs[pos + 5]), ref[f"k{k}_psi1"]
        s = np.sign(a @ b)
        assert np.abs(a - s * b).max() < 1e-7
        pos += 5 + int(ref[f"k{k}_meval"][0])


def test_c2_dropin_levels_match_oracle(tmp_path):
    """The c2 deck (ABC with hashed PES noise, J = 2, N3 ~ 7400-7700 per k)
    through the executable and its example POTV; IWAVE off (all N3 levels of
    three blocks would be tens of GB)."""
    path = os.path.join(H.GOLD, "c2_oracle.json")
    if not os.path.exists(path):
        pytest.skip("c2 oracle golden not generated")
    gold = json.load(open(path))["k"]
    out = _run_exe(tmp_path, "c2", "abc_c2.par",
                   edit=lambda t: t.replace("ztran=.true.", "ztran=.false."))
    blocks = _levels_per_k(out)
    assert len(blocks) == 3
    for k, b in enumerate(blocks):
        ref = np.array(gold[str(k)]["eval_au"])
        assert len(b) == len(ref)
        assert np.abs(b - ref).max() * O.AUTOCM < 1e-4, k


@pytest.mark.parametrize("name", ["h3p_J0", "h3p_J1"])
def test_h3p_reference_jobs_through_executable(tmp_path, name):
    """The reference's own H3+ Jacobi decks (IDIA=2, zr2r1=F; J=0 odd parity
    with EZERO, and J=1 with zembed=F over k = 0, 1) through the executable,
    with the reference's H3+ POTV supplied as a table of its values on every
    k block's grid (kind 3 of the example POTV): levels per k and the IWAVE
    records against the reference run of the same decks."""
    from dvr3d_amd import host
    g = json.load(open(os.path.join(H.GOLD, "h3p_ref.json")))[name]
    V = np.load(os.path.join(H.GOLD, "h3p_vgrid.npz"))
    deck = os.path.join(H.JOBS, name + ".job")
    job = host.read_job(deck)
    sz = host.sizes(job)
    ks = list(range(sz.kd, sz.ku + 1))
    rows = []
    for k in ks:
        b = host.Basis(job, k)
        x = b.xalf[: b.sz.npnta]
        if job.zr2r1:
            R1, R2, X = np.meshgrid(b.r1, b.r2, x, indexing="ij")
        else:
            R2, R1, X = np.meshgrid(b.r2, b.r1, x, indexing="ij")
        rows.append(np.stack([R1.ravel(), R2.ravel(), X.ravel(), V[f"{name}_k{k}"].ravel()], axis=1))
    tab = np.concatenate(rows)
    with open(tmp_path / "table.bin", "wb") as f:
        np.array([len(tab)], dtype="<f8").tofile(f)
        tab.astype("<f8").tofile(f)
    (tmp_path / "pes.par").write_text("3\n" + str(tmp_path / "table.bin") + "\n")
    shutil.copy(deck, tmp_path / (name + ".job"))
    with open(tmp_path / (name + ".job")) as fi, open(tmp_path / "out.txt", "w") as fo:
        r = subprocess.run([EXE], stdin=fi, stdout=fo, stderr=subprocess.STDOUT, cwd=tmp_path,
                           timeout=300)
    out = open(tmp_path / "out.txt").read()
    assert r.returncode == 0, out[-2000:]
    blocks = _levels_per_k(out)
    assert len(blocks) == len(ks)
    for k, blk in zip(ks, blocks):
        ref = np.array(g[str(k)]["eval_au"])
        assert len(blk) == len(ref) == 400
        assert np.abs(blk - ref).max() * O.AUTOCM < 1e-4, k
    ref = np.load(os.path.join(H.GOLD, name + "_iwave_ref.npz"))
    recs = F.read_records(tmp_path / "fort.26")
    assert len(recs) == int(ref["nrec"][0])
    assert np.array_equal(F.ints(recs[0]), ref["rec1"])
    assert np.array_equal(np.frombuffer(recs[1][12:52], "<f8"),
                          np.frombuffer(ref["rec2"][12:52].tobytes(), "<f8"))
    assert np.array_equal(F.doubles(recs[2]), ref["rec3"])
    r4, r4ref = np.frombuffer(recs[3], np.uint8), ref["rec4_raw"]
    assert len(r4) == len(r4ref)
    if len(ks) > 1:       # the r**-2 array of a zrot run
        assert np.allclose(np.frombuffer(recs[3], "<f8"), np.frombuffer(r4ref.tobytes(), "<f8"),
                           rtol=1e-15, atol=0)
    else:                 # mbass, jstart, nang, mbass
        assert np.array_equal(r4, r4ref)
    assert np.array_equal(F.doubles(recs[4]), ref["rec5_r1"])
    assert np.array_equal(F.doubles(recs[5]), ref["rec6_r2"])
    pos = 6
    for i, k in enumerate(ks):
        assert np.abs(F.doubles(recs[pos]) - ref[f"k{i}_xalf"]).max() < 1e-14
        assert np.array_equal(F.ints(recs[pos + 1]), ref[f"k{i}_kzrec"])
        assert np.abs(F.doubles(recs[pos + 2]) - ref[f"k{i}_plegw"]).max() < 1e-13
        assert np.array_equal(F.ints(recs[pos + 3]), ref[f"k{i}_meval"])
        assert np.abs(F.doubles(recs[pos + 4]) - ref[f"k{i}_eval"]).max() < 1e-9
        a, b2 = F.doubles(recs[pos + 5]), ref[f"k{i}_psi1"]
        assert np.abs(a - np.sign(a @ b2) * b2).max() < 1e-7
        pos += 5 + int(ref[f"k{i}_meval"][0])


@pytest.mark.parametrize("name", ["water_J0", "c0r"])
def test_symmetrised_radau_decks_through_executable(tmp_path, name):
    """IDIA = -2 decks through the executable: the reference's own water deck
    (water/jobs/dvr0.job) with the reference's water POTV supplied as a table
    of its values on the deck's grid (kind 3 of the example POTV), and the c0
    surface in symmetrised Radau coordinates.  The per-angle selections, the
    printed levels and every IWAVE record transr writes (D:3555-3646) against
    the reference run of the same decks."""
    from oracle import nfmain_oracle as NF
    g = json.load(open(os.path.join(H.GOLD, "nf_ref.json")))[name]
    ref = np.load(os.path.join(H.GOLD, "nf_iwave_ref.npz"))
    deck = os.path.join(H.JOBS, name + ".job")
    if name == "water_J0":
        b = NF.nf_basis(O.load_job(deck))
        s = b.sizes
        V = np.load(os.path.join(H.GOLD, "water_vgrid.npz"))["v"]
        G, B, A = np.meshgrid(np.arange(s.nalf), np.arange(s.npnt), np.arange(s.npnt),
                              indexing="ij")
        tab = np.stack([b.r[A.ravel()], b.r[B.ravel()], b.xalf[G.ravel()], V.ravel()], axis=1)
        with open(tmp_path / "table.bin", "wb") as f:
            np.array([len(tab)], dtype="<f8").tofile(f)
            tab.astype("<f8").tofile(f)
        (tmp_path / "pes.par").write_text("3\n" + str(tmp_path / "table.bin") + "\n")
    else:
        shutil.copy(os.path.join(H.JOBS, "ab2_c0.par"), tmp_path / "pes.par")
    shutil.copy(deck, tmp_path / "in.job")
    with open(tmp_path / "in.job") as fi, open(tmp_path / "out.txt", "w") as fo:
        r = subprocess.run([EXE], stdin=fi, stdout=fo, stderr=subprocess.STDOUT, cwd=tmp_path,
                           timeout=300)
    out = open(tmp_path / "out.txt").read()
    assert r.returncode == 0, out[-2000:]
    iv2 = [int(m) for m in re.findall(r"for gamma =\s*\d+ selected\s+(\d+) energies", out)]
    assert iv2 == g["iv2"]
    assert int(re.search(r"total =\s*(\d+)", out).group(1)) == g["nham3"]
    lev = _levels_per_k(out)
    assert len(lev) == 1 and len(lev[0]) == len(g["eval_au"])
    assert np.abs(lev[0] - np.array(g["eval_au"])).max() * O.AUTOCM < 1e-4
    recs = F.read_records(tmp_path / "fort.26")
    assert len(recs) == int(ref[name + "_nrec"][0])
    assert np.array_equal(F.ints(recs[0]), ref[name + "_rec1"])
    for i in (2, 4, 7):     # logicals and masses; mbass ...; the angular k block
        assert np.array_equal(np.frombuffer(recs[i - 1], np.uint8), ref[f"{name}_hdr{i}"]), i
    for i in (3, 5, 6, 8):  # Morse parameters; r; the angular points; pleg * w
        a = np.frombuffer(recs[i - 1], "<f8")
        b2 = np.frombuffer(ref[f"{name}_hdr{i}"].tobytes(), "<f8")
        assert len(a) == len(b2) and np.abs(a - b2).max() <= 1e-13 * max(1.0, np.abs(b2).max()), i
    assert np.array_equal(F.ints(recs[8]), ref[name + "_iang_nbass"])
    assert np.array_equal(F.ints(recs[9]), ref[name + "_iv2"])
    assert np.array_equal(F.ints(recs[10]), ref[name + "_neval"])
    assert np.abs(F.doubles(recs[11]) - ref[name + "_eval"]).max() * O.AUTOCM < 1e-4
    for k, key in enumerate(("_psi1", "_psi2")):
        a, b2 = F.doubles(recs[12 + k]), ref[name + key]
        assert np.abs(a - np.sign(a @ b2) * b2).max() < 1e-7
