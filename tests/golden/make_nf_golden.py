"""Golden fixtures of the symmetrised-Radau driver nfmain (IDIA = -2) from
reference runs.

The reference DVR3DRJZ built by oracle/Makefile (from the sources where they
lie) is run here on:
  water_J0  the reference's own water deck water/jobs/dvr0.job (copied as
            data to dvr3d_amd/jobs/water_J0.job), linked with the
            reference's water POTV (water/source/pot_water.f,
            `make -C oracle water`);
  c0r       the c0 grid and surface of dvr3d_amd/jobs/ab2_c0.par in
            symmetrised Radau coordinates (dvr3d_amd/jobs/c0r.job), linked
            with oracle/potv_synth.f90.
Written:
  tests/golden/nf_ref.json      per deck: 2D functions chosen per angle
                                (iv2), nham3 and the printed eigenvalues
                                (hartree)
  tests/golden/water_vgrid.npz  the water POTV on the deck's grid,
                                V[g, b, a] = V(r_a, r_b, x_g) (data: the GPU
                                tests need no PES code)
  tests/golden/nf_iwave_ref.npz the records transr writes to fort.26 for
                                both decks (angle count, basis size, the
                                per-angle counts, the levels and the first
                                two wavefunctions; records 2-8 as raw bytes)
The grid points come from oracle/nfmain_oracle.py, whose set-up reproduces
the reference's quadratures bit for bit (tests/test_oracle.py).

usage: python tests/golden/make_nf_golden.py
"""
import json
import os
import re
import shutil
import subprocess
import sys
import tempfile

import numpy as np

REPO = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, REPO)

from oracle import dvr3d_oracle as O  # noqa: E402
from oracle import nfmain_oracle as NF  # noqa: E402
from tests import fortran_io as F  # noqa: E402
from tests.golden.make_goldens import _floats  # noqa: E402

JOBS = os.path.join(REPO, "dvr3d_amd", "jobs")
REFD = os.path.join(REPO, "oracle", "_ref")
DECKS = {"water_J0": ("dvr3drjz_water", None), "c0r": ("dvr3drjz_ref", "ab2_c0.par")}


def parse(text):
    iv2 = [int(m) for m in re.findall(r"for gamma =\s*\d+ selected\s+(\d+) energies", text)]
    m = re.search(r"total =\s*(\d+)", text)
    i = text.index("eigenvalues in hartrees")
    vals = []
    for line in text[i:].split("\n")[1:]:
        if "wavenumbers" in line:
            break
        if line.strip():
            vals += _floats(line)
    return dict(iv2=iv2, nham3=int(m.group(1)), eval_au=vals)


def water_grid(job):
    b = NF.nf_basis(job)
    s = b.sizes
    G, B, A = np.meshgrid(np.arange(s.nalf), np.arange(s.npnt), np.arange(s.npnt), indexing="ij")
    pts = np.stack([b.r[A.ravel()], b.r[B.ravel()], b.xalf[G.ravel()]], axis=1)
    c = b.consts
    hdr = np.array([len(pts)], dtype="<i4").tobytes()
    mass = np.array(list(job["xmass"]) + [c.g1, c.g2] + list(job["xmassr"]), "<f8")
    inp = hdr + mass.tobytes() + np.ascontiguousarray(pts, "<f8").tobytes()
    r = subprocess.run([os.path.join(REFD, "water_potgrid")], input=inp, stdout=subprocess.PIPE,
                       check=True)
    return np.frombuffer(r.stdout, "<f8").copy().reshape(s.nalf, s.npnt, s.npnt)


def main():
    subprocess.run(["make", "-s", "-C", os.path.join(REPO, "oracle"), "all", "water"], check=True,
                   stdout=subprocess.DEVNULL)
    gold, iw = {}, {}
    for name, (exe, par) in DECKS.items():
        d = tempfile.mkdtemp(prefix="dvr3d_nf_")
        shutil.copy(os.path.join(JOBS, name + ".job"), os.path.join(d, "in.job"))
        if par:
            shutil.copy(os.path.join(JOBS, par), os.path.join(d, "pes.par"))
        with open(os.path.join(d, "in.job")) as fi:
            r = subprocess.run([os.path.join(REFD, exe)], stdin=fi, cwd=d, capture_output=True,
                               text=True, check=True, timeout=600)
        res = parse(r.stdout)
        res["source"] = f"oracle/_ref/{exe} on dvr3d_amd/jobs/{name}.job"
        gold[name] = res
        recs = F.read_records(os.path.join(d, "fort.26"))
        # set-up records 1-3 (setcon), r1 (D:843), angular records of the k
        # block (D:973-977), then testiv (iang, nbass; iv2) and transr
        # (neval; levels; one record per level)
        iw[name + "_nrec"] = np.array([len(recs)])
        iw[name + "_rec1"] = F.ints(recs[0])
        k = 8   # records 1-3, 4 (mbass ...), 5 (r1), 6-8 angular (no r2 record)
        for i in range(1, k):
            iw[f"{name}_hdr{i + 1}"] = np.frombuffer(recs[i], np.uint8).copy()
        iw[name + "_iang_nbass"] = F.ints(recs[k])
        iw[name + "_iv2"] = F.ints(recs[k + 1])
        iw[name + "_neval"] = F.ints(recs[k + 2])
        iw[name + "_eval"] = F.doubles(recs[k + 3])
        iw[name + "_psi1"] = F.doubles(recs[k + 4])
        iw[name + "_psi2"] = F.doubles(recs[k + 5])
        shutil.rmtree(d, ignore_errors=True)
        print(name, "nham3", res["nham3"], "levels", len(res["eval_au"]), "records", len(recs))
    job = O.load_job(os.path.join(JOBS, "water_J0.job"))
    np.savez_compressed(os.path.join(REPO, "tests", "golden", "water_vgrid.npz"), v=water_grid(job))
    np.savez_compressed(os.path.join(REPO, "tests", "golden", "nf_iwave_ref.npz"), **iw)
    with open(os.path.join(REPO, "tests", "golden", "nf_ref.json"), "w") as f:
        json.dump(gold, f, indent=1)


if __name__ == "__main__":
    main()
