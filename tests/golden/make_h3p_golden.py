"""Golden fixtures from the reference's own H3+ Jacobi jobs
(H3+Jacobi/jobs/dvrJ0.job, dvrJ1.job, copied as data to
dvr3d_amd/jobs/h3p_J0.job, h3p_J1.job).

The reference DVR3DRJZ linked with the reference's own H3+ POTV
(H3+Jacobi/source/pot_h3p.f90) is built by `make -C oracle h3p` from the
sources where they lie, and run here on both jobs.  Written:
  tests/golden/h3p_ref.json   per job and per k block: nham2, iv2, nham3 and
                              the printed eigenvalues (hartree); plus the
                              levels (cm-1) of the reference's shipped
                              H3+Jacobi/results/result.H3p.J0.dvr
  tests/golden/h3p_vgrid.npz  the same POTV on every k block's DVR grid,
                              V[c, b, a] (data: the GPU tests need no PES code)
  tests/golden/h3p_J0_iwave_ref.npz, h3p_J1_iwave_ref.npz
                              the set-up records of the reference's fort.26
                              and, per k, the angular records, the
                              eigenvalues and the first wavefunction
The grid points come from the oracle's set-up, which reproduces the
reference's quadratures bit for bit (tests/test_oracle.py).

usage: python tests/golden/make_h3p_golden.py
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
from tests.golden.make_goldens import _floats, write_iwave_fixture_k  # noqa: E402

REF = os.path.join(REPO, "oracle", "_ref", "dvr3drjz_h3p")
GRID = os.path.join(REPO, "oracle", "_ref", "h3p_potgrid")
JOBS = ("h3p_J0", "h3p_J1")
REFDIR = "/root/reference"


def potv_on(points, consts, job):
    """The reference POTV at points (n, 3) = (r1, r2, xcos)."""
    hdr = np.array([len(points)], dtype="<i4").tobytes()
    mass = np.array(list(job["xmass"]) + [consts.g1, consts.g2] + list(job["xmassr"]), "<f8")
    inp = hdr + mass.tobytes() + np.ascontiguousarray(points, "<f8").tobytes()
    r = subprocess.run([GRID], input=inp, stdout=subprocess.PIPE, check=True)
    return np.frombuffer(r.stdout, "<f8").copy()


def grid_of(basis):
    s = basis.sizes
    x = basis.xalf[: s.npnta]
    if s.zr2r1:
        R1, R2, X = np.meshgrid(basis.r1, basis.r2, x, indexing="ij")
    else:
        R2, R1, X = np.meshgrid(basis.r2, basis.r1, x, indexing="ij")
    return np.stack([R1.ravel(), R2.ravel(), X.ravel()], axis=1), (s.npntc, s.npntb, s.npnta)


def parse_blocks(text):
    """Per k block: nham2 list, iv2 list, nham3, eigenvalues (hartree)."""
    parts = re.split(r"point Gauss-associated Legendre integration with", text)[1:]
    out = []
    for p in parts:
        nham2 = [int(m) for m in re.findall(r"nham2 =\s*(\d+)", p)]
        iv2 = [int(m) for m in re.findall(r"no\. of eigenvectors =\s*(\d+)", p)]
        m = re.search(r"(\d+) eigenvalues selected from", p)
        i = p.index("eigenvalues in hartrees")
        vals = []
        for line in p[i:].split("\n")[1:]:
            if "wavenumbers" in line:
                break
            if line.strip():
                vals += _floats(line)
        out.append(dict(nham2=nham2, iv2=iv2, nham3=int(m.group(1)) if m else None,
                        eval_au=vals))
    return out


def main():
    subprocess.run(["make", "-C", os.path.join(REPO, "oracle"), "h3p"], check=True,
                   stdout=subprocess.DEVNULL)
    gold, grids = {}, {}
    for name in JOBS:
        jobfile = os.path.join(REPO, "dvr3d_amd", "jobs", name + ".job")
        job = O.load_job(jobfile)
        rundir = os.path.join(tempfile.mkdtemp(prefix="dvr3d_ref_"), name)
        os.makedirs(rundir, exist_ok=True)
        shutil.copy(jobfile, os.path.join(rundir, name + ".job"))
        with open(os.path.join(rundir, name + ".job")) as fi, \
                open(os.path.join(rundir, name + ".out"), "w") as fo:
            subprocess.run([REF], stdin=fi, stdout=fo, cwd=rundir, check=True)
        blocks = parse_blocks(open(os.path.join(rundir, name + ".out")).read())
        ks = O.k_blocks(job)
        assert len(blocks) == len(ks), (name, len(blocks), ks)
        write_iwave_fixture_k(name, os.path.join(rundir, "fort.26"), len(ks))
        gold[name] = {}
        for k, b in zip(ks, blocks):
            basis = O.setup_basis(job, k)
            pts, shape = grid_of(basis)
            grids[f"{name}_k{k}"] = potv_on(pts, basis.consts, job).reshape(shape)
            gold[name][str(k)] = b
            print(name, "k", k, "nham3", b["nham3"], "levels", len(b["eval_au"]), flush=True)
    shipped = open(os.path.join(REFDIR, "H3+Jacobi", "results", "result.H3p.J0.dvr")).read()
    i = shipped.index("eigenvalues in wavenumbers")
    vals = []
    for line in shipped[i:].split("\n")[1:]:
        if "Band origins" in line:
            break
        if line.strip():
            vals += _floats(line)
    gold["shipped_J0_cm"] = vals
    gold["source"] = ("oracle/_ref/dvr3drjz_h3p (reference DVR3DRJZ + H3+Jacobi/source/"
                      "pot_h3p.f90) on dvr3d_amd/jobs/h3p_J0.job, h3p_J1.job")
    with open(os.path.join(REPO, "tests", "golden", "h3p_ref.json"), "w") as f:
        json.dump(gold, f)
    np.savez_compressed(os.path.join(REPO, "tests", "golden", "h3p_vgrid.npz"), **grids)


if __name__ == "__main__":
    main()
