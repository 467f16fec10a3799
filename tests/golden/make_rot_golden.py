"""Golden fixtures of the rotational step ROTLEV3 (source/rotlev3.f90) from
reference runs.

The reference DVR3DRJZ (built by oracle/Makefile from the sources where they
lie) writes the IWAVE file of a zrot run, and the reference ROTLEV3
(`make -C oracle rot`, source/rotlev3.f90 + source/f02fjf.f, MKL) reads it:
  h3p  the reference's own H3+ Jacobi jobs: H3+Jacobi/jobs/dvrJ1.job
       (dvr3d_amd/jobs/h3p_J1.job; IDIA=2, zembed=F, J=1, k = 0, 1) with the
       reference's H3+ POTV, then H3+Jacobi/jobs/rot.job
       (dvr3d_amd/jobs/h3p_rot.job: KMIN=2, IBASS=700 of 800; the f set of
       J=1 is DVR3DRJZ's own k=1 block, DSTORE1)
  k2s  the J=2 ABC deck dvr3d_amd/jobs/k2s.job (IDIA=1, zembed=T) with the
       synthetic surface abc_k2s.par, then dvr3d_amd/jobs/k2s_rot.job
       (KMIN=2, IBASS=70 of 90: e set over k = 0, 1, 2 and f set over
       k = 1, 2)
ILEV output (zpfun) is left off: the reference's ILEV append writes after an
end-of-file read, which the amdflang runtime refuses (rotlev3.f90:1104-1113,
its "machine dependent" backspace card).
Written:
  tests/golden/rot_ref.json     per deck: the selection (mvib per k), and per
                                parity set the levels (hartree) printed on
                                stdout
  tests/golden/rot_vec_ref.npz  per deck and set: the KVEC records (header
                                records 1-7 raw, the level record, the first
                                two transformed vectors of every k block) and
                                the JVEC header records and first two columns
                                of every k block

usage: python tests/golden/make_rot_golden.py
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

from tests import fortran_io as F  # noqa: E402
from tests.golden.make_goldens import _floats  # noqa: E402

JOBS = os.path.join(REPO, "dvr3d_amd", "jobs")
REFD = os.path.join(REPO, "oracle", "_ref")
DECKS = {"h3p": ("dvr3drjz_h3p", "h3p_J1.job", None, "h3p_rot.job"),
         "k2s": ("dvr3drjz_ref", "k2s.job", "abc_k2s.par", "k2s_rot.job")}


def levels(text):
    """Every 'EIGENVALUES IN HARTREES' block of a ROTLEV3 report."""
    out = []
    for part in text.split("EIGENVALUES IN HARTREES")[1:]:
        vals = []
        for line in part.split("\n")[1:]:
            if "wavenumbers" in line:
                break
            if line.strip():
                vals += _floats(line)
        out.append(vals)
    return out


def run_reference(name, d):
    exe, job, par, rot = DECKS[name]
    shutil.copy(os.path.join(JOBS, job), os.path.join(d, "dvr.job"))
    if par:
        shutil.copy(os.path.join(JOBS, par), os.path.join(d, "pes.par"))
    with open(os.path.join(d, "dvr.job")) as fi:
        subprocess.run([os.path.join(REFD, exe)], stdin=fi, cwd=d, stdout=subprocess.DEVNULL,
                       check=True, timeout=900)
    with open(os.path.join(JOBS, rot)) as fi:
        r = subprocess.run([os.path.join(REFD, "rotlev3_ref")], stdin=fi, cwd=d,
                           capture_output=True, text=True, check=True, timeout=900)
    return r.stdout


def main():
    subprocess.run(["make", "-s", "-C", os.path.join(REPO, "oracle"), "all", "h3p", "rot"],
                   check=True, stdout=subprocess.DEVNULL)
    gold, vec = {}, {}
    for name in DECKS:
        d = tempfile.mkdtemp(prefix="dvr3d_rot_")
        out = run_reference(name, d)
        lev = levels(out)
        sel = [int(b) - int(a) + 1 for a, b in re.findall(r"I RUNS FROM\s+(\d+) TO\s+(\d+)", out)]
        gold[name] = dict(mvib=sel, sets=[dict(eval_au=v) for v in lev],
                          source=f"oracle/_ref/rotlev3_ref on the fort.26 of "
                                 f"oracle/_ref/{DECKS[name][0]} {DECKS[name][1]}")
        for s, (kv, jv) in enumerate((("fort.8", "fort.3"), ("fort.9", "fort.2"))):
            p = os.path.join(d, kv)
            if not os.path.exists(p) or os.path.getsize(p) == 0:
                continue
            recs = F.read_records(p)
            key = f"{name}_s{s}"
            # DSTORE (R:1201-1341): 8 header records, then one record per k
            # block (nkbas x nval); the J=1f copy DSTORE1 (R:1344-1475): 11
            # header records (the k=1 block's angular records added), then
            # one record per level
            j1f = int(F.ints(recs[0])[5]) == 1 and int(F.ints(recs[0])[6]) == 0
            nh = 11 if j1f else 8
            for i in range(nh - 1):
                vec[f"{key}_k{i + 1}"] = np.frombuffer(recs[i], np.uint8).copy()
            vec[f"{key}_keval"] = F.doubles(recs[nh - 1]).copy()
            nval = int(F.ints(recs[nh - 2])[0])
            vec[f"{key}_nrec"] = np.array([len(recs), nh])
            blocks = recs[nh:]
            if j1f:
                firsts = [np.stack([F.doubles(r) for r in blocks[:2]])]
            else:
                firsts = [F.doubles(r).reshape(nval, -1)[:2].copy() for r in blocks]
            for b, f in enumerate(firsts):
                vec[f"{key}_d{b}"] = f
            vec[f"{key}_ndblk"] = np.array([len(firsts)])
            jp = os.path.join(d, jv)
            if os.path.exists(jp) and os.path.getsize(jp) > 0:
                jr = F.read_records(jp)
                vec[f"{key}_j1"] = F.ints(jr[0]).copy()
                vec[f"{key}_j2"] = F.ints(jr[1]).copy()
                vec[f"{key}_jeval"] = F.doubles(jr[2]).copy()
                for b, r in enumerate(jr[3:]):
                    a = F.doubles(r)
                    vec[f"{key}_jc{b}"] = a.reshape(nval, len(a) // nval)[:2].copy()
                vec[f"{key}_njblk"] = np.array([len(jr) - 3])
        shutil.rmtree(d, ignore_errors=True)
        print(name, "mvib", sel, "levels per set", [len(v) for v in lev])
    np.savez_compressed(os.path.join(REPO, "tests", "golden", "rot_vec_ref.npz"), **vec)
    with open(os.path.join(REPO, "tests", "golden", "rot_ref.json"), "w") as f:
        json.dump(gold, f, indent=1)


if __name__ == "__main__":
    main()
