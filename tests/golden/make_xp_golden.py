"""Golden fixtures of the expectation-value step XPECT3 (source/xpect3.f90)
from reference runs.

Each case is a chain of reference programs (built by oracle/Makefile from the
sources where they lie): DVR3DRJZ writes the IWAVE file (and ILEV), ROTLEV3
turns the J > 0 ones into KVEC files, and XPECT3 reads one of them:
  water  the reference's own water J = 0 test (water/jobs/dvr0.job, an
         IDIA = -2 nfmain run with the reference's water POTV, then
         water/jobs/xpect0.job with its water/source/prop_water.f90: zfit,
         differences from the ground state, ILEV's lowest level as zero)
  c0     the c0 deck (IDIA = -1 jhmain, J = 0) with ILEV present, synthetic
         properties (oracle/prop_synth.f90), xp_synth.job
  k2s    the J = 2 e set of ROTLEV3 on k2s (IDIA = 1), xp_synth.job
  k2f    the J = 2 f set of the same run, levels 1-12 of three properties
  h3p    the J = 1 e set of the reference's H3+ rot.job (IDIA = 2,
         zembed = F), 42 quadrature points (xp_h3p.job)
Written:
  tests/golden/xp_ref.json         per case: the stdout table (energies in
                                   cm-1 and the property columns) and the
                                   numbers of the ITRA file (and of ITRA0 for
                                   water)
  tests/golden/water_propgrid.npz  the reference's water property routine
                                   (iprop 1-4) at every (r_a, r_b, x) of the
                                   water deck's grid, so that the GPU tests
                                   need no property code

usage: python tests/golden/make_xp_golden.py
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

from oracle import xpect3_oracle as X  # noqa: E402

JOBS = os.path.join(REPO, "dvr3d_amd", "jobs")
REFD = os.path.join(REPO, "oracle", "_ref")
# wavefunction runs: DVR3DRJZ executable, deck, surface parameters, ROTLEV3 deck
RUNS = {"water": ("dvr3drjz_water", "water_J0.job", None, None),
        "c0": ("dvr3drjz_ref", "c0.job", "ab2_c0.par", None),
        "k2s": ("dvr3drjz_ref", "k2s.job", "abc_k2s.par", "k2s_rot.job"),
        "h3pJ1": ("dvr3drjz_h3p", "h3p_J1.job", None, "h3p_rot.job")}
# XPECT3 runs: executable, deck, (run, file)
CASES = {"water": ("xpect3_water", "water_xp0.job", ("water", "fort.26")),
         "c0": ("xpect3_synth", "xp_synth.job", ("c0", "fort.26")),
         "k2s": ("xpect3_synth", "xp_synth.job", ("k2s", "fort.8")),
         "k2f": ("xpect3_synth", "xp_synth3.job", ("k2s", "fort.9")),
         "h3p": ("xpect3_synth", "xp_h3p.job", ("h3pJ1", "fort.8"))}


def built():
    need = {r[0] for r in RUNS.values()} | {c[0] for c in CASES.values()} | {"rotlev3_ref",
                                                                          "water_propgrid"}
    return all(os.path.exists(os.path.join(REFD, e)) for e in need)


def run_wave(name, d):
    exe, job, par, rot = RUNS[name]
    os.makedirs(d, exist_ok=True)
    if par:
        shutil.copy(os.path.join(JOBS, par), os.path.join(d, "pes.par"))
    with open(os.path.join(JOBS, job)) as fi:
        subprocess.run([os.path.join(REFD, exe)], stdin=fi, cwd=d, stdout=subprocess.DEVNULL,
                       stderr=subprocess.DEVNULL, check=True, timeout=900)
    if rot:
        with open(os.path.join(JOBS, rot)) as fi:
            subprocess.run([os.path.join(REFD, "rotlev3_ref")], stdin=fi, cwd=d,
                           stdout=subprocess.DEVNULL, stderr=subprocess.DEVNULL, check=True,
                           timeout=900)
    return d


def table(text):
    """The rows of OUTPT2's table (format 308): energy and the values."""
    rows = [l.split() for l in text.split("energy   property")[-1].splitlines()
            if re.match(r"^\s*\d+ [ab][12]\s", l)]
    return [[float(v.replace("D", "E")) for v in r[3:]] for r in rows]


def numbers(path):
    """Every number of a formatted file (list-directed and format 308)."""
    out = []
    for tok in open(path).read().replace("D", "E").split():
        try:
            out.append(float(tok))
        except ValueError:
            pass
    return out


def run_xpect(name, root, waves):
    exe, job, (wr, wf) = CASES[name]
    d = os.path.join(root, "xp_" + name)
    os.makedirs(d, exist_ok=True)
    shutil.copy(os.path.join(waves[wr], wf), os.path.join(d, "fort.11"))
    if wf == "fort.26" and os.path.exists(os.path.join(waves[wr], "fort.14")):
        shutil.copy(os.path.join(waves[wr], "fort.14"), os.path.join(d, "fort.14"))
    with open(os.path.join(JOBS, job)) as fi:
        r = subprocess.run([os.path.join(REFD, exe)], stdin=fi, cwd=d, capture_output=True,
                           text=True, check=True, timeout=900)
    g = dict(table=table(r.stdout), itra=numbers(os.path.join(d, "fort.12")))
    if os.path.exists(os.path.join(d, "fort.28")):
        g["itra0"] = numbers(os.path.join(d, "fort.28"))
    g["source"] = (f"oracle/_ref/{exe} < dvr3d_amd/jobs/{job} on the {wf} of the "
                   f"{RUNS[wr][1]} run")
    return d, g


def water_property_table(iwave):
    """prop_water (oracle/_ref/water_propgrid) at every (r_a, r_b, x_g)."""
    k = X.read_ket(iwave)
    R1, R2, XX = np.meshgrid(k.r1, k.r1, k.xalf, indexing="ij")
    pts = np.stack([R1.ravel(), R2.ravel(), XX.ravel()], axis=1)
    inb = (np.array([len(pts), 4, 29], "<i4").tobytes()
           + np.concatenate([k.xmass, [k.g1, k.g2]]).astype("<f8").tobytes()
           + np.array([1, 1], "<i4").tobytes() + np.array([1, 2, 3, 4], "<i4").tobytes()
           + pts.astype("<f8").tobytes())
    out = subprocess.run([os.path.join(REFD, "water_propgrid")], input=inb, capture_output=True,
                         check=True).stdout
    return dict(r=k.r1, x=k.xalf, v=np.frombuffer(out, "<f8").reshape(R1.shape + (4,)))


def main():
    subprocess.run(["make", "-s", "-C", os.path.join(REPO, "oracle"), "all", "h3p", "water", "rot",
                    "xpect"], check=True, stdout=subprocess.DEVNULL)
    root = tempfile.mkdtemp(prefix="dvr3d_xp_")
    waves = {n: run_wave(n, os.path.join(root, n)) for n in RUNS}
    gold = {}
    for name in CASES:
        _, g = run_xpect(name, root, waves)
        gold[name] = g
        print(name, "rows", len(g["table"]), "itra numbers", len(g["itra"]))
    tab = water_property_table(os.path.join(waves["water"], "fort.26"))
    np.savez_compressed(os.path.join(REPO, "tests", "golden", "water_propgrid.npz"), **tab)
    with open(os.path.join(REPO, "tests", "golden", "xp_ref.json"), "w") as f:
        json.dump(gold, f, indent=0)
    shutil.rmtree(root, ignore_errors=True)


if __name__ == "__main__":
    main()
