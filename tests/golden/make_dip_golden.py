"""Golden fixtures of the line-strength step DIPOLE3
(source/dipole3_segmented.f90) from reference runs.

Each case is a chain of reference programs (built by oracle/Makefile from the
sources where they lie): DVR3DRJZ writes the IWAVE file of every J, ROTLEV3
turns the J > 0 ones into the KVEC files DIPOLE3 reads, and DIPOLE3 pairs a
ket file with a bra file:
  h3p  the reference's own H3+ Jacobi test (H3+Jacobi/jobs: dvrJ0.job ->
       ket, dvrJ1.job + rot.job -> bra fort.8, dip.job; the reference's H3+
       POTV and dipole surface H3+Jacobi/source/dip_h3p.f): J = 0 odd ->
       J = 1 e, IDIA = 2, zembed = F
  k01  ABC deck (dvr3d_amd/jobs/k0s.job, abc_k2s.par) J = 0 -> J = 1 e
       (k1s.job + k1s_rot.job), synthetic dipole oracle/dipd_synth.f90
  k12  J = 1 e (levels 3-12) -> J = 2 e (k2s.job + k2s_rot.job)
  k22  J = 2 e -> J = 2 f (the e -> f pairing, same J; odd npot rounded up)
Written:
  tests/golden/dip_ref.json   per case: the SPECTRA record header (j1, j2,
                              kmin1, kmin2, ne1, ne2, idia, ipar1, ipar2,
                              ezero, zembed, ibase1, ibase2), the ket and
                              bra energies and the line strengths sint
                              (a.u., ne1 x ne2) of the reference's fort.13,
                              and the block lines of its stdout
  tests/golden/h3p_dipgrid.npz  the reference's H3+ dipole surface on the
                              integration grid of the h3p case (bra radial
                              grids x the 40 positive points of the 80-point
                              Gauss-Legendre rule, nu = 0 and 1), so that the
                              GPU tests need no surface code

usage: python tests/golden/make_dip_golden.py
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

from oracle import dipole3_oracle as P  # noqa: E402

JOBS = os.path.join(REPO, "dvr3d_amd", "jobs")
REFD = os.path.join(REPO, "oracle", "_ref")

# wavefunction runs: DVR3DRJZ executable, deck, surface parameters, ROTLEV3 deck
RUNS = {"h3pJ0": ("dvr3drjz_h3p", "h3p_J0.job", None, None),
        "h3pJ1": ("dvr3drjz_h3p", "h3p_J1.job", None, "h3p_rot.job"),
        "k0s": ("dvr3drjz_ref", "k0s.job", "abc_k2s.par", None),
        "k1s": ("dvr3drjz_ref", "k1s.job", "abc_k2s.par", "k1s_rot.job"),
        "k2s": ("dvr3drjz_ref", "k2s.job", "abc_k2s.par", "k2s_rot.job")}
# DIPOLE3 runs: executable, deck, (run, file) of the ket, of the bra
CASES = {"h3p": ("dipole3_h3p", "h3p_dip.job", ("h3pJ0", "fort.26"), ("h3pJ1", "fort.8")),
         "k01": ("dipole3_synth", "dip_k01.job", ("k0s", "fort.26"), ("k1s", "fort.8")),
         "k12": ("dipole3_synth", "dip_k12.job", ("k1s", "fort.8"), ("k2s", "fort.8")),
         "k22": ("dipole3_synth", "dip_k22.job", ("k2s", "fort.8"), ("k2s", "fort.9"))}


def built():
    need = {r[0] for r in RUNS.values()} | {c[0] for c in CASES.values()} | {"rotlev3_ref"}
    return all(os.path.exists(os.path.join(REFD, e)) for e in need)


def run_wave(name, d):
    """One DVR3DRJZ (+ ROTLEV3) run in directory d."""
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


def run_dipole(name, root, waves):
    """The reference DIPOLE3 of case `name` in root/<name>; waves maps run
    names to their directories.  Returns (stdout, fort.13 records)."""
    exe, job, (kr, kf), (br, bf) = CASES[name]
    d = os.path.join(root, "dip_" + name)
    os.makedirs(d, exist_ok=True)
    shutil.copy(os.path.join(waves[kr], kf), os.path.join(d, "fort.11"))
    shutil.copy(os.path.join(waves[br], bf), os.path.join(d, "fort.12"))
    with open(os.path.join(JOBS, job)) as fi:
        r = subprocess.run([os.path.join(REFD, exe)], stdin=fi, cwd=d, capture_output=True,
                           text=True, check=True, timeout=900)
    return d, r.stdout, P.read_records(os.path.join(d, "fort.13"))


def spectra_record(recs):
    """The OUTPT records (P:2000-2006)."""
    h = recs[0]
    ints = np.frombuffer(h[:36], "<i4").tolist()
    gz = float(np.frombuffer(h[36:44], "<f8")[0])
    tail = np.frombuffer(h[44:56], "<i4").tolist()
    e1 = P._d(recs[1]).tolist()
    e2 = P._d(recs[2]).tolist()
    sint = np.stack([P._d(r) for r in recs[3:]], axis=1)
    return dict(header=ints, ezero=gz, zembed=bool(tail[0]), ibase=tail[1:], e1=e1, e2=e2,
                sint=sint.tolist())


def h3p_dipole_table(bra_file):
    """The reference H3+ DIPD (oracle/_ref/h3p_dipgrid) on the h3p grid."""
    w = P.read_wave(bra_file)
    xd, _, _, _ = P.gauss_legendre(80)
    ipot = P.ipot_of(w.idia, 80)
    R2, R1, X, NU = np.meshgrid(w.r2, w.r1, xd[:ipot], [0.0, 1.0], indexing="ij")
    pts = np.stack([R1.ravel(), R2.ravel(), X.ravel(), NU.ravel()], axis=1)
    inb = (np.array([len(pts)], "<i4").tobytes()
           + np.concatenate([w.xmass, [w.g1, w.g2]]).astype("<f8").tobytes()
           + np.array([0, 0], "<i4").tobytes() + pts.astype("<f8").tobytes())
    out = subprocess.run([os.path.join(REFD, "h3p_dipgrid")], input=inb, capture_output=True,
                         check=True).stdout
    v = np.frombuffer(out, "<f8").reshape(R1.shape)
    return dict(r1=w.r1, r2=w.r2, x=xd[:ipot], v=np.moveaxis(v, 3, 0))   # (nu, i2, i1, j)


def main():
    subprocess.run(["make", "-s", "-C", os.path.join(REPO, "oracle"), "all", "h3p", "rot", "dip"],
                   check=True, stdout=subprocess.DEVNULL)
    root = tempfile.mkdtemp(prefix="dvr3d_dip_")
    waves = {n: run_wave(n, os.path.join(root, n)) for n in RUNS}
    gold = {}
    for name in CASES:
        d, out, recs = run_dipole(name, root, waves)
        g = spectra_record(recs)
        g["blocks"] = re.findall(r"Block\s+\d+ k1 =\s*\d+ to k2 =\s*\d+ (?:completed|skipped)", out)
        g["source"] = (f"oracle/_ref/{CASES[name][0]} < dvr3d_amd/jobs/{CASES[name][1]} on "
                       f"{CASES[name][2]} and {CASES[name][3]}")
        gold[name] = g
        print(name, "sint", np.array(g["sint"]).shape, "blocks", len(g["blocks"]))
    tab = h3p_dipole_table(os.path.join(waves["h3pJ1"], "fort.8"))
    np.savez_compressed(os.path.join(REPO, "tests", "golden", "h3p_dipgrid.npz"), **tab)
    with open(os.path.join(REPO, "tests", "golden", "dip_ref.json"), "w") as f:
        json.dump(gold, f, indent=1)
    shutil.rmtree(root, ignore_errors=True)


if __name__ == "__main__":
    main()
