"""FP64 oracle levels of the c4 sweep instances that bench.py checks: the
c1 problem with every PES parameter scaled by (1 + 0.01 N(0,1)) (seed
1000 + i, bench.perturbed_par), lowest 300 levels, for the first two
instances of the dense part and the first two of the learned-preconditioner
part of the sweep.

usage: python tests/golden/make_c4_golden.py
"""
import json
import os
import sys

import numpy as np

REPO = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, REPO)

from oracle import dvr3d_oracle as O  # noqa: E402
import bench  # noqa: E402

CHECKED = (0, 1, bench.C4_TRAIN, bench.C4_TRAIN + 1)


def main():
    lines = [l for l in open(os.path.join(REPO, "dvr3d_amd", "jobs", "ab2_c0.par")).read().split("\n")
             if l.strip()]
    de, a, re_ = (float(v) for v in lines[1].split()[:3])
    kb, te, c = (float(v) for v in lines[2].split()[:3])
    base = np.array([de, a, re_, kb, te, c])
    job = O.load_job(os.path.join(REPO, "dvr3d_amd", "jobs", "c1.job"))
    job["neval"] = bench.C4_LEVELS
    out = {}
    for i in CHECKED:
        p = base * bench.perturbed_par(base, 1000 + i)
        r = O.run(job, O.PES(1, *p))
        out[str(i)] = r["eval"][: bench.C4_LEVELS].tolist()
        print("instance", i, "nham3", r["nham3"], flush=True)
    with open(os.path.join(REPO, "tests", "golden", "c4_oracle.json"), "w") as f:
        json.dump({"source": "oracle/dvr3d_oracle.py on dvr3d_amd/jobs/c1.job, perturbed PES",
                   "levels": bench.C4_LEVELS, "eval_au": out}, f)


if __name__ == "__main__":
    main()
