"""Golden levels from the oracle (numpy restatement) for configs where the
reference executable is too slow to run as shipped (its dead r2**-2 loop,
D:2464-2475, costs ~4e14 flops at c3).  The oracle itself is pinned against
the reference executable on c0 and c1 (tests/test_oracle.py).

usage: python tests/golden/make_oracle_golden.py c3
       python tests/golden/make_oracle_golden.py c2     (J > 0: one entry per k)
"""
import json
import os
import sys
import time

REPO = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, REPO)

from oracle import dvr3d_oracle as O  # noqa: E402

PES_OF = {"c0": "ab2_c0.par", "c1": "ab2_c0.par", "c3": "ab2_c3.par", "c2": "abc_c2.par"}


def main(cfg):
    job = O.load_job(os.path.join(REPO, "dvr3d_amd", "jobs", cfg + ".job"))
    pes = O.PES.from_par(open(os.path.join(REPO, "dvr3d_amd", "jobs", PES_OF[cfg])).read())
    t0 = time.time()
    ks = O.k_blocks(job)
    if len(ks) > 1:
        blocks = {}
        for k in ks:
            r = O.run(job, pes, kz=k)
            blocks[str(k)] = dict(nham2=r["ndim2d"].tolist(), iv2=r["iv2"].tolist(),
                                  nham3=int(r["nham3"]), eval_au=r["eval"].tolist())
            print(cfg, "k", k, "nham3", r["nham3"], "levels", len(r["eval"]), flush=True)
        out = dict(source=f"oracle/dvr3d_oracle.py on dvr3d_amd/jobs/{cfg}.job", jrot=job["jrot"],
                   k=blocks, oracle_wall_s=time.time() - t0)
        with open(os.path.join(REPO, "tests", "golden", cfg + "_oracle.json"), "w") as f:
            json.dump(out, f)
        return
    r = O.run(job, pes)
    out = dict(source=f"oracle/dvr3d_oracle.py on dvr3d_amd/jobs/{cfg}.job",
               nham2=r["ndim2d"].tolist(), iv2=r["iv2"].tolist(), nham3=int(r["nham3"]),
               eval_au=r["eval"].tolist(), eval_cm=(r["eval"] * O.AUTOCM).tolist(),
               oracle_wall_s=time.time() - t0)
    with open(os.path.join(REPO, "tests", "golden", cfg + "_oracle.json"), "w") as f:
        json.dump(out, f, indent=1)
    print(cfg, "levels", len(out["eval_cm"]), "nham3", out["nham3"], "wall", out["oracle_wall_s"])


if __name__ == "__main__":
    main(sys.argv[1])
