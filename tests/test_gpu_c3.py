"""GPU parity of the headline workload c3 and of the c4 sweep's code path.

c3: AB2 J=0, NPNT1 = NPNT2 = 64, NALF = 128, final dimension N3 = 20000,
lowest 1000 levels (the bench workload).  The FP64 oracle's truncation
(nham2 per r1 block, iv2, N3) and levels (tests/golden/c3_oracle.json) are
the reference: the final solve replaces diag3d's dsyev (source/
dvr3drjz_segmented.f90:2661-2764 -> diag D:2638-2659), the truncation is
choose (D:2767-2867).

c4: the sweep of bench.py on its checked instances (0, 1: dense solves that
train the learned preconditioner; 32, 33: block-Davidson with it, dense
fallback allowed) against tests/golden/c4_oracle.json.
"""
import json
import os

import numpy as np
import pytest

from oracle import dvr3d_oracle as O
from tests import helpers as H

pytestmark = pytest.mark.gpu

DAVIDSON = 1


@pytest.fixture(scope="module")
def c3():
    job, pes = H.job_and_pes("c3")
    with open(os.path.join(H.GOLD, "c3_oracle.json")) as f:
        gold = json.load(f)
    return job, pes, gold


def _check_truncation(ctx, gold):
    iv1, iv2, n3 = ctx.sizes()
    assert n3 == gold["nham3"] == 20000
    assert iv2.tolist() == gold["iv2"]
    # nham2 of r1 block c = sum over r2 points b of iv1(c, b) (D:2021-2025)
    assert iv1.sum(axis=1).tolist() == gold["nham2"]
    assert len(ctx.eigenvalues_2d()) == sum(gold["nham2"])


def test_c3_chosen_form_levels_and_truncation(c3):
    """Chosen form (default settings): truncation identical to the oracle and
    1000/1000 levels within 1e-3 cm^-1, certified, no dense fallback."""
    job, pes, gold = c3
    ctx = H.run_device(H.device_inputs(job, pes, solver=DAVIDSON))
    _check_truncation(ctx, gold)
    st = ctx.solver_stats()
    assert st["status"] == 0 and st["dense_fallback"] == 0, st
    assert st["certified"] >= 1000
    ev = ctx.eigenvalues() * O.AUTOCM
    ref = np.array(gold["eval_cm"])
    assert len(ev) == len(ref) == 1000
    err = np.abs(ev - ref).max()
    print(f"c3 chosen form: max |dE| = {err:.2e} cm-1, {st['iterations']} iterations")
    assert err <= 1e-3


def test_c3_plain_dense_levels(c3):
    """Plain FP64 dense path: 1000 levels within 1e-4 cm^-1 of the oracle."""
    job, pes, gold = c3
    ctx = H.run_device(H.device_inputs(job, pes, solver=0))
    _check_truncation(ctx, gold)
    ev = ctx.eigenvalues() * O.AUTOCM
    err = np.abs(ev - np.array(gold["eval_cm"])).max()
    print(f"c3 plain dense: max |dE| = {err:.2e} cm-1")
    assert err <= 1e-4


def test_c3_split32_assembly(c3, monkeypatch):
    """Split-precision FP32 assembly with FP64 refinement: within 1e-3 cm^-1."""
    monkeypatch.setenv("DVR_ASSEMBLY", "split32")
    job, pes, gold = c3
    ctx = H.run_device(H.device_inputs(job, pes, solver=DAVIDSON))
    assert ctx.solver_stats()["status"] == 0
    ev = ctx.eigenvalues() * O.AUTOCM
    err = np.abs(ev - np.array(gold["eval_cm"])).max()
    print(f"c3 split32: max |dE| = {err:.2e} cm-1")
    assert err <= 1e-3


def test_c4_sweep_checked_instances(monkeypatch):
    """The c4 sweep's code path on instances 0, 1 (dense, Hamiltonian kept,
    learned preconditioner trained on their eigenpairs) and 32, 33
    (block-Davidson with the learned preconditioner, iteration cap 100, dense
    fallback): 300 levels each within 1e-4 cm^-1 of the FP64 oracle."""
    import bench
    from dvr3d_amd import _lib
    monkeypatch.setenv("DVR_DAV_MAXIT", "100")
    with open(os.path.join(H.GOLD, "c4_oracle.json")) as f:
        gold = json.load(f)["eval_au"]
    b, xa, context, params = bench.c4_setup()
    dense = context(_lib.DVR_SOLVER_DENSE, bench.C4_LEVELS)
    dense.keep_hamiltonian(True)
    mlp = _lib.mlp_init(0)
    for i in (0, 1):
        dense.eval_potential_ab2(b.rc, b.rb, xa, params(i))
        dense.run()
        err = np.abs(dense.eigenvalues() - np.array(gold[str(i)])).max() * O.AUTOCM
        print(f"c4 instance {i} (dense): max |dE| = {err:.2e} cm-1")
        assert err <= 1e-4
        mlp, hist = dense.mlp_train(mlp, steps=40, batch=32, seed=i, lr=2e-3)
        assert np.all(np.isfinite(hist))
    learned = context(_lib.DVR_SOLVER_DAVIDSON, 0)
    learned.set_preconditioner(_lib.PREC_LEARNED, mlp)
    for i in (bench.C4_TRAIN, bench.C4_TRAIN + 1):
        learned.eval_potential_ab2(b.rc, b.rb, xa, params(i))
        learned.run()
        st = learned.solver_stats()
        err = np.abs(learned.eigenvalues() - np.array(gold[str(i)])).max() * O.AUTOCM
        print(f"c4 instance {i} (learned Davidson): status {st['status']}, "
              f"fallback {st['dense_fallback']}, max |dE| = {err:.2e} cm-1")
        assert len(learned.eigenvalues()) == bench.C4_LEVELS
        assert err <= 1e-4
