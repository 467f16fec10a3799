"""GPU parity of the rotational (J > 0) k-block loop of ccmain
(source/dvr3drjz_segmented.f90:857-1007): every k block runs the theta-first
jhmain path with its own associated-Legendre basis and the rotational term of
mkham1 (D:2159-2173).

k2s: a small non-symmetric ABC molecule in Jacobi coordinates (IDIA=1), J=2,
KMIN=1 (k = 0, 1, 2), checked against the reference executable's levels
(tests/golden/k2s_ref.json).  c2: the J=2 configuration with N3 ~ 10^4 per k
and hashed PES noise, checked against the FP64 oracle
(tests/golden/c2_oracle.json, written by tests/golden/make_oracle_golden.py).
"""
import json
import os

import numpy as np
import pytest

from oracle import dvr3d_oracle as O
from tests import helpers as H

pytestmark = pytest.mark.gpu


def _run_k(job, pes, k, solver):
    inp = H.device_inputs(job, pes, solver=solver, kz=k)
    basis = inp["basis"]
    pes.g1 = basis.consts.g1
    v = O.potential_grid(basis, pes)
    return H.run_device(inp, analytic=False, vgrid=np.ascontiguousarray(v).ravel())


@pytest.mark.parametrize("solver", [0, 1])
def test_k2s_every_k_matches_reference(solver):
    job, pes = H.job_and_pes("k2s")
    with open(os.path.join(H.GOLD, "k2s_ref.json")) as f:
        gold = json.load(f)["eval_au"]
    ks = O.k_blocks(job)
    assert ks == [0, 1, 2]
    for k in ks:
        ctx = _run_k(job, pes, k, solver)
        ev = ctx.eigenvalues()
        ref = np.array(gold[str(k)])
        assert len(ev) == len(ref)
        assert np.abs(ev - ref).max() * O.AUTOCM <= 1e-4, k


def test_c2_every_k_matches_oracle():
    path = os.path.join(H.GOLD, "c2_oracle.json")
    if not os.path.exists(path):
        pytest.skip("c2 oracle golden not generated")
    with open(path) as f:
        gold = json.load(f)["k"]
    job, pes = H.job_and_pes("c2")
    for k in O.k_blocks(job):
        ctx = _run_k(job, pes, k, 0)
        iv1, iv2, n3 = ctx.sizes()
        assert n3 == gold[str(k)]["nham3"]
        assert iv2.tolist() == gold[str(k)]["iv2"]
        ev = ctx.eigenvalues()
        ref = np.array(gold[str(k)]["eval_au"])
        assert len(ev) == len(ref)
        assert np.abs(ev - ref).max() * O.AUTOCM <= 1e-4, k
