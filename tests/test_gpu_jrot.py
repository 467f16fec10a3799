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


def _h3p():
    with open(os.path.join(H.GOLD, "h3p_ref.json")) as f:
        g = json.load(f)
    return g, np.load(os.path.join(H.GOLD, "h3p_vgrid.npz"))


@pytest.mark.parametrize("name,solver", [("h3p_J0", 0), ("h3p_J0", 1), ("h3p_J1", 0),
                                         ("h3p_J1", 1)])
def test_h3p_reference_jobs(name, solver):
    """The reference's own H3+ Jacobi jobs (IDIA=2, zr2r1=F; J=0 odd parity,
    J=1 zembed=F over k = 0, 1) with the reference H3+ POTV tabulated on each
    k block's grid: truncation identical and 400 levels per block within
    1e-4 cm-1 of the reference executable, dense and block-Davidson."""
    g, V = _h3p()
    job = O.load_job(os.path.join(H.JOBS, name + ".job"))
    for k in O.k_blocks(job):
        inp = H.device_inputs(job, None, solver=solver, kz=k)
        ctx = H.run_device(inp, analytic=False, vgrid=np.ascontiguousarray(V[f"{name}_k{k}"]).ravel())
        gb = g[name][str(k)]
        iv1, iv2, n3 = ctx.sizes()
        assert iv1.sum(axis=1).tolist() == gb["nham2"]
        assert iv2.tolist() == gb["iv2"] and n3 == gb["nham3"]
        ev = ctx.eigenvalues()
        ref = np.array(gb["eval_au"])
        assert len(ev) == len(ref) == 400
        assert np.abs(ev - ref).max() * O.AUTOCM <= 1e-4, k


def test_k_blocks_run_concurrently_match_sequential():
    """dvr_run_blocks on the two k blocks of the H3+ J=1 job gives exactly the
    levels of running each block alone."""
    from dvr3d_amd import _lib
    g, V = _h3p()
    job = O.load_job(os.path.join(H.JOBS, "h3p_J1.job"))
    ctxs, alone = [], []
    for k in O.k_blocks(job):
        inp = H.device_inputs(job, None, kz=k)
        v = np.ascontiguousarray(V[f"h3p_J1_k{k}"]).ravel()
        alone.append(H.run_device(inp, analytic=False, vgrid=v).eigenvalues())
        c = _lib.Context(inp["dims"])
        c.set_basis(inp["xlmatr"], inp["xkb"], inp["xkc"], inp["wsum"], inp["wterm"],
                    inp["emax1"], inp["emax2"])
        c.set_potential(v)
        ctxs.append(c)
    _lib.run_blocks(ctxs)
    for c, a in zip(ctxs, alone):
        assert np.array_equal(c.eigenvalues(), a)
