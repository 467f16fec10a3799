"""GPU parity of the block-Davidson final solve (chosen form) and of its
shift-and-invert kernel.

The replaced call is the dense diag3d solve (source/dvr3drjz_segmented.f90:
2661-2764 -> 2638-2659).  Levels are checked against the reference
executable's goldens (tests/golden/*_ref.json) and against the plain dense GPU
path on the same Hamiltonian; the LDL^T inertia and solves against LAPACK.
"""
import numpy as np
import pytest

from oracle import dvr3d_oracle as O
from tests import helpers as H

pytestmark = pytest.mark.gpu

DAVIDSON = 1


def _spd_like(n, rng):
    """Symmetric matrix with a spread diagonal and dense coupling, like H3."""
    a = rng.standard_normal((n, n)) * 0.05
    a = 0.5 * (a + a.T)
    a[np.diag_indices(n)] += np.linspace(0.0, 2.0, n)
    return a


@pytest.mark.parametrize("n", [1, 7, 64, 65, 200, 700, 1500, 2203])
def test_ldl_inertia_and_solve(n):
    """Inertia and solves of the unpivoted LDL^T (ragged last panel at 2203)."""
    from dvr3d_amd import _lib
    rng = np.random.default_rng(n)
    a = _spd_like(n, rng)
    w = np.linalg.eigvalsh(a)
    k = n // 3
    sigma = 0.5 * (w[k] + w[k + 1]) if n > 1 else w[0] - 0.1
    r = rng.standard_normal((n, 5))
    nu, t = _lib.ldl_solve(a, sigma, r)
    assert nu == int((w < sigma).sum())
    res = (a - sigma * np.eye(n)) @ t - r
    cond = np.abs(w - sigma).max() / np.abs(w - sigma).min()
    assert np.abs(res).max() <= 1e-13 * cond * np.abs(r).max() * max(n, 1) ** 0.5


def test_ldl_panel_widths_agree(monkeypatch):
    """Panels of 2 and 4 sub-panels of 64 columns (trailing updates with
    K = 128 and 256) give the same factorisation: same inertia, solves equal
    to rounding."""
    from dvr3d_amd import _lib
    n = 1700
    rng = np.random.default_rng(5)
    a = _spd_like(n, rng)
    w = np.linalg.eigvalsh(a)
    sigma = 0.5 * (w[600] + w[601])
    r = rng.standard_normal((n, 3))
    out = {}
    for sub in ("1", "2", "3", "4"):
        monkeypatch.setenv("DVR_LDL_SUB", sub)
        out[sub] = _lib.ldl_solve(a, sigma, r)
    t0 = out["2"][1]
    for sub, (nu, t) in out.items():
        assert nu == 601, sub
        assert np.abs(t - t0).max() <= 1e-9 * np.abs(t0).max(), sub
    # the XCD-aware order of the trailing-update tiles changes only which
    # work-group computes which tile: bit-identical results
    monkeypatch.setenv("DVR_LDL_SUB", "4")
    monkeypatch.setenv("DVR_LDL_XCD", "0")
    nu0, tx0 = _lib.ldl_solve(a, sigma, r)
    monkeypatch.setenv("DVR_LDL_XCD", "1")
    nu1, tx1 = _lib.ldl_solve(a, sigma, r)
    assert nu0 == nu1 == 601 and np.array_equal(tx0, tx1)


@pytest.mark.parametrize("n", [200, 1700, 2203, 4133])
def test_ldl_persistent_update_same_bits(monkeypatch, n):
    """The trailing update on the persistent wavefront-tile kernel (32, 28 and
    5 work-groups per XCD; K = 256, 128, 64 from panels of 4, 2, 1
    sub-panels; its M % 64 ragged rows on the 64 x 64 kernel; tile orders in
    super-tiles of 16, 7 and 24 tiles) sums in the order of the 64 x 64
    kernel: inertia (LAPACK's count) and solves bit-identical to
    DVR_LDL_WIDE=0.  (The shift sits mid-spectrum, where
    the unpivoted factorisation has element growth: the accuracy of the
    solves is tested on the cases above and below.)"""
    from dvr3d_amd import _lib
    rng = np.random.default_rng(n + 3)
    a = _spd_like(n, rng)
    w = np.linalg.eigvalsh(a)
    sigma = 0.5 * (w[n // 2] + w[n // 2 + 1])
    r = rng.standard_normal((n, 3))
    for sub in ("4", "2", "1"):
        monkeypatch.setenv("DVR_LDL_SUB", sub)
        monkeypatch.setenv("DVR_LDL_WIDE", "0")
        nu0, t0 = _lib.ldl_solve(a, sigma, r)
        assert nu0 == n // 2 + 1
        for wgs in ("32", "28", "5"):
            monkeypatch.setenv("DVR_LDL_WIDE", wgs)
            nu, t = _lib.ldl_solve(a, sigma, r)
            assert nu == nu0 and np.array_equal(t, t0), (sub, wgs)
        # other super-tile widths of the tile order: same tiles, same sums
        monkeypatch.setenv("DVR_LDL_WIDE", "32")
        for st in ("7", "24"):
            monkeypatch.setenv("DVR_LDL_ST", st)
            nu, t = _lib.ldl_solve(a, sigma, r)
            assert nu == nu0 and np.array_equal(t, t0), (sub, st)
        monkeypatch.delenv("DVR_LDL_ST")


@pytest.mark.parametrize("sb", ["512", "1024", "2048"])
def test_ldl_super_block_sizes(monkeypatch, sb):
    """Triangular solves over super-blocks of 512 / 1024 / 2048 rows (the
    full super-blocks' inverses formed together as one strided batch, the
    ragged last one alone) agree with the dense solve."""
    from dvr3d_amd import _lib
    monkeypatch.setenv("DVR_LDL_SB", sb)
    n = 2203
    rng = np.random.default_rng(11)
    a = _spd_like(n, rng)
    w = np.linalg.eigvalsh(a)
    sigma = 0.5 * (w[700] + w[701])
    r = rng.standard_normal((n, 4))
    nu, t = _lib.ldl_solve(a, sigma, r)
    assert nu == 701
    ref = np.linalg.solve(a - sigma * np.eye(n), r)
    cond = np.abs(w - sigma).max() / np.abs(w - sigma).min()
    assert np.abs(t - ref).max() <= 1e-13 * cond * np.abs(ref).max() * n ** 0.5


def test_ldl_inertia_on_c0_hamiltonian():
    """Eigenvalue counts of the c0 H3 at shifts between its levels."""
    from dvr3d_amd import _lib
    job, pes = H.job_and_pes("c0")
    ref = O.run(job, pes, keep_vectors=True, neval_vectors=1)
    h3 = ref["H3"]
    w = np.linalg.eigvalsh(h3)
    for k in (0, 99, 400):
        sigma = 0.5 * (w[k] + w[k + 1])
        nu, _ = _lib.ldl_solve(h3, sigma)
        assert nu == k + 1


@pytest.mark.parametrize("cfg", ["c0", "c1"])
def test_davidson_levels_match_reference(cfg):
    job, pes = H.job_and_pes(cfg)
    inp = H.device_inputs(job, pes, solver=DAVIDSON)
    ctx = H.run_device(inp)
    st = ctx.solver_stats()
    assert st["status"] == 0 and st["dense_fallback"] == 0, st
    assert st["certified"] >= len(ctx.eigenvalues())
    ev = ctx.eigenvalues() * O.AUTOCM
    gold = np.array(H.golden(cfg)["eval_cm"])
    assert len(ev) == len(gold)
    assert np.abs(ev - gold).max() <= 1e-4


def test_davidson_block_wider_than_one_gram_factorisation(monkeypatch):
    """Correction blocks of 128 (orthonormalised in two chunks of 64): c1
    levels unchanged against the reference."""
    monkeypatch.setenv("DVR_DAV_BLOCK", "128")
    monkeypatch.setenv("DVR_DAV_MMAX", "512")
    monkeypatch.setenv("DVR_DAV_KEEP", "320")
    job, pes = H.job_and_pes("c1")
    ctx = H.run_device(H.device_inputs(job, pes, solver=DAVIDSON))
    st = ctx.solver_stats()
    assert st["status"] == 0 and st["dense_fallback"] == 0, st
    ev = ctx.eigenvalues() * O.AUTOCM
    gold = np.array(H.golden("c1")["eval_cm"])
    assert len(ev) == len(gold) and np.abs(ev - gold).max() <= 1e-4


@pytest.mark.parametrize("early,slice_,first", [("0", "256", "6"), ("1e-3", "96", "6"),
                                                 ("1e-1", "64", "6"), ("1e-1", "256", "6"),
                                                 ("1e-4", "256", "0"), ("1e-4", "256", "30")])
def test_davidson_early_shift_paths(early, slice_, first, monkeypatch):
    """The next shift factored on a second stream while a slice converges
    (DVR_DAV_EARLY = residual that triggers it; 0 = after the slice, on the
    solver's stream).  With rough Ritz values (1e-1) the shift lands anywhere:
    the slice then takes every level below it, or, when the locked space fills
    first, is certified at its top as without the early shift.  c1 levels
    unchanged against the reference in every case.  DVR_DAV_FIRST = F > 0
    also factors the first shift on the second stream, with diagonal
    corrections for the first F iterations; 0 factors it before the first
    iteration."""
    monkeypatch.setenv("DVR_DAV_EARLY", early)
    monkeypatch.setenv("DVR_DAV_FIRST", first)
    monkeypatch.setenv("DVR_DAV_SLICE", slice_)
    job, pes = H.job_and_pes("c1")
    ctx = H.run_device(H.device_inputs(job, pes, solver=DAVIDSON))
    st = ctx.solver_stats()
    assert st["status"] == 0 and st["dense_fallback"] == 0, st
    ev = ctx.eigenvalues() * O.AUTOCM
    gold = np.array(H.golden("c1")["eval_cm"])
    assert st["certified"] >= len(gold)
    assert len(ev) == len(gold) and np.abs(ev - gold).max() <= 1e-4


def test_davidson_wavefunctions_match_dense(monkeypatch):
    monkeypatch.setenv("DVR_DAV_TOL", "1e-10")
    job, pes = H.job_and_pes("c0")
    nvec = 20
    dense = H.run_device(H.device_inputs(job, pes, nvec=nvec, solver=0))
    dav = H.run_device(H.device_inputs(job, pes, nvec=nvec, solver=DAVIDSON))
    assert dav.solver_stats()["status"] == 0
    pd = dense.wavefunctions(nvec).reshape(nvec, -1)
    pv = dav.wavefunctions(nvec).reshape(nvec, -1)
    for k in range(nvec):
        s = np.sign(pd[k] @ pv[k])
        assert np.abs(pd[k] - s * pv[k]).max() <= 1e-5 * np.abs(pd[k]).max()


@pytest.mark.parametrize("cfg", ["c0", "c1"])
def test_split32_assembly_refined_in_fp64(cfg, monkeypatch):
    """H3 assembled with FP32 split (hi + lo) products, Davidson on it, then the
    Rayleigh-Ritz refinement with the exact FP64 operator: levels back within
    1e-4 cm^-1 of the reference."""
    monkeypatch.setenv("DVR_ASSEMBLY", "split32")
    job, pes = H.job_and_pes(cfg)
    ctx = H.run_device(H.device_inputs(job, pes, solver=DAVIDSON))
    assert ctx.solver_stats()["status"] == 0
    ev = ctx.eigenvalues() * O.AUTOCM
    gold = np.array(H.golden(cfg)["eval_cm"])
    assert np.abs(ev - gold).max() <= 1e-4


@pytest.mark.parametrize("cfg", ["c0", "c1"])
def test_split32_dense_refined_by_rayleigh_quotients(cfg, monkeypatch):
    """Every level wanted (dense solver): H3 assembled with FP32 split
    products, solved densely, each level refined by its Rayleigh quotient
    with the exact FP64 operator: levels within 1e-4 cm^-1 of the reference,
    and the eigenvectors (in the refined order) span the FP64 dense run's:
    |overlap| of each pair 1 to 1e-6 where the level is 0.01 cm^-1 away from
    its neighbours."""
    job, pes = H.job_and_pes(cfg)
    gold = np.array(H.golden(cfg)["eval_cm"])
    nvec = 20
    monkeypatch.setenv("DVR_ASSEMBLY", "fp64")
    ref = H.run_device(H.device_inputs(job, pes, nvec=nvec, solver=0))
    monkeypatch.setenv("DVR_ASSEMBLY", "split32")
    ctx = H.run_device(H.device_inputs(job, pes, nvec=nvec, solver=0))
    ev = ctx.eigenvalues() * O.AUTOCM
    err = np.abs(ev - gold)
    print(cfg, "max |split32 + Rayleigh quotients - reference| cm^-1:", err.max())
    assert len(ev) == len(gold) and err.max() <= 1e-4
    assert np.all(np.diff(ev) >= 0)
    # the wavefunctions are an isometric image of the eigenvectors
    za = ref.wavefunctions(nvec).reshape(nvec, -1)
    zb = ctx.wavefunctions(nvec).reshape(nvec, -1)
    ov = np.abs(np.einsum("ij,ij->i", za, zb))
    e = gold[:nvec + 1]
    iso = np.ones(nvec, bool)
    iso[1:] &= np.diff(e[:nvec]) > 1e-2
    iso &= np.diff(e) > 1e-2
    assert np.abs(ov[iso] - 1.0).max() <= 1e-6, ov


def test_graph_replay_and_tuner_match_eager(monkeypatch):
    """The shift-and-invert solves replayed from HIP graphs give the same levels
    as eager launches (same kernels, bit for bit), and the tuned GEMM shapes
    agree with the fixed rule to rounding."""
    job, pes = H.job_and_pes("c1")
    inp = H.device_inputs(job, pes, solver=DAVIDSON)
    runs = {}
    for graph, tune in ((0, 0), (1, 0), (1, 1)):
        monkeypatch.setenv("DVR_GRAPH", str(graph))
        monkeypatch.setenv("DVR_TUNE", str(tune))
        ctx = H.run_device(inp)
        assert ctx.solver_stats()["status"] == 0
        runs[(graph, tune)] = ctx.eigenvalues() * O.AUTOCM
        # a second solve in the same context replays the captured graphs
        ctx.run()
        d = np.abs(ctx.eigenvalues() * O.AUTOCM - runs[(graph, tune)]).max()
        print(f"graph {graph} tune {tune}: rerun differs by {d:.3e} cm-1")
        assert np.array_equal(ctx.eigenvalues() * O.AUTOCM, runs[(graph, tune)])
    assert np.array_equal(runs[(0, 0)], runs[(1, 0)])
    assert np.abs(runs[(1, 1)] - runs[(0, 0)]).max() <= 1e-6


N3_RERUN = (784, 621)


def test_davidson_context_rerun_with_smaller_basis():
    """One Davidson context run twice on a zcut=T deck (c0 grid, EMAX2 =
    30000 cm^-1) with two potentials whose final dimensions differ: the
    second solve must not reuse anything sized or captured for the
    first (device buffers do not shrink, so the pointers stay the same).
    Levels within 1e-4 cm^-1 of the oracle, certified, without fallback."""
    import copy
    job, pes = H.job_and_pes("c0")
    job = copy.deepcopy(job)
    job["nml"]["zcut"] = True
    job["emax2"] = 30000.0
    pes2 = copy.deepcopy(pes)
    pes2.kb *= 1.3
    pes2.de *= 1.1
    inp = H.device_inputs(job, pes, solver=DAVIDSON)
    ctx = H.run_device(inp)
    for k, p in enumerate((pes, pes2)):
        if k:
            ctx.eval_potential_ab2(inp["rc"], inp["rb"], inp["xa"],
                                   np.array([p.de, p.a, p.re, p.kb, p.cte, p.c]))
            ctx.run()
        ref = O.run(job, p)
        st = ctx.solver_stats()
        n3 = ctx.sizes()[2]
        print(f"potential {k}: nham3 {n3}, {st['iterations']} iterations, status {st['status']}")
        assert n3 == ref["nham3"] == (N3_RERUN[k] if N3_RERUN else n3)
        assert st["status"] == 0 and st["dense_fallback"] == 0, st
        assert st["iterations"] < 100
        ev = ctx.eigenvalues()
        assert np.abs(ev - ref["eval"][: len(ev)]).max() * O.AUTOCM <= 1e-4


@pytest.mark.parametrize("k", [1, 2, 3, 4])
@pytest.mark.parametrize("first", ["0", "6"])
def test_davidson_recovers_from_ambiguous_inertia(k, first, monkeypatch):
    """A factorisation whose inertia count is ambiguous (a near-zero pivot:
    the shift almost on a level) is repeated at a shift nudged within the
    same gap instead of handing the problem to the dense path.  The k-th
    count of the c1 solve is forced ambiguous from the host
    (DVR_DAV_FORCE_AMBIG = k; k = 1 the first shift, later ones the next
    shifts and the certification): the solve still certifies every level,
    with one factorisation more, and matches the reference's levels; with the
    first shift factored beside the first iterations (DVR_DAV_FIRST = 6) as
    well."""
    monkeypatch.setenv("DVR_DAV_FIRST", first)
    job, pes = H.job_and_pes("c1")
    monkeypatch.setenv("DVR_DAV_FORCE_AMBIG", "0")
    base = H.run_device(H.device_inputs(job, pes, solver=DAVIDSON)).solver_stats()
    monkeypatch.setenv("DVR_DAV_FORCE_AMBIG", str(k))
    ctx = H.run_device(H.device_inputs(job, pes, solver=DAVIDSON))
    st = ctx.solver_stats()
    print(k, first, "factorisations", base["factorizations"], "->", st["factorizations"])
    assert st["status"] == 0 and st["dense_fallback"] == 0, st
    ev = ctx.eigenvalues() * O.AUTOCM
    gold = np.array(H.golden("c1")["eval_cm"])
    assert len(ev) == len(gold) and np.abs(ev - gold).max() <= 1e-4
    if k <= base["factorizations"]:
        assert st["factorizations"] == base["factorizations"] + 1, (base, st)


def test_ldl_solve_zero_pivot_reported():
    """A shift on an eigenvalue gives a zero pivot: the inertia is ambiguous
    and dvr_ldl_solve reports DVR_ERR_NOCONV instead of a count."""
    from dvr3d_amd import _lib
    a = np.diag(np.arange(1.0, 9.0))
    with pytest.raises(_lib.DvrError) as e:
        _lib.ldl_solve(a, 3.0)
    assert e.value.code == _lib.DVR_ERR_NOCONV
    nu, _ = _lib.ldl_solve(a, 3.5)
    assert nu == 3


def test_davidson_too_small_problem_goes_dense():
    """N3 = 244 with 100 levels wanted (zcut=T, EMAX2 = 18000 cm^-1 on the c0
    grid) leaves no room for the search space: the dense path solves it
    (status 5) and the levels match the oracle."""
    import copy
    job, pes = H.job_and_pes("c0")
    job = copy.deepcopy(job)
    job["nml"]["zcut"] = True
    job["emax2"] = 18000.0
    ctx = H.run_device(H.device_inputs(job, pes, solver=DAVIDSON))
    st = ctx.solver_stats()
    assert st["status"] == 5 and st["dense_fallback"] == 1, st
    ref = O.run(job, pes)
    assert ctx.sizes()[2] == ref["nham3"] == 244
    ev = ctx.eigenvalues()
    assert np.abs(ev - ref["eval"][: len(ev)]).max() * O.AUTOCM <= 1e-4
