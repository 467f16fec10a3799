"""CPU checks that pin the oracle (oracle/dvr3d_oracle.py).

Each check is chosen so that a plausible slip in the restatement fails it:
the reference executable's own output (goldens), closed forms of the
quadratures and of the L matrix, and a brute-force direct-product DVR that
the contraction must reproduce exactly when nothing is truncated.
"""
import math
import os

import numpy as np
import pytest
import scipy.special as sp

from oracle import dvr3d_oracle as O
from tests import helpers as H


def test_c0_oracle_matches_reference_golden():
    job, pes = H.job_and_pes("c0")
    r = O.run(job, pes)
    g = H.golden("c0")
    assert r["ndim2d"].tolist() == g["nham2"]
    assert r["iv2"].tolist() == g["iv2"]
    assert r["nham3"] == g["nham3"]
    ev = r["eval"] * O.AUTOCM
    # the reference prints 12 significant digits (5d24.12)
    assert np.abs(ev - np.array(g["eval_cm"])).max() < 1e-6
    assert np.abs(r["eval"] - np.array(g["eval_au"])).max() < 1e-12
    # the printed grids
    b = r["basis"]
    assert np.abs(b.r1 - np.array(g["r1"])).max() < 6e-6
    assert np.abs(b.xalf[-20:][::-1] - np.array(g["xalf_half"])).max() < 1e-15
    assert b.cswalf == pytest.approx(g["cswalf"], rel=0, abs=1e-14)


@pytest.mark.parametrize("n,alf", [(20, 44.0), (40, 44.0), (7, 3.0)])
def test_gauss_laguerre_against_scipy(n, alf):
    dn, cc = O.norms(n, alf)
    y, a, csx, csa, tsx = O.laguer(n, alf, cc)
    ys, ws = sp.roots_genlaguerre(n, alf)
    assert np.allclose(y, ys, rtol=1e-12, atol=0)
    assert np.allclose(a / a.sum(), ws / ws.sum(), rtol=1e-9, atol=0)
    assert abs(csx - tsx) / tsx < 1e-12          # sum of roots = n(n+alf)
    assert abs(csa * dn[0] ** 2 - 1.0) < 1e-9     # exact weight sum 1/dnorm0^2


@pytest.mark.parametrize("n,k", [(40, 0), (80, 0), (36, 0), (30, 1), (20, 2)])
def test_gauss_jacobi_against_library(n, k):
    tsw = O.tswalf_exact(k)
    x, w, csa = O.jacobi(n, (n + 1) // 2, float(k), float(k), tsw)
    xs, ws = sp.roots_jacobi(n, k, k)
    xs = np.sort(xs)[::-1][: (n + 1) // 2]
    ws = ws[np.argsort(-sp.roots_jacobi(n, k, k)[0])][: (n + 1) // 2]
    assert np.abs(x - xs).max() < 1e-13
    # the reference forms each weight from the derivative at the point before
    # the last Newton step (D:1716-1722), which leaves up to ~1e-8 relative
    # error in the outermost weights; the oracle keeps that on purpose
    assert np.abs(w / ws - 1.0).max() < 1e-7
    # exact integral of (1-x^2)^k over [-1,1]
    assert tsw == pytest.approx(2 ** (2 * k + 1) * math.factorial(k) ** 2 / math.factorial(2 * k + 1))


@pytest.mark.parametrize("nalf,idia,ipar", [(40, -1, 0), (24, 2, 0), (24, 2, 1)])
def test_lmatrix_spectrum_is_j_j_plus_1(nalf, idia, ipar):
    """L(a,a') = w w' sum_j j(j+1) P_j P_j' has eigenvalues j(j+1) over the
    retained j (all j for idia=-1; j of parity ipar for idia=2)."""
    job, _ = H.job_and_pes("c0")
    job = dict(job)
    job["nalf"], job["idia"], job["ipar"] = nalf, idia, ipar
    b = O.setup_basis(job)
    ev = np.sort(np.linalg.eigvalsh(b.xlmatr))
    if idia == 2:
        js = np.arange(ipar, nalf, 2)
    else:
        js = np.arange(nalf)
    ref = np.sort(js * (js + 1.0))
    # exact up to the reference's quadrature-weight rounding (~1e-11 relative
    # at nalf=40; see test_gauss_jacobi_against_library)
    assert (np.abs(ev - ref) / np.maximum(ref, 1.0)).max() < 1e-9


def test_morse_radial_dvr_closed_form():
    """1D Morse oscillator in its own Morse-like DVR: eigenvalues close to
    we(n+1/2) - we^2 (n+1/2)^2 / (4 De) (exact when A = 4De/we is an integer)."""
    mu, de, we, re = 1800.0, 0.2, 0.02, 1.8
    beta = we * math.sqrt(0.5 * mu / de)
    a = 4.0 * de / we           # = 40, an integer: the basis is exact
    iu = int(a + 0.5)
    rad = O.lagpt(40, re, beta, a, iu)
    fke = beta * beta / (8.0 * mu)
    xk = O.k1k2(O.keints(fke, 39, iu), rad["bass"], np.sqrt(rad["wt"]))
    v = de * (1.0 - np.exp(-beta * (rad["r"] - re))) ** 2
    ev = np.linalg.eigvalsh(xk + np.diag(v))
    n = np.arange(4) + 0.5
    exact = we * n - we * we * n * n / (4.0 * de)
    # the DVR quadrature error grows with n; the lowest levels are exact
    assert np.abs(ev[:4] - exact).max() < 1e-11


def test_contraction_without_truncation_equals_direct_product():
    """zall=T keeps every 1D and 2D function, so the contracted 3D spectrum
    must equal the brute-force direct-product DVR Hamiltonian's."""
    job, pes = H.job_and_pes("c0")
    job = dict(job)
    job["npnt1"] = job["npnt2"] = 5
    job["nalf"] = 6
    job["nml"] = dict(job["nml"], zall=True)
    job["neval"] = 20
    r = O.run(job, pes)
    b = r["basis"]
    s = b.sizes
    na, nb, nc = s.npnta, s.npntb, s.npntc
    V = r["vgrid"]                                  # (c, b, a)
    c = b.consts
    w1 = 0.5 / (b.r1 ** 2 * c.ur1)
    w2 = 0.5 / (b.r2 ** 2 * c.ur2)
    N = na * nb * nc
    Hf = np.zeros((N, N))
    idx = lambda ic, ib, ia: (ic * nb + ib) * na + ia
    for ic in range(nc):
        for ib in range(nb):
            for ia in range(na):
                i = idx(ic, ib, ia)
                Hf[i, i] += V[ic, ib, ia]
                for ja in range(na):
                    Hf[i, idx(ic, ib, ja)] += (w1[ic] + w2[ib]) * b.xlmatr[ia, ja]
                for jb in range(nb):
                    Hf[i, idx(ic, jb, ia)] += b.xk2[ib, jb]
                for jc in range(nc):
                    Hf[i, idx(jc, ib, ia)] += b.xk1[ic, jc]
    ref = np.linalg.eigvalsh(Hf)[:20]
    assert np.abs(r["eval"] - ref).max() < 1e-11


def test_choose_prefers_lower_block_on_ties():
    e = [np.array([0.1, 0.3]), np.array([0.1, 0.2]), np.array([0.05])]
    iv2, low3d = O.choose(e, 3)
    assert low3d == 3 and iv2.tolist() == [1, 1, 1]
    iv2, low3d = O.choose(e, 4)
    assert iv2.tolist() == [1, 2, 1]
    iv2, low3d = O.choose(e, 10)
    assert low3d == 5 and iv2.tolist() == [2, 2, 1]


def test_c1_oracle_matches_reference_golden():
    """The full c1 workload (1600 angular problems of size 80, nham3 = 4000):
    the oracle's 500 levels against the reference executable's."""
    job, pes = H.job_and_pes("c1")
    r = O.run(job, pes)
    g = H.golden("c1")
    assert r["ndim2d"].tolist() == g["nham2"]
    assert r["iv2"].tolist() == g["iv2"]
    ev = r["eval"] * O.AUTOCM
    assert len(ev) == len(g["eval_cm"]) == 500
    assert np.abs(ev - np.array(g["eval_cm"])).max() < 1e-6


def test_jrot_k_blocks_match_reference_golden():
    """J = 2 with zrot: the oracle's k loop (k = 0, 1, 2) against the levels
    the reference executable prints for every k block (k2s deck)."""
    job, pes = H.job_and_pes("k2s")
    assert O.k_blocks(job) == [0, 1, 2]
    g = H.golden("k2s")["eval_au"]
    for r, k in zip(O.run_all_k(job, pes), O.k_blocks(job)):
        ref = np.array(g[str(k)])
        assert len(r["eval"]) == len(ref)
        assert np.abs(r["eval"] - ref).max() < 1e-12            # 12 printed digits


def test_abc_surface_distance_is_cartesian_ab_distance():
    """The A-B distance of the kind-2 surface equals the Cartesian distance
    between atom 1 (A) and atom 2 (B) in Jacobi coordinates: B and C on the
    z axis r1 apart with their centre of mass at the origin (B at
    z = -(1 - g1) r1, C on the +z side), A at r2 (sin t, 0, cos t)."""
    pes = O.PES(2, 0.30, 1.30, 2.18, 0.05, 2.01, 0.0)
    m2, m3 = 12.0, 14.00307
    pes.g1 = m2 / (m2 + m3)
    rng = np.random.default_rng(3)
    r1, r2, x = rng.uniform(1.5, 3, 50), rng.uniform(1, 5, 50), rng.uniform(-1, 1, 50)
    zb = -m3 / (m2 + m3) * r1                      # B's z (centre of mass at 0)
    za, xa = r2 * x, r2 * np.sqrt(1 - x * x)
    d = np.sqrt(xa ** 2 + (za - zb) ** 2)
    v = pes(r1, r2, x)
    expect = 0.30 * ((1 - np.exp(-1.3 * (r1 - 2.18))) ** 2 + (1 - np.exp(-1.3 * (d - 2.01))) ** 2) \
        + 0.05 * (x - 1) ** 2
    assert np.abs(v - expect).max() < 1e-14


def test_noise_is_standard_normal_and_keyed_by_point():
    rng = np.random.default_rng(11)
    r1, r2, x = rng.uniform(1, 4, 200000), rng.uniform(1, 6, 200000), rng.uniform(-1, 1, 200000)
    n = O.noise_normal(r1, r2, x, O.PES.NOISE_SEED)
    assert abs(n.mean()) < 0.01 and abs(n.std() - 1.0) < 0.01
    assert abs((n ** 4).mean() - 3.0) < 0.05                 # normal kurtosis
    assert np.array_equal(n[:10], O.noise_normal(r1[:10], r2[:10], x[:10], O.PES.NOISE_SEED))
    assert not np.array_equal(n[:10], O.noise_normal(r1[:10], r2[:10], x[:10], 1))


def _h3p():
    import json
    import os
    with open(os.path.join(H.GOLD, "h3p_ref.json")) as f:
        g = json.load(f)
    return g, np.load(os.path.join(H.GOLD, "h3p_vgrid.npz"))


def test_reference_build_reproduces_shipped_h3p_result():
    """The reference build used for the goldens prints, for the reference's
    own H3+ J=0 job, the levels of its shipped result file."""
    g, _ = _h3p()
    ref = np.array(g["h3p_J0"]["0"]["eval_au"]) * O.AUTOCM
    assert np.abs(ref - np.array(g["shipped_J0_cm"])).max() < 1e-6


@pytest.mark.parametrize("name", ["h3p_J0", "h3p_J1"])
def test_h3p_jobs_match_reference(name):
    """The reference's own H3+ Jacobi jobs (IDIA=2, zr2r1=F; J=0 odd parity;
    J=1 with zembed=F over k = 0, 1), on the reference H3+ POTV tabulated on
    each k block's grid: truncation identical, 400 levels per block to the
    printed digits."""
    g, V = _h3p()
    job = O.load_job(os.path.join(H.JOBS, name + ".job"))
    for k in O.k_blocks(job):
        r = O.run(job, None, kz=k, vgrid=V[f"{name}_k{k}"])
        gb = g[name][str(k)]
        assert r["ndim2d"].tolist() == gb["nham2"]
        assert r["iv2"].tolist() == gb["iv2"] and r["nham3"] == gb["nham3"]
        assert np.abs(r["eval"] - np.array(gb["eval_au"])).max() < 1e-12


@pytest.mark.skipif(not os.path.exists(os.path.join(H.REPO, "oracle", "_ref", "dvr3drjz_fast")),
                    reason="reference executables not built (oracle/Makefile)")
def test_fast_reference_build_equals_as_shipped(tmp_path):
    """The fast reference build (rm2t = 0 loops skipped, dsyev with its
    optimal workspace; the CPU baseline of bench.py) against the as-shipped
    build on c0: stdout levels and ILEV identical, IWAVE eigenvalues within
    1e-9 cm^-1 and wavefunctions within 1e-10 (sign-aligned)."""
    import shutil
    import subprocess
    from tests import fortran_io as F
    outs = {}
    for name in ("dvr3drjz_ref", "dvr3drjz_fast"):
        d = tmp_path / name
        d.mkdir()
        shutil.copy(os.path.join(H.JOBS, "ab2_c0.par"), d / "pes.par")
        with open(os.path.join(H.JOBS, "c0.job")) as fi:
            r = subprocess.run([os.path.join(H.REPO, "oracle", "_ref", name)], stdin=fi,
                               cwd=d, capture_output=True, text=True, timeout=300, check=True)
        outs[name] = (r.stdout, (d / "fort.14").read_text(), F.read_records(d / "fort.26"))
    (sa, la, wa), (sb, lb, wb) = outs["dvr3drjz_ref"], outs["dvr3drjz_fast"]
    strip = lambda t: [l for l in t.splitlines() if "secs" not in l and "time" not in l]
    assert strip(sa) == strip(sb)
    assert la == lb
    assert len(wa) == len(wb)
    assert wa[:10] == wb[:10]
    assert np.abs(F.doubles(wa[10]) - F.doubles(wb[10])).max() * O.AUTOCM < 1e-9
    for x, y in zip(wa[11:], wb[11:]):
        x, y = F.doubles(x), F.doubles(y)
        assert np.abs(x - np.sign(x @ y) * y).max() < 1e-10
