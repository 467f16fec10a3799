"""CPU tests of the host side: the C ABI library loads and exports every
declared entry point, and the Fortran host reproduces the oracle's basis."""
import os
import re
import subprocess
import sys

import numpy as np
import pytest

from oracle import dvr3d_oracle as O
from tests import helpers as H

REPO = H.REPO


@pytest.fixture(scope="session", autouse=True)
def built():
    from dvr3d_amd import build as B
    need = [os.path.join(REPO, "dvr3d_amd", f) for f in ("libdvr3d_amd.so", "libdvr3d_host.so")]
    if not all(os.path.exists(p) for p in need):
        B.build()


def declared_functions():
    text = open(os.path.join(REPO, "include", "dvr3d_amd.h")).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(dvr_\w+)\s*\(", text)))


def test_library_exports_every_declared_symbol():
    import ctypes
    from dvr3d_amd import _lib
    lib = ctypes.CDLL(_lib.LIB_PATH)
    names = declared_functions()
    assert len(names) >= 18
    for n in names:
        assert hasattr(lib, n), n
    assert set(names) == set(_lib.EXPORTS)


def test_load_without_gpu_only_binds():
    from dvr3d_amd import _lib
    lib = _lib.load()
    assert lib.dvr_last_error(None) == b"null context"


@pytest.mark.parametrize("cfg", ["c0", "c1"])
def test_fortran_host_basis_matches_oracle(cfg):
    from dvr3d_amd import host
    job = host.read_job(os.path.join(H.JOBS, cfg + ".job"))
    b = host.Basis(job)
    ob = O.setup_basis(O.load_job(os.path.join(H.JOBS, cfg + ".job")))
    for name in ("r1", "r2", "xalf", "walf", "xlmatr", "pleg"):
        assert np.array_equal(np.asarray(getattr(b, name)), np.asarray(getattr(ob, name))), name
    # k1k2 sums in a different order: a few ulps
    for name in ("xk1", "xk2"):
        x, y = getattr(b, name), getattr(ob, name)
        assert np.abs(x - y).max() <= 4e-16 * np.abs(y).max(), name
    c, oc = b.consts, ob.consts
    for f in ("g1", "g2", "ur1", "ur2", "urr1", "urr2", "beta1", "beta2", "a1", "a2"):
        assert getattr(c, f) == getattr(oc, f), f
    assert (c.iu1, c.iu2) == (oc.iu1, oc.iu2)
    assert c.cswalf == ob.cswalf
    s, os_ = b.sz, ob.sizes
    assert (s.npnta, s.npntb, s.npntc, s.max2d, s.max3d, s.neval, s.idvr) == \
        (os_.npnta, os_.npntb, os_.npntc, os_.max2d, os_.max3d, os_.neval, os_.idvr)
    # (w1 + w2) at every (r1, r2) point, mkham1 D:2155-2157
    w1 = 0.5 / (ob.r1 * ob.r1 * oc.ur1)
    w2 = 0.5 / (ob.r2 * ob.r2 * oc.ur2)
    assert np.array_equal(b.wsum, (w1[:, None] + w2[None, :]).ravel())
    assert not b.wterm.any()


def test_job_reader_fields():
    from dvr3d_amd import host
    job = host.read_job(os.path.join(H.JOBS, "c1.job"))
    assert (job.npnt1, job.npnt2, job.nalf, job.max2d, job.max3d, job.neval, job.idia) == \
        (40, 40, 80, 3200, 4000, 500, -1)
    assert job.ztheta == 1 and job.zr2r1 == 1 and job.ztran == 1 and job.zpfun == 1
    assert job.emax1 == 26336.95 and list(job.xmass) == [1.00783, 1.00783, 15.99491]
    assert job.title.decode().startswith(" c1: AB2")
    oj = O.load_job(os.path.join(H.JOBS, "c1.job"))
    assert (oj["re1"], oj["diss1"], oj["we1"]) == (job.re1, job.diss1, job.we1)


def test_sizes_reject_paths_outside_theta_first(tmp_path):
    from dvr3d_amd import host
    text = open(os.path.join(H.JOBS, "c0.job")).read()
    text = text.replace("   -1    0   20", "   -2    0   20")   # idia = -2 -> nfmain
    p = tmp_path / "nf.job"
    p.write_text(text)
    job = host.read_job(str(p))
    assert job.idia == -2
    with pytest.raises(ValueError):
        host.sizes(job)


def test_bench_refuses_several_gpus():
    r = subprocess.run([sys.executable, os.path.join(REPO, "bench.py"), "--gpus", "2"],
                       stdout=subprocess.PIPE, stderr=subprocess.PIPE, text=True)
    assert r.returncode == 2 and "one GPU" in r.stdout


@pytest.mark.parametrize("k", [0, 1, 2])
def test_fortran_host_jrot_blocks_match_oracle(k):
    """J = 2, zrot, each k block of the ccmain loop (D:857-1007): the
    associated-Legendre basis and the rotational term of mkham1 (D:2159-2173)
    from the Fortran host equal the oracle's."""
    from dvr3d_amd import host
    job = host.read_job(os.path.join(H.JOBS, "k2s.job"))
    sz = host.sizes(job)
    assert (sz.kd, sz.ku, sz.kmin, sz.zrot) == (0, 2, 1, 1)
    b = host.Basis(job, k)
    ojob = O.load_job(os.path.join(H.JOBS, "k2s.job"))
    ob = O.setup_basis(ojob, k)
    for name in ("xalf", "walf"):
        assert np.abs(np.asarray(getattr(b, name)) - getattr(ob, name)).max() < 1e-15, name
    for name in ("xlmatr", "pleg"):
        x, y = np.asarray(getattr(b, name)), np.asarray(getattr(ob, name))
        assert np.abs(x - y).max() <= 1e-13 * max(1.0, np.abs(y).max()), name
    _, pes = H.job_and_pes("k2s")
    inp = H.device_inputs(ojob, pes, kz=k)
    assert np.abs(b.wterm - inp["wterm"]).max() <= 1e-15 * np.abs(inp["wterm"]).max()
    assert np.array_equal(b.wsum, inp["wsum"])


def test_example_potv_noise_equals_oracle_noise(tmp_path):
    """The hashed PES noise of the example POTV (C, linked into the
    executable) and the oracle's numpy restatement give the same numbers."""
    import ctypes
    import subprocess
    so = str(tmp_path / "noise.so")
    subprocess.run(["gcc", "-O2", "-shared", "-fPIC",
                    os.path.join(H.REPO, "dvr3d_amd", "host", "potv_noise.c"), "-o", so, "-lm"],
                   check=True)
    lib = ctypes.CDLL(so)
    f = lib.dvr_synth_normal
    f.restype = ctypes.c_double
    f.argtypes = [ctypes.c_double] * 3
    rng = np.random.default_rng(7)
    r1, r2, x = rng.uniform(1, 4, 500), rng.uniform(1, 6, 500), rng.uniform(-1, 1, 500)
    a = np.array([f(*t) for t in zip(r1, r2, x)])
    b = O.noise_normal(r1, r2, x, O.PES.NOISE_SEED)
    assert np.abs(a - b).max() < 1e-14


@pytest.mark.parametrize("name,k", [("h3p_J0", 0), ("h3p_J1", 0), ("h3p_J1", 1)])
def test_fortran_host_h3p_blocks_match_oracle(name, k):
    """The reference's H3+ jobs (IDIA=2, zr2r1=F, zembed=F at J=1): grids,
    L matrix and the (w1 + w2) / rotational terms of the Fortran host equal
    the oracle's."""
    from dvr3d_amd import host
    job = host.read_job(os.path.join(H.JOBS, name + ".job"))
    b = host.Basis(job, k)
    ojob = O.load_job(os.path.join(H.JOBS, name + ".job"))
    ob = O.setup_basis(ojob, k)
    for nm in ("r1", "r2", "xalf", "walf"):
        assert np.abs(np.asarray(getattr(b, nm)) - getattr(ob, nm)).max() < 1e-15, nm
    x, y = np.asarray(b.xlmatr), ob.xlmatr
    assert np.abs(x - y).max() <= 1e-13 * np.abs(y).max()
    inp = H.device_inputs(ojob, None, kz=k)
    assert np.abs(b.wsum - inp["wsum"]).max() <= 1e-15 * np.abs(inp["wsum"]).max()
    if inp["wterm"] is None:
        assert not b.wterm.any()
    else:
        assert np.abs(b.wterm - inp["wterm"]).max() <= 1e-15 * np.abs(inp["wterm"]).max()


def test_host_potential_grid_equals_oracle_abc_surface(monkeypatch):
    """The example POTV through dvrh_potential_grid (the loop of mkham1 in
    the Fortran host) on the c2 grid of every k block equals the oracle's
    ABC surface with its hashed noise."""
    from dvr3d_amd import host
    monkeypatch.setenv("DVR_PES", os.path.join(H.JOBS, "abc_c2.par"))
    job = host.read_job(os.path.join(H.JOBS, "c2.job"))
    ojob, pes = H.job_and_pes("c2")
    for k in O.k_blocks(ojob):
        b = host.Basis(job, k)
        v = b.potential_grid()
        ob = O.setup_basis(ojob, k)
        pes.g1 = ob.consts.g1
        ref = O.potential_grid(ob, pes)
        assert v.shape == ref.shape
        assert np.abs(v - ref).max() <= 1e-14 * np.abs(ref).max(), k


def test_run_blocks_rejects_bad_arguments_without_a_gpu():
    """dvr_run_blocks checks its arguments before touching the device: a null
    array, a null context or the same context twice."""
    import ctypes
    from dvr3d_amd import _lib
    lib = _lib.load()
    assert lib.dvr_run_blocks(None, 1) == _lib.DVR_ERR_ARG
    arr = (ctypes.c_void_p * 2)(None, None)
    assert lib.dvr_run_blocks(arr, 2) == _lib.DVR_ERR_ARG
    fake = ctypes.c_void_p(0x1000)
    arr = (ctypes.c_void_p * 2)(fake, fake)
    assert lib.dvr_run_blocks(arr, 2) == _lib.DVR_ERR_ARG
    assert lib.dvr_run_blocks(arr, -1) == _lib.DVR_ERR_ARG
