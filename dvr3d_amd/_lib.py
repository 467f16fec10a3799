"""ctypes binding of libdvr3d_amd.so (include/dvr3d_amd.h).

The library is built in-tree by ``dvr3d_amd.build.build()``; loading fails
loudly when it is missing -- there is no CPU fallback.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libdvr3d_amd.so")

DVR_OK = 0
DVR_ERR_ARG = 1
DVR_ERR_HIP = 2
DVR_ERR_CORE_2D = 3
DVR_ERR_CORE_3D = 4
DVR_ERR_NOCONV = 5
DVR_SOLVER_DENSE = 0
DVR_SOLVER_DAVIDSON = 1
DVR_NSTAGE = 7
STAGE_NAMES = ("1d", "2d_build", "2d_eig", "2d_vec", "3d_build", "3d_eig", "3d_vec")

EXPORTS = (
    "dvr_ctx_create", "dvr_ctx_destroy", "dvr_last_error", "dvr_set_basis",
    "dvr_set_potential", "dvr_eval_potential_ab2", "dvr_run", "dvr_run_blocks", "dvr_get_sizes",
    "dvr_get_eigenvalues", "dvr_get_1d_eigenvalues", "dvr_get_2d_eigenvalues",
    "dvr_get_wavefunctions", "dvr_get_vectors", "dvr_get_stage_times",
    "dvr_diag", "dvr_syev_batch", "dvr_syev_batch_noalloc", "dvr_nfmain", "dvr_rotlev3",
    "dvr_dipole3", "dvr_xpect3", "dvr_set_profiling",
    "dvr_get_profile",
    "dvr_get_solver_stats", "dvr_ldl_solve", "dvr_keep_hamiltonian", "dvr_set_preconditioner",
    "dvr_mlp_train", "dvr_mlp_forward", "dvr_mlp_backward",
)
MLP_NPARAM = 17537
PREC_SHIFT_INVERT, PREC_DIAGONAL, PREC_LEARNED, PREC_SI_LEARNED = 0, 1, 2, 3
SOLVER_STATS = ("status", "iterations", "matvecs", "solves", "factorizations", "slices",
                "certified", "missing", "dense_fallback", "rr_us", "si_us", "expand_us",
                "factor_us", "start_us", "twostage")


class NfDims(C.Structure):
    _fields_ = [(n, C.c_int32) for n in (
        "npnt", "nalf", "ipar", "max3d", "neval", "zcut", "nvec", "jrot", "kz", "kmin",
        "zrot", "zlin")]


class RotDims(C.Structure):
    _fields_ = [(n, C.c_int32) for n in (
        "nblk", "kfirst", "jrot", "idia", "ipar", "npnt1", "npnt2", "zembed", "zquad2", "nlim",
        "nsets")] + [("neval", C.c_int32 * 2), ("nvec", C.c_int32)]


class RotBlock(C.Structure):
    _fields_ = [(n, C.c_int32) for n in ("maxleg", "idvr", "lincr", "mvib")]


class DipDims(C.Structure):
    _fields_ = [(n, C.c_int32) for n in ("nrad", "ipot", "ne1", "ne2", "nblk1", "nblk2", "npair")]


class DipBlock(C.Structure):
    _fields_ = [("nin", C.c_int32), ("vec", C.POINTER(C.c_double)),
                ("ang", C.POINTER(C.c_double))]


class DipPair(C.Structure):
    _fields_ = [("b1", C.c_int32), ("b2", C.c_int32), ("comp", C.c_int32),
                ("factor", C.c_double)]


class DipStats(C.Structure):
    _fields_ = [("transform_ms", C.c_double), ("scale_ms", C.c_double),
                ("trans_ms", C.c_double), ("ntransform", C.c_int32), ("ntrans", C.c_int32)]


class XpDims(C.Structure):
    _fields_ = [(n, C.c_int32) for n in ("npts", "npot", "ne", "nblk", "nprop")]


class XpBlock(C.Structure):
    _fields_ = [("nin", C.c_int32), ("ldv", C.c_int64), ("vec", C.POINTER(C.c_double)),
                ("ang", C.POINTER(C.c_double))]


class Dims(C.Structure):
    _fields_ = [(n, C.c_int32) for n in (
        "npnta", "npntb", "npntc", "max2d", "max3d", "neval", "zcut", "zall",
        "zr2r1", "idvr", "npnt1", "npnt2", "nvec", "solver")]


class DvrError(RuntimeError):
    def __init__(self, code: int, msg: str):
        super().__init__(f"[{code}] {msg}")
        self.code = code


_lib = None


def load(path: str = LIB_PATH) -> C.CDLL:
    """Load the HIP library (raises if it has not been built)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(path):
        raise FileNotFoundError(
            f"{path} not built: run dvr3d_amd.build.build() (HIP library required, no CPU fallback)")
    lib = C.CDLL(path)
    P = C.c_void_p
    D = C.POINTER(C.c_double)
    I = C.POINTER(C.c_int32)
    lib.dvr_ctx_create.argtypes = [C.POINTER(Dims), C.c_int32, C.POINTER(P)]
    lib.dvr_ctx_destroy.argtypes = [P]
    lib.dvr_ctx_destroy.restype = None
    lib.dvr_last_error.argtypes = [P]
    lib.dvr_last_error.restype = C.c_char_p
    lib.dvr_set_basis.argtypes = [P, D, D, D, D, D, C.c_double, C.c_double]
    lib.dvr_set_potential.argtypes = [P, D]
    lib.dvr_eval_potential_ab2.argtypes = [P, D, D, D, D]
    lib.dvr_run.argtypes = [P]
    lib.dvr_run_blocks.argtypes = [C.POINTER(P), C.c_int32]
    lib.dvr_get_sizes.argtypes = [P, I, I, I]
    lib.dvr_get_eigenvalues.argtypes = [P, D, I]
    lib.dvr_get_1d_eigenvalues.argtypes = [P, D]
    lib.dvr_get_2d_eigenvalues.argtypes = [P, D, I]
    lib.dvr_get_wavefunctions.argtypes = [P, D]
    lib.dvr_get_vectors.argtypes = [P, D]
    lib.dvr_get_stage_times.argtypes = [P, D]
    lib.dvr_diag.argtypes = [D, C.c_int32, C.c_int32, D]
    lib.dvr_syev_batch.argtypes = [C.c_int32, I, D, I, D, D]
    lib.dvr_syev_batch_noalloc.argtypes = [C.c_int32, I, D, I, D, D]
    lib.dvr_nfmain.argtypes = [C.POINTER(NfDims), D, D, D, D, C.c_double, C.c_double, D,
                               C.c_double, I, I, I, D, D]
    lib.dvr_rotlev3.argtypes = [C.POINTER(RotDims), C.POINTER(RotBlock), D, D, D, D, I, D, D,
                                D]
    lib.dvr_dipole3.argtypes = [C.POINTER(DipDims), C.POINTER(DipBlock), C.POINTER(DipBlock), D,
                                C.POINTER(DipPair), D, D, C.POINTER(DipStats)]
    lib.dvr_xpect3.argtypes = [C.POINTER(XpDims), C.POINTER(XpBlock), D, D]
    lib.dvr_set_profiling.argtypes = [P, C.c_int32]
    lib.dvr_get_profile.argtypes = [P, C.c_int32, C.c_char_p, C.c_int32, I, D, D, D]
    lib.dvr_get_solver_stats.argtypes = [P, I, C.c_int32]
    lib.dvr_ldl_solve.argtypes = [D, C.c_int32, C.c_double, D, C.c_int32, D, I]
    Fp = C.POINTER(C.c_float)
    lib.dvr_keep_hamiltonian.argtypes = [P, C.c_int32]
    lib.dvr_set_preconditioner.argtypes = [P, C.c_int32, Fp]
    lib.dvr_mlp_train.argtypes = [P, C.c_int32, C.c_int32, C.c_uint64, C.c_float, Fp, D]
    lib.dvr_mlp_forward.argtypes = [Fp, C.c_int64, Fp, Fp, C.c_int32]
    lib.dvr_mlp_backward.argtypes = [Fp, C.c_int64, Fp, Fp, Fp]
    for name in EXPORTS:
        if name not in ("dvr_ctx_destroy", "dvr_last_error"):
            getattr(lib, name).restype = C.c_int
    _lib = lib
    return lib


def _dp(a: np.ndarray | None):
    if a is None:
        return None
    return a.ctypes.data_as(C.POINTER(C.c_double))


def _ip(a: np.ndarray | None):
    if a is None:
        return None
    return a.ctypes.data_as(C.POINTER(C.c_int32))


def _f64(a) -> np.ndarray:
    return np.ascontiguousarray(a, dtype=np.float64)


class Context:
    """One device context: the jhmain stage sequence behind dvr_run."""

    def __init__(self, dims: dict, device: int = 0):
        self.lib = load()
        self.dims = dict(dims)
        d = Dims(**{k: int(v) for k, v in dims.items()})
        self._d = d
        h = C.c_void_p()
        rc = self.lib.dvr_ctx_create(C.byref(d), device, C.byref(h))
        if rc != DVR_OK:
            raise DvrError(rc, "dvr_ctx_create failed")
        self.h = h

    def close(self):
        if getattr(self, "h", None):
            self.lib.dvr_ctx_destroy(self.h)
            self.h = None

    def __del__(self):
        self.close()

    def _check(self, rc: int):
        if rc != DVR_OK:
            msg = self.lib.dvr_last_error(self.h)
            raise DvrError(rc, msg.decode() if msg else "")

    def set_basis(self, xlmatr, xkb, xkc, wsum, wterm, emax1_au, emax2_au):
        self._keep = [_f64(x) for x in (xlmatr, xkb, xkc, wsum)]
        wt = None if wterm is None else _f64(wterm)
        self._check(self.lib.dvr_set_basis(
            self.h, *[_dp(x) for x in self._keep], _dp(wt), float(emax1_au), float(emax2_au)))

    def set_potential(self, v):
        v = _f64(v)
        self._check(self.lib.dvr_set_potential(self.h, _dp(v)))

    def eval_potential_ab2(self, rc, rb, xa, par):
        arrs = [_f64(x) for x in (rc, rb, xa, par)]
        self._check(self.lib.dvr_eval_potential_ab2(self.h, *[_dp(x) for x in arrs]))

    def run(self):
        self._check(self.lib.dvr_run(self.h))

    def sizes(self):
        dd = self.dims
        iv1 = np.zeros(dd["npntc"] * dd["npntb"], dtype=np.int32)
        iv2 = np.zeros(dd["npntc"], dtype=np.int32)
        n3 = C.c_int32()
        self._check(self.lib.dvr_get_sizes(self.h, _ip(iv1), _ip(iv2), C.byref(n3)))
        # iv1 comes back as the reference's iv1(ndimc, ndimb), column-major
        return iv1.reshape(dd["npntb"], dd["npntc"]).T.copy(), iv2, int(n3.value)

    def eigenvalues(self):
        m = C.c_int32()
        self._check(self.lib.dvr_get_eigenvalues(self.h, None, C.byref(m)))
        out = np.zeros(m.value)
        self._check(self.lib.dvr_get_eigenvalues(self.h, _dp(out), C.byref(m)))
        return out

    def eigenvalues_1d(self):
        dd = self.dims
        out = np.zeros(dd["npntc"] * dd["npntb"] * dd["npnta"])
        self._check(self.lib.dvr_get_1d_eigenvalues(self.h, _dp(out)))
        return out.reshape(dd["npntc"], dd["npntb"], dd["npnta"])

    def eigenvalues_2d(self):
        t = C.c_int32()
        self._check(self.lib.dvr_get_2d_eigenvalues(self.h, None, C.byref(t)))
        out = np.zeros(t.value)
        self._check(self.lib.dvr_get_2d_eigenvalues(self.h, _dp(out), C.byref(t)))
        return out

    def wavefunctions(self, nvec: int):
        dd = self.dims
        out = np.zeros(nvec * dd["idvr"] * dd["npnt1"] * dd["npnt2"])
        self._check(self.lib.dvr_get_wavefunctions(self.h, _dp(out)))
        # record l is psi(idvr, npnt1, npnt2) in Fortran order
        return out.reshape(nvec, dd["npnt2"], dd["npnt1"], dd["idvr"])

    def vectors(self, n3: int, nvec: int):
        out = np.zeros(n3 * nvec)
        self._check(self.lib.dvr_get_vectors(self.h, _dp(out)))
        return out.reshape(nvec, n3).T

    def set_profiling(self, on: bool):
        self._check(self.lib.dvr_set_profiling(self.h, 1 if on else 0))

    def kernel_profile(self):
        """Per-kernel stats of the last profiled run: name -> dict."""
        out = {}
        i = 0
        buf = C.create_string_buffer(128)
        while True:
            n = C.c_int32()
            sec, by, fl = C.c_double(), C.c_double(), C.c_double()
            rc = self.lib.dvr_get_profile(self.h, i, buf, 128, C.byref(n), C.byref(sec),
                                          C.byref(by), C.byref(fl))
            if rc != DVR_OK:
                break
            out[buf.value.decode()] = dict(launches=n.value, seconds=sec.value,
                                           bytes=by.value, flops=fl.value)
            i += 1
        return out

    def stage_times(self):
        out = np.zeros(DVR_NSTAGE)
        self._check(self.lib.dvr_get_stage_times(self.h, _dp(out)))
        return dict(zip(STAGE_NAMES, out.tolist()))

    def keep_hamiltonian(self, on=True):
        self._check(self.lib.dvr_keep_hamiltonian(self.h, int(on)))

    def set_preconditioner(self, kind, params=None):
        p = None
        if params is not None:
            params = np.ascontiguousarray(params, dtype=np.float32)
            self._keep_mlp = params
            p = params.ctypes.data_as(C.POINTER(C.c_float))
        self._check(self.lib.dvr_set_preconditioner(self.h, int(kind), p))

    def mlp_train(self, params, steps, batch=32, seed=0, lr=2e-3):
        """Trains params (float32, MLP_NPARAM) in place; returns the loss history."""
        params = np.ascontiguousarray(params, dtype=np.float32)
        hist = np.zeros(steps)
        self._check(self.lib.dvr_mlp_train(self.h, steps, batch, seed, lr,
                                           params.ctypes.data_as(C.POINTER(C.c_float)),
                                           _dp(hist)))
        return params, hist

    def solver_stats(self):
        out = np.zeros(len(SOLVER_STATS), dtype=np.int32)
        self._check(self.lib.dvr_get_solver_stats(
            self.h, out.ctypes.data_as(C.POINTER(C.c_int32)), len(SOLVER_STATS)))
        return dict(zip(SOLVER_STATS, out.tolist()))


def run_blocks(ctxs):
    """Run several contexts (the k blocks of one J > 0 run) concurrently."""
    lib = load()
    arr = (C.c_void_p * len(ctxs))(*[c.h for c in ctxs])
    rc = lib.dvr_run_blocks(arr, len(ctxs))
    for c in ctxs:
        if rc != DVR_OK:
            msg = lib.dvr_last_error(c.h)
            if msg:
                raise DvrError(rc, msg.decode())
    if rc != DVR_OK:
        raise DvrError(rc, "dvr_run_blocks failed")


def syev_batch(mats, nvec=None, vectors=True, noalloc=False):
    """Lowest eigenpairs of several symmetric matrices (lower triangle read);
    noalloc: through dvr_syev_batch_noalloc (orders <= 128, static device
    workspace)."""
    lib = load()
    n = np.array([m.shape[0] for m in mats], dtype=np.int32)
    nv = n.copy() if nvec is None else np.array(nvec, dtype=np.int32)
    a = np.concatenate([np.asfortranarray(m, dtype=np.float64).ravel(order="F") for m in mats])
    w = np.zeros(int(nv.sum()))
    z = np.zeros(int((n.astype(np.int64) * nv).sum())) if vectors else None
    fn = lib.dvr_syev_batch_noalloc if noalloc else lib.dvr_syev_batch
    rc = fn(len(mats), _ip(n), _dp(a), _ip(nv), _dp(w), _dp(z))
    if rc != DVR_OK:
        raise DvrError(rc, "dvr_syev_batch%s failed" % ("_noalloc" if noalloc else ""))
    ws, zs, wo, zo = [], [], 0, 0
    for p in range(len(mats)):
        ws.append(w[wo:wo + nv[p]])
        wo += nv[p]
        if vectors:
            zs.append(z[zo:zo + n[p] * nv[p]].reshape(nv[p], n[p]).T)
            zo += n[p] * nv[p]
    return (ws, zs) if vectors else ws


def _fp(a):
    return a.ctypes.data_as(C.POINTER(C.c_float))


MLP_NF, MLP_WIDTH = 6, 128


def mlp_init(seed=0):
    """Initial parameters of the learned preconditioner in the device layout
    (float32: W1 [NF x 128] row-major, b1, W2 [128 x 128] row-major, b2, W3,
    b3): W1, W2 with N(0, 1/fan-in) entries, biases zero, W3 = 0 and b3 = 1,
    so that the untrained network is the classical Davidson preconditioner."""
    rng = np.random.default_rng(seed)
    w1 = rng.standard_normal((MLP_NF, MLP_WIDTH)) / np.sqrt(MLP_NF)
    w2 = rng.standard_normal((MLP_WIDTH, MLP_WIDTH)) / np.sqrt(MLP_WIDTH)
    out = np.concatenate([w1.ravel(), np.zeros(MLP_WIDTH), w2.ravel(), np.zeros(MLP_WIDTH),
                          np.zeros(MLP_WIDTH), np.ones(1)]).astype(np.float32)
    assert out.size == MLP_NPARAM
    return out


def mlp_forward(params, feats, bf16=False):
    lib = load()
    p = np.ascontiguousarray(params, dtype=np.float32)
    f = np.ascontiguousarray(feats, dtype=np.float32)
    w = np.zeros(f.shape[0], dtype=np.float32)
    rc = lib.dvr_mlp_forward(_fp(p), f.shape[0], _fp(f), _fp(w), int(bf16))
    if rc != DVR_OK:
        raise DvrError(rc, "dvr_mlp_forward failed")
    return w


def mlp_backward(params, feats, gw):
    lib = load()
    p = np.ascontiguousarray(params, dtype=np.float32)
    f = np.ascontiguousarray(feats, dtype=np.float32)
    g = np.ascontiguousarray(gw, dtype=np.float32)
    out = np.zeros(MLP_NPARAM, dtype=np.float32)
    rc = lib.dvr_mlp_backward(_fp(p), f.shape[0], _fp(f), _fp(g), _fp(out))
    if rc != DVR_OK:
        raise DvrError(rc, "dvr_mlp_backward failed")
    return out


def ldl_solve(a: np.ndarray, sigma: float, r: np.ndarray | None = None):
    """dvr_ldl_solve: inertia of a - sigma I and (a - sigma I)^-1 r."""
    lib = load()
    h = np.asfortranarray(a, dtype=np.float64)
    n = h.shape[0]
    nu = np.zeros(1, dtype=np.int32)
    if r is None:
        rc = lib.dvr_ldl_solve(_dp(h), n, float(sigma), None, 0, None,
                               nu.ctypes.data_as(C.POINTER(C.c_int32)))
        t = None
    else:
        rr = np.asfortranarray(r.reshape(n, -1), dtype=np.float64)
        t = np.zeros_like(rr, order="F")
        rc = lib.dvr_ldl_solve(_dp(h), n, float(sigma), _dp(rr), rr.shape[1], _dp(t),
                               nu.ctypes.data_as(C.POINTER(C.c_int32)))
    if rc != DVR_OK:
        raise DvrError(rc, "dvr_ldl_solve failed")
    return int(nu[0]), t


def diag(ham: np.ndarray):
    """dvr_diag on a square column-major copy of ham: returns (eig, vectors)."""
    lib = load()
    h = np.asfortranarray(ham, dtype=np.float64).copy(order="F")
    n = h.shape[0]
    e = np.zeros(n)
    rc = lib.dvr_diag(h.ctypes.data_as(C.POINTER(C.c_double)), n, n, _dp(e))
    if rc != DVR_OK:
        raise DvrError(rc, "dvr_diag failed")
    return e, h


def nfmain(dims: dict, hr, htheta, r, xalf, ur1, urr1, vgrid, emax2_au=0.0):
    """dvr_nfmain: the symmetrised-Radau driver for one k block.  vgrid is
    V[g, b, a] = V(r_a, r_b, x_g).  Returns a dict with iv2, nham3, iang,
    eval and (nvec > 0) wvfunc (nvec x npair*iang)."""
    lib = load()
    d = NfDims(**{k: int(dims.get(k, 0)) for k, _ in NfDims._fields_})
    npnt, nalf, ipar = d.npnt, d.nalf, d.ipar
    npair = npnt * (npnt + 1 - 2 * ipar) // 2
    iv2 = np.zeros(nalf, dtype=np.int32)
    n3 = np.zeros(1, dtype=np.int32)
    iang = np.zeros(1, dtype=np.int32)
    ev = np.zeros(d.neval)
    wv = np.zeros(max(1, d.nvec * npair * nalf)) if d.nvec > 0 else None
    args = [_f64(np.asfortranarray(hr).ravel(order="F")),
            _f64(np.asfortranarray(htheta).ravel(order="F")), _f64(r), _f64(xalf)]
    vg = _f64(np.ascontiguousarray(vgrid).ravel())
    rc = lib.dvr_nfmain(C.byref(d), *[_dp(x) for x in args], float(ur1), float(urr1), _dp(vg),
                        float(emax2_au), _ip(iv2), _ip(n3), _ip(iang), _dp(ev), _dp(wv))
    if rc != DVR_OK:
        raise DvrError(rc, "dvr_nfmain failed")
    n = int(n3[0])
    out = dict(iv2=iv2, nham3=n, iang=int(iang[0]), eval=ev[: min(d.neval, n)])
    if d.nvec > 0:
        out["wvfunc"] = wv[: d.nvec * npair * int(iang[0])].reshape(d.nvec, -1)
    return out


def rotlev3(dims: dict, blocks, pleg, dvrvec, evib, rm2):
    """dvr_rotlev3: ROTLEV3's Hamiltonian and solve on the device.  blocks:
    per k block dict(maxleg, idvr, lincr, mvib); pleg: per block the
    (maxleg+1, idvr) pleg*w array; dvrvec: per block mvib x (idvr*npnt1*npnt2)
    vectors (angle fastest, as IWAVE stores them); evib: per block its mvib
    levels.  dvrvec may also be one 1-D array holding every block's vectors
    one block after another.  Returns per set dict(eval, vec, d) (d: per
    block nkbas x neval) and the per-block (jstart, nang)."""
    lib = load()
    nb = len(blocks)
    ne = list(dims.get("neval", (0, 0))) + [0, 0]
    d = RotDims(**{k: int(dims.get(k, 0)) for k, _ in RotDims._fields_ if k != "neval"})
    d.neval[0], d.neval[1] = int(ne[0]), int(ne[1])
    B = (RotBlock * nb)(*[RotBlock(int(b["maxleg"]), int(b["idvr"]), int(b["lincr"]),
                                   int(b["mvib"])) for b in blocks])
    pl = _f64(np.concatenate([np.asarray(p, dtype=np.float64).ravel(order="F") for p in pleg]))
    if isinstance(dvrvec, np.ndarray) and dvrvec.ndim == 1:
        xv = _f64(dvrvec)          # already concatenated block after block
    else:
        xs = [np.asarray(x, dtype=np.float64).reshape(-1) for x in dvrvec]
        xv = _f64(np.concatenate(xs)) if sum(len(x) for x in xs) else np.zeros(1)
    ev = _f64(np.concatenate([np.asarray(e, dtype=np.float64)[: b["mvib"]]
                              for e, b in zip(evib, blocks)]))
    rm = _f64(rm2)
    nrad = d.npnt1 * d.npnt2
    jdia = max(d.idia, 1)
    jn = np.zeros(2 * nb, dtype=np.int32)
    # sizes of the outputs (the angular count of block b as JTRAN forms it)
    nang = []
    for i, b in enumerate(blocks):
        k = d.kfirst + i
        js = k + (1 if k % jdia != d.ipar else 0)
        nang.append((b["maxleg"] + b["lincr"] - js) // jdia + 1)
    mv = [int(b["mvib"]) for b in blocks]
    nbass = sum(mv)
    sets = [(0, nbass, d.neval[0])] + ([(1, nbass - mv[0], d.neval[1])] if d.nsets == 2 else [])
    eval_ = np.zeros(sum(x[2] for x in sets))
    nv = sum(x[1] * x[2] for x in sets)
    ndv = sum(sum(nang[b] * nrad * x[2] for b in range(x[0], nb)) for x in sets)
    vec = np.zeros(max(1, nv)) if d.nvec else None
    dvec = np.zeros(max(1, ndv)) if d.nvec else None
    rc = lib.dvr_rotlev3(C.byref(d), B, _dp(pl), _dp(xv), _dp(ev), _dp(rm), _ip(jn), _dp(eval_),
                         _dp(vec), _dp(dvec))
    if rc != DVR_OK:
        raise DvrError(rc, "dvr_rotlev3 failed")
    out, eo, vo, do = [], 0, 0, 0
    for first, n, nev in sets:
        r = dict(eval=eval_[eo:eo + nev].copy())
        eo += nev
        if d.nvec:
            r["vec"] = vec[vo:vo + n * nev].reshape(nev, n).T.copy()
            vo += n * nev
            r["d"] = []
            for b in range(first, nb):
                nk = nang[b] * nrad
                r["d"].append(dvec[do:do + nk * nev].reshape(nev, nk).T.copy())
                do += nk * nev
        out.append(r)
    return out, jn.reshape(nb, 2)


def _arr_ptr(x):
    """(pointer, element count, keep-alive) of a float64 host array or of a
    device tensor (anything with data_ptr()/numel(), read in place)."""
    if hasattr(x, "data_ptr"):
        return C.cast(C.c_void_p(x.data_ptr()), C.POINTER(C.c_double)), int(x.numel()), x
    a = _f64(np.asarray(x, dtype=np.float64))
    return _dp(a), int(a.size), a


def dipole3(nrad: int, ipot: int, ket, bra, dipol, pairs, tz=None, tx=None, stats=None):
    """dvr_dipole3: DIPOLE3's transition moments on the device.  ket, bra:
    per k block (vec, ang) with vec the block's ne vectors of nin*nrad values
    (an (ne, nin*nrad) array, index a + nin*mn) and ang the (ipot, nin)
    values of the nin input functions at the ipot integration points (a
    numpy array, or a device tensor already holding them column-major), or
    None for a block no pair reads; dipol: (2, nrad*ipot) z and x dipole values with the weights;
    pairs: (b1, b2, comp, factor) tuples.  vec, ang and dipol may be numpy
    arrays or device tensors (read in place).  Returns (tz, tx), (ne1, ne2)
    each, added to the given start values; stats, if a dict, receives the
    call's device phase times."""
    lib = load()
    keep = []

    def blocks(side):
        out, ne = [], 0
        for b in side:
            if b is None:
                out.append(DipBlock(0, None, None))
                continue
            vec, ang = b
            if not hasattr(ang, "data_ptr"):
                ang = np.asarray(ang, dtype=np.float64).ravel(order="F")
            pa, na, ka = _arr_ptr(ang)
            pv, nv, kv = _arr_ptr(vec)
            keep.extend([ka, kv])
            nin = na // ipot
            ne = nv // (nin * nrad)
            out.append(DipBlock(nin, pv, pa))
        return (DipBlock * len(out))(*out), ne

    K, ne1 = blocks(ket)
    B, ne2 = blocks(bra)
    P = (DipPair * max(1, len(pairs)))(*[DipPair(int(b1), int(b2), int(c), float(f))
                                          for b1, b2, c, f in pairs])
    d = DipDims(nrad, ipot, ne1, ne2, len(ket), len(bra), len(pairs))
    if not hasattr(dipol, "data_ptr"):
        dipol = np.asarray(dipol, dtype=np.float64).reshape(-1)
    pd, _, kd = _arr_ptr(dipol)
    z = np.zeros((ne1, ne2), order="F") if tz is None else np.array(tz, dtype=np.float64, order="F")
    x = np.zeros((ne1, ne2), order="F") if tx is None else np.array(tx, dtype=np.float64, order="F")
    st = DipStats()
    rc = lib.dvr_dipole3(C.byref(d), K, B, pd, P, _dp(z.reshape(-1, order="F")),
                         _dp(x.reshape(-1, order="F")), C.byref(st))
    if rc != DVR_OK:
        raise DvrError(rc, "dvr_dipole3 failed")
    if isinstance(stats, dict):
        stats.update({k: getattr(st, k) for k, _ in DipStats._fields_})
    return z, x


def xpect3(npts: int, npot: int, ne: int, blocks, prop, out=None):
    """dvr_xpect3: XPECT3's expectation values on the device.  blocks: per k
    block (nin, ldv, vec, ang) -- nin = 0: vec holds each level's values at
    the npts grid points (level i at row i of an (ne, ldv) array), ang None;
    nin > 0: vec the (ne, ldv = nin*nrad) FBR coefficients and ang the
    (npot, nin) sqrt(w) P_l values -- or None for an empty block; prop: the
    (npts, nprop) property values.  vec, ang and prop may be numpy arrays or
    device tensors.  Returns (ne, nprop), added to the given start values."""
    lib = load()
    keep, bl = [], []
    for b in blocks:
        if b is None:
            bl.append(XpBlock(0, 0, None, None))
            continue
        nin, ldv, vec, ang = b
        pv, _, kv = _arr_ptr(vec)
        keep.append(kv)
        pa = None
        if ang is not None:
            if not hasattr(ang, "data_ptr"):
                ang = np.asarray(ang, dtype=np.float64).ravel(order="F")
            pa, _, ka = _arr_ptr(ang)
            keep.append(ka)
        bl.append(XpBlock(int(nin), int(ldv), pv, pa))
    B = (XpBlock * len(bl))(*bl)
    if not hasattr(prop, "data_ptr"):
        prop = np.asarray(prop, dtype=np.float64).ravel(order="F")
    pp, nprop_tot, kp = _arr_ptr(prop)
    nprop = nprop_tot // npts
    d = XpDims(int(npts), int(npot), int(ne), len(bl), int(nprop))
    o = np.zeros((ne, nprop), order="F") if out is None else np.array(out, dtype=np.float64,
                                                                      order="F")
    rc = lib.dvr_xpect3(C.byref(d), B, pp, _dp(o.reshape(-1, order="F")))
    if rc != DVR_OK:
        raise DvrError(rc, "dvr_xpect3 failed")
    return o
