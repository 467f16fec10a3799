"""ctypes binding of the Fortran host (libdvr3d_host.so, host/dvrh_setup.f90).

The Fortran host reads the DVR3DRJZ .job deck and builds the basis of one k
block exactly as the drop-in executable does; this module hands the arrays
to Python callers (bench, tests) without any Python restatement of the
algorithms.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
HOST_LIB = os.path.join(HERE, "libdvr3d_host.so")
AUTOCM = 2.19474624e05

_JOB_INTS = ("ncoord npnt1 npnt2 jrot neval nalf max2d max3d idia kmin ipar max3d2 "
             "zpham zprad zpvec zrot zladd zembed zmors1 zmors2 zpmin zvec zquad2 zdiag "
             "zlmat zcut zall zlin zp1d zp2d zr2r1 ztheta ztran ztwod zperp zx zs0 zs1 zpfun "
             "idiag1 idiag2 iout1 iout2 iwave ilev has_ezero").split()


class Job(C.Structure):
    _fields_ = ([(n, C.c_int32) for n in _JOB_INTS]
                + [("fixcos", C.c_double), ("xmass", C.c_double * 3), ("xmassr", C.c_double * 3),
                   ("emax1", C.c_double), ("emax2", C.c_double), ("re1", C.c_double),
                   ("diss1", C.c_double), ("we1", C.c_double), ("re2", C.c_double),
                   ("diss2", C.c_double), ("we2", C.c_double), ("ezero", C.c_double),
                   ("title", C.c_char * 72)])


class Sizes(C.Structure):
    _fields_ = [(n, C.c_int32) for n in (
        "npnta npntb npntc idvr nalf2 maxleg max2d max3d neval "
        "jrot kmin idia ipar zrot zembed zvec kd ku").split()]


class Consts(C.Structure):
    _fields_ = ([(n, C.c_double) for n in (
        "g1 g2 ur1 ur2 urr1 urr2 beta1 beta2 a1 a2 csx1 csa1 tsx1 tsa1 "
        "csx2 csa2 tsx2 tsa2 cswalf tswalf").split()]
                + [("iu1", C.c_int32), ("iu2", C.c_int32)])


_lib = None


def load() -> C.CDLL:
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(HOST_LIB):
        raise FileNotFoundError(f"{HOST_LIB} not built: run dvr3d_amd.build.build()")
    lib = C.CDLL(HOST_LIB)
    D = C.POINTER(C.c_double)
    lib.dvrh_read_job.argtypes = [C.c_char_p, C.POINTER(Job)]
    lib.dvrh_read_job.restype = C.c_int32
    lib.dvrh_sizes.argtypes = [C.POINTER(Job), C.POINTER(Sizes)]
    lib.dvrh_sizes.restype = C.c_int32
    lib.dvrh_setup_basis.argtypes = [C.POINTER(Job), C.POINTER(Sizes), C.c_int32,
                                     D, D, D, D, D, D, D, D, C.POINTER(Consts)]
    lib.dvrh_setup_basis.restype = C.c_int32
    lib.dvrh_wsum.argtypes = [C.POINTER(Job), C.POINTER(Sizes), C.POINTER(Consts), C.c_int32,
                              D, D, D, D]
    lib.dvrh_wsum.restype = C.c_int32
    lib.dvrh_potential_grid.argtypes = [C.POINTER(Job), C.POINTER(Sizes), C.POINTER(Consts),
                                        D, D, D, D]
    lib.dvrh_potential_grid.restype = C.c_int32
    _lib = lib
    return lib


def _dp(a):
    return a.ctypes.data_as(C.POINTER(C.c_double))


def read_job(path: str) -> Job:
    lib = load()
    job = Job()
    rc = lib.dvrh_read_job(os.fsencode(path), C.byref(job))
    if rc != 0:
        raise ValueError(f"dvrh_read_job({path}) failed with code {rc}")
    return job


def sizes(job: Job) -> Sizes:
    lib = load()
    sz = Sizes()
    rc = lib.dvrh_sizes(C.byref(job), C.byref(sz))
    if rc != 0:
        raise ValueError(f"job outside the theta-first jhmain path (code {rc})")
    return sz


class Basis:
    """Basis arrays of one k block, as computed by the Fortran host."""

    def __init__(self, job: Job, kz: int | None = None):
        lib = load()
        self.job = job
        self.sz = sz = sizes(job)
        self.kz = sz.kmin if kz is None else kz
        n1, n2, nd, ml = job.npnt1, job.npnt2, sz.idvr, sz.maxleg
        self.r1 = np.zeros(n1)
        self.r2 = np.zeros(n2)
        self.xk1 = np.zeros((n1, n1), order="F")
        self.xk2 = np.zeros((n2, n2), order="F")
        self.xalf = np.zeros(nd)
        self.walf = np.zeros(nd)
        self.pleg = np.zeros((ml + 1, nd), order="F")
        self.xlmatr = np.zeros((nd, nd), order="F")
        self.consts = Consts()
        rc = lib.dvrh_setup_basis(C.byref(job), C.byref(sz), self.kz, _dp(self.r1), _dp(self.r2),
                                  _dp(self.xk1), _dp(self.xk2), _dp(self.xalf), _dp(self.walf),
                                  _dp(self.pleg), _dp(self.xlmatr), C.byref(self.consts))
        if rc != 0:
            raise RuntimeError(f"points & weights in error (code {rc})")
        nbc = sz.npntb * sz.npntc
        self.wsum = np.zeros(nbc)
        self.wterm = np.zeros(nbc)
        lib.dvrh_wsum(C.byref(job), C.byref(sz), C.byref(self.consts), self.kz,
                      _dp(self.r1), _dp(self.r2), _dp(self.wsum), _dp(self.wterm))

    @property
    def rc(self):
        return self.r1 if self.job.zr2r1 else self.r2

    @property
    def rb(self):
        return self.r2 if self.job.zr2r1 else self.r1

    @property
    def xkb(self):
        return self.xk2 if self.job.zr2r1 else self.xk1

    @property
    def xkc(self):
        return self.xk1 if self.job.zr2r1 else self.xk2

    def potential_grid(self) -> np.ndarray:
        """The linked example POTV (potv_example.f90; surface chosen by the
        parameter file named in DVR_PES, default pes.par) on this block's
        grid, V[c, b, a]."""
        sz = self.sz
        v = np.zeros((sz.npntc, sz.npntb, sz.npnta))
        rc = load().dvrh_potential_grid(C.byref(self.job), C.byref(sz), C.byref(self.consts),
                                        _dp(self.r1), _dp(self.r2), _dp(self.xalf), _dp(v))
        if rc != 0:
            raise RuntimeError(f"dvrh_potential_grid failed (code {rc})")
        return v

    def dims(self, nvec: int = 0, solver: int = 0) -> dict:
        sz, job = self.sz, self.job
        return dict(npnta=sz.npnta, npntb=sz.npntb, npntc=sz.npntc, max2d=sz.max2d,
                    max3d=sz.max3d, neval=sz.neval, zcut=job.zcut, zall=job.zall,
                    zr2r1=job.zr2r1, idvr=sz.idvr, npnt1=job.npnt1, npnt2=job.npnt2,
                    nvec=nvec, solver=solver)
