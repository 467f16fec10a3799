"""SFIT's potential fit to vibrational term values (SFIT_DVR3D/H2S/dvr3d_sfit.f90,
cited as S:<line>) in one process on the GPU.

The reference's SFIT runs the DVR3DRJZ, ROTLEV3B and XPECT3 executables
through system calls at every iteration and reads their files back (S:741-
880); the energy derivatives with respect to the potential parameters are
Hellmann-Feynman expectation values <psi_n | dV/dp_j | psi_n> computed by
XPECT3 with a property routine that returns dV/dp_j (S:825-910).  Here one
iteration is one dvr_run (potential, 1D/2D/3D stages, levels and
wavefunctions, all in HBM) and one dvr_xpect3 call (every level against
every parameter derivative on the DVR grid as one GEMM), and the normal
equations of the weighted least-squares step are those of S:1060-1130:

    eps_n  = obs_n - (E_n - E_0)                       (term values, cm^-1)
    J_nj   = d(E_n - E_0)/dp_j = <dV/dp_j>_n - <dV/dp_j>_0      (S:902)
    A = J^T W J,  b = J^T W eps,  dp = lstsq(A, b)  (dgelss, rcond 1e-12)

with W the normalised weights (S:593-606).  The loop stops when the weighted
standard deviation or its relative change falls below the reference's
thresholds (stadev_best = 1e-4, stab_best = 1e-12, S:23) or after itmax
iterations.  J = 0 levels of one parity (the vibrational fit); the surface is
any callable with parameter derivatives, e.g. AB2Surface.
"""
from __future__ import annotations

from dataclasses import dataclass, field

import numpy as np

from dvr3d_amd import _lib, host

AUTOCM = 2.19474624e05
STADEV_BEST = 1e-4        # S:23
STAB_BEST = 1e-12         # S:23


class AB2Surface:
    """V = De[(1-e^{-a(r1-re)})^2 + (1-e^{-a(r2-re)})^2] + kb (x - cte)^2
           + c (r1-re)(r2-re),   params = (De, a, re, kb, cte, c), hartree/bohr.
    grad returns dV/dp for every parameter (the properties of the fit)."""
    names = ("De", "a", "re", "kb", "cte", "c")

    @staticmethod
    def value(p, r1, r2, x):
        De, a, re, kb, cte, c = p
        y1 = 1.0 - np.exp(-a * (r1 - re))
        y2 = 1.0 - np.exp(-a * (r2 - re))
        return De * (y1 * y1 + y2 * y2) + kb * (x - cte) ** 2 + c * (r1 - re) * (r2 - re)

    @staticmethod
    def grad(p, r1, r2, x):
        De, a, re, kb, cte, c = p
        e1, e2 = np.exp(-a * (r1 - re)), np.exp(-a * (r2 - re))
        y1, y2 = 1.0 - e1, 1.0 - e2
        g = [y1 * y1 + y2 * y2,
             2.0 * De * (y1 * e1 * (r1 - re) + y2 * e2 * (r2 - re)),
             -2.0 * De * a * (y1 * e1 + y2 * e2) - c * ((r1 - re) + (r2 - re)),
             (x - cte) ** 2,
             -2.0 * kb * (x - cte),
             (r1 - re) * (r2 - re)]
        return np.stack(g, axis=-1)


@dataclass
class FitStep:
    """One iteration: parameters used, calculated term values, obs - calc,
    the weighted standard deviation and rms (cm^-1)."""
    params: np.ndarray
    calc: np.ndarray
    eps: np.ndarray
    stadev: float
    rms: float
    jacobian: np.ndarray = field(repr=False, default=None)


class VibFit:
    """In-process SFIT for J = 0 term values of one DVR3DRJZ deck (theta-
    first jhmain path).  obs: {level index n >= 1 (0 = ground): term value
    in cm^-1}; weights: per observed level (default 1)."""

    def __init__(self, job_path: str, surface, params, ivar, obs: dict, weights=None):
        self.job = host.read_job(job_path)
        self.basis = host.Basis(self.job)
        b = self.basis
        self.surface = surface
        self.params = np.array(params, dtype=float)
        self.ivar = np.array(ivar, dtype=bool)
        self.levels = np.array(sorted(obs))
        if self.levels.min() < 1:
            raise ValueError("observed levels are term values above level 0")
        self.obs = np.array([obs[n] for n in self.levels], dtype=float)
        w = np.ones(len(self.levels)) if weights is None else np.asarray(weights, dtype=float)
        self.w = w / w.sum()                                  # S:593-606
        self.nvec = int(self.levels.max()) + 1
        sz = b.sz
        if sz.neval < self.nvec:
            raise ValueError("the deck's NEVAL is below the highest observed level")
        # grids: V[c, b, a] for dvr_set_potential, and the wavefunction
        # records psi(idvr, npnt1, npnt2) for the derivatives
        xa = b.xalf[:sz.npnta]
        rc, rb = b.rc, b.rb
        self._C, self._B, self._A = np.meshgrid(rc, rb, xa, indexing="ij")
        r1c, r2c = (self._C, self._B) if self.job.zr2r1 else (self._B, self._C)
        self._v_grid = (r1c, r2c, self._A)
        R2, R1, X = np.meshgrid(b.r2, b.r1, b.xalf[:sz.idvr], indexing="ij")
        self._psi_grid = (R1, R2, X)
        self.ctx = _lib.Context(b.dims(nvec=self.nvec, solver=_lib.DVR_SOLVER_DENSE))
        self.ctx.set_basis(b.xlmatr[:sz.npnta, :sz.npnta], b.xkb, b.xkc, b.wsum, None,
                           self.job.emax1 / host.AUTOCM, self.job.emax2 / host.AUTOCM)

    def levels_and_derivatives(self, p):
        """dvr_run at parameters p, then <dV/dp_j> of levels 0 .. nvec-1 by
        dvr_xpect3 on the DVR grid.  Returns (E (hartree), D (nvec x npar,
        hartree per parameter unit))."""
        r1, r2, x = self._v_grid
        self.ctx.set_potential(np.ascontiguousarray(self.surface.value(p, r1, r2, x)))
        self.ctx.run()
        ev = self.ctx.eigenvalues()[:self.nvec].copy()
        psi = self.ctx.wavefunctions(self.nvec).reshape(self.nvec, -1)
        G = self.surface.grad(p, *self._psi_grid).reshape(-1, len(p))
        D = _lib.xpect3(G.shape[0], 0, self.nvec, [(0, psi.shape[1], psi, None)], G)
        return ev, D

    def step(self, p):
        ev, D = self.levels_and_derivatives(p)
        calc = (ev[self.levels] - ev[0]) * AUTOCM
        eps = self.obs - calc
        J = (D[self.levels] - D[0]) * AUTOCM                  # S:902
        ssq = float(np.sum(eps * eps * self.w))
        nused = int((self.w > 0).sum())
        npar = int(self.ivar.sum())
        stadev = float(np.sqrt(ssq * nused / max(1, nused - npar)))
        rms = float(np.sqrt(np.mean(eps * eps)))
        return FitStep(p.copy(), calc, eps, stadev, rms, J)

    def fit(self, itmax: int = 10):
        """Gauss-Newton iterations of the weighted normal equations (S:1060-
        1130); returns the list of FitStep (the last one at the final
        parameters, not followed by an update)."""
        p = self.params.copy()
        hist = []
        stadev_old = 1e10
        for it in range(itmax + 1):
            s = self.step(p)
            hist.append(s)
            stability = abs(s.stadev - stadev_old) / max(s.stadev, 1e-300)
            if it == itmax or s.stadev < STADEV_BEST or stability < STAB_BEST:
                break
            stadev_old = s.stadev
            Jv = s.jacobian[:, self.ivar]
            A = Jv.T @ (self.w[:, None] * Jv)
            b = Jv.T @ (self.w * s.eps)
            dx, *_ = np.linalg.lstsq(A, b, rcond=1e-12)
            p = p.copy()
            p[self.ivar] += dx
        self.params = p
        return hist
