"""FP64 CPU restatement of DVR3DRJZ's symmetrised-Radau driver nfmain
(IDIA = -2), for parity tests only.

Test infrastructure: imported by tests/, never by the product path.

Reference: source/dvr3drjz_segmented.f90 (cited as D:<line>)
  insize clamps for IDIA = -2      D:246-262, 333-345
  nfmain (2D loop, 3D, final)      D:3006-3122
  blc2d1 (2D Hamiltonian per angle) D:3125-3220
  choosr / cut2dr (selection)      D:3223-3377
  testiv (bisector embedding)      D:3380-3452
  bloc3d / blc2d2 / vecmul (3D)    D:3455-3595
  transr (wavefunctions)           D:3599-3646
The set-up (masses, radial Gauss-Laguerre DVR, kinetic matrix, angular
Gauss-(associated)-Legendre DVR, L matrix) is the one restated and pinned in
oracle/dvr3d_oracle.py; IDIA = -2 uses the full angular range like IDIA = -1
(allpts, D:1747-1772) and the Radau g1, g2 (D:1076-1080).  The Schwenke /
NBO correction switches of the segmented file (xp0, xp1, xp2) are off by
default and contribute nothing; they are not restated.
"""
from __future__ import annotations

import math
from dataclasses import dataclass

import numpy as np

from oracle import dvr3d_oracle as O

AUTOCM = O.AUTOCM


@dataclass
class NfSizes:
    npnt: int
    nalf: int
    nalf2: int
    maxleg: int
    npair: int       # dimension of every 2D problem (max2d of D:333-340)
    max3d: int
    neval: int
    jrot: int
    kmin: int
    ipar: int
    zcut: bool
    zrot: bool
    zlin: bool
    ztran: bool


def nf_sizes(job: dict) -> NfSizes:
    """insize for IDIA = -2: npnt1 = npnt2, idvr = nalf, ztheta forced F
    (D:246-252); 2D dimension npnt (npnt + 1 - 2 ipar) / 2 and max3d clamp
    (D:333-345)."""
    n = dict(job["nml"])
    jrot, kmin, ipar = job["jrot"], job["kmin"], job["ipar"]
    if job["idia"] != -2:
        raise ValueError("nfmain is the IDIA = -2 driver")
    if jrot == 0:                                   # D:241-245
        kmin = 0
        n["zrot"] = False
    if n["zperp"] or n["ztwod"]:
        raise NotImplementedError("zperp / ztwod take other drivers")
    npnt = job["npnt2"]                             # npnt1 = npnt2 (D:248)
    nalf = job["nalf"]
    if n["zrot"] and jrot + kmin > 1:
        raise NotImplementedError("both parities of one k block (kmin = 2) not restated")
    npair = npnt * (npnt + 1 - 2 * ipar) // 2       # D:338
    max3d = job["max3d"]
    max3d = npair * nalf if n["zall"] else min(max3d, npair * nalf)
    neval = job["neval"] if job["neval"] > 0 else 10
    neval = min(max3d, neval)
    return NfSizes(npnt=npnt, nalf=nalf, nalf2=(nalf + 1) // 2, maxleg=nalf - 1, npair=npair,
                   max3d=max3d, neval=neval, jrot=jrot, kmin=kmin, ipar=ipar,
                   zcut=bool(n["zcut"]), zrot=bool(n["zrot"]), zlin=bool(n["zlin"]),
                   ztran=bool(n["ztran"]))


@dataclass
class NfBasis:
    sizes: NfSizes
    consts: O.Consts
    r: np.ndarray         # radial points (both bonds)
    hr: np.ndarray        # radial DVR kinetic matrix (k1k2)
    xalf: np.ndarray      # angular points, full range
    walf: np.ndarray
    htheta: np.ndarray    # L matrix (lmatrx)
    kz: int


def nf_basis(job: dict, kz: int = 0) -> NfBasis:
    """ccmain set-up for IDIA = -2 (D:750-982): r1 radial basis, xk1,
    full-range angular points and the L matrix of j = kz, kz + 1, ..."""
    s = nf_sizes(job)
    c = O.setcon(job)
    rad = O.lagpt(s.npnt, job["re1"], c.beta1, c.a1, c.iu1)
    fke = c.beta1 * c.beta1 / (8.0 * c.ur1)
    hbl = O.keints(fke, s.npnt - 1, c.iu1)
    hr = O.k1k2(hbl, rad["bass"], np.sqrt(rad["wt"]))
    tsw = O.tswalf_exact(kz)
    xh, wh, csw = O.jacobi(s.nalf, s.nalf2, float(kz), float(kz), tsw)
    if abs((csw - tsw) / tsw) > 1.0e-8:
        raise RuntimeError("angular points & weights in error (D:939-943)")
    xalf, walf = O.allpts(xh, wh, s.nalf, s.nalf2, -2)
    pleg = O.asleg(s.maxleg, xalf, kz, kz)
    xl = O.lmatrx(pleg, walf, kz, 0, kz, -2, s.maxleg)
    return NfBasis(sizes=s, consts=c, r=rad["r"], hr=hr, xalf=xalf, walf=walf, htheta=xl, kz=kz)


def pairs(npnt: int, ipar: int):
    """(alpha, beta) of every symmetrised radial function, in the order of
    blc2d1's loops (beta outer, alpha = 1 .. beta - ipar; D:3180-3183)."""
    return [(a, b) for b in range(npnt) for a in range(b + 1 - ipar)]


def rot_terms(jrot: int, kz: int, kmin: int, zrot: bool):
    """term, term2, term3 of nfmain (D:3056-3064)."""
    term = term2 = term3 = 0.0
    if jrot != 0:
        term = (jrot * jrot + jrot - kz * kz) / 8.0
        term2 = (jrot * jrot + jrot - 3 * kz * kz) / 4.0
        if abs(kz) == 1:
            term3 = (jrot * jrot + jrot) / 16.0
            if kmin >= 1 and zrot:
                term3 = -term3
    return term, term2, term3


def blc2d1(basis: NfBasis, xcos: float, vfun) -> np.ndarray:
    """Symmetrised radial Hamiltonian at one angle, D:3125-3220: kinetic
    terms of both bonds, the potential at (r_a, r_b) and (r_b, r_a) (sign q
    for the odd combination), 1/sqrt(2) for the r1 = r2 functions, and for
    J > 0 the rotational diagonal (which the r1 = r2 function of each beta
    receives twice, as the loop structure of D:3159-3175 adds it)."""
    s = basis.sizes
    r, hr = basis.r, basis.hr
    P = pairs(s.npnt, s.ipar)
    n = len(P)
    H = np.zeros((n, n))
    q = -1.0 if s.ipar == 1 else 1.0
    f2 = math.sqrt(0.5)
    if s.jrot != 0:
        term, term2, term3 = rot_terms(s.jrot, basis.kz, s.kmin, s.zrot)
        fact = term + term2 / (1.0 - xcos)
        if basis.kz == 1:
            fact = fact + term3 * (1.0 + xcos) / (1.0 - xcos)
        urr1 = basis.consts.urr1
        for p, (a, b) in enumerate(P):
            wsum = (0.5 / (r[a] * r[a]) + 0.5 / (r[b] * r[b])) * fact / urr1
            H[p, p] += wsum
            if s.ipar == 0 and a == b:
                H[p, p] += wsum
    for pp, (ap, bp) in enumerate(P):
        for p, (a, b) in enumerate(P):
            h = H[pp, p]
            if b == bp:
                h += hr[ap, a]
            if b == ap:
                h += q * hr[bp, a]
            if a == bp:
                h += q * hr[ap, b]
            if a == ap:
                h += hr[bp, b]
            if a == ap and b == bp:
                h += vfun(r[a], r[b], xcos)
            if a == bp and b == ap:
                h += q * vfun(r[b], r[a], xcos)
            if ap == bp:
                h *= f2
            if a == b:
                h *= f2
            H[pp, p] = h
    return H


def choosr(eigs: list, max3d: int):
    """Global selection of choosr (D:3223-3320): emax = the max3d-th lowest
    2D level over all angles (or the highest, when there are fewer), then
    every level <= emax of every angle, in ascending order.  Returns the
    number chosen per angle."""
    allv = np.sort(np.concatenate(eigs))
    emax = allv[len(allv) - 1] if len(allv) <= max3d else allv[max3d - 1]
    return [int(np.count_nonzero(e <= emax)) for e in eigs], float(emax)


def cut2dr(eigs: list, emax2_cm: float, max3d: int):
    """Energy cut of cut2dr (D:3323-3377): levels strictly below EMAX2,
    stop when more than max3d are kept."""
    emaxau = emax2_cm / AUTOCM
    nv = [int(np.count_nonzero(e < emaxau)) for e in eigs]
    if sum(nv) > max3d:
        raise RuntimeError("number of 2d eigenvalues greater than max3d (D:3359-3362)")
    return nv


def testiv(nv: list, s: NfSizes):
    """Bisector-embedding pruning for J > 0 (D:3380-3452): functions at
    angles beyond the first empty one past the middle are removed, and with
    zlin the theta = 0 ones as well (without zlin their presence stops the
    reference run).  Returns the pruned counts and iang (angles kept)."""
    nv = list(nv)
    if s.jrot > 0:
        nalf = s.nalf
        iend = nalf
        for i in range(nalf // 2, nalf + 1):        # Fortran iend = nalf/2 .. nalf
            if nv[i - 1] == 0:
                iend = i
                break
        for i in range(iend + 1, nalf + 1):
            nv[i - 1] = 0
        if nv[nalf - 1] > 0:
            if not s.zlin:
                raise RuntimeError("bisector embedding: wavefunction has amplitude for theta = 0")
            nv[nalf - 1] = 0
    iang = sum(1 for x in nv if x > 0)
    return nv, iang


def angular_weights(basis: NfBasis) -> np.ndarray:
    """Diagonal of blc2d2 per symmetrised radial function (D:3509-3567):
    (1/2)(1/r_a^2 + 1/r_b^2)/mu; the L matrix element multiplies it."""
    r = basis.r
    mu = basis.consts.ur1
    return np.array([0.5 / (r[a] * r[a] * mu) + 0.5 / (r[b] * r[b] * mu)
                     for a, b in pairs(basis.sizes.npnt, basis.sizes.ipar)])


def run_nf(job: dict, pes=None, vgrid: np.ndarray | None = None, kz: int = 0,
           keep_vectors: bool = False):
    """The whole nfmain for one k block.  vgrid (optional): the potential on
    the grid, V[g, b, a] = V(r_a, r_b, x_g); otherwise pes(r1, r2, x)."""
    basis = nf_basis(job, kz)
    s = basis.sizes
    r = basis.r
    if vgrid is None:
        vfun_of = lambda g: (lambda r1, r2, x: float(pes(r1, r2, x)))
    else:
        idx = {float(v): i for i, v in enumerate(r)}
        vfun_of = lambda g: (lambda r1, r2, x: float(vgrid[g, idx[float(r2)], idx[float(r1)]]))
    eigs, vecs = [], []
    for g in range(s.nalf):
        H = blc2d1(basis, basis.xalf[g], vfun_of(g))
        e, v = np.linalg.eigh(H)
        eigs.append(e)
        vecs.append(v)
    if s.zcut:
        nv = cut2dr(eigs, job["emax2"], s.max3d)
    else:
        nv, _ = choosr(eigs, s.max3d)
    nv, iang = testiv(nv, s)
    n3 = sum(nv)
    w = angular_weights(basis)
    col0 = np.concatenate([[0], np.cumsum(nv)[:-1]]).astype(int)
    C = np.zeros((s.npair, n3))
    e2 = np.zeros(n3)
    blk = np.zeros(n3, dtype=int)
    for g in range(s.nalf):
        k = nv[g]
        C[:, col0[g]:col0[g] + k] = np.sqrt(w)[:, None] * vecs[g][:, :k]
        e2[col0[g]:col0[g] + k] = eigs[g][:k]
        blk[col0[g]:col0[g] + k] = g
    H3 = (C.T @ C) * basis.htheta[np.ix_(blk, blk)] + np.diag(e2)
    if keep_vectors:
        ev, Z = np.linalg.eigh(H3)
    else:
        ev = np.linalg.eigvalsh(H3)
        Z = None
    out = dict(basis=basis, iv2=np.array(nv), iang=iang, nham3=n3, eval=ev[:s.neval],
               eigs2d=eigs, H3=H3)
    if keep_vectors:
        out["Z"] = Z
        out["wvfunc"] = transr(nv, vecs, Z, s)
    return out


def transr(nv, vecs, Z, s: NfSizes) -> np.ndarray:
    """Wavefunction records of transr (D:3599-3646): for every level, index
    running over the symmetrised radial functions (outer) and the angles
    with functions (inner), amplitude sum_l V_g(irs, l) Z(col(g, l), level)."""
    gs = [g for g in range(s.nalf) if nv[g] > 0]
    col0 = np.concatenate([[0], np.cumsum(nv)[:-1]]).astype(int)
    out = np.zeros((s.neval, s.npair, len(gs)))
    for j, g in enumerate(gs):
        k = nv[g]
        out[:, :, j] = (vecs[g][:, :k] @ Z[col0[g]:col0[g] + k, :s.neval]).T
    return out.reshape(s.neval, -1)
