"""CPU restatement of XPECT3, the expectation values of user properties over
the vib-rot wavefunctions of one run (source/xpect3.f90, cited as X:<line>).

TEST INFRASTRUCTURE ONLY: imported by tests/ and the golden scripts as the
checker of the device path, never by the product.

What is restated:
  read_ket    INSIZE X:76-295, GENIND X:320-360, X3MAIN X:380-421  headers,
              block table, grids and energies of a KVEC file (ROTLEV3) or of
              a J = 0 IWAVE file (DVR3DRJZ, IDIA = -1 jhmain or IDIA = -2
              nfmain records)
  quadrature  LAGPT X:566-603, JACOBI/ROOT/RECUR X:659-758  npot-point
              Gauss-Legendre by the sign scan, mirrored and reflected
  properties  X3MAIN X:438-505 (J = 0, on the DVR grid), LAGPT X:605-648
              (J > 0, on the quadrature grid), PROP averaged over the
              homonuclear exchange as the reference does
  values      DSRD X:763-845, REST X:848-925  FBR coefficients -> the
              quadrature points, times sqrt(w)
  expect      TRANS X:998-1082  props(i, n) = sum over the grid of
              P_n * |psi_i|^2, per k block
  report      OUTPT2 X:1086-1280  the table and the ITRA records

Parity pinned by tests/test_xp_oracle.py against the reference XPECT3 built
from its sources (oracle/Makefile target xpect) on the reference's own water
J = 0 test (water/jobs/dvr0.job + xpect0.job with its prop_water.f90), on a
J = 0 jhmain run and on ROTLEV3 output of J = 1 and J = 2, and against the
norm of the wavefunctions (property 1 = 1) and the exchange symmetry of the
homonuclear cases.
"""
from __future__ import annotations

from dataclasses import dataclass, field

import numpy as np

from oracle.dipole3_oracle import _d, _i, asleg, read_records

AUTOCM = 2.19474624e05      # X:1108


@dataclass
class Ket:
    idia: int
    ipar: int
    lmax: int
    npnt1: int
    npnt2: int
    jrot: int
    kmin: int
    neval: int
    zembed: bool
    zncor: bool
    xmass: np.ndarray
    g1: float
    g2: float
    jk: int
    lmin: np.ndarray
    nbass: np.ndarray
    r1: np.ndarray
    r2: np.ndarray
    energies: np.ndarray
    recs: list = field(default_factory=list)
    first: int = 0        # record of the first vector (KVEC block or IWAVE level)
    xalf: np.ndarray = None
    iv: np.ndarray = None  # IDIA = -2: 2D functions kept at each angle


def read_ket(path) -> Ket:
    """The file as INSIZE, GENIND and X3MAIN read it (X:180-182, 335, 380-421)."""
    r = read_records(path)
    h = _i(r[0])
    idia, ipar, lmax, n1, n2, jrot, kmin, neval = (int(x) for x in h[:8])
    lg = _i(r[1][:12])
    xm = _d(r[1][12:36])
    g1, g2 = _d(r[1][36:52])
    zncor = bool(_i(r[1][52:56])[0])
    if jrot == 0:
        kmin = 1                                     # X:255
    jk = 1 if zncor else jrot + kmin                 # X:256-261
    t = _i(r[3])
    if len(t) != 1 + 3 * jk:
        raise ValueError("block record does not match J + KMIN")
    lmin = t[1:1 + jk].copy()
    nbass = t[1 + 2 * jk:].copy()
    k = Ket(idia, ipar, lmax, n1, n2, jrot, kmin, neval, bool(lg[0]) if idia > -2 else True,
            zncor, xm.copy(), float(g1), float(g2), jk, lmin, nbass, _d(r[4]).copy(),
            np.zeros(0), np.zeros(0), r)
    if idia > -2:
        k.r2 = _d(r[5]).copy()
        pos = 6
    else:
        k.r2 = k.r1.copy()                           # one radial grid for both bonds
        pos = 5
    if jk <= 1:
        if idia > -2 and idia >= 0:
            # X:405-411 read the energy-count record as the angular grid here
            raise ValueError("J = 0 files of IDIA > 0 are not readable by XPECT3")
        k.xalf = _d(r[pos]).copy()
        if idia == -2:
            k.iv = _i(r[pos + 4]).copy()
            pos += 5
        else:
            pos += 3
        k.energies = _d(r[pos + 1]).copy()
        k.first = pos + 2
    else:
        k.energies = _d(r[pos + 1]).copy()
        k.first = pos + 2
    return k


@dataclass
class XpInput:
    """Namelist PRT (X:18-19, 68-71) and the cards of INSIZE (X:141-184)."""
    lpot: int
    npropin: int
    nprt: int
    nv1: int = 0
    iprop: tuple = ()
    zfit: bool = False


def sizes(k: Ket, inp: XpInput):
    """lpot, npot and the levels taken (X:144-170, 281-291)."""
    nv1 = inp.nv1
    if nv1 > k.neval:
        raise ValueError("NV1 > NEVAL")
    if nv1 == 0:
        nv1 = k.neval
    lpot = inp.lpot
    if lpot == 0:
        lpot = k.lmax + k.jrot + (k.lmax + k.jrot) % 2
    if k.jrot == 0:
        lpot = k.lmax
    npot = lpot
    if k.jrot > 0:
        ltop = lpot
        if k.idia == 2 and ltop % 2 != k.ipar:
            ltop += 1
        npot = ((ltop + 2) // 2) * 2
    return nv1, lpot, npot


# ---------------------------------------------------------------- quadrature
_STEP = float(np.float32(0.0001))   # X:702: a single-precision literal


def _recur(x, nn, c):
    p1, p = 1.0, x
    dp1, dp = 0.0, 1.0
    for j in range(2, nn + 1):
        q = x * p - c[j] * p1
        dq = x * dp + p - c[j] * dp1
        p1, p, dp1, dp = p, q, dp, dq
    return p, dp, p1


def quadrature(npot: int):
    """LAGPT and JACOBI (X:566-603, 659-715): the positive zeros found by a
    sign scan from 1 down in steps of 1e-4 (single precision), Newton steps
    with the weight from the derivative before the last one, the negative
    mirror added and every point reflected; points ascending from -x1 to x1.
    Returns (xd, wtd, csa, tsa)."""
    nn = npot
    nn2 = nn // 2
    c = np.zeros(nn + 1)
    for i in range(2, nn + 1):
        xi = float(i)
        c[i] = (4.0 * (xi - 1.0) * (xi - 1.0) * (xi - 1.0) * (xi - 1.0)
                / ((2 * xi - 1.0) * (2 * xi - 2.0) * (2 * xi - 2.0) * (2 * xi - 3.0)))
    cc = 2.0
    tsa = cc
    for i in range(2, nn + 1):
        cc = cc * c[i]
    xt = []
    pm1, xxx = 1.0, 1.0
    while True:
        p, _, _ = _recur(xxx, nn, c)
        if pm1 * p < 0.0:
            pm1 = -pm1
            xt.append(xxx)
        if len(xt) == nn2:
            break
        xxx = xxx - _STEP
        if not xxx > -0.0002:
            raise ValueError("incorrect number of zeros found in JACOBI")
    x = np.zeros(nn2)
    a = np.zeros(nn2)
    csa = 0.0
    for i in range(nn2):
        t, it = xt[i], 0
        while True:
            it += 1
            p, dp, pn1 = _recur(t, nn, c)
            d = p / dp
            t = t - d
            if not (abs(d) > 1e-12 and it < 10):
                break
        x[i] = t
        a[i] = cc / (dp * pn1)
        csa = csa + a[i] + a[i]
    if 2 * nn2 != nn:
        csa -= a[nn2 - 1]
    xd = np.concatenate([x, -x[::-1]])
    wd = np.concatenate([a, a[::-1]])
    return -xd, wd, csa, tsa


# ---------------------------------------------------------------- properties
def property_grid(k: Ket, prop, npot: int, xd=None):
    """The property values on the grid the wavefunctions are given on,
    (K, npropin): J = 0 on the DVR grid (X:438-505), J > 0 on the quadrature
    grid (X:622-648).  prop(r1, r2, x) returns (..., npropin) for arrays."""
    if k.jk <= 1:
        x = k.xalf
        if k.idia == -2:
            ia = np.flatnonzero(k.iv[:len(x)] != 0)
            rows = []
            for i2 in range(k.npnt2):
                for i1 in range(i2 + 1 - k.ipar):
                    a = prop(np.full(len(ia), k.r1[i1]), np.full(len(ia), k.r1[i2]), x[ia])
                    b = prop(np.full(len(ia), k.r1[i2]), np.full(len(ia), k.r1[i1]), x[ia])
                    rows.append(0.5 * (a + b))
            return np.concatenate(rows)
        R2, R1, X = np.meshgrid(k.r2, k.r1, x[:k.lmax], indexing="ij")
        v = prop(R1, R2, X)
        return v.reshape(-1, v.shape[-1])
    R2, R1, X = np.meshgrid(k.r2, k.r1, xd[:npot], indexing="ij")
    v = prop(R1, R2, X)
    if k.idia == 2:
        v = 0.5 * (v + prop(R1, R2, -X))
    return v.reshape(-1, v.shape[-1])


def values(k: Ket, blk: int, nv1: int, xd, wtd, npot: int):
    """DSRD + REST (X:763-925): block blk's nv1 vectors at the quadrature
    points times sqrt(w), (nv1, nrad*npot)."""
    nb = int(k.nbass[blk])
    rec = _d(k.recs[k.first + blk]).reshape(-1, nb)[:nv1]
    jdia = max(1, k.idia)
    nang2 = k.lmax
    kz = blk + 1 - k.kmin
    pl = asleg(nang2 * jdia - 1, xd[:npot], kz)            # (npot, nang2*jdia)
    if k.idia == 2:
        jj = int(k.lmin[kz]) - kz
        pl = pl[:, jj::2][:, :nang2]
    else:
        pl = pl[:, :nang2]
    pl = pl * np.sqrt(wtd[:npot])[:, None]
    nrad = nb // nang2
    return (rec.reshape(nv1, nrad, nang2) @ pl.T).reshape(nv1, nrad * npot)


def run(k: Ket, inp: XpInput, prop):
    """TRANS over every block: props (nv1, npropin) = sum over the grid of
    P_n |psi_i|^2 (X:1048-1078).  Returns dict(props, eh, npot, xd, wtd)."""
    nv1, lpot, npot = sizes(k, inp)
    if k.jk <= 1:
        P = property_grid(k, prop, npot)
        K = P.shape[0]
        psi = np.stack([_d(k.recs[k.first + l])[:K] for l in range(nv1)])
        return dict(props=(psi * psi) @ P, eh=k.energies, npot=npot, xd=None, wtd=None, nv1=nv1)
    xd, wtd, csa, tsa = quadrature(npot)
    if abs((csa - tsa) / tsa) > 1e-8:
        raise ValueError("gauss-legendre weights in error")
    P = property_grid(k, prop, npot, xd)
    props = np.zeros((nv1, P.shape[1]))
    for b in range(k.jk):
        if k.nbass[b] == 0:
            continue
        d = values(k, b, nv1, xd, wtd, npot)
        props += (d * d) @ P
    return dict(props=props, eh=k.energies, npot=npot, xd=xd, wtd=wtd, nv1=nv1)


def symmetry_label(ipar: int, kmin: int) -> str:
    """OUTPT2 X:1121-1129."""
    return {(0, 1): "a1", (1, 0): "b2", (1, 1): "a2", (0, 0): "b1"}[(ipar, kmin)]


def device_inputs(k: Ket, inp: XpInput, prop):
    """The inputs dvr_xpect3 takes: (npts, npot, nv1, blocks, P) with blocks
    per k block (nin, ldv, vec (nv1, ldv), ang (npot, nin) or None) and P the
    (npts, npropin) property values on the grid."""
    nv1, lpot, npot = sizes(k, inp)
    if k.jk <= 1:
        P = property_grid(k, prop, npot)
        vec = np.stack([_d(k.recs[k.first + l]) for l in range(nv1)])
        return P.shape[0], 0, nv1, [(0, vec.shape[1], vec, None)], P
    xd, wtd, _, _ = quadrature(npot)
    P = property_grid(k, prop, npot, xd)
    jdia = max(1, k.idia)
    blocks = []
    for b in range(k.jk):
        nb = int(k.nbass[b])
        if nb == 0:
            blocks.append(None)
            continue
        kz = b + 1 - k.kmin
        pl = asleg(k.lmax * jdia - 1, xd[:npot], kz)
        if k.idia == 2:
            jj = int(k.lmin[kz]) - kz
            pl = pl[:, jj::2][:, :k.lmax]
        else:
            pl = pl[:, :k.lmax]
        ang = pl * np.sqrt(wtd[:npot])[:, None]
        vec = _d(k.recs[k.first + b]).reshape(-1, nb)[:nv1]
        blocks.append((k.lmax, nb, vec, ang))
    return P.shape[0], npot, nv1, blocks, P
