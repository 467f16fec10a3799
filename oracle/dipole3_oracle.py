"""CPU restatement of DIPOLE3, the line-strength step that reads the
wavefunction files of DVR3DRJZ (J = 0) and ROTLEV3 (source/dipole3_segmented.f90,
cited as P:<line>).

TEST INFRASTRUCTURE ONLY: imported by tests/ and the golden scripts as the
checker of the device path, never by the product.

What is restated (coordinates with IDIA > -2; the symmetrised Radau branches
of IDIA = -2 belong to ROTLEV3B input and are not part of this build):
  read_wave   INSIZE P:198-620, GENIND P:656-704    headers, block sizes
  gauss_legendre  LAGPT P:1141-1204, JACOBI/ROOT/RECUR P:1311-1409
                                  npot-point Gauss-Legendre, Stroud-Secrest
  dipole_grid LAGPT P:1206-1230   DIPD at (r1, r2, x) times the weights
  asleg       ASLEG P:1234-1308   normalised associated Legendre functions
  jtran       JTRAN P:1564-1645   DVR vectors of a J = 0 IWAVE -> FBR
  transform   DSRD  P:1412-1562   FBR coefficients -> the integration points
  threej      THREEJ P:1778-1822  3-j symbol from the binomial table (SETFAC)
  run         DMAIN P:708-1115, TRANS P:1648-1776  the k -> k' block sums
  spect       SPECT P:1825-1953, OUTPT P:1956-2002  frequencies, moments,
                                  line strengths, Einstein A, the SPECTRA file

Parity pinned by tests/test_dip_oracle.py against the reference DIPOLE3 built
from its sources (oracle/Makefile target dip) on the reference's own H3+ chain
(H3+Jacobi jobs dvrJ0, dvrJ1, rot, dip) and on ABC decks with a synthetic
dipole surface, and against what the mathematics fixes (Gauss-Legendre
points, the orthogonality of the 3-j symbols, the overlap limit of a constant
dipole).
"""
from __future__ import annotations

import struct
from dataclasses import dataclass, field

import numpy as np

AUTOCM = 2.19474624e05      # P:1851
AUTODE = 2.5417662          # P:1852
DETOSEC = 3.136186e-07      # P:1853


def read_records(path):
    """Fortran sequential unformatted records (4-byte markers)."""
    recs = []
    with open(path, "rb") as f:
        data = f.read()
    pos = 0
    while pos < len(data):
        (n,) = struct.unpack_from("<i", data, pos)
        pos += 4
        recs.append(data[pos:pos + n])
        pos += n + 4
    return recs


def _i(rec):
    return np.frombuffer(rec, "<i4")


def _d(rec):
    return np.frombuffer(rec, "<f8")


@dataclass
class Wave:
    """One wavefunction file: the KVEC file of ROTLEV3 (DSTORE R:1290-1323)
    or the IWAVE file of a J = 0 DVR3DRJZ run (D:1178-1184, 973-977, 3733-3841)."""
    idia: int
    ipar: int
    lmax: int          # angular DVR points (idvr)
    npnt1: int
    npnt2: int
    jrot: int
    kmin: int
    neval: int
    zembed: bool
    zmors1: bool
    zmors2: bool
    xmass: np.ndarray
    g1: float
    g2: float
    znco: bool
    morse: np.ndarray
    jk: int
    nbass: np.ndarray  # basis functions per k block (GENIND)
    r1: np.ndarray
    r2: np.ndarray
    energies: np.ndarray
    recs: list = field(default_factory=list)
    first: int = 0     # record index of the first vector record (KVEC) or of xalf (IWAVE)

    @property
    def iwave(self):
        return self.jk <= 1


def read_wave(path) -> Wave:
    """Header records as INSIZE reads them (P:267-275, 442-443), the block
    table of GENIND (P:680) and the energies (P:817-828)."""
    r = read_records(path)
    h = _i(r[0])
    lg = _i(r[1][:12])
    xm = _d(r[1][12:36])
    g1, g2 = _d(r[1][36:52])
    znco = bool(_i(r[1][52:56])[0])
    idia, ipar, lmax, n1, n2, jrot, kmin, neval = (int(x) for x in h[:8])
    # INSIZE's KMIN: 0 for a non-Coriolis file paired with a coupled one
    # (P:485-503), 1 for J = 0 (P:516); jk = J + KMIN blocks (P:517), which
    # is also the length of GENIND's block record
    t = _i(r[3])
    if (len(t) - 1) % 3 != 0:
        raise ValueError("malformed block record")
    jk = (len(t) - 1) // 3
    kmin = jk - jrot
    if kmin not in (0, 1):
        raise ValueError("block record does not match J")
    nb = t[1 + 2 * jk:].copy()
    w = Wave(idia, ipar, lmax, n1, n2, jrot, kmin, neval, bool(lg[0]), bool(lg[1]), bool(lg[2]),
             xm.copy(), float(g1), float(g2), znco, _d(r[2]).copy(), jk, nb,
             _d(r[4]).copy(), _d(r[5]).copy(), np.zeros(0), r)
    if w.iwave:
        # records 7-9 angular data, 10 meval, 11 energies, 12... vectors
        w.energies = _d(r[10]).copy()
        w.first = 6
    else:
        w.energies = _d(r[7]).copy()
        w.first = 8
    return w


@dataclass
class DipInput:
    """Namelist PRT (P:145-146, 192-194) and the cards of INSIZE/SPECT
    (P:564-567, 1884)."""
    npot: int
    nv1: int = 0
    nv2: int = 0
    ibase1: int = 0
    ibase2: int = 0
    ezero: float = 0.0
    nblock: int = 1000


# ---------------------------------------------------------------- quadrature
def _recur(x, nn, alf, bta, b, c):
    """P:1389-1409."""
    p1, p = 1.0, x + (alf - bta) / (alf + bta + 2.0)
    dp1, dp = 0.0, 1.0
    for j in range(2, nn + 1):
        q = (x - b[j]) * p - c[j] * p1
        dq = (x - b[j]) * dp + p - c[j] * dp1
        p1, p, dp1, dp = p, q, dp, dq
    return p, dp, p1


def _root(x, nn, alf, bta, b, c, eps=1e-12):
    """Newton steps P:1370-1386; the derivative returned is the one of the
    last step's starting point."""
    it = 0
    while True:
        it += 1
        p, dp, pn1 = _recur(x, nn, alf, bta, b, c)
        d = p / dp
        x = x - d
        if not (abs(d) > eps and it < 10):
            return x, dp, pn1


# single-precision literals of JACOBI (0.768, 0.452, 0.156, 0.012 without a
# kind suffix), promoted to double where they meet double operands
_F32_0768 = float(np.float32(0.768))
_F32_0452 = float(np.float32(0.452))
_F32_0156 = float(np.float32(0.156))
_F32_0012 = float(np.float32(0.012))


def gauss_legendre(npot: int):
    """npot-point Gauss-Legendre points and weights as LAGPT and JACOBI
    form them (Stroud-Secrest start values, alf = bta = 0): the npot/2
    positive points largest first, then their mirror images
    (P:1166-1204, 1311-1367).  Returns (xd, wtd, csa, tsa)."""
    alf = bta = 0.0
    nn = npot
    b = np.zeros(nn + 1)
    c = np.zeros(nn + 1)
    for i in range(2, nn + 1):
        xi = float(i)
        b[i] = (alf + bta) * (bta - alf) / ((alf + bta + 2 * xi) * (alf + bta + 2 * xi - 2))
        c[i] = (4.0 * (xi - 1.0) * (alf + xi - 1.0) * (bta + xi - 1.0) * (alf + bta + xi - 1.0)
                / ((alf + bta + 2 * xi - 1) * (alf + bta + 2 * xi - 2)
                   * (alf + bta + 2 * xi - 2) * (alf + bta + 2 * xi - 3)))
    fn = float(nn)
    nn2 = nn // 2
    beta = 1.0
    cc = beta * 2.0 ** (alf + bta + 1.0)
    tsa = cc / 2.0
    for i in range(2, nn + 1):
        cc = cc * c[i]
    x = np.zeros(nn)
    a = np.zeros(nn)
    csa = 0.0
    xt = 0.0
    for i in range(1, nn2 + 1):
        an, bn = alf / fn, bta / fn
        if i == 1:
            r1 = (1.0 + alf) * (2.78 / (4.0 + fn * fn) + _F32_0768 * an / fn)
            r2 = 1.0 + 1.48 * an + 0.96 * bn + _F32_0452 * an * an + 0.83 * an * bn
            xt = 1.0 - r1 / r2
        elif i == 2:
            r1 = (4.1 + alf) / ((1.0 + alf) * (1.0 + _F32_0156 * alf))
            r2 = 1.0 + 0.06 * (fn - 8.0) * (1.0 + 0.12 * alf) / fn
            r3 = 1.0 + _F32_0012 * bta * (1.0 + abs(alf) / 4.0) / fn
            xt = xt - r1 * r2 * r3 * (1.0 - xt)
        elif i == 3:
            r1 = (1.67 + 0.28 * alf) / (1.0 + 0.37 * alf)
            r2 = 1.0 + 0.22 * (fn - 8.0) / fn
            r3 = 1.0 + 8.0 * bta / ((6.28 + bta) * fn * fn)
            xt = xt - r1 * r2 * r3 * (x[0] - xt)
        else:
            xt = 3.0 * x[i - 2] - 3.0 * x[i - 3] + x[i - 4]
        xt, dpn, pn1 = _root(float(xt), nn, alf, bta, b, c)
        x[i - 1] = xt
        a[i - 1] = cc / (dpn * pn1)
        csa += a[i - 1]
    for i in range(1, nn2 + 1):                      # P:1194-1204
        j = npot + 1 - i
        x[j - 1] = -x[i - 1]
        a[j - 1] = a[i - 1]
    return x, a, csa, tsa


def ipot_of(idia: int, npot: int) -> int:
    """Angular integration points used (P:609-618)."""
    return npot if idia < 2 else npot // 2


def dipole_grid(dipd, w: Wave, xd, wtd, nu: int, ipot: int) -> np.ndarray:
    """LAGPT P:1206-1230: d0[j + ipot (i1 + npnt1 i2)] = wtd_j DIPD(r1_i1,
    r2_i2, x_j, nu), doubled for IDIA = 2.  dipd(r1, r2, x, nu) is the
    surface (vectorised over numpy arrays)."""
    jdia = max(1, w.idia)
    R2, R1, X = np.meshgrid(w.r2, w.r1, xd[:ipot], indexing="ij")
    v = np.asarray(dipd(R1, R2, X, nu), dtype=float)
    if jdia == 2:
        v = 2.0 * v
    return (wtd[:ipot][None, None, :] * v).reshape(-1)


def asleg(lmax: int, x, m: int) -> np.ndarray:
    """P:1234-1308: (len(x), lmax+1) array, column ll = normalised
    P_{m+ll}^m(x) including the (-1)^m phase of the recurrence."""
    x = np.asarray(x, dtype=float)
    out = np.zeros((len(x), lmax + 2))
    for i, xi in enumerate(x):
        pmm, fact = 1.0, 1.0
        somx2 = np.sqrt((1.0 - xi) * (1.0 + xi))
        for _ in range(m):
            pmm = -pmm * fact * somx2
            fact += 2.0
        out[i, 0] = pmm
        pmmp1 = xi * float(m + m + 1) * pmm
        out[i, 1] = pmmp1
        ll = 1
        for l in range(2 + m, lmax + m + 1):
            pll = (xi * float(l + l - 1) * pmmp1 - float(l + m - 1) * pmm) / float(l - m)
            pmm, pmmp1 = pmmp1, pll
            ll += 1
            out[i, ll] = pll
    pnorm = np.zeros(lmax + 1)
    for jj, j in enumerate(range(m, lmax + m + 1)):
        fact = 1.0
        for i in range(j - m + 1, j + m + 1):
            fact *= float(i)
        pnorm[jj] = np.sqrt(float(j + j + 1) / (fact + fact))
    return out[:, :lmax + 1] * pnorm[None, :]


# ------------------------------------------------------------ wavefunctions
def _jstart(kz, jdia, ipar, maxleg):
    """First l and count of the FBR functions of a block (JTRAN P:1578-1600)."""
    jstart, jj0 = kz, -jdia
    if jdia == 2 and jstart % 2 != ipar:
        jj0 += 1
        jstart += 1
    nang = (maxleg + kz - jstart) // jdia + 1
    return jstart, jj0, nang


def jtran(w: Wave, ne: int, ibase: int, jay_ipar: int, ipar1: int):
    """The first ne vectors (from ibase) of a J = 0 IWAVE file in the FBR
    basis: coef[ie, j + nang mn] = sum_a psi_ie[a, mn] pleg[jj_j, a]
    (DSRD P:1481-1507, JTRAN P:1564-1645).  Returns (coef, kz, nang)."""
    r = w.recs
    kz, maxleg, nidvr = (int(v) for v in _i(r[7])[:3])
    pleg = _d(r[8]).reshape(nidvr, maxleg + 1).T            # (0:maxleg, idvr)
    jdia = max(w.idia, 1)
    par = ipar1 if w.zembed else jay_ipar
    _, jj0, nang = _jstart(kz, jdia, par, maxleg)
    jj = jj0 + jdia * np.arange(1, nang + 1)
    P = pleg[jj, :]                                         # (nang, idvr)
    nrad = w.npnt1 * w.npnt2
    coef = np.empty((ne, nang * nrad))
    for l in range(ne):
        X = _d(r[11 + ibase - 1 + l]).reshape(nrad, nidvr).T
        coef[l] = (P @ X).T.reshape(-1)
    return coef, kz, nang


def fbr_block(w: Wave, blk: int, ne: int, ibase: int):
    """Block blk (0-based) of a KVEC file: coef[ie, j] for the ne levels
    from ibase (DSRD P:1478-1479)."""
    nb = int(w.nbass[blk])
    rec = _d(w.recs[w.first + blk])
    return rec.reshape(-1, nb)[ibase - 1:ibase - 1 + ne].copy()


def transform(w: Wave, blk: int, ne: int, ibase: int, jay_ipar: int, ipar1: int,
              xd, ipot: int):
    """DSRD: the block's wavefunctions at the angular integration points,
    d[ie, p + ipot mn] = sum_l coef[ie, l + nang mn] P_l(x_p) (P:1511-1548)."""
    nrad = w.npnt1 * w.npnt2
    jdia = max(1, w.idia)
    kk = blk + 1 - w.kmin                          # the block's k
    if w.iwave:
        coef, kz, nang = jtran(w, ne, ibase, jay_ipar, ipar1)
        kk = kz
    else:
        coef = fbr_block(w, blk, ne, ibase)
        nang = int(w.nbass[blk]) // nrad
    pl = asleg(nang * jdia - 1, xd[:ipot], kk)
    if w.idia == 2:
        jfirst = (kk + ipar1) % 2 if w.zembed else (kk + jay_ipar) % 2
        pl = pl[:, jfirst::2][:, :nang]
    C = coef.reshape(ne, nrad, nang)
    return (C @ pl.T).reshape(ne, nrad * ipot)


# --------------------------------------------------------------- 3-j symbol
def binomials(nbin: int) -> np.ndarray:
    """SETFAC P:1118-1138: binom[i, j] = C(i, j) (0-based)."""
    b = np.zeros((nbin, nbin))
    b[0, 0] = b[1, 0] = b[1, 1] = 1.0
    for i in range(2, nbin):
        b[i, 0] = b[i, i] = 1.0
        for j in range(1, i):
            b[i, j] = b[i - 1, j - 1] + b[i - 1, j]
    return b


def threej(j1, j2, j3, m1, m2, m3, binom) -> float:
    """P:1778-1822 (Racah's sum over the binomial table, 1-based indices
    of the reference mapped to 0-based rows/columns)."""
    B = lambda i, j: binom[i - 1, j - 1]   # noqa: E731
    if m1 + m2 + m3 != 0:
        return 0.0
    i1, i2, i3 = -j1 + j2 + j3 + 1, j1 - j2 + j3 + 1, j1 + j2 - j3 + 1
    k1, k2, k3 = j1 + m1 + 1, j2 + m2 + 1, j3 + m3 + 1
    l1, l2, l3 = j1 - m1 + 1, j2 - m2 + 1, j3 - m3 + 1
    if min(i1, i2, i3, k1, k2, k3, l1, l2, l3) <= 0:
        return 0.0
    n1, n2, n3 = -j1 - m2 + j3, m1 - j2 + j3, j1 - j2 + m3
    imin = max(-n1, -n2, 0) + 1
    imax = min(l1, k2, i3)
    if imin > imax:
        return 0.0
    s, t = 1.0, 0.0
    for i in range(imin, imax + 1):
        s = -s
        t += s * B(i1, n1 + i) * B(i2, n2 + i) * B(i3, i)
    t *= np.sqrt(B(j2 + j2 + 1, i3) * B(j1 + j1 + 1, i2)
                 / (B(j1 + j2 + j3 + 2, i3) * float(j3 + j3 + 1)
                    * B(j1 + j1 + 1, l1) * B(j2 + j2 + 1, l2) * B(j3 + j3 + 1, l3)))
    if (n3 + imin) % 2 != 0:
        t = -t
    return t


# ------------------------------------------------------------- block sums
@dataclass
class Pair:
    """One k -> k' block of DMAIN: ket block b1, bra block b2 (0-based),
    component (0: z, nu = 0; 1: x, nu = +-1), nu and the factor
    threej * xfac * (-1)^kk1 that TRANS applies (P:1699-1700)."""
    b1: int
    b2: int
    comp: int
    nu: int
    factor: float


def check(ket: Wave, bra: Wave):
    """The consistency checks of INSIZE that decide what is run
    (P:282-352, 485-520, 590-595)."""
    if ket.idia != bra.idia:
        raise ValueError("diatomic mismatch")
    if ket.idia <= -2:
        raise ValueError("IDIA = -2 (symmetrised Radau) files are not supported")
    if ket.jrot == 0 and bra.jrot == 0:
        raise ValueError("j = 0 -> 0 not allowed")
    if ket.jrot == 0 or bra.jrot == 0:
        zembed = bra.zembed if ket.jrot == 0 else ket.zembed
    else:
        if ket.zembed != bra.zembed:
            raise ValueError("embedding mismatch")
        zembed = ket.zembed
    if ket.idia == 2 and ((ket.ipar != bra.ipar) if zembed else (ket.ipar == bra.ipar)):
        raise ValueError("parity mismatch, spin forbidden")
    if ket.znco and bra.znco:
        raise ValueError("non-Coriolis pairs are not supported")
    if (ket.znco and abs(ket.jrot) > 1) or (bra.znco and abs(bra.jrot) > 1):
        raise ValueError("attempt to mix no coriolis and coupled calcs")   # P:485-503
    itot = ket.jrot + bra.jrot + 1 + ket.kmin + bra.kmin
    if itot % 2 != 0 or abs(ket.jrot - bra.jrot) > 1:
        raise ValueError("selection rules violated")
    return zembed


def schedule(ket: Wave, bra: Wave, zembed: bool, nblock: int = 1000):
    """The k -> k' blocks of DMAIN in the reference's order (P:877-913 for
    nu = 0, P:933-1049 for nu = +-1), each with its factor; blocks with an
    empty side are listed with factor 0 (the reference skips them but counts
    them, P:884-888)."""
    jk1, jk2 = ket.jk, bra.jk
    j1, j2 = jk1 - ket.kmin, jk2 - bra.kmin
    binom = binomials(ket.jrot + bra.jrot + 3)
    out = []

    def add(k1, k2, nu, xfac, comp):
        kk1, kk2 = k1 - ket.kmin, k2 - bra.kmin
        empty = ket.nbass[k1 - 1] == 0 or bra.nbass[k2 - 1] == 0
        f = threej(j1, 1, j2, kk1, nu, -kk2, binom) * xfac
        if kk1 % 2 != 0:
            f = -f
        out.append(Pair(k1 - 1, k2 - 1, comp, nu, 0.0 if empty else f))

    ks1 = 2 if ket.kmin > bra.kmin else 1
    for k1 in range(ks1, jk1 + 1):                    # nu = 0
        k2 = k1 - ket.kmin + bra.kmin
        if k2 > jk2:
            continue
        add(k1, k2, 0, 1.0, 0)
    flip = (not zembed) and ket.idia < 0
    for k1 in range(1, jk1 + 1):                      # nu = +-1
        kk1 = k1 - ket.kmin
        if ket.nbass[k1 - 1] == 0:
            continue
        if jk2 > 1:
            kk2 = kk1 + 1
            k2 = kk2 + bra.kmin
            if k2 <= jk2:
                xfac = -1.0 if kk1 == 0 else -1.0 / np.sqrt(2.0)
                add(k1, k2, 1, -xfac if flip else xfac, 1)
        kk2 = kk1 - 1
        k2 = kk2 + bra.kmin
        if k2 >= 1:
            xfac = 1.0 if kk2 == 0 else 1.0 / np.sqrt(2.0)
            add(k1, k2, -1, -xfac if flip else xfac, 1)
    return out[:nblock]


def run(ket: Wave, bra: Wave, inp: DipInput, dipd):
    """DMAIN + TRANS: the z and x transition moments tz, tx (ne1 x ne2) of
    every ket/bra pair, plus what SPECT prints."""
    zembed = check(ket, bra)
    ib1 = inp.ibase1 if 0 < inp.ibase1 <= ket.neval else 1      # P:566-567
    ib2 = inp.ibase2 if 0 < inp.ibase2 <= bra.neval else 1
    ne1 = min(inp.nv1, ket.neval - ib1 + 1) if inp.nv1 > 0 else ket.neval - ib1 + 1
    ne2 = min(inp.nv2, bra.neval - ib2 + 1) if inp.nv2 > 0 else bra.neval - ib2 + 1
    npot = inp.npot + (inp.npot % 2)
    ipot = ipot_of(ket.idia, npot)
    xd, wtd, csa, tsa = gauss_legendre(npot)
    if abs((csa - tsa) / tsa) > 1e-8:
        raise ValueError("gauss-legendre weights in error")
    dip = [dipole_grid(dipd, bra, xd, wtd, nu, ipot) for nu in (0, 1)]
    pairs = schedule(ket, bra, zembed, inp.nblock)
    tz = np.zeros((ne1, ne2))
    tx = np.zeros((ne1, ne2))
    cache = {}

    def block(side, w, b, ne, ib, jay):
        key = (side, b)
        if key not in cache:
            cache[key] = transform(w, b, ne, ib, jay, ket.ipar, xd, ipot)
        return cache[key]

    for p in pairs:
        if p.factor == 0.0:
            continue
        d1 = block(1, ket, p.b1, ne1, ib1, ket.ipar)
        d2 = block(2, bra, p.b2, ne2, ib2, bra.ipar)
        t = tz if p.comp == 0 else tx
        t += p.factor * (d1 * dip[p.comp][None, :]) @ d2.T
    e1 = ket.energies[ib1 - 1:ib1 - 1 + ne1]
    e2 = bra.energies[ib2 - 1:ib2 - 1 + ne2]
    return dict(tz=tz, tx=tx, e1=e1, e2=e2, pairs=pairs, zembed=zembed, xd=xd, wtd=wtd,
                ipot=ipot, ibase=(ib1, ib2), j=(ket.jk - ket.kmin, bra.jk - bra.kmin))


def spect(res, ezero: float = 0.0, zbisc: bool = False):
    """SPECT P:1898-1946: per (ie1, ie2) the ket and bra energies and the
    frequency (cm^-1), z and x moments and |dipole| (debye), S(f-i)
    (debye^2), Einstein A (s^-1), and sint (a.u.) for OUTPT."""
    j1, j2 = res["j"]
    xf = np.sqrt(float((2 * j1 + 1) * (2 * j2 + 1)))
    xe1 = res["e1"] * AUTOCM - ezero
    xe2 = res["e2"] * AUTOCM - ezero
    dd = xe2[None, :] - xe1[:, None]
    tz, tx = res["tz"], res["tx"]
    sint = (tz + tx) ** 2 * xf * xf
    tzd = tz * AUTODE * xf
    txd = tx * AUTODE * xf
    if not zbisc and res["zembed"]:
        txd = -txd
    t = np.abs(tzd + txd)
    sxd = t * t
    dd3 = np.abs(dd ** 3)
    a = np.where(dd > 0, sxd * dd3 * DETOSEC / float(2 * j2 + 1),
                 sxd * dd3 * DETOSEC / float(2 * j1 + 1))
    return dict(xe1=xe1, xe2=xe2, freq=dd, tzd=tzd, txd=txd, dipole=t, s=sxd, a=a, sint=sint)


def device_blocks(w: Wave, ne: int, ibase: int, jay_ipar: int, ipar1: int, xd, ipot: int):
    """The inputs dvr_dipole3 takes for every k block of a file: (vec, ang)
    with vec (ne, nin*nrad) the block's vectors as stored (FBR coefficients
    of a KVEC record, or the DVR amplitudes of J = 0 IWAVE records) and ang
    (ipot, nin) the input functions at the integration points (the selected
    associated Legendre values, times JTRAN's P_sel for DVR input)."""
    nrad = w.npnt1 * w.npnt2
    jdia = max(1, w.idia)
    out = []
    for blk in range(w.jk):
        kk = blk + 1 - w.kmin
        if w.iwave:
            r = w.recs
            kz, maxleg, nidvr = (int(v) for v in _i(r[7])[:3])
            pleg = _d(r[8]).reshape(nidvr, maxleg + 1).T
            par = ipar1 if w.zembed else jay_ipar
            _, jj0, nang = _jstart(kz, jdia, par, maxleg)
            Psel = pleg[jj0 + jdia * np.arange(1, nang + 1), :]       # (nang, idvr)
            vec = np.stack([_d(r[11 + ibase - 1 + l]) for l in range(ne)])
            kk = kz
        else:
            if int(w.nbass[blk]) == 0:
                out.append(None)
                continue
            vec = fbr_block(w, blk, ne, ibase)
            nang = int(w.nbass[blk]) // nrad
            Psel = None
        pl = asleg(nang * jdia - 1, xd[:ipot], kk)
        if w.idia == 2:
            jfirst = (kk + ipar1) % 2 if w.zembed else (kk + jay_ipar) % 2
            pl = pl[:, jfirst::2][:, :nang]
        ang = pl if Psel is None else pl @ Psel
        out.append((vec, ang))
    return out
