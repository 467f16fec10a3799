"""CPU restatement of ROTLEV3, the rotational step that reads the IWAVE file
of a DVR3DRJZ zrot run (source/rotlev3.f90, cited as R:<line>).

TEST INFRASTRUCTURE ONLY: imported by tests/ and the golden scripts as the
checker of the device path, never by the product.

What is restated (in-core path, zdcore=T; the iterative F02FJF path solves
the same matrix):
  select     SELECT  R:301-478   basis selection by energy ordering
  jtran      JTRAN   R:839-904   DVR vectors -> associated-Legendre FBR
  hamiltonian SOLRT  R:671-836   diagonal = vibrational levels, Coriolis
                                 blocks between consecutive k (dger loop)
             LOADH   R:907-972   dense matrix of one parity set
  solve      DGROT   R:1046-1197 lowest NEVAL eigenpairs (DSYEVX 'I')
  dstore     DSTORE  R:1201-1341 vectors back in the FBR basis (for DIPOLE3)

Parity pinned by tests/test_rot_oracle.py against the reference ROTLEV3
built from its sources (oracle/Makefile target rot) on the reference's own
H3+ J=1 rot.job and on a J=2 ABC deck, and against closed forms (coupling
constants, the Coriolis-free limit).
"""
from __future__ import annotations

import struct
from dataclasses import dataclass, field

import numpy as np

AUTOCM = 2.19474624e05       # R:1068
SQRT2 = 1.4142135623731      # R:693, the truncated constant of the reference


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
class KBlock:
    """One k block of the DVR3DRJZ zrot IWAVE file (D:973-977, 3841)."""
    xalf: np.ndarray
    kz: int
    maxleg: int
    idvr: int
    lincr: int
    pleg: np.ndarray          # (maxleg+1, idvr): pleg(l, k) * walf(k)
    eval: np.ndarray          # meval levels
    vecs: list                # meval records of idvr*npnt1*npnt2 amplitudes


@dataclass
class Iwave:
    idia: int
    ipar: int
    npntt: int
    npnt1: int
    npnt2: int
    jrot: int
    kmin0: int
    meval: int
    nlim: int
    zembed: bool
    zmors1: bool
    zmors2: bool
    xmass: np.ndarray
    g1: float
    g2: float
    z1da: bool
    zquad2: bool
    morse: np.ndarray         # re1, diss1, we1, re2, diss2, we2
    rm2: np.ndarray
    r1: np.ndarray
    r2: np.ndarray
    blocks: list = field(default_factory=list)


def read_iwave(path) -> Iwave:
    """The IWAVE file of a DVR3DRJZ zrot run (R:155-159, 325-374)."""
    r = read_records(path)
    h = _i(r[0])
    lg = _i(r[1][:12])
    xm = _d(r[1][12:36])
    g1, g2 = _d(r[1][36:52])
    z1da, zq2 = _i(r[1][52:60])
    w = Iwave(idia=int(h[0]), ipar=int(h[1]), npntt=int(h[2]), npnt1=int(h[3]), npnt2=int(h[4]),
              jrot=int(h[5]), kmin0=int(h[6]), meval=int(h[7]), nlim=int(h[8]),
              zembed=bool(lg[0]), zmors1=bool(lg[1]), zmors2=bool(lg[2]), xmass=xm.copy(),
              g1=float(g1), g2=float(g2), z1da=bool(z1da), zquad2=bool(zq2),
              morse=_d(r[2]).copy(), rm2=_d(r[3]).copy(), r1=_d(r[4]).copy(),
              r2=_d(r[5]).copy())
    pos = 6
    while pos < len(r):
        xalf = _d(r[pos]).copy()
        kz, maxleg, idvr, lincr = (int(x) for x in _i(r[pos + 1]))
        pleg = _d(r[pos + 2]).reshape(idvr, maxleg + 1).T.copy()
        meval = int(_i(r[pos + 3])[0])
        ev = _d(r[pos + 4]).copy()
        vecs = [r[pos + 5 + i] for i in range(meval)]
        w.blocks.append(KBlock(xalf, kz, maxleg, idvr, lincr, pleg, ev, vecs))
        pos += 5 + meval
    return w


@dataclass
class RotInput:
    """Namelist PRT and the cards of INSIZE (R:185-187, 208-209, 496-499)."""
    nvib: int
    neval: int
    kmin: int
    ibass: int = 0
    neval2: int = 0
    ezero: float = 0.0
    ztran: bool = False
    zvec: bool = False


@dataclass
class Selection:
    nblk: int
    nvib: int                 # largest per-block count after selection
    mvib: np.ndarray          # functions kept per block
    nbass: int
    neval: int
    neval2: int
    evib: list                # kept vibrational levels per block
    blocks: list              # the KBlock of each block used (kz = K ...)
    emin: float = 0.0
    emax: float = 0.0


def select(w: Iwave, inp: RotInput) -> Selection:
    """INSIZE sizes (R:185-196) and SELECT (R:301-478)."""
    meval = w.meval
    nvib = min(inp.nvib, meval)
    nblk = w.jrot + (1 if inp.kmin > 0 else 0)
    nbass = nblk * nvib
    if inp.ibass > 0:
        nbass = min(nbass, inp.ibass)
    neval = inp.neval if inp.neval > 0 else 10
    neval = min(neval, nbass)
    neval2 = inp.neval2 if inp.neval2 > 0 else neval
    if inp.kmin > 0 and w.kmin0 != 1:
        raise ValueError("KMIN > 0 needs a DVR3DRJZ run with KMIN = 1 (R:290-295)")
    # blocks in file order, k = 0 skipped for an f-only run (R:346-354)
    blocks = [b for b in w.blocks if not (b.kz == 0 and inp.kmin == 0)][:nblk]
    if len(blocks) < nblk:
        raise ValueError("IWAVE ends before the last k block")
    mvib = np.array([min(nvib, len(b.eval)) for b in blocks], dtype=np.int64)
    evib = [b.eval[:m].copy() for b, m in zip(blocks, mvib)]
    ipt = int(mvib.sum())
    ibass = inp.ibass
    sel = Selection(nblk=nblk, nvib=nvib, mvib=mvib.copy(), nbass=nbass, neval=neval,
                    neval2=neval2, evib=evib, blocks=blocks)
    if ibass <= 0 or ibass >= ipt:
        sel.nbass = ipt
        return sel
    # the ibass lowest functions over all blocks; on equal energies the
    # earlier block wins (strict < at R:415)
    iv = [0] * nblk
    p = 0
    evibr = 0.0
    for _ in range(ibass):
        while iv[p] >= mvib[p]:
            p += 1
        evibr = evib[p][iv[p]]
        jpt = p
        for j in range(p + 1, nblk):
            if iv[j] >= mvib[j]:
                continue
            if evib[j][iv[j]] >= evibr:
                continue
            evibr = evib[j][iv[j]]
            jpt = j
        iv[jpt] += 1
    sel.emin = min(e[0] for e in evib)
    sel.emax = evibr
    sel.mvib = np.array(iv, dtype=np.int64)
    sel.nvib = int(sel.mvib.max())
    sel.evib = [e[:m] for e, m in zip(evib, sel.mvib)]
    sel.nbass = ibass
    return sel


def _jstart(kz, jdia, ipar, maxleg, lincr):
    """First l and number of associated Legendre functions of a block
    (JTRAN R:879-885)."""
    jstart, jj0 = kz, -jdia
    if jstart % jdia != ipar:
        jj0 += 1
        jstart += 1
    nang = (maxleg + lincr - jstart) // jdia + 1
    return jstart, jj0, nang


def jtran(w: Iwave, b: KBlock, k: int, m: int):
    """The first m vibrational vectors of block b in the FBR basis,
    coef[j + mn*nang, l] = sum_a dvrvec_l[a, mn] pleg[jj_j, a]  (R:886-899).
    Returns (coef (nang*nrad, m), jstart, nang)."""
    jdia = max(w.idia, 1)
    nrad = w.npnt1 * w.npnt2
    jstart, jj0, nang = _jstart(k, jdia, w.ipar, b.maxleg, b.lincr)
    jj = jj0 + jdia * np.arange(1, nang + 1)
    if jj.min() < 0 or jj.max() > b.maxleg:
        raise ValueError("associated Legendre index outside the stored range")
    P = b.pleg[jj, :]                              # (nang, idvr)
    coef = np.empty((nang * nrad, m))
    for l in range(m):
        X = _d(b.vecs[l]).reshape(nrad, b.idvr).T  # dvrvec(idvr, nrad)
        coef[:, l] = (P @ X).T.reshape(-1)         # index j + mn*nang
    return coef, jstart, nang


def coriolis_terms(w: Iwave, k: int, jst1: int, maxlg1: int, jst2: int, maxlg2: int,
                   nblk1: int):
    """The (bra row, ket row, weight) triples of the rank-one updates that
    SOLRT accumulates into the block between k-1 and k (R:776-810):
    OFF(a, b) += cor * coef2[j, a] * coef1[i, b].  0-based rows."""
    jdia = max(w.idia, 1)
    jjp1 = w.jrot * (w.jrot + 1)
    kkp1 = k * (k - 1)
    n2max = 1 if w.zquad2 else w.npnt2
    J, I, C = [], [], []
    ii = j = 0
    ijump = int((jst2 - jst1) / jdia)              # Fortran integer division
    jjump = int(maxlg2 / jdia) - int(maxlg1 / jdia)
    for n1 in range(1, w.npnt2 + 1):
        for m1 in range(1, w.npnt1 + 1):
            ii += ijump
            for l1 in range(jst2, maxlg1 + 1, jdia):
                llp1 = l1 * (l1 + 1)
                cjlp = -np.sqrt(float((llp1 - kkp1) * (jjp1 - kkp1)))
                if k == 1:
                    cjlp = SQRT2 * cjlp
                ii += 1
                j += 1
                i = ii if w.zquad2 else (ii - 1) % nblk1 + 1
                for n2 in range(1, n2max + 1):
                    if w.zquad2:
                        i = ii
                        cor = cjlp * (w.rm2[n1 - 1] if w.zembed else w.rm2[m1 - 1])
                    else:
                        cor = cjlp * w.rm2[(n1 - 1) * w.npnt2 + n2 - 1]
                    J.append(j - 1)
                    I.append(i - 1)
                    C.append(cor)
                    i += nblk1
            j += jjump
    return np.array(J, np.int64), np.array(I, np.int64), np.array(C)


@dataclass
class RotHam:
    H: np.ndarray                 # nbass x nbass, both triangles
    offsets: np.ndarray           # first row of every block
    coef: list                    # FBR vectors of every block (nkbas x mvib)
    jstart: list
    nang: list
    offdiag: list                 # OFF of blocks (k-1, k), mvib[k] x mvib[k-1]


def hamiltonian(w: Iwave, sel: Selection, kmin: int) -> RotHam:
    """SOLRT (R:671-836) and LOADH (R:907-972) for all blocks."""
    jdia = max(w.idia, 1)
    nrad = w.npnt1 * w.npnt2
    mv = sel.mvib
    off = np.concatenate([[0], np.cumsum(mv)])
    nb = int(off[-1])
    H = np.zeros((nb, nb))
    H[np.arange(nb), np.arange(nb)] = np.concatenate(sel.evib) if nb else []
    k = 1 if kmin == 0 else 0
    coefs, jst, nangs, offd = [], [], [], []
    maxlg1 = jst1 = nblk1 = 0
    coef1 = None
    for bi, b in enumerate(sel.blocks):
        maxlg2 = b.maxleg + b.lincr
        if maxlg2 % jdia != w.ipar:
            maxlg2 -= 1
        ibass2 = b.idvr * nrad
        coef2, jst2, nang = jtran(w, b, k, int(mv[bi]))
        coefs.append(coef2)
        jst.append(jst2)
        nangs.append(nang)
        if bi > 0 and mv[bi] * mv[bi - 1] > 0:
            J, I, C = coriolis_terms(w, k, jst1, maxlg1, jst2, maxlg2, nblk1)
            O = (coef2[J, :] * C[:, None]).T @ coef1[I, :]     # mvib[bi] x mvib[bi-1]
            offd.append(O)
            H[off[bi]:off[bi + 1], off[bi - 1]:off[bi]] = O
            H[off[bi - 1]:off[bi], off[bi]:off[bi + 1]] = O.T
        elif bi > 0:
            offd.append(np.zeros((int(mv[bi]), int(mv[bi - 1]))))
        k += 1
        maxlg1, jst1, nblk1, coef1 = maxlg2, jst2, ibass2 // w.npnt2, coef2
    return RotHam(H=H, offsets=off, coef=coefs, jstart=jst, nang=nangs, offdiag=offd)


def solve(H: np.ndarray, neval: int):
    """Lowest neval eigenpairs (DGROT with DSYEVX 'V','I', R:1073-1085)."""
    e, v = np.linalg.eigh(H)
    return e[:neval].copy(), v[:, :neval].copy()


def dstore(rh: RotHam, first: int, vec: np.ndarray, mvib: np.ndarray):
    """Vectors in the FBR basis per block, d_k = coef_k C_k (R:1297-1323)."""
    out = []
    row = 0
    for bi in range(first, len(rh.coef)):
        m = int(mvib[bi])
        out.append(rh.coef[bi][:, :m] @ vec[row:row + m, :])
        row += m
    return out


def run(w: Iwave, inp: RotInput):
    """The parity sets of one ROTLEV3 run (VRMAIN R:482-629): set 1 over all
    blocks (e parity when kmin >= 1, f when kmin = 0), and for kmin > 1 set 2
    over the blocks k >= 1 (f parity).  J = 1 f sets are DVR3DRJZ's own k = 1
    levels (DSTORE1, R:1344-1475)."""
    sel = select(w, inp)
    rh = hamiltonian(w, sel, inp.kmin)
    sets = []
    if w.jrot == 1 and inp.kmin == 0:
        b = sel.blocks[0]
        sets.append(dict(kind="j1f", eval=b.eval[:sel.neval].copy()))
        return sel, rh, sets
    e, v = solve(rh.H, sel.neval)
    sets.append(dict(kind="solve", first=0, ibass=sel.nbass, eval=e, vec=v,
                     d=dstore(rh, 0, v, sel.mvib)))
    if inp.kmin > 1:
        ib2 = sel.nbass - int(sel.mvib[0])
        neval2 = min(sel.neval2, ib2)
        if w.jrot == 1:
            b = sel.blocks[1]
            sets.append(dict(kind="j1f", eval=b.eval[:neval2].copy()))
        else:
            m0 = int(sel.mvib[0])
            e2, v2 = solve(rh.H[m0:, m0:], neval2)
            sets.append(dict(kind="solve", first=1, ibass=ib2, eval=e2, vec=v2,
                             d=dstore(rh, 1, v2, sel.mvib)))
    return sel, rh, sets
