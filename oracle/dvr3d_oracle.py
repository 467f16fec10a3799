"""CPU restatement of the DVR3DRJZ theta-first contraction path (FP64).

TEST INFRASTRUCTURE ONLY.  This module is the parity oracle: only tests/,
``__graft_entry__.smoke()`` and the ``cpu_baseline`` leg of ``bench.py`` may
import it, and only to check or time against.  The product path (the HIP
library behind ``include/dvr3d_amd.h`` and the Fortran host) never calls it.

Every function cites the reference routine it restates, as
``D:<lines>`` = ``source/dvr3drjz_segmented.f90:<lines>``.  The setup
routines (quadrature, kinetic matrices, Legendre functions, L-matrix) are
restated as scalar loops in the same operation order as the reference so
that the grid arrays agree with the reference executable to the last bit
or nearly so; the Hamiltonian stages are vectorised numpy and use LAPACK
(numpy.linalg.eigh) where the reference calls ``dsyev`` (D:2638-2659).

Scope: the ``jhmain`` driver (D:1915-2118) with ztheta=T, idia in
{1, 2, -1}, J=0 (or J>0 with zrot=F at fixed k), zquad2=T, Morse radial
bases (zmors1=zmors2=T), zcut T or F.  The reference's dead r2**-2 loop
(D:2464-2475, multiplied by rm2t=0 when zquad2=T) adds exactly +-0 and is
not executed here.
"""
from __future__ import annotations

import math
import re
from dataclasses import dataclass, field

import numpy as np

AMTOAU = 1.8228883e03          # D:1037
AUTOCM = 2.19474624e05         # D:2681, 2789, 2895, 2964


# ----------------------------------------------------------------------------
# .job input (D:65 namelist &PRT, D:232-239 cards, D:372 title,
# D:1042-1095 masses / cut-offs / radial basis cards; docum.tex:72-235)
# ----------------------------------------------------------------------------
NAMELIST_DEFAULTS = dict(       # block data, D:153-167
    zpham=False, zprad=False, zpvec=False, zrot=True, zladd=True,
    zembed=True, zmors2=True, zpmin=False, zvec=False, zquad2=True,
    zcut=False, zdiag=True, zlmat=False, zall=False, zp1d=False,
    zp2d=False, zr2r1=True, ztheta=True, zmors1=True, ztran=False,
    ztwod=False, zperp=False, zx=False, zs0=False, zs1=False,
    zlin=False, zpfun=False, ieigs1=7, ivecs1=3, ieigs2=2, ivecs2=4,
    ivint=17, iband=15, intvec=16, idiag1=20, idiag2=21, iout1=24,
    iout2=25, iwave=26, ilev=14)


def _fortran_value(tok: str):
    t = tok.strip().lower()
    if t in (".true.", ".t.", "t", "true", ".true"):
        return True
    if t in (".false.", ".f.", "f", "false", ".false"):
        return False
    try:
        return int(t)
    except ValueError:
        return float(t.replace("d", "e"))


def _ints(line: str, width: int = 5, n: int = 15):
    """Fortran (nIw) list: blank fields read as 0."""
    out = []
    for i in range(n):
        f = line[i * width:(i + 1) * width]
        out.append(int(f) if f.strip() else 0)
    return out


def _reals(line: str, width: int = 20, n: int = 3):
    """Fortran (nF20.0): blank fields read as 0.0; 'd' exponents accepted."""
    out = []
    for i in range(n):
        f = line[i * width:(i + 1) * width].strip().lower().replace("d", "e")
        out.append(float(f) if f else 0.0)
    return out


def parse_job(text: str) -> dict:
    """Parse a DVR3DRJZ .job deck (namelist + fixed-format cards)."""
    m = re.search(r"&prt(.*?)(/|&end)", text, re.I | re.S)
    if not m:
        raise ValueError("no &PRT namelist")
    nml = dict(NAMELIST_DEFAULTS)
    body = m.group(1)
    for key, val in re.findall(r"(\w+)\s*=\s*([^,\s/]+)", body):
        k = key.lower()
        if k not in nml:
            raise ValueError(f"unknown namelist item {key}")
        nml[k] = _fortran_value(val)
    rest = text[m.end():].split("\n")
    # the namelist record ends at the terminator; cards start on the next line
    rest = rest[1:]
    j = dict(nml=nml)
    j["ncoord"] = _ints(rest[0])[0]
    (j["npnt2"], j["jrot"], j["neval"], j["nalf"], j["max2d"], j["max3d"],
     j["idia"], j["kmin"], j["npnt1"], j["ipar"], j["max3d2"]) = _ints(rest[1])[:11]
    j["title"] = rest[2][:72]
    j["fixcos"] = _reals(rest[3], n=1)[0]
    j["xmass"] = _reals(rest[4])
    j["xmassr"] = _reals(rest[5])
    j["emax1"], j["emax2"] = _reals(rest[6], n=2)
    j["re1"], j["diss1"], j["we1"] = _reals(rest[7])
    j["re2"], j["diss2"], j["we2"] = _reals(rest[8])
    j["ezero_card"] = rest[9] if len(rest) > 9 else ""
    return j


# ----------------------------------------------------------------------------
# sizes (insize, D:171-638) and constants (setcon, D:1016-1188)
# ----------------------------------------------------------------------------
@dataclass
class Sizes:
    npnt1: int
    npnt2: int
    nalf: int
    nalf2: int
    idvr: int
    maxleg: int
    nmax1: int
    nmax2: int
    npnta: int
    npntb: int
    npntc: int
    max2d: int
    max3d: int
    neval: int
    jrot: int
    kmin: int
    idia: int
    ipar: int
    zr2r1: bool
    ztheta: bool
    zcut: bool
    zall: bool
    zrot: bool
    zembed: bool
    ztran: bool
    zvec: bool


def insize(job: dict) -> Sizes:
    """Derived sizes and clamps, D:232-355 (J=0 overrides D:242-246)."""
    n = dict(job["nml"])
    jrot, kmin, idia, ipar = job["jrot"], job["kmin"], job["idia"], job["ipar"]
    if jrot == 0:                                  # D:242-246
        n["zembed"] = True
        kmin = 0
        n["zrot"] = False
    if idia == -2:
        raise NotImplementedError("idia=-2 (nfmain) is outside the jhmain path")
    if n["ztran"]:                                 # D:264
        n["zvec"] = True
    if n["ztwod"] or n["zperp"]:
        raise NotImplementedError("ztwod / zperp are outside the jhmain path")
    if not n["ztheta"]:
        raise NotImplementedError("ztheta=F (r-first order) not restated")
    if not n["zquad2"] or not (n["zmors1"] and n["zmors2"]):
        raise NotImplementedError("only zquad2=T with Morse radial bases")
    nalf = job["nalf"]
    nalf2 = (nalf + 1) // 2                        # D:282
    maxleg = nalf - 1
    npnt1, npnt2 = job["npnt1"], job["npnt2"]
    npnta = nalf // max(1, idia)                   # D:311
    if n["zr2r1"]:
        npntb, npntc = npnt2, npnt1
    else:
        npntb, npntc = npnt1, npnt2
    max2d, max3d = job["max2d"], job["max3d"]
    if n["zall"]:                                  # D:344-350
        max2d = npnta * npntb
        max3d = max2d * npntc
    else:
        max2d = min(max2d, npnta * npntb)
        max3d = min(max3d, npnta * npntb * npntc)
    neval = job["neval"]
    if neval <= 0:
        neval = 10
    neval = min(max3d, neval)                      # D:358
    if abs(idia) == 2:                             # D:526-541
        if ipar not in (0, 1):
            raise ValueError("illegal ipar for idia=+-2")
        if idia == 2:
            idvr = nalf2
            if 2 * idvr != nalf:
                raise ValueError("nalf must be even when idia=2")
        else:
            idvr = nalf
    else:
        idvr = nalf
        ipar = 0
    return Sizes(npnt1=npnt1, npnt2=npnt2, nalf=nalf, nalf2=nalf2, idvr=idvr,
                 maxleg=maxleg, nmax1=npnt1 - 1, nmax2=npnt2 - 1, npnta=npnta,
                 npntb=npntb, npntc=npntc, max2d=max2d, max3d=max3d,
                 neval=neval, jrot=jrot, kmin=kmin, idia=idia, ipar=ipar,
                 zr2r1=n["zr2r1"], ztheta=n["ztheta"], zcut=n["zcut"],
                 zall=n["zall"], zrot=n["zrot"], zembed=n["zembed"],
                 ztran=n["ztran"], zvec=n["zvec"])


@dataclass
class Consts:
    g1: float
    g2: float
    ur1: float
    ur2: float
    urr1: float
    urr2: float
    beta1: float
    beta2: float
    a1: float
    a2: float
    iu1: int
    iu2: int


def setcon(job: dict) -> Consts:
    """Masses, g1/g2, reduced masses, Morse constants: D:1066-1126."""
    xm = job["xmass"]
    xmr = job["xmassr"] if job["xmassr"][0] > 0.0 else xm    # D:1049
    idia = job["idia"]
    if idia >= 1:                                  # D:1071-1074
        g1 = xm[1] / (xm[1] + xm[2])
        g2 = 0.0
    else:                                          # D:1076-1080
        a = math.sqrt(xm[2] / (xm[0] + xm[1] + xm[2]))
        b = xm[1] / (xm[0] + xm[1])
        g1 = 1.0 - a / (a + b - a * b)
        g2 = 1.0 - a / (1.0 - b + a * b)
    ur1 = AMTOAU / (g2 * g2 / xm[0] + 1.0 / xm[1] + (1.0 - g2) ** 2 / xm[2])
    ur2 = AMTOAU / (1.0 / xm[0] + g1 * g1 / xm[1] + (1.0 - g1) ** 2 / xm[2])
    urr1 = AMTOAU / (g2 * g2 / xmr[0] + 1.0 / xmr[1] + (1.0 - g2) ** 2 / xmr[2])
    urr2 = AMTOAU / (1.0 / xmr[0] + g1 * g1 / xmr[1] + (1.0 - g1) ** 2 / xmr[2])
    beta1 = job["we1"] * math.sqrt(0.5 * ur1 / job["diss1"])   # D:1123
    a1 = 4.0 * job["diss1"] / job["we1"]
    iu1 = int(a1 + 0.5)
    if job["diss2"] > 0.0 and job["we2"] > 0.0:
        beta2 = job["we2"] * math.sqrt(0.5 * ur2 / job["diss2"])   # D:1147
        a2 = 4.0 * job["diss2"] / job["we2"]
        iu2 = int(a2 + 0.5)
    else:
        # no r2 basis card (IDIA = -2 decks use the r1 basis for both bonds)
        beta2, a2, iu2 = 0.0, 0.0, 0
    return Consts(g1, g2, ur1, ur2, urr1, urr2, beta1, beta2, a1, a2, iu1, iu2)


# ----------------------------------------------------------------------------
# radial Gauss-Laguerre DVR (setfac/norms/lagpt/laguer/lgroot/lgrecr,
# D:1191-1492), kinetic integrals keints (D:1495-1535) and k1k2 (D:1603-1632)
# ----------------------------------------------------------------------------
def norms(npnt: int, alf: float):
    """dnorm(0:nmax) and cc = exact pseudo-weight sum, D:1261-1287."""
    fact = [1.0]
    count = 1.0
    for _ in range(npnt):                          # setfac D:1222-1227
        fact.append(count * fact[-1])
        count += 1.0
    count = float(npnt) + alf
    cc = fact[npnt - 1] / count
    npt1 = npnt + 1
    bin_ = [0.0] * (npt1 + 1)
    bin_[npt1] = 1.0
    for i in range(1, npnt + 1):
        nn = npt1 - i
        bin_[nn] = bin_[nn + 1] * count / float(i)
        count -= 1.0
    npt1 += 1
    dnorm = [math.sqrt(bin_[i + 1] * fact[i] * fact[npt1 - i - 2])
             for i in range(npnt)]
    return np.array(dnorm), cc


def _lgrecr(x, nn, alf, b, c):
    """Laguerre recurrence with derivative, D:1466-1492."""
    p1 = 1.0
    p = x - alf - 1.0
    dp1 = 0.0
    dp = 1.0
    for j in range(2, nn + 1):
        q = (x - b[j]) * p - c[j] * p1
        dq = (x - b[j]) * dp - c[j] * dp1 + p
        p1, p, dp1, dp = p, q, dp, dq
    return p, dp, p1


def _lgroot(x, nn, alf, b, c, eps):
    """Newton refinement, D:1448-1464 (at most 10 iterations)."""
    it = 0
    while True:
        it += 1
        p, dpn, pn1 = _lgrecr(x, nn, alf, b, c)
        d = p / dpn
        x = x - d
        if abs(d / x) <= eps or it >= 10:
            return x, dpn, pn1


def laguer(nn: int, alf: float, cc: float):
    """Gauss-Laguerre points and pseudo-weights, Stroud & Secrest search,
    D:1379-1446.  Returns (x, a, csx, csa, tsx)."""
    eps = 1.0e-12
    b = [0.0] * (nn + 2)
    c = [0.0] * (nn + 2)
    fa = alf + 1.0
    b[1] = fa
    c[1] = 0.0
    fn = 1.0
    for j in range(2, nn + 1):
        fa = fa + 2.0
        b[j] = fa
        c[j] = fn * (alf + fn)
        fn = fn + 1.0
    tsx = fn * (alf + fn)
    xt1 = 0.0
    xt = (1.0 + alf) * (2.0 + alf) / (1.0 + 3.0 * fn + 2.0 * alf)
    step = 3.0 * (1.0 + alf) / (1.0 + 3.0 * fn + alf)
    pt, _, _ = _lgrecr(xt, nn, alf, b, c)
    x = np.zeros(nn)
    a = np.zeros(nn)
    csx = 0.0
    csa = 0.0
    xt2 = 0.0
    for i in range(1, nn + 1):
        if i <= 2:                                 # brute-force bracket
            while True:
                xt2 = xt + step
                pt2, _, _ = _lgrecr(xt2, nn, alf, b, c)
                if math.copysign(1.0, pt) * math.copysign(1.0, pt2) > 0.0:
                    pt = pt2
                    xt = xt2
                    continue
                pt = pt2
                xt = 0.5 * (xt + xt2)
                break
        else:                                      # Stroud & Secrest step
            fi = float(i - 2)
            r1 = (1.0 + 2.55 * fi) / (1.9 * fi)
            r2 = 1.26 * fi * alf / (1.0 + 3.5 * fi)
            ratio = (r1 + r2) / (1.0 + 0.3 * alf)
            xt = xt + ratio * (xt - xt2)
        xt, dpn, pn1 = _lgroot(xt, nn, alf, b, c, eps)
        xt2 = xt1
        xt1 = xt
        x[i - 1] = xt
        a[i - 1] = cc / dpn / pn1
        csx += xt
        csa += a[i - 1]
    return x, a, csx, csa, tsx


def lagpt(npnt: int, re_: float, beta: float, a: float, iu: int):
    """Points r, sqrt-able weights wt and normalised Laguerre basis
    bass(0:nmax, npnt), D:1290-1376 (Morse branch, zmorse=T)."""
    alf = float(iu)
    dnorm, cc = norms(npnt, alf)
    y, wt, csx, csa, tsx = laguer(npnt, alf, cc)
    tsa = 1.0 / (dnorm[0] * dnorm[0])
    nmax = npnt - 1
    r = np.array([re_ + math.log(a / y[i]) / beta for i in range(npnt)])
    bass = np.zeros((nmax + 1, npnt))
    alfm1 = alf - 1.0
    for i in range(npnt):
        bass[0, i] = 1.0
        if nmax >= 1:
            amx = alf + 1.0 - y[i]
            bass[1, i] = amx
            en = 1.0
            for n in range(2, nmax + 1):
                en += 1.0
                amx += 2.0
                bass[n, i] = (amx * bass[n - 1, i] - (alfm1 + en) * bass[n - 2, i]) / en
        for n2 in range(nmax + 1):
            bass[n2, i] = bass[n2, i] * dnorm[n2]
    if abs((csx - tsx) / tsx) > 1.0e-8 or abs((csa - tsa) / tsa) > 1.0e-8:
        raise RuntimeError("points & weights in error (D:1357-1365)")
    return dict(y=y, r=r, wt=wt, bass=bass, csx=csx, csa=csa, tsx=tsx, tsa=tsa)


def keints(fke: float, nmax: int, iu: int) -> np.ndarray:
    """Analytic Morse-like kinetic energy integrals (packed lower), D:1495-1535.
    Returned as a full symmetric (nmax+1)^2 matrix."""
    n = nmax + 1
    h = np.zeros((n, n))
    for n2 in range(1, n + 1):
        for n1 in range(1, n2 + 1):
            if n1 == n2:
                v = fke * float(2 * (n2 - 1) * (n2 + iu) + iu + 1)
            elif n1 + 2 == n2:
                v = -fke * math.sqrt(float((iu + n1 + 1) * (iu + n1)) * float((n1 + 1) * n1))
            else:
                v = 0.0
            h[n2 - 1, n1 - 1] = v
            h[n1 - 1, n2 - 1] = v
    return h


def k1k2(hbl: np.ndarray, bass: np.ndarray, wt: np.ndarray) -> np.ndarray:
    """DVR kinetic matrix xk(k,kp) = sum_{m,mp} w_k w_kp B(m,k) T(m,mp) B(mp,kp),
    D:1603-1632, accumulated in the reference's order."""
    nmax1, npnt = bass.shape
    xk = np.zeros((npnt, npnt))
    for k in range(npnt):
        for kp in range(k + 1):
            wtkkp = wt[k] * wt[kp]
            s = 0.0
            for m in range(nmax1):
                t = bass[m, k] * wtkkp
                for mp in range(nmax1):
                    s = s + (hbl[m, mp] * t * bass[mp, kp])
            xk[k, kp] = s
            xk[kp, k] = s
    return xk


# ----------------------------------------------------------------------------
# angular Gauss-(associated) Legendre DVR: jacobi/root/recur (D:1635-1744),
# allpts (D:1747-1772), asleg (D:1775-1843), lmatrx (D:1846-1911)
# ----------------------------------------------------------------------------
def _recur(x, nn, alf, bta, b, c):
    """Jacobi polynomial recurrence, D:1725-1744."""
    p1 = 1.0
    p = x + (alf - bta) / (alf + bta + 2.0)
    dp1 = 0.0
    dp = 1.0
    for j in range(2, nn + 1):
        q = (x - b[j]) * p - c[j] * p1
        dq = (x - b[j]) * dp + p - c[j] * dp1
        p1, p, dp1, dp = p, q, dp, dq
    return p, dp, p1


def jacobi(nn: int, nn2: int, alf: float, bta: float, tsa: float):
    """Half-range Gauss-Jacobi points/weights by the 1e-4 sign scan + Newton,
    D:1635-1706.  Returns (x[nn2], a[nn2], csa)."""
    eps = 1.0e-12
    b = [0.0] * (nn + 1)
    c = [0.0] * (nn + 1)
    for i in range(2, nn + 1):
        xi = float(i)
        c[i] = (4.0 * (xi - 1.0) * (alf + xi - 1.0) * (bta + xi - 1.0) * (alf + bta + xi - 1.0)
                / ((alf + bta + 2.0 * xi - 1.0) * (alf + bta + 2.0 * xi - 2.0)
                   * (alf + bta + 2.0 * xi - 2.0) * (alf + bta + 2.0 * xi - 3.0)))
    cc = tsa
    for i in range(2, nn + 1):
        cc = cc * c[i]
    # scan step is the default-real literal 0.0001 (D:1685): its double value
    step = float(np.float32(0.0001))
    xt = []
    pm1 = 1.0
    xxx = 1.0
    while True:
        p, _, _ = _recur(xxx, nn, alf, bta, b, c)
        if pm1 * p < 0.0:
            pm1 = -pm1
            xt.append(xxx)
        if len(xt) == nn2:
            break
        xxx = xxx - step
        if not xxx > -0.0002:
            raise RuntimeError(f"incorrect number {len(xt) - 1} of zeros found in JACOBI")
    x = np.zeros(nn2)
    a = np.zeros(nn2)
    csa = 0.0
    for i in range(nn2):
        xr = xt[i]
        it = 0
        while True:                                # root, D:1708-1723
            it += 1
            p, dp, pn1 = _recur(xr, nn, alf, bta, b, c)
            d = p / dp
            xr = xr - d
            if abs(d) <= eps or it >= 10:
                break
        dpn = dp
        x[i] = xr
        a[i] = cc / (dpn * pn1)
        csa = csa + a[i] + a[i]
    if 2 * nn2 != nn:
        csa = csa - a[nn2 - 1]
    return x, a, csa


def tswalf_exact(kz: int) -> float:
    """Exact sum of Gauss-Jacobi weights, D:888-892."""
    t = 2.0 ** (kz + kz + 1) / float(kz + 1)
    for ia in range(1, kz + 1):
        t = t * float(ia) / float(kz + ia + 1)
    return t


def allpts(xh, wh, nang: int, nang2: int, idia: int):
    """Full-range points and sqrt weights, D:1747-1772."""
    scale = float(max(1, idia))
    if idia == 2:
        xalf = np.array(xh[:nang2], dtype=float)
        walf = np.sqrt(scale * np.asarray(wh[:nang2]))
        return xalf, walf
    xalf = np.zeros(nang)
    walf = np.zeros(nang)
    for i in range(nang2):
        w = math.sqrt(scale * wh[i])
        walf[i] = w
        xalf[nang - 1 - i] = xh[i]
        xalf[i] = -xh[i]
        walf[nang - 1 - i] = w
    return xalf, walf


def asleg(lmax: int, x, kz: int, lincr: int) -> np.ndarray:
    """Normalised associated Legendre polynomial parts pleg(0:lmax, n), D:1775-1843."""
    m = kz
    nn2 = len(x)
    pleg = np.zeros((lmax + 1 + lincr, nn2))
    for i in range(nn2):
        pmm = 1.0
        fact = 1.0
        for _ in range(m):
            pmm = -pmm * fact
            fact = fact + 2.0
        pleg[0, i] = pmm
        pmmp1 = x[i] * (m + m + 1) * pmm
        if lmax + lincr >= 1:
            pleg[1, i] = pmmp1
        ll = 1
        for l in range(2 + m, lmax + lincr + 1):
            r2lm1 = float(l + l - 1)
            rlpmm1 = float(l + m - 1)
            rlmm = float(l - m)
            pll = (x[i] * r2lm1 * pmmp1 - rlpmm1 * pmm) / rlmm
            pmm = pmmp1
            pmmp1 = pll
            ll += 1
            pleg[ll, i] = pll
    pnorm = []
    for j in range(m, lmax + lincr + 1):
        fact = 1.0
        for i in range(j - m + 1, j + m + 1):
            fact = fact * float(i)
        rj = float(j)
        pnorm.append(math.sqrt((rj + rj + 1.0) / (fact + fact)))
    for i in range(nn2):
        for jj in range(len(pnorm)):
            pleg[jj, i] = pleg[jj, i] * pnorm[jj]
    return pleg[:lmax + 1]


def lmatrx(pleg, walf, kz: int, ipar0: int, lincr: int, idia: int, maxleg: int):
    """Transformed angular-momentum matrix L(a,a'), D:1846-1911 (zperp=F)."""
    nidvr = len(walf)
    jstart = kz
    iparam = lincr
    jdia = max(1, idia)
    jj0 = -jdia
    if jstart % jdia != ipar0:
        jj0 += 1
        jstart += 1
    xl = np.zeros((nidvr, nidvr))
    js = list(range(jstart, maxleg + iparam + 1, jdia))
    for k in range(nidvr):
        term = walf[k]
        for kp in range(k, nidvr):
            s = 0.0
            jj = jj0
            for j in js:
                jj += jdia
                s = s + pleg[jj, k] * pleg[jj, kp] * float((j + 1) * j)
            v = s * term * walf[kp]
            xl[kp, k] = v
            xl[k, kp] = v
    return xl


@dataclass
class Basis:
    sizes: Sizes
    consts: Consts
    r1: np.ndarray
    r2: np.ndarray
    xk1: np.ndarray
    xk2: np.ndarray
    xalf: np.ndarray
    walf: np.ndarray
    pleg: np.ndarray
    xlmatr: np.ndarray
    kz: int
    lincr: int
    nidvr: int
    cswalf: float
    tswalf: float
    rad1: dict = field(default_factory=dict)
    rad2: dict = field(default_factory=dict)


def setup_basis(job: dict, kz: int | None = None) -> Basis:
    """ccmain set-up, D:750-982, for one k block (J=0: kz=0)."""
    s = insize(job)
    c = setcon(job)
    rad1 = lagpt(s.npnt1, job["re1"], c.beta1, c.a1, c.iu1)
    rad2 = lagpt(s.npnt2, job["re2"], c.beta2, c.a2, c.iu2)
    fke1 = c.beta1 * c.beta1 / (8.0 * c.ur1)       # D:745
    fke2 = c.beta2 * c.beta2 / (8.0 * c.ur2)       # D:739
    hbl1 = keints(fke1, s.nmax1, c.iu1)
    hbl2 = keints(fke2, s.nmax2, c.iu2)
    wt1 = np.sqrt(rad1["wt"])                      # D:797-798
    wt2 = np.sqrt(rad2["wt"])
    xk1 = k1k2(hbl1, rad1["bass"], wt1)
    xk2 = k1k2(hbl2, rad2["bass"], wt2)
    if kz is None:
        kz = s.kmin if not s.zrot else 0
    tsw = tswalf_exact(kz)
    nidvr, nang, nang2, lincr = s.idvr, s.nalf, s.nalf2, kz   # zladd branch D:895-899
    xh, wh, csw = jacobi(nang, nang2, float(kz), float(kz), tsw)
    if abs((csw - tsw) / tsw) > 1.0e-8:
        raise RuntimeError("angular points & weights in error (D:939-943)")
    xalf, walf = allpts(xh, wh, nang, nang2, s.idia)
    pleg = asleg(s.maxleg, xalf, kz, lincr)
    ipar0 = 1 if (s.idia == 2 and s.ipar == 1) else 0
    xl = lmatrx(pleg, walf, kz, ipar0, lincr, s.idia, s.maxleg)
    return Basis(sizes=s, consts=c, r1=rad1["r"], r2=rad2["r"], xk1=xk1, xk2=xk2,
                 xalf=xalf, walf=walf, pleg=pleg, xlmatr=xl, kz=kz, lincr=lincr,
                 nidvr=nidvr, cswalf=csw, tswalf=tsw, rad1=rad1, rad2=rad2)


# ----------------------------------------------------------------------------
# synthetic potentials (the POTV plug-in, docum.tex:31-36)
# ----------------------------------------------------------------------------
@dataclass
class PES:
    kind: int
    de: float
    a: float
    re: float
    kb: float
    te_deg: float
    c: float

    @property
    def cte(self) -> float:
        return math.cos(self.te_deg * (4.0 * math.atan(1.0)) / 180.0)

    @staticmethod
    def from_par(text: str) -> "PES":
        lines = [l for l in text.split("\n") if l.strip()]
        kind = int(lines[0].split()[0])
        de, a, re_ = (float(v) for v in lines[1].split()[:3])
        kb, te, c = (float(v) for v in lines[2].split()[:3])
        return PES(kind, de, a, re_, kb, te, c)

    # kind 2 (non-symmetric ABC in Jacobi coordinates, IDIA=1): the fields
    # are read as de, a, re = r1 Morse minimum, kb, te_deg = A-B Morse
    # minimum (bohr), c = noise amplitude (Eh); g1 is the Jacobi mass factor
    # m2/(m2+m3) and NOISE_SEED keys the noise.
    g1: float = 0.0
    NOISE_SEED = 20251203

    def __call__(self, r1, r2, x):
        if self.kind == 2:
            return self._abc(r1, r2, x)
        if self.kind != 1:
            raise ValueError("unknown PES kind")
        m1 = (1.0 - np.exp(-self.a * (r1 - self.re))) ** 2
        m2 = (1.0 - np.exp(-self.a * (r2 - self.re))) ** 2
        return self.de * (m1 + m2) + self.kb * (x - self.cte) ** 2 + self.c * ((r1 - self.re) * (r2 - self.re))

    def _abc(self, r1, r2, x):
        """Two Morse terms, in r1 (the B-C distance, atoms 2-3) and in the A-B
        distance d = |atom 1 - atom 2| = sqrt(yy^2 + r2^2 + 2 yy r2 x),
        yy = r1 (1 - g1) (the water convention of the reference's
        pot_water.f for g2 = 0), plus kb (x - 1)^2 and index-free noise: a
        standard normal drawn from a hash of the bit patterns of (r1, r2, x)
        and NOISE_SEED, times c."""
        r1, r2, x = np.broadcast_arrays(np.asarray(r1, float), np.asarray(r2, float),
                                        np.asarray(x, float))
        yy = r1 * (1.0 - self.g1)
        d = np.sqrt(yy * yy + r2 * r2 + 2.0 * yy * r2 * x)
        m1 = (1.0 - np.exp(-self.a * (r1 - self.re))) ** 2
        m2 = (1.0 - np.exp(-self.a * (d - self.te_deg))) ** 2
        v = self.de * (m1 + m2) + self.kb * (x - 1.0) ** 2
        if self.c != 0.0:
            v = v + self.c * noise_normal(r1, r2, x, self.NOISE_SEED)
        return v


_M64 = (1 << 64) - 1


def _splitmix(z):
    """splitmix64 finaliser on uint64 arrays (wrapping arithmetic)."""
    z = (z + np.uint64(0x9E3779B97F4A7C15)) & np.uint64(_M64)
    z = ((z ^ (z >> np.uint64(30))) * np.uint64(0xBF58476D1CE4E5B9)) & np.uint64(_M64)
    z = ((z ^ (z >> np.uint64(27))) * np.uint64(0x94D049BB133111EB)) & np.uint64(_M64)
    return z ^ (z >> np.uint64(31))


def noise_normal(r1, r2, x, seed):
    """Standard normal numbers keyed by the IEEE bit patterns of (r1, r2, x):
    h = mix(bits(x) ^ mix(bits(r2) ^ mix(bits(r1) ^ seed))), two uniforms
    from h and mix(h), Box-Muller (cosine branch)."""
    with np.errstate(over="ignore"):
        b1 = np.ascontiguousarray(r1, dtype=np.float64).view(np.uint64)
        b2 = np.ascontiguousarray(r2, dtype=np.float64).view(np.uint64)
        b3 = np.ascontiguousarray(x, dtype=np.float64).view(np.uint64)
        h = _splitmix(b1 ^ np.uint64(seed))
        h = _splitmix(b2 ^ h)
        h = _splitmix(b3 ^ h)
        h2 = _splitmix(h)
    u1 = ((h >> np.uint64(11)).astype(np.float64) + 1.0) * (1.0 / 9007199254740992.0)
    u2 = (h2 >> np.uint64(11)).astype(np.float64) * (1.0 / 9007199254740992.0)
    return np.sqrt(-2.0 * np.log(u1)) * np.cos(2.0 * math.pi * u2)


def potential_grid(basis: Basis, pes) -> np.ndarray:
    """V[c, b, a] on the (npntc, npntb, npnta) grid in theta-first order."""
    s = basis.sizes
    x = basis.xalf[: s.npnta]
    if s.zr2r1:
        R1, R2, X = np.meshgrid(basis.r1, basis.r2, x, indexing="ij")   # c=r1, b=r2
    else:
        R2, R1, X = np.meshgrid(basis.r2, basis.r1, x, indexing="ij")   # c=r2, b=r1
    return np.asarray(pes(R1, R2, X), dtype=np.float64)


# ----------------------------------------------------------------------------
# the jhmain contraction stages, D:1915-2118
# ----------------------------------------------------------------------------
def mkham1_all(basis: Basis, vgrid: np.ndarray) -> np.ndarray:
    """All 1D angular Hamiltonians H1[c, b] (npnta x npnta, full symmetric),
    D:2121-2225 with ztheta=T: diag(V + wterm) + (w1+w2) L."""
    s, c = basis.sizes, basis.consts
    rc = basis.r1 if s.zr2r1 else basis.r2
    rb = basis.r2 if s.zr2r1 else basis.r1
    if s.zr2r1:
        w1 = 0.5 / (rc * rc * c.ur1)               # w1gama over c (r1)
        w2 = 0.5 / (rb * rb * c.ur2)               # w2beta over b (r2)
        wsum = w1[:, None] + w2[None, :]
        w1g = np.broadcast_to(w1[:, None], wsum.shape)
        w2b = np.broadcast_to(w2[None, :], wsum.shape)
    else:
        w1 = 0.5 / (rb * rb * c.ur1)               # r1 index is b
        w2 = 0.5 / (rc * rc * c.ur2)
        wsum = w1[None, :] + w2[:, None]
        w1g = np.broadcast_to(w1[None, :], wsum.shape)
        w2b = np.broadcast_to(w2[:, None], wsum.shape)
    term = float(s.jrot * s.jrot + s.jrot - 2 * basis.kz * basis.kz)
    wterm = np.zeros_like(wsum)
    if s.jrot > 0:                                 # D:2159-2173 (zx term off)
        wterm = term * w2b * c.ur2 / c.urr2 if s.zembed else term * w1g * c.ur1 / c.urr1
    na = s.npnta
    L = basis.xlmatr[:na, :na]
    H = wsum[:, :, None, None] * L[None, None, :, :]
    idx = np.arange(na)
    H[:, :, idx, idx] += vgrid + wterm[:, :, None]
    return H


def solve_1d(basis: Basis, vgrid: np.ndarray, emax1_cm: float):
    """diag + cut1d over every (c, b) point, D:1975-2003 and D:2870-2938.
    Returns per-c lists of (eigs1d, vecs1d) and iv1[c, b]."""
    s = basis.sizes
    H = mkham1_all(basis, vgrid)
    e, v = np.linalg.eigh(H)                       # columns = eigenvectors
    emaxau = emax1_cm / AUTOCM                     # D:2891
    iv1 = np.zeros((s.npntc, s.npntb), dtype=np.int64)
    eigs1d, vecs1d = [], []
    for ic in range(s.npntc):
        es, vs = [], []
        for ib in range(s.npntb):
            n = s.npnta if s.zall else int(np.count_nonzero(
                np.cumprod(e[ic, ib] <= emaxau)))
            iv1[ic, ib] = n
            es.append(e[ic, ib, :n])
            vs.append(v[ic, ib][:, :n].T)          # rows = kept functions
        ec = np.concatenate(es)
        if ec.size > s.max2d and not s.zall:
            raise RuntimeError("**** core exceeded: reduce cut-off emax1 ****")
        eigs1d.append(ec)
        vecs1d.append(np.concatenate(vs, axis=0) if ec.size else np.zeros((0, s.npnta)))
    return dict(e1=e, v1=v, iv1=iv1, eigs1d=eigs1d, vecs1d=vecs1d)


def mkham2(basis: Basis, eigs1d, vecs1d, iv1_row) -> np.ndarray:
    """2D Hamiltonian for one c block, D:2228-2298:
    H2 = (V1 V1^T) * xk2-blocks + diag(E1)."""
    s = basis.sizes
    if s.zr2r1:
        xk = basis.xk2
    else:
        xk = basis.xk1
    V = vecs1d
    S = V @ V.T
    blk = np.repeat(np.arange(s.npntb), iv1_row)
    H = S * xk[np.ix_(blk, blk)]
    H[np.diag_indices_from(H)] += eigs1d
    return H


def choose(eigs2: list, max3d: int):
    """Global selection of the lowest max3d 2D levels, D:2767-2867.
    Strict '<' keeps the lowest block index on ties (D:2808-2813)."""
    ndim = [len(e) for e in eigs2]
    low3d = min(max3d, sum(ndim))
    keys = []
    for ib, e in enumerate(eigs2):
        for i, val in enumerate(e):
            keys.append((val, ib, i))
    keys.sort()
    iv2 = np.zeros(len(eigs2), dtype=np.int64)
    for val, ib, i in keys[:low3d]:
        iv2[ib] += 1
    return iv2, low3d


def cut2d(eig2: np.ndarray, emax2au: float, zall: bool) -> int:
    """Energy cut of one block's 2D levels, D:2941-3003."""
    if zall:
        return len(eig2)
    return int(np.count_nonzero(np.cumprod(eig2 <= emax2au)))


def cint_block(basis: Basis, vecs1d, iv1_row, c2) -> np.ndarray:
    """Expansion of the chosen 2D vectors on the (angle, b) grid, D:2362-2411:
    cint[(a-1)*npntb + b, j] = sum_{i in b} c2[i, j] * vecs1d[i, a]."""
    s = basis.sizes
    G = s.npnta * s.npntb
    out = np.zeros((G, c2.shape[1]))
    off = 0
    for ib in range(s.npntb):
        n = iv1_row[ib]
        if n:
            blk = vecs1d[off:off + n].T @ c2[off:off + n]    # (npnta, m)
            out[ib::s.npntb, :] = blk
        off += n
    return out


def k_blocks(job: dict):
    """The k values of the ccmain rotational loop, D:811-819 and 857-865:
    zrot -> k = 1-min(kmin,1) .. jrot (kmin = 2 adds a second k = 1 block,
    not restated); otherwise the single block k = kmin."""
    s = insize(job)
    if not s.zrot:
        return [s.kmin]
    if s.kmin >= 2:
        raise NotImplementedError("kmin=2 (both parities of k=1) not restated")
    return list(range(1 - min(s.kmin, 1), s.jrot + 1))


def run_all_k(job: dict, pes, **kw):
    """jhmain for every k block of the run (one result dict per k)."""
    return [run(job, pes, kz=k, **kw) for k in k_blocks(job)]


def run(job: dict, pes, keep_vectors: bool = False, neval_vectors: int | None = None,
        kz: int | None = None, vgrid: np.ndarray | None = None):
    """Full theta-first jhmain contraction for one k block (default: the
    first k of the run).  vgrid, when given, is the potential V[c, b, a] on
    this block's grid (tabulated from another POTV); pes is then unused.
    Returns a dict of every intermediate that parity tests compare."""
    basis = setup_basis(job, kz)
    s = basis.sizes
    if vgrid is None:
        if getattr(pes, "kind", 1) == 2:
            pes.g1 = basis.consts.g1
        vgrid = potential_grid(basis, pes)
    else:
        vgrid = np.asarray(vgrid, dtype=np.float64).reshape(s.npntc, s.npntb, s.npnta)
    one = solve_1d(basis, vgrid, job["emax1"])
    iv1 = one["iv1"]
    eig2_all, vec2_all = [], []
    ndim2d = iv1.sum(axis=1)
    for ic in range(s.npntc):
        if ndim2d[ic] == 0:
            eig2_all.append(np.zeros(0))
            vec2_all.append(np.zeros((0, 0)))
            continue
        H2 = mkham2(basis, one["eigs1d"][ic], one["vecs1d"][ic], iv1[ic])
        e2, v2 = np.linalg.eigh(H2)
        eig2_all.append(e2)
        vec2_all.append(v2)
    if s.zcut or s.zall:
        iv2 = np.array([cut2d(e, job["emax2"] / AUTOCM, s.zall) for e in eig2_all])
        low3d = int(iv2.sum())
        if low3d > s.max3d and not s.zall:
            raise RuntimeError("**** core exceeded: reduce cut-off emax2 ****")
    else:
        iv2, low3d = choose(eig2_all, s.max3d)
    nham3 = low3d
    # 3D assembly, mkham3 loop 81 (D:2459-2463) + loadh (D:2554-2609)
    xk3 = basis.xk1 if s.zr2r1 else basis.xk2
    cints = []
    for ic in range(s.npntc):
        m = int(iv2[ic])
        if m == 0:
            cints.append(np.zeros((s.npnta * s.npntb, 0)))
            continue
        cints.append(cint_block(basis, one["vecs1d"][ic], iv1[ic], vec2_all[ic][:, :m]))
    C = np.concatenate(cints, axis=1)
    blk3 = np.repeat(np.arange(s.npntc), iv2)
    H3 = (C.T @ C) * xk3[np.ix_(blk3, blk3)]
    e2sel = np.concatenate([eig2_all[ic][: iv2[ic]] for ic in range(s.npntc)])
    H3[np.diag_indices_from(H3)] += e2sel
    meval = min(s.neval, nham3)
    if keep_vectors:
        e3, v3 = np.linalg.eigh(H3)
    else:
        import scipy.linalg as sl
        e3 = sl.eigh(H3, eigvals_only=True, subset_by_index=(0, meval - 1),
                     driver="evr")
        v3 = None
    out = dict(basis=basis, vgrid=vgrid, iv1=iv1, e1=one["e1"], ndim2d=ndim2d,
               eig2=eig2_all, iv2=iv2, nham3=nham3, meval=meval,
               eval=e3[:meval], H3=H3 if keep_vectors else None,
               cint=C if keep_vectors else None,
               e2sel=e2sel, vec2=vec2_all if keep_vectors else None,
               one=one if keep_vectors else None)
    if keep_vectors:
        nv = meval if neval_vectors is None else min(neval_vectors, meval)
        out["evec"] = v3[:, :nv]
        out["psi"] = trans(basis, C, iv2, v3[:, :nv])
    return out


def trans(basis: Basis, C: np.ndarray, iv2, c3: np.ndarray) -> np.ndarray:
    """Wavefunction amplitudes psi(idvr, npnt1, npnt2) per level, D:3650-3845
    (theta-first branches D:3771-3803).  C[(a,b), j] is phi(j, a, b)."""
    s = basis.sizes
    nlev = c3.shape[1]
    psi = np.zeros((nlev, s.npnt2, s.npnt1, s.idvr))   # C order: (l, r2, r1, a)
    off = 0
    for ic in range(s.npntc):
        m = int(iv2[ic])
        if m == 0:
            continue
        blk = C[:, off:off + m] @ c3[off:off + m]     # ((a,b), l)
        blk = blk.reshape(s.npnta, s.npntb, nlev)      # (a, b, l)
        if s.zr2r1:                                    # psi(a, ic=r1, ib=r2)
            psi[:, :, ic, : s.npnta] += np.transpose(blk, (2, 1, 0))
        else:                                          # psi(a, ib=r1, ic=r2)
            psi[:, ic, :, : s.npnta] += np.transpose(blk, (2, 1, 0))
        off += m
    return psi


def load_job(path: str) -> dict:
    with open(path) as f:
        return parse_job(f.read())
