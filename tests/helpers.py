"""Shared helpers of the parity tests (test infrastructure)."""
import json
import os

import numpy as np

from oracle import dvr3d_oracle as O

REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
JOBS = os.path.join(REPO, "dvr3d_amd", "jobs")
GOLD = os.path.join(REPO, "tests", "golden")
PES_OF = {"c0": "ab2_c0.par", "c1": "ab2_c0.par", "c3": "ab2_c3.par", "c2": "abc_c2.par",
          "k2s": "abc_k2s.par"}


def job_and_pes(cfg):
    job = O.load_job(os.path.join(JOBS, cfg + ".job"))
    pes = O.PES.from_par(open(os.path.join(JOBS, PES_OF[cfg])).read())
    return job, pes


def golden(cfg):
    with open(os.path.join(GOLD, cfg + "_ref.json")) as f:
        return json.load(f)


def device_inputs(job, pes, nvec=0, solver=0, basis=None, kz=None):
    """Inputs of the C ABI computed from the oracle's basis arrays (one k
    block; kz defaults to the first block of the run)."""
    b = basis or O.setup_basis(job, kz)
    s, c = b.sizes, b.consts
    if s.zr2r1:
        rc, rb = b.r1, b.r2
        w1 = 0.5 / (rc * rc * c.ur1)
        w2 = 0.5 / (rb * rb * c.ur2)
        wsum = (w1[:, None] + w2[None, :]).ravel()
        xkb, xkc = b.xk2, b.xk1
    else:
        rc, rb = b.r2, b.r1
        w1 = 0.5 / (rb * rb * c.ur1)
        w2 = 0.5 / (rc * rc * c.ur2)
        wsum = (w1[None, :] + w2[:, None]).ravel()
        xkb, xkc = b.xk1, b.xk2
    # rotational term of mkham1 for J > 0 (D:2159-2173), same (c, b) layout
    wterm = None
    if s.jrot > 0:
        term = float(s.jrot * s.jrot + s.jrot - 2 * b.kz * b.kz)
        if s.zr2r1:
            w1g = np.broadcast_to(w1[:, None], (len(rc), len(rb)))
            w2b = np.broadcast_to(w2[None, :], (len(rc), len(rb)))
        else:
            w1g = np.broadcast_to(w1[None, :], (len(rc), len(rb)))
            w2b = np.broadcast_to(w2[:, None], (len(rc), len(rb)))
        wterm = np.ascontiguousarray(
            (term * w2b * c.ur2 / c.urr2) if s.zembed else (term * w1g * c.ur1 / c.urr1)).ravel()
    dims = dict(npnta=s.npnta, npntb=s.npntb, npntc=s.npntc, max2d=s.max2d,
                max3d=s.max3d, neval=s.neval, zcut=int(s.zcut), zall=int(s.zall),
                zr2r1=int(s.zr2r1), idvr=s.idvr, npnt1=s.npnt1, npnt2=s.npnt2,
                nvec=nvec, solver=solver)
    na = s.npnta
    xl = np.ascontiguousarray(b.xlmatr[:na, :na])
    par = None if pes is None else np.array([pes.de, pes.a, pes.re, pes.kb, pes.cte, pes.c])
    return dict(dims=dims, xlmatr=xl, xkb=np.asfortranarray(xkb).ravel(order="F"),
                xkc=np.asfortranarray(xkc).ravel(order="F"), wsum=wsum, rc=rc, rb=rb,
                xa=b.xalf[:na], par=par, emax1=job["emax1"] / O.AUTOCM,
                emax2=job["emax2"] / O.AUTOCM, basis=b, wterm=wterm)


def run_device(inp, analytic=True, vgrid=None):
    from dvr3d_amd import _lib
    ctx = _lib.Context(inp["dims"])
    ctx.set_basis(inp["xlmatr"], inp["xkb"], inp["xkc"], inp["wsum"], inp.get("wterm"),
                  inp["emax1"], inp["emax2"])
    if analytic:
        ctx.eval_potential_ab2(inp["rc"], inp["rb"], inp["xa"], inp["par"])
    else:
        ctx.set_potential(vgrid)
    ctx.run()
    return ctx


def write_h3p_table(tmp, name):
    """pes.par (kind 3 of the example POTV) and its table: the reference's
    H3+ POTV on every k block's grid of deck `name` (tests/golden/h3p_vgrid.npz)."""
    from dvr3d_amd import host
    V = np.load(os.path.join(GOLD, "h3p_vgrid.npz"))
    job = host.read_job(os.path.join(JOBS, name + ".job"))
    sz = host.sizes(job)
    rows = []
    for k in range(sz.kd, sz.ku + 1):
        b = host.Basis(job, k)
        x = b.xalf[: b.sz.npnta]
        if job.zr2r1:
            R1, R2, X = np.meshgrid(b.r1, b.r2, x, indexing="ij")
        else:
            R2, R1, X = np.meshgrid(b.r2, b.r1, x, indexing="ij")
        rows.append(np.stack([R1.ravel(), R2.ravel(), X.ravel(), V[f"{name}_k{k}"].ravel()], axis=1))
    tab = np.concatenate(rows)
    with open(os.path.join(tmp, "table.bin"), "wb") as f:
        np.array([len(tab)], dtype="<f8").tofile(f)
        tab.astype("<f8").tofile(f)
    with open(os.path.join(tmp, "pes.par"), "w") as f:
        f.write("3\n" + os.path.join(str(tmp), "table.bin") + "\n")


def write_water_potv_table(tmp):
    """pes.par (kind 3 of the example POTV) and its table: the reference's
    water POTV on the grid of the water J=0 deck (tests/golden/water_vgrid.npz)."""
    from oracle import dvr3d_oracle as O
    from oracle import nfmain_oracle as NF
    b = NF.nf_basis(O.load_job(os.path.join(JOBS, "water_J0.job")))
    s = b.sizes
    V = np.load(os.path.join(GOLD, "water_vgrid.npz"))["v"]
    G, B, A = np.meshgrid(np.arange(s.nalf), np.arange(s.npnt), np.arange(s.npnt), indexing="ij")
    tab = np.stack([b.r[A.ravel()], b.r[B.ravel()], b.xalf[G.ravel()], V.ravel()], axis=1)
    with open(os.path.join(tmp, "table.bin"), "wb") as f:
        np.array([len(tab)], dtype="<f8").tofile(f)
        tab.astype("<f8").tofile(f)
    with open(os.path.join(tmp, "pes.par"), "w") as f:
        f.write("3\n" + os.path.join(str(tmp), "table.bin") + "\n")


def write_water_prop_table(tmp):
    """prop.par (kind 3 of the example PROP) and its table: the reference's
    water property routine (properties 1-4) at every (r_a, r_b, x) of the
    water J=0 deck's grid (tests/golden/water_propgrid.npz)."""
    t = np.load(os.path.join(GOLD, "water_propgrid.npz"))
    r, x, v = t["r"], t["x"], t["v"]
    R1, R2, X = np.meshgrid(r, r, x, indexing="ij")
    tab = np.concatenate([np.stack([R1.ravel(), R2.ravel(), X.ravel()], axis=1),
                          v.reshape(-1, v.shape[-1])], axis=1)
    with open(os.path.join(tmp, "ptable.bin"), "wb") as f:
        np.array([len(tab), v.shape[-1]], dtype="<f8").tofile(f)
        tab.astype("<f8").tofile(f)
    with open(os.path.join(tmp, "prop.par"), "w") as f:
        f.write("3\n" + os.path.join(str(tmp), "ptable.bin") + "\n")
