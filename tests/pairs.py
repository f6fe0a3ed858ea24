"""Shared set-up of the parity tests: one oracle state and one GPU DataBlock built from
the same configuration and the same initial data, plus the comparison helpers."""
import numpy as np

from oracle import oracle as O
from oracle.oracle import OracleGrid  # noqa: F401

# FP64 parity bar: 1e-12 relative per cell (absolute floor 1e-12 x field scale)
RTOL = 1e-12


def cfg_kwargs(dim, nx, mhd=True, recon="plm", solver="hlld", emf="uct_contact",
               bc="periodic", gamma=5.0 / 3.0, cfl=0.6, nstages=2, xbeg=(0, 0, 0),
               xend=(1, 1, 1), first_dt=1e-4, comps=None, ntracer=0, shock_flattening=None,
               csiso=None):
    return dict(dim=dim, nx=nx, mhd=mhd, recon=recon, solver=solver, emf=emf, bc=bc,
                gamma=gamma, cfl=cfl, nstages=nstages, xbeg=xbeg, xend=xend, first_dt=first_dt,
                comps=dim if comps is None else comps, ntracer=ntracer,
                shock_flattening=shock_flattening, csiso=csiso)


def make_oracle(kw):
    cfg = O.make_cfg(kw["dim"], kw.get("comps", kw["dim"]), mhd=kw["mhd"], recon=kw["recon"], solver=kw["solver"],
                     emf=kw["emf"], nstages=kw["nstages"], nx=kw["nx"], xbeg=kw["xbeg"],
                     xend=kw["xend"], bc=kw["bc"], gamma=kw["gamma"], cfl=kw["cfl"],
                     first_dt=kw["first_dt"], ntracer=kw.get("ntracer", 0),
                     shock_flattening=kw.get("shock_flattening"), csiso=kw.get("csiso"))
    return O.Oracle(cfg)


def make_pair(kw, ic, precision="fp64"):
    from idefix_amd.datablock import DataBlock
    o = make_oracle(kw)
    g = DataBlock(kw["dim"], kw["nx"], xbeg=kw["xbeg"], xend=kw["xend"], mhd=kw["mhd"],
                  comps=kw.get("comps", kw["dim"]),
                  recon=kw["recon"], solver=kw["solver"], emf=kw["emf"], bc=kw["bc"],
                  gamma=kw["gamma"], cfl=kw["cfl"], nstages=kw["nstages"],
                  first_dt=kw["first_dt"], precision=precision, ntracer=kw.get("ntracer", 0),
                  shock_flattening=kw.get("shock_flattening"), csiso=kw.get("csiso"))
    og = OracleGrid(o)
    for d in range(3):
        assert np.array_equal(og.x[d], g.x[d]), "grid coordinates differ"
    Vc, Vs = ic(og)
    o.Vc[...] = Vc
    if kw["mhd"]:
        o.Vs[...] = Vs
    g.upload(Vc, Vs)
    return o, g


def rel_err(a, b):
    """max over cells of |a-b| / max(|b|, 1e-12 max|b|)."""
    a = np.asarray(a, dtype=np.float64)
    b = np.asarray(b, dtype=np.float64)
    scale = max(np.max(np.abs(b)), 1e-300)
    den = np.maximum(np.abs(b), 1e-12 * scale)
    return float(np.max(np.abs(a - b) / den)) if a.size else 0.0


def interior_faces(o, Vs):
    """Interior faces of each Vs component: [beg, end] along its own direction."""
    out = []
    b, e = o.beg, o.end
    for c in range(Vs.shape[0]):
        sl = [slice(b[2], e[2]), slice(b[1], e[1]), slice(b[0], e[0])]
        ax = 2 - c
        sl[ax] = slice(b[c], e[c] + 1)
        out.append(Vs[c][tuple(sl)])
    return out


def compare_state(o, g, rtol=RTOL, what=""):
    Vc_g, Vs_g = g.download()
    errs = {}
    ic_o = o.interior(o.Vc)
    ic_g = o.interior(Vc_g)
    for n in range(o.nvar):
        errs["Vc%d" % n] = rel_err(ic_g[n], ic_o[n])
    if o.cfg.mhd:
        fo = interior_faces(o, o.Vs)
        fg = interior_faces(o, Vs_g)
        for c in range(len(fo)):
            errs["Vs%d" % c] = rel_err(fg[c], fo[c])
    worst = max(errs.values())
    assert worst <= rtol, "%s parity failed: %s" % (what, errs)
    return errs
